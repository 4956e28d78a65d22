// Drives the C++ mirror of the reference interface (mdacp_amd/host) the way
// the reference's Benchmark project drives MDManager (benchmark.cc:13-52):
// inject particles, MakePairList, then Calculate() steps with observations.
//
// usage: shim_driver <state.bin> <steps> <observe_every> <out_prefix> [T tau]
// prints one line per observation: step KE PE virial constructions zeta;
// with T and tau the steps run under the Nose-Hoover heat bath
// (ControlTemperature, AimedTemperature = T, HeatbathTau = tau).
// writes <out_prefix>_q.f64 / _p.f64 (by particle id) at the end.
//        shim_driver save <state.bin> <steps> <dir>
// MakePairList and <steps> steps, then in <dir> the reference's output
// files: SaveAsCdviewSequential (conf0000.cd), SaveConfiguration
// (conf000.cnf) and the step timers' all/force/comm/pair00000.dat.
#include <unistd.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../mdacp_amd/host/mdacp_gpu.h"

int main(int argc, char **argv) {
  const bool save = argc >= 2 && std::strcmp(argv[1], "save") == 0;
  if (save) {
    argv++;
    argc--;
  }
  if (argc < 4 || (!save && argc < 5)) {
    fprintf(stderr, "usage: %s state.bin steps every out_prefix\n", argv[0]);
    return 2;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  int64_t n = 0;
  double L[3], dt;
  if (fread(&n, 8, 1, f) != 1 || fread(L, 8, 3, f) != 3 || fread(&dt, 8, 1, f) != 1) return 2;
  std::vector<double> q(3 * n), p(3 * n);
  if (fread(q.data(), 8, 3 * n, f) != (size_t)(3 * n) || fread(p.data(), 8, 3 * n, f) != (size_t)(3 * n))
    return 2;
  fclose(f);
  const int steps = atoi(argv[2]), every = atoi(argv[3]);
  try {
    mdx::SimulationInfo sinfo(L[0], L[1], L[2], dt);
    if (argc >= 7) {
      sinfo.ControlTemperature = true;
      sinfo.HeatbathType = mdx::HT_NOSEHOOVER;
      sinfo.AimedTemperature = atof(argv[5]);
      sinfo.HeatbathTau = atof(argv[6]);
    }
    mdx::MDUnit unit(&sinfo);
    for (int64_t i = 0; i < n; i++) {
      double x[mdx::D] = {q[3 * i], q[3 * i + 1], q[3 * i + 2], 0};
      double v[mdx::D] = {p[3 * i], p[3 * i + 1], p[3 * i + 2], 0};
      unit.AddParticle(x, v, static_cast<int>(i));
    }
    unit.MakePairList();
    if (save) {
      if (chdir(argv[3]) != 0) return 2;
      unit.EnableStopWatches(0);
      for (int s = 0; s < atoi(argv[2]); s++) unit.Calculate();
      mdx::SaveAsCdviewSequential(unit);
      unit.SaveConfiguration(0);
      return 0;   // the unit's timers write their files as it goes out of scope
    }
    mdx::KineticEnergyObserver ke;
    mdx::PotentialEnergyObserver pe(&sinfo);
    mdx::VirialObserver vir;
    for (int s = 1; s <= steps; s++) {
      unit.Calculate();
      if (s % every == 0) {
        unit.GetMeshList()->RefreshImages(unit.GetVariables());
        printf("%d %.17g %.17g %.17g %d %.17g\n", s, unit.ObserveDouble(&ke),
               unit.ObserveDouble(&pe), unit.ObserveDouble(&vir),
               unit.GetMeshList()->GetNumberOfConstructions(), unit.GetVariables()->Zeta);
      }
    }
    mdx::Variables *vars = unit.GetVariables();
    vars->SyncToHost();
    std::vector<double> qo(3 * n), po(3 * n);
    for (int i = 0; i < vars->GetParticleNumber(); i++)
      for (int d = 0; d < 3; d++) {
        qo[3 * vars->type[i] + d] = vars->q[i][d];
        po[3 * vars->type[i] + d] = vars->p[i][d];
      }
    std::string pre = argv[4];
    FILE *fq = fopen((pre + "_q.f64").c_str(), "wb");
    FILE *fp = fopen((pre + "_p.f64").c_str(), "wb");
    if (!fq || !fp) return 2;
    fwrite(qo.data(), 8, qo.size(), fq);
    fwrite(po.data(), 8, po.size(), fp);
    fclose(fq);
    fclose(fp);
  } catch (const mdx::Error &e) {
    fprintf(stderr, "mdx error %d: %s\n", e.code, e.what());
    return 3;
  }
  return 0;
}
