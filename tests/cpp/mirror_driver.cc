// Uses the C++ mirror of the reference interface (mdacp_amd/host) the way
// the reference's own projects and observers use MDUnit and MeshList:
// particles enter through an Executor (mdunit.h:22-25), observers read the
// half list through MeshList's CSR getters and Variables' host arrays with
// the ghosts (observer.cc:18-79), MDUnit::ChangeScale rescales the box
// (mdunit.cc:217-222, as MDManager::ChangeScale drives it, mdmanager.cc:
// 572-584) and SortParticle reorders the particles (meshlist.cc:140-173).
//
// usage: mirror_driver observe <state.bin>
//          PE and virial from the half list, pair counts, CSR checks
//        mirror_driver scale <state.bin> <alpha>
//          ChangeScale(alpha), then PE/N, virial/N and the canonical pair count
//        mirror_driver run <state.bin> <steps> <sort 0|1> <out_prefix>
//          <steps> steps with SortParticle on or off; final q, p by type
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../mdacp_amd/host/mdacp_gpu.h"

namespace {

struct State {
  int64_t n = 0;
  double L[3] = {0, 0, 0}, dt = 0.001;
  std::vector<double> q, p;
};

bool Read(const char *path, State &s) {
  FILE *f = fopen(path, "rb");
  if (!f) return false;
  bool ok = fread(&s.n, 8, 1, f) == 1 && fread(s.L, 8, 3, f) == 3 && fread(&s.dt, 8, 1, f) == 1;
  s.q.resize(3 * s.n);
  s.p.resize(3 * s.n);
  ok = ok && fread(s.q.data(), 8, 3 * s.n, f) == (size_t)(3 * s.n) &&
       fread(s.p.data(), 8, 3 * s.n, f) == (size_t)(3 * s.n);
  fclose(f);
  return ok;
}

// particles enter the unit as the reference's configuration makers do
class Loader : public mdx::Executor {
 public:
  explicit Loader(const State &s) : st(s) {}
  void Execute(mdx::MDUnit *mdu) override {
    for (int64_t i = 0; i < st.n; i++) {
      double x[mdx::D] = {st.q[3 * i], st.q[3 * i + 1], st.q[3 * i + 2], 0.0};
      double v[mdx::D] = {st.p[3 * i], st.p[3 * i + 1], st.p[3 * i + 2], 0.0};
      mdu->AddParticle(x, v, static_cast<int>(i));
    }
  }
  const State &st;
};

// potential energy and virial summed over the half list, a pair with a
// ghost counted half (it is seen from both of its real ends)
class HalfListPotential : public mdx::DoubleObserver {
 public:
  double Observe(mdx::Variables *vars, mdx::MeshList *mesh) override {
    const double CL2 = mdx::CUTOFF_LENGTH * mdx::CUTOFF_LENGTH;
    const double C0 = vars->GetC0(), C2 = vars->GetC2();
    const int pn = vars->GetParticleNumber();
    const int pairs = mesh->GetPairNumber();
    int(*kp)[2] = mesh->GetKeyPartnerPairs();
    double pe = 0.0, w = 0.0;
    for (int k = 0; k < pairs; k++) {
      const int a = kp[k][mdx::MeshList::KEY], b = kp[k][mdx::MeshList::PARTNER];
      double r2 = 0.0;
      for (int d = 0; d < 3; d++) {
        const double dd = vars->q[b][d] - vars->q[a][d];
        r2 += dd * dd;
      }
      if (r2 > CL2) continue;
      const double r6 = r2 * r2 * r2;
      const double weight = (a >= pn || b >= pn) ? 0.5 : 1.0;
      pe += weight * 4.0 * (1.0 / (r6 * r6) - 1.0 / r6 + C2 * r2 + C0);
      w += weight * ((24.0 * r6 - 48.0) / (r6 * r6 * r2) + 8.0 * C2) * r2;
    }
    virial = w / 3.0;
    return pe;
  }
  double virial = 0.0;
};

// CSR consistency: key_pointer is the prefix sum of number_of_partners,
// sorted_list holds the partners of key_partner_pairs in order, and every
// key is the smaller index of its pair.  Returns the number of violations.
int CheckCsr(mdx::MeshList *mesh, int tn) {
  int bad = 0, pos = 0, k = 0;
  int(*kp)[2] = mesh->GetKeyPartnerPairs();
  const int *sl = mesh->GetSortedList();
  const int *kpp = mesh->GetKeyPointerP();
  const int *np = mesh->GetNumberOfPartners();
  for (int i = 0; i < tn; i++) {
    if (kpp[i] != pos || mesh->GetKeyPointer(i) != pos) bad++;
    if (np[i] != mesh->GetPartnerNumber(i)) bad++;
    for (int m = 0; m < np[i]; m++, k++) {
      if (kp[k][0] != i || kp[k][1] != sl[k] || !(i < sl[k])) bad++;
    }
    pos += np[i];
  }
  if (pos != mesh->GetPairNumber()) bad++;
  return bad;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s observe|scale|run state.bin ...\n", argv[0]);
    return 2;
  }
  const std::string mode = argv[1];
  State st;
  if (!Read(argv[2], st)) return 2;
  try {
    mdx::SimulationInfo sinfo(st.L[0], st.L[1], st.L[2], st.dt);
    mdx::ParaInfo pinfo;
    if (mode == "run" && argc >= 5) sinfo.SortParticle = atoi(argv[4]) != 0;
    mdx::MDUnit unit(0, &sinfo, &pinfo);
    Loader loader(st);
    unit.Execute(&loader);
    mdx::ParticleNumberObserver pno;
    if (unit.IntegerDouble(&pno) != st.n) return 4;
    mdx::Variables *vars = unit.GetVariables();
    mdx::MeshList *mesh = unit.GetMeshList();
    if (mode == "observe") {
      unit.AdjustPeriodicBoundary();   // the positions are inside: a no-op
      unit.MakePairList();
      mesh->RefreshImages(vars);
      vars->SyncToHost();
      HalfListPotential pot;
      const double pe = unit.ObserveDouble(&pot);
      const int tn = unit.GetTotalParticleNumber();
      printf("pe %.17g\nvirial %.17g\npairs %d\ncanonical %ld\ntotal %d\ncsr_bad %d\n", pe,
             pot.virial, mesh->GetPairNumber(), mesh->GetCanonicalPairNumber(vars), tn,
             CheckCsr(mesh, tn));
      unit.SetTotalParticleNumber(unit.GetParticleNumber());   // images dropped
      printf("total_after_drop %d\n", unit.GetTotalParticleNumber());
    } else if (mode == "scale") {
      if (argc < 4) return 2;
      const double alpha = atof(argv[3]);
      unit.MakePairList();
      // MDManager::ChangeScale (mdmanager.cc:572-584)
      for (int d = 0; d < 3; d++) sinfo.L[d] *= alpha;
      unit.ChangeScale(alpha);
      unit.MakePairList();
      mesh->RefreshImages(vars);
      vars->SyncToHost();
      HalfListPotential pot;
      const double pe = unit.ObserveDouble(&pot);
      const double n = static_cast<double>(unit.GetParticleNumber());
      printf("pe_per_particle %.17g\nvirial_per_particle %.17g\ncanonical %ld\n"
             "potential_per_particle %.17g\nbox %.17g %.17g %.17g\n",
             pe / n, pot.virial / n, mesh->GetCanonicalPairNumber(vars), unit.PotentialEnergy(),
             unit.GetRect()->GetWidth(mdx::X), unit.GetRect()->GetWidth(mdx::Y),
             unit.GetRect()->GetWidth(mdx::Z));
    } else if (mode == "run") {
      if (argc < 6) return 2;
      const int steps = atoi(argv[3]);
      for (int s = 0; s < steps; s++) unit.Calculate();
      vars->SyncToHost();
      std::vector<double> qo(3 * st.n), po(3 * st.n);
      for (int i = 0; i < vars->GetParticleNumber(); i++)
        for (int d = 0; d < 3; d++) {
          qo[3 * vars->type[i] + d] = vars->q[i][d];
          po[3 * vars->type[i] + d] = vars->p[i][d];
        }
      const std::string pre = argv[5];
      FILE *fq = fopen((pre + "_q.f64").c_str(), "wb");
      FILE *fp = fopen((pre + "_p.f64").c_str(), "wb");
      if (!fq || !fp) return 2;
      fwrite(qo.data(), 8, qo.size(), fq);
      fwrite(po.data(), 8, po.size(), fp);
      fclose(fq);
      fclose(fp);
      printf("constructions %d\n", mesh->GetNumberOfConstructions());
    } else {
      return 2;
    }
  } catch (const mdx::Error &e) {
    fprintf(stderr, "mdx error %d: %s\n", e.code, e.what());
    return 3;
  }
  return 0;
}
