// ref_harness.cc — drives the reference's own CPU path (compiled from the
// reference tree by oracle/Makefile into oracle/_ref/) on an injected state.
//
// TEST INFRASTRUCTURE ONLY: it writes golden vectors for tests/golden/ and
// times the reference for bench.py's cpu_baseline.  It links the reference's
// MDManager / MDUnit / Variables / MeshList / ForceCalculator objects
// unchanged and reaches them only through their public interface:
//   * particles are injected through MDManager::ExecuteAll with an Executor
//     (mdunit.h:22-25) calling MDUnit::AddParticle (mdunit.cc:61-65), the
//     particle id travelling in type[] (migration keeps it, mdunit.cc:103);
//   * the pairlist and forces are read through a DoubleObserver
//     (observer.h:13-16), the one hook that receives MeshList*;
//   * forces are ForceCalculator::CalculateForce (fcalculator.cc:33-50) run
//     on p = 0 with TimeStep = 1, p restored afterwards.
//
// usage: ref_harness <mode> <state.bin> <out_prefix> [args]
//   pairs      one domain: canonical pair set, forces, PE, virial, KE
//   forces     any number of OpenMP domains: forces, PE, virial, KE
//   run K M    K steps of MDManager::Calculate, observables every M steps,
//              final state; with MDX_REF_HEATBATH=NoseHoover:T:tau in the
//              environment the steps run under the reference's Nose-Hoover
//              heat bath (ControlTemperature, simulationinfo.cc:11-23) and
//              Variables::Zeta is recorded with the observables
//   bench K W  W warm-up steps then K steps, each timed alone: atom-steps/s,
//              and the steps that rebuilt the pairlist and their time
//   save K     MakePairList and K steps, then in <out_prefix> (a directory)
//              MDManager::SaveAsCdviewSequential (conf0000.cd) and
//              MDManager::SaveConfiguration (conf000.cnf)
#include <mpi.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <fstream>
#include <string>
#include <vector>

#include "fcalculator.h"
#include "mdmanager.h"
#include "observer.h"

namespace {

struct State {
  long n = 0;
  double L[3] = {0, 0, 0};
  double dt = 0.001;
  std::vector<double> q, p;
};

bool ReadState(const char *path, State &s) {
  FILE *f = fopen(path, "rb");
  if (!f) return false;
  int64_t n = 0;
  bool ok = fread(&n, 8, 1, f) == 1 && fread(s.L, 8, 3, f) == 3 && fread(&s.dt, 8, 1, f) == 1;
  s.n = n;
  s.q.resize(3 * n);
  s.p.resize(3 * n);
  ok = ok && fread(s.q.data(), 8, 3 * n, f) == (size_t)(3 * n);
  ok = ok && fread(s.p.data(), 8, 3 * n, f) == (size_t)(3 * n);
  fclose(f);
  return ok;
}

template <class T>
void WriteArray(const std::string &path, const std::vector<T> &v) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) {
    fprintf(stderr, "cannot write %s\n", path.c_str());
    exit(2);
  }
  fwrite(v.data(), sizeof(T), v.size(), f);
  fclose(f);
}

class Loader : public Executor {
 public:
  explicit Loader(const State *s) : st(s) {}
  void Execute(MDUnit *mdu) override {
    for (long i = 0; i < st->n; i++) {
      double x[D] = {st->q[3 * i], st->q[3 * i + 1], st->q[3 * i + 2], 0.0};
      double v[D] = {st->p[3 * i], st->p[3 * i + 1], st->p[3 * i + 2], 0.0};
      mdu->AddParticle(x, v, static_cast<int>(i));
    }
  }
  const State *st;
};

// Collects q, p of the real particles by id.
class StateCollector : public DoubleObserver {
 public:
  explicit StateCollector(long n) : q(3 * n), p(3 * n) {}
  double Observe(Variables *vars, MeshList *) override {
    const int pn = vars->GetParticleNumber();
    for (int i = 0; i < pn; i++) {
      const int t = vars->type[i];
      for (int d = 0; d < 3; d++) {
        q[3 * t + d] = vars->q[i][d];
        p[3 * t + d] = vars->p[i][d];
      }
    }
    return 0.0;
  }
  std::vector<double> q, p;
};

// Forces through the reference kick: p := 0, CalculateForce at TimeStep 1.
class ForceCollector : public DoubleObserver {
 public:
  ForceCollector(long n, SimulationInfo *si) : f(3 * n), sinfo(si) {}
  double Observe(Variables *vars, MeshList *mesh) override {
    const int pn = vars->GetParticleNumber();
    const int tn = vars->GetTotalParticleNumber();
    std::vector<double> save(static_cast<size_t>(tn) * D);
    for (int i = 0; i < tn; i++)
      for (int d = 0; d < D; d++) {
        save[i * D + d] = vars->p[i][d];
        vars->p[i][d] = 0.0;
      }
    ForceCalculator::CalculateForce(vars, mesh, sinfo);
    for (int i = 0; i < pn; i++) {
      const int t = vars->type[i];
      for (int d = 0; d < 3; d++) f[3 * t + d] = vars->p[i][d];
    }
    for (int i = 0; i < tn; i++)
      for (int d = 0; d < D; d++) vars->p[i][d] = save[i * D + d];
    return 0.0;
  }
  std::vector<double> f;
  SimulationInfo *sinfo;
};

// Variables::Zeta summed over the units (one unit: its zeta)
class ZetaReader : public DoubleObserver {
 public:
  double Observe(Variables *vars, MeshList *) override { return vars->Zeta; }
};

class ConstructionCounter : public DoubleObserver {
 public:
  double Observe(Variables *, MeshList *mesh) override {
    return static_cast<double>(mesh->GetNumberOfConstructions());
  }
};

// One domain only: rebuild the image bookkeeping exactly as
// MDManager::MakePairList creates the images (mdmanager.cc:395-401,
// mdunit.cc:141-192, mdrect.cc:24-47), verify every image position
// bitwise, then emit the canonical pair set.
class PairCollector : public DoubleObserver {
 public:
  PairCollector(SimulationInfo *si, MDRect *r) : sinfo(si), rect(r) {}
  double Observe(Variables *vars, MeshList *mesh) override {
    const int pn = vars->GetParticleNumber();
    const int tn = vars->GetTotalParticleNumber();
    std::vector<int> src(tn), shx(tn, 0), shy(tn, 0), shz(tn, 0);
    for (int i = 0; i < pn; i++) src[i] = i;
    int cur = pn;
    const int dsh[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int dir = 0; dir < MAX_DIR; dir++) {
      std::vector<int> found;
      for (int i = 0; i < cur; i++)
        if (rect->IsInsideEdge(dir, vars->q[i], sinfo)) found.push_back(i);
      for (size_t k = 0; k < found.size(); k++) {
        const int j = found[k];
        const int g = cur + static_cast<int>(k);
        if (g >= tn) {
          bad++;
          continue;
        }
        src[g] = src[j];
        shx[g] = shx[j] + dsh[dir][0];
        shy[g] = shy[j] + dsh[dir][1];
        shz[g] = shz[j] + dsh[dir][2];
        const double diff[3] = {dsh[dir][0] * sinfo->L[X], dsh[dir][1] * sinfo->L[Y],
                                dsh[dir][2] * sinfo->L[Z]};
        for (int d = 0; d < 3; d++)
          if (vars->q[g][d] != vars->q[j][d] + diff[d]) bad++;
      }
      cur += static_cast<int>(found.size());
    }
    if (cur != tn) bad++;
    int(*kp)[2] = mesh->GetKeyPartnerPairs();
    const int np = mesh->GetPairNumber();
    raw_pairs = np;
    triples.reserve(3 * static_cast<size_t>(np));
    for (int k = 0; k < np; k++) {
      const int a = kp[k][0], b = kp[k][1];
      const int ta = vars->type[src[a]], tb = vars->type[src[b]];
      int rel[3] = {shx[b] - shx[a], shy[b] - shy[a], shz[b] - shz[a]};
      if (ta > tb) {
        for (int d = 0; d < 3; d++) rel[d] = -rel[d];
      }
      const int code = (rel[0] + 1) + 3 * (rel[1] + 1) + 9 * (rel[2] + 1);
      triples.push_back(std::min(ta, tb));
      triples.push_back(std::max(ta, tb));
      triples.push_back(code);
    }
    // sort + unique over triples
    const size_t m = triples.size() / 3;
    std::vector<size_t> idx(m);
    for (size_t i = 0; i < m; i++) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](size_t x, size_t y) {
      for (int d = 0; d < 3; d++) {
        if (triples[3 * x + d] != triples[3 * y + d]) return triples[3 * x + d] < triples[3 * y + d];
      }
      return false;
    });
    std::vector<int32_t> out;
    for (size_t k = 0; k < m; k++) {
      const size_t i = idx[k];
      if (!out.empty()) {
        const size_t l = out.size() - 3;
        if (out[l] == triples[3 * i] && out[l + 1] == triples[3 * i + 1] &&
            out[l + 2] == triples[3 * i + 2])
          continue;
      }
      out.push_back(triples[3 * i]);
      out.push_back(triples[3 * i + 1]);
      out.push_back(triples[3 * i + 2]);
    }
    triples.swap(out);
    ghosts = tn - pn;
    return 0.0;
  }
  SimulationInfo *sinfo;
  MDRect *rect;
  std::vector<int32_t> triples;
  long raw_pairs = 0;
  int ghosts = 0;
  int bad = 0;
};

class RectGetter : public Executor {
 public:
  void Execute(MDUnit *mdu) override { rect = *mdu->GetRect(); }
  MDRect rect;
};

std::string WriteConfig(const State &s) {
  char path[] = "/tmp/mdacp_ref_cfgXXXXXX";
  int fd = mkstemp(path);
  if (fd < 0) exit(2);
  FILE *f = fdopen(fd, "w");
  fprintf(f, "Mode=Harness\n");
  fprintf(f, "SystemSizeX=%.17g\nSystemSizeY=%.17g\nSystemSizeZ=%.17g\n", s.L[0], s.L[1], s.L[2]);
  fprintf(f, "TimeStep=%.17g\n", s.dt);
  const char *hb = getenv("MDX_REF_HEATBATH");
  double T = 1.0, tau = 0.1;
  if (hb && sscanf(hb, "NoseHoover:%lf:%lf", &T, &tau) == 2) {
    fprintf(f, "ControlTemperature=yes\nHeatbathType=NoseHoover\n");
    fprintf(f, "AimedTemperature=%.17g\nHeatbathTau=%.17g\n", T, tau);
  } else {
    fprintf(f, "ControlTemperature=no\n");
  }
  fclose(f);
  return path;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s pairs|forces|run|bench state.bin out_prefix [args]\n", argv[0]);
    return 2;
  }
  const std::string mode = argv[1];
  State st;
  if (!ReadState(argv[2], st)) {
    fprintf(stderr, "cannot read %s\n", argv[2]);
    return 2;
  }
  const std::string out = argv[3];
  const std::string cfg = WriteConfig(st);
  std::string a0 = "mdacp", a1 = "-i", a2 = cfg;
  char *fake[] = {&a0[0], &a1[0], &a2[0], nullptr};
  int fargc = 3;
  char **fargv = fake;
  MDManager *mdm = new MDManager(fargc, fargv);
  unlink(cfg.c_str());
  if (!mdm->IsValid()) {
    fprintf(stderr, "invalid decomposition\n");
    delete mdm;
    return 3;
  }
  Loader loader(&st);
  mdm->ExecuteAll(&loader);
  if (static_cast<long>(mdm->GetTotalParticleNumber()) != st.n) {
    fprintf(stderr, "particle count mismatch %lu != %ld\n", mdm->GetTotalParticleNumber(), st.n);
    delete mdm;
    return 3;
  }
  int grid_size[D];
  mdm->GetGridSize(grid_size);
  SimulationInfo sinfo_unit(*mdm->GetParameter(), grid_size);
  sinfo_unit.TimeStep = 1.0;
  const double N = static_cast<double>(st.n);
  int rc = 0;

  if (mode == "pairs" || mode == "forces") {
    mdm->MakePairList();
    if (mode == "pairs") {
      if (mdm->GetTotalUnits() != 1) {
        fprintf(stderr, "pairs mode needs one domain (OMP_NUM_THREADS=1)\n");
        delete mdm;
        return 3;
      }
      RectGetter rg;
      mdm->ExecuteAll(&rg);
      PairCollector pc(&sinfo_unit, &rg.rect);
      mdm->ObserveDouble(&pc);
      if (pc.bad) {
        fprintf(stderr, "image bookkeeping mismatch (%d)\n", pc.bad);
        rc = 4;
      }
      WriteArray(out + "_pairs.i32", pc.triples);
      printf("raw_pairs %ld\nunique_pairs %zu\nghosts %d\n", pc.raw_pairs, pc.triples.size() / 3,
             pc.ghosts);
    }
    ForceCollector fc(st.n, &sinfo_unit);
    mdm->ObserveDouble(&fc);
    WriteArray(out + "_forces.f64", fc.f);
    std::vector<double> obs = {mdm->KineticEnergy() * N, mdm->PotentialEnergy() * N,
                               mdm->ConfigurationTemperature() * N};
    WriteArray(out + "_obs.f64", obs);
    printf("ke %.17g\npe %.17g\nvirial %.17g\n", obs[0], obs[1], obs[2]);
  } else if (mode == "run") {
    const int K = argc > 4 ? atoi(argv[4]) : 1000;
    const int M = argc > 5 ? atoi(argv[5]) : 100;
    ConstructionCounter cc;
    ZetaReader zr;
    std::vector<double> series;  // step, KE, PE, virial, constructions
    std::vector<double> zeta;
    for (int i = 0; i < K; i++) {
      mdm->Calculate();
      if ((i + 1) % M == 0) {
        // observables with images at the current positions
        mdm->SendBorderParticles();
        series.push_back(i + 1);
        series.push_back(mdm->KineticEnergy() * N);
        series.push_back(mdm->PotentialEnergy() * N);
        series.push_back(mdm->ConfigurationTemperature() * N);
        series.push_back(mdm->ObserveDouble(&cc));
        zeta.push_back(mdm->ObserveDouble(&zr));
      }
    }
    StateCollector sc(st.n);
    mdm->ObserveDouble(&sc);
    WriteArray(out + "_series.f64", series);
    WriteArray(out + "_zeta.f64", zeta);
    WriteArray(out + "_q.f64", sc.q);
    WriteArray(out + "_p.f64", sc.p);
    printf("steps %d\n", K);
  } else if (mode == "save") {
    const int K = argc > 4 ? atoi(argv[4]) : 0;
    mdm->MakePairList();
    for (int i = 0; i < K; i++) mdm->Calculate();
    if (chdir(out.c_str()) != 0) {
      fprintf(stderr, "cannot enter %s\n", out.c_str());
      rc = 2;
    } else {
      mdm->SaveAsCdviewSequential();
      mdm->SaveConfiguration();
    }
  } else if (mode == "bench") {
    // Each timed step is clocked alone and the pairlist constructions are
    // read after it (ConstructionCounter sums over the units), so the caller
    // can separate steps that rebuilt from steps that did not and amortize
    // the rebuild cost over any interval.
    const int K = argc > 4 ? atoi(argv[4]) : 20;
    const int W = argc > 5 ? atoi(argv[5]) : 2;
    const double units = static_cast<double>(mdm->GetTotalUnits());
    ConstructionCounter cc;
    mdm->MakePairList();
    for (int i = 0; i < W; i++) mdm->Calculate();
    double built = mdm->ObserveDouble(&cc) / units;
    double sec = 0.0, sec_rebuild = 0.0;
    int rebuild_steps = 0;
    for (int i = 0; i < K; i++) {
      const auto t0 = std::chrono::steady_clock::now();
      mdm->Calculate();
      const auto t1 = std::chrono::steady_clock::now();
      const double dt = std::chrono::duration<double>(t1 - t0).count();
      const double now = mdm->ObserveDouble(&cc) / units;
      sec += dt;
      if (now > built) {
        rebuild_steps++;
        sec_rebuild += dt;
      }
      built = now;
    }
    printf("{\"steps\": %d, \"seconds\": %.6f, \"atom_steps_per_s\": %.6e, \"units\": %d, "
           "\"threads\": %d, \"rebuild_steps\": %d, \"seconds_rebuild_steps\": %.6f}\n",
           K, sec, N * K / sec, mdm->GetTotalUnits(), omp_get_max_threads(), rebuild_steps,
           sec_rebuild);
  } else {
    fprintf(stderr, "unknown mode %s\n", mode.c_str());
    rc = 2;
  }
  delete mdm;
  return rc;
}
