// ref_benchmark.cc — drives the reference's own Benchmark-mode pieces
// (compiled from the reference tree by oracle/Makefile into oracle/_ref/).
//
// TEST INFRASTRUCTURE ONLY: it writes golden vectors for tests/golden/.  It
// links the reference's MDManager / MDUnit / ExtractConfigurationMaker
// unchanged and reaches them only through their public interface.  The
// domain units are MPI processes x OpenMP threads (mpirun -np P with
// OMP_NUM_THREADS=T); "conf" takes one process.
//
// usage: ref_benchmark conf <input.cfg> <out.bin>
//          ExtractConfigurationMaker (confmaker.cc:174-255) and
//          MDManager::SetInitialVelocity (variables.cc:87-113) on every unit,
//          as Benchmark::Run starts (benchmark.cc:15-19); writes, by unit
//          id: int32 id, int32 n, then n x (q xyz, p xyz) doubles
//        ref_benchmark run <input.cfg> <out.txt>
//          then the loop of Benchmark::Run (benchmark.cc:24-41): the
//          thermalizing steps and the timed loop's observations, written as
//          "time T P E" lines at full precision
#include <mpi.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>

#include <map>
#include <string>
#include <vector>

#include "confmaker.h"
#include "mdmanager.h"

namespace {

class Dumper : public Executor {
 public:
  void Execute(MDUnit *mdu) override {
    Variables *v = mdu->GetVariables();
    std::vector<double> rec;
    const int n = v->GetParticleNumber();
    for (int i = 0; i < n; i++) {
      for (int d = 0; d < 3; d++) rec.push_back(v->q[i][d]);
      for (int d = 0; d < 3; d++) rec.push_back(v->p[i][d]);
    }
#pragma omp critical
    units[mdu->GetID()].swap(rec);
  }
  std::map<int, std::vector<double>> units;
};

}  // namespace

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s conf|run input.cfg out\n", argv[0]);
    return 2;
  }
  const std::string mode = argv[1];
  std::string a0 = "mdacp", a1 = "-i", a2 = argv[2];
  char *fake[] = {&a0[0], &a1[0], &a2[0], nullptr};
  int fargc = 3;
  char **fargv = fake;
  MDManager *mdm = new MDManager(fargc, fargv);
  if (!mdm->IsValid()) {
    fprintf(stderr, "invalid decomposition\n");
    delete mdm;
    return 3;
  }
  Parameter *param = mdm->GetParameter();
  ExtractConfigurationMaker c(param);
  mdm->ExecuteAll(&c);
  mdm->SetInitialVelocity(param->GetDoubleDef("InitialVelocity", 1.0));
  int rc = 0;
  if (mode == "conf") {
    Dumper d;
    mdm->ExecuteAll(&d);
    FILE *f = fopen(argv[3], "wb");
    if (!f) return 2;
    for (auto &u : d.units) {
      const int32_t id = u.first, n = static_cast<int32_t>(u.second.size() / 6);
      fwrite(&id, 4, 1, f);
      fwrite(&n, 4, 1, f);
      fwrite(u.second.data(), sizeof(double), u.second.size(), f);
    }
    fclose(f);
  } else if (mode == "run") {
    const int T_LOOP = param->GetIntegerDef("ThermalizeLoop", 150);
    const int LOOP = param->GetIntegerDef("TotalLoop", 1000);
    const int OBSERVE_LOOP = param->GetIntegerDef("ObserveLoop", 100);
    int rank = 0;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    FILE *f = rank == 0 ? fopen(argv[3], "w") : fopen("/dev/null", "w");
    if (!f) return 2;
    fprintf(f, "# N = %lu\n", mdm->GetTotalParticleNumber());
    mdm->MakePairList();
    for (int i = 0; i < T_LOOP; i++) mdm->Calculate();
    for (int i = 0; i < LOOP; i++) {
      mdm->Calculate();
      if (i % OBSERVE_LOOP == 0) {
        const double t = mdm->GetSimulationTime();
        const double T = mdm->Temperature();
        const double P = mdm->Pressure();
        const double E = mdm->TotalEnergy();
        fprintf(f, "%.17g %.17g %.17g %.17g\n", t, T, P, E);
      }
    }
    fclose(f);
  } else {
    rc = 2;
  }
  delete mdm;
  return rc;
}
