// confmaker.h — the reference's Benchmark-mode initial conditions, for one
// system holding every domain unit of a reference run.
//
// The reference splits the box into units (MPI processes x OpenMP threads,
// ParaInfo, parainfo.cc:16-172), fills each unit's brick with
// ExtractConfigurationMaker (confmaker.cc:174-255: an FCC lattice with
// randomly deleted sites, the deletions drawn by std::mt19937 seeded with
// the unit id) and gives it velocities with Variables::SetInitialVelocity
// (variables.cc:87-113: fixed speed V0, isotropic direction, the same
// std::mt19937 seed, unit mean removed).  The functions here restate those
// steps with the same <random> engines and distributions, so one GPU can
// start from the exact state a reference run of P processes x T threads
// starts from, the units concatenated in id order.
#ifndef MDACP_CONFMAKER_H
#define MDACP_CONFMAKER_H

#include <vector>

namespace mdx {

// ParaInfo's unit grid (parainfo.cc:16-63, SetGrid :146-172): the grid
// size and each unit id's position in it
struct UnitGrid {
  UnitGrid(int num_procs, int num_threads);
  int size[3];
  std::vector<int> x, y, z;   // by unit id
  int units() const { return static_cast<int>(x.size()); }
};

// ExtractConfigurationMaker::Execute for the unit brick [s, s + w): the
// positions it adds, in the order it adds them (xyz triples)
void ExtractConfiguration(int id, const double s[3], const double w[3], double density,
                          std::vector<double> &q);

// Variables::SetInitialVelocity for n particles of unit id (xyz triples)
void InitialVelocity(double V0, int id, int n, std::vector<double> &p);

// Benchmark-mode state of a whole reference run: box L = unit_length x the
// unit grid, every unit filled and given velocities, concatenated by unit id.
struct BenchmarkState {
  double L[3];
  std::vector<double> q, p;       // [n][3]
  std::vector<int> unit_of;       // unit id of each particle
};
BenchmarkState MakeBenchmarkState(int num_procs, int num_threads, double unit_length,
                                  double density, double V0);

}  // namespace mdx

extern "C" {
// C entry for harnesses that are not C++ (ctypes): writes up to cap
// particles of MakeBenchmarkState into q3/p3 ([cap][3]) and the box into L3;
// *n receives the full count (call with cap = 0 to size the buffers).
// Returns 0, or -1 on a bad argument.
int mdacp_benchmark_state(int num_procs, int num_threads, double unit_length, double density,
                          double v0, double *q3, double *p3, long cap, long *n, double *L3);
}

#endif
