// confmaker.cpp — Benchmark-mode initial conditions of the reference
// (see confmaker.h).  Built with the reference's floating-point flags (the
// same contraction of a + b c into one rounding), so positions and momenta
// match a reference run bit for bit.
#include "confmaker.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <stdexcept>

namespace mdx {

UnitGrid::UnitGrid(int num_procs, int num_threads) {
  if (num_procs < 1 || num_threads < 1) throw std::invalid_argument("unit grid: counts must be >= 1");
  // halving rounds, y z x for processes and x z y for threads
  // (parainfo.cc:27-58)
  int g[3] = {1, 1, 1}, l[3] = {1, 1, 1};
  for (int n = num_procs; n > 1;) {
    const int order[3] = {1, 2, 0};
    for (int k = 0; k < 3 && n > 1; k++) {
      g[order[k]] *= 2;
      n /= 2;
    }
  }
  for (int n = num_threads; n > 1;) {
    const int order[3] = {0, 2, 1};
    for (int k = 0; k < 3 && n > 1; k++) {
      l[order[k]] *= 2;
      n /= 2;
    }
  }
  // the reference rejects counts that do not fill the grid (parainfo.cc:71-90)
  if (g[0] * g[1] * g[2] != num_procs || l[0] * l[1] * l[2] != num_threads)
    throw std::invalid_argument("unit grid: counts must be powers of two");
  for (int d = 0; d < 3; d++) size[d] = g[d] * l[d];
  // unit id = thread + process x threads at global grid position
  // (process position x local size + thread position), parainfo.cc:146-172
  for (int i = 0; i < num_procs; i++) {
    const int gp[3] = {i % g[0], (i / g[0]) % g[1], i / g[0] / g[1]};
    for (int j = 0; j < num_threads; j++) {
      const int lp[3] = {j % l[0], (j / l[0]) % l[1], j / l[0] / l[1]};
      x.push_back(gp[0] * l[0] + lp[0]);
      y.push_back(gp[1] * l[1] + lp[1]);
      z.push_back(gp[2] * l[2] + lp[2]);
    }
  }
}

void ExtractConfiguration(int id, const double s[3], const double w[3], double density,
                          std::vector<double> &q) {
  // lattice with at least the brick's share of sites (confmaker.cc:184-190)
  const double a0 = 1.0 / std::pow(density * 0.25, 1.0 / 3.0);
  const int nx = static_cast<int>(static_cast<unsigned long>(w[0] / a0)) + 1;
  const int ny = static_cast<int>(static_cast<unsigned long>(w[1] / a0)) + 1;
  const int nz = static_cast<int>(static_cast<unsigned long>(w[2] / a0)) + 1;
  const int n_full = nx * ny * nz * 4;
  const int n_exact = static_cast<int>(w[0] * w[1] * w[2] * density);
  // distinct sites to delete, drawn in order (confmaker.cc:191-199); a
  // repeat is redrawn
  std::vector<int> order;
  std::vector<char> taken(static_cast<size_t>(n_full), 0);
  std::mt19937 mt(id);
  std::uniform_real_distribution<double> ud(0.0, 1.0);
  while (static_cast<int>(order.size()) < n_full - n_exact) {
    const int v = static_cast<int>(static_cast<unsigned long>(n_full * ud(mt)));
    if (!taken[v]) {
      taken[v] = 1;
      order.push_back(v);
    }
  }
  // sites in z, y, x order, four per cubic cell (confmaker.cc:201-253)
  const double a = w[0] / nx;
  const double hs = a * 0.5;
  int site = 0;
  for (int iz = 0; iz < nz; iz++) {
    for (int iy = 0; iy < ny; iy++) {
      for (int ix = 0; ix < nx; ix++) {
        const double bx = s[0] + static_cast<double>(ix) * a;
        const double by = s[1] + static_cast<double>(iy) * a;
        const double bz = s[2] + static_cast<double>(iz) * a;
        const double cand[4][3] = {{bx, by, bz},
                                   {bx, s[1] + static_cast<double>(iy) * a + hs,
                                    s[2] + static_cast<double>(iz) * a + hs},
                                   {s[0] + static_cast<double>(ix) * a + hs, by,
                                    s[2] + static_cast<double>(iz) * a + hs},
                                   {s[0] + static_cast<double>(ix) * a + hs,
                                    s[1] + static_cast<double>(iy) * a + hs, bz}};
        for (int b = 0; b < 4; b++, site++) {
          if (taken[site]) continue;
          q.push_back(cand[b][0]);
          q.push_back(cand[b][1]);
          q.push_back(cand[b][2]);
        }
      }
    }
  }
}

void InitialVelocity(double V0, int id, int n, std::vector<double> &p) {
  // variables.cc:87-113: direction uniform on the sphere, then the unit's
  // mean velocity removed
  std::mt19937 mt(id);
  std::uniform_real_distribution<double> ud(0.0, 1.0);
  const size_t base = p.size();
  p.resize(base + 3 * static_cast<size_t>(n));
  double *v = p.data() + base;
  for (int i = 0; i < n; i++) {
    const double z = ud(mt) * 2.0 - 1.0;
    const double phi = ud(mt) * 2.0 * M_PI;
    v[3 * i] = V0 * std::sqrt(1 - z * z) * std::cos(phi);
    v[3 * i + 1] = V0 * std::sqrt(1 - z * z) * std::sin(phi);
    v[3 * i + 2] = V0 * z;
  }
  double av[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) av[d] += v[3 * i + d];
  for (int d = 0; d < 3; d++) av[d] /= static_cast<double>(n);
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) v[3 * i + d] -= av[d];
}

BenchmarkState MakeBenchmarkState(int num_procs, int num_threads, double unit_length,
                                  double density, double V0) {
  const UnitGrid grid(num_procs, num_threads);
  BenchmarkState st;
  // L = UnitLength x grid (simulationinfo.cc:35-39); each unit a brick of
  // width L / grid (mdunit.cc:17-28)
  for (int d = 0; d < 3; d++) st.L[d] = unit_length * grid.size[d];
  for (int id = 0; id < grid.units(); id++) {
    const int pos[3] = {grid.x[id], grid.y[id], grid.z[id]};
    double s[3], w[3];
    for (int d = 0; d < 3; d++) {
      const double ul = st.L[d] / static_cast<double>(grid.size[d]);
      s[d] = ul * static_cast<double>(pos[d]);
      w[d] = (s[d] + ul) - s[d];   // MDRect::GetWidth of [s, s + ul)
    }
    const size_t before = st.q.size() / 3;
    ExtractConfiguration(id, s, w, density, st.q);
    const int n = static_cast<int>(st.q.size() / 3 - before);
    InitialVelocity(V0, id, n, st.p);
    st.unit_of.insert(st.unit_of.end(), static_cast<size_t>(n), id);
  }
  return st;
}

}  // namespace mdx

extern "C" int mdacp_benchmark_state(int num_procs, int num_threads, double unit_length,
                                     double density, double v0, double *q3, double *p3, long cap,
                                     long *n, double *L3) {
  if (!n || cap < 0 || (cap > 0 && (!q3 || !p3)) || !(unit_length > 0.0) || !(density > 0.0))
    return -1;
  mdx::BenchmarkState st;
  try {
    st = mdx::MakeBenchmarkState(num_procs, num_threads, unit_length, density, v0);
  } catch (const std::exception &) {
    return -1;
  }
  *n = static_cast<long>(st.q.size() / 3);
  if (L3)
    for (int d = 0; d < 3; d++) L3[d] = st.L[d];
  const long m = std::min(cap, *n);
  std::copy(st.q.begin(), st.q.begin() + 3 * m, q3 ? q3 : st.q.data());
  std::copy(st.p.begin(), st.p.begin() + 3 * m, p3 ? p3 : st.p.data());
  return 0;
}
