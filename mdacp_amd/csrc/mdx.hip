// mdx.hip — MI355X (gfx950) short-range Lennard-Jones step of MDACP.
//
// Device-resident state of one periodic domain (all in HBM):
//   pos   double4[n + g]  cell-ordered; reals in [0, n), images in [n, n+g)
//   mom   double4[n]      momenta of the reals (same order)
//   tag   int[n]          particle id of each real
//   gsrc  int[g], gcode   image -> its real (cell-ordered index) and shift
//   scan_in/cnt_in, scan_out/cnt_out over the (ms+2)^3 padded cell grid:
//         interior cells hold reals only, the outer layer images only
//   tiles BlkTile[]       <= kBlk reals of a block of 3 x 3 cell rows, with
//         the slots of their staged partners in sidx and the real of each
//         list column in treal
//   list16 u16[tile][K][kBlk]  full neighbour list of every real (partners
//         real or image), as 24 x the partner's index in its tile's LDS
//         staging of interleaved (x, y, z); nn[n]
//
// The reference keeps q/p as AoS double[N][4] on the host, materialises
// images with an MPI halo exchange and builds a half list on the CPU
// (mdmanager.cc:388-489, meshlist.cc:68-415).  Here images are formed on
// the device from the same rule (mdrect.cc:24-47, mdunit.cc:153-192), the
// cell sort is a counting sort with LDS-staged scans, and the force kernel
// is reactionless over a full list: each pair is evaluated from both
// sides, nothing is written by two lanes, and results are deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mdx.h"
#include "mdx_common.h"
#include "mdx_tiles.h"

namespace {

using namespace mdxk;
constexpr int kScanItems = 4;                       // items per thread in a scan block
constexpr int kScanTile = kBlock * kScanItems;      // 1024 elements per scan block

enum StatusBits : int {
  ST_INVALID_POS = 1,
  ST_GHOST_OVERFLOW = 2,
  ST_LIST_OVERFLOW = 4,
  ST_CELL_INDEX = 8,
  ST_STAGE_OVERFLOW = 16,
  ST_TILE_OVERFLOW = 32,
};

// Small block of device results read back by the host in one copy.
struct DevStatus {
  int flags;
  int ghost_total;
  int ntiles;
  int cand_count;
  int pair_hit;
  int stage_total;                 // staged particles of all tiles
  double vmax2;                    // max |p|^2 over the reals
  double dmax2;                    // max displacement^2 since the last build
  double kmax_d;                   // longest neighbour list (as a double)
  double stage_max;                // largest tile staging (particles)
  long long canon_count;
  double sums[4];                  // KE, PE, virial, spare
};


// Block maximum of a non-negative value written to partial[blockIdx.x];
// every thread of the block must call it.  A one-block kernel folds the
// partials (k_max_partials), so no address sees more than one writer.  It
// ends with a barrier, so calls may follow one another.
__device__ __forceinline__ void block_max_to(double v, double *__restrict__ partial) {
  __shared__ double lds[16];   // up to 1024 threads
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = lds[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, lds[w]);
    partial[blockIdx.x] = m;
  }
  __syncthreads();
}

// Block maximum of an int, returned to every thread; every thread calls it.
__device__ __forceinline__ int block_max_int(int v) {
  __shared__ int lds[16];   // up to 1024 threads
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  int m = lds[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = max(m, lds[w]);
  __syncthreads();
  return m;
}

__global__ void k_max_partials(const double *__restrict__ partial, int nb, double *__restrict__ out) {
  __shared__ double lds[kBlock];
  double m = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) m = fmax(m, partial[b]);
  lds[threadIdx.x] = m;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = lds[0];
}

// Device side of the pairlist velocity budget (pairlist.cc:48-65), so that
// steps can be queued without a host round trip each.  The guard kernel
// ending step b folds |p|max^2 of that step's kick, records it in ring[b],
// and applies the budget update the reference makes at the start of the
// next step.  Once the budget is spent the reference checks displacements
// at the start of every step (pairlist.cc:40-43).  Steps from index
// first_checked on carry that check on the device (k_chk_* below): the
// guard queues it when the budget is surely negative.  Otherwise, if the
// budget could be spent (below a margin that covers rounding of the device
// sqrt) or a particle is too fast, the guard raises halt and every later
// kernel of the batch returns at once; the check raises halt when the list
// expired, and the host then rebuilds it.  The host replays the budget
// exactly from ring[] (steps_done entries).
constexpr int kRing = 64;
struct StepCtl {
  double budget, dt, buffer;
  int halt, steps_done, too_fast, first_checked;
  int check, cand, expired, checks_passed;
  double cd2;
  double ring[kRing];
};

// Maximum of nb non-negative partials by one block of kGuard threads: up
// to kFoldLoads loads per thread go out together (one memory round trip for
// grids of up to kGuard x kFoldLoads blocks), then wavefront and block
// maxima.  Every thread calls it; the result is valid in thread 0.
constexpr int kGuard = 1024;
constexpr int kFoldLoads = 16;
__device__ __forceinline__ double fold_max(const double *__restrict__ partial, int nb) {
  __shared__ double lds[kGuard / 64];
  double m[kFoldLoads];
#pragma unroll
  for (int u = 0; u < kFoldLoads; u++) {
    const int b = u * kGuard + (int)threadIdx.x;
    m[u] = b < nb ? partial[b] : 0.0;
  }
  for (int b = kFoldLoads * kGuard + (int)threadIdx.x; b < nb; b += kGuard) m[0] = fmax(m[0], partial[b]);
#pragma unroll
  for (int w = kFoldLoads / 2; w > 0; w >>= 1)
#pragma unroll
    for (int u = 0; u < w; u++) m[u] = fmax(m[u], m[u + w]);
  const double v = wave_max(m[0]);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kGuard / 64; w++) r = fmax(r, lds[w]);
  return r;
}

// Sum of nb partials in a fixed order by one block of kGuard threads (the
// loads as fold_max); the result is valid in thread 0.
__device__ __forceinline__ double fold_sum(const double *__restrict__ partial, int nb) {
  __shared__ double lds[kGuard / 64];
  double m[kFoldLoads];
#pragma unroll
  for (int u = 0; u < kFoldLoads; u++) {
    const int b = u * kGuard + (int)threadIdx.x;
    m[u] = b < nb ? partial[b] : 0.0;
  }
  double v = 0.0;
#pragma unroll
  for (int u = 0; u < kFoldLoads; u++) v += m[u];
  for (int b = kFoldLoads * kGuard + (int)threadIdx.x; b < nb; b += kGuard) v += partial[b];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kGuard / 64; w++) r += lds[w];
  return r;
}

__device__ __forceinline__ void step_guard(const double *__restrict__ partial, int nb,
                                           StepCtl *__restrict__ ctl) {
  if (ctl->halt) return;
  const double fold = fold_max(partial, nb);
  if (threadIdx.x == 0) {
    const double v2 = fold;
    const int k = ctl->steps_done;
    ctl->ring[k] = v2;
    ctl->steps_done = k + 1;
    double b = ctl->budget;
    if (b > 0.0) {
      const double vmax = sqrt(v2);
      if (vmax * 2.0 * ctl->dt > ctl->buffer) {
        ctl->too_fast = 1;
        ctl->halt = 1;
      }
      b = b - vmax * 2.0 * ctl->dt;
      ctl->budget = b;
    }
    if (k + 1 >= kRing) {
      ctl->halt = 1;
    } else if (b < 1e-9) {
      if (k + 1 >= ctl->first_checked && b < -1e-9 && !ctl->too_fast) ctl->check = 1;
      else ctl->halt = 1;
    }
  }
}

// The displacement check (PairList::CheckByDisplacement, pairlist.cc:68-121)
// at the start of a step queued by the guard, in four kernels: the
// displacement of every real (end of the last step, in pos) and image (its
// mid-step refresh, in mid) from its place at the build, with block maxima;
// the threshold; the candidates; the decision.  Each returns at once unless
// a check is queued and the batch runs.  Arithmetic as the host path's:
// disp2 for the distances, correctly rounded sqrt and no contraction for
// the threshold.
__device__ __forceinline__ bool check_queued(const StepCtl *ctl) { return ctl->check && !ctl->halt; }

// candidate list the device check can hold (CheckListLength of the
// reference configuration, simulationinfo.cc:29, is 500); a larger one
// keeps the check on the host
constexpr int kMaxCheck = 512;

__device__ __forceinline__ double4 check_pos(const double4 *pos, const double4 *mid, int n, int i) {
  return i < n ? pos[i] : mid[i];
}

__global__ void k_chk_disp(const StepCtl *__restrict__ ctl, const double4 *__restrict__ pos,
                           const double4 *__restrict__ mid, const double4 *__restrict__ ref, int n,
                           int tn, double *__restrict__ disp, double *__restrict__ partial) {
  if (!check_queued(ctl)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  if (i < tn) {
    d2 = disp2(ref[i], check_pos(pos, mid, n, i));
    disp[i] = d2;
  }
  block_max_to(d2, partial);
}

__global__ void __launch_bounds__(kGuard) k_chk_fold(StepCtl *__restrict__ ctl,
                                                     const double *__restrict__ partial, int nb) {
#pragma clang fp contract(off)
  if (!check_queued(ctl)) return;
  const double dmax2 = fold_max(partial, nb);
  if (threadIdx.x == 0) {
    const double c_distance = ctl->buffer - __dsqrt_rn(dmax2);
    ctl->cd2 = c_distance * c_distance;
    ctl->cand = 0;
  }
}

__global__ void k_chk_cand(StepCtl *__restrict__ ctl, const double *__restrict__ disp, int tn,
                           int cap, int *__restrict__ cand) {
  if (!check_queued(ctl)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= tn || !(disp[i] > ctl->cd2)) return;
  const int k = atomicAdd(&ctl->cand, 1);
  if (k < cap) cand[k] = i;
}

__global__ void __launch_bounds__(kGuard) k_chk_decide(StepCtl *__restrict__ ctl,
                                                       const int *__restrict__ cand, int cap,
                                                       const double4 *__restrict__ pos,
                                                       const double4 *__restrict__ mid,
                                                       const double4 *__restrict__ ref, int n,
                                                       double SL2, double CL2) {
  __shared__ int hit;
  __shared__ double4 cref[kMaxCheck], cnow[kMaxCheck];
  if (!check_queued(ctl)) return;
  const int l = ctl->cand;
  int expired = 0;
  if (l >= cap) {
    expired = 1;
  } else if (l > 1) {
    // the candidates' places at the build and now, staged once
    for (int a = threadIdx.x; a < l; a += kGuard) {
      cref[a] = ref[cand[a]];
      cnow[a] = check_pos(pos, mid, n, cand[a]);
    }
    if (threadIdx.x == 0) hit = 0;
    __syncthreads();
    for (int t = threadIdx.x; t < l * l; t += kGuard) {
      const int a = t / l, b = t % l;
      if (b <= a) continue;
      if (disp2(cref[a], cref[b]) < SL2) continue;
      if (disp2(cnow[a], cnow[b]) < CL2) hit = 1;
    }
    __syncthreads();
    expired = hit;
  }
  if (threadIdx.x == 0) {
    ctl->check = 0;
    if (expired) {
      ctl->expired = 1;
      ctl->halt = 1;
    } else {
      ctl->checks_passed = ctl->checks_passed + 1;
    }
  }
}

// ---------------------------------------------------------------- scans
// Exclusive scan of an int array, three passes, each block staging a
// 1024-element tile in LDS.  Deterministic.
__device__ int block_exclusive_scan(int v, int *lds, int *total) {
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    int x = (t >= off) ? lds[t - off] : 0;
    __syncthreads();
    lds[t] += x;
    __syncthreads();
  }
  const int incl = lds[t];
  *total = lds[kBlock - 1];
  __syncthreads();
  return incl - v;
}

__global__ void k_scan_tiles(const int *__restrict__ in, int *__restrict__ out,
                             int *__restrict__ tile_sums, int n) {
  __shared__ int lds[kBlock];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
  for (int k = 0; k < kScanItems; k++) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int total;
  int ex = block_exclusive_scan(s, lds, &total);
  for (int k = 0; k < kScanItems; k++) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// one block scans the tile sums in chunks, carrying the running total
__global__ void k_scan_tile_sums(int *__restrict__ sums, int ntiles, int *__restrict__ grand) {
  __shared__ int lds[kBlock];
  int carry = 0;
  for (int base = 0; base < ntiles; base += kBlock) {
    const int i = base + threadIdx.x;
    const int v = i < ntiles ? sums[i] : 0;
    int total;
    const int ex = block_exclusive_scan(v, lds, &total);
    if (i < ntiles) sums[i] = ex + carry;
    carry += total;
  }
  if (threadIdx.x == 0 && grand) *grand = carry;
}

__global__ void k_scan_add(int *__restrict__ out, const int *__restrict__ tile_sums, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += tile_sums[i / kScanTile];
}

// ---------------------------------------------------------------- geometry

// ------------------------------------------------------------ rebuild
// Wrap reals into the box exactly as Variables::AdjustPeriodicBoundary
// (variables.cc:64-84), bin them into interior cells, count their images
// (mdrect.cc:24-47: q in (0, SL) -> image at q+L; q in (L-SL, L) -> q-L).
__global__ void k_wrap_bin_count(double4 *__restrict__ pos, int n, Geom g, double SL,
                                 int *__restrict__ cellid, int *__restrict__ slot,
                                 int *__restrict__ cell_count, int *__restrict__ nimg,
                                 int *__restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 q = pos[i];
  double x[3] = {q.x, q.y, q.z};
  bool bad = false;
  int c[3], cnt = 1;
  for (int d = 0; d < 3; d++) {
    if (x[d] < 0.0) x[d] += g.L[d];
    else if (x[d] > g.L[d]) x[d] -= g.L[d];
    if (!(x[d] >= 0.0 && x[d] <= g.L[d])) bad = true;   // also catches NaN
    c[d] = bad ? 1 : interior_coord(x[d], g.inv_w[d], g.ms[d]);
    const bool left = x[d] < SL && x[d] > 0.0;
    const bool right = x[d] > g.L[d] - SL && x[d] < g.L[d];
    cnt *= (left || right) ? 2 : 1;
  }
  if (bad) atomicOr(flags, ST_INVALID_POS);
  q.x = x[0]; q.y = x[1]; q.z = x[2];
  pos[i] = q;
  const int cl = cell_linear(g, c[0], c[1], c[2]);
  cellid[i] = cl;
  slot[i] = wave_grouped_add(cell_count, cl);
  nimg[i] = cnt - 1;
}


// Variables::AdjustPeriodicBoundary (variables.cc:64-84) alone: q < 0 ->
// q + L, q > L -> q - L per axis, then "Invalid position" if still outside.
__global__ void k_wrap_reals(double4 *__restrict__ pos, int n, Geom g, int *__restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 q = pos[i];
  double x[3] = {q.x, q.y, q.z};
  bool bad = false;
  for (int d = 0; d < 3; d++) {
    if (x[d] < 0.0) x[d] += g.L[d];
    else if (x[d] > g.L[d]) x[d] -= g.L[d];
    if (!(x[d] >= 0.0 && x[d] <= g.L[d])) bad = true;
  }
  if (bad) atomicOr(flags, ST_INVALID_POS);
  q.x = x[0];
  q.y = x[1];
  q.z = x[2];
  pos[i] = q;
}

// BubbleHist::AnalyseSub (cavitation.cc:117-141) for one domain at the
// origin: grid box (ix, iy, iz) = trunc(q * inv) of each real, counted at
// ix + iy gx + iz gx gy; an index outside the grid is the reference's
// failed assert (flagged here)
__global__ void k_density_count(const double4 *__restrict__ pos, int n, double inv, int gx, int gy,
                                int total, int *__restrict__ hist, int *__restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 q = pos[i];
  const int ix = (int)(q.x * inv), iy = (int)(q.y * inv), iz = (int)(q.z * inv);
  const long long k = (long long)ix + (long long)iy * gx + (long long)iz * gx * gy;
  if (k < 0 || k >= total) {
    atomicOr(flags, ST_INVALID_POS);
    return;
  }
  atomicAdd(hist + k, 1);
}

// the counts as the reference's unsigned char cells hold them (modulo 256)
__global__ void k_density_bytes(const int *__restrict__ hist, int total,
                                unsigned char *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < total) out[k] = (unsigned char)(hist[k] & 0xff);
}

// Variables::ChangeScale (variables.cc:138-145): q *= alpha for the reals
__global__ void k_scale_reals(double4 *__restrict__ pos, int n, double alpha) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 q = pos[i];
  q.x *= alpha;
  q.y *= alpha;
  q.z *= alpha;
  pos[i] = q;
}

// Images of each real, in a fixed (sz, sy, sx) enumeration; their cells are
// fixed by the shift (outer layer) so no image can land in an interior cell.
__global__ void k_make_images(const double4 *__restrict__ pos, const int *__restrict__ cellid,
                              const int *__restrict__ nimg, const int *__restrict__ goff, int n,
                              int gcap, Geom g, double SL, double4 *__restrict__ gpos,
                              int *__restrict__ gsrc, int *__restrict__ gcode,
                              int *__restrict__ gcell, int *__restrict__ gslot,
                              int *__restrict__ cell_count, int *__restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || nimg[i] == 0) return;
  const double4 q = pos[i];
  const double x[3] = {q.x, q.y, q.z};
  int sh[3];
  for (int d = 0; d < 3; d++) {
    const bool left = x[d] < SL && x[d] > 0.0;
    const bool right = x[d] > g.L[d] - SL && x[d] < g.L[d];
    sh[d] = left ? 1 : (right ? -1 : 0);
  }
  const int cl = cellid[i];
  const int cx = cl % g.m[0], cy = (cl / g.m[0]) % g.m[1], cz = cl / (g.m[0] * g.m[1]);
  int k = goff[i];
  for (int bz = 0; bz <= (sh[2] != 0); bz++)
    for (int by = 0; by <= (sh[1] != 0); by++)
      for (int bx = 0; bx <= (sh[0] != 0); bx++) {
        if (bx == 0 && by == 0 && bz == 0) continue;
        const int s[3] = {bx ? sh[0] : 0, by ? sh[1] : 0, bz ? sh[2] : 0};
        if (k >= gcap) {
          atomicOr(flags, ST_GHOST_OVERFLOW);
          return;
        }
        double4 r;
        r.x = shifted(x[0], s[0], g.L[0]);
        r.y = shifted(x[1], s[1], g.L[1]);
        r.z = shifted(x[2], s[2], g.L[2]);
        r.w = 0.0;
        gpos[k] = r;
        gsrc[k] = i;
        gcode[k] = (s[0] + 1) + 3 * (s[1] + 1) + 9 * (s[2] + 1);
        const int gx = s[0] > 0 ? g.m[0] - 1 : (s[0] < 0 ? 0 : cx);
        const int gy = s[1] > 0 ? g.m[1] - 1 : (s[1] < 0 ? 0 : cy);
        const int gz = s[2] > 0 ? g.m[2] - 1 : (s[2] < 0 ? 0 : cz);
        const int gcl = cell_linear(g, gx, gy, gz);
        gcell[k] = gcl;
        gslot[k] = wave_grouped_add(cell_count, gcl);
        k++;
      }
}

// split cell counts into interior / outer-layer arrays for two scans
__global__ void k_split_counts(const int *__restrict__ cell_count, int ncell, Geom g,
                               int *__restrict__ cnt_in, int *__restrict__ cnt_out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int x = c % g.m[0], y = (c / g.m[0]) % g.m[1], z = c / (g.m[0] * g.m[1]);
  const bool inner = x >= 1 && x <= g.ms[0] && y >= 1 && y <= g.ms[1] && z >= 1 && z <= g.ms[2];
  cnt_in[c] = inner ? cell_count[c] : 0;
  cnt_out[c] = inner ? 0 : cell_count[c];
}

__global__ void k_cell_start(const int *__restrict__ cnt_in, const int *__restrict__ scan_in,
                             const int *__restrict__ scan_out, int ncell, int n,
                             int *__restrict__ cell_start) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  // a cell is either interior (reals, [0,n)) or outer (images, [n, n+g))
  cell_start[c] = cnt_in[c] > 0 ? scan_in[c] : n + scan_out[c];
}

// place pre-sort index k (reals 0..n-1, images n..n+g-1) into its cell
__global__ void k_scatter(const int *__restrict__ cellid, const int *__restrict__ slot,
                          const int *__restrict__ cell_start, int count, int key_base,
                          int *__restrict__ order) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  order[cell_start[cellid[k]] + slot[k]] = key_base + k;
}

// atomics fixed the cell, not the order inside it: sort each cell's
// segment by (x, pre-sort index), which leaves every cell row sorted by x
// (the list build searches it), deterministically (one wavefront per cell);
// pre-sort index k < n is real k (pos), else image k - n (gpos)
__global__ void k_cell_sort(int *__restrict__ order, const int *__restrict__ cell_start,
                            const int *__restrict__ cell_count, int ncell, int n,
                            const double4 *__restrict__ pos, const double4 *__restrict__ gpos) {
  __shared__ int buf[kBlock / 64][kSortBuf];
  __shared__ double xbuf[kBlock / 64][kSortBuf];
  const int w = threadIdx.x >> 6;
  const auto xof = [&](int k) { return k < n ? pos[k].x : gpos[k - n].x; };
  for (int c = blockIdx.x * (kBlock / 64) + w; c < ncell; c += gridDim.x * (kBlock / 64))
    wave_rank_sort_x(order + cell_start[c], cell_count[c], buf[w], xbuf[w], xof);
}

__global__ void k_gather_reals(const int *__restrict__ order, int n,
                               const double4 *__restrict__ pos_in, const double4 *__restrict__ mom_in,
                               const int *__restrict__ tag_in, double4 *__restrict__ pos,
                               double4 *__restrict__ mom, int *__restrict__ tag,
                               int *__restrict__ inv) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int k = order[s];
  pos[s] = pos_in[k];
  mom[s] = mom_in[k];
  tag[s] = tag_in[k];
  inv[k] = s;
}

__global__ void k_gather_images(const int *__restrict__ order, int n, int g,
                                const double4 *__restrict__ gpos_in, const int *__restrict__ gsrc_in,
                                const int *__restrict__ gcode_in, const int *__restrict__ inv,
                                double4 *__restrict__ pos, int *__restrict__ gsrc,
                                int *__restrict__ gcode) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g) return;
  const int pre = order[n + k] - n;
  pos[n + k] = gpos_in[pre];
  gsrc[k] = inv[gsrc_in[pre]];
  gcode[k] = gcode_in[pre];
}

// ---------------------------------------------------------------- tiles
// Block tiles (mdx_tiles.h) over the whole box: the system's cells from
// cell 0, images at slots n + scan_out.

// XCD-aware tile order: the 8 XCDs each take one contiguous range of tiles
// (workgroup b runs on XCD b % 8), so tiles of neighbouring blocks, which
// stage the same rows, share an L2.
__device__ __forceinline__ int tile_of_block(int b, int ntiles) {
  const int chunk = (ntiles + 7) >> 3;
  return (b & 7) * chunk + (b >> 3);
}
__host__ __device__ inline int tile_grid(int ntiles) { return ((ntiles + 7) >> 3) << 3; }

// tiles of each block of br x br rows, one workgroup per block: each within
// kBlk reals and cap staged particles
__global__ void k_blk_count(const double4 *__restrict__ pos, Geom g, Cells c, int br, int target,
                            int cap, int *__restrict__ blk_nt) {
  __shared__ int hist[kMaxCellsX + 1];
  const int b = blockIdx.x, BY = blk_count(g, 1, br);
  const int by = b % BY, bz = b / BY;
  blk_hist(g, c, by, bz, br, hist);
  const int nt = blk_ntiles(g, c, pos, hist, by, bz, br, target, cap);
  if (threadIdx.x == 0) blk_nt[b] = nt;
}

// The tiles of one block per workgroup, one wavefront per tile; tile_S gets
// each tile's staging size.  A staging above cap raises ST_STAGE_OVERFLOW,
// a tile above kBlk reals ST_TILE_OVERFLOW; block maxima of the stagings
// go to partial.
__global__ void k_blk_tiles(const double4 *__restrict__ pos, Geom g, Cells c, int br,
                            const int *__restrict__ blk_nt, const int *__restrict__ blk_toff,
                            int cap, BlkTile *__restrict__ tiles, BlkInfo *__restrict__ info,
                            int *__restrict__ tile_S, double *__restrict__ partial,
                            int *__restrict__ flags) {
  __shared__ int hist[kMaxCellsX + 1];
  const int b = blockIdx.x, BY = blk_count(g, 1, br);
  const int by = b % BY, bz = b / BY;
  const int nt = blk_nt[b], lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  blk_hist(g, c, by, bz, br, hist);
  double smax = 0.0;
  for (int k = w; k < nt; k += (int)(blockDim.x >> 6)) {
    const int t = blk_toff[b] + k;
    int nreal;
    const int S = blk_make_tile(g, c, pos, hist, by, bz, br, k, nt, 0, tiles + t, info + t, &nreal);
    if (lane == 0) {
      tile_S[t] = S;
      if (S > cap) atomicOr(flags, ST_STAGE_OVERFLOW);
      if (nreal > kBlk) atomicOr(flags, ST_TILE_OVERFLOW);
    }
    smax = fmax(smax, (double)S);
  }
  block_max_to(smax, partial);
}

// sidx of every tile (one workgroup per tile) and the tile's first entry in it
__global__ void k_blk_sidx(Geom g, Cells c, int ntiles, const int *__restrict__ s0,
                           BlkTile *__restrict__ tiles, const BlkInfo *__restrict__ info,
                           int *__restrict__ sidx) {
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  if (threadIdx.x == 0) tiles[t].s0 = s0[t];
  blk_write_sidx(g, c, info + t, sidx + s0[t]);
}

// The neighbour lists of every tile (build_tile_lists), one workgroup per
// tile; block maxima of the list lengths go to partial, a list longer than
// kcap raises ST_LIST_OVERFLOW.
__global__ void __launch_bounds__(kBlk) k_build_tiles(
    const double4 *__restrict__ pos, const BlkTile *__restrict__ tiles,
    const BlkInfo *__restrict__ info, const int *__restrict__ sidx, int ntiles, Geom g, int kcap,
    unsigned short *__restrict__ list_nat, unsigned short *__restrict__ list,
    int *__restrict__ nn, int *__restrict__ treal, unsigned *__restrict__ tmeta,
    double *__restrict__ partial, int *__restrict__ flags) {
  extern __shared__ double stage_mem[];
  const int tile = tile_of_block(blockIdx.x, ntiles);
  if (tile >= ntiles) {
    block_max_to(0.0, partial);
    return;
  }
  const BlkTile t = tiles[tile];
  const size_t region = (size_t)tile * kcap * kBlk;
  const int cnt = build_tile_lists(pos, t, info + tile, sidx + t.s0, g, kcap, list_nat + region,
                                   list + region, nn, treal + (size_t)tile * kBlk,
                                   tmeta + (size_t)tile * kBlk, stage_mem);
  if (cnt > kcap) atomicOr(flags, ST_LIST_OVERFLOW);
  block_max_to((double)cnt, partial);
}

// ---------------------------------------------------------------- step
// ForceCalculator::UpdatePositionHalf (fcalculator.cc:18-30)
__global__ void k_drift(double4 *__restrict__ pos, const double4 *__restrict__ mom, int n,
                        double dt2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  pos[i] = drift_half(pos[i], mom[i], dt2);
}

// image refresh (MDUnit::MakeBufferForBorderParticles, mdunit.cc:181-191)
__global__ void k_refresh_images(double4 *__restrict__ pos, const int *__restrict__ gsrc,
                                 const int *__restrict__ gcode, int n, int g, double Lx, double Ly,
                                 double Lz) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g) return;
  const double L[3] = {Lx, Ly, Lz};
  pos[n + k] = image_of(pos[gsrc[k]], gcode[k], L);
}

// Drift by half a step fused with the image refresh, out of place:
// out[i] = q_i + p_i dt/2 for the reals (fcalculator.cc:26-28) and, for
// image k, the same drifted source plus its shift (mdunit.cc:181-191), so
// the images equal a refresh after the drift bit for bit.  Images are
// also written back into pos, which keeps them at these mid-step positions
// until the next refresh, as the reference's ghosts are.
__global__ void k_drift_refresh(double4 *__restrict__ pos, const double4 *__restrict__ mom,
                                const int *__restrict__ gsrc, const int *__restrict__ gcode, int n,
                                int g, double dt2, double Lx, double Ly, double Lz,
                                double4 *__restrict__ out, const int *__restrict__ halt) {
  if (halt && *halt) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n + g) return;
  const int src = i < n ? i : gsrc[i - n];
  double4 q = drift_half(pos[src], mom[src], dt2);
  if (i >= n) {
    const double L[3] = {Lx, Ly, Lz};
    q = image_of(q, gcode[i - n], L);
  }
  q.w = 0.0;
  out[i] = q;
  if (i >= n) pos[i] = q;
}


// Parameters of the differentiable form V(r) = eps v(r / sigma), v the
// reference's shifted LJ cut at r / sigma = 3 (so the cutoff is 3 sigma).
// With s = r^2 and u = s / sigma^2:
//   phi(s)  = V'(r) / r = (eps / sigma^2) phi0(u),  phi0 = 24 u^-4 - 48 u^-7 + 8 C2
//   phi'(s)             = (eps / sigma^4) phi0'(u), phi0' = -96 u^-5 + 336 u^-8
//   V       = eps v0(u),  v0 = 4 (u^-6 - u^-3 + C2 u + C0)
//   g = r V'(r) = s phi (virial term)
// and the parameter derivatives
//   dphi/deps = phi / eps,   dphi/dsigma = -(2 / sigma) (phi + s phi')
//   dV/deps   = V / eps,     dV/dsigma   = -g / sigma
//   dg/deps   = g / eps,     dg/dsigma   = -(2 / sigma) s (phi + s phi').
// At eps = sigma = 1 the force term equals pair_df exactly.
struct LJ {
  double eps, sigma, inv_sig2, es2, es4, cl2, C0, C2;
};

__device__ __forceinline__ double pair_df_lj(double s, const LJ &p, double dt) {
  const double u = s * p.inv_sig2;
  const double u6 = u * u * u;
  const double df = p.es2 * ((24.0 * u6 - 48.0) * recip(u6 * u6 * u) + 8.0 * p.C2) * dt;
  return s > p.cl2 ? 0.0 : df;
}

// phi, phi' and V of one pair (zero beyond the cutoff)
__device__ __forceinline__ void lj_terms(double s, const LJ &p, double *phi, double *dphi,
                                         double *V) {
  if (s > p.cl2) {
    *phi = *dphi = *V = 0.0;
    return;
  }
  const double u = s * p.inv_sig2;
  const double iu = recip(u);
  const double iu2 = iu * iu, iu3 = iu2 * iu, iu4 = iu2 * iu2;
  const double iu6 = iu3 * iu3, iu7 = iu6 * iu, iu8 = iu4 * iu4;
  *phi = p.es2 * (24.0 * iu4 - 48.0 * iu7 + 8.0 * p.C2);
  *dphi = p.es4 * (-96.0 * iu4 * iu + 336.0 * iu8);
  *V = p.eps * 4.0 * (iu6 - iu3 + p.C2 * u + p.C0);
}

// Force kick over the tile lists: lane per real, loop over its full list
// with partners read from LDS (no reaction, so no write conflicts).
//   MODE 0  p += F dt in place; block max |p|^2 to partial
//   MODE 1  out[i] = F (dt = 1); p untouched
//   MODE 2  MODE 0, plus the trailing half drift out of place:
//           out[i] = q_i + p_i dt/2 (fcalculator.cc:26-28)
//   MODE 3  MODE 2, plus the next step's leading half drift of the reals:
//           qn[i] = out[i] + p_i dt/2
//   SCALED  the pair term of the differentiable form (pair_df_lj)
// Nose-Hoover state on the device: Variables::Zeta (variables.h:29) at the
// end of the last step, and inside a batch of steps zeta after the first
// update of the step running (zeta_mid) with its momentum factor
// exp(-zeta_mid dt/2) (scale)
struct HeatDev {
  double zeta, zeta_mid, scale;
};

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): ten rounds of two
// 32x32 -> 64-bit products and xors, the key bumped by the Weyl constants
__device__ __forceinline__ uint4 philox4x32(uint4 c, uint2 k) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const unsigned long long a = (unsigned long long)0xD2511F53u * c.x;
    const unsigned long long b = (unsigned long long)0xCD9E8D57u * c.z;
    c = make_uint4((unsigned)(b >> 32) ^ c.y ^ k.x, (unsigned)b, (unsigned)(a >> 32) ^ c.w ^ k.y,
                   (unsigned)a);
    k.x += 0x9E3779B9u;
    k.y += 0xBB67AE85u;
  }
  return c;
}

// (0, 1] from 53 bits of two words
__device__ __forceinline__ double uniform53(unsigned a, unsigned b) {
  const unsigned long long m = ((unsigned long long)(a >> 5) << 26) | (unsigned long long)(b >> 6);
  return ((double)m + 1.0) * (1.0 / 9007199254740992.0);
}

// four standard normals of (seed, id, k): two Philox blocks, counter (id,
// k lo, k hi, 0 / 1), key (seed lo, seed hi), Box-Muller on 53-bit uniforms
__device__ __forceinline__ void langevin_normals(unsigned long long seed, unsigned id,
                                                 unsigned long long k, double nrm[4]) {
  const uint2 key = make_uint2((unsigned)seed, (unsigned)(seed >> 32));
  const uint4 w0 = philox4x32(make_uint4(id, (unsigned)k, (unsigned)(k >> 32), 0u), key);
  const uint4 w1 = philox4x32(make_uint4(id, (unsigned)k, (unsigned)(k >> 32), 1u), key);
  const double two_pi = 6.283185307179586;
  {
    const double r = sqrt(-2.0 * log(uniform53(w0.x, w0.y)));
    const double a = two_pi * uniform53(w0.z, w0.w);
    nrm[0] = r * cos(a);
    nrm[1] = r * sin(a);
  }
  {
    const double r = sqrt(-2.0 * log(uniform53(w1.x, w1.y)));
    const double a = two_pi * uniform53(w1.z, w1.w);
    nrm[2] = r * cos(a);
    nrm[3] = r * sin(a);
  }
}

// p += (-gamma p + D n) dt on all four lanes (ForceCalculator::Langevin,
// fcalculator.cc:1001-1017)
__device__ __forceinline__ void langevin_update(double4 &p, const double nrm[4], double gamma,
                                                double D, double dt) {
  p.x += (-gamma * p.x + D * nrm[0]) * dt;
  p.y += (-gamma * p.y + D * nrm[1]) * dt;
  p.z += (-gamma * p.z + D * nrm[2]) * dt;
  p.w += (-gamma * p.w + D * nrm[3]) * dt;
}

// The heat bath a batched kick applies around the force (MDManager::
// CalculateNoseHoover / CalculateLangevin, mdmanager.cc:255-315):
//   Nose-Hoover: p <- s (s p + F dt), s = hd->scale, all four lanes, and
//     the block's kinetic energy of the new momenta into ke_partial;
//   Langevin: p <- p + F dt, then the noise update of kick number k, and
//     the reals' end-of-step positions are the mid-step ones (no trailing
//     half drift).
struct HeatKick {
  int type = 0;   // mdx_heatbath type
  const HeatDev *hd = nullptr;
  double *ke_partial = nullptr;
  const int *tag = nullptr;
  unsigned long long seed = 0, k = 0;
  double gamma = 0.0, D = 0.0;
};

// Fixed-order sum over the block of K values per thread into partial[block
// * K + k]: a butterfly within each wavefront, then the wavefronts' sums in
// order by thread 0 (up to 1024 threads).  Every thread calls it; it ends
// with a barrier.
template <int K>
__device__ __forceinline__ void block_sums_to(const double (&v)[K], double *__restrict__ partial) {
  __shared__ double red[K][16];
  double w[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    w[k] = v[k];
    for (int off = 32; off > 0; off >>= 1) w[k] += __shfl_xor(w[k], off);
  }
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < K; k++) red[k][threadIdx.x >> 6] = w[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < K; k++) {
      double s = red[k][0];
      for (int q = 1; q < (int)(blockDim.x >> 6); q++) s += red[k][q];
      partial[(size_t)blockIdx.x * K + k] = s;
    }
  }
  __syncthreads();
}

// How far ahead of its arithmetic the kick reads a partner from LDS, and
// how many list groups it loads ahead.
constexpr int kKickAhead = 1, kKickPrefetch = 2;

// The force kick over a tile: one lane per list column, the ranked reals of
// the tile; the tile's partners are staged once, and each lane walks its
// whole list (list_groups over [0, nn)).
//   MODE 0: p += F dt;  MODE 1: out = F (dt = 1);  MODE 2: p += F dt and
//   out = q + p dt/2;  MODE 3 (batched step): as 2, plus the next mid-step
//   position qn = out + p dt/2 and the heat bath's epilogue.
// Per-block maxima of |p|^2 go to partial (MODE != 1).
template <int MODE, bool SCALED = false>
__global__ void __launch_bounds__(kBlk) k_kick_tiles(
    const double4 *__restrict__ pos, double4 *__restrict__ mom, const BlkTile *__restrict__ tiles,
    int ntiles, const int *__restrict__ treal, const int *__restrict__ sidx, int kcap,
    const unsigned short *__restrict__ list, const unsigned *__restrict__ tmeta, PairK pk, double dt,
    double CL2, double4 *__restrict__ out, double *__restrict__ partial, LJ lj = LJ{},
    const int *__restrict__ halt = nullptr, double4 *__restrict__ qn = nullptr,
    HeatKick hk = HeatKick{}) {
  extern __shared__ double stage_mem[];
  if (halt && *halt) return;
  const bool nh = MODE == 3 && hk.type == MDX_HEATBATH_NOSE_HOOVER;
  const bool lgv = MODE == 3 && hk.type == MDX_HEATBATH_LANGEVIN;
  const int tile = tile_of_block(blockIdx.x, ntiles);
  if (tile >= ntiles) {
    if (MODE != 1 && threadIdx.x == 0) partial[blockIdx.x] = 0.0;
    if (nh && threadIdx.x == 0) hk.ke_partial[blockIdx.x] = 0.0;
    return;
  }
  const BlkTile t = tiles[tile];
  const int lane = threadIdx.x;
  const bool act = col_active(lane, t.nreal);
  // list column -> real, and the real's staging index and list length:
  // loads of consecutive columns, where the real's own record would be a
  // gather (the columns hold the tile's reals ranked by list length)
  const size_t c = (size_t)tile * kBlk + lane;
  const int i = act ? treal[c] : 0;
  const unsigned meta = act ? tmeta[c] : 0u;
  const int m = (int)(meta >> 16), self = (int)(meta & 0xffffu);
  // the momentum goes out before the staging, which hides it
  double4 pm = make_double4(0.0, 0.0, 0.0, 0.0);
  if (act && MODE != 1) pm = mom[i];
  stage_tile<kBlk>(t.S, sidx + t.s0, stage_mem, [&](int slot) { return pos[slot]; });
  // the real's position from its staged copy (pos[i] bit for bit; w is 0)
  double4 qi = make_double4(0.0, 0.0, 0.0, 0.0);
  if (act) qi = make_double4(stage_mem[3 * self], stage_mem[3 * self + 1], stage_mem[3 * self + 2], 0.0);
  const unsigned short *lst = list + (size_t)tile * kcap * kBlk + (size_t)lane * kILP;
  // kChains independent partner chains per lane: entry k of the list feeds
  // chain k % kChains, and the chains are summed in a fixed order at the
  // end.  A sentinel entry adds exactly +0.0 (the accumulators are never
  // -0.0).
  double ax[kChains], ay[kChains], az[kChains];
#pragma unroll
  for (int u = 0; u < kChains; u++) ax[u] = ay[u] = az[u] = 0.0;
  const double xi = qi.x, yi = qi.y, zi = qi.z;
  if (act) {
    list_groups<kBlk, kKickAhead, kKickPrefetch>(lst, 0, m, stage_mem,
                                                  [&](int u, double xj, double yj, double zj) {
      if (SCALED) {
        const double dx = xj - xi, dy = yj - yi, dz = zj - zi;
        const double df = pair_df_lj(dx * dx + dy * dy + dz * dz, lj, dt);
        ax[u % kChains] += df * dx;
        ay[u % kChains] += df * dy;
        az[u % kChains] += df * dz;
      } else {
        kick_term(xj, yj, zj, xi, yi, zi, pk, CL2, ax[u % kChains], ay[u % kChains],
                  az[u % kChains]);
      }
    });
  }
  double v2 = 0.0, ke = 0.0;
  if (act) {
    double fx = 0.0, fy = 0.0, fz = 0.0;
#pragma unroll
    for (int u = 0; u < kChains; u++) {
      fx += ax[u];
      fy += ay[u];
      fz += az[u];
    }
    if (MODE == 1) {
      out[i] = make_double4(fx, fy, fz, 0.0);
    } else {
      double4 p = pm;
      if (nh) {
        // HeatbathMomenta, kick, HeatbathMomenta (mdmanager.cc:255-287)
        const double sc = hk.hd->scale;
        p.x *= sc;
        p.y *= sc;
        p.z *= sc;
        p.w *= sc;
        p.x += fx;
        p.y += fy;
        p.z += fz;
        p.x *= sc;
        p.y *= sc;
        p.z *= sc;
        p.w *= sc;
        ke = 0.5 * p.x * p.x + 0.5 * p.y * p.y + 0.5 * p.z * p.z + 0.5 * p.w * p.w;
      } else {
        p.x += fx;
        p.y += fy;
        p.z += fz;
        if (lgv) {
          double nrm[4];
          langevin_normals(hk.seed, (unsigned)hk.tag[i], hk.k, nrm);
          langevin_update(p, nrm, hk.gamma, hk.D, dt);
        }
      }
      mom[i] = p;
      v2 = speed2(p);
      if (MODE == 2) out[i] = drift_half(qi, p, dt * 0.5);
      if (MODE == 3) {
        // Langevin steps end at the mid-step positions (no trailing drift)
        const double4 q = lgv ? qi : drift_half(qi, p, dt * 0.5);
        out[i] = q;
        qn[i] = drift_half(q, p, dt * 0.5);
      }
    }
  }
  if (MODE != 1) block_max_to(v2, partial);
  if (nh) {
    const double kv[1] = {ke};
    block_sums_to<1>(kv, hk.ke_partial);
  }
}

// The guard ending a step (block 0) and the image refresh of the next
// step's mid-step buffer (the other blocks, kGuard images each) in one
// launch.  The refresh blocks may read halt before or after block 0 raises
// it: a refresh of a step that will not run is never read.
// Nose-Hoover zeta around a batched step (CalculateNoseHoover,
// mdmanager.cc:255-287): the step's second update from the temperature of
// its new momenta, HeatbathZeta (fcalculator.cc:979-985) with the
// temperature of MDManager::Temperature (mdmanager.cc:531-548), gives
// zeta at the end of the step; the next step's first update sees the same
// momenta, so it adds the same increment.
struct HeatGuard {
  int type = 0;
  const double *ke_partial = nullptr;
  int n = 0;
  double T0 = 0.0, tau = 0.0, dt = 0.0;
  HeatDev *hd = nullptr;
};

__global__ void __launch_bounds__(kGuard) k_guard_refresh(const double *__restrict__ partial, int nb,
                                                          StepCtl *__restrict__ ctl,
                                                          double4 *__restrict__ mid,
                                                          const int *__restrict__ gsrc,
                                                          const int *__restrict__ gcode, int n,
                                                          int g, Geom geom,
                                                          HeatGuard hg = HeatGuard{}) {
  if (blockIdx.x == 0) {
    const bool ran = !ctl->halt;   // read by every thread before the guard writes it
    step_guard(partial, nb, ctl);
    if (ran && hg.type == MDX_HEATBATH_NOSE_HOOVER) {
      const double ke = fold_sum(hg.ke_partial, nb);
      if (threadIdx.x == 0) {
        const double t = ke / (double)hg.n / 1.5;
        const double t1 = (t - hg.T0) / (hg.tau * hg.tau);
        const double zeta = hg.hd->zeta_mid + t1 * (hg.dt * 0.5);
        const double next = zeta + t1 * (hg.dt * 0.5);
        hg.hd->zeta = zeta;
        hg.hd->zeta_mid = next;
        hg.hd->scale = exp(-(hg.dt * 0.5) * next);
      }
    }
    return;
  }
  if (__hip_atomic_load(&ctl->halt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const int k = (blockIdx.x - 1) * kGuard + threadIdx.x;
  if (k >= g) return;
  mid[n + k] = image_of(mid[gsrc[k]], gcode[k], geom.L);
}

// After a batch: the images of pos at the mid-step refresh of its last
// step, which read mid[(steps_done - 1) & 1] (the reference's ghosts keep
// their last refresh until the next one).
__global__ void k_images_from_mid(double4 *__restrict__ pos, const double4 *__restrict__ mid0,
                                  const double4 *__restrict__ mid1, int n, int g,
                                  const StepCtl *__restrict__ ctl) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g) return;
  const double4 *mid = ((ctl->steps_done - 1) & 1) ? mid1 : mid0;
  pos[n + k] = mid[n + k];
}

__global__ void k_vmax(const double4 *__restrict__ mom, int n, double *__restrict__ partial) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v2 = 0.0;
  if (i < n) {
    v2 = speed2(mom[i]);
  }
  block_max_to(v2, partial);
}

// ------------------------------------------------------------ heat baths

// Fixed-order block sum (kBlock threads) to partial[blockIdx.x]
__device__ __forceinline__ void block_sum_to(double v, double *__restrict__ partial) {
  __shared__ double lds[kBlock];
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

// KineticEnergyObserver (observer.cc:5-15) per block: 0.5 p^2 over all
// four lanes
__global__ void k_ke_partials(const double4 *__restrict__ mom, int n, double *__restrict__ partial) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (i < n) {
    const double4 p = mom[i];
    e = 0.5 * p.x * p.x + 0.5 * p.y * p.y + 0.5 * p.z * p.z + 0.5 * p.w * p.w;
  }
  block_sum_to(e, partial);
}

// ForceCalculator::HeatbathZeta (fcalculator.cc:979-985) with the
// temperature MDManager::Temperature = KE / N / 1.5 (mdmanager.cc:531-548)
// of the block sums (folded in a fixed order), or with t given (nb = 0).
__global__ void k_nh_zeta(const double *__restrict__ partial, int nb, int n, double t_given,
                          double T0, double tau, double dt, HeatDev *__restrict__ hd) {
  __shared__ double lds[kBlock];
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) acc += partial[b];
  lds[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double t = nb > 0 ? lds[0] / (double)n / 1.5 : t_given;
    const double t1 = (t - T0) / (tau * tau);
    hd->zeta += t1 * (dt * 0.5);
  }
}

// The first zeta update of a batch's first step (as k_nh_zeta) into
// zeta_mid, with its momentum factor; zeta itself stays the end-of-step value
__global__ void k_nh_begin(const double *__restrict__ partial, int nb, int n, double T0,
                           double tau, double dt, HeatDev *__restrict__ hd) {
  __shared__ double lds[kBlock];
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) acc += partial[b];
  lds[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double t = lds[0] / (double)n / 1.5;
    const double t1 = (t - T0) / (tau * tau);
    const double zm = hd->zeta + t1 * (dt * 0.5);
    hd->zeta_mid = zm;
    hd->scale = exp(-(dt * 0.5) * zm);
  }
}

// ForceCalculator::HeatbathMomenta (fcalculator.cc:987-999), all four lanes
__global__ void k_nh_scale(double4 *__restrict__ mom, int n, double dt,
                           const HeatDev *__restrict__ hd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double exp1 = exp(-(dt * 0.5) * hd->zeta);
  double4 p = mom[i];
  p.x *= exp1;
  p.y *= exp1;
  p.z *= exp1;
  p.w *= exp1;
  mom[i] = p;
}

// ForceCalculator::Langevin (fcalculator.cc:1001-1017) for the k-th kick:
// p += (-gamma p + r) dt on all four lanes, r = hb_D n with n four standard
// normals of (seed, id, k): two Philox blocks, counter (id, k lo, k hi,
// 0 / 1), key (seed lo, seed hi), Box-Muller on 53-bit uniforms.
__global__ void k_langevin(double4 *__restrict__ mom, const int *__restrict__ tag, int n,
                           unsigned long long seed, unsigned long long k, double gamma, double hb_D,
                           double dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double nrm[4];
  langevin_normals(seed, (unsigned)tag[i], k, nrm);
  double4 p = mom[i];
  langevin_update(p, nrm, gamma, hb_D, dt);
  mom[i] = p;
}

// ------------------------------------------------------------ adjoint
template <int K>
__global__ void k_sum_partials_k(const double *__restrict__ partial, int nblocks,
                                 double *__restrict__ out, double scale) {
  __shared__ double lds[K][kBlock];
  const int t = threadIdx.x;
  double acc[K];
  for (int k = 0; k < K; k++) acc[k] = 0.0;
  for (int b = t; b < nblocks; b += kBlock)
    for (int k = 0; k < K; k++) acc[k] += partial[(size_t)b * K + k];
  for (int k = 0; k < K; k++) lds[k][t] = acc[k];
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (t < off)
      for (int k = 0; k < K; k++) lds[k][t] += lds[k][t + off];
    __syncthreads();
  }
  if (t < K) out[t] = lds[t][0] * scale;
}

// Observables of the differentiable form at one state, per tile:
//   GRAD = false: block sums (KE, PE, W), W = sum r V'(r) / 3 over pairs;
//   GRAD = true:  adds the loss sensitivity of this observation,
//     aq_i += cw gW_i + ce gPE_i,   gW_i = -(2/3) sum_j (phi + s phi') d_ij,
//                                   gPE_i = -sum_j phi d_ij = -F_i,
//     ap_i += ck p_i,
//   and block sums (dPE/deps, dPE/dsigma, dW/deps, dW/dsigma).
template <bool GRAD>
__global__ void __launch_bounds__(kBlk) k_lj_observe_tiles(
    const double4 *__restrict__ pos, const double4 *__restrict__ mom,
    const BlkTile *__restrict__ tiles, int ntiles, const int *__restrict__ treal,
    const int *__restrict__ sidx, int kcap, const unsigned short *__restrict__ list,
    const int *__restrict__ nn, LJ lj, double cw, double ce, double ck, double4 *__restrict__ aq,
    double4 *__restrict__ ap, double *__restrict__ partial) {
  extern __shared__ double stage_mem[];
  constexpr int K = GRAD ? 4 : 3;
  const int tile = tile_of_block(blockIdx.x, ntiles);
  if (tile >= ntiles) {
    if (threadIdx.x < K) partial[(size_t)blockIdx.x * K + threadIdx.x] = 0.0;
    return;
  }
  const BlkTile t = tiles[tile];
  double *sx = stage_mem, *sy = stage_mem + 1, *sz = stage_mem + 2;
  const int lane = threadIdx.x;
  const bool act = col_active(lane, t.nreal);
  const int i = act ? treal[(size_t)tile * kBlk + lane] : 0;
  const double4 z4 = make_double4(0.0, 0.0, 0.0, 0.0);
  const double4 qi = act ? pos[i] : z4;
  const double4 p = act ? mom[i] : z4;
  const unsigned short *lst = list + (size_t)tile * kcap * kBlk + (size_t)lane * kILP;
  double acc[K];
  for (int k = 0; k < K; k++) acc[k] = 0.0;
  double pe = 0.0, w = 0.0, dpe_s = 0.0, dw_s = 0.0;
  double gwx = 0.0, gwy = 0.0, gwz = 0.0, gex = 0.0, gey = 0.0, gez = 0.0;
  stage_tile<kBlk>(t.S, sidx + t.s0, stage_mem, [&](int slot) { return pos[slot]; });
  for_each_partner(lst, act ? nn[i] : 0, (unsigned)(kEntryBytes * t.S), [&](int o) {
      const double dx = sx[o] - qi.x, dy = sy[o] - qi.y, dz = sz[o] - qi.z;
      const double s2 = dx * dx + dy * dy + dz * dz;
      double phi, dphi, V;
      lj_terms(s2, lj, &phi, &dphi, &V);
      const double gterm = s2 * phi;
      pe += V;
      w += gterm;
      if (GRAD) {
        const double h = phi + s2 * dphi;
        dpe_s += -gterm;            // sigma dV/dsigma
        dw_s += -2.0 * s2 * h;      // sigma dg/dsigma
        gwx += h * dx;
        gwy += h * dy;
        gwz += h * dz;
        gex += phi * dx;
        gey += phi * dy;
        gez += phi * dz;
      }
  });
  if (act) {
    if constexpr (!GRAD) {
      acc[0] = 0.5 * p.x * p.x + 0.5 * p.y * p.y + 0.5 * p.z * p.z;
      acc[1] = 0.5 * pe;
      acc[2] = 0.5 * w / 3.0;
    } else {
      // dPE/deps = PE/eps, dW/deps = W/eps; sigma terms divided by sigma
      acc[0] = 0.5 * pe / lj.eps;
      acc[1] = 0.5 * dpe_s / lj.sigma;
      acc[2] = 0.5 * w / 3.0 / lj.eps;
      acc[3] = 0.5 * dw_s / 3.0 / lj.sigma;
      double4 a = aq[i];
      a.x += cw * (-2.0 / 3.0) * gwx - ce * gex;
      a.y += cw * (-2.0 / 3.0) * gwy - ce * gey;
      a.z += cw * (-2.0 / 3.0) * gwz - ce * gez;
      aq[i] = a;
      double4 b = ap[i];
      b.x += ck * p.x;
      b.y += ck * p.y;
      b.z += ck * p.z;
      ap[i] = b;
    }
  }
  block_sums_to<K>(acc, partial);
}

// Before the HVP of a step: ext[] = ap + aq dt/2 over reals and images (an
// image carries its source's adjoint, as it carries its source's position).
// ap itself is left alone here (images read it); k_adj_hvp_tiles writes the
// reals' new ap from ext[].
__global__ void k_adj_pre(const double4 *__restrict__ ap, const double4 *__restrict__ aq,
                          const int *__restrict__ gsrc, int n, int g, double dt2,
                          double4 *__restrict__ ext) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n + g) return;
  const int src = i < n ? i : gsrc[i - n];
  double4 a = ap[src];
  const double4 b = aq[src];
  a.x += b.x * dt2;
  a.y += b.y * dt2;
  a.z += b.z * dt2;
  a.w = 0.0;
  ext[i] = a;
}

// Adjoint of the kick p' = p + F(q_h) dt, then of the half drift into q_h:
//   aq_i += dt (J ap)_i,  (J a)_i = sum_j [phi (a_j - a_i) + 2 phi' (d.(a_j - a_i)) d],
//   ap_i += aq_i dt/2,
//   block sums dt sum_i ap_i . dF_i/d(eps, sigma).
// Partner adjoints come from ext[] through the staged slot index.
__global__ void __launch_bounds__(kBlk) k_adj_hvp_tiles(
    const double4 *__restrict__ pos, const BlkTile *__restrict__ tiles, int ntiles,
    const int *__restrict__ treal, const int *__restrict__ sidx, int kcap,
    const unsigned short *__restrict__ list, const int *__restrict__ nn, LJ lj, double dt,
    const double4 *__restrict__ ext, double4 *__restrict__ aq, double4 *__restrict__ ap,
    double *__restrict__ partial) {
  extern __shared__ double stage_mem[];
  const int tile = tile_of_block(blockIdx.x, ntiles);
  if (tile >= ntiles) {
    if (threadIdx.x < 2) partial[(size_t)blockIdx.x * 2 + threadIdx.x] = 0.0;
    return;
  }
  const BlkTile t = tiles[tile];
  double *sx = stage_mem, *sy = stage_mem + 1, *sz = stage_mem + 2;
  const int *tsidx = sidx + t.s0;   // global slot of each staged particle
  const int lane = threadIdx.x;
  const bool act = col_active(lane, t.nreal);
  const int i = act ? treal[(size_t)tile * kBlk + lane] : 0;
  const double4 z4 = make_double4(0.0, 0.0, 0.0, 0.0);
  const double4 qi = act ? pos[i] : z4;
  const double4 ai = act ? ext[i] : z4;
  const unsigned short *lst = list + (size_t)tile * kcap * kBlk + (size_t)lane * kILP;
  double acc[2] = {0.0, 0.0};
  double jx = 0.0, jy = 0.0, jz = 0.0, te = 0.0, ts = 0.0;
  stage_tile<kBlk>(t.S, tsidx, stage_mem, [&](int slot) { return pos[slot]; });
  for_each_partner(lst, act ? nn[i] : 0, (unsigned)(kEntryBytes * t.S), [&](int o) {
      const double dx = sx[o] - qi.x, dy = sy[o] - qi.y, dz = sz[o] - qi.z;
      const double s2 = dx * dx + dy * dy + dz * dz;
      double phi, dphi, V;
      lj_terms(s2, lj, &phi, &dphi, &V);
      const double4 aj = ext[tsidx[o / 3]];
      const double wx = aj.x - ai.x, wy = aj.y - ai.y, wz = aj.z - ai.z;
      const double dw = 2.0 * dphi * (dx * wx + dy * wy + dz * wz);
      jx += phi * wx + dw * dx;
      jy += phi * wy + dw * dy;
      jz += phi * wz + dw * dz;
      const double ad = ai.x * dx + ai.y * dy + ai.z * dz;
      te += phi * ad;                         // eps dF/deps . a
      ts += (phi + s2 * dphi) * ad;           // -(sigma/2) dF/dsigma . a
  });
  if (act) {
    acc[0] = dt * te / lj.eps;
    acc[1] = dt * (-2.0 / lj.sigma) * ts;
    double4 a = aq[i];
    a.x += dt * jx;
    a.y += dt * jy;
    a.z += dt * jz;
    aq[i] = a;
    double4 b = ai;
    b.x += a.x * (dt * 0.5);
    b.y += a.y * (dt * 0.5);
    b.z += a.z * (dt * 0.5);
    ap[i] = b;
  }
  block_sums_to<2>(acc, partial);
}

__global__ void k_scatter4_by_tag(const double4 *__restrict__ src, const int *__restrict__ tag, int n,
                                  double4 *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[tag[i]] = src[i];
}

__global__ void k_gather4_by_tag(const double4 *__restrict__ by_id, const int *__restrict__ tag, int n,
                                 double4 *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = by_id[tag[i]];
}

// ------------------------------------------------------------ observers
// Per tile: KE = sum 0.5 p^2 (observer.cc:5-15), and over the full list
// PE = sum 0.5 * 4 (r^-12 - r^-6 + C2 r^2 + C0) (observer.cc:18-47) and
// virial = sum 0.5 * r V'(r) (observer.cc:50-79); every pair sits in two
// lists, hence the halves.  Fixed-order tree within the block.
__global__ void __launch_bounds__(kBlk) k_observe_tiles(
    const double4 *__restrict__ pos, const double4 *__restrict__ mom,
    const BlkTile *__restrict__ tiles, int ntiles, const int *__restrict__ treal,
    const int *__restrict__ sidx, int kcap, const unsigned short *__restrict__ list,
    const int *__restrict__ nn, double CL2, double C0, double C2, double *__restrict__ partial) {
  extern __shared__ double stage_mem[];
  const int tile = tile_of_block(blockIdx.x, ntiles);
  if (tile >= ntiles) {
    if (threadIdx.x < 3) partial[(size_t)blockIdx.x * 3 + threadIdx.x] = 0.0;
    return;
  }
  const BlkTile t = tiles[tile];
  double *sx = stage_mem, *sy = stage_mem + 1, *sz = stage_mem + 2;
  const int lane = threadIdx.x;
  const bool act = col_active(lane, t.nreal);
  const int i = act ? treal[(size_t)tile * kBlk + lane] : 0;
  const double4 qi = act ? pos[i] : make_double4(0.0, 0.0, 0.0, 0.0);
  const unsigned short *lst = list + (size_t)tile * kcap * kBlk + (size_t)lane * kILP;
  double ke = 0.0, pe = 0.0, vir = 0.0;
  stage_tile<kBlk>(t.S, sidx + t.s0, stage_mem, [&](int slot) { return pos[slot]; });
  for_each_partner(lst, act ? nn[i] : 0, (unsigned)(kEntryBytes * t.S), [&](int o) {
    const double dx = qi.x - sx[o], dy = qi.y - sy[o], dz = qi.z - sz[o];
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 <= CL2) observe_term(r2, C0, C2, pe, vir);
  });
  if (act) {
    // all four lanes, as the reference's observer (the fourth is zero
    // unless the Langevin bath ran)
    const double4 p = mom[i];
    ke = 0.5 * p.x * p.x + 0.5 * p.y * p.y + 0.5 * p.z * p.z + 0.5 * p.w * p.w;
    pe *= 0.5;
    vir *= 0.5;
  }
  const double v[3] = {ke, pe, vir};
  block_sums_to<3>(v, partial);
}

__global__ void k_sum_partials(const double *__restrict__ partial, int nblocks,
                               double *__restrict__ out) {
  __shared__ double lds[3][kBlock];
  const int t = threadIdx.x;
  double s0 = 0, s1 = 0, s2 = 0;
  for (int b = t; b < nblocks; b += kBlock) {
    s0 += partial[3 * (size_t)b];
    s1 += partial[3 * (size_t)b + 1];
    s2 += partial[3 * (size_t)b + 2];
  }
  lds[0][t] = s0;
  lds[1][t] = s1;
  lds[2][t] = s2;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (t < off) {
      lds[0][t] += lds[0][t + off];
      lds[1][t] += lds[1][t + off];
      lds[2][t] += lds[2][t + off];
    }
    __syncthreads();
  }
  if (t < 3) out[t] = lds[t][0];
}

// ------------------------------------------------------------ expiry
// PairList::CheckByDisplacement (pairlist.cc:68-121) in three kernels.
__global__ void k_disp(const double4 *__restrict__ pos, const double4 *__restrict__ ref, int tn,
                       double *__restrict__ disp, double *__restrict__ partial) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  if (i < tn) {
    d2 = disp2(ref[i], pos[i]);
    disp[i] = d2;
  }
  block_max_to(d2, partial);
}

__global__ void k_candidates(const double *__restrict__ disp, int tn, double c_d2, int cap,
                             int *__restrict__ cand, int *__restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= tn || !(disp[i] > c_d2)) return;
  const int k = atomicAdd(count, 1);
  if (k < cap) cand[k] = i;
}

__global__ void k_candidate_pairs(const double4 *__restrict__ pos, const double4 *__restrict__ ref,
                                  const int *__restrict__ cand, int l, double SL2, double CL2,
                                  int *__restrict__ hit) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)l * l) return;
  const int a = (int)(t / l), b = (int)(t % l);
  if (b <= a) return;
  const int i1 = cand[a], i2 = cand[b];
  if (disp2(ref[i1], ref[i2]) < SL2) return;
  if (disp2(pos[i1], pos[i2]) < CL2) atomicOr(hit, 1);
}

// ------------------------------------------------------------ export
__device__ __forceinline__ void partner_id(int j, int n, const int *tag, const int *gsrc,
                                           const int *gcode, int *id, int *code) {
  if (j < n) {
    *id = tag[j];
    *code = kCodeNone;
  } else {
    *id = tag[gsrc[j - n]];
    *code = gcode[j - n];
  }
}

// each pair sits in both lists; emit it from the side with the smaller id.
// List entries map back to global slots through the tile's sidx.
// Counts and offsets are indexed by particle id, so after an exclusive scan
// in id order the triples of id_lo = t occupy one run starting at offset[t],
// and the runs are in ascending id_lo.
// CSR (the reference's half list, meshlist.cc:68-104, 556-583): indices in
// the reference's space, reals by id and image k as n + k; a real-real pair
// from the side with the smaller id, a real-image pair always from the
// real's side (a pair across the box then appears once from each side, as
// the reference's real-ghost pairs do).  The partner index is written.
template <bool WRITE, bool CSR = false>
__global__ void __launch_bounds__(kBlk) k_canonical_tiles(
    const BlkTile *__restrict__ tiles, int ntiles, const int *__restrict__ treal,
    const int *__restrict__ sidx, int n, int kcap, const unsigned short *__restrict__ list,
    const int *__restrict__ nn, const int *__restrict__ tag, const int *__restrict__ gsrc,
    const int *__restrict__ gcode, const int *__restrict__ offset, int *__restrict__ count,
    int *__restrict__ out) {
  const int tile = blockIdx.x;
  if (tile >= ntiles) return;
  const BlkTile t = tiles[tile];
  const int lane = threadIdx.x;
  if (!col_active(lane, t.nreal)) return;
  const int i = treal[(size_t)tile * kBlk + lane];
  const int ti = tag[i];
  const int *tsidx = sidx + t.s0;
  const unsigned short *lst = list + (size_t)tile * kcap * kBlk + (size_t)lane * kILP;
  int cnum = 0;
  int w = WRITE ? offset[ti] : 0;
  for_each_partner(lst, nn[i], (unsigned)(kEntryBytes * t.S), [&](int o3) {
    const int j = tsidx[o3 / 3];
    if (CSR) {
      const int hj = j < n ? tag[j] : j;
      if (j >= n || ti < hj) {
        if (WRITE) out[w++] = hj;
        cnum++;
      }
      return;
    }
    int tj, code;
    partner_id(j, n, tag, gsrc, gcode, &tj, &code);
    if (ti < tj) {
      if (WRITE) {
        out[3 * (size_t)w + 0] = ti;
        out[3 * (size_t)w + 1] = tj;
        out[3 * (size_t)w + 2] = code;
        w++;
      }
      cnum++;
    }
  });
  if (!WRITE) count[ti] = cnum;
}

// Sorts each id's run of triples by (id_hi, code), one thread per id, in
// place (insertion sort; a run holds one particle's partners with a larger
// id, a few dozen).  With the runs already in ascending id_lo this leaves
// the whole array in the canonical (id_lo, id_hi, code) order.
__global__ void k_sort_runs(const int *__restrict__ offset, const int *__restrict__ count, int n,
                            int32_t *__restrict__ tri) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int32_t *run = tri + 3 * (size_t)offset[t];
  const int len = count[t];
  for (int a = 1; a < len; a++) {
    const int32_t lo = run[3 * a], hi = run[3 * a + 1], code = run[3 * a + 2];
    int b = a - 1;
    while (b >= 0 && (run[3 * b + 1] > hi || (run[3 * b + 1] == hi && run[3 * b + 2] > code))) {
      run[3 * (b + 1)] = run[3 * b];
      run[3 * (b + 1) + 1] = run[3 * b + 1];
      run[3 * (b + 1) + 2] = run[3 * b + 2];
      b--;
    }
    run[3 * (b + 1)] = lo;
    run[3 * (b + 1) + 1] = hi;
    run[3 * (b + 1) + 2] = code;
  }
}

__global__ void k_scatter_by_tag(const double4 *__restrict__ src, const int *__restrict__ tag, int n,
                                 double *__restrict__ out3) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 v = src[i];
  const int t = tag[i];
  out3[3 * (size_t)t + 0] = v.x;
  out3[3 * (size_t)t + 1] = v.y;
  out3[3 * (size_t)t + 2] = v.z;
}

__global__ void k_load_state(const double *__restrict__ q3, const double *__restrict__ p3, int n,
                             double4 *__restrict__ pos, double4 *__restrict__ mom,
                             int *__restrict__ tag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  pos[i] = make_double4(q3[3 * (size_t)i], q3[3 * (size_t)i + 1], q3[3 * (size_t)i + 2], 0.0);
  mom[i] = make_double4(p3[3 * (size_t)i], p3[3 * (size_t)i + 1], p3[3 * (size_t)i + 2], 0.0);
  tag[i] = i;
}

__global__ void k_iota(int *__restrict__ v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

inline int nblk(long long n) { return (int)((n + kBlock - 1) / kBlock); }

thread_local std::string g_create_error;

// Staging of mdx_pairs_canonical_noalloc: the canonical triples of at most
// kNoallocPairs pairs (12 MiB), in a module-scope device array that every
// GPU the library runs on holds until the process ends.  Host code never
// passes the array's name to a kernel or a copy (on the host it is a
// stand-in, not a device address): mdx_create looks up each device's own
// address once, with that device current, and keeps it in g_scratch_dev.
// All systems on one GPU share that GPU's array, so allocation-free
// exports that overlap on one GPU race.
constexpr long long kNoallocPairs = MDX_NOALLOC_PAIRS;
__device__ __align__(256) int32_t scratch[3 * kNoallocPairs];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t *g_scratch_dev[kMaxDevices] = {};

// this device's address of `scratch` (the device must be current)
hipError_t scratch_address(int device, int32_t **out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[device] = static_cast<int32_t *>(p);
  }
  *out = g_scratch_dev[device];
  return hipSuccess;
}

}  // namespace

// ====================================================================== host
struct mdx_system {
  mdx_params prm{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  Geom geom{};
  double SL = 0, C0 = 0, C2 = 0;
  int ncell = 0, cell_cap = 0;
  // sizes
  int n = 0, g = 0, cap = 0, gcap = 0, kcap = 0;
  size_t tot_elems = 0;     // length of pos / ref_pos / cellid / ... arrays
  // device buffers
  double4 *pos = nullptr, *pos_alt = nullptr, *pos_mid = nullptr, *pos_tmp = nullptr, *ref_pos = nullptr, *mom = nullptr,
          *mom_tmp = nullptr, *gpos_tmp = nullptr, *vec_out = nullptr;
  int *tag = nullptr, *tag_tmp = nullptr, *gsrc = nullptr, *gcode = nullptr, *gsrc_tmp = nullptr,
      *gcode_tmp = nullptr;
  int *cellid = nullptr, *slot = nullptr, *order = nullptr, *inv = nullptr, *nimg = nullptr,
      *goff = nullptr;
  int *cell_count = nullptr, *cell_start = nullptr, *cnt_in = nullptr, *cnt_out = nullptr,
      *scan_in = nullptr, *scan_out = nullptr, *scan_aux = nullptr;
  int *nn = nullptr;   // padded list length
  // tiles over blocks of cell rows and their 16-bit lists
  BlkTile *tiles = nullptr;
  BlkInfo *tinfo = nullptr;
  int *treal = nullptr;                        // [tiles_cap][kBlk] real of each ranked column
  int *tile_S = nullptr, *tile_s0 = nullptr;   // [tiles_cap] staging sizes, their prefix sums
  unsigned *tmeta = nullptr;   // [tiles_cap][kBlk] per column: staging index | list length << 16
  int *sidx = nullptr;                         // global slot of every staged particle
  size_t sidx_cap = 0;
  int *blk_nt = nullptr, *blk_toff = nullptr;  // [blk_cap] tiles of each block, their prefix
  int blk_cap = 0, blk_level = 0;   // block shape in use (kBlkLadder)
  bool blk_stage_check = false;     // tile counts also bound the staging (after an overflow)
  int ntiles = 0, tiles_cap = 0, stage_max = 0;
  unsigned short *list16 = nullptr;     // ranked columns, as the list walkers read them
  unsigned short *list_nat = nullptr;   // the build's columns, in the tile's row order
  size_t list16_elems = 0;
  size_t lds_limit = 65536;
  int partial_blocks = 0;
  double *disp = nullptr, *partial = nullptr, *pmax = nullptr;
  int *cand = nullptr;
  double *io_buf = nullptr;   // [cap][3] staging for host transfers
  int *i_buf = nullptr;
  DevStatus *dstat = nullptr;
  DevStatus *hstat = nullptr;  // pinned
  // pairlist bookkeeping (PairList, pairlist.h:10-31)
  bool have_state = false, list_valid = false;
  double buffer_length = 0;
  bool vmax_valid = false;
  double vmax2_host = 0.0;          // max |p|^2 after the last kick
  StepCtl *dctl = nullptr, *hctl = nullptr;
  int check_steps = 4;              // device-checked steps queued after a spent budget
  hipEvent_t evk[2 * kRing] = {};
  // heat bath (SimulationInfo::ControlTemperature ..., simulationinfo.cc:11-23)
  mdx_heatbath hb{};           // zeta here is stale; the device holds it
  HeatDev *dhb = nullptr;
  unsigned long long langevin_k = 0;   // Langevin kicks since mdx_set_state
  // timing
  bool timing = false;
  hipEvent_t ev[8] = {};
  double t_build = 0, t_force = 0, t_other = 0;
  long long n_build = 0, n_steps = 0, n_kick = 0;
  double t_kick_only = 0;
  long long n_kick_timed = 0;
  int32_t *scratch_dev = nullptr;   // this device's `scratch` (mdx_pairs_canonical_noalloc)
  std::string err;
};

namespace {

int fail(mdx_system *s, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (s) s->err = buf;
  else g_create_error = buf;
  return code;
}

#define HIPCHK(s, call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail((s), MDX_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));  \
  } while (0)

template <class T>
int dev_alloc(mdx_system *s, T **p, size_t count) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (count == 0) count = 1;
  HIPCHK(s, hipMalloc((void **)p, count * sizeof(T)));
  return MDX_OK;
}

int launch_check(mdx_system *s, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(s, MDX_ERR_HIP, "launch %s: %s", what, hipGetErrorString(e));
  return MDX_OK;
}

int sync_status(mdx_system *s) {
  HIPCHK(s, hipMemcpyAsync(s->hstat, s->dstat, sizeof(DevStatus), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return MDX_OK;
}

// exclusive scan helper (tiles staged in LDS); grand total to *grand if set
int scan(mdx_system *s, const int *in, int *out, int count, int *grand) {
  const int ntiles = (count + kScanTile - 1) / kScanTile;
  if (count > 0) {
    k_scan_tiles<<<ntiles, kBlock, 0, s->stream>>>(in, out, s->scan_aux, count);
    k_scan_tile_sums<<<1, kBlock, 0, s->stream>>>(s->scan_aux, ntiles, grand);
    k_scan_add<<<nblk(count), kBlock, 0, s->stream>>>(out, s->scan_aux, count);
  }
  return launch_check(s, "scan");
}

int ensure_partials(mdx_system *s, int blocks);

int ensure_capacity(mdx_system *s, int n) {
  if (n <= s->cap && s->cap > 0) return MDX_OK;
  int cap = std::max(n, s->prm.capacity);
  int rc;
  if ((rc = dev_alloc(s, &s->mom, cap)) || (rc = dev_alloc(s, &s->mom_tmp, cap)) ||
      (rc = dev_alloc(s, &s->vec_out, cap)) ||
      (rc = dev_alloc(s, &s->tag, cap)) || (rc = dev_alloc(s, &s->tag_tmp, cap)) ||
      (rc = dev_alloc(s, &s->inv, cap)) || (rc = dev_alloc(s, &s->nimg, cap)) ||
      (rc = dev_alloc(s, &s->goff, cap)) || (rc = dev_alloc(s, &s->nn, cap)) ||
      (rc = dev_alloc(s, &s->io_buf, 3 * (size_t)cap)) || (rc = dev_alloc(s, &s->i_buf, cap)))
    return rc;
  s->cap = cap;
  const int tiles = (std::max(cap, s->ncell) + kScanTile - 1) / kScanTile + 1;
  if ((rc = dev_alloc(s, &s->scan_aux, tiles))) return rc;
  return MDX_OK;
}

// Arrays indexed over reals + images.  pos keeps its first n entries (the
// reals) across a reallocation; everything else is rebuilt by the caller.
int ensure_image_capacity(mdx_system *s, int gcap) {
  const size_t tot = (size_t)s->cap + std::max(gcap, s->gcap);
  if (gcap <= s->gcap && tot <= s->tot_elems) return MDX_OK;
  gcap = std::max(gcap, s->gcap);
  double4 *npos = nullptr;
  HIPCHK(s, hipMalloc((void **)&npos, tot * sizeof(double4)));
  if (s->pos) {
    const size_t keep = std::min(tot, s->tot_elems);
    HIPCHK(s, hipMemcpyAsync(npos, s->pos, keep * sizeof(double4), hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    (void)hipFree(s->pos);
  }
  s->pos = npos;
  int rc;
  if ((rc = ensure_partials(s, nblk((long long)tot))) ||
      (rc = dev_alloc(s, &s->ref_pos, tot)) || (rc = dev_alloc(s, &s->pos_alt, tot)) ||
      (rc = dev_alloc(s, &s->pos_tmp, tot)) ||
      (rc = dev_alloc(s, &s->pos_mid, tot)) ||
      (rc = dev_alloc(s, &s->gpos_tmp, gcap)) || (rc = dev_alloc(s, &s->gsrc, gcap)) ||
      (rc = dev_alloc(s, &s->gcode, gcap)) || (rc = dev_alloc(s, &s->gsrc_tmp, gcap)) ||
      (rc = dev_alloc(s, &s->gcode_tmp, gcap)) || (rc = dev_alloc(s, &s->cellid, tot)) ||
      (rc = dev_alloc(s, &s->slot, tot)) || (rc = dev_alloc(s, &s->order, tot)) ||
      (rc = dev_alloc(s, &s->disp, tot)) || (rc = dev_alloc(s, &s->cand, tot)))
    return rc;
  s->gcap = gcap;
  s->tot_elems = tot;
  s->list_valid = false;
  return MDX_OK;
}

int ensure_list_capacity(mdx_system *s, int kcap) {
  // padded by one group of entries, which the kick reads past a list's end
  kcap = (kcap + kILP - 1) / kILP * kILP;   // whole groups per column
  const size_t need = (size_t)std::max(s->ntiles, 1) * kcap * kBlk + (size_t)kILP * kBlk;
  if (need <= s->list16_elems && kcap <= s->kcap) return MDX_OK;
  int rc = dev_alloc(s, &s->list16, need);
  if (rc || (rc = dev_alloc(s, &s->list_nat, need))) return rc;
  s->list16_elems = need;
  s->kcap = kcap;
  return MDX_OK;
}

// dynamic LDS of the tile kernels: the staged particles plus the sentinel
size_t stage_bytes(const mdx_system *s) { return (size_t)(s->stage_max + 1) * 3 * sizeof(double); }

// the largest staging a tile may have: its sentinel's list entry fits 16
// bits, and the staging fits the LDS beside the kernels' own (the list
// build's, the largest)
int stage_cap(const mdx_system *s) {
  const long long lds =
      ((long long)s->lds_limit - kBuildStaticLds) / (3 * (long long)sizeof(double)) - 1;
  return (int)std::min<long long>(kMaxStaged, lds);
}

Cells cells_of(const mdx_system *s) {
  return Cells{s->scan_in, s->cnt_in, s->scan_out, s->cnt_out, 0, s->n};
}

// per-block partial buffers for every grid the library launches
int ensure_partials(mdx_system *s, int blocks) {
  if (blocks <= s->partial_blocks) return MDX_OK;
  int rc;
  if ((rc = dev_alloc(s, &s->pmax, (size_t)blocks)) ||
      (rc = dev_alloc(s, &s->partial, 4 * (size_t)blocks)))
    return rc;
  s->partial_blocks = blocks;
  return MDX_OK;
}

void ev_record(mdx_system *s, int k) {
  if (s->timing) (void)hipEventRecord(s->ev[k], s->stream);
}

float ev_ms(mdx_system *s, int a, int b) {
  float ms = 0.f;
  if (s->timing) (void)hipEventElapsedTime(&ms, s->ev[a], s->ev[b]);
  return ms;
}

// The cell grid of MeshList::ChangeScale (meshlist.cc:41-65) for box L:
// ms = floor(L / SL) interior cells per axis of width L / ms, plus one image
// layer each side; the arrays sized by it grow when it does.  The pairlist
// is void afterwards.
int set_geometry(mdx_system *s, const double L[3]) {
  const double SL = s->SL;
  for (int d = 0; d < 3; d++)
    if (!(L[d] >= 2.0 * SL))
      return fail(s, MDX_ERR_BOX, "box edge %g < 2 x search length %g", L[d], SL);
  if (!(L[0] / SL < (double)kMaxCellsX))
    return fail(s, MDX_ERR_BOX, "box edge %g in x holds more than %d cells", L[0], kMaxCellsX);
  Geom g{};
  for (int d = 0; d < 3; d++) {
    g.L[d] = L[d];
    g.ms[d] = (int)(L[d] / SL);
    const double w = L[d] / (double)g.ms[d];
    g.inv_w[d] = 1.0 / w;
    g.m[d] = g.ms[d] + 2;
  }
  g.SL2 = SL * SL;
  g.CL2 = s->prm.cutoff * s->prm.cutoff;
  const int ncell = g.m[0] * g.m[1] * g.m[2], blocks = blk_count(g, 1, 1) * blk_count(g, 2, 1);
  int rc;
  if (ncell > s->cell_cap) {
    if ((rc = dev_alloc(s, &s->cell_count, ncell)) || (rc = dev_alloc(s, &s->cell_start, ncell)) ||
        (rc = dev_alloc(s, &s->cnt_in, ncell)) || (rc = dev_alloc(s, &s->cnt_out, ncell)) ||
        (rc = dev_alloc(s, &s->scan_in, ncell)) || (rc = dev_alloc(s, &s->scan_out, ncell)))
      return rc;
    s->cell_cap = ncell;
    if (s->cap > 0) {
      const int tiles = (std::max(s->cap, ncell) + kScanTile - 1) / kScanTile + 1;
      if ((rc = dev_alloc(s, &s->scan_aux, tiles))) return rc;
    }
  }
  if (blocks > s->blk_cap) {
    if ((rc = dev_alloc(s, &s->blk_nt, blocks)) || (rc = dev_alloc(s, &s->blk_toff, blocks))) return rc;
    s->blk_cap = blocks;
  }
  s->geom = g;
  s->ncell = ncell;
  for (int d = 0; d < 3; d++) s->prm.L[d] = L[d];
  s->list_valid = false;
  return MDX_OK;
}

int status_error(mdx_system *s) {
  const int f = s->hstat->flags;
  if (f & ST_INVALID_POS) return fail(s, MDX_ERR_INVALID_POS, "Invalid position (non-finite or outside the box)");
  if (f & ST_CELL_INDEX) return fail(s, MDX_ERR_INVALID_POS, "Invalid cell index");
  return MDX_OK;
}

int build_pairlist(mdx_system *s) {
  if (!s->have_state) return fail(s, MDX_ERR_STATE, "no particles loaded");
  const int n = s->n;
  ev_record(s, 0);
  HIPCHK(s, hipMemsetAsync(s->dstat, 0, sizeof(DevStatus), s->stream));
  for (int attempt = 0; attempt < 3; attempt++) {
    HIPCHK(s, hipMemsetAsync(s->cell_count, 0, sizeof(int) * s->ncell, s->stream));
    HIPCHK(s, hipMemsetAsync(&s->dstat->flags, 0, sizeof(int), s->stream));
    // 1. wrap, bin, count images
    k_wrap_bin_count<<<nblk(n), kBlock, 0, s->stream>>>(s->pos, n, s->geom, s->SL, s->cellid,
                                                        s->slot, s->cell_count, s->nimg,
                                                        &s->dstat->flags);
    int rc = scan(s, s->nimg, s->goff, n, &s->dstat->ghost_total);
    if (rc) return rc;
    // 2. images with their cells
    k_make_images<<<nblk(n), kBlock, 0, s->stream>>>(
        s->pos, s->cellid, s->nimg, s->goff, n, s->gcap, s->geom, s->SL, s->gpos_tmp, s->gsrc_tmp,
        s->gcode_tmp, s->cellid + n, s->slot + n, s->cell_count, &s->dstat->flags);
    if ((rc = launch_check(s, "images"))) return rc;
    if ((rc = sync_status(s))) return rc;
    if ((rc = status_error(s))) return rc;
    if (s->hstat->flags & ST_GHOST_OVERFLOW) {
      const int need = s->hstat->ghost_total + s->hstat->ghost_total / 4 + 1024;
      if ((rc = ensure_image_capacity(s, need))) return rc;
      continue;  // positions are already wrapped; redo from a clean count
    }
    s->g = s->hstat->ghost_total;
    break;
  }
  if (s->hstat->flags & ST_GHOST_OVERFLOW) return fail(s, MDX_ERR_CAPACITY, "image capacity");
  const int g = s->g, tn = n + g;
  // 3. counting sort: interior cells -> [0,n), outer cells -> [n, n+g)
  k_split_counts<<<nblk(s->ncell), kBlock, 0, s->stream>>>(s->cell_count, s->ncell, s->geom,
                                                            s->cnt_in, s->cnt_out);
  int rc = scan(s, s->cnt_in, s->scan_in, s->ncell, nullptr);
  if (rc) return rc;
  if ((rc = scan(s, s->cnt_out, s->scan_out, s->ncell, nullptr))) return rc;
  k_cell_start<<<nblk(s->ncell), kBlock, 0, s->stream>>>(s->cnt_in, s->scan_in, s->scan_out,
                                                          s->ncell, n, s->cell_start);
  k_scatter<<<nblk(n), kBlock, 0, s->stream>>>(s->cellid, s->slot, s->cell_start, n, 0, s->order);
  if (g > 0)
    k_scatter<<<nblk(g), kBlock, 0, s->stream>>>(s->cellid + n, s->slot + n, s->cell_start, g, n,
                                                  s->order);
  k_cell_sort<<<std::min(4096, (s->ncell + 3) / 4), kBlock, 0, s->stream>>>(
      s->order, s->cell_start, s->cell_count, s->ncell, n, s->pos, s->gpos_tmp);
  // 4. gather into cell order, from pos / mom / tag into their spare
  // buffers, which then take their place (no copy of the pre-sort state)
  k_gather_reals<<<nblk(n), kBlock, 0, s->stream>>>(s->order, n, s->pos, s->mom, s->tag,
                                                     s->pos_tmp, s->mom_tmp, s->tag_tmp, s->inv);
  if (g > 0)
    k_gather_images<<<nblk(g), kBlock, 0, s->stream>>>(s->order, n, g, s->gpos_tmp, s->gsrc_tmp,
                                                        s->gcode_tmp, s->inv, s->pos_tmp, s->gsrc,
                                                        s->gcode);
  std::swap(s->pos, s->pos_tmp);
  std::swap(s->mom, s->mom_tmp);
  std::swap(s->tag, s->tag_tmp);
  if ((rc = launch_check(s, "sort"))) return rc;
  // 5. tiles over blocks of cell rows, their staging sizes, then the
  // staging index; once a staging has not fitted, the tile counts bound the
  // stagings too (more tiles per block), and smaller blocks follow if even
  // that fails
  const Cells cells = cells_of(s);
  const int cap = stage_cap(s);
  for (;;) {
    const BlkShape shape = kBlkLadder[s->blk_level];
    const int br = shape.rows;
    const int nblocks = blk_count(s->geom, 1, br) * blk_count(s->geom, 2, br);
    k_blk_count<<<nblocks, kBlock, 0, s->stream>>>(s->pos, s->geom, cells, br, shape.target,
                                                  s->blk_stage_check ? cap : INT_MAX, s->blk_nt);
    if ((rc = scan(s, s->blk_nt, s->blk_toff, nblocks, &s->dstat->ntiles))) return rc;
    if ((rc = sync_status(s))) return rc;
    const int ntiles = s->hstat->ntiles;
    if (ntiles > s->tiles_cap) {
      const size_t tc = (size_t)ntiles + ntiles / 8 + 64;
      if ((rc = dev_alloc(s, &s->tiles, tc)) || (rc = dev_alloc(s, &s->tinfo, tc)) ||
          (rc = dev_alloc(s, &s->treal, tc * kBlk)) || (rc = dev_alloc(s, &s->tmeta, tc * kBlk)) ||
          (rc = dev_alloc(s, &s->tile_S, tc)) ||
          (rc = dev_alloc(s, &s->tile_s0, tc)))
        return rc;
      s->tiles_cap = (int)tc;
    }
    if ((rc = ensure_partials(s, std::max(tile_grid(ntiles), nblocks)))) return rc;
    HIPCHK(s, hipMemsetAsync(&s->dstat->flags, 0, sizeof(int), s->stream));
    k_blk_tiles<<<nblocks, kBlock, 0, s->stream>>>(s->pos, s->geom, cells, br, s->blk_nt,
                                                     s->blk_toff, cap, s->tiles, s->tinfo,
                                                     s->tile_S, s->pmax, &s->dstat->flags);
    k_max_partials<<<1, kBlock, 0, s->stream>>>(s->pmax, nblocks, &s->dstat->stage_max);
    if ((rc = scan(s, s->tile_S, s->tile_s0, ntiles, &s->dstat->stage_total))) return rc;
    if ((rc = launch_check(s, "tiles"))) return rc;
    if ((rc = sync_status(s))) return rc;
    if (s->hstat->flags & ST_TILE_OVERFLOW)
      return fail(s, MDX_ERR_CAPACITY, "more than %d reals share one x in a block of cell rows", kBlk);
    if (!(s->hstat->flags & ST_STAGE_OVERFLOW)) {
      s->ntiles = ntiles;
      s->stage_max = (int)s->hstat->stage_max;
      break;
    }
    if (!s->blk_stage_check) {   // first more tiles per block, then smaller blocks
      s->blk_stage_check = true;
      continue;
    }
    if (s->blk_level + 1 >= kBlkLevels)
      return fail(s, MDX_ERR_CAPACITY,
                  "a tile of %d reals of one cell row stages %d particles, more than %d (16-bit "
                  "list offsets, LDS)",
                  shape.target, (int)s->hstat->stage_max, cap);
    s->blk_level++;
  }
  const size_t staged = (size_t)std::max(s->hstat->stage_total, 1);
  if (staged > s->sidx_cap) {
    const size_t sc = staged + staged / 8 + 4096;
    if ((rc = dev_alloc(s, &s->sidx, sc))) return rc;
    s->sidx_cap = sc;
  }
  k_blk_sidx<<<s->ntiles, kBlock, 0, s->stream>>>(s->geom, cells, s->ntiles, s->tile_s0,
                                                   s->tiles, s->tinfo, s->sidx);
  if ((rc = ensure_list_capacity(s, s->kcap))) return rc;
  // 6. neighbour lists, one pass at the current width, again if it overflows
  for (int attempt = 0; attempt < 3; attempt++) {
    HIPCHK(s, hipMemsetAsync(&s->dstat->flags, 0, sizeof(int), s->stream));
    k_build_tiles<<<tile_grid(s->ntiles), kBlk, stage_bytes(s), s->stream>>>(
        s->pos, s->tiles, s->tinfo, s->sidx, s->ntiles, s->geom, s->kcap, s->list_nat, s->list16,
        s->nn, s->treal, s->tmeta, s->pmax, &s->dstat->flags);
    k_max_partials<<<1, kBlock, 0, s->stream>>>(s->pmax, tile_grid(s->ntiles), &s->dstat->kmax_d);
    if ((rc = launch_check(s, "list"))) return rc;
    if ((rc = sync_status(s))) return rc;
    if (!(s->hstat->flags & ST_LIST_OVERFLOW)) break;
    if ((rc = ensure_list_capacity(s, (int)s->hstat->kmax_d + 8))) return rc;
  }
  if (s->hstat->flags & ST_LIST_OVERFLOW) return fail(s, MDX_ERR_CAPACITY, "list capacity");
  // 7. PairList::Init: snapshot of reals and images
  HIPCHK(s, hipMemcpyAsync(s->ref_pos, s->pos, sizeof(double4) * tn, hipMemcpyDeviceToDevice, s->stream));
  ev_record(s, 1);
  s->buffer_length = s->prm.buffer_length;
  s->list_valid = true;
  if (s->timing) {
    HIPCHK(s, hipEventSynchronize(s->ev[1]));
    s->t_build += ev_ms(s, 0, 1);
  }
  s->n_build++;
  return MDX_OK;
}

int kick(mdx_system *s) {
  const int grid = tile_grid(s->ntiles);
  ev_record(s, 2);
  k_kick_tiles<0><<<grid, kBlk, stage_bytes(s), s->stream>>>(
      s->pos, s->mom, s->tiles, s->ntiles, s->treal, s->sidx, s->kcap, s->list16, s->tmeta,
      make_pairk(s->C2, s->prm.dt), s->prm.dt, s->geom.CL2, nullptr, s->pmax);
  ev_record(s, 3);
  s->vmax_valid = false;   // recomputed from p when the expiry check needs it
  s->n_kick++;
  return launch_check(s, "kick");
}

// Up to `want` NVE steps after the expiry check of the first, queued
// without a host round trip.  The first step's drift + image refresh goes
// into pos_alt (images also written back to pos); then per step the kick
// stages the current mid-step buffer (pos_alt / pos_mid, reals and images),
// writes the reals' end-of-step positions into pos and the next step's
// mid-step reals into the other buffer, whose images are then refreshed,
// and the guard checks the budget.  After the batch the images of pos get
// the mid-step refresh of its last step, so pos ends as the reference's
// state after Calculate(): reals at the end of the step, images at their
// mid-step refresh.  Same arithmetic as drift(); refresh(); kick(); drift().
// *ran receives the number of steps the device completed (>= 1).
int run_batch(mdx_system *s, int want, int first_checked, int *ran) {
  const int n = s->n, g = s->g, grid = tile_grid(s->ntiles), tn = n + g;
  const int B = std::max(1, std::min(want, kRing));
  const int Nl = tn > s->prm.check_list_length ? s->prm.check_list_length : tn;
  const double dt = s->prm.dt;
  StepCtl *h = s->hctl;
  h->budget = s->buffer_length;
  h->dt = dt;
  h->buffer = s->prm.buffer_length;
  h->halt = 0;
  h->steps_done = 0;
  h->too_fast = 0;
  // steps from first_checked on check displacements on the device (not at
  // all when the candidate list is longer than the device check holds)
  h->first_checked = Nl <= kMaxCheck ? std::max(1, first_checked) : kRing;
  h->check = 0;
  h->cand = 0;
  h->expired = 0;
  h->checks_passed = 0;
  h->cd2 = 0.0;
  HIPCHK(s, hipMemcpyAsync(s->dctl, h, offsetof(StepCtl, ring), hipMemcpyHostToDevice, s->stream));
  // the heat bath the kicks apply (MDManager::CalculateNoseHoover /
  // CalculateLangevin, mdmanager.cc:255-315)
  HeatKick hk;
  HeatGuard hgd;
  hk.type = hgd.type = s->hb.type;
  if (s->hb.type == MDX_HEATBATH_NOSE_HOOVER) {
    int rc = ensure_partials(s, std::max(grid, nblk(n)));
    if (rc) return rc;
    hk.hd = s->dhb;
    hk.ke_partial = s->partial;
    hgd.ke_partial = s->partial;
    hgd.n = n;
    hgd.T0 = s->hb.aimed_temperature;
    hgd.tau = s->hb.tau;
    hgd.dt = dt;
    hgd.hd = s->dhb;
  } else if (s->hb.type == MDX_HEATBATH_LANGEVIN) {
    hk.tag = s->tag;
    hk.seed = s->hb.seed;
    hk.gamma = s->hb.gamma;
    hk.D = std::sqrt(2.0 * s->hb.gamma * s->hb.aimed_temperature / dt);
  }
  ev_record(s, 4);
  k_drift_refresh<<<nblk(n + g), kBlock, 0, s->stream>>>(s->pos, s->mom, s->gsrc, s->gcode, n, g,
                                                          dt * 0.5, s->geom.L[0], s->geom.L[1],
                                                          s->geom.L[2], s->pos_alt, &s->dctl->halt);
  if (s->hb.type == MDX_HEATBATH_NOSE_HOOVER) {
    // the first step's first zeta update, from the momenta as they stand
    k_ke_partials<<<nblk(n), kBlock, 0, s->stream>>>(s->mom, n, s->partial);
    k_nh_begin<<<1, kBlock, 0, s->stream>>>(s->partial, nblk(n), n, s->hb.aimed_temperature,
                                            s->hb.tau, dt, s->dhb);
  }
  for (int b = 0; b < B; b++) {
    const double4 *cur = (b & 1) ? s->pos_mid : s->pos_alt;
    double4 *nxt = (b & 1) ? s->pos_alt : s->pos_mid;
    if (b >= h->first_checked) {
      // displacement check at the start of step b: reals at the end of step
      // b - 1 (pos), images at its mid-step refresh (its mid-step buffer)
      const double4 *prev = ((b - 1) & 1) ? s->pos_mid : s->pos_alt;
      k_chk_disp<<<nblk(tn), kBlock, 0, s->stream>>>(s->dctl, s->pos, prev, s->ref_pos, n, tn,
                                                      s->disp, s->pmax);
      k_chk_fold<<<1, kGuard, 0, s->stream>>>(s->dctl, s->pmax, nblk(tn));
      k_chk_cand<<<nblk(tn), kBlock, 0, s->stream>>>(s->dctl, s->disp, tn, Nl, s->cand);
      k_chk_decide<<<1, kGuard, 0, s->stream>>>(s->dctl, s->cand, Nl, s->pos, prev, s->ref_pos, n,
                                                s->geom.SL2, s->geom.CL2);
    }
    if (s->timing) (void)hipEventRecord(s->evk[2 * b], s->stream);
    hk.k = s->langevin_k + (unsigned long long)b;   // Langevin kicks since mdx_set_state
    k_kick_tiles<3><<<grid, kBlk, stage_bytes(s), s->stream>>>(
        cur, s->mom, s->tiles, s->ntiles, s->treal, s->sidx, s->kcap, s->list16, s->tmeta,
        make_pairk(s->C2, dt), dt, s->geom.CL2, s->pos, s->pmax, LJ{}, &s->dctl->halt, nxt, hk);
    if (s->timing) (void)hipEventRecord(s->evk[2 * b + 1], s->stream);
    const int refresh_blocks = (g > 0 && b + 1 < B) ? (g + kGuard - 1) / kGuard : 0;
    k_guard_refresh<<<1 + refresh_blocks, kGuard, 0, s->stream>>>(
        s->pmax, grid, s->dctl, nxt, s->gsrc, s->gcode, n, g, s->geom, hgd);
  }
  if (g > 0)
    k_images_from_mid<<<nblk(g), kBlock, 0, s->stream>>>(s->pos, s->pos_alt, s->pos_mid, n, g,
                                                          s->dctl);
  ev_record(s, 5);
  int rc = launch_check(s, "step");
  if (rc) return rc;
  HIPCHK(s, hipMemcpyAsync(h, s->dctl, sizeof(StepCtl), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  *ran = h->steps_done;
  if (*ran < 1 || *ran > B) return fail(s, MDX_ERR_STATE, "step guard ran %d of %d steps", *ran, B);
  if (s->timing) {
    double k = 0.0;
    for (int b = 0; b < *ran; b++) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, s->evk[2 * b], s->evk[2 * b + 1]);
      k += ms;
    }
    const float all = ev_ms(s, 4, 5);
    s->t_force += k;
    s->t_kick_only += k;
    s->n_kick_timed += *ran;
    s->t_other += all - k;
  }
  s->n_kick += *ran;
  if (s->hb.type == MDX_HEATBATH_LANGEVIN) s->langevin_k += (unsigned long long)*ran;
  return MDX_OK;
}

int drift(mdx_system *s) {
  k_drift<<<nblk(s->n), kBlock, 0, s->stream>>>(s->pos, s->mom, s->n, s->prm.dt * 0.5);
  return launch_check(s, "drift");
}

int refresh(mdx_system *s) {
  if (s->g > 0)
    k_refresh_images<<<nblk(s->g), kBlock, 0, s->stream>>>(s->pos, s->gsrc, s->gcode, s->n, s->g,
                                                            s->geom.L[0], s->geom.L[1],
                                                            s->geom.L[2]);
  return launch_check(s, "refresh");
}

// ForceCalculator::HeatbathZeta with the current temperature, on the device
int nh_zeta(mdx_system *s) {
  const int n = s->n;
  int rc = ensure_partials(s, nblk(n));
  if (rc) return rc;
  k_ke_partials<<<nblk(n), kBlock, 0, s->stream>>>(s->mom, n, s->partial);
  k_nh_zeta<<<1, kBlock, 0, s->stream>>>(s->partial, nblk(n), n, 0.0, s->hb.aimed_temperature,
                                         s->hb.tau, s->prm.dt, s->dhb);
  return launch_check(s, "heatbath zeta");
}

int nh_scale(mdx_system *s) {
  k_nh_scale<<<nblk(s->n), kBlock, 0, s->stream>>>(s->mom, s->n, s->prm.dt, s->dhb);
  s->vmax_valid = false;
  return launch_check(s, "heatbath momenta");
}

int langevin(mdx_system *s) {
  const double hb_D = std::sqrt(2.0 * s->hb.gamma * s->hb.aimed_temperature / s->prm.dt);
  k_langevin<<<nblk(s->n), kBlock, 0, s->stream>>>(s->mom, s->tag, s->n, s->hb.seed, s->langevin_k,
                                                   s->hb.gamma, hb_D, s->prm.dt);
  s->langevin_k++;
  s->vmax_valid = false;
  return launch_check(s, "langevin");
}



// PairList::IsPairListExpired (pairlist.cc:26-45)
int expired(mdx_system *s, int *out) {
  if (!s->list_valid) {
    *out = 1;
    return MDX_OK;
  }
  int rc;
  if (s->buffer_length > 0.0) {
    // UpdatePairListValidity (pairlist.cc:48-65): max |p| of the reals
    if (!s->vmax_valid) {
      k_vmax<<<nblk(s->n), kBlock, 0, s->stream>>>(s->mom, s->n, s->pmax);
      k_max_partials<<<1, kBlock, 0, s->stream>>>(s->pmax, nblk(s->n), &s->dstat->vmax2);
      if ((rc = sync_status(s))) return rc;
      s->vmax2_host = s->hstat->vmax2;
      s->vmax_valid = true;
    }
    const double vmax = std::sqrt(s->vmax2_host);
    if (vmax * 2.0 * s->prm.dt > s->prm.buffer_length)
      return fail(s, MDX_ERR_TOO_FAST, "Too fast particles exists. max_velocity = %f", vmax);
    s->buffer_length = s->buffer_length - vmax * 2.0 * s->prm.dt;
  }
  *out = 0;
  if (!(s->buffer_length < 0.0)) return MDX_OK;
  // CheckByDisplacement (pairlist.cc:68-121) over reals and images
  const int tn = s->n + s->g;
  HIPCHK(s, hipMemsetAsync(&s->dstat->cand_count, 0, 2 * sizeof(int), s->stream));
  k_disp<<<nblk(tn), kBlock, 0, s->stream>>>(s->pos, s->ref_pos, tn, s->disp, s->pmax);
  k_max_partials<<<1, kBlock, 0, s->stream>>>(s->pmax, nblk(tn), &s->dstat->dmax2);
  if ((rc = sync_status(s))) return rc;
  const int Nl = tn > s->prm.check_list_length ? s->prm.check_list_length : tn;
  const double c_distance = s->prm.buffer_length - std::sqrt(s->hstat->dmax2);
  const double c_d2 = c_distance * c_distance;
  k_candidates<<<nblk(tn), kBlock, 0, s->stream>>>(s->disp, tn, c_d2, Nl, s->cand, &s->dstat->cand_count);
  if ((rc = sync_status(s))) return rc;
  const int l = s->hstat->cand_count;
  if (l >= Nl) {
    *out = 1;
    return MDX_OK;
  }
  if (l > 1) {
    k_candidate_pairs<<<nblk((long long)l * l), kBlock, 0, s->stream>>>(
        s->pos, s->ref_pos, s->cand, l, s->geom.SL2, s->geom.CL2, &s->dstat->pair_hit);
    if ((rc = sync_status(s))) return rc;
    *out = s->hstat->pair_hit ? 1 : 0;
  }
  return launch_check(s, "expiry");
}

int reduce_observe(mdx_system *s, double *ke, double *pe, double *vir) {
  if (!s->list_valid) return fail(s, MDX_ERR_STATE, "observe before a pairlist exists");
  const int nb = tile_grid(s->ntiles);
  k_observe_tiles<<<nb, kBlk, stage_bytes(s), s->stream>>>(
      s->pos, s->mom, s->tiles, s->ntiles, s->treal, s->sidx, s->kcap, s->list16, s->nn,
      s->geom.CL2, s->C0, s->C2, s->partial);
  k_sum_partials<<<1, kBlock, 0, s->stream>>>(s->partial, nb, s->dstat->sums);
  int rc = launch_check(s, "observe");
  if (rc) return rc;
  if ((rc = sync_status(s))) return rc;
  if (ke) *ke = s->hstat->sums[0];
  if (pe) *pe = s->hstat->sums[1];
  if (vir) *vir = s->hstat->sums[2] / 3.0;
  return MDX_OK;
}


// ---------------------------------------------------------------- adjoint host
// Device scratch of one differentiable-trajectory call, freed on return.
struct AdjScratch {
  mdx_system *s = nullptr;
  std::vector<void *> bufs;
  template <class T>
  int get(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) return fail(s, MDX_ERR_HIP, "hipMalloc(%zu B): %s", count * sizeof(T), hipGetErrorString(e));
    bufs.push_back(*p);
    return MDX_OK;
  }
  ~AdjScratch() {
    for (void *b : bufs) (void)hipFree(b);
  }
};

LJ make_lj(const mdx_system *s, double eps, double sigma) {
  LJ p;
  p.eps = eps;
  p.sigma = sigma;
  p.inv_sig2 = 1.0 / (sigma * sigma);
  p.es2 = eps / (sigma * sigma);
  p.es4 = eps / (sigma * sigma * sigma * sigma);
  const double rc = s->prm.cutoff * sigma;
  p.cl2 = rc * rc;
  p.C0 = s->C0;
  p.C2 = s->C2;
  return p;
}


// one step of the differentiable form: drift + image refresh into mid,
// kick + drift back into pos
int lj_step(mdx_system *s, const LJ &lj, double4 *mid) {
  const int n = s->n, g = s->g, grid = tile_grid(s->ntiles);
  k_drift_refresh<<<nblk(n + g), kBlock, 0, s->stream>>>(s->pos, s->mom, s->gsrc, s->gcode, n, g,
                                                          s->prm.dt * 0.5, s->geom.L[0],
                                                          s->geom.L[1], s->geom.L[2], mid, nullptr);
  k_kick_tiles<2, true><<<grid, kBlk, stage_bytes(s), s->stream>>>(
      mid, s->mom, s->tiles, s->ntiles, s->treal, s->sidx, s->kcap, s->list16, s->tmeta,
      make_pairk(s->C2, s->prm.dt), s->prm.dt, s->geom.CL2, s->pos, s->pmax, lj);
  return launch_check(s, "lj step");
}

// the largest displacement of a real since the last build must stay below
// half the skin, so the list of the checkpoint covers the whole segment
int check_segment_list(mdx_system *s, const LJ &lj) {
  k_disp<<<nblk(s->n), kBlock, 0, s->stream>>>(s->pos, s->ref_pos, s->n, s->disp, s->pmax);
  k_max_partials<<<1, kBlock, 0, s->stream>>>(s->pmax, nblk(s->n), &s->dstat->dmax2);
  int rc = sync_status(s);
  if (rc) return rc;
  const double skin = s->SL - s->prm.cutoff * lj.sigma;
  if (!(2.0 * std::sqrt(s->hstat->dmax2) < skin))
    return fail(s, MDX_ERR_STATE,
                "checkpoint interval too long: a particle moved %g, more than half the skin %g",
                std::sqrt(s->hstat->dmax2), skin);
  return MDX_OK;
}

int lj_trajectory(mdx_system *s, double eps, double sigma, int nsteps, int nobs, int seg,
                  double p_target, double pe_target, bool grad, mdx_lj_result *res,
                  double *dq0) {
  if (!s->have_state) return fail(s, MDX_ERR_STATE, "no state");
  if (nsteps < 1 || nobs < 1 || nobs > nsteps || seg < 1 || !(eps > 0) || !(sigma > 0))
    return fail(s, MDX_ERR_ARG, "bad trajectory arguments");
  if (!(s->prm.cutoff * sigma < s->SL))
    return fail(s, MDX_ERR_ARG, "cutoff 3 sigma must stay inside the search length");
  const LJ lj = make_lj(s, eps, sigma);
  const int n = s->n;
  const int nseg = (nsteps + seg - 1) / seg;
  const double V = s->geom.L[0] * s->geom.L[1] * s->geom.L[2];
  AdjScratch sc;
  sc.s = s;
  double4 *ck_pos, *ck_mom, *fin_q, *fin_p;
  int *ck_tag;
  double *obs;
  int rc;
  if ((rc = sc.get(&ck_pos, (size_t)nseg * n)) || (rc = sc.get(&ck_mom, (size_t)nseg * n)) ||
      (rc = sc.get(&ck_tag, (size_t)nseg * n)) || (rc = sc.get(&obs, 3 * (size_t)nsteps)) ||
      (rc = sc.get(&fin_q, (size_t)n)) || (rc = sc.get(&fin_p, (size_t)n)))
    return rc;
  HIPCHK(s, hipMemsetAsync(obs, 0, sizeof(double) * 3 * nsteps, s->stream));
  const auto t0 = std::chrono::steady_clock::now();
  // ---- forward, a list and a checkpoint every seg steps
  for (int sgi = 0; sgi < nseg; sgi++) {
    if ((rc = build_pairlist(s))) return rc;
    HIPCHK(s, hipMemcpyAsync(ck_pos + (size_t)sgi * n, s->pos, sizeof(double4) * n, hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(s, hipMemcpyAsync(ck_mom + (size_t)sgi * n, s->mom, sizeof(double4) * n, hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(s, hipMemcpyAsync(ck_tag + (size_t)sgi * n, s->tag, sizeof(int) * n, hipMemcpyDeviceToDevice, s->stream));
    const int k1 = std::min(nsteps, (sgi + 1) * seg);
    for (int k = sgi * seg; k < k1; k++) {
      if ((rc = lj_step(s, lj, s->pos_alt))) return rc;
      if (k >= nsteps - nobs) {
        if ((rc = refresh(s))) return rc;
        const int nb = tile_grid(s->ntiles);
        k_lj_observe_tiles<false><<<nb, kBlk, stage_bytes(s), s->stream>>>(
            s->pos, s->mom, s->tiles, s->ntiles, s->treal, s->sidx, s->kcap, s->list16, s->nn, lj,
            0.0, 0.0, 0.0, nullptr, nullptr, s->partial);
        k_sum_partials_k<3><<<1, kBlock, 0, s->stream>>>(s->partial, nb, obs + 3 * (size_t)k, 1.0);
      }
    }
    if ((rc = check_segment_list(s, lj))) return rc;
  }
  k_scatter4_by_tag<<<nblk(n), kBlock, 0, s->stream>>>(s->pos, s->tag, n, fin_q);
  k_scatter4_by_tag<<<nblk(n), kBlock, 0, s->stream>>>(s->mom, s->tag, n, fin_p);
  std::vector<double> ho(3 * (size_t)nsteps);
  HIPCHK(s, hipMemcpyAsync(ho.data(), obs, sizeof(double) * 3 * nsteps, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  double Psum = 0.0, Esum = 0.0;
  for (int k = nsteps - nobs; k < nsteps; k++) {
    const double ke = ho[3 * k], pe = ho[3 * k + 1], w = ho[3 * k + 2];
    Psum += (2.0 * ke / 3.0 - w) / V;   // MDManager::Pressure (mdmanager.cc:563-570)
    Esum += pe / n;
  }
  const double Pm = Psum / nobs, Em = Esum / nobs;
  res->pressure_mean = Pm;
  res->pe_mean = Em;
  res->loss = (Pm - p_target) * (Pm - p_target) + (Em - pe_target) * (Em - pe_target);
  res->dloss_deps = res->dloss_dsigma = 0.0;
  const auto t1 = std::chrono::steady_clock::now();
  res->forward_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  res->backward_ms = 0.0;
  if (grad) {
    // ---- reverse sweep: recompute each segment from its checkpoint, then
    // walk its steps backwards
    const int tot = n + std::max(s->g, 0);
    const size_t totcap = (size_t)s->cap + s->gcap;
    double4 *st_mid, *st_end, *st_mom, *aq, *ap, *aext, *aid_q, *aid_p;
    double *theta;
    const int nslots = nsteps + nobs;
    if ((rc = sc.get(&st_mid, (size_t)seg * totcap)) || (rc = sc.get(&st_end, (size_t)seg * totcap)) ||
        (rc = sc.get(&st_mom, (size_t)seg * n)) || (rc = sc.get(&aq, (size_t)n)) ||
        (rc = sc.get(&ap, (size_t)n)) || (rc = sc.get(&aext, totcap)) ||
        (rc = sc.get(&aid_q, (size_t)n)) || (rc = sc.get(&aid_p, (size_t)n)) ||
        (rc = sc.get(&theta, 4 * (size_t)nslots)))
      return rc;
    (void)tot;
    HIPCHK(s, hipMemsetAsync(aid_q, 0, sizeof(double4) * n, s->stream));
    HIPCHK(s, hipMemsetAsync(aid_p, 0, sizeof(double4) * n, s->stream));
    HIPCHK(s, hipMemsetAsync(theta, 0, sizeof(double) * 4 * nslots, s->stream));
    const double cP = 2.0 * (Pm - p_target) / nobs, cE = 2.0 * (Em - pe_target) / nobs;
    const double cw = -cP / V, ce = cE / n, ck = cP * 2.0 / (3.0 * V);
    int slot = 0;
    for (int sgi = nseg - 1; sgi >= 0; sgi--) {
      HIPCHK(s, hipMemcpyAsync(s->pos, ck_pos + (size_t)sgi * n, sizeof(double4) * n, hipMemcpyDeviceToDevice, s->stream));
      HIPCHK(s, hipMemcpyAsync(s->mom, ck_mom + (size_t)sgi * n, sizeof(double4) * n, hipMemcpyDeviceToDevice, s->stream));
      HIPCHK(s, hipMemcpyAsync(s->tag, ck_tag + (size_t)sgi * n, sizeof(int) * n, hipMemcpyDeviceToDevice, s->stream));
      // the checkpoint is already in cell order: the stable sort keeps it and
      // the list, tiles and images come out as in the forward pass
      if ((rc = build_pairlist(s))) return rc;
      const int T = s->n + s->g;
      k_gather4_by_tag<<<nblk(n), kBlock, 0, s->stream>>>(aid_q, s->tag, n, aq);
      k_gather4_by_tag<<<nblk(n), kBlock, 0, s->stream>>>(aid_p, s->tag, n, ap);
      const int k0 = sgi * seg, k1 = std::min(nsteps, (sgi + 1) * seg);
      for (int k = k0; k < k1; k++) {
        double4 *mid = st_mid + (size_t)(k - k0) * totcap;
        if ((rc = lj_step(s, lj, mid))) return rc;
        if (k >= nsteps - nobs) {
          if ((rc = refresh(s))) return rc;
          HIPCHK(s, hipMemcpyAsync(st_end + (size_t)(k - k0) * totcap, s->pos, sizeof(double4) * T, hipMemcpyDeviceToDevice, s->stream));
          HIPCHK(s, hipMemcpyAsync(st_mom + (size_t)(k - k0) * n, s->mom, sizeof(double4) * n, hipMemcpyDeviceToDevice, s->stream));
        }
      }
      const int nb = tile_grid(s->ntiles);
      for (int k = k1 - 1; k >= k0; k--) {
        if (k >= nsteps - nobs) {
          // sensitivity of the observation after step k
          k_lj_observe_tiles<true><<<nb, kBlk, stage_bytes(s), s->stream>>>(
              st_end + (size_t)(k - k0) * totcap, st_mom + (size_t)(k - k0) * n, s->tiles,
              s->ntiles, s->treal, s->sidx, s->kcap, s->list16, s->nn, lj, cw, ce, ck, aq, ap,
              s->partial);
          k_sum_partials_k<4><<<1, kBlock, 0, s->stream>>>(s->partial, nb, theta + 4 * (size_t)slot++, 1.0);
        }
        k_adj_pre<<<nblk(T), kBlock, 0, s->stream>>>(ap, aq, s->gsrc, n, s->g, s->prm.dt * 0.5, aext);
        k_adj_hvp_tiles<<<nb, kBlk, stage_bytes(s), s->stream>>>(
            st_mid + (size_t)(k - k0) * totcap, s->tiles, s->ntiles, s->treal, s->sidx, s->kcap,
            s->list16, s->nn, lj, s->prm.dt, aext, aq, ap, s->partial);
        k_sum_partials_k<2><<<1, kBlock, 0, s->stream>>>(s->partial, nb, theta + 4 * (size_t)slot++, 1.0);
        if ((rc = launch_check(s, "adjoint step"))) return rc;
      }
      k_scatter4_by_tag<<<nblk(n), kBlock, 0, s->stream>>>(aq, s->tag, n, aid_q);
      k_scatter4_by_tag<<<nblk(n), kBlock, 0, s->stream>>>(ap, s->tag, n, aid_p);
    }
    std::vector<double> ht(4 * (size_t)nslots);
    HIPCHK(s, hipMemcpyAsync(ht.data(), theta, sizeof(double) * 4 * nslots, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    // slots in the order written: per step backwards, [observation], hvp
    double ge = 0.0, gs = 0.0;
    slot = 0;
    for (int sgi = nseg - 1; sgi >= 0; sgi--) {
      const int k0 = sgi * seg, k1 = std::min(nsteps, (sgi + 1) * seg);
      for (int k = k1 - 1; k >= k0; k--) {
        if (k >= nsteps - nobs) {
          const double *v = &ht[4 * (size_t)slot++];
          ge += ce * v[0] + cw * v[2];
          gs += ce * v[1] + cw * v[3];
        }
        const double *v = &ht[4 * (size_t)slot++];
        ge += v[0];
        gs += v[1];
      }
    }
    res->dloss_deps = ge;
    res->dloss_dsigma = gs;
    if (dq0) {
      std::vector<double4> hq(n);
      HIPCHK(s, hipMemcpyAsync(hq.data(), aid_q, sizeof(double4) * n, hipMemcpyDeviceToHost, s->stream));
      HIPCHK(s, hipStreamSynchronize(s->stream));
      for (int i = 0; i < n; i++) {
        dq0[3 * (size_t)i] = hq[i].x;
        dq0[3 * (size_t)i + 1] = hq[i].y;
        dq0[3 * (size_t)i + 2] = hq[i].z;
      }
    }
    res->backward_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
  }
  // leave the system at the end of the trajectory (ids in order, no list)
  HIPCHK(s, hipMemcpyAsync(s->pos, fin_q, sizeof(double4) * n, hipMemcpyDeviceToDevice, s->stream));
  HIPCHK(s, hipMemcpyAsync(s->mom, fin_p, sizeof(double4) * n, hipMemcpyDeviceToDevice, s->stream));
  k_iota<<<nblk(n), kBlock, 0, s->stream>>>(s->tag, n);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  s->list_valid = false;
  s->vmax_valid = false;
  return launch_check(s, "trajectory");
}

}  // namespace

// ==================================================================== C ABI
extern "C" {

void mdx_default_params(mdx_params *prm) {
  if (!prm) return;
  std::memset(prm, 0, sizeof *prm);
  prm->dt = 0.001;
  prm->cutoff = 3.0;
  prm->buffer_length = 0.3;
  prm->check_list_length = 500;
}

int mdx_create(const mdx_params *prm, mdx_system **out) {
  if (!prm || !out) return fail(nullptr, MDX_ERR_ARG, "null argument");
  *out = nullptr;
  const double SL = prm->cutoff + prm->buffer_length;
  for (int d = 0; d < 3; d++) {
    if (!(prm->L[d] >= 2.0 * SL))
      return fail(nullptr, MDX_ERR_BOX, "box edge %g < 2 x search length %g", prm->L[d], SL);
  }
  if (!(prm->dt > 0.0) || !(prm->cutoff > 0.0) || !(prm->buffer_length > 0.0))
    return fail(nullptr, MDX_ERR_ARG, "dt, cutoff and buffer_length must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= prm->device)
    return fail(nullptr, MDX_ERR_HIP, "HIP device %d not available", prm->device);
  if (hipSetDevice(prm->device) != hipSuccess)
    return fail(nullptr, MDX_ERR_HIP, "hipSetDevice(%d) failed", prm->device);
  mdx_system *s = new mdx_system();
  s->prm = *prm;
  s->SL = SL;
  {
    const hipError_t e = scratch_address(prm->device, &s->scratch_dev);
    if (e != hipSuccess) {
      delete s;
      return fail(nullptr, MDX_ERR_HIP, "address of the export staging on device %d: %s",
                  prm->device, hipGetErrorString(e));
    }
  }
  // quadratic-shift constants (variables.cc:21-27)
  {
    const double s2 = 1.0 / (prm->cutoff * prm->cutoff);
    const double s6 = s2 * s2 * s2, s8 = s6 * s2, s12 = s6 * s6, s14 = s12 * s2;
    s->C2 = 6.0 * s14 - 3.0 * s8;
    s->C0 = -s12 + s6 - s->C2 / s2;
  }
  int rc = MDX_OK;
  if (prm->stream) {
    s->stream = (hipStream_t)prm->stream;
  } else {
    if (hipStreamCreate(&s->stream) != hipSuccess) rc = MDX_ERR_HIP;
    s->own_stream = true;
  }
  if (!rc) {
    for (auto &e : s->ev)
      if (hipEventCreate(&e) != hipSuccess) rc = MDX_ERR_HIP;
  }
  if (!rc && ((rc = set_geometry(s, prm->L)) || (rc = dev_alloc(s, &s->dstat, 1))))
    ;
  if (!rc) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, prm->device) == hipSuccess) s->lds_limit = prop.sharedMemPerBlock;
  }
  if (!rc && hipHostMalloc((void **)&s->hstat, sizeof(DevStatus)) != hipSuccess) rc = MDX_ERR_HIP;
  if (!rc && hipHostMalloc((void **)&s->hctl, sizeof(StepCtl)) != hipSuccess) rc = MDX_ERR_HIP;
  if (!rc && hipMalloc((void **)&s->dctl, sizeof(StepCtl)) != hipSuccess) rc = MDX_ERR_HIP;
  if (!rc && (rc = dev_alloc(s, &s->dhb, 1)) == MDX_OK &&
      hipMemsetAsync(s->dhb, 0, sizeof(HeatDev), s->stream) != hipSuccess)
    rc = MDX_ERR_HIP;
  // SimulationInfo defaults (simulationinfo.cc:11-23); ControlTemperature = no
  s->hb.type = MDX_HEATBATH_NONE;
  s->hb.aimed_temperature = 1.0;
  s->hb.tau = 0.1;
  s->hb.gamma = 0.1;
  s->hb.zeta = 0.0;
  s->hb.seed = 0;
  if (!rc) {
    for (auto &e : s->evk)
      if (hipEventCreate(&e) != hipSuccess) rc = MDX_ERR_HIP;
  }
  if (!rc && prm->capacity > 0) rc = ensure_capacity(s, prm->capacity);
  if (rc) {
    g_create_error = s->err.empty() ? "allocation failed" : s->err;
    mdx_destroy(s);
    return rc;
  }
  *out = s;
  return MDX_OK;
}

int mdx_destroy(mdx_system *s) {
  if (!s) return MDX_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  void *bufs[] = {s->pos, s->pos_alt, s->pos_mid, s->tinfo, s->treal, s->tmeta, s->tile_S, s->tile_s0, s->sidx,
                  s->blk_nt, s->blk_toff, s->pmax, s->tiles, s->list16, s->list_nat,
                  s->pos_tmp, s->ref_pos, s->mom, s->mom_tmp, s->gpos_tmp, s->vec_out,
                  s->tag, s->tag_tmp, s->gsrc, s->gcode, s->gsrc_tmp, s->gcode_tmp, s->cellid,
                  s->slot, s->order, s->inv, s->nimg, s->goff, s->cell_count, s->cell_start,
                  s->cnt_in, s->cnt_out, s->scan_in, s->scan_out, s->scan_aux, s->nn,
                  s->disp, s->partial, s->cand, s->io_buf, s->i_buf, s->dstat, s->dhb};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (s->hstat) (void)hipHostFree(s->hstat);
  if (s->hctl) (void)hipHostFree(s->hctl);
  if (s->dctl) (void)hipFree(s->dctl);
  for (auto &e : s->evk)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : s->ev)
    if (e) (void)hipEventDestroy(e);
  if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return MDX_OK;
}

const char *mdx_last_error(const mdx_system *s) {
  return s ? s->err.c_str() : g_create_error.c_str();
}

int mdx_set_state(mdx_system *s, int n, const double *q3, const double *p3) {
  if (!s || n <= 0 || !q3 || !p3) return fail(s, MDX_ERR_ARG, "bad arguments to mdx_set_state");
  int rc = ensure_capacity(s, n);
  if (rc) return rc;
  s->n = n;
  s->langevin_k = 0;
  // initial image capacity: the SL shell around the box at the mean
  // density, doubled, plus slack; grown on demand
  const double V = s->geom.L[0] * s->geom.L[1] * s->geom.L[2];
  const double Vx = (s->geom.L[0] + 2 * s->SL) * (s->geom.L[1] + 2 * s->SL) * (s->geom.L[2] + 2 * s->SL);
  const int gguess = (int)std::min(8.0 * n, 2.0 * n * (Vx - V) / V + 4096);
  if ((rc = ensure_image_capacity(s, gguess))) return rc;
  const double rho = n / V;
  const int kguess = (int)(4.18879 * s->SL * s->SL * s->SL * rho * 1.35) + 16 + (kILP - 1);
  s->kcap = std::max((kguess + kILP - 1) / kILP * kILP, s->kcap);
  // positions staged in io_buf, momenta in vec_out (double4[cap] >= 3n doubles)
  double *pbuf = reinterpret_cast<double *>(s->vec_out);
  HIPCHK(s, hipMemcpyAsync(s->io_buf, q3, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s->stream));
  HIPCHK(s, hipMemcpyAsync(pbuf, p3, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s->stream));
  k_load_state<<<nblk(n), kBlock, 0, s->stream>>>(s->io_buf, pbuf, n, s->pos, s->mom, s->tag);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  s->g = 0;
  s->have_state = true;
  s->list_valid = false;
  s->vmax_valid = false;
  return launch_check(s, "load");
}

int mdx_get_state(mdx_system *s, double *q3, double *p3) {
  if (!s) return fail(nullptr, MDX_ERR_ARG, "null system");
  if (!s->have_state) return fail(s, MDX_ERR_STATE, "no state");
  if (q3) {
    k_scatter_by_tag<<<nblk(s->n), kBlock, 0, s->stream>>>(s->pos, s->tag, s->n, s->io_buf);
    HIPCHK(s, hipMemcpyAsync(q3, s->io_buf, sizeof(double) * 3 * s->n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
  }
  if (p3) {
    k_scatter_by_tag<<<nblk(s->n), kBlock, 0, s->stream>>>(s->mom, s->tag, s->n, s->io_buf);
    HIPCHK(s, hipMemcpyAsync(p3, s->io_buf, sizeof(double) * 3 * s->n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
  }
  return launch_check(s, "get_state");
}

// images [k0, k0 + m) into host-transfer rows: positions, source momenta
// (3 doubles each) and source ids, shift codes
__global__ void k_image_rows(const double4 *__restrict__ pos, const double4 *__restrict__ mom,
                             const int *__restrict__ gsrc, const int *__restrict__ gcode,
                             const int *__restrict__ tag, int n, int k0, int m, int what,
                             double *__restrict__ out3, int *__restrict__ outi) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const int k = k0 + j, src = gsrc[k];
  if (what == 0 || what == 1) {
    const double4 v = what == 0 ? pos[n + k] : mom[src];
    out3[3 * (size_t)j + 0] = v.x;
    out3[3 * (size_t)j + 1] = v.y;
    out3[3 * (size_t)j + 2] = v.z;
  } else {
    outi[j] = what == 2 ? tag[src] : gcode[k];
  }
}

int mdx_get_images(mdx_system *s, double *q3, double *p3, int32_t *id, int32_t *code) {
  if (!s) return fail(nullptr, MDX_ERR_ARG, "null system");
  if (!s->have_state) return fail(s, MDX_ERR_STATE, "no state");
  if (!s->list_valid) return fail(s, MDX_ERR_STATE, "no pairlist (no images yet)");
  // in chunks of the host-transfer buffers (cap particles)
  for (int k0 = 0; k0 < s->g; k0 += s->cap) {
    const int m = std::min(s->cap, s->g - k0);
    void *dst[4] = {q3 ? q3 + 3 * (size_t)k0 : nullptr, p3 ? p3 + 3 * (size_t)k0 : nullptr,
                    id ? id + k0 : nullptr, code ? code + k0 : nullptr};
    for (int what = 0; what < 4; what++) {
      if (!dst[what]) continue;
      k_image_rows<<<nblk(m), kBlock, 0, s->stream>>>(s->pos, s->mom, s->gsrc, s->gcode, s->tag,
                                                        s->n, k0, m, what, s->io_buf, s->i_buf);
      if (what < 2)
        HIPCHK(s, hipMemcpyAsync(dst[what], s->io_buf, sizeof(double) * 3 * m, hipMemcpyDeviceToHost, s->stream));
      else
        HIPCHK(s, hipMemcpyAsync(dst[what], s->i_buf, sizeof(int) * m, hipMemcpyDeviceToHost, s->stream));
      HIPCHK(s, hipStreamSynchronize(s->stream));
    }
  }
  return launch_check(s, "get_images");
}

int mdx_synchronize(mdx_system *s) {
  if (!s) return MDX_ERR_ARG;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return launch_check(s, "synchronize");
}

int mdx_particle_number(const mdx_system *s, int *pn, int *tn) {
  if (!s) return MDX_ERR_ARG;
  if (pn) *pn = s->n;
  if (tn) *tn = s->n + s->g;
  return MDX_OK;
}

int mdx_make_pairlist(mdx_system *s) {
  if (!s) return MDX_ERR_ARG;
  return build_pairlist(s);
}

int mdx_is_pairlist_expired(mdx_system *s, int *out) {
  if (!s || !out) return fail(s, MDX_ERR_ARG, "null argument");
  return expired(s, out);
}

int mdx_update_position_half(mdx_system *s) {
  if (!s) return fail(nullptr, MDX_ERR_ARG, "null system");
  if (!s->have_state) return fail(s, MDX_ERR_STATE, "no state");
  return drift(s);
}

int mdx_refresh_images(mdx_system *s) {
  if (!s) return fail(nullptr, MDX_ERR_ARG, "null system");
  if (!s->list_valid) return fail(s, MDX_ERR_STATE, "no pairlist");
  return refresh(s);
}

int mdx_calculate_force(mdx_system *s) {
  if (!s) return fail(nullptr, MDX_ERR_ARG, "null system");
  if (!s->list_valid) return fail(s, MDX_ERR_STATE, "no pairlist");
  int rc = kick(s);
  if (rc) return rc;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  if (s->timing) {
    s->t_kick_only += ev_ms(s, 2, 3);
    s->n_kick_timed++;
  }
  return MDX_OK;
}

int mdx_set_heatbath(mdx_system *s, const mdx_heatbath *hb) {
  if (!s || !hb) return fail(s, MDX_ERR_ARG, "bad arguments to mdx_set_heatbath");
  if (hb->type < MDX_HEATBATH_NONE || hb->type > MDX_HEATBATH_LANGEVIN)
    return fail(s, MDX_ERR_ARG, "unknown heat bath type %d", hb->type);
  if (!std::isfinite(hb->aimed_temperature) || hb->aimed_temperature < 0.0 ||
      !std::isfinite(hb->zeta))
    return fail(s, MDX_ERR_ARG, "heat bath temperature and zeta must be finite, temperature >= 0");
  if (hb->type == MDX_HEATBATH_NOSE_HOOVER && !(hb->tau > 0.0 && std::isfinite(hb->tau)))
    return fail(s, MDX_ERR_ARG, "Nose-Hoover needs a finite tau > 0");
  if (hb->type == MDX_HEATBATH_LANGEVIN && !(hb->gamma >= 0.0 && std::isfinite(hb->gamma)))
    return fail(s, MDX_ERR_ARG, "Langevin needs a finite gamma >= 0");
  s->hb = *hb;
  HeatDev h{hb->zeta};
  HIPCHK(s, hipMemcpyAsync(s->dhb, &h, sizeof h, hipMemcpyHostToDevice, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return MDX_OK;
}

int mdx_get_heatbath(mdx_system *s, mdx_heatbath *hb) {
  if (!s || !hb) return fail(s, MDX_ERR_ARG, "bad arguments to mdx_get_heatbath");
  HeatDev h{};
  HIPCHK(s, hipMemcpyAsync(&h, s->dhb, sizeof h, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  *hb = s->hb;
  hb->zeta = h.zeta;
  return MDX_OK;
}

int mdx_heatbath_zeta(mdx_system *s, double current_temperature) {
  if (!s) return MDX_ERR_ARG;
  if (!(s->hb.tau > 0.0)) return fail(s, MDX_ERR_STATE, "heat bath tau must be > 0");
  k_nh_zeta<<<1, kBlock, 0, s->stream>>>(s->partial, 0, 1, current_temperature,
                                         s->hb.aimed_temperature, s->hb.tau, s->prm.dt, s->dhb);
  return launch_check(s, "heatbath zeta");
}

int mdx_heatbath_momenta(mdx_system *s) {
  if (!s) return fail(nullptr, MDX_ERR_ARG, "null system");
  if (!s->have_state) return fail(s, MDX_ERR_STATE, "no state");
  return nh_scale(s);
}

int mdx_langevin(mdx_system *s) {
  if (!s) return fail(nullptr, MDX_ERR_ARG, "null system");
  if (!s->have_state) return fail(s, MDX_ERR_STATE, "no state");
  return langevin(s);
}

int mdx_calculate(mdx_system *s, int nsteps, int *rebuilds) {
  if (!s) return fail(nullptr, MDX_ERR_ARG, "null system");
  if (!s->have_state) return fail(s, MDX_ERR_STATE, "no state");
  int nreb = 0, rc;
  int done = 0;
  bool device_expired = false;   // the last batch ended on a device check that expired the list
  while (done < nsteps) {
    // start of step `done`: PairList::IsPairListExpired, then MakePairList.
    // After a device check that found the list expired the host has the
    // verdict already (the budget is spent, so the check is all the
    // reference does at this point) and rebuilds at once.
    int exp = 0;
    if (device_expired) {
      exp = 1;
      device_expired = false;
    } else if ((rc = expired(s, &exp))) {
      return rc;
    }
    if (exp) {
      if ((rc = build_pairlist(s))) return rc;
      nreb++;
    }
    // queue as many steps as the budget is expected to last at the last
    // |p|max, then a few steps that check displacements on the device (as
    // many as the last pairlist survived, plus one), so that the batch runs
    // on into the checks and ends near the expiry
    int want = nsteps - done, first_checked = kRing;
    if (!(s->buffer_length > 0.0)) {
      want = std::min(want, s->check_steps + 1);
      first_checked = 1;
    } else if (s->vmax_valid && s->vmax2_host > 0.0) {
      const double per = 2.0 * std::sqrt(s->vmax2_host) * s->prm.dt;
      const int budget_steps = (int)std::min(s->buffer_length / per + 1.0, (double)kRing);
      first_checked = budget_steps;
      want = std::min(want, std::min(budget_steps + s->check_steps, kRing));
    }
    int ran = 0;
    if ((rc = run_batch(s, want, first_checked, &ran))) return rc;
    // replay the budget at the starts of the batch's later steps exactly;
    // the device only went on while the budget stayed positive, or past a
    // displacement check it passed
    const int checked_from = s->hctl->first_checked;
    for (int b = 0; b + 1 < ran; b++) {
      const double vmax = std::sqrt(s->hctl->ring[b]);
      if (s->buffer_length > 0.0) {
        if (vmax * 2.0 * s->prm.dt > s->prm.buffer_length)
          return fail(s, MDX_ERR_TOO_FAST, "Too fast particles exists. max_velocity = %f", vmax);
        s->buffer_length = s->buffer_length - vmax * 2.0 * s->prm.dt;
      }
      if (s->buffer_length < 0.0 && b + 1 < checked_from)
        return fail(s, MDX_ERR_STATE, "step guard ran past a spent pairlist budget");
    }
    device_expired = s->hctl->expired != 0;
    if (s->hctl->expired)
      s->check_steps = std::min(16, std::max(2, s->hctl->checks_passed + 2));
    else if (s->hctl->checks_passed > 0 && ran == want)
      s->check_steps = std::min(16, 2 * s->check_steps);
    s->vmax2_host = s->hctl->ring[ran - 1];
    s->vmax_valid = true;
    done += ran;
    s->n_steps += ran;
  }
  if (rebuilds) *rebuilds = nreb;
  return MDX_OK;
}

int mdx_observe(mdx_system *s, double *ke, double *pe, double *vir) {
  if (!s) return MDX_ERR_ARG;
  return reduce_observe(s, ke, pe, vir);
}

int mdx_forces(mdx_system *s, double *f3) {
  if (!s || !f3) return fail(s, MDX_ERR_ARG, "null argument");
  if (!s->list_valid) return fail(s, MDX_ERR_STATE, "no pairlist");
  k_kick_tiles<1><<<tile_grid(s->ntiles), kBlk, stage_bytes(s), s->stream>>>(
      s->pos, s->mom, s->tiles, s->ntiles, s->treal, s->sidx, s->kcap, s->list16, s->tmeta,
      make_pairk(s->C2, 1.0), 1.0, s->geom.CL2, s->vec_out, nullptr);
  k_scatter_by_tag<<<nblk(s->n), kBlock, 0, s->stream>>>(s->vec_out, s->tag, s->n, s->io_buf);
  HIPCHK(s, hipMemcpyAsync(f3, s->io_buf, sizeof(double) * 3 * s->n, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return launch_check(s, "forces");
}

int mdx_pairs_canonical(mdx_system *s, int32_t *out, int64_t cap, int64_t *count) {
  if (!s) return MDX_ERR_ARG;
  if (!s->list_valid) return fail(s, MDX_ERR_STATE, "no pairlist");
  const int n = s->n;
  k_canonical_tiles<false><<<s->ntiles, kBlk, 0, s->stream>>>(
      s->tiles, s->ntiles, s->treal, s->sidx, n, s->kcap, s->list16, s->nn, s->tag, s->gsrc,
      s->gcode, nullptr, s->i_buf, nullptr);
  int rc = scan(s, s->i_buf, s->goff, n, &s->dstat->ghost_total);
  if (rc) return rc;
  if ((rc = sync_status(s))) return rc;
  const long long total = s->hstat->ghost_total;
  if (count) *count = total;
  if (!out) return MDX_OK;
  int32_t *dout = nullptr;
  HIPCHK(s, hipMalloc((void **)&dout, sizeof(int32_t) * 3 * (size_t)std::max(total, 1LL)));
  k_canonical_tiles<true><<<s->ntiles, kBlk, 0, s->stream>>>(
      s->tiles, s->ntiles, s->treal, s->sidx, n, s->kcap, s->list16, s->nn, s->tag, s->gsrc,
      s->gcode, s->goff, nullptr, dout);
  std::vector<int32_t> h(3 * (size_t)total);
  hipError_t e = hipMemcpyAsync(h.data(), dout, sizeof(int32_t) * 3 * (size_t)total,
                                hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(dout);
  if (e != hipSuccess) return fail(s, MDX_ERR_HIP, "canonical copy: %s", hipGetErrorString(e));
  std::vector<long long> idx(total);
  for (long long k = 0; k < total; k++) idx[k] = k;
  std::sort(idx.begin(), idx.end(), [&](long long a, long long b) {
    for (int d = 0; d < 3; d++)
      if (h[3 * a + d] != h[3 * b + d]) return h[3 * a + d] < h[3 * b + d];
    return false;
  });
  const long long w = std::min<long long>(cap, total);
  for (long long k = 0; k < w; k++)
    for (int d = 0; d < 3; d++) out[3 * k + d] = h[3 * idx[k] + d];
  return MDX_OK;
}

int mdx_pairs_csr(mdx_system *s, int32_t *key_pointer, int32_t *number_of_partners,
                  int32_t *sorted_list, int64_t cap, int64_t *count) {
  if (!s) return MDX_ERR_ARG;
  if (cap < 0) return fail(s, MDX_ERR_ARG, "negative capacity");
  if (!s->list_valid) return fail(s, MDX_ERR_STATE, "no pairlist");
  const int n = s->n, tn = n + s->g;
  k_canonical_tiles<false, true><<<s->ntiles, kBlk, 0, s->stream>>>(
      s->tiles, s->ntiles, s->treal, s->sidx, n, s->kcap, s->list16, s->nn, s->tag, s->gsrc,
      s->gcode, nullptr, s->i_buf, nullptr);
  int rc = scan(s, s->i_buf, s->goff, n, &s->dstat->ghost_total);
  if (rc) return rc;
  if ((rc = sync_status(s))) return rc;
  const long long total = s->hstat->ghost_total;
  if (count) *count = total;
  // counts and offsets by id; the images key no pair (0 partners, their
  // offset the total, as the reference's prefix sum over all indices)
  if (number_of_partners || key_pointer) {
    if (number_of_partners)
      HIPCHK(s, hipMemcpyAsync(number_of_partners, s->i_buf, sizeof(int32_t) * n,
                               hipMemcpyDeviceToHost, s->stream));
    if (key_pointer)
      HIPCHK(s, hipMemcpyAsync(key_pointer, s->goff, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                               s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    for (int k = n; k < tn; k++) {
      if (number_of_partners) number_of_partners[k] = 0;
      if (key_pointer) key_pointer[k] = (int32_t)total;
    }
  }
  if (!sorted_list || cap == 0 || total == 0) return launch_check(s, "csr export");
  int32_t *dout = nullptr;
  HIPCHK(s, hipMalloc((void **)&dout, sizeof(int32_t) * (size_t)total));
  k_canonical_tiles<true, true><<<s->ntiles, kBlk, 0, s->stream>>>(
      s->tiles, s->ntiles, s->treal, s->sidx, n, s->kcap, s->list16, s->nn, s->tag, s->gsrc,
      s->gcode, s->goff, nullptr, dout);
  const long long w = std::min<long long>(cap, total);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(sorted_list, dout, sizeof(int32_t) * (size_t)w, hipMemcpyDeviceToHost,
                       s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(dout);
  if (e != hipSuccess) return fail(s, MDX_ERR_HIP, "csr export: %s", hipGetErrorString(e));
  return MDX_OK;
}

int mdx_adjust_periodic_boundary(mdx_system *s) {
  if (!s) return fail(nullptr, MDX_ERR_ARG, "null system");
  if (!s->have_state) return fail(s, MDX_ERR_STATE, "no particles loaded");
  HIPCHK(s, hipMemsetAsync(&s->dstat->flags, 0, sizeof(int), s->stream));
  k_wrap_reals<<<nblk(s->n), kBlock, 0, s->stream>>>(s->pos, s->n, s->geom, &s->dstat->flags);
  s->list_valid = false;   // reals may have moved across the box
  int rc = launch_check(s, "wrap");
  if (rc || (rc = sync_status(s))) return rc;
  return status_error(s);
}

int mdx_scale_positions(mdx_system *s, double alpha) {
  if (!s) return fail(nullptr, MDX_ERR_ARG, "null system");
  if (!s->have_state) return fail(s, MDX_ERR_STATE, "no particles loaded");
  if (!(alpha > 0.0) || !std::isfinite(alpha)) return fail(s, MDX_ERR_ARG, "scale %g", alpha);
  k_scale_reals<<<nblk(s->n), kBlock, 0, s->stream>>>(s->pos, s->n, alpha);
  s->list_valid = false;
  int rc = launch_check(s, "scale");
  if (rc) return rc;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return MDX_OK;
}

int mdx_set_box(mdx_system *s, const double *L3) {
  if (!s || !L3) return fail(s, MDX_ERR_ARG, "null argument");
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return set_geometry(s, L3);
}

int mdx_density_grid(mdx_system *s, double grid_size, int32_t *dims, uint8_t *data, int64_t cap) {
  if (!s || !dims) return fail(s, MDX_ERR_ARG, "null argument");
  if (!s->have_state) return fail(s, MDX_ERR_STATE, "no particles loaded");
  if (!(grid_size > 0.0) || !std::isfinite(grid_size))
    return fail(s, MDX_ERR_ARG, "grid size %g", grid_size);
  // BubbleHist::BubbleHist (cavitation.cc:38-80): boxes per axis rounded
  // from the widths, one box size from the x width
  int g[3];
  for (int d = 0; d < 3; d++) g[d] = (int)(s->geom.L[d] / grid_size + 0.5);
  if (g[0] < 1 || g[1] < 1 || g[2] < 1) return fail(s, MDX_ERR_ARG, "grid size %g exceeds the box", grid_size);
  const long long total = (long long)g[0] * g[1] * g[2];
  for (int d = 0; d < 3; d++) dims[d] = g[d];
  if (!data) return MDX_OK;
  if (cap < total || total > (1LL << 31) - 1)
    return fail(s, MDX_ERR_CAPACITY, "density grid of %lld boxes, buffer of %lld", total, (long long)cap);
  const double box = s->geom.L[0] / (double)g[0];
  const double inv = 1.0 / box;
  int *hist = nullptr;
  unsigned char *bytes = nullptr;
  HIPCHK(s, hipMalloc((void **)&hist, sizeof(int) * (size_t)total));
  hipError_t e = hipMalloc((void **)&bytes, (size_t)total);
  if (e == hipSuccess) e = hipMemsetAsync(hist, 0, sizeof(int) * (size_t)total, s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(&s->dstat->flags, 0, sizeof(int), s->stream);
  if (e == hipSuccess) {
    k_density_count<<<nblk(s->n), kBlock, 0, s->stream>>>(s->pos, s->n, inv, g[0], g[1], (int)total,
                                                          hist, &s->dstat->flags);
    k_density_bytes<<<nblk(total), kBlock, 0, s->stream>>>(hist, (int)total, bytes);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(data, bytes, (size_t)total, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(s->hstat, s->dstat, sizeof(DevStatus), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(hist);
  (void)hipFree(bytes);
  if (e != hipSuccess) return fail(s, MDX_ERR_HIP, "density grid: %s", hipGetErrorString(e));
  if (s->hstat->flags & ST_INVALID_POS)
    return fail(s, MDX_ERR_INVALID_POS, "a particle lies outside the density grid");
  return MDX_OK;
}

int mdx_get_cell_order(mdx_system *s, int32_t *ids) {
  if (!s || !ids) return fail(s, MDX_ERR_ARG, "null argument");
  if (!s->list_valid) return fail(s, MDX_ERR_STATE, "no pairlist");
  HIPCHK(s, hipMemcpyAsync(ids, s->tag, sizeof(int32_t) * s->n, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return MDX_OK;
}

int mdx_pairs_canonical_noalloc(mdx_system *s, int32_t *out, int64_t cap, int64_t *count) {
  if (!s) return MDX_ERR_ARG;
  if (cap < 0) return fail(s, MDX_ERR_ARG, "negative capacity");
  if (!s->list_valid) return fail(s, MDX_ERR_STATE, "no pairlist");
  if (!s->scratch_dev) return fail(s, MDX_ERR_STATE, "no export staging on this device");
  const int n = s->n;
  k_canonical_tiles<false><<<s->ntiles, kBlk, 0, s->stream>>>(
      s->tiles, s->ntiles, s->treal, s->sidx, n, s->kcap, s->list16, s->nn, s->tag, s->gsrc,
      s->gcode, nullptr, s->i_buf, nullptr);
  int rc = scan(s, s->i_buf, s->goff, n, &s->dstat->ghost_total);
  if (rc) return rc;
  if ((rc = sync_status(s))) return rc;
  const long long total = s->hstat->ghost_total;
  if (count) *count = total;
  if (total > kNoallocPairs)
    return fail(s, MDX_ERR_CAPACITY,
                "canonical pair set of %lld pairs exceeds the %lld pairs the allocation-free "
                "export stages (mdx_pairs_canonical has no limit)",
                total, kNoallocPairs);
  if (!out || cap == 0 || total == 0) return MDX_OK;
  k_canonical_tiles<true><<<s->ntiles, kBlk, 0, s->stream>>>(
      s->tiles, s->ntiles, s->treal, s->sidx, n, s->kcap, s->list16, s->nn, s->tag, s->gsrc,
      s->gcode, s->goff, nullptr, s->scratch_dev);
  k_sort_runs<<<nblk(n), kBlock, 0, s->stream>>>(s->goff, s->i_buf, n, s->scratch_dev);
  if ((rc = launch_check(s, "canonical export"))) return rc;
  const long long w = std::min<long long>(cap, total);
  HIPCHK(s, hipMemcpyAsync(out, s->scratch_dev, sizeof(int32_t) * 3 * (size_t)w,
                           hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return MDX_OK;
}

int mdx_lj_trajectory(mdx_system *s, double eps, double sigma, int nsteps, int nobs,
                      int checkpoint_every, double p_target, double pe_target, int grad,
                      mdx_lj_result *out, double *dloss_dq0) {
  if (!s || !out) return fail(s, MDX_ERR_ARG, "null argument");
  return lj_trajectory(s, eps, sigma, nsteps, nobs, checkpoint_every, p_target, pe_target,
                       grad != 0, out, dloss_dq0);
}

int mdx_pair_number(mdx_system *s, int64_t *count) {
  return mdx_pairs_canonical(s, nullptr, 0, count);
}

int mdx_timing(mdx_system *s, double *build_ms, double *force_ms, double *other_ms,
               int64_t *builds, int64_t *steps) {
  if (!s) return MDX_ERR_ARG;
  if (build_ms) *build_ms = s->t_build;
  if (force_ms) *force_ms = s->t_force;
  if (other_ms) *other_ms = s->t_other;
  if (builds) *builds = s->n_build;
  if (steps) *steps = s->n_steps;
  return MDX_OK;
}

int mdx_timing_reset(mdx_system *s, int enable) {
  if (!s) return MDX_ERR_ARG;
  s->timing = enable != 0;
  s->t_build = s->t_force = s->t_other = s->t_kick_only = 0;
  s->n_build = s->n_steps = s->n_kick = s->n_kick_timed = 0;
  return MDX_OK;
}

int mdx_kick_profile(mdx_system *s, double *avg_ms, double *bytes) {
  if (!s) return MDX_ERR_ARG;
  if (avg_ms) *avg_ms = s->n_kick_timed > 0 ? s->t_kick_only / (double)s->n_kick_timed : 0.0;
  if (bytes) {
    // Algorithmic bytes of one fused kick + half-drift launch, per real:
    // read q and p, write p and q (24 B each), the list's offset and count
    // (4 B each: the CSR key_pointer and number_of_partners of the
    // reference), and 4 B per half-list partner (each pair once, as the
    // reference's half list); partner positions are re-reads of q and
    // images re-reads of their sources, counted once in q.  That is 104 B
    // per real plus 4 B per pair of the current list.
    int64_t pairs = 0;
    if (s->list_valid) {
      int rc = mdx_pair_number(s, &pairs);
      if (rc) return rc;
    }
    *bytes = (double)s->n * (4.0 * 24.0 + 8.0) + 4.0 * (double)pairs;
  }
  return MDX_OK;
}

}  // extern "C"
