// mdx_tiles.h — block tiles, the unit of work of the list build and of
// every kernel that walks the neighbour lists.  Shared by the single system
// (mdx.hip) and the replica ensemble (ensemble.hip), so that both form the
// same tiles from the same cells and build every real's list the same way.
//
// A tile is up to kBlk reals from a block of at most 3 x 3 interior cell
// rows (y0 .. y0+ny-1, z0 .. z0+nz-1): the reals of the block's rows between
// two x cuts shared by all of them.  Every partner of a tile's reals lies in
// the (ny+2) x (nz+2) cell rows around the block, within the x window
// [xa - h, xb + h] of the tile's reals (h a little over the search length).
// Each cell row is sorted by x as a whole (image cell 0, real cells 1..ms,
// image cell ms+1, or image cells only), so its window is one run of that
// order, found by two binary searches.  The runs are staged into LDS once
// per tile, row after row (staged row gr = gy + 5 gz), as interleaved (x, y,
// z); sidx holds the slot of every staged particle.  A neighbour list entry
// is 24 x the partner's staging index (its byte offset), so a tile stages
// at most kMaxStaged particles and the sentinel after them.
//
// Against tiles of one cell row staged a z-plane at a time, a block tile
// stages 4.5 instead of 11 particles per real at the same LDS per
// wavefront, and each list is walked in one pass: the wavefront's lanes
// (list columns, sorted by list length) run for the longest list of the
// wavefront instead of the longest per plane.
#ifndef MDX_TILES_H
#define MDX_TILES_H

#include <climits>

#include "mdx_common.h"

namespace mdxk {

constexpr int kBlk = 512;                                 // reals (list columns) per tile
constexpr int kBlkRows = 3;                               // most cell rows per block and axis
constexpr int kSRow = kBlkRows + 2;                       // staged rows per axis, at most
constexpr int kSRows = kSRow * kSRow;                     // staged rows per tile, at most
constexpr int kRealRows = kBlkRows * kBlkRows;            // real rows per block, at most

// Block shapes and reals aimed at per tile, tried in this order while some
// tile's staging overflows: 3 x 3 rows of cells suit the usual densities
// (about 2,300 staged particles for 500 reals at rho 0.5); dense boxes of
// wide cells need smaller blocks and tiles.
struct BlkShape {
  int rows, target;
};
constexpr BlkShape kBlkLadder[] = {{3, 503}, {3, 402}, {3, 321}, {3, 256}, {2, 224}, {2, 179},
                                   {2, 143}, {2, 114}, {1, 96},  {1, 76},  {1, 61},  {1, 48},
                                   {1, 38},  {1, 32}};
constexpr int kBlkLevels = (int)(sizeof kBlkLadder / sizeof kBlkLadder[0]);

// The list columns of a tile hold its reals ranked by list length, 64 ranks
// to a wavefront.  The wavefronts of a workgroup go round-robin to the CU's
// four SIMDs, so rank groups 0..3 take wavefronts 0..3 and groups 4..7
// wavefronts 7..4: each SIMD runs one long and one short wavefront of the
// tile.  The map is its own inverse.
__host__ __device__ __forceinline__ int col_of_rank(int r) {
  const int g = r >> 6;
  return (g < 4 ? g : 11 - g) * 64 + (r & 63);
}
// whether list column col of a tile of nreal reals holds a real
__device__ __forceinline__ bool col_active(int col, int nreal) { return col_of_rank(col) < nreal; }

// per tile, as every list walker reads it
struct BlkTile {
  int nreal;   // reals, list columns 0 .. nreal-1
  int S;       // staged particles; the sentinel sits at S
  int s0;      // first entry of the tile in sidx
  int pad;
};

// per tile, for the list build and the staging index
struct BlkInfo {
  int y0, z0, ny, nz;
  double xa, xb;              // x range of the tile's reals
  int sp[kRealRows + 1];      // place of the first real of real row r = ry + 3 rz among the tile's
  int sfirst[kRealRows];      // slot of that real
  int selfoff[kRealRows];     // staging index of real i of row r is i + selfoff[r]
  int rs[kSRows + 1];         // staging index of the first particle of staged row gr
  int rlo[kSRows];            // that particle's index in its row's x-sorted run
};

// Cells of one periodic system: its padded grid at cells [cbase, cbase +
// m^3) of the arrays; the reals of an interior cell cl sit at slots
// scan_in[cl] .., the images of an outer cell at img0 + scan_out[cl] ..
struct Cells {
  const int *scan_in, *cnt_in, *scan_out, *cnt_out;
  int cbase, img0;
};
__device__ __forceinline__ int cell_at(const Geom &g, const Cells &c, int x, int y, int z) {
  return c.cbase + cell_linear(g, x, y, z);
}

// The cell row (y, z) of the padded grid in x order: slot ranges a (image
// cell x = 0, or every cell of an image row), b (real cells 1..ms), c
// (image cell ms + 1).
struct RowRun {
  int a0, na, b0, nb, c0, nc;
};
__device__ __forceinline__ RowRun row_run(const Geom &g, const Cells &c, int y, int z) {
  const int cl0 = cell_at(g, c, 0, y, z), cl1 = cell_at(g, c, g.m[0] - 1, y, z);
  RowRun r;
  if (y >= 1 && y <= g.ms[1] && z >= 1 && z <= g.ms[2]) {
    r.a0 = c.img0 + c.scan_out[cl0];
    r.na = c.cnt_out[cl0];
    r.b0 = c.scan_in[cl0 + 1];
    r.nb = c.scan_in[cl1 - 1] + c.cnt_in[cl1 - 1] - r.b0;
    r.c0 = c.img0 + c.scan_out[cl1];
    r.nc = c.cnt_out[cl1];
  } else {
    r.a0 = c.img0 + c.scan_out[cl0];
    r.na = c.scan_out[cl1] + c.cnt_out[cl1] - c.scan_out[cl0];
    r.b0 = r.c0 = r.a0;
    r.nb = r.nc = 0;
  }
  return r;
}
__device__ __forceinline__ int run_len(const RowRun &r) { return r.na + r.nb + r.nc; }
__device__ __forceinline__ int run_slot(const RowRun &r, int t) {
  return t < r.na ? r.a0 + t : (t < r.na + r.nb ? r.b0 + (t - r.na) : r.c0 + (t - r.na - r.nb));
}
// first index of the run whose x is >= v (ABOVE: > v)
template <bool ABOVE>
__device__ __forceinline__ int run_bound(const RowRun &r, const double4 *__restrict__ pos, double v) {
  int lo = 0, hi = run_len(r);
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const double x = pos[run_slot(r, mid)].x;
    if (ABOVE ? x <= v : x < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// blocks of br x br cell rows: by fastest, bz-major
__host__ __device__ inline int blk_count(const Geom &g, int axis, int br) {
  return (g.ms[axis] + br - 1) / br;
}

// reals of real row r of block (by, bz) of br x br rows (ry = r % 3, rz =
// r / 3; none outside the block or the grid)
__device__ __forceinline__ void blk_row(const Geom &g, const Cells &c, int by, int bz, int br,
                                        int r, int *first, int *cnt) {
  const int ry = r % kBlkRows, rz = r / kBlkRows;
  const int y = br * by + ry + 1, z = br * bz + rz + 1;
  *first = 0;
  *cnt = 0;
  if (ry >= br || rz >= br || y > g.ms[1] || z > g.ms[2]) return;
  const int a = cell_at(g, c, 1, y, z), b = cell_at(g, c, g.ms[0], y, z);
  *first = c.scan_in[a];
  *cnt = c.scan_in[b] + c.cnt_in[b] - *first;
}

// Cumulative count of a block's reals over its x cells by one workgroup into
// LDS: hist[c] = reals of the block in x cells 1..c, hist[0] = 0.  Every
// thread calls it; it ends with a barrier.
constexpr int kMaxCellsX = 4096;
__device__ inline void blk_hist(const Geom &g, const Cells &c, int by, int bz, int br, int *hist) {
  const int ms = g.ms[0];
  for (int cx = (int)threadIdx.x + 1; cx <= ms; cx += (int)blockDim.x) {
    int s = 0;
    for (int ry = 0; ry < br; ry++)
      for (int rz = 0; rz < br; rz++) {
        const int y = br * by + ry + 1, z = br * bz + rz + 1;
        if (y <= g.ms[1] && z <= g.ms[2]) s += c.cnt_in[cell_at(g, c, cx, y, z)];
      }
    hist[cx] = s;
  }
  if (threadIdx.x == 0) hist[0] = 0;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int cx = 1; cx <= ms; cx++) hist[cx] += hist[cx - 1];
  __syncthreads();
}

// Where tile k of nt starts in real row r of the block: the first real of
// the row (slots [first, first + cnt), sorted by x) at or past cut k, the x
// at which the block's cumulative count, taken as linear within each x
// cell, reaches k M / nt.  Cut 0 is the row's start, cut nt its end, so the
// tiles of a block share their x cuts across its rows and cover every real
// once.
__device__ inline int blk_cut_slot(const Geom &g, const int *hist, const double4 *__restrict__ pos,
                                   int first, int cnt, int k, int nt) {
  if (k <= 0) return first;
  if (k >= nt) return first + cnt;
  const int ms = g.ms[0];
  const double cum = (double)k * (double)hist[ms] / (double)nt;
  int lo = 1, hi = ms;   // first x cell whose cumulative count exceeds cum
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((double)hist[mid] > cum) hi = mid;
    else lo = mid + 1;
  }
  const double f = (cum - (double)hist[lo - 1]) / (double)max(hist[lo] - hist[lo - 1], 1);
  const double X = ((double)(lo - 1) + f) / g.inv_w[0];
  int a = 0, b = cnt;
  while (a < b) {
    const int mid = (a + b) >> 1;
    if (pos[first + mid].x < X) a = mid + 1;
    else b = mid;
  }
  return first + a;
}

// The cut of tile k of nt of block (by, bz), by one wavefront (hist filled
// by blk_hist): real row r (lanes 0..8) gives the tile the slots between
// its cuts.  Lanes 0..24 then cut the staged rows to the window [xa - h,
// xb + h] of the tile's reals, h = SL (1 + 1e-10) + eps, wider than any
// |dx| of a pair with r^2 <= SL^2 as rounded, eps covering the rounding of
// positions near the box (the images' shift).
struct TileCut {
  int y0, z0, ny, nz;
  int lo, len;       // lanes 0..8: the tile's slots of real row `lane`
  int sp;            // lanes 0..9: reals of the rows before
  int nreal;         // every lane
  double xa, xb;     // every lane: x range of the tile's reals
  RowRun run;        // lanes 0..24: staged row `lane`
  int wlo, wlen;     // its window: first index in the run, length
  int rs;            // lanes 0..25: staged particles of the rows before
  int S;             // every lane: the staging size
};
__device__ inline void blk_cut_tile(const Geom &g, const Cells &c, const double4 *__restrict__ pos,
                                    const int *hist, int by, int bz, int br, int k, int nt,
                                    TileCut &t) {
  const int lane = threadIdx.x & 63;
  t.y0 = br * by + 1;
  t.z0 = br * bz + 1;
  t.ny = min(br, g.ms[1] - t.y0 + 1);
  t.nz = min(br, g.ms[2] - t.z0 + 1);
  const double h = sqrt(g.SL2) * (1.0 + 1e-10) + 1e-9 * (g.L[0] + g.L[1] + g.L[2]);
  t.lo = 0;
  t.len = 0;
  if (lane < kRealRows) {
    int rf, rm;
    blk_row(g, c, by, bz, br, lane, &rf, &rm);
    t.lo = blk_cut_slot(g, hist, pos, rf, rm, k, nt);
    t.len = blk_cut_slot(g, hist, pos, rf, rm, k + 1, nt) - t.lo;
  }
  double xa = kFar, xb = -kFar;
  if (t.len > 0) {
    xa = pos[t.lo].x;
    xb = pos[t.lo + t.len - 1].x;
  }
  for (int off = 32; off > 0; off >>= 1) {
    xa = fmin(xa, __shfl_xor(xa, off));
    xb = fmax(xb, __shfl_xor(xb, off));
  }
  t.xa = xa;
  t.xb = xb;
  int sp = 0;
  for (int r = 0; r < kRealRows; r++) {
    const int lr = __shfl(t.len, r);
    if (r < lane) sp += lr;
  }
  t.sp = sp;
  t.nreal = __shfl(sp, kRealRows);   // lane 9 holds the total
  t.wlo = 0;
  t.wlen = 0;
  t.run = RowRun{};
  const int gy = lane % kSRow, gz = lane / kSRow;
  if (lane < kSRows && gy < t.ny + 2 && gz < t.nz + 2 && t.nreal > 0) {
    t.run = row_run(g, c, t.y0 - 1 + gy, t.z0 - 1 + gz);
    t.wlo = run_bound<false>(t.run, pos, xa - h);
    t.wlen = run_bound<true>(t.run, pos, xb + h) - t.wlo;
  }
  int rs = 0;
  for (int r = 0; r < kSRows; r++) {
    const int lr = __shfl(t.wlen, r);
    if (r < lane) rs += lr;
  }
  t.rs = rs;
  t.S = __shfl(rs, kSRows);   // lane 25 holds the total
}

// Tiles of block (by, bz), by one workgroup (hist filled by blk_hist): nt =
// ceil(reals / target), raised until no tile holds more than kBlk reals and
// none stages more than cap particles (the shared x cuts follow the block's
// x histogram, so a tile's count is near M / nt, and a narrower x window
// stages fewer; cap = INT_MAX skips that check); none for an empty block.
// Every thread calls it and gets nt; it ends with a barrier.  A tile that
// still overflows after the last raise is left to the tile writer to
// report.
__device__ inline int blk_ntiles(const Geom &g, const Cells &c, const double4 *__restrict__ pos,
                                 const int *hist, int by, int bz, int br, int target, int cap) {
  __shared__ int rfirst[kRealRows], rcnt[kRealRows];
  constexpr int kChunk = kBlock / kRealRows;   // tiles counted per round of kBlock threads
  __shared__ int len[kChunk][kRealRows];
  __shared__ int over;
  const int tid = threadIdx.x, w = tid >> 6, nw = (int)(blockDim.x >> 6);
  if (tid < kRealRows) blk_row(g, c, by, bz, br, tid, &rfirst[tid], &rcnt[tid]);
  __syncthreads();
  const int M = hist[g.ms[0]];
  int nt = (M + target - 1) / target;
  if (cap < INT_MAX && M > 0) {
    // start from the count whose tiles would just fit the cap if the
    // block's reals were spread evenly: (ny + 2)(nz + 2) staged rows of
    // `line` particles per unit length over a tile's x extent T / (ny nz
    // line) plus 2 h, so that the checks below seldom raise it again
    const int ny = min(br, g.ms[1] - br * by), nz = min(br, g.ms[2] - br * bz);
    const double line = (double)M / ((double)(ny * nz) * g.L[0]);
    const double ext = 0.95 * (double)cap / ((double)((ny + 2) * (nz + 2)) * line) - 2.0 * sqrt(g.SL2);
    const double tmax = ext * (double)(ny * nz) * line;   // reals of such a tile
    if (tmax >= 16.0) nt = max(nt, (int)ceil((double)M / tmax));
  }
  for (int pass = 0; pass < 64 && nt > 0; pass++) {
    if (tid == 0) over = 0;
    __syncthreads();
    for (int k0 = 0; k0 < nt; k0 += kChunk) {
      const int kk = tid / kRealRows, r = tid % kRealRows, k = k0 + kk;
      if (kk < kChunk && k < nt)
        len[kk][r] = blk_cut_slot(g, hist, pos, rfirst[r], rcnt[r], k + 1, nt) -
                     blk_cut_slot(g, hist, pos, rfirst[r], rcnt[r], k, nt);
      __syncthreads();
      if (tid < kChunk && k0 + tid < nt) {
        int s = 0;
        for (int q = 0; q < kRealRows; q++) s += len[tid][q];
        if (s > kBlk) atomicOr(&over, 1);
      }
      __syncthreads();
    }
    if (!over && cap < INT_MAX) {
      // the stagings, one wavefront per tile
      for (int k = w; k < nt; k += nw) {
        TileCut tc;
        blk_cut_tile(g, c, pos, hist, by, bz, br, k, nt, tc);
        if ((tid & 63) == 0 && tc.S > cap) atomicOr(&over, 1);
      }
      __syncthreads();
    }
    if (!over) break;
    nt += max(1, nt / 8);
    __syncthreads();
  }
  __syncthreads();
  return nt;
}

// Tile k of nt of block (by, bz), by one wavefront (hist filled): its cut
// (blk_cut_tile) into *I, and *T with s0 = 0 and the given pad, in one
// store; returns the staging size S (every lane).  *nreal gets the tile's
// reals.
__device__ inline int blk_make_tile(const Geom &g, const Cells &c, const double4 *__restrict__ pos,
                                    const int *hist, int by, int bz, int br, int k, int nt,
                                    int pad, BlkTile *T, BlkInfo *I, int *nreal_out) {
  const int lane = threadIdx.x & 63;
  TileCut t;
  blk_cut_tile(g, c, pos, hist, by, bz, br, k, nt, t);
  if (lane == 0) {
    *T = BlkTile{t.nreal, t.S, 0, pad};
    I->y0 = t.y0;
    I->z0 = t.z0;
    I->ny = t.ny;
    I->nz = t.nz;
    I->xa = t.xa;
    I->xb = t.xb;
  }
  if (lane <= kRealRows) I->sp[lane] = t.sp;
  if (lane < kRealRows) I->sfirst[lane] = t.lo;
  if (lane <= kSRows) I->rs[lane] = t.rs;
  if (lane < kSRows) I->rlo[lane] = t.wlo;
  // the real rows' self offsets: real row r sits in staged row (ry + 1) + 5 (rz + 1)
  const int gy = lane % kSRow, gz = lane / kSRow;
  if (lane < kSRows && gy >= 1 && gy <= t.ny && gz >= 1 && gz <= t.nz) {
    const int r = (gy - 1) + kBlkRows * (gz - 1);
    I->selfoff[r] = t.rs + t.run.na - t.run.b0 - t.wlo;
  }
  *nreal_out = t.nreal;
  return t.S;
}

// The slot of every staged particle of a tile into sidx[0, S), by one
// workgroup (every thread calls it).
__device__ inline void blk_write_sidx(const Geom &g, const Cells &c, const BlkInfo *I,
                                      int *__restrict__ sidx) {
  __shared__ RowRun runs[kSRows];
  __shared__ int rs[kSRows + 1], rlo[kSRows];
  const int tid = threadIdx.x;
  if (tid < kSRows) {
    const int gy = tid % kSRow, gz = tid / kSRow;
    RowRun r{};
    if (gy < I->ny + 2 && gz < I->nz + 2) r = row_run(g, c, I->y0 - 1 + gy, I->z0 - 1 + gz);
    runs[tid] = r;
    rlo[tid] = I->rlo[tid];
  }
  if (tid <= kSRows) rs[tid] = I->rs[tid];
  __syncthreads();
  const int S = rs[kSRows];
  for (int u = tid; u < S; u += blockDim.x) {
    int gr = 0;
    while (u >= rs[gr + 1]) gr++;
    sidx[u] = run_slot(runs[gr], rlo[gr] + (u - rs[gr]));
  }
  __syncthreads();
}

// Stage a tile: its S particles, load(sidx[u]) for staged index u, into s3
// as interleaved (x, y, z), and the sentinel after them.  kB loads per
// thread go out together; kB NT covers the largest staging.  Every thread
// calls it; it ends with a barrier.
template <int NT, class Load>
__device__ __forceinline__ void stage_tile(int S, const int *__restrict__ sidx, double *s3,
                                           Load load) {
  constexpr int kB = (kMaxStaged + NT) / NT;
  for (int u0 = 0; u0 < S; u0 += kB * NT) {
    int sl[kB];
#pragma unroll
    for (int j = 0; j < kB; j++) {
      const int u = u0 + j * NT + (int)threadIdx.x;
      sl[j] = sidx[u < S ? u : 0];
    }
    double4 q[kB];
#pragma unroll
    for (int j = 0; j < kB; j++) q[j] = load(sl[j]);
#pragma unroll
    for (int j = 0; j < kB; j++) {
      const int u = u0 + j * NT + (int)threadIdx.x;
      if (u < S) {
        s3[3 * u] = q[j].x;
        s3[3 * u + 1] = q[j].y;
        s3[3 * u + 2] = q[j].z;
      }
    }
  }
  if (threadIdx.x == 0) s3[3 * S] = s3[3 * S + 1] = s3[3 * S + 2] = kFar;
  __syncthreads();
}

// Partners of one list column (entries [0, m), column base lst, kBlk
// columns) in list order: f(o) with o = 3 x the partner's staging index;
// sentinel entries (>= lim = 24 S) are skipped.  One 16-byte load per group.
template <class F>
__device__ __forceinline__ void for_each_partner(const unsigned short *__restrict__ lst, int m,
                                                 unsigned lim, F f) {
  for (int k = 0; k < m; k += kILP) {
    const uint4 w = *reinterpret_cast<const uint4 *>(lst + list_index<kBlk>(k));
    const unsigned h[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int u = 0; u < kILP; u++) {
      const unsigned e = (h[u / 2] >> (16 * (u % 2))) & 0xffffu;
      if (e < lim) f((int)(e >> 3));
    }
  }
}

// Partners of a real at qi among staged candidates [lo, hi) (candidate o
// at s3 + 3 o): r^2 <= SL2 (list_r2) and not the real itself, put in
// candidate order.  Four candidates are tested at a time so that their LDS
// reads are in flight together.
template <class Writer>
__device__ __forceinline__ void range_partners(const double *s3, const double4 &qi, double SL2,
                                               int lo, int hi, int self, Writer &gw) {
  constexpr int kU = 4;
  for (int o = lo; o < hi; o += kU) {
    bool ok[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const double *q = s3 + 3 * min(o + u, hi - 1);
      ok[u] = o + u < hi && list_r2(q[0], q[1], q[2], qi.x, qi.y, qi.z) <= SL2 && o + u != self;
    }
#pragma unroll
    for (int u = 0; u < kU; u++)
      if (ok[u]) gw.put(staged_entry(o + u));
  }
}

// The x window of the cell row (dy, dz) around a real at qi of cell row
// (y, z): every candidate of that row with list_r2 <= SL2 has |x - qi.x| <=
// h.  gy, gz are lower bounds of its distance to the row in y and z (the
// cell's faces (c - 1) w, c w, less eps for the rounding of the faces and of
// the binning), and the slack 1e-10 SL2 covers the roundings of r^2, so h
// is only ever too wide.  Returns false when no candidate of the row can be
// in range.
__device__ __forceinline__ bool row_window(const Geom &g, const double4 &qi, int y, int z,
                                           int dy, int dz, double eps, double *h) {
  double gy = 0.0, gz = 0.0;
  if (dy < 0) gy = qi.y - (double)(y - 1) / g.inv_w[1];
  if (dy > 0) gy = (double)y / g.inv_w[1] - qi.y;
  if (dz < 0) gz = qi.z - (double)(z - 1) / g.inv_w[2];
  if (dz > 0) gz = (double)z / g.inv_w[2] - qi.z;
  gy = fmax(gy - eps, 0.0);
  gz = fmax(gz - eps, 0.0);
  const double h2 = g.SL2 * (1.0 + 1e-10) - gy * gy - gz * gz;
  if (h2 < 0.0) return false;
  *h = sqrt(h2) + eps;
  return true;
}

// Rank of this thread's key among the keys of the workgroup's active
// threads: the number of greater keys plus the number of equal keys held
// by lower threads (keys below kRankBins).  Per wavefront, a ballot per
// distinct key counts the lanes holding it and gives each lane its place
// among them; the counts, summed over the wavefronts before and over the
// greater keys, make the rank.  Every thread calls it.
constexpr int kRankBins = 64;
__device__ __forceinline__ int block_rank(bool act, int key, int (*cw)[kRankBins], int *base) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int e = threadIdx.x; e < nw * kRankBins; e += blockDim.x) cw[e / kRankBins][e % kRankBins] = 0;
  __syncthreads();
  int tie = 0;
  unsigned long long todo = __ballot(act);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int lkey = __shfl(key, leader);
    const bool mine = act && key == lkey;
    const unsigned long long same = __ballot(mine);
    if (mine) tie = __popcll(same & ((1ull << lane) - 1ull));
    if (lane == leader) cw[w][lkey] = __popcll(same);
    todo &= ~same;
  }
  __syncthreads();
  if ((int)threadIdx.x < kRankBins) {   // exclusive prefix over wavefronts; bin totals
    const int b = threadIdx.x;
    int s = 0;
    for (int q = 0; q < nw; q++) {
      const int c = cw[q][b];
      cw[q][b] = s;
      s += c;
    }
    base[b] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {   // keys above each bin
    int s = 0;
    for (int b = kRankBins - 1; b >= 0; b--) {
      const int c = base[b];
      base[b] = s;
      s += c;
    }
  }
  __syncthreads();
  return act ? base[key] + cw[w][key] + tie : 0;
}

// Bucket starts of a tile's staged rows: the staging window [x0, x0 + E)
// (x0 = xa - h of the tile, E its width) cut into kBuckets equal buckets;
// btab[gr][b] is the first staged index of row gr whose x falls in bucket b
// or later (the row's end if none), btab[gr][kBuckets] the row's end.  x
// order within a row makes a bucket's particles one run, so the run of a
// window [v0, v1] is contained in btab[gr][bucket(v0)] ..
// btab[gr][bucket(v1) + 1]: two independent LDS reads in place of two
// binary searches.  bucket() is monotonic in x, so the range only ever
// holds more candidates than the window, never fewer.
constexpr int kBuckets = 64;
// LDS the list build takes beside the staging (ranks, columns, bucket
// starts), rounded up: what a tile's staging may use is the rest
constexpr int kBuildStaticLds = 20 * 1024;
struct BucketMap {
  double x0, inv;   // bucket(x) = floor((x - x0) inv), clamped
  __device__ __forceinline__ int operator()(double x) const {
    const double f = (x - x0) * inv;
    return f <= 0.0 ? 0 : (f >= (double)(kBuckets - 1) ? kBuckets - 1 : (int)f);
  }
};

// Fill btab for a tile staged at s3 with row starts rs (every thread calls
// it; it ends with a barrier).
__device__ inline void fill_buckets(const double *s3, const int *rs, const BucketMap &bk,
                                    int (*btab)[kBuckets + 1]) {
  const int S = rs[kSRows];
  for (int e = threadIdx.x; e < kSRows * (kBuckets + 1); e += blockDim.x)
    btab[e / (kBuckets + 1)][e % (kBuckets + 1)] = rs[e / (kBuckets + 1) + 1];
  __syncthreads();
  // particle u writes the buckets from its predecessor's (exclusive) to its
  // own: each bucket gets the first particle at or past it
  for (int u = threadIdx.x; u < S; u += blockDim.x) {
    int lo = 0, hi = kSRows;   // the row holding u: rs[gr] <= u < rs[gr + 1]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (rs[mid] <= u) lo = mid;
      else hi = mid;
    }
    const int gr = lo;
    const int bu = bk(s3[3 * u]);
    const int bp = u == rs[gr] ? -1 : bk(s3[3 * (u - 1)]);
    for (int b = bp + 1; b <= bu; b++) btab[gr][b] = u;
  }
  __syncthreads();
}

// Full neighbour list of every real of a tile, by a workgroup of kBlk
// threads (every thread calls it): every partner j != i (real or image) of
// its 27 cells with r^2 <= SL^2, r^2 formed as the reference's AVX2 search,
// fma(dx,dx, fma(dy,dy, dz*dz)) with dq = q_j - q_i (meshlist.cc:354-362).
// The partners go plane by plane (dz = -1, 0, +1), row by row (dy = -1, 0,
// +1), in x order within a row; the list is padded at its end with the
// sentinel to a multiple of kILP, and nn[i] is the padded length.
//
// The lanes take the tile's reals grouped by their place in their cell's
// y-z cross-section (4 x 4 classes), which sets the x windows of the
// off-centre rows, so a wavefront walks windows of similar length; which
// lane builds a real changes nothing in its list.  Each list goes first to
// the lane's column of the build region (list_nat); then the columns are
// ranked by list length, longest first, ties by the real's place in the
// tile's rows (slot order, so that a wavefront of the kick loads and stores
// runs of consecutive reals), and each column of the tile's region (list)
// takes its rank's list; treal gets the real of each column and tmeta (if
// given) its staging index and list length (low and high 16 bits).  pos holds
// every staged particle at its slot; stage_mem is the workgroup's dynamic
// LDS, at least 24 (S + 1) bytes.  Returns the thread's list length (before
// the kcap cut; 0 for a lane without a real).
__device__ inline int build_tile_lists(const double4 *__restrict__ pos, const BlkTile &t,
                                       const BlkInfo *I, const int *__restrict__ tsidx,
                                       const Geom &g, int kcap, unsigned short *__restrict__ lnat,
                                       unsigned short *__restrict__ lst, int *__restrict__ nn,
                                       int *__restrict__ treal,
                                       unsigned *__restrict__ tmeta, double *stage_mem) {
  __shared__ int keys[kBlk];
  __shared__ int rank_cw[kBlk / 64][kRankBins], rank_base[kRankBins];
  __shared__ int rs[kSRows + 1];
  __shared__ int csrc[kBlk], glen[kBlk];
  __shared__ int btab[kSRows][kBuckets + 1];
  const int lane = threadIdx.x;
  if (lane <= kSRows) rs[lane] = I->rs[lane];
  const bool act = lane < t.nreal;
  int r = 0;
  while (r < kRealRows - 1 && lane >= I->sp[r + 1]) r++;
  int key = 0;
  if (act) {
    const double4 q0 = pos[I->sfirst[r] + (lane - I->sp[r])];
    const int y0r = I->y0 + r % kBlkRows - 1, z0r = I->z0 + r / kBlkRows - 1;
    const int cy = (int)((q0.y * g.inv_w[1] - (double)y0r) * 4.0);
    const int cz = (int)((q0.z * g.inv_w[2] - (double)z0r) * 4.0);
    key = min(max(cy, 0), 3) * 4 + min(max(cz, 0), 3);
  }
  const int place = block_rank(act, key, rank_cw, rank_base);
  if (act) keys[place] = lane;
  __syncthreads();
  int src = 0;
  if (act) {
    src = keys[lane];   // this lane's real, by its place in the tile's rows
    r = 0;
    while (r < kRealRows - 1 && src >= I->sp[r + 1]) r++;
  }
  const int i = act ? I->sfirst[r] + (src - I->sp[r]) : 0;
  const double4 qi = act ? pos[i] : make_double4(0.0, 0.0, 0.0, 0.0);
  const int self = act ? i + I->selfoff[r] : -1;
  const int ry = r % kBlkRows, rz = r / kBlkRows, y = I->y0 + ry, z = I->z0 + rz;
  stage_tile<kBlk>(t.S, tsidx, stage_mem, [&](int slot) { return pos[slot]; });
  // rounding allowance of the cell faces and the binning, far above it
  const double eps = 1e-9 * (g.L[0] + g.L[1] + g.L[2]);
  // the tile's staging window, as blk_cut_tile cut it, in kBuckets buckets
  BucketMap bk{0.0, 0.0};
  if (t.nreal > 0) {
    const double ht = sqrt(g.SL2) * (1.0 + 1e-10) + eps;
    bk.x0 = I->xa - ht;
    bk.inv = (double)kBuckets / (I->xb + ht - bk.x0);
  }
  fill_buckets(stage_mem, rs, bk, btab);
  GroupWriter<kBlk> gw(lnat + (size_t)lane * kILP, kcap);
  if (act) {
    for (int p = 0; p < kPlanes; p++)
      for (int dy = -1; dy <= 1; dy++) {
        const int gr = (ry + 1 + dy) + kSRow * (rz + p);
        double h;
        if (!row_window(g, qi, y, z, dy, p - 1, eps, &h)) continue;
        const int lo = btab[gr][bk(qi.x - h)], hi = btab[gr][bk(qi.x + h) + 1];
        range_partners(stage_mem, qi, g.SL2, lo, hi, self, gw);
      }
    gw.pad_end(t.S);
  }
  const int cnt = gw.cnt, m = min(cnt, kcap);
  if (act) nn[i] = m;
  // rank by list length: thread k ranks the real at place k, and the lane
  // that built it hands its column to the rank's column
  if (act) keys[src] = m;
  __syncthreads();
  const bool actk = lane < t.nreal;
  const int mk = actk ? keys[lane] : 0;
  int rank = 0;
  if (kcap / kILP < kRankBins) {
    rank = block_rank(actk, mk / kILP, rank_cw, rank_base);
  } else if (actk) {
    for (int l = 0; l < t.nreal; l++) {
      const int kl = keys[l];
      rank += (kl > mk || (kl == mk && l < lane)) ? 1 : 0;
    }
  }
  __syncthreads();   // every rank read its keys
  if (actk) keys[lane] = col_of_rank(rank);
  __syncthreads();
  // Each column takes its list from the column its real was built in:
  // reads spread over the tile's rows (the workgroup just wrote them, they
  // sit in L2), stores in whole runs, no barrier between the groups.  Every
  // column once: a real's rank's column, or an unused rank's.
  const int col = act ? keys[src] : col_of_rank(lane);
  csrc[col] = lane;
  glen[col] = act ? m / kILP : 0;
  if (act) treal[col] = i;
  if (act && tmeta) tmeta[col] = (unsigned)self | ((unsigned)m << 16);
  __syncthreads();
  {
    const uint4 *from = reinterpret_cast<const uint4 *>(lnat) + csrc[lane];
    uint4 *to = reinterpret_cast<uint4 *>(lst) + lane;
    const int ng = glen[lane];
    for (int gr = 0; gr < ng; gr++) to[(size_t)gr * kBlk] = from[(size_t)gr * kBlk];
  }
  __syncthreads();
  return act ? cnt : 0;
}

}  // namespace mdxk

#endif
