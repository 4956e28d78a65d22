// ensemble.hip — many independent periodic LJ systems ("replicas") of
// uneven sizes, densities and temperatures, advanced together on one GPU.
//
// Every replica follows exactly the step of a single system (the same
// binning, image rule, neighbour-list order, kick arithmetic and pairlist
// expiry), so each replica's trajectory is the one it would have run alone.
// What changes is the schedule: one captured graph per step advances all
// replicas, and a replica whose pairlist expired is rebuilt inside that
// graph while the others keep stepping.
//
// Layout in HBM: fixed per-replica regions of shared arrays.
//   reals    slots [rbase, rbase + n)                        of NR = sum n
//   images   slots NR + [ibase, ibase + 7n)  (a real has at most 7 images)
//   cells    [cbase, cbase + m^3)           padded cell grid of the replica
//   tiles    [tbase, tbase + rows + ceil(n / T))    T = tile size in reals
//   lists    [lbase, lbase + tcap * kcap * T)  u16 LDS offsets per tile
// so a rebuild of one replica writes its own regions only.
//
// Per step (one graph launch):
//   rebuild   nine small kernels over the replicas flagged by the last guard
//             (wrap, bin, images, counting sort, tiles, lists); they return
//             at once when no replica is flagged
//   step      one persistent kernel over the tiles of all replicas: stage the
//             tile's partners in LDS (images formed from their source's
//             mid-step position on the fly), kick, trailing half drift, and
//             the next step's leading half drift into the other mid-step
//             buffer
//   guard     one workgroup per replica: |p|max, the velocity budget and
//             CheckByDisplacement (pairlist.cc:26-121) for the next step
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mdx.h"
#include "mdx_common.h"
#include "mdx_tiles.h"

namespace {

using namespace mdxk;

constexpr int kRebuildBlocks = 1024;   // grid of the per-particle rebuild kernels
constexpr int kReplicaThreads = 1024;  // one workgroup per replica: scans, guard
constexpr int kMaxCand = 1024;         // CheckByDisplacement candidates held in LDS
constexpr int kMaxImages = 7;          // images per real
constexpr int kScanPer = 8;            // items per thread in the per-replica scans

enum EnsFlags : int {
  EF_INVALID_POS = 1, EF_STAGE = 2, EF_LIST = 4, EF_TOO_FAST = 8, EF_IMAGES = 16, EF_TILES = 32
};

struct Rep {
  Geom g;
  size_t lbase;      // first list entry of its tiles
  size_t sbase;      // first staging index entry of its tiles
  int n, rbase;      // reals: slots [rbase, rbase + n)
  int ibase, gcap;   // images: slots NR + [ibase, ibase + gcap)
  int cbase, ncell;  // padded cell grid
  int tbase, tcap;   // tile slots
  int kcap;          // list entries per real
  int level;         // block shape and reals per tile (kBlkLadder)
  int bbase, bcap;   // per-block tile counts: [bbase, bbase + bcap)
};

// Device-resident state machine of one replica (PairList, pairlist.h:10-31,
// plus the step bookkeeping).  Written by the rebuild, step and guard
// kernels; read back by the host between calls.
struct RepCtl {
  double budget;     // PairList::buffer_length
  double vmax2;      // max |p|^2 after the last step
  double cd2;        // candidate threshold of the running displacement check
  int need_rebuild;  // pairlist expired: rebuild before the next step
  int eflags;        // failure bits (EnsFlags) of the last rebuild / guard
  int error;         // mdx_status that stopped the replica, 0 while running
  int g;             // images in use
  int ntiles, stage_max, kmax;
  int steps, target;  // steps done since mdx_ensemble_set_state, steps requested
  int rebuilds;       // pairlist constructions since mdx_ensemble_set_state
  int par;            // mid-step buffer (qh0 / qh1) that holds the next step's q + p dt/2
  int stepped;        // a step ran since the last rebuild
  int cand;           // candidates of the last displacement check
  int check;          // the budget is spent: a displacement check runs after this step
  int check_slot;     // its position in the check list
  int mid;            // mid-step buffer of the step just done
  int phase;          // 1 if flagged for a rebuild at the start of this step
  unsigned long long dmax_bits;   // max squared displacement of the running check (bits of a
                                  // non-negative double, which order as integers)
};

struct EnsDev {
  const Rep *rep;
  RepCtl *ctl;
  int *meta;      // meta[0]: replicas flagged for a rebuild in this step; meta[1]: checks of
                  // the others; meta[2]: the flagged count of the last rebuild; meta[3]:
                  // checks of the rebuilt ones
  int *flagged;   // their ids
  int *checklist; // [2R] replicas that run a displacement check after this step: those not
                  // rebuilt in it first, the rebuilt ones from R on
  int R, NR, NI, NT;
  int stage_cap;  // particles one tile may stage (dynamic LDS / 24 B)
  int check_list_length;
  double dt, buffer, SL;
  PairK pk;       // pair-term constants with dt folded in
  double4 *qe;    // [NR + NI] reals at the end of the step; images as built
  double4 *qh0, *qh1;   // [NR] mid-step positions q + p dt/2 (two buffers)
  double4 *mom, *ref;   // [NR], [NR + NI] snapshot at the last build
  int *tag;
  int *gsrc, *gcode;    // [NI] image -> real slot, shift code
  double4 *pos_tmp, *mom_tmp, *gpos_tmp;
  int *tag_tmp, *gsrc_tmp, *gcode_tmp;
  int *cellid, *slot, *nimg, *goff, *order, *inv;
  int *cell_count, *cnt_in, *cnt_out, *scan_in, *scan_out, *cell_start;
  BlkTile *btiles;      // [NT] tiles, each replica's at [tbase, tbase + tcap); pad = replica
  BlkInfo *binfo;       // [NT]
  int *treal;           // [NT][kBlk] real of each list column
  int *tile_S;          // [NT] staged particles of each tile
  int *sidx;            // staging index, each replica's from sbase on (tile s0 relative to it)
  int *blk_nt, *blk_toff;   // per block of cell rows: tiles, first tile
  unsigned short *list;     // lists in ranked columns
  unsigned short *list_nat; // the build's columns
  int *nn;
  double *partial;      // [NT] max |p|^2 per tile
  double *disp;         // [NR + NI] squared displacement since the build
  int *foff;            // [kOffKinds][R + 1] prefix sums over the flagged / checked replicas
  int *cand_buf;        // [R][kMaxCand] candidates of each check
  double *obs_partial;  // [3 NT] KE, PE, virial per tile
  int *tile_kmax;       // [NT]
};

// Index spaces of the flagged replicas (those rebuilding in this step) laid
// end to end, so that one launch covers every flagged particle, image, cell
// or tile with all its workgroups: foff[kind][f] is the first index of the
// f-th flagged replica (a replica stopped by a failure counts zero), and
// the last entry the total.  OFF_CHECK (OFF_CHECK2) is the same over the
// replicas not rebuilt (rebuilt) in this step whose
// displacement check runs in this step (reals and images).
enum OffKind {
  OFF_REALS = 0, OFF_IMAGES = 1, OFF_CELLS = 2, OFF_TILES = 3, OFF_CHECK = 4, OFF_CHECK2 = 5,
  OFF_BLOCKS = 6, kOffKinds = 7
};

__device__ __forceinline__ const int *offsets(const EnsDev &d, int kind) {
  return d.foff + (size_t)kind * (d.R + 1);
}

// the replica (position f in its list) holding flat index t: the last f
// with off[f] <= t (replicas of size zero are passed over)
__device__ __forceinline__ int replica_at(const int *off, int nf, int t) {
  int lo = 0, hi = nf;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

// flagged for a rebuild in this step: expired and about to step (a replica
// that has finished its steps keeps its state until it is asked for more)
__device__ __forceinline__ bool rebuild_due(const RepCtl *c) {
  return c->need_rebuild && !c->error && c->steps < c->target;
}

// The expiry check runs in two groups that may overlap in time: group 1
// the replicas not flagged for a rebuild at the start of the step, group 2
// those flagged (c->phase, set by k_e_begin).  Each group has its own
// count, list and offsets.
__device__ __forceinline__ bool in_check_group(const RepCtl *c, int grp) {
  return (c->phase != 0) == (grp == 2);
}
__device__ __forceinline__ int check_kind(int grp) { return grp == 2 ? OFF_CHECK2 : OFF_CHECK; }
__device__ __forceinline__ int *check_count(const EnsDev &d, int kind) {
  return d.meta + (kind == OFF_CHECK2 ? 3 : 1);
}
__device__ __forceinline__ int *check_list(const EnsDev &d, int kind) {
  return d.checklist + (kind == OFF_CHECK2 ? d.R : 0);
}

__device__ __forceinline__ bool stepping(const RepCtl *c) {
  return !c->error && !c->need_rebuild && c->steps < c->target;
}

// ------------------------------------------------------------ block helpers
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(v, o);
    if (lane >= o) v += y;
  }
  return v;
}

// exclusive scan over the block (blockDim a multiple of 64, <= 1024)
__device__ int block_excl_scan(int v, int *total) {
  __shared__ int wsum[16];
  __shared__ int tot;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int incl = wave_incl_scan(v);
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  if (w == 0) {
    const int s = lane < nw ? wsum[lane] : 0;
    const int si = wave_incl_scan(s);
    if (lane < nw) wsum[lane] = si - s;
    if (lane == 63) tot = si;
  }
  __syncthreads();
  const int ex = incl - v + wsum[w];
  *total = tot;
  __syncthreads();
  return ex;
}

__device__ double block_max(double v) {
  __shared__ double wm[16];
  __shared__ double res;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_max(v);
  if (lane == 0) wm[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wm[0];
    for (int k = 1; k < nw; k++) m = fmax(m, wm[k]);
    res = m;
  }
  __syncthreads();
  const double r = res;
  __syncthreads();
  return r;
}

__device__ int block_max_int(int v) {
  __shared__ int wm[16];
  __shared__ int res;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
  if (lane == 0) wm[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = wm[0];
    for (int k = 1; k < nw; k++) m = max(m, wm[k]);
    res = m;
  }
  __syncthreads();
  const int r = res;
  __syncthreads();
  return r;
}

// ------------------------------------------------------------ tiles
// Block tiles (mdx_tiles.h) of one replica: its cells from rp.cbase, reals
// and images at the absolute slots the scans hold.
__device__ __forceinline__ Cells rep_cells(const EnsDev &d, const Rep &rp) {
  return Cells{d.scan_in, d.cnt_in, d.scan_out, d.cnt_out, rp.cbase, 0};
}
__device__ __forceinline__ int rep_blocks(const Rep &rp) {
  const int br = kBlkLadder[rp.level].rows;
  return blk_count(rp.g, 1, br) * blk_count(rp.g, 2, br);
}

// The position a tile stages for slot s: a real's from `reals`; an image's
// from `mid`, its source's mid-step position plus the shift (the image
// refresh of MDUnit::MakeBufferForBorderParticles, mdunit.cc:181-191, done
// while staging), or, with mid == nullptr, as stored at the build.
__device__ __forceinline__ double4 staged_pos(const EnsDev &d, const Rep &rp, const double4 *reals,
                                              const double4 *mid, int s) {
  if (s < d.NR) return reals[s];
  if (!mid) return d.qe[s];
  const int gi = s - d.NR;
  return image_of(mid[d.gsrc[gi]], d.gcode[gi], rp.g.L);
}

// Persistent tile loop: workgroup b runs on XCD b % 8; the workgroups of one
// XCD walk one contiguous eighth of the tile slots, so neighbouring rows of
// a replica share that XCD's L2.
struct SlotWalk {
  int s, end, stride;
};
__device__ __forceinline__ SlotWalk slot_walk(int NT) {
  const int chunk = (NT + 7) >> 3;
  const int x = blockIdx.x & 7, j = blockIdx.x >> 3, nw = gridDim.x >> 3;
  SlotWalk w;
  w.s = x * chunk + j;
  w.end = min(NT, (x + 1) * chunk);
  w.stride = nw;
  return w;
}

// ------------------------------------------------------------ rebuild
// MDManager::MakePairList (mdmanager.cc:388-417) for the flagged replicas.
__global__ void k_e_begin(EnsDev d) {
  const int r = blockIdx.x;
  RepCtl *c = d.ctl + r;
  const bool due = rebuild_due(c);
  // the step kernels split the replicas by this flag: those not rebuilding
  // step while the rebuild runs, the rebuilt ones after it
  if (threadIdx.x == 0) c->phase = due ? 1 : 0;
  if (!due) return;
  const Rep &rp = d.rep[r];
  for (int k = threadIdx.x; k < rp.ncell; k += blockDim.x) d.cell_count[rp.cbase + k] = 0;
  if (threadIdx.x == 0) {
    c->eflags = 0;
    d.flagged[atomicAdd(d.meta, 1)] = r;
  }
}

// The offsets of the kinds in `mask` (one workgroup): a block scan of the
// sizes in list order.  A replica whose rebuild has failed counts zero.
__global__ void __launch_bounds__(kReplicaThreads) k_e_offsets(EnsDev d, int mask) {
  for (int kind = 0; kind < kOffKinds; kind++) {
    if (!(mask & (1 << kind))) continue;
    const bool chk = kind == OFF_CHECK || kind == OFF_CHECK2;
    const int nf = chk ? *check_count(d, kind) : d.meta[0];
    const int *list = chk ? check_list(d, kind) : d.flagged;
    int *off = d.foff + (size_t)kind * (d.R + 1);
    int carry = 0;
    for (int base = 0; base < nf; base += blockDim.x) {
      const int f = base + threadIdx.x;
      int v = 0;
      if (f < nf) {
        const int r = list[f];
        const RepCtl *c = d.ctl + r;
        const Rep &rp = d.rep[r];
        if (chk) v = rp.n + c->g;
        else if (!c->eflags)
          v = kind == OFF_REALS    ? rp.n
              : kind == OFF_IMAGES ? c->g
              : kind == OFF_CELLS  ? rp.ncell
              : kind == OFF_BLOCKS ? rep_blocks(rp)
                                   : c->ntiles;
      }
      int tot;
      const int ex = block_excl_scan(v, &tot);
      if (f < nf) off[f] = carry + ex;
      carry += tot;
    }
    if (threadIdx.x == 0) off[nf] = carry;
    __syncthreads();
  }
}

// The offsets of one kind as the workgroup reads them: copied into LDS when
// they fit (the binary search of every item then reads LDS, not a chain of
// global loads), else the global array.  Every thread calls it.
constexpr int kOffLds = 256;
struct OffView {
  const int *off;
  int nf, total;
};
__device__ __forceinline__ OffView load_offsets(const EnsDev &d, int kind, int *lds) {
  const int nf = kind == OFF_CHECK || kind == OFF_CHECK2 ? *check_count(d, kind) : d.meta[0];
  const int *g = offsets(d, kind);
  OffView v{g, nf, g[nf]};
  if (nf + 1 <= kOffLds && v.total > 0) {
    for (int f = threadIdx.x; f <= nf; f += blockDim.x) lds[f] = g[f];
    __syncthreads();
    v.off = lds;
  }
  return v;
}

// Loop over the flat index space of one offsets kind: body(r, k) for item k
// of replica r, every workgroup taking a share.  Every thread calls it.
template <class Body>
__device__ __forceinline__ void for_flagged(const EnsDev &d, int kind, int *lds, Body body) {
  const int *list = kind == OFF_CHECK || kind == OFF_CHECK2 ? check_list(d, kind) : d.flagged;
  const OffView v = load_offsets(d, kind, lds);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < v.total; t += gridDim.x * blockDim.x) {
    const int f = replica_at(v.off, v.nf, t);
    body(list[f], t - v.off[f]);
  }
}

// wrap into the box (variables.cc:64-84), bin, count images (mdrect.cc:24-47);
// the pre-sort copies go to the *_tmp arrays
__global__ void k_e_wrap(EnsDev d) {
  __shared__ int soff[kOffLds];
  for_flagged(d, OFF_REALS, soff, [&](int r, int k) {
    {
      const Rep &rp = d.rep[r];
      const Geom &g = rp.g;
      const int i = rp.rbase + k;
      double4 q = d.qe[i];
      double x[3] = {q.x, q.y, q.z};
      bool bad = false;
      int c[3], cnt = 1;
      for (int a = 0; a < 3; a++) {
        if (x[a] < 0.0) x[a] += g.L[a];
        else if (x[a] > g.L[a]) x[a] -= g.L[a];
        if (!(x[a] >= 0.0 && x[a] <= g.L[a])) bad = true;
        c[a] = bad ? 1 : interior_coord(x[a], g.inv_w[a], g.ms[a]);
        const bool left = x[a] < d.SL && x[a] > 0.0;
        const bool right = x[a] > g.L[a] - d.SL && x[a] < g.L[a];
        cnt *= (left || right) ? 2 : 1;
      }
      if (bad) atomicOr(&d.ctl[r].eflags, EF_INVALID_POS);
      q.x = x[0];
      q.y = x[1];
      q.z = x[2];
      d.pos_tmp[i] = q;
      d.mom_tmp[i] = d.mom[i];
      d.tag_tmp[i] = d.tag[i];
      const int cl = rp.cbase + cell_linear(g, c[0], c[1], c[2]);
      d.cellid[i] = cl;
      d.slot[i] = wave_grouped_add(d.cell_count, cl);
      d.nimg[i] = cnt - 1;
    }
  });
}

// image offsets: exclusive scan of the image counts of each flagged replica
__global__ void __launch_bounds__(kReplicaThreads) k_e_img_scan(EnsDev d) {
  const int r = blockIdx.x;
  RepCtl *c = d.ctl + r;
  // the flag of k_e_begin, not need_rebuild: the check of the replicas not
  // rebuilding may run at the same time and set theirs
  if (!c->phase || c->eflags) return;
  const Rep &rp = d.rep[r];
  // kScanPer consecutive reals per thread: a block scan per kScanPer x 1024
  int carry = 0;
  for (int base = 0; base < rp.n; base += blockDim.x * kScanPer) {
    const int k0 = base + threadIdx.x * kScanPer;
    int v[kScanPer], sum = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; j++) {
      v[j] = k0 + j < rp.n ? d.nimg[rp.rbase + k0 + j] : 0;
      sum += v[j];
    }
    int tot;
    int run = carry + block_excl_scan(sum, &tot);
#pragma unroll
    for (int j = 0; j < kScanPer; j++) {
      if (k0 + j < rp.n) d.goff[rp.rbase + k0 + j] = run;
      run += v[j];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) {
    c->g = carry;
    if (carry > rp.gcap) c->eflags |= EF_IMAGES;
  }
}

// images of each real in the fixed (sz, sy, sx) enumeration, with their
// outer-layer cells (MDUnit::FindBorderParticles, mdunit.cc:153-192)
__global__ void k_e_images(EnsDev d) {
  __shared__ int soff[kOffLds];
  for_flagged(d, OFF_REALS, soff, [&](int r, int k) {
    if (d.ctl[r].eflags) return;
    const Rep &rp = d.rep[r];
    const Geom &g = rp.g;
    {
      const int i = rp.rbase + k;
      if (d.nimg[i] == 0) return;
      const double4 q = d.pos_tmp[i];
      const double x[3] = {q.x, q.y, q.z};
      int sh[3];
      for (int a = 0; a < 3; a++) {
        const bool left = x[a] < d.SL && x[a] > 0.0;
        const bool right = x[a] > g.L[a] - d.SL && x[a] < g.L[a];
        sh[a] = left ? 1 : (right ? -1 : 0);
      }
      const int cl = d.cellid[i] - rp.cbase;
      const int cx = cl % g.m[0], cy = (cl / g.m[0]) % g.m[1], cz = cl / (g.m[0] * g.m[1]);
      int idx = rp.ibase + d.goff[i];
      for (int bz = 0; bz <= (sh[2] != 0); bz++)
        for (int by = 0; by <= (sh[1] != 0); by++)
          for (int bx = 0; bx <= (sh[0] != 0); bx++) {
            if (bx == 0 && by == 0 && bz == 0) continue;
            const int s[3] = {bx ? sh[0] : 0, by ? sh[1] : 0, bz ? sh[2] : 0};
            double4 im;
            im.x = shifted(x[0], s[0], g.L[0]);
            im.y = shifted(x[1], s[1], g.L[1]);
            im.z = shifted(x[2], s[2], g.L[2]);
            im.w = 0.0;
            d.gpos_tmp[idx] = im;
            d.gsrc_tmp[idx] = i;
            d.gcode_tmp[idx] = (s[0] + 1) + 3 * (s[1] + 1) + 9 * (s[2] + 1);
            const int gx = s[0] > 0 ? g.m[0] - 1 : (s[0] < 0 ? 0 : cx);
            const int gy = s[1] > 0 ? g.m[1] - 1 : (s[1] < 0 ? 0 : cy);
            const int gz = s[2] > 0 ? g.m[2] - 1 : (s[2] < 0 ? 0 : cz);
            const int gcl = rp.cbase + cell_linear(g, gx, gy, gz);
            d.cellid[d.NR + idx] = gcl;
            d.slot[d.NR + idx] = wave_grouped_add(d.cell_count, gcl);
            idx++;
          }
    }
  });
}

// Per flagged replica: split the cell counts into reals (interior) and
// images (outer layer) and scan each into absolute slots (the counting sort
// of MeshList::MakeMesh, meshlist.cc:176-228).
__global__ void __launch_bounds__(kReplicaThreads) k_e_cells(EnsDev d) {
  const int r = blockIdx.x;
  RepCtl *c = d.ctl + r;
  if (!c->phase || c->eflags) return;   // as k_e_img_scan
  const Rep &rp = d.rep[r];
  const Geom &g = rp.g;
  int carry_in = 0, carry_out = 0;
  // kScanPer consecutive cells per thread: two block scans per kScanPer x 1024
  for (int base = 0; base < rp.ncell; base += blockDim.x * kScanPer) {
    const int k0 = base + threadIdx.x * kScanPer;
    int ci[kScanPer], co[kScanPer], sin_ = 0, sout = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; j++) {
      const int k = k0 + j;
      ci[j] = co[j] = 0;
      if (k < rp.ncell) {
        const int x = k % g.m[0], y = (k / g.m[0]) % g.m[1], z = k / (g.m[0] * g.m[1]);
        const bool inner = x >= 1 && x <= g.ms[0] && y >= 1 && y <= g.ms[1] && z >= 1 && z <= g.ms[2];
        const int cnt = d.cell_count[rp.cbase + k];
        ci[j] = inner ? cnt : 0;
        co[j] = inner ? 0 : cnt;
      }
      sin_ += ci[j];
      sout += co[j];
    }
    int tin, tout;
    int run_in = carry_in + block_excl_scan(sin_, &tin);
    int run_out = carry_out + block_excl_scan(sout, &tout);
#pragma unroll
    for (int j = 0; j < kScanPer; j++) {
      const int k = k0 + j;
      if (k < rp.ncell) {
        const int cl = rp.cbase + k;
        const int si = rp.rbase + run_in;
        const int so = d.NR + rp.ibase + run_out;
        d.cnt_in[cl] = ci[j];
        d.cnt_out[cl] = co[j];
        d.scan_in[cl] = si;
        d.scan_out[cl] = so;
        d.cell_start[cl] = ci[j] > 0 ? si : so;
      }
      run_in += ci[j];
      run_out += co[j];
    }
    carry_in += tin;
    carry_out += tout;
  }
}

// place pre-sort key k (real slot, or NR + image index) into its cell
__global__ void k_e_scatter(EnsDev d) {
  __shared__ int soff[kOffLds];
  auto place = [&](int key) { d.order[d.cell_start[d.cellid[key]] + d.slot[key]] = key; };
  for_flagged(d, OFF_REALS, soff, [&](int r, int k) {
    if (!d.ctl[r].eflags) place(d.rep[r].rbase + k);
  });
  __syncthreads();   // soff is reloaded for the images
  for_flagged(d, OFF_IMAGES, soff, [&](int r, int k) { place(d.NR + d.rep[r].ibase + k); });
}

// atomics fixed the cell, not the order in it: sort each cell by (x,
// pre-sort key), the single-system order (rows sorted by x as a whole)
__global__ void k_e_cell_sort(EnsDev d) {
  __shared__ int buf[kBlock / 64][kSortBuf];
  __shared__ double xbuf[kBlock / 64][kSortBuf];
  __shared__ int soff[kOffLds];
  const OffView v = load_offsets(d, OFF_CELLS, soff);
  const int w = threadIdx.x >> 6, nw = gridDim.x * (kBlock / 64);
  const auto xof = [&](int k) { return k < d.NR ? d.pos_tmp[k].x : d.gpos_tmp[k - d.NR].x; };
  for (int t = blockIdx.x * (kBlock / 64) + w; t < v.total; t += nw) {
    const int f = replica_at(v.off, v.nf, t);   // the same for the whole wavefront
    const int cl = d.rep[d.flagged[f]].cbase + (t - v.off[f]);
    wave_rank_sort_x(d.order + d.cell_start[cl], d.cell_count[cl], buf[w], xbuf[w], xof);
  }
}

// reals into cell order, snapshot (PairList::Init, pairlist.cc:11-23), and
// the first half drift of the coming step into the current mid-step buffer
__global__ void k_e_gather_reals(EnsDev d) {
  __shared__ int soff[kOffLds];
  const double dt2 = d.dt * 0.5;
  for_flagged(d, OFF_REALS, soff, [&](int r, int k) {
    if (d.ctl[r].eflags) return;
    const Rep &rp = d.rep[r];
    double4 *mid = d.ctl[r].par ? d.qh1 : d.qh0;
    {
      const int s = rp.rbase + k;
      const int key = d.order[s];
      const double4 q = d.pos_tmp[key];
      const double4 p = d.mom_tmp[key];
      d.qe[s] = q;
      d.ref[s] = q;
      d.mom[s] = p;
      d.tag[s] = d.tag_tmp[key];
      d.inv[key] = s;
      mid[s] = drift_half(q, p, dt2);
    }
  });
}

__global__ void k_e_gather_images(EnsDev d) {
  __shared__ int soff[kOffLds];
  for_flagged(d, OFF_IMAGES, soff, [&](int r, int k) {
    const int s = d.NR + d.rep[r].ibase + k;
    const int pre = d.order[s] - d.NR;
    const double4 q = d.gpos_tmp[pre];
    d.qe[s] = q;
    d.ref[s] = q;
    d.gsrc[s - d.NR] = d.inv[d.gsrc_tmp[pre]];
    d.gcode[s - d.NR] = d.gcode_tmp[pre];
  });
}

// ---- block tiles of the flagged replicas, after their sort and gathers
// (the tiles cut the x-sorted cell rows).  A replica whose rebuild already
// failed is passed over: its flags are read once per workgroup, from
// kernels before.

// Flags of replica r as one workgroup reads them (one load, then shared).
__device__ __forceinline__ int rep_flags(const EnsDev &d, int r) {
  __shared__ int f;
  __syncthreads();
  if (threadIdx.x == 0) f = d.ctl[r].eflags;
  __syncthreads();
  return f;
}

// tiles of every block of the flagged replicas, one workgroup per block
__global__ void k_e_blk_count(EnsDev d) {
  __shared__ int hist[kMaxCellsX + 1];
  __shared__ int soff[kOffLds];
  const OffView v = load_offsets(d, OFF_BLOCKS, soff);
  for (int item = blockIdx.x; item < v.total; item += gridDim.x) {
    const int f = replica_at(v.off, v.nf, item);
    const int r = d.flagged[f];
    if (rep_flags(d, r)) continue;
    const Rep &rp = d.rep[r];
    const int b = item - v.off[f];
    const int br = kBlkLadder[rp.level].rows, BY = blk_count(rp.g, 1, br);
    const Cells c = rep_cells(d, rp);
    blk_hist(rp.g, c, b % BY, b / BY, br, hist);
    const int nt = blk_ntiles(rp.g, c, d.qe, hist, b % BY, b / BY, br, kBlkLadder[rp.level].target,
                              d.stage_cap);
    if (threadIdx.x == 0) d.blk_nt[rp.bbase + b] = nt;
  }
}

// per flagged replica: each block's first tile, the tile count against the
// replica's tile slots (EF_TILES), and its unused slots emptied
__global__ void __launch_bounds__(kReplicaThreads) k_e_blk_scan(EnsDev d) {
  const int r = blockIdx.x;
  RepCtl *c = d.ctl + r;
  if (!c->phase || c->eflags) return;   // as k_e_img_scan
  const Rep &rp = d.rep[r];
  const int nb = rep_blocks(rp);
  int carry = 0;
  for (int base = 0; base < nb; base += blockDim.x) {
    const int b = base + threadIdx.x;
    const int v = b < nb ? d.blk_nt[rp.bbase + b] : 0;
    int tot;
    const int ex = block_excl_scan(v, &tot);
    if (b < nb) d.blk_toff[rp.bbase + b] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    c->ntiles = carry;
    c->stage_max = 0;
    if (carry > rp.tcap) c->eflags |= EF_TILES;
  }
  for (int s = carry + threadIdx.x; s < rp.tcap; s += blockDim.x) d.btiles[rp.tbase + s] = BlkTile{0, 0, 0, r};
}

// the tiles of every block of the flagged replicas, one workgroup per
// block, one wavefront per tile; a staging above the LDS or a tile above
// kBlk reals raises EF_STAGE with the largest staging
__global__ void k_e_blk_tiles(EnsDev d) {
  __shared__ int hist[kMaxCellsX + 1];
  __shared__ int soff[kOffLds];
  const OffView v = load_offsets(d, OFF_BLOCKS, soff);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int item = blockIdx.x; item < v.total; item += gridDim.x) {
    const int f = replica_at(v.off, v.nf, item);
    const int r = d.flagged[f];
    if (rep_flags(d, r) & ~EF_STAGE) continue;
    RepCtl *c = d.ctl + r;
    const Rep &rp = d.rep[r];
    const int b = item - v.off[f];
    const int br = kBlkLadder[rp.level].rows, BY = blk_count(rp.g, 1, br);
    const Cells cl = rep_cells(d, rp);
    blk_hist(rp.g, cl, b % BY, b / BY, br, hist);
    const int nt = d.blk_nt[rp.bbase + b], t0 = d.blk_toff[rp.bbase + b];
    for (int k = w; k < nt; k += nw) {
      const int s = rp.tbase + t0 + k;
      int nreal;
      // the replica's id goes with the tile in one store: the step of the
      // other replicas, running meanwhile, reads it to pass the tile over
      const int S = blk_make_tile(rp.g, cl, d.qe, hist, b % BY, b / BY, br, k, nt, r, d.btiles + s,
                                  d.binfo + s, &nreal);
      if (lane == 0) {
        d.tile_S[s] = S;
        if (S > d.stage_cap || nreal > kBlk) {
          atomicOr(&c->eflags, EF_STAGE);
          atomicMax(&c->stage_max, S);
        }
      }
    }
  }
}

// per flagged replica: each tile's first staging index entry (relative to
// the replica's)
__global__ void __launch_bounds__(kReplicaThreads) k_e_tile_scan(EnsDev d) {
  const int r = blockIdx.x;
  RepCtl *c = d.ctl + r;
  if (!c->phase || c->eflags) return;
  const Rep &rp = d.rep[r];
  int carry = 0;
  for (int base = 0; base < c->ntiles; base += blockDim.x) {
    const int k = base + threadIdx.x;
    const int v = k < c->ntiles ? d.tile_S[rp.tbase + k] : 0;
    int tot;
    const int ex = block_excl_scan(v, &tot);
    if (k < c->ntiles) d.btiles[rp.tbase + k].s0 = carry + ex;
    carry += tot;
  }
}

// the staging index of every tile of the flagged replicas, one workgroup per tile
__global__ void k_e_blk_sidx(EnsDev d) {
  __shared__ int soff[kOffLds];
  const OffView v = load_offsets(d, OFF_TILES, soff);
  for (int item = blockIdx.x; item < v.total; item += gridDim.x) {
    const int f = replica_at(v.off, v.nf, item);
    const int r = d.flagged[f];
    if (rep_flags(d, r)) continue;
    const Rep &rp = d.rep[r];
    const int s = rp.tbase + item - v.off[f];
    blk_write_sidx(rp.g, rep_cells(d, rp), d.binfo + s, d.sidx + rp.sbase + d.btiles[s].s0);
  }
}

// The neighbour lists of every tile of the flagged replicas
// (build_tile_lists, as the single system builds them), one workgroup per
// tile; the longest list of each tile to tile_kmax.
__global__ void __launch_bounds__(kBlk, 4) k_e_build(EnsDev d) {
  extern __shared__ double stage_mem[];
  __shared__ int soff[kOffLds];
  const OffView v = load_offsets(d, OFF_TILES, soff);
  for (int item = blockIdx.x; item < v.total; item += gridDim.x) {
    const int f = replica_at(v.off, v.nf, item);
    const int r = d.flagged[f];
    if (rep_flags(d, r)) continue;
    const Rep &rp = d.rep[r];
    const int k = item - v.off[f], s = rp.tbase + k;
    const BlkTile t = d.btiles[s];
    const size_t region = rp.lbase + (size_t)k * rp.kcap * kBlk;
    const int cnt = build_tile_lists(d.qe, t, d.binfo + s, d.sidx + rp.sbase + t.s0, rp.g, rp.kcap,
                                     d.list_nat + region, d.list + region, d.nn,
                                     d.treal + (size_t)s * kBlk, nullptr, stage_mem);
    const int km = block_max_int(cnt);
    if (threadIdx.x == 0) d.tile_kmax[s] = km;
  }
}

// close the rebuild: longest list against the capacity, then the reset of
// the velocity budget (pairlist.cc:11-23)
__global__ void k_e_end(EnsDev d) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    d.meta[2] = d.meta[0];   // the flagged count, for the step of the rebuilt replicas
    d.meta[0] = 0;
  }
  const int r = blockIdx.x;
  RepCtl *c = d.ctl + r;
  if (!c->phase) return;   // as k_e_img_scan
  const Rep &rp = d.rep[r];
  int km = 0;
  if (!c->eflags)
    for (int s = threadIdx.x; s < c->ntiles; s += blockDim.x) km = max(km, d.tile_kmax[rp.tbase + s]);
  km = block_max_int(km);
  if (threadIdx.x == 0) {
    int ef = c->eflags;
    c->kmax = km;
    if (km > rp.kcap) ef |= EF_LIST;
    c->eflags = ef;
    if (ef) {
      c->error = (ef & EF_INVALID_POS) ? MDX_ERR_INVALID_POS : MDX_ERR_CAPACITY;
    } else {
      c->budget = d.buffer;
      c->need_rebuild = 0;
      c->rebuilds++;
      c->stepped = 0;
    }
  }
}

// ------------------------------------------------------------ step
// One NVE step (MDManager::Calculate, mdmanager.cc:154-186) of every tile
// of every stepping replica: partners staged from the mid-step positions
// q + p dt/2 (images formed on the fly), p += F dt over the tile lists,
// q_end = q_mid + p dt/2 (fcalculator.cc:18-30), and the next step's
// q_end + p dt/2 into the other mid-step buffer.  The arithmetic is the
// single-system kick's (kick_term, kILP chains, drift_half) over the same
// lists, so every replica's sums are the single system's bit for bit.
__global__ void __launch_bounds__(kBlk) k_e_step(EnsDev d, int mode) {
  extern __shared__ double stage_mem[];
  const double dt2 = d.dt * 0.5;
  auto tile_step = [&](int slot) {
    const BlkTile t = d.btiles[slot];
    if (t.nreal == 0) return;
    const int r = t.pad;
    const RepCtl *c = d.ctl + r;
    // mode 1: only replicas not rebuilding in this step (they run beside
    // the rebuild, whose kernels touch only the rebuilding replicas);
    // mode 2: only the replicas rebuilt in this step
    if (mode != 0 && (c->phase != 0) != (mode == 2)) return;
    if (!stepping(c)) return;
    const Rep &rp = d.rep[r];
    const double4 *qh = c->par ? d.qh1 : d.qh0;
    double4 *qn = c->par ? d.qh0 : d.qh1;
    const int lane = threadIdx.x;
    const bool act = col_active(lane, t.nreal);
    const int i = act ? d.treal[(size_t)slot * kBlk + lane] : 0;
    double4 qi = make_double4(0.0, 0.0, 0.0, 0.0), pm = qi;
    int m = 0;
    if (act) {
      qi = qh[i];
      m = d.nn[i];
      pm = d.mom[i];
    }
    stage_tile<kBlk>(t.S, d.sidx + rp.sbase + t.s0, stage_mem,
                     [&](int s) { return staged_pos(d, rp, qh, qh, s); });
    const unsigned short *lst =
        d.list + rp.lbase + (size_t)(slot - rp.tbase) * rp.kcap * kBlk + (size_t)lane * kILP;
    const double CL2 = rp.g.CL2, xi = qi.x, yi = qi.y, zi = qi.z;
    // the single-system kick's sums: entry k feeds chain k % kChains
    double ax[kChains], ay[kChains], az[kChains];
#pragma unroll
    for (int u = 0; u < kChains; u++) ax[u] = ay[u] = az[u] = 0.0;
    if (act) {
      list_groups<kBlk>(lst, 0, m, stage_mem, [&](int u, double xj, double yj, double zj) {
        kick_term(xj, yj, zj, xi, yi, zi, d.pk, CL2, ax[u % kChains], ay[u % kChains],
                  az[u % kChains]);
      });
    }
    double v2 = 0.0;
    if (act) {
      double fx = 0.0, fy = 0.0, fz = 0.0;
#pragma unroll
      for (int u = 0; u < kChains; u++) {
        fx += ax[u];
        fy += ay[u];
        fz += az[u];
      }
      double4 p = pm;
      p.x += fx;
      p.y += fy;
      p.z += fz;
      d.mom[i] = p;
      v2 = speed2(p);
      const double4 q = drift_half(qi, p, dt2);
      d.qe[i] = q;
      qn[i] = drift_half(q, p, dt2);
    }
    v2 = block_max(v2);
    if (threadIdx.x == 0) d.partial[slot] = v2;
    __syncthreads();
  };
  if (mode == 2) {
    // the tiles of the replicas rebuilt in this step, as the list build
    // spreads them (meta[2]: how many were flagged)
    const int nf = d.meta[2];
    const int *off = offsets(d, OFF_TILES);
    const int total = off[nf];
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
      const int f = replica_at(off, nf, item);
      tile_step(d.rep[d.flagged[f]].tbase + item - off[f]);
    }
  } else {
    const SlotWalk w = slot_walk(d.NT);
    for (int slot = w.s; slot < w.end; slot += w.stride) tile_step(slot);
  }
}

// Expiry check for the next step: PairList::IsPairListExpired /
// UpdatePairListValidity / CheckByDisplacement (pairlist.cc:26-121), in five
// kernels so that the displacement and candidate passes of a large replica
// spread over many workgroups:
//   budget   per replica: |p|max of the step, the velocity budget, the
//            Too-fast test; a spent budget queues a displacement check
//   disp     all workgroups: |old - now|^2 of every real and image of the
//            queued replicas, with block maxima
//   dfold    per replica: max displacement -> candidate threshold
//   cand     all workgroups: candidates above the threshold
//   decide   per replica: too many candidates, or the pair test
// Images enter at their mid-step positions, as the reference's ghosts do
// after a step.  The budget arithmetic is rounded operation by operation as
// the host does it, and sqrt is the correctly rounded one.
__global__ void k_e_budget(EnsDev d, int grp) {
#pragma clang fp contract(off)
  const int r = blockIdx.x;
  RepCtl *c = d.ctl + r;
  if (!in_check_group(c, grp)) return;
  if (!stepping(c)) {
    // no step, no check: the later kernels of this step leave it alone
    __syncthreads();
    if (threadIdx.x == 0) c->check = 0;
    return;
  }
  const Rep &rp = d.rep[r];
  double v = 0.0;
  for (int s = threadIdx.x; s < c->ntiles; s += blockDim.x) v = fmax(v, d.partial[rp.tbase + s]);
  const double v2max = block_max(v);
  if (threadIdx.x != 0) return;
  const int par = c->par;
  double budget = c->budget;
  c->steps = c->steps + 1;
  c->mid = par;
  c->par = par ^ 1;
  c->stepped = 1;
  c->vmax2 = v2max;
  c->check = 0;
  if (budget > 0.0) {
    const double vmax = __dsqrt_rn(v2max);
    const double move = vmax * 2.0 * d.dt;
    if (move > d.buffer) {
      c->eflags |= EF_TOO_FAST;
      c->error = MDX_ERR_TOO_FAST;
      return;
    }
    budget = budget - move;
  }
  c->budget = budget;
  c->need_rebuild = 0;
  if (budget < 0.0) {
    c->check = 1;
    c->cand = 0;
    c->dmax_bits = 0ull;
    const int kind = check_kind(grp);
    const int f = atomicAdd(check_count(d, kind), 1);
    c->check_slot = f;
    check_list(d, kind)[f] = r;
  }
}

// position a displacement check sees for slot s of replica rp
__device__ __forceinline__ double4 check_pos(const EnsDev &d, const Rep &rp, const double4 *mid, int s) {
  return s < d.NR ? d.qe[s] : image_of(mid[d.gsrc[s - d.NR]], d.gcode[s - d.NR], rp.g.L);
}

__device__ __forceinline__ int check_slot_of(const EnsDev &d, const Rep &rp, int k) {
  return k < rp.n ? rp.rbase + k : d.NR + rp.ibase + (k - rp.n);
}

// displacement of every real and image of the checked replicas; the maximum
// per replica goes through one 64-bit atomic max per replica a wavefront
// touches (a maximum does not depend on the order it is taken in)
__global__ void k_e_disp(EnsDev d, int grp) {
  __shared__ int soff[kOffLds];
  const int kind = check_kind(grp);
  const int *list = check_list(d, kind);
  const OffView v = load_offsets(d, kind, soff);
  const int *off = v.off;
  const int nf = v.nf, total = v.total;
  const int lane = threadIdx.x & 63;
  const int stride = gridDim.x * blockDim.x;
  // every lane of a wavefront runs the same number of passes (the ballots
  // below need them all), idle lanes with r = -1
  for (int t0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63); t0 < total; t0 += stride) {
    const int t = t0 + lane;
    int r = -1;
    double d2 = 0.0;
    if (t < total) {
      const int f = replica_at(off, nf, t);
      r = list[f];
      const RepCtl *c = d.ctl + r;
      const Rep &rp = d.rep[r];
      const double4 *mid = c->mid ? d.qh1 : d.qh0;
      const int sl = check_slot_of(d, rp, t - off[f]);
      d2 = disp2(d.ref[sl], check_pos(d, rp, mid, sl));
      d.disp[sl] = d2;
    }
    unsigned long long todo = __ballot(r >= 0);
    while (todo) {
      const int leader = __ffsll((long long)todo) - 1;
      const int rl = __shfl(r, leader);
      const unsigned long long same = __ballot(r == rl) & todo;
      const double m = wave_max(r == rl ? d2 : 0.0);
      if (lane == leader)
        atomicMax(&d.ctl[rl].dmax_bits, (unsigned long long)__double_as_longlong(m));
      todo &= ~same;
    }
  }
}

__global__ void k_e_dfold(EnsDev d, int grp) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= d.R) return;
  RepCtl *c = d.ctl + r;
  if (!in_check_group(c, grp) || c->error || !c->check) return;
  const double dmax2 = __longlong_as_double((long long)c->dmax_bits);
  {
    const double c_distance = d.buffer - __dsqrt_rn(dmax2);
    c->cd2 = c_distance * c_distance;
  }
}

__global__ void k_e_cand(EnsDev d, int grp) {
  __shared__ int soff[kOffLds];
  for_flagged(d, check_kind(grp), soff, [&](int r, int k) {
    RepCtl *c = d.ctl + r;
    const int sl = check_slot_of(d, d.rep[r], k);
    if (d.disp[sl] > c->cd2) {
      const int at = atomicAdd(&c->cand, 1);
      if (at < kMaxCand) d.cand_buf[(size_t)r * kMaxCand + at] = sl;
    }
  });
}

__global__ void __launch_bounds__(kReplicaThreads) k_e_decide(EnsDev d, int grp) {
  __shared__ int hit;
  __shared__ double old_q[3 * kMaxCand], now_q[3 * kMaxCand];
  if (blockIdx.x == 0 && threadIdx.x == 0) *check_count(d, check_kind(grp)) = 0;
  const int r = blockIdx.x;
  RepCtl *c = d.ctl + r;
  if (!in_check_group(c, grp) || c->error || !c->check) return;
  const Rep &rp = d.rep[r];
  const int tn = rp.n + c->g;
  const int Nl = tn > d.check_list_length ? d.check_list_length : tn;
  const int l = c->cand;
  int expired = 0;
  if (l >= Nl) {
    expired = 1;
  } else if (l > 1) {
    // the candidates' build-time and present positions, staged in LDS for
    // the pair test (l < Nl <= kMaxCand)
    const double4 *mid = c->mid ? d.qh1 : d.qh0;
    const int *cand = d.cand_buf + (size_t)r * kMaxCand;
    if (threadIdx.x == 0) hit = 0;
    for (int a = threadIdx.x; a < l; a += blockDim.x) {
      const int sl = cand[a];
      const double4 o = d.ref[sl], w = check_pos(d, rp, mid, sl);
      old_q[3 * a] = o.x;
      old_q[3 * a + 1] = o.y;
      old_q[3 * a + 2] = o.z;
      now_q[3 * a] = w.x;
      now_q[3 * a + 1] = w.y;
      now_q[3 * a + 2] = w.z;
    }
    __syncthreads();
    auto at = [](const double *v, int a) { return make_double4(v[3 * a], v[3 * a + 1], v[3 * a + 2], 0.0); };
    // pairs (a, b > a) in 32 x 32 blocks of the upper triangle, thread
    // (ty, tx) of the 1024 taking pair (32 A + ty, 32 B + tx)
    const int nblk32 = (l + 31) / 32;
    const int ty = (int)threadIdx.x >> 5, tx = (int)threadIdx.x & 31;
    for (int A = 0; A < nblk32; A++)
      for (int B = A; B < nblk32; B++) {
        const int a = 32 * A + ty, b = 32 * B + tx;
        if (b <= a || b >= l) continue;
        if (disp2(at(old_q, a), at(old_q, b)) < rp.g.SL2) continue;
        if (disp2(at(now_q, a), at(now_q, b)) < rp.g.CL2) hit = 1;
      }
    __syncthreads();
    expired = hit;
  }
  if (threadIdx.x == 0) c->need_rebuild = expired;
}

// ------------------------------------------------------------ observe
// KE, PE and virial per tile (observer.cc:5-79 over the full list, halves
// for pairs counted from both sides); images at the positions the reference's
// ghosts hold: the last step's mid-step refresh, or the build if no step ran.
__global__ void __launch_bounds__(kBlk) k_e_observe(EnsDev d, double C0, double C2) {
  extern __shared__ double stage_mem[];
  __shared__ double red[3][kBlk / 64];
  const int slot = blockIdx.x;
  const BlkTile t = d.btiles[slot];
  const int r = t.pad;
  const RepCtl *c = d.ctl + r;
  const int lane = threadIdx.x;
  if (t.nreal == 0 || c->error || c->need_rebuild) {
    if (lane < 3) d.obs_partial[3 * (size_t)slot + lane] = 0.0;
    return;
  }
  const Rep &rp = d.rep[r];
  // reals at the end of the step; images at their mid-step refresh when a
  // step ran since the build, else as built
  const double4 *mid = c->stepped ? (c->par ? d.qh0 : d.qh1) : nullptr;
  double *sx = stage_mem, *sy = stage_mem + 1, *sz = stage_mem + 2;
  const bool act = col_active(lane, t.nreal);
  const int i = act ? d.treal[(size_t)slot * kBlk + lane] : 0;
  stage_tile<kBlk>(t.S, d.sidx + rp.sbase + t.s0, stage_mem,
                   [&](int s) { return staged_pos(d, rp, d.qe, mid, s); });
  double ke = 0.0, pe = 0.0, vir = 0.0;
  const int m = act ? d.nn[i] : 0;
  const double4 qi = act ? d.qe[i] : make_double4(0.0, 0.0, 0.0, 0.0);
  const unsigned short *lst =
      d.list + rp.lbase + (size_t)(slot - rp.tbase) * rp.kcap * kBlk + (size_t)lane * kILP;
  for_each_partner(lst, m, (unsigned)(kEntryBytes * t.S), [&](int o) {
    const double dx = qi.x - sx[o], dy = qi.y - sy[o], dz = qi.z - sz[o];
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 <= rp.g.CL2) observe_term(r2, C0, C2, pe, vir);
  });
  if (act) {
    const double4 p = d.mom[i];
    ke = 0.5 * p.x * p.x + 0.5 * p.y * p.y + 0.5 * p.z * p.z;
    pe *= 0.5;
    vir *= 0.5;
  }
  // fixed order: a butterfly in each wavefront, then the wavefronts in turn
  double v[3] = {ke, pe, vir};
  for (int k = 0; k < 3; k++)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off);
  if ((lane & 63) == 0)
    for (int k = 0; k < 3; k++) red[k][lane >> 6] = v[k];
  __syncthreads();
  if (lane < 3) {
    double s = red[lane][0];
    for (int q = 1; q < kBlk / 64; q++) s += red[lane][q];
    d.obs_partial[3 * (size_t)slot + lane] = s;
  }
}

// per replica, the tile sums in tile order (fixed tree)
__global__ void k_e_observe_fold(EnsDev d, double *__restrict__ out) {
  __shared__ double lds[3][kBlock];
  const int r = blockIdx.x, t = threadIdx.x;
  const Rep &rp = d.rep[r];
  const int nt = d.ctl[r].ntiles;
  double s0 = 0, s1 = 0, s2 = 0;
  for (int k = t; k < nt; k += kBlock) {
    const size_t b = 3 * (size_t)(rp.tbase + k);
    s0 += d.obs_partial[b];
    s1 += d.obs_partial[b + 1];
    s2 += d.obs_partial[b + 2];
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
  if (t < 3) out[3 * (size_t)r + t] = lds[t][0];
}

// ------------------------------------------------------------ state i/o
__global__ void k_e_load(const double *__restrict__ q3, const double *__restrict__ p3, int n,
                         double4 *__restrict__ qe, double4 *__restrict__ mom, int *__restrict__ tag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  qe[i] = make_double4(q3[3 * (size_t)i], q3[3 * (size_t)i + 1], q3[3 * (size_t)i + 2], 0.0);
  mom[i] = make_double4(p3[3 * (size_t)i], p3[3 * (size_t)i + 1], p3[3 * (size_t)i + 2], 0.0);
  tag[i] = i;
}

__global__ void k_e_scatter_by_tag(const double4 *__restrict__ src, const int *__restrict__ tag, int n,
                                   double *__restrict__ out3) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 v = src[i];
  const int t = tag[i];
  out3[3 * (size_t)t + 0] = v.x;
  out3[3 * (size_t)t + 1] = v.y;
  out3[3 * (size_t)t + 2] = v.z;
}

inline int nblk(long long n) { return (int)((n + kBlock - 1) / kBlock); }

thread_local std::string g_ens_create_error;

}  // namespace

// ====================================================================== host
struct mdx_ensemble {
  mdx_params prm{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // second stream of a step: the replicas not rebuilding step on it while
  // the rebuild runs on `stream` (default priority, like every stream here)
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  double SL = 0, C0 = 0, C2 = 0;
  int R = 0, NR = 0, NI = 0, NC = 0, NT = 0;
  int tile = 512;          // reals aimed at per tile at first: 512 or 256 (kBlkLadder level 0 or 3)
  int grid_mult = 2;       // persistent step grid = grid_mult * resident workgroups
  int step_grid = 0;
  int stage_cap = 0;       // particles a tile may stage (dynamic LDS / 24 B)
  size_t lds_limit = 65536;
  int num_cu = 256;
  std::vector<Rep> hrep;
  std::vector<RepCtl> hctl;
  std::vector<int> kcap;   // per replica list capacity
  Rep *rep = nullptr;
  RepCtl *ctl = nullptr;
  int *meta = nullptr, *flagged = nullptr, *checklist = nullptr, *cand_buf = nullptr;
  double *disp = nullptr;
  int *foff = nullptr;
  double4 *qe = nullptr, *qh0 = nullptr, *qh1 = nullptr, *mom = nullptr, *ref = nullptr,
          *pos_tmp = nullptr, *mom_tmp = nullptr, *gpos_tmp = nullptr, *init_q = nullptr,
          *init_p = nullptr;
  int *tag = nullptr, *tag_tmp = nullptr, *gsrc = nullptr, *gcode = nullptr, *gsrc_tmp = nullptr,
      *gcode_tmp = nullptr, *cellid = nullptr, *slot = nullptr, *nimg = nullptr, *goff = nullptr,
      *order = nullptr, *inv = nullptr;
  int *cell_count = nullptr, *cnt_in = nullptr, *cnt_out = nullptr, *scan_in = nullptr,
      *scan_out = nullptr, *cell_start = nullptr;
  BlkTile *btiles = nullptr;
  BlkInfo *binfo = nullptr;
  int *treal = nullptr, *tile_S = nullptr, *sidx = nullptr, *blk_nt = nullptr, *blk_toff = nullptr;
  unsigned short *list = nullptr, *list_nat = nullptr;
  size_t list_elems = 0, sidx_elems = 0;
  int NB = 0;              // per-block slots over all replicas
  int *nn = nullptr, *tile_kmax = nullptr;
  double *partial = nullptr, *obs_partial = nullptr, *obs_out = nullptr, *io = nullptr;
  bool have_state = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graph_dirty = true;
  hipEvent_t ev[2] = {};
  double last_ms = 0.0;
  long long last_steps = 0;
  std::string err;
};

namespace {

int efail(mdx_ensemble *e, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (e) e->err = buf;
  else g_ens_create_error = buf;
  return code;
}

#define EHIP(e, call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return efail((e), MDX_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));   \
  } while (0)

template <class T>
int ealloc(mdx_ensemble *e, T **p, size_t count) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (count == 0) count = 1;
  EHIP(e, hipMalloc((void **)p, count * sizeof(T)));
  return MDX_OK;
}

EnsDev dev_view(const mdx_ensemble *e) {
  EnsDev d{};
  d.rep = e->rep;
  d.ctl = e->ctl;
  d.meta = e->meta;
  d.flagged = e->flagged;
  d.checklist = e->checklist;
  d.disp = e->disp;
  d.foff = e->foff;
  d.cand_buf = e->cand_buf;
  d.R = e->R;
  d.NR = e->NR;
  d.NI = e->NI;
  d.NT = e->NT;
  d.stage_cap = e->stage_cap;
  d.check_list_length = e->prm.check_list_length;
  d.dt = e->prm.dt;
  d.buffer = e->prm.buffer_length;
  d.SL = e->SL;
  d.pk = make_pairk(e->C2, e->prm.dt);
  d.qe = e->qe;
  d.qh0 = e->qh0;
  d.qh1 = e->qh1;
  d.mom = e->mom;
  d.ref = e->ref;
  d.tag = e->tag;
  d.gsrc = e->gsrc;
  d.gcode = e->gcode;
  d.pos_tmp = e->pos_tmp;
  d.mom_tmp = e->mom_tmp;
  d.gpos_tmp = e->gpos_tmp;
  d.tag_tmp = e->tag_tmp;
  d.gsrc_tmp = e->gsrc_tmp;
  d.gcode_tmp = e->gcode_tmp;
  d.cellid = e->cellid;
  d.slot = e->slot;
  d.nimg = e->nimg;
  d.goff = e->goff;
  d.order = e->order;
  d.inv = e->inv;
  d.cell_count = e->cell_count;
  d.cnt_in = e->cnt_in;
  d.cnt_out = e->cnt_out;
  d.scan_in = e->scan_in;
  d.scan_out = e->scan_out;
  d.cell_start = e->cell_start;
  d.btiles = e->btiles;
  d.binfo = e->binfo;
  d.treal = e->treal;
  d.tile_S = e->tile_S;
  d.sidx = e->sidx;
  d.blk_nt = e->blk_nt;
  d.blk_toff = e->blk_toff;
  d.list = e->list;
  d.list_nat = e->list_nat;
  d.nn = e->nn;
  d.partial = e->partial;
  d.obs_partial = e->obs_partial;
  d.tile_kmax = e->tile_kmax;
  return d;
}

// dynamic LDS of the tile kernels: a tile's staging plus the sentinel
size_t stage_bytes(const mdx_ensemble *e) { return (size_t)(e->stage_cap + 1) * 3 * sizeof(double); }

int launch_check(mdx_ensemble *e, const char *what) {
  hipError_t er = hipGetLastError();
  if (er != hipSuccess) return efail(e, MDX_ERR_HIP, "launch %s: %s", what, hipGetErrorString(er));
  return MDX_OK;
}

// Tile slots of a replica at its block shape: one per block plus 1.25 x its
// reals over the reals aimed at per tile, room for the tiles that the
// 512-real limit (a crystal's rows are uneven) adds.  A dense replica whose
// blocks the staging limit cuts into narrower tiles grows its slots at its
// first build (EF_TILES), before it steps.  (Twice the reals measured 2 %
// slower over 256 replicas, 1.1 x grew slots mid-run and 7 % slower.)
int tile_slots(const Rep &rp) {
  const BlkShape sh = kBlkLadder[rp.level];
  const int nb = blk_count(rp.g, 1, sh.rows) * blk_count(rp.g, 2, sh.rows);
  return nb + (int)((5LL * rp.n + 4LL * sh.target - 1) / (4LL * sh.target)) + 8;
}

// Regions of every replica (tiles, staging index, lists, block counts) from
// hrep's sizes, in new buffers.  With `old` given, a replica whose tile
// slots and list width are unchanged keeps its tiles and lists (copied);
// the others are rebuilt before they step again.
int layout(mdx_ensemble *e, const std::vector<Rep> *old) {
  size_t tb = 0, lb = 0, sb = 0, bb = 0;
  for (int r = 0; r < e->R; r++) {
    Rep &rp = e->hrep[r];
    rp.tbase = (int)tb;
    rp.kcap = e->kcap[r];
    rp.lbase = lb;
    rp.sbase = sb;
    rp.bbase = (int)bb;
    rp.bcap = rp.g.ms[1] * rp.g.ms[2];
    tb += rp.tcap;
    lb += (size_t)rp.tcap * rp.kcap * kBlk;
    sb += (size_t)rp.tcap * (kMaxStaged + 1);
    bb += rp.bcap;
  }
  if (tb > (size_t)0x7fffffff || bb > (size_t)0x7fffffff)
    return efail(e, MDX_ERR_CAPACITY, "too many tile slots");
  BlkTile *nt = nullptr;
  BlkInfo *ni = nullptr;
  int *ntr = nullptr, *nts = nullptr, *nsx = nullptr, *nkm = nullptr;
  unsigned short *nl = nullptr;
  EHIP(e, hipMalloc((void **)&nt, std::max<size_t>(tb, 1) * sizeof(BlkTile)));
  EHIP(e, hipMalloc((void **)&ni, std::max<size_t>(tb, 1) * sizeof(BlkInfo)));
  EHIP(e, hipMalloc((void **)&ntr, std::max<size_t>(tb * kBlk, 1) * sizeof(int)));
  EHIP(e, hipMalloc((void **)&nts, std::max<size_t>(tb, 1) * sizeof(int)));
  EHIP(e, hipMalloc((void **)&nsx, std::max<size_t>(sb, 1) * sizeof(int)));
  EHIP(e, hipMalloc((void **)&nkm, std::max<size_t>(tb, 1) * sizeof(int)));
  EHIP(e, hipMalloc((void **)&nl, std::max<size_t>(lb, 1) * sizeof(unsigned short)));
  EHIP(e, hipMemsetAsync(nt, 0, std::max<size_t>(tb, 1) * sizeof(BlkTile), e->stream));
  if (old) {
    for (int r = 0; r < e->R; r++) {
      const Rep &o = (*old)[r], &n = e->hrep[r];
      if (o.tcap != n.tcap || o.kcap != n.kcap) continue;
      EHIP(e, hipMemcpyAsync(nt + n.tbase, e->btiles + o.tbase, sizeof(BlkTile) * o.tcap, hipMemcpyDeviceToDevice, e->stream));
      EHIP(e, hipMemcpyAsync(ni + n.tbase, e->binfo + o.tbase, sizeof(BlkInfo) * o.tcap, hipMemcpyDeviceToDevice, e->stream));
      EHIP(e, hipMemcpyAsync(ntr + (size_t)n.tbase * kBlk, e->treal + (size_t)o.tbase * kBlk, sizeof(int) * o.tcap * kBlk, hipMemcpyDeviceToDevice, e->stream));
      EHIP(e, hipMemcpyAsync(nts + n.tbase, e->tile_S + o.tbase, sizeof(int) * o.tcap, hipMemcpyDeviceToDevice, e->stream));
      EHIP(e, hipMemcpyAsync(nkm + n.tbase, e->tile_kmax + o.tbase, sizeof(int) * o.tcap, hipMemcpyDeviceToDevice, e->stream));
      EHIP(e, hipMemcpyAsync(nsx + n.sbase, e->sidx + o.sbase, sizeof(int) * o.tcap * (kMaxStaged + 1), hipMemcpyDeviceToDevice, e->stream));
      EHIP(e, hipMemcpyAsync(nl + n.lbase, e->list + o.lbase, sizeof(unsigned short) * o.tcap * o.kcap * kBlk, hipMemcpyDeviceToDevice, e->stream));
    }
  }
  EHIP(e, hipStreamSynchronize(e->stream));
  void *oldbufs[] = {e->btiles, e->binfo, e->treal, e->tile_S, e->sidx, e->tile_kmax, e->list};
  for (void *b : oldbufs)
    if (b) (void)hipFree(b);
  e->btiles = nt;
  e->binfo = ni;
  e->treal = ntr;
  e->tile_S = nts;
  e->sidx = nsx;
  e->tile_kmax = nkm;
  e->list = nl;
  e->list_elems = lb;
  e->sidx_elems = sb;
  e->NT = (int)tb;
  e->NB = (int)bb;
  int rc;
  if ((rc = ealloc(e, &e->list_nat, lb)) || (rc = ealloc(e, &e->partial, tb)) ||
      (rc = ealloc(e, &e->obs_partial, 3 * tb)) || (rc = ealloc(e, &e->blk_nt, bb)) ||
      (rc = ealloc(e, &e->blk_toff, bb)))
    return rc;
  EHIP(e, hipMemcpyAsync(e->rep, e->hrep.data(), sizeof(Rep) * e->R, hipMemcpyHostToDevice, e->stream));
  EHIP(e, hipStreamSynchronize(e->stream));
  e->graph_dirty = true;
  return MDX_OK;
}

// first block shape of a tile target: 512 reals (kBlkLadder level 0) or 256 (level 3)
int tile_level(int tile) { return tile == 512 ? 0 : 3; }

// Every replica at block shape `level` (or the one it already needed, if
// larger), new regions, nothing kept.
int layout_level(mdx_ensemble *e, int level) {
  for (auto &rp : e->hrep) {
    rp.level = std::max(rp.level, level);
    rp.tcap = tile_slots(rp);
  }
  return layout(e, nullptr);
}

// The largest tile staging the LDS allows: the list build's own LDS beside
// it, the sentinel's list entry within 16 bits.
int max_stage_cap(const mdx_ensemble *e) {
  const long long lds = ((long long)e->lds_limit - (long long)kBuildStaticLds) / (3 * (long long)sizeof(double)) - 1;
  return (int)std::min<long long>(kMaxStaged, lds);
}

// the masked rebuild after k_e_begin: bin, images, sort and gather, then
// the block tiles of the sorted rows and their lists
void enqueue_rebuild_body(mdx_ensemble *e, const EnsDev &d) {
  k_e_offsets<<<1, kReplicaThreads, 0, e->stream>>>(d, 1 << OFF_REALS);
  k_e_wrap<<<kRebuildBlocks, kBlock, 0, e->stream>>>(d);
  k_e_img_scan<<<e->R, kReplicaThreads, 0, e->stream>>>(d);
  k_e_offsets<<<1, kReplicaThreads, 0, e->stream>>>(d, 1 << OFF_IMAGES);
  k_e_images<<<kRebuildBlocks, kBlock, 0, e->stream>>>(d);
  k_e_cells<<<e->R, kReplicaThreads, 0, e->stream>>>(d);
  k_e_offsets<<<1, kReplicaThreads, 0, e->stream>>>(d, (1 << OFF_IMAGES) | (1 << OFF_CELLS) | (1 << OFF_BLOCKS));
  k_e_scatter<<<kRebuildBlocks, kBlock, 0, e->stream>>>(d);
  k_e_cell_sort<<<kRebuildBlocks, kBlock, 0, e->stream>>>(d);
  k_e_gather_reals<<<kRebuildBlocks, kBlock, 0, e->stream>>>(d);
  k_e_gather_images<<<kRebuildBlocks, kBlock, 0, e->stream>>>(d);
  k_e_blk_count<<<kRebuildBlocks, kBlock, 0, e->stream>>>(d);
  k_e_blk_scan<<<e->R, kReplicaThreads, 0, e->stream>>>(d);
  k_e_blk_tiles<<<kRebuildBlocks, kBlock, 0, e->stream>>>(d);
  k_e_tile_scan<<<e->R, kReplicaThreads, 0, e->stream>>>(d);
  k_e_offsets<<<1, kReplicaThreads, 0, e->stream>>>(d, 1 << OFF_TILES);
  k_e_blk_sidx<<<kRebuildBlocks, kBlock, 0, e->stream>>>(d);
  k_e_build<<<e->step_grid, kBlk, stage_bytes(e), e->stream>>>(d);
  k_e_end<<<e->R, kBlock, 0, e->stream>>>(d);
}

int enqueue_rebuild(mdx_ensemble *e) {
  const EnsDev d = dev_view(e);
  k_e_begin<<<e->R, kBlock, 0, e->stream>>>(d);
  enqueue_rebuild_body(e, d);
  return launch_check(e, "rebuild");
}

// the expiry check of one group of replicas on stream st
void enqueue_check(mdx_ensemble *e, const EnsDev &d, hipStream_t st, int grp) {
  k_e_budget<<<e->R, kBlock, 0, st>>>(d, grp);
  k_e_offsets<<<1, kReplicaThreads, 0, st>>>(d, 1 << (grp == 2 ? OFF_CHECK2 : OFF_CHECK));
  k_e_disp<<<kRebuildBlocks, kBlock, 0, st>>>(d, grp);
  k_e_dfold<<<(e->R + kBlock - 1) / kBlock, kBlock, 0, st>>>(d, grp);
  k_e_cand<<<kRebuildBlocks, kBlock, 0, st>>>(d, grp);
  k_e_decide<<<e->R, kReplicaThreads, 0, st>>>(d, grp);
}

// One step: k_e_begin flags the replicas to rebuild; then on `side` the
// others step and run their expiry check, while on `stream` the flagged
// ones rebuild, step and run theirs; the two join before the next step.
// Each replica's kernels run in the same order as one after another, so
// its arithmetic is unchanged; the two branches touch disjoint replicas and
// keep their own check lists.
int enqueue_step(mdx_ensemble *e) {
  const EnsDev d = dev_view(e);
  k_e_begin<<<e->R, kBlock, 0, e->stream>>>(d);
  EHIP(e, hipEventRecord(e->fork, e->stream));
  EHIP(e, hipStreamWaitEvent(e->side, e->fork, 0));
  k_e_step<<<e->step_grid, kBlk, stage_bytes(e), e->side>>>(d, 1);
  enqueue_check(e, d, e->side, 1);
  EHIP(e, hipEventRecord(e->join, e->side));
  enqueue_rebuild_body(e, d);
  k_e_step<<<e->step_grid, kBlk, stage_bytes(e), e->stream>>>(d, 2);
  enqueue_check(e, d, e->stream, 2);
  EHIP(e, hipStreamWaitEvent(e->stream, e->join, 0));
  return launch_check(e, "step");
}

// resident workgroups of the step kernel times grid_mult, a multiple of 8
int size_step_grid(mdx_ensemble *e) {
  int per_cu = 0;
  hipError_t er = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_e_step, kBlk, stage_bytes(e));
  if (er != hipSuccess || per_cu < 1) per_cu = 1;
  long long g = (long long)per_cu * e->num_cu * e->grid_mult;
  g = std::min<long long>(g, ((long long)e->NT + 7) / 8 * 8);
  g = std::max<long long>(8, (g + 7) / 8 * 8);
  e->step_grid = (int)g;
  return MDX_OK;
}

int read_ctl(mdx_ensemble *e) {
  EHIP(e, hipMemcpyAsync(e->hctl.data(), e->ctl, sizeof(RepCtl) * e->R, hipMemcpyDeviceToHost, e->stream));
  EHIP(e, hipStreamSynchronize(e->stream));
  return MDX_OK;
}

int write_ctl(mdx_ensemble *e) {
  EHIP(e, hipMemcpyAsync(e->ctl, e->hctl.data(), sizeof(RepCtl) * e->R, hipMemcpyHostToDevice, e->stream));
  return MDX_OK;
}

// Replicas whose rebuild failed on a capacity that can grow: grow it (a
// longer list width, more tile slots, or a smaller block shape when a
// tile's staging overflows) and clear the failure, so the rebuild is
// retried; the other replicas keep their tiles and lists.  Returns whether
// any was fixed.
int fix_capacity(mdx_ensemble *e, bool *fixed) {
  *fixed = false;
  const std::vector<Rep> old = e->hrep;
  bool any = false;
  for (int r = 0; r < e->R; r++) {
    RepCtl &c = e->hctl[r];
    if (c.error != MDX_ERR_CAPACITY) continue;
    if (c.eflags & EF_IMAGES) return efail(e, MDX_ERR_CAPACITY, "replica %d: image capacity", r);
    Rep &rp = e->hrep[r];
    if (c.eflags & EF_LIST) e->kcap[r] = ((std::max(e->kcap[r], c.kmax) * 5 / 4 + 16) + 7) & ~7;
    if (c.eflags & EF_STAGE) {
      if (rp.level + 1 >= kBlkLevels)
        return efail(e, MDX_ERR_CAPACITY,
                     "replica %d: a tile of one cell row stages %d particles, more than %d", r,
                     c.stage_max, e->stage_cap);
      rp.level++;
      rp.tcap = std::max(rp.tcap, tile_slots(rp));
    }
    if (c.eflags & EF_TILES) rp.tcap = std::max(rp.tcap + rp.tcap / 2, c.ntiles + 8);
    any = true;
  }
  if (!any) return MDX_OK;
  int rc;
  if ((rc = layout(e, &old)) || (rc = size_step_grid(e))) return rc;
  for (int r = 0; r < e->R; r++) {
    RepCtl &c = e->hctl[r];
    if (c.error == MDX_ERR_CAPACITY) {
      c.error = 0;
      c.eflags = 0;
      c.need_rebuild = 1;
    }
  }
  if ((rc = write_ctl(e))) return rc;
  *fixed = true;
  return MDX_OK;
}

int capture(mdx_ensemble *e) {
  if (!e->graph_dirty && e->exec) return MDX_OK;
  if (e->exec) {
    (void)hipGraphExecDestroy(e->exec);
    e->exec = nullptr;
  }
  if (e->graph) {
    (void)hipGraphDestroy(e->graph);
    e->graph = nullptr;
  }
  EHIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_step(e);
  hipGraph_t g = nullptr;
  hipError_t er = hipStreamEndCapture(e->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (er != hipSuccess) return efail(e, MDX_ERR_HIP, "graph capture: %s", hipGetErrorString(er));
  e->graph = g;
  EHIP(e, hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
  e->graph_dirty = false;
  return MDX_OK;
}

// the loaded state again, every pairlist expired, then the first build
int reset_state(mdx_ensemble *e) {
  EHIP(e, hipMemcpyAsync(e->qe, e->init_q, sizeof(double4) * e->NR, hipMemcpyDeviceToDevice, e->stream));
  EHIP(e, hipMemcpyAsync(e->mom, e->init_p, sizeof(double4) * e->NR, hipMemcpyDeviceToDevice, e->stream));
  std::vector<int> ids(e->NR);
  for (int i = 0; i < e->NR; i++) ids[i] = i;
  EHIP(e, hipMemcpyAsync(e->tag, ids.data(), sizeof(int) * e->NR, hipMemcpyHostToDevice, e->stream));
  for (int r = 0; r < e->R; r++) {
    RepCtl c{};
    c.budget = e->prm.buffer_length;
    c.need_rebuild = 1;
    c.target = 1;   // admits the build; no step runs here
    e->hctl[r] = c;
  }
  int rc = write_ctl(e);
  if (rc) return rc;
  EHIP(e, hipMemsetAsync(e->meta, 0, sizeof(int) * 4, e->stream));
  // the first step's pairlist construction, retried while a capacity grows
  for (int attempt = 0; attempt < 2 * kBlkLevels; attempt++) {
    if ((rc = enqueue_rebuild(e))) return rc;
    if ((rc = read_ctl(e))) return rc;
    bool fixed = false;
    if ((rc = fix_capacity(e, &fixed))) return rc;
    if (!fixed) break;
  }
  for (int r = 0; r < e->R; r++) {
    RepCtl &c = e->hctl[r];
    if (c.error) return efail(e, c.error, "replica %d: first pairlist build failed (flags %d)", r, c.eflags);
    c.target = 0;
  }
  return write_ctl(e);
}

// graph launches between two reads of the replicas' state, and the most
// capacity growth rounds one call makes
constexpr int kLaunchChunk = 64;
constexpr int kMaxFixRounds = 4 * kBlkLevels;

}  // namespace

extern "C" {

int mdx_ensemble_create(const mdx_params *prm, int nrep, const int *n, const double *L3,
                        mdx_ensemble **out) {
  if (!prm || !out || !n || !L3 || nrep < 1) return efail(nullptr, MDX_ERR_ARG, "bad arguments");
  *out = nullptr;
  const double SL = prm->cutoff + prm->buffer_length;
  if (!(prm->dt > 0.0) || !(prm->cutoff > 0.0) || !(prm->buffer_length > 0.0))
    return efail(nullptr, MDX_ERR_ARG, "dt, cutoff and buffer_length must be positive");
  if (prm->check_list_length < 1 || prm->check_list_length > kMaxCand)
    return efail(nullptr, MDX_ERR_ARG, "check_list_length must be in [1, %d]", kMaxCand);
  long long NR = 0, NI = 0;
  for (int r = 0; r < nrep; r++) {
    if (n[r] < 1) return efail(nullptr, MDX_ERR_ARG, "replica %d has no particles", r);
    for (int a = 0; a < 3; a++)
      if (!(L3[3 * r + a] >= 2.0 * SL))
        return efail(nullptr, MDX_ERR_BOX, "replica %d: box edge %g < 2 x search length %g", r,
                     L3[3 * r + a], SL);
    if (!(L3[3 * r] / SL < (double)kMaxCellsX))
      return efail(nullptr, MDX_ERR_BOX, "replica %d: box edge %g in x holds more than %d cells", r,
                   L3[3 * r], kMaxCellsX);
    NR += n[r];
    NI += (long long)kMaxImages * n[r];
  }
  if (NR + NI > 0x7fffffffLL) return efail(nullptr, MDX_ERR_CAPACITY, "too many particles");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= prm->device)
    return efail(nullptr, MDX_ERR_HIP, "HIP device %d not available", prm->device);
  if (hipSetDevice(prm->device) != hipSuccess)
    return efail(nullptr, MDX_ERR_HIP, "hipSetDevice(%d) failed", prm->device);
  mdx_ensemble *e = new mdx_ensemble();
  e->prm = *prm;
  e->SL = SL;
  {
    const double s2 = 1.0 / (prm->cutoff * prm->cutoff);
    const double s6 = s2 * s2 * s2, s8 = s6 * s2, s12 = s6 * s6, s14 = s12 * s2;
    e->C2 = 6.0 * s14 - 3.0 * s8;
    e->C0 = -s12 + s6 - e->C2 / s2;
  }
  e->R = nrep;
  e->NR = (int)NR;
  e->NI = (int)NI;
  e->hrep.resize(nrep);
  e->hctl.resize(nrep);
  e->kcap.resize(nrep);
  int rb = 0, ib = 0, cb = 0;
  for (int r = 0; r < nrep; r++) {
    Rep &rp = e->hrep[r];
    rp = Rep{};
    double V = 1.0;
    for (int a = 0; a < 3; a++) {
      const double L = L3[3 * r + a];
      rp.g.L[a] = L;
      rp.g.ms[a] = (int)(L / SL);   // MeshList::ChangeScale (meshlist.cc:41-65)
      rp.g.inv_w[a] = 1.0 / (L / (double)rp.g.ms[a]);
      rp.g.m[a] = rp.g.ms[a] + 2;
      V *= L;
    }
    rp.g.SL2 = SL * SL;
    rp.g.CL2 = prm->cutoff * prm->cutoff;
    rp.n = n[r];
    rp.rbase = rb;
    rp.ibase = ib;
    rp.gcap = kMaxImages * n[r];
    rp.cbase = cb;
    rp.ncell = rp.g.m[0] * rp.g.m[1] * rp.g.m[2];
    rb += n[r];
    ib += rp.gcap;
    cb += rp.ncell;
    // partners within SL at the mean density, with room for fluctuations
    const double mean = 4.18879 * SL * SL * SL * (n[r] / V);
    e->kcap[r] = ((int)(mean * 1.4) + 24 + (kILP - 1) + 7) & ~7;
    rp.level = 0;
  }
  e->NC = cb;
  int rc = MDX_OK;
  if (prm->stream) {
    e->stream = (hipStream_t)prm->stream;
  } else {
    if (hipStreamCreate(&e->stream) != hipSuccess) rc = efail(e, MDX_ERR_HIP, "hipStreamCreate failed");
    e->own_stream = true;
  }
  hipDeviceProp_t prop;
  if (!rc && hipGetDeviceProperties(&prop, prm->device) == hipSuccess) {
    e->lds_limit = prop.sharedMemPerBlock;
    e->num_cu = prop.multiProcessorCount;
  }
  for (auto &ev : e->ev)
    if (!rc && hipEventCreate(&ev) != hipSuccess) rc = efail(e, MDX_ERR_HIP, "hipEventCreate failed");
  if (!rc && (hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking) != hipSuccess ||
              hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&e->join, hipEventDisableTiming) != hipSuccess))
    rc = efail(e, MDX_ERR_HIP, "second stream or its events");
  const size_t tot = (size_t)NR + NI;
  if (!rc && ((rc = ealloc(e, &e->rep, nrep)) || (rc = ealloc(e, &e->ctl, nrep)) ||
              (rc = ealloc(e, &e->meta, 4)) || (rc = ealloc(e, &e->flagged, nrep)) ||
              (rc = ealloc(e, &e->checklist, 2 * (size_t)nrep)) || (rc = ealloc(e, &e->disp, tot)) ||
              (rc = ealloc(e, &e->foff, (size_t)kOffKinds * (nrep + 1))) ||
              (rc = ealloc(e, &e->cand_buf, (size_t)nrep * kMaxCand)) ||
              (rc = ealloc(e, &e->qe, tot)) || (rc = ealloc(e, &e->ref, tot)) ||
              (rc = ealloc(e, &e->qh0, NR)) || (rc = ealloc(e, &e->qh1, NR)) ||
              (rc = ealloc(e, &e->mom, NR)) || (rc = ealloc(e, &e->pos_tmp, NR)) ||
              (rc = ealloc(e, &e->mom_tmp, NR)) || (rc = ealloc(e, &e->gpos_tmp, NI)) ||
              (rc = ealloc(e, &e->init_q, NR)) || (rc = ealloc(e, &e->init_p, NR)) ||
              (rc = ealloc(e, &e->tag, NR)) || (rc = ealloc(e, &e->tag_tmp, NR)) ||
              (rc = ealloc(e, &e->gsrc, NI)) || (rc = ealloc(e, &e->gcode, NI)) ||
              (rc = ealloc(e, &e->gsrc_tmp, NI)) || (rc = ealloc(e, &e->gcode_tmp, NI)) ||
              (rc = ealloc(e, &e->cellid, tot)) || (rc = ealloc(e, &e->slot, tot)) ||
              (rc = ealloc(e, &e->order, tot)) || (rc = ealloc(e, &e->nimg, NR)) ||
              (rc = ealloc(e, &e->goff, NR)) || (rc = ealloc(e, &e->inv, NR)) ||
              (rc = ealloc(e, &e->nn, NR)) ||
              (rc = ealloc(e, &e->cell_count, cb)) ||
              (rc = ealloc(e, &e->cnt_in, cb)) || (rc = ealloc(e, &e->cnt_out, cb)) ||
              (rc = ealloc(e, &e->scan_in, cb)) || (rc = ealloc(e, &e->scan_out, cb)) ||
              (rc = ealloc(e, &e->cell_start, cb)) || (rc = ealloc(e, &e->obs_out, 3 * (size_t)nrep)) ||
              (rc = ealloc(e, &e->io, 3 * (size_t)NR))))
    ;
  if (!rc && hipMemsetAsync(e->meta, 0, 4 * sizeof(int), e->stream) != hipSuccess)
    rc = efail(e, MDX_ERR_HIP, "hipMemsetAsync failed");
  // every tile may stage as much as the LDS allows (two workgroups of kBlk
  // threads fit a CU at any staging up to it)
  if (!rc) {
    e->stage_cap = max_stage_cap(e);
    if (e->stage_cap < 1024) rc = efail(e, MDX_ERR_CAPACITY, "LDS of %zu B per workgroup", e->lds_limit);
  }
  if (!rc) rc = layout_level(e, 0);
  if (!rc) rc = size_step_grid(e);
  if (rc) {
    g_ens_create_error = e->err.empty() ? "allocation failed" : e->err;
    mdx_ensemble_destroy(e);
    return rc;
  }
  *out = e;
  return MDX_OK;
}

int mdx_ensemble_destroy(mdx_ensemble *e) {
  if (!e) return MDX_OK;
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  void *bufs[] = {e->rep, e->ctl, e->meta, e->flagged, e->checklist, e->disp, e->foff, e->cand_buf, e->qe, e->qh0, e->qh1, e->mom, e->ref,
                  e->pos_tmp, e->mom_tmp, e->gpos_tmp, e->init_q, e->init_p, e->tag, e->tag_tmp,
                  e->gsrc, e->gcode, e->gsrc_tmp, e->gcode_tmp, e->cellid, e->slot, e->nimg, e->goff,
                  e->order, e->inv, e->cell_count, e->cnt_in, e->cnt_out, e->scan_in, e->scan_out,
                  e->cell_start, e->btiles, e->binfo, e->treal, e->tile_S, e->sidx, e->blk_nt,
                  e->blk_toff, e->list, e->list_nat, e->nn, e->tile_kmax, e->partial,
                  e->obs_partial, e->obs_out, e->io};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (auto &ev : e->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->join) (void)hipEventDestroy(e->join);
  if (e->side) (void)hipStreamDestroy(e->side);
  if (e->own_stream && e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return MDX_OK;
}

const char *mdx_ensemble_last_error(const mdx_ensemble *e) {
  return e ? e->err.c_str() : g_ens_create_error.c_str();
}

int mdx_ensemble_size(const mdx_ensemble *e, int *nrep, int *ntotal) {
  if (!e) return MDX_ERR_ARG;
  if (nrep) *nrep = e->R;
  if (ntotal) *ntotal = e->NR;
  return MDX_OK;
}

int mdx_ensemble_set_state(mdx_ensemble *e, const double *q3, const double *p3) {
  if (!e || !q3 || !p3) return efail(e, MDX_ERR_ARG, "bad arguments to mdx_ensemble_set_state");
  double *pbuf = reinterpret_cast<double *>(e->pos_tmp);   // >= 3 NR doubles
  EHIP(e, hipMemcpyAsync(e->io, q3, sizeof(double) * 3 * (size_t)e->NR, hipMemcpyHostToDevice, e->stream));
  EHIP(e, hipMemcpyAsync(pbuf, p3, sizeof(double) * 3 * (size_t)e->NR, hipMemcpyHostToDevice, e->stream));
  k_e_load<<<nblk(e->NR), kBlock, 0, e->stream>>>(e->io, pbuf, e->NR, e->init_q, e->init_p, e->tag);
  int rc = launch_check(e, "load");
  if (rc) return rc;
  e->have_state = true;
  return reset_state(e);
}

int mdx_ensemble_get_state(mdx_ensemble *e, double *q3, double *p3) {
  if (!e) return efail(nullptr, MDX_ERR_ARG, "null ensemble");
  if (!e->have_state) return efail(e, MDX_ERR_STATE, "no state");
  if (q3) {
    k_e_scatter_by_tag<<<nblk(e->NR), kBlock, 0, e->stream>>>(e->qe, e->tag, e->NR, e->io);
    EHIP(e, hipMemcpyAsync(q3, e->io, sizeof(double) * 3 * (size_t)e->NR, hipMemcpyDeviceToHost, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
  }
  if (p3) {
    k_e_scatter_by_tag<<<nblk(e->NR), kBlock, 0, e->stream>>>(e->mom, e->tag, e->NR, e->io);
    EHIP(e, hipMemcpyAsync(p3, e->io, sizeof(double) * 3 * (size_t)e->NR, hipMemcpyDeviceToHost, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
  }
  return launch_check(e, "get_state");
}

int mdx_ensemble_calculate(mdx_ensemble *e, int nsteps, int *rebuilds) {
  if (!e || nsteps < 0) return efail(e, MDX_ERR_ARG, "bad arguments to mdx_ensemble_calculate");
  if (!e->have_state) return efail(e, MDX_ERR_STATE, "no state");
  int rc;
  if ((rc = read_ctl(e))) return rc;
  for (auto &c : e->hctl)
    if (!c.error) c.target = c.steps + nsteps;
  if ((rc = write_ctl(e))) return rc;
  if ((rc = capture(e))) return rc;
  EHIP(e, hipEventRecord(e->ev[0], e->stream));
  // The step graph goes out in chunks; between chunks the replicas stopped
  // on a capacity get it grown and rebuild in the next chunk, so one falls
  // at most a chunk behind the others, and the launches after the last
  // chunk, which step it alone, are at most that many.  The growth rounds
  // are bounded: a capacity that never fits ends the call.
  int launches = nsteps, fixes = 0;
  while (launches > 0) {
    const int chunk = std::min(launches, kLaunchChunk);
    for (int k = 0; k < chunk; k++) EHIP(e, hipGraphLaunch(e->exec, e->stream));
    if ((rc = read_ctl(e))) return rc;
    bool fixed = false;
    if ((rc = fix_capacity(e, &fixed))) return rc;
    if (fixed && ++fixes > kMaxFixRounds) break;
    if ((rc = capture(e))) return rc;
    launches = 0;
    for (const auto &c : e->hctl)
      if (!c.error) launches = std::max(launches, c.target - c.steps);
  }
  EHIP(e, hipEventRecord(e->ev[1], e->stream));
  EHIP(e, hipEventSynchronize(e->ev[1]));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e->ev[0], e->ev[1]);
  e->last_ms = ms;
  e->last_steps = nsteps;
  int worst = MDX_OK, first_bad = -1;
  for (int r = 0; r < e->R; r++) {
    const RepCtl &c = e->hctl[r];
    if (rebuilds) rebuilds[r] = c.rebuilds;
    if (c.error && first_bad < 0) {
      first_bad = r;
      worst = c.error;
    }
  }
  if (first_bad >= 0) {
    const RepCtl &c = e->hctl[first_bad];
    if (c.error == MDX_ERR_TOO_FAST)
      return efail(e, c.error, "replica %d: Too fast particles exists. max_velocity = %f", first_bad,
                   std::sqrt(c.vmax2));
    return efail(e, c.error, "replica %d stopped (status %d, flags %d)", first_bad, c.error, c.eflags);
  }
  // every running replica has done its steps, or the capacity fixes did not
  // converge within the launch rounds: say so rather than return short
  for (int r = 0; r < e->R; r++) {
    const RepCtl &c = e->hctl[r];
    if (c.steps < c.target)
      return efail(e, MDX_ERR_CAPACITY,
                   "replica %d completed %d of %d steps: its capacity was grown %d times without "
                   "the rebuild fitting (mdx_ensemble_status gives the counts)",
                   r, c.steps - (c.target - nsteps), nsteps, kMaxFixRounds);
  }
  return worst;
}

int mdx_ensemble_status(mdx_ensemble *e, int *status, int *steps) {
  if (!e) return MDX_ERR_ARG;
  int rc = read_ctl(e);
  if (rc) return rc;
  for (int r = 0; r < e->R; r++) {
    if (status) status[r] = e->hctl[r].error;
    if (steps) steps[r] = e->hctl[r].steps;
  }
  return MDX_OK;
}

int mdx_ensemble_observe(mdx_ensemble *e, double *ke, double *pe, double *vir) {
  if (!e) return efail(nullptr, MDX_ERR_ARG, "null ensemble");
  if (!e->have_state) return efail(e, MDX_ERR_STATE, "no state");
  const EnsDev d = dev_view(e);
  k_e_observe<<<e->NT, kBlk, stage_bytes(e), e->stream>>>(d, e->C0, e->C2);
  k_e_observe_fold<<<e->R, kBlock, 0, e->stream>>>(d, e->obs_out);
  int rc = launch_check(e, "observe");
  if (rc) return rc;
  std::vector<double> h(3 * (size_t)e->R);
  EHIP(e, hipMemcpyAsync(h.data(), e->obs_out, sizeof(double) * h.size(), hipMemcpyDeviceToHost, e->stream));
  EHIP(e, hipStreamSynchronize(e->stream));
  for (int r = 0; r < e->R; r++) {
    if (ke) ke[r] = h[3 * r];
    if (pe) pe[r] = h[3 * r + 1];
    if (vir) vir[r] = h[3 * r + 2] / 3.0;
  }
  return MDX_OK;
}

int mdx_ensemble_get_config(const mdx_ensemble *e, mdx_ensemble_config *cfg) {
  if (!e || !cfg) return MDX_ERR_ARG;
  cfg->tile = e->tile;
  cfg->grid_mult = e->grid_mult;
  cfg->step_grid = e->step_grid;
  cfg->stage_cap = e->stage_cap;
  cfg->ms_per_step = e->last_steps > 0 ? e->last_ms / e->last_steps : 0.0;
  return MDX_OK;
}

// Reals aimed at per tile and persistent grid.  Changing the tile target
// re-tiles every replica, so it is accepted only before the first step
// after mdx_ensemble_set_state, and rebuilds from the loaded state.
int mdx_ensemble_set_config(mdx_ensemble *e, int tile, int grid_mult) {
  if (!e || (tile != 256 && tile != 512) || grid_mult < 1 || grid_mult > 16)
    return efail(e, MDX_ERR_ARG, "tile must be 256 or 512, grid_mult in [1, 16]");
  int rc;
  const bool retile = tile != e->tile;
  if (retile && e->have_state) {
    if ((rc = read_ctl(e))) return rc;
    for (const auto &c : e->hctl)
      if (c.steps > 0) return efail(e, MDX_ERR_STATE, "tile size changes only before the first step");
  }
  if (retile) {
    e->tile = tile;
    for (auto &rp : e->hrep) rp.level = 0;
    if ((rc = layout_level(e, tile_level(tile)))) return rc;
  }
  e->grid_mult = grid_mult;
  if ((rc = size_step_grid(e))) return rc;
  e->graph_dirty = true;
  if (retile && e->have_state && (rc = reset_state(e))) return rc;
  return MDX_OK;
}

// Times trial_steps steps for every tile target and grid multiple, each from
// the loaded state (after one untimed trial per target), keeps the fastest,
// and leaves the loaded state (with its first pairlist) in place.  Results
// do not depend on the choice: the tiles and the grid change which
// workgroup computes a real, not its arithmetic or the order of its
// partners.
int mdx_ensemble_autotune(mdx_ensemble *e, int trial_steps, mdx_ensemble_config *best) {
  if (!e || trial_steps < 1) return efail(e, MDX_ERR_ARG, "bad arguments to mdx_ensemble_autotune");
  if (!e->have_state) return efail(e, MDX_ERR_STATE, "autotune needs a loaded state");
  int rc;
  if ((rc = read_ctl(e))) return rc;
  for (const auto &c : e->hctl)
    if (c.steps > 0) return efail(e, MDX_ERR_STATE, "autotune runs only before the first step");
  const int tiles[2] = {512, 256};
  const int mults[3] = {1, 2, 4};
  double best_ms = 1e300;
  int bt = e->tile, bm = e->grid_mult;
  for (int T : tiles) {
    if (T != e->tile) {
      e->tile = T;
      for (auto &rp : e->hrep) rp.level = 0;
      if ((rc = layout_level(e, tile_level(T)))) return rc;
    }
    // an untimed trial first: the capacities this target needs grow in it
    e->grid_mult = mults[0];
    if ((rc = size_step_grid(e)) || (rc = reset_state(e))) return rc;
    e->graph_dirty = true;
    if ((rc = mdx_ensemble_calculate(e, trial_steps, nullptr))) return rc;
    for (int gm : mults) {
      e->grid_mult = gm;
      if ((rc = size_step_grid(e)) || (rc = reset_state(e))) return rc;
      e->graph_dirty = true;
      if ((rc = mdx_ensemble_calculate(e, trial_steps, nullptr))) return rc;
      const double ms = e->last_ms / trial_steps;
      if (ms < best_ms) {
        best_ms = ms;
        bt = T;
        bm = gm;
      }
    }
  }
  if (bt != e->tile) {
    e->tile = bt;
    for (auto &rp : e->hrep) rp.level = 0;
    if ((rc = layout_level(e, tile_level(bt)))) return rc;
  }
  e->grid_mult = bm;
  if ((rc = size_step_grid(e)) || (rc = reset_state(e))) return rc;
  e->graph_dirty = true;
  if (best) {
    mdx_ensemble_get_config(e, best);
    best->ms_per_step = best_ms;
  }
  return MDX_OK;
}

}  // extern "C"
