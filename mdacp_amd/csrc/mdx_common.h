// mdx_common.h — device definitions shared by the single-system step
// (mdx.hip) and the replica ensemble (ensemble.hip): constants, cell
// geometry, the pair term, tile descriptors and wave reductions.
#ifndef MDX_COMMON_H
#define MDX_COMMON_H

#include <hip/hip_runtime.h>

namespace mdxk {

constexpr int kBlock = 256;

constexpr int kCodeNone = 13;                       // image code of "no shift"

// The helpers below round operation by operation, as written: contraction
// into FMA is switched off in each (the library is built to honour the
// pragma; otherwise the compiler fuses a + b c where it sees fit, which can
// differ from kernel to kernel).  Every kernel that uses them then gets
// identical bits.

// |p|^2 rounded as written (pairlist.cc:52), identical in every kernel
__device__ __forceinline__ double speed2(const double4 &p) {
#pragma clang fp contract(off)
  return p.x * p.x + p.y * p.y + p.z * p.z;
}

__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  return v;
}

// 1/x from the hardware estimate r and one third-order correction:
// with e = 1 - x r, r (1 + e + e^2) = (1 - e^3) / x, so the relative error
// of the estimate is cubed (two Newton steps raise it to the fourth power
// at one more operation).  Within an ulp of the IEEE quotient.
__device__ __forceinline__ double recip(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  const double e = __fma_rn(-x, r, 1.0);
  return __fma_rn(r, __fma_rn(e, e, e), r);
}

struct Geom {
  double L[3];
  double inv_w[3];
  int ms[3];   // interior cells per axis
  int m[3];    // with the image layer on both sides
  double SL2, CL2;
};

__device__ __forceinline__ int cell_linear(const Geom &g, int ix, int iy, int iz) {
  return ix + g.m[0] * (iy + g.m[1] * iz);
}

__device__ __forceinline__ int interior_coord(double x, double inv_w, int ms) {
  int c = (int)floor(x * inv_w) + 1;
  return c < 1 ? 1 : (c > ms ? ms : c);
}

__device__ __forceinline__ double shifted(double x, int s, double L) {
  // the reference's image q + diff (mdunit.cc:183), one rounding: s L is
  // exactly 0 or +-L, so x + s L is x, x + L or x - L bit for bit (for any
  // x but -0.0, which positions never are), without a branch
  return __fma_rn((double)s, L, x);
}

constexpr int kILP = 8;   // list entries per group (one 16-byte load)
// independent accumulator chains per lane of a kick: entry k of a list
// feeds chain k % kChains, and the chains are summed in a fixed order
constexpr int kChains = 4;
static_assert(kILP % kChains == 0, "a group's entries map onto whole chains");

// far beyond any cutoff: the staged sentinel after a tile's partners, which
// padding entries of a list point at
constexpr double kFar = 1.0e100;

// A real's partners lie in three z-planes of cell rows (dz = -1, 0, +1),
// which its list walks in order.
constexpr int kPlanes = 3;

// List entries.  An entry is the partner's byte offset in its tile's LDS
// staging (kEntryBytes x its index there), so a kernel reads the partner
// at staging + entry with no index arithmetic.  A list is padded at its
// end to a multiple of kILP with the sentinel's offset (the slot after the
// last staged particle, which holds kFar), so no group of kILP entries
// needs a mask.  A tile may stage at most kMaxStaged particles: the
// sentinel's offset must fit 16 bits.
constexpr int kEntryBytes = 3 * (int)sizeof(double);
constexpr int kMaxStaged = 65535 / kEntryBytes - 1;

__device__ __forceinline__ unsigned short staged_entry(int index) {
  return (unsigned short)(kEntryBytes * index);
}

// xyz of the partner at entry e of a tile staged at s3
__device__ __forceinline__ const double *staged_at(const double *s3, unsigned e) {
  return reinterpret_cast<const double *>(reinterpret_cast<const char *>(s3) + e);
}

// Grouped list layout: a tile's list entries sit in groups of kILP per
// column, group g of column l at [(g S + l) kILP, (g S + l + 1) kILP) of
// the tile's region (S columns), so one lane's group is one 16-byte load
// and a wavefront's is 1 KiB in a row.  A column's base is the region plus
// l kILP; its entry k sits at base[list_index<S>(k)].
static_assert(kILP == 8, "a group of entries is one 16-byte load");
template <int S>
__host__ __device__ constexpr size_t list_index(int k) {
  return (size_t)(k / kILP) * S * kILP + k % kILP;
}

// Writes one column of a list under construction in the grouped layout:
// the place of the next entry advances by one within a group and on to the
// column's next group after the last.  Entries past kcap (a multiple of
// kILP) are counted, not stored.
template <int S>
struct GroupWriter {
  unsigned short *ptr;   // the next entry's place
  int kcap, cnt = 0;
  __device__ GroupWriter(unsigned short *c, int kc) : ptr(c), kcap(kc) {}
  __device__ __forceinline__ void put(unsigned e) {
    if (cnt < kcap) *ptr = (unsigned short)e;
    ++cnt;
    ptr += (cnt % kILP) ? 1 : S * kILP - (kILP - 1);
  }
  // the list's end: the sentinel of a staging of len particles up to a
  // whole group
  __device__ __forceinline__ void pad_end(int len) {
    while (cnt % kILP) put(staged_entry(len));
  }
};

// Entries [kb, ke) of one lane's list (column base lst), both multiples of
// kILP, S columns, partners staged at s3: entry k goes to term(k % kILP, xj,
// yj, zj) in list order.  The list streams from memory, so the loads run
// PF groups ahead in a ring of registers: group j computes from ring slot
// j % PF, which then loads group j + PF (clamped to the last group: a
// reload that is never used).  The ring turns once per loop pass, so no
// register copies; the groups past the last full turn compute from the
// ring as it stands.  Partner positions are read from LDS AHEAD entries
// ahead: the reads of entry k + AHEAD (in the next group after a group's
// last entries, the clamped reload past the end) are issued before the
// arithmetic of entry k, and scheduling barriers keep them there, so their
// latency runs under that arithmetic.  The loads and reads change nothing
// in the sums.
constexpr int kPrefetch = 3;
template <int S, int AHEAD = 1, int PF = kPrefetch, class Term>
__device__ __forceinline__ void list_groups(const unsigned short *__restrict__ lst, int kb, int ke,
                                             const double *s3, Term term) {
  static_assert(AHEAD == 1 || AHEAD == 2, "partners are read one or two entries ahead");
  static_assert(PF >= 2, "the group after the one in hand must be in the ring");
  if (kb >= ke) return;
  // every k loaded is a group start (a multiple of kILP, as kb and ke are),
  // where list_index<S>(k) = k S: one unsigned 32-bit byte offset
  auto load = [&](int k) {
    const unsigned off = (unsigned)k * (unsigned)(S * sizeof(unsigned short));
    return *reinterpret_cast<const uint4 *>(reinterpret_cast<const char *>(lst) + off);
  };
  auto entry = [&](const uint4 &w, int u) {
    const unsigned h = u < 2 ? w.x : (u < 4 ? w.y : (u < 6 ? w.z : w.w));
    return staged_at(s3, (h >> (16 * (u % 2))) & 0xffffu);
  };
  double x, y, z;      // the partner of the entry in hand
  double x1, y1, z1;   // and of the next one (AHEAD 2)
  auto run = [&](const uint4 &w, const uint4 &next) {
#pragma unroll
    for (int u = 0; u < kILP; u++) {
      const double *q = u + AHEAD < kILP ? entry(w, u + AHEAD) : entry(next, u + AHEAD - kILP);
      const double nx = q[0], ny = q[1], nz = q[2];
      __builtin_amdgcn_sched_barrier(0);
      term(u, x, y, z);
      __builtin_amdgcn_sched_barrier(0);
      if (AHEAD == 1) {
        x = nx;
        y = ny;
        z = nz;
      } else {
        x = x1;
        y = y1;
        z = z1;
        x1 = nx;
        y1 = ny;
        z1 = nz;
      }
    }
  };
  const int last = ke - kILP, ng = (ke - kb) / kILP;
  uint4 ring[PF];
#pragma unroll
  for (int d = 0; d < PF; d++) ring[d] = load(min(kb + d * kILP, last));
  {
    const double *q = entry(ring[0], 0);
    x = q[0];
    y = q[1];
    z = q[2];
    if (AHEAD == 2) {
      const double *q1 = entry(ring[0], 1);
      x1 = q1[0];
      y1 = q1[1];
      z1 = q1[2];
    }
  }
  int k = kb;
  for (int j = 0; j < ng / PF; j++, k += PF * kILP) {
#pragma unroll
    for (int d = 0; d < PF; d++) {
      run(ring[d], ring[(d + 1) % PF]);
      ring[d] = load(min(k + (PF + d) * kILP, last));
    }
  }
  const int rem = ng % PF;
#pragma unroll
  for (int d = 0; d < PF - 1; d++)
    if (d < rem) run(ring[d], ring[d + 1]);
}

// Order of the interior cell rows (y, z) in the tile array: strips of h
// consecutive z, walked y-major with z fastest inside a strip, h = ceil(ms_z
// / 8) so that each XCD's eighth of the tiles is about one strip.  The rows
// a tile stages (y +- 1, z +- 1) then belong to tiles running on the same
// XCD at about the same time, and come from its L2.  Row p -> (y, z), both
// 1-based.
__host__ __device__ inline int row_strip(const Geom &g) { return (g.ms[2] + 7) / 8; }
__host__ __device__ inline void row_of(const Geom &g, int p, int *y, int *z) {
  const int h = row_strip(g), ny = g.ms[1];
  const int st = p / (h * ny), z0 = st * h;
  const int hs = g.ms[2] - z0 < h ? g.ms[2] - z0 : h;
  const int q = p - st * h * ny;
  *y = q / hs + 1;
  *z = z0 + q % hs + 1;
}

// Constants of the kick's pair term with the time step folded in.
struct PairK {
  double a, b, c;   // 24 dt, -48 dt, 8 C2 dt
};
inline PairK make_pairk(double C2, double dt) { return PairK{24.0 * dt, -48.0 * dt, 8.0 * C2 * dt}; }

// Pair term of ForceCalculator::CalculateForceSorted (fcalculator.cc:203-217):
// df = ((24 r^6 - 48) / (r^6 r^6 r^2) + 8 C2) dt, zero beyond the cutoff,
// evaluated as r^-8 (24 dt - 48 dt r^-6) + 8 C2 dt from one reciprocal of
// r^2 (recip()): the same quantity in fewer dependent operations.  Beyond
// the cutoff only the high word of df is cleared (one 32-bit select, not
// two): what is left is a positive subnormal below 2^-1042 (2.3e-314), so a
// term beyond the cutoff adds at most 2.3e-314 |dq| (2.3e-214 for the 1e100
// sentinel), which any accumulator of size above 1e-198 absorbs exactly,
// and never -0.0.
__device__ __forceinline__ double pair_df(double r2, const PairK &k, double CL2) {
  const double ir2 = recip(r2);
  const double ir6 = ir2 * ir2 * ir2;
  const double df = __fma_rn(ir6 * ir2, __fma_rn(k.b, ir6, k.a), k.c);
  const long long bits = __double_as_longlong(df);
  const int hi = r2 > CL2 ? 0 : (int)(bits >> 32);
  return __hiloint2double(hi, (int)(unsigned)bits);
}


// ------------------------------------------------------------------
// Per-partner and per-lane arithmetic shared by every kernel that must
// produce identical bits (single system and replica ensemble): one inline
// definition gives one expression tree, so the compiler forms the same
// fused operations wherever it is used.

// r^2 of the list build, as the reference's AVX2 search forms it:
// fma(dx,dx, fma(dy,dy, dz*dz)) with dq = q_j - q_i (meshlist.cc:354-362)
__device__ __forceinline__ double list_r2(double xj, double yj, double zj, double xi, double yi,
                                          double zi) {
  const double dx = xj - xi, dy = yj - yi, dz = zj - zi;
  return __fma_rn(dx, dx, __fma_rn(dy, dy, __dmul_rn(dz, dz)));
}

// one partner of the kick: a += df (q_j - q_i)   (fcalculator.cc:203-217).
// r^2 and the accumulations are written as the fused operations they
// compile to, so that no kernel is left a choice of which product to fuse:
// dx dx + dy dy + dz dz contracts to fma(dz,dz, fma(dx,dx, dy dy)) in one
// kernel and fma(dz,dz, fma(dy,dy, dx dx)) in another, which round
// differently.  r^2 is the list build's form (list_r2).
__device__ __forceinline__ void kick_term(double xj, double yj, double zj, double xi, double yi,
                                          double zi, const PairK &k, double CL2, double &ax,
                                          double &ay, double &az) {
  const double dx = xj - xi, dy = yj - yi, dz = zj - zi;
  const double s2 = __fma_rn(dx, dx, __fma_rn(dy, dy, dz * dz));
  const double df = pair_df(s2, k, CL2);
  ax = __fma_rn(df, dx, ax);
  ay = __fma_rn(df, dy, ay);
  az = __fma_rn(df, dz, az);
}

// q + p dt/2 (ForceCalculator::UpdatePositionHalf, fcalculator.cc:26-28)
__device__ __forceinline__ double4 drift_half(double4 q, const double4 &p, double dt2) {
#pragma clang fp contract(off)
  q.x = q.x + p.x * dt2;
  q.y = q.y + p.y * dt2;
  q.z = q.z + p.z * dt2;
  return q;
}

// image of a position under shift code (sx+1) + 3(sy+1) + 9(sz+1)
__device__ __forceinline__ double4 image_of(double4 q, int code, const double *L) {
  q.x = shifted(q.x, code % 3 - 1, L[0]);
  q.y = shifted(q.y, (code / 3) % 3 - 1, L[1]);
  q.z = shifted(q.z, code / 9 - 1, L[2]);
  q.w = 0.0;
  return q;
}

// squared displacement old - new (pairlist.cc:73-80), also the pair
// distances of the candidate test (pairlist.cc:98-112)
__device__ __forceinline__ double disp2(const double4 &a, const double4 &b) {
#pragma clang fp contract(off)
  const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return dx * dx + dy * dy + dz * dz;
}

// observer terms of one partner within the cutoff: potential energy
// 4 (r^-12 - r^-6 + C2 r^2 + C0) (observer.cc:18-47) and r V'(r)
// (observer.cc:50-79)
__device__ __forceinline__ void observe_term(double r2, double C0, double C2, double &pe,
                                             double &vir) {
  const double r6 = r2 * r2 * r2;
  pe += 4.0 * (1.0 / (r2 * r2 * r2 * r2 * r2 * r2) - 1.0 / r6 + C2 * r2 + C0);
  vir += ((24.0 * r6 - 48.0) / (r6 * r6 * r2) + C2 * 8.0) * r2;
}

// Sort one cell's entries in place by (x, entry), one wavefront per cell,
// x = xof(entry) the x coordinate of the particle an entry names: entries
// and their x go to LDS, and each entry is written to its rank (the number
// of entries before it in that order).  Cells are monotonic in x (the cell
// of x is floor(x / w), and floor of a rounded product is monotonic; an
// image cell holds x - L < 0 or x + L > L), so a row of cells sorted this
// way is sorted by x as a whole.  The wave's lanes run in lockstep, so the
// LDS copy is complete before any lane reads it back; cells larger than the
// LDS buffer (not met in practice) fall back to an insertion sort by one
// lane.
constexpr int kSortBuf = 256;
__device__ __forceinline__ bool x_before(double xa, int a, double xb, int b) {
  return xa < xb || (xa == xb && a < b);
}
template <class X>
__device__ __forceinline__ void wave_rank_sort_x(int *seg, int len, int *buf, double *xb,
                                                 X xof) {
  const int lane = threadIdx.x & 63;
  if (len <= 1) return;
  if (len > kSortBuf) {
    if (lane == 0) {
      for (int a = 1; a < len; a++) {
        const int v = seg[a];
        const double xv = xof(v);
        int b = a - 1;
        while (b >= 0 && x_before(xv, v, xof(seg[b]), seg[b])) {
          seg[b + 1] = seg[b];
          b--;
        }
        seg[b + 1] = v;
      }
    }
    return;
  }
  for (int a = lane; a < len; a += 64) {
    buf[a] = seg[a];
    xb[a] = xof(seg[a]);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int a = lane; a < len; a += 64) {
    const int v = buf[a];
    const double xv = xb[a];
    int rank = 0;
    for (int b = 0; b < len; b++) rank += x_before(xb[b], buf[b], xv, v) ? 1 : 0;
    seg[rank] = v;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Slot of this lane in counter[key], with one atomicAdd per distinct key of
// the wavefront: lanes that share a key take consecutive slots in lane
// order.  Every active lane of the wavefront calls it (it is safe under
// divergence: the ballots see the active lanes only).
__device__ __forceinline__ int wave_grouped_add(int *counter, int key) {
  const int lane = threadIdx.x & 63;
  unsigned long long todo = __ballot(1);
  int slot = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int lkey = __shfl(key, leader);
    const unsigned long long same = __ballot(key == lkey) & todo;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter + lkey, __popcll(same));
    base = __shfl(base, leader);
    if (key == lkey) slot = base + __popcll(same & ((1ull << lane) - 1ull));
    todo &= ~same;
  }
  return slot;
}

}  // namespace mdxk

#endif
