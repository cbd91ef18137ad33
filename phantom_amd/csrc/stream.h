// stream.h -- particle groups, their candidate leaves and the per-lane
// traversal of the particles a lane can reach, shared by the density and
// force list kernels.
//
// A group is 64 consecutive entries of the active list (tree order), handled
// by one wavefront with one particle per lane.  k_group_walk walks the tree
// once per group and lists every leaf that can hold a neighbour of any
// particle of the group: a node is opened when the distance from the group's
// bounding box to the node centre is below size_node + rcut, rcut being the
// group's search radius (or radkern*hmax of the node itself in the force
// walk).  Since |x_i - x_j| >= dist(box, xcen_n) - size_n for every particle
// i in the box and j in node n (also with periodic images), the candidate
// leaves are a superset of every lane's getneigh candidates
// (kdtree.F90:1288-1292); the neighbour sets themselves are decided per pair
// by the reference's q^2 tests, so they do not change.
//
// build_list then narrows the candidates per lane: the wavefront stages a
// chunk of candidate leaves and their particles in LDS, each lane marks the
// leaves its own particle can reach and visits only their particles.  Per candidate
// leaf every lane computes the periodic offset that get_sep
// (kdtree.F90:1468-1501) gives for its own leaf centre against the
// candidate's centre, so a separation rounds exactly as the reference's
// xpartveci - xyzcache (dens.F90:653-670).
#pragma once

#include "ctx.h"

constexpr int GROUP = 64;                         // particles per group = wavefront size
constexpr int WALK_STACK = 4096;                  // node stack of the group walk (LDS)
constexpr int HTRAJ = 16;                         // h after each density iteration kept per particle (self-gravity)

// separation code of a candidate leaf for a lane: 0..26 periodic offset
// (sx+1)+3(sy+1)+9(sz+1) with the minimum image; SEP_PLAIN = no offset and
// no minimum image possible; SEP_WRAP0 = no offset but the minimum image may
// apply
constexpr int SEP_PLAIN = 13;
constexpr int SEP_WRAP0 = 27;

// Blocks are dispatched to the 8 XCDs round robin; give each XCD a contiguous
// range of groups so that neighbouring groups share its L2.  The grid must
// have a multiple of 8 blocks; blocks mapping beyond the last group exit.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int per = nb >> 3;
  return (b & 7) * per + (b >> 3);
}
static inline unsigned xcd_grid(int ngroups) { return (unsigned)(((ngroups + 7) / 8) * 8); }

// The list kernels run LW independent wavefronts per block, one group each,
// so that more waves share a CU; each wavefront stages through its own LDS
// and synchronises only itself.  LW (2 or 4) and the chunk length CL of
// build_list (8 or 16 candidate leaves) are picked per context by the
// autotuner (tune_density_iter, density.hip); every choice gives
// bit-identical results.
template <int LW>
static inline unsigned list_grid(int ngroups) { return xcd_grid((ngroups + LW - 1) / LW); }
template <int LW>
__device__ __forceinline__ int list_group(int nb) {
  return xcd_block(blockIdx.x, nb) * LW + (int)(threadIdx.x >> 6);
}
// orders a wavefront's LDS stores before its later LDS loads
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}



__device__ __forceinline__ unsigned long long lanemask_lt() {
  const unsigned lane = __lane_id();
  return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}

// Separation code of candidate leaf L for a lane whose own leaf has centre
// (lcx,lcy,lcz) and size lsz: the get_sep offset, and whether any pair of
// the two leaves can need the minimum-image correction.
__device__ __forceinline__ int leaf_sep_code(const DevConst &dc, double ccx, double ccy, double ccz, double csz,
                                             double lcx, double lcy, double lcz, double lsz) {
  if (!dc.periodic) return SEP_PLAIN;
  double dx = xsub(lcx, ccx), dy = xsub(lcy, ccy), dz = xsub(lcz, ccz);
  int sx = 0, sy = 0, sz = 0;
  if (fabs(dx) > dc.hdlx) { sx = (dx > 0.) ? 1 : -1; dx = xsub(dx, copysign(dc.dxbound, dx)); }
  if (fabs(dy) > dc.hdly) { sy = (dy > 0.) ? 1 : -1; dy = xsub(dy, copysign(dc.dybound, dy)); }
  if (fabs(dz) > dc.hdlz) { sz = (dz > 0.) ? 1 : -1; dz = xsub(dz, copysign(dc.dzbound, dz)); }
  const int code = off_code(sx, sy, sz);
  // |x_i - x_j - offset| <= |dx| + lsz + size_L: below half the box no
  // separation of the two leaves is ever corrected
  const double reach = (lsz + csz) * (1. + 1e-9);
  const double m = 1. - 1e-9;
  const bool wrap = (fabs(dx) + reach >= m * dc.hdlx) || (fabs(dy) + reach >= m * dc.hdly) ||
                    (fabs(dz) + reach >= m * dc.hdlz);
  if (code == SEP_PLAIN) return wrap ? SEP_WRAP0 : SEP_PLAIN;
  return code;
}

// separation x_i - (x_j + offset) with the minimum image, rounded as the
// reference (dens.F90:653-670, force.F90:1250-1270)
__device__ __forceinline__ void sep_of(const DevConst &dc, int code, double xi, double yi, double zi, double xj,
                                       double yj, double zj, double &dx, double &dy, double &dz) {
  if (code == SEP_PLAIN) {
    dx = xi - xj;
    dy = yi - yj;
    dz = zi - zj;
    return;
  }
  double ox = 0., oy = 0., oz = 0.;
  if (code != SEP_WRAP0) {
    int sx, sy, sz;
    off_decode(code, sx, sy, sz);
    ox = (sx == 0) ? 0. : copysign(dc.dxbound, (double)sx);
    oy = (sy == 0) ? 0. : copysign(dc.dybound, (double)sy);
    oz = (sz == 0) ? 0. : copysign(dc.dzbound, (double)sz);
  }
  dx = xsub(xi, xadd(xj, ox));
  dy = xsub(yi, xadd(yj, oy));
  dz = xsub(zi, xadd(zj, oz));
  dx = min_image(dx, dc.dxbound, 0.5 * dc.dxbound);
  dy = min_image(dy, dc.dybound, 0.5 * dc.dybound);
  dz = min_image(dz, dc.dzbound, 0.5 * dc.dzbound);
}

// The hmax a leaf ends densityiterate with where the reference lists its
// neighbours again (compute_hmax, dens.F90:1275-1289, then set_hmaxcell):
// the cell's iterations are replayed from each particle's h after every
// Newton update (htraj: iteration k of particle p at (k-1)*hstride + p, up
// to HTRAJ iterations) and at convergence (hconv); the
// cell's hmax after iteration k is 1.01 max h, and a relist happens when it
// exceeds the previous hmax while the cell is not yet converged.  Returns 1
// with *hset the hmax of the last relist, 0 when the leaf is never relisted,
// -1 when a cell needed more than HTRAJ + 1 iterations.
__device__ __forceinline__ int leaf_relist_hmax(int lo, int hi, double hcell, const int32_t *__restrict__ p_nits,
                                                const double *__restrict__ htraj, size_t hstride,
                                                const double *__restrict__ hconv, double *hset) {
  int nits = 0;
  for (int p = lo; p <= hi; p++) nits = max(nits, p_nits[p]);
  if (nits <= 1) return 0;  // converged at the first iteration: no relist
  if (nits - 1 > HTRAJ) return -1;
  int set = 0;
  for (int k = 1; k < nits; k++) {  // not yet converged after iteration k
    double hmx = 0.;
    for (int p = lo; p <= hi; p++) {
      const int np_ = p_nits[p];
      if (np_ < 1) continue;
      const double v = (k < np_) ? htraj[(size_t)(k - 1) * hstride + p] : hconv[p];
      hmx = fmax(hmx, v);
    }
    const double hm = 1.01 * hmx;
    if (hm > hcell) {
      set = 1;
      *hset = hm;
    }
    hcell = hm;
  }
  return set;
}

// host side (density.hip)
// Candidate leaves of each group of 64 active particles, into c->gcand at
// c->goff_p (exact per-group offsets after an overflow, else nullptr and the
// uniform capacity c->caps.gcand).
int walk_groups(phsph_ctx *c, int mode, double rfac, const double *d_hsrc, const double *d_hmaxn);

// Neighbour lists: entry k of active slot t lives in segment k/16 of that
// slot, a run of 16 entries (64 bytes) that the slot's lane fills on its own;
// the segments of one index k/16 are consecutive over t.  A lane thus writes
// whole 64-byte pieces (no partial-line write-backs), and a wavefront reading
// entry k of 64 consecutive slots reads 64 adjacent segments.
constexpr int LSEG = 16;
__device__ __forceinline__ size_t list_at(int nact, int t, int k) {
  return ((size_t)(k / LSEG) * (size_t)nact + (size_t)t) * LSEG + (size_t)(k % LSEG);
}
// The same lists with a capacity per group of 64 slots, for lengths that vary
// strongly between groups: group g owns entries off[g] .. off[g+1]-1 (a
// multiple of 64*16), its slots' segments interleaved as above.  With
// off == nullptr every slot has the uniform capacity of list_at.
__device__ __forceinline__ size_t list_at_g(const int64_t *off, int nact, int t, int k) {
  if (!off) return list_at(nact, t, k);
  return (size_t)off[t / GROUP] + ((size_t)(k / LSEG) * GROUP + (size_t)(t % GROUP)) * LSEG + (size_t)(k % LSEG);
}
__device__ __forceinline__ int list_cap_g(const int64_t *off, int t, int ucap) {
  return off ? (int)((off[t / GROUP + 1] - off[t / GROUP]) / GROUP) : ucap;
}
// candidate leaves of group g: uniform capacity capg, or exact offsets
__device__ __forceinline__ size_t cand_base(const int64_t *goff, int g, int capg) {
  return goff ? (size_t)goff[g] : (size_t)g * capg;
}
__device__ __forceinline__ int cand_cap(const int64_t *goff, int g, int capg) {
  return goff ? (int)(goff[g + 1] - goff[g]) : capg;
}

// Device memory a uniform list capacity may take (phsph_ctx::list_budget,
// 16 GiB by default) before the lists switch to exact per-group sizes;
// uniform capacities grow by doubling and are kept for later calls while
// they fit.
static inline bool list_fits(const phsph_ctx *c, size_t entries) { return entries * sizeof(uint32_t) <= c->list_budget; }

// exclusive prefix sum of n counts into off[0..n] (off[n] = total), one block
template <class T>
__global__ void __launch_bounds__(1024) k_scan_offsets(const T *__restrict__ cnt, int n, int64_t scale, int round,
                                                       int64_t *__restrict__ off) {
  __shared__ int64_t part[1024];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < n; base += 1024) {
    const int i = base + threadIdx.x;
    int64_t v = 0;
    if (i < n) {
      int64_t c = (int64_t)cnt[i];
      if (round > 1) c = (c + round - 1) / round * round;
      v = c * scale;
    }
    part[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const int64_t a = (threadIdx.x >= (unsigned)o) ? part[threadIdx.x - o] : 0;
      __syncthreads();
      part[threadIdx.x] += a;
      __syncthreads();
    }
    if (i < n) off[i] = carry + part[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) off[n] = carry;
}

static inline int list_cap(int cap) { return (cap + LSEG - 1) / LSEG * LSEG; }

// longest list of each group of 64 slots (t_nl over the group's live slots)
static __global__ __attribute__((unused)) void k_group_max(const int32_t *__restrict__ t_nl, int nact, int ngroups,
                                   int32_t *__restrict__ gmax) {
  const int g = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int t = g * GROUP + (threadIdx.x & 63);
  int v = (g < ngroups && t < nact) ? t_nl[t] : 0;
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  if (g < ngroups && (threadIdx.x & 63) == 0) gmax[g] = v;
}

// Neighbour-list building over a group's candidate leaves, CL leaves at a
// time through LDS.  For each chunk the wavefront loads the leaves' node
// data and their particles' records with parallel loads into LDS, each lane
// marks the leaves of the chunk its particle can reach (reach(L) with the
// leaf's centre, size and hmax in LDS), and then visits only the staged
// particles of its own reachable leaves, reading them from LDS:
// visit(j, code, r2, w) with the leaf's separation code for the lane, the
// squared separation from (xi, yi, zi) rounded as sep_of and r2_exact, and
// the record's fourth word.  S is the wavefront's own stage; only the
// wavefront synchronises (wave_lds_sync).
template <int CLn>
struct ListStage {
  static constexpr int CL = CLn;                // candidate leaves per chunk
  static constexpr int CS = CLn * PH_MINPART;  // staged particles per chunk
  int lo[CL], start[CL + 1];
  double cx[CL], cy[CL], cz[CL], sz[CL], hm[CL];
  double x[CS], y[CS], z[CS], hinv[CS];
};

// xyzh-like records: x, y, z and (HINV) 1/h of particle j
template <class Rec>
__device__ __forceinline__ void rec_xyzi(const Rec *__restrict__ r, int j, double &x, double &y, double &z,
                                         double &w) {
  const double4 v = *reinterpret_cast<const double4 *>(r + j);
  x = v.x; y = v.y; z = v.z; w = v.w;
}

template <int CL, class Rec, class Reach, class Visit>
__device__ __forceinline__ void build_list(ListStage<CL> &S, const DevConst &dc, const uint32_t *__restrict__ gc, int nc,
                                           const phsph_kdnode *__restrict__ nd, const int32_t *__restrict__ nd_lo,
                                           const int32_t *__restrict__ nd_hi, const double *__restrict__ hmaxn,
                                           const Rec *__restrict__ rec, double xi, double yi, double zi,
                                           double lcx, double lcy, double lcz, double lsz, Reach reach,
                                           Visit visit) {
  const int lane = threadIdx.x & 63;
  for (int k0 = 0; k0 < nc; k0 += CL) {
    const int kn = min(CL, nc - k0);
    int cnt = 0;
    if (lane < kn) {
      const int n = (int)gc[k0 + lane];
      const int lo = nd_lo[n];
      cnt = min(nd_hi[n] - lo + 1, PH_MINPART);
      S.lo[lane] = lo;
      S.cx[lane] = nd[n].xcen[0];
      S.cy[lane] = nd[n].xcen[1];
      S.cz[lane] = nd[n].xcen[2];
      S.sz[lane] = nd[n].size;
      S.hm[lane] = hmaxn ? hmaxn[n] : 0.;
    }
    int incl = cnt;
    for (int o = 1; o < CL; o <<= 1) {
      const int v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    if (lane < kn) S.start[lane] = incl - cnt;
    if (lane == kn - 1) S.start[kn] = incl;
    const int total = __shfl(incl, kn - 1, 64);
    wave_lds_sync();
    for (int s = lane; s < total; s += 64) {
      int L = 0;  // leaf of slot s: last L with start[L] <= s
      for (int b = CL / 2; b > 0; b >>= 1)
        if (L + b < kn && S.start[L + b] <= s) L += b;
      rec_xyzi(rec, S.lo[L] + (s - S.start[L]), S.x[s], S.y[s], S.z[s], S.hinv[s]);
    }
    wave_lds_sync();
    unsigned mask = 0u;
    for (int L = 0; L < kn; L++)
      if (reach(S.cx[L], S.cy[L], S.cz[L], S.sz[L], S.hm[L])) mask |= 1u << L;
    while (mask != 0u) {
      const int L = __builtin_ctz(mask);
      mask &= mask - 1u;
      const int code = leaf_sep_code(dc, S.cx[L], S.cy[L], S.cz[L], S.sz[L], lcx, lcy, lcz, lsz);
      // the separation of sep_of with the leaf's offset taken out of the
      // loop: x_i - (x_j + o), then the minimum image where it may apply
      // (x_j + 0 = x_j, so the plain case rounds identically)
      double ox = 0., oy = 0., oz = 0.;
      const bool wrap = code != SEP_PLAIN;
      if (wrap && code != SEP_WRAP0) {
        int sx, sy, sz;
        off_decode(code, sx, sy, sz);
        ox = (sx == 0) ? 0. : copysign(dc.dxbound, (double)sx);
        oy = (sy == 0) ? 0. : copysign(dc.dybound, (double)sy);
        oz = (sz == 0) ? 0. : copysign(dc.dzbound, (double)sz);
      }
      const int s1 = S.start[L + 1], j0 = S.lo[L] - S.start[L];
      if (!wrap) {  // x_i - (x_j + 0) = x_i - x_j: no offset add
        for (int s = S.start[L]; s < s1; s++) {
          const double dx = xsub(xi, S.x[s]), dy = xsub(yi, S.y[s]), dz = xsub(zi, S.z[s]);
          visit(j0 + s, code, r2_exact(dx, dy, dz), S.hinv[s]);
        }
      } else {
        for (int s = S.start[L]; s < s1; s++) {
          double dx = xsub(xi, xadd(S.x[s], ox)), dy = xsub(yi, xadd(S.y[s], oy)), dz = xsub(zi, xadd(S.z[s], oz));
          dx = min_image(dx, dc.dxbound, 0.5 * dc.dxbound);
          dy = min_image(dy, dc.dybound, 0.5 * dc.dybound);
          dz = min_image(dz, dc.dzbound, 0.5 * dc.dzbound);
          visit(j0 + s, code, r2_exact(dx, dy, dz), S.hinv[s]);
        }
      }
    }
    wave_lds_sync();
  }
}

// the reach test of a lane: |x_i - xcen| (minimum image) against the leaf
// size plus the lane's search radius, with a margin for rounding
__device__ __forceinline__ bool leaf_in_reach(const DevConst &dc, double xi, double yi, double zi, double ccx,
                                              double ccy, double ccz, double csz, double rad) {
  double ex = xi - ccx, ey = yi - ccy, ez = zi - ccz;
  if (dc.periodic) {
    ex = min_image(ex, dc.dxbound, dc.hdlx);
    ey = min_image(ey, dc.dybound, dc.hdly);
    ez = min_image(ez, dc.dzbound, dc.hdlz);
  }
  const double sr = (csz + rad) * (1. + 1e-10);
  return ex * ex + ey * ey + ez * ez < sr * sr;
}
