// stream.h -- particle groups and their candidate streams, shared by the
// density and force pair kernels.
//
// A group is 64 consecutive entries of the active list (tree order), handled
// by one wavefront with one particle per lane.  k_group_walk walks the tree
// once per group and lists every leaf that can hold a neighbour of any
// particle of the group: a node is opened when the distance from the group's
// bounding box to the node centre is below size_node + rcut, rcut being the
// group's search radius (or radkern*hmax of the node itself in the force
// walk).  Since |x_i - x_j| >= dist(box, xcen_n) - size_n for every particle
// i in the box and j in node n (also with periodic images), the stream is a
// superset of every lane's getneigh candidates (kdtree.F90:1288-1292); the
// neighbour sets themselves are decided per pair by the reference's q^2
// tests, so they do not change.
//
// The pair kernels pull the stream through LDS, NL leaves (at most
// NL*minpart particles) at a time: every staged quantity of a leaf's
// particles is loaded once per group with coalesced loads.  Each lane tests
// every staged particle against its own (positions read as LDS broadcasts),
// keeps the hits as a bit mask, and then evaluates its hits reading the j
// data from LDS, so no pair ever gathers from global memory and no neighbour
// list is stored.  Per candidate leaf every lane computes the periodic
// offset that get_sep (kdtree.F90:1468-1501) gives for its own leaf centre
// against the candidate's centre, so a separation rounds exactly as the
// reference's xpartveci - xyzcache (dens.F90:653-670).
#pragma once

#include "ctx.h"

constexpr int GROUP = 64;                         // particles per group = wavefront size
constexpr int WALK_STACK = 2048;                  // node stack of the group walk (LDS)
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

__device__ __forceinline__ unsigned long long lanemask_lt() {
  const unsigned lane = __lane_id();
  return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}

// Metadata of one batch of NL candidate leaves staged in LDS: first tree
// position, start of each leaf's particles in the batch, centre and size.
template <int NL>
struct BatchMeta {
  int lo[NL];
  int start[NL + 1];
  double cx[NL], cy[NL], cz[NL], sz[NL];
};

// Load the metadata of candidate leaves gc[b0 .. b0+nb) and return the
// number of staged particles (uniform).  Every lane of the wave must call it.
template <int NL>
__device__ __forceinline__ int batch_meta(BatchMeta<NL> &M, const uint32_t *__restrict__ gc, int b0, int nc,
                                          const phsph_kdnode *__restrict__ nd, const int32_t *__restrict__ nd_lo,
                                          const int32_t *__restrict__ nd_hi, int lane, int &nb_out) {
  const int nb = min(NL, nc - b0);
  int cnt = 0;
  if (lane < nb) {
    const int n = (int)gc[b0 + lane];
    const int lo = nd_lo[n];
    cnt = min(nd_hi[n] - lo + 1, PH_MINPART);
    M.lo[lane] = lo;
    M.cx[lane] = nd[n].xcen[0];
    M.cy[lane] = nd[n].xcen[1];
    M.cz[lane] = nd[n].xcen[2];
    M.sz[lane] = nd[n].size;
  }
  int incl = cnt;
  for (int o = 1; o < NL; o <<= 1) {
    const int v = __shfl_up(incl, o, 64);
    if (lane >= o) incl += v;
  }
  if (lane < nb) M.start[lane] = incl - cnt;
  if (lane == nb - 1) M.start[nb] = incl;
  nb_out = nb;
  const int total = __shfl(incl, nb - 1, 64);
  __syncthreads();
  return total;
}

// leaf of staged particle s (M.start is increasing, nb <= NL leaves)
template <int NL>
__device__ __forceinline__ int batch_leaf_of(const BatchMeta<NL> &M, int nb, int s) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (M.start[mid] <= s) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// per-lane separation codes of up to 16 staged leaves, 5 bits each
struct SepCodes {
  unsigned long long lo = 0ull, hi = 0ull;
  __device__ __forceinline__ void set(int L, int code) {
    if (L < 12) lo |= (unsigned long long)code << (5 * L);
    else hi |= (unsigned long long)code << (5 * (L - 12));
  }
  __device__ __forceinline__ int get(int L) const {
    return (int)(((L < 12) ? (lo >> (5 * L)) : (hi >> (5 * (L - 12)))) & 31ull);
  }
};

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

// host side (density.hip)
int walk_groups(phsph_ctx *c, int mode, double rfac, const double *d_hsrc, const double *d_hmaxn);
