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
// The pair kernels pull the stream through LDS, SB_LEAVES leaves at a time:
// each leaf's particles are loaded once per group with vector loads and read
// by all lanes as LDS broadcasts.  Per candidate leaf every lane computes the
// periodic offset that get_sep (kdtree.F90:1468-1501) gives for its own leaf
// centre against the candidate's centre, so a separation rounds exactly as
// the reference's xpartveci - xyzcache (dens.F90:653-670).
//
// Neighbour lists are kept per lane in blocks of 8 entries: entry k of the
// lane of active index t is nlist[((k/8)*nrows + t)*8 + k%8].  A lane stages
// 8 entries in LDS and writes them as two 16-byte stores.
#pragma once

#include "ctx.h"

constexpr int GROUP = 64;                         // particles per group = wavefront size
constexpr int SB_LEAVES = 32;                     // candidate leaves staged per batch
constexpr int SB_PARTS = SB_LEAVES * PH_MINPART;  // leaves hold at most minpart particles
constexpr int WALK_STACK = 2048;                  // node stack of the group walk (LDS)

// list entry: tree position j (27 bits) and a 5-bit separation code:
// 0..26 periodic offset (sx+1)+3(sy+1)+9(sz+1) with minimum image,
// 13 = no offset and no minimum image possible, 27 = no offset but the
// minimum image may apply
constexpr int SEP_PLAIN = 13;
constexpr int SEP_WRAP0 = 27;

__device__ __forceinline__ uint32_t list_pack(int j, int code) { return (uint32_t)j | ((uint32_t)code << 27); }

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

struct StageLDS {
  double4 pos[SB_PARTS];   // x, y, z, w of the staged candidate particles
  int jpos[SB_PARTS];      // their tree positions
  int lstart[SB_LEAVES + 1];
  double cx[SB_LEAVES], cy[SB_LEAVES], cz[SB_LEAVES], sz[SB_LEAVES];
};

// Stage one batch of candidate leaves (gc[b0 .. b0+nb)) and their particles
// from a tree-ordered double4 array.  Returns the number of staged leaves.
// Every lane of the wave must call it.
__device__ __forceinline__ int stage_batch(StageLDS &S, const uint32_t *__restrict__ gc, int b0, int nc,
                                           const phsph_kdnode *__restrict__ nd, const int32_t *__restrict__ nd_lo,
                                           const int32_t *__restrict__ nd_hi, const double4 *__restrict__ src,
                                           int lane) {
  const int nb = min(SB_LEAVES, nc - b0);
  int lo = 0, cnt = 0;
  if (lane < nb) {
    const int n = (int)gc[b0 + lane];
    lo = nd_lo[n];
    cnt = min(nd_hi[n] - lo + 1, PH_MINPART);
    S.cx[lane] = nd[n].xcen[0];
    S.cy[lane] = nd[n].xcen[1];
    S.cz[lane] = nd[n].xcen[2];
    S.sz[lane] = nd[n].size;
  }
  int incl = cnt;
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o, 64);
    if (lane >= o) incl += v;
  }
  const int excl = incl - cnt;
  if (lane < nb) S.lstart[lane] = excl;
  if (lane == nb - 1) S.lstart[nb] = incl;
  for (int k = 0; k < cnt; k++) {
    S.pos[excl + k] = src[lo + k];
    S.jpos[excl + k] = lo + k;
  }
  __syncthreads();
  return nb;
}

// Separation code of candidate leaf L for a lane whose own leaf has centre
// (lcx,lcy,lcz) and size lsz: the get_sep offset, and whether any pair of
// the two leaves can need the minimum-image correction.
__device__ __forceinline__ int leaf_sep_code(const DevConst &dc, const StageLDS &S, int L, double lcx, double lcy,
                                             double lcz, double lsz) {
  if (!dc.periodic) return SEP_PLAIN;
  double dx = xsub(lcx, S.cx[L]), dy = xsub(lcy, S.cy[L]), dz = xsub(lcz, S.cz[L]);
  int sx = 0, sy = 0, sz = 0;
  if (fabs(dx) > dc.hdlx) { sx = (dx > 0.) ? 1 : -1; dx = xsub(dx, copysign(dc.dxbound, dx)); }
  if (fabs(dy) > dc.hdly) { sy = (dy > 0.) ? 1 : -1; dy = xsub(dy, copysign(dc.dybound, dy)); }
  if (fabs(dz) > dc.hdlz) { sz = (dz > 0.) ? 1 : -1; dz = xsub(dz, copysign(dc.dzbound, dz)); }
  const int code = off_code(sx, sy, sz);
  // |x_i - x_j - offset| <= |dx| + lsz + size_L: below half the box no
  // separation of the two leaves is ever corrected
  const double reach = (lsz + S.sz[L]) * (1. + 1e-9);
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

// per-lane neighbour list writer (blocked layout, see above)
struct ListWriter {
  uint32_t *nlist;
  size_t nrows;
  int t, capl, n;
  uint32_t (*buf)[GROUP];
  int lane;
  __device__ __forceinline__ void flush_block(int blk) {
    const uint4 a = make_uint4(buf[0][lane], buf[1][lane], buf[2][lane], buf[3][lane]);
    const uint4 b = make_uint4(buf[4][lane], buf[5][lane], buf[6][lane], buf[7][lane]);
    uint4 *dst = (uint4 *)(nlist + (((size_t)blk * nrows + (size_t)t) << 3));
    dst[0] = a;
    dst[1] = b;
  }
  __device__ __forceinline__ void push(uint32_t e) {
    buf[n & 7][lane] = e;
    n++;
    if ((n & 7) == 0 && n <= capl) flush_block((n >> 3) - 1);
  }
  __device__ __forceinline__ void finish() {
    if ((n & 7) != 0 && n <= capl) flush_block(n >> 3);
  }
};

__device__ __forceinline__ uint32_t list_get(const uint32_t *__restrict__ nlist, size_t nrows, int t, int k) {
  return nlist[((((size_t)(k >> 3)) * nrows + (size_t)t) << 3) + (k & 7)];
}

// host side (density.hip)
int walk_groups(phsph_ctx *c, int mode, double rfac, const double *d_hsrc, const double *d_hmaxn);
