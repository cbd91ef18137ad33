// gravity.hip -- kd-tree multipole self-gravity (the GRAVITY build of the
// reference): node second moments, the node hmax the force walk sees, the
// dual-tree walk per leaf with M2L/L2L (kdtree.F90:1357-1840) and the force
// pair loop with softened and point-mass gravity (force.F90:1286-2053,
// 2909-2927).
//
//  * k_quads: second moments of every node about its centre of mass
//    (kdtree.F90:724-750), one wavefront per node of at most TREE_SUB_MAX
//    particles (the tree build forms those of larger nodes chunk by chunk).
//  * k_hmax_replay: the reference raises a leaf's hmax (and its ancestors')
//    whenever a density iteration relists the cell (set_hmaxcell,
//    neigh_kdtree.f90:115-131; dens.F90:338-373).  The density kernel keeps
//    each particle's h after every iteration; this kernel replays the cell's
//    compute_hmax sequence (dens.F90:1275-1289) to give the force walk the
//    same node hmax, which decides which node pairs are well separated.
//  * k_gw_level: getneigh_dual node by node, one level of destination nodes
//    per launch, one wavefront per node: well-separated candidate sources
//    add their quadrupole expansion (compute_M2L) to the node's
//    coefficients, the rest are opened for the node's children; the
//    children's incoming expansions are shifted from the node's
//    (propagate_fnode_to_node).  At a leaf the sources are opened down to
//    leaves, which form its near-field list.
//  * k_force_grav: one thread per active particle over its leaf's
//    near-field list: SPH pair terms plus softened gravity for neighbours,
//    1/r^2 for the other trial particles, then the far field by Taylor
//    expansion about the leaf centre (expand_fgrav_in_taylor_series).
#include "records.h"

constexpr int LENFGRAV = 20;
constexpr int GMAXDEPTH = 32;   // kdtree.F90:42
constexpr int GSTACK = 2048;    // (source, level) stack of the dual walk (LDS)
constexpr int GW_MASKS = 64;    // candidates whose first-pass result is kept: 64 * GW_MASKS

enum { GFL_STACK = 1, GFL_DEPTH = 2, GFL_P2P = 4, GFL_TRAJ = 8, GFL_CAND = 16, GFL_DTC = 32, GFL_NEAR = 64 };

// ids 2n, 2n+1 are the children of n: a slot is a node when every ancestor
// is an internal node (leaf_flag == 0)
__device__ __forceinline__ bool node_valid(int n, const int32_t *__restrict__ leaf_flag) {
  while (n > 1) {
    n >>= 1;
    if (leaf_flag[n] != 0) return false;
  }
  return true;
}

// ------------------------------------------------------------------------
__global__ void __launch_bounds__(64)
k_quads(DevConst dc, int64_t ncells, const phsph_kdnode *__restrict__ nd, const int32_t *__restrict__ nd_lo,
        const int32_t *__restrict__ nd_hi, const int32_t *__restrict__ leaf_flag, const double *__restrict__ tx,
        const double *__restrict__ ty, const double *__restrict__ tz, double *__restrict__ quads) {
  const int64_t n = (int64_t)blockIdx.x + 1;
  if (n > ncells || !node_valid((int)n, leaf_flag)) return;
  if (nd_hi[n] - nd_lo[n] + 1 > TREE_SUB_MAX) return;  // formed by the tree build's large-node phase
  const int lane = threadIdx.x;
  const double c0 = nd[n].xcen[0], c1 = nd[n].xcen[1], c2 = nd[n].xcen[2];
  const double pm = dc.pmass;
  double q[6] = {0., 0., 0., 0., 0., 0.};
  for (int p = nd_lo[n] + lane; p <= nd_hi[n]; p += 64) {
    const double dx = tx[p] - c0, dy = ty[p] - c1, dz = tz[p] - c2;
    q[0] += pm * (dx * dx); q[1] += pm * (dx * dy); q[2] += pm * (dx * dz);
    q[3] += pm * (dy * dy); q[4] += pm * (dy * dz); q[5] += pm * (dz * dz);
  }
  for (int k = 0; k < 6; k++) q[k] = wave_sum(q[k]);
  if (lane == 0)
    for (int k = 0; k < 6; k++) quads[6 * n + k] = q[k];
}

__global__ void k_init_ghmax(int64_t nslots, const phsph_kdnode *__restrict__ nd,
                             unsigned long long *__restrict__ ghmax) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < nslots) ghmax[n] = (unsigned long long)__double_as_longlong(nd[n].hmax);
}

// ------------------------------------------------------------------------
// node hmax after densityiterate (compute_hmax + set_hmaxcell)
// ------------------------------------------------------------------------
__global__ void k_hmax_replay(int nleaves, const int32_t *__restrict__ leaves, const int32_t *__restrict__ leaf_flag,
                              const int32_t *__restrict__ nd_lo, const int32_t *__restrict__ nd_hi,
                              const phsph_kdnode *__restrict__ nd, const int32_t *__restrict__ p_nits,
                              const double *__restrict__ htraj, size_t hstride, const double *__restrict__ hconv,
                              unsigned long long *__restrict__ ghmax, int32_t *__restrict__ flags) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nleaves) return;
  const int c = leaves[l];
  if (leaf_flag[c] <= 0) return;
  double hset = 0.;
  const int set = leaf_relist_hmax(nd_lo[c], nd_hi[c], __longlong_as_double((long long)ghmax[c]), p_nits, htraj,
                                   hstride, hconv, &hset);
  if (set < 0) {
    atomicOr(flags, GFL_TRAJ);
    return;
  }
  if (!set) return;
  const unsigned long long v = (unsigned long long)__double_as_longlong(hset);
  ghmax[c] = v;
  // a node already at >= v was raised by a thread that carries a value >= v
  // further up
  for (int n = nd[c].parent; n != 0; n = nd[n].parent)
    if (atomicMax(&ghmax[n], v) >= v) break;
}

// ------------------------------------------------------------------------
// expansions (kdtree.F90:1527-1560, 1702-1781, 1799-1840)
// ------------------------------------------------------------------------
__device__ __forceinline__ void m2l(double dx, double dy, double dz, double dr1, double q0, const double *quads,
                                    double *f) {
  const double dr12 = dr1 * dr1;
  const double dx2 = dx * dx, dx3 = dx * dx2, dy2 = dy * dy, dy3 = dy * dy2, dz2 = dz * dz, dz3 = dz * dz2;
  const double g0 = -dr1;
  const double g1 = 1. * dr12 * g0;
  const double g2 = -3. * dr12 * g1;
  const double g3 = -5. * dr12 * g2;
  const double g2dx = g2 * dx, g2dy = g2 * dy, g2dz = g2 * dz;
  double D3[10], D2[6];
  D3[0] = 3. * g2dx + g3 * dx3;
  D3[1] = g2dy + g3 * dx2 * dy;
  D3[2] = g2dz + g3 * dx2 * dz;
  D3[3] = g2dx + g3 * dy2 * dx;
  D3[4] = g3 * dx * dy * dz;
  D3[5] = g2dx + g3 * dz2 * dx;
  D3[6] = 3. * g2dy + g3 * dy3;
  D3[7] = g2dz + g3 * dy2 * dz;
  D3[8] = g2dy + g3 * dz2 * dy;
  D3[9] = 3. * g2dz + g3 * dz3;
  D2[0] = g1 + g2 * dx2;
  D2[1] = g2dx * dy;
  D2[2] = g2dx * dz;
  D2[3] = g1 + g2 * dy2;
  D2[4] = g2dy * dz;
  D2[5] = g1 + g2 * dz2;
  const double qxx = quads[0], qxy = quads[1], qxz = quads[2], qyy = quads[3], qyz = quads[4], qzz = quads[5];
  f[0] = g1 * dx * q0 + 0.5 * (D3[0] * qxx + 2. * (D3[1] * qxy + D3[2] * qxz + D3[4] * qyz) + D3[3] * qyy + D3[5] * qzz);
  f[1] = g1 * dy * q0 + 0.5 * (D3[1] * qxx + 2. * (D3[3] * qxy + D3[4] * qxz + D3[7] * qyz) + D3[6] * qyy + D3[8] * qzz);
  f[2] = g1 * dz * q0 + 0.5 * (D3[2] * qxx + 2. * (D3[4] * qxy + D3[5] * qxz + D3[8] * qyz) + D3[7] * qyy + D3[9] * qzz);
  for (int k = 0; k < 6; k++) f[3 + k] = D2[k] * q0;
  for (int k = 0; k < 10; k++) f[9 + k] = D3[k] * q0;
  f[19] = g0 * q0 - 0.5 * (D2[0] * qxx + D2[3] * qyy + D2[5] * qzz + 2 * (D2[1] * qxy + D2[2] * qxz + D2[4] * qyz));
}

__device__ __forceinline__ void l2l(double *f, const double *fs, double dx, double dy, double dz) {
  f[0] = fs[0] + dx * (fs[3] + 0.5 * (dx * fs[9] + dy * fs[10] + dz * fs[11])) +
         dy * (fs[4] + 0.5 * (dx * fs[10] + dy * fs[12] + dz * fs[13])) +
         dz * (fs[5] + 0.5 * (dx * fs[11] + dy * fs[13] + dz * fs[14]));
  f[1] = fs[1] + dx * (fs[4] + 0.5 * (dx * fs[10] + dy * fs[12] + dz * fs[13])) +
         dy * (fs[6] + 0.5 * (dx * fs[12] + dy * fs[15] + dz * fs[16])) +
         dz * (fs[7] + 0.5 * (dx * fs[13] + dy * fs[16] + dz * fs[17]));
  f[2] = fs[2] + dx * (fs[5] + 0.5 * (dx * fs[11] + dy * fs[13] + dz * fs[14])) +
         dy * (fs[7] + 0.5 * (dx * fs[13] + dy * fs[16] + dz * fs[17])) +
         dz * (fs[8] + 0.5 * (dx * fs[14] + dy * fs[17] + dz * fs[18]));
  f[3] = fs[3] + dx * fs[9] + dy * fs[10] + dz * fs[11];
  f[4] = fs[4] + dx * fs[10] + dy * fs[12] + dz * fs[13];
  f[5] = fs[5] + dx * fs[11] + dy * fs[13] + dz * fs[14];
  f[6] = fs[6] + dx * fs[12] + dy * fs[15] + dz * fs[16];
  f[7] = fs[7] + dx * fs[13] + dy * fs[16] + dz * fs[17];
  f[8] = fs[8] + dx * fs[14] + dy * fs[17] + dz * fs[18];
  for (int k = 9; k < 19; k++) f[k] = fs[k];
  f[19] = fs[19] + dx * (fs[0] + 0.5 * (dx * fs[3] + dy * fs[4] + dz * fs[5])) +
          dy * (fs[1] + 0.5 * (dx * fs[4] + dy * fs[6] + dz * fs[7])) +
          dz * (fs[2] + 0.5 * (dx * fs[5] + dy * fs[7] + dz * fs[8]));
}

// the reference's expansion, including its dfxy in the z term of the first
// potential bracket (kdtree.F90:1836)
__device__ __forceinline__ void l2p(const double *fn, double dx, double dy, double dz, double &fx, double &fy,
                                    double &fz, double &pot) {
  double fxi = fn[0], fyi = fn[1], fzi = fn[2];
  const double dfxx = fn[3], dfxy = fn[4], dfxz = fn[5], dfyy = fn[6], dfyz = fn[7], dfzz = fn[8];
  const double d2fxxx = fn[9], d2fxxy = fn[10], d2fxxz = fn[11], d2fxyy = fn[12], d2fxyz = fn[13];
  const double d2fxzz = fn[14], d2fyyy = fn[15], d2fyyz = fn[16], d2fyzz = fn[17], d2fzzz = fn[18];
  double poti = fn[19];
  fxi = fxi + dx * (dfxx + 0.5 * (dx * d2fxxx + dy * d2fxxy + dz * d2fxxz)) +
        dy * (dfxy + 0.5 * (dx * d2fxxy + dy * d2fxyy + dz * d2fxyz)) +
        dz * (dfxz + 0.5 * (dx * d2fxxz + dy * d2fxyz + dz * d2fxzz));
  fyi = fyi + dx * (dfxy + 0.5 * (dx * d2fxxy + dy * d2fxyy + dz * d2fxyz)) +
        dy * (dfyy + 0.5 * (dx * d2fxyy + dy * d2fyyy + dz * d2fyyz)) +
        dz * (dfyz + 0.5 * (dx * d2fxyz + dy * d2fyyz + dz * d2fyzz));
  fzi = fzi + dx * (dfxz + 0.5 * (dx * d2fxxz + dy * d2fxyz + dz * d2fxzz)) +
        dy * (dfyz + 0.5 * (dx * d2fxyz + dy * d2fyyz + dz * d2fyzz)) +
        dz * (dfzz + 0.5 * (dx * d2fxzz + dy * d2fyzz + dz * d2fzzz));
  poti = poti - (dx * (fxi - 0.5 * (dx * dfxx + dy * dfxy + dz * dfxy)) +
                 dy * (fyi - 0.5 * (dx * dfxy + dy * dfyy + dz * dfyz)) +
                 dz * (fzi - 0.5 * (dx * dfxz + dy * dfyz + dz * dfzz)));
  fx = fxi; fy = fyi; fz = fzi; pot = poti;
}

// get_sep between two node centres (kdtree.F90:1468-1501): separation and
// the periodic offset code
__device__ __forceinline__ double node_sep(const DevConst &dc, const double *a, const double *b, double &dx,
                                           double &dy, double &dz, int &code) {
  dx = xsub(a[0], b[0]); dy = xsub(a[1], b[1]); dz = xsub(a[2], b[2]);
  int sx = 0, sy = 0, sz = 0;
  if (dc.periodic) {
    if (fabs(dx) > dc.hdlx) { sx = (dx > 0.) ? 1 : -1; dx = xsub(dx, copysign(dc.dxbound, dx)); }
    if (fabs(dy) > dc.hdly) { sy = (dy > 0.) ? 1 : -1; dy = xsub(dy, copysign(dc.dybound, dy)); }
    if (fabs(dz) > dc.hdlz) { sz = (dz > 0.) ? 1 : -1; dz = xsub(dz, copysign(dc.dzbound, dz)); }
  }
  code = off_code(sx, sy, sz);
  return r2_exact(dx, dy, dz);
}

// ------------------------------------------------------------------------
// dual-tree walk, node by node (getneigh_dual, kdtree.F90:1357-1460;
// node_interaction :1653-1694; open_nodes :1591-1645).
//
// The reference walks once per leaf: (source, destination) pairs start at
// (root, root) and the destination runs down the leaf's branch.  The pairs
// whose destination is a node A depend only on the path to A, so here every
// node processes them once, level by level.  A node's candidate sources are
// those its parent opened: a source equal to the destination, or not well
// separated from it, is opened into its children (a leaf source stays
// itself) for both children of the destination.  Well-separated sources add
// their expansion (compute_M2L) to the node's coefficients.  A node's
// incoming expansion is its parent's, plus the parent's own coefficients,
// shifted to the node's centre (propagate_fnode_to_node): the same
// operations as the reference's per-leaf downward pass.  At a leaf the
// candidates are resolved completely, as the reference does once its
// destination is the leaf: sources that are not well separated are opened
// until they are leaves, which form the leaf's near-field list.
// ------------------------------------------------------------------------
struct GwLevel {
  const int32_t *node;   // destination node ids
  const int64_t *coff;   // candidate list offset
  const int32_t *ccnt;   // candidate list length
  const double *H;       // incoming expansion (LENFGRAV per entry)
  const uint32_t *cand;  // candidate source node ids
};
struct GwNext {
  int32_t *node;
  int64_t *coff;
  int32_t *ccnt;
  double *H;
  uint32_t *cand;
  int64_t cand_cap;
  unsigned long long *cand_used;  // candidate entries taken (may exceed cand_cap)
  int32_t *nent;                  // entries of the next level
};

// node_interaction: separation (dst - src) and whether the pair is well
// separated
__device__ __forceinline__ bool gw_wellsep(const DevConst &dc, const phsph_kdnode *__restrict__ nd,
                                           const unsigned long long *__restrict__ ghmax, int dst, int src,
                                           double &dx, double &dy, double &dz, double &r2, int &code) {
  r2 = node_sep(dc, nd[dst].xcen, nd[src].xcen, dx, dy, dz, code);
  const double R = dc.radkern;
  const double rcut_dst = __longlong_as_double((long long)ghmax[dst]) * R;
  const double rcut_src = __longlong_as_double((long long)ghmax[src]) * R;
  const double rcut = fmax(rcut_dst, rcut_src);
  const double ss = nd[dst].size + nd[src].size;
  const double rcut2 = (ss + rcut) * (ss + rcut);
  return (dc.tree_acc2 * r2 > ss * ss) && (r2 > rcut2);
}

__device__ __forceinline__ bool gw_active(int n, const int32_t *__restrict__ leaf_flag,
                                          const int32_t *__restrict__ act) {
  const int lf = leaf_flag[n];
  return (lf != 0) ? (lf > 0) : (act == nullptr || act[n] != 0);
}

// ancestors of active leaves (block timesteps): the walk skips the rest
__global__ void k_gw_mark(int nleaves, const int32_t *__restrict__ leaves, const int32_t *__restrict__ leaf_flag,
                          const phsph_kdnode *__restrict__ nd, int32_t *act) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nleaves) return;
  const int c = leaves[l];
  if (leaf_flag[c] <= 0) return;
  for (int n = nd[c].parent; n != 0; n = nd[n].parent)
    if (atomicOr(&act[n], 1) != 0) break;
}

__global__ void k_gw_fill(int32_t *a, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// level 0: the root with itself as its one candidate
__global__ void k_gw_init(int32_t *node, int64_t *coff, int32_t *ccnt, double *H, uint32_t *cand) {
  const int k = threadIdx.x;
  if (k < LENFGRAV) H[k] = 0.;
  if (k == 0) {
    node[0] = 1;
    coff[0] = 0;
    ccnt[0] = 1;
    cand[0] = 1u;
  }
}

__global__ void __launch_bounds__(64)
k_gw_level(DevConst dc, int nent, GwLevel in, GwNext out, const int32_t *__restrict__ leaf_flag,
           const int32_t *__restrict__ act, const phsph_kdnode *__restrict__ nd,
           const unsigned long long *__restrict__ ghmax, const double *__restrict__ quads,
           const int32_t *__restrict__ nd_lo, const int32_t *__restrict__ pleaf, double *__restrict__ fnode_out,
           uint32_t *__restrict__ p2p, int32_t *__restrict__ np2p, const int64_t *poff, int ucapp,
           int32_t *__restrict__ flags) {
  __shared__ uint32_t stk[GSTACK];
  __shared__ unsigned long long omask[GW_MASKS];  // opened candidates of the first pass, 64 per word
  // consecutive entries (nodes near each other, sharing candidate sources)
  // on one XCD and its L2
  const int e = xcd_block(blockIdx.x, gridDim.x);
  if (e >= nent) return;
  const int lane = threadIdx.x;
  const int A = in.node[e];
  const int64_t c0 = in.coff[e];
  const int nc = in.ccnt[e];
  const int lfA = leaf_flag[A];
  double facc[LENFGRAV];
  for (int k = 0; k < LENFGRAV; k++) facc[k] = 0.;
  if (lfA == 0) {
    // internal destination: expansions of the well-separated sources, and
    // the opened sources as the children's candidates
    // The opened sources are staged in LDS (the leaf stack, which internal
    // nodes do not use) in the order the list is written, and copied out
    // once the list has its place; a list longer than the stack is formed
    // again in a second pass over the candidates
    int ntot = 0;
    for (int b = 0; b < nc; b += 64) {
      const int k = b + lane;
      int npush = 0;
      uint32_t e0 = 0u, e1 = 0u;
      if (k < nc) {
        const int src = (int)in.cand[c0 + k];
        bool opened = (src == A);
        if (!opened) {
          double dx, dy, dz, r2;
          int code;
          if (gw_wellsep(dc, nd, ghmax, A, src, dx, dy, dz, r2, code)) {
            double f[LENFGRAV];
            m2l(dx, dy, dz, 1. / sqrt(r2), nd[src].mass, quads + 6 * (size_t)src, f);
            for (int q = 0; q < LENFGRAV; q++) facc[q] += f[q];
          } else {
            opened = true;
          }
        }
        if (opened) {
          if (leaf_flag[src] != 0) {
            e0 = (uint32_t)src;
            npush = 1;
          } else {
            const int il = nd[src].leftchild, ir = nd[src].rightchild;
            if (il != 0) { e0 = (uint32_t)il; npush = 1; }
            if (ir != 0) {
              if (npush == 0) e0 = (uint32_t)ir;
              else e1 = (uint32_t)ir;
              npush++;
            }
          }
        }
      }
      const unsigned long long m = __ballot(npush > 0);
      if (lane == 0 && (b >> 6) < GW_MASKS) omask[b >> 6] = m;
      int incl = npush;
      for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
      }
      const int at = ntot + incl - npush;
      if (npush > 0 && at < GSTACK) stk[at] = e0;
      if (npush > 1 && at + 1 < GSTACK) stk[at + 1] = e1;
      ntot += __shfl(incl, 63, 64);
    }
    wave_lds_sync();
    unsigned long long base = 0ull;
    if (lane == 0) base = atomicAdd(out.cand_used, (unsigned long long)ntot);
    base = __shfl(base, 0, 64);
    const bool fits = (int64_t)(base + (unsigned long long)ntot) <= out.cand_cap;
    if (fits && ntot <= GSTACK) {
      for (int k = lane; k < ntot; k += 64) out.cand[(size_t)base + k] = stk[k];
    } else if (fits) {
      int w = 0;
      for (int b = 0; b < nc; b += 64) {
        const int k = b + lane;
        int npush = 0;
        uint32_t e0 = 0u, e1 = 0u;
        if (k < nc) {
          const int src = (int)in.cand[c0 + k];
          bool opened;
          if ((b >> 6) < GW_MASKS) {
            opened = (omask[b >> 6] >> lane) & 1ull;
          } else {
            opened = (src == A);
            if (!opened) {
              double dx, dy, dz, r2;
              int code;
              opened = !gw_wellsep(dc, nd, ghmax, A, src, dx, dy, dz, r2, code);
            }
          }
          if (opened) {
            if (leaf_flag[src] != 0) {
              e0 = (uint32_t)src;
              npush = 1;
            } else {
              const int il = nd[src].leftchild, ir = nd[src].rightchild;
              if (il != 0) { e0 = (uint32_t)il; npush = 1; }
              if (ir != 0) {
                if (npush == 0) e0 = (uint32_t)ir;
                else e1 = (uint32_t)ir;
                npush++;
              }
            }
          }
        }
        int incl = npush;
        for (int o = 1; o < 64; o <<= 1) {
          const int v = __shfl_up(incl, o, 64);
          if (lane >= o) incl += v;
        }
        const size_t at = (size_t)base + w + incl - npush;
        if (npush > 0) out.cand[at] = e0;
        if (npush > 1) out.cand[at + 1] = e1;
        w += __shfl(incl, 63, 64);
      }
    }
    double fs[LENFGRAV];
    for (int q = 0; q < LENFGRAV; q++) fs[q] = in.H[(size_t)e * LENFGRAV + q] + wave_sum(facc[q]);
    // the children's incoming expansions: lane 0 the left child, lane 1 the right
    const int il = nd[A].leftchild, ir = nd[A].rightchild;
    const bool ul = il != 0 && gw_active(il, leaf_flag, act);
    const bool ur = ir != 0 && gw_active(ir, leaf_flag, act);
    int slot = 0;
    if (lane == 0 && (ul || ur)) slot = atomicAdd(out.nent, (ul ? 1 : 0) + (ur ? 1 : 0));
    slot = __shfl(slot, 0, 64);
    if (lane < 2) {
      const bool mine = (lane == 0) ? ul : ur;
      if (mine) {
        const int ch = (lane == 0) ? il : ir;
        const int at = slot + ((lane == 1 && ul) ? 1 : 0);
        double dx, dy, dz;
        int cd;
        node_sep(dc, nd[ch].xcen, nd[A].xcen, dx, dy, dz, cd);
        double fc[LENFGRAV];
        l2l(fc, fs, dx, dy, dz);
        for (int q = 0; q < LENFGRAV; q++) out.H[(size_t)at * LENFGRAV + q] = fc[q];
        out.node[at] = ch;
        out.coff[at] = (int64_t)base;
        out.ccnt[at] = fits ? ntot : 0;
      }
    }
    if (!fits && lane == 0) atomicOr(flags, GFL_CAND);
    return;
  }
  if (lfA < 0) return;  // inactive leaf: no walk
  // active leaf: resolve its candidates completely
  const int li = pleaf[nd_lo[A]];
  const size_t pbase = poff ? (size_t)poff[li] : (size_t)li * ucapp;
  const int capp = poff ? (int)(poff[li + 1] - poff[li]) : ucapp;
  int top = 0, npp = 0, b = 0;
  bool ovf = false;
  while (true) {
    int take;
    uint32_t ent = 0u;
    if (b < nc) {
      take = min(64, nc - b);
      if (lane < take) ent = in.cand[c0 + b + lane];
      b += take;
    } else if (top > 0) {
      take = min(top, 64);
      if (lane < take) ent = stk[top - take + lane];
      top -= take;
    } else {
      break;
    }
    wave_lds_sync();
    bool p2p_hit = false;
    int code = SEP_PLAIN, npush = 0;
    uint32_t e0 = 0u, e1 = 0u;
    const int src = (int)ent;
    if (lane < take) {
      bool opened = (src == A);
      if (!opened) {
        double dx, dy, dz, r2;
        if (gw_wellsep(dc, nd, ghmax, A, src, dx, dy, dz, r2, code)) {
          double f[LENFGRAV];
          m2l(dx, dy, dz, 1. / sqrt(r2), nd[src].mass, quads + 6 * (size_t)src, f);
          for (int q = 0; q < LENFGRAV; q++) facc[q] += f[q];
        } else {
          opened = true;
        }
      }
      if (opened) {
        if (leaf_flag[src] != 0) {
          p2p_hit = true;
        } else {
          const int il = nd[src].leftchild, ir = nd[src].rightchild;
          if (il != 0) { e0 = (uint32_t)il; npush = 1; }
          if (ir != 0) {
            if (npush == 0) e0 = (uint32_t)ir;
            else e1 = (uint32_t)ir;
            npush++;
          }
        }
      }
    }
    // near-field leaves, with the separation code of the pair
    const unsigned long long mp = __ballot(p2p_hit);
    if (p2p_hit) {
      const int k = npp + __popcll(mp & lanemask_lt());
      int sc = SEP_PLAIN;
      if (dc.periodic) sc = (code == SEP_PLAIN) ? SEP_WRAP0 : code;
      if (k < capp) p2p[pbase + k] = cand_pack(src, sc);
    }
    npp += __popcll(mp);
    int incl = npush;
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    const int total = __shfl(incl, 63, 64);
    if (top + total > GSTACK) {
      ovf = true;
      break;
    }
    const int at = top + incl - npush;
    if (npush > 0) stk[at] = e0;
    if (npush > 1) stk[at + 1] = e1;
    top += total;
    wave_lds_sync();
  }
  for (int q = 0; q < LENFGRAV; q++) {
    const double v = in.H[(size_t)e * LENFGRAV + q] + wave_sum(facc[q]);
    if (lane == q) fnode_out[(size_t)li * LENFGRAV + q] = v;
  }
  if (lane == 0) {
    np2p[li] = npp;
    if (ovf) atomicOr(flags, GFL_STACK);
    if (npp > capp) atomicOr(flags, GFL_P2P);
  }
}

// the walk over all levels; flags of the walk in counters[0]
static int grav_walk(phsph_ctx *c, const int64_t *poff) {
  hipStream_t s = c->stream;
  const int nl = c->nleaves;
  const size_t ecap = (size_t)nl + 2;  // nodes of one level <= leaves
  for (int k = 0; k < 2; k++) {
    PH_HIPCHECK(c, c->gw_node[k].ensure(ecap));
    PH_HIPCHECK(c, c->gw_ccnt[k].ensure(ecap));
    PH_HIPCHECK(c, c->gw_coff[k].ensure(ecap));
    PH_HIPCHECK(c, c->gw_H[k].ensure(ecap * LENFGRAV));
    PH_HIPCHECK(c, c->gw_cand[k].ensure((size_t)64 * ecap + 4096));
  }
  PH_HIPCHECK(c, c->gw_used.ensure(1));
  hipLaunchKernelGGL(k_gw_fill, dim3(nblocks(nl + 1, 256)), dim3(256), 0, s, c->np2p.p, (int64_t)nl + 1, 0);
  const int32_t *act = nullptr;
  if (c->dc.ind_timesteps) {
    PH_HIPCHECK(c, c->gw_act.ensure(c->nslots));
    hipLaunchKernelGGL(k_gw_fill, dim3(nblocks(c->nslots, 256)), dim3(256), 0, s, c->gw_act.p, c->nslots, 0);
    hipLaunchKernelGGL(k_gw_mark, dim3(nblocks(nl, 256)), dim3(256), 0, s, nl, c->leaves.p, c->leaf_flag.p,
                       c->nd.p, c->gw_act.p);
    act = c->gw_act.p;
  }
  hipLaunchKernelGGL(k_gw_init, dim3(1), dim3(64), 0, s, c->gw_node[0].p, c->gw_coff[0].p, c->gw_ccnt[0].p,
                     c->gw_H[0].p, c->gw_cand[0].p);
  PH_HIPCHECK(c, hipGetLastError());
  // the root takes part if it is an active leaf or has one below
  int32_t root_flag = 0, root_act = 1;
  PH_HIPCHECK(c, hipMemcpyAsync(&root_flag, c->leaf_flag.p + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  if (act) PH_HIPCHECK(c, hipMemcpyAsync(&root_act, act + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  PH_HIPCHECK(c, hipStreamSynchronize(s));
  int nent = (root_flag != 0) ? (root_flag > 0 ? 1 : 0) : (root_act != 0 ? 1 : 0);
  int cur = 0, level = 0;
  while (nent > 0) {
    if (level >= GMAXDEPTH) {  // a branch of more than GMAXDEPTH nodes
      int32_t fl = GFL_DEPTH;
      PH_HIPCHECK(c, dev_set(c, c->counters.p, &fl, sizeof(int32_t)));
      break;
    }
    const int nxt = cur ^ 1;
    int32_t zero = 0;
    unsigned long long zero64 = 0ull;
    PH_HIPCHECK(c, dev_set(c, c->counters.p + 1, &zero, sizeof(int32_t)));
    PH_HIPCHECK(c, dev_set(c, c->gw_used.p, &zero64, sizeof(zero64)));
    GwLevel in{c->gw_node[cur].p, c->gw_coff[cur].p, c->gw_ccnt[cur].p, c->gw_H[cur].p, c->gw_cand[cur].p};
    GwNext out{c->gw_node[nxt].p, c->gw_coff[nxt].p, c->gw_ccnt[nxt].p, c->gw_H[nxt].p, c->gw_cand[nxt].p,
               (int64_t)c->gw_cand[nxt].cap, c->gw_used.p, c->counters.p + 1};
    hipLaunchKernelGGL(k_gw_level, dim3(xcd_grid(nent)), dim3(64), 0, s, c->dc, nent, in, out, c->leaf_flag.p, act, c->nd.p,
                       c->ghmax.p, c->quads.p, c->nd_lo.p, c->pleaf.p, c->fnode.p, c->p2p.p, c->np2p.p, poff,
                       c->caps.p2p, c->counters.p);
    PH_HIPCHECK(c, hipGetLastError());
    int32_t cnt[2];
    unsigned long long used = 0ull;
    PH_HIPCHECK(c, hipMemcpyAsync(cnt, c->counters.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipMemcpyAsync(&used, c->gw_used.p, sizeof(used), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    if (cnt[0] & GFL_CAND) {
      // the level's candidate lists did not fit: a larger buffer, the level again
      PH_HIPCHECK(c, c->gw_cand[nxt].ensure((size_t)used + used / 4 + 1024));
      int32_t fl = cnt[0] & ~GFL_CAND;
      PH_HIPCHECK(c, dev_set(c, c->counters.p, &fl, sizeof(int32_t)));
      continue;
    }
    if (cnt[0] & GFL_STACK) break;
    nent = cnt[1];
    cur = nxt;
    level++;
  }
  return PHSPH_OK;
}

// ------------------------------------------------------------------------
// force with self-gravity: hydro pair terms (grad-h, Cullen-Dehnen
// viscosity, conductivity) plus gravity, one thread per active particle
// ------------------------------------------------------------------------
struct GravArgs {
  const double *fnode;     // 20 coefficients per leaf (leaf index)
  const uint32_t *p2p;     // near-field leaves per leaf
  const int32_t *np2p;
  const int64_t *poff;     // per-leaf offsets into p2p, or null: l*capp
  int capp;
  const float *gsoft;      // gradh(2) in tree order
  float *poten;            // caller order (may be null)
  uint32_t *nbuf;          // neighbours per active particle (list_at_g layout)
  int capn;                // its uniform capacity per particle (noff == null)
  const int64_t *noff;     // per-group offsets into nbuf, or null
  int32_t *nnb;            // neighbours met per active particle
  int32_t *nmax;           // longest neighbour list met (after an overflow)
};

// 1/sqrt(r2) for the mixed-precision form: an FP32 reciprocal square root
// refined by one FP64 Newton step (relative error ~1e-14), or the FP64
// routine where r2 is outside the normal FP32 range
__device__ __forceinline__ double rsqrt_mixed(double r2) {
  const float r2f = (float)r2;
  if (!(r2f >= 1.17549435e-38f && r2f <= 3.4e38f)) return rsqrt(r2);
  const double y = (double)rsqrtf(r2f);
  return y * (1.5 - 0.5 * r2 * y * y);
}

// MIXED: the mixed-precision form.  Positions, the neighbour tests, the
// multipole expansions and every sum stay FP64; 1/r comes from
// rsqrt_mixed, and a neighbour's quantities other than position, velocity
// and u from its FP32 record (rec32; P/rho^2 as an FP32 pair of FP64
// accuracy) with position and velocity from the packed FP64 arrays
template <int K, bool THERMAL, bool MIXED>
__global__ void __launch_bounds__(128)
k_force_grav(DevConst dc, const int32_t *__restrict__ act, int nact, const int32_t *__restrict__ pleaf,
             const int32_t *__restrict__ leaves, const int32_t *__restrict__ nd_lo, const int32_t *__restrict__ nd_hi,
             const phsph_kdnode *__restrict__ nd, const int32_t *__restrict__ iparts, const FRec *__restrict__ rec,
             const double4 *__restrict__ fpos, const double4 *__restrict__ fvel, const FRec32 *__restrict__ rec32,
             GravArgs G, ForceIOArgs out, const int8_t *__restrict__ tbin_old, int icall,
             unsigned long long *__restrict__ dtred, int32_t *__restrict__ flags) {
  // consecutive blocks (neighbouring leaves, overlapping near-field lists)
  // on one XCD and its L2
  const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  double dtc_o = PH_HUGE64, dtf_o = PH_HUGE64;
  int nl = 0, bmax = -1;
  if (t < nact) {
    const int p = act[t];
    const int l = pleaf[p];
    const int c = leaves[l];
    const double pm = dc.pmass;
    const double R2 = dc.radkern2;
    const FRec Ri = rec[p];
    const double xi = Ri.x, yi = Ri.y, zi = Ri.z, hi = Ri.h;
    // rhoanddhdrho (part.F90:805-816)
    const double hi1 = 1. / fabs(hi);
    const double tt = dc.hfact * hi1;
    const double rhoi = pm * (tt * tt * tt);
    const double rho1i = 1. / rhoi;
    const double hi21 = xmul(hi1, hi1), hi41 = hi21 * hi21;
    double gradhi = (double)Ri.gradh;
    if (!(gradhi > 0.)) gradhi = 1.;  // force.F90:2616-2621
    const double gradsofti = (double)G.gsoft[p];
    const double hfacgrkern = hi41 * dc.cnormk * gradhi;
    const double vxi = Ri.vx, vyi = Ri.vy, vzi = Ri.vz, eni = Ri.u;
    const double pri = Ri.pr, csi = Ri.cs, alphai = (double)Ri.alpha;
    const double vwavei = csi;
    const double pro2i = pri * rho1i * rho1i;
    const double mrhoi5 = 0.5 * pm * rho1i;
    const double auterm = mrhoi5 * dc.alphau;
    double fx = 0., fy = 0., fz = 0., pot = 0., drhodt = 0., dudtdiss = 0., dendtdiss = 0., vsigmax = 0.;
    const bool ind = out.ibin != nullptr;
    int ibin_neighi = 0;
    const int8_t ibinnow_m1 = (int8_t)(out.ibinnow - 1);
    const size_t pbase = G.poff ? (size_t)G.poff[l] : (size_t)l * G.capp;
    const int ne = min(G.np2p[l], G.poff ? (int)(G.poff[l + 1] - G.poff[l]) : G.capp);
    int nn = 0;  // neighbours met in the first pass
    const int capl = list_cap_g(G.noff, t, G.capn);
    // first pass, over the near-field leaves: the 1/r^2 terms of the trial
    // particles outside both kernels, and the list of the neighbours.  List
    // entries are read two leaves ahead and leaf ranges one ahead, and each
    // particle's position record while the previous one is summed
    uint32_t en1 = (ne > 0) ? G.p2p[pbase] : 0u;
    uint32_t en2 = (ne > 1) ? G.p2p[pbase + 1] : 0u;
    int nlo = 0, nhi = -1;
    if (ne > 0) {
      nlo = nd_lo[cand_node(en1)];
      nhi = nd_hi[cand_node(en1)];
    }
    for (int e = 0; e < ne; e++) {
      const int code = cand_code(en1);
      const int j0 = nlo, j1 = nhi;
      en1 = en2;
      if (e + 2 < ne) en2 = G.p2p[pbase + e + 2];
      if (e + 1 < ne) {
        nlo = nd_lo[cand_node(en1)];
        nhi = nd_hi[cand_node(en1)];
      }
      // positions from the packed array (four to a cache line), not from
      // the 128-byte records
      double4 pn = make_double4(0., 0., 0., 0.);
      if (j0 <= j1) pn = fpos[j0];
      for (int j = j0; j <= j1; j++) {
        const double4 pj = pn;
        if (j + 1 <= j1) pn = fpos[j + 1];
        if (j == p) continue;
        double dx, dy, dz;
        sep_of(dc, code, xi, yi, zi, pj.x, pj.y, pj.z, dx, dy, dz);
        const double rij2 = r2_exact(dx, dy, dz);
        const double q2i = xmul(rij2, hi21);
        const double hj1 = pj.w;
        const double hj21 = xmul(hj1, hj1);
        const double q2j = xmul(rij2, hj21);
        if (!(q2i < R2 || q2j < R2)) {
          // trial neighbour outside both kernels: 1/r^2 (force.F90:1991-2053).
          // 1/sqrt(r^2) as one reciprocal square root (within 2 ulp of the
          // reference's sqrt and division); the neighbour tests above use r^2
          const double rij1 = MIXED ? rsqrt_mixed(rij2) : rsqrt(rij2);
          const double fgravj = rij1 * rij1 * rij1 * pm;
          fx = fx - dx * fgravj;
          fy = fy - dy * fgravj;
          fz = fz - dz * fgravj;
          pot = pot + pm * (-rij1);
          continue;
        }
        // a neighbour: its pair terms are summed in the second pass, where
        // every lane of the wavefront has a neighbour to sum
        if (nn < capl) G.nbuf[list_at_g(G.noff, nact, t, nn)] = cand_pack(j, code);
        nn++;
      }
    }
    G.nnb[t] = nn;
    if (nn > capl) {
      atomicOr(flags, GFL_NEAR);
      atomicMax(G.nmax, nn);
    }
    // second pass: the neighbours (force.F90:1286-1990)
    const int nk = min(nn, capl);
    for (int kk = 0; kk < nk; kk++) {
      const uint32_t ek = G.nbuf[list_at_g(G.noff, nact, t, kk)];
      const int j = cand_node(ek), code = cand_code(ek);
      const FRec &Rj = rec[j];
      double4 pj;
      if constexpr (MIXED) pj = fpos[j];
      else pj = make_double4(Rj.x, Rj.y, Rj.z, Rj.hinv);
      double dx, dy, dz;
      sep_of(dc, code, xi, yi, zi, pj.x, pj.y, pj.z, dx, dy, dz);
      const double rij2 = r2_exact(dx, dy, dz);
      const double q2i = xmul(rij2, hi21);
      const double hj1 = pj.w;
      const double hj21 = xmul(hj1, hj1);
      const double q2j = xmul(rij2, hj21);
      nl++;
      if (ind) {  // force.F90:1346-1352
        ibin_neighi = max(ibin_neighi, (int)tbin_old[j]);
        const int o = abs(iparts[j]) - 1;
        if (out.ibin_wake[o] < ibinnow_m1) out.ibin_wake[o] = ibinnow_m1;
      }
      double rij1, qi;
      if (rij2 > PH_TINY64) {
        rij1 = MIXED ? rsqrt_mixed(rij2) : 1. / sqrt(rij2);
        qi = (rij2 * rij1) * hi1;
      } else {
        rij1 = 0.;
        qi = 0.;
      }
      double grkerni = 0., phii, fgravi;
      if (q2i < R2) {
        grkerni = kern_g<K, double>(q2i, qi) * hfacgrkern;
        double fmi;
        kern_soft<K>(q2i, qi, phii, fmi);
        phii = phii * hi1;
        fmi = fmi * hi21;
        fgravi = fmi + gradsofti * grkerni;
      } else {
        phii = -rij1;
        fgravi = rij1 * rij1;
      }
      // what the pair terms read of j
      double vxj, vyj, vzj, uj, gradhj, rho1j_, pro2j_, vwavej_, alphaj;
      if constexpr (MIXED) {
        const double4 vj = fvel[j];
        const FRec32 A = rec32[j];
        vxj = vj.x; vyj = vj.y; vzj = vj.z; uj = vj.w;
        gradhj = (double)A.gradh;
        rho1j_ = (double)A.rho1;
        pro2j_ = (double)A.pro2 + (double)A.pro2lo;
        vwavej_ = (double)A.vwave;
        alphaj = (double)A.alpha;
      } else {
        vxj = Rj.vx; vyj = Rj.vy; vzj = Rj.vz; uj = Rj.u;
        gradhj = (double)Rj.gradh;
        rho1j_ = Rj.rho1;
        pro2j_ = Rj.pro2;
        vwavej_ = Rj.vwave;
        alphaj = (double)Rj.alpha;
      }
      double grkernj = 0., fgravj;
      bool usej = false;
      if (q2j < R2) {
        const double qj = (rij2 * rij1) * hj1;
        grkernj = kern_g<K, double>(q2j, qj) * hj21 * hj21 * dc.cnormk * gradhj;
        double phij, fmj;
        kern_soft<K>(q2j, qj, phij, fmj);
        fmj = fmj * hj21;
        fgravj = fmj + (double)G.gsoft[j] * grkernj;
        usej = true;
      } else {
        fgravj = rij1 * rij1;
      }
      const double fgrav = 0.5 * (pm * fgravi + pm * fgravj);
      const double runix = dx * rij1, runiy = dy * rij1, runiz = dz * rij1;
      const double dvx = vxi - vxj, dvy = vyi - vyj, dvz = vzi - vzj;
      const double projv = dvx * runix + dvy * runiy + dvz * runiz;
      const double denij = THERMAL ? (eni - uj) : 0.;
      const double vsigi = fmax(vwavei - dc.beta * projv, 0.);
      const double vsigavi = fmax(alphai * vwavei - dc.beta * projv, 0.);
      vsigmax = fmax(vsigmax, vsigi);
      double rho1j = 0., pro2j = 0., vwavej = 0., vsigavj = 0., autermj = 0.;
      if (usej) {
        rho1j = rho1j_;
        pro2j = pro2j_;
        vwavej = vwavej_;
        autermj = 0.5 * pm * rho1j * dc.alphau;
        vsigmax = fmax(vsigmax, fmax(vwavej - dc.beta * projv, 0.));
        vsigavj = fmax(alphaj * vwavej - dc.beta * projv, 0.);
      }
      double qrho2i = 0., qrho2j = 0.;
      if (projv < 0.) {
        qrho2i = -0.5 * rho1i * vsigavi * projv;
        if (usej) qrho2j = -0.5 * rho1j * vsigavj * projv;
      }
      const double dudtdissi = pm * qrho2i * projv * grkerni;
      const double gradpi = pm * (pro2i + qrho2i) * grkerni;
      const double gradpj = usej ? pm * (pro2j + qrho2j) * grkernj : 0.;
      if (THERMAL) {
        const double vsigu = fabs(projv);  // with self-gravity (force.F90:1606-1624)
        dudtdiss += dudtdissi;
        dendtdiss += vsigu * denij * (auterm * grkerni + autermj * grkernj);
      }
      const double gradp = gradpi + gradpj;
      fx = fx - runix * (gradp + fgrav);
      fy = fy - runiy * (gradp + fgrav);
      fz = fz - runiz * (gradp + fgrav);
      pot = pot + pm * phii;
      drhodt += projv * grkerni;
    }
    // --- finish_cell_and_store_results with gravity (force.F90:2909-2927) ---
    const int o = abs(iparts[p]) - 1;
    double potensoft0, dum;
    kern_soft<K>(0., 0., potensoft0, dum);
    double epoti = 0.5 * pm * (pot + pm * potensoft0 * hi1);
    double gx, gy, gz, poti;
    const double *fn = G.fnode + (size_t)l * LENFGRAV;
    l2p(fn, xi - nd[c].xcen[0], yi - nd[c].xcen[1], zi - nd[c].xcen[2], gx, gy, gz, poti);
    fx = fx + gx;
    fy = fy + gy;
    fz = fz + gz;
    epoti = epoti + 0.5 * pm * poti;
    if (G.poten) G.poten[o] = (float)epoti;
    const double f2i = fx * fx + fy * fy + fz * fz;
    double *fo = out.fxyzu + (size_t)o * out.ldf;
    fo[0] = fx; fo[1] = fy; fo[2] = fz;
    const double drhodti = pm * drhodt;
    const double divvi = -drhodti * rho1i;
    out.divcurlv[(size_t)o * out.lddcv] = (float)divvi;
    double dtc = dc.dtmax, dtf = PH_BIGNUMBER;
    double fxyz4 = 0.;
    if (THERMAL) {
      const double pdv_work = pri * rho1i * rho1i * drhodti;
      if (dc.ipdv_heating > 0) fxyz4 += pdv_work;
      if (dc.ishock_heating > 0) fxyz4 += dudtdiss;
      fxyz4 += dendtdiss;
    }
    const double vsigdtc = fmax(vsigmax, vwavei);
    if (vsigdtc > PH_TINY64) dtc = dc.C_cour * hi / (vsigdtc * fmax(dc.alpha, 1.0));
    if (THERMAL) {
      if (eni + dtc * fxyz4 < PH_EPS64 && eni > PH_EPS64) fxyz4 = fxyz4 / (1. - dtc * fxyz4 / eni);
      fo[3] = fxyz4;
    }
    if (fabs(f2i) > PH_EPS64) dtf = dc.C_force * sqrt(hi / sqrt(f2i));
    if (!(dtc > PH_TINY64) || !(dtc < PH_HUGE64)) atomicOr(flags, GFL_DTC);
    dtc_o = dtc;
    dtf_o = dtf;
    if (ind) {
      double dti = dtc;
      if (dtf < dti + PH_TINY64 && dtf < dc.dtmax) dti = dtf;
      int b = get_newbin_d(dti, dc.dtmax, (int)out.ibin[o], icall < 2, out.nbinmax, out.istepfrac);
      b = max(b, ibin_neighi - 1);
      out.ibin[o] = (int8_t)b;
      bmax = b;
      dtc_o = PH_HUGE64;
      dtf_o = PH_HUGE64;
    }
  }
  reduce_force_block(dtc_o, dtf_o, nl, bmax, dtred, flags);
}

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------
// after maketree: node second moments and the node hmax mirror
int grav_tree(phsph_ctx *c) {
  hipStream_t s = c->stream;
  // the walk stack packs (node id << 6 | level) into 32 bits
  if (c->nslots > ((int64_t)1 << 26)) PH_RET(c, PHSPH_ERR_CAPACITY, "maketree: too many tree nodes for self-gravity");
  PH_HIPCHECK(c, c->quads.ensure((size_t)c->nslots * 6));
  PH_HIPCHECK(c, c->ghmax.ensure((size_t)c->nslots));
  if (c->ncells > 0) {
    LaunchTimer lt(c, "tree_quads");
    hipLaunchKernelGGL(k_quads, dim3((unsigned)c->ncells), dim3(64), 0, s, c->dc, c->ncells, c->nd.p, c->nd_lo.p,
                       c->nd_hi.p, c->leaf_flag.p, c->tx.p, c->ty.p, c->tz.p, c->quads.p);
  }
  hipLaunchKernelGGL(k_init_ghmax, dim3(nblocks(c->nslots, 256)), dim3(256), 0, s, (int64_t)c->nslots, c->nd.p,
                     c->ghmax.p);
  PH_HIPCHECK(c, hipGetLastError());
  return PHSPH_OK;
}

// after densityiterate: raise node hmax as the reference's relists do
int grav_hmax_replay(phsph_ctx *c, int icall) {
  if (icall == 0 || c->nleaves == 0) return PHSPH_OK;
  hipStream_t s = c->stream;
  int32_t zero = 0;
  PH_HIPCHECK(c, dev_set(c, c->counters.p + 4, &zero, sizeof(int32_t)));
  hipLaunchKernelGGL(k_hmax_replay, dim3(nblocks(c->nleaves, 128)), dim3(128), 0, s, c->nleaves, c->leaves.p,
                     c->leaf_flag.p, c->nd_lo.p, c->nd_hi.p, c->nd.p, c->p_nits.p, c->htraj.p, (size_t)c->htraj_stride, c->hconv.p,
                     c->ghmax.p, c->counters.p + 4);
  PH_HIPCHECK(c, hipGetLastError());
  int32_t fl = 0;
  PH_HIPCHECK(c, hipMemcpyAsync(&fl, c->counters.p + 4, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  PH_HIPCHECK(c, hipStreamSynchronize(s));
  if (fl & GFL_TRAJ) PH_RET(c, PHSPH_ERR_CAPACITY, "densityiterate: more than 17 iterations in a cell with self-gravity");
  return PHSPH_OK;
}

int grav_force(phsph_ctx *c, int icall, const ForceIOArgs &a, const FRec *rec, const FRec32 *rec32,
               const float *gsoft, double *dtc_out, double *dtf_out) {
  hipStream_t s = c->stream;
  const int nl = c->nleaves, nact = c->nact;
  PH_HIPCHECK(c, c->fnode.ensure((size_t)(nl + 1) * LENFGRAV));
  PH_HIPCHECK(c, c->np2p.ensure(nl + 1));
  // near-field lists at a uniform capacity, doubled (and kept) while it fits
  // in the list budget; beyond that the walk runs again with each leaf's list
  // sized from its count
  const int64_t *poff = nullptr;
  while (c->caps.p2p > 16 && !list_fits(c, (size_t)(nl + 1) * c->caps.p2p)) c->caps.p2p /= 2;
  for (int attempt = 0; attempt < 12; attempt++) {
    if (!poff) PH_HIPCHECK(c, c->p2p.ensure((size_t)(nl + 1) * c->caps.p2p));
    int32_t zero = 0;
    PH_HIPCHECK(c, dev_set(c, c->counters.p, &zero, sizeof(int32_t)));
    if (nl > 0) {
      LaunchTimer lt(c, "grav_walk");
      int rc = grav_walk(c, poff);
      if (rc) return rc;
    }
    PH_HIPCHECK(c, hipGetLastError());
    int32_t fl = 0;
    PH_HIPCHECK(c, hipMemcpyAsync(&fl, c->counters.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    if (fl & GFL_DEPTH) PH_RET(c, PHSPH_ERR_CAPACITY, "force: tree deeper than 32 levels for the gravity walk");
    if (fl & GFL_STACK) PH_RET(c, PHSPH_ERR_CAPACITY, "force: gravity walk stack overflow");
    if (!(fl & GFL_P2P)) break;
    if (poff) PH_RET(c, PHSPH_ERR_CAPACITY, "force: near-field list capacity exceeded");
    if (list_fits(c, (size_t)(nl + 1) * c->caps.p2p * 2)) {  // a larger uniform capacity, kept
      c->caps.p2p *= 2;
      continue;
    }
    PH_HIPCHECK(c, c->poff.ensure((size_t)nl + 1));
    hipLaunchKernelGGL(k_scan_offsets<int32_t>, dim3(1), dim3(1024), 0, s, c->np2p.p, nl, (int64_t)1, 1, c->poff.p);
    PH_HIPCHECK(c, hipGetLastError());
    int64_t total = 0;
    PH_HIPCHECK(c, hipMemcpyAsync(&total, c->poff.p + nl, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    PH_HIPCHECK(c, c->p2p.ensure((size_t)total + 1));
    poff = c->poff.p;
  }
  PH_HIPCHECK(c, c->dtred.ensure(3));
  unsigned long long init[3];
  double big = PH_HUGE64;
  memcpy(&init[0], &big, 8);
  init[1] = init[0];
  init[2] = 0ull;
  int32_t fl2[2] = {0, 0};
  unsigned long long red[3];
  // the second pass's neighbour lists: a capacity per particle, kept for
  // later calls; after an overflow the call runs again at the longest list,
  // or with per-group capacities when that does not fit in the list budget.
  // The kernel moves ibin to the new bin in place, so a re-run starts from
  // the bins the call was given.
  int capn = c->caps.gnear;
  const int64_t *noff = nullptr;
  PH_HIPCHECK(c, c->gnear_n.ensure((size_t)(nact > 0 ? nact : 1)));
  if (a.ibin && a.np > 0) {
    PH_HIPCHECK(c, c->ibin_save.ensure((size_t)a.np));
    PH_HIPCHECK(c, hipMemcpyAsync(c->ibin_save.p, a.ibin, (size_t)a.np, hipMemcpyDeviceToDevice, s));
  }
  for (int attempt = 0; attempt < 2; attempt++) {
    if (attempt > 0 && a.ibin && a.np > 0)
      PH_HIPCHECK(c, hipMemcpyAsync(a.ibin, c->ibin_save.p, (size_t)a.np, hipMemcpyDeviceToDevice, s));
    if (!noff) PH_HIPCHECK(c, c->gnear.ensure((size_t)list_cap(capn) * (size_t)(nact > 0 ? nact : 1)));
    PH_HIPCHECK(c, dev_set(c, c->dtred.p, init, sizeof(init)));
    const int32_t zero3[3] = {0, 0, 0};
    PH_HIPCHECK(c, dev_set(c, c->counters.p, zero3, 2 * sizeof(int32_t)));
    PH_HIPCHECK(c, dev_set(c, c->counters.p + 5, zero3, sizeof(int32_t)));
    GravArgs G{c->fnode.p, c->p2p.p, c->np2p.p, poff, c->caps.p2p, gsoft, c->poten_out,
               c->gnear.p, list_cap(capn), noff, c->gnear_n.p, c->counters.p + 5};
    if (nact > 0) {
      LaunchTimer lt(c, "force");
      const dim3 gr(xcd_grid(nblocks(nact, 128))), b(128);
      const bool th = c->p.maxvxyzu >= 4;
#define GL(KK, T, M)                                                                                            \
  hipLaunchKernelGGL((k_force_grav<KK, T, M>), gr, b, 0, s, c->dc, c->act.p, nact, c->pleaf.p, c->leaves.p,     \
                     c->nd_lo.p, c->nd_hi.p, c->nd.p, c->iparts.p, rec, c->fpos.p, c->fvel.p, rec32, G, a,      \
                     c->tbin_old.p, icall, c->dtred.p, c->counters.p)
#define GLK(KK, M) \
  if (th) GL(KK, true, M); else GL(KK, false, M);
      if (rec32) {
        if (c->p.kernel == PHSPH_KERNEL_CUBIC) { GLK(PHSPH_KERNEL_CUBIC, true) } else { GLK(PHSPH_KERNEL_QUINTIC, true) }
      } else {
        if (c->p.kernel == PHSPH_KERNEL_CUBIC) { GLK(PHSPH_KERNEL_CUBIC, false) } else { GLK(PHSPH_KERNEL_QUINTIC, false) }
      }
#undef GLK
#undef GL
    }
    PH_HIPCHECK(c, hipGetLastError());
    int32_t nmax = 0;
    PH_HIPCHECK(c, hipMemcpyAsync(fl2, c->counters.p, sizeof(fl2), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipMemcpyAsync(red, c->dtred.p, sizeof(red), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipMemcpyAsync(&nmax, c->counters.p + 5, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    if (!(fl2[0] & GFL_NEAR)) break;
    if (attempt == 1) PH_RET(c, PHSPH_ERR_CAPACITY, "force: neighbour list capacity exceeded");
    capn = nmax;
    if (list_fits(c, (size_t)list_cap(capn) * (size_t)nact)) {
      c->caps.gnear = list_cap(capn);
      continue;
    }
    // one particle with many neighbours would size every list: give each
    // group of 64 the longest list among its own particles instead
    const int ng = (nact + GROUP - 1) / GROUP;
    PH_HIPCHECK(c, c->gmax.ensure((size_t)ng));
    PH_HIPCHECK(c, c->gnear_off.ensure((size_t)ng + 1));
    hipLaunchKernelGGL(k_group_max, dim3(nblocks((size_t)ng * 64, 256)), dim3(256), 0, s, c->gnear_n.p, nact, ng,
                       c->gmax.p);
    hipLaunchKernelGGL(k_scan_offsets<int32_t>, dim3(1), dim3(1024), 0, s, c->gmax.p, ng, (int64_t)GROUP, LSEG,
                       c->gnear_off.p);
    PH_HIPCHECK(c, hipGetLastError());
    int64_t total = 0;
    PH_HIPCHECK(c, hipMemcpyAsync(&total, c->gnear_off.p + ng, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    PH_HIPCHECK(c, c->gnear.ensure((size_t)total + 1));
    noff = c->gnear_off.p;
  }
  if (fl2[0] & GFL_DTC) PH_RET(c, PHSPH_ERR_INPUT, "force: invalid dtc");
  if (a.ibin) c->nbinmax_new = fl2[1];
  double dtc, dtf;
  memcpy(&dtc, &red[0], 8);
  memcpy(&dtf, &red[1], 8);
  c->force_npairs = (int64_t)red[2];
  if (a.ibin) dtc = dtf = PH_BIGNUMBER;
  if (dtc_out) *dtc_out = (nact > 0) ? dtc : PH_BIGNUMBER;
  if (dtf_out) *dtf_out = (nact > 0) ? fmin(dtf, PH_BIGNUMBER) : PH_BIGNUMBER;
  return PHSPH_OK;
}
