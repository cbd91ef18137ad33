// gravity.hip -- kd-tree multipole self-gravity (the GRAVITY build of the
// reference): node second moments, the node hmax the force walk sees, the
// dual-tree walk per leaf with M2L/L2L (kdtree.F90:1357-1840) and the force
// pair loop with softened and point-mass gravity (force.F90:1286-2053,
// 2909-2927).
//
//  * k_quads: second moments of every node about its centre of mass
//    (kdtree.F90:724-750), one wavefront per node.
//  * k_hmax_replay: the reference raises a leaf's hmax (and its ancestors')
//    whenever a density iteration relists the cell (set_hmaxcell,
//    neigh_kdtree.f90:115-131; dens.F90:338-373).  The density kernel keeps
//    each particle's h after every iteration; this kernel replays the cell's
//    compute_hmax sequence (dens.F90:1275-1289) to give the force walk the
//    same node hmax, which decides which node pairs are well separated.
//  * k_grav_walk: one wavefront per leaf runs getneigh_dual breadth-parallel:
//    up to 64 (source, level) pairs per step; well-separated pairs add their
//    quadrupole expansion (compute_M2L) to the level's coefficients in LDS,
//    the rest open the source; leaf-leaf pairs that are not well separated
//    become the leaf's near-field list.  Lane 0 then shifts the per-level
//    expansions down the branch to the leaf (propagate_fnode_to_node).
//  * k_force_grav: one thread per active particle over its leaf's
//    near-field list: SPH pair terms plus softened gravity for neighbours,
//    1/r^2 for the other trial particles, then the far field by Taylor
//    expansion about the leaf centre (expand_fgrav_in_taylor_series).
#include "records.h"

constexpr int LENFGRAV = 20;
constexpr int GMAXDEPTH = 32;   // kdtree.F90:42
constexpr int GSTACK = 2048;    // (source, level) stack of the dual walk (LDS)

enum { GFL_STACK = 1, GFL_DEPTH = 2, GFL_P2P = 4, GFL_TRAJ = 8, GFL_DTC = 32 };

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
                              const double *__restrict__ htraj, const double *__restrict__ hconv,
                              unsigned long long *__restrict__ ghmax, int32_t *__restrict__ flags) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nleaves) return;
  const int c = leaves[l];
  if (leaf_flag[c] <= 0) return;
  const int lo = nd_lo[c], hi = nd_hi[c];
  int nits = 0;
  for (int p = lo; p <= hi; p++) nits = max(nits, p_nits[p]);
  if (nits <= 1) return;  // converged at the first iteration: no relist
  if (nits - 1 > HTRAJ) {
    atomicOr(flags, GFL_TRAJ);
    return;
  }
  double hcell = __longlong_as_double((long long)ghmax[c]);
  bool set = false;
  double hset = 0.;
  for (int k = 1; k < nits; k++) {  // not yet converged after iteration k
    double hmx = 0.;
    for (int p = lo; p <= hi; p++) {
      const int np_ = p_nits[p];
      if (np_ < 1) continue;
      const double v = (k < np_) ? htraj[(size_t)p * HTRAJ + k - 1] : hconv[p];
      hmx = fmax(hmx, v);
    }
    const double hm = 1.01 * hmx;
    if (hm > hcell) {
      set = true;
      hset = hm;
    }
    hcell = hm;
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
// dual-tree walk, one wavefront per leaf
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(64)
k_grav_walk(DevConst dc, int nleaves, const int32_t *__restrict__ leaves, const int32_t *__restrict__ leaf_flag,
            const phsph_kdnode *__restrict__ nd, const unsigned long long *__restrict__ ghmax,
            const double *__restrict__ quads, double *__restrict__ fnode_out, uint32_t *__restrict__ p2p,
            int32_t *__restrict__ np2p, const int64_t *poff, int ucapp, int32_t *__restrict__ flags) {
  __shared__ uint32_t stk[GSTACK];       // source id << 6 | level
  __shared__ double fb[GMAXDEPTH + 1][LENFGRAV];
  __shared__ int branch[GMAXDEPTH + 1];
  __shared__ int s_nparents;
  const int l = blockIdx.x;
  if (l >= nleaves) return;
  const int icell = leaves[l];
  if (leaf_flag[icell] <= 0) {
    if (threadIdx.x == 0) np2p[l] = 0;
    return;
  }
  // near-field list of this leaf: uniform capacity, or exact offsets
  const size_t pbase = poff ? (size_t)poff[l] : (size_t)l * ucapp;
  const int capp = poff ? (int)(poff[l + 1] - poff[l]) : ucapp;
  const int lane = threadIdx.x;
  for (int k = lane; k < (GMAXDEPTH + 1) * LENFGRAV; k += 64) (&fb[0][0])[k] = 0.;
  if (lane == 0) {
    int np_ = 1, j = icell;
    branch[1] = icell;
    bool ok = true;
    while (nd[j].parent != 0) {
      j = nd[j].parent;
      if (np_ + 1 > GMAXDEPTH) { ok = false; break; }
      branch[++np_] = j;
    }
    if (!ok) atomicOr(flags, GFL_DEPTH);
    s_nparents = ok ? np_ : 0;
    stk[0] = (1u << 6) | (uint32_t)np_;
  }
  __syncthreads();
  const int nparents = s_nparents;
  if (nparents == 0) return;
  const double R = dc.radkern;
  int top = 1, nc = 0;
  bool ovf = false;
  while (top > 0) {
    const int take = min(top, 64);
    uint32_t ent = 0u;
    if (lane < take) ent = stk[top - take + lane];
    __syncthreads();
    top -= take;
    bool stackit = false, p2p_hit = false;
    int code = 13, src = 0, lev = 0, npush = 0, c0 = 0, c1 = 0;
    if (lane < take) {
      src = (int)(ent >> 6);
      lev = (int)(ent & 63u);
      const int dst = branch[lev];
      if (dst == src) {
        stackit = true;
      } else {
        // node_interaction (kdtree.F90:1653-1694)
        double dx, dy, dz;
        const double r2 = node_sep(dc, nd[dst].xcen, nd[src].xcen, dx, dy, dz, code);
        const double rcut_dst = __longlong_as_double((long long)ghmax[dst]) * R;
        const double rcut_src = __longlong_as_double((long long)ghmax[src]) * R;
        const double rcut = fmax(rcut_dst, rcut_src);
        const double ss = nd[dst].size + nd[src].size;
        const double rcut2 = (ss + rcut) * (ss + rcut);
        const bool wellsep = (dc.tree_acc2 * r2 > ss * ss) && (r2 > rcut2);
        if (wellsep) {
          double f[LENFGRAV];
          m2l(dx, dy, dz, 1. / sqrt(r2), nd[src].mass, quads + 6 * (size_t)src, f);
          for (int k = 0; k < LENFGRAV; k++) atomicAdd(&fb[lev][k], f[k]);
        } else {
          stackit = true;
        }
      }
      if (stackit) {  // open_nodes (kdtree.F90:1591-1645)
        const bool dstleaf = !(lev - 1 > 0);
        const int levn = dstleaf ? lev : lev - 1;
        if (leaf_flag[src] != 0) {
          if (dstleaf) {
            p2p_hit = true;
          } else {
            c0 = (src << 6) | levn;
            npush = 1;
          }
        } else {
          const int il = nd[src].leftchild, ir = nd[src].rightchild;
          if (il != 0) { c0 = (il << 6) | levn; npush = 1; }
          if (ir != 0) {
            if (npush == 0) c0 = (ir << 6) | levn;
            else c1 = (ir << 6) | levn;
            npush++;
          }
        }
      }
    }
    // near-field leaves, with the separation code of the pair
    const unsigned long long mp = __ballot(p2p_hit);
    if (p2p_hit) {
      const int k = nc + __popcll(mp & lanemask_lt());
      int sc = SEP_PLAIN;
      if (dc.periodic) sc = (code == SEP_PLAIN) ? SEP_WRAP0 : code;
      if (k < capp) p2p[pbase + k] = cand_pack(src, sc);
    }
    nc += __popcll(mp);
    // pushes: exclusive scan of the per-lane counts
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
    if (npush > 0) stk[at] = (uint32_t)c0;
    if (npush > 1) stk[at + 1] = (uint32_t)c1;
    top += total;
    __syncthreads();
  }
  __syncthreads();
  if (lane == 0) {
    np2p[l] = nc;
    if (ovf) atomicOr(flags, GFL_STACK);
    if (nc > capp) atomicOr(flags, GFL_P2P);
    // downward pass: shift each level's expansion to the next centre
    double facc[LENFGRAV], ftmp[LENFGRAV];
    for (int k = 0; k < LENFGRAV; k++) facc[k] = 0.;
    for (int i = nparents; i >= 2; i--) {
      double dx, dy, dz;
      int cd;
      node_sep(dc, nd[branch[i - 1]].xcen, nd[branch[i]].xcen, dx, dy, dz, cd);
      for (int k = 0; k < LENFGRAV; k++) ftmp[k] = facc[k] + fb[i][k];
      l2l(facc, ftmp, dx, dy, dz);
    }
    for (int k = 0; k < LENFGRAV; k++) fnode_out[(size_t)l * LENFGRAV + k] = facc[k] + fb[1][k];
  }
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
};

template <int K, bool THERMAL>
__global__ void __launch_bounds__(128)
k_force_grav(DevConst dc, const int32_t *__restrict__ act, int nact, const int32_t *__restrict__ pleaf,
             const int32_t *__restrict__ leaves, const int32_t *__restrict__ nd_lo, const int32_t *__restrict__ nd_hi,
             const phsph_kdnode *__restrict__ nd, const int32_t *__restrict__ iparts, const FRec *__restrict__ rec,
             GravArgs G, ForceIOArgs out, const int8_t *__restrict__ tbin_old, int icall,
             unsigned long long *__restrict__ dtred, int32_t *__restrict__ flags) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
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
    for (int e = 0; e < ne; e++) {
      const uint32_t en = G.p2p[pbase + e];
      const int cn = cand_node(en), code = cand_code(en);
      const int j1 = nd_hi[cn];
      for (int j = nd_lo[cn]; j <= j1; j++) {
        if (j == p) continue;
        const FRec &Rj = rec[j];
        double dx, dy, dz;
        sep_of(dc, code, xi, yi, zi, Rj.x, Rj.y, Rj.z, dx, dy, dz);
        const double rij2 = r2_exact(dx, dy, dz);
        const double q2i = xmul(rij2, hi21);
        const double hj1 = Rj.hinv;
        const double hj21 = xmul(hj1, hj1);
        const double q2j = xmul(rij2, hj21);
        if (!(q2i < R2 || q2j < R2)) {
          // trial neighbour outside both kernels: 1/r^2 (force.F90:1991-2053)
          const double rij1 = 1. / sqrt(rij2);
          const double fgravj = rij1 * rij1 * rij1 * pm;
          fx = fx - dx * fgravj;
          fy = fy - dy * fgravj;
          fz = fz - dz * fgravj;
          pot = pot + pm * (-rij1);
          continue;
        }
        nl++;
        if (ind) {  // force.F90:1346-1352
          ibin_neighi = max(ibin_neighi, (int)tbin_old[j]);
          const int o = abs(iparts[j]) - 1;
          if (out.ibin_wake[o] < ibinnow_m1) out.ibin_wake[o] = ibinnow_m1;
        }
        double rij1, qi;
        if (rij2 > PH_TINY64) {
          rij1 = 1. / sqrt(rij2);
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
        double grkernj = 0., fgravj;
        bool usej = false;
        if (q2j < R2) {
          const double qj = (rij2 * rij1) * hj1;
          grkernj = kern_g<K, double>(q2j, qj) * hj21 * hj21 * dc.cnormk * (double)Rj.gradh;
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
        const double dvx = vxi - Rj.vx, dvy = vyi - Rj.vy, dvz = vzi - Rj.vz;
        const double projv = dvx * runix + dvy * runiy + dvz * runiz;
        const double denij = THERMAL ? (eni - Rj.u) : 0.;
        const double vsigi = fmax(vwavei - dc.beta * projv, 0.);
        const double vsigavi = fmax(alphai * vwavei - dc.beta * projv, 0.);
        vsigmax = fmax(vsigmax, vsigi);
        double rho1j = 0., pro2j = 0., vwavej = 0., vsigavj = 0., autermj = 0.;
        if (usej) {
          rho1j = Rj.rho1;
          pro2j = Rj.pro2;
          vwavej = Rj.vwave;
          autermj = 0.5 * pm * rho1j * dc.alphau;
          vsigmax = fmax(vsigmax, fmax(vwavej - dc.beta * projv, 0.));
          vsigavj = fmax((double)Rj.alpha * vwavej - dc.beta * projv, 0.);
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
  PH_HIPCHECK(c, hipMemcpyAsync(c->counters.p + 4, &zero, sizeof(int32_t), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_hmax_replay, dim3(nblocks(c->nleaves, 128)), dim3(128), 0, s, c->nleaves, c->leaves.p,
                     c->leaf_flag.p, c->nd_lo.p, c->nd_hi.p, c->nd.p, c->p_nits.p, c->htraj.p, c->hconv.p,
                     c->ghmax.p, c->counters.p + 4);
  PH_HIPCHECK(c, hipGetLastError());
  int32_t fl = 0;
  PH_HIPCHECK(c, hipMemcpyAsync(&fl, c->counters.p + 4, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  PH_HIPCHECK(c, hipStreamSynchronize(s));
  if (fl & GFL_TRAJ) PH_RET(c, PHSPH_ERR_CAPACITY, "densityiterate: more than 17 iterations in a cell with self-gravity");
  return PHSPH_OK;
}

int grav_force(phsph_ctx *c, int icall, const ForceIOArgs &a, const FRec *rec, const float *gsoft,
               double *dtc_out, double *dtf_out) {
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
    PH_HIPCHECK(c, hipMemcpyAsync(c->counters.p, &zero, sizeof(int32_t), hipMemcpyHostToDevice, s));
    if (nl > 0) {
      LaunchTimer lt(c, "grav_walk");
      hipLaunchKernelGGL(k_grav_walk, dim3(nl), dim3(64), 0, s, c->dc, nl, c->leaves.p, c->leaf_flag.p, c->nd.p,
                         c->ghmax.p, c->quads.p, c->fnode.p, c->p2p.p, c->np2p.p, poff, c->caps.p2p,
                         c->counters.p);
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
  PH_HIPCHECK(c, hipMemcpyAsync(c->dtred.p, init, sizeof(init), hipMemcpyHostToDevice, s));
  int32_t zero2[2] = {0, 0};
  PH_HIPCHECK(c, hipMemcpyAsync(c->counters.p, zero2, sizeof(zero2), hipMemcpyHostToDevice, s));
  GravArgs G{c->fnode.p, c->p2p.p, c->np2p.p, poff, c->caps.p2p, gsoft, c->poten_out};
  if (nact > 0) {
    LaunchTimer lt(c, "force");
    const dim3 gr(nblocks(nact, 128)), b(128);
    const bool th = c->p.maxvxyzu >= 4;
#define GL(KK, T)                                                                                               \
  hipLaunchKernelGGL((k_force_grav<KK, T>), gr, b, 0, s, c->dc, c->act.p, nact, c->pleaf.p, c->leaves.p,        \
                     c->nd_lo.p, c->nd_hi.p, c->nd.p, c->iparts.p, rec, G, a, c->tbin_old.p, icall, c->dtred.p,  \
                     c->counters.p)
    if (c->p.kernel == PHSPH_KERNEL_CUBIC) {
      if (th) GL(PHSPH_KERNEL_CUBIC, true); else GL(PHSPH_KERNEL_CUBIC, false);
    } else {
      if (th) GL(PHSPH_KERNEL_QUINTIC, true); else GL(PHSPH_KERNEL_QUINTIC, false);
    }
#undef GL
  }
  PH_HIPCHECK(c, hipGetLastError());
  int32_t fl2[2] = {0, 0};
  unsigned long long red[3];
  PH_HIPCHECK(c, hipMemcpyAsync(fl2, c->counters.p, sizeof(fl2), hipMemcpyDeviceToHost, s));
  PH_HIPCHECK(c, hipMemcpyAsync(red, c->dtred.p, sizeof(red), hipMemcpyDeviceToHost, s));
  PH_HIPCHECK(c, hipStreamSynchronize(s));
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
