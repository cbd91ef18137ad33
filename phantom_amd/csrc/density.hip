// density.hip -- neighbour search and the h-rho density iteration
// (densityiterate, dens.F90:117-568; get_density_sums :578-857;
// finish_cell :1382-1466; finish_rhosum :1470-1507; store_results
// :1511-1681).
//
// Work decomposition:
//  * k_group_walk: one wavefront per group of 64 active particles walks the
//    tree breadth-parallel (64 nodes per step) and lists the candidate leaves
//    of the whole group (stream.h).
//  * k_dens_iter: one wavefront per group, one particle per lane.  Each lane
//    visits the candidate leaves its particle can reach (build_list),
//    keeps the particles within rfac*radkern*h of its own particle in a
//    neighbour list, then runs its own Newton-Raphson
//    iteration on h over that list with the two sums the step needs.
//    Per-particle iteration gives the reference's result: in a cell the
//    reference keeps re-evaluating converged particles at their unchanged h,
//    so each particle's h sequence depends only on its own sums.  A particle
//    whose h grows past its list radius (or whose list overflows) is run
//    again with a wider search.
//  * k_dens_final: one thread per particle; every sum of get_density_sums
//    at the h of the last iteration over the list, then store_results.
//  * k_walk: the reference's per-leaf walk (getneigh, kdtree.F90:1221-1347),
//    used only for the trial-neighbour statistics.
//
// Every neighbour membership test computes r^2 = ((dx*dx + dy*dy) + dz*dz)
// and q^2 = r^2 * (1/h)^2 with separately rounded FP64 operations, exactly
// as the reference, so neighbour counts agree bit for bit.
#include "stream.h"

#include <algorithm>
#include <vector>
#include <cstdlib>

enum { DFL_CAND = 1, DFL_LIST = 2, DFL_REWALK = 4, DFL_NOCONV = 8, DFL_NEGH = 16, DFL_STACK = 32 };

struct DensIn {
  const double4 *tp;  // x, y, z (treecache) and h at entry
  const double4 *tv;  // vx, vy, vz
  const double4 *tf;  // fxyzu + fext (null: zero)
  const double4 *tB;  // Bevol x, y, z and rho(h at entry) (MHD)
};
struct DensOut {
  double *xyzh;  int ldx;
  float *divcurlv; int lddcv;
  float *divcurlB;
  float *alphaind;
  float *gradh; int ldgradh;
  float *dvdx;
};

// ------------------------------------------------------------------------
// per-leaf walk of the reference (trial-neighbour statistics only)
// ------------------------------------------------------------------------
// hmax of each leaf at the last neighbour list the reference makes for it:
// the relisted hmax of the replay, else the tree's (get_cell_location)
__global__ void k_trial_hmax(int nleaves, const int32_t *__restrict__ leaves, const int32_t *__restrict__ leaf_flag,
                             const int32_t *__restrict__ nd_lo, const int32_t *__restrict__ nd_hi,
                             const phsph_kdnode *__restrict__ nd, const int32_t *__restrict__ p_nits,
                             const double *__restrict__ htraj, size_t hstride, const double *__restrict__ hconv,
                             double *__restrict__ lhmax, int32_t *__restrict__ flags) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nleaves) return;
  const int c = leaves[l];
  double h = nd[c].hmax;
  if (leaf_flag[c] > 0) {
    double hset = 0.;
    const int set = leaf_relist_hmax(nd_lo[c], nd_hi[c], h, p_nits, htraj, hstride, hconv, &hset);
    if (set < 0) atomicOr(flags, DFL_NOCONV);
    if (set > 0) h = hset;
  }
  lhmax[l] = h;
}

// lhmax: hmax of each leaf by leaf index (nullptr: the node's)
__global__ void k_walk(DevConst dc, double rfac, const phsph_kdnode *__restrict__ nd,
                       const int32_t *__restrict__ leaf_flag, const int32_t *__restrict__ nd_lo,
                       const int32_t *__restrict__ nd_hi, const int32_t *__restrict__ leaves, int nleaves,
                       const double *__restrict__ lhmax, int32_t *__restrict__ ntrial, int32_t *__restrict__ flags) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nleaves) return;
  const int c = leaves[l];
  if (leaf_flag[c] <= 0) {
    ntrial[l] = 0;
    return;
  }
  const double xp0 = nd[c].xcen[0], xp1 = nd[c].xcen[1], xp2 = nd[c].xcen[2];
  const double xsizei = nd[c].size;
  // rcut = radkern*hmax (get_cell_location, neigh_kdtree.f90:363)
  const double rcuti = xmul(xmul(rfac, dc.radkern), lhmax ? lhmax[l] : nd[c].hmax);
  int stack[64];
  int ns = 1;
  stack[0] = 1;
  int ntr = 0;
  while (ns > 0) {
    const int n = stack[--ns];
    const phsph_kdnode e = nd[n];
    double dx = xsub(xp0, e.xcen[0]), dy = xsub(xp1, e.xcen[1]), dz = xsub(xp2, e.xcen[2]);
    if (dc.periodic) {
      if (fabs(dx) > dc.hdlx) dx = xsub(dx, copysign(dc.dxbound, dx));
      if (fabs(dy) > dc.hdly) dy = xsub(dy, copysign(dc.dybound, dy));
      if (fabs(dz) > dc.hdlz) dz = xsub(dz, copysign(dc.dzbound, dz));
    }
    const double r2 = r2_exact(dx, dy, dz);
    const double s = xadd(xadd(xsizei, e.size), rcuti);
    if (r2 < xmul(s, s)) {
      if (leaf_flag[n] != 0) {
        ntr += nd_hi[n] - nd_lo[n] + 1;
      } else {
        if (ns + 2 > 64) {
          atomicOr(flags, DFL_CAND);
          break;
        }
        if (e.leftchild != 0) stack[ns++] = e.leftchild;
        if (e.rightchild != 0) stack[ns++] = e.rightchild;
      }
    }
  }
  ntrial[l] = ntr;
}

// ------------------------------------------------------------------------
// group walk: candidate leaves of 64 consecutive active particles
// mode 0: rcut = rfac*radkern*max(h of the group)
// mode 1: rcut = max(that, radkern*hmax of the node)      (force, get_hj)
// mode 0 with hmaxn: rcut = rfac*radkern*max(h of the group, hmax of the node)
//   (density lists that also hold every j whose own radius reaches i)
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(64)
k_group_walk(DevConst dc, int mode, double rfac, const phsph_kdnode *__restrict__ nd,
             const double *__restrict__ hmaxn, const int32_t *__restrict__ leaf_flag,
             const int32_t *__restrict__ act, int nact, const double *__restrict__ tx,
             const double *__restrict__ ty, const double *__restrict__ tz, const double *__restrict__ hsrc,
             int ngroups, uint32_t *__restrict__ gcand, int32_t *__restrict__ gncand, const int64_t *goff,
             int capg, int32_t *__restrict__ flags) {
  __shared__ int stk[WALK_STACK];
  const int g = xcd_block(blockIdx.x, gridDim.x);
  if (g >= ngroups) return;
  const size_t gbase = cand_base(goff, g, capg);
  const int gcap = cand_cap(goff, g, capg);
  const int lane = threadIdx.x;
  const int t = g * GROUP + lane;
  double xl = PH_HUGE64, yl = PH_HUGE64, zl = PH_HUGE64, xh = -PH_HUGE64, yh = -PH_HUGE64, zh = -PH_HUGE64;
  double hm = 0.;
  if (t < nact && act[t] >= 0) {
    const int p = act[t];
    xl = xh = tx[p];
    yl = yh = ty[p];
    zl = zh = tz[p];
    hm = hsrc[p];
  }
  xl = wave_min(xl); yl = wave_min(yl); zl = wave_min(zl);
  xh = wave_max(xh); yh = wave_max(yh); zh = wave_max(zh);
  hm = wave_max(hm);
  const double bx = 0.5 * (xl + xh), by = 0.5 * (yl + yh), bz = 0.5 * (zl + zh);
  // half extents, rounded up so that every particle lies inside
  const double ex = 0.5 * (xh - xl) * (1. + 1e-12) + 1e-300, ey = 0.5 * (yh - yl) * (1. + 1e-12) + 1e-300,
               ez = 0.5 * (zh - zl) * (1. + 1e-12) + 1e-300;
  const double rcg = rfac * dc.radkern * hm;
  if (lane == 0) stk[0] = 1;
  int top = 1, nc = 0;
  bool ovf = false;
  __syncthreads();
  while (top > 0) {
    const int take = min(top, 64);
    int node = 0;
    if (lane < take) node = stk[top - take + lane];
    __syncthreads();
    top -= take;
    bool open = false, isleaf = false;
    int lc = 0, rc = 0;
    if (node > 0) {
      const phsph_kdnode e = nd[node];
      double dx = e.xcen[0] - bx, dy = e.xcen[1] - by, dz = e.xcen[2] - bz;
      if (dc.periodic) {
        if (fabs(dx) > dc.hdlx) dx -= copysign(dc.dxbound, dx);
        if (fabs(dy) > dc.hdly) dy -= copysign(dc.dybound, dy);
        if (fabs(dz) > dc.hdlz) dz -= copysign(dc.dzbound, dz);
      }
      const double qx = fmax(fabs(dx) - ex, 0.), qy = fmax(fabs(dy) - ey, 0.), qz = fmax(fabs(dz) - ez, 0.);
      const double d2 = qx * qx + qy * qy + qz * qz;
      double rc_ = rcg;
      if (hmaxn) rc_ = fmax(rc_, (mode == 1 ? 1. : rfac) * dc.radkern * hmaxn[node]);
      const double s = (e.size + rc_) * (1. + 1e-10);
      open = d2 < s * s;
      isleaf = leaf_flag[node] != 0;
      lc = e.leftchild;
      rc = e.rightchild;
    }
    const unsigned long long ml = __ballot(open && isleaf);
    if (open && isleaf) {
      const int k = nc + __popcll(ml & lanemask_lt());
      if (k < gcap) gcand[gbase + k] = (uint32_t)node;
    }
    nc += __popcll(ml);
    const bool push = open && !isleaf;
    const unsigned long long mi = __ballot(push);
    const int npush = 2 * __popcll(mi);
    if (top + npush > WALK_STACK) {
      ovf = true;
      break;
    }
    if (push) {
      const int k = top + 2 * __popcll(mi & lanemask_lt());
      stk[k] = lc;
      stk[k + 1] = rc;
    }
    top += npush;
    __syncthreads();
  }
  if (lane == 0) {
    gncand[g] = nc;
    if (nc > gcap) atomicOr(flags, DFL_CAND);
    if (ovf) atomicOr(flags, DFL_STACK);
  }
}

// ------------------------------------------------------------------------
// gathers: caller layout (original order) -> tree order
// ------------------------------------------------------------------------
__global__ void k_gather_dens(DevConst dc, int nlive, const int32_t *__restrict__ iparts,
                              const double *__restrict__ tx, const double *__restrict__ ty,
                              const double *__restrict__ tz, const double *__restrict__ xyzh, int ldx,
                              const double *__restrict__ vxyzu, int ldv, const double *__restrict__ fxyzu, int ldf,
                              const double *__restrict__ fext, const double *__restrict__ Bevol,
                              const int8_t *__restrict__ iphase, double *th_in, double4 *tp, double4 *tv,
                              double4 *tf, double4 *tB, int32_t *actflag) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nlive) return;
  int ip = iparts[p];
  int o = abs(ip) - 1;
  const double h = xyzh[(size_t)o * ldx + 3];
  th_in[p] = h;
  tp[p] = make_double4(tx[p], ty[p], tz[p], h);
  tv[p] = make_double4(vxyzu[(size_t)o * ldv + 0], vxyzu[(size_t)o * ldv + 1], vxyzu[(size_t)o * ldv + 2], 0.);
  if (tf) {
    double a0 = 0., a1 = 0., a2 = 0.;
    if (fxyzu) {
      a0 = fxyzu[(size_t)o * ldf + 0]; a1 = fxyzu[(size_t)o * ldf + 1]; a2 = fxyzu[(size_t)o * ldf + 2];
    }
    if (fext) {
      a0 = xadd(a0, fext[(size_t)o * 3 + 0]); a1 = xadd(a1, fext[(size_t)o * 3 + 1]);
      a2 = xadd(a2, fext[(size_t)o * 3 + 2]);
    }
    tf[p] = make_double4(a0, a1, a2, 0.);
  }
  if (Bevol)
    tB[p] = make_double4(Bevol[(size_t)o * 4 + 0], Bevol[(size_t)o * 4 + 1], Bevol[(size_t)o * 4 + 2],
                         rhoh_d(h, dc.hfact, dc.pmass));
  // start_cell (dens.F90:1310-1331): skip inactive particles
  actflag[p] = (ip > 0) && (iphase == nullptr || iphase[o] > 0);
}

// stable compaction of flagged tree positions into act[]
__global__ void k_compact_count(const int32_t *flag, int n, int32_t *bcount) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int b = (i < n) ? flag[i] : 0;
  unsigned long long m = __ballot(b);
  __shared__ int shi[4];
  if ((threadIdx.x & 63) == 0) shi[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) bcount[blockIdx.x] = shi[0] + shi[1] + shi[2] + shi[3];
}
__global__ void k_compact_fill(const int32_t *flag, int n, const int32_t *boff, int32_t *out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int b = (i < n) ? flag[i] : 0;
  unsigned long long m = __ballot(b);
  __shared__ int shi[4];
  int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (l == 0) shi[w] = __popcll(m);
  __syncthreads();
  int off = boff[blockIdx.x];
  for (int k = 0; k < w; k++) off += shi[k];
  if (b) out[off + __popcll(m & ((1ull << l) - 1ull))] = i;
}
// exclusive scan of a[0..n) in place, total to *total: one block of 1024
// threads, four consecutive items per thread
__global__ void __launch_bounds__(1024) k_scan_i32(int32_t *a, int n, int32_t *total) {
  __shared__ int wsum[16];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int carry = 0;
  for (int base = 0; base < n; base += 4096) {
    const int i0 = base + tid * 4;
    int v[4];
    for (int k = 0; k < 4; k++) v[k] = (i0 + k < n) ? a[i0 + k] : 0;
    const int t = v[0] + v[1] + v[2] + v[3];
    int incl = t;
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(incl, o, 64);
      if (lane >= o) incl += u;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int woff = 0, btot = 0;
    for (int k = 0; k < 16; k++) {
      const int x = wsum[k];
      if (k < w) woff += x;
      btot += x;
    }
    int run = carry + woff + incl - t;
    for (int k = 0; k < 4; k++) {
      if (i0 + k < n) a[i0 + k] = run;
      run += v[k];
    }
    carry += btot;
    __syncthreads();
  }
  if (tid == 0) *total = carry;
}

// ------------------------------------------------------------------------
// density, part 1: neighbour list and Newton-Raphson iteration on h.
// One wavefront per group of 64 active particles, one particle per lane.
// Each lane visits the candidate leaves of its group that its particle can
// reach and keeps the particles within its own radius rfac*radkern*h0 (a
// superset of every kernel support the iteration can reach) in its list
// (list_at: 64-byte segments per lane).
// The iteration itself evaluates only the sums the
// Newton step needs (rho and dOmega); the full set of get_density_sums is
// taken once, at the h of the last iteration, by part 2.
// one neighbour's terms of the h iteration's density and gradh sums
// (get_density_sums, dens.F90:646-692) at 1/h = hi1, 1/h^2 = hi21
template <int K>
__device__ __forceinline__ void dens_iter_term(double r2, double hi1, double hi21, double R2, double pm,
                                               double &srho, double &sgh) {
  const double q2 = xmul(r2, hi21);
  if (q2 < R2) {
    const double q = sqrt(r2) * hi1;
    double wk, gk;
    kern_wg<K, double>(q2, q, wk, gk);
    srho += wk * pm;
    sgh += (-q * gk - 3. * wk) * pm;
  }
}

// ------------------------------------------------------------------------
template <int K, int CL, int LW, bool STORE_R2, bool QUEUE>
__global__ void __launch_bounds__(64 * LW) __attribute__((amdgpu_waves_per_eu(5)))
k_dens_iter(DevConst dc, int icall, double rfac, const int32_t *__restrict__ act, int nact, int ngroups,
            const int32_t *__restrict__ pleaf, const int32_t *__restrict__ leaves,
            const uint32_t *__restrict__ gcand, const int32_t *__restrict__ gncand, const int64_t *goff, int capg,
            const int32_t *__restrict__ nd_lo, const int32_t *__restrict__ nd_hi,
            const phsph_kdnode *__restrict__ nd, const double4 *__restrict__ tp, const double *__restrict__ hmax0,
            double lmargin, uint32_t *__restrict__ nlist, double *__restrict__ nr2, int capl,
            int32_t *__restrict__ t_nl,
            int32_t *__restrict__ t_nlj,
            double *__restrict__ t_hsum, double *__restrict__ t_hfin,
            int32_t *__restrict__ p_nits, int32_t *__restrict__ leaf_nits, int32_t *__restrict__ retry,
            double *__restrict__ htraj, size_t hstride, double *__restrict__ hconv, int32_t *__restrict__ flags,
            int32_t *__restrict__ wq) {
  __shared__ ListStage<CL> Sw[LW];
  ListStage<CL> &S = Sw[threadIdx.x >> 6];
  // one group of 64 active particles: its lists and h iteration
  auto group = [&](const int g) {
  const int lane = threadIdx.x & 63;
  const int t = g * GROUP + lane;
  const bool live = t < nact && act[t] >= 0;  // act[t] < 0: padding of a partial group
  int p = -1, l = 0;
  double xi = 0., yi = 0., zi = 0., h0 = 1., lcx = 0., lcy = 0., lcz = 0., lsz = 0.;
  if (live) {
    p = act[t];
    l = pleaf[p];
    const int c = leaves[l];
    const double4 r = tp[p];
    xi = r.x; yi = r.y; zi = r.z; h0 = r.w;
    lcx = nd[c].xcen[0]; lcy = nd[c].xcen[1]; lcz = nd[c].xcen[2]; lsz = nd[c].size;
  }
  const double rcut = rfac * dc.radkern * h0;
  // list radius: rcut (times a margin), and with hmax0 also every j whose own
  // radius rfac*radkern*h0_j reaches i, so that the list holds the force
  // neighbours (q_i < R or q_j < R at the converged h) as well
  const double rl = rfac * lmargin * dc.radkern;
  const double rcl = rl * h0, rc2 = rcl * rcl;
  // --- neighbour list over the leaves this particle can reach ---
  // entries within its own radius fill the list from the front (the density
  // sums read only those), entries only within the neighbour's radius fill
  // it from the back (for the force pass)
  int nl = 0, nlj = 0;
  // The sums of the first Newton iteration (h = h0) are taken while the list
  // is built: the own entries arrive in list order with the separations a
  // pass over the list computes, so these are that pass's sums.
  const double pm = dc.pmass;
  const double R2 = dc.radkern2;
  const double h01 = 1. / h0, h021 = xmul(h01, h01);
  double srho0 = 0., sgh0 = 0.;
  build_list(
      S, dc, gcand + cand_base(goff, g, capg), gncand[g], nd, nd_lo, nd_hi, hmax0, tp, xi, yi, zi, lcx, lcy, lcz, lsz,
      [&](double ccx, double ccy, double ccz, double csz, double chm) {
        return live && leaf_in_reach(dc, xi, yi, zi, ccx, ccy, ccz, csz, fmax(rcl, rl * chm));
      },
      [&](int j, int code, double r2, double hj0) {
        const bool own = r2 < rc2;
        if (j != p && (own || (hmax0 && r2 < (rl * hj0) * (rl * hj0)))) {
          if (nl + nlj < capl) {
            const size_t a = list_at(nact, t, own ? nl : capl - 1 - nlj);
            nlist[a] = cand_pack(j, code);
            if (STORE_R2 && own) nr2[a] = r2;  // the separation the iteration needs, exact as get_density_sums takes it
          }
          if (own) dens_iter_term<K>(r2, h01, h021, R2, pm, srho0, sgh0);
          nl += own ? 1 : 0;
          nlj += own ? 0 : 1;
        }
      });
  if (!live) return;
  if (nl + nlj > capl) {
    atomicOr(flags, DFL_LIST);
    retry[p] = 1;
    return;
  }
  // --- Newton-Raphson on h (finish_cell, dens.F90:1382-1466) ---
  double hi = h0, hfin = h0;
  int it = 0;
  for (;;) {
    it++;
    const double hi1 = 1. / hi, hi21 = xmul(hi1, hi1);
    double srho = 0., sgh = 0.;
    auto term = [&](double r2) { dens_iter_term<K>(r2, hi1, hi21, R2, pm, srho, sgh); };
    if (it == 1) {
      srho = srho0;
      sgh = sgh0;
    } else if constexpr (STORE_R2) {
      // the squared separations the search stored with the list, two
      // entries (16 bytes of the slot's segment) per load, read ahead
      double2 rn = make_double2(0., 0.);
      if (nl > 0) rn = *reinterpret_cast<const double2 *>(nr2 + list_at(nact, t, 0));
      for (int k = 0; k < nl; k += 2) {
        const double2 rr = rn;
        if (k + 2 < nl) rn = *reinterpret_cast<const double2 *>(nr2 + list_at(nact, t, k + 2));
        term(rr.x);
        if (k + 1 < nl) term(rr.y);
      }
    } else {
      // the separation again from the positions: the record of entry k+1
      // and the list entry k+2 are read while entry k is summed
      uint32_t en = 0u, en2 = 0u;
      double4 rn = make_double4(0., 0., 0., 0.);
      if (nl > 0) {
        en = nlist[list_at(nact, t, 0)];
        rn = tp[cand_node(en)];
      }
      if (nl > 1) en2 = nlist[list_at(nact, t, 1)];
      for (int k = 0; k < nl; k++) {
        const uint32_t e = en;
        const double4 r = rn;
        if (k + 1 < nl) {
          en = en2;
          rn = tp[cand_node(en)];
          if (k + 2 < nl) en2 = nlist[list_at(nact, t, k + 2)];
        }
        double dx, dy, dz;
        sep_of(dc, cand_code(e), xi, yi, zi, r.x, r.y, r.z, dx, dy, dz);
        term(r2_exact(dx, dy, dz));
      }
    }
    const double hi31 = hi1 * hi21, hi41 = hi21 * hi21;
    const double rhoi = dc.cnormk * (srho + dc.wab0 * pm) * hi31;
    const double gradhi = dc.cnormk * (sgh + dc.gradh0 * pm) * hi41;
    const double rhohi = rhoh_d(hi, dc.hfact, pm);
    const double dhdrhoi = -hi / (3. * rhohi);
    const double omegai = 1. - dhdrhoi * gradhi;
    const double func = rhohi - rhoi;
    const double dfdh1 = (omegai > PH_TINY64) ? dhdrhoi / omegai : dhdrhoi / fabs(omegai + PH_EPS64);
    double hnew = hi - func * dfdh1;
    if (hnew > 1.2 * hi) hnew = 1.2 * hi;
    else if (hnew < 0.8 * hi) hnew = 0.8 * hi;
    const bool conv = (fabs(hnew - hi) / h0 < dc.tolh) && omegai > 0. && hi > 0.;
    if (icall == 0) {  // one pass: store at hnew with the sums taken at h0 (dens.F90:341)
      hfin = conv ? hi : hnew;
      break;
    }
    if (conv) {
      hfin = hi;
      break;
    }
    if (it >= PH_MAXDENSITS) {
      atomicOr(flags, DFL_NOCONV);
      retry[p] = 1;
      return;
    }
    hi = hnew;
    // iteration-major: the lanes still iterating write adjacent words
    if (htraj && it <= HTRAJ) htraj[(size_t)(it - 1) * hstride + p] = hi;
    if (dc.radkern * hi * (1. + 1e-12) > rcut) {  // outgrew the list: run again with a wider one
      atomicOr(flags, DFL_REWALK);
      retry[p] = 1;
      return;
    }
  }
  if (hconv) hconv[p] = hi;  // the h of the last summation (cell%h once converged)
  t_nl[t] = nl;
  if (t_nlj) t_nlj[t] = nlj;
  t_hsum[t] = hi;
  t_hfin[t] = hfin;
  p_nits[p] = (icall == 0) ? 1 : it;
  atomicMax(&leaf_nits[l], (icall == 0) ? 1 : it);
  };
  if constexpr (!QUEUE) {
    // one group per wavefront, XCD-contiguous
    const int g = list_group<LW>(gridDim.x);
    if (g < ngroups) group(g);
  } else {
    // persistent wavefronts: the groups are split into 8 contiguous ranges,
    // one per XCD (blocks are dispatched to the XCDs round robin); each
    // wavefront takes the next group of its XCD's range from that range's
    // counter and, once the range is exhausted, of the other ranges in
    // turn.  Counters only grow, so every wavefront ends after at most
    // 8 exhausted draws plus the groups it ran.
    const int x = blockIdx.x & 7, per = (ngroups + 7) >> 3;
    for (int k = 0; k < 8; k++) {
      const int r = (x + k) & 7;
      for (;;) {
        int tk = 0;
        if ((threadIdx.x & 63) == 0) tk = atomicAdd(&wq[r], 1);
        tk = __shfl(tk, 0, 64);
        const int g = r * per + tk;
        if (tk >= per || g >= ngroups) break;
        group(g);
      }
    }
  }
}

// store_results (dens.F90:1511-1681) at the final h from the sums taken at
// the h of the last iteration
template <bool MHD>
__device__ __forceinline__ void dens_store(const DevConst &dc, int p, double hfin, double rs_rho, double rs_gh,
                                           double rs_gs, double rs_divv, double rs_divB, const double *dv,
                                           const double *da, const double *rr, const double *dB,
                                           const int32_t *__restrict__ iparts, double *__restrict__ th_out,
                                           const DensOut &out, int32_t *__restrict__ p_nneigh, int nneigh,
                                           int32_t *__restrict__ flags) {
  const double pm = dc.pmass;
  const double hf1 = 1. / hfin, hf21 = hf1 * hf1, hf31 = hf1 * hf21, hf41 = hf21 * hf21;
  const double rhoi = dc.cnormk * (rs_rho + dc.wab0 * pm) * hf31;
  const double gradhs = dc.cnormk * (rs_gh + dc.gradh0 * pm) * hf41;
  const double dhdrhoi = -hfin / (3. * rhoh_d(hfin, dc.hfact, pm));
  const double omegai = 1. - dhdrhoi * gradhs;
  const double gradhi = 1. / omegai;
  const double gradsofti = ((rs_gs + dc.dphidh0 * pm) * hf21) * dhdrhoi;
  const double hnew = hrho_d(rhoi, dc.hfact, pm);
  if (hnew < 0.) atomicOr(flags, DFL_NEGH);
  const int o = abs(iparts[p]) - 1;
  out.xyzh[(size_t)o * out.ldx + 3] = hnew;
  th_out[p] = hnew;
  const float gh32 = (float)gradhi;
  out.gradh[(size_t)o * out.ldgradh] = gh32;
  if (dc.gravity && out.ldgradh > 1) out.gradh[(size_t)o * out.ldgradh + 1] = (float)gradsofti;
  const double rho1i = 1. / rhoi;
  const double term = dc.cnormk * (double)gh32 * rho1i * hf41;
  out.divcurlv[(size_t)o * out.lddcv] = (float)(-rs_divv * term);
  // exact-linear gradients (dens.F90:866-967, 997-1052, 1086-1101)
  const double rxx = rr[0], rxy = rr[1], rxz = rr[2], ryy = rr[3], ryz = rr[4], rzz = rr[5];
  const double denom = rxx * ryy * rzz + 2. * rxy * rxz * ryz - rxx * ryz * ryz - ryy * rxz * rxz - rzz * rxy * rxy;
  const double rm0 = ryy * rzz - ryz * ryz, rm1 = rxz * ryz - rzz * rxy, rm2 = rxy * ryz - rxz * ryy;
  const double rm3 = rzz * rxx - rxz * rxz, rm4 = rxy * rxz - rxx * ryz, rm5 = rxx * ryy - rxy * rxy;
  double gv[9], div_a;
  if (fabs(denom) > PH_TINY64) {
    const double dd = 1. / denom;
    for (int cpt = 0; cpt < 3; cpt++) {
      const double ax = dv[3 * cpt], ay = dv[3 * cpt + 1], az = dv[3 * cpt + 2];
      gv[3 * cpt] = -((ax * rm0 + ay * rm1 + az * rm2) * dd);
      gv[3 * cpt + 1] = -((ax * rm1 + ay * rm3 + az * rm4) * dd);
      gv[3 * cpt + 2] = -((ax * rm2 + ay * rm4 + az * rm5) * dd);
    }
    const double gax = (da[0] * rm0 + da[1] * rm1 + da[2] * rm2) * dd;
    const double gay = (da[3] * rm1 + da[4] * rm3 + da[5] * rm4) * dd;
    const double gaz = (da[6] * rm2 + da[7] * rm4 + da[8] * rm5) * dd;
    div_a = -(gax + gay + gaz);
  } else {
    for (int k = 0; k < 9; k++) gv[k] = -dv[k] * term;
    div_a = -term * (da[0] + da[4] + da[8]);
  }
  const double d5 = div_a - (gv[0] * gv[0] + gv[4] * gv[4] + gv[8] * gv[8] +
                             2. * (gv[1] * gv[3] + gv[2] * gv[6] + gv[5] * gv[7]));
  if (out.alphaind) out.alphaind[(size_t)o * 3 + 2] = (float)d5;
  if (out.dvdx)
    for (int k = 0; k < 9; k++) out.dvdx[(size_t)o * 9 + k] = (float)gv[k];
  if (MHD && out.divcurlB) {
    out.divcurlB[(size_t)o * 4 + 0] = (float)(-rs_divB * term);
    out.divcurlB[(size_t)o * 4 + 1] = (float)(-(dB[7] - dB[5]) * term);
    out.divcurlB[(size_t)o * 4 + 2] = (float)(-(dB[2] - dB[6]) * term);
    out.divcurlB[(size_t)o * 4 + 3] = (float)(-(dB[3] - dB[1]) * term);
  }
  p_nneigh[p] = nneigh;
}

// ------------------------------------------------------------------------
// density, part 2: every sum of get_density_sums (dens.F90:578-857) at the h
// of the last iteration, over the particle's list, and store_results
// (dens.F90:1511-1681) at the final h.  One thread per active particle.
// ------------------------------------------------------------------------
template <int K, bool MHD>
__global__ void __launch_bounds__(256)
k_dens_final(DevConst dc, const int32_t *__restrict__ act, int nact, const uint32_t *__restrict__ nlist,
             const int32_t *__restrict__ t_nl, const double *__restrict__ t_hsum, const double *__restrict__ t_hfin,
             const int32_t *__restrict__ retry, const int32_t *__restrict__ iparts, DensIn in,
             double *__restrict__ th_out, DensOut out, int32_t *__restrict__ p_nneigh, int32_t *__restrict__ flags) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= nact) return;
  const int p = act[t];
  if (p < 0 || retry[p]) return;
  const double pm = dc.pmass;
  const double4 ri = in.tp[p];
  const double xi = ri.x, yi = ri.y, zi = ri.z;
  const double hi = t_hsum[t], hfin = t_hfin[t];
  const int nl = t_nl[t];
  const bool have_f = in.tf != nullptr;
  const double4 vi = in.tv[p];
  double4 fi = make_double4(0., 0., 0., 0.);
  if (have_f) fi = in.tf[p];
  const double hi1 = 1. / hi, hi21 = xmul(hi1, hi1);
  double Bxi = 0., Byi = 0., Bzi = 0.;
  if (MHD) {
    const double4 Bev = in.tB[p];
    const double rhoiB = rhoh_d(hi, dc.hfact, pm);
    Bxi = Bev.x * rhoiB; Byi = Bev.y * rhoiB; Bzi = Bev.z * rhoiB;
  }
  double rs_rho = 0., rs_gh = 0., rs_gs = 0., rs_divv = 0., rs_divB = 0.;
  double dv[9], da[9], rr[6], dB[9];
  for (int k = 0; k < 9; k++) { dv[k] = 0.; da[k] = 0.; dB[k] = 0.; }
  for (int k = 0; k < 6; k++) rr[k] = 0.;
  int nneigh = 1;
  if constexpr (MHD) {
    for (int k = 0; k < nl; k++) {
      const uint32_t e = nlist[list_at(nact, t, k)];
      const int j = cand_node(e);
      const double4 r = in.tp[j];
      double dx, dy, dz;
      sep_of(dc, cand_code(e), xi, yi, zi, r.x, r.y, r.z, dx, dy, dz);
      const double r2 = r2_exact(dx, dy, dz);
      const double q2 = xmul(r2, hi21);
      if (!(q2 < dc.radkern2)) continue;
      const double rij = sqrt(r2);
      const double q = rij * hi1;
      double wk, gk;
      kern_wg<K, double>(q2, q, wk, gk);
      rs_rho += wk * pm;
      rs_gh += (-q * gk - 3. * wk) * pm;
      if (dc.gravity) rs_gs += kern_dphidh<K, double>(q2, q) * pm;
      nneigh++;
      const double rij1 = 1. / (rij + PH_EPS64);
      const double rg = rij1 * gk;
      const double runix = dx * rg * pm, runiy = dy * rg * pm, runiz = dz * rg * pm;
      const double4 vj = in.tv[j];
      const double dvx = vi.x - vj.x, dvy = vi.y - vj.y, dvz = vi.z - vj.z;
      rs_divv += dvx * runix + dvy * runiy + dvz * runiz;
      dv[0] += dvx * runix; dv[1] += dvx * runiy; dv[2] += dvx * runiz;
      dv[3] += dvy * runix; dv[4] += dvy * runiy; dv[5] += dvy * runiz;
      dv[6] += dvz * runix; dv[7] += dvz * runiy; dv[8] += dvz * runiz;
      if (have_f) {
        const double4 fj = in.tf[j];
        const double dax = fi.x - fj.x, day = fi.y - fj.y, daz = fi.z - fj.z;
        da[0] += dax * runix; da[1] += dax * runiy; da[2] += dax * runiz;
        da[3] += day * runix; da[4] += day * runiy; da[5] += day * runiz;
        da[6] += daz * runix; da[7] += daz * runiy; da[8] += daz * runiz;
      }
      rr[0] -= dx * runix; rr[1] -= dx * runiy; rr[2] -= dx * runiz;
      rr[3] -= dy * runiy; rr[4] -= dy * runiz; rr[5] -= dz * runiz;
      // B_j = Bevol_j * rho_j with rho_j from the h it had on entry
      const double4 B = in.tB[j];
      const double dBx = Bxi - B.x * B.w, dBy = Byi - B.y * B.w, dBz = Bzi - B.z * B.w;
      rs_divB += dBx * runix + dBy * runiy + dBz * runiz;
      dB[0] += dBx * runix; dB[1] += dBx * runiy; dB[2] += dBx * runiz;
      dB[3] += dBy * runix; dB[4] += dBy * runiy; dB[5] += dBy * runiz;
      dB[6] += dBz * runix; dB[7] += dBz * runiy; dB[8] += dBz * runiz;
    }
  } else {
    // the records of entry k+1 are loaded while entry k is summed (velocity
    // and force also for entries that turn out to lie outside the kernel)
    const double4 zero4 = make_double4(0., 0., 0., 0.);
    uint32_t e = 0u;
    double4 r = zero4, vj = zero4, fj = zero4;
    if (nl > 0) {
      e = nlist[list_at(nact, t, 0)];
      const int j = cand_node(e);
      r = in.tp[j];
      vj = in.tv[j];
      if (have_f) fj = in.tf[j];
    }
    for (int k = 0; k < nl; k++) {
      uint32_t en = 0u;
      double4 rn = zero4, vn = zero4, fn = zero4;
      if (k + 1 < nl) {
        en = nlist[list_at(nact, t, k + 1)];
        const int jn = cand_node(en);
        rn = in.tp[jn];
        vn = in.tv[jn];
        if (have_f) fn = in.tf[jn];
      }
      double dx, dy, dz;
      sep_of(dc, cand_code(e), xi, yi, zi, r.x, r.y, r.z, dx, dy, dz);
      const double r2 = r2_exact(dx, dy, dz);
      const double q2 = xmul(r2, hi21);
      if (q2 < dc.radkern2) {
        const double rij = sqrt(r2);
        const double q = rij * hi1;
        double wk, gk;
        kern_wg<K, double>(q2, q, wk, gk);
        rs_rho += wk * pm;
        rs_gh += (-q * gk - 3. * wk) * pm;
        if (dc.gravity) rs_gs += kern_dphidh<K, double>(q2, q) * pm;
        nneigh++;
        const double rij1 = 1. / (rij + PH_EPS64);
        const double rg = rij1 * gk;
        const double runix = dx * rg * pm, runiy = dy * rg * pm, runiz = dz * rg * pm;
        const double dvx = vi.x - vj.x, dvy = vi.y - vj.y, dvz = vi.z - vj.z;
        rs_divv += dvx * runix + dvy * runiy + dvz * runiz;
        dv[0] += dvx * runix; dv[1] += dvx * runiy; dv[2] += dvx * runiz;
        dv[3] += dvy * runix; dv[4] += dvy * runiy; dv[5] += dvy * runiz;
        dv[6] += dvz * runix; dv[7] += dvz * runiy; dv[8] += dvz * runiz;
        if (have_f) {
          const double dax = fi.x - fj.x, day = fi.y - fj.y, daz = fi.z - fj.z;
          da[0] += dax * runix; da[1] += dax * runiy; da[2] += dax * runiz;
          da[3] += day * runix; da[4] += day * runiy; da[5] += day * runiz;
          da[6] += daz * runix; da[7] += daz * runiy; da[8] += daz * runiz;
        }
        rr[0] -= dx * runix; rr[1] -= dx * runiy; rr[2] -= dx * runiz;
        rr[3] -= dy * runiy; rr[4] -= dy * runiz; rr[5] -= dz * runiz;
      }
      e = en; r = rn; vj = vn; fj = fn;
    }
  }
  dens_store<MHD>(dc, p, hfin, rs_rho, rs_gh, rs_gs, rs_divv, rs_divB, dv, da, rr, dB, iparts, th_out, out,
                  p_nneigh, nneigh, flags);
}

// Mixed precision (no MHD): the neighbour test q^2 < R^2 on the FP64
// separation as in k_dens_final (the same neighbours).  The sums that fix
// h and rho (kernel, dOmega, softening) stay FP64, with r from an FP32
// reciprocal square root refined in FP64; the velocity and acceleration
// differences are taken in FP64 (they are small against the values in
// smooth flow) and the gradient-tensor products in FP32, summed in FP64.
template <int K>
__global__ void __launch_bounds__(256)
k_dens_final_mixed(DevConst dc, const int32_t *__restrict__ act, int nact, const uint32_t *__restrict__ nlist,
                   const int32_t *__restrict__ t_nl, const double *__restrict__ t_hsum,
                   const double *__restrict__ t_hfin, const int32_t *__restrict__ retry,
                   const int32_t *__restrict__ iparts, const double4 *__restrict__ tp,
                   const double4 *__restrict__ tv, const double4 *__restrict__ tf, double *__restrict__ th_out,
                   DensOut out, int32_t *__restrict__ p_nneigh, int32_t *__restrict__ flags) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= nact) return;
  const int p = act[t];
  if (p < 0 || retry[p]) return;
  const double pm = dc.pmass;
  const double4 ri = tp[p];
  const double xi = ri.x, yi = ri.y, zi = ri.z;
  const double hi = t_hsum[t], hfin = t_hfin[t];
  const int nl = t_nl[t];
  const bool have_f = tf != nullptr;
  const double4 vi = tv[p];
  double4 fi = make_double4(0., 0., 0., 0.);
  if (have_f) fi = tf[p];
  const double hi1 = 1. / hi, hi21 = xmul(hi1, hi1);
  const float pmf = (float)pm;
  double rs_rho = 0., rs_gh = 0., rs_gs = 0., rs_divv = 0.;
  double dv[9], da[9], rr[6];
  for (int k = 0; k < 9; k++) { dv[k] = 0.; da[k] = 0.; }
  for (int k = 0; k < 6; k++) rr[k] = 0.;
  int nneigh = 1;
  for (int k = 0; k < nl; k++) {
    const uint32_t e = nlist[list_at(nact, t, k)];
    const int j = cand_node(e);
    const double4 r = tp[j];
    double dx, dy, dz;
    sep_of(dc, cand_code(e), xi, yi, zi, r.x, r.y, r.z, dx, dy, dz);
    const double r2 = r2_exact(dx, dy, dz);
    const double q2 = xmul(r2, hi21);
    if (!(q2 < dc.radkern2)) continue;
    // r from an FP32 reciprocal square root refined by one FP64 Newton step
    double rij1 = 0.;
    const float r2f = (float)r2;
    if (r2f > 0.f) {
      const double y = (double)rsqrtf(r2f);
      rij1 = y * (1.5 - 0.5 * r2 * y * y);
    }
    const double rij = r2 * rij1;
    const double q = rij * hi1;
    double wk, gk;
    kern_wg<K, double>(q2, q, wk, gk);
    rs_rho += wk * pm;
    rs_gh += (-q * gk - 3. * wk) * pm;
    if (dc.gravity) rs_gs += kern_dphidh<K, double>(q2, q) * pm;
    nneigh++;
    const float rg = (float)(gk * rij1);
    const float runix = (float)dx * rg * pmf, runiy = (float)dy * rg * pmf, runiz = (float)dz * rg * pmf;
    const double4 vj = tv[j];
    const float dvx = (float)(vi.x - vj.x), dvy = (float)(vi.y - vj.y), dvz = (float)(vi.z - vj.z);
    rs_divv += (double)(dvx * runix + dvy * runiy + dvz * runiz);
    dv[0] += (double)(dvx * runix); dv[1] += (double)(dvx * runiy); dv[2] += (double)(dvx * runiz);
    dv[3] += (double)(dvy * runix); dv[4] += (double)(dvy * runiy); dv[5] += (double)(dvy * runiz);
    dv[6] += (double)(dvz * runix); dv[7] += (double)(dvz * runiy); dv[8] += (double)(dvz * runiz);
    if (have_f) {
      const double4 fj = tf[j];
      const float dax = (float)(fi.x - fj.x), day = (float)(fi.y - fj.y), daz = (float)(fi.z - fj.z);
      da[0] += (double)(dax * runix); da[1] += (double)(dax * runiy); da[2] += (double)(dax * runiz);
      da[3] += (double)(day * runix); da[4] += (double)(day * runiy); da[5] += (double)(day * runiz);
      da[6] += (double)(daz * runix); da[7] += (double)(daz * runiy); da[8] += (double)(daz * runiz);
    }
    const float dxf = (float)dx, dyf = (float)dy, dzf = (float)dz;
    rr[0] -= (double)(dxf * runix); rr[1] -= (double)(dxf * runiy); rr[2] -= (double)(dxf * runiz);
    rr[3] -= (double)(dyf * runiy); rr[4] -= (double)(dyf * runiz); rr[5] -= (double)(dzf * runiz);
  }
  const double dB[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
  dens_store<false>(dc, p, hfin, rs_rho, rs_gh, rs_gs, rs_divv, 0., dv, da, rr, dB, iparts, th_out, out, p_nneigh,
                    nneigh, flags);
}

// actual-neighbour statistics and density evaluations (dens.F90:1670-1679)
__global__ void __launch_bounds__(256)
k_dens_stats(int nact, const int32_t *act, const int32_t *pleaf, const int32_t *p_nneigh,
             const int32_t *leaf_nits, const int32_t *ntrial, long long *acc) {
  long long nn = 0, nit = 0, ntr = 0, mx = 0, mxt = 0;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nact; t += gridDim.x * blockDim.x) {
    const int p = act[t];
    const int l = pleaf[p];
    if (p_nneigh) {
      const long long v = p_nneigh[p];
      nn += v;
      nit += leaf_nits[l];
      mx = mx > v ? mx : v;
    }
    if (ntrial) {
      const long long v = ntrial[l];
      ntr += v;
      mxt = mxt > v ? mxt : v;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    nn += __shfl_xor(nn, o, 64);
    nit += __shfl_xor(nit, o, 64);
    ntr += __shfl_xor(ntr, o, 64);
    long long a = __shfl_xor(mx, o, 64), b = __shfl_xor(mxt, o, 64);
    mx = mx > a ? mx : a;
    mxt = mxt > b ? mxt : b;
  }
  __shared__ long long sh[4][5];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[w][0] = nn; sh[w][1] = nit; sh[w][2] = ntr; sh[w][3] = mx; sh[w][4] = mxt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; k++) {
      nn += sh[k][0]; nit += sh[k][1]; ntr += sh[k][2];
      mx = mx > sh[k][3] ? mx : sh[k][3];
      mxt = mxt > sh[k][4] ? mxt : sh[k][4];
    }
    atomicAdd((unsigned long long *)&acc[0], (unsigned long long)nn);
    atomicAdd((unsigned long long *)&acc[1], (unsigned long long)nit);
    atomicAdd((unsigned long long *)&acc[2], (unsigned long long)ntr);
    atomicMax(&acc[3], mx);
    atomicMax(&acc[4], mxt);
  }
}

static inline unsigned stats_grid(int n) { return std::min(nblocks(n, 256), 512u); }

__global__ void k_zero_i32(int32_t *a, int n, int32_t v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_scatter_pstats(int nact, const int32_t *act, const int32_t *iparts, const int32_t *p_nneigh,
                                 const int32_t *p_nits, int32_t *nneigh, int32_t *nits) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nact) return;
  const int p = act[t];
  int o = abs(iparts[p]) - 1;
  if (nneigh) nneigh[o] = p_nneigh[p];
  if (nits) nits[o] = p_nits[p];
}

int dens_particle_stats(phsph_ctx *c, int32_t *d_nneigh, int32_t *d_nits) {
  if (c->nact_dens > 0)
    hipLaunchKernelGGL(k_scatter_pstats, dim3(nblocks(c->nact_dens, 256)), dim3(256), 0, c->stream, c->nact_dens,
                       c->act_dens.p, c->iparts.p, c->p_nneigh.p, c->p_nits.p, d_nneigh, d_nits);
  PH_HIPCHECK(c, hipGetLastError());
  return PHSPH_OK;
}

// ------------------------------------------------------------------------
// k_dens_iter configurations and the autotuner
// ------------------------------------------------------------------------
struct IterLaunch {
  phsph_ctx *c;
  hipStream_t s;
  int ng, icall;
  double rfac;
  int na;
  double lmargin;
  int dlist;
  bool widen;
  uint32_t *lb;
  double *r2b;
  int32_t *tnl;
  double *ths, *thf;
};

template <int K, int CL, int LW, bool R2, bool Q>
static void launch_iter_t(const IterLaunch &L) {
  phsph_ctx *c = L.c;
  unsigned grid = list_grid<LW>(L.ng);
  if (Q) {
    // persistent: as many blocks as are resident at once (never more than
    // one group per wavefront), their queue counters cleared first
    int bpc = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, reinterpret_cast<const void *>(
                                                                &k_dens_iter<K, CL, LW, R2, true>),
                                                     64 * LW, 0) != hipSuccess || bpc < 1)
      bpc = 1;
    const unsigned res = xcd_grid(bpc * (c->ncu > 0 ? c->ncu : 256));
    if (res < grid) grid = res;
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, L.s, c->wq.p, 8, 0);
  }
  hipLaunchKernelGGL((k_dens_iter<K, CL, LW, R2, Q>), dim3(grid), dim3(64 * LW), 0, L.s, c->dc, L.icall,
                     L.rfac, c->act.p, L.na, L.ng, c->pleaf.p, c->leaves.p, c->gcand.p, c->gncand.p, c->goff_p,
                     c->caps.gcand, c->nd_lo.p, c->nd_hi.p, c->nd.p, c->tp4.p, L.widen ? c->hmax0.p : nullptr,
                     L.lmargin, L.lb, L.r2b, L.dlist, L.tnl, L.widen ? c->t_nlj.p : nullptr, L.ths, L.thf,
                     c->p_nits.p, c->leaf_nits.p, c->tmpi.p, c->htraj.p, (size_t)c->htraj_stride, c->hconv.p,
                     c->counters.p, c->wq.p);
}

template <int K, int CL, int LW>
static void launch_iter_r(const ListTune &t, const IterLaunch &L) {
  if (t.r2) {
    if (t.q) launch_iter_t<K, CL, LW, true, true>(L); else launch_iter_t<K, CL, LW, true, false>(L);
  } else {
    if (t.q) launch_iter_t<K, CL, LW, false, true>(L); else launch_iter_t<K, CL, LW, false, false>(L);
  }
}

template <int K>
static void launch_iter_k(const ListTune &t, const IterLaunch &L) {
  if (t.cl == 8) {
    if (t.lw == 2) launch_iter_r<K, 8, 2>(t, L); else launch_iter_r<K, 8, 4>(t, L);
  } else {
    if (t.lw == 2) launch_iter_r<K, 16, 2>(t, L); else launch_iter_r<K, 16, 4>(t, L);
  }
}

void launch_dens_iter(const ListTune &t, const IterLaunch &L) {
  if (L.c->p.kernel == PHSPH_KERNEL_CUBIC) launch_iter_k<PHSPH_KERNEL_CUBIC>(t, L);
  else launch_iter_k<PHSPH_KERNEL_QUINTIC>(t, L);
}

// The autotuner: on the first density call of a context with at least
// TUNE_MIN_ACTIVE active particles, k_dens_iter runs once in every
// configuration (candidate leaves staged per chunk, wavefronts per block,
// squared separations stored with the lists or recomputed from the
// positions), timed with HIP events on the library stream, and the fastest
// is kept for the context.  Every configuration visits the same leaves in
// the same order and writes the same lists and sums, so the outputs of the
// last run are those of any other.
int tune_density_iter(phsph_ctx *c, const IterLaunch &L) {
  // the separations buffer is needed by the configurations that store them
  PH_HIPCHECK(c, c->nr2.ensure((size_t)list_cap(L.dlist) * L.na + 2));
  IterLaunch L2 = L;
  L2.r2b = c->nr2.p;
  hipEvent_t e0, e1;
  PH_HIPCHECK(c, hipEventCreate(&e0));
  PH_HIPCHECK(c, hipEventCreate(&e1));
  float best = 1e30f;
  int nt = 0;
  ListTune bt = c->tune;
  for (int cl : {8, 16})
    for (int lw : {2, 4})
      for (int r2 : {0, 1})
        for (int q : {0, 1}) {
        const ListTune t{cl, lw, r2, q};
        PH_HIPCHECK(c, hipEventRecord(e0, L.s));
        launch_dens_iter(t, L2);
        PH_HIPCHECK(c, hipEventRecord(e1, L.s));
        PH_HIPCHECK(c, hipEventSynchronize(e1));
        float ms = 0.f;
        PH_HIPCHECK(c, hipEventElapsedTime(&ms, e0, e1));
        if (nt < TUNE_MAX_CONFIGS) {
          c->tune_tried[nt] = t;
          c->tune_ms[nt] = ms;
          nt++;
        }
        if (ms < best) {
          best = ms;
          bt = t;
        }
      }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  c->tune_ntried = nt;
  c->tune = bt;
  c->tune_mode = TUNE_DONE;
  c->tune_nact = L.na;
  return PHSPH_OK;
}

// trial neighbours of the reference's per-leaf walk at radkern*hmax, over
// the active particles of the last densityiterate (computed on request).
// The hmax is that of the leaf's last neighbour list: the tree's, or the
// one the relists of the iteration raised it to (dens.F90:338-373, the
// count of the last get_neighbour_list is the cell's nneightry,
// dens.F90:1267).
int dens_trial_stats(phsph_ctx *c) {
  if (!c->trial_stale) return PHSPH_OK;
  hipStream_t s = c->stream;
  PH_HIPCHECK(c, c->ntrial.ensure(c->nleaves + 1));
  PH_HIPCHECK(c, c->trial_hmax.ensure(c->nleaves + 1));
  int32_t zero = 0;
  PH_HIPCHECK(c, dev_set(c, c->counters.p, &zero, sizeof(int32_t)));
  const bool replay = c->dens_icall != 0 && c->htraj.p != nullptr;
  if (c->nleaves > 0 && replay)
    hipLaunchKernelGGL(k_trial_hmax, dim3(nblocks(c->nleaves, 128)), dim3(128), 0, s, c->nleaves, c->leaves.p,
                       c->leaf_flag.p, c->nd_lo.p, c->nd_hi.p, c->nd.p, c->p_nits.p, c->htraj.p, (size_t)c->htraj_stride, c->hconv.p,
                       c->trial_hmax.p, c->counters.p);
  if (c->nleaves > 0)
    hipLaunchKernelGGL(k_walk, dim3(nblocks(c->nleaves, 128)), dim3(128), 0, s, c->dc, 1.0, c->nd.p,
                       c->leaf_flag.p, c->nd_lo.p, c->nd_hi.p, c->leaves.p, c->nleaves,
                       replay ? c->trial_hmax.p : nullptr, c->ntrial.p, c->counters.p);
  PH_HIPCHECK(c, hipGetLastError());
  PH_HIPCHECK(c, c->stats_i.ensure(8));
  long long z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  PH_HIPCHECK(c, dev_set(c, c->stats_i.p, z, sizeof(z)));
  if (c->nact_dens > 0)
    hipLaunchKernelGGL(k_dens_stats, dim3(stats_grid(c->nact_dens)), dim3(256), 0, s, c->nact_dens,
                       c->act_dens.p, c->pleaf.p, nullptr, nullptr, c->ntrial.p, c->stats_i.p);
  long long acc[8];
  int32_t fl = 0;
  PH_HIPCHECK(c, hipMemcpyAsync(acc, c->stats_i.p, sizeof(acc), hipMemcpyDeviceToHost, s));
  PH_HIPCHECK(c, hipMemcpyAsync(&fl, c->counters.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  PH_HIPCHECK(c, hipStreamSynchronize(s));
  if (fl & DFL_CAND) PH_RET(c, PHSPH_ERR_CAPACITY, "neighbour statistics: walk stack overflow");
  if (fl & DFL_NOCONV)
    PH_RET(c, PHSPH_ERR_CAPACITY, "neighbour statistics: a cell iterated more often than the h record holds");
  if (c->nact_dens > 0) {
    c->trialmean = (double)acc[2] / (double)c->nact_dens;
    c->maxtrial = (int)acc[4];
  }
  c->trial_stale = false;
  return PHSPH_OK;
}

// ------------------------------------------------------------------------
int walk_groups(phsph_ctx *c, int mode, double rfac, const double *d_hsrc, const double *d_hmaxn) {
  hipStream_t s = c->stream;
  const int ng = (c->nact + GROUP - 1) / GROUP;
  c->ngroups = ng;
  c->goff_p = nullptr;
  while (c->caps.gcand > 16 && !list_fits(c, (size_t)(ng + 1) * c->caps.gcand)) c->caps.gcand /= 2;
  // a uniform capacity per group, doubled (and kept) while it fits in the
  // list budget; beyond that every group is walked again into a list sized
  // from the counts of the previous walk
  bool exact = false;
  for (int attempt = 0; attempt < 12; attempt++) {
    const int64_t *goff = exact ? c->goff.p : nullptr;
    if (!exact) PH_HIPCHECK(c, c->gcand.ensure((size_t)(ng + 1) * c->caps.gcand));
    PH_HIPCHECK(c, c->gncand.ensure(ng + 1));
    int32_t zero = 0;
    PH_HIPCHECK(c, dev_set(c, c->counters.p, &zero, sizeof(int32_t)));
    if (ng > 0) {
      LaunchTimer lt(c, mode == 0 ? "walk_density" : "walk_force");
      hipLaunchKernelGGL(k_group_walk, dim3(xcd_grid(ng)), dim3(64), 0, s, c->dc, mode, rfac, c->nd.p, d_hmaxn,
                         c->leaf_flag.p, c->act.p, c->nact, c->tx.p, c->ty.p, c->tz.p, d_hsrc, ng, c->gcand.p,
                         c->gncand.p, goff, c->caps.gcand, c->counters.p);
    }
    PH_HIPCHECK(c, hipGetLastError());
    int32_t fl = 0;
    PH_HIPCHECK(c, hipMemcpyAsync(&fl, c->counters.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    if (getenv("PHSPH_DEBUG")) {
      size_t fr = 0, tot = 0;
      (void)hipMemGetInfo(&fr, &tot);
      fprintf(stderr, "phsph walk: mode %d groups %d cap %d exact %d flags %d free %.1f GiB\n", mode, ng,
              c->caps.gcand, (int)exact, fl, fr / 1073741824.0);
    }
    if (fl & DFL_STACK) PH_RET(c, PHSPH_ERR_CAPACITY, "neighbour walk: node stack exceeded");
    if (!(fl & DFL_CAND)) {
      c->goff_p = goff;
      return PHSPH_OK;
    }
    if (exact) break;
    if (list_fits(c, (size_t)(ng + 1) * c->caps.gcand * 2)) {
      c->caps.gcand *= 2;
      continue;
    }
    // exact offsets from the counts (gncand holds every group's full count)
    exact = true;
    PH_HIPCHECK(c, c->goff.ensure((size_t)ng + 1));
    hipLaunchKernelGGL(k_scan_offsets<int32_t>, dim3(1), dim3(1024), 0, s, c->gncand.p, ng, (int64_t)1, 1,
                       c->goff.p);
    PH_HIPCHECK(c, hipGetLastError());
    int64_t total = 0;
    PH_HIPCHECK(c, hipMemcpyAsync(&total, c->goff.p + ng, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    PH_HIPCHECK(c, c->gcand.ensure((size_t)total + 1));
  }
  PH_RET(c, PHSPH_ERR_CAPACITY, "neighbour walk: candidate list capacity exceeded");
}

int compact_active(phsph_ctx *c, const int32_t *flag, int n) {
  hipStream_t s = c->stream;
  const unsigned nb = nblocks(n, 256);
  PH_HIPCHECK(c, c->tmpi2.ensure((size_t)nb + 1));
  PH_HIPCHECK(c, c->act.ensure(n + 1));
  hipLaunchKernelGGL(k_compact_count, dim3(nb), dim3(256), 0, s, flag, n, c->tmpi2.p);
  hipLaunchKernelGGL(k_scan_i32, dim3(1), dim3(1024), 0, s, c->tmpi2.p, (int)nb, c->counters.p + 1);
  hipLaunchKernelGGL(k_compact_fill, dim3(nb), dim3(256), 0, s, flag, n, c->tmpi2.p, c->act.p);
  int32_t na = 0;
  PH_HIPCHECK(c, hipMemcpyAsync(&na, c->counters.p + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  PH_HIPCHECK(c, hipStreamSynchronize(s));
  c->nact = na;
  return PHSPH_OK;
}

// Groups of the active list.  64 consecutive active particles form a group,
// except that a group closes early (the rest of its 64 slots are padding,
// act = -1) where the next active particle lies far from the previous one:
// more than PAD_GAP tree positions after it (sparse active sets: block
// timesteps, retried particles), or in a leaf whose centre is more than
// GROUP_JUMP times the two leaves' sizes away.  Tree order is depth-first,
// so consecutive leaves are neighbours except where the order crosses from
// one subtree to the next; a group across such a jump has a bounding box
// spanning both sides, and its candidate leaves cover the space between.
// The padded length becomes c->nact.  PAD_GAP is wide (about 100 leaves):
// with a random 1/64 of the particles active, a gap of 64 positions closed
// most groups after two or three particles, and the padded slots cost the
// pair kernels as much as live ones (128^3 MHD + dust substep at 1/64
// active: 12.7 ms with 64, 9.3 ms with 1024; the spatial test above still
// closes groups across jumps).
constexpr int PAD_GAP = 1024;
constexpr double GROUP_JUMP = 4.;

__global__ void k_fill_act(int n, int32_t *a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = -1;
}

__global__ void k_grp_breaks(int na, const int32_t *__restrict__ act, const int32_t *__restrict__ pleaf,
                             const int32_t *__restrict__ leaves, const phsph_kdnode *__restrict__ nd,
                             int32_t *__restrict__ brk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= na) return;
  int b = 1;
  if (i > 0) {
    const int p = act[i], q = act[i - 1];
    b = (p - q > PAD_GAP) ? 1 : 0;
    if (!b) {
      const int lp = pleaf[p], lq = pleaf[q];
      if (lp != lq) {
        const phsph_kdnode &A = nd[leaves[lp]];
        const phsph_kdnode &B = nd[leaves[lq]];
        const double dx = A.xcen[0] - B.xcen[0], dy = A.xcen[1] - B.xcen[1], dz = A.xcen[2] - B.xcen[2];
        const double sl = GROUP_JUMP * (A.size + B.size);
        b = (dx * dx + dy * dy + dz * dz > sl * sl) ? 1 : 0;
      }
    }
  }
  brk[i] = b;
}

// groups of each run (block sums): a run is rs[r] .. rs[r+1]-1
__global__ void k_grp_bsum(int na, const int32_t *__restrict__ rs, const int32_t *__restrict__ nruns_p,
                           int32_t *__restrict__ bsum) {
  __shared__ int shi[4];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int nruns = *nruns_p;
  int g = 0;
  if (r < nruns) {
    const int len = ((r + 1 < nruns) ? rs[r + 1] : na) - rs[r];
    g = (len + GROUP - 1) / GROUP;
  }
  const int v = wave_sum_i(g);
  if ((threadIdx.x & 63) == 0) shi[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) bsum[blockIdx.x] = shi[0] + shi[1] + shi[2] + shi[3];
}
// first group of each run
__global__ void k_grp_place(int na, const int32_t *__restrict__ rs, const int32_t *__restrict__ nruns_p,
                            const int32_t *__restrict__ boff, int32_t *__restrict__ gr) {
  __shared__ int shi[4];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int nruns = *nruns_p;
  int g = 0;
  if (r < nruns) {
    const int len = ((r + 1 < nruns) ? rs[r + 1] : na) - rs[r];
    g = (len + GROUP - 1) / GROUP;
  }
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  int incl = g;
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(incl, o, 64);
    if (l >= o) incl += u;
  }
  if (l == 63) shi[w] = incl;
  __syncthreads();
  int off = boff[blockIdx.x];
  for (int k = 0; k < w; k++) off += shi[k];
  if (r < nruns) gr[r] = off + incl - g;
}
// each active particle to its slot: its run by bisection over the run starts
__global__ void k_grp_scatter(int na, const int32_t *__restrict__ act, const int32_t *__restrict__ rs,
                              const int32_t *__restrict__ nruns_p, const int32_t *__restrict__ gr,
                              int32_t *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= na) return;
  int lo = 0, hi = *nruns_p - 1;  // largest r with rs[r] <= i (rs[0] = 0)
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (rs[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  const int o = i - rs[lo];
  out[(size_t)GROUP * (gr[lo] + o / GROUP) + (o % GROUP)] = act[i];
}

int pad_groups(phsph_ctx *c) {
  const int na = c->nact;
  if (na == 0) return PHSPH_OK;
  hipStream_t s = c->stream;
  const unsigned nb = nblocks(na, 256);
  PH_HIPCHECK(c, c->grp_brk.ensure(na + 1));
  PH_HIPCHECK(c, c->grp_rs.ensure(na + 1));
  PH_HIPCHECK(c, c->grp_gr.ensure(na + 1));
  PH_HIPCHECK(c, c->grp_bc.ensure((size_t)nb + 1));
  int32_t *nruns = c->counters.p + 6, *ngrp = c->counters.p + 7;
  hipLaunchKernelGGL(k_grp_breaks, dim3(nb), dim3(256), 0, s, na, c->act.p, c->pleaf.p, c->leaves.p, c->nd.p,
                     c->grp_brk.p);
  hipLaunchKernelGGL(k_compact_count, dim3(nb), dim3(256), 0, s, c->grp_brk.p, na, c->grp_bc.p);
  hipLaunchKernelGGL(k_scan_i32, dim3(1), dim3(1024), 0, s, c->grp_bc.p, (int)nb, nruns);
  hipLaunchKernelGGL(k_compact_fill, dim3(nb), dim3(256), 0, s, c->grp_brk.p, na, c->grp_bc.p, c->grp_rs.p);
  PH_HIPCHECK(c, hipGetLastError());
  int32_t nr = 0;
  PH_HIPCHECK(c, hipMemcpyAsync(&nr, nruns, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  PH_HIPCHECK(c, hipStreamSynchronize(s));
  if (nr <= 1) return PHSPH_OK;  // one run: 64 consecutive active particles per group
  hipLaunchKernelGGL(k_grp_bsum, dim3(nb), dim3(256), 0, s, na, c->grp_rs.p, nruns, c->grp_bc.p);
  hipLaunchKernelGGL(k_scan_i32, dim3(1), dim3(1024), 0, s, c->grp_bc.p, (int)nb, ngrp);
  hipLaunchKernelGGL(k_grp_place, dim3(nb), dim3(256), 0, s, na, c->grp_rs.p, nruns, c->grp_bc.p, c->grp_gr.p);
  PH_HIPCHECK(c, hipGetLastError());
  int32_t ng = 0;
  PH_HIPCHECK(c, hipMemcpyAsync(&ng, ngrp, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  PH_HIPCHECK(c, hipStreamSynchronize(s));
  const int64_t slots = (int64_t)GROUP * ng;
  if (slots == na) return PHSPH_OK;  // no group closed early
  if (slots > INT32_MAX - 1) PH_RET(c, PHSPH_ERR_CAPACITY, "active list: too many padded groups");
  PH_HIPCHECK(c, c->act2.ensure((size_t)slots + 1));
  hipLaunchKernelGGL(k_fill_act, dim3(nblocks((int)slots, 256)), dim3(256), 0, s, (int)slots, c->act2.p);
  hipLaunchKernelGGL(k_grp_scatter, dim3(nb), dim3(256), 0, s, na, c->act.p, c->grp_rs.p, nruns, c->grp_gr.p,
                     c->act2.p);
  PH_HIPCHECK(c, hipGetLastError());
  std::swap(c->act.p, c->act2.p);
  std::swap(c->act.cap, c->act2.cap);
  c->nact = (int)slots;
  return PHSPH_OK;
}

// The first attempt's lists hold, for a particle i, every j within
// hf*radkern*max(h0_i, h0_j).  They contain the force neighbours (q_i < R or
// q_j < R at the final h) only if no final h exceeds hf*h0: flag otherwise
// (a particle whose h grew past its first search).
__global__ void k_check_hbound(int nact, const int32_t *__restrict__ act, const double *__restrict__ h0,
                               const double *__restrict__ hfin, double hf, int32_t *__restrict__ flag) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  if (t < nact) {
    const int p = act[t];
    bad = !(hfin[p] * (1. + 1e-12) <= hf * h0[p]);
  }
  if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}
int dens_run(phsph_ctx *c, int icall, const DensIOArgs &a) {
  hipStream_t s = c->stream;
  const int n = c->nlive;
  const bool mhd = c->p.mhd != 0;
  const bool have_f = (a.fxyzu != nullptr) || (a.fext != nullptr);
  PH_HIPCHECK(c, c->th_in.ensure(n));
  PH_HIPCHECK(c, c->tp4.ensure(n));
  PH_HIPCHECK(c, c->tv4.ensure(n));
  if (have_f) PH_HIPCHECK(c, c->tf4.ensure(n));
  if (mhd) PH_HIPCHECK(c, c->tB4.ensure(n));
  PH_HIPCHECK(c, c->tmpi.ensure(n));
  PH_HIPCHECK(c, c->wq.ensure(16));
  // mixed precision covers the hydro sums; with MHD the pass stays FP64
  const bool mixed = c->p.precision == PHSPH_PREC_MIXED && !mhd;
  c->dens_prec_used = mixed ? PHSPH_PREC_MIXED : PHSPH_PREC_FP64;
  {
    LaunchTimer lt(c, "dens_gather");
    hipLaunchKernelGGL(k_gather_dens, dim3(nblocks(n, 256)), dim3(256), 0, s, c->dc, n, c->iparts.p, c->tx.p,
                       c->ty.p, c->tz.p, a.xyzh, a.ldx, a.vxyzu, a.ldv, a.fxyzu, a.ldf, a.fext,
                       mhd ? a.Bevol : nullptr, a.iphase, c->th_in.p, c->tp4.p, c->tv4.p,
                       have_f ? c->tf4.p : nullptr, c->tB4.p, c->tmpi.p);
  }
  PH_HIPCHECK(c, hipGetLastError());
  int rc = compact_active(c, c->tmpi.p, n);
  if (rc) return rc;
  const int nact = c->nact;
  PH_HIPCHECK(c, c->act_dens.ensure(nact + 1));
  if (nact > 0)
    PH_HIPCHECK(c, hipMemcpyAsync(c->act_dens.p, c->act.p, sizeof(int32_t) * nact, hipMemcpyDeviceToDevice, s));
  c->nact_dens = nact;
  c->trial_stale = true;
  PH_HIPCHECK(c, c->p_nneigh.ensure(n + 1));
  PH_HIPCHECK(c, c->p_nits.ensure(n + 1));
  PH_HIPCHECK(c, c->leaf_nits.ensure(c->nleaves + 1));
  PH_HIPCHECK(c, c->tmpi.ensure(n + 1));
  DensIn in{c->tp4.p, c->tv4.p, have_f ? c->tf4.p : nullptr, c->tB4.p};
  DensOut out{a.xyzh, a.ldx, a.divcurlv, a.lddcv, a.divcurlB, a.alphaind, a.gradh, a.ldgradh, a.dvdx};
  const bool grav = c->p.gravity != 0;
  // h after every Newton update: the relists of the reference are replayed
  // from it (node hmax for the gravity walk, trial-neighbour statistics)
  PH_HIPCHECK(c, c->htraj.ensure((size_t)n * HTRAJ + 1));
  c->htraj_stride = n;
  PH_HIPCHECK(c, c->hconv.ensure(n + 1));
  c->dens_icall = icall;
  hipLaunchKernelGGL(k_zero_i32, dim3(nblocks(c->nleaves, 256)), dim3(256), 0, s, c->leaf_nits.p, c->nleaves, 0);
  hipLaunchKernelGGL(k_zero_i32, dim3(nblocks(n, 256)), dim3(256), 0, s, c->p_nits.p, n, -1);
  hipLaunchKernelGGL(k_zero_i32, dim3(nblocks(n, 256)), dim3(256), 0, s, c->p_nneigh.p, n, -1);
  const bool dbg = getenv("PHSPH_DEBUG") != nullptr;
  // Particles whose h outgrows their neighbour list, or whose list does not
  // fit, are handed back and run again (alone) with a wider search or a
  // longer list.  The first search radius (units of radkern*h at entry) is
  // adapted from call to call (end of this function): the narrowest level
  // whose lists rarely need a wider search.  Neighbour sets and iteration
  // counts do not depend on it; the sums agree to round-off (the group walk
  // lists the candidate leaves in an order that depends on the radius).
  const double rfac0 = RFAC_LEVELS[c->rfac_level];
  double rfac = rfac0;
  bool rewalk0 = false;
  // Lists the force pass can use as its neighbour lists (no self-gravity):
  // they also cover every j whose radius reaches i, with a margin for the
  // converged h (within tolh of the h of the last summation, which is at
  // most rfac*h0 without a retry).  The lists of the first attempt are kept;
  // particles run again there are listed by dretry0 and get force lists of
  // their own.
  const bool reusable = !grav && c->p.tolh <= 1e-3;
  const double lmargin = reusable ? 1. + 10. * c->p.tolh : 1.;
  c->dlist_reuse = false;
  if (reusable) {
    PH_HIPCHECK(c, c->hmax0.ensure(c->ncells + 2));
    rc = leaf_hmax_run(c, c->th_in.p, c->hmax0.p);
    if (rc) return rc;
  }
  rc = pad_groups(c);
  if (rc) return rc;
  // list capacity: it grows for the particles run again (the list memory is
  // that of the retried particles); a grown capacity is kept for later calls
  // while the lists of every active particle fit in the list budget
  // 12 bytes per entry: the entry and its squared separation
  while (c->caps.dlist > 16 && !list_fits(c, (size_t)list_cap(c->caps.dlist) * nact * 3)) c->caps.dlist /= 2;
  int dlist = c->caps.dlist;
  const int dlist0 = dlist;
  bool tuned_now = false, retune = false;
  for (int attempt = 0; attempt < 12; attempt++) {
    const int na = c->nact;  // active slots, padding included
    if (na == 0) break;
    // the first attempt writes the lists the force pass may use; retries
    // write buffers of their own so that those stay intact
    DBuf<uint32_t> &LB = attempt == 0 ? c->nlist : c->nlist_r;
    DBuf<int32_t> &TNL = attempt == 0 ? c->t_nl : c->t_nl_r;
    DBuf<double> &THS = attempt == 0 ? c->t_hsum : c->t_hsum_r;
    DBuf<double> &THF = attempt == 0 ? c->t_hfin : c->t_hfin_r;
    PH_HIPCHECK(c, TNL.ensure((size_t)na + 1));
    if (attempt == 0 && reusable) PH_HIPCHECK(c, c->t_nlj.ensure((size_t)na + 1));
    PH_HIPCHECK(c, THS.ensure((size_t)na + 1));
    PH_HIPCHECK(c, THF.ensure((size_t)na + 1));
    // only the first attempt's lists serve the force pass
    const bool widen = attempt == 0 && reusable;
    rc = walk_groups(c, 0, rfac * lmargin, c->th_in.p, widen ? c->hmax0.p : nullptr);
    if (rc) return rc;
    const int ng = c->ngroups;
    PH_HIPCHECK(c, LB.ensure((size_t)list_cap(dlist) * na + 1));
    DBuf<double> &R2B = attempt == 0 ? c->nr2 : c->nr2_r;
    if (c->tune.r2 || c->tune_mode == TUNE_PENDING) PH_HIPCHECK(c, R2B.ensure((size_t)list_cap(dlist) * na + 2));
    hipLaunchKernelGGL(k_zero_i32, dim3(nblocks(n, 256)), dim3(256), 0, s, c->tmpi.p, n, 0);
    int32_t zero = 0;
    PH_HIPCHECK(c, dev_set(c, c->counters.p, &zero, sizeof(int32_t)));
    {
      const IterLaunch L{c, s, ng, icall, rfac, na, lmargin, dlist, widen, LB.p, R2B.p, TNL.p, THS.p, THF.p};
      if (attempt == 0 && c->tune_mode == TUNE_PENDING && na >= TUNE_MIN_ACTIVE) {
        tuned_now = true;
        rc = tune_density_iter(c, L);  // times every configuration, keeps the fastest
        if (rc) return rc;
      } else {
        LaunchTimer lt(c, "density_iter");
        launch_dens_iter(c->tune, L);
      }
    }
    {
      LaunchTimer lt(c, "density_sums");
#define DFIN(KK, MM)                                                                                           \
  hipLaunchKernelGGL((k_dens_final<KK, MM>), dim3(xcd_grid(nblocks(na, 256))), dim3(256), 0, s, c->dc, c->act.p,  \
                     na, LB.p, TNL.p, THS.p, THF.p, c->tmpi.p, c->iparts.p, in, c->th.p, out,                       \
                     c->p_nneigh.p, c->counters.p)
#define DMIX(KK)                                                                                                \
  hipLaunchKernelGGL((k_dens_final_mixed<KK>), dim3(xcd_grid(nblocks(na, 256))), dim3(256), 0, s, c->dc, c->act.p, \
                     na, LB.p, TNL.p, THS.p, THF.p, c->tmpi.p, c->iparts.p, c->tp4.p,                               \
                     c->tv4.p, have_f ? c->tf4.p : nullptr, c->th.p, out, c->p_nneigh.p, c->counters.p)
      if (mixed) {
        if (c->p.kernel == PHSPH_KERNEL_CUBIC) DMIX(PHSPH_KERNEL_CUBIC); else DMIX(PHSPH_KERNEL_QUINTIC);
      } else if (c->p.kernel == PHSPH_KERNEL_CUBIC) {
        if (mhd) DFIN(PHSPH_KERNEL_CUBIC, true); else DFIN(PHSPH_KERNEL_CUBIC, false);
      } else {
        if (mhd) DFIN(PHSPH_KERNEL_QUINTIC, true); else DFIN(PHSPH_KERNEL_QUINTIC, false);
      }
#undef DMIX
#undef DFIN
    }
    PH_HIPCHECK(c, hipGetLastError());
    int32_t fl = 0;
    PH_HIPCHECK(c, hipMemcpyAsync(&fl, c->counters.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    if (dbg) {
      size_t fr = 0, tot = 0;
      (void)hipMemGetInfo(&fr, &tot);
      fprintf(stderr, "phsph densityiterate: attempt %d particles %d groups %d rfac %g list %d flags %d free %.1f GiB\n",
              attempt, na, ng, rfac, dlist, fl, fr / 1073741824.0);
    }
    if (fl & DFL_NOCONV) PH_RET(c, PHSPH_ERR_NOCONVERGE, "densityiterate: could not converge in density");
    // a tuning attempt whose lists overflowed (the first call on a context
    // grows the list capacity) times unrepresentative work: tune again on
    // the next large call if that call's capacity is larger
    if (attempt == 0 && tuned_now && (fl & DFL_LIST)) retune = true;
    if (fl & DFL_NEGH) PH_RET(c, PHSPH_ERR_INPUT, "densityiterate: setting negative h from hrho");
    if (attempt == 0) {
      rewalk0 = (fl & DFL_REWALK) != 0;
      // the lists of the first attempt, and which particles they do not
      // cover (those run again)
      c->dlist_nact = na;
      c->dlist_cap = dlist;
      c->dlist_nretry = 0;
      if (reusable) {
        PH_HIPCHECK(c, c->dretry0.ensure((size_t)n + 1));
        PH_HIPCHECK(c, hipMemcpyAsync(c->dretry0.p, c->tmpi.p, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
      }
    }
    if (!(fl & (DFL_REWALK | DFL_LIST))) {
      c->nact = 0;
      break;
    }
    if (fl & DFL_REWALK) rfac *= 1.5;
    if (fl & DFL_LIST) dlist *= 2;
    rc = compact_active(c, c->tmpi.p, n);
    if (rc) return rc;
    if (attempt == 0) c->dlist_nretry = c->nact;
    rc = pad_groups(c);
    if (rc) return rc;
  }
  if (c->nact != 0) PH_RET(c, PHSPH_ERR_CAPACITY, "densityiterate: neighbour search could not be widened enough");
  c->dlist_reuse = reusable;
  if (reusable && nact > 0) {
    int32_t zero = 0, bad = 0;
    PH_HIPCHECK(c, dev_set(c, c->counters.p + 2, &zero, sizeof(int32_t)));
    hipLaunchKernelGGL(k_check_hbound, dim3(nblocks(nact, 256)), dim3(256), 0, s, nact, c->act_dens.p, c->th_in.p,
                       c->th.p, rfac0 * lmargin, c->counters.p + 2);
    PH_HIPCHECK(c, hipGetLastError());
    PH_HIPCHECK(c, hipMemcpyAsync(&bad, c->counters.p + 2, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    if (bad) c->dlist_reuse = false;
  }
  // the next call's first search radius: one level wider (and never again
  // this narrow) when more than 1/200 of the particles outgrew their lists
  // or the lists could not serve the force pass, one level narrower when
  // none outgrew them
  if (nact > 0) {
    const bool outgrown = rewalk0 && c->dlist_nretry * 200 > nact;
    if (outgrown || (reusable && !c->dlist_reuse)) {
      if (c->rfac_level < 2) {
        c->rfac_floor = std::max(c->rfac_floor, c->rfac_level + 1);
        c->rfac_level++;
      }
    } else if (!rewalk0 && c->rfac_level > c->rfac_floor) {
      c->rfac_level--;
    }
  }
  while (dlist > c->caps.dlist && !list_fits(c, (size_t)list_cap(dlist) * nact * 3)) dlist /= 2;
  if (dlist > c->caps.dlist) c->caps.dlist = dlist;
  // where the list budget holds the capacity, the overflow recurs on every
  // call and the tuned configuration stands
  if (retune && c->caps.dlist > dlist0) c->tune_mode = TUNE_PENDING;
  if (grav) {
    rc = grav_hmax_replay(c, icall);
    if (rc) return rc;
  }
  // statistics (dens.F90:1670-1679); trial counts are computed on request
  PH_HIPCHECK(c, c->stats_i.ensure(8));
  long long z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  PH_HIPCHECK(c, dev_set(c, c->stats_i.p, z, sizeof(z)));
  if (nact > 0)
    hipLaunchKernelGGL(k_dens_stats, dim3(stats_grid(nact)), dim3(256), 0, s, nact, c->act_dens.p, c->pleaf.p,
                       c->p_nneigh.p, c->leaf_nits.p, nullptr, c->stats_i.p);
  PH_HIPCHECK(c, hipGetLastError());
  long long acc[8];
  PH_HIPCHECK(c, hipMemcpyAsync(acc, c->stats_i.p, sizeof(acc), hipMemcpyDeviceToHost, s));
  // the active list again holds every active particle of this call (the
  // padding of a retry set may have left c->act with a smaller buffer)
  PH_HIPCHECK(c, c->act.ensure((size_t)nact + 1));
  if (nact > 0)
    PH_HIPCHECK(c, hipMemcpyAsync(c->act.p, c->act_dens.p, sizeof(int32_t) * nact, hipMemcpyDeviceToDevice, s));
  PH_HIPCHECK(c, hipStreamSynchronize(s));
  c->nact = nact;
  if (nact > 0) {
    c->actualmean = (double)acc[0] / (double)nact;
    c->nactualtot = acc[0];
    c->nrhocalc = acc[1];
    c->maxactual = (int)acc[3];
  } else {
    c->actualmean = 0.;
    c->nactualtot = 0;
    c->nrhocalc = 0;
    c->maxactual = 0;
  }
  return PHSPH_OK;
}
