// density.hip -- leaf neighbour search and the h-rho density iteration
// (densityiterate, dens.F90:117-568; get_density_sums :578-857;
// finish_cell :1382-1466; finish_rhosum :1470-1507; store_results
// :1511-1681).
//
// Work decomposition:
//  * k_walk: one thread per leaf walks the tree exactly as getneigh does
//    (kdtree.F90:1221-1347): open a node when r^2 < (size_i+size_j+rcut)^2,
//    with the minimum-image offset of get_sep (:1468-1501).  The output is
//    a per-leaf list of candidate leaves with their periodic offsets, in
//    walk order.  Leaves are contiguous in the tree-ordered particle arrays,
//    so a candidate leaf is a contiguous range.
//  * k_density: one thread per active particle.  It first compacts the
//    candidates within the walk radius into a per-particle neighbour list,
//    then runs its own Newton-Raphson iterations on rho and Omega only,
//    and finally makes one pass with every sum (div v, the velocity,
//    acceleration and exact-linear matrices, div B) at the converged h.
//    Per-particle iteration gives the reference's result: in a cell the
//    reference keeps re-evaluating converged particles at their unchanged
//    h, so each particle's h sequence depends only on its own sums.
//
// Every neighbour membership test computes r^2 = ((dx*dx + dy*dy) + dz*dz)
// and q^2 = r^2 * (1/h)^2 with separately rounded FP64 operations, exactly
// as the reference, so neighbour counts agree bit for bit.
#include "ctx.h"

enum { DFL_CAND = 1, DFL_LIST = 2, DFL_REWALK = 4, DFL_NOCONV = 8, DFL_NEGH = 16 };

struct DensIn {
  const double *tx, *ty, *tz, *th_in;
  const double *vx, *vy, *vz;
  const double *fx, *fy, *fz;    // fxyzu + fext (null: zero)
  const double *Bx, *By, *Bz;    // Bevol (MHD)
};
struct DensOut {
  double *xyzh;  int ldx;
  float *divcurlv; int lddcv;
  float *divcurlB;
  float *alphaind;
  float *gradh; int ldgradh;
  float *dvdx;
};

__device__ __forceinline__ void offsets_of(const DevConst &dc, int code, double &ox, double &oy, double &oz) {
  int sx, sy, sz;
  off_decode(code, sx, sy, sz);
  ox = (sx == 0) ? 0. : copysign(dc.dxbound, (double)sx);
  oy = (sy == 0) ? 0. : copysign(dc.dybound, (double)sy);
  oz = (sz == 0) ? 0. : copysign(dc.dzbound, (double)sz);
}

// separation i - j with the candidate offset and minimum image
// (dens.F90:653-673)
__device__ __forceinline__ double sep2(const DevConst &dc, double xi, double yi, double zi, double xj, double yj,
                                       double zj, double ox, double oy, double oz, double &dx, double &dy,
                                       double &dz) {
  dx = xsub(xi, xadd(xj, ox));
  dy = xsub(yi, xadd(yj, oy));
  dz = xsub(zi, xadd(zj, oz));
  if (dc.periodic) {
    dx = min_image(dx, dc.dxbound, 0.5 * dc.dxbound);
    dy = min_image(dy, dc.dybound, 0.5 * dc.dybound);
    dz = min_image(dz, dc.dzbound, 0.5 * dc.dzbound);
  }
  return r2_exact(dx, dy, dz);
}

// ------------------------------------------------------------------------
// tree walk per leaf (getneigh)
// mode 0: rcut = rfac*radkern*hmax_i; mode 1: rcut = max(that, radkern*hmax_j)
// ------------------------------------------------------------------------
__global__ void k_walk(DevConst dc, int mode, double rfac, const phsph_kdnode *__restrict__ nd,
                       const double *__restrict__ hmaxn, const int32_t *__restrict__ leaf_flag,
                       const int32_t *__restrict__ nd_lo, const int32_t *__restrict__ nd_hi,
                       const int32_t *__restrict__ leaves, int nleaves, uint32_t *__restrict__ cand,
                       int32_t *__restrict__ ncand, int32_t *__restrict__ ntrial, int capc,
                       int32_t *__restrict__ flags) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nleaves) return;
  const int c = leaves[l];
  if (leaf_flag[c] <= 0) {
    ncand[l] = 0;
    ntrial[l] = 0;
    return;
  }
  const double xp0 = nd[c].xcen[0], xp1 = nd[c].xcen[1], xp2 = nd[c].xcen[2];
  const double xsizei = nd[c].size;
  // rcut = radkern*hmax (get_cell_location, neigh_kdtree.f90:363), rounded
  // on its own before it enters (size_i + size_j + rcut)^2
  const double rcuti = xmul(xmul(rfac, dc.radkern), hmaxn ? hmaxn[c] : nd[c].hmax);
  int stack[64];
  int ns = 1;
  stack[0] = 1;
  int k = 0, ntr = 0;
  while (ns > 0) {
    const int n = stack[--ns];
    const phsph_kdnode e = nd[n];
    double dx = xsub(xp0, e.xcen[0]), dy = xsub(xp1, e.xcen[1]), dz = xsub(xp2, e.xcen[2]);
    int sx = 0, sy = 0, sz = 0;
    if (dc.periodic) {
      if (fabs(dx) > dc.hdlx) { sx = (dx > 0.) ? 1 : -1; dx = xsub(dx, copysign(dc.dxbound, dx)); }
      if (fabs(dy) > dc.hdly) { sy = (dy > 0.) ? 1 : -1; dy = xsub(dy, copysign(dc.dybound, dy)); }
      if (fabs(dz) > dc.hdlz) { sz = (dz > 0.) ? 1 : -1; dz = xsub(dz, copysign(dc.dzbound, dz)); }
    }
    const double r2 = r2_exact(dx, dy, dz);
    double rcut = rcuti;
    if (mode == 1) rcut = fmax(rcuti, xmul(dc.radkern, hmaxn ? hmaxn[n] : e.hmax));
    const double s = xadd(xadd(xsizei, e.size), rcut);
    if (r2 < xmul(s, s)) {
      if (leaf_flag[n] != 0) {
        if (k < capc) cand[(size_t)l * capc + k] = cand_pack(n, off_code(sx, sy, sz));
        k++;
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
  ncand[l] = k;
  ntrial[l] = ntr;
  if (k > capc) atomicOr(flags, DFL_CAND);
}

// ------------------------------------------------------------------------
// gathers: caller layout (original order) -> tree order
// ------------------------------------------------------------------------
__global__ void k_gather_dens(int nlive, const int32_t *__restrict__ iparts, const double *__restrict__ xyzh,
                              int ldx, const double *__restrict__ vxyzu, int ldv, const double *__restrict__ fxyzu,
                              int ldf, const double *__restrict__ fext, const double *__restrict__ Bevol,
                              const int8_t *__restrict__ iphase, double *th_in, double *vx, double *vy, double *vz,
                              double *fx, double *fy, double *fz, double *Bx, double *By, double *Bz,
                              int32_t *actflag) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nlive) return;
  int ip = iparts[p];
  int o = abs(ip) - 1;
  th_in[p] = xyzh[(size_t)o * ldx + 3];
  vx[p] = vxyzu[(size_t)o * ldv + 0];
  vy[p] = vxyzu[(size_t)o * ldv + 1];
  vz[p] = vxyzu[(size_t)o * ldv + 2];
  double a0 = 0., a1 = 0., a2 = 0.;
  if (fxyzu) {
    a0 = fxyzu[(size_t)o * ldf + 0]; a1 = fxyzu[(size_t)o * ldf + 1]; a2 = fxyzu[(size_t)o * ldf + 2];
  }
  if (fext) {
    a0 = xadd(a0, fext[(size_t)o * 3 + 0]); a1 = xadd(a1, fext[(size_t)o * 3 + 1]); a2 = xadd(a2, fext[(size_t)o * 3 + 2]);
  }
  fx[p] = a0; fy[p] = a1; fz[p] = a2;
  if (Bevol) {
    Bx[p] = Bevol[(size_t)o * 4 + 0]; By[p] = Bevol[(size_t)o * 4 + 1]; Bz[p] = Bevol[(size_t)o * 4 + 2];
  }
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
__global__ void k_scan_i32(int32_t *a, int n, int32_t *total) {
  __shared__ int carry;
  __shared__ int tmp[256];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < n; base += 256) {
    int i = base + threadIdx.x;
    int v = (i < n) ? a[i] : 0;
    tmp[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
      int t = (threadIdx.x >= o) ? tmp[threadIdx.x - o] : 0;
      __syncthreads();
      tmp[threadIdx.x] += t;
      __syncthreads();
    }
    int incl = tmp[threadIdx.x];
    int tot = tmp[255];
    if (i < n) a[i] = carry + incl - v;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// ------------------------------------------------------------------------
// density: one thread per active particle
// ------------------------------------------------------------------------
template <int K, bool MHD>
__global__ void __launch_bounds__(256)
k_density(DevConst dc, int icall, double rfac, const int32_t *__restrict__ act, int nact,
          const int32_t *__restrict__ pleaf, const int32_t *__restrict__ leaves, const uint32_t *__restrict__ cand,
          const int32_t *__restrict__ ncand, int capc, const int32_t *__restrict__ nd_lo,
          const int32_t *__restrict__ nd_hi, const phsph_kdnode *__restrict__ nd,
          const int32_t *__restrict__ iparts, DensIn in, uint32_t *__restrict__ nlist, int capl,
          double *__restrict__ th_out, DensOut out, int32_t *__restrict__ p_nneigh, int32_t *__restrict__ p_nits,
          int32_t *__restrict__ leaf_nits, int32_t *__restrict__ flags) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nact) return;
  const int p = act[t];
  const int l = pleaf[p];
  const int c = leaves[l];
  const double xi = in.tx[p], yi = in.ty[p], zi = in.tz[p];
  const double h0 = in.th_in[p];
  const double pm = dc.pmass;
  const double rcut = rfac * dc.radkern * nd[c].hmax;
  const double rc2 = xmul(rcut, rcut);
  // --- neighbour list: candidates within the walk radius ---
  int nl = 0;
  const int nc = ncand[l];
  if (nc > capc) {
    atomicOr(flags, DFL_CAND);
    return;
  }
  for (int k = 0; k < nc; k++) {
    const uint32_t e = cand[(size_t)l * capc + k];
    const int cn = cand_node(e), code = cand_code(e);
    double ox, oy, oz;
    offsets_of(dc, code, ox, oy, oz);
    const int j1 = nd_hi[cn];
    for (int j = nd_lo[cn]; j <= j1; j++) {
      if (j == p) continue;
      double dx, dy, dz;
      const double r2 = sep2(dc, xi, yi, zi, in.tx[j], in.ty[j], in.tz[j], ox, oy, oz, dx, dy, dz);
      if (r2 < rc2) {
        if (nl < capl) nlist[(size_t)nl * nact + t] = (uint32_t)j | ((uint32_t)code << 27);
        nl++;
      }
    }
  }
  if (nl > capl) {
    atomicOr(flags, DFL_LIST);
    return;
  }
  // --- Newton-Raphson on h with rho and Omega only ---
  double hi = h0, hfin = h0;
  int it = 0;
  for (;;) {
    it++;
    const double hi1 = 1. / hi, hi21 = xmul(hi1, hi1);
    double srho = 0., sgh = 0.;
    for (int k = 0; k < nl; k++) {
      const uint32_t e = nlist[(size_t)k * nact + t];
      const int j = (int)(e & 0x07ffffffu);
      double ox, oy, oz, dx, dy, dz;
      offsets_of(dc, (int)(e >> 27), ox, oy, oz);
      const double r2 = sep2(dc, xi, yi, zi, in.tx[j], in.ty[j], in.tz[j], ox, oy, oz, dx, dy, dz);
      const double q2 = xmul(r2, hi21);
      if (q2 < dc.radkern2) {
        const double q = sqrt(r2) * hi1;
        double w, g;
        kern_wg<K, double>(q2, q, w, g);
        srho += w * pm;
        sgh += (-q * g - 3. * w) * pm;
      }
    }
    // finish_rhosum + finish_cell
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
      hfin = hi;
      break;
    }
    hi = hnew;
    if (dc.radkern * hi > rcut) {  // list no longer covers 2h: walk again wider
      atomicOr(flags, DFL_REWALK);
      return;
    }
  }
  // --- all sums at h = hi (the h of the last summation) ---
  const double hs1 = 1. / hi, hs21 = xmul(hs1, hs1);
  double rs_rho = 0., rs_gh = 0., rs_gs = 0., rs_divv = 0.;
  double dv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, da[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double rr[6] = {0, 0, 0, 0, 0, 0};
  double rs_divB = 0., dB[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const double vxi = in.vx[p], vyi = in.vy[p], vzi = in.vz[p];
  const double fxi = in.fx ? in.fx[p] : 0., fyi = in.fx ? in.fy[p] : 0., fzi = in.fx ? in.fz[p] : 0.;
  double Bxi = 0., Byi = 0., Bzi = 0., rhoiB = 0.;
  if (MHD) {
    rhoiB = rhoh_d(hi, dc.hfact, pm);
    Bxi = in.Bx[p] * rhoiB; Byi = in.By[p] * rhoiB; Bzi = in.Bz[p] * rhoiB;
  }
  int nneigh = 1;
  for (int k = 0; k < nl; k++) {
    const uint32_t e = nlist[(size_t)k * nact + t];
    const int j = (int)(e & 0x07ffffffu);
    double ox, oy, oz, dx, dy, dz;
    offsets_of(dc, (int)(e >> 27), ox, oy, oz);
    const double r2 = sep2(dc, xi, yi, zi, in.tx[j], in.ty[j], in.tz[j], ox, oy, oz, dx, dy, dz);
    const double q2 = xmul(r2, hs21);
    if (!(q2 < dc.radkern2)) continue;
    const double rij = sqrt(r2);
    const double q = rij * hs1;
    double w, g;
    kern_wg<K, double>(q2, q, w, g);
    rs_rho += w * pm;
    rs_gh += (-q * g - 3. * w) * pm;
    if (dc.gravity) rs_gs += kern_dphidh<K, double>(q2, q) * pm;
    nneigh++;
    const double rij1 = 1. / (rij + PH_EPS64);
    const double rg = rij1 * g;
    const double runix = dx * rg * pm, runiy = dy * rg * pm, runiz = dz * rg * pm;
    const double dvx = vxi - in.vx[j], dvy = vyi - in.vy[j], dvz = vzi - in.vz[j];
    rs_divv += dvx * runix + dvy * runiy + dvz * runiz;
    dv[0] += dvx * runix; dv[1] += dvx * runiy; dv[2] += dvx * runiz;
    dv[3] += dvy * runix; dv[4] += dvy * runiy; dv[5] += dvy * runiz;
    dv[6] += dvz * runix; dv[7] += dvz * runiy; dv[8] += dvz * runiz;
    if (in.fx) {
      const double dax = fxi - in.fx[j], day = fyi - in.fy[j], daz = fzi - in.fz[j];
      da[0] += dax * runix; da[1] += dax * runiy; da[2] += dax * runiz;
      da[3] += day * runix; da[4] += day * runiy; da[5] += day * runiz;
      da[6] += daz * runix; da[7] += daz * runiy; da[8] += daz * runiz;
    }
    rr[0] -= dx * runix; rr[1] -= dx * runiy; rr[2] -= dx * runiz;
    rr[3] -= dy * runiy; rr[4] -= dy * runiz; rr[5] -= dz * runiz;
    if (MHD) {
      const double rhoj = rhoh_d(in.th_in[j], dc.hfact, pm);
      const double dBx = Bxi - in.Bx[j] * rhoj, dBy = Byi - in.By[j] * rhoj, dBz = Bzi - in.Bz[j] * rhoj;
      rs_divB += dBx * runix + dBy * runiy + dBz * runiz;
      dB[0] += dBx * runix; dB[1] += dBx * runiy; dB[2] += dBx * runiz;
      dB[3] += dBy * runix; dB[4] += dBy * runiy; dB[5] += dBy * runiz;
      dB[6] += dBz * runix; dB[7] += dBz * runiy; dB[8] += dBz * runiz;
    }
  }
  // --- store_results at h = hfin ---
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
  p_nneigh[t] = nneigh;
  p_nits[t] = (icall == 0) ? 1 : it;
  atomicMax(&leaf_nits[l], (icall == 0) ? 1 : it);
}

// per-leaf statistics: active particle count x iterations, trial counts
__global__ void k_dens_stats(int nact, const int32_t *act, const int32_t *pleaf, const int32_t *p_nneigh,
                             const int32_t *leaf_nits, const int32_t *ntrial, long long *acc) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  long long nn = 0, nit = 0, ntr = 0, mx = 0, mxt = 0;
  if (t < nact) {
    int l = pleaf[act[t]];
    nn = p_nneigh[t];
    nit = leaf_nits[l];
    ntr = ntrial[l];
    mx = nn;
    mxt = ntr;
  }
  for (int o = 32; o > 0; o >>= 1) {
    nn += __shfl_xor(nn, o, 64);
    nit += __shfl_xor(nit, o, 64);
    ntr += __shfl_xor(ntr, o, 64);
    long long a = __shfl_xor(mx, o, 64), b = __shfl_xor(mxt, o, 64);
    mx = mx > a ? mx : a;
    mxt = mxt > b ? mxt : b;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd((unsigned long long *)&acc[0], (unsigned long long)nn);
    atomicAdd((unsigned long long *)&acc[1], (unsigned long long)nit);
    atomicAdd((unsigned long long *)&acc[2], (unsigned long long)ntr);
    atomicMax(&acc[3], mx);
    atomicMax(&acc[4], mxt);
  }
}

__global__ void k_zero_i32(int32_t *a, int n, int32_t v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_scatter_pstats(int nact, const int32_t *act, const int32_t *iparts, const int32_t *p_nneigh,
                                 const int32_t *p_nits, int32_t *nneigh, int32_t *nits) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nact) return;
  int o = abs(iparts[act[t]]) - 1;
  if (nneigh) nneigh[o] = p_nneigh[t];
  if (nits) nits[o] = p_nits[t];
}

int dens_particle_stats(phsph_ctx *c, int32_t *d_nneigh, int32_t *d_nits) {
  if (c->nact > 0)
    hipLaunchKernelGGL(k_scatter_pstats, dim3(nblocks(c->nact, 256)), dim3(256), 0, c->stream, c->nact, c->act.p,
                       c->iparts.p, c->p_nneigh.p, c->p_nits.p, d_nneigh, d_nits);
  PH_HIPCHECK(c, hipGetLastError());
  return PHSPH_OK;
}

// ------------------------------------------------------------------------
int walk_leaves(phsph_ctx *c, int mode, double rfac, const double *hmaxn) {
  hipStream_t s = c->stream;
  for (int attempt = 0; attempt < 4; attempt++) {
    PH_HIPCHECK(c, c->cand.ensure((size_t)c->nleaves * c->caps.cand));
    PH_HIPCHECK(c, c->ncand.ensure(c->nleaves));
    PH_HIPCHECK(c, c->ntrial.ensure(c->nleaves));
    int32_t zero = 0;
    PH_HIPCHECK(c, hipMemcpyAsync(c->counters.p, &zero, sizeof(int32_t), hipMemcpyHostToDevice, s));
    {
      LaunchTimer lt(c, mode == 0 ? "walk_density" : "walk_force");
      hipLaunchKernelGGL(k_walk, dim3(nblocks(c->nleaves, 128)), dim3(128), 0, s, c->dc, mode, rfac, c->nd.p,
                         hmaxn, c->leaf_flag.p, c->nd_lo.p, c->nd_hi.p, c->leaves.p, c->nleaves, c->cand.p,
                         c->ncand.p, c->ntrial.p, c->caps.cand, c->counters.p);
    }
    PH_HIPCHECK(c, hipGetLastError());
    int32_t fl = 0;
    PH_HIPCHECK(c, hipMemcpyAsync(&fl, c->counters.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    if (!(fl & DFL_CAND)) return PHSPH_OK;
    c->caps.cand *= 2;
  }
  PH_RET(c, PHSPH_ERR_CAPACITY, "neighbour walk: candidate list capacity exceeded");
}

int compact_active(phsph_ctx *c, const int32_t *flag, int n) {
  hipStream_t s = c->stream;
  const unsigned nb = nblocks(n, 256);
  PH_HIPCHECK(c, c->tmpi2.ensure((size_t)nb + 1));
  PH_HIPCHECK(c, c->act.ensure(n + 1));
  hipLaunchKernelGGL(k_compact_count, dim3(nb), dim3(256), 0, s, flag, n, c->tmpi2.p);
  hipLaunchKernelGGL(k_scan_i32, dim3(1), dim3(256), 0, s, c->tmpi2.p, (int)nb, c->counters.p + 1);
  hipLaunchKernelGGL(k_compact_fill, dim3(nb), dim3(256), 0, s, flag, n, c->tmpi2.p, c->act.p);
  int32_t na = 0;
  PH_HIPCHECK(c, hipMemcpyAsync(&na, c->counters.p + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  PH_HIPCHECK(c, hipStreamSynchronize(s));
  c->nact = na;
  return PHSPH_OK;
}

int dens_run(phsph_ctx *c, int icall, const DensIOArgs &a) {
  hipStream_t s = c->stream;
  const int n = c->nlive;
  PH_HIPCHECK(c, c->th_in.ensure(n));
  PH_HIPCHECK(c, c->tvx.ensure(n)); PH_HIPCHECK(c, c->tvy.ensure(n)); PH_HIPCHECK(c, c->tvz.ensure(n));
  PH_HIPCHECK(c, c->tfx.ensure(n)); PH_HIPCHECK(c, c->tfy.ensure(n)); PH_HIPCHECK(c, c->tfz.ensure(n));
  if (c->p.mhd) {
    PH_HIPCHECK(c, c->tBx.ensure(n)); PH_HIPCHECK(c, c->tBy.ensure(n)); PH_HIPCHECK(c, c->tBz.ensure(n));
  }
  PH_HIPCHECK(c, c->tmpi.ensure(n));
  const bool have_f = (a.fxyzu != nullptr) || (a.fext != nullptr);
  {
    LaunchTimer lt(c, "dens_gather");
    hipLaunchKernelGGL(k_gather_dens, dim3(nblocks(n, 256)), dim3(256), 0, s, n, c->iparts.p, a.xyzh, a.ldx,
                       a.vxyzu, a.ldv, a.fxyzu, a.ldf, a.fext, c->p.mhd ? a.Bevol : nullptr, a.iphase, c->th_in.p,
                       c->tvx.p, c->tvy.p, c->tvz.p, c->tfx.p, c->tfy.p, c->tfz.p, c->tBx.p, c->tBy.p, c->tBz.p,
                       c->tmpi.p);
  }
  PH_HIPCHECK(c, hipGetLastError());
  int rc = compact_active(c, c->tmpi.p, n);
  if (rc) return rc;
  const int nact = c->nact;
  PH_HIPCHECK(c, c->p_nneigh.ensure(nact + 1));
  PH_HIPCHECK(c, c->p_nits.ensure(nact + 1));
  PH_HIPCHECK(c, c->leaf_nits.ensure(c->nleaves + 1));
  DensIn in{c->tx.p, c->ty.p, c->tz.p, c->th_in.p, c->tvx.p, c->tvy.p, c->tvz.p,
            have_f ? c->tfx.p : nullptr, c->tfy.p, c->tfz.p, c->tBx.p, c->tBy.p, c->tBz.p};
  DensOut out{a.xyzh, a.ldx, a.divcurlv, a.lddcv, a.divcurlB, a.alphaind, a.gradh, a.ldgradh, a.dvdx};
  double rfac = 1.2;
  for (int attempt = 0; attempt < 8; attempt++) {
    rc = walk_leaves(c, 0, rfac, nullptr);
    if (rc) return rc;
    PH_HIPCHECK(c, c->nlist.ensure((size_t)c->caps.nlist * (nact + 1)));
    hipLaunchKernelGGL(k_zero_i32, dim3(nblocks(c->nleaves, 256)), dim3(256), 0, s, c->leaf_nits.p,
                       c->nleaves, 0);
    int32_t zero = 0;
    PH_HIPCHECK(c, hipMemcpyAsync(c->counters.p, &zero, sizeof(int32_t), hipMemcpyHostToDevice, s));
    if (nact > 0) {
      LaunchTimer lt(c, "density");
      const dim3 g(nblocks(nact, 256)), b(256);
#define DLAUNCH(KK, MM)                                                                                      \
  hipLaunchKernelGGL((k_density<KK, MM>), g, b, 0, s, c->dc, icall, rfac, c->act.p, nact, c->pleaf.p,          \
                     c->leaves.p, c->cand.p, c->ncand.p, c->caps.cand, c->nd_lo.p, c->nd_hi.p, c->nd.p,        \
                     c->iparts.p, in, c->nlist.p, c->caps.nlist, c->th.p, out, c->p_nneigh.p, c->p_nits.p,     \
                     c->leaf_nits.p, c->counters.p)
      if (c->p.kernel == PHSPH_KERNEL_CUBIC) {
        if (c->p.mhd) DLAUNCH(PHSPH_KERNEL_CUBIC, true); else DLAUNCH(PHSPH_KERNEL_CUBIC, false);
      } else {
        if (c->p.mhd) DLAUNCH(PHSPH_KERNEL_QUINTIC, true); else DLAUNCH(PHSPH_KERNEL_QUINTIC, false);
      }
#undef DLAUNCH
    }
    PH_HIPCHECK(c, hipGetLastError());
    int32_t fl = 0;
    PH_HIPCHECK(c, hipMemcpyAsync(&fl, c->counters.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    if (fl & DFL_NOCONV) PH_RET(c, PHSPH_ERR_NOCONVERGE, "densityiterate: could not converge in density");
    if (fl & DFL_NEGH) PH_RET(c, PHSPH_ERR_INPUT, "densityiterate: setting negative h from hrho");
    if (fl & DFL_CAND) { c->caps.cand *= 2; continue; }
    if (fl & DFL_LIST) { c->caps.nlist = c->caps.nlist * 3 / 2; continue; }
    if (fl & DFL_REWALK) { rfac *= 1.5; continue; }
    // statistics (dens.F90:1670-1679): actual neighbours, density calcs and
    // trial neighbours of a walk at the reference's search radius
    rc = walk_leaves(c, 0, 1.0, nullptr);
    if (rc) return rc;
    PH_HIPCHECK(c, c->stats_i.ensure(8));
    long long z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    PH_HIPCHECK(c, hipMemcpyAsync(c->stats_i.p, z, sizeof(z), hipMemcpyHostToDevice, s));
    if (nact > 0)
      hipLaunchKernelGGL(k_dens_stats, dim3(nblocks(nact, 256)), dim3(256), 0, s, nact, c->act.p, c->pleaf.p,
                         c->p_nneigh.p, c->leaf_nits.p, c->ntrial.p, c->stats_i.p);
    long long acc[8];
    PH_HIPCHECK(c, hipMemcpyAsync(acc, c->stats_i.p, sizeof(acc), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    if (nact > 0) {
      c->actualmean = (double)acc[0] / (double)nact;
      c->nactualtot = acc[0];
      c->nrhocalc = acc[1];
      c->trialmean = (double)acc[2] / (double)nact;
      c->maxactual = (int)acc[3];
      c->maxtrial = (int)acc[4];
    }
    return PHSPH_OK;
  }
  PH_RET(c, PHSPH_ERR_CAPACITY, "densityiterate: neighbour list capacity could not be satisfied");
}
