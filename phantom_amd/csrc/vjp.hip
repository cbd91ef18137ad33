// vjp.hip -- adjoints (vector-Jacobian products) of the density summation
// and of the pressure force with respect to particle positions.
//
// Phantom has no adjoint; the forward sums differentiated here are
// get_density_sums (dens.F90:578-857, rho = sum_j m W(r_ij, h_i) with h
// solved from rho = m (hfact/h)^3) and the pressure part of compute_forces
// (force.F90:1286-1600, a_i = -sum_j m [A_i dW_i + A_j dW_j] r_hat with
// A = P/(Omega rho^2), P = c rho for a barotropic closure: c = polyk for
// ieos 1, (gamma-1) u with u held fixed for ieos 2).
//
// h-rho constraint.  F_i(h_i, x) = rho_sum_i(h_i, x) - m (hfact/h_i)^3 = 0
// gives dF/dh = (3 rho/h) Omega with Omega = 1 + h/(3 rho) d(rho_sum)/dh, so
// dh_i/dx = -(h_i/(3 rho_i Omega_i)) d(rho_sum_i)/dx and
// drho_i/dx = (1/Omega_i) d(rho_sum_i)/dx (implicit differentiation; the
// forward must have converged h).  Every adjoint therefore ends in one
// gather of the form
//   dL/dx_k = sum_j m r_hat_kj [G_k W'(r, h_k) + G_j W'(r, h_j)
//                               + Db_k Z(r, h_k) + Db_j Z(r, h_j)]
// with per-particle weights G (adjoint of rho_sum) and Db (adjoint of
// D = d(rho_sum)/dh, through Omega), W' = dW/dr and Z = d^2W/dr dh.
//
// Neighbour lists are recomputed (the force lists: {j : q_i < R or
// q_j < R}) instead of storing pair data.  One thread per particle, FP64.
#include "records.h"

// w, dw/dq, d2w/dq2 of the kernels (kernel_cubic.f90:34-60,
// kernel_quintic.f90:34-75; the second derivative differentiates the
// same polynomials once more)
template <int K>
__device__ __forceinline__ void kern_w012(double q, double &w, double &g, double &g2) {
  if (K == PHSPH_KERNEL_CUBIC) {
    if (q < 1.) {
      w = 0.75 * q * q * q - 1.5 * q * q + 1.;
      g = q * (2.25 * q - 3.);
      g2 = 4.5 * q - 3.;
    } else if (q < 2.) {
      const double t = q - 2.;
      w = -0.25 * t * t * t;
      g = -0.75 * t * t;
      g2 = -1.5 * t;
    } else {
      w = g = g2 = 0.;
    }
  } else {
    if (q < 1.) {
      const double q2 = q * q, q4 = q2 * q2;
      w = -10. * q4 * q + 30. * q4 - 60. * q2 + 66.;
      g = q * (-50. * q2 * q + 120. * q2 - 120.);
      g2 = -200. * q2 * q + 360. * q2 - 120.;
    } else if (q < 2.) {
      const double a = q - 3., b = q - 2.;
      const double a2 = a * a, b2 = b * b;
      w = -(a2 * a2 * a) + 6. * (b2 * b2 * b);
      g = -5. * (a2 * a2) + 30. * (b2 * b2);
      g2 = -20. * a2 * a + 120. * b2 * b;
    } else if (q < 3.) {
      const double a = q - 3.;
      const double a2 = a * a;
      w = -(a2 * a2 * a);
      g = -5. * (a2 * a2);
      g2 = -20. * a2 * a;
    } else {
      w = g = g2 = 0.;
    }
  }
}

// kernel derivatives of W(r, h) = cnormk/h^3 w(r/h) at one (r, h):
//   W'   = dW/dr        = cnormk/h^4 w'
//   W''  = d2W/dr2      = cnormk/h^5 w''
//   Wh   = dW/dh        = -cnormk/h^4 (3 w + q w')
//   Z    = d2W/dr dh    = -cnormk/h^5 (4 w' + q w'')
//   Whh  = d2W/dh2      = cnormk/h^5 (12 w + 8 q w' + q^2 w'')
struct KD {
  double w1, w2, wh, z, whh;
};
template <int K>
__device__ __forceinline__ KD kern_derivs(const DevConst &dc, double r, double h1) {
  const double q = r * h1;
  double w, g, g2;
  kern_w012<K>(q, w, g, g2);
  const double h2 = h1 * h1, h4 = h2 * h2, h5 = h4 * h1;
  KD d;
  d.w1 = dc.cnormk * h4 * g;
  d.w2 = dc.cnormk * h5 * g2;
  d.wh = -dc.cnormk * h4 * (3. * w + q * g);
  d.z = -dc.cnormk * h5 * (4. * g + q * g2);
  d.whh = dc.cnormk * h5 * (12. * w + 8. * q * g + q * q * g2);
  return d;
}

// per-particle state of the adjoint (tree order)
struct __align__(16) VjpState {
  double omega, A, E, rho;  // Omega, P/(Omega rho^2), d2(rho_sum)/dh2, rho(h)
};

// positions (tree order, wrapped), 1/h and h from the caller's converged h,
// c = P/rho of the barotropic closure, adjoint inputs gathered to tree order
__global__ void k_vjp_prep(DevConst dc, int nlive, const int32_t *__restrict__ iparts,
                           const double *__restrict__ tx, const double *__restrict__ ty,
                           const double *__restrict__ tz, const double *__restrict__ xyzh, int ldx,
                           const double *__restrict__ vxyzu, int ldv, const double *__restrict__ fbar, int ldfb,
                           double4 *__restrict__ fpos, double *__restrict__ fh, double *__restrict__ cfac,
                           double4 *__restrict__ abar, int32_t *__restrict__ flag) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nlive) return;
  const int o = abs(iparts[p]) - 1;
  const double h = xyzh[(size_t)o * ldx + 3];
  fpos[p] = make_double4(tx[p], ty[p], tz[p], 1. / h);
  fh[p] = h;
  double cc = dc.polyk;
  if (dc.ieos == 2) cc = (dc.gamma - 1.) * ((vxyzu && dc.maxvxyzu >= 4) ? vxyzu[(size_t)o * ldv + 3] : 0.);
  cfac[p] = cc;
  if (abar) {
    if (fbar) {
      const double *f = fbar + (size_t)o * ldfb;
      abar[p] = make_double4(f[0], f[1], f[2], 0.);
    } else {
      abar[p] = make_double4(0., 0., 0., 0.);
    }
  }
  flag[p] = 1;
}

// Omega, A and d2(rho_sum)/dh2 from the sums over j with q_i < R, the self
// term included (dens.F90:735-740 for d(rho_sum)/dh)
template <int K>
__global__ void __launch_bounds__(256)
k_vjp_omega(DevConst dc, const int32_t *__restrict__ act, int nact, const uint32_t *__restrict__ nlist,
            const int64_t *loff, const int32_t *__restrict__ t_nl, const double4 *__restrict__ fpos,
            const double *__restrict__ cfac, VjpState *__restrict__ st) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= nact || act[t] < 0) return;
  const int p = act[t];
  const double4 ri = fpos[p];
  const double h1 = ri.w, h = 1. / h1;
  const double pm = dc.pmass;
  KD s0 = kern_derivs<K>(dc, 0., h1);
  double D = pm * s0.wh, E = pm * s0.whh;
  const int nl = t_nl[t];
  for (int k = 0; k < nl; k++) {
    const uint32_t e = nlist[list_at_g(loff, nact, t, k)];
    const int j = cand_node(e);
    const double4 rj = fpos[j];
    double dx, dy, dz;
    sep_of(dc, cand_code(e), ri.x, ri.y, ri.z, rj.x, rj.y, rj.z, dx, dy, dz);
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r * h1 >= dc.radkern) continue;
    const KD d = kern_derivs<K>(dc, r, h1);
    D += pm * d.wh;
    E += pm * d.whh;
  }
  const double rho = rhoh_d(h, dc.hfact, pm);
  const double omega = 1. + h * D / (3. * rho);
  VjpState S;
  S.omega = omega;
  S.A = cfac[p] / (omega * rho);
  S.E = E;
  S.rho = rho;
  st[p] = S;
}

// the pressure acceleration a_i = -sum_j m [A_i W'(r,h_i) + A_j W'(r,h_j)] r_hat
// (force.F90:1595-1596 with q = 0 viscosity), FP64 Omega
template <int K>
__global__ void __launch_bounds__(256)
k_vjp_accel(DevConst dc, const int32_t *__restrict__ act, int nact, const uint32_t *__restrict__ nlist,
            const int64_t *loff, const int32_t *__restrict__ t_nl, const int32_t *__restrict__ iparts,
            const double4 *__restrict__ fpos, const VjpState *__restrict__ st, double *__restrict__ accel, int lda) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= nact || act[t] < 0) return;
  const int p = act[t];
  const double4 ri = fpos[p];
  const double Ai = st[p].A;
  const double pm = dc.pmass;
  double ax = 0., ay = 0., az = 0.;
  const int nl = t_nl[t];
  for (int k = 0; k < nl; k++) {
    const uint32_t e = nlist[list_at_g(loff, nact, t, k)];
    const int j = cand_node(e);
    const double4 rj = fpos[j];
    double dx, dy, dz;
    sep_of(dc, cand_code(e), ri.x, ri.y, ri.z, rj.x, rj.y, rj.z, dx, dy, dz);
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double r1 = 1. / r;
    const KD di = kern_derivs<K>(dc, r, ri.w);
    const KD dj = kern_derivs<K>(dc, r, rj.w);
    const double phi = -pm * (Ai * di.w1 + st[j].A * dj.w1);
    ax += phi * dx * r1;
    ay += phi * dy * r1;
    az += phi * dz * r1;
  }
  const int o = abs(iparts[p]) - 1;
  double *a = accel + (size_t)o * lda;
  a[0] = ax; a[1] = ay; a[2] = az;
}

// adjoint of the pressure acceleration through the explicit separations and
// through A_i and h_i as they appear in the pair terms:
//   grad_i  = sum_j J_ij (abar_i - abar_j),  J = phi' rr^T + (phi/r)(I - rr^T)
//   Abar_i  = -sum_j m W'(r,h_i) r_hat.(abar_i - abar_j)
//   hbar_i  = -sum_j m A_i Z(r,h_i) r_hat.(abar_i - abar_j)
template <int K>
__global__ void __launch_bounds__(256)
k_vjp_force_pairs(DevConst dc, const int32_t *__restrict__ act, int nact, const uint32_t *__restrict__ nlist,
                  const int64_t *loff, const int32_t *__restrict__ t_nl, const double4 *__restrict__ fpos,
                  const VjpState *__restrict__ st, const double4 *__restrict__ abar, double4 *__restrict__ grad,
                  double2 *__restrict__ ahbar) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= nact || act[t] < 0) return;
  const int p = act[t];
  const double4 ri = fpos[p];
  const double Ai = st[p].A;
  const double4 ai = abar[p];
  const double pm = dc.pmass;
  double gx = 0., gy = 0., gz = 0., Ab = 0., hb = 0.;
  const int nl = t_nl[t];
  for (int k = 0; k < nl; k++) {
    const uint32_t e = nlist[list_at_g(loff, nact, t, k)];
    const int j = cand_node(e);
    const double4 rj = fpos[j];
    double dx, dy, dz;
    sep_of(dc, cand_code(e), ri.x, ri.y, ri.z, rj.x, rj.y, rj.z, dx, dy, dz);
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double r1 = 1. / r;
    const double ux = dx * r1, uy = dy * r1, uz = dz * r1;
    const KD di = kern_derivs<K>(dc, r, ri.w);
    const KD dj = kern_derivs<K>(dc, r, rj.w);
    const double Aj = st[j].A;
    const double phi = -pm * (Ai * di.w1 + Aj * dj.w1);
    const double dphi = -pm * (Ai * di.w2 + Aj * dj.w2);
    const double4 aj = abar[j];
    const double vx = ai.x - aj.x, vy = ai.y - aj.y, vz = ai.z - aj.z;
    const double s = ux * vx + uy * vy + uz * vz;
    const double pr = phi * r1;
    gx += dphi * ux * s + pr * (vx - ux * s);
    gy += dphi * uy * s + pr * (vy - uy * s);
    gz += dphi * uz * s + pr * (vz - uz * s);
    Ab += -pm * di.w1 * s;
    hb += -pm * Ai * di.z * s;
  }
  grad[p] = make_double4(gx, gy, gz, 0.);
  ahbar[p] = make_double2(Ab, hb);
}

// per-particle chain rule to the weights of the final gather:
//   rho = m (hfact/h)^3, Omega = 1 + h D/(3 rho), A = c/(Omega rho)
//   Hbar = hbar_pairs + Abar dA/dh|x + hbar_user - rhobar_user 3 rho/h
//   G    = -Hbar h/(3 rho Omega)          (adjoint of rho_sum through h(x))
//   Db   = Abar (-A/Omega) h/(3 rho)      (adjoint of D through Omega)
// with dA/dh|x = (A/rho)(3 rho/h) - (A/Omega)(4 (Omega-1)/h + h E/(3 rho)).
__global__ void k_vjp_chain(int nlive, const int32_t *__restrict__ iparts, const double *__restrict__ fh,
                            const VjpState *__restrict__ st, const double2 *__restrict__ ahbar,
                            const double *__restrict__ rhobar, const double *__restrict__ hbar,
                            double2 *__restrict__ gdb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nlive) return;
  const int o = abs(iparts[p]) - 1;
  const VjpState S = st[p];
  const double h = fh[p];
  double Hb = 0., Db = 0.;
  if (ahbar) {
    const double2 ab = ahbar[p];
    const double dAdh = (S.A / S.rho) * (3. * S.rho / h) -
                        (S.A / S.omega) * (4. * (S.omega - 1.) / h + h * S.E / (3. * S.rho));
    Hb = ab.y + ab.x * dAdh;
    Db = ab.x * (-S.A / S.omega) * h / (3. * S.rho);
  }
  if (hbar) Hb += hbar[o];
  if (rhobar) Hb += rhobar[o] * (-3. * S.rho / h);
  gdb[p] = make_double2(-Hb * h / (3. * S.rho * S.omega), Db);
}

// the gather dL/dx_k = sum_j m r_hat [G_k W'_k + G_j W'_j + Db_k Z_k + Db_j Z_j]
// added to the pair part (grad, may be null) and written in caller order
template <int K>
__global__ void __launch_bounds__(256)
k_vjp_gather(DevConst dc, const int32_t *__restrict__ act, int nact, const uint32_t *__restrict__ nlist,
             const int64_t *loff, const int32_t *__restrict__ t_nl, const int32_t *__restrict__ iparts,
             const double4 *__restrict__ fpos, const double2 *__restrict__ gdb, const double4 *__restrict__ grad,
             double *__restrict__ dxyz, int ldd) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= nact || act[t] < 0) return;
  const int p = act[t];
  const double4 ri = fpos[p];
  const double2 wi = gdb[p];
  const double pm = dc.pmass;
  double gx = 0., gy = 0., gz = 0.;
  if (grad) {
    const double4 g0 = grad[p];
    gx = g0.x; gy = g0.y; gz = g0.z;
  }
  const int nl = t_nl[t];
  for (int k = 0; k < nl; k++) {
    const uint32_t e = nlist[list_at_g(loff, nact, t, k)];
    const int j = cand_node(e);
    const double4 rj = fpos[j];
    double dx, dy, dz;
    sep_of(dc, cand_code(e), ri.x, ri.y, ri.z, rj.x, rj.y, rj.z, dx, dy, dz);
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double r1 = 1. / r;
    const KD di = kern_derivs<K>(dc, r, ri.w);
    const KD dj = kern_derivs<K>(dc, r, rj.w);
    const double2 wj = gdb[j];
    const double sc = pm * (wi.x * di.w1 + wj.x * dj.w1 + wi.y * di.z + wj.y * dj.z) * r1;
    gx += sc * dx;
    gy += sc * dy;
    gz += sc * dz;
  }
  const int o = abs(iparts[p]) - 1;
  double *d = dxyz + (size_t)o * ldd;
  d[0] = gx; d[1] = gy; d[2] = gz;
}

__global__ void k_zero_rows(int np, double *a, int ld, int ncomp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  for (int c = 0; c < ncomp; c++) a[(size_t)i * ld + c] = 0.;
}

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------
static inline VjpState *VST(phsph_ctx *c) { return reinterpret_cast<VjpState *>(c->vjp_st.p); }
static_assert(sizeof(VjpState) == sizeof(double4), "VjpState is stored in a double4 buffer");
int force_lists_all(phsph_ctx *c, DBuf<uint32_t> &L, DBuf<int32_t> &TNL, const int64_t **loff,
                    const char *tname);

// Lists and per-particle state over every live particle of the current
// tree at the caller's converged h (xyzh, device, caller order)
static int vjp_setup(phsph_ctx *c, const double *xyzh, int ldx, const double *vxyzu, int ldv, const double *fbar,
                     int ldfb, const int64_t **loff) {
  hipStream_t s = c->stream;
  const int n = c->nlive;
  PH_HIPCHECK(c, c->fpos.ensure(n + 1));
  PH_HIPCHECK(c, c->fh.ensure(n + 1));
  PH_HIPCHECK(c, c->vjp_c.ensure(n + 1));
  PH_HIPCHECK(c, c->vjp_st.ensure(n + 1));
  PH_HIPCHECK(c, c->vjp_gdb.ensure(n + 1));
  PH_HIPCHECK(c, c->tmpi.ensure(n + 1));
  if (fbar) {
    PH_HIPCHECK(c, c->vjp_abar.ensure(n + 1));
    PH_HIPCHECK(c, c->vjp_grad.ensure(n + 1));
    PH_HIPCHECK(c, c->vjp_ahb.ensure(n + 1));
  }
  hipLaunchKernelGGL(k_vjp_prep, dim3(nblocks(n, 256)), dim3(256), 0, s, c->dc, n, c->iparts.p, c->tx.p, c->ty.p,
                     c->tz.p, xyzh, ldx, vxyzu, ldv, fbar, ldfb, c->fpos.p, c->fh.p, c->vjp_c.p,
                     fbar ? c->vjp_abar.p : nullptr, c->tmpi.p);
  PH_HIPCHECK(c, hipGetLastError());
  int rc = compact_active(c, c->tmpi.p, n);
  if (rc) return rc;
  rc = pad_groups(c);
  if (rc) return rc;
  PH_HIPCHECK(c, c->hmaxf.ensure(c->ncells + 2));
  rc = leaf_hmax_run(c, c->fh.p, c->hmaxf.p);
  if (rc) return rc;
  // the lists below take the place of the density pass's lists
  c->dlist_reuse = false;
  rc = force_lists_all(c, c->vjp_nl, c->vjp_tnl, loff, "vjp_list");
  if (rc) return rc;
  const int nact = c->nact;
  if (nact > 0) {
    LaunchTimer lt(c, "vjp_omega");
    const dim3 gr(xcd_grid(nblocks(nact, 256))), b(256);
    if (c->p.kernel == PHSPH_KERNEL_CUBIC)
      hipLaunchKernelGGL(k_vjp_omega<PHSPH_KERNEL_CUBIC>, gr, b, 0, s, c->dc, c->act.p, nact, c->vjp_nl.p, *loff,
                         c->vjp_tnl.p, c->fpos.p, c->vjp_c.p, VST(c));
    else
      hipLaunchKernelGGL(k_vjp_omega<PHSPH_KERNEL_QUINTIC>, gr, b, 0, s, c->dc, c->act.p, nact, c->vjp_nl.p, *loff,
                         c->vjp_tnl.p, c->fpos.p, c->vjp_c.p, VST(c));
  }
  PH_HIPCHECK(c, hipGetLastError());
  return PHSPH_OK;
}

static int vjp_check(phsph_ctx *c, int np) {
  if (!c->tree_valid) PH_RET(c, PHSPH_ERR_STATE, "vjp: no current tree (call maketree and densityiterate first)");
  if (np != c->np) PH_RET(c, PHSPH_ERR_INPUT, "vjp: np differs from the tree's");
  if (c->p.mhd || c->p.gravity || (c->p.use_dustfrac && c->p.ndustsmall > 0))
    PH_RET(c, PHSPH_ERR_INPUT, "vjp: hydrodynamics only (no MHD, dust or self-gravity)");
  return PHSPH_OK;
}

// density adjoint: dxyz = d/dx of sum_i (rhobar_i rho_i + hbar_i h_i)
int vjp_density_run(phsph_ctx *c, int np, const double *xyzh, int ldx, const double *rhobar, const double *hbar,
                    double *dxyz, int ldd) {
  int rc = vjp_check(c, np);
  if (rc) return rc;
  hipStream_t s = c->stream;
  const int64_t *loff = nullptr;
  hipLaunchKernelGGL(k_zero_rows, dim3(nblocks(np, 256)), dim3(256), 0, s, np, dxyz, ldd, 3);
  rc = vjp_setup(c, xyzh, ldx, nullptr, 0, nullptr, 0, &loff);
  if (rc) return rc;
  const int n = c->nlive, nact = c->nact;
  hipLaunchKernelGGL(k_vjp_chain, dim3(nblocks(n, 256)), dim3(256), 0, s, n, c->iparts.p, c->fh.p, VST(c),
                     nullptr, rhobar, hbar, c->vjp_gdb.p);
  if (nact > 0) {
    LaunchTimer lt(c, "vjp_gather");
    const dim3 gr(xcd_grid(nblocks(nact, 256))), b(256);
    if (c->p.kernel == PHSPH_KERNEL_CUBIC)
      hipLaunchKernelGGL(k_vjp_gather<PHSPH_KERNEL_CUBIC>, gr, b, 0, s, c->dc, c->act.p, nact, c->vjp_nl.p, loff,
                         c->vjp_tnl.p, c->iparts.p, c->fpos.p, c->vjp_gdb.p, nullptr, dxyz, ldd);
    else
      hipLaunchKernelGGL(k_vjp_gather<PHSPH_KERNEL_QUINTIC>, gr, b, 0, s, c->dc, c->act.p, nact, c->vjp_nl.p, loff,
                         c->vjp_tnl.p, c->iparts.p, c->fpos.p, c->vjp_gdb.p, nullptr, dxyz, ldd);
  }
  PH_HIPCHECK(c, hipGetLastError());
  return PHSPH_OK;
}

// pressure acceleration (accel, may be null) and its adjoint
// (fbar null: forward only)
int vjp_force_run(phsph_ctx *c, int np, const double *xyzh, int ldx, const double *vxyzu, int ldv,
                  const double *fbar, int ldfb, double *accel, int lda, double *dxyz, int ldd) {
  int rc = vjp_check(c, np);
  if (rc) return rc;
  if (c->p.ieos == 2 && (!vxyzu || c->p.maxvxyzu < 4)) PH_RET(c, PHSPH_ERR_INPUT, "vjp: ieos 2 needs u in vxyzu");
  hipStream_t s = c->stream;
  const int64_t *loff = nullptr;
  if (accel) hipLaunchKernelGGL(k_zero_rows, dim3(nblocks(np, 256)), dim3(256), 0, s, np, accel, lda, 3);
  if (fbar && dxyz) hipLaunchKernelGGL(k_zero_rows, dim3(nblocks(np, 256)), dim3(256), 0, s, np, dxyz, ldd, 3);
  const double *fb = (fbar && dxyz) ? fbar : nullptr;
  rc = vjp_setup(c, xyzh, ldx, vxyzu, ldv, fb, ldfb, &loff);
  if (rc) return rc;
  const int n = c->nlive, nact = c->nact;
  const dim3 gr(xcd_grid(nblocks(nact > 0 ? nact : 1, 256))), b(256);
  const bool cubic = c->p.kernel == PHSPH_KERNEL_CUBIC;
  if (accel && nact > 0) {
    LaunchTimer lt(c, "vjp_accel");
    if (cubic)
      hipLaunchKernelGGL(k_vjp_accel<PHSPH_KERNEL_CUBIC>, gr, b, 0, s, c->dc, c->act.p, nact, c->vjp_nl.p, loff,
                         c->vjp_tnl.p, c->iparts.p, c->fpos.p, VST(c), accel, lda);
    else
      hipLaunchKernelGGL(k_vjp_accel<PHSPH_KERNEL_QUINTIC>, gr, b, 0, s, c->dc, c->act.p, nact, c->vjp_nl.p, loff,
                         c->vjp_tnl.p, c->iparts.p, c->fpos.p, VST(c), accel, lda);
  }
  if (fb && nact > 0) {
    {
      LaunchTimer lt(c, "vjp_force_pairs");
      if (cubic)
        hipLaunchKernelGGL(k_vjp_force_pairs<PHSPH_KERNEL_CUBIC>, gr, b, 0, s, c->dc, c->act.p, nact, c->vjp_nl.p,
                           loff, c->vjp_tnl.p, c->fpos.p, VST(c), c->vjp_abar.p, c->vjp_grad.p, c->vjp_ahb.p);
      else
        hipLaunchKernelGGL(k_vjp_force_pairs<PHSPH_KERNEL_QUINTIC>, gr, b, 0, s, c->dc, c->act.p, nact, c->vjp_nl.p,
                           loff, c->vjp_tnl.p, c->fpos.p, VST(c), c->vjp_abar.p, c->vjp_grad.p, c->vjp_ahb.p);
    }
    hipLaunchKernelGGL(k_vjp_chain, dim3(nblocks(n, 256)), dim3(256), 0, s, n, c->iparts.p, c->fh.p, VST(c),
                       c->vjp_ahb.p, nullptr, nullptr, c->vjp_gdb.p);
    LaunchTimer lt(c, "vjp_gather");
    if (cubic)
      hipLaunchKernelGGL(k_vjp_gather<PHSPH_KERNEL_CUBIC>, gr, b, 0, s, c->dc, c->act.p, nact, c->vjp_nl.p, loff,
                         c->vjp_tnl.p, c->iparts.p, c->fpos.p, c->vjp_gdb.p, c->vjp_grad.p, dxyz, ldd);
    else
      hipLaunchKernelGGL(k_vjp_gather<PHSPH_KERNEL_QUINTIC>, gr, b, 0, s, c->dc, c->act.p, nact, c->vjp_nl.p, loff,
                         c->vjp_tnl.p, c->iparts.p, c->fpos.p, c->vjp_gdb.p, c->vjp_grad.p, dxyz, ldd);
  }
  PH_HIPCHECK(c, hipGetLastError());
  return PHSPH_OK;
}

// ---- one kick-drift-kick step and its adjoint ---------------------------
// x, v: device, caller order.  v_half = v + dt/2 a(x); x' = x + dt v_half;
// v' = v_half + dt/2 a(x').
__global__ void k_vjp_axpy3(int np, double *y, int ldy, const double *x, int ldx, double s, const double *z,
                            int ldz) {
  // y = x + s z (three components)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  for (int c = 0; c < 3; c++) y[(size_t)i * ldy + c] = x[(size_t)i * ldx + c] + s * z[(size_t)i * ldz + c];
}
__global__ void k_vjp_copy_rows(int np, double *y, int ldy, const double *x, int ldx, int ncomp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  for (int c = 0; c < ncomp; c++) y[(size_t)i * ldy + c] = x[(size_t)i * ldx + c];
}

static int vjp_density_at(phsph_ctx *c, int np, double *X, const double *V, int ldv) {
  int rc = tree_build(c, X, 4, np, nullptr);
  if (rc) return rc;
  DensIOArgs d;
  memset(&d, 0, sizeof(d));
  d.xyzh = X; d.ldx = 4;
  d.vxyzu = V; d.ldv = ldv;
  d.divcurlv = c->vjp_dcv.p; d.lddcv = 1;
  d.alphaind = c->vjp_alpha.p;
  d.gradh = c->vjp_gradh.p; d.ldgradh = 1;
  d.dvdx = c->vjp_dvdx.p;
  return dens_run(c, 1, d);
}

int vjp_step_run(phsph_ctx *c, int np, double dt, const double *xyzh, int ldx, const double *vxyzu, int ldv,
                 double *xyzh_new, int ldxn, double *vxyzu_new, int ldvn, const double *xbar, int ldxb,
                 const double *vbar, int ldvb, double *dx, int lddx, double *dv, int lddv) {
  if (c->p.mhd || c->p.gravity || (c->p.use_dustfrac && c->p.ndustsmall > 0))
    PH_RET(c, PHSPH_ERR_INPUT, "step_vjp: hydrodynamics only (no MHD, dust or self-gravity)");
  if (c->p.ieos == 2 && c->p.maxvxyzu < 4) PH_RET(c, PHSPH_ERR_INPUT, "step_vjp: ieos 2 needs u in vxyzu");
  hipStream_t s = c->stream;
  const size_t nn = (size_t)np + 1;
  const int nv = c->p.maxvxyzu;
  PH_HIPCHECK(c, c->vjp_x0.ensure(nn * 4));
  PH_HIPCHECK(c, c->vjp_x1.ensure(nn * 4));
  PH_HIPCHECK(c, c->vjp_vh.ensure(nn * nv));
  PH_HIPCHECK(c, c->vjp_a.ensure(nn * 3));
  PH_HIPCHECK(c, c->vjp_t3.ensure(nn * 3));
  PH_HIPCHECK(c, c->vjp_t3b.ensure(nn * 3));
  PH_HIPCHECK(c, c->vjp_dcv.ensure(nn));
  PH_HIPCHECK(c, c->vjp_alpha.ensure(nn * 3));
  PH_HIPCHECK(c, c->vjp_gradh.ensure(nn));
  PH_HIPCHECK(c, c->vjp_dvdx.ensure(nn * 9));
  const unsigned nb = nblocks(np, 256);
  double *X0 = c->vjp_x0.p, *X1 = c->vjp_x1.p, *VH = c->vjp_vh.p, *A = c->vjp_a.p;
  hipLaunchKernelGGL(k_vjp_copy_rows, dim3(nb), dim3(256), 0, s, np, X0, 4, xyzh, ldx, 4);
  hipLaunchKernelGGL(k_vjp_copy_rows, dim3(nb), dim3(256), 0, s, np, VH, nv, vxyzu, ldv, nv);
  PH_HIPCHECK(c, hipGetLastError());
  // forward: a(x), v_half, x'
  int rc = vjp_density_at(c, np, X0, vxyzu, ldv);
  if (rc) return rc;
  rc = vjp_force_run(c, np, X0, 4, vxyzu, ldv, nullptr, 0, A, 3, nullptr, 0);
  if (rc) return rc;
  hipLaunchKernelGGL(k_vjp_axpy3, dim3(nb), dim3(256), 0, s, np, VH, nv, vxyzu, ldv, 0.5 * dt, A, 3);
  hipLaunchKernelGGL(k_vjp_copy_rows, dim3(nb), dim3(256), 0, s, np, X1, 4, X0, 4, 4);  // h: start from x's
  hipLaunchKernelGGL(k_vjp_axpy3, dim3(nb), dim3(256), 0, s, np, X1, 4, X0, 4, dt, VH, nv);
  PH_HIPCHECK(c, hipGetLastError());
  // a(x'), v'
  rc = vjp_density_at(c, np, X1, VH, nv);
  if (rc) return rc;
  const bool back = xbar || vbar;
  double *T3 = c->vjp_t3.p, *T3b = c->vjp_t3b.p;
  // abar(x') = dt/2 vbar
  if (back) {
    hipLaunchKernelGGL(k_zero_rows, dim3(nb), dim3(256), 0, s, np, T3, 3, 3);
    if (vbar) hipLaunchKernelGGL(k_vjp_axpy3, dim3(nb), dim3(256), 0, s, np, T3, 3, T3, 3, 0.5 * dt, vbar, ldvb);
  }
  rc = vjp_force_run(c, np, X1, 4, VH, nv, back ? T3 : nullptr, 3, A, 3, back ? T3b : nullptr, 3);
  if (rc) return rc;
  if (xyzh_new) hipLaunchKernelGGL(k_vjp_copy_rows, dim3(nb), dim3(256), 0, s, np, xyzh_new, ldxn, X1, 4, 4);
  if (vxyzu_new) {
    hipLaunchKernelGGL(k_vjp_copy_rows, dim3(nb), dim3(256), 0, s, np, vxyzu_new, ldvn, VH, nv, nv);
    hipLaunchKernelGGL(k_vjp_axpy3, dim3(nb), dim3(256), 0, s, np, vxyzu_new, ldvn, VH, nv, 0.5 * dt, A, 3);
  }
  PH_HIPCHECK(c, hipGetLastError());
  if (!back) return PHSPH_OK;
  // Xbar' = xbar + VJP_a(x', dt/2 vbar)                      (in T3b)
  if (xbar) hipLaunchKernelGGL(k_vjp_axpy3, dim3(nb), dim3(256), 0, s, np, T3b, 3, T3b, 3, 1.0, xbar, ldxb);
  // Vbar_half = vbar + dt Xbar'                               (in dv)
  hipLaunchKernelGGL(k_zero_rows, dim3(nb), dim3(256), 0, s, np, dv, lddv, 3);
  if (vbar) hipLaunchKernelGGL(k_vjp_axpy3, dim3(nb), dim3(256), 0, s, np, dv, lddv, dv, lddv, 1.0, vbar, ldvb);
  hipLaunchKernelGGL(k_vjp_axpy3, dim3(nb), dim3(256), 0, s, np, dv, lddv, dv, lddv, dt, T3b, 3);
  // abar(x) = dt/2 Vbar_half; Xbar = Xbar' + VJP_a(x, abar) (lists and h
  // recomputed at x)
  hipLaunchKernelGGL(k_zero_rows, dim3(nb), dim3(256), 0, s, np, T3, 3, 3);
  hipLaunchKernelGGL(k_vjp_axpy3, dim3(nb), dim3(256), 0, s, np, T3, 3, T3, 3, 0.5 * dt, dv, lddv);
  PH_HIPCHECK(c, hipGetLastError());
  hipLaunchKernelGGL(k_vjp_copy_rows, dim3(nb), dim3(256), 0, s, np, X0, 4, xyzh, ldx, 4);
  rc = vjp_density_at(c, np, X0, vxyzu, ldv);
  if (rc) return rc;
  rc = vjp_force_run(c, np, X0, 4, vxyzu, ldv, T3, 3, nullptr, 0, dx, lddx);
  if (rc) return rc;
  hipLaunchKernelGGL(k_vjp_axpy3, dim3(nb), dim3(256), 0, s, np, dx, lddx, dx, lddx, 1.0, T3b, 3);
  PH_HIPCHECK(c, hipGetLastError());
  return PHSPH_OK;
}
