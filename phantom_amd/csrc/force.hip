// force.hip -- primitive variables and the force pair loop
// (cons2prim_everything, cons2prim.f90:274-456; force, force.F90:193-905;
// compute_forces :914-2060; get_stress :2068-2168; start_cell
// :2172-2514; finish_cell_and_store_results :2649-3330).
//
// Physics covered: hydro with the grad-h terms, Monaghan viscosity with the
// Cullen-Dehnen alpha, artificial conductivity, ideal MHD (anisotropic
// stress, Borve div B source subtraction, artificial resistivity, psi
// cleaning) and one-fluid dust (sqrt(eps/(1-eps)) evolution, Delta v).
//
// One wavefront per group of 64 active particles, one particle per lane
// (stream.h).  Pass 1 pulls the group's candidate stream through LDS and
// collects each lane's symmetric neighbour set {j : q_i < R or q_j < R}
// (force.F90:1287) with the same separately rounded FP64 tests as the
// reference; pass 2 evaluates the pair terms.  Quantities of particle j that
// the reference recomputes per pair (rho_j, P/rho^2, wave speed, MHD stress,
// stopping time) are evaluated once per particle by k_force_prep with the
// same arithmetic and packed into one 128-byte record per particle, so a
// pair reads one cache line of j data.
#include "records.h"

enum { FFL_DTC = 32 };

// per-particle j terms (force.F90:1455-1509 with get_stress)
__global__ void k_force_prep(DevConst dc, int nlive, const int32_t *__restrict__ iparts,
                             const double *__restrict__ tx, const double *__restrict__ ty,
                             const double *__restrict__ tz, const double *__restrict__ xyzh, int ldx,
                             const double *__restrict__ vxyzu, int ldv, const float *__restrict__ gradh,
                             int ldgradh, const float *__restrict__ alphaind, const double *__restrict__ eos,
                             int ldeos, const double *__restrict__ Bevol, const double *__restrict__ dustfrac,
                             double *__restrict__ th, FRec *__restrict__ rec,
                             FRecB *__restrict__ recB, FRecD *__restrict__ recD,
                             const int8_t *__restrict__ ibin_old, int8_t *__restrict__ tbin_old,
                             float *__restrict__ gsoft) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nlive) return;
  const int o = abs(iparts[p]) - 1;
  const double pm = dc.pmass;
  const double hh = xyzh[(size_t)o * ldx + 3];
  th[p] = hh;
  if (ibin_old) tbin_old[p] = ibin_old[o];
  if (gsoft) gsoft[p] = gradh[(size_t)o * ldgradh + 1];
  const double hj1 = 1. / hh;  // xyzcache(4) = 1/treecache(4) (kdtree.F90:1202)
  FRec R;
  R.x = tx[p]; R.y = ty[p]; R.z = tz[p];
  R.hinv = hj1;
  R.vx = vxyzu[(size_t)o * ldv + 0];
  R.vy = vxyzu[(size_t)o * ldv + 1];
  R.vz = vxyzu[(size_t)o * ldv + 2];
  R.u = (dc.maxvxyzu >= 4) ? vxyzu[(size_t)o * ldv + 3] : 0.;
  R.gradh = gradh[(size_t)o * ldgradh];
  R.alpha = alphaind[(size_t)o * 3];
  const double prj = eos[(size_t)o * ldeos], csj = eos[(size_t)o * ldeos + 1];
  R.pr = prj;
  R.cs = csj;
  R.h = hh;
  const double hj = 1. / hj1;
  const double rhoj = rhoh_d(hj, dc.hfact, pm);
  const double rho1j = 1. / rhoj;
  R.rho = rhoj;
  R.rho1 = rho1j;
  if (dc.mhd) {
    FRecB M;
    const double rhom = rhoh_d(hh, dc.hfact, pm);  // force.F90:1439-1443
    const double *Be = Bevol + (size_t)o * 4;
    M.Bex = Be[0]; M.Bey = Be[1]; M.Bez = Be[2];
    const double Bxj = Be[0] * rhom, Byj = Be[1] * rhom, Bzj = Be[2] * rhom;
    M.Bx = Bxj; M.By = Byj; M.Bz = Bzj;
    M.psi = Be[3];
    // get_stress MHD branch (force.F90:2136-2158)
    const double Brx = Bxj * rho1j, Bry = Byj * rho1j, Brz = Bzj * rho1j;
    const double Bro2 = Brx * Brx + Bry * Bry + Brz * Brz;
    const double va2 = Bro2 * rhoj;
    R.vwave = sqrt(csj * csj + va2);
    M.sxx = -(pm * Brx * Brx); M.sxy = -(pm * Brx * Bry); M.sxz = -(pm * Brx * Brz);
    M.syy = -(pm * Bry * Bry); M.syz = -(pm * Bry * Brz); M.szz = -(pm * Brz * Brz);
    M.pad0 = M.pad1 = M.pad2 = 0.;
    R.pro2 = prj * rho1j * rho1j + 0.5 * Bro2;
    recB[p] = M;
  } else {
    R.vwave = csj;
    R.pro2 = prj * rho1j * rho1j;
  }
  rec[p] = R;
  if (dc.dust) {
    FRecD D;
    const double e = dustfrac[o];
    D.eps = e;
    D.sqeps = sqrt(e / (1. - e));
    double tsj = get_ts_d(dc, rhoj * (1. - e), rhoj * e);
    if (dc.ilimitdustflux) tsj = fmin(tsj, hj / csj);
    D.ts = tsj;
    D.pad = 0.;
    recD[p] = D;
  }
}

// LDS copy of the j quantities of one batch of candidate leaves, as arrays
// (the lanes read them at independent positions)
template <int NP, bool MHD, bool DUST, bool THERMAL>
struct ForceStage {
  static constexpr int NB = MHD ? NP : 1, ND = DUST ? NP : 1, NT = THERMAL ? NP : 1;
  double x[NP], y[NP], z[NP], hinv[NP];
  double vx[NP], vy[NP], vz[NP], u[NT];
  double rho[NP], rho1[NP], pr[NP], pro2[NP], vwave[NP];
  float gradh[NP], alpha[NP];
  double Bx[NB], By[NB], Bz[NB], psi[NB], sxx[NB], sxy[NB], sxz[NB], syy[NB], syz[NB], szz[NB];
  double eps[ND], sqeps[ND], ts[ND];
  int j[NP];
  unsigned char leaf[NP];
  int8_t bold[NP];  // ibin_old (block timesteps)
};

// Force: one wavefront per group of 64 active particles, one per lane.
template <int K, bool MHD, bool DUST, bool THERMAL>
__global__ void __launch_bounds__(64)
k_force(DevConst dc, const int32_t *__restrict__ act, int nact, int ngroups, const int32_t *__restrict__ pleaf,
        const int32_t *__restrict__ leaves, const uint32_t *__restrict__ gcand, const int32_t *__restrict__ gncand,
        int capg, const int32_t *__restrict__ nd_lo, const int32_t *__restrict__ nd_hi,
        const phsph_kdnode *__restrict__ nd, const int32_t *__restrict__ iparts, const FRec *__restrict__ rec,
        const FRecB *__restrict__ recB, const FRecD *__restrict__ recD, ForceIOArgs out,
        const int8_t *__restrict__ tbin_old, int icall, unsigned long long *__restrict__ dtred,
        int32_t *__restrict__ flags) {
  constexpr int NL = (MHD || DUST) ? 8 : 16, NP = NL * PH_MINPART, NM = (NP + 63) / 64;
  __shared__ BatchMeta<NL> M;
  __shared__ ForceStage<NP, MHD, DUST, THERMAL> S;
  const int g = xcd_block(blockIdx.x, gridDim.x);
  if (g >= ngroups) return;
  const int lane = threadIdx.x;
  const int t = g * GROUP + lane;
  const bool live = t < nact;
  const int p = live ? act[t] : act[0];
  const double pm = dc.pmass;
  const double R2 = dc.radkern2;
  const int c = leaves[pleaf[p]];
  const double lcx = nd[c].xcen[0], lcy = nd[c].xcen[1], lcz = nd[c].xcen[2], lsz = nd[c].size;
  const FRec Ri = rec[p];
  const double xi = Ri.x, yi = Ri.y, zi = Ri.z;
  const double hi = Ri.h;
  // rhoanddhdrho (part.F90:805-816)
  const double hi1 = 1. / fabs(hi);
  const double tt = dc.hfact * hi1;
  const double rhoi = pm * (tt * tt * tt);
  const double rho1i = 1. / rhoi;
  const double hi21 = xmul(hi1, hi1), hi41 = hi21 * hi21;
  double gradhi = (double)Ri.gradh;
  if (!(gradhi > 0.)) gradhi = 1.;  // force.F90:2616-2621
  const double hfacgrkern = hi41 * dc.cnormk * gradhi;
  const double vxi = Ri.vx, vyi = Ri.vy, vzi = Ri.vz, eni = Ri.u;
  const double pri = Ri.pr, csi = Ri.cs, alphai = (double)Ri.alpha;
  // i terms: get_stress with rho from rhoanddhdrho
  double Bxi = 0., Byi = 0., Bzi = 0., psii = 0., pro2i, vwavei;
  double sxi = 0., sxyi = 0., sxzi = 0., syi = 0., syzi = 0., szi = 0.;
  if (MHD) {
    const FRecB Mi = recB[p];
    Bxi = Mi.Bex * rhoi; Byi = Mi.Bey * rhoi; Bzi = Mi.Bez * rhoi;
    psii = Mi.psi;
    const double Brx = Bxi * rho1i, Bry = Byi * rho1i, Brz = Bzi * rho1i;
    const double Bro2 = Brx * Brx + Bry * Bry + Brz * Brz;
    vwavei = sqrt(csi * csi + Bro2 * rhoi);
    sxi = -(pm * Brx * Brx); sxyi = -(pm * Brx * Bry); sxzi = -(pm * Brx * Brz);
    syi = -(pm * Bry * Bry); syzi = -(pm * Bry * Brz); szi = -(pm * Brz * Brz);
    pro2i = pri * rho1i * rho1i + 0.5 * Bro2;
  } else {
    vwavei = csi;
    pro2i = pri * rho1i * rho1i;
  }
  double epsi = 0., tsi = 0., epstsi = 0., sqepsi = 0.;
  if (DUST) {
    epsi = recD[p].eps;
    tsi = get_ts_d(dc, rhoi * (1. - epsi), rhoi * epsi);
    if (dc.ilimitdustflux) tsi = fmin(tsi, hi / csi);
    epstsi = epsi * tsi;
    sqepsi = sqrt(epsi / (1. - epsi));
  }
  const double rho21i = rho1i * rho1i;
  const double mrhoi5 = 0.5 * pm * rho1i;
  const double auterm = mrhoi5 * dc.alphau;
  const double avBterm = mrhoi5 * dc.alphaB * rho1i;
  // --- pair terms over the candidate stream (force.F90:1227-2056) ---
  double fx = 0., fy = 0., fz = 0., drhodt = 0., dudtdiss = 0., dendtdiss = 0.;
  double divBsym = 0., dBx_s = 0., dBy_s = 0., dBz_s = 0., divBdiff = 0.;
  double ddust = 0., dudtdust = 0., dvx_s = 0., dvy_s = 0., dvz_s = 0.;
  double vsigmax = 0.;
  const bool usej_always = MHD || DUST || (THERMAL && !dc.gravity);
  int nl = 0;
  const bool ind = out.ibin != nullptr;
  int ibin_neighi = 0;
  const int8_t ibinnow_m1 = (int8_t)(out.ibinnow - 1);
  const int nc = gncand[g];
  const uint32_t *gc = gcand + (size_t)g * capg;
  for (int b0 = 0; b0 < nc; b0 += NL) {
    int nb;
    const int total = batch_meta<NL>(M, gc, b0, nc, nd, nd_lo, nd_hi, lane, nb);
    for (int s = lane; s < total; s += GROUP) {
      const int L = batch_leaf_of<NL>(M, nb, s);
      const int j = M.lo[L] + (s - M.start[L]);
      const FRec R = rec[j];
      S.x[s] = R.x; S.y[s] = R.y; S.z[s] = R.z; S.hinv[s] = R.hinv;
      S.vx[s] = R.vx; S.vy[s] = R.vy; S.vz[s] = R.vz;
      if (THERMAL) S.u[s] = R.u;
      S.rho[s] = R.rho; S.rho1[s] = R.rho1; S.pr[s] = R.pr; S.pro2[s] = R.pro2; S.vwave[s] = R.vwave;
      S.gradh[s] = R.gradh; S.alpha[s] = R.alpha;
      if (MHD) {
        const FRecB B = recB[j];
        S.Bx[s] = B.Bx; S.By[s] = B.By; S.Bz[s] = B.Bz; S.psi[s] = B.psi;
        S.sxx[s] = B.sxx; S.sxy[s] = B.sxy; S.sxz[s] = B.sxz; S.syy[s] = B.syy; S.syz[s] = B.syz; S.szz[s] = B.szz;
      }
      if (DUST) {
        const FRecD D = recD[j];
        S.eps[s] = D.eps; S.sqeps[s] = D.sqeps; S.ts[s] = D.ts;
      }
      S.j[s] = j;
      S.leaf[s] = (unsigned char)L;
      if (ind) S.bold[s] = tbin_old[j];
    }
    __syncthreads();
    // symmetric neighbour test {q_i < R or q_j < R} (force.F90:1250-1290)
    unsigned long long mk[NM];
    for (int w = 0; w < NM; w++) mk[w] = 0ull;
    SepCodes codes;
    for (int L = 0; L < nb; L++) {
      const int code = leaf_sep_code(dc, M.cx[L], M.cy[L], M.cz[L], M.sz[L], lcx, lcy, lcz, lsz);
      codes.set(L, code);
      const int s1 = M.start[L + 1];
      for (int s = M.start[L]; s < s1; s++) {
        double dx, dy, dz;
        sep_of(dc, code, xi, yi, zi, S.x[s], S.y[s], S.z[s], dx, dy, dz);
        const double r2 = r2_exact(dx, dy, dz);
        const double hj1 = S.hinv[s];
        const bool hit = live && S.j[s] != p && (xmul(r2, hi21) < R2 || xmul(r2, xmul(hj1, hj1)) < R2);
        const unsigned long long bit = (unsigned long long)hit << (s & 63);
#pragma unroll
        for (int w = 0; w < NM; w++)
          if ((s >> 6) == w) mk[w] |= bit;
      }
    }
    if (ind) {
      // force.F90:1346-1352: every neighbour of an active particle is
      // flagged to be woken; the limiter needs the largest neighbour bin
#pragma unroll
      for (int w = 0; w < NM; w++) {
        unsigned long long m = mk[w], u = mk[w];
        while (m != 0ull) {
          const int s = w * 64 + (int)__builtin_ctzll(m);
          m &= m - 1ull;
          ibin_neighi = max(ibin_neighi, (int)S.bold[s]);
        }
        for (int off = 32; off > 0; off >>= 1) u |= __shfl_xor(u, off, 64);
        const int s = w * 64 + lane;
        if ((u >> lane) & 1ull) {
          const int o = abs(iparts[S.j[s]]) - 1;
          if (out.ibin_wake[o] < ibinnow_m1) out.ibin_wake[o] = ibinnow_m1;
        }
      }
    }
#pragma unroll
    for (int w = 0; w < NM; w++) {
      unsigned long long m = mk[w];
      nl += __popcll(m);
      while (m != 0ull) {
        const int s = w * 64 + (int)__builtin_ctzll(m);
        m &= m - 1ull;
        double dx, dy, dz;
        sep_of(dc, codes.get(S.leaf[s]), xi, yi, zi, S.x[s], S.y[s], S.z[s], dx, dy, dz);
        const double rij2 = r2_exact(dx, dy, dz);
        const double q2i = xmul(rij2, hi21);
        const double hj1 = S.hinv[s];
        const double hj21 = xmul(hj1, hj1);
        const double q2j = xmul(rij2, hj21);
        double rij1, qi;
        if (rij2 > PH_TINY64) {
          rij1 = 1. / sqrt(rij2);
          qi = (rij2 * rij1) * hi1;
        } else {
          rij1 = 0.;
          qi = 0.;
        }
        const double grkerni = (q2i < R2) ? kern_g<K, double>(q2i, qi) * hfacgrkern : 0.;
        double grkernj = 0.;
        bool usej = false;
        if (q2j < R2) {
          const double qj = (rij2 * rij1) * hj1;
          grkernj = kern_g<K, double>(q2j, qj) * hj21 * hj21 * dc.cnormk * (double)S.gradh[s];
          usej = true;
        }
        if (usej_always) usej = true;
        const double runix = dx * rij1, runiy = dy * rij1, runiz = dz * rij1;
        const double dvx = vxi - S.vx[s], dvy = vyi - S.vy[s], dvz = vzi - S.vz[s];
        const double projv = dvx * runix + dvy * runiy + dvz * runiz;
        const double denij = THERMAL ? (eni - S.u[s]) : 0.;
        const double vsigi = fmax(vwavei - dc.beta * projv, 0.);
        const double vsigavi = fmax(alphai * vwavei - dc.beta * projv, 0.);
        vsigmax = fmax(vsigmax, vsigi);
        double rho1j = 0., rho21j = 0., pro2j = 0., vwavej = 0., vsigavj = 0., autermj = 0., avBtermj = 0.;
        double rhoj = 0., prj = 0.;
        if (usej) {
          rhoj = S.rho[s];
          rho1j = S.rho1[s];
          rho21j = rho1j * rho1j;
          prj = S.pr[s];
          pro2j = S.pro2[s];
          vwavej = S.vwave[s];
          const double mrhoj5 = 0.5 * pm * rho1j;
          autermj = mrhoj5 * dc.alphau;
          avBtermj = mrhoj5 * dc.alphaB * rho1j;
          vsigmax = fmax(vsigmax, fmax(vwavej - dc.beta * projv, 0.));
          vsigavj = fmax((double)S.alpha[s] * vwavej - dc.beta * projv, 0.);
        }
        double qrho2i = 0., qrho2j = 0.;
        if (projv < 0.) {
          qrho2i = -0.5 * rho1i * vsigavi * projv;
          if (usej) qrho2j = -0.5 * rho1j * vsigavj * projv;
        }
        const double dudtdissi = pm * qrho2i * projv * grkerni;
        const double gradpi = pm * (pro2i + qrho2i) * grkerni;
        const double gradpj = usej ? pm * (pro2j + qrho2j) * grkernj : 0.;
        double dendissterm = 0.;
        if (THERMAL) {
          double vsigu;
          if (dc.gravity) vsigu = fabs(projv);
          else vsigu = sqrt(fabs(pri - prj) * (2. / (rhoi + rhoj)));
          dendissterm = vsigu * denij * (auterm * grkerni + autermj * grkernj);
        }
        double projsx = 0., projsy = 0., projsz = 0., dudtresist = 0.;
        if (MHD) {
          const double Bxj = S.Bx[s], Byj = S.By[s], Bzj = S.Bz[s], psij = S.psi[s];
          const double dBx = Bxi - Bxj, dBy = Byi - Byj, dBz = Bzi - Bzj;
          const double projBi = Bxi * runix + Byi * runiy + Bzi * runiz;
          const double projBj = Bxj * runix + Byj * runiy + Bzj * runiz;
          const double projdB = dBx * runix + dBy * runiy + dBz * runiz;
          const double dB2 = dBx * dBx + dBy * dBy + dBz * dBz;
          divBdiff += -pm * projdB * grkerni;
          const double ax = dvx - projv * runix, ay = dvy - projv * runiy, az = dvz - projv * runiz;
          const double vsigB = sqrt(ax * ax + ay * ay + az * az);
          const double dBdissterm = (avBterm * grkerni + avBtermj * grkernj) * vsigB;
          if (dc.iresistive_heating > 0) dudtresist = -0.5 * dB2 * dBdissterm;
          const double pmi = pm * rho21i * grkerni;
          const double pmj = pm * rho21j * grkernj;
          const double termi = pmi * projBi;
          divBsym += termi + pmj * projBj;
          const double dpsiterm = dc.overcleanfac * (pmi * psii * vwavei + pmj * psij * vwavej);
          dBx_s += -termi * dvx + dBdissterm * dBx - dpsiterm * runix;
          dBy_s += -termi * dvy + dBdissterm * dBy - dpsiterm * runiy;
          dBz_s += -termi * dvz + dBdissterm * dBz - dpsiterm * runiz;
          projsx = (sxi * runix + sxyi * runiy + sxzi * runiz) * grkerni +
                   (S.sxx[s] * runix + S.sxy[s] * runiy + S.sxz[s] * runiz) * grkernj;
          projsy = (sxyi * runix + syi * runiy + syzi * runiz) * grkerni +
                   (S.sxy[s] * runix + S.syy[s] * runiy + S.syz[s] * runiz) * grkernj;
          projsz = (sxzi * runix + syzi * runiy + szi * runiz) * grkerni +
                   (S.sxz[s] * runix + S.syz[s] * runiy + S.szz[s] * runiz) * grkernj;
        }
        const double gradp = gradpi + gradpj;
        fx = fx - runix * gradp - projsx;
        fy = fy - runiy * gradp - projsy;
        fz = fz - runiz * gradp - projsz;
        drhodt += projv * grkerni;
        if (THERMAL) {
          dudtdiss += dudtdissi + dudtresist;
          dendtdiss += dendissterm;
        }
        if (DUST) {
          const double epsj = S.eps[s];
          const double tsj = S.ts[s];
          const double epstsj = epsj * tsj;
          if (epsi > 0. || epsj > 0.) {
            const double grkernav = 0.5 * (grkerni + grkernj);
            const double dft = pm * S.sqeps[s] * rho1j *
                               ((tsi - epstsi) * (1. - epsi) / (1. - epsi) + (tsj - epstsj) * (1. - epsj) / (1. - epsj)) *
                               (pri - prj) * grkernav * rij1;
            ddust -= dft;
            if (THERMAL) dudtdust -= sqepsi * dft * denij;
          }
          if (epsi < 1.) {
            const double term = tsi / (1. - epsi) * pm * (pro2i * grkerni + pro2j * grkernj);
            dvx_s += term * runix; dvy_s += term * runiy; dvz_s += term * runiz;
          }
        }
  
      }
    }
    __syncthreads();
  }
  // --- finish_cell_and_store_results ---
  double dtc_o = PH_HUGE64, dtf_o = PH_HUGE64;
  if (live) {
      const int o = abs(iparts[p]) - 1;
      if (MHD) {
        const double B2i = Bxi * Bxi + Byi * Byi + Bzi * Bzi;
        double frac = 0.;
        if (B2i > 0.) {
          const double betai = 2. * pri / B2i;
          frac = (betai < 2.) ? 1. : (betai < 10.) ? (10. - betai) * 0.125 : 0.;
        }
        fx = fx - Bxi * divBsym * frac;
        fy = fy - Byi * divBsym * frac;
        fz = fz - Bzi * divBsym * frac;
        if (out.divBsymm) out.divBsymm[o] = (float)(rhoi * divBsym);
      }
      const double f2i = fx * fx + fy * fy + fz * fz;
      double *fo = out.fxyzu + (size_t)o * out.ldf;
      fo[0] = fx; fo[1] = fy; fo[2] = fz;
      const double drhodti = pm * drhodt;
      const double divvi = -drhodti * rho1i;
      out.divcurlv[(size_t)o * out.lddcv] = (float)divvi;
      double dtc = dc.dtmax, dtf = PH_BIGNUMBER, dtclean = PH_BIGNUMBER, dtdusti = PH_BIGNUMBER;
      double fxyz4 = 0.;
      if (THERMAL) {
        const double rhogasi = DUST ? rhoi * (1. - epsi) : rhoi;
        const double fac = (rhogasi > PH_TINY64) ? rhoi / rhogasi : 0.;
        const double pdv_work = fac * pri * rho1i * rho1i * drhodti;
        if (dc.ipdv_heating > 0) fxyz4 += pdv_work;
        if (dc.ishock_heating > 0) fxyz4 += fac * dudtdiss;
        fxyz4 += fac * dendtdiss;
        if (DUST) fxyz4 += 0.5 * fac * rho1i * dudtdust;
      }
      if (MHD) {
        double *dB = out.dBevol + (size_t)o * 4;
        dB[0] = dBx_s; dB[1] = dBy_s; dB[2] = dBz_s;
        if (dc.psidecayfac > 0.) {
          const double vclean = dc.overcleanfac * vwavei;
          const double dtau = dc.psidecayfac * vclean * hi1;
          dB[3] = -vclean * divBdiff * rho1i - psii * dtau - 0.5 * psii * divvi;
          dtclean = dc.C_cour * hi / (vclean + PH_TINY64);
        }
      }
      if (DUST) {
        const double om = 1. - epsi;
        out.ddustevol[o] = 0.5 * (ddust * rho1i / (om * om));
        out.deltav[(size_t)o * 3 + 0] = dvx_s;
        out.deltav[(size_t)o * 3 + 1] = dvy_s;
        out.deltav[(size_t)o * 3 + 2] = dvz_s;
        if (out.tstop) out.tstop[o] = tsi;
        if (epsi > 0. && csi > 0. && epsi > PH_EPS64) {
          const double tseff = (1. - epsi) / epsi * (epsi * tsi);
          const double den = epsi * tseff * csi * csi;
          if (den > PH_TINY64) dtdusti = dc.C_force * hi * hi / den;
        }
      }
      const double vsigdtc = fmax(vsigmax, vwavei);
      if (vsigdtc > PH_TINY64) dtc = dc.C_cour * hi / (vsigdtc * fmax(dc.alpha, 1.0));
      if (THERMAL) {
        if (eni + dtc * fxyz4 < PH_EPS64 && eni > PH_EPS64) fxyz4 = fxyz4 / (1. - dtc * fxyz4 / eni);
        fo[3] = fxyz4;
      }
      if (fabs(f2i) > PH_EPS64) dtf = dc.C_force * sqrt(hi / sqrt(f2i));
      if (!(dtc > PH_TINY64) || !(dtc < PH_HUGE64)) atomicOr(flags, FFL_DTC);
      dtc_o = dtc;
      dtf_o = fmin(fmin(dtf, dtdusti), dtclean);
      if (ind) {
        // force.F90:3272-3310: the particle's own timestep and its new bin
        double dti = dtc;
        const double dtitmp = fmin(fmin(dtf, dtclean), dtdusti);
        if (dtitmp < dti + PH_TINY64 && dtitmp < dc.dtmax) dti = dtitmp;
        int b = get_newbin_d(dti, dc.dtmax, (int)out.ibin[o], icall < 2, out.nbinmax, out.istepfrac);
        b = max(b, ibin_neighi - 1);  // Saitoh-Makino limiter
        out.ibin[o] = (int8_t)b;
        atomicMax(&flags[1], b);
        dtc_o = PH_HUGE64;
        dtf_o = PH_HUGE64;
      }
  }
  unsigned long long np64 = (unsigned long long)nl;
  for (int off = 32; off > 0; off >>= 1) np64 += __shfl_xor(np64, off, 64);
  // global minima (force.F90:3312-3327): positive doubles order as integers
  unsigned long long a = __double_as_longlong(dtc_o), b = __double_as_longlong(dtf_o);
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long a2 = __shfl_xor(a, off, 64), b2 = __shfl_xor(b, off, 64);
    a = a < a2 ? a : a2;
    b = b < b2 ? b : b2;
  }
  if (lane == 0) {
    atomicMin(&dtred[0], a);
    atomicMin(&dtred[1], b);
    atomicAdd(&dtred[2], np64);
  }
}

__global__ void k_force_flags(DevConst dc, int nlive, const int32_t *__restrict__ iparts,
                              const int8_t *__restrict__ iphase, int32_t *actflag) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nlive) return;
  int ip = iparts[p];
  int o = abs(ip) - 1;
  actflag[p] = (ip > 0) && (iphase == nullptr || iphase[o] > 0);
}

// node hmax from the current smoothing lengths: leaves first, then every
// ancestor takes the maximum (positive doubles order as their bit patterns)
__global__ void k_leaf_hmax(int nleaves, const int32_t *leaves, const int32_t *nd_lo, const int32_t *nd_hi,
                            const phsph_kdnode *nd, const double *h, unsigned long long *hmaxn) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nleaves) return;
  int c = leaves[l];
  double hm = 0.;
  for (int p = nd_lo[c]; p <= nd_hi[c]; p++) hm = fmax(hm, h[p]);
  unsigned long long v = __double_as_longlong(hm);
  int n = c;
  while (n != 0) {
    atomicMax(&hmaxn[n], v);
    n = nd[n].parent;
  }
}
__global__ void k_zero_u64(unsigned long long *a, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0ull;
}

// ------------------------------------------------------------------------
// cons2prim_everything for ieos 1 and 2 (cons2prim.f90:274-456)
// ------------------------------------------------------------------------
__global__ void k_cons2prim(DevConst dc, int np, const double *__restrict__ xyzh, int ldx,
                            const double *__restrict__ vxyzu, int ldv, const float *__restrict__ dvdx,
                            float *__restrict__ alphaind, const double *__restrict__ dustevol,
                            double *__restrict__ dustfrac, double *__restrict__ eos, int ldeos) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const double hi = xyzh[(size_t)i * ldx + 3];
  if (hi < PH_TINY64) return;
  const double rhoi = rhoh_d(hi, dc.hfact, dc.pmass);
  double rhogas = rhoi;
  if (dc.dust && dustevol) {
    const double sv = dustevol[i];
    const double e = sv * sv / (1. + sv * sv);
    dustfrac[i] = e;
    rhogas = rhoi * (1. - e);
  }
  double ponrho, cs;
  if (dc.ieos == 1) {  // eos.f90:184-194
    ponrho = dc.polyk;
    cs = sqrt(ponrho);
  } else {  // eos.f90:196-240
    if (dc.maxvxyzu >= 4) {
      const double u = vxyzu[(size_t)i * ldv + 3];
      ponrho = (dc.gamma > 1.0001) ? (dc.gamma - 1.) * u : 2. / 3. * u;
    } else {
      ponrho = dc.polyk * pow(rhogas, dc.gamma - 1.);
    }
    cs = sqrt(dc.gamma * ponrho);
  }
  eos[(size_t)i * ldeos + 0] = ponrho * rhogas;
  eos[(size_t)i * ldeos + 1] = cs;
  if (ldeos > 2) eos[(size_t)i * ldeos + 2] = 0.;
  if (alphaind && dvdx) {
    // xi_limiter and get_alphaloc (shock_capturing.f90:131-178)
    const float *d = dvdx + (size_t)i * 9;
    const double divv = (double)d[0] + (double)d[4] + (double)d[8];
    const double cx = (double)d[7] - (double)d[5], cy = (double)d[2] - (double)d[6], cz = (double)d[3] - (double)d[1];
    const double m = fmax(-divv, 0.);
    const double fac = m * m;
    const double tr = cx * cx + cy * cy + cz * cz;
    const double xil = (fac + tr > PH_EPS64) ? fac / (fac + tr) : 1.;
    const double divvdt = (double)alphaind[(size_t)i * 3 + 2];
    const double source = 10. * (hi * hi) * xil * fmax(-divvdt, 0.);
    const double temp = cs * cs;
    double aloc = dc.alpha;
    if (temp > PH_EPS64) aloc = fmax(fmin(source / temp, dc.alphamax), dc.alpha);
    alphaind[(size_t)i * 3 + 1] = (float)aloc;
  }
}

int cons2prim_run(phsph_ctx *c, int np, const double *xyzh, int ldx, const double *vxyzu, int ldv,
                  const float *dvdx, float *alphaind, const double *dustevol, double *dustfrac, double *eos,
                  int ldeos) {
  LaunchTimer lt(c, "cons2prim");
  hipLaunchKernelGGL(k_cons2prim, dim3(nblocks(np, 256)), dim3(256), 0, c->stream, c->dc, np, xyzh, ldx, vxyzu,
                     ldv, dvdx, alphaind, dustevol, dustfrac, eos, ldeos);
  PH_HIPCHECK(c, hipGetLastError());
  return PHSPH_OK;
}

// ------------------------------------------------------------------------
int force_run(phsph_ctx *c, int icall, const ForceIOArgs &a, double *dtc_out, double *dtf_out) {
  (void)icall;
  hipStream_t s = c->stream;
  const int n = c->nlive;
  const bool mhd = c->p.mhd, dust = c->p.use_dustfrac && c->p.ndustsmall > 0;
  PH_HIPCHECK(c, c->frec.ensure((size_t)n * sizeof(FRec)));
  if (mhd) PH_HIPCHECK(c, c->frecB.ensure((size_t)n * sizeof(FRecB)));
  if (dust) PH_HIPCHECK(c, c->frecD.ensure((size_t)n * sizeof(FRecD)));
  PH_HIPCHECK(c, c->fh.ensure(n));
  FRec *rec = (FRec *)c->frec.p;
  FRecB *recB = mhd ? (FRecB *)c->frecB.p : nullptr;
  FRecD *recD = dust ? (FRecD *)c->frecD.p : nullptr;
  if (a.ibin) PH_HIPCHECK(c, c->tbin_old.ensure(n));
  if (c->p.gravity) PH_HIPCHECK(c, c->gsoft.ensure(n));
  {
    LaunchTimer lt(c, "force_prep");
    hipLaunchKernelGGL(k_force_prep, dim3(nblocks(n, 256)), dim3(256), 0, s, c->dc, n, c->iparts.p, c->tx.p,
                       c->ty.p, c->tz.p, a.xyzh, a.ldx, a.vxyzu, a.ldv, a.gradh, a.ldgradh, a.alphaind, a.eos_vars,
                       a.ldeos, a.Bevol, a.dustfrac, c->fh.p, rec, recB, recD, a.ibin ? a.ibin_old : nullptr,
                       c->tbin_old.p, c->p.gravity ? c->gsoft.p : nullptr);
  }
  PH_HIPCHECK(c, hipGetLastError());
  PH_HIPCHECK(c, c->tmpi.ensure(n));
  hipLaunchKernelGGL(k_force_flags, dim3(nblocks(n, 256)), dim3(256), 0, s, c->dc, n, c->iparts.p, a.iphase,
                     c->tmpi.p);
  int rc = compact_active(c, c->tmpi.p, n);
  if (rc) return rc;
  const int nact = c->nact;
  if (c->p.gravity) return grav_force(c, icall, a, rec, c->gsoft.p, dtc_out, dtf_out);
  // node hmax of the final smoothing lengths for the h_j search radius
  PH_HIPCHECK(c, c->hmaxf.ensure(c->ncells + 2));
  hipLaunchKernelGGL(k_zero_u64, dim3(nblocks(c->ncells + 2, 256)), dim3(256), 0, s,
                     (unsigned long long *)c->hmaxf.p, (int64_t)(c->ncells + 2));
  hipLaunchKernelGGL(k_leaf_hmax, dim3(nblocks(c->nleaves, 256)), dim3(256), 0, s, c->nleaves, c->leaves.p,
                     c->nd_lo.p, c->nd_hi.p, c->nd.p, c->fh.p, (unsigned long long *)c->hmaxf.p);
  rc = walk_groups(c, 1, 1.0, c->fh.p, c->hmaxf.p);
  if (rc) return rc;
  const int ng = c->ngroups;
  PH_HIPCHECK(c, c->dtred.ensure(3));
  {
    unsigned long long init[3];
    double big = PH_HUGE64;
    memcpy(&init[0], &big, 8);
    init[1] = init[0];
    init[2] = 0ull;
    PH_HIPCHECK(c, hipMemcpyAsync(c->dtred.p, init, sizeof(init), hipMemcpyHostToDevice, s));
    int32_t zero2[2] = {0, 0};
    PH_HIPCHECK(c, hipMemcpyAsync(c->counters.p, zero2, sizeof(zero2), hipMemcpyHostToDevice, s));
    if (nact > 0) {
      LaunchTimer lt(c, "force");
      const dim3 gr(xcd_grid(ng)), b(64);
      const bool th = c->p.maxvxyzu >= 4;
#define FL(KK, M, D, T)                                                                                        \
  hipLaunchKernelGGL((k_force<KK, M, D, T>), gr, b, 0, s, c->dc, c->act.p, nact, ng, c->pleaf.p, c->leaves.p,    \
                     c->gcand.p, c->gncand.p, c->caps.gcand, c->nd_lo.p, c->nd_hi.p, c->nd.p, c->iparts.p,     \
                     rec, recB, recD, a, c->tbin_old.p, icall, c->dtred.p, c->counters.p)
#define FLK(KK)                                          \
  if (mhd && dust) { if (th) FL(KK, true, true, true); else FL(KK, true, true, false); }     \
  else if (mhd) { if (th) FL(KK, true, false, true); else FL(KK, true, false, false); }      \
  else if (dust) { if (th) FL(KK, false, true, true); else FL(KK, false, true, false); }     \
  else { if (th) FL(KK, false, false, true); else FL(KK, false, false, false); }
      if (c->p.kernel == PHSPH_KERNEL_CUBIC) { FLK(PHSPH_KERNEL_CUBIC) } else { FLK(PHSPH_KERNEL_QUINTIC) }
#undef FLK
#undef FL
    }
    PH_HIPCHECK(c, hipGetLastError());
    int32_t fl2[2] = {0, 0};
    unsigned long long red[3];
    PH_HIPCHECK(c, hipMemcpyAsync(fl2, c->counters.p, sizeof(fl2), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipMemcpyAsync(red, c->dtred.p, sizeof(red), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    const int32_t fl = fl2[0];
    if (fl & FFL_DTC) PH_RET(c, PHSPH_ERR_INPUT, "force: invalid dtc");
    if (a.ibin) c->nbinmax_new = fl2[1];
    double dtc, dtf;
    memcpy(&dtc, &red[0], 8);
    memcpy(&dtf, &red[1], 8);
    c->force_npairs = (int64_t)red[2];
    if (a.ibin) dtc = dtf = PH_BIGNUMBER;  // per-particle bins instead (force.F90:3312)
    if (dtc_out) *dtc_out = (nact > 0) ? dtc : PH_BIGNUMBER;
    if (dtf_out) *dtf_out = (nact > 0) ? fmin(dtf, PH_BIGNUMBER) : PH_BIGNUMBER;
    return PHSPH_OK;
  }
}

// ------------------------------------------------------------------------
// set_active_particles (utils_indtimesteps.f90:114-185)
// ------------------------------------------------------------------------
__global__ void k_set_active(int np, const double *__restrict__ xyzh, int ldx, int8_t *__restrict__ iphase,
                             const int8_t *__restrict__ ibin, int nbinmax, int istepfrac, int32_t *__restrict__ acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int na = 0, nl = 0, bad = 0;
  if (i < np && xyzh[(size_t)i * ldx + 3] > 0.) {
    nl = 1;
    const int b = ibin[i];
    if (b > nbinmax || b < 0) {
      bad = 1;
    } else {
      const bool act = (istepfrac % (1 << (nbinmax - b))) == 0;
      na = act ? 1 : 0;
      iphase[i] = act ? (int8_t)1 : (int8_t)-1;  // isetphase(igas, iactive)
    }
  }
  na = wave_sum_i(na);
  nl = wave_sum_i(nl);
  bad = wave_sum_i(bad);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&acc[0], na);
    atomicAdd(&acc[1], nl);
    atomicAdd(&acc[2], bad);
  }
}

int set_active_run(phsph_ctx *c, int np, const double *xyzh, int ldx, int8_t *iphase, const int8_t *ibin,
                   int nbinmax, int istepfrac, int32_t *nactive, int32_t *nalive) {
  hipStream_t s = c->stream;
  PH_HIPCHECK(c, c->counters.ensure(64));
  int32_t z[3] = {0, 0, 0};
  PH_HIPCHECK(c, hipMemcpyAsync(c->counters.p, z, sizeof(z), hipMemcpyHostToDevice, s));
  if (np > 0) {
    LaunchTimer lt(c, "set_active");
    hipLaunchKernelGGL(k_set_active, dim3(nblocks(np, 256)), dim3(256), 0, s, np, xyzh, ldx, iphase, ibin, nbinmax,
                       istepfrac, c->counters.p);
  }
  PH_HIPCHECK(c, hipGetLastError());
  int32_t acc[3];
  PH_HIPCHECK(c, hipMemcpyAsync(acc, c->counters.p, sizeof(acc), hipMemcpyDeviceToHost, s));
  PH_HIPCHECK(c, hipStreamSynchronize(s));
  if (acc[2] > 0) PH_RET(c, PHSPH_ERR_INPUT, "set_active_particles: timestep bin exceeds max bins");
  *nactive = acc[0];
  *nalive = acc[1];
  return PHSPH_OK;
}
