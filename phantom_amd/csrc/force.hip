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
// k_flist (one wavefront per group of 64 active particles, one particle per
// lane) collects each particle's symmetric neighbour set
// {j : q_i < R or q_j < R} (force.F90:1287) from the group's candidate leaves
// with the same separately rounded FP64 tests as the reference; k_force (one
// thread per particle) evaluates the pair terms over that list.  Quantities
// of particle j that the reference recomputes per pair (rho_j, P/rho^2, wave
// speed, MHD stress, stopping time) are evaluated once per particle by
// k_force_prep with the same arithmetic and packed into one 128-byte record
// per particle, so a pair reads one cache line of j data.
#include "records.h"

enum { FFL_LIST = 2, FFL_DTC = 32 };

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
                             float *__restrict__ gsoft, double4 *__restrict__ fpos, double4 *__restrict__ fvel,
                             FRec32 *__restrict__ rec32) {
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
  fpos[p] = make_double4(R.x, R.y, R.z, R.hinv);
  if (rec32) {
    FRec32 A;
    A.rho1 = (float)R.rho1;
    A.pro2 = (float)R.pro2;
    A.pro2lo = (float)(R.pro2 - (double)A.pro2);
    A.vwave = (float)R.vwave;
    A.alpha = R.alpha; A.gradh = R.gradh; A.rho = (float)R.rho;
    A.pr = (float)R.pr;
    A.prlo = (float)(R.pr - (double)A.pr);
    A.pad0 = A.pad1 = A.pad2 = 0.f;
    rec32[p] = A;
    fvel[p] = make_double4(R.vx, R.vy, R.vz, R.u);
  }
  if (dc.dust) {
    FRecD D;
    const double e = dustfrac[o];
    D.eps = e;
    D.sqeps = sqrt(e / (1. - e));
    double tsj = get_ts_d(dc, rhoj * (1. - e), rhoj * e, csj);
    if (dc.ilimitdustflux) tsj = fmin(tsj, hj / csj);
    D.ts = tsj;
    D.pad = 0.;
    recD[p] = D;
  }
}

// Force, part 1: the symmetric neighbour set {j : q_i < R or q_j < R}
// (force.F90:1287) of every active particle, with the reference's separately
// rounded FP64 tests.  One wavefront per group of 64 active particles; each
// lane visits the candidate leaves it can reach, a leaf's particles lying
// within its size of its centre and having h at most the leaf's hmax.
template <int CL, int LW>
__global__ void __launch_bounds__(64 * LW)
k_flist(DevConst dc, const int32_t *__restrict__ act, int nact, int ngroups, const int32_t *__restrict__ pleaf,
        const int32_t *__restrict__ leaves, const uint32_t *__restrict__ gcand, const int32_t *__restrict__ gncand,
        const int64_t *goff, int capg, const int32_t *__restrict__ nd_lo, const int32_t *__restrict__ nd_hi,
        const phsph_kdnode *__restrict__ nd, const double *__restrict__ hmaxn, const double4 *__restrict__ fpos,
        uint32_t *__restrict__ nlist, const int64_t *loff, int ucap, int32_t *__restrict__ t_nl,
        int32_t *__restrict__ flags) {
  __shared__ ListStage<CL> Sw[LW];
  ListStage<CL> &S = Sw[threadIdx.x >> 6];
  const int g = list_group<LW>(gridDim.x);
  if (g >= ngroups) return;
  const int lane = threadIdx.x & 63;
  const int t = g * GROUP + lane;
  const bool live = t < nact && act[t] >= 0;  // act[t] < 0: padding of a partial group
  const double R2 = dc.radkern2;
  int p = -1;
  double xi = 0., yi = 0., zi = 0., hi = 0., hi1 = 1., lcx = 0., lcy = 0., lcz = 0., lsz = 0.;
  if (live) {
    p = act[t];
    const int c = leaves[pleaf[p]];
    const double4 r = fpos[p];
    xi = r.x; yi = r.y; zi = r.z;
    hi1 = r.w;       // 1/h as the pair kernel computes it
    hi = 1. / hi1;   // search radius only (the reach test has a margin)
    lcx = nd[c].xcen[0]; lcy = nd[c].xcen[1]; lcz = nd[c].xcen[2]; lsz = nd[c].size;
  }
  const double hi21 = xmul(hi1, hi1);
  const int capl = list_cap_g(loff, g * GROUP, ucap);
  int nl = 0;
  build_list(
      S, dc, gcand + cand_base(goff, g, capg), gncand[g], nd, nd_lo, nd_hi, hmaxn, fpos, xi, yi, zi, lcx, lcy, lcz,
      lsz,
      [&](double ccx, double ccy, double ccz, double csz, double chm) {
        return live && leaf_in_reach(dc, xi, yi, zi, ccx, ccy, ccz, csz, dc.radkern * fmax(hi, chm));
      },
      [&](int j, int code, double r2, double hj1) {
        if (j != p && (xmul(r2, hi21) < R2 || xmul(r2, xmul(hj1, hj1)) < R2)) {
          if (nl < capl) nlist[list_at_g(loff, nact, t, nl)] = cand_pack(j, code);
          nl++;
        }
      });
  if (t < nact) t_nl[t] = nl;  // 0 for padding
  if (!live) return;
  if (nl > capl) atomicOr(flags, FFL_LIST);
}

// Force, part 2: the pair terms (force.F90:1227-2056) over each particle's
// list and finish_cell_and_store_results.  One thread per active particle.
template <int K, bool MHD, bool DUST, bool THERMAL>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu((MHD || DUST) ? 2 : (THERMAL ? 3 : 4))))
k_force(DevConst dc, const int32_t *__restrict__ act, int nact, const int32_t *__restrict__ skip,
        const uint32_t *__restrict__ nlist, const int64_t *loff, const int32_t *__restrict__ t_nl,
        const int32_t *__restrict__ t_nlj, int capb, const int32_t *__restrict__ iparts,
        const double4 *__restrict__ fpos, const FRec *__restrict__ rec,
        const FRecB *__restrict__ recB, const FRecD *__restrict__ recD, ForceIOArgs out,
        const int8_t *__restrict__ tbin_old, int icall, unsigned long long *__restrict__ dtred,
        int32_t *__restrict__ flags) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  // act[t] < 0: padding; skip: particles whose lists are elsewhere
  const bool live = t < nact && act[t] >= 0 && !(skip && skip[act[t]]);
  const int p = live ? act[t] : act[0];
  const double pm = dc.pmass;
  const double R2 = dc.radkern2;
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
    tsi = get_ts_d(dc, rhoi * (1. - epsi), rhoi * epsi, csi);
    if (dc.ilimitdustflux) tsi = fmin(tsi, hi / csi);
    epstsi = epsi * tsi;
    sqepsi = sqrt(epsi / (1. - epsi));
  }
  const double rho21i = rho1i * rho1i;
  const double mrhoi5 = 0.5 * pm * rho1i;
  const double auterm = mrhoi5 * dc.alphau;
  const double avBterm = mrhoi5 * dc.alphaB * rho1i;
  // --- pair terms over the neighbour list (force.F90:1227-2056) ---
  double fx = 0., fy = 0., fz = 0., drhodt = 0., dudtdiss = 0., dendtdiss = 0.;
  double divBsym = 0., dBx_s = 0., dBy_s = 0., dBz_s = 0., divBdiff = 0.;
  double ddust = 0., dudtdust = 0., dvx_s = 0., dvy_s = 0., dvz_s = 0.;
  double vsigmax = 0.;
  const bool usej_always = MHD || DUST || (THERMAL && !dc.gravity);
  const bool ind = out.ibin != nullptr;
  int ibin_neighi = 0;
  const int8_t ibinnow_m1 = (int8_t)(out.ibinnow - 1);
  // the list's front entries, then (density lists) its back entries
  const int nl = live ? t_nl[t] : 0;
  const int ntot = nl + ((live && t_nlj) ? t_nlj[t] : 0);
  int npair = 0;  // interacting pairs (entries that pass the neighbour test)
  // Three ways to read the neighbour records ahead (same pairs, same order,
  // same arithmetic):
  //  - hydro, cubic kernel (TEST_FIRST): entry kk+1's neighbour test is made
  //    from its 32-byte position record while entry kk is summed, and the
  //    other 96 bytes of its record are read only when it is a force
  //    neighbour (a density list also holds particles that are not); entry
  //    kk+2's position record and list entry kk+3 are read at the same time;
  //  - hydro, quintic kernel (WHOLE): the whole 128-byte record of entry kk+1
  //    and list entry kk+2 are read while entry kk is summed (its pairs cost
  //    more arithmetic, and the second neighbour test costs more than the
  //    reads it saves);
  //  - MHD and dust (more registers live): entry kk+1's position record
  //    ahead, the rest of the record after the neighbour test.
  constexpr bool WHOLE = !MHD && !DUST;
  constexpr bool TEST_FIRST = WHOLE && K == PHSPH_KERNEL_CUBIC;
  auto slot = [&](int kk) { return kk < nl ? kk : capb - 1 - (kk - nl); };
  auto is_pair = [&](uint32_t e, const double4 &pj) {
    double dx, dy, dz;
    sep_of(dc, cand_code(e), xi, yi, zi, pj.x, pj.y, pj.z, dx, dy, dz);
    const double rij2 = r2_exact(dx, dy, dz);
    return xmul(rij2, hi21) < R2 || xmul(rij2, xmul(pj.w, pj.w)) < R2;
  };
  auto tail_of = [&](int j) { return *reinterpret_cast<const FTail *>(&rec[j].vx); };
  uint32_t en = 0u, en2 = 0u, en3 = 0u;
  double4 pn = make_double4(0., 0., 0., 0.), pn2 = pn;
  FTail Tn;
  FRec Rn;
  bool okn = false;
  if (ntot > 0) {
    en = nlist[list_at_g(loff, nact, t, slot(0))];
    if constexpr (TEST_FIRST) {
      pn = fpos[cand_node(en)];
      okn = is_pair(en, pn);
      if (okn) Tn = tail_of(cand_node(en));
      if (ntot > 1) {
        en2 = nlist[list_at_g(loff, nact, t, slot(1))];
        pn2 = fpos[cand_node(en2)];
      }
      if (ntot > 2) en3 = nlist[list_at_g(loff, nact, t, slot(2))];
    } else if constexpr (WHOLE) {
      Rn = rec[cand_node(en)];
    } else {
      pn = fpos[cand_node(en)];
    }
  }
  if (!TEST_FIRST && ntot > 1) en2 = nlist[list_at_g(loff, nact, t, slot(1))];
  for (int kk = 0; kk < ntot; kk++) {
    const uint32_t e = en;
    const int j = cand_node(e);
    double4 pj;
    FTail Tw;
    if constexpr (TEST_FIRST) {
      pj = pn;
      const bool okw = okn;
      Tw = Tn;
      if (kk + 1 < ntot) {
        en = en2;
        pn = pn2;
        okn = is_pair(en, pn);
        if (okn) Tn = tail_of(cand_node(en));
        if (kk + 2 < ntot) {
          en2 = en3;
          pn2 = fpos[cand_node(en2)];
          if (kk + 3 < ntot) en3 = nlist[list_at_g(loff, nact, t, slot(kk + 3))];
        }
      }
      if (!okw) continue;
    } else if constexpr (WHOLE) {
      const FRec Rw = Rn;
      pj = make_double4(Rw.x, Rw.y, Rw.z, Rw.hinv);
      Tw = *reinterpret_cast<const FTail *>(&Rw.vx);
      if (kk + 1 < ntot) {
        en = en2;
        Rn = rec[cand_node(en)];
        if (kk + 2 < ntot) en2 = nlist[list_at_g(loff, nact, t, slot(kk + 2))];
      }
    } else {
      pj = pn;
      if (kk + 1 < ntot) {
        en = en2;
        pn = fpos[cand_node(en)];
        if (kk + 2 < ntot) en2 = nlist[list_at_g(loff, nact, t, slot(kk + 2))];
      }
    }
    double dx, dy, dz;
    sep_of(dc, cand_code(e), xi, yi, zi, pj.x, pj.y, pj.z, dx, dy, dz);
    const double rij2 = r2_exact(dx, dy, dz);
    const double q2i = xmul(rij2, hi21);
    const double hj1 = pj.w;
    const double hj21 = xmul(hj1, hj1);
    const double q2j = xmul(rij2, hj21);
    // a density list also holds particles that are not force neighbours
    if (!(q2i < R2 || q2j < R2)) continue;
    npair++;
    const FTail &Rj = WHOLE ? Tw : *reinterpret_cast<const FTail *>(&rec[j].vx);
    if (ind) {
      // force.F90:1346-1352: every neighbour of an active particle is
      // flagged to be woken; the limiter needs the largest neighbour bin
      ibin_neighi = max(ibin_neighi, (int)tbin_old[j]);
      const int oj = abs(iparts[j]) - 1;
      if (out.ibin_wake[oj] < ibinnow_m1) out.ibin_wake[oj] = ibinnow_m1;
    }
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
      grkernj = kern_g<K, double>(q2j, qj) * hj21 * hj21 * dc.cnormk * (double)Rj.gradh;
      usej = true;
    }
    if (usej_always) usej = true;
    const double runix = dx * rij1, runiy = dy * rij1, runiz = dz * rij1;
    const double dvx = vxi - Rj.vx, dvy = vyi - Rj.vy, dvz = vzi - Rj.vz;
    const double projv = dvx * runix + dvy * runiy + dvz * runiz;
    const double denij = THERMAL ? (eni - Rj.u) : 0.;
    const double vsigi = fmax(vwavei - dc.beta * projv, 0.);
    const double vsigavi = fmax(alphai * vwavei - dc.beta * projv, 0.);
    vsigmax = fmax(vsigmax, vsigi);
    double rho1j = 0., rho21j = 0., pro2j = 0., vwavej = 0., vsigavj = 0., autermj = 0., avBtermj = 0.;
    double rhoj = 0., prj = 0.;
    if (usej) {
      rhoj = Rj.rho;
      rho1j = Rj.rho1;
      rho21j = rho1j * rho1j;
      prj = Rj.pr;
      pro2j = Rj.pro2;
      vwavej = Rj.vwave;
      const double mrhoj5 = 0.5 * pm * rho1j;
      autermj = mrhoj5 * dc.alphau;
      avBtermj = mrhoj5 * dc.alphaB * rho1j;
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
    double dendissterm = 0.;
    if (THERMAL) {
      double vsigu;
      if (dc.gravity) vsigu = fabs(projv);
      else vsigu = sqrt(fabs(pri - prj) * (2. / (rhoi + rhoj)));
      dendissterm = vsigu * denij * (auterm * grkerni + autermj * grkernj);
    }
    double projsx = 0., projsy = 0., projsz = 0., dudtresist = 0.;
    if (MHD) {
      const FRecB &Mj = recB[j];
      const double Bxj = Mj.Bx, Byj = Mj.By, Bzj = Mj.Bz, psij = Mj.psi;
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
               (Mj.sxx * runix + Mj.sxy * runiy + Mj.sxz * runiz) * grkernj;
      projsy = (sxyi * runix + syi * runiy + syzi * runiz) * grkerni +
               (Mj.sxy * runix + Mj.syy * runiy + Mj.syz * runiz) * grkernj;
      projsz = (sxzi * runix + syzi * runiy + szi * runiz) * grkerni +
               (Mj.sxz * runix + Mj.syz * runiy + Mj.szz * runiz) * grkernj;
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
      const FRecD &Dj = recD[j];
      const double epsj = Dj.eps;
      const double tsj = Dj.ts;
      const double epstsj = epsj * tsj;
      if (epsi > 0. || epsj > 0.) {
        const double grkernav = 0.5 * (grkerni + grkernj);
        const double dft = pm * Dj.sqeps * rho1j *
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
  // --- finish_cell_and_store_results ---
  double dtc_o = PH_HUGE64, dtf_o = PH_HUGE64;
  int bmax = -1;
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
        bmax = b;
        dtc_o = PH_HUGE64;
        dtf_o = PH_HUGE64;
      }
  }
  // global minima (force.F90:3312-3327), pair count, highest bin
  reduce_force_block(dtc_o, dtf_o, npair, bmax, dtred, flags);
}

// Mixed precision, hydro (grad-h, Cullen-Dehnen viscosity, conductivity).
// The neighbour tests, separations and q^2 are those of k_force (the same
// neighbour sets and kernel branches).  The j quantities other than
// position, velocity and u are stored in FP32 (P/rho^2 and P as FP32 pairs
// carrying FP64 accuracy), 1/r is an FP32 reciprocal square root refined by
// one FP64 Newton step, and the conductivity signal speed takes its square
// root and division in FP32; the rest is FP64, because the pressure and
// viscous terms cancel to a small net force in smooth flow.  j data: 32-byte
// FP64 position and velocity records and the 48-byte FRec32 instead of the
// 128-byte record.
template <int K, bool THERMAL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(THERMAL ? 3 : 4)))
k_force_mixed(DevConst dc, const int32_t *__restrict__ act, int nact, const int32_t *__restrict__ skip,
              const uint32_t *__restrict__ nlist, const int64_t *loff, const int32_t *__restrict__ t_nl,
              const int32_t *__restrict__ t_nlj, int capb, const int32_t *__restrict__ iparts, const FRec *__restrict__ rec,
              const double4 *__restrict__ fpos, const double4 *__restrict__ fvel, const FRec32 *__restrict__ rec32,
              ForceIOArgs out, const int8_t *__restrict__ tbin_old, int icall,
              unsigned long long *__restrict__ dtred, int32_t *__restrict__ flags) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  // act[t] < 0: padding; skip: particles whose lists are elsewhere
  const bool live = t < nact && act[t] >= 0 && !(skip && skip[act[t]]);
  const int p = live ? act[t] : act[0];
  const double pm = dc.pmass;
  const double R2 = dc.radkern2;
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
  const double vwavei = Ri.cs;
  const double pri = Ri.pr, eni = Ri.u;
  const double pro2i = pri * rho1i * rho1i;
  const double vxi = Ri.vx, vyi = Ri.vy, vzi = Ri.vz;
  const double alphai = (double)Ri.alpha;
  // FP32 copies of the i terms of the conductivity
  const float pmf = (float)pm, rhoif = (float)rhoi;
  const float autermf = (float)(0.5 * pm * rho1i * dc.alphau), alphauf = (float)dc.alphau;
  double fx = 0., fy = 0., fz = 0., drhodt = 0., dudtdiss = 0., dendtdiss = 0.;
  double vsigmax = 0.;
  const bool ind = out.ibin != nullptr;
  int ibin_neighi = 0;
  const int8_t ibinnow_m1 = (int8_t)(out.ibinnow - 1);
  // the list's front entries, then (density lists) its back entries
  const int nl = live ? t_nl[t] : 0;
  const int ntot = nl + ((live && t_nlj) ? t_nlj[t] : 0);
  int npair = 0;  // interacting pairs (entries that pass the neighbour test)
  // Cubic kernel (TEST_FIRST): entry kk+1's neighbour test is made from its
  // position record while entry kk is summed, and its FP32 record and
  // velocity record are read then only when it is a force neighbour; entry
  // kk+2's position record and list entry kk+3 are read at the same time.
  // Quintic kernel: entry kk+1's position record ahead, the rest after the
  // test (its pairs carry more arithmetic, and the second test costs more
  // than the reads it saves).
  constexpr bool TEST_FIRST = K == PHSPH_KERNEL_CUBIC;
  auto slot = [&](int kk) { return kk < nl ? kk : capb - 1 - (kk - nl); };
  auto is_pair = [&](uint32_t e, const double4 &pj) {
    double sx, sy, sz;
    sep_of(dc, cand_code(e), xi, yi, zi, pj.x, pj.y, pj.z, sx, sy, sz);
    const double s2 = r2_exact(sx, sy, sz);
    return xmul(s2, hi21) < R2 || xmul(s2, xmul(pj.w, pj.w)) < R2;
  };
  uint32_t en = 0u, en2 = 0u, en3 = 0u;
  double4 pn = make_double4(0., 0., 0., 0.), pn2 = pn, vn = pn;
  FRec32 An;
  bool okn = false;
  if (ntot > 0) {
    en = nlist[list_at_g(loff, nact, t, slot(0))];
    pn = fpos[cand_node(en)];
    if constexpr (TEST_FIRST) {
      okn = is_pair(en, pn);
      if (okn) {
        An = rec32[cand_node(en)];
        vn = fvel[cand_node(en)];
      }
      if (ntot > 1) {
        en2 = nlist[list_at_g(loff, nact, t, slot(1))];
        pn2 = fpos[cand_node(en2)];
      }
      if (ntot > 2) en3 = nlist[list_at_g(loff, nact, t, slot(2))];
    } else if (ntot > 1) {
      en2 = nlist[list_at_g(loff, nact, t, slot(1))];
    }
  }
  for (int kk = 0; kk < ntot; kk++) {
    const uint32_t e = en;
    const int j = cand_node(e);
    const double4 pj = pn;
    FRec32 A;
    double4 vj;
    if constexpr (TEST_FIRST) {
      const bool okw = okn;
      A = An;
      vj = vn;
      if (kk + 1 < ntot) {
        en = en2;
        pn = pn2;
        okn = is_pair(en, pn);
        if (okn) {
          An = rec32[cand_node(en)];
          vn = fvel[cand_node(en)];
        }
        if (kk + 2 < ntot) {
          en2 = en3;
          pn2 = fpos[cand_node(en2)];
          if (kk + 3 < ntot) en3 = nlist[list_at_g(loff, nact, t, slot(kk + 3))];
        }
      }
      if (!okw) continue;
    } else {
      if (kk + 1 < ntot) {
        en = en2;
        pn = fpos[cand_node(en)];
        if (kk + 2 < ntot) en2 = nlist[list_at_g(loff, nact, t, slot(kk + 2))];
      }
    }
    double dx, dy, dz;
    sep_of(dc, cand_code(e), xi, yi, zi, pj.x, pj.y, pj.z, dx, dy, dz);
    const double rij2 = r2_exact(dx, dy, dz);
    const double q2i = xmul(rij2, hi21);
    const double hj1 = pj.w;
    const double hj21 = xmul(hj1, hj1);
    const double q2j = xmul(rij2, hj21);
    // a density list also holds particles that are not force neighbours
    if (!(q2i < R2 || q2j < R2)) continue;
    npair++;
    if (ind) {  // force.F90:1346-1352
      ibin_neighi = max(ibin_neighi, (int)tbin_old[j]);
      const int oj = abs(iparts[j]) - 1;
      if (out.ibin_wake[oj] < ibinnow_m1) out.ibin_wake[oj] = ibinnow_m1;
    }
    const float r2f = (float)rij2;
    double rij1 = 0.;
    if (r2f > 0.f) {
      const double y = (double)rsqrtf(r2f);
      rij1 = y * (1.5 - 0.5 * rij2 * y * y);
    }
    const double rij = rij2 * rij1;
    const double grkerni = (q2i < R2) ? kern_g<K, double>(q2i, rij * hi1) * hfacgrkern : 0.;
    if constexpr (!TEST_FIRST) A = rec32[j];
    double grkernj = 0.;
    bool usej = THERMAL;
    if (q2j < R2) {
      grkernj = kern_g<K, double>(q2j, rij * hj1) * hj21 * hj21 * dc.cnormk * (double)A.gradh;
      usej = true;
    }
    const double runix = dx * rij1, runiy = dy * rij1, runiz = dz * rij1;
    if constexpr (!TEST_FIRST) vj = fvel[j];
    const double dvx = vxi - vj.x, dvy = vyi - vj.y, dvz = vzi - vj.z;
    const double projv = dvx * runix + dvy * runiy + dvz * runiz;
    // signal speeds and artificial viscosity (force.F90:1416-1427, 1581-1596)
    const double vsigi = fmax(vwavei - dc.beta * projv, 0.);
    const double vsigavi = fmax(alphai * vwavei - dc.beta * projv, 0.);
    vsigmax = fmax(vsigmax, vsigi);
    double rho1j = 0., vsigavj = 0., pro2j = 0.;
    if (usej) {
      rho1j = (double)A.rho1;
      pro2j = (double)A.pro2 + (double)A.pro2lo;
      const double vwavej = (double)A.vwave;
      vsigmax = fmax(vsigmax, fmax(vwavej - dc.beta * projv, 0.));
      vsigavj = fmax((double)A.alpha * vwavej - dc.beta * projv, 0.);
    }
    double qrho2i = 0., qrho2j = 0.;
    if (projv < 0.) {
      qrho2i = -0.5 * rho1i * vsigavi * projv;
      if (usej) qrho2j = -0.5 * rho1j * vsigavj * projv;
    }
    const double gradp = pm * ((pro2i + qrho2i) * grkerni + (usej ? (pro2j + qrho2j) * grkernj : 0.));
    fx -= runix * gradp;
    fy -= runiy * gradp;
    fz -= runiz * gradp;
    drhodt += projv * grkerni;
    if (THERMAL) {
      // conductivity: the square root and division in FP32 on FP64
      // differences (force.F90:1606-1624)
      const double dpr = pri - ((double)A.pr + (double)A.prlo);
      const float vsigu = sqrtf((float)fabs(dpr) * (2.f / (rhoif + A.rho)));
      const float denij = (float)(eni - vj.w);
      dudtdiss += pm * qrho2i * projv * grkerni;
      dendtdiss += (double)(vsigu * denij * (autermf * (float)grkerni + 0.5f * pmf * (float)rho1j * alphauf * (float)grkernj));
    }
  }
  // --- finish_cell_and_store_results (hydro) ---
  double dtc_o = PH_HUGE64, dtf_o = PH_HUGE64;
  int bmax = -1;
  if (live) {
    const int o = abs(iparts[p]) - 1;
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
    if (!(dtc > PH_TINY64) || !(dtc < PH_HUGE64)) atomicOr(flags, FFL_DTC);
    dtc_o = dtc;
    dtf_o = dtf;
    if (ind) {  // force.F90:3272-3310
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
  reduce_force_block(dtc_o, dtf_o, npair, bmax, dtred, flags);
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
  const unsigned long long v = __double_as_longlong(hm);
  hmaxn[c] = v;  // each leaf has one thread
  // A node already at >= v was raised by a thread carrying a value >= v,
  // which goes on raising the nodes above it: stop there.
  for (int n = nd[c].parent; n != 0; n = nd[n].parent)
    if (atomicMax(&hmaxn[n], v) >= v) break;
}
__global__ void k_zero_u64(unsigned long long *a, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0ull;
}

// node hmax (as the bit patterns of the positive doubles) of h in tree order
int leaf_hmax_run(phsph_ctx *c, const double *d_h, double *d_hmaxn) {
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(k_zero_u64, dim3(nblocks(c->ncells + 2, 256)), dim3(256), 0, s, (unsigned long long *)d_hmaxn,
                     (int64_t)(c->ncells + 2));
  if (c->nleaves > 0)
    hipLaunchKernelGGL(k_leaf_hmax, dim3(nblocks(c->nleaves, 256)), dim3(256), 0, s, c->nleaves, c->leaves.p,
                       c->nd_lo.p, c->nd_hi.p, c->nd.p, d_h, (unsigned long long *)d_hmaxn);
  PH_HIPCHECK(c, hipGetLastError());
  return PHSPH_OK;
}

// The density lists serve as force lists only if the active set, every
// position and every h are those the density pass left: flag otherwise.
__global__ void k_check_reuse(int nlive, int nact, const int32_t *__restrict__ act,
                              const int32_t *__restrict__ act_dens, const double4 *__restrict__ fpos,
                              const double *__restrict__ fh, const double4 *__restrict__ tp,
                              const double *__restrict__ th, int32_t *__restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  if (i < nact && act[i] != act_dens[i]) bad = true;
  if (i < nlive) {
    const double4 f = fpos[i], d = tp[i];
    if (f.x != d.x || f.y != d.y || f.z != d.z || fh[i] != th[i]) bad = true;
  }
  if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
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
// Force neighbour lists of the active slots c->act[0..c->nact) into L (list
// lengths into TNL): a uniform capacity, doubled (and kept) while the lists of
// every active particle fit in the list budget; beyond that the lists are
// built again with a capacity per group of 64 (its longest list), so that a
// few particles with many neighbours do not size the whole buffer.  Needs
// c->hmaxf of the final h.
static int force_lists(phsph_ctx *c, DBuf<uint32_t> &L, DBuf<int32_t> &TNL, const int64_t **loff_out,
                       const char *tname) {
  hipStream_t s = c->stream;
  const int nact = c->nact;
  int rc = walk_groups(c, 1, 1.0, c->fh.p, c->hmaxf.p);
  if (rc) return rc;
  const int ng = c->ngroups;
  PH_HIPCHECK(c, TNL.ensure((size_t)nact + 1));
  const int64_t *loff = nullptr;
  while (c->caps.flist > 16 && !list_fits(c, (size_t)list_cap(c->caps.flist) * nact)) c->caps.flist /= 2;
  for (int attempt = 0; attempt < 12; attempt++) {
    const int ucap = list_cap(c->caps.flist);
    if (!loff) PH_HIPCHECK(c, L.ensure((size_t)ucap * nact + 1));
    int32_t zero = 0;
    PH_HIPCHECK(c, dev_set(c, c->counters.p + 3, &zero, sizeof(int32_t)));
    if (nact > 0) {
      LaunchTimer lt(c, tname);
#define FLIST(CC, LL)                                                                                         \
  hipLaunchKernelGGL((k_flist<CC, LL>), dim3(list_grid<LL>(ng)), dim3(64 * LL), 0, s, c->dc, c->act.p, nact, ng,   \
                     c->pleaf.p, c->leaves.p, c->gcand.p, c->gncand.p, c->goff_p, c->caps.gcand, c->nd_lo.p,       \
                     c->nd_hi.p, c->nd.p, c->hmaxf.p, c->fpos.p, L.p, loff, ucap, TNL.p, c->counters.p + 3)
      // the chunk length and wavefronts per block the density autotuner chose
      if (c->tune.cl == 8) { if (c->tune.lw == 2) FLIST(8, 2); else FLIST(8, 4); }
      else { if (c->tune.lw == 2) FLIST(16, 2); else FLIST(16, 4); }
#undef FLIST
    }
    PH_HIPCHECK(c, hipGetLastError());
    int32_t fl = 0;
    PH_HIPCHECK(c, hipMemcpyAsync(&fl, c->counters.p + 3, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    if (!(fl & FFL_LIST)) break;
    if (loff) PH_RET(c, PHSPH_ERR_CAPACITY, "force: neighbour list capacity exceeded");
    if (list_fits(c, (size_t)list_cap(2 * c->caps.flist) * nact)) {  // a larger uniform capacity, kept
      c->caps.flist *= 2;
      continue;
    }
    // per-group capacity: the group's longest list rounded to 16, times 64 slots
    PH_HIPCHECK(c, c->gmax.ensure((size_t)ng + 1));
    PH_HIPCHECK(c, c->loff.ensure((size_t)ng + 1));
    hipLaunchKernelGGL(k_group_max, dim3(nblocks((size_t)ng * 64, 256)), dim3(256), 0, s, TNL.p, nact, ng,
                       c->gmax.p);
    hipLaunchKernelGGL(k_scan_offsets<int32_t>, dim3(1), dim3(1024), 0, s, c->gmax.p, ng, (int64_t)GROUP, LSEG,
                       c->loff.p);
    PH_HIPCHECK(c, hipGetLastError());
    int64_t total = 0;
    PH_HIPCHECK(c, hipMemcpyAsync(&total, c->loff.p + ng, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    PH_HIPCHECK(c, L.ensure((size_t)total + 1));
    loff = c->loff.p;
  }
  *loff_out = loff;
  return PHSPH_OK;
}

int force_lists_all(phsph_ctx *c, DBuf<uint32_t> &L, DBuf<int32_t> &TNL, const int64_t **loff,
                    const char *tname) {
  return force_lists(c, L, TNL, loff, tname);
}

struct ForcePass {
  const FRec *rec;
  const FRecB *recB;
  const FRecD *recD;
  const FRec32 *rec32;
  bool mixed, mhd, dust;
};

// One launch of the pair kernel over the active slots c->act[0..c->nact)
// with lists L (layout loff) and lengths TNL; particles flagged in skip are
// left out.  dt minima, pair count and flags accumulate in c->dtred and
// c->counters.
static int force_pass(phsph_ctx *c, int icall, const ForceIOArgs &a, const ForcePass &f, const uint32_t *L,
                      const int64_t *loff, const int32_t *TNL, const int32_t *TNLJ, int capb, const int32_t *skip) {
  hipStream_t s = c->stream;
  const int nact = c->nact;
  if (nact == 0) return PHSPH_OK;
  LaunchTimer lt(c, "force");
  const dim3 gr(xcd_grid(nblocks(nact, 256))), b(256);
  const bool th = c->p.maxvxyzu >= 4;
  const bool mhd = f.mhd, dust = f.dust;
#define FL(KK, M, D, T)                                                                                        \
  hipLaunchKernelGGL((k_force<KK, M, D, T>), gr, b, 0, s, c->dc, c->act.p, nact, skip, L, loff, TNL, TNLJ, capb,   \
                     c->iparts.p, c->fpos.p, f.rec, f.recB, f.recD, a, c->tbin_old.p, icall, c->dtred.p, c->counters.p)
#define FLK(KK)                                          \
  if (mhd && dust) { if (th) FL(KK, true, true, true); else FL(KK, true, true, false); }     \
  else if (mhd) { if (th) FL(KK, true, false, true); else FL(KK, true, false, false); }      \
  else if (dust) { if (th) FL(KK, false, true, true); else FL(KK, false, true, false); }     \
  else { if (th) FL(KK, false, false, true); else FL(KK, false, false, false); }
#define FM(KK, T)                                                                                              \
  hipLaunchKernelGGL((k_force_mixed<KK, T>), gr, b, 0, s, c->dc, c->act.p, nact, skip, L, loff, TNL, TNLJ, capb,  \
                     c->iparts.p, f.rec, c->fpos.p, c->fvel.p, f.rec32, a, c->tbin_old.p, icall, c->dtred.p,         \
                     c->counters.p)
  if (f.mixed) {
    if (c->p.kernel == PHSPH_KERNEL_CUBIC) {
      if (th) FM(PHSPH_KERNEL_CUBIC, true); else FM(PHSPH_KERNEL_CUBIC, false);
    } else {
      if (th) FM(PHSPH_KERNEL_QUINTIC, true); else FM(PHSPH_KERNEL_QUINTIC, false);
    }
  } else if (c->p.kernel == PHSPH_KERNEL_CUBIC) { FLK(PHSPH_KERNEL_CUBIC) } else { FLK(PHSPH_KERNEL_QUINTIC) }
#undef FM
#undef FLK
#undef FL
  PH_HIPCHECK(c, hipGetLastError());
  return PHSPH_OK;
}

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
  // mixed precision covers the hydro and self-gravity pair terms; MHD and
  // dust keep their FP64 pair kernels
  const bool mixed = c->p.precision == PHSPH_PREC_MIXED && !mhd && !dust;
  c->force_prec_used = mixed ? PHSPH_PREC_MIXED : PHSPH_PREC_FP64;
  PH_HIPCHECK(c, c->fpos.ensure(n));
  if (mixed) {
    PH_HIPCHECK(c, c->frec32.ensure((size_t)n * sizeof(FRec32)));
    PH_HIPCHECK(c, c->fvel.ensure(n));
  }
  FRec32 *rec32 = mixed ? (FRec32 *)c->frec32.p : nullptr;
  {
    LaunchTimer lt(c, "force_prep");
    hipLaunchKernelGGL(k_force_prep, dim3(nblocks(n, 256)), dim3(256), 0, s, c->dc, n, c->iparts.p, c->tx.p,
                       c->ty.p, c->tz.p, a.xyzh, a.ldx, a.vxyzu, a.ldv, a.gradh, a.ldgradh, a.alphaind, a.eos_vars,
                       a.ldeos, a.Bevol, a.dustfrac, c->fh.p, rec, recB, recD, a.ibin ? a.ibin_old : nullptr,
                       c->tbin_old.p, c->p.gravity ? c->gsoft.p : nullptr, c->fpos.p, c->fvel.p, rec32);
  }
  PH_HIPCHECK(c, hipGetLastError());
  PH_HIPCHECK(c, c->tmpi.ensure(n));
  hipLaunchKernelGGL(k_force_flags, dim3(nblocks(n, 256)), dim3(256), 0, s, c->dc, n, c->iparts.p, a.iphase,
                     c->tmpi.p);
  int rc = compact_active(c, c->tmpi.p, n);
  if (rc) return rc;
  if (c->p.gravity) return grav_force(c, icall, a, rec, rec32, c->gsoft.p, dtc_out, dtf_out);
  const int nact_real = c->nact;
  // the lists of the last densityiterate, if they still describe these particles
  bool reuse = c->dlist_reuse && c->nact == c->nact_dens;
  if (reuse) {
    int32_t zero = 0;
    PH_HIPCHECK(c, dev_set(c, c->counters.p + 2, &zero, sizeof(int32_t)));
    const int m = std::max(n, c->nact);
    hipLaunchKernelGGL(k_check_reuse, dim3(nblocks(m, 256)), dim3(256), 0, s, n, c->nact, c->act.p, c->act_dens.p,
                       c->fpos.p, c->fh.p, c->tp4.p, c->th.p, c->counters.p + 2);
    PH_HIPCHECK(c, hipGetLastError());
    int32_t bad = 0;
    PH_HIPCHECK(c, hipMemcpyAsync(&bad, c->counters.p + 2, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    reuse = (bad == 0);
  }
  rc = pad_groups(c);
  if (rc) return rc;
  if (reuse && c->nact != c->dlist_nact) reuse = false;
  c->force_reused_lists = reuse;
  // node hmax of the final smoothing lengths for the h_j search radius
  const bool need_lists = !reuse || c->dlist_nretry > 0;
  if (need_lists) {
    PH_HIPCHECK(c, c->hmaxf.ensure(c->ncells + 2));
    rc = leaf_hmax_run(c, c->fh.p, c->hmaxf.p);
    if (rc) return rc;
  }
  PH_HIPCHECK(c, c->dtred.ensure(3));
  {
    unsigned long long init[3];
    double big = PH_HUGE64;
    memcpy(&init[0], &big, 8);
    init[1] = init[0];
    init[2] = 0ull;
    PH_HIPCHECK(c, dev_set(c, c->dtred.p, init, sizeof(init)));
  }
  int32_t zero2[2] = {0, 0};
  int32_t fl_first[2] = {0, 0};  // flags and bin maximum of a first pass
  PH_HIPCHECK(c, dev_set(c, c->counters.p, zero2, sizeof(zero2)));
  ForcePass fp{rec, recB, recD, rec32, mixed, mhd, dust};
  if (reuse) {
    // the density lists, without the particles the density pass ran again
    rc = force_pass(c, icall, a, fp, c->nlist.p, nullptr, c->t_nl.p, c->t_nlj.p, c->dlist_cap,
                    c->dlist_nretry > 0 ? c->dretry0.p : nullptr);
    if (rc) return rc;
    if (c->dlist_nretry > 0) {  // those particles get lists of their own
      // the walk and the compaction below use the flag counters: keep this
      // pass's flags and bin maximum
      PH_HIPCHECK(c, hipMemcpyAsync(fl_first, c->counters.p, sizeof(fl_first), hipMemcpyDeviceToHost, s));
      PH_HIPCHECK(c, hipStreamSynchronize(s));
      rc = compact_active(c, c->dretry0.p, n);
      if (rc) return rc;
      rc = pad_groups(c);
      if (rc) return rc;
      const int64_t *loff = nullptr;
      rc = force_lists(c, c->nlist_r, c->t_nl_r, &loff, "force_list_rerun");
      if (rc) return rc;
      PH_HIPCHECK(c, dev_set(c, c->counters.p, zero2, sizeof(zero2)));
      rc = force_pass(c, icall, a, fp, c->nlist_r.p, loff, c->t_nl_r.p, nullptr, 0, nullptr);
      if (rc) return rc;
    }
  } else {
    c->dlist_reuse = false;  // the lists built here take the density lists' place
    const int64_t *loff = nullptr;
    rc = force_lists(c, c->nlist, c->t_nl, &loff, "force_list");
    if (rc) return rc;
    rc = force_pass(c, icall, a, fp, c->nlist.p, loff, c->t_nl.p, nullptr, 0, nullptr);
    if (rc) return rc;
  }
  int32_t fl2[2] = {0, 0};
  unsigned long long red[3];
  PH_HIPCHECK(c, hipMemcpyAsync(fl2, c->counters.p, sizeof(fl2), hipMemcpyDeviceToHost, s));
  PH_HIPCHECK(c, hipMemcpyAsync(red, c->dtred.p, sizeof(red), hipMemcpyDeviceToHost, s));
  PH_HIPCHECK(c, hipStreamSynchronize(s));
  fl2[0] |= fl_first[0];
  fl2[1] = std::max(fl2[1], fl_first[1]);
  const int32_t fl = fl2[0];
  if (fl & FFL_DTC) PH_RET(c, PHSPH_ERR_INPUT, "force: invalid dtc");
  // with no active particle nbinmax is kept (force.F90:748-756)
  if (a.ibin) c->nbinmax_new = (nact_real > 0) ? fl2[1] : a.nbinmax;
  double dtc, dtf;
  memcpy(&dtc, &red[0], 8);
  memcpy(&dtf, &red[1], 8);
  c->force_npairs = (int64_t)red[2];
  if (a.ibin) dtc = dtf = PH_BIGNUMBER;  // per-particle bins instead (force.F90:3312)
  if (dtc_out) *dtc_out = (nact_real > 0) ? dtc : PH_BIGNUMBER;
  if (dtf_out) *dtf_out = (nact_real > 0) ? fmin(dtf, PH_BIGNUMBER) : PH_BIGNUMBER;
  return PHSPH_OK;
}

// ------------------------------------------------------------------------
// set_active_particles (utils_indtimesteps.f90:114-185)
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_set_active(int np, const double *__restrict__ xyzh, int ldx, int8_t *__restrict__ iphase,
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
  __shared__ int sh[4][3];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[w][0] = na; sh[w][1] = nl; sh[w][2] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; k++) {
      na += sh[k][0]; nl += sh[k][1]; bad += sh[k][2];
    }
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
  PH_HIPCHECK(c, dev_set(c, c->counters.p, z, sizeof(z)));
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
