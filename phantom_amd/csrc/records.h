// records.h -- packed per-particle records of the force pair loops (tree
// order), shared by the hydro and the self-gravity force kernels.
#pragma once

#include "stream.h"

// j quantities of the hydro pair terms (tree order); cs and h are read for i
struct __align__(16) FRec {
  double x, y, z, hinv;
  double vx, vy, vz, u;
  double rho, rho1, pr, pro2;
  double vwave;
  float gradh, alpha;
  double cs, h;
};
static_assert(sizeof(FRec) == 128, "FRec is one 128-byte line");
// the record after its 32-byte position part (what a pair reads once the
// neighbour test has passed)
struct __align__(16) FTail {
  double vx, vy, vz, u;
  double rho, rho1, pr, pro2;
  double vwave;
  float gradh, alpha;
  double cs, h;
};
static_assert(sizeof(FTail) == 96 && offsetof(FRec, vx) == 32, "FTail is the last 96 bytes of FRec");
// mixed precision (hydro): the j quantities of the dissipation terms in
// FP32 and P/rho^2 as an unevaluated FP32 pair; the FP64 position, 1/h,
// velocity and u of j come from 32-byte position and velocity records
struct __align__(16) FRec32 {
  float rho1, pro2, pro2lo, vwave;  // pro2 = (double)pro2 + pro2lo to FP64 accuracy
  float alpha, gradh, rho, pr;
  float prlo, pad0, pad1, pad2;     // pr = (double)pr + prlo
};
static_assert(sizeof(FRec32) == 48, "FRec32 is 48 bytes");
// MHD: B = Bevol*rho, psi, the stress -m B(x)B/rho^2, and Bevol itself (i)
struct __align__(16) FRecB {
  double Bx, By, Bz, psi;
  double sxx, sxy, sxz, syy;
  double syz, szz, Bex, Bey;
  double Bez, pad0, pad1, pad2;
};
static_assert(sizeof(FRecB) == 128, "FRecB is one 128-byte line");
// one-fluid dust
struct __align__(16) FRecD {
  double eps, sqeps, ts, pad;
};

// get_ts (dust.f90:161-276) with dv2 = 0, as the one-fluid force calls it
// (force.F90:1793): idrag 1 (Epstein or Stokes by the Knudsen number, no
// supersonic correction at dv2 = 0), 2 (constant K) and 3 (constant t_s)
__device__ __forceinline__ double get_ts_d(const DevConst &dc, double rhogas, double rhodust, double spsoundgas) {
  if (dc.idrag == 1) {
    const double rhosum = rhogas + rhodust;
    const double lambda = dc.seff / rhogas;
    const double kn_eff = (dc.grainsize > 0.) ? 9. * lambda / (4. * dc.grainsize) : PH_HUGE64;
    double dragcoeff, f;
    if (kn_eff >= 1.) {  // Epstein
      dragcoeff = (dc.graindens > PH_TINY64) ? dc.coeff_gei_1 * spsoundgas / (dc.graindens * dc.grainsize) : PH_HUGE64;
      f = 1.;
    } else {  // Stokes at |dv| = 0: Re_dust = 0 (NaN when the viscosity is 0)
      const double viscmol_nu = dc.cste_mu * lambda * spsoundgas;
      const double re_dust = 2. * dc.grainsize * 0. / viscmol_nu;
      if (re_dust <= 1.) {
        dragcoeff = 4.5 * viscmol_nu / (dc.graindens * dc.grainsize * dc.grainsize);
        f = 1.;
      } else if (re_dust <= 800.) {
        dragcoeff = 9. / (dc.graindens * dc.grainsize * pow(re_dust, 0.6));
        f = 0.;
      } else {
        dragcoeff = 0.163075 / (dc.graindens * dc.grainsize);
        f = 0.;
      }
    }
    const double ts1 = (dragcoeff == PH_HUGE64) ? PH_HUGE64 : dragcoeff * f * rhosum;
    return (ts1 > 0.) ? 1. / ts1 : PH_HUGE64;
  }
  if (dc.idrag == 2) {
    if (dc.K_code > 0.) return rhogas * rhodust / (dc.K_code * (rhogas + rhodust));
    return PH_HUGE64;
  }
  if (dc.idrag == 3) return dc.K_code;
  return 0.;
}

// get_newbin (utils_indtimesteps.f90:230-287)
__device__ __forceinline__ int get_newbin_d(double dti, double dtmax, int ibin_oldi, bool allow_decrease,
                                            int nbinmax, int istepfrac) {
  const double dlog2 = 1.4426950408889634;  // 1/log(2)
  int ibin_newi;
  if (dti > dtmax) ibin_newi = 0;
  else if (dti < PH_TINY64) ibin_newi = 30;  // maxbins (utils_indtimesteps.f90:29)
  else ibin_newi = max((int)(log(2. * dtmax / dti) * dlog2 - PH_EPS64), 0);
  if (ibin_newi > ibin_oldi) return ibin_newi;
  if (ibin_newi < ibin_oldi && ibin_oldi <= nbinmax && allow_decrease) {
    // up only when the bins are synchronised, one bin at a time
    if (istepfrac % (1 << (nbinmax - (ibin_oldi - 1))) == 0) return ibin_oldi - 1;
  }
  return ibin_oldi;
}

// Block-wide reduction of the force results that are global: the timestep
// minima (positive doubles order as their bit patterns), the interacting
// pair count and the highest new bin (bmax < 0: none).  Wavefront shuffles,
// then one set of atomics per block.  Every thread of the block calls it.
__device__ __forceinline__ void reduce_force_block(double dtc, double dtf, int npairs, int bmax,
                                                   unsigned long long *dtred, int32_t *flags) {
  __shared__ unsigned long long s_a[16], s_b[16], s_n[16];
  __shared__ int s_m[16];
  unsigned long long a = (unsigned long long)__double_as_longlong(dtc);
  unsigned long long b = (unsigned long long)__double_as_longlong(dtf);
  unsigned long long n = (unsigned long long)npairs;
  int m = bmax;
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long a2 = __shfl_xor(a, off, 64), b2 = __shfl_xor(b, off, 64);
    a = a < a2 ? a : a2;
    b = b < b2 ? b : b2;
    n += __shfl_xor(n, off, 64);
    m = max(m, __shfl_xor(m, off, 64));
  }
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_a[w] = a; s_b[w] = b; s_n[w] = n; s_m[w] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < nw; k++) {
      a = a < s_a[k] ? a : s_a[k];
      b = b < s_b[k] ? b : s_b[k];
      n += s_n[k];
      m = max(m, s_m[k]);
    }
    atomicMin(&dtred[0], a);
    atomicMin(&dtred[1], b);
    atomicAdd(&dtred[2], n);
    if (m >= 0) atomicMax(&flags[1], m);
  }
}
