// common.h -- device-side building blocks shared by the tree, density and
// force kernels: SPH kernel functions, particle helpers and exact-rounding
// arithmetic for the places where the result must match the reference's
// FP64 evaluation order bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/phsph.h"

#define PH_PI 3.1415926536                 // physcon.f90:23 (10 digits, as the reference)
#define PH_EPS64 2.220446049250313e-16     // epsilon(1d0)
#define PH_TINY64 2.2250738585072014e-308  // tiny(1d0)
#define PH_HUGE64 1.7976931348623157e308   // huge(1d0)
#define PH_BIGNUMBER 1.e29                 // timestep.f90:40
#define PH_MAXDENSITS 100                  // dens.F90:101
#define PH_MINPART 10                      // config.F90:113

// Runtime constants derived from phsph_params, passed by value to kernels.
struct DevConst {
  int kernel;        // PHSPH_KERNEL_*
  int periodic;
  int maxvxyzu;
  int mhd, dust, gravity, ind_timesteps, ieos;
  int ishock_heating, ipdv_heating, iresistive_heating, ilimitdustflux, idrag;
  double radkern, radkern2, cnormk, wab0, gradh0, dphidh0;
  double dxbound, dybound, dzbound, hdlx, hdly, hdlz;
  double xmin, xmax, ymin, ymax, zmin, zmax;
  double hfact, tolh, pmass, gamma, polyk;
  double alpha, alphamax, alphau, alphaB, beta;
  double C_cour, C_force, dtmax, psidecayfac, overcleanfac, K_code;
  double tree_acc2;  // tree_accuracy^2 (opening angle of the gravity walk)
  double grainsize, graindens, seff, cste_mu, coeff_gei_1;  // idrag = 1 (dust.f90:78-97)
};

// ---- exact (non-contracted) FP64 arithmetic ----------------------------
// The reference's x86-64 build rounds every product and sum separately; the
// HIP compiler contracts a*b+c into one FMA (HIP's __dmul_rn is a plain
// multiply, so it does not prevent this).  Where bit-exact agreement is
// required (tree centres and sizes, neighbour membership tests) every product
// passes through an empty asm statement that pins the rounded value in a
// register, so no later add can be fused with it.  The statement is not
// volatile: a volatile one would also keep the compiler from moving loads
// across it and serialise every load in the loops that use it.
__device__ __forceinline__ double xmul(double a, double b) {
  double p = a * b;
  __asm__("" : "+v"(p));
  return p;
}
__device__ __forceinline__ double xadd(double a, double b) { return a + b; }
__device__ __forceinline__ double xsub(double a, double b) { return a - b; }
__device__ __forceinline__ double r2_exact(double dx, double dy, double dz) {
  return xadd(xadd(xmul(dx, dx), xmul(dy, dy)), xmul(dz, dz));
}

// minimum-image correction of a separation (dens.F90:666-670,
// force.F90:1266-1270): if |dx| > L/2 then dx -= L*sign(dx)
__device__ __forceinline__ double min_image(double dx, double L, double halfL) {
  return (fabs(dx) > halfL) ? xsub(dx, copysign(L, dx)) : dx;
}

// ---- SPH kernels (kernel_cubic.f90:34-120, kernel_quintic.f90:34-152) ---
template <int K, class R>
__device__ __forceinline__ void kern_wg(R q2, R q, R &w, R &g) {
  if (K == PHSPH_KERNEL_CUBIC) {
    if (q < R(1)) {
      w = R(0.75) * q2 * q - R(1.5) * q2 + R(1);
      g = q * (R(2.25) * q - R(3));
    } else if (q < R(2)) {
      R t = q - R(2);
      w = R(-0.25) * (t * t * t);
      g = R(-0.75) * (t * t);
    } else {
      w = R(0); g = R(0);
    }
  } else {
    if (q < R(1)) {
      R q4 = q2 * q2;
      w = R(-10) * q4 * q + R(30) * q4 - R(60) * q2 + R(66);
      g = q * (R(-50) * q2 * q + R(120) * q2 - R(120));
    } else if (q < R(2)) {
      R a = q - R(3), b = q - R(2);
      R a2 = a * a, b2 = b * b;
      w = -(a2 * a2 * a) + R(6) * (b2 * b2 * b);
      g = R(-5) * (a2 * a2) + R(30) * (b2 * b2);
    } else if (q < R(3)) {
      R a = q - R(3);
      R a2 = a * a;
      w = -(a2 * a2 * a);
      g = R(-5) * (a2 * a2);
    } else {
      w = R(0); g = R(0);
    }
  }
}

template <int K, class R>
__device__ __forceinline__ R kern_g(R q2, R q) {
  R w, g;
  kern_wg<K, R>(q2, q, w, g);
  return g;
}

// dphidh of get_kernel_grav1 (kernel_cubic.f90:80-101, kernel_quintic.f90:90-122)
template <int K, class R>
__device__ __forceinline__ R kern_dphidh(R q2, R q) {
  R q4 = q2 * q2;
  if (K == PHSPH_KERNEL_CUBIC) {
    if (q < R(1)) return R(-0.6) * q4 * q + R(1.5) * q4 - R(2) * q2 + R(1.4);
    if (q < R(2)) return R(0.2) * q4 * q - R(1.5) * q4 + R(4) * q2 * q - R(4) * q2 + R(1.6);
    return R(0);
  } else {
    R q6 = q4 * q2;
    if (q < R(1)) return q6 * q / R(21) - q6 / R(6) + q4 / R(2) - R(11) * q2 / R(10) + R(239.0 / 210.0);
    if (q < R(2))
      return -q6 * q / R(42) + q6 / R(4) - q4 * q + R(7) * q4 / R(4) - R(5) * q2 * q / R(6) - R(17) * q2 / R(20) +
             R(473.0 / 420.0);
    if (q < R(3))
      return q6 * q / R(210) - q6 / R(12) + R(3) * q4 * q / R(5) - R(9) * q4 / R(4) + R(9) * q2 * q / R(2) -
             R(81) * q2 / R(20) + R(243.0 / 140.0);
    return R(0);
  }
}

// ---- particle helpers (part.F90:779-858) ------------------------------
__device__ __forceinline__ double rhoh_d(double h, double hfact, double m) {
  double t = hfact / fabs(h);
  return m * (t * t * t);
}
__device__ __forceinline__ double hrho_d(double rho, double hfact, double m) {
  return hfact * pow(m / fabs(rho), 1.0 / 3.0);
}

// periodic offset code of a candidate leaf: (sx+1) + 3 (sy+1) + 9 (sz+1)
__device__ __forceinline__ int off_code(int sx, int sy, int sz) { return (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1); }
__device__ __forceinline__ void off_decode(int code, int &sx, int &sy, int &sz) {
  sx = code % 3 - 1;
  sy = (code / 3) % 3 - 1;
  sz = code / 9 - 1;
}

// candidate leaf entry: node id (<2^27) and offset code (<27) packed
__device__ __forceinline__ uint32_t cand_pack(int node, int code) { return (uint32_t)node | ((uint32_t)code << 27); }
__device__ __forceinline__ int cand_node(uint32_t c) { return (int)(c & 0x07ffffffu); }
__device__ __forceinline__ int cand_code(uint32_t c) { return (int)(c >> 27); }

// ---- wave-level reductions ---------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// softened potential and force of a unit mass at q = r/h
// (kernel_softening, kernel_cubic.f90:103-121, kernel_quintic.f90:124-152)
template <int K>
__device__ __forceinline__ void kern_soft(double q2, double q, double &pot, double &fs) {
  if (K == PHSPH_KERNEL_CUBIC) {
    if (q < 1.) {
      const double q4 = q2 * q2;
      pot = q4 * q / 10. - 3. * q4 / 10. + 2. * q2 / 3. - 7. / 5.;
      fs = q * (15. * q2 * q - 36. * q2 + 40.) / 30.;
    } else if (q < 2.) {
      const double q4 = q2 * q2, q6 = q4 * q2;
      pot = (q * (-q4 * q + 9. * q4 - 30. * q2 * q + 40. * q2 - 48.) + 2.) / (30. * q);
      fs = (-5. * q6 + 36. * q4 * q - 90. * q4 + 80. * q2 * q - 2.) / (30. * q2);
    } else {
      pot = -1. / q;
      fs = 1. / q2;
    }
  } else {
    if (q < 1.) {
      const double q4 = q2 * q2, q6 = q4 * q2;
      pot = -q6 * q / 168. + q6 / 42. - q4 / 10. + 11. * q2 / 30. - 239. / 210.;
      fs = q * (-35. * q4 * q + 120. * q4 - 336. * q2 + 616.) / 840.;
    } else if (q < 2.) {
      const double q4 = q2 * q2, q6 = q4 * q2, q8 = q6 * q2;
      pot = (q * (5. * q6 * q - 60. * q6 + 280. * q4 * q - 588. * q4 + 350. * q2 * q + 476. * q2 - 1892.) - 5.) /
            (1680. * q);
      fs = (35. * q8 - 360. * q6 * q + 1400. * q6 - 2352. * q4 * q + 1050. * q4 + 952. * q2 * q + 5.) /
           (1680. * q2);
    } else if (q < 3.) {
      const double q4 = q2 * q2, q6 = q4 * q2, q8 = q6 * q2;
      pot = (q * (-q6 * q + 20. * q6 - 168. * q4 * q + 756. * q4 - 1890. * q2 * q + 2268. * q2 - 2916.) + 507.) /
            (1680. * q);
      fs = (-7. * q8 + 120. * q6 * q - 840. * q6 + 3024. * q4 * q - 5670. * q4 + 4536. * q2 * q - 507.) /
           (1680. * q2);
    } else {
      pot = -1. / q;
      fs = 1. / q2;
    }
  }
}
