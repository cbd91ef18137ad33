"""Algorithmic work of the density and force kernels.

Floating-point operations per interacting pair, counted from the kernel
source (phantom_amd/csrc/density.hip, force.hip) with FMA = 2 flops and
sqrt / division = 1 flop each.  Rejected trial pairs (outside the kernel
support) are not counted: their number depends on the neighbour-search
implementation, while the interacting pairs are fixed by the physics
(N_ngb^dens = #{j : q_i < R}, N_ngb^force = #{j : q_i < R or q_j < R}).

Per interacting pair:
  F_RHO_ITER  h-rho iteration pass: separation with periodic offset (6),
              r^2 (5), q^2 (1), q = sqrt(r^2)/h (2), kernel W and dW/dq
              (13), rho and Omega sums (7)                           = 34
  F_DENS      final pass: F_RHO_ITER + 1/(r+eps) (2), unit vector times
              grad W m (7), dv (3), div v (5), grad v (18), dv/dt
              differences and grad a (21), exact-linear matrix (12)   = 102
  F_DENS_MHD  extra for div B / grad B with rho_j = m (hfact/h_j)^3   = 34
  F_FORCE     hydro + Cullen-Dehnen viscosity + conductivity: separation
              (11), q_i^2, q_j^2 (3), 1/r and q (4), dW_i and dW_j with
              grad-h factors (25), unit vector, dv, v.r (11), signal
              speeds (16), j terms (8), viscous pressure (14), pressure
              gradients (6), conductivity (12), accumulation (14)     = 124
  F_FORCE_MHD stress projections, induction, resistivity, cleaning    = 110
  F_FORCE_DUST one-fluid dust terms                                   = 40
"""

F_RHO_ITER = 34
F_DENS = 102
F_DENS_MHD = 34
F_FORCE = 124
F_FORCE_MHD = 110
F_FORCE_DUST = 40

# per-particle finish (Newton update, exact-linear inversion, stores)
F_FINISH_ITER = 30
F_FINISH = 150


def density_iter_flops(nneigh_dens, nits):
    """The h-rho iterations: rho and Omega sums over the neighbours at every
    iteration plus the Newton update."""
    import numpy as np
    nn = np.asarray(nneigh_dens, dtype=np.float64) - 1.0
    it = np.asarray(nits, dtype=np.float64)
    return float(np.sum(nn * it * F_RHO_ITER + it * F_FINISH_ITER))


def density_sums_flops(nneigh_dens, mhd=False):
    """One pass of every density sum at the final h, and store_results."""
    import numpy as np
    nn = np.asarray(nneigh_dens, dtype=np.float64) - 1.0
    f_pair = F_DENS + (F_DENS_MHD if mhd else 0)
    return float(np.sum(nn * f_pair + F_FINISH))


def density_flops(nneigh_dens, nits, mhd=False):
    """nneigh_dens, nits: per-particle arrays (neighbour count including
    self, iterations to converge)."""
    import numpy as np
    nn = np.asarray(nneigh_dens, dtype=np.float64) - 1.0
    it = np.asarray(nits, dtype=np.float64)
    f_pair = F_DENS + (F_DENS_MHD if mhd else 0)
    return float(np.sum(nn * it * F_RHO_ITER + it * F_FINISH_ITER + nn * f_pair + F_FINISH))


def force_flops(npairs_force, mhd=False, dust=False):
    f = F_FORCE + (F_FORCE_MHD if mhd else 0) + (F_FORCE_DUST if dust else 0)
    return float(npairs_force) * f


# compulsory HBM bytes per particle (read inputs once, write outputs once)
BYTES_DENS = 8 * (4 + 3 + 3) + 4 * (1 + 1 + 9 + 1) + 8   # xyzh, v, f; gradh, divv, dvdx, alpha(3); h
BYTES_FORCE = 8 * (4 + 4 + 2) + 4 * (1 + 1) + 8 * 4 + 4  # xyzh, vxyzu, P, cs; gradh, alpha; fxyzu; divv

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def step_loop_bytes(loop, nv=4, mhd=False, dust=False, ind=False):
    """Compulsory HBM bytes per particle of the integrator loops
    (step.hip): every array the loop reads is read once and every array it
    updates is written once, in the caller's FP64 / FP32 / int8 layout."""
    b = 0
    if loop == "step_predictor":
        b = 8 + 8 * nv * 3                      # h; v read + write; f
        b += (32 * 2 + 32) if mhd else 0        # Bevol read + write, dBevol
        b += (8 * 2 + 8) if dust else 0
        b += (8 + 1 + 1 + 1) if ind else 0      # twas, ibin, ibin_old, iphase
    elif loop == "substep_sph":
        b = 8 + 24 * 2 + 24                     # h; x read + write; v
    elif loop == "step_predict_sph":
        b = 8 * 2 + 4 + 8 * nv * 3 + 8 + 4 * 2 + 4  # h rw, div v, v, f, vpred, c_s, alpha(1:2) r, alpha(1) w
        b += 32 * 3 if mhd else 0
        b += 8 * 3 if dust else 0
        b += 8 if ind else 0
    elif loop == "step_corrector":
        b = 8 + 8 * nv * 3                      # h; v read + write; f
        b += (8 * 2 + 1 * 2 + 1 * 2 + 1) if ind else 24  # twas, ibin, ibin_wake, iphase | vpred(1:3)
        b += (32 * 2 + 32) if mhd else 0
        b += (8 * 2 + 8) if dust else 0
    return b
