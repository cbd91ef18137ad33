"""Mixed precision (FP32 pair arithmetic, FP64 sums) against the FP64 oracle.

The neighbour tests stay FP64, so neighbour counts and iteration counts
must be bit-exact.  h and rho must agree with the FP64 oracle to 1e-6
(max difference over the rms of the quantity); each component of dv/dt to
1e-6 of the rms of |dv/dt| (a vector: a component that is only noise, like
the transverse force of a shock tube, is measured against the force, not
against its own noise), du/dt to 1e-6 of its rms, and the float32 velocity
gradients (div v, dv/dx) to 1e-5 of the rms of |dv/dx|.  The error of each
quantity is printed.  Self-gravity has a mixed form (FP64 multipoles,
far-field 1/r^2 sums with a refined FP32 reciprocal square root, FP32
neighbour records); MHD and dust keep their FP64 pair kernels, so with
mixed precision requested they must agree with the oracle to the FP64
tolerance."""
import numpy as np
import pytest

from phantom_amd import abi, setups
from tests import parity

pytestmark = pytest.mark.gpu

TOL_MIXED = 1e-6


def _mixed(s):
    s["params"].precision = abi.PREC_MIXED
    return s


def _cases():
    return {
        "shock_tube_ny8_quintic": lambda: setups.shock_tube(ny=8),
        "box24_iso_cubic": lambda: setups.uniform_box(nx=24, seed=11),
        "random_8k": lambda: setups.random_box(n=8000, seed=3),
        "shock_tube_bins": lambda: setups.with_bins(setups.shock_tube(ny=8), nbinmax=3, istepfrac=2),
        # self-gravity, tree opening angle 0.5
        "polytrope_4k_cubic": lambda: setups.polytrope(n=4000, theta=0.5),
        "polytrope_3k_quintic_bins": lambda: setups.with_bins(
            setups.polytrope(n=3000, theta=0.5, kernel=abi.KERNEL_QUINTIC), nbinmax=3, istepfrac=2),
    }


@pytest.fixture(scope="module", params=list(_cases().keys()))
def pair(request, gpu):
    setup = _mixed(_cases()[request.param]())
    print(f"\n[case {request.param}] n={setup['xyzh'].shape[0]} mixed", flush=True)
    g = parity.run_gpu(setup)
    o = parity.run_oracle(setup)
    # every case here has a mixed form of both passes
    assert g["ctx"].get_precision() == (abi.PREC_MIXED, abi.PREC_MIXED)
    return request.param, setup, g, o


def test_neighbours_bit_exact(pair):
    name, setup, g, o = pair
    np.testing.assert_array_equal(g["nneigh"], o["dens"]["nneigh"])
    np.testing.assert_array_equal(g["nits"], o["dens"]["nits"])


def test_density(pair):
    name, setup, g, o = pair
    h_g, h_o = g["xyzh"][:, 3], o["xyzh"][:, 3]
    pm = setup["params"].massoftype_gas
    hf = setup["params"].hfact
    rho_g, rho_o = pm * (hf / np.abs(h_g)) ** 3, pm * (hf / np.abs(h_o)) ** 3
    e_h = parity.rel_rms(h_g, h_o)
    e_rho = parity.rel_rms(rho_g, rho_o)
    print(f"  mixed vs FP64: h {e_h:.2e}  rho {e_rho:.2e}", flush=True)
    assert e_h <= TOL_MIXED and e_rho <= TOL_MIXED
    dvo = np.asarray(o["dens"]["dvdx"], np.float64)
    scale = np.sqrt(np.mean(np.sum(dvo * dvo, axis=1)))
    for key in ("divcurlv", "dvdx"):
        d = np.asarray(g["dens"][key], np.float64) - np.asarray(o["dens"][key], np.float64)
        if scale == 0.0:  # fluid at rest: the gradients are exactly zero
            assert np.max(np.abs(d)) == 0.0, key
            continue
        e = float(np.max(np.abs(d))) / scale
        print(f"  {key} {e:.2e}", flush=True)
        assert e <= 1e-5, key


def test_force(pair):
    name, setup, g, o = pair
    gf, of = g["force"], o["force"]
    fo = of["fxyzu"][:, :3]
    scale = np.sqrt(np.mean(np.sum(fo * fo, axis=1)))
    for c in range(3):
        e = float(np.max(np.abs(gf["fxyzu"][:, c] - fo[:, c]))) / scale
        print(f"  dv/dt[{c}] {e:.2e}", flush=True)
        assert e <= TOL_MIXED, (name, c)
    if setup["params"].maxvxyzu >= 4:
        e = parity.rel_rms(gf["fxyzu"][:, 3], of["fxyzu"][:, 3])
        print(f"  du/dt {e:.2e}", flush=True)
        assert e <= TOL_MIXED, (name, "du/dt")
    if setup.get("bins") is None:
        assert abs(gf["dtcourant"] - of["dtcourant"]) <= 1e-6 * of["dtcourant"]
        assert abs(gf["dtforce"] - of["dtforce"]) <= 1e-6 * of["dtforce"]
    else:
        # a new bin only moves where log2 of the timestep sits within
        # rounding of an integer
        assert np.mean(gf["ibin"] == of["ibin"]) >= 0.999


def test_fp64_kernels_unchanged_with_mhd_dust(gpu):
    """MHD and dust pair terms keep FP64 under the mixed setting."""
    setup = _mixed(setups.mhd_dust_box(nx=12, jitter=0.05))
    g = parity.run_gpu(setup)
    o = parity.run_oracle(setup)
    for c in range(3):
        assert parity.rel_rms(g["force"]["fxyzu"][:, c], o["force"]["fxyzu"][:, c]) <= 1e-12
