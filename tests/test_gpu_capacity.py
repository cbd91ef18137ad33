"""The second search passes: with tiny initial capacities every group's
candidate leaves, density and force neighbour lists and gravity near-field
lists overflow.  With the default list budget the library doubles its
uniform capacities; with a 1 KiB budget it takes its exact-size paths
(per-group candidate offsets, per-group force list capacities, per-leaf
near-field offsets; density retries with longer lists in both), and the
self-gravity force pass's neighbour lists overflow and run again at the
longest list.  Either way
it must agree with the oracle to the parity tolerances: neighbour and
iteration counts bit-exact, FP64 results within 1e-12 of the rms."""
import os

import numpy as np
import pytest

from phantom_amd import setups
from tests import parity
from tests.test_gpu_gravity import _moving, _run_gpu as run_gpu_grav, _run_oracle as run_oracle_grav

pytestmark = pytest.mark.gpu

SPECS = {"doubling": "gcand=4,dlist=16,flist=16,p2p=2,gnear=4",
         "exact": "gcand=4,dlist=16,flist=16,p2p=2,gnear=4,budget_kib=1"}


def _with_caps(spec, fn, *a):
    old = os.environ.get("PHSPH_CAPS")
    os.environ["PHSPH_CAPS"] = spec
    try:
        return fn(*a)
    finally:
        if old is None:
            del os.environ["PHSPH_CAPS"]
        else:
            os.environ["PHSPH_CAPS"] = old


CASES = {
    "shock_tube_ny8": lambda: setups.shock_tube(ny=8),
    "mhd_dust_box": lambda: setups.mhd_dust_box(nx=12, jitter=0.05),
    "box_bins": lambda: setups.with_bins(setups.uniform_box(nx=16, seed=4), nbinmax=3, istepfrac=4),
}


@pytest.mark.parametrize("mode", list(SPECS))
@pytest.mark.parametrize("name", list(CASES))
def test_tiny_capacities_hydro(name, mode, gpu):
    setup = CASES[name]()
    print(f"\n[case {name}] n={setup['xyzh'].shape[0]} caps {SPECS[mode]}", flush=True)
    g = _with_caps(SPECS[mode], parity.run_gpu, setup)
    o = parity.run_oracle(setup)
    np.testing.assert_array_equal(g["nneigh"], o["dens"]["nneigh"])
    np.testing.assert_array_equal(g["nits"], o["dens"]["nits"])
    assert parity.rel_rms(g["xyzh"][:, 3], o["xyzh"][:, 3]) <= 1e-12
    ncomp = 4 if setup["params"].maxvxyzu >= 4 else 3
    for c in range(ncomp):
        assert parity.rel_rms(g["force"]["fxyzu"][:, c], o["force"]["fxyzu"][:, c]) <= 1e-12, c
    if setup["params"].mhd:
        for c in range(4):
            assert parity.rel_rms(g["force"]["dBevol"][:, c], o["force"]["dBevol"][:, c]) <= 1e-12, c
    if setup.get("bins") is not None:
        np.testing.assert_array_equal(g["force"]["ibin"], o["force"]["ibin"])
    g["ctx"].close()


@pytest.mark.parametrize("mode", list(SPECS))
def test_tiny_capacities_gravity(mode, gpu):
    setup = setups.polytrope(n=3000, theta=0.5)
    print(f"\n[case polytrope] n={setup['xyzh'].shape[0]} caps {SPECS[mode]}", flush=True)
    g = _with_caps(SPECS[mode], run_gpu_grav, setup)
    o = run_oracle_grav(setup)
    np.testing.assert_array_equal(g["nneigh"], o["dens"]["nneigh"])
    assert parity.rel_rms(g["xyzh"][:, 3], o["xyzh"][:, 3]) <= 1e-12
    for c in range(4):
        assert parity.rel_rms(g["force"]["fxyzu"][:, c], o["force"]["fxyzu"][:, c]) <= 1e-12, c
    pg = np.asarray(g["force"]["poten"], np.float64)
    po = np.asarray(o["force"]["poten"], np.float64)
    assert parity.rel_rms(pg, po) <= 2e-7


@pytest.mark.parametrize("mode", list(SPECS))
def test_tiny_capacities_gravity_bins(mode, gpu):
    """The self-gravity force pass runs again after its neighbour lists
    overflow; with block timesteps that run must start from the bins the
    call was given (the kernel moves ibin in place), so ibin, ibin_wake and
    nbinmax equal the oracle's.  With the 1 KiB budget the re-run takes
    per-group list capacities."""
    setup = setups.with_bins(_moving(setups.polytrope(n=3000, theta=0.5)), nbinmax=3, istepfrac=2, dtmax=0.05)
    print(f"\n[case polytrope bins] n={setup['xyzh'].shape[0]} caps {SPECS[mode]}", flush=True)
    g = _with_caps(SPECS[mode], run_gpu_grav, setup)
    o = run_oracle_grav(setup)
    np.testing.assert_array_equal(g["nneigh"], o["dens"]["nneigh"])
    for c in range(4):
        assert parity.rel_rms(g["force"]["fxyzu"][:, c], o["force"]["fxyzu"][:, c]) <= 1e-12, c
    np.testing.assert_array_equal(g["force"]["ibin"], o["force"]["ibin"])
    np.testing.assert_array_equal(g["force"]["ibin_wake"], o["force"]["ibin_wake"])
    assert g["force"]["nbinmax"] == o["force"]["nbinmax"]
