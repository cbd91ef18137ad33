"""Run-to-run determinism: the derivative chain (tree, density, cons2prim,
force) run twice, each time in a fresh context on the same inputs, must
give bit-identical outputs.  Work is handed out through atomic counters in
several places (the gravity walk's level entries and list space, the
density retries, the per-XCD queues), so the order in which wavefronts
finish differs from run to run; every sum is taken in an order fixed by
the tree and the lists, so no output may depend on it."""
import numpy as np
import pytest

from phantom_amd import abi, setups
from tests import parity

pytestmark = pytest.mark.gpu


def _mixed(s):
    s["params"].precision = abi.PREC_MIXED
    return s


CASES = {
    "box32_fp64": lambda: setups.uniform_box(nx=32, jitter=0.1, seed=5, isothermal=False, vamp=0.5),
    "mhd_dust_16": lambda: setups.mhd_dust_box(nx=16, jitter=0.05),
    "polytrope_20k_theta05_bins": lambda: setups.with_bins(setups.polytrope(n=20000, theta=0.5), nbinmax=3,
                                                           istepfrac=2, dtmax=0.05),
    "shock_tube_ny8_mixed": lambda: _mixed(setups.shock_tube(ny=8)),
}


def _flatten(out):
    """every array and scalar of a run_gpu result, by name"""
    flat = {"h": out["xyzh"][:, 3], "nneigh": out["nneigh"], "nits": out["nits"], "alphaind": out["alphaind"]}
    for group in ("dens", "c2p", "force"):
        for k, v in out[group].items():
            if v is None:
                continue
            flat[f"{group}.{k}"] = np.asarray(v)
    return flat


@pytest.mark.parametrize("name", list(CASES))
def test_two_runs_bit_identical(name, gpu):
    runs = []
    for r in range(2):
        setup = CASES[name]()
        print(f"\n[case {name}] run {r + 1} n={setup['xyzh'].shape[0]}", flush=True)
        g = parity.run_gpu(setup)
        g["ctx"].close()
        runs.append(_flatten(g))
    a, b = runs
    assert a.keys() == b.keys()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{name} {k}")
