"""Tree parity at sizes where the large-node phase cuts each top-level node
into many chunks (more than one wavefront's worth per node): leaf
membership, node ranges, children, hmax and the in-leaf order bit-exact
with the oracle's sequential build, centres and sizes within a few ulps.
The lattice shock tube puts particles exactly on split planes, so its top
nodes take the exact sequential-sum path; the random box does not."""
import numpy as np
import pytest

from oracle import oracle
from phantom_amd import setups, sph
from tests import parity

pytestmark = pytest.mark.gpu

CASES = {
    "shock_tube_ny40": lambda: setups.shock_tube(ny=40),
    "random_300k": lambda: setups.random_box(n=300000, seed=17),
    "box_64_jitter": lambda: setups.uniform_box(nx=64, seed=5),
}


@pytest.mark.parametrize("name", list(CASES))
def test_large_tree_bit_exact(name, gpu):
    setup = CASES[name]()
    p = setup["params"]
    n = setup["xyzh"].shape[0]
    print(f"\n[case {name}] n={n}", flush=True)
    ctx = sph.Context(p)
    xg = setup["xyzh"].copy()
    leaf, ncells = ctx.maketree(xg)
    gt = ctx.get_tree()
    # a second build of the same particles gives the same tree
    xg2 = setup["xyzh"].copy()
    ctx.maketree(xg2)
    gt2 = ctx.get_tree()
    ctx.close()
    o = oracle.Oracle(p)
    xo = setup["xyzh"].copy()
    o.maketree(xo)
    ot = o.get_tree()
    assert ncells == ot["ncells"]
    np.testing.assert_array_equal(leaf, ot["leaf_is_active"])
    used = parity.reachable_nodes(ot)
    np.testing.assert_array_equal(gt["inoderange"][used], ot["inoderange"][used])
    np.testing.assert_array_equal(gt["inodeparts"], ot["inodeparts"])
    for f in ("leftchild", "rightchild", "parent"):
        np.testing.assert_array_equal(gt["node"][f][used], ot["node"][f][used])
    np.testing.assert_array_equal(gt["node"]["hmax"][used], ot["node"]["hmax"][used])
    # a centre summed in parallel differs from the sequential sum by at most
    # 2 gamma_{n-1} sum|m x| / M <= 2 (n-1) u / (1 - (n-1) u) max|x| (u = 2^-53)
    u = 2.0 ** -53
    nn = (ot["inoderange"][used, 1] - ot["inoderange"][used, 0] + 1).astype(np.float64)
    xmax = np.max(np.abs(setup["xyzh"][:, :3]))
    tol = 2.0 * nn * u / (1.0 - nn * u) * xmax + 4 * u * xmax
    dx = np.max(np.abs(gt["node"]["xcen"][used] - ot["node"]["xcen"][used]), axis=1)
    assert np.all(dx <= tol), (np.max(dx / tol))
    ds = np.abs(gt["node"]["size"][used] - ot["node"]["size"][used])
    assert np.all(ds <= 2 * tol + 1e-15), np.max(ds / (2 * tol + 1e-15))
    np.testing.assert_array_equal(gt2["inodeparts"], gt["inodeparts"])
    np.testing.assert_array_equal(gt2["node"][used], gt["node"][used])
