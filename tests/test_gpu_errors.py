"""Error paths of the C-ABI on the GPU: the reference stops with fatal()
where these calls return a status code (and a message), and every kernel
still runs to its end.  Each failure is forced from the host side:

* a NaN position: maketree returns PHSPH_ERR_NAN (kdtree.F90:391);
* no live particle (every h <= 0): PHSPH_ERR_NODATA (kdtree.F90:162);
* tolh = 1e-300: the h-rho iteration reaches maxdensits and
  densityiterate returns PHSPH_ERR_NOCONVERGE (dens.F90:1443); the oracle
  fails on the same particles;
* force or densityiterate before maketree: PHSPH_ERR_STATE;
* a context stays usable after an error: the next valid call succeeds."""
import numpy as np
import pytest

from oracle import oracle
from phantom_amd import abi, setups, sph

pytestmark = pytest.mark.gpu


def _code(excinfo):
    return excinfo.value.code


def test_nan_position(gpu):
    s = setups.uniform_box(nx=8, seed=2)
    xyzh = s["xyzh"].copy()
    xyzh[17, 1] = np.nan
    ctx = sph.Context(s["params"])
    with pytest.raises(sph.PhsphError) as e:
        ctx.maketree(xyzh)
    assert _code(e) == abi.ERR_NAN
    assert "NaN" in str(e.value)
    # the same context builds a tree of valid particles afterwards
    ctx.maketree(s["xyzh"].copy())
    ctx.close()


def test_no_live_particles(gpu):
    s = setups.uniform_box(nx=8, seed=2)
    xyzh = s["xyzh"].copy()
    xyzh[:, 3] = -1.0  # every particle accreted (h < 0)
    ctx = sph.Context(s["params"])
    with pytest.raises(sph.PhsphError) as e:
        ctx.maketree(xyzh)
    assert _code(e) == abi.ERR_NODATA
    ctx.close()


def test_density_no_convergence(gpu):
    s = setups.random_box(n=2000, seed=3)
    s["params"].tolh = 1e-300
    with pytest.raises(Exception, match="NOCONVERGE"):
        oracle.run_all(s)
    ctx = sph.Context(s["params"])
    xyzh = s["xyzh"].copy()
    ctx.maketree(xyzh)
    with pytest.raises(sph.PhsphError) as e:
        ctx.densityiterate(1, xyzh, s["vxyzu"], alphaind=s["alphaind"].copy())
    assert _code(e) == abi.ERR_NOCONVERGE
    ctx.close()


def test_calls_before_the_tree(gpu):
    s = setups.uniform_box(nx=8, seed=2)
    ctx = sph.Context(s["params"])
    xyzh = s["xyzh"].copy()
    with pytest.raises(sph.PhsphError) as e:
        ctx.densityiterate(1, xyzh, s["vxyzu"], alphaind=s["alphaind"].copy())
    assert _code(e) == abi.ERR_STATE
    n = xyzh.shape[0]
    with pytest.raises(sph.PhsphError) as e:
        ctx.force(1, xyzh, s["vxyzu"], np.ones((n, 1), np.float32), s["alphaind"].copy(), np.ones((n, 3)))
    assert _code(e) == abi.ERR_STATE
    ctx.close()
