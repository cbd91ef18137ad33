"""phsph_step_corrector_noalloc (scratch in a fixed device array, nothing
allocated during the call) against phsph_step_corrector on the same
device-resident state: every array bit-identical and the reductions equal,
on two calls in a row, the second with fewer particles than the first, for
a global and a block-timestep configuration (MHD + dust).  Host arrays are
refused by the no-allocation entry point."""
import numpy as np
import pytest

from phantom_amd import device, integrator, sph
from tests.test_gpu_step import ALL, _params, _random_state

pytestmark = pytest.mark.gpu

CASES = {
    "global_thermal": dict(maxvxyzu=4),
    "ind_mhd_dust": dict(maxvxyzu=3, ind=True, mhd=True, dust=True),
}


def _to_device(s):
    """A copy of an integrator state with every array in HBM."""
    import copy
    d = copy.copy(s)
    for k, v in vars(s).items():
        if isinstance(v, np.ndarray):
            setattr(d, k, device.DeviceArray.from_numpy(np.ascontiguousarray(v)))
    return d


def _host(d, keys):
    out = {}
    for k in keys:
        v = getattr(d, k, None)
        out[k] = None if v is None else v.numpy()
    return out


def _corrector(fn, s, t, dt, ufloor):
    a = integrator._arrays(s)
    return fn(t + dt, dt, s.xyzh, s.vxyzu, s.fxyzu, vpred=s.vpred, ibin=s.ibin, ibin_wake=s.ibin_wake,
              nbinmax=s.nbinmax, twas=s.twas, ufloor=ufloor, **a)


@pytest.mark.parametrize("name", list(CASES))
def test_noalloc_corrector_matches(name, gpu):
    p = _params(**CASES[name])
    ctx = sph.Context(p)
    ufloor = 0.5 if p.maxvxyzu == 4 else 0.0
    t, dt = 1.0, 0.0125
    for n, seed in ((3000, 1), (1000, 2)):  # the second call is smaller
        print(f"\n[case {name}] n={n} seed={seed}", flush=True)
        s = _random_state(p, n=n, seed=seed)
        d_ref, d_na = _to_device(s), _to_device(s)
        r_ref = _corrector(ctx.step_corrector, d_ref, t, dt, ufloor)
        r_na = _corrector(ctx.step_corrector_noalloc, d_na, t, dt, ufloor)
        assert r_na == r_ref
        h_ref, h_na = _host(d_ref, ALL), _host(d_na, ALL)
        for k in ALL:
            if h_ref[k] is None:
                assert h_na[k] is None
            else:
                np.testing.assert_array_equal(h_na[k], h_ref[k], err_msg=k)
    ctx.close()


def test_noalloc_refuses_host_arrays(gpu):
    p = _params(maxvxyzu=4)
    ctx = sph.Context(p)
    s = _random_state(p, n=500, seed=3)
    with pytest.raises(sph.PhsphError, match="device memory"):
        _corrector(ctx.step_corrector_noalloc, s, 1.0, 0.0125, 0.0)
    ctx.close()
