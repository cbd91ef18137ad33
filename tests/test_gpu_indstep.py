"""Block (individual) timesteps on the GPU against the oracle: the active
set (set_active_particles), and one force call with per-particle bins
(get_newbin + the Saitoh-Makino limiter), the neighbour wake flags and the
new nbinmax, all bit-exact; the forces of the active particles within the
FP64 tolerance."""
import numpy as np
import pytest

from oracle import oracle
from phantom_amd import setups, sph
from tests import parity

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12


def _cases():
    return {
        "box16_bins3_step2": lambda: setups.with_bins(setups.uniform_box(nx=16, seed=7, isothermal=False),
                                                      nbinmax=3, istepfrac=2, seed=5),
        "mhd_dust_12_bins4_step4": lambda: setups.with_bins(setups.mhd_dust_box(nx=12, jitter=0.05),
                                                            nbinmax=4, istepfrac=4, seed=9),
    }


@pytest.fixture(scope="module", params=list(_cases().keys()))
def pair(request, gpu):
    setup = _cases()[request.param]()
    print(f"\n[case {request.param}] n={setup['xyzh'].shape[0]}", flush=True)
    g = parity.run_gpu(setup)
    o = parity.run_oracle(setup)
    return request.param, setup, g, o


def test_set_active_particles(pair):
    name, setup, g, o = pair
    b = setup["bins"]
    ctx = g["ctx"]
    iph = np.zeros(setup["xyzh"].shape[0], np.int8)
    na, nl, bnow = ctx.set_active_particles(setup["xyzh"], iph, b["ibin_old"], b["nbinmax"], b["istepfrac"])
    iph_o, na_o, nl_o, bnow_o = oracle.set_active_particles(setup["xyzh"], b["ibin_old"], b["nbinmax"],
                                                            b["istepfrac"])
    np.testing.assert_array_equal(iph, iph_o)
    np.testing.assert_array_equal(iph, setup["iphase"])
    assert (na, nl, bnow) == (na_o, nl_o, bnow_o)
    assert 0 < na < nl


def test_bins_bit_exact(pair):
    name, setup, g, o = pair
    gf, of = g["force"], o["force"]
    np.testing.assert_array_equal(gf["ibin"], of["ibin"])
    np.testing.assert_array_equal(gf["ibin_wake"], of["ibin_wake"])
    assert gf["nbinmax"] == of["nbinmax"]
    assert gf["dtcourant"] == 1e29 and gf["dtforce"] == 1e29
    # the bins moved: some particles changed bin and some neighbours woke
    assert np.any(gf["ibin"] != setup["bins"]["ibin_old"])
    assert np.any(gf["ibin_wake"] != 0)


def test_active_particles_match(pair):
    name, setup, g, o = pair
    act = setup["iphase"] > 0
    np.testing.assert_array_equal(g["nneigh"][act], o["dens"]["nneigh"][act])
    assert np.all(g["nneigh"][~act] == 0)
    p = setup["params"]
    ncomp = 4 if p.maxvxyzu >= 4 else 3
    for c in range(ncomp):
        assert parity.rel_rms(g["force"]["fxyzu"][act, c], o["force"]["fxyzu"][act, c]) <= TOL_FP64, c
    assert parity.rel_rms(g["xyzh"][act, 3], o["xyzh"][act, 3]) <= TOL_FP64
