"""The autotuner of the neighbour-list kernels: on the first large density
call it times every configuration (candidate leaves per chunk, wavefronts
per block, squared separations stored or recomputed, one group per
wavefront or persistent wavefronts on per-XCD work queues) and keeps the
fastest.
Every configuration must give bit-identical outputs, so the choice never
changes a result: each one is fixed in turn and the whole derivative chain
(tree, density, cons2prim, force) compared bit for bit with the first; the
oracle agreement of the tuned run is the parity tests' own."""
import numpy as np
import pytest

from phantom_amd import setups, sph
from tests import parity

pytestmark = pytest.mark.gpu

CONFIGS = [(cl, lw, r2, q) for cl in (8, 16) for lw in (2, 4) for r2 in (0, 1) for q in (0, 1)]


def _chain(setup, cfg=None, calls=1):
    """tree + density (calls times from the same start: the first call on a
    context may grow the list capacity, which postpones tuning) + cons2prim
    + force"""
    p = setup["params"]
    ctx = sph.Context(p)
    if cfg is not None:
        ctx.set_tuning(*cfg)
    for _ in range(calls):
        xyzh = setup["xyzh"].copy()
        alphaind = setup["alphaind"].copy()
        ctx.maketree(xyzh)
        d = ctx.densityiterate(1, xyzh, setup["vxyzu"], alphaind=alphaind)
    nneigh, nits = ctx.get_particle_stats(xyzh.shape[0])
    c = ctx.cons2prim(xyzh, setup["vxyzu"], d["dvdx"], alphaind)
    f = ctx.force(1, xyzh, setup["vxyzu"], d["gradh"], alphaind, c["eos_vars"])
    tune = ctx.get_tuning()
    ctx.close()
    out = dict(h=xyzh[:, 3], gradh=d["gradh"], divv=d["divcurlv"], dvdx=d["dvdx"], nneigh=nneigh, nits=nits,
               f=f["fxyzu"], dtc=f["dtcourant"], dtf=f["dtforce"])
    return out, tune


@pytest.fixture(scope="module", params=["box64_cubic", "shock_tube_ny24_quintic"])
def case(request, gpu):
    if request.param == "box64_cubic":
        s = setups.uniform_box(nx=64, isothermal=False)
    else:
        s = setups.shock_tube(ny=24)
    print(f"\n[case {request.param}] n={s['xyzh'].shape[0]}", flush=True)
    return request.param, s


def test_autotune_picks_fastest(case):
    name, s = case
    _, tune = _chain(s, calls=2)
    assert tune["tuned"]
    tried = tune["tried"]
    assert len(tried) == len(CONFIGS)
    assert {(t["cl"], t["lw"], t["r2"], t["q"]) for t in tried} == set(CONFIGS)
    assert all(t["ms"] > 0 for t in tried)
    best = min(tried, key=lambda t: t["ms"])
    assert (tune["cl"], tune["lw"], tune["r2"], tune["q"]) == (best["cl"], best["lw"], best["r2"], best["q"])
    print(f"[{name}] tuned " + ", ".join(f"cl{t['cl']}/lw{t['lw']}/r2{t['r2']}/q{t['q']}:{t['ms']:.3f}ms"
                                         for t in tried), flush=True)


def test_configurations_bit_identical(case):
    name, s = case
    ref, t0 = _chain(s, CONFIGS[0])
    assert not t0["tuned"] and (t0["cl"], t0["lw"], t0["r2"], t0["q"]) == CONFIGS[0]
    for cfg in CONFIGS[1:]:
        out, _ = _chain(s, cfg)
        for k in ref:
            np.testing.assert_array_equal(np.asarray(out[k]), np.asarray(ref[k]), err_msg=f"{cfg} {k}")
    # and the tuned run against the oracle's neighbour counts
    o = parity.run_oracle(s)
    np.testing.assert_array_equal(ref["nneigh"], o["dens"]["nneigh"])


def test_set_tuning_rejects_unknown(gpu):
    ctx = sph.Context(setups.uniform_box(nx=8)["params"])
    with pytest.raises(sph.PhsphError):
        ctx.set_tuning(12, 4, 0)
    with pytest.raises(sph.PhsphError):
        ctx.set_tuning(16, 3, 0)
    with pytest.raises(sph.PhsphError):
        ctx.set_tuning(16, 4, 0, 2)
    ctx.set_tuning(0)  # tune again on the next large call
    assert not ctx.get_tuning()["tuned"]
    ctx.close()


def test_search_radius_adapts_results_unchanged(gpu):
    """The first search radius of the density pass narrows from call to call
    while no particle outgrows its list.  Neighbour counts and iteration
    counts stay those of the first call exactly; h, gradh, the velocity
    gradients and the forces agree to round-off (the group walk lists the
    candidate leaves in an order that depends on the radius, so the sums add
    the same terms in another order)."""
    s = setups.uniform_box(nx=48, isothermal=False)
    p = s["params"]
    ctx = sph.Context(p)
    outs, radii = [], []
    for k in range(4):
        radii.append(ctx.get_tuning()["search_radius"])
        xyzh = s["xyzh"].copy()
        alphaind = s["alphaind"].copy()
        ctx.maketree(xyzh)
        d = ctx.densityiterate(1, xyzh, s["vxyzu"], alphaind=alphaind)
        nneigh, nits = ctx.get_particle_stats(xyzh.shape[0])
        c = ctx.cons2prim(xyzh, s["vxyzu"], d["dvdx"], alphaind)
        f = ctx.force(1, xyzh, s["vxyzu"], d["gradh"], alphaind, c["eos_vars"])
        outs.append(dict(h=xyzh[:, 3].copy(), gradh=d["gradh"].copy(), divv=d["divcurlv"].copy(),
                         dvdx=d["dvdx"].copy(), nneigh=nneigh, nits=nits, f=f["fxyzu"].copy()))
    ctx.close()
    print(f"\n  search radii {radii}", flush=True)
    assert radii[0] == 1.1 and min(radii) < 1.1
    for o in outs[1:]:
        for k in ("nneigh", "nits"):
            np.testing.assert_array_equal(o[k], outs[0][k], err_msg=k)
        assert np.max(np.abs(o["h"] - outs[0]["h"]) / outs[0]["h"]) <= 1e-14
        for k in ("gradh", "divv", "dvdx", "f"):
            assert parity.rel_rms(np.asarray(o[k], np.float64), np.asarray(outs[0][k], np.float64)) <= 1e-7, k


def test_tuned_once_when_budget_holds_capacity(gpu, monkeypatch):
    """A list budget that holds the capacity below the longest trial list:
    the first attempt's lists overflow on every call, so a tuning run over
    them cannot be bettered by a later call; it is kept, not repeated on
    every call, and the outputs are those of an unconstrained context."""
    s = setups.uniform_box(nx=64, isothermal=False)
    n = s["xyzh"].shape[0]
    # 16-entry lists (3 words per entry) fit, 32-entry lists do not
    monkeypatch.setenv("PHSPH_CAPS", f"dlist=16,budget_kib={16 * n * 3 * 4 // 1024}")
    p = s["params"]
    ctx = sph.Context(p)
    tried = []
    for _ in range(3):
        xyzh = s["xyzh"].copy()
        alphaind = s["alphaind"].copy()
        ctx.maketree(xyzh)
        d = ctx.densityiterate(1, xyzh, s["vxyzu"], alphaind=alphaind)
        t = ctx.get_tuning()
        assert t["tuned"]
        tried.append([(u["cl"], u["lw"], u["r2"], u["q"], u["ms"]) for u in t["tried"]])
    ctx.close()
    assert tried[0] == tried[1] == tried[2]
    monkeypatch.delenv("PHSPH_CAPS")
    ref, _ = _chain(s, calls=1)
    # the overflowing particles are walked again into longer lists (and the
    # search radius may have narrowed since the first call): h agrees to
    # round-off, the sums to their summation order
    assert np.max(np.abs(xyzh[:, 3] - ref["h"]) / ref["h"]) <= 1e-14
    assert parity.rel_rms(np.asarray(d["gradh"], np.float64), np.asarray(ref["gradh"], np.float64)) <= 1e-7
