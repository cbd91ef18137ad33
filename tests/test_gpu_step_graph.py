"""Whole steps inside the library (phsph_step_graph) against the same steps
driven from the host through the individual entry points
(integrator.evolve_ind / evolve_global): block timesteps with the
reference's substep cadence (evolve.f90:200-222, utils_indtimesteps.f90:
114-223) and global steps (evolve.f90:105-160), each step's predictor,
drift, predict_sph, derivs and corrector (step_leapfrog.f90:91-843).

Every state array, the clock (time, tlast, nbinmax, istepfrac, ibinnow, dt)
and every step's counters must agree bit for bit after every step, with
the fixed loop chains replayed as HIP graphs and launched directly; the
graphs are instantiated once per context and updated in place afterwards.
The host-driven path is itself checked against the oracle by
test_gpu_step.py."""
import numpy as np
import pytest

from phantom_amd import integrator, setups, sph

pytestmark = pytest.mark.gpu

KEYS = ("its", "nactive", "nwake", "nvfloorp", "nvfloorps", "nvfloorc")


def _arrays_equal(se, sd, label):
    for k in integrator.State.ARRAYS:
        a, b = getattr(se, k), getattr(sd, k)
        assert (a is None) == (b is None), (label, k)
        if a is None:
            continue
        a = a.numpy() if not isinstance(a, np.ndarray) else a
        b = b.numpy() if not isinstance(b, np.ndarray) else b
        np.testing.assert_array_equal(a, b, err_msg=f"{label} {k}")


def _ind_state(setup):
    p = setup["params"]
    s = integrator.State(p, setup["xyzh"], setup["vxyzu"], setup["alphaind"], Bevol=setup.get("Bevol"),
                         dustevol=setup.get("dustevol"))
    s.enable_bins(0)
    ctx = sph.Context(p)
    integrator.init_ind(ctx, s)
    ctx.close()
    return s


def _level_bins(s, setup, nbinmax):
    """replace the bins of an initialised state by the setup's levels (twas
    as init_step sets it for those bins at time 0)"""
    b = setup["bins"]["ibin_old"]
    s.ibin[:] = b
    s.ibin_old[:] = b
    s.ibin_wake[:] = 0
    s.iphase[:] = 1
    s.nbinmax = nbinmax
    dtmax = s.params.dtmax
    s.twas[:] = 0.5 * (dtmax / 2.0 ** b.astype(np.float64))


def _cases():
    return {
        # bins from the first force call (velocity field, jittered box)
        "box16_natural_bins": lambda: (setups.uniform_box(nx=16, jitter=0.1, seed=21, isothermal=False, vamp=0.5),
                                       0.05, None),
        # MHD + dust with four populated bins (active fractions 1/64 .. 1)
        "mhd_dust_12_levels4": lambda: (setups.with_levels(setups.mhd_dust_box(nx=12, jitter=0.05), setups.LEVELS_4,
                                                           nbinmax=6, istepfrac=0), None, 6),
    }


@pytest.mark.parametrize("name", list(_cases()))
def test_block_substeps_bit_exact(name, gpu):
    setup, dtmax, levels = _cases()[name]()
    p = setup["params"]
    p.ind_timesteps = 1
    if dtmax is not None:
        p.dtmax = dtmax
    s0 = _ind_state(setup)
    if levels is not None:
        _level_bins(s0, setup, levels)
    nsub = 12
    se = s0.copy(device=True)   # host-driven path (device arrays: no staging)
    sg = s0.copy(device=True)   # graphs
    sl = s0.copy(device=True)   # direct launches
    ce, cg, cl = sph.Context(p), sph.Context(p), sph.Context(p)
    te = tle = tg = tlg = tl = tll = 0.0
    ninst = 0
    print(f"\n[case {name}] n={s0.n} nbinmax={s0.nbinmax}", flush=True)
    for k in range(nsub):
        te, tle, oe = integrator.evolve_ind(ce, se, 1, time=te, tlast=tle)
        tg, tlg, _, og = integrator.evolve_graph(cg, sg, 1, time=tg, tlast=tlg, graphs=True)
        tl, tll, _, ol = integrator.evolve_graph(cl, sl, 1, time=tl, tlast=tll, graphs=False)
        label = f"substep {k + 1} (active {oe[0]['nactive']}, its {oe[0]['its']})"
        print(f"  {label}", flush=True)
        for o in (og[0], ol[0]):
            assert all(o[key] == oe[0][key] for key in KEYS), (label, o, oe[0])
            assert o["dtnew"] == oe[0]["dtnew"] and o["dterr"] == oe[0]["dterr"], label
        assert (tg, tlg) == (te, tle) and (tl, tll) == (te, tle), label
        for s in (sg, sl):
            assert (s.nbinmax, s.istepfrac, s.ibinnow) == (se.nbinmax, se.istepfrac, se.ibinnow), label
        _arrays_equal(se, sg, label + " graphs")
        _arrays_equal(se, sl, label + " direct")
        assert og[0]["ngraph"] == 1 + og[0]["its"] and ol[0]["ngraph"] == 0
        ninst += og[0]["ninst"]
    # one instantiation per chain; every later step updates them in place
    assert ninst == 2
    for c in (ce, cg, cl):
        c.close()


def test_block_substeps_full_size_mhd_dust(gpu):
    """The 128^3 MHD + one-fluid dust box (2,097,152 particles) on four
    populated bins (nbinmax 6): eight substeps, active fractions from 1/64
    to about 1/8, graph replay against the host-driven loop."""
    setup = setups.with_levels(setups.mhd_dust_box(nx=128, jitter=0.05), setups.LEVELS_4, nbinmax=6, istepfrac=0)
    s0 = _ind_state(setup)
    _level_bins(s0, setup, 6)
    se, sg = s0.copy(device=True), s0.copy(device=True)
    del s0
    p = setup["params"]
    ce, cg = sph.Context(p), sph.Context(p)
    te = tle = tg = tlg = 0.0
    fractions = []
    for k in range(8):
        te, tle, oe = integrator.evolve_ind(ce, se, 1, time=te, tlast=tle)
        tg, tlg, _, og = integrator.evolve_graph(cg, sg, 1, time=tg, tlast=tlg, graphs=True)
        label = f"substep {k + 1} (active {oe[0]['nactive']})"
        print(f"  {label}", flush=True)
        assert all(og[0][key] == oe[0][key] for key in KEYS), (label, og[0], oe[0])
        assert (tg, tlg) == (te, tle), label
        assert (sg.nbinmax, sg.istepfrac, sg.ibinnow) == (se.nbinmax, se.istepfrac, se.ibinnow), label
        _arrays_equal(se, sg, label)
        fractions.append(oe[0]["nactive"] / setup["xyzh"].shape[0])
    # substep 1: bin 6 only (1/64); substep 8: bin 3 (7/64) and what remains in bins 3..6
    assert min(fractions) <= 1 / 64 + 1e-3 and max(fractions) >= 7 / 64
    ce.close()
    cg.close()


def test_global_steps_bit_exact(gpu):
    setup = setups.shock_tube(ny=8)
    p = setup["params"]
    s0 = integrator.State(p, setup["xyzh"], setup["vxyzu"], setup["alphaind"])
    ctx = sph.Context(p)
    integrator.init_global(ctx, s0)
    ctx.close()
    se, sg = s0.copy(device=True), s0.copy(device=True)
    ce, cg = sph.Context(p), sph.Context(p)
    te = tg = 0.0
    dte = dtg = None
    for k in range(4):
        te, dte, oe = integrator.evolve_global(ce, se, 1, time=te, dt=dte)
        tg, _, dtg, og = integrator.evolve_graph(cg, sg, 1, time=tg, dt=dtg)
        label = f"step {k + 1}"
        assert all(og[0][key] == oe[0][key] for key in ("its", "nwake", "nvfloorp", "nvfloorps", "nvfloorc")), label
        assert og[0]["dterr"] == oe[0]["dterr"] and og[0]["dtnew"] == oe[0]["dtnew"], label
        assert (tg, dtg) == (te, dte), label
        assert (sg.dtcourant, sg.dtforce) == (se.dtcourant, se.dtforce), label
        _arrays_equal(se, sg, label)
    ce.close()
    cg.close()


def test_reevaluation_bit_exact(gpu):
    """A velocity error above tolv: the reset and derivs(icall = 2) run
    after the corrector (maxits = 1 leaves the loop without an error)."""
    setup = setups.uniform_box(nx=12, jitter=0.1, seed=3, isothermal=False, vamp=0.5)
    p = setup["params"]
    s0 = integrator.State(p, setup["xyzh"], setup["vxyzu"], setup["alphaind"])
    ctx = sph.Context(p)
    integrator.init_global(ctx, s0)
    ctx.close()
    se, sg = s0.copy(device=True), s0.copy(device=True)
    ce, cg = sph.Context(p), sph.Context(p)
    dt = min(s0.dtforce, s0.dtcourant, p.dtmax)
    te, dte, oe = integrator.evolve_global(ce, se, 2, time=0.0, dt=dt, tolv=1e-30, maxits=1)
    tg, _, dtg, og = integrator.evolve_graph(cg, sg, 2, time=0.0, dt=dt, tolv=1e-30, maxits=1)
    assert (tg, dtg) == (te, dte)
    for a, b in zip(oe, og):
        assert a["its"] == b["its"] == 1 and a["dterr"] == b["dterr"]
    _arrays_equal(se, sg, "re-evaluated")
    ce.close()
    cg.close()


def test_step_graph_needs_device_arrays(gpu):
    setup = setups.uniform_box(nx=8)
    p = setup["params"]
    s = integrator.State(p, setup["xyzh"], setup["vxyzu"], setup["alphaind"])
    ctx = sph.Context(p)
    with pytest.raises(ValueError):
        integrator.evolve_graph(ctx, s, 1)
    # a host pointer smuggled into the state is refused by the library
    sd = s.copy(device=True)
    st = integrator.step_state(sd)
    st.vpred = s.vpred.ctypes.data
    from phantom_amd import abi
    k = abi.StepClock()
    k.dt = 1e-4
    with pytest.raises(sph.PhsphError, match="device memory"):
        ctx.step_graph(st, k, 1)
    ctx.close()
