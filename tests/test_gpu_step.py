"""Parity of the device integrator loops (predictor, drift, predict_sph,
corrector, reset) with the oracle, and of whole leapfrog steps.

Each loop is a per-particle update whose products are rounded before they
are summed, so the device result must equal the oracle's bit for bit; the
velocity-error sum v2sum is reduced in a different order and is compared to
1e-13.  Whole steps (tree, density, cons2prim, force, loops) follow the FP64
tolerance of the other parity tests, accumulated over the steps taken.
The reference's own step test (test_step.F90:63-160) runs on the device with
the arrays resident in HBM."""
import numpy as np
import pytest

from phantom_amd import abi, device, integrator, setups, sph
from tests import stepping

pytestmark = pytest.mark.gpu


def _params(maxvxyzu=4, ind=False, mhd=False, dust=False, dtmax=0.1):
    p = abi.default_params(abi.KERNEL_CUBIC)
    p.maxvxyzu = maxvxyzu
    p.ieos = 2 if maxvxyzu == 4 else 1
    p.ind_timesteps = 1 if ind else 0
    p.mhd = 1 if mhd else 0
    if dust:
        p.use_dustfrac, p.ndustsmall, p.idrag, p.K_code = 1, 1, 2, 1.0
    p.dtmax = dtmax
    p.massoftype_gas = 1e-3
    return p


def _random_state(p, n=3000, seed=1, nbinmax=4):
    rng = np.random.default_rng(seed)
    xyzh = np.empty((n, 4))
    xyzh[:, :3] = rng.uniform(-0.5, 0.5, (n, 3))
    xyzh[:, 3] = rng.uniform(0.01, 0.05, n)
    xyzh[rng.random(n) < 0.05, 3] = 0.0   # dead particles
    xyzh[rng.random(n) < 0.02, 3] = -0.02  # accreted
    v = np.zeros((n, p.maxvxyzu))
    v[:, :3] = rng.normal(0.0, 1.0, (n, 3))
    if p.maxvxyzu == 4:
        v[:, 3] = rng.uniform(0.0, 2.0, n)
    alpha = np.zeros((n, 3), np.float32)
    alpha[:, 0] = rng.uniform(0.0, 1.0, n)
    alpha[:, 1] = np.where(rng.random(n) < 0.5, 0.0, rng.uniform(0.0, 1.0, n))
    B = rng.normal(0.0, 1.0, (n, 4)) if p.mhd else None
    de = rng.uniform(0.0, 1.0, n) if p.use_dustfrac else None
    s = integrator.State(p, xyzh, v, alpha, Bevol=B, dustevol=de)
    s.fxyzu[...] = rng.normal(0.0, 3.0, s.fxyzu.shape)
    s.vpred[...] = s.vxyzu + rng.normal(0.0, 1e-3, s.vxyzu.shape)
    s.divcurlv[:, 0] = rng.normal(0.0, 2.0, n).astype(np.float32)
    s.eos_vars[:, 1] = rng.uniform(0.5, 2.0, n)
    if p.mhd:
        s.dBevol[...] = rng.normal(0.0, 1.0, (n, 4))
        s.Bpred[...] = s.Bevol
    if p.use_dustfrac:
        s.ddustevol[...] = rng.normal(0.0, 1.0, n)
        s.dustpred[...] = s.dustevol
    if p.ind_timesteps:
        s.enable_bins(nbinmax, ibin=rng.integers(0, nbinmax + 1, n))
        s.iphase[:] = np.where(rng.random(n) < 0.5, 1, -1)
        s.twas[:] = rng.uniform(0.9, 1.1, n)
        s.ibin_old[:] = rng.integers(0, nbinmax + 1, n)
        wake = rng.random(n) < 0.2
        s.ibin_wake[wake] = rng.integers(0, nbinmax + 3, int(wake.sum()))
    return s


def _copy_state(s):
    import copy
    c = copy.copy(s)
    for k, v in vars(s).items():
        if isinstance(v, np.ndarray):
            setattr(c, k, v.copy())
    return c


def _same(a, b, keys):
    for k in keys:
        x, y = getattr(a, k), getattr(b, k)
        if x is None:
            assert y is None
            continue
        np.testing.assert_array_equal(x, y, err_msg=k)


CASES = {
    "global_thermal": dict(maxvxyzu=4),
    "global_isothermal": dict(maxvxyzu=3),
    "global_mhd_dust": dict(maxvxyzu=4, mhd=True, dust=True),
    "ind_thermal": dict(maxvxyzu=4, ind=True),
    "ind_mhd_dust": dict(maxvxyzu=3, ind=True, mhd=True, dust=True),
}
ALL = ["xyzh", "vxyzu", "vpred", "alphaind", "Bevol", "Bpred", "dustevol", "dustpred", "ibin", "ibin_old",
       "ibin_wake", "twas"]


@pytest.fixture(scope="module", params=list(CASES.keys()))
def case(request, gpu):
    p = _params(**CASES[request.param])
    print(f"\n[case {request.param}]", flush=True)
    return request.param, p


def test_loops_bit_exact(case):
    name, p = case
    sg = _random_state(p)
    so = _copy_state(sg)
    ctx = sph.Context(p)
    o = stepping.OracleBackend(p)
    ufloor = 0.5 if p.maxvxyzu == 4 else 0.0
    t, dt = 1.0, 0.0125
    for side, b, s in (("gpu", ctx, sg), ("oracle", o, so)):
        a = integrator._arrays(s)
        s.r_pred = b.step_predictor(t, dt, s.xyzh, s.vxyzu, s.fxyzu, ibin=s.ibin, ibin_old=s.ibin_old, twas=s.twas,
                                    ufloor=ufloor, **a)
    _same(sg, so, ALL)
    assert sg.r_pred == so.r_pred
    if p.maxvxyzu == 4:
        assert sg.r_pred > 0  # the floor was exercised
    for b, s in ((ctx, sg), (o, so)):
        b.substep_sph(dt, s.xyzh, s.vxyzu)
    _same(sg, so, ALL)
    for b, s in ((ctx, sg), (o, so)):
        s.r_ps = b.step_predict_sph(t + dt, dt, s.xyzh, s.vxyzu, s.fxyzu, s.vpred, s.divcurlv, s.eos_vars,
                                    s.alphaind, Bevol=s.Bevol, dBevol=s.dBevol, Bpred=s.Bpred,
                                    dustevol=s.dustevol, ddustevol=s.ddustevol, dustpred=s.dustpred,
                                    iphase=s.iphase, twas=s.twas, ufloor=ufloor)
    _same(sg, so, ALL)
    assert sg.r_ps == so.r_ps
    for b, s in ((ctx, sg), (o, so)):
        a = integrator._arrays(s)
        s.r_corr = b.step_corrector(t + dt, dt, s.xyzh, s.vxyzu, s.fxyzu, vpred=s.vpred, ibin=s.ibin,
                                    ibin_wake=s.ibin_wake, nbinmax=s.nbinmax, twas=s.twas, ufloor=ufloor, **a)
    _same(sg, so, ALL)
    rg, ro = sg.r_corr, so.r_corr
    assert (rg["np"], rg["nwake"], rg["nvfloor"]) == (ro["np"], ro["nwake"], ro["nvfloor"])
    assert rg["errmax"] == ro["errmax"]
    assert abs(rg["v2sum"] - ro["v2sum"]) <= 1e-13 * max(ro["v2sum"], 1e-300)
    if p.ind_timesteps:
        assert rg["nwake"] > 0 and rg["np"] == 0
    else:
        assert rg["np"] == int(np.sum(sg.xyzh[:, 3] >= np.finfo(np.float64).tiny))
    for b, s in ((ctx, sg), (o, so)):
        b.step_reset(dt, s.vxyzu, s.fxyzu, s.vpred, Bevol=s.Bevol, dBevol=s.dBevol, Bpred=s.Bpred,
                     dustevol=s.dustevol, ddustevol=s.ddustevol, dustpred=s.dustpred, iphase=s.iphase)
    _same(sg, so, ALL)
    ctx.close()


def test_loops_on_device_arrays(gpu):
    """The same loops on arrays resident in HBM (used in place)."""
    p = _params(maxvxyzu=4, mhd=True)
    sh = _random_state(p, seed=5)
    so = _copy_state(sh)
    dev = {k: device.DeviceArray.from_numpy(v) for k, v in vars(sh).items()
           if isinstance(v, np.ndarray) and k in ("xyzh", "vxyzu", "fxyzu", "vpred", "Bevol", "dBevol", "Bpred",
                                                 "divcurlv", "eos_vars", "alphaind")}
    ctx = sph.Context(p)
    o = stepping.OracleBackend(p)
    t, dt = 0.0, 0.01
    ctx.step_predictor(t, dt, dev["xyzh"], dev["vxyzu"], dev["fxyzu"], Bevol=dev["Bevol"], dBevol=dev["dBevol"])
    o.step_predictor(t, dt, so.xyzh, so.vxyzu, so.fxyzu, Bevol=so.Bevol, dBevol=so.dBevol)
    ctx.substep_sph(dt, dev["xyzh"], dev["vxyzu"])
    o.substep_sph(dt, so.xyzh, so.vxyzu)
    ctx.step_predict_sph(t + dt, dt, dev["xyzh"], dev["vxyzu"], dev["fxyzu"], dev["vpred"], dev["divcurlv"],
                         dev["eos_vars"], dev["alphaind"], Bevol=dev["Bevol"], dBevol=dev["dBevol"],
                         Bpred=dev["Bpred"])
    o.step_predict_sph(t + dt, dt, so.xyzh, so.vxyzu, so.fxyzu, so.vpred, so.divcurlv, so.eos_vars, so.alphaind,
                       Bevol=so.Bevol, dBevol=so.dBevol, Bpred=so.Bpred)
    rg = ctx.step_corrector(t + dt, dt, dev["xyzh"], dev["vxyzu"], dev["fxyzu"], vpred=dev["vpred"],
                            Bevol=dev["Bevol"], dBevol=dev["dBevol"])
    ro = o.step_corrector(t + dt, dt, so.xyzh, so.vxyzu, so.fxyzu, vpred=so.vpred, Bevol=so.Bevol, dBevol=so.dBevol)
    assert rg["errmax"] == ro["errmax"] and rg["np"] == ro["np"]
    for k in ("xyzh", "vxyzu", "vpred", "Bevol", "Bpred", "alphaind"):
        np.testing.assert_array_equal(dev[k].numpy(), getattr(so, k), err_msg=k)
    ctx.close()


def test_reference_test_step_lattice_on_device(gpu):
    """test_step.F90:63-160 (32^3 lattice, state resident in HBM)."""
    nx = 32
    pos, dx = setups._cubic_lattice(nx, nx, nx, -0.5, 0.5, -0.5, 0.5, -0.5, 0.5)
    p = _params(maxvxyzu=4)
    p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax = -0.5, 0.5, -0.5, 0.5, -0.5, 0.5
    p.gamma, p.polyk = 5.0 / 3.0, 0.0
    p.alpha = p.alphau = p.alphaB = 0.0
    p.tolh = 1e-5
    n = pos.shape[0]
    rhozero = 7.5
    p.massoftype_gas = rhozero / n
    p.maxp = n
    hzero = p.hfact * (p.massoftype_gas / rhozero) ** (1.0 / 3.0)
    nsteps, dt = 10, 0.2
    p.dtmax = dt
    xyzh = np.empty((n, 4))
    xyzh[:, :3] = pos
    xyzh[:, 3] = p.hfact * dx
    v = np.zeros((n, 4))
    v[:, :3] = 1.0
    s = integrator.State(p, xyzh, v, np.zeros((n, 3), np.float32), device=True)
    ctx = sph.Context(p)
    integrator.init_global(ctx, s)
    t = 0.0
    for k in range(nsteps):
        r = integrator.step(ctx, s, t, dt)
        t = t + dt
        h = s.xyzh.numpy()[:, 3]
        f = s.fxyzu.numpy()
        print(f"  step {k + 1}: max |h/hzero - 1| {np.max(np.abs(h - hzero)) / hzero:.2e}", flush=True)
        assert np.max(np.abs(h - hzero) / hzero) <= 3e-4
        assert np.all(np.abs(f) <= np.finfo(np.float64).tiny)
        assert r["its"] == 1
    x = s.xyzh.numpy()[:, :3]
    wrapped = (x + 0.5) % 1.0 - 0.5
    d = np.abs(wrapped - pos)
    d = np.minimum(d, 1.0 - d)
    assert d.max() <= 1e-12
    ctx.close()


def _err(g, o):
    """max |g - o| over the rms of o; components 1-3 of a vector array are
    measured against the rms length of the vector."""
    g = np.asarray(g, np.float64)
    o = np.asarray(o, np.float64)
    out = []
    if o.ndim == 2 and o.shape[1] >= 3:
        scale = np.sqrt(np.mean(np.sum(o[:, :3] ** 2, axis=1)))
        out.append(np.max(np.abs(g[:, :3] - o[:, :3])) / (scale if scale > 0 else 1.0))
        rest = range(3, o.shape[1])
    else:
        rest = range(o.shape[1]) if o.ndim == 2 else [None]
    for c in rest:
        a = o if c is None else o[:, c]
        b = g if c is None else g[:, c]
        scale = np.sqrt(np.mean(a * a))
        out.append(np.max(np.abs(b - a)) / (scale if scale > 0 else 1.0))
    return max(out)


def _compare_states(sg, so, tol, label):
    for k in ("xyzh", "vxyzu", "fxyzu"):
        e = _err(getattr(sg, k), getattr(so, k))
        print(f"  {label} {k}: {e:.2e}", flush=True)
        assert e <= tol, (label, k)
    ag, ao = sg.alphaind[:, 0].astype(np.float64), so.alphaind[:, 0].astype(np.float64)
    assert np.max(np.abs(ag - ao)) <= 1e-6, label


def test_shock_tube_steps_match_oracle(gpu):
    """Three global steps of the shock tube: tree, density, force and the
    loops, device against oracle."""
    setup = setups.shock_tube(ny=8)
    p = setup["params"]
    sg = integrator.State(p, setup["xyzh"], setup["vxyzu"], setup["alphaind"])
    so = integrator.State(p, setup["xyzh"], setup["vxyzu"], setup["alphaind"])
    ctx = sph.Context(p)
    o = stepping.OracleBackend(p)
    dtg = integrator.init_global(ctx, sg)
    dto = integrator.init_global(o, so)
    assert abs(dtg - dto) <= 1e-12 * dto
    t = 0.0
    for k in range(3):
        dt = min(so.dtforce, so.dtcourant, p.dtmax)
        rg = integrator.step(ctx, sg, t, dt)
        ro = integrator.step(o, so, t, dt)
        t = t + dt
        assert rg["its"] == ro["its"]
        _compare_states(sg, so, 1e-10, f"step {k + 1}")
        assert abs(sg.dtcourant - so.dtcourant) <= 1e-10 * so.dtcourant
    ctx.close()


def test_block_timestep_cycle_matches_oracle(gpu):
    """One dtmax of block timesteps on a jittered box with a velocity
    field: bins from force, active sets, wakes, twas."""
    setup = setups.uniform_box(nx=16, jitter=0.1, seed=21, isothermal=False, vamp=0.5)
    p = setup["params"]
    p.ind_timesteps = 1
    p.dtmax = 0.05
    sg = integrator.State(p, setup["xyzh"], setup["vxyzu"], setup["alphaind"])
    sg.enable_bins(0)
    so = _copy_state(sg)
    ctx = sph.Context(p)
    o = stepping.OracleBackend(p)
    integrator.init_ind(ctx, sg)
    integrator.init_ind(o, so)
    assert sg.nbinmax == so.nbinmax
    np.testing.assert_array_equal(sg.ibin, so.ibin)
    print(f"  nbinmax {sg.nbinmax}", flush=True)
    time = tlast = 0.0
    for k in range(64):
        tg, tlg, outg = integrator.evolve_ind(ctx, sg, 1, time=time, tlast=tlast)
        to, tlo, outo = integrator.evolve_ind(o, so, 1, time=time, tlast=tlast)
        assert (tg, tlg) == (to, tlo)
        time, tlast = tg, tlg
        assert outg[0]["nactive"] == outo[0]["nactive"] and outg[0]["nwake"] == outo[0]["nwake"]
        np.testing.assert_array_equal(sg.ibin, so.ibin)
        np.testing.assert_array_equal(sg.ibin_wake, so.ibin_wake)
        np.testing.assert_array_equal(sg.iphase, so.iphase)
        np.testing.assert_array_equal(sg.twas, so.twas)
        _compare_states(sg, so, 1e-10, f"substep {k + 1} (active {outg[0]['nactive']})")
        if sg.istepfrac == 0:
            break
    ctx.close()
