"""The oracle's leapfrog loops pinned by the reference's test and by the
mathematics of the integrator (CPU only).

* test_step.F90 (:63-160): a cubic lattice moving at v = (1,1,1) through a
  periodic unit box with zero pressure and no dissipation keeps h within
  3e-4 of hzero and zero force over 10 steps of dt = 0.2 (here also: the
  lattice comes back onto itself after two box crossings);
* constant acceleration: kick-drift-kick is exact for a constant force, so
  x = x0 + v0 t + a t^2/2 and v = v0 + a t to rounding, on one global step
  size and on block timesteps (each bin reaches the parabola at the end of
  its own step);
* harmonic force f = -x: the step is the kick-drift-kick map
  M = [[1 - dt^2/2, dt], [-dt (1 - dt^2/4), 1 - dt^2/2]] per component;
  n steps equal M^n computed independently, on one step size and per bin;
* predict_sph: h follows dh/dt = (h/3) div v to first order
  (h(1 + dt div v/3) exactly as an identity of dhdrho*rho = -h/3), and the
  Cullen-Dehnen decay relaxes alpha by 1/(1 + dt 0.1 c_s/h) per step;
* waking (step_leapfrog.f90:553-559): a woken particle moves to
  min(ibin_wake, nbinmax) and its twas is the half step of that bin.
"""
import numpy as np
import pytest

from phantom_amd import abi, integrator, setups
from tests import stepping


def _params(maxvxyzu=4, ind=False, dtmax=0.1):
    p = abi.default_params(abi.KERNEL_CUBIC)
    p.maxvxyzu = maxvxyzu
    p.ieos = 2 if maxvxyzu == 4 else 1
    p.ind_timesteps = 1 if ind else 0
    p.dtmax = dtmax
    p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax = -0.5, 0.5, -0.5, 0.5, -0.5, 0.5
    p.massoftype_gas = 1e-3
    return p


def _free_state(p, n=40, seed=3):
    rng = np.random.default_rng(seed)
    xyzh = np.empty((n, 4))
    xyzh[:, :3] = rng.uniform(-0.5, 0.5, (n, 3))
    xyzh[:, 3] = 0.05
    v = np.zeros((n, p.maxvxyzu))
    v[:, :3] = rng.normal(0.0, 1.0, (n, 3))
    if p.maxvxyzu == 4:
        v[:, 3] = rng.uniform(1.0, 2.0, n)
    s = integrator.State(p, xyzh, v, np.zeros((n, 3), np.float32))
    return s


def test_reference_test_step_lattice():
    """test_step.F90:63-160 on a 16^3 lattice (the reference uses 50^3)."""
    nx = 16
    pos, dx = setups._cubic_lattice(nx, nx, nx, -0.5, 0.5, -0.5, 0.5, -0.5, 0.5)
    p = _params(maxvxyzu=4)
    p.gamma = 5.0 / 3.0
    p.polyk = 0.0
    p.alpha = p.alphau = p.alphaB = 0.0
    p.tolh = 1e-5
    n = pos.shape[0]
    rhozero = 7.5
    p.massoftype_gas = rhozero / n  # totmass = rhozero/(dxbound*dybound*dzbound) in a unit box
    hzero = p.hfact * (p.massoftype_gas / rhozero) ** (1.0 / 3.0)
    nsteps = 10
    dt = 2.0 / nsteps
    p.dtmax = dt
    p.maxp = n
    xyzh = np.empty((n, 4))
    xyzh[:, :3] = pos
    xyzh[:, 3] = p.hfact * dx
    v = np.zeros((n, 4))
    v[:, :3] = 1.0
    s = integrator.State(p, xyzh, v, np.zeros((n, 3), np.float32))
    o = stepping.OracleBackend(p)
    integrator.init_global(o, s)
    t = 0.0
    for _ in range(nsteps):
        r = integrator.step(o, s, t, dt)
        t = t + dt
        assert np.max(np.abs(s.xyzh[:, 3] - hzero) / hzero) <= 3e-4
        assert np.all(np.abs(s.fxyzu) <= np.finfo(np.float64).tiny)
        assert r["its"] == 1
    # two crossings of the unit box: the lattice maps onto itself
    wrapped = (s.xyzh[:, :3] + 0.5) % 1.0 - 0.5
    d = np.abs(wrapped - pos)
    d = np.minimum(d, 1.0 - d)
    assert d.max() <= 1e-12


@pytest.mark.parametrize("maxvxyzu", [3, 4])
def test_constant_acceleration_global(maxvxyzu):
    p = _params(maxvxyzu=maxvxyzu, dtmax=0.1)
    s = _free_state(p)
    x0, v0 = s.xyzh[:, :3].copy(), s.vxyzu.copy()
    a = np.array([0.3, -1.1, 0.7])
    o = stepping.OracleBackend(p)
    fn = stepping.constant_force(a, ufdot=0.25)
    integrator.init_global(o, s, derivs_fn=fn)
    dt, nsteps = 0.1, 7
    t, _, out = integrator.evolve_global(o, s, nsteps, dt=dt, derivs_fn=fn)
    T = nsteps * dt
    assert all(r["its"] == 1 for r in out)
    np.testing.assert_allclose(s.xyzh[:, :3], x0 + v0[:, :3] * T + 0.5 * a * T * T, rtol=0, atol=1e-13)
    np.testing.assert_allclose(s.vxyzu[:, :3], v0[:, :3] + a * T, rtol=0, atol=1e-13)
    if maxvxyzu == 4:
        np.testing.assert_allclose(s.vxyzu[:, 3], v0[:, 3] + 0.25 * T, rtol=0, atol=1e-13)


def _kdk(dt, n):
    m = np.array([[1 - dt * dt / 2, dt], [-dt * (1 - dt * dt / 4), 1 - dt * dt / 2]])
    return np.linalg.matrix_power(m, n)


def test_harmonic_global_is_kick_drift_kick():
    p = _params(maxvxyzu=3)
    s = _free_state(p)
    x0, v0 = s.xyzh[:, :3].copy(), s.vxyzu[:, :3].copy()
    o = stepping.OracleBackend(p)
    fn = stepping.harmonic_force()
    integrator.init_global(o, s, derivs_fn=fn)
    dt, nsteps = 0.05, 40
    t = 0.0
    for _ in range(nsteps):
        r = integrator.step(o, s, t, dt, derivs_fn=fn)
        assert r["its"] == 1
        t = t + dt
    M = _kdk(dt, nsteps)
    np.testing.assert_allclose(s.xyzh[:, :3], M[0, 0] * x0 + M[0, 1] * v0, rtol=0, atol=1e-12)
    np.testing.assert_allclose(s.vxyzu[:, :3], M[1, 0] * x0 + M[1, 1] * v0, rtol=0, atol=1e-12)


def _bins_state(p, nbinmax, seed=9):
    s = _free_state(p, n=48, seed=seed)
    rng = np.random.default_rng(seed)
    s.enable_bins(nbinmax, ibin=rng.integers(0, nbinmax + 1, s.n))
    return s


@pytest.mark.parametrize("force", ["constant", "harmonic"])
def test_block_timesteps_reach_each_bins_solution(force):
    nbinmax, dtmax = 3, 0.2
    p = _params(maxvxyzu=4, ind=True, dtmax=dtmax)
    s = _bins_state(p, nbinmax)
    ibin = s.ibin.copy()
    x0, v0 = s.xyzh[:, :3].copy(), s.vxyzu[:, :3].copy()
    o = stepping.OracleBackend(p)
    a = np.array([0.5, 0.25, -2.0])
    fn = stepping.constant_force(a) if force == "constant" else stepping.harmonic_force()
    t0 = 1.0  # init_step moves every particle to nbinmax only at time 0
    integrator.init_ind(o, s, time=t0, derivs_fn=fn)
    assert s.nbinmax == nbinmax
    np.testing.assert_array_equal(s.ibin, ibin)
    # twas starts at the half step of each bin (init_step)
    np.testing.assert_array_equal(s.twas, t0 + 0.5 * (dtmax / 2.0 ** ibin))
    time, tlast, out = integrator.evolve_ind(o, s, 2 ** nbinmax, time=t0, tlast=t0, derivs_fn=fn)
    assert abs(time - (t0 + dtmax)) <= 1e-15 and s.istepfrac == 0
    assert [r["nactive"] for r in out] == [int(np.sum(k % 2 ** (nbinmax - ibin) == 0)) for k in range(1, 9)]
    T = dtmax
    if force == "constant":
        np.testing.assert_allclose(s.xyzh[:, :3], x0 + v0 * T + 0.5 * a * T * T, rtol=0, atol=1e-13)
        np.testing.assert_allclose(s.vxyzu[:, :3], v0 + a * T, rtol=0, atol=1e-13)
    else:
        for b in range(nbinmax + 1):
            sel = ibin == b
            dtb = dtmax / 2 ** b
            M = _kdk(dtb, 2 ** b)
            np.testing.assert_allclose(s.xyzh[sel, :3], M[0, 0] * x0[sel] + M[0, 1] * v0[sel], rtol=0, atol=1e-13)
            np.testing.assert_allclose(s.vxyzu[sel, :3], M[1, 0] * x0[sel] + M[1, 1] * v0[sel], rtol=0,
                                       atol=1e-13)
    # every particle ends its step synchronised: twas at the next half step
    np.testing.assert_allclose(s.twas, t0 + T + 0.5 * dtmax / 2.0 ** ibin, rtol=0, atol=1e-14)


def test_predict_sph_h_and_alpha_closed_forms():
    p = _params(maxvxyzu=4)
    n = 64
    rng = np.random.default_rng(4)
    s = _free_state(p, n=n, seed=4)
    s.xyzh[:, 3] = rng.uniform(0.02, 0.08, n)
    s.divcurlv[:, 0] = rng.normal(0.0, 2.0, n).astype(np.float32)
    cs = rng.uniform(0.5, 2.0, n)
    s.eos_vars[:, 1] = cs
    s.alphaind[:, 0] = rng.uniform(0.0, 1.0, n).astype(np.float32)
    s.alphaind[:, 1] = np.where(np.arange(n) % 2 == 0, 0.0, 1.5).astype(np.float32)
    h0 = s.xyzh[:, 3].copy()
    a0 = s.alphaind[:, 0].astype(np.float64)
    o = stepping.OracleBackend(p)
    dt = 0.01
    o.step_predict_sph(dt, dt, s.xyzh, s.vxyzu, s.fxyzu, s.vpred, s.divcurlv, s.eos_vars, s.alphaind)
    divv = s.divcurlv[:, 0].astype(np.float64)
    np.testing.assert_allclose(s.xyzh[:, 3], h0 * (1.0 + dt * divv / 3.0), rtol=1e-14, atol=0)
    hnew = s.xyzh[:, 3]
    decay = a0 / (1.0 + dt * 0.1 * cs / hnew)
    even = np.arange(n) % 2 == 0
    np.testing.assert_allclose(s.alphaind[even, 0], decay[even], rtol=2e-7)
    np.testing.assert_array_equal(s.alphaind[~even, 0], np.float32(1.5))
    # v is predicted half a step ahead with the old force
    np.testing.assert_array_equal(s.vpred, s.vxyzu + 0.5 * dt * s.fxyzu)


def test_wake_moves_to_the_bin_half_step():
    nbinmax, dtmax = 3, 0.25  # binary fractions: every time below is exact
    p = _params(maxvxyzu=3, ind=True, dtmax=dtmax)
    s = _bins_state(p, nbinmax, seed=2)
    s.ibin[:] = 0
    s.ibin[:6] = 1
    o = stepping.OracleBackend(p)
    fn = stepping.constant_force([0.0, 0.0, 0.0])
    t0 = 0.5
    integrator.init_ind(o, s, time=t0, derivs_fn=fn)
    # wake requests at the synchronisation point t0 + dtmax/2 (istepfrac 4 of 8)
    time, tlast, _ = integrator.evolve_ind(o, s, 3, time=t0, tlast=t0, derivs_fn=fn)
    s.ibin_wake[:] = 0
    s.ibin_wake[10:14] = [2, 3, 5, 1]  # particles 10-13 are on bin 0; 5 is above nbinmax
    res = {}

    def fn_record(ctx, icall, st, timei, dt):
        res["timei"] = timei
        return fn(ctx, icall, st, timei, dt)

    time, _, out = integrator.evolve_ind(o, s, 1, time=time, tlast=tlast, derivs_fn=fn_record)
    timei = res["timei"]
    assert timei == t0 + 0.5 * dtmax
    assert out[0]["nwake"] == 4
    np.testing.assert_array_equal(s.ibin[10:14], [2, 3, 3, 1])
    assert np.all(s.ibin_wake == 0)
    # timei is a multiple of every bin's step: twas is half a step ahead
    for k, b in zip(range(10, 14), [2, 3, 3, 1]):
        assert s.twas[k] == timei + 0.5 * dtmax / 2 ** b


def _drag_force(k):
    """derivs replacement: f = -k v at the predicted velocity, so the
    corrected velocity differs from the prediction and the corrector has to
    iterate when k dt is large."""

    def fn(ctx, icall, s, timei, dt):
        act = np.ones(s.n, bool) if s.iphase is None else (s.iphase > 0)
        s.fxyzu[act, :3] = -k * s.vpred[act, :3]
        if s.fxyzu.shape[1] >= 4:
            s.fxyzu[act, 3] = 0.0
        s.dtcourant = s.dtforce = 1.e29
        return s.params.dtmax

    return fn


def test_dterr_after_iterated_step():
    """check_velocity_error sets timestep:dterr on every call, and finite
    only on the first iteration (step_leapfrog.f90:807-838): an iterated
    step leaves it huge, so the next dt is dtmax, not the first iteration's
    dterr.  A step that converges at once keeps the finite dterr."""
    p = _params(maxvxyzu=4, dtmax=0.1)
    dt = 0.1
    huge = np.finfo(np.float64).max
    # strong drag: the first corrector iteration misses tolv
    s = _free_state(p)
    o = stepping.OracleBackend(p)
    fn = _drag_force(5.0)
    integrator.init_global(o, s, derivs_fn=fn)
    _, dtnext, out = integrator.evolve_global(o, s, 1, dt=dt, derivs_fn=fn)
    assert out[0]["its"] >= 2
    assert out[0]["dterr"] == huge
    assert dtnext == p.dtmax + np.finfo(np.float64).eps
    # weak drag: converged at the first iteration with a finite dterr, which
    # sets dt = dt sqrt(tolv/errmax)
    s = _free_state(p)
    fn = _drag_force(0.05)
    integrator.init_global(o, s, derivs_fn=fn)
    _, dtnext, out = integrator.evolve_global(o, s, 1, dt=dt, derivs_fn=fn)
    assert out[0]["its"] == 1
    assert out[0]["dterr"] < huge
    assert dtnext == min(out[0]["dterr"], p.dtmax + np.finfo(np.float64).eps)
