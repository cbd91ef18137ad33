"""The adjoint kernels (phsph_density_vjp, phsph_force_vjp, phsph_step_vjp)
and the torch harness over them, on the GPU.

Phantom has no adjoint, so the gradients are checked against (a) central
differences (eps = 1e-6, FP64) of the library's own forward on 64 seeded
particles, and (b) an independent FP64 reference gradient: the dense torch
restatement in tests/adjoint_ref.py, differentiated by torch.autograd
through an unrolled Newton solve of the h-rho constraint.  The forward of
the pressure acceleration is also checked against phsph_force with the
dissipation switched off (its 1/Omega is stored in float32, so they agree
to float rounding).  At the differentiable-relaxation size (the n = 3
polytrope of 1,048,576 particles with sigma = 0.02 h noise) the density
adjoint is checked against central differences on 64 particles.
"""
import numpy as np
import pytest

from phantom_amd import abi, adjoint, setups, sph
from tests import parity

pytestmark = pytest.mark.gpu

EPS = 1e-6
TOL_FD = 1e-6     # relative to the largest gradient component
TOL_REF = 1e-8    # against the dense torch reference


def _cloud(n=300, seed=5, kernel=abi.KERNEL_CUBIC, ieos=1):
    rng = np.random.default_rng(seed)
    p = abi.default_params(kernel)
    p.periodic = 0
    p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax = -4.0, 4.0, -4.0, 4.0, -4.0, 4.0
    p.massoftype_gas = 1.0 / n
    p.maxp = n
    p.tolh = 1e-13
    p.alpha = p.alphau = p.beta = 0.0
    p.ieos = ieos
    p.polyk = 1.0
    p.maxvxyzu = 3 if ieos == 1 else 4
    x = rng.normal(0.0, 0.3, size=(n, 3))
    xyzh = np.empty((n, 4))
    xyzh[:, :3] = x
    xyzh[:, 3] = 0.2
    v = np.zeros((n, p.maxvxyzu))
    v[:, :3] = rng.normal(0.0, 0.1, size=(n, 3))
    if ieos == 2:
        v[:, 3] = rng.uniform(0.5, 1.5, n)
    return p, xyzh, v


def _solve(ctx, xyzh, v):
    a = xyzh.copy()
    ctx.maketree(a, leaf=False)
    ctx.densityiterate(1, a, v)
    return a


def _rho(p, xyzh):
    return p.massoftype_gas * (p.hfact / xyzh[:, 3]) ** 3


def _torch_ref():
    import torch

    from tests import adjoint_ref as R
    return torch, R


def _fd_particles(n, k=64, seed=9):
    return np.random.default_rng(seed).choice(n, size=min(k, n), replace=False)


def _max_rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def cloud(gpu):
    p, xyzh, v = _cloud()
    ctx = sph.Context(p)
    yield p, xyzh, v, ctx
    ctx.close()


def test_density_vjp_vs_reference_and_fd(cloud):
    torch, R = _torch_ref()
    p, xyzh, v, ctx = cloud
    n = xyzh.shape[0]
    w = np.random.default_rng(1).normal(size=n)
    a = _solve(ctx, xyzh, v)
    g = ctx.density_vjp(a, rhobar=w)
    # the dense reference through autograd
    # (its unclamped Newton iteration starts from the library's converged h)
    x = torch.tensor(xyzh[:, :3], requires_grad=True)
    rho, _ = R.density(x, torch.tensor(a[:, 3]), p.massoftype_gas, p.hfact, R.Kernel(True))
    gr = torch.autograd.grad((torch.tensor(w) * rho).sum(), x)[0].numpy()
    np.testing.assert_allclose(_rho(p, a), rho.detach().numpy(), rtol=1e-12)
    err_ref = _max_rel(g, gr)
    # central differences of the library's forward
    fd = np.zeros((n, 3))
    ks = _fd_particles(n)
    for k in ks:
        for c in range(3):
            xp = xyzh.copy()
            xp[k, c] += EPS
            xm = xyzh.copy()
            xm[k, c] -= EPS
            fd[k, c] = (w * (_rho(p, _solve(ctx, xp, v)) - _rho(p, _solve(ctx, xm, v)))).sum() / (2 * EPS)
    err_fd = float(np.abs(g[ks] - fd[ks]).max() / np.abs(g).max())
    print(f"\n[density_vjp] vs torch reference {err_ref:.2e}, vs central differences {err_fd:.2e}", flush=True)
    assert err_ref <= TOL_REF
    assert err_fd <= TOL_FD


def test_pressure_accel_forward(cloud):
    """phsph_force_vjp's forward against the dense reference (FP64 Omega)
    and against phsph_force with alpha = beta = alphau = 0 (float32
    1/Omega), and momentum conservation."""
    torch, R = _torch_ref()
    p, xyzh, v, ctx = cloud
    a = _solve(ctx, xyzh, v)
    acc, _ = ctx.force_vjp(a, v)
    x = torch.tensor(xyzh[:, :3], requires_grad=True)
    _, h = R.density(x, torch.tensor(a[:, 3]), p.massoftype_gas, p.hfact, R.Kernel(True))
    ar = R.pressure_accel(x, h, p.massoftype_gas, p.hfact, R.Kernel(True), p.polyk).detach().numpy()
    assert _max_rel(acc, ar) <= 1e-10
    assert np.abs(acc.sum(0)).max() <= 1e-12 * np.abs(acc).sum(0).max()
    d = ctx.densityiterate(1, a, v)
    c = ctx.cons2prim(a, v, d["dvdx"], d["alphaind"])
    f = ctx.force(1, a, v, d["gradh"], d["alphaind"], c["eos_vars"])
    for k in range(3):
        assert parity.rel_rms(acc[:, k], f["fxyzu"][:, k]) <= 1e-6


def test_force_vjp_vs_reference_and_fd(cloud):
    torch, R = _torch_ref()
    p, xyzh, v, ctx = cloud
    n = xyzh.shape[0]
    wb = np.random.default_rng(2).normal(size=(n, 3))
    a = _solve(ctx, xyzh, v)
    _, g = ctx.force_vjp(a, v, fbar=wb)
    x = torch.tensor(xyzh[:, :3], requires_grad=True)
    K = R.Kernel(True)
    _, h = R.density(x, torch.tensor(a[:, 3]), p.massoftype_gas, p.hfact, K)
    ar = R.pressure_accel(x, h, p.massoftype_gas, p.hfact, K, p.polyk)
    gr = torch.autograd.grad((torch.tensor(wb) * ar).sum(), x)[0].numpy()
    err_ref = _max_rel(g, gr)
    fd = np.zeros((n, 3))
    ks = _fd_particles(n)
    for k in ks:
        for c in range(3):
            acc = []
            for s in (1.0, -1.0):
                xx = xyzh.copy()
                xx[k, c] += s * EPS
                aa, _ = ctx.force_vjp(_solve(ctx, xx, v), v)
                acc.append(aa)
            fd[k, c] = (wb * (acc[0] - acc[1])).sum() / (2 * EPS)
    err_fd = float(np.abs(g[ks] - fd[ks]).max() / np.abs(g).max())
    print(f"\n[force_vjp] vs torch reference {err_ref:.2e}, vs central differences {err_fd:.2e}", flush=True)
    assert err_ref <= TOL_REF
    assert err_fd <= TOL_FD


@pytest.mark.parametrize("ieos", [1, 2])
def test_step_vjp_vs_reference_and_fd(ieos, gpu):
    torch, R = _torch_ref()
    p, xyzh, v = _cloud(n=250, seed=8, ieos=ieos)
    ctx = sph.Context(p)
    n = xyzh.shape[0]
    dt = 0.02
    rng = np.random.default_rng(3)
    xb, vb = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    xn, vn, gx, gv = ctx.step_vjp(dt, xyzh, v, xbar=xb, vbar=vb)
    # forward and gradients of the dense reference
    x = torch.tensor(xyzh[:, :3], requires_grad=True)
    vt = torch.tensor(v[:, :3], requires_grad=True)
    K = R.Kernel(True)
    cfac = p.polyk if ieos == 1 else torch.tensor((p.gamma - 1.0) * v[:, 3])
    hc = _solve(ctx, xyzh, v)[:, 3]  # the reference's Newton starts from converged h
    x1, v1 = R.step(x, vt, torch.tensor(hc), dt, p.massoftype_gas, p.hfact, K, cfac)
    assert _max_rel(xn[:, :3], x1.detach().numpy()) <= 1e-12
    assert _max_rel(vn[:, :3], v1.detach().numpy()) <= 1e-10
    L = (torch.tensor(xb) * x1).sum() + (torch.tensor(vb) * v1).sum()
    grx, grv = torch.autograd.grad(L, (x, vt))
    err_ref = max(_max_rel(gx, grx.numpy()), _max_rel(gv, grv.numpy()))
    # central differences of the step's own forward
    ks = _fd_particles(n)
    fdx = np.zeros((n, 3))
    for k in ks:
        for c in range(3):
            out = []
            for s in (1.0, -1.0):
                xx = xyzh.copy()
                xx[k, c] += s * EPS
                a, b, _, _ = ctx.step_vjp(dt, xx, v)
                out.append((a, b))
            fdx[k, c] = ((xb * (out[0][0][:, :3] - out[1][0][:, :3])).sum() +
                         (vb * (out[0][1][:, :3] - out[1][1][:, :3])).sum()) / (2 * EPS)
    err_fd = float(np.abs(gx[ks] - fdx[ks]).max() / np.abs(gx).max())
    print(f"\n[step_vjp ieos {ieos}] vs torch reference {err_ref:.2e}, vs central differences {err_fd:.2e}",
          flush=True)
    ctx.close()
    assert err_ref <= TOL_REF
    assert err_fd <= TOL_FD


def test_torch_autograd_harness(cloud):
    """adjoint.density and adjoint.leapfrog_step inside a torch graph: the
    gradient of a loss built on them equals the dense reference's."""
    torch, R = _torch_ref()
    p, xyzh, v, ctx = cloud
    K = R.Kernel(True)
    h0 = torch.tensor(_solve(ctx, xyzh, v)[:, 3])
    target = adjoint.profile_target(np.linspace(0.0, 2.0, 200), 2.0 * np.exp(-np.linspace(0.0, 2.0, 200) ** 2))
    x = torch.tensor(xyzh[:, :3], requires_grad=True)
    loss = adjoint.density_loss(x, h0, ctx, target)
    loss.backward()
    xr = torch.tensor(xyzh[:, :3], requires_grad=True)
    rho, _ = R.density(xr, h0, p.massoftype_gas, p.hfact, K)
    rt = target(xr)
    lr = p.massoftype_gas * (((rho - rt) / rt) ** 2).sum()
    lr.backward()
    assert abs(float(loss) - float(lr)) <= 1e-12 * abs(float(lr))
    assert _max_rel(x.grad.numpy(), xr.grad.numpy()) <= TOL_REF
    # a step inside the graph
    xs = torch.tensor(xyzh[:, :3], requires_grad=True)
    vs = torch.tensor(v[:, :3], requires_grad=True)
    x1, v1 = adjoint.leapfrog_step(xs, vs, h0, 0.02, ctx)
    (x1 * x1).sum().add((v1 * v1).sum()).backward()
    xr = torch.tensor(xyzh[:, :3], requires_grad=True)
    vr = torch.tensor(v[:, :3], requires_grad=True)
    x2, v2 = R.step(xr, vr, h0, 0.02, p.massoftype_gas, p.hfact, K, p.polyk)
    (x2 * x2).sum().add((v2 * v2).sum()).backward()
    assert _max_rel(xs.grad.numpy(), xr.grad.numpy()) <= TOL_REF
    assert _max_rel(vs.grad.numpy(), vr.grad.numpy()) <= TOL_REF


def test_gradient_relaxation_lowers_density_error(gpu):
    """relax_to_density on a noisy n = 3 polytrope (the relaxation workload at
    20k particles): the loss falls monotonically and by a clear factor."""
    s = setups.relaxation_target(n=20000, sigma=0.3)
    p = s["params"]
    p.tolh = 1e-10
    ctx = sph.Context(p)
    target = adjoint.profile_target(*s["target"])
    xyzh, hist = adjoint.relax_to_density(ctx, s["xyzh"], target, niter=8)
    ctx.close()
    print(f"\n[relax_to_density] loss {hist[0]:.4e} -> {hist[-1]:.4e}", flush=True)
    assert all(b < a for a, b in zip(hist, hist[1:]))
    assert hist[-1] < 0.8 * hist[0]


def test_density_vjp_relaxation_size_fd(gpu):
    """The differentiable-relaxation workload at its size (n = 3 polytrope,
    1,048,576 requested particles, Gaussian noise 0.02 h): the density
    adjoint of L = sum w rho against central differences (eps = 1e-6) on 64
    seeded particles; each difference is summed particle by particle so
    that the unaffected densities cancel exactly."""
    s = setups.relaxation_target(n=1048576)
    p = s["params"]
    p.tolh = 1e-12
    n = s["xyzh"].shape[0]
    print(f"\n[relaxation size] n={n}", flush=True)
    ctx = sph.Context(p)
    v = s["vxyzu"]
    a = _solve(ctx, s["xyzh"], v)
    base = a.copy()  # converged h: the start of every perturbed solve
    w = np.random.default_rng(4).normal(size=n)
    g = ctx.density_vjp(a, rhobar=w)
    ks = _fd_particles(n)
    err = 0.0
    gscale = np.abs(g[ks]).max()
    for k in ks:
        for c in range(3):
            rr = []
            for sgn in (1.0, -1.0):
                xx = base.copy()
                xx[k, c] += sgn * EPS
                rr.append(_rho(p, _solve(ctx, xx, v)))
            fd = float((w * (rr[0] - rr[1])).sum() / (2 * EPS))
            err = max(err, abs(fd - g[k, c]) / gscale)
    ctx.close()
    print(f"[relaxation size] density_vjp vs central differences {err:.2e}", flush=True)
    assert err <= TOL_FD
