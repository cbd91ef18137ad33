"""The dense torch reference of the adjoint tests (tests/adjoint_ref.py)
pinned on the CPU: its autograd gradients against central differences of
itself, its density against the h-rho fixed point, and its pressure
acceleration against momentum conservation and against the oracle's force
(dissipation off, float32 1/Omega aside)."""
import numpy as np
import torch

from oracle import oracle
from phantom_amd import abi
from tests import adjoint_ref as R


def _cloud(n=120, seed=3):
    rng = np.random.default_rng(seed)
    return torch.tensor(rng.normal(0.0, 0.3, (n, 3))), n


def test_density_fixed_point_and_gradient():
    x0, n = _cloud()
    m, hfact, K = 1.0 / n, 1.2, R.Kernel(True)
    h0 = torch.full((n,), 0.2, dtype=torch.float64)
    x = x0.clone().requires_grad_(True)
    rho, h = R.density(x, h0, m, hfact, K)
    # the solved h satisfies rho_sum(h) = m (hfact/h)^3
    rs = R.rho_sum(x, h, m, K)
    assert float(((rs - rho) / rho).abs().max().detach()) < 1e-13
    w = torch.tensor(np.random.default_rng(1).normal(size=n))
    g = torch.autograd.grad((w * rho).sum(), x)[0]
    eps = 1e-6
    for k in range(8):
        for c in range(3):
            xp, xm = x0.clone(), x0.clone()
            xp[k, c] += eps
            xm[k, c] -= eps
            fd = float((w * (R.density(xp, h0, m, hfact, K)[0] - R.density(xm, h0, m, hfact, K)[0])).sum()) / (2 * eps)
            assert abs(fd - float(g[k, c])) <= 1e-8 * float(g.abs().max())


def test_pressure_accel_momentum_and_oracle():
    x0, n = _cloud(n=200, seed=4)
    p = abi.default_params(abi.KERNEL_CUBIC)
    p.periodic = 0
    p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax = -4.0, 4.0, -4.0, 4.0, -4.0, 4.0
    p.massoftype_gas = 1.0 / n
    p.maxp = n
    p.tolh = 1e-13
    p.alpha = p.alphau = p.beta = 0.0
    p.ieos, p.polyk, p.maxvxyzu = 1, 1.0, 3
    K = R.Kernel(True)
    x = x0.clone().requires_grad_(True)
    _, h = R.density(x, torch.full((n,), 0.2, dtype=torch.float64), p.massoftype_gas, p.hfact, K)
    a = R.pressure_accel(x, h, p.massoftype_gas, p.hfact, K, 1.0).detach().numpy()
    assert np.abs(a.sum(0)).max() <= 1e-12 * np.abs(a).sum(0).max()
    o = oracle.Oracle(p)
    xyzh = np.empty((n, 4))
    xyzh[:, :3] = x0.numpy()
    xyzh[:, 3] = 0.2
    v = np.zeros((n, 3))
    o.maketree(xyzh)
    d = o.densityiterate(xyzh, v)
    np.testing.assert_allclose(xyzh[:, 3], h.detach().numpy(), rtol=1e-11)
    c = o.cons2prim(xyzh, v, d["dvdx"], d["alphaind"])
    f = o.force(xyzh, v, d["gradh"], d["alphaind"], c["eos_vars"])["fxyzu"][:, :3]
    assert np.abs(f - a).max() <= 1e-6 * np.abs(a).max()
