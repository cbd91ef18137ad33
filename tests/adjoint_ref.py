"""Dense FP64 torch restatement of the differentiated forward sums, for
checking the adjoint kernels (test infrastructure).

Every pair (O(N^2), open boundaries) is summed; h is solved by Newton's
method on rho_sum(h) = m (hfact/h)^3 unrolled for a fixed number of
iterations, the Newton slope itself taken by autograd, and torch.autograd
differentiates through the iterations: at a converged fixed point the
derivative of the unrolled solve equals the implicit derivative, so the
implicit-function formula the HIP adjoint uses is not restated here.  The
forward follows get_density_sums (dens.F90:578-857, self term included) and
the pressure part of compute_forces (force.F90:1595-1596, with
Omega = 1 - dh/drho d(rho_sum)/dh, dens.F90:1382-1466).
"""
import torch

PI = 3.1415926536  # physcon.f90:23


def _w_cubic(q):
    """kernel_cubic.f90:34-60, w(q)."""
    w1 = 0.75 * q ** 3 - 1.5 * q ** 2 + 1.0
    w2 = -0.25 * (q - 2.0) ** 3
    return torch.where(q < 1.0, w1, torch.where(q < 2.0, w2, torch.zeros_like(q)))


def _w_quintic(q):
    """kernel_quintic.f90:34-75, w(q)."""
    w1 = -10.0 * q ** 5 + 30.0 * q ** 4 - 60.0 * q ** 2 + 66.0
    w2 = -(q - 3.0) ** 5 + 6.0 * (q - 2.0) ** 5
    w3 = -(q - 3.0) ** 5
    return torch.where(q < 1.0, w1, torch.where(q < 2.0, w2, torch.where(q < 3.0, w3, torch.zeros_like(q))))


class Kernel:
    def __init__(self, cubic):
        self.w = _w_cubic if cubic else _w_quintic
        self.cnormk = 1.0 / PI if cubic else 1.0 / (120.0 * PI)


def _sep(x):
    d = x[:, None, :] - x[None, :, :]
    n = x.shape[0]
    eye = torch.eye(n, dtype=torch.bool)
    r2 = (d * d).sum(-1)
    r = torch.sqrt(torch.where(eye, torch.ones_like(r2), r2))
    return d, torch.where(eye, torch.zeros_like(r), r), eye


def rho_sum(x, h, m, K):
    """sum_j m W(r_ij, h_i), self term included."""
    _, r, _ = _sep(x)
    return m * (K.cnormk / h ** 3 * K.w(r / h[:, None]).sum(1))


def _live(t):
    return t if t.requires_grad else t.detach().requires_grad_(True)


def solve_h(x, h0, m, hfact, K, iters=30):
    """Newton on F(h) = rho_sum(h) - m (hfact/h)^3 = 0, unrolled; F_i depends
    on h_i only, so the slope is the gradient of sum F."""
    h = _live(h0 + 0.0 * x[:, 0])
    for _ in range(iters):
        f = rho_sum(x, h, m, K) - m * (hfact / h) ** 3
        dfdh = torch.autograd.grad(f.sum(), h, create_graph=True)[0]
        h = h - f / dfdh
    return h


def density(x, h0, m, hfact, K, iters=30):
    """(rho, h) at the solved h."""
    h = solve_h(x, h0, m, hfact, K, iters)
    return m * (hfact / h) ** 3, h


def _wprime(q, K):
    """w'(q), differentiable in q: autograd of the kernel polynomial."""
    q = _live(q)
    return torch.autograd.grad(K.w(q).sum(), q, create_graph=True)[0]


def pressure_accel(x, h, m, hfact, K, cfac):
    """a_i = -sum_j m [A_i dW(r,h_i)/dr + A_j dW(r,h_j)/dr] r_hat with
    A = c/(Omega rho) and Omega = 1 + h/(3 rho) d(rho_sum)/dh; dW/dr and
    d(rho_sum)/dh are taken by autograd from W itself."""
    d, r, eye = _sep(x)
    h = _live(h)
    rho = m * (hfact / h) ** 3
    dsdh = torch.autograd.grad(rho_sum(x, h, m, K).sum(), h, create_graph=True)[0]
    omega = 1.0 + h / (3.0 * rho) * dsdh
    A = cfac / (omega * rho)
    gi = K.cnormk / h[:, None] ** 4 * _wprime(r / h[:, None], K)
    gj = K.cnormk / h[None, :] ** 4 * _wprime(r / h[None, :], K)
    phi = -m * (A[:, None] * gi + A[None, :] * gj)
    rinv = torch.where(eye, torch.zeros_like(r), 1.0 / torch.where(eye, torch.ones_like(r), r))
    return ((phi * rinv)[:, :, None] * d).sum(1)


def step(x, v, h0, dt, m, hfact, K, cfac, iters=30):
    """v_half = v + dt/2 a(x); x' = x + dt v_half; v' = v_half + dt/2 a(x')."""
    _, h = density(x, h0, m, hfact, K, iters)
    vh = v + 0.5 * dt * pressure_accel(x, h, m, hfact, K, cfac)
    x1 = x + dt * vh
    _, h1 = density(x1, h.detach(), m, hfact, K, iters)
    v1 = vh + 0.5 * dt * pressure_accel(x1, h1, m, hfact, K, cfac)
    return x1, v1
