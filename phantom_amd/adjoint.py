"""Differentiable density and leapfrog step over the C-ABI, for PyTorch.

Phantom has no adjoint; this module exposes the library's hand-written
adjoint kernels (phsph_density_vjp, phsph_step_vjp; the forward sums are
get_density_sums, dens.F90:578-857, and the pressure force of
compute_forces, force.F90:1595-1596) as torch.autograd.Function objects:

    rho = density(x, h0, ctx)                 # h-rho solved exactly
    x1, v1 = leapfrog_step(x, v, h0, dt, ctx) # density -> force -> kick

Tensors are float64 on the host (the library stages host arrays through its
own device buffers; torch's bundled HIP runtime is a different one from the
library's, so device tensors are not exchanged).  Nothing is stored between
forward and backward except the inputs: the backward pass rebuilds the tree,
the converged h and the neighbour lists at the saved positions.

relax_to_density() drives gradient-based relaxation of particle positions
towards a target density profile rho_t(r) with the loss
sum_i m (rho_i - rho_t(r_i))^2 / rho_t(r_i)^2, where rho_t(r) is a torch
function so that its dependence on the positions is differentiated too.
"""
import numpy as np


def _torch():
    import torch
    return torch


def _xyzh(x, h):
    a = np.empty((x.shape[0], 4), dtype=np.float64)
    a[:, :3] = x
    a[:, 3] = h
    return a


def solve_density(ctx, x, h0, vxyzu=None):
    """maketree + densityiterate(icall 1) at positions x (n,3) from initial h0:
    returns the (n,4) xyzh with the converged h (wrapped for periodic boxes)."""
    xyzh = _xyzh(np.asarray(x, dtype=np.float64), np.asarray(h0, dtype=np.float64))
    n = xyzh.shape[0]
    if vxyzu is None:
        vxyzu = np.zeros((n, ctx.params.maxvxyzu))
    ctx.maketree(xyzh, leaf=False)
    ctx.densityiterate(1, xyzh, vxyzu)
    return xyzh


def rho_of_h(params, h):
    """rhoh (part.F90:779-790): m (hfact/|h|)^3."""
    return params.massoftype_gas * (params.hfact / np.abs(h)) ** 3


class _Density:
    """torch.autograd.Function (built lazily so that importing this module
    does not import torch)."""
    fn = None

    @classmethod
    def get(cls):
        if cls.fn is None:
            torch = _torch()

            class DensityFn(torch.autograd.Function):
                @staticmethod
                def forward(fctx, x, h0, ctx):
                    xyzh = solve_density(ctx, x.detach().numpy(), h0.detach().numpy())
                    fctx.ctx = ctx
                    fctx.save_for_backward(x, h0)
                    return torch.from_numpy(rho_of_h(ctx.params, xyzh[:, 3]))

                @staticmethod
                def backward(fctx, grad_rho):
                    x, h0 = fctx.saved_tensors
                    ctx = fctx.ctx
                    xyzh = solve_density(ctx, x.detach().numpy(), h0.detach().numpy())
                    g = np.ascontiguousarray(grad_rho.detach().numpy(), dtype=np.float64)
                    dx = ctx.density_vjp(xyzh, rhobar=g)
                    return torch.from_numpy(dx), None, None

            cls.fn = DensityFn
        return cls.fn


class _Step:
    fn = None

    @classmethod
    def get(cls):
        if cls.fn is None:
            torch = _torch()

            class StepFn(torch.autograd.Function):
                @staticmethod
                def forward(fctx, x, v, h0, u, dt, ctx):
                    xyzh = _xyzh(x.detach().numpy(), h0.detach().numpy())
                    vxyzu = _vxyzu(ctx, v.detach().numpy(), u)
                    xn, vn, _, _ = ctx.step_vjp(dt, xyzh, vxyzu)
                    fctx.ctx = ctx
                    fctx.dt = dt
                    fctx.u = u
                    fctx.save_for_backward(x, v, h0)
                    return torch.from_numpy(xn[:, :3].copy()), torch.from_numpy(vn[:, :3].copy())

                @staticmethod
                def backward(fctx, gx, gv):
                    x, v, h0 = fctx.saved_tensors
                    ctx = fctx.ctx
                    xyzh = _xyzh(x.detach().numpy(), h0.detach().numpy())
                    vxyzu = _vxyzu(ctx, v.detach().numpy(), fctx.u)
                    n = xyzh.shape[0]
                    xb = np.zeros((n, 3)) if gx is None else np.ascontiguousarray(gx.detach().numpy())
                    vb = np.zeros((n, 3)) if gv is None else np.ascontiguousarray(gv.detach().numpy())
                    _, _, dx, dv = ctx.step_vjp(fctx.dt, xyzh, vxyzu, xbar=xb, vbar=vb, want_state=False)
                    return torch.from_numpy(dx), torch.from_numpy(dv), None, None, None, None

            cls.fn = StepFn
        return cls.fn


def _vxyzu(ctx, v, u):
    n = v.shape[0]
    out = np.zeros((n, ctx.params.maxvxyzu))
    out[:, :3] = v
    if ctx.params.maxvxyzu >= 4 and u is not None:
        out[:, 3] = u
    return out


def density(x, h0, ctx):
    """rho_i (torch, float64) at positions x (n,3) with the h-rho constraint
    solved from the initial guess h0; differentiable in x."""
    return _Density.get().apply(x, h0, ctx)


def leapfrog_step(x, v, h0, dt, ctx, u=None):
    """One kick-drift-kick step with the pressure acceleration (barotropic:
    P = polyk rho, or (gamma-1) u rho with u fixed); returns (x', v'),
    differentiable in x and v."""
    return _Step.get().apply(x, v, h0, None if u is None else np.asarray(u, np.float64), float(dt), ctx)


def density_loss(x, h0, ctx, rho_target):
    """sum_i m (rho_i - rho_t(r_i))^2 / rho_t(r_i)^2 with rho_t a torch
    function of the positions."""
    rho = density(x, h0, ctx)
    rt = rho_target(x)
    return ctx.params.massoftype_gas * (((rho - rt) / rt) ** 2).sum()


def relax_to_density(ctx, xyzh, rho_target, niter=50, step=None, maxshift=0.1, log=None):
    """Gradient relaxation of positions towards rho_target (a torch function
    of the (n,3) positions): steepest descent on density_loss with steps
    scaled so that no particle moves by more than maxshift * h per
    iteration, halved when the loss does not decrease.  Returns
    (xyzh, history of the loss)."""
    torch = _torch()
    x = torch.tensor(np.ascontiguousarray(xyzh[:, :3]), dtype=torch.float64)
    h = np.ascontiguousarray(xyzh[:, 3]).copy()
    hist = []
    scale = step
    for it in range(niter):
        xv = x.clone().requires_grad_(True)
        loss = density_loss(xv, torch.from_numpy(h), ctx, rho_target)
        loss.backward()
        g = xv.grad.detach()
        hist.append(float(loss))
        if log:
            log(it, float(loss))
        # converged h of this iterate: the start of the next solve
        h = solve_density(ctx, x.numpy(), h)[:, 3].copy()
        gn = torch.sqrt((g * g).sum(1))
        if scale is None:
            scale = float(maxshift * np.min(h) / max(float(gn.max()), 1e-300))
        while True:
            xt = x - scale * g
            lt = float(density_loss(xt, torch.from_numpy(h), ctx, rho_target))
            if lt < hist[-1] or scale < 1e-12:
                break
            scale *= 0.5
        x = xt.detach()
        scale *= 1.5
    out = solve_density(ctx, x.numpy(), h)
    return out, hist


def profile_target(r_tab, rho_tab, center=(0.0, 0.0, 0.0)):
    """rho_t(|x - center|) by linear interpolation of a table (clamped at
    the ends, as table_utils:yinterp), as a torch function of positions,
    differentiable in x."""
    torch = _torch()
    rt = torch.as_tensor(np.asarray(r_tab, np.float64))
    yt = torch.as_tensor(np.asarray(rho_tab, np.float64))
    c = torch.as_tensor(np.asarray(center, np.float64))

    def fn(x):
        r = torch.sqrt(((x - c) ** 2).sum(1) + 1e-300)
        rc = torch.clamp(r, float(rt[0]), float(rt[-1]))
        k = torch.clamp(torch.searchsorted(rt, rc.detach(), right=True) - 1, 0, rt.shape[0] - 2)
        t = (rc - rt[k]) / (rt[k + 1] - rt[k])
        return yt[k] + t * (yt[k + 1] - yt[k])

    return fn
