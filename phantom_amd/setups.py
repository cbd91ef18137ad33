"""Synthetic particle sets for the hot-path workloads.

Every generator returns a dict of numpy arrays in the reference's storage
layout: an array a(ncomp, n) of Fortran is a C-contiguous numpy array of
shape (n, ncomp).  Keys: ``xyzh`` (n,4) f8, ``vxyzu`` (n,maxvxyzu) f8,
``alphaind`` (n,3) f4, optional ``Bevol`` (n,4) f8, ``dustevol`` (n,) f8,
plus ``params`` (a phantom_amd.abi.Params) and ``info`` (a dict).

Lattices follow set_unifdis('cubic') (src/setup/set_unifdis.f90:182-222):
particle (i,j,k) sits at xmin + (i-1/2) dx.  Smoothing lengths start at
hfact * dx, as set_unifdis sets them.
"""
import math

import numpy as np

from . import abi

PI_PH = 3.1415926536  # physcon.f90:23


def _cubic_lattice(nx, ny, nz, x0, x1, y0, y1, z0, z1):
    dx = (x1 - x0) / nx
    dy = (y1 - y0) / ny
    dz = (z1 - z0) / nz
    xs = x0 + (np.arange(nx) + 0.5) * dx
    ys = y0 + (np.arange(ny) + 0.5) * dy
    zs = z0 + (np.arange(nz) + 0.5) * dz
    z, y, x = np.meshgrid(zs, ys, xs, indexing="ij")  # x fastest, as set_unifdis
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1), dx


def _wrap(p, lo, hi):
    L = hi - lo
    return np.where(p < lo, p + L, np.where(p > hi, p - L, p))


def uniform_box(nx=64, ny=None, nz=None, jitter=0.05, seed=1234, kernel=abi.KERNEL_CUBIC,
                isothermal=True, box=(-0.5, 0.5, -0.5, 0.5, -0.5, 0.5), rho0=1.0, vamp=0.1,
                u0=1.5, alpha_seed=None):
    """Jittered cubic lattice in a periodic box with v_x = vamp sin(2 pi y).

    With the defaults this is the 64^3 isothermal box: rho = 1, c_s = 1,
    every coordinate jittered by U(-0.05 dx, 0.05 dx).
    """
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    x0, x1, y0, y1, z0, z1 = box
    pos, dx = _cubic_lattice(nx, ny, nz, x0, x1, y0, y1, z0, z1)
    rng = np.random.default_rng(seed)
    n = pos.shape[0]
    if jitter:
        pos = pos + rng.uniform(-jitter * dx, jitter * dx, size=pos.shape)
        pos[:, 0] = _wrap(pos[:, 0], x0, x1)
        pos[:, 1] = _wrap(pos[:, 1], y0, y1)
        pos[:, 2] = _wrap(pos[:, 2], z0, z1)
    vol = (x1 - x0) * (y1 - y0) * (z1 - z0)
    p = abi.default_params(kernel)
    p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax = box
    p.massoftype_gas = rho0 * vol / n
    p.maxp = n
    if isothermal:
        p.maxvxyzu = 3
        p.ieos = 1
        p.polyk = 1.0
    xyzh = np.empty((n, 4))
    xyzh[:, :3] = pos
    xyzh[:, 3] = p.hfact * dx
    vxyzu = np.zeros((n, p.maxvxyzu))
    vxyzu[:, 0] = vamp * np.sin(2.0 * PI_PH * (pos[:, 1] - y0) / (y1 - y0))
    if not isothermal:
        vxyzu[:, 3] = u0
    alphaind = np.zeros((n, 3), dtype=np.float32)
    arng = np.random.default_rng(seed + 1 if alpha_seed is None else alpha_seed)
    alphaind[:, 0] = arng.uniform(0.0, 1.0, size=n).astype(np.float32)
    return dict(xyzh=xyzh, vxyzu=vxyzu, alphaind=alphaind, params=p,
                info=dict(workload="uniform_box", n=n, nx=nx, ny=ny, nz=nz, dx=dx, jitter=jitter,
                          seed=seed, kernel=kernel, isothermal=isothermal))


def shock_tube(ny=58, seed=4321, kernel=abi.KERNEL_QUINTIC):
    """3D shock tube in [-0.5,0.5]x[-0.05,0.05]^2, periodic, equal masses.

    Left half on a lattice of spacing dx (rho_L = 1, P_L = 1), right half on
    spacing 2 dx (rho_R = 1/8, P_R = 0.1), gamma = 5/3, v = 0.  With ny = 58
    lattice rows across the tube the particle count is 290*58*58 +
    145*29*29 = 1,097,505, the lattice count nearest 2^20 that a 2:1 spacing
    ratio allows.
    """
    gamma = 5.0 / 3.0
    dx = 0.1 / ny
    nxl = int(round(0.5 / dx))
    left, _ = _cubic_lattice(nxl, ny, ny, -0.5, 0.0, -0.05, 0.05, -0.05, 0.05)
    right, _ = _cubic_lattice(nxl // 2, ny // 2, ny // 2, 0.0, 0.5, -0.05, 0.05, -0.05, 0.05)
    pos = np.concatenate([left, right])
    n = pos.shape[0]
    mtot = 1.0 * 0.5 * 0.01 + 0.125 * 0.5 * 0.01
    p = abi.default_params(kernel)
    p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax = -0.5, 0.5, -0.05, 0.05, -0.05, 0.05
    p.massoftype_gas = mtot / n
    p.maxp = n
    p.maxvxyzu = 4
    p.ieos = 2
    p.gamma = gamma
    xyzh = np.empty((n, 4))
    xyzh[:, :3] = pos
    nl = left.shape[0]
    xyzh[:nl, 3] = p.hfact * dx
    xyzh[nl:, 3] = p.hfact * 2.0 * dx
    vxyzu = np.zeros((n, 4))
    # u = P / ((gamma-1) rho)
    vxyzu[:nl, 3] = 1.0 / ((gamma - 1.0) * 1.0)
    vxyzu[nl:, 3] = 0.1 / ((gamma - 1.0) * 0.125)
    rng = np.random.default_rng(seed)
    alphaind = np.zeros((n, 3), dtype=np.float32)
    alphaind[:, 0] = rng.uniform(0.0, 1.0, size=n).astype(np.float32)
    return dict(xyzh=xyzh, vxyzu=vxyzu, alphaind=alphaind, params=p,
                info=dict(workload="shock_tube", n=n, ny=ny, dx=dx, seed=seed, kernel=kernel))


def mhd_dust_box(nx=128, seed=777, kernel=abi.KERNEL_QUINTIC, jitter=0.0, K_code=None):
    """Orszag-Tang-type MHD box with one-fluid dust in a periodic unit cube.

    rho = 25/(36 pi), P = 5/(12 pi), gamma = 5/3,
    v = (-sin 2 pi y, sin 2 pi x, 0.01), B = (-sin 2 pi y, sin 4 pi x, 0)/sqrt(4 pi)
    stored as Bevol = B/rho with psi = 0.  Dust fraction eps is log-uniform in
    [0.01, 0.5] (seeded) and is passed as dustevol = sqrt(eps/(1-eps))
    (cons2prim.f90:350 inverts it).  Stopping times use the constant drag
    coefficient law t_s = rho_g rho_d / (K rho) (dust.f90:249-256).
    """
    pos, dx = _cubic_lattice(nx, nx, nx, -0.5, 0.5, -0.5, 0.5, -0.5, 0.5)
    rng = np.random.default_rng(seed)
    n = pos.shape[0]
    if jitter:
        pos = pos + rng.uniform(-jitter * dx, jitter * dx, size=pos.shape)
        for c in range(3):
            pos[:, c] = _wrap(pos[:, c], -0.5, 0.5)
    gamma = 5.0 / 3.0
    rho0 = 25.0 / (36.0 * PI_PH)
    P0 = 5.0 / (12.0 * PI_PH)
    p = abi.default_params(kernel)
    p.massoftype_gas = rho0 / n
    p.maxp = n
    p.mhd = 1
    p.use_dustfrac = 1
    p.ndustsmall = 1
    p.idrag = 2
    # K chosen so that t_s = eps (1-eps) rho / K spans ~0.01 .. 0.5
    p.K_code = rho0 * 0.0099 / 0.01 if K_code is None else K_code
    p.gamma = gamma
    x, y = pos[:, 0] + 0.5, pos[:, 1] + 0.5
    xyzh = np.empty((n, 4))
    xyzh[:, :3] = pos
    xyzh[:, 3] = p.hfact * dx
    vxyzu = np.zeros((n, 4))
    vxyzu[:, 0] = -np.sin(2 * PI_PH * y)
    vxyzu[:, 1] = np.sin(2 * PI_PH * x)
    vxyzu[:, 2] = 0.01
    eps = np.exp(rng.uniform(math.log(0.01), math.log(0.5), size=n))
    rhogas = rho0 * (1.0 - eps)
    vxyzu[:, 3] = P0 / ((gamma - 1.0) * rhogas)
    B = np.zeros((n, 3))
    B[:, 0] = -np.sin(2 * PI_PH * y) / math.sqrt(4 * PI_PH)
    B[:, 1] = np.sin(4 * PI_PH * x) / math.sqrt(4 * PI_PH)
    Bevol = np.zeros((n, 4))
    Bevol[:, :3] = B / rho0
    dustevol = np.sqrt(eps / (1.0 - eps))
    alphaind = np.zeros((n, 3), dtype=np.float32)
    alphaind[:, 0] = rng.uniform(0.0, 1.0, size=n).astype(np.float32)
    return dict(xyzh=xyzh, vxyzu=vxyzu, alphaind=alphaind, Bevol=Bevol, dustevol=dustevol,
                params=p, info=dict(workload="mhd_dust_box", n=n, nx=nx, dx=dx, seed=seed,
                                    kernel=kernel, jitter=jitter))


def derivs_test_box(nx=100, kernel=abi.KERNEL_CUBIC, rhozero=5.0):
    """The lattice of Phantom's derivative unit test (test_derivs.f90:128-160):
    psep = dxbound/100 in the default box, rho = 5, with the analytic
    velocity and thermal energy fields vx, vy, vz, utherm
    (test_derivs.f90:1216-1242, 1604-1613)."""
    pos, dx = _cubic_lattice(nx, nx, nx, -0.5, 0.5, -0.5, 0.5, -0.5, 0.5)
    n = pos.shape[0]
    p = abi.default_params(kernel)
    p.massoftype_gas = rhozero * 1.0 / n
    p.maxp = n
    p.tolh = 1.e-5
    x, y, z = pos[:, 0] + 0.5, pos[:, 1] + 0.5, pos[:, 2] + 0.5
    xyzh = np.empty((n, 4))
    xyzh[:, :3] = pos
    xyzh[:, 3] = p.hfact * dx
    vxyzu = np.zeros((n, 4))
    vxyzu[:, 0] = 0.5 / PI_PH * np.sin(2 * PI_PH * x)
    vxyzu[:, 1] = 0.5 / PI_PH * np.sin(2 * PI_PH * x) - 0.5 / PI_PH * np.sin(2 * PI_PH * z)
    vxyzu[:, 2] = 0.05 / PI_PH * np.cos(4 * PI_PH * y)
    vxyzu[:, 3] = 0.5 / PI_PH * (3.0 + np.sin(2 * PI_PH * x) + np.cos(2 * PI_PH * y) + np.sin(2 * PI_PH * z))
    alphaind = np.zeros((n, 3), dtype=np.float32)
    hzero = p.hfact * (p.massoftype_gas / rhozero) ** (1.0 / 3.0)
    return dict(xyzh=xyzh, vxyzu=vxyzu, alphaind=alphaind, params=p,
                info=dict(workload="derivs_test_box", n=n, nx=nx, dx=dx, hzero=hzero, rhozero=rhozero))


def random_box(n=20000, seed=99, kernel=abi.KERNEL_CUBIC, hmin_frac=0.5, hmax_frac=2.0):
    """Uniform random positions with random h (for neighbour-count checks in
    the spirit of test_neigh.f90:122-396)."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.5, 0.5, size=(n, 3))
    p = abi.default_params(kernel)
    p.massoftype_gas = 1.0 / n
    p.maxp = n
    dx = n ** (-1.0 / 3.0)
    xyzh = np.empty((n, 4))
    xyzh[:, :3] = pos
    xyzh[:, 3] = p.hfact * dx * rng.uniform(hmin_frac, hmax_frac, size=n)
    vxyzu = np.zeros((n, 4))
    vxyzu[:, :3] = rng.normal(0.0, 0.1, size=(n, 3))
    vxyzu[:, 3] = 1.0 + 0.1 * rng.uniform(size=n)
    alphaind = np.zeros((n, 3), dtype=np.float32)
    alphaind[:, 0] = rng.uniform(0.0, 1.0, size=n).astype(np.float32)
    return dict(xyzh=xyzh, vxyzu=vxyzu, alphaind=alphaind, params=p,
                info=dict(workload="random_box", n=n, seed=seed, kernel=kernel))


def with_bins(setup, nbinmax=3, istepfrac=2, seed=5, dtmax=0.016):
    """Block-timestep version of a setup: ibin_old (= ibin on entry) set in
    slabs along x (bin k in the k-th of nbinmax+1 slabs) with 10% of the
    particles redrawn uniformly, ibin_wake = 0, and iphase set as
    set_active_particles does (utils_indtimesteps.f90:114-185): a particle is
    active when mod(istepfrac, 2**(nbinmax - ibin)) == 0.  dtmax places the
    particles' timesteps a few bins below it."""
    s = dict(setup)
    p = s["params"]
    p.ind_timesteps = 1
    p.dtmax = dtmax
    n = s["xyzh"].shape[0]
    rng = np.random.default_rng(seed)
    fx = (s["xyzh"][:, 0] - p.xmin) / (p.xmax - p.xmin)
    ibin_old = np.minimum((fx * (nbinmax + 1)).astype(np.int64), nbinmax)
    redraw = rng.random(n) < 0.1
    ibin_old[redraw] = rng.integers(0, nbinmax + 1, size=int(redraw.sum()))
    ibin_old = ibin_old.astype(np.int8)
    active = (istepfrac % (2 ** (nbinmax - ibin_old.astype(np.int64)))) == 0
    s["iphase"] = np.where(active, 1, -1).astype(np.int8)
    ibinnow = nbinmax
    for k in range(nbinmax):
        if istepfrac % 2 ** (nbinmax - k) == 0:
            ibinnow = k
            break
    s["bins"] = dict(ibin=ibin_old.copy(), ibin_old=ibin_old, ibin_wake=np.zeros(n, np.int8), nbinmax=nbinmax,
                     istepfrac=istepfrac, ibinnow=ibinnow)
    return s


def with_levels(setup, levels, nbinmax, istepfrac, seed=5, dtmax=0.016):
    """Block-timestep version of a setup with a chosen population per bin:
    levels = [(ibin, fraction), ...] (fractions summing to 1) assigns each
    bin round(fraction * n) particles drawn by a seeded permutation, so the
    active set of substep istepfrac (mod(istepfrac, 2**(nbinmax - ibin))
    == 0, utils_indtimesteps.f90:114-185) has an exact, spatially random
    fraction of the particles.  ibin_wake = 0; iphase and ibinnow as
    set_active_particles leaves them."""
    s = dict(setup)
    p = s["params"]
    p.ind_timesteps = 1
    p.dtmax = dtmax
    n = s["xyzh"].shape[0]
    perm = np.random.default_rng(seed).permutation(n)
    ibin_old = np.zeros(n, np.int8)
    start = 0
    for k, (b, frac) in enumerate(levels):
        stop = n if k == len(levels) - 1 else start + int(round(frac * n))
        ibin_old[perm[start:stop]] = b
        start = stop
    active = (istepfrac % (2 ** (nbinmax - ibin_old.astype(np.int64)))) == 0
    s["iphase"] = np.where(active, 1, -1).astype(np.int8)
    ibinnow = nbinmax
    for k in range(nbinmax):
        if istepfrac % 2 ** (nbinmax - k) == 0:
            ibinnow = k
            break
    s["bins"] = dict(ibin=ibin_old.copy(), ibin_old=ibin_old, ibin_wake=np.zeros(n, np.int8), nbinmax=nbinmax,
                     istepfrac=istepfrac, ibinnow=ibinnow)
    return s


# Four bins whose active sets are 1, 1/2, 1/8 and 1/64 of the particles at
# substeps istepfrac = 64, 32, 8 and 1 of a nbinmax = 6 hierarchy.
LEVELS_4 = [(0, 32 / 64), (1, 24 / 64), (3, 7 / 64), (6, 1 / 64)]
# Ten bins 0..9 holding 2^-(k+1) of the particles (bin 9: 2^-9): the active
# fraction at istepfrac = 2^m of a nbinmax = 9 hierarchy is 2^-(9-m).
LEVELS_10 = [(k, 2.0 ** -(k + 1)) for k in range(9)] + [(9, 2.0 ** -9)]


def lane_emden(nindex, dxi=1e-4):
    """Lane-Emden solution theta(xi) for polytropic index nindex, integrated
    with RK4 from the centre to the first zero xi1.  Returns (xi, theta,
    dtheta/dxi) up to and including xi1."""
    def rhs(xi, y):
        th, dth = y
        return np.array([dth, -max(th, 0.0) ** nindex - 2.0 * dth / xi])
    xi = dxi
    y = np.array([1.0 - xi * xi / 6.0, -xi / 3.0])
    xs, ths, dths = [0.0, xi], [1.0, y[0]], [0.0, y[1]]
    while y[0] > 0.0:
        k1 = rhs(xi, y)
        k2 = rhs(xi + 0.5 * dxi, y + 0.5 * dxi * k1)
        k3 = rhs(xi + 0.5 * dxi, y + 0.5 * dxi * k2)
        k4 = rhs(xi + dxi, y + dxi * k3)
        ynew = y + dxi / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
        if ynew[0] <= 0.0:  # linear interpolation to the surface
            f = y[0] / (y[0] - ynew[0])
            xi = xi + f * dxi
            y = y + f * (ynew - y)
            y[0] = 0.0
        else:
            xi, y = xi + dxi, ynew
        xs.append(xi); ths.append(y[0]); dths.append(y[1])
    return np.array(xs), np.array(ths), np.array(dths)


def polytrope(n=8000, nindex=1.5, seed=42, kernel=abi.KERNEL_CUBIC, jitter=0.1, gamma=5.0 / 3.0,
              pressureless=False, theta=0.5):
    """Self-gravitating polytrope, M = R = G = 1: a jittered cubic lattice
    cut to the unit sphere and stretched radially so that the enclosed mass
    follows the Lane-Emden profile of index nindex; u = K rho^(1/n) n (the
    polytropic relation P = K rho^(1+1/n) with gamma = 1 + 1/n for the
    thermal energy).  pressureless=True sets an isothermal gas with c_s = 0
    (pure gravity, for comparisons with direct summation)."""
    rng = np.random.default_rng(seed)
    # lattice with about n points inside the unit sphere
    nside = int(round((n * 6.0 / math.pi) ** (1.0 / 3.0)))
    dx = 2.0 / nside
    g = (np.arange(nside) + 0.5) * dx - 1.0
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    pos += rng.uniform(-jitter * dx, jitter * dx, size=pos.shape)
    r0 = np.sqrt((pos * pos).sum(axis=1))
    pos = pos[r0 < 1.0]
    r0 = r0[r0 < 1.0]
    npart = pos.shape[0]
    # stretch map: mass fraction r0^3 -> radius with the same enclosed mass
    xi, th, dth = lane_emden(nindex)
    xi1 = xi[-1]
    menc = -(xi * xi) * dth            # proportional to the enclosed mass
    menc = menc / menc[-1]
    r_of_m = np.interp(r0 ** 3, menc, xi / xi1)
    pos = pos * (r_of_m / np.maximum(r0, 1e-300))[:, None]
    rho_c = 1.0 / (4.0 * math.pi * (1.0 / xi1) ** 3 * xi1 * xi1 * (-dth[-1]))
    alpha = 1.0 / xi1
    K = 4.0 * math.pi * alpha * alpha * rho_c ** (1.0 - 1.0 / nindex) / (nindex + 1.0)
    rho = rho_c * np.interp(r_of_m * xi1, xi, th) ** nindex
    p = abi.default_params(kernel)
    p.periodic = 0
    p.gravity = 1
    p.tree_accuracy = theta
    p.massoftype_gas = 1.0 / npart
    p.maxp = npart
    p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax = -2.0, 2.0, -2.0, 2.0, -2.0, 2.0
    xyzh = np.empty((npart, 4))
    xyzh[:, :3] = pos
    rho_est = np.maximum(rho, 1e-3 * rho_c)
    xyzh[:, 3] = p.hfact * (p.massoftype_gas / rho_est) ** (1.0 / 3.0)
    if pressureless:
        p.ieos = 1
        p.polyk = 0.0
        p.maxvxyzu = 3
        vxyzu = np.zeros((npart, 3))
    else:
        p.ieos = 2
        p.gamma = gamma
        p.maxvxyzu = 4
        vxyzu = np.zeros((npart, 4))
        vxyzu[:, 3] = K * rho_est ** (1.0 / nindex) / (gamma - 1.0)
    alphaind = np.zeros((npart, 3), dtype=np.float32)
    return dict(xyzh=xyzh, vxyzu=vxyzu, alphaind=alphaind, params=p,
                info=dict(workload="polytrope", n=npart, nindex=nindex, seed=seed, K=K, rho_c=rho_c))


def two_stars(n_each=10000, seed=11, separation=20.0, theta=0.5, kernel=abi.KERNEL_CUBIC):
    """The two-star configuration of Phantom's symmetric-FMM momentum test
    (test_gravity.f90:634-736): two uniform random spheres of radius 1, the
    second centred at (separation, 0, 0), n_each particles each, total mass
    1 in equal masses, hfact 1.2, zero thermal energy (ideal gas with
    u = 0, so pure gravity), open boundaries, tree_accuracy theta.  The
    initial h is hfact times the mean interparticle spacing in a sphere."""
    rng = np.random.default_rng(seed)
    pts = []
    for x0 in (0.0, separation):
        r = rng.uniform(0.0, 1.0, n_each) ** (1.0 / 3.0)
        mu = rng.uniform(-1.0, 1.0, n_each)
        ph = rng.uniform(0.0, 2.0 * math.pi, n_each)
        s = np.sqrt(1.0 - mu * mu)
        pts.append(np.stack([x0 + r * s * np.cos(ph), r * s * np.sin(ph), r * mu], axis=1))
    pos = np.concatenate(pts)
    n = pos.shape[0]
    p = abi.default_params(kernel)
    p.hfact = 1.2
    p.periodic = 0
    p.gravity = 1
    p.tree_accuracy = theta
    p.ieos = 2
    p.gamma = 5.0 / 3.0
    p.maxvxyzu = 4
    p.massoftype_gas = 1.0 / n
    p.maxp = n
    p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax = -2.0, separation + 2.0, -2.0, 2.0, -2.0, 2.0
    xyzh = np.empty((n, 4))
    xyzh[:, :3] = pos
    xyzh[:, 3] = p.hfact * (4.0 / 3.0 * math.pi / n_each) ** (1.0 / 3.0)
    vxyzu = np.zeros((n, 4))
    alphaind = np.zeros((n, 3), dtype=np.float32)
    return dict(xyzh=xyzh, vxyzu=vxyzu, alphaind=alphaind, params=p,
                info=dict(workload="two_stars", n=n, seed=seed, separation=separation))


def polytrope_profile(nindex=1.5, nt=2000):
    """Tabulated (r, rho, P) of the M = R = G = 1 polytrope of index nindex
    (Lane-Emden), the table a stellar relaxation starts from: rho =
    rho_c theta^n(xi1 r), P = K rho^(1+1/n)."""
    xi, th, dth = lane_emden(nindex)
    xi1 = xi[-1]
    rho_c = 1.0 / (4.0 * math.pi * (1.0 / xi1) ** 3 * xi1 * xi1 * (-dth[-1]))
    alpha = 1.0 / xi1
    K = 4.0 * math.pi * alpha * alpha * rho_c ** (1.0 - 1.0 / nindex) / (nindex + 1.0)
    r = np.linspace(0.0, 1.0, nt)
    # theta floored just above zero at the surface so that the entropy
    # P/rho^gamma = K is defined at every table point
    rho = rho_c * np.maximum(np.interp(r * xi1, xi, th), 1e-6) ** nindex
    return r, rho, K * rho ** (1.0 + 1.0 / nindex)


def relaxation_target(n=1048576, nindex=3.0, sigma=0.02, seed=2024, kernel=abi.KERNEL_CUBIC):
    """The differentiable-relaxation workload: an n = 3 polytrope (M = R = 1)
    sampled by the radial stretch map of polytrope() plus seeded Gaussian
    noise of sigma * h on every coordinate; hydrodynamics only (the adjoint
    kernels have no self-gravity), ideal gas with gamma = 1 + 1/n and
    u = K rho^(1/n) n held per particle, open boundaries.  Returns the setup
    and the target profile table (r, rho_t) of the same polytrope."""
    s = polytrope(n=n, nindex=nindex, seed=seed, kernel=kernel, gamma=1.0 + 1.0 / nindex)
    p = s["params"]
    p.gravity = 0
    rng = np.random.default_rng(seed + 1)
    h = s["xyzh"][:, 3]
    s["xyzh"][:, :3] += rng.normal(0.0, 1.0, size=(h.shape[0], 3)) * (sigma * h)[:, None]
    r, rho, _ = polytrope_profile(nindex)
    s["target"] = (r, rho)
    return s
