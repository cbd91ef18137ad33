"""Host driver of one leapfrog step, mirroring the reference's ``step``
(step_leapfrog.f90:91-760) and ``derivs`` (deriv.f90:1-197) over the
library's entry points.

The particle loops run on the device (``Context.step_*``); what stays on the
host is what the reference also keeps scalar: the order of the loops, the
velocity-error test of the corrector (check_velocity_error,
step_leapfrog.f90:769-843), the iteration that re-evaluates forces
(derivs with icall = 2) when it fails, and the timestep bookkeeping of
block timesteps (evol_prestep, evolve.f90:165-257; update_time_and_dt and
change_nbinmax, evolve_utils.F90, utils_indtimesteps.f90:186-223).

``State`` holds the particle arrays of part.F90 that the step touches, in
the reference layout, as numpy arrays (host) or DeviceArrays (HBM); with
device arrays no particle data crosses PCIe during a step.  The backend is
any object with the call surface of ``phantom_amd.sph.Context``.
"""
import numpy as np

from .sph import _zeros_like_layout

BIGNUMBER = 1.e29  # timestep.f90:40
MAXITS = 30        # step_leapfrog.f90:147


class State:
    """Particle arrays of part.F90 used by one step.  Arrays for physics
    that is off are None."""

    def __init__(self, params, xyzh, vxyzu, alphaind, Bevol=None, dustevol=None, device=False):
        p = params
        n = xyzh.shape[0]
        self.params = p
        self.n = n
        nv = p.maxvxyzu

        def put(a, shape, dtype):
            if a is None:
                a = np.zeros(shape, dtype)
            a = np.ascontiguousarray(np.asarray(a, dtype).reshape(shape))
            if device:
                from .device import DeviceArray
                return DeviceArray.from_numpy(a)
            return a.copy()

        def new(shape, dtype):
            ref = self.xyzh
            return _zeros_like_layout(ref, shape, dtype)

        self.xyzh = put(xyzh, (n, 4), np.float64)
        self.vxyzu = put(vxyzu[:, :nv], (n, nv), np.float64)
        self.vpred = new((n, nv), np.float64)
        self.fxyzu = new((n, nv), np.float64)
        self.alphaind = put(alphaind, (n, 3), np.float32)
        self.divcurlv = new((n, 1), np.float32)
        self.gradh = new((n, 2 if p.gravity else 1), np.float32)
        self.dvdx = new((n, 9), np.float32)
        self.eos_vars = new((n, 3), np.float64)
        mhd = bool(p.mhd)
        dust = bool(p.use_dustfrac and p.ndustsmall > 0)
        self.Bevol = put(Bevol, (n, 4), np.float64) if mhd else None
        self.Bpred = new((n, 4), np.float64) if mhd else None
        self.dBevol = new((n, 4), np.float64) if mhd else None
        self.divcurlB = new((n, 4), np.float32) if mhd else None
        self.divBsymm = new((n,), np.float32) if mhd else None
        self.dustevol = put(dustevol, (n,), np.float64) if dust else None
        self.dustpred = new((n,), np.float64) if dust else None
        self.ddustevol = new((n,), np.float64) if dust else None
        self.dustfrac = new((n,), np.float64) if dust else None
        self.deltav = new((n, 3), np.float64) if dust else None
        self.tstop = new((n,), np.float64) if dust else None
        self.iphase = None
        # block timesteps (timestep_ind, part:ibin/ibin_old/ibin_wake/twas)
        self.ibin = self.ibin_old = self.ibin_wake = self.twas = None
        self.nbinmax = 0
        self.istepfrac = 0
        self.ibinnow = 0
        # global timestep constraints of the last force call
        self.dtcourant = BIGNUMBER
        self.dtforce = BIGNUMBER

    ARRAYS = ("xyzh", "vxyzu", "vpred", "fxyzu", "alphaind", "divcurlv", "gradh", "dvdx", "eos_vars", "Bevol",
              "Bpred", "dBevol", "divcurlB", "divBsymm", "dustevol", "dustpred", "ddustevol", "dustfrac", "deltav",
              "tstop", "iphase", "ibin", "ibin_old", "ibin_wake", "twas")
    SCALARS = ("nbinmax", "istepfrac", "ibinnow", "dtcourant", "dtforce")

    def copy(self, device=None):
        """A copy of the state: device=True puts every array in HBM, False
        on the host, None keeps each where it is."""
        from .device import DeviceArray
        c = object.__new__(State)
        c.params, c.n = self.params, self.n
        for k in self.ARRAYS:
            a = getattr(self, k)
            on_dev = a is not None and not isinstance(a, np.ndarray)
            want = on_dev if device is None else device
            if a is None:
                b = None
            elif want:
                if on_dev:
                    b = DeviceArray(a.shape, a.dtype)
                    b.copy_from(a)
                else:
                    b = DeviceArray.from_numpy(a)
            else:
                b = a.numpy() if on_dev else a.copy()
            setattr(c, k, b)
        for k in self.SCALARS:
            setattr(c, k, getattr(self, k))
        return c

    def enable_bins(self, nbinmax, ibin=None):
        """Allocate the block-timestep arrays (host side; the library stages
        them).  ibin defaults to nbinmax for every particle."""
        n = self.n
        self.ibin = np.full(n, nbinmax, np.int8) if ibin is None else np.asarray(ibin, np.int8).copy()
        self.ibin_old = self.ibin.copy()
        self.ibin_wake = np.zeros(n, np.int8)
        self.twas = np.zeros(n, np.float64)
        self.iphase = np.ones(n, np.int8)
        self.nbinmax = nbinmax


def derivs(ctx, icall, s, timei=0.0, dt=0.0):
    """derivs (deriv.f90:1-197) on the predicted state: with icall = 1 the
    tree and densities are rebuilt; cons2prim and force always run.
    Returns dtnew (deriv.f90:186-191)."""
    p = s.params
    if icall == 1:
        ctx.maketree(s.xyzh, iphase=s.iphase, leaf=False)
        ctx.densityiterate(1, s.xyzh, s.vpred, fxyzu=s.fxyzu, Bevol=s.Bpred, iphase=s.iphase,
                           divcurlv=s.divcurlv, divcurlB=s.divcurlB, alphaind=s.alphaind, gradh=s.gradh,
                           dvdx=s.dvdx)
    ctx.cons2prim(s.xyzh, s.vpred, s.dvdx, s.alphaind, dustevol=s.dustpred, dustfrac=s.dustfrac,
                  eos_vars=s.eos_vars)
    if p.ind_timesteps:
        ctx.set_bins(s.ibin, s.ibin_old, s.ibin_wake, s.nbinmax, s.istepfrac, s.ibinnow)
    f = ctx.force(icall, s.xyzh, s.vpred, s.gradh, s.alphaind, s.eos_vars, divcurlB=s.divcurlB, Bevol=s.Bpred,
                  dustfrac=s.dustfrac, iphase=s.iphase, fxyzu=s.fxyzu, divcurlv=s.divcurlv, dBevol=s.dBevol,
                  divBsymm=s.divBsymm, ddustevol=s.ddustevol, deltav=s.deltav, tstop=s.tstop)
    if p.ind_timesteps:
        # force leaves timestep_ind:nbinmax at the largest new bin of the
        # active particles (force.F90:748-792)
        s.nbinmax = ctx.get_nbinmax()
        return p.dtmax / 2.0 ** s.nbinmax
    s.dtcourant, s.dtforce = f["dtcourant"], f["dtforce"]
    return min(s.dtforce, s.dtcourant, p.dtmax)


def check_velocity_error(errmax, v2sum, npc, its, tolv, dt, dtcourant, dtforce, ind_timesteps):
    """check_velocity_error (step_leapfrog.f90:769-843) without MPI and
    damping.  Returns (converged, dterr, normalised errmax)."""
    v2mean = v2sum / npc if npc > 0 else 0.0
    if v2mean > np.finfo(np.float64).tiny:
        errmax = errmax / np.sqrt(v2mean)
    else:
        errmax = 0.0
    errtol = tolv
    dterr = np.finfo(np.float64).max
    if tolv < 1.e2:
        if not ind_timesteps:
            dtf = min(dtcourant, dtforce)
            if dt < dtf < BIGNUMBER:
                errtol = errtol * (dt / dtf) ** 2
            if its == 1 and errtol > np.finfo(np.float64).tiny and errmax > np.finfo(np.float64).eps:
                dterr = dt * np.sqrt(errtol / errmax)
        converged = errmax < tolv
    else:
        converged = True
    return converged, dterr, errmax


def _arrays(s):
    return dict(Bevol=s.Bevol, dBevol=s.dBevol, dustevol=s.dustevol, ddustevol=s.ddustevol, iphase=s.iphase)


def step(ctx, s, t, dtsph, ufloor=0.0, tolv=1.e-2, maxits=MAXITS, derivs_fn=None):
    """One leapfrog step (step_leapfrog.f90:91-760) without sinks, external
    forces or cooling.  ``derivs_fn(ctx, icall, s, timei, dt)`` replaces
    derivs (e.g. a prescribed force field); it returns dtnew.  Returns a dict
    with dtnew, the number of corrector iterations and the summary
    counters."""
    derivs_fn = derivs if derivs_fn is None else derivs_fn
    p = s.params
    ind = bool(p.ind_timesteps)
    a = _arrays(s)
    nvfloorp = ctx.step_predictor(t, dtsph, s.xyzh, s.vxyzu, s.fxyzu, ibin=s.ibin, ibin_old=s.ibin_old,
                                  twas=s.twas, ufloor=ufloor, **a)
    ctx.substep_sph(dtsph, s.xyzh, s.vxyzu)
    timei = t + dtsph
    nvfloorps = ctx.step_predict_sph(timei, dtsph, s.xyzh, s.vxyzu, s.fxyzu, s.vpred, s.divcurlv, s.eos_vars,
                                     s.alphaind, Bevol=s.Bevol, dBevol=s.dBevol, Bpred=s.Bpred,
                                     dustevol=s.dustevol, ddustevol=s.ddustevol, dustpred=s.dustpred,
                                     iphase=s.iphase, twas=s.twas, ufloor=ufloor)
    dtnew = derivs_fn(ctx, 1, s, timei, dtsph)
    its, converged = 0, False
    nwake = nvfloorc = 0
    dterr = np.finfo(np.float64).max
    while its < maxits and not converged:
        its += 1
        r = ctx.step_corrector(timei, dtsph, s.xyzh, s.vxyzu, s.fxyzu, vpred=s.vpred, ibin=s.ibin,
                               ibin_wake=s.ibin_wake, nbinmax=s.nbinmax, twas=s.twas, ufloor=ufloor, **a)
        nwake += r["nwake"]
        nvfloorc += r["nvfloor"]
        converged, dterr_i, _ = check_velocity_error(r["errmax"], r["v2sum"], r["np"], its, tolv, dtsph,
                                                     s.dtcourant, s.dtforce, ind)
        # timestep:dterr is overwritten by every call (step_leapfrog.f90:807),
        # so after an iterated step it is huge and does not limit dt
        dterr = dterr_i
        if not converged:
            ctx.step_reset(dtsph, s.vxyzu, s.fxyzu, s.vpred, Bevol=s.Bevol, dBevol=s.dBevol, Bpred=s.Bpred,
                           dustevol=s.dustevol, ddustevol=s.ddustevol, dustpred=s.dustpred, iphase=s.iphase)
            derivs_fn(ctx, 2, s, timei, dtsph)
    if maxits > 1 and its >= maxits and not converged:
        raise RuntimeError("step: VELOCITY ITERATIONS NOT CONVERGED")
    return dict(dtnew=dtnew, its=its, dterr=dterr, nwake=nwake, nvfloorp=nvfloorp, nvfloorps=nvfloorps,
                nvfloorc=nvfloorc)


def init_global(ctx, s, time=0.0, derivs_fn=None):
    """The initial derivs call of a run with a global timestep
    (initial.F90 calls derivs before the first step): vpred = vxyzu and
    the state's forces, densities and dt constraints are computed.
    Returns the first dt."""
    _copy_v_to_vpred(s)
    return (derivs if derivs_fn is None else derivs_fn)(ctx, 1, s, time, 0.0)


def _copy_v_to_vpred(s):
    if isinstance(s.vxyzu, np.ndarray):
        s.vpred[...] = s.vxyzu
        if s.Bpred is not None:
            s.Bpred[...] = s.Bevol
        if s.dustpred is not None:
            s.dustpred[...] = s.dustevol
    else:
        s.vpred.copy_from(s.vxyzu)
        if s.Bpred is not None:
            s.Bpred.copy_from(s.Bevol)
        if s.dustpred is not None:
            s.dustpred.copy_from(s.dustevol)


def evolve_global(ctx, s, nsteps, time=0.0, dt=None, **kw):
    """nsteps global-timestep steps (evolve.f90:105-160 with
    update_time_and_dt's dt = min(dtforce, dtcourant, dterr, dtmax)).
    The state must hold the derivatives of the current time (init_global).
    Returns (time, dt, the list of step summaries)."""
    p = s.params
    out = []
    if dt is None:
        dt = min(s.dtforce, s.dtcourant, p.dtmax)
    for _ in range(nsteps):
        r = step(ctx, s, time, dt, **kw)
        out.append(r)
        time = time + dt
        dt = min(s.dtforce, s.dtcourant, r["dterr"], p.dtmax + np.finfo(np.float64).eps)
    return time, dt, out


def init_ind(ctx, s, time=0.0, derivs_fn=None):
    """Start of a run on block timesteps: init_step (step_leapfrog.f90:44-83)
    and the first derivs with every particle active.  Bins then follow the
    first force call.  Returns dt = dtmax/2**nbinmax."""
    p = s.params
    _copy_v_to_vpred(s)
    s.iphase[:] = 1
    s.istepfrac = 0
    s.ibinnow = 0
    (derivs if derivs_fn is None else derivs_fn)(ctx, 1, s, time, 0.0)
    ctx.init_step(time, s.ibin, s.nbinmax, s.twas)
    return p.dtmax / 2.0 ** s.nbinmax


def evolve_ind(ctx, s, nsub, time=0.0, tlast=0.0, **kw):
    """nsub substeps on block timesteps: evol_prestep's istepfrac and
    set_active_particles (evolve.f90:200-222), step, and update_time_and_dt's
    time and nbinmax update (evolve_utils.F90, change_nbinmax
    utils_indtimesteps.f90:186-223); istepfrac restarts at the end of each
    dtmax (update_dump_counters, evolve.f90:434-440).
    Returns (time, tlast, the list of step summaries)."""
    p = s.params
    out = []
    for _ in range(nsub):
        nbinmaxprev = s.nbinmax
        s.istepfrac += 1
        if s.istepfrac > 2 ** s.nbinmax:
            raise RuntimeError("evolve: error in individual timesteps")
        nact, nalive, s.ibinnow = ctx.set_active_particles(s.xyzh, s.iphase, s.ibin, s.nbinmax, s.istepfrac)
        dt = p.dtmax / 2.0 ** s.nbinmax
        r = step(ctx, s, time, dt, **kw)
        r["nactive"] = nact
        out.append(r)
        time = tlast + s.istepfrac / float(2 ** nbinmaxprev) * p.dtmax
        if s.nbinmax != nbinmaxprev:  # change_nbinmax (utils_indtimesteps.f90:186-223)
            if s.nbinmax < nbinmaxprev and s.istepfrac % 2 ** (nbinmaxprev - s.nbinmax) != 0:
                raise RuntimeError("change_nbinmax: istepfrac not a multiple of the new bin ratio")
            if s.nbinmax > nbinmaxprev:
                s.istepfrac = s.istepfrac * 2 ** (s.nbinmax - nbinmaxprev)
            else:
                s.istepfrac = s.istepfrac // 2 ** (nbinmaxprev - s.nbinmax)
        if s.istepfrac == 2 ** s.nbinmax:
            s.istepfrac = 0
            tlast = time
    return time, tlast, out


def step_state(s):
    """The phsph_step_state of a state whose arrays are all in HBM."""
    from . import abi
    st = abi.StepState()

    def ptr(a):
        if a is None:
            return None
        if isinstance(a, np.ndarray):
            raise ValueError("step_graph needs device arrays (State.copy(device=True))")
        return a.ptr

    for k in State.ARRAYS:
        setattr(st, k, ptr(getattr(s, k)))
    st.poten = None
    st.np = s.n
    st.ldxyzh = 4
    st.ldvxyzu = s.vxyzu.shape[1]
    st.ldfxyzu = s.fxyzu.shape[1]
    st.lddivcurlv = s.divcurlv.shape[1]
    st.ldgradh = s.gradh.shape[1]
    st.ldeos = s.eos_vars.shape[1]
    return st


def evolve_graph(ctx, s, nsub, time=0.0, tlast=0.0, dt=None, graphs=True, ufloor=0.0, tolv=1.e-2, maxits=MAXITS):
    """nsub steps of evolve_ind (block timesteps) or evolve_global inside
    the library (phsph_step_graph), on a state whose arrays are all in HBM.
    Returns (time, tlast, dt, the list of step summaries); dt is the next
    global timestep (global timestepping) and the state's bin counters and
    dt constraints are updated as evolve_ind / evolve_global leave them."""
    from . import abi
    p = s.params
    k = abi.StepClock()
    k.time, k.tlast = time, tlast
    k.dtcourant, k.dtforce = s.dtcourant, s.dtforce
    k.dt = min(s.dtforce, s.dtcourant, p.dtmax) if dt is None else dt
    k.nbinmax, k.istepfrac, k.ibinnow = s.nbinmax, s.istepfrac, s.ibinnow
    out = ctx.step_graph(step_state(s), k, nsub, ufloor=ufloor, tolv=tolv, maxits=maxits, graphs=graphs)
    s.nbinmax, s.istepfrac, s.ibinnow = k.nbinmax, k.istepfrac, k.ibinnow
    s.dtcourant, s.dtforce = k.dtcourant, k.dtforce
    return k.time, k.tlast, k.dt, out
