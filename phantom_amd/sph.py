"""Host-side interface to libphsph.so, mirroring the reference's calls.

``Context`` keeps one library context (one HIP stream, the device-resident
tree and particle mirror) and exposes the reference's entry points with the
reference's names and argument meaning:

    maketree(xyzh, iphase)                     kdtree.F90:117
    densityiterate(icall, xyzh, vxyzu, ...)    dens.F90:117
    cons2prim(xyzh, vxyzu, dvdx, alphaind,...) cons2prim.f90:274
    force(icall, xyzh, vxyzu, ...)             force.F90:193
    get_neighbour_stats()                      dens.F90:1137
    step_predictor / substep_sph / step_predict_sph / step_corrector /
    step_reset                                 step_leapfrog.f90:150-696,
                                               substepping.F90:241

Arrays are numpy arrays (host memory) or phantom_amd.device.DeviceArray
(HBM), in the reference layout: Fortran a(ncomp, n) is a C-contiguous array of shape
(n, ncomp).  Errors raise PhsphError carrying the C status code, as the
reference would call fatal().  The library must be built (python -m
phantom_amd.build); there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libphsph.so")

EXPORTED = [
    "phsph_default_params", "phsph_create", "phsph_destroy", "phsph_set_params", "phsph_last_error",
    "phsph_maketree", "phsph_get_tree", "phsph_densityiterate", "phsph_cons2prim", "phsph_force",
    "phsph_get_neighbour_stats", "phsph_get_particle_stats", "phsph_get_force_stats", "phsph_get_timings",
    "phsph_set_bins", "phsph_get_nbinmax", "phsph_set_active_particles", "phsph_set_poten",
    "phsph_get_quads", "phsph_init_step", "phsph_step_predictor", "phsph_substep_sph", "phsph_step_predict_sph",
    "phsph_step_corrector", "phsph_step_reset", "phsph_get_precision", "phsph_density_vjp", "phsph_force_vjp",
    "phsph_step_vjp", "phsph_step_corrector_noalloc", "phsph_get_tree_grav", "phsph_set_tuning",
    "phsph_get_tuning", "phsph_step_graph",
]

_lib = None


class PhsphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{abi.ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


def load_library(path=LIB_PATH):
    """Load libphsph.so and declare the C signatures.  Raises if missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -m phantom_amd.build`")
    L = C.CDLL(path)
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    P = C.POINTER
    L.phsph_default_params.argtypes = [P(abi.Params), i32]
    L.phsph_default_params.restype = None
    L.phsph_create.argtypes = [P(abi.Params), i32, P(vp)]
    L.phsph_destroy.argtypes = [vp]
    L.phsph_set_params.argtypes = [vp, P(abi.Params)]
    L.phsph_last_error.argtypes = [vp]
    L.phsph_last_error.restype = C.c_char_p
    L.phsph_maketree.argtypes = [vp, vp, i32, i32, vp, vp, P(i64)]
    L.phsph_get_tree.argtypes = [vp, vp, vp, vp, i64, P(i32)]
    L.phsph_densityiterate.argtypes = [vp, i32, i32, vp, i32, vp, i32, vp, i32, vp, vp, vp, vp, i32, vp, vp, vp,
                                       i32, vp]
    L.phsph_cons2prim.argtypes = [vp, i32, vp, i32, vp, i32, vp, vp, vp, vp, vp, i32, vp]
    L.phsph_force.argtypes = [vp, i32, i32, vp, i32, vp, i32, vp, i32, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                              i32, vp, vp, i32, vp, P(dbl), P(dbl)]
    L.phsph_get_neighbour_stats.argtypes = [vp, P(dbl), P(dbl), P(i32), P(i32), P(i64), P(i64)]
    L.phsph_get_particle_stats.argtypes = [vp, i32, vp, vp]
    L.phsph_get_force_stats.argtypes = [vp, P(i64)]
    L.phsph_get_precision.argtypes = [vp, P(i32), P(i32)]
    L.phsph_density_vjp.argtypes = [vp, i32, vp, i32, vp, vp, vp, i32]
    L.phsph_force_vjp.argtypes = [vp, i32, vp, i32, vp, i32, vp, i32, vp, i32, vp, i32]
    L.phsph_step_vjp.argtypes = [vp, i32, dbl, vp, i32, vp, i32, vp, vp, vp, i32, vp, i32, vp, vp, i32]
    L.phsph_get_timings.argtypes = [vp, P(C.c_char_p), P(dbl), i32]
    L.phsph_set_bins.argtypes = [vp, vp, vp, vp, i32, i32, i32]
    L.phsph_get_nbinmax.argtypes = [vp, P(i32)]
    L.phsph_set_poten.argtypes = [vp, vp]
    L.phsph_get_quads.argtypes = [vp, vp, i64]
    L.phsph_get_tree_grav.argtypes = [vp, vp, i64]
    L.phsph_set_tuning.argtypes = [vp, i32, i32, i32, i32]
    L.phsph_get_tuning.argtypes = [vp, P(i32), P(i32), P(i32), P(i32), P(i32), vp, vp, P(i32), P(dbl)]
    L.phsph_set_active_particles.argtypes = [vp, i32, vp, i32, vp, vp, i32, i32, P(i32), P(i32), P(i32)]
    L.phsph_init_step.argtypes = [vp, i32, dbl, vp, i32, vp]
    L.phsph_step_predictor.argtypes = [vp, i32, dbl, dbl, vp, i32, vp, i32, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp,
                                       dbl, P(i32)]
    L.phsph_substep_sph.argtypes = [vp, i32, dbl, vp, i32, vp, i32]
    L.phsph_step_predict_sph.argtypes = [vp, i32, dbl, dbl, vp, i32, vp, i32, vp, i32, vp, vp, vp, vp, vp, vp, vp,
                                         vp, i32, vp, i32, vp, vp, vp, dbl, P(i32)]
    L.phsph_step_corrector.argtypes = [vp, i32, dbl, dbl, vp, i32, vp, i32, vp, i32, vp, vp, vp, vp, vp, vp, vp,
                                       vp, i32, vp, dbl, P(dbl), P(dbl), P(i32), P(i32), P(i32)]
    L.phsph_step_corrector_noalloc.argtypes = L.phsph_step_corrector.argtypes
    L.phsph_step_graph.argtypes = [vp, vp, vp, i32, dbl, dbl, i32, i32, vp]
    L.phsph_step_reset.argtypes = [vp, i32, dbl, vp, i32, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp]
    for name in EXPORTED:
        if name != "phsph_default_params" and name != "phsph_last_error":
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _ptr(a):
    """(pointer, leading dimension) of a numpy array (host memory) or a
    phantom_amd.device.DeviceArray (HBM)."""
    if a is None:
        return None, 0
    if isinstance(a, np.ndarray):
        if not a.flags.c_contiguous:
            raise ValueError("arrays must be C-contiguous (Fortran column-major a(ncomp,n))")
        ld = a.shape[1] if a.ndim == 2 else 1
        return C.c_void_p(a.ctypes.data), ld
    ptr = getattr(a, "ptr", None)
    if ptr is None:
        raise TypeError(f"unsupported array type {type(a)}")
    ld = a.shape[1] if len(a.shape) == 2 else 1
    return C.c_void_p(ptr), ld


def _zeros_like_layout(ref, shape, dtype):
    """Allocate an output array on the same side (host/device) as ``ref``."""
    if isinstance(ref, np.ndarray):
        return np.zeros(shape, dtype=dtype)
    from .device import DeviceArray
    return DeviceArray(shape, dtype)


class Context:
    """One MI355X context for the SPH hot path."""

    def __init__(self, params, device=0):
        self.params = params
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.phsph_create(C.byref(params), device, C.byref(h))
        self._h = h
        if rc != 0:
            msg = self.lib.phsph_last_error(h).decode() if h.value else "create failed"
            if h.value:
                self.lib.phsph_destroy(h)
                self._h = None
            raise PhsphError(rc, msg)
        self.np = 0
        self._poten = None

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.phsph_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise PhsphError(rc, self.lib.phsph_last_error(self._h).decode())

    def set_params(self, params):
        self._check(self.lib.phsph_set_params(self._h, C.byref(params)))
        self.params = params

    # -- maketree -----------------------------------------------------------
    def maketree(self, xyzh, iphase=None, leaf=True):
        """Build the tree; xyzh is wrapped in place (periodic).  Returns
        (leaf_is_active (ncells,), ncells); with leaf=False the leaf flags
        are not copied out and (None, ncells) is returned."""
        n = xyzh.shape[0]
        px, ldx = _ptr(xyzh)
        pph, _ = _ptr(iphase)
        maxp = max(self.params.maxp, n)
        leaf_arr = np.zeros(2 * maxp + 1, dtype=np.int32) if leaf else None
        ncells = C.c_int64()
        self._check(self.lib.phsph_maketree(self._h, px, ldx, n, pph,
                                            leaf_arr.ctypes.data_as(C.c_void_p) if leaf else None,
                                            C.byref(ncells)))
        self.np = n
        self._ncells = ncells.value
        return (leaf_arr[:ncells.value] if leaf else None), ncells.value

    def get_tree(self):
        """Mirror of node(1:ncells), inoderange(2,ncells) (1-based positions),
        inodeparts(nlive) (signed 1-based particle indices) and maxlevel."""
        ml = C.c_int32()
        node_dt = np.dtype([("xcen", "f8", 3), ("size", "f8"), ("hmax", "f8"), ("mass", "f8"),
                            ("leftchild", "i4"), ("rightchild", "i4"), ("parent", "i4"), ("idum", "i4")])
        nc = self._ncells
        node = np.zeros(nc, dtype=node_dt)
        inoderange = np.zeros((nc, 2), dtype=np.int32)
        # the library fills the first nlive entries; particle indices are never 0
        inodeparts = np.zeros(self.np, dtype=np.int32)
        self._check(self.lib.phsph_get_tree(self._h, node.ctypes.data_as(C.c_void_p),
                                            inoderange.ctypes.data_as(C.c_void_p),
                                            inodeparts.ctypes.data_as(C.c_void_p), nc, C.byref(ml)))
        nlive = int(np.count_nonzero(inodeparts))
        return dict(node=node, inoderange=inoderange, inodeparts=inodeparts[:nlive], maxlevel=ml.value, ncells=nc)

    def set_tuning(self, cl=0, lw=0, r2=0, q=0):
        """Fix the list-kernel configuration (cl 8|16, lw 2|4, r2 0|1, q
        0|1), or with cl = 0 ask the autotuner to time them on the next
        large call."""
        self._check(self.lib.phsph_set_tuning(self._h, cl, lw, r2, q))

    def get_tuning(self):
        """Current configuration, whether it was tuned, and the time (ms) of
        each configuration the autotuner tried."""
        cl, lw, r2, q, tuned, nt = (C.c_int32() for _ in range(6))
        rfac = C.c_double()
        ms = np.zeros(16, np.float32)
        cfg = np.zeros(64, np.int32)
        self._check(self.lib.phsph_get_tuning(self._h, C.byref(cl), C.byref(lw), C.byref(r2), C.byref(q),
                                              C.byref(tuned), ms.ctypes.data_as(C.c_void_p),
                                              cfg.ctypes.data_as(C.c_void_p), C.byref(nt), C.byref(rfac)))
        tried = [dict(cl=int(cfg[4 * k]), lw=int(cfg[4 * k + 1]), r2=int(cfg[4 * k + 2]), q=int(cfg[4 * k + 3]),
                      ms=float(ms[k])) for k in range(nt.value)]
        return dict(cl=cl.value, lw=lw.value, r2=r2.value, q=q.value, tuned=bool(tuned.value), tried=tried,
                    search_radius=rfac.value)

    def get_tree_grav(self):
        """node(1:ncells) in the GRAVITY layout of type(kdnode) (with
        tobecached, cached and quads) as a numpy structured array."""
        nodes = (abi.KdNodeGrav * max(self._ncells, 1))()
        self._check(self.lib.phsph_get_tree_grav(self._h, nodes, self._ncells))
        dt = np.dtype([("xcen", "<f8", (3,)), ("size", "<f8"), ("hmax", "<f8"), ("mass", "<f8"),
                       ("leftchild", "<i4"), ("rightchild", "<i4"), ("parent", "<i4"), ("idum", "<i4"),
                       ("tobecached", "<i4"), ("cached", "<i4"), ("quads", "<f8", (6,))])
        assert dt.itemsize == C.sizeof(abi.KdNodeGrav)
        return np.frombuffer(nodes, dtype=dt, count=self._ncells).copy()

    def get_quads(self):
        """node(1:ncells)%quads: second moments about each node's centre of
        mass (self-gravity only)."""
        q = np.zeros((self._ncells, 6))
        self._check(self.lib.phsph_get_quads(self._h, q.ctypes.data_as(C.c_void_p), self._ncells))
        return q

    # -- densityiterate ------------------------------------------------------
    def densityiterate(self, icall, xyzh, vxyzu, fxyzu=None, fext=None, Bevol=None, iphase=None,
                       divcurlv=None, divcurlB=None, alphaind=None, gradh=None, dvdx=None):
        """densityiterate (dens.F90:117).  Outputs not passed in are
        allocated on the side of xyzh.  Returns a dict of the outputs."""
        n = xyzh.shape[0]
        p = self.params
        if divcurlv is None:
            divcurlv = _zeros_like_layout(xyzh, (n, 1), np.float32)
        if p.mhd and divcurlB is None:
            divcurlB = _zeros_like_layout(xyzh, (n, 4), np.float32)
        if alphaind is None:
            alphaind = _zeros_like_layout(xyzh, (n, 3), np.float32)
        if gradh is None:
            gradh = _zeros_like_layout(xyzh, (n, 2 if p.gravity else 1), np.float32)
        if dvdx is None:
            dvdx = _zeros_like_layout(xyzh, (n, 9), np.float32)
        px, ldx = _ptr(xyzh)
        pv, ldv = _ptr(vxyzu)
        pf, ldf = _ptr(fxyzu)
        pe, _ = _ptr(fext)
        pB, _ = _ptr(Bevol)
        pph, _ = _ptr(iphase)
        pdc, lddc = _ptr(divcurlv)
        pdB, _ = _ptr(divcurlB)
        pa, _ = _ptr(alphaind)
        pg, ldg = _ptr(gradh)
        pdv, _ = _ptr(dvdx)
        self._check(self.lib.phsph_densityiterate(self._h, icall, n, px, ldx, pv, ldv, pf, ldf, pe, pB, pph, pdc,
                                                  lddc, pdB, pa, pg, ldg, pdv))
        return dict(divcurlv=divcurlv, divcurlB=divcurlB, alphaind=alphaind, gradh=gradh, dvdx=dvdx)

    def get_neighbour_stats(self):
        tm, am = C.c_double(), C.c_double()
        mt, ma = C.c_int32(), C.c_int32()
        nr, na = C.c_int64(), C.c_int64()
        self._check(self.lib.phsph_get_neighbour_stats(self._h, C.byref(tm), C.byref(am), C.byref(mt),
                                                       C.byref(ma), C.byref(nr), C.byref(na)))
        return dict(trialmean=tm.value, actualmean=am.value, maxtrial=mt.value, maxactual=ma.value,
                    nrhocalc=nr.value, nactualtot=na.value)

    def get_particle_stats(self, n):
        nn = np.zeros(n, dtype=np.int32)
        it = np.zeros(n, dtype=np.int32)
        self._check(self.lib.phsph_get_particle_stats(self._h, n, nn.ctypes.data_as(C.c_void_p),
                                                      it.ctypes.data_as(C.c_void_p)))
        return nn, it

    # -- cons2prim -----------------------------------------------------------
    def cons2prim(self, xyzh, vxyzu, dvdx, alphaind, dustevol=None, dustfrac=None, eos_vars=None):
        n = xyzh.shape[0]
        p = self.params
        if eos_vars is None:
            eos_vars = _zeros_like_layout(xyzh, (n, 3), np.float64)
        dust = p.use_dustfrac and p.ndustsmall > 0
        if dust and dustfrac is None:
            dustfrac = _zeros_like_layout(xyzh, (n,), np.float64)
        px, ldx = _ptr(xyzh)
        pv, ldv = _ptr(vxyzu)
        pdv, _ = _ptr(dvdx)
        pa, _ = _ptr(alphaind)
        pde, _ = _ptr(dustevol if dust else None)
        pdf, _ = _ptr(dustfrac if dust else None)
        pe, lde = _ptr(eos_vars)
        self._check(self.lib.phsph_cons2prim(self._h, n, px, ldx, pv, ldv, pdv, pa, pde, pdf, pe, lde, None))
        return dict(eos_vars=eos_vars, dustfrac=dustfrac)

    # -- force ---------------------------------------------------------------
    def force(self, icall, xyzh, vxyzu, gradh, alphaind, eos_vars, divcurlB=None, Bevol=None, dustfrac=None,
              iphase=None, fxyzu=None, divcurlv=None, dBevol=None, divBsymm=None, ddustevol=None, deltav=None,
              tstop=None):
        """force (force.F90:193).  Returns a dict of the outputs and the
        global dtcourant / dtforce."""
        n = xyzh.shape[0]
        p = self.params
        dust = p.use_dustfrac and p.ndustsmall > 0
        if fxyzu is None:
            fxyzu = _zeros_like_layout(xyzh, (n, 4), np.float64)
        if divcurlv is None:
            divcurlv = _zeros_like_layout(xyzh, (n, 1), np.float32)
        if p.mhd:
            if dBevol is None:
                dBevol = _zeros_like_layout(xyzh, (n, 4), np.float64)
            if divBsymm is None:
                divBsymm = _zeros_like_layout(xyzh, (n,), np.float32)
        if dust:
            if ddustevol is None:
                ddustevol = _zeros_like_layout(xyzh, (n,), np.float64)
            if deltav is None:
                deltav = _zeros_like_layout(xyzh, (n, 3), np.float64)
            if tstop is None:
                tstop = _zeros_like_layout(xyzh, (n,), np.float64)
        if p.gravity and (self._poten is None or self._poten.shape[0] != n):
            self.set_poten(_zeros_like_layout(xyzh, (n,), np.float32))
        dtc, dtf = C.c_double(), C.c_double()
        px, ldx = _ptr(xyzh)
        pv, ldv = _ptr(vxyzu)
        pf, ldf = _ptr(fxyzu)
        pdc, lddc = _ptr(divcurlv)
        pdB, _ = _ptr(divcurlB if p.mhd else None)
        pB, _ = _ptr(Bevol if p.mhd else None)
        pdBe, _ = _ptr(dBevol)
        pdBs, _ = _ptr(divBsymm)
        pdf, _ = _ptr(dustfrac if dust else None)
        pdd, _ = _ptr(ddustevol)
        pdl, _ = _ptr(deltav)
        pts, _ = _ptr(tstop)
        pg, ldg = _ptr(gradh)
        pa, _ = _ptr(alphaind)
        pe, lde = _ptr(eos_vars)
        pph, _ = _ptr(iphase)
        self._check(self.lib.phsph_force(self._h, icall, n, px, ldx, pv, ldv, pf, ldf, pdc, lddc, pdB, pB, pdBe,
                                         pdBs, pdf, pdd, pdl, pts, pg, ldg, pa, pe, lde, pph, C.byref(dtc),
                                         C.byref(dtf)))
        return dict(fxyzu=fxyzu, divcurlv=divcurlv, dBevol=dBevol, divBsymm=divBsymm, ddustevol=ddustevol,
                    deltav=deltav, tstop=tstop, dtcourant=dtc.value, dtforce=dtf.value,
                    poten=self._poten if p.gravity else None)

    def set_poten(self, poten):
        """Register part:poten (float32, np): the potential energy that force
        writes per active particle while params.gravity is set."""
        self._poten = poten
        pp, _ = _ptr(poten)
        self._check(self.lib.phsph_set_poten(self._h, pp))

    # -- block timesteps ------------------------------------------------------
    def set_bins(self, ibin, ibin_old, ibin_wake, nbinmax, istepfrac, ibinnow):
        """Register part:ibin / ibin_old / ibin_wake (int8, updated in place by
        force while params.ind_timesteps is set) and the timestep_ind state."""
        self._bins = (ibin, ibin_old, ibin_wake)  # keep the arrays alive
        pb, _ = _ptr(ibin)
        po, _ = _ptr(ibin_old)
        pw, _ = _ptr(ibin_wake)
        self._check(self.lib.phsph_set_bins(self._h, pb, po, pw, nbinmax, istepfrac, ibinnow))

    def get_nbinmax(self):
        v = C.c_int32()
        self._check(self.lib.phsph_get_nbinmax(self._h, C.byref(v)))
        return v.value

    def set_active_particles(self, xyzh, iphase, ibin, nbinmax, istepfrac):
        """set_active_particles (utils_indtimesteps.f90:114-185); iphase is
        set in place.  Returns (nactive, nalive, ibinnow)."""
        n = xyzh.shape[0]
        px, ldx = _ptr(xyzh)
        pph, _ = _ptr(iphase)
        pb, _ = _ptr(ibin)
        na, nl, bn = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.phsph_set_active_particles(self._h, n, px, ldx, pph, pb, nbinmax, istepfrac,
                                                        C.byref(na), C.byref(nl), C.byref(bn)))
        return na.value, nl.value, bn.value

    # -- leapfrog integrator loops (step_leapfrog.f90) ------------------------
    def init_step(self, time, ibin, nbinmax, twas):
        """init_step (step_leapfrog.f90:44-83): twas = time + dt(ibin)/2,
        all bins set to nbinmax first when time is 0."""
        pb, _ = _ptr(ibin)
        pt, _ = _ptr(twas)
        self._check(self.lib.phsph_init_step(self._h, ibin.shape[0], time, pb, nbinmax, pt))

    def step_predictor(self, t, dtsph, xyzh, vxyzu, fxyzu, Bevol=None, dBevol=None, dustevol=None,
                       ddustevol=None, iphase=None, ibin=None, ibin_old=None, twas=None, ufloor=0.0):
        """Predictor loop of step (step_leapfrog.f90:150-237).  Returns the
        number of floored thermal energies."""
        n = xyzh.shape[0]
        px, ldx = _ptr(xyzh)
        pv, ldv = _ptr(vxyzu)
        pf, ldf = _ptr(fxyzu)
        nfl = C.c_int32()
        self._check(self.lib.phsph_step_predictor(
            self._h, n, t, dtsph, px, ldx, pv, ldv, pf, ldf, _ptr(Bevol)[0], _ptr(dBevol)[0], _ptr(dustevol)[0],
            _ptr(ddustevol)[0], _ptr(iphase)[0], _ptr(ibin)[0], _ptr(ibin_old)[0], _ptr(twas)[0], ufloor,
            C.byref(nfl)))
        return nfl.value

    def substep_sph(self, dt, xyzh, vxyzu):
        """substep_sph (substepping.F90:241-264): x += dt*v."""
        px, ldx = _ptr(xyzh)
        pv, ldv = _ptr(vxyzu)
        self._check(self.lib.phsph_substep_sph(self._h, xyzh.shape[0], dt, px, ldx, pv, ldv))

    def step_predict_sph(self, timei, dtsph, xyzh, vxyzu, fxyzu, vpred, divcurlv, eos_vars, alphaind, Bevol=None,
                         dBevol=None, Bpred=None, dustevol=None, ddustevol=None, dustpred=None, iphase=None,
                         twas=None, ufloor=0.0):
        """predict_sph loop of step (step_leapfrog.f90:307-400).  Returns the
        number of floored predicted thermal energies."""
        n = xyzh.shape[0]
        px, ldx = _ptr(xyzh)
        pv, ldv = _ptr(vxyzu)
        pf, ldf = _ptr(fxyzu)
        pvp, ldvp = _ptr(vpred)
        if ldvp != ldv:
            raise ValueError("vpred must have the layout of vxyzu")
        pdc, lddc = _ptr(divcurlv)
        pe, lde = _ptr(eos_vars)
        nfl = C.c_int32()
        self._check(self.lib.phsph_step_predict_sph(
            self._h, n, timei, dtsph, px, ldx, pv, ldv, pf, ldf, pvp, _ptr(Bevol)[0], _ptr(dBevol)[0],
            _ptr(Bpred)[0], _ptr(dustevol)[0], _ptr(ddustevol)[0], _ptr(dustpred)[0], pdc, lddc, pe, lde,
            _ptr(alphaind)[0], _ptr(iphase)[0], _ptr(twas)[0], ufloor, C.byref(nfl)))
        return nfl.value

    def step_corrector(self, timei, dtsph, xyzh, vxyzu, fxyzu, vpred=None, Bevol=None, dBevol=None,
                       dustevol=None, ddustevol=None, iphase=None, ibin=None, ibin_wake=None, nbinmax=0,
                       twas=None, ufloor=0.0):
        """One pass of the corrector loop (step_leapfrog.f90:482-625).
        Returns errmax, v2sum, np (the check_velocity_error inputs), nwake
        and nvfloor."""
        n = xyzh.shape[0]
        px, ldx = _ptr(xyzh)
        pv, ldv = _ptr(vxyzu)
        pf, ldf = _ptr(fxyzu)
        em, v2 = C.c_double(), C.c_double()
        npc, nw, nfl = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.phsph_step_corrector(
            self._h, n, timei, dtsph, px, ldx, pv, ldv, pf, ldf, _ptr(vpred)[0], _ptr(Bevol)[0], _ptr(dBevol)[0],
            _ptr(dustevol)[0], _ptr(ddustevol)[0], _ptr(iphase)[0], _ptr(ibin)[0], _ptr(ibin_wake)[0], nbinmax,
            _ptr(twas)[0], ufloor, C.byref(em), C.byref(v2), C.byref(npc), C.byref(nw), C.byref(nfl)))
        return dict(errmax=em.value, v2sum=v2.value, np=npc.value, nwake=nw.value, nvfloor=nfl.value)

    def step_corrector_noalloc(self, timei, dtsph, xyzh, vxyzu, fxyzu, vpred=None, Bevol=None, dBevol=None,
                               dustevol=None, ddustevol=None, iphase=None, ibin=None, ibin_wake=None, nbinmax=0,
                               twas=None, ufloor=0.0):
        """step_corrector through phsph_step_corrector_noalloc: every array
        must be a device array (phantom_amd.device.DeviceArray); the library
        allocates nothing during the call."""
        n = xyzh.shape[0]
        px, ldx = _ptr(xyzh)
        pv, ldv = _ptr(vxyzu)
        pf, ldf = _ptr(fxyzu)
        em, v2 = C.c_double(), C.c_double()
        npc, nw, nfl = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.phsph_step_corrector_noalloc(
            self._h, n, timei, dtsph, px, ldx, pv, ldv, pf, ldf, _ptr(vpred)[0], _ptr(Bevol)[0], _ptr(dBevol)[0],
            _ptr(dustevol)[0], _ptr(ddustevol)[0], _ptr(iphase)[0], _ptr(ibin)[0], _ptr(ibin_wake)[0], nbinmax,
            _ptr(twas)[0], ufloor, C.byref(em), C.byref(v2), C.byref(npc), C.byref(nw), C.byref(nfl)))
        return dict(errmax=em.value, v2sum=v2.value, np=npc.value, nwake=nw.value, nvfloor=nfl.value)

    def step_reset(self, dtsph, vxyzu, fxyzu, vpred, Bevol=None, dBevol=None, Bpred=None, dustevol=None,
                   ddustevol=None, dustpred=None, iphase=None):
        """until_converged loop of step (step_leapfrog.f90:650-696)."""
        pv, ldv = _ptr(vxyzu)
        pf, ldf = _ptr(fxyzu)
        self._check(self.lib.phsph_step_reset(
            self._h, vxyzu.shape[0], dtsph, pv, ldv, pf, ldf, _ptr(vpred)[0], _ptr(Bevol)[0], _ptr(dBevol)[0],
            _ptr(Bpred)[0], _ptr(dustevol)[0], _ptr(ddustevol)[0], _ptr(dustpred)[0], _ptr(iphase)[0]))

    def step_graph(self, state, clock, nsub, ufloor=0.0, tolv=1.e-2, maxits=30, graphs=True):
        """nsub steps of the evolve loop inside the library
        (phsph_step_graph) on device-resident state.  state: abi.StepState,
        clock: abi.StepClock (updated in place).  Returns one dict per step
        (dt, dtnew, dterr, its, nactive, nwake, nvfloorp/ps/c)."""
        out = (abi.StepSummary * max(nsub, 1))()
        self._check(self.lib.phsph_step_graph(self._h, C.byref(state), C.byref(clock), nsub, ufloor, tolv, maxits,
                                              1 if graphs else 0, out))
        return [{k: getattr(out[i], k) for k, _ in abi.StepSummary._fields_} for i in range(nsub)]

    # -- adjoints --------------------------------------------------------------
    def density_vjp(self, xyzh, rhobar=None, hbar=None, dxyz=None):
        """dL/dx for L = sum rhobar.rho + hbar.h at the converged h of the
        last densityiterate on these positions (tree current)."""
        n = xyzh.shape[0]
        if dxyz is None:
            dxyz = _zeros_like_layout(xyzh, (n, 3), np.float64)
        px, ldx = _ptr(xyzh)
        pr, _ = _ptr(rhobar)
        ph, _ = _ptr(hbar)
        pd, ldd = _ptr(dxyz)
        self._check(self.lib.phsph_density_vjp(self._h, n, px, ldx, pr, ph, pd, ldd))
        return dxyz

    def force_vjp(self, xyzh, vxyzu=None, fbar=None, fxyz=None, dxyz=None, want_accel=True):
        """The pressure acceleration (barotropic closure) and, with fbar, its
        adjoint dL/dx of L = sum fbar.a.  Returns (a, dxyz)."""
        n = xyzh.shape[0]
        if want_accel and fxyz is None:
            fxyz = _zeros_like_layout(xyzh, (n, 3), np.float64)
        if fbar is not None and dxyz is None:
            dxyz = _zeros_like_layout(xyzh, (n, 3), np.float64)
        px, ldx = _ptr(xyzh)
        pv, ldv = _ptr(vxyzu)
        pb, ldb = _ptr(fbar)
        pf, ldf = _ptr(fxyz)
        pd, ldd = _ptr(dxyz)
        self._check(self.lib.phsph_force_vjp(self._h, n, px, ldx, pv, ldv, pb, ldb, pf, ldf, pd, ldd))
        return fxyz, dxyz

    def step_vjp(self, dt, xyzh, vxyzu, xbar=None, vbar=None, want_state=True):
        """One kick-drift-kick step with the pressure acceleration; returns
        (xyzh_new, vxyzu_new, dL/dx, dL/dv) with L = xbar.x' + vbar.v' (the
        gradients are None without xbar and vbar)."""
        n = xyzh.shape[0]
        xn = _zeros_like_layout(xyzh, xyzh.shape, np.float64) if want_state else None
        vn = _zeros_like_layout(xyzh, vxyzu.shape, np.float64) if want_state else None
        back = xbar is not None or vbar is not None
        dx = _zeros_like_layout(xyzh, (n, 3), np.float64) if back else None
        dv = _zeros_like_layout(xyzh, (n, 3), np.float64) if back else None
        px, ldx = _ptr(xyzh)
        pv, ldv = _ptr(vxyzu)
        pxb, ldxb = _ptr(xbar)
        pvb, ldvb = _ptr(vbar)
        pxn, _ = _ptr(xn)
        pvn, _ = _ptr(vn)
        pdx, ldd = _ptr(dx)
        pdv, _ = _ptr(dv)
        self._check(self.lib.phsph_step_vjp(self._h, n, float(dt), px, ldx, pv, ldv, pxn, pvn, pxb, ldxb, pvb, ldvb,
                                            pdx, pdv, ldd if back else 3))
        return xn, vn, dx, dv

    def get_precision(self):
        """(density, force) precision the last calls ran in (abi.PREC_*,
        -1 before the first call)."""
        d, f = C.c_int32(), C.c_int32()
        self._check(self.lib.phsph_get_precision(self._h, C.byref(d), C.byref(f)))
        return d.value, f.value

    def get_force_stats(self):
        n = C.c_int64()
        self._check(self.lib.phsph_get_force_stats(self._h, C.byref(n)))
        return dict(npairs=n.value)

    def timings(self):
        """Kernel times (ms) of the last call, measured with HIP events."""
        cap = 4096
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        n = self.lib.phsph_get_timings(self._h, names, ms, cap)
        return [(names[k].decode(), ms[k]) for k in range(min(n, cap))]
