"""ctypes wrapper of the CPU oracle (liboracle.so built from sph_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the CPU
baseline leg of bench.py use it as the checker.  Nothing in the product
package imports it.

The wrapper takes numpy arrays in the reference layout (see
phantom_amd/setups.py) and returns numpy arrays.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from phantom_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
# PHSPH_ORACLE_LIB names another build of the same source (the sanitized one
# of tests/test_oracle_sanitized.py)
_LIB = os.environ.get("PHSPH_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force=False):
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-B" if force else "liboracle.so"], cwd=_HERE, check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P = C.POINTER
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [P(abi.Params)]
        L.orc_destroy.argtypes = [vp]
        L.orc_last_error.restype = C.c_char_p
        L.orc_last_error.argtypes = [vp]
        L.orc_maketree.argtypes = [vp, vp, C.c_int32, vp]
        L.orc_get_tree.argtypes = [vp, vp, vp, vp, vp, C.c_int64, P(C.c_int64), P(C.c_int32)]
        L.orc_densityiterate.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, C.c_int32, vp, C.c_int32, vp, vp,
                                         vp, vp, C.c_int32, vp, vp, vp, C.c_int32, vp, vp, vp]
        L.orc_get_neighbour_stats.argtypes = [vp, P(C.c_double), P(C.c_double), P(C.c_int32), P(C.c_int32),
                                              P(C.c_int64), P(C.c_int64)]
        L.orc_cons2prim.argtypes = [vp, C.c_int32, vp, vp, C.c_int32, vp, vp, vp, vp, vp, C.c_int32, vp]
        L.orc_force.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, C.c_int32, vp, C.c_int32, vp, C.c_int32, vp,
                                vp, vp, vp, vp, vp, vp, vp, vp, C.c_int32, vp, vp, C.c_int32, vp,
                                P(C.c_double), P(C.c_double)]
        L.orc_kernel.argtypes = [C.c_int32, C.c_double, P(C.c_double), P(C.c_double), P(C.c_double)]
        L.orc_get_newbin.argtypes = [C.c_double, C.c_double, P(C.c_int8), C.c_int, C.c_int32, C.c_int32]
        L.orc_get_newbin.restype = None
        L.orc_set_bins.argtypes = [vp, vp, vp, vp, C.c_int32, C.c_int32, C.c_int32]
        L.orc_set_bins.restype = None
        L.orc_get_nbinmax.argtypes = [vp]
        L.orc_get_nbinmax.restype = C.c_int32
        L.orc_get_max_force_trial.argtypes = [vp]
        L.orc_get_max_force_trial.restype = C.c_int64
        L.orc_set_active_particles.argtypes = [C.c_int32, vp, vp, vp, C.c_int32, C.c_int32, P(C.c_int32),
                                               P(C.c_int32)]
        L.orc_set_active_particles.restype = C.c_int32
        L.orc_set_poten.argtypes = [vp, vp]
        L.orc_set_poten.restype = None
        L.orc_get_quads.argtypes = [vp, vp, C.c_int64]
        L.orc_kernel_softening.argtypes = [C.c_int32, C.c_double, P(C.c_double), P(C.c_double)]
        i32, dbl = C.c_int32, C.c_double
        L.orc_init_step.argtypes = [vp, i32, dbl, vp, i32, vp]
        L.orc_step_predictor.argtypes = [vp, i32, dbl, dbl, vp, vp, i32, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp,
                                         dbl, P(i32)]
        L.orc_substep_sph.argtypes = [vp, i32, dbl, vp, vp, i32]
        L.orc_step_predict_sph.argtypes = [vp, i32, dbl, dbl, vp, vp, i32, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp,
                                           i32, vp, i32, vp, vp, vp, dbl, P(i32)]
        L.orc_step_corrector.argtypes = [vp, i32, dbl, dbl, vp, vp, i32, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp,
                                         i32, vp, dbl, P(dbl), P(dbl), P(i32), P(i32), P(i32)]
        L.orc_step_reset.argtypes = [vp, i32, dbl, vp, i32, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{abi.ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


def get_newbin(dti, dtmax, ibin, allow_decrease=True, nbinmax=30, istepfrac=0):
    """get_newbin (utils_indtimesteps.f90:230-287): the bin after the step."""
    b = C.c_int8(ibin)
    lib().orc_get_newbin(float(dti), float(dtmax), C.byref(b), int(bool(allow_decrease)), nbinmax, istepfrac)
    return b.value


def set_active_particles(xyzh, ibin, nbinmax, istepfrac):
    """set_active_particles (utils_indtimesteps.f90:114-185): returns
    (iphase, nactive, nalive, ibinnow)."""
    n = xyzh.shape[0]
    iphase = np.ones(n, dtype=np.int8)
    na, nl = C.c_int32(), C.c_int32()
    ibin = np.ascontiguousarray(ibin, dtype=np.int8)
    ibinnow = lib().orc_set_active_particles(n, _ptr(np.ascontiguousarray(xyzh)), _ptr(iphase), _ptr(ibin),
                                             nbinmax, istepfrac, C.byref(na), C.byref(nl))
    if ibinnow < 0:
        raise OracleError(-1, "set_active_particles: timestep bin exceeds max bins")
    return iphase, na.value, nl.value, ibinnow


def kernel(kind, q):
    w, g, d = C.c_double(), C.c_double(), C.c_double()
    lib().orc_kernel(kind, q, C.byref(w), C.byref(g), C.byref(d))
    return w.value, g.value, d.value


def kernel_softening(kind, q):
    pot, fs = C.c_double(), C.c_double()
    lib().orc_kernel_softening(kind, q, C.byref(pot), C.byref(fs))
    return pot.value, fs.value


class Oracle:
    """One oracle context: maketree -> densityiterate -> cons2prim -> force."""

    def __init__(self, params):
        self.params = params
        self._ctx = lib().orc_create(C.byref(params))

    def __del__(self):
        if getattr(self, "_ctx", None):
            lib().orc_destroy(self._ctx)
            self._ctx = None

    def _check(self, rc):
        if rc != 0:
            raise OracleError(rc, lib().orc_last_error(self._ctx).decode())

    def maketree(self, xyzh, iphase=None, leaf=True):
        """xyzh (n,4) float64 C-contiguous, modified in place (periodic wrap).
        leaf is accepted for the call surface of phantom_amd.sph.Context; the
        tree is read with get_tree."""
        assert xyzh.dtype == np.float64 and xyzh.flags.c_contiguous and xyzh.shape[1] == 4
        self.n = xyzh.shape[0]
        self._check(lib().orc_maketree(self._ctx, _ptr(xyzh), self.n, _ptr(iphase)))

    def get_tree(self):
        ncells = C.c_int64()
        maxlevel = C.c_int32()
        lib().orc_get_tree(self._ctx, None, None, None, None, 0, C.byref(ncells), C.byref(maxlevel))
        nc = ncells.value
        node = np.zeros(nc, dtype=np.dtype([("xcen", "f8", 3), ("size", "f8"), ("hmax", "f8"),
                                             ("mass", "f8"), ("leftchild", "i4"), ("rightchild", "i4"),
                                             ("parent", "i4"), ("idum", "i4")]))
        inoderange = np.zeros((nc, 2), dtype=np.int32)
        leaf = np.zeros(nc, dtype=np.int32)
        inodeparts = np.zeros(self.n, dtype=np.int32)
        self._check(lib().orc_get_tree(self._ctx, _ptr(node), _ptr(inoderange), _ptr(inodeparts), _ptr(leaf),
                                       nc, C.byref(ncells), C.byref(maxlevel)))
        return dict(node=node, inoderange=inoderange, inodeparts=inodeparts, leaf_is_active=leaf,
                    ncells=nc, maxlevel=maxlevel.value)

    def densityiterate(self, xyzh, vxyzu, fxyzu=None, fext=None, Bevol=None, iphase=None, icall=1,
                       alphaind=None):
        n = xyzh.shape[0]
        p = self.params
        divcurlv = np.zeros((n, 1), dtype=np.float32)
        divcurlB = np.zeros((n, 4), dtype=np.float32) if p.mhd else None
        if alphaind is None:
            alphaind = np.zeros((n, 3), dtype=np.float32)
        ngradh = 2 if p.gravity else 1
        gradh = np.zeros((n, ngradh), dtype=np.float32)
        dvdx = np.zeros((n, 9), dtype=np.float32)
        nits = np.zeros(n, dtype=np.int32)
        nneigh = np.zeros(n, dtype=np.int32)
        ldv = vxyzu.shape[1]
        ldf = fxyzu.shape[1] if fxyzu is not None else 0
        self._check(lib().orc_densityiterate(self._ctx, icall, n, _ptr(xyzh), _ptr(vxyzu), ldv, _ptr(fxyzu), ldf,
                                             _ptr(fext), _ptr(Bevol), _ptr(iphase), _ptr(divcurlv), 1,
                                             _ptr(divcurlB), _ptr(alphaind), _ptr(gradh), ngradh, _ptr(dvdx),
                                             _ptr(nits), _ptr(nneigh)))
        return dict(divcurlv=divcurlv, divcurlB=divcurlB, alphaind=alphaind, gradh=gradh, dvdx=dvdx, nits=nits,
                    nneigh=nneigh)

    def neighbour_stats(self):
        tm, am = C.c_double(), C.c_double()
        mt, ma = C.c_int32(), C.c_int32()
        nr, na = C.c_int64(), C.c_int64()
        lib().orc_get_neighbour_stats(self._ctx, C.byref(tm), C.byref(am), C.byref(mt), C.byref(ma),
                                      C.byref(nr), C.byref(na))
        return dict(trialmean=tm.value, actualmean=am.value, maxtrial=mt.value, maxactual=ma.value,
                    nrhocalc=nr.value, nactualtot=na.value)

    def set_bins(self, ibin, ibin_old, ibin_wake, nbinmax, istepfrac, ibinnow):
        """Register the block-timestep arrays used by force (part:ibin,
        ibin_old, ibin_wake).  ibin and ibin_wake are updated in place."""
        self._bins = (ibin, ibin_old, ibin_wake)
        lib().orc_set_bins(self._ctx, _ptr(ibin), _ptr(ibin_old), _ptr(ibin_wake), nbinmax, istepfrac, ibinnow)

    def nbinmax(self):
        return lib().orc_get_nbinmax(self._ctx)

    def max_force_trial(self):
        """Longest cell neighbour list of the last force call."""
        return lib().orc_get_max_force_trial(self._ctx)

    def cons2prim(self, xyzh, vxyzu, dvdx, alphaind, dustevol=None, iphase=None):
        n = xyzh.shape[0]
        eos_vars = np.zeros((n, 3))
        dustfrac = np.zeros(n) if dustevol is not None else None
        self._check(lib().orc_cons2prim(self._ctx, n, _ptr(xyzh), _ptr(vxyzu), vxyzu.shape[1], _ptr(dvdx),
                                        _ptr(alphaind), _ptr(dustevol), _ptr(dustfrac), _ptr(eos_vars), 3,
                                        _ptr(iphase)))
        return dict(eos_vars=eos_vars, dustfrac=dustfrac)

    def force(self, xyzh, vxyzu, gradh, alphaind, eos_vars, divcurlB=None, Bevol=None, dustfrac=None,
              iphase=None, icall=1):
        n = xyzh.shape[0]
        p = self.params
        fxyzu = np.zeros((n, 4))
        divcurlv = np.zeros((n, 1), dtype=np.float32)
        dBevol = np.zeros((n, 4)) if p.mhd else None
        divBsymm = np.zeros(n, dtype=np.float32) if p.mhd else None
        dust = p.use_dustfrac and p.ndustsmall > 0
        ddustevol = np.zeros(n) if dust else None
        deltav = np.zeros((n, 3)) if dust else None
        tstop = np.zeros(n) if dust else None
        dtc, dtf = C.c_double(), C.c_double()
        poten = np.zeros(n, dtype=np.float32) if p.gravity else None
        lib().orc_set_poten(self._ctx, _ptr(poten))
        self._check(lib().orc_force(self._ctx, icall, n, _ptr(xyzh), _ptr(vxyzu), vxyzu.shape[1], _ptr(fxyzu), 4,
                                    _ptr(divcurlv), 1, _ptr(divcurlB), _ptr(Bevol), _ptr(dBevol), _ptr(divBsymm),
                                    _ptr(dustfrac), _ptr(ddustevol), _ptr(deltav), _ptr(tstop), _ptr(gradh),
                                    gradh.shape[1], _ptr(alphaind), _ptr(eos_vars), eos_vars.shape[1],
                                    _ptr(iphase), C.byref(dtc), C.byref(dtf)))
        lib().orc_set_poten(self._ctx, None)
        return dict(fxyzu=fxyzu, divcurlv=divcurlv, dBevol=dBevol, divBsymm=divBsymm, ddustevol=ddustevol,
                    deltav=deltav, tstop=tstop, dtcourant=dtc.value, dtforce=dtf.value, poten=poten)

    # -- leapfrog integrator loops, same calls as phantom_amd.sph.Context -----
    @staticmethod
    def _ld(a):
        return a.shape[1] if a.ndim == 2 else 1

    def init_step(self, time, ibin, nbinmax, twas):
        self._check(lib().orc_init_step(self._ctx, ibin.shape[0], time, _ptr(ibin), nbinmax, _ptr(twas)))

    def step_predictor(self, t, dtsph, xyzh, vxyzu, fxyzu, Bevol=None, dBevol=None, dustevol=None,
                       ddustevol=None, iphase=None, ibin=None, ibin_old=None, twas=None, ufloor=0.0):
        nfl = C.c_int32()
        self._check(lib().orc_step_predictor(
            self._ctx, xyzh.shape[0], t, dtsph, _ptr(xyzh), _ptr(vxyzu), self._ld(vxyzu), _ptr(fxyzu),
            self._ld(fxyzu), _ptr(Bevol), _ptr(dBevol), _ptr(dustevol), _ptr(ddustevol), _ptr(iphase), _ptr(ibin),
            _ptr(ibin_old), _ptr(twas), ufloor, C.byref(nfl)))
        return nfl.value

    def substep_sph(self, dt, xyzh, vxyzu):
        self._check(lib().orc_substep_sph(self._ctx, xyzh.shape[0], dt, _ptr(xyzh), _ptr(vxyzu), self._ld(vxyzu)))

    def step_predict_sph(self, timei, dtsph, xyzh, vxyzu, fxyzu, vpred, divcurlv, eos_vars, alphaind, Bevol=None,
                         dBevol=None, Bpred=None, dustevol=None, ddustevol=None, dustpred=None, iphase=None,
                         twas=None, ufloor=0.0):
        nfl = C.c_int32()
        self._check(lib().orc_step_predict_sph(
            self._ctx, xyzh.shape[0], timei, dtsph, _ptr(xyzh), _ptr(vxyzu), self._ld(vxyzu), _ptr(fxyzu),
            self._ld(fxyzu), _ptr(vpred), _ptr(Bevol), _ptr(dBevol), _ptr(Bpred), _ptr(dustevol), _ptr(ddustevol),
            _ptr(dustpred), _ptr(divcurlv), self._ld(divcurlv), _ptr(eos_vars), self._ld(eos_vars), _ptr(alphaind),
            _ptr(iphase), _ptr(twas), ufloor, C.byref(nfl)))
        return nfl.value

    def step_corrector(self, timei, dtsph, xyzh, vxyzu, fxyzu, vpred=None, Bevol=None, dBevol=None,
                       dustevol=None, ddustevol=None, iphase=None, ibin=None, ibin_wake=None, nbinmax=0,
                       twas=None, ufloor=0.0):
        em, v2 = C.c_double(), C.c_double()
        npc, nw, nfl = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(lib().orc_step_corrector(
            self._ctx, xyzh.shape[0], timei, dtsph, _ptr(xyzh), _ptr(vxyzu), self._ld(vxyzu), _ptr(fxyzu),
            self._ld(fxyzu), _ptr(vpred), _ptr(Bevol), _ptr(dBevol), _ptr(dustevol), _ptr(ddustevol), _ptr(iphase),
            _ptr(ibin), _ptr(ibin_wake), nbinmax, _ptr(twas), ufloor, C.byref(em), C.byref(v2), C.byref(npc),
            C.byref(nw), C.byref(nfl)))
        return dict(errmax=em.value, v2sum=v2.value, np=npc.value, nwake=nw.value, nvfloor=nfl.value)

    def step_reset(self, dtsph, vxyzu, fxyzu, vpred, Bevol=None, dBevol=None, Bpred=None, dustevol=None,
                   ddustevol=None, dustpred=None, iphase=None):
        self._check(lib().orc_step_reset(
            self._ctx, vxyzu.shape[0], dtsph, _ptr(vxyzu), self._ld(vxyzu), _ptr(fxyzu), self._ld(fxyzu),
            _ptr(vpred), _ptr(Bevol), _ptr(dBevol), _ptr(Bpred), _ptr(dustevol), _ptr(ddustevol), _ptr(dustpred),
            _ptr(iphase)))

    def get_quads(self):
        """Second moments quads(6) of every node (node ids 1..ncells)."""
        nc = self.get_tree()["ncells"]
        q = np.zeros((nc, 6))
        self._check(lib().orc_get_quads(self._ctx, _ptr(q), nc))
        return q


def run_all(setup, icall=1):
    """maketree + densityiterate + cons2prim + force on a copy of a setup."""
    p = setup["params"]
    o = Oracle(p)
    xyzh = setup["xyzh"].copy()
    vxyzu = setup["vxyzu"]
    alphaind = setup["alphaind"].copy()
    o.maketree(xyzh, iphase=setup.get("iphase"))
    tree = o.get_tree()
    d = o.densityiterate(xyzh, vxyzu, Bevol=setup.get("Bevol"), icall=icall, alphaind=alphaind,
                         iphase=setup.get("iphase"))
    c = o.cons2prim(xyzh, vxyzu, d["dvdx"], alphaind, dustevol=setup.get("dustevol"))
    bins = setup.get("bins")
    if bins is not None:
        ibin = bins["ibin"].copy()
        ibin_wake = bins["ibin_wake"].copy()
        o.set_bins(ibin, bins["ibin_old"], ibin_wake, bins["nbinmax"], bins["istepfrac"], bins["ibinnow"])
    f = o.force(xyzh, vxyzu, d["gradh"], alphaind, c["eos_vars"], divcurlB=d["divcurlB"],
                Bevol=setup.get("Bevol"), dustfrac=c["dustfrac"], iphase=setup.get("iphase"))
    if bins is not None:
        f.update(ibin=ibin, ibin_wake=ibin_wake, nbinmax=o.nbinmax())
    f["max_force_trial"] = o.max_force_trial()
    return dict(oracle=o, xyzh=xyzh, tree=tree, dens=d, c2p=c, force=f, stats=o.neighbour_stats(),
                alphaind=alphaind)
