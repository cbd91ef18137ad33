"""ctypes mirror of the C-ABI types declared in include/phsph.h.

The field order, widths and defaults follow ``phsph_params`` exactly; each
default is the reference default it carries (options.f90:69-128,
shock_capturing.f90:47-65, timestep.f90:52-62, eos.f90:1888-1905,
dust.f90:26-33, kdtree.F90:46).
"""
import ctypes as C
import math

KERNEL_CUBIC = 0
KERNEL_QUINTIC = 1
PREC_FP64 = 0
PREC_MIXED = 1

OK = 0
ERR_INPUT = -1
ERR_NAN = -2
ERR_NOCONVERGE = -3
ERR_CAPACITY = -4
ERR_HIP = -5
ERR_STATE = -6
ERR_NODATA = -7

ERROR_NAMES = {
    ERR_INPUT: "PHSPH_ERR_INPUT",
    ERR_NAN: "PHSPH_ERR_NAN",
    ERR_NOCONVERGE: "PHSPH_ERR_NOCONVERGE",
    ERR_CAPACITY: "PHSPH_ERR_CAPACITY",
    ERR_HIP: "PHSPH_ERR_HIP",
    ERR_STATE: "PHSPH_ERR_STATE",
    ERR_NODATA: "PHSPH_ERR_NODATA",
}

_INT_FIELDS = [
    "kernel", "periodic", "maxvxyzu", "mhd", "use_dustfrac", "ndustsmall",
    "gravity", "ind_timesteps", "ieos", "maxp", "precision", "idrag",
    "ishock_heating", "ipdv_heating", "iresistive_heating", "ilimitdustflux",
]
_DBL_FIELDS = [
    "xmin", "xmax", "ymin", "ymax", "zmin", "zmax", "hfact", "tolh",
    "massoftype_gas", "gamma", "polyk", "alpha", "alphamax", "alphau",
    "alphaB", "beta", "C_cour", "C_force", "dtmax", "psidecayfac",
    "overcleanfac", "K_code", "tree_accuracy", "grainsize", "graindens", "seff",
]


class Params(C.Structure):
    """phsph_params (include/phsph.h)."""
    _fields_ = [(n, C.c_int32) for n in _INT_FIELDS] + [(n, C.c_double) for n in _DBL_FIELDS]

    def __repr__(self):
        return "Params(" + ", ".join(f"{n}={getattr(self, n)}" for n, _ in self._fields_) + ")"


class KdNode(C.Structure):
    """type(kdnode) of a non-GRAVITY build (dtype_kdtree.F90:53-68)."""
    _fields_ = [
        ("xcen", C.c_double * 3), ("size", C.c_double), ("hmax", C.c_double),
        ("mass", C.c_double), ("leftchild", C.c_int32), ("rightchild", C.c_int32),
        ("parent", C.c_int32), ("idum", C.c_int32),
    ]


assert C.sizeof(KdNode) == 64


class KdNodeGrav(C.Structure):
    """type(kdnode) of a GRAVITY build (dtype_kdtree.F90:53-68)."""
    _fields_ = KdNode._fields_ + [("tobecached", C.c_int32), ("cached", C.c_int32),
                                  ("quads", C.c_double * 6)]


assert C.sizeof(KdNodeGrav) == 120


class StepState(C.Structure):
    """phsph_step_state: device pointers of the arrays one step touches."""
    _fields_ = [(k, C.c_void_p) for k in (
        "xyzh", "vxyzu", "fxyzu", "vpred", "Bevol", "dBevol", "Bpred", "dustevol", "ddustevol", "dustpred",
        "dustfrac", "deltav", "tstop", "divcurlv", "divcurlB", "divBsymm", "alphaind", "gradh", "dvdx", "eos_vars",
        "poten", "iphase", "ibin", "ibin_old", "ibin_wake", "twas")] + \
        [(k, C.c_int32) for k in ("np", "ldxyzh", "ldvxyzu", "ldfxyzu", "lddivcurlv", "ldgradh", "ldeos", "pad")]


class StepClock(C.Structure):
    """phsph_step_clock: evolve loop time and bin bookkeeping."""
    _fields_ = [("time", C.c_double), ("tlast", C.c_double), ("dt", C.c_double), ("dtcourant", C.c_double),
                ("dtforce", C.c_double), ("nbinmax", C.c_int32), ("istepfrac", C.c_int32), ("ibinnow", C.c_int32),
                ("pad", C.c_int32)]


class StepSummary(C.Structure):
    """phsph_step_summary: one step's dt, dtnew, dterr and counters."""
    _fields_ = [("dt", C.c_double), ("dtnew", C.c_double), ("dterr", C.c_double), ("its", C.c_int32),
                ("nactive", C.c_int32), ("nwake", C.c_int32), ("nvfloorp", C.c_int32), ("nvfloorps", C.c_int32),
                ("nvfloorc", C.c_int32), ("ngraph", C.c_int32), ("ninst", C.c_int32)]


assert C.sizeof(StepState) == 26 * 8 + 8 * 4
assert C.sizeof(StepClock) == 5 * 8 + 4 * 4
assert C.sizeof(StepSummary) == 3 * 8 + 8 * 4

PI_PH = 3.1415926536            # physcon.f90:23
MASS_PROTON_CGS = 1.67262158e-24  # physcon.f90:41
CROSS_SECTION_H2_CGS = 2.367e-15  # physcon.f90:43


def seff_code_units(umass, udist):
    """Effective surface density of the gas mean free path, lambda =
    seff/rho_gas, in code units (init_drag, dust.f90:94-97)."""
    cste_seff = PI_PH / math.sqrt(2.0) * 5.0 / 64.0
    mass_mol_gas = (2.0 * MASS_PROTON_CGS) / umass
    cross_section_gas = CROSS_SECTION_H2_CGS / (udist * udist)
    return cste_seff * mass_mol_gas / cross_section_gas


def default_params(kernel=KERNEL_CUBIC, **overrides):
    """Reference defaults; ``hfact`` is the kernel's hfact_default
    (kernel_cubic.f90:29, kernel_quintic.f90:29)."""
    p = Params()
    p.kernel = kernel
    p.periodic = 1
    p.maxvxyzu = 4
    p.mhd = 0
    p.use_dustfrac = 0
    p.ndustsmall = 0
    p.gravity = 0
    p.ind_timesteps = 0
    p.ieos = 2
    p.maxp = 0
    p.precision = PREC_FP64
    p.idrag = 2
    p.ishock_heating = 1
    p.ipdv_heating = 1
    p.iresistive_heating = 1
    p.ilimitdustflux = 0
    p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax = -0.5, 0.5, -0.5, 0.5, -0.5, 0.5
    p.hfact = 1.2 if kernel == KERNEL_CUBIC else 1.0
    p.tolh = 1.e-4
    p.massoftype_gas = 1.0
    p.gamma = 5.0 / 3.0
    p.polyk = 1.0
    p.alpha = 0.0
    p.alphamax = 1.0
    p.alphau = 1.0
    p.alphaB = 1.0
    p.beta = 2.0
    p.C_cour = 0.3
    p.C_force = 0.25
    p.dtmax = 1.e29
    p.psidecayfac = 1.0
    p.overcleanfac = 1.0
    p.K_code = 1.0
    p.tree_accuracy = 0.5
    p.grainsize = 0.1   # grainsizecgs (dust.f90:38) in units with udist = 1 cm
    p.graindens = 3.0   # graindenscgs (dust.f90:39) in units with unit_density = 1 g/cm^3
    p.seff = seff_code_units(1.0, 1.0)
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def params_to_dict(p):
    return {n: getattr(p, n) for n, _ in p._fields_}
