"""Benchmark of the SPH hot path on one MI355X.

One step = densityiterate(icall=1) + cons2prim + force over every particle
of the workload, on arrays resident in HBM.  Before each step the smoothing
lengths are reset to their initial values (one device-to-device copy of
xyzh), so every step does the same work; the kd-tree is built outside the
timed region and its build time is reported separately, with the combined
tree + density + force figure.

The default workload is the 8,388,608-particle adiabatic hydro box (cubic
kernel, hfact 1.2) in FP64: `value`.  The same box then runs in mixed
precision (the ratio and the per-quantity errors of mixed against FP64),
and the 1,097,505-particle 3D shock tube runs in both precisions.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from phantom_amd import abi, setups, work_model  # noqa: E402

WORKLOADS = {
    # the configuration the headline metric is quoted on: 3D shock tube,
    # quintic kernel, Cullen-Dehnen viscosity + conductivity
    "shock_tube_1M": lambda: setups.shock_tube(ny=58),
    "box64_isothermal": lambda: setups.uniform_box(nx=64),
    "mhd_dust_128": lambda: setups.mhd_dust_box(nx=128, jitter=0.05),
    "hydro_box_8M": lambda: setups.uniform_box(nx=256, ny=256, nz=128, box=(-0.5, 0.5, -0.5, 0.5, -0.25, 0.25),
                                                isothermal=False, seed=8),
    # self-gravitating n=1.5 polytrope, multipole walk at theta = 0.5
    "polytrope_1M": lambda: setups.polytrope(n=1048576, theta=0.5),
    "polytrope_8M": lambda: setups.polytrope(n=8388608, theta=0.5),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="hydro_box_8M", choices=sorted(WORKLOADS))
    ap.add_argument("--quick", action="store_true",
                    help="the named workload and precision only (no mixed/FP64 pair, no 1M shock tube)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--mode", default="derivs", choices=["derivs", "step", "substep"],
                    help="derivs: densityiterate + cons2prim + force per step (the headline metric); "
                         "step: whole leapfrog steps (tree, density, force, integrator loops)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a 256 MiB device copy (reference for the HBM byte counters)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.calibrate_copy:
        return calibrate_copy()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is a design note, "
                                   "not an executable path", "n_gpus": args.gpus}))
        return 2
    if args.mode == "step":
        return step_bench(args)
    if args.mode == "substep":
        print(json.dumps(substep_bench()), flush=True)
        return 0
    main_run = run_derivs(args.workload, args.precision, args.steps, args.warmup, keep=not args.quick)
    p = main_run["params"]
    n = main_run["n"]
    kt = main_run["kt"]

    def kmean(name):
        return float(np.mean(kt[name])) * 1e-3 if name in kt else 0.0

    # algorithmic flops per launch of each pair kernel (work_model); the
    # list builders and walks do search work only and are not credited
    flops = {
        "density_iter": work_model.density_iter_flops(main_run["nneigh"], main_run["nits"]),
        "density_sums": work_model.density_sums_flops(main_run["nneigh"], mhd=bool(p.mhd)),
        "force": work_model.force_flops(main_run["npairs"], mhd=bool(p.mhd), dust=bool(p.use_dustfrac)),
        "force_list": 0.0, "density_list": 0.0,
    }
    step_kernels = [nm for nm in kt if nm not in ("force_prep", "dens_gather")]
    dominant = max(step_kernels, key=lambda nm: kmean(nm) * len(kt[nm]))
    dens_prec, force_prec = main_run["precision_used"]
    kprec = force_prec if dominant.startswith("force") else dens_prec
    peak = work_model.PEAK_FP64_TFLOPS if kprec == abi.PREC_FP64 else work_model.PEAK_FP32_TFLOPS
    t_dom = kmean(dominant)
    ach = flops.get(dominant, 0.0) / t_dom / 1e12 if t_dom > 0 else 0.0
    roofline = {"bound": "valu", "achieved": round(ach, 4), "peak": peak, "unit": "TFLOP/s",
                "frac": round(ach / peak, 6), "traffic": measured_traffic(dominant, args.workload),
                "kernel": dominant, "launch_ms": round(1e3 * t_dom, 4),
                "note": "vector ALU pair arithmetic (no dense contraction, no MFMA); peak is MI355X vector "
                        + ("FP64" if peak == work_model.PEAK_FP64_TFLOPS else "FP32")
                        + "; algorithmic flops per work_model (interacting pairs only: rejected trials and the "
                          "neighbour search are not credited); launch time from HIP events on the library stream"}
    kernel_tflops = {nm: round(flops[nm] / kmean(nm) / 1e12, 4) for nm in flops if nm in kt and kmean(nm) > 0}

    extra = {}
    if not args.quick:
        # the same workload in the other precision: speed ratio and the error
        # of mixed precision against the FP64 path (which equals the oracle
        # to 1e-12 in the parity tests)
        other = "mixed" if args.precision == "fp64" else "fp64"
        o_run = run_derivs(args.workload, other, args.steps, args.warmup, keep=True)
        fp, mx = (main_run, o_run) if args.precision == "fp64" else (o_run, main_run)
        extra["mixed"] = precision_pair(fp, mx)
        # the 1M shock tube, the configuration the metric's 1M half is quoted on
        if args.workload != "shock_tube_1M":
            t64 = run_derivs("shock_tube_1M", "fp64", args.steps, args.warmup, keep=True)
            t32 = run_derivs("shock_tube_1M", "mixed", args.steps, args.warmup, keep=True)
            extra["shock_tube_1M"] = dict(n_particles=t64["n"], fp64=summary(t64), **precision_pair(t64, t32))
        extra["substep"] = substep_bench()
    tree_ms = main_run["tree_ms"]
    combined_ms = tree_ms + main_run["ms_step"]

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(main_run["setup"], args)

    out = {
        "metric": "density+force particle updates/s",
        "value": main_run["value"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_run["ms_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-pair/f64-accum",
        "data": "synthetic",
        "config": {"workload": args.workload, "n_particles": n, "kernel": "cubic" if p.kernel == 0 else "quintic",
                   "precision": args.precision, "precision_used": {"density": dens_prec, "force": force_prec},
                   "mhd": bool(p.mhd), "dust": bool(p.use_dustfrac), "start_h": main_run["start_h"],
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "tree": {"build_ms": round(tree_ms, 4), "first_call_ms": round(main_run["tree_first_ms"], 4),
                 "kernels_ms": {nm: round(ms, 4) for nm, ms in main_run["tree_kernels"].items()},
                 "combined_ms_per_step": round(combined_ms, 4),
                 "combined_updates_per_s": n / (combined_ms * 1e-3)},
        "breakdown_ms": dict(main_run["breakdown"], kernels={nm: round(1e3 * kmean(nm), 4) for nm in sorted(kt)}),
        "kernel_tflops": kernel_tflops,
        "tuning": main_run["tuning"],
        "neighbours": {"dens_mean": main_run["stats"]["actualmean"], "dens_max": main_run["stats"]["maxactual"],
                       "force_pairs_per_particle": main_run["npairs"] / n, "nrhocalc": main_run["stats"]["nrhocalc"]},
    }
    out.update(extra)
    print(json.dumps(out), flush=True)
    return 0


def summary(r):
    return {"value": r["value"], "ms_per_step": r["ms_step"], "tree_build_ms": round(r["tree_ms"], 4),
            "precision_used": {"density": r["precision_used"][0], "force": r["precision_used"][1]}}


def precision_pair(fp, mx):
    """Mixed against FP64 on the same particles: updates/s of each, their
    ratio and the error of mixed precision in rho (from h through
    rho = m (hfact/h)^3) and in dv/dt.  rho: max over particles of the
    relative error.  dv/dt: max over particles and components of
    |a_mixed - a_fp64| / rms|a_fp64| (per-particle relative errors are not
    meaningful where the acceleration passes through zero, as in the
    transverse components of the shock tube)."""
    out = {"mixed": summary(mx), "fp64_value": fp["value"], "mixed_value": mx["value"],
           "mixed_over_fp64": mx["value"] / fp["value"]}
    if fp.get("h") is not None and mx.get("h") is not None:
        h64, h32 = fp["h"], mx["h"]
        rho64, rho32 = (1.0 / h64) ** 3, (1.0 / h32) ** 3
        a64, a32 = fp["f"][:, :3], mx["f"][:, :3]
        rms = float(np.sqrt(np.mean(np.sum(a64 * a64, axis=1))))
        out["max_rel_err_rho"] = float(np.max(np.abs(rho32 - rho64) / rho64))
        out["max_err_dvdt_over_rms"] = float(np.max(np.abs(a32 - a64)) / rms) if rms > 0 else 0.0
        if fp["f"].shape[1] >= 4:
            u64, u32 = fp["f"][:, 3], mx["f"][:, 3]
            urms = float(np.sqrt(np.mean(u64 * u64)))
            out["max_err_dudt_over_rms"] = float(np.max(np.abs(u32 - u64)) / urms) if urms > 0 else 0.0
        out["nneigh_identical"] = bool(np.array_equal(fp["nneigh"], mx["nneigh"]))
    return out


def run_derivs(workload, precision, steps, warmup, keep=False):
    """Time `steps` density + cons2prim + force steps of one workload in one
    precision on HBM-resident arrays (after `warmup` untimed steps).  Every
    step starts from the same positions and h (one device copy before it).
    The tree is built outside the timed region: its time is that of the
    second build on the context (buffers allocated, as in a run)."""
    from phantom_amd import device, sph

    setup = WORKLOADS[workload]()
    p = setup["params"]
    p.precision = abi.PREC_MIXED if precision == "mixed" else abi.PREC_FP64
    n = setup["xyzh"].shape[0]
    log(f"workload {workload} n={n} precision={precision}")
    D = device.DeviceArray
    xyzh0 = D.from_numpy(setup["xyzh"])
    xyzh = D.from_numpy(setup["xyzh"])
    vxyzu = D.from_numpy(setup["vxyzu"])
    alphaind = D.from_numpy(setup["alphaind"])
    Bevol = D.from_numpy(setup["Bevol"]) if "Bevol" in setup else None
    dustevol = D.from_numpy(setup["dustevol"]) if "dustevol" in setup else None
    divcurlv = D((n, 1), np.float32)
    divcurlB = D((n, 4), np.float32) if p.mhd else None
    gradh = D((n, 2 if p.gravity else 1), np.float32)
    dvdx = D((n, 9), np.float32)
    eos = D((n, 3), np.float64)
    dustfrac = D((n,), np.float64) if p.use_dustfrac else None
    fxyzu = D((n, 4), np.float64)
    dBevol = D((n, 4), np.float64) if p.mhd else None
    divBsymm = D((n,), np.float32) if p.mhd else None
    ddustevol = D((n,), np.float64) if p.use_dustfrac else None
    deltav = D((n, 3), np.float64) if p.use_dustfrac else None
    tstop = D((n,), np.float64) if p.use_dustfrac else None

    ctx = sph.Context(p)
    t0 = time.perf_counter()
    ctx.maketree(xyzh, leaf=False)
    tree_first_ms = 1e3 * (time.perf_counter() - t0)
    start_h = "setup"
    if p.gravity:
        # the polytrope's analytic density falls to zero at the surface, so
        # its h there is far from what the density iteration converges to; a
        # running simulation starts each step from the h the last step
        # converged to, so one untimed density call provides the start
        ctx.densityiterate(1, xyzh, vxyzu, Bevol=Bevol, divcurlv=divcurlv, divcurlB=divcurlB, alphaind=alphaind,
                           gradh=gradh, dvdx=dvdx)
        setup["xyzh"] = xyzh.numpy()
        start_h = "converged"
    xyzh0.copy_from(xyzh)  # wrapped positions, h at the start of every step
    # further builds on the same context: buffers exist, as in a time step
    tms = []
    tree_kernels = {}
    for _ in range(3):
        device.synchronize()
        t0 = time.perf_counter()
        ctx.maketree(xyzh, leaf=False)
        tms.append(1e3 * (time.perf_counter() - t0))
        tree_kernels = {}
        for nm, ms in ctx.timings():
            tree_kernels[nm] = tree_kernels.get(nm, 0.0) + ms
    tree_ms = float(np.median(tms))
    log(f"tree built in {tree_first_ms:.2f} ms (first call), {tree_ms:.2f} ms (median of 3 later builds)")

    kt = {}  # per-launch kernel times (HIP events on the library's stream), by name
    phase = {"density": 0.0, "cons2prim": 0.0, "force": 0.0}

    def record_timings():
        for nm, ms in ctx.timings():
            kt.setdefault(nm, []).append(ms)

    def step(record):
        xyzh.copy_from(xyzh0)
        a = time.perf_counter()
        ctx.densityiterate(1, xyzh, vxyzu, Bevol=Bevol, divcurlv=divcurlv, divcurlB=divcurlB, alphaind=alphaind,
                           gradh=gradh, dvdx=dvdx)
        if record:
            record_timings()
        b = time.perf_counter()
        ctx.cons2prim(xyzh, vxyzu, dvdx, alphaind, dustevol=dustevol, dustfrac=dustfrac, eos_vars=eos)
        c = time.perf_counter()
        ctx.force(1, xyzh, vxyzu, gradh, alphaind, eos, divcurlB=divcurlB, Bevol=Bevol, dustfrac=dustfrac,
                  fxyzu=fxyzu, divcurlv=divcurlv, dBevol=dBevol, divBsymm=divBsymm, ddustevol=ddustevol,
                  deltav=deltav, tstop=tstop)
        if record:
            record_timings()
        d = time.perf_counter()
        if record:
            phase["density"] += b - a
            phase["cons2prim"] += c - b
            phase["force"] += d - c

    for k in range(warmup):
        step(False)
        log(f"warmup step {k + 1}/{warmup} done")
    device.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step(True)
    device.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"timed {steps} steps in {elapsed:.3f} s ({n * steps / elapsed / 1e6:.1f} M updates/s)")
    t_dens = sum(float(np.sum(kt.get(nm, []))) for nm in ("density_iter", "density_sums")) * 1e-3 / steps
    t_force = sum(float(np.sum(kt.get(nm, []))) for nm in ("force_list", "force")) * 1e-3 / steps
    out = dict(
        n=n, params=p, setup=setup, start_h=start_h, kt=kt, value=n * steps / elapsed,
        ms_step=1e3 * elapsed / steps, tree_ms=tree_ms, tree_first_ms=tree_first_ms, tree_kernels=tree_kernels,
        stats=ctx.get_neighbour_stats(), npairs=ctx.get_force_stats()["npairs"],
        precision_used=ctx.get_precision(), tuning=ctx.get_tuning(),
        breakdown={"density_call": 1e3 * phase["density"] / steps,
                   "cons2prim_call": 1e3 * phase["cons2prim"] / steps,
                   "force_call": 1e3 * phase["force"] / steps,
                   "density_kernels": 1e3 * t_dens, "force_kernels": 1e3 * t_force})
    out["nneigh"], out["nits"] = ctx.get_particle_stats(n)
    if keep:
        out["h"] = xyzh.numpy()[:, 3].copy()
        out["f"] = fxyzu.numpy()
    ctx.close()
    for a in (xyzh0, xyzh, vxyzu, alphaind, Bevol, dustevol, divcurlv, divcurlB, gradh, dvdx, eos, dustfrac, fxyzu,
              dBevol, divBsymm, ddustevol, deltav, tstop):
        if a is not None:
            a.free()
    return out


class _Recorder:
    """Forwards the calls of a library context and keeps the kernel times
    (HIP events on the library stream) of every call."""

    def __init__(self, ctx):
        self._ctx = ctx
        self.kt = {}
        self.on = False

    def __getattr__(self, name):
        f = getattr(self._ctx, name)
        if not callable(f) or name.startswith("_") or name == "timings":
            return f

        def call(*a, **k):
            r = f(*a, **k)
            if self.on:
                for nm, ms in self._ctx.timings():
                    self.kt.setdefault(nm, []).append(ms)
            return r

        return call


def step_bench(args):
    """Whole leapfrog steps (step_leapfrog.f90:91-760 through
    phantom_amd.integrator) with the particle state resident in HBM: per
    step one tree build, densityiterate, cons2prim, force and the predictor,
    drift, predict_sph and corrector loops.  dt is the global Courant/force
    timestep.  Reports updates/s of whole steps, the per-kernel times and
    the HBM roofline of the integrator loops."""
    from phantom_amd import device, integrator, sph

    setup = WORKLOADS[args.workload]()
    p = setup["params"]
    p.precision = abi.PREC_MIXED if args.precision == "mixed" else abi.PREC_FP64
    n = setup["xyzh"].shape[0]
    log(f"workload {args.workload} n={n} precision={args.precision} mode=step")
    s = integrator.State(p, setup["xyzh"], setup["vxyzu"], setup["alphaind"], Bevol=setup.get("Bevol"),
                         dustevol=setup.get("dustevol"), device=True)
    ctx = sph.Context(p)
    rec = _Recorder(ctx)
    integrator.init_global(rec, s)
    t = 0.0

    def one():
        nonlocal t
        dt = min(s.dtforce, s.dtcourant, p.dtmax)
        r = integrator.step(rec, s, t, dt)
        t = t + dt
        return r

    for k in range(args.warmup):
        one()
        log(f"warmup step {k + 1}/{args.warmup} done")
    device.synchronize()
    rec.on = True
    its = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        its.append(one()["its"])
    device.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"timed {args.steps} steps in {elapsed:.3f} s")
    kt = rec.kt

    def kmean(name):
        return float(np.mean(kt[name])) * 1e-3 if name in kt else 0.0

    def ktot(name):
        return float(np.sum(kt.get(name, []))) * 1e-3 / args.steps

    loops = {}
    for nm in ("step_predictor", "substep_sph", "step_predict_sph", "step_corrector"):
        if nm not in kt:
            continue
        b = n * work_model.step_loop_bytes(nm, nv=p.maxvxyzu, mhd=bool(p.mhd),
                                           dust=bool(p.use_dustfrac and p.ndustsmall), ind=bool(p.ind_timesteps))
        gbs = b / kmean(nm) / 1e9
        loops[nm] = {"bound": "hbm", "bytes": b, "launch_ms": round(1e3 * kmean(nm), 4),
                     "achieved": round(gbs, 1), "peak": work_model.PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(gbs / work_model.PEAK_HBM_GBS, 4)}
    out = {
        "metric": "leapfrog step particle updates/s",
        "value": n * args.steps / elapsed,
        "unit": "particle-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-pair/f64-accum",
        "data": "synthetic",
        "config": {"workload": args.workload, "n_particles": n, "mode": "step", "precision": args.precision,
                   "parallelism": "single GPU"},
        "corrector_iterations": its,
        "kernels_ms_per_step": {nm: round(1e3 * ktot(nm), 4) for nm in sorted(kt)},
        "integrator_loops": loops,
    }
    print(json.dumps(out), flush=True)
    return 0


def substep_bench(nx=128, reps=3):
    """Time of one block-timestep substep (evol_prestep, step, time update:
    evolve.f90:200-257) on the 128^3 MHD + one-fluid dust box with four
    populated bins (nbinmax 6), at the substeps whose active fractions are
    1, 1/2, 1/8 and 1/64 (istepfrac 64, 32, 8, 1).  Three drivers of the
    same substep, state resident in HBM: the host loop over the individual
    entry points (integrator.evolve_ind), the library's own loop with
    direct launches, and the library's loop with the fixed loop chains
    replayed as HIP graphs (phsph_step_graph).  Each time is the median of
    `reps` runs from the same saved state, after one untimed run."""
    from phantom_amd import device, integrator, sph

    setup = setups.with_levels(setups.mhd_dust_box(nx=nx, jitter=0.05), setups.LEVELS_4, nbinmax=6, istepfrac=0)
    p = setup["params"]
    n = setup["xyzh"].shape[0]
    log(f"substep timing: mhd_dust_box_{nx} n={n}, bins {setups.LEVELS_4}")
    s = integrator.State(p, setup["xyzh"], setup["vxyzu"], setup["alphaind"], Bevol=setup.get("Bevol"),
                         dustevol=setup.get("dustevol"))
    s.enable_bins(0)
    ctx = sph.Context(p)
    integrator.init_ind(ctx, s)
    b = setup["bins"]["ibin_old"]
    s.ibin[:] = b
    s.ibin_old[:] = b
    s.ibin_wake[:] = 0
    s.iphase[:] = 1
    s.nbinmax = 6
    s.twas[:] = 0.5 * (p.dtmax / 2.0 ** b.astype(np.float64))
    save = s.copy(device=True)
    work = save.copy()
    del s

    def restore():
        for k in integrator.State.ARRAYS:
            a = getattr(save, k)
            if a is not None:
                getattr(work, k).copy_from(a)
        for k in integrator.State.SCALARS:
            setattr(work, k, getattr(save, k))

    def host(t):
        return integrator.evolve_ind(ctx, work, 1, time=t, tlast=0.0)[2][0]

    def library(t):
        return integrator.evolve_graph(ctx, work, 1, time=t, tlast=0.0, graphs=False)[3][0]

    def graphs(t):
        return integrator.evolve_graph(ctx, work, 1, time=t, tlast=0.0, graphs=True)[3][0]

    res = {}
    for m, frac in ((64, "1"), (32, "1/2"), (8, "1/8"), (1, "1/64")):
        row = {}
        for name, fn in (("host", host), ("library", library), ("graphs", graphs)):
            ts = []
            for r in range(reps + 1):
                restore()
                work.istepfrac = m - 1
                t = (m - 1) / 64.0 * p.dtmax
                device.synchronize()
                t0 = time.perf_counter()
                o = fn(t)
                device.synchronize()
                ts.append(time.perf_counter() - t0)
            row[f"{name}_ms"] = round(1e3 * float(np.median(ts[1:])), 3)
            row["nactive"] = o["nactive"]
            row["its"] = o["its"]
            if name == "graphs":
                # kernel times of the last run (HIP events; the graph-replayed
                # loop chains are not among them)
                kt = {}
                for nm, ms in ctx.timings():
                    kt[nm] = kt.get(nm, 0.0) + ms
                row["kernels_ms"] = {nm: round(v, 3) for nm, v in sorted(kt.items(), key=lambda x: -x[1])[:8]}
        row["active_fraction"] = row["nactive"] / n
        log(f"substep at active fraction {frac}: {row}")
        res[frac] = row
    ctx.close()
    return {"workload": f"mhd_dust_box_{nx}", "n_particles": n, "nbinmax": 6, "levels": setups.LEVELS_4,
            "fractions": res,
            "note": "ms per substep (set_active_particles, predictor, drift, predict_sph, tree, density, "
                    "cons2prim, force, corrector); state resident in HBM"}


def measured_traffic(kernel, workload):
    """HBM bytes per launch of a kernel from the committed rocprofv3 FETCH_SIZE
    and WRITE_SIZE runs (profiles/traffic.json), each scaled by what the
    counter reads for a device copy of known size; None when not measured."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    ent = rec.get(workload, {}).get(kernel)
    return None if ent is None else ent.get("bytes")


def calibrate_copy():
    """A plain device-to-device copy of 256 MiB, three times: the reference
    point that scales FETCH_SIZE / WRITE_SIZE to bytes."""
    from phantom_amd import device
    nbytes = 256 * 1024 * 1024
    a = device.DeviceArray((nbytes // 8,), np.float64)
    b = device.DeviceArray((nbytes // 8,), np.float64)
    for _ in range(3):
        b.copy_from(a)
    device.synchronize()
    print(json.dumps({"copy_bytes_each_way": nbytes, "copies": 3}), flush=True)
    return 0


def host_cores():
    """CPUs this process may run on (sched_getaffinity) and the machine's
    physical cores and sockets as lscpu reports them."""
    import subprocess
    info = {"affinity_cpus": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = {}
        for line in txt.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        cps, sk = int(kv.get("Core(s) per socket", "0")), int(kv.get("Socket(s)", "0"))
        info.update(model=kv.get("Model name"), sockets=sk, cores_per_socket=cps, physical_cores=cps * sk,
                    threads_per_core=int(kv.get("Thread(s) per core", "0")))
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(setup, args):
    """The oracle (a C restatement of the reference algorithm, OpenMP over the
    host cores) on the same particles: maketree once, then density +
    cons2prim + force, one warm-up and the median of three calls.  Threads:
    the process's OpenMP thread count (OMP_NUM_THREADS: the CPU share of one
    GPU on a shared host), else every CPU of its affinity mask."""
    from oracle import oracle

    cores = host_cores()
    env_threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = args.cpu_threads or env_threads or cores["affinity_cpus"]
    threads = min(threads, cores["affinity_cpus"])
    os.environ["OMP_NUM_THREADS"] = str(threads)
    p = setup["params"]
    p.precision = abi.PREC_FP64
    o = oracle.Oracle(p)
    xyzh0 = setup["xyzh"].copy()
    t0 = time.perf_counter()
    o.maketree(xyzh0)
    tree_s = time.perf_counter() - t0
    n = xyzh0.shape[0]
    times = []
    for rep in range(4):
        xyzh = xyzh0.copy()
        alphaind = setup["alphaind"].copy()
        t0 = time.perf_counter()
        d = o.densityiterate(xyzh, setup["vxyzu"], Bevol=setup.get("Bevol"), alphaind=alphaind)
        c = o.cons2prim(xyzh, setup["vxyzu"], d["dvdx"], alphaind, dustevol=setup.get("dustevol"))
        o.force(xyzh, setup["vxyzu"], d["gradh"], alphaind, c["eos_vars"], divcurlB=d["divcurlB"],
                Bevol=setup.get("Bevol"), dustfrac=c["dustfrac"])
        if rep > 0:
            times.append(time.perf_counter() - t0)
        log(f"cpu baseline call {rep + 1}/4: {time.perf_counter() - t0:.3f} s")
    dt = float(np.median(times))
    log(f"cpu baseline: median {dt:.3f} s per call on {threads} threads")
    return {"value": n / dt, "unit": "particle-updates/s", "cores": threads, "kind": "port",
            "sample": f"median of 3 density+cons2prim+force calls (after one warm-up) over all {n} particles "
                      f"of the same workload, OpenMP with {threads} threads; maketree (serial) {tree_s:.2f} s",
            "host": dict(cores, omp_num_threads_env=env_threads)}


if __name__ == "__main__":
    sys.exit(main())
