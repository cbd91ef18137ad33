// substep.hip -- whole steps on device-resident state: the reference's
// evolve loop restricted to the SPH path, with its control flow in the
// library instead of the caller.
//
//   block timesteps   evol_prestep (evolve.f90:200-222): istepfrac += 1,
//                     set_active_particles (utils_indtimesteps.f90:114-185);
//                     step at dt = dtmax/2**nbinmax;
//                     update_time_and_dt: time = tlast + istepfrac/2**nbinmaxprev
//                     * dtmax, change_nbinmax (utils_indtimesteps.f90:186-223),
//                     istepfrac restarts at the end of each dtmax
//                     (evolve.f90:434-440)
//   global timestep   step at dt; dt = min(dtforce, dtcourant, dterr,
//                     dtmax + epsilon) (evolve.f90:105-160)
//   step              predictor, substep_sph drift, predict_sph, derivs(1),
//                     then the corrector and check_velocity_error until the
//                     velocity error converges, with the reset and derivs(2)
//                     between iterations (step_leapfrog.f90:91-843)
//   derivs            maketree, densityiterate, cons2prim, force
//                     (deriv.f90:114-189)
//
// The loops whose launch sequence is fixed -- predictor, drift and
// predict_sph with their reductions, and the corrector with its reduction --
// are captured from the library stream into two HIP graphs.  Each step
// captures them again with its own times and folds the capture into the
// instantiated graph with hipGraphExecUpdate (topology unchanged), so a
// step costs two graph launches for these loops.  The tree build, the
// density iteration and the force pass keep their data-dependent control
// (retries, list capacities, overflow re-runs) and run as the individual
// entry points run them.  Both launch forms give bit-identical states.
#include "ctx.h"

#include <cfloat>
#include <cmath>

namespace {

constexpr double BIGNUMBER = 1.e29;  // timestep.f90:40

// Captures the launches of `enqueue` on the library stream into graph slot
// `slot` and launches it: the first capture is instantiated, later ones
// update the instantiated graph in place (re-instantiated if the update is
// refused).  With use_graph false the launches are made directly.
template <class F>
int run_chain(phsph_ctx *c, bool use_graph, int slot, F &&enqueue) {
  hipStream_t s = c->stream;
  if (!use_graph) {
    enqueue();
    PH_HIPCHECK(c, hipGetLastError());
    return PHSPH_OK;
  }
  PH_HIPCHECK(c, hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
  enqueue();
  const hipError_t le = hipGetLastError();
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(s, &g);
  if (le != hipSuccess || ce != hipSuccess) {
    if (g) hipGraphDestroy(g);
    PH_HIPCHECK(c, le);
    PH_HIPCHECK(c, ce);
  }
  hipGraphExec_t &x = c->step_graph[slot];
  if (x) {
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult r = hipGraphExecUpdateSuccess;
    if (hipGraphExecUpdate(x, g, &bad, &r) != hipSuccess || r != hipGraphExecUpdateSuccess) {
      (void)hipGetLastError();
      hipGraphExecDestroy(x);
      x = nullptr;
    } else {
      c->step_graph_upd++;
    }
  }
  if (!x) {
    const hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      x = nullptr;
      hipGraphDestroy(g);
      PH_HIPCHECK(c, ie);
    }
    c->step_graph_inst++;
  }
  hipGraphDestroy(g);
  PH_HIPCHECK(c, hipGraphLaunch(x, s));
  c->step_graph_launch++;
  return PHSPH_OK;
}

bool dust_on(const phsph_ctx *c) { return c->p.use_dustfrac && c->p.ndustsmall > 0; }

// the argument sets of the individual loop entry points (api.hip), so that
// each kernel sees exactly what it sees there
StepArgs args_predictor(const phsph_ctx *c, const phsph_step_state &S) {
  const bool mhd = c->p.mhd, dust = dust_on(c), ind = c->p.ind_timesteps;
  StepArgs a;
  memset(&a, 0, sizeof(a));
  a.xyzh = S.xyzh; a.ldx = S.ldxyzh;
  a.vxyzu = S.vxyzu; a.ldv = S.ldvxyzu;
  a.fxyzu = S.fxyzu; a.ldf = S.ldfxyzu;
  a.Bevol = mhd ? S.Bevol : nullptr;
  a.dBevol = mhd ? S.dBevol : nullptr;
  a.dustevol = dust ? S.dustevol : nullptr;
  a.ddustevol = dust ? S.ddustevol : nullptr;
  a.iphase = S.iphase;
  if (ind) {
    a.ibin_in = S.ibin;
    a.ibin_old = S.ibin_old;
    a.twas = S.twas;
  }
  return a;
}

StepArgs args_predict_sph(const phsph_ctx *c, const phsph_step_state &S) {
  const bool mhd = c->p.mhd, dust = dust_on(c), ind = c->p.ind_timesteps;
  StepArgs a;
  memset(&a, 0, sizeof(a));
  a.xyzh = S.xyzh; a.ldx = S.ldxyzh;
  a.vxyzu = S.vxyzu; a.ldv = S.ldvxyzu;
  a.fxyzu = S.fxyzu; a.ldf = S.ldfxyzu;
  a.vpred = S.vpred;
  a.Bevol = mhd ? S.Bevol : nullptr;
  a.dBevol = mhd ? S.dBevol : nullptr;
  a.Bpred = mhd ? S.Bpred : nullptr;
  a.dustevol = dust ? S.dustevol : nullptr;
  a.ddustevol = dust ? S.ddustevol : nullptr;
  a.dustpred = dust ? S.dustpred : nullptr;
  a.divcurlv = S.divcurlv; a.lddcv = S.lddivcurlv;
  a.eos_vars = S.eos_vars; a.ldeos = S.ldeos;
  a.alphaind = S.alphaind;
  a.iphase = S.iphase;
  if (ind) a.twas = S.twas;
  return a;
}

StepArgs args_corrector(const phsph_ctx *c, const phsph_step_state &S) {
  const bool mhd = c->p.mhd, dust = dust_on(c), ind = c->p.ind_timesteps;
  StepArgs a;
  memset(&a, 0, sizeof(a));
  a.xyzh = S.xyzh; a.ldx = S.ldxyzh;
  a.vxyzu = S.vxyzu; a.ldv = S.ldvxyzu;
  a.fxyzu = S.fxyzu; a.ldf = S.ldfxyzu;
  a.vpred = ind ? nullptr : S.vpred;
  a.Bevol = mhd ? S.Bevol : nullptr;
  a.dBevol = mhd ? S.dBevol : nullptr;
  a.dustevol = dust ? S.dustevol : nullptr;
  a.ddustevol = dust ? S.ddustevol : nullptr;
  a.iphase = S.iphase;
  if (ind) {
    a.ibin = S.ibin;
    a.ibin_wake = S.ibin_wake;
    a.twas = S.twas;
  }
  return a;
}

StepArgs args_reset(const phsph_ctx *c, const phsph_step_state &S) {
  const bool mhd = c->p.mhd, dust = dust_on(c);
  StepArgs a;
  memset(&a, 0, sizeof(a));
  a.vxyzu = S.vxyzu; a.ldv = S.ldvxyzu;
  a.fxyzu = S.fxyzu; a.ldf = S.ldfxyzu;
  a.vpred = S.vpred;
  a.Bevol = mhd ? S.Bevol : nullptr;
  a.dBevol = mhd ? S.dBevol : nullptr;
  a.Bpred = mhd ? S.Bpred : nullptr;
  a.dustevol = dust ? S.dustevol : nullptr;
  a.ddustevol = dust ? S.ddustevol : nullptr;
  a.dustpred = dust ? S.dustpred : nullptr;
  a.iphase = S.iphase;
  return a;
}

// derivs (deriv.f90:114-189) on the predicted state; returns dtnew
// (deriv.f90:186-191) in *dtnew
int derivs(phsph_ctx *c, int icall, const phsph_step_state &S, phsph_step_clock &k, double *dtnew) {
  const int np = S.np;
  const bool mhd = c->p.mhd, dust = dust_on(c), ind = c->p.ind_timesteps;
  int rc;
  if (icall == 1) {
    c->dlist_reuse = false;
    rc = tree_build(c, S.xyzh, S.ldxyzh, np, S.iphase);
    if (rc) return rc;
    DensIOArgs d;
    memset(&d, 0, sizeof(d));
    d.xyzh = S.xyzh; d.ldx = S.ldxyzh;
    d.vxyzu = S.vpred; d.ldv = S.ldvxyzu;
    d.fxyzu = S.fxyzu; d.ldf = S.ldfxyzu;
    d.Bevol = mhd ? S.Bpred : nullptr;
    d.iphase = S.iphase;
    d.divcurlv = S.divcurlv; d.lddcv = S.lddivcurlv;
    d.divcurlB = mhd ? S.divcurlB : nullptr;
    d.alphaind = S.alphaind;
    d.gradh = S.gradh; d.ldgradh = S.ldgradh;
    d.dvdx = S.dvdx;
    rc = dens_run(c, 1, d);
    if (rc) return rc;
  }
  rc = cons2prim_run(c, np, S.xyzh, S.ldxyzh, S.vpred, S.ldvxyzu, S.dvdx, S.alphaind, dust ? S.dustpred : nullptr,
                     dust ? S.dustfrac : nullptr, S.eos_vars, S.ldeos);
  if (rc) return rc;
  ForceIOArgs f;
  memset(&f, 0, sizeof(f));
  f.np = np;
  f.xyzh = S.xyzh; f.ldx = S.ldxyzh;
  f.vxyzu = S.vpred; f.ldv = S.ldvxyzu;
  f.fxyzu = S.fxyzu; f.ldf = S.ldfxyzu;
  f.divcurlv = S.divcurlv; f.lddcv = S.lddivcurlv;
  f.divcurlB = mhd ? S.divcurlB : nullptr;
  f.Bevol = mhd ? S.Bpred : nullptr;
  f.dBevol = mhd ? S.dBevol : nullptr;
  f.divBsymm = mhd ? S.divBsymm : nullptr;
  f.dustfrac = dust ? S.dustfrac : nullptr;
  f.ddustevol = dust ? S.ddustevol : nullptr;
  f.deltav = dust ? S.deltav : nullptr;
  f.tstop = dust ? S.tstop : nullptr;
  f.gradh = S.gradh; f.ldgradh = S.ldgradh;
  f.alphaind = S.alphaind;
  f.eos_vars = S.eos_vars; f.ldeos = S.ldeos;
  f.iphase = S.iphase;
  if (ind) {
    f.ibin = S.ibin;
    f.ibin_old = S.ibin_old;
    f.ibin_wake = S.ibin_wake;
    f.nbinmax = k.nbinmax;
    f.istepfrac = k.istepfrac;
    f.ibinnow = k.ibinnow;
  }
  c->poten_out = c->p.gravity ? S.poten : nullptr;
  double dtc = 0., dtf = 0.;
  rc = force_run(c, icall, f, &dtc, &dtf);
  c->poten_out = nullptr;
  if (rc) return rc;
  if (ind) {
    // force leaves nbinmax at the largest new bin of the active particles
    // (force.F90:748-792)
    k.nbinmax = c->nbinmax_new;
    *dtnew = c->p.dtmax / std::ldexp(1.0, k.nbinmax);
    return PHSPH_OK;
  }
  k.dtcourant = dtc;
  k.dtforce = dtf;
  double m = dtf;
  if (dtc < m) m = dtc;
  if (c->p.dtmax < m) m = c->p.dtmax;
  *dtnew = m;
  return PHSPH_OK;
}

// check_velocity_error (step_leapfrog.f90:769-843) without MPI and damping
bool check_velocity_error(double errmax, double v2sum, int npc, int its, double tolv, double dt, double dtcourant,
                          double dtforce, bool ind, double *dterr) {
  const double v2mean = npc > 0 ? v2sum / npc : 0.0;
  if (v2mean > DBL_MIN)
    errmax = errmax / std::sqrt(v2mean);
  else
    errmax = 0.0;
  double errtol = tolv;
  *dterr = DBL_MAX;
  if (!(tolv < 1.e2)) return true;
  if (!ind) {
    const double dtf = dtcourant < dtforce ? dtcourant : dtforce;
    if (dt < dtf && dtf < BIGNUMBER) errtol = errtol * std::pow(dt / dtf, 2);
    if (its == 1 && errtol > DBL_MIN && errmax > DBL_EPSILON) *dterr = dt * std::sqrt(errtol / errmax);
  }
  return errmax < tolv;
}

// one step (step_leapfrog.f90:91-760) from time t over dtsph
int step_once(phsph_ctx *c, bool graphs, const phsph_step_state &S, phsph_step_clock &k, double t, double dtsph,
              double ufloor, double tolv, int maxits, phsph_step_summary &o) {
  const int np = S.np;
  const bool ind = c->p.ind_timesteps;
  int rc = step_chain_prepare(c, np);
  if (rc) return rc;
  const StepArgs pred = args_predictor(c, S), psph = args_predict_sph(c, S), corr = args_corrector(c, S);
  rc = run_chain(c, graphs, 0, [&] { step_chain_pre(c, np, t, dtsph, pred, psph, ufloor); });
  if (rc) return rc;
  c->tree_valid = false;  // the drift moved the particles
  const double timei = t + dtsph;
  double dtnew = 0.;
  rc = derivs(c, 1, S, k, &dtnew);
  if (rc) return rc;
  StepSums sums[3];
  int its = 0, nwake = 0, nvfloorc = 0;
  bool converged = false;
  double dterr = DBL_MAX;
  bool first = true;
  while (its < maxits && !converged) {
    its++;
    // the step_red buffer was sized for np above and nothing in derivs uses it
    rc = run_chain(c, graphs, 1, [&] { step_chain_post(c, np, timei, dtsph, corr, k.nbinmax, ufloor); });
    if (rc) return rc;
    rc = step_chain_read(c, np, sums);
    if (rc) return rc;
    if (first) {
      o.nvfloorp = sums[0].nfloor;
      o.nvfloorps = sums[1].nfloor;
      first = false;
    }
    if (sums[2].nbad > 0) PH_RET(c, PHSPH_ERR_INPUT, "step_graph: ibin outside 0..30");
    nwake += sums[2].nwake;
    nvfloorc += sums[2].nfloor;
    double dterr_i;
    converged = check_velocity_error(sums[2].emax, sums[2].v2, sums[2].np, its, tolv, dtsph, k.dtcourant, k.dtforce,
                                     ind, &dterr_i);
    // timestep:dterr is overwritten by every call (step_leapfrog.f90:807)
    dterr = dterr_i;
    if (!converged) {
      rc = step_reset_run(c, np, dtsph, args_reset(c, S));
      if (rc) return rc;
      double unused;
      rc = derivs(c, 2, S, k, &unused);
      if (rc) return rc;
    }
  }
  if (maxits > 1 && its >= maxits && !converged)
    PH_RET(c, PHSPH_ERR_NOCONVERGE, "step_graph: velocity iterations not converged");
  o.dt = dtsph;
  o.dtnew = dtnew;
  o.dterr = dterr;
  o.its = its;
  o.nwake = nwake;
  o.nvfloorc = nvfloorc;
  return PHSPH_OK;
}

}  // namespace

int step_graph_run(phsph_ctx *c, const phsph_step_state &S, phsph_step_clock &k, int nsub, double ufloor,
                   double tolv, int maxits, bool graphs, phsph_step_summary *out) {
  const bool ind = c->p.ind_timesteps;
  const double dtmax = c->p.dtmax;
  for (int n = 0; n < nsub; n++) {
    ctx_clear_timings(c);
    phsph_step_summary o;
    memset(&o, 0, sizeof(o));
    const int64_t launch0 = c->step_graph_launch, inst0 = c->step_graph_inst;
    int rc;
    if (ind) {
      const int nbinmaxprev = k.nbinmax;
      k.istepfrac += 1;
      if (k.istepfrac > (1 << k.nbinmax)) PH_RET(c, PHSPH_ERR_STATE, "step_graph: error in individual timesteps");
      int32_t na = 0, nl = 0;
      rc = set_active_run(c, S.np, S.xyzh, S.ldxyzh, S.iphase, S.ibin, k.nbinmax, k.istepfrac, &na, &nl);
      if (rc) return rc;
      // the largest active bin (utils_indtimesteps.f90:170-176)
      int bnow = k.nbinmax;
      for (int b = 0; bnow == k.nbinmax && b < k.nbinmax; b++)
        if (k.istepfrac % (1 << (k.nbinmax - b)) == 0) bnow = b;
      k.ibinnow = bnow;
      o.nactive = na;
      const double dt = dtmax / std::ldexp(1.0, k.nbinmax);
      rc = step_once(c, graphs, S, k, k.time, dt, ufloor, tolv, maxits, o);
      if (rc) return rc;
      k.time = k.tlast + (double)k.istepfrac / std::ldexp(1.0, nbinmaxprev) * dtmax;
      if (k.nbinmax != nbinmaxprev) {  // change_nbinmax
        if (k.nbinmax < nbinmaxprev && k.istepfrac % (1 << (nbinmaxprev - k.nbinmax)) != 0)
          PH_RET(c, PHSPH_ERR_STATE, "step_graph: istepfrac not a multiple of the new bin ratio");
        if (k.nbinmax > nbinmaxprev)
          k.istepfrac = k.istepfrac * (1 << (k.nbinmax - nbinmaxprev));
        else
          k.istepfrac = k.istepfrac / (1 << (nbinmaxprev - k.nbinmax));
      }
      if (k.istepfrac == (1 << k.nbinmax)) {
        k.istepfrac = 0;
        k.tlast = k.time;
      }
    } else {
      const double dt = k.dt;
      o.nactive = S.np;
      rc = step_once(c, graphs, S, k, k.time, dt, ufloor, tolv, maxits, o);
      if (rc) return rc;
      k.time = k.time + dt;
      double m = k.dtforce;
      if (k.dtcourant < m) m = k.dtcourant;
      if (o.dterr < m) m = o.dterr;
      const double cap = dtmax + DBL_EPSILON;
      if (cap < m) m = cap;
      k.dt = m;
    }
    o.ngraph = (int32_t)(c->step_graph_launch - launch0);
    o.ninst = (int32_t)(c->step_graph_inst - inst0);
    if (out) out[n] = o;
  }
  return PHSPH_OK;
}
