// step.hip -- the particle loops of the leapfrog integrator
// (step_leapfrog.f90:91-760, substepping.F90:241-264): predictor, drift,
// prediction to the full step, corrector and the reset before a
// velocity-dependent iteration.  Each loop is a streaming kernel with one
// thread per particle on the caller's arrays (no tree order, no staging when
// the arrays are already in HBM).  Counts and the corrector's velocity-error
// sums are reduced per block into partials and then by one block in a fixed
// order, so the returned sums do not depend on the launch schedule.
//
// Every product that feeds a sum passes through xmul, so the device
// rounds each operation as the reference's FP64 build does and the updated
// state is bit-identical to a sequential evaluation.
#include "ctx.h"

#include <mutex>

namespace {

constexpr int SBS = 256;          // threads per block of the step kernels
constexpr int MAXBINS = 30;       // utils_indtimesteps.f90:29
constexpr double AVDECAYCONST = 0.1;  // shock_capturing.f90:32

// Per-bin half step and wake time of one step: ibin_dts(ithdt,:) and
// ibin_dts(ittwas,:) (step_leapfrog.f90:73-80, 156-161).
struct BinTable {
  double hdt[MAXBINS + 1];
  double twas[MAXBINS + 1];
};

// get_dt (utils_indtimesteps.f90:45-51): dtmax/2**ibin, the power formed by
// repeated doubling as the reference's real**integer(1) is.
double get_dt_host(double dtmax, int ibin) {
  double p = 1.;
  for (int k = 0; k < ibin; k++) p = p * 2.;
  return dtmax / p;
}

BinTable make_bin_table(double dtmax, double time_now) {
  BinTable t;
  for (int b = 0; b <= MAXBINS; b++) {
    const double dtb = get_dt_host(dtmax, b);
    const double dt1 = 1.0 / dtb;
    t.hdt[b] = 0.5 * dtb;
    t.twas[b] = ((double)(int64_t)(time_now * dt1) + 0.5) * dtb;
  }
  return t;
}

// Block partial of the step reductions.
struct StepPart {
  double emax, v2;
  int np, nwake, nfloor, nbad;  // nbad: bins outside 0..MAXBINS
};

struct StepFlags {
  int nv, mhd, dust, ind, nbinmax;
  double hfact, pmass, ufloor;
};

__device__ __forceinline__ bool is_live(double h) { return !(h < PH_TINY64); }
__device__ __forceinline__ bool is_active(const int8_t *iphase, int i) { return !iphase || iphase[i] > 0; }
// a + b*c with the product rounded first
__device__ __forceinline__ double axpy(double a, double b, double c) { return a + xmul(b, c); }

// reduce a block's partial and write it to parts[blockIdx.x]
__device__ void block_part(StepPart v, StepPart *parts) {
  __shared__ double se[SBS / 64], sv[SBS / 64];
  __shared__ int sn[SBS / 64], sw[SBS / 64], sf[SBS / 64], sb[SBS / 64];
  double e = wave_max(v.emax), s = wave_sum(v.v2);
  int n = wave_sum_i(v.np), w = wave_sum_i(v.nwake), f = wave_sum_i(v.nfloor), b = wave_sum_i(v.nbad);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    se[wid] = e; sv[wid] = s; sn[wid] = n; sw[wid] = w; sf[wid] = f; sb[wid] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    StepPart r = {0., 0., 0, 0, 0, 0};
    for (int k = 0; k < SBS / 64; k++) {
      r.emax = fmax(r.emax, se[k]);
      r.v2 += sv[k];
      r.np += sn[k]; r.nwake += sw[k]; r.nfloor += sf[k]; r.nbad += sb[k];
    }
    parts[blockIdx.x] = r;
  }
}

// predictor loop (step_leapfrog.f90:181-237)
__global__ void __launch_bounds__(SBS) k_predictor(StepArgs a, StepFlags fl, double timei, double hdtsph, int np,
                                                    StepPart *parts) {
  const int i = blockIdx.x * SBS + threadIdx.x;
  StepPart r = {0., 0., 0, 0, 0, 0};
  if (i < np && is_live(a.xyzh[(size_t)i * a.ldx + 3])) {
    double hdti;
    if (fl.ind) {
      if (is_active(a.iphase, i)) a.ibin_old[i] = a.ibin_in[i];
      hdti = a.twas[i] - timei;
    } else {
      hdti = hdtsph;
    }
    double *v = a.vxyzu + (size_t)i * a.ldv;
    const double *f = a.fxyzu + (size_t)i * a.ldf;
    for (int k = 0; k < fl.nv; k++) v[k] = axpy(v[k], hdti, f[k]);
    if (fl.ufloor > 0. && fl.nv >= 4 && v[3] < fl.ufloor) {
      v[3] = fl.ufloor;
      r.nfloor = 1;
    }
    if (fl.mhd) {
      double *B = a.Bevol + (size_t)i * 4;
      const double *dB = a.dBevol + (size_t)i * 4;
      for (int k = 0; k < 4; k++) B[k] = axpy(B[k], hdti, dB[k]);
    }
    if (fl.dust) a.dustevol[i] = fabs(axpy(a.dustevol[i], hdti, a.ddustevol[i]));
  }
  block_part(r, parts);
}

// substep_sph (substepping.F90:241-264)
__global__ void __launch_bounds__(SBS) k_drift(double *xyzh, int ldx, const double *vxyzu, int ldv, double dt,
                                                int np) {
  const int i = blockIdx.x * SBS + threadIdx.x;
  if (i >= np) return;
  double *x = xyzh + (size_t)i * ldx;
  if (!is_live(x[3])) return;
  const double *v = vxyzu + (size_t)i * ldv;
  x[0] = axpy(x[0], dt, v[0]);
  x[1] = axpy(x[1], dt, v[1]);
  x[2] = axpy(x[2], dt, v[2]);
}

// predict_sph loop (step_leapfrog.f90:307-400), gas with the Cullen-Dehnen
// switch (nalpha = 3, config.F90:161)
__global__ void __launch_bounds__(SBS) k_predict_sph(StepArgs a, StepFlags fl, double timei, double dtsph,
                                                      double hdtsph, int np, StepPart *parts) {
  const int i = blockIdx.x * SBS + threadIdx.x;
  StepPart r = {0., 0., 0, 0, 0, 0};
  double *x = a.xyzh + (size_t)i * a.ldx;
  if (i < np && is_live(x[3])) {
    // h prediction (:332): h - dtsph*dhdrho(h)*rhoh(h)*div v
    const double h0 = x[3];
    const double rho = rhoh_d(h0, fl.hfact, fl.pmass);
    const double dhdrho = -h0 / (3. * rho);
    const double hi = h0 - xmul(xmul(xmul(dtsph, dhdrho), rho), (double)a.divcurlv[(size_t)i * a.lddcv]);
    x[3] = hi;
    const double hdti = fl.ind ? timei - a.twas[i] : hdtsph;
    const double *v = a.vxyzu + (size_t)i * a.ldv;
    const double *f = a.fxyzu + (size_t)i * a.ldf;
    double *vp = a.vpred + (size_t)i * a.ldv;
    for (int k = 0; k < fl.nv; k++) vp[k] = axpy(v[k], hdti, f[k]);
    if (fl.ufloor > 0. && fl.nv >= 4 && vp[3] < fl.ufloor) {
      vp[3] = fl.ufloor;
      r.nfloor = 1;
    }
    if (fl.mhd) {
      const double *B = a.Bevol + (size_t)i * 4;
      const double *dB = a.dBevol + (size_t)i * 4;
      double *Bp = a.Bpred + (size_t)i * 4;
      for (int k = 0; k < 4; k++) Bp[k] = axpy(B[k], hdti, dB[k]);
    }
    if (fl.dust) a.dustpred[i] = axpy(a.dustevol[i], hdti, a.ddustevol[i]);
    // viscosity switch (:377-396)
    const double spsoundi = a.eos_vars[(size_t)i * a.ldeos + 1];
    const double tdecay1 = xmul(AVDECAYCONST, spsoundi) / hi;
    const double ddenom = 1. / (1. + xmul(dtsph, tdecay1));
    float *al = a.alphaind + (size_t)i * 3;
    const double alphaloci = (double)al[1];
    const double a1 = (double)al[0];
    if (a1 < alphaloci)
      al[0] = (float)alphaloci;
    else
      al[0] = (float)xmul(a1 + xmul(xmul(dtsph, alphaloci), tdecay1), ddenom);
  }
  block_part(r, parts);
}

// one pass of the corrector loop (step_leapfrog.f90:482-625)
__global__ void __launch_bounds__(SBS) k_corrector(StepArgs a, StepFlags fl, BinTable tab, double timei,
                                                    double hdtsph, int np, StepPart *parts) {
  const int i = blockIdx.x * SBS + threadIdx.x;
  StepPart r = {0., 0., 0, 0, 0, 0};
  if (i < np && is_live(a.xyzh[(size_t)i * a.ldx + 3])) {
    double *v = a.vxyzu + (size_t)i * a.ldv;
    const double *f = a.fxyzu + (size_t)i * a.ldf;
    double *B = fl.mhd ? a.Bevol + (size_t)i * 4 : nullptr;
    const double *dB = fl.mhd ? a.dBevol + (size_t)i * 4 : nullptr;
    if (fl.ind) {
      int ib = a.ibin[i];
      if (ib < 0 || ib > MAXBINS) {  // outside the bin table: reported as an error
        r.nbad = 1;
        ib = ib < 0 ? 0 : MAXBINS;
      }
      double tw = a.twas[i];
      if (is_active(a.iphase, i)) {
        a.ibin_wake[i] = 0;
        const double dti = (timei - tw) + tab.hdt[ib];
        for (int k = 0; k < fl.nv; k++) v[k] = axpy(v[k], dti, f[k]);
        if (fl.mhd)
          for (int k = 0; k < 4; k++) B[k] = axpy(B[k], dti, dB[k]);
        if (fl.dust) a.dustevol[i] = axpy(a.dustevol[i], dti, a.ddustevol[i]);
        tw = tw + dti;
      }
      const double hdti = timei - tw;
      for (int k = 0; k < fl.nv; k++) v[k] = axpy(v[k], hdti, f[k]);
      if (fl.ufloor > 0. && fl.nv >= 4 && v[3] < fl.ufloor) {
        v[3] = fl.ufloor;
        r.nfloor = 1;
      }
      if (fl.mhd)
        for (int k = 0; k < 4; k++) B[k] = axpy(B[k], hdti, dB[k]);
      if (fl.dust) a.dustevol[i] = axpy(a.dustevol[i], hdti, a.ddustevol[i]);
      const int wk = a.ibin_wake[i];
      if (wk > ib) {  // wake for the next step (:553-559)
        const int w = wk < fl.nbinmax ? wk : fl.nbinmax;
        r.nwake = 1;
        tw = tab.twas[w];
        a.ibin[i] = (int8_t)w;
        a.ibin_wake[i] = 0;
      }
      a.twas[i] = tw;
    } else {
      const double *vp = a.vpred + (size_t)i * a.ldv;
      const double vxi = axpy(v[0], hdtsph, f[0]);
      const double vyi = axpy(v[1], hdtsph, f[1]);
      const double vzi = axpy(v[2], hdtsph, f[2]);
      const double dx = vxi - vp[0], dy = vyi - vp[1], dz = vzi - vp[2];
      r.emax = r2_exact(dx, dy, dz);
      r.v2 = r2_exact(vxi, vyi, vzi);
      r.np = 1;
      if (fl.nv >= 4) v[3] = axpy(v[3], hdtsph, f[3]);
      v[0] = vxi;
      v[1] = vyi;
      v[2] = vzi;
      if (fl.mhd)
        for (int k = 0; k < 4; k++) B[k] = axpy(B[k], hdtsph, dB[k]);
      if (fl.dust) a.dustevol[i] = axpy(a.dustevol[i], hdtsph, a.ddustevol[i]);
    }
  }
  block_part(r, parts);
}

// until_converged loop (step_leapfrog.f90:650-696); no dead-particle test,
// as in the reference
__global__ void __launch_bounds__(SBS) k_reset(StepArgs a, StepFlags fl, double hdtsph, int np) {
  const int i = blockIdx.x * SBS + threadIdx.x;
  if (i >= np) return;
  if (fl.ind && !is_active(a.iphase, i)) return;
  double *v = a.vxyzu + (size_t)i * a.ldv;
  double *vp = a.vpred + (size_t)i * a.ldv;
  for (int k = 0; k < fl.nv; k++) vp[k] = v[k];
  if (fl.mhd)
    for (int k = 0; k < 4; k++) a.Bpred[(size_t)i * 4 + k] = a.Bevol[(size_t)i * 4 + k];
  if (fl.dust) a.dustpred[i] = a.dustevol[i];
  if (fl.ind) return;
  const double *f = a.fxyzu + (size_t)i * a.ldf;
  for (int k = 0; k < fl.nv; k++) v[k] = v[k] - xmul(hdtsph, f[k]);
  if (fl.mhd) {
    double *B = a.Bevol + (size_t)i * 4;
    const double *dB = a.dBevol + (size_t)i * 4;
    for (int k = 0; k < 4; k++) B[k] = B[k] - xmul(hdtsph, dB[k]);
  }
  if (fl.dust) a.dustevol[i] = a.dustevol[i] - xmul(hdtsph, a.ddustevol[i]);
}

// init_step (step_leapfrog.f90:44-83)
__global__ void __launch_bounds__(SBS) k_init_step(int8_t *ibin, double *twas, double time, double dtmax,
                                                    int reset, int nbinmax, int np) {
  const int i = blockIdx.x * SBS + threadIdx.x;
  if (i >= np) return;
  int b = ibin[i];
  if (reset) {
    b = nbinmax;
    ibin[i] = (int8_t)b;
  }
  double p = 1.;
  for (int k = 0; k < b; k++) p = p * 2.;
  twas[i] = time + 0.5 * (dtmax / p);
}

// final reduction of the block partials, in block order
__global__ void __launch_bounds__(SBS) k_step_reduce(const StepPart *parts, int nparts, StepPart *out) {
  __shared__ double se[SBS], sv[SBS];
  __shared__ int sn[SBS], sw[SBS], sf[SBS], sb[SBS];
  StepPart r = {0., 0., 0, 0, 0, 0};
  for (int k = threadIdx.x; k < nparts; k += SBS) {
    const StepPart p = parts[k];
    r.emax = fmax(r.emax, p.emax);
    r.v2 += p.v2;
    r.np += p.np; r.nwake += p.nwake; r.nfloor += p.nfloor; r.nbad += p.nbad;
  }
  se[threadIdx.x] = r.emax; sv[threadIdx.x] = r.v2;
  sn[threadIdx.x] = r.np; sw[threadIdx.x] = r.nwake; sf[threadIdx.x] = r.nfloor; sb[threadIdx.x] = r.nbad;
  __syncthreads();
  for (int s = SBS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      se[threadIdx.x] = fmax(se[threadIdx.x], se[threadIdx.x + s]);
      sv[threadIdx.x] += sv[threadIdx.x + s];
      sn[threadIdx.x] += sn[threadIdx.x + s];
      sw[threadIdx.x] += sw[threadIdx.x + s];
      sf[threadIdx.x] += sf[threadIdx.x + s];
      sb[threadIdx.x] += sb[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    StepPart o = {se[0], sv[0], sn[0], sw[0], sf[0], sb[0]};
    *out = o;
  }
}

StepFlags make_flags(const phsph_ctx *c, double ufloor, int nbinmax) {
  StepFlags f;
  f.nv = c->p.maxvxyzu;
  f.mhd = c->p.mhd ? 1 : 0;
  f.dust = (c->p.use_dustfrac && c->p.ndustsmall > 0) ? 1 : 0;
  f.ind = c->p.ind_timesteps ? 1 : 0;
  f.nbinmax = nbinmax;
  f.hfact = c->p.hfact;
  f.pmass = c->p.massoftype_gas;
  f.ufloor = ufloor;
  return f;
}

// room for the block partials and the final result
int ensure_parts(phsph_ctx *c, int nb, StepPart **parts, StepPart **out) {
  const size_t words = (size_t)(nb + 1) * (sizeof(StepPart) / sizeof(double));
  PH_HIPCHECK(c, c->step_red.ensure(words));
  *parts = reinterpret_cast<StepPart *>(c->step_red.p);
  *out = *parts + nb;
  return PHSPH_OK;
}

int finish_reduce(phsph_ctx *c, int nb, StepPart *parts, StepPart *out, StepPart *host) {
  hipLaunchKernelGGL(k_step_reduce, dim3(1), dim3(SBS), 0, c->stream, parts, nb, out);
  PH_HIPCHECK(c, hipGetLastError());
  PH_HIPCHECK(c, hipMemcpyAsync(host, out, sizeof(StepPart), hipMemcpyDeviceToHost, c->stream));
  return PHSPH_OK;
}

}  // namespace

int step_init_run(phsph_ctx *c, int np, double time, int8_t *ibin, int nbinmax, double *twas) {
  if (np > 0) {
    LaunchTimer lt(c, "init_step");
    hipLaunchKernelGGL(k_init_step, dim3(nblocks(np, SBS)), dim3(SBS), 0, c->stream, ibin, twas, time, c->p.dtmax,
                       time < PH_TINY64 ? 1 : 0, nbinmax, np);
  }
  PH_HIPCHECK(c, hipGetLastError());
  PH_HIPCHECK(c, hipStreamSynchronize(c->stream));
  return PHSPH_OK;
}

int step_predictor_run(phsph_ctx *c, int np, double t, double dtsph, const StepArgs &s, double ufloor,
                       int32_t *nvfloor) {
  const int nb = np > 0 ? (int)nblocks(np, SBS) : 1;
  StepPart *parts, *out, res;
  int rc = ensure_parts(c, nb, &parts, &out);
  if (rc) return rc;
  {
    LaunchTimer lt(c, "step_predictor");
    hipLaunchKernelGGL(k_predictor, dim3(nb), dim3(SBS), 0, c->stream, s, make_flags(c, ufloor, 0), t, 0.5 * dtsph,
                       np, parts);
  }
  PH_HIPCHECK(c, hipGetLastError());
  rc = finish_reduce(c, nb, parts, out, &res);
  if (rc) return rc;
  PH_HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (nvfloor) *nvfloor = res.nfloor;
  return PHSPH_OK;
}

int step_drift_run(phsph_ctx *c, int np, double dt, double *xyzh, int ldx, const double *vxyzu, int ldv) {
  if (np > 0) {
    LaunchTimer lt(c, "substep_sph");
    hipLaunchKernelGGL(k_drift, dim3(nblocks(np, SBS)), dim3(SBS), 0, c->stream, xyzh, ldx, vxyzu, ldv, dt, np);
  }
  PH_HIPCHECK(c, hipGetLastError());
  PH_HIPCHECK(c, hipStreamSynchronize(c->stream));
  return PHSPH_OK;
}

int step_predict_sph_run(phsph_ctx *c, int np, double timei, double dtsph, const StepArgs &s, double ufloor,
                         int32_t *nvfloor) {
  const int nb = np > 0 ? (int)nblocks(np, SBS) : 1;
  StepPart *parts, *out, res;
  int rc = ensure_parts(c, nb, &parts, &out);
  if (rc) return rc;
  {
    LaunchTimer lt(c, "step_predict_sph");
    hipLaunchKernelGGL(k_predict_sph, dim3(nb), dim3(SBS), 0, c->stream, s, make_flags(c, ufloor, 0), timei, dtsph,
                       0.5 * dtsph, np, parts);
  }
  PH_HIPCHECK(c, hipGetLastError());
  rc = finish_reduce(c, nb, parts, out, &res);
  if (rc) return rc;
  PH_HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (nvfloor) *nvfloor = res.nfloor;
  return PHSPH_OK;
}

int step_corrector_run(phsph_ctx *c, int np, double timei, double dtsph, const StepArgs &s, int nbinmax,
                       double ufloor, double *errmax, double *v2sum, int32_t *npc, int32_t *nwake,
                       int32_t *nvfloor) {
  const int nb = np > 0 ? (int)nblocks(np, SBS) : 1;
  StepPart *parts, *out, res;
  int rc = ensure_parts(c, nb, &parts, &out);
  if (rc) return rc;
  const BinTable tab = make_bin_table(c->p.dtmax, timei);
  {
    LaunchTimer lt(c, "step_corrector");
    hipLaunchKernelGGL(k_corrector, dim3(nb), dim3(SBS), 0, c->stream, s, make_flags(c, ufloor, nbinmax), tab,
                       timei, 0.5 * dtsph, np, parts);
  }
  PH_HIPCHECK(c, hipGetLastError());
  rc = finish_reduce(c, nb, parts, out, &res);
  if (rc) return rc;
  PH_HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (res.nbad > 0) PH_RET(c, PHSPH_ERR_INPUT, "step_corrector: ibin outside 0..30");
  if (errmax) *errmax = res.emax;
  if (v2sum) *v2sum = res.v2;
  if (npc) *npc = res.np;
  if (nwake) *nwake = res.nwake;
  if (nvfloor) *nvfloor = res.nfloor;
  return PHSPH_OK;
}

// ---- the corrector without allocation --------------------------------------
// Scratch memory of phsph_step_corrector_noalloc: the per-block partials of
// the corrector's reductions and their combined result, in a fixed device
// array instead of context buffers.  Each GPU holds its own copy for the
// life of the process; every call on a GPU shares it.
constexpr int NOALLOC_MAX_BLOCKS = 65536;  // np <= NOALLOC_MAX_BLOCKS * SBS = 16,777,216
static __device__ __align__(256) StepPart scratch[NOALLOC_MAX_BLOCKS + 1];
constexpr int NOALLOC_MAX_DEVICES = 64;
static StepPart *g_scratch[NOALLOC_MAX_DEVICES];  // device address of scratch per device
static std::mutex g_scratch_mu;

// Looks up the device address of scratch on the current device (the
// caller's context device), once per device; called by phsph_create.
int step_scratch_init(phsph_ctx *c) {
  if (c->device < 0 || c->device >= NOALLOC_MAX_DEVICES) PH_RET(c, PHSPH_ERR_INPUT, "device index beyond 63");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch[c->device]) return PHSPH_OK;
  void *ptr = nullptr;
  PH_HIPCHECK(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  g_scratch[c->device] = static_cast<StepPart *>(ptr);
  return PHSPH_OK;
}

int step_corrector_noalloc_run(phsph_ctx *c, int np, double timei, double dtsph, const StepArgs &s, int nbinmax,
                               double ufloor, double *errmax, double *v2sum, int32_t *npc, int32_t *nwake,
                               int32_t *nvfloor) {
  const int nb = np > 0 ? (int)nblocks(np, SBS) : 1;
  if (nb > NOALLOC_MAX_BLOCKS)
    PH_RET(c, PHSPH_ERR_CAPACITY, "step_corrector_noalloc: np above 16,777,216 (the scratch limit)");
  StepPart *parts;
  {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    parts = (c->device >= 0 && c->device < NOALLOC_MAX_DEVICES) ? g_scratch[c->device] : nullptr;
  }
  if (!parts) PH_RET(c, PHSPH_ERR_INPUT, "step_corrector_noalloc: scratch not set up on this device");
  StepPart *out = parts + nb, res;
  const BinTable tab = make_bin_table(c->p.dtmax, timei);
  // every block writes its partial and k_step_reduce overwrites out, so
  // nothing left by an earlier call is read
  hipLaunchKernelGGL(k_corrector, dim3(nb), dim3(SBS), 0, c->stream, s, make_flags(c, ufloor, nbinmax), tab, timei,
                     0.5 * dtsph, np, parts);
  PH_HIPCHECK(c, hipGetLastError());
  int rc = finish_reduce(c, nb, parts, out, &res);
  if (rc) return rc;
  PH_HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (res.nbad > 0) PH_RET(c, PHSPH_ERR_INPUT, "step_corrector_noalloc: ibin outside 0..30");
  if (errmax) *errmax = res.emax;
  if (v2sum) *v2sum = res.v2;
  if (npc) *npc = res.np;
  if (nwake) *nwake = res.nwake;
  if (nvfloor) *nvfloor = res.nfloor;
  return PHSPH_OK;
}

int step_reset_run(phsph_ctx *c, int np, double dtsph, const StepArgs &s) {
  if (np > 0) {
    LaunchTimer lt(c, "step_reset");
    hipLaunchKernelGGL(k_reset, dim3(nblocks(np, SBS)), dim3(SBS), 0, c->stream, s, make_flags(c, 0., 0),
                       0.5 * dtsph, np);
  }
  PH_HIPCHECK(c, hipGetLastError());
  PH_HIPCHECK(c, hipStreamSynchronize(c->stream));
  return PHSPH_OK;
}

// ---- enqueue-only chains of whole steps ------------------------------------
// The reduction buffer holds the block partials (reused by each loop in
// stream order) and three result slots: predictor, predict_sph, corrector.

static_assert(sizeof(StepSums) == sizeof(StepPart), "StepSums mirrors StepPart");

int step_chain_prepare(phsph_ctx *c, int np) {
  const int nb = np > 0 ? (int)nblocks(np, SBS) : 1;
  const size_t words = (size_t)(nb + 3) * (sizeof(StepPart) / sizeof(double));
  PH_HIPCHECK(c, c->step_red.ensure(words));
  return PHSPH_OK;
}

void step_chain_pre(phsph_ctx *c, int np, double t, double dtsph, const StepArgs &pred, const StepArgs &psph,
                    double ufloor) {
  const int nb = np > 0 ? (int)nblocks(np, SBS) : 1;
  StepPart *parts = reinterpret_cast<StepPart *>(c->step_red.p), *out = parts + nb;
  hipStream_t s = c->stream;
  const StepFlags fl = make_flags(c, ufloor, 0);
  hipLaunchKernelGGL(k_predictor, dim3(nb), dim3(SBS), 0, s, pred, fl, t, 0.5 * dtsph, np, parts);
  hipLaunchKernelGGL(k_step_reduce, dim3(1), dim3(SBS), 0, s, parts, nb, out);
  if (np > 0)
    hipLaunchKernelGGL(k_drift, dim3(nb), dim3(SBS), 0, s, psph.xyzh, psph.ldx, psph.vxyzu, psph.ldv, dtsph, np);
  const double timei = t + dtsph;
  hipLaunchKernelGGL(k_predict_sph, dim3(nb), dim3(SBS), 0, s, psph, fl, timei, dtsph, 0.5 * dtsph, np, parts);
  hipLaunchKernelGGL(k_step_reduce, dim3(1), dim3(SBS), 0, s, parts, nb, out + 1);
}

void step_chain_post(phsph_ctx *c, int np, double timei, double dtsph, const StepArgs &corr, int nbinmax,
                     double ufloor) {
  const int nb = np > 0 ? (int)nblocks(np, SBS) : 1;
  StepPart *parts = reinterpret_cast<StepPart *>(c->step_red.p), *out = parts + nb;
  const BinTable tab = make_bin_table(c->p.dtmax, timei);
  hipLaunchKernelGGL(k_corrector, dim3(nb), dim3(SBS), 0, c->stream, corr, make_flags(c, ufloor, nbinmax), tab,
                     timei, 0.5 * dtsph, np, parts);
  hipLaunchKernelGGL(k_step_reduce, dim3(1), dim3(SBS), 0, c->stream, parts, nb, out + 2);
}

int step_chain_read(phsph_ctx *c, int np, StepSums *out3) {
  const int nb = np > 0 ? (int)nblocks(np, SBS) : 1;
  const StepPart *out = reinterpret_cast<const StepPart *>(c->step_red.p) + nb;
  PH_HIPCHECK(c, hipMemcpyAsync(out3, out, 3 * sizeof(StepPart), hipMemcpyDeviceToHost, c->stream));
  PH_HIPCHECK(c, hipStreamSynchronize(c->stream));
  return PHSPH_OK;
}
