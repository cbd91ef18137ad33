// ctx.h -- host-side context of the library: device buffers that persist
// between calls (the tree and the tree-ordered particle mirror), staging of
// caller arrays, error reporting and per-launch timing.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"

// Growable device buffer.  Contents are not preserved on growth.
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t cap = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  hipError_t ensure(size_t n) {
    if (n <= cap) return hipSuccess;
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = n < 64 ? 64 : n;
    hipError_t e = hipMalloc((void **)&p, want * sizeof(T));
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Nodes of at most this many particles are built inside one workgroup;
// larger nodes take the chunked large-node phase (which also forms their
// second moments under self-gravity).
constexpr int TREE_SUB_MAX = 1024;

// One node of the tree build's large-node phase (more particles than one
// workgroup builds in LDS): its range, split decision and chunk span.
struct BigNode {
  double com[3];   // centre of mass (parallel sums, or sequential after amb)
  double tot;      // node mass
  double hmax;
  double delta;    // bound on |sequential - parallel| centre on the split axis
  double pivot;    // com[iaxis]
  int lo, hi;      // particle range (tree positions, inclusive)
  int nnode;       // node id
  int iaxis;       // split axis
  int chunk0, nchunk;  // chunks of this node (0 for a small node)
  int nl;          // particles in the left child (0: no children)
  int part;        // the two-pointer partition moves particles
  int nbad;        // swaps of the partition
  int pad_;
};

struct TimedLaunch {
  std::string name;
  hipEvent_t t0, t1;
};

// Capacities of the candidate and neighbour lists.  They grow (and the
// affected work is re-run) when a launch reports overflow.
struct Caps {
  int gcand = 512;   // candidate leaves per group of 64 particles
  int p2p = 256;     // near-field leaves per leaf (self-gravity)
  int dlist = 256;   // density neighbour list per particle
  int flist = 256;   // force neighbour list per particle
  int gnear = 128;   // neighbours per particle in the self-gravity force pass
};

// list-kernel configuration picked by the autotuner (density.hip):
// candidate leaves staged per chunk (8 or 16), wavefronts per block (2 or
// 4), squared separations stored with the density lists (1) or recomputed
// from the positions in the h iteration (0), and the density list kernel
// launched one group per wavefront (0) or as persistent wavefronts drawing
// groups from per-XCD work queues (1)
struct ListTune {
  int cl = 16, lw = 4, r2 = 0, q = 0;
};
enum { TUNE_FIXED = 0, TUNE_PENDING = 1, TUNE_DONE = 2 };
constexpr double RFAC_LEVELS[3] = {1.03, 1.05, 1.1};
constexpr int TUNE_MIN_ACTIVE = 32768;  // smallest density call the autotuner times
constexpr int TUNE_MAX_CONFIGS = 16;

struct phsph_ctx {
  phsph_params p;
  DBuf<double> hmaxf;                 // node hmax of the final h (force walk)
  DBuf<unsigned long long> dtred;     // global dt minima
  DevConst dc;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  Caps caps;
  size_t list_budget = (size_t)16 << 30;  // bytes a uniform list capacity may take

  // ---- tree (valid after maketree) ----
  int np = 0;            // particles passed to maketree
  int nlive = 0;         // live particles (h > 0) = tree positions
  int64_t ncells = 0;    // reference ncells
  int64_t nslots = 0;    // allocated node slots (1-based ids < nslots)
  int maxlevel = 0, minlevel = 0, maxlevel_indexed = 0;
  int nleaves = 0;       // leaves (leaf_is_active != 0), in node-id order
  int tree_fallbacks = 0; // top-level nodes re-summed sequentially
  bool tree_valid = false;

  DBuf<phsph_kdnode> nd;     // node(1:ncells), index = node id
  DBuf<int32_t> nd_lo, nd_hi;  // inoderange (0-based positions, inclusive)
  DBuf<int32_t> leaf_flag;   // leaf_is_active
  DBuf<int32_t> leaves;      // leaf node ids
  DBuf<int32_t> pleaf;       // tree position -> index into leaves
  DBuf<int32_t> iparts;      // inodeparts: signed 1-based original index
  DBuf<double> tx, ty, tz, th;  // treecache x, y, z, h (tree order)
  DBuf<double> th_in;        // h at densityiterate entry (tree order)
  DBuf<double> tmpd;         // scratch (tree build)
  DBuf<int32_t> tmpi, tmpi2;  // scratch (tree build)
  DBuf<int32_t> frontier_a, frontier_b;  // node ids of a tree level
  DBuf<double> fbox_a, fbox_b;           // their bounding boxes (6 each)
  DBuf<BigNode> bnode;       // large nodes of a level (one per frontier entry)
  DBuf<int32_t> ch_node;     // chunk -> frontier entry
  DBuf<int32_t> ch_i;        // per-chunk counts (4 each)
  DBuf<int32_t> ch_off;      // per-chunk partition offsets (2 each)
  DBuf<double> ch_red;       // per-chunk partial sums and bounds (12 each)
  DBuf<int32_t> counters;    // small device counters
  DBuf<double> dred;         // small device reductions

  // ---- neighbour search ----
  DBuf<uint32_t> gcand;      // candidate leaves per group [ngroups*caps.gcand]
  DBuf<int32_t> gncand;      // count per group
  DBuf<int64_t> goff;        // exact per-group offsets into gcand (after an overflow)
  const int64_t *goff_p = nullptr;  // goff.p when the last walk used exact offsets
  DBuf<int64_t> loff;        // per-group offsets of the force lists (after an overflow)
  DBuf<int32_t> gmax;        // longest list per group
  DBuf<int64_t> poff;        // per-leaf offsets of the near-field lists (after an overflow)
  DBuf<double> hmax0;        // node hmax of h at densityiterate entry (bit patterns of positive doubles)
  // the density lists also serve as force lists while positions, h and the
  // active set are those of the last densityiterate
  bool dlist_reuse = false;
  int dlist_nact = 0, dlist_cap = 0;
  int dlist_nretry = 0;             // particles of the first attempt run again
  // first search radius of densityiterate (units of radkern*h at entry):
  // index into RFAC_LEVELS, adapted after every call; it never goes below
  // rfac_floor, which rises above a level that had to be widened
  int rfac_level = 2, rfac_floor = 0;
  DBuf<int32_t> dretry0;            // their flags, by tree position
  DBuf<uint32_t> nlist_r;           // lists of retries (density) and of those particles (force)
  DBuf<double> nr2, nr2_r;          // squared separations of the own entries of the density lists
  DBuf<int32_t> t_nl_r;
  DBuf<double> t_hsum_r, t_hfin_r;
  bool force_reused_lists = false;  // the last force call used the density lists
  int ngroups = 0;
  DBuf<int32_t> ntrial;      // trial particles per leaf (statistics)
  DBuf<uint32_t> nlist;      // neighbour lists (list_at layout; node | code << 27)
  DBuf<int32_t> t_nl;        // list length per active slot
  DBuf<int32_t> t_nlj;       // density lists: entries at the back (only within the neighbour's radius)
  DBuf<double> t_hsum, t_hfin;  // density: h of the last summation, h to store

  // ---- tree-ordered particle mirror ----
  DBuf<double4> tp4;                     // density: x, y, z, h at entry
  DBuf<double4> tv4, tf4, tB4;           // density: v, f + fext, Bevol and rho(h)
  DBuf<char> frec, frecB, frecD;         // force: packed j records
  DBuf<char> frec32;                     // force: FP32 j records (mixed precision)
  DBuf<double4> fpos;                    // force: x, y, z, 1/h (tree order)
  DBuf<double4> fvel;                    // force: vx, vy, vz, u (mixed precision)
  DBuf<double> fh;                       // force: h (tree order)
  DBuf<int32_t> act;                     // active tree positions (-1: group padding)
  DBuf<int32_t> act2;                    // padded active list being built
  DBuf<int32_t> grp_brk, grp_rs, grp_gr, grp_bc;  // group formation: breaks, run starts, run groups, block sums
  int nact = 0;
  DBuf<int32_t> act_dens;                // active positions of the last densityiterate
  int nact_dens = 0;
  bool trial_stale = true;               // trial statistics not yet computed
  // autotuner state (density.hip: tune_density_iter)
  ListTune tune;
  int tune_mode = TUNE_PENDING;
  int tune_ntried = 0, tune_nact = 0;
  ListTune tune_tried[TUNE_MAX_CONFIGS];
  float tune_ms[TUNE_MAX_CONFIGS] = {0.f};
  int dens_icall = 1;                    // icall of the last densityiterate

  // per-particle / per-leaf outputs of density kept for statistics
  DBuf<int32_t> p_nneigh, p_nits, leaf_nits;
  DBuf<double> stats_d;
  DBuf<long long> stats_i;

  // neighbour statistics of the last densityiterate (dens.F90:1137)
  double trialmean = -1, actualmean = -1;
  int maxtrial = -1, maxactual = -1;
  int64_t nrhocalc = -1, nactualtot = -1;
  int64_t force_npairs = -1;  // interacting pairs of the last force call
  // precision the pair sums of the last densityiterate / force ran in
  // (PHSPH_PREC_*; -1 before the first call)
  int32_t dens_prec_used = -1, force_prec_used = -1;

  // block timesteps registered by phsph_set_bins (host or device pointers)
  int8_t *bins_ibin = nullptr, *bins_ibin_wake = nullptr;
  const int8_t *bins_ibin_old = nullptr;
  int32_t bins_nbinmax = 0, bins_istepfrac = 0, bins_ibinnow = 0;
  bool bins_set = false;
  int32_t nbinmax_new = -1;   // result of the last force call with bins
  DBuf<int8_t> tbin_old;      // ibin_old in tree order

  // self-gravity
  DBuf<double> quads;                 // node second moments, 6 per slot
  DBuf<unsigned long long> ghmax;     // node hmax as the force walk sees it
  DBuf<double> htraj, hconv;          // h per density iteration, converged h
  int64_t htraj_stride = 0;           // htraj: iteration k of particle p at (k-1)*htraj_stride + p
  DBuf<double> trial_hmax;            // hmax of each leaf at its last neighbour list (trial statistics)
  DBuf<double> fnode;                 // far-field expansion per leaf (20)
  DBuf<uint32_t> gnear;               // self-gravity force: neighbours per active particle
  DBuf<int32_t> gnear_n;              // their count per active particle
  DBuf<int64_t> gnear_off;            // per-group offsets into gnear (beyond the list budget)
  DBuf<int8_t> ibin_save;             // ibin at force entry, restored before a re-run
  // node-by-node dual walk: destination nodes of a level (ping-pong), their
  // candidate source lists (shared by siblings) and incoming expansions
  DBuf<int32_t> gw_node[2], gw_ccnt[2];
  DBuf<int64_t> gw_coff[2];
  DBuf<double> gw_H[2];
  DBuf<uint32_t> gw_cand[2];
  DBuf<int32_t> gw_act;               // node has an active leaf below (block timesteps)
  DBuf<unsigned long long> gw_used;   // candidate entries taken by the level
  DBuf<uint32_t> p2p;                 // near-field leaves per leaf
  DBuf<int32_t> np2p;
  DBuf<float> gsoft;                  // gradh(2) in tree order
  float *poten_user = nullptr;        // registered potential energy output (host or device)

  // adjoints (vjp.hip): per-particle state and weights in tree order, the
  // recomputed neighbour lists, and the state of the step adjoint
  DBuf<double> vjp_c;                 // P/rho of the barotropic closure
  DBuf<double4> vjp_st;               // Omega, A, d2(rho_sum)/dh2, rho
  DBuf<double2> vjp_gdb, vjp_ahb;     // gather weights; Abar, hbar
  DBuf<double4> vjp_abar, vjp_grad;   // adjoint of the acceleration; pair part of the gradient
  DBuf<uint32_t> vjp_nl;
  DBuf<int32_t> vjp_tnl;
  DBuf<double> vjp_x0, vjp_x1, vjp_vh, vjp_a, vjp_t3, vjp_t3b;
  DBuf<float> vjp_dcv, vjp_alpha, vjp_gradh, vjp_dvdx;

  // work-queue counters of the persistent list kernels (8 per kernel)
  DBuf<int32_t> wq;
  int ncu = 0;  // compute units of the device

  // integrator reductions (block partials of the step loops)
  DBuf<double> step_red;
  // whole steps (substep.hip): the captured loop chains (0: predictor,
  // drift, predict_sph; 1: corrector), how often they were instantiated,
  // updated in place and launched
  hipGraphExec_t step_graph[2] = {nullptr, nullptr};
  int64_t step_graph_inst = 0, step_graph_upd = 0, step_graph_launch = 0;
  float *poten_out = nullptr;         // device view of poten_user during a force call

  // staging buffers for host arrays
  DBuf<char> stage[24];

  std::vector<TimedLaunch> timings;
  std::vector<std::string> timing_names;
  bool timing_enabled = true;
};

// ---- error helpers -----------------------------------------------------
#define PH_HIPCHECK(ctx, call)                                                     \
  do {                                                                             \
    hipError_t _e = (call);                                                        \
    if (_e != hipSuccess) {                                                        \
      (ctx)->err = std::string("HIP error ") + hipGetErrorString(_e) + " at " +    \
                   __FILE__ + ":" + std::to_string(__LINE__) + " (" #call ")";    \
      return PHSPH_ERR_HIP;                                                        \
    }                                                                              \
  } while (0)

#define PH_RET(ctx, code, msg)   \
  do {                           \
    (ctx)->err = (msg);          \
    return (code);               \
  } while (0)

// Launch wrapper: records a HIP event pair around the launch on the
// context's stream so that kernel time can be reported per call.
struct LaunchTimer {
  phsph_ctx *c;
  TimedLaunch tl;
  bool on;
  LaunchTimer(phsph_ctx *ctx, const char *name) : c(ctx), on(ctx->timing_enabled) {
    if (on) {
      tl.name = name;
      hipEventCreate(&tl.t0);
      hipEventCreate(&tl.t1);
      hipEventRecord(tl.t0, c->stream);
    }
  }
  ~LaunchTimer() {
    if (on) {
      hipEventRecord(tl.t1, c->stream);
      c->timings.push_back(tl);
    }
  }
};

void ctx_clear_timings(phsph_ctx *c);
// copies a few host bytes (a multiple of 4, at most 128) to device memory on
// the library stream with a one-block kernel whose arguments carry them:
// no host staging buffer, and a capturable kernel node instead of a copy
hipError_t dev_set(phsph_ctx *c, void *dst, const void *src, size_t bytes);
DevConst make_devconst(const phsph_params &p);

static inline unsigned nblocks(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// tree.hip
int tree_build(phsph_ctx *c, double *d_xyzh, int ldxyzh, int np, const int8_t *d_iphase);
// density.hip
struct DensIOArgs {  // device pointers in the caller's layout
  double *xyzh; int ldx;
  const double *vxyzu; int ldv;
  const double *fxyzu; int ldf;
  const double *fext;
  const double *Bevol;
  const int8_t *iphase;
  float *divcurlv; int lddcv;
  float *divcurlB;
  float *alphaind;
  float *gradh; int ldgradh;
  float *dvdx;
};
int dens_run(phsph_ctx *c, int icall, const DensIOArgs &a);
int dens_trial_stats(phsph_ctx *c);
int compact_active(phsph_ctx *c, const int32_t *flag, int n);
int pad_groups(phsph_ctx *c);
int dens_particle_stats(phsph_ctx *c, int32_t *d_nneigh, int32_t *d_nits);
// force.hip
struct ForceIOArgs {  // device pointers in the caller's layout
  const double *xyzh; int ldx;
  const double *vxyzu; int ldv;
  double *fxyzu; int ldf;
  float *divcurlv; int lddcv;
  const float *divcurlB;
  const double *Bevol;
  double *dBevol;
  float *divBsymm;
  const double *dustfrac;
  double *ddustevol, *deltav, *tstop;
  const float *gradh; int ldgradh;
  const float *alphaind;
  const double *eos_vars; int ldeos;
  const int8_t *iphase;
  // block timesteps (null ibin: global timestep)
  int8_t *ibin; const int8_t *ibin_old; int8_t *ibin_wake;
  int nbinmax, istepfrac, ibinnow;
  int np;  // length of the caller's particle arrays
};
int force_run(phsph_ctx *c, int icall, const ForceIOArgs &a, double *dtc, double *dtf);
int leaf_hmax_run(phsph_ctx *c, const double *d_h, double *d_hmaxn);
// step.hip
struct StepArgs {  // device pointers in the caller's layout
  double *xyzh; int ldx;
  double *vxyzu; int ldv;
  const double *fxyzu; int ldf;
  double *vpred;
  double *Bevol; const double *dBevol; double *Bpred;
  double *dustevol; const double *ddustevol; double *dustpred;
  const float *divcurlv; int lddcv;
  const double *eos_vars; int ldeos;
  float *alphaind;
  const int8_t *iphase;
  const int8_t *ibin_in; int8_t *ibin; int8_t *ibin_old; int8_t *ibin_wake;
  double *twas;
};
int step_init_run(phsph_ctx *c, int np, double time, int8_t *ibin, int nbinmax, double *twas);
int step_predictor_run(phsph_ctx *c, int np, double t, double dtsph, const StepArgs &s, double ufloor,
                       int32_t *nvfloor);
int step_drift_run(phsph_ctx *c, int np, double dt, double *xyzh, int ldx, const double *vxyzu, int ldv);
int step_predict_sph_run(phsph_ctx *c, int np, double timei, double dtsph, const StepArgs &s, double ufloor,
                         int32_t *nvfloor);
int step_corrector_run(phsph_ctx *c, int np, double timei, double dtsph, const StepArgs &s, int nbinmax,
                       double ufloor, double *errmax, double *v2sum, int32_t *npc, int32_t *nwake,
                       int32_t *nvfloor);
int step_reset_run(phsph_ctx *c, int np, double dtsph, const StepArgs &s);
// Enqueue-only forms of the step loops for whole steps (substep.hip): no
// timer, no synchronisation and no allocation, so they can be captured
// into a HIP graph; step_chain_prepare sizes the reduction buffer first.
// The reductions of the predictor, predict_sph and the corrector land in
// three consecutive result slots that step_chain_read copies back.
struct StepSums {
  double emax, v2;
  int np, nwake, nfloor, nbad;
};
int step_chain_prepare(phsph_ctx *c, int np);
void step_chain_pre(phsph_ctx *c, int np, double t, double dtsph, const StepArgs &pred, const StepArgs &psph,
                    double ufloor);
void step_chain_post(phsph_ctx *c, int np, double timei, double dtsph, const StepArgs &corr, int nbinmax,
                     double ufloor);
int step_chain_read(phsph_ctx *c, int np, StepSums *out3);
// substep.hip
int step_graph_run(phsph_ctx *c, const phsph_step_state &S, phsph_step_clock &k, int nsub, double ufloor,
                   double tolv, int maxits, bool graphs, phsph_step_summary *out);
int step_scratch_init(phsph_ctx *c);
int step_corrector_noalloc_run(phsph_ctx *c, int np, double timei, double dtsph, const StepArgs &s, int nbinmax,
                               double ufloor, double *errmax, double *v2sum, int32_t *npc, int32_t *nwake,
                               int32_t *nvfloor);
struct FRec;
struct FRec32;
int grav_tree(phsph_ctx *c);
int grav_hmax_replay(phsph_ctx *c, int icall);
int grav_force(phsph_ctx *c, int icall, const ForceIOArgs &a, const FRec *rec, const FRec32 *rec32,
               const float *gsoft, double *dtc_out, double *dtf_out);
int set_active_run(phsph_ctx *c, int np, const double *xyzh, int ldx, int8_t *iphase, const int8_t *ibin,
                   int nbinmax, int istepfrac, int32_t *nactive, int32_t *nalive);
// vjp.hip
int vjp_density_run(phsph_ctx *c, int np, const double *xyzh, int ldx, const double *rhobar, const double *hbar,
                    double *dxyz, int ldd);
int vjp_force_run(phsph_ctx *c, int np, const double *xyzh, int ldx, const double *vxyzu, int ldv,
                  const double *fbar, int ldfb, double *accel, int lda, double *dxyz, int ldd);
int vjp_step_run(phsph_ctx *c, int np, double dt, const double *xyzh, int ldx, const double *vxyzu, int ldv,
                 double *xyzh_new, int ldxn, double *vxyzu_new, int ldvn, const double *xbar, int ldxb,
                 const double *vbar, int ldvb, double *dx, int lddx, double *dv, int lddv);
int cons2prim_run(phsph_ctx *c, int np, const double *xyzh, int ldx, const double *vxyzu, int ldv,
                  const float *dvdx, float *alphaind, const double *dustevol, double *dustfrac,
                  double *eos, int ldeos);
