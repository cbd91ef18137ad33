/*
 * phsph.h -- C-ABI of the MI355X-native SPH hot path (kd-tree build, density
 * iteration, primitive-variable update and force evaluation).
 *
 * The entry points replace, one for one, the three calls that Phantom's
 * derivative driver makes on its hot path (src/main/deriv.f90:114,135,189):
 *
 *   phsph_maketree       <- maketree        (src/main/kdtree.F90:117-128,
 *                                            reached via build_tree,
 *                                            src/main/neigh_kdtree.f90:161)
 *   phsph_densityiterate <- densityiterate  (src/main/dens.F90:117-152)
 *   phsph_cons2prim      <- cons2prim_everything (src/main/cons2prim.f90:274)
 *   phsph_force          <- force           (src/main/force.F90:193-263)
 *   phsph_get_neighbour_stats <- get_neighbour_stats (src/main/dens.F90:1137)
 *
 * and the particle loops of the leapfrog step around it
 * (src/main/step_leapfrog.f90:91-760, src/main/substepping.F90:241-264):
 * phsph_step_predictor, phsph_substep_sph, phsph_step_predict_sph,
 * phsph_step_corrector, phsph_step_reset, phsph_init_step.
 *
 * Conventions
 *  - Every particle array is a Fortran column-major array a(ncomp, maxp)
 *    passed as a pointer to a(1,1) plus its leading dimension `ld` (>= ncomp).
 *    Particle i (0-based here, i+1 in Fortran) component c is a[i*ld + c].
 *  - Pointers may be host memory or HIP device memory; the library detects
 *    which (hipPointerGetAttributes) and stages host arrays through its own
 *    device buffers.  Device pointers must be on the context's device.
 *  - The storage kind of every array is the one the reference uses in its
 *    default FP64 build (src/main/part.F90:47-542): positions, velocities,
 *    B/rho, forces are double; divcurlv, divcurlB, alphaind, gradh, dvdx,
 *    divBsymm are float (real(kind=4)); iphase is int8.
 *  - Every call returns PHSPH_OK (0) or a negative error code; it never
 *    aborts the process.  phsph_last_error() returns the message of the most
 *    recent failure on a context.  The Fortran side maps a non-zero code to
 *    Phantom's fatal(), which keeps the reference's behaviour
 *    (src/main/io.F90:525-551).
 *  - Calls on one context are not re-entrant (the reference's module state is
 *    not either); each context owns one HIP stream and returns synchronously.
 *    That stream synchronises with the null stream, so device arrays the
 *    caller filled by work on the null stream are ready when the library's
 *    kernels start; work on other streams must be complete before the call.
 */
#ifndef PHSPH_H
#define PHSPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------- */
#define PHSPH_OK              0
#define PHSPH_ERR_INPUT      -1  /* invalid argument or parameter; negative h
                                    from hrho (dens.F90:1598); invalid dtc */
#define PHSPH_ERR_NAN        -2  /* NaN in particle positions (kdtree.F90:391) */
#define PHSPH_ERR_NOCONVERGE -3  /* h-rho iteration hit maxdensits (dens.F90:1443) */
#define PHSPH_ERR_CAPACITY   -4  /* tree depth / list capacity exceeded */
#define PHSPH_ERR_HIP        -5  /* a HIP runtime call failed */
#define PHSPH_ERR_STATE      -6  /* call order violated (e.g. force before tree) */
#define PHSPH_ERR_NODATA     -7  /* no live particles (kdtree.F90:162) */

/* ---- compile-time choices of the reference, made runtime here -------- */
#define PHSPH_KERNEL_CUBIC    0  /* kernel_cubic.f90   (M4, radkern 2) */
#define PHSPH_KERNEL_QUINTIC  1  /* kernel_quintic.f90 (M6, radkern 3) */

#define PHSPH_PREC_FP64       0  /* all pair arithmetic in FP64 */
#define PHSPH_PREC_MIXED      1  /* FP32 pair arithmetic, FP64 accumulation */

/*
 * Runtime parameters.  Each field names the reference variable it carries;
 * defaults are set by phsph_default_params() to the reference defaults
 * (options.f90:69-128, shock_capturing.f90:47-65, timestep.f90:52-62,
 * eos.f90:1888-1905, dust.f90:26-33, kdtree.F90:46).
 */
typedef struct phsph_params {
  int32_t kernel;              /* KERNEL= (build/Makefile:213-218)            */
  int32_t periodic;            /* dim:periodic                                */
  int32_t maxvxyzu;            /* 3 with ISOTHERMAL=yes, else 4               */
  int32_t mhd;                 /* dim:mhd (ideal MHD + div B cleaning)        */
  int32_t use_dustfrac;        /* one-fluid dust (options:use_dustfrac)       */
  int32_t ndustsmall;          /* part:ndustsmall (0 or 1 supported)          */
  int32_t gravity;             /* dim:gravity                                 */
  int32_t ind_timesteps;       /* dim:ind_timesteps                           */
  int32_t ieos;                /* eos:ieos, 1 (isothermal) or 2 (ideal gas)   */
  int32_t maxp;                /* dim:maxp; sets ncellsmax = 2*maxp           */
  int32_t precision;           /* PHSPH_PREC_*                                */
  int32_t idrag;               /* dust:idrag, 1 (Epstein/Stokes), 2 (const K) */
                               /* or 3 (const t_s)                             */
  int32_t ishock_heating;      /* eos:ishock_heating                          */
  int32_t ipdv_heating;        /* eos:ipdv_heating                            */
  int32_t iresistive_heating;  /* eos:iresistive_heating                      */
  int32_t ilimitdustflux;      /* dust:ilimitdustflux                         */
  double xmin, xmax, ymin, ymax, zmin, zmax;   /* boundary (boundary.f90:63) */
  double hfact;                /* part:hfact                                  */
  double tolh;                 /* part:tolh                                   */
  double massoftype_gas;       /* part:massoftype(igas)                       */
  double gamma;                /* eos:gamma                                   */
  double polyk;                /* eos:polyk (isothermal c_s^2)                */
  double alpha;                /* options:alpha (minimum, Cullen-Dehnen)      */
  double alphamax;             /* options:alphamax                            */
  double alphau;               /* options:alphau (conductivity)               */
  double alphaB;               /* options:alphaB (resistivity)                */
  double beta;                 /* options:beta                                */
  double C_cour, C_force;      /* timestep:C_cour, C_force                    */
  double dtmax;                /* timestep:dtmax                              */
  double psidecayfac;          /* timestep:psidecayfac                        */
  double overcleanfac;         /* timestep:overcleanfac                       */
  double K_code;               /* dust:K_code(1)                              */
  double tree_accuracy;        /* kdtree:tree_accuracy (opening angle)        */
  /* idrag = 1 (dust.f90:188-245): grain size and intrinsic density in code
   * units (part:grainsize(1) = grainsizecgs/udist, graindens(1) =
   * graindenscgs/unit_density) and the effective surface density of the
   * mean free path lambda = seff/rho_gas (dust:seff of init_drag,
   * dust.f90:94-97: pi/sqrt(2)*5/64 * (2 m_p/umass) / (sigma_H2/udist^2)) */
  double grainsize;
  double graindens;
  double seff;
} phsph_params;

/* Reference tree node, layout of type(kdnode) in a non-GRAVITY build
 * (src/main/dtype_kdtree.F90:53-68, sequence type, 64 bytes). */
typedef struct phsph_kdnode {
  double  xcen[3];
  double  size;
  double  hmax;
  double  mass;
  int32_t leftchild;
  int32_t rightchild;
  int32_t parent;
  int32_t idum;
} phsph_kdnode;

/* Reference tree node of a GRAVITY build (dtype_kdtree.F90:53-68 with the
 * GRAVITY fields, 120 bytes): the 64-byte node followed by tobecached,
 * cached (a default-kind logical, 4 bytes) and quads(6). */
typedef struct phsph_kdnode_grav {
  double  xcen[3];
  double  size;
  double  hmax;
  double  mass;
  int32_t leftchild;
  int32_t rightchild;
  int32_t parent;
  int32_t idum;
  int32_t tobecached;
  int32_t cached;
  double  quads[6];
} phsph_kdnode_grav;

typedef struct phsph_ctx phsph_ctx;

/* Fill *p with the reference defaults for a given kernel (hfact_default
 * from the kernel file, tolh 1e-4, alpha 0 / alphamax 1 / beta 2,
 * alphau 1, alphaB 1, C_cour 0.3, C_force 0.25, psidecayfac 1, ...). */
void phsph_default_params(phsph_params *p, int32_t kernel);

/* Create a context on HIP device `device`. */
int phsph_create(const phsph_params *p, int32_t device, phsph_ctx **ctx);
int phsph_destroy(phsph_ctx *ctx);
int phsph_set_params(phsph_ctx *ctx, const phsph_params *p);
const char *phsph_last_error(const phsph_ctx *ctx);

/*
 * maketree (kdtree.F90:117): build the kd-tree over the live particles
 * (h > 0) of xyzh(4,np).  xyzh is inout: periodic particles outside the box
 * are wrapped (construct_root_node -> cross_boundary, boundary.f90:123).
 * iphase(np) (may be NULL when every particle is an active gas particle)
 * marks active (>0) / inactive (<0) particles (part.F90:1013-1075).
 * Outputs (each may be NULL):
 *   leaf_is_active(ncellsmax+1): 1 active leaf, -1 inactive leaf, 0 other
 *   ncells: number of node slots (2^(maxlevel+1)-1 when indexed)
 * The node array, inoderange and inodeparts are retrieved with
 * phsph_get_tree.
 */
int phsph_maketree(phsph_ctx *ctx, double *xyzh, int32_t ldxyzh, int32_t np,
                   const int8_t *iphase, int32_t *leaf_is_active,
                   int64_t *ncells);

/*
 * Mirror the tree of the last phsph_maketree / phsph_densityiterate into
 * host arrays in the reference's layout (kdtree.F90:31-32,
 * neigh_kdtree.f90:30-36): node(ncells), inoderange(2,ncells) and
 * inodeparts(np) (1-based particle indices, negative = inactive).
 * Any pointer may be NULL.  `cap_nodes` is the length of node/inoderange.
 */
int phsph_get_tree(phsph_ctx *ctx, phsph_kdnode *node, int32_t *inoderange,
                   int32_t *inodeparts, int64_t cap_nodes, int32_t *maxlevel);

/* Second moments quads(6) = sum m dx dx (xx, xy, xz, yy, yz, zz) about each
 * node's centre of mass, for node ids 1..min(cap_nodes, ncells), row-major
 * (6 per node).  The reference stores them in node(i)%quads under GRAVITY
 * (kdtree.F90:724-750, dtype_kdtree.F90:66).  Unreachable node slots hold
 * unspecified values.  Only with params.gravity. */
int phsph_get_quads(phsph_ctx *ctx, double *quads, int64_t cap_nodes);

/* The node array of a GRAVITY build of the reference (maketree's node(:)
 * out-argument, dtype_kdtree.F90:53-68) for node ids
 * 1..min(cap_nodes, ncells): the fields of phsph_get_tree's nodes (hmax as
 * raised by the density relists), tobecached = 1 and cached = .false. as
 * construct_node leaves them (kdtree.F90:788-789), and quads as
 * phsph_get_quads.  Only with params.gravity. */
int phsph_get_tree_grav(phsph_ctx *ctx, phsph_kdnode_grav *node, int64_t cap_nodes);

/*
 * densityiterate (dens.F90:117-152).  icall 0 = one summation pass without
 * h iteration, 1 = converge h-rho (Newton-Raphson, tolh, clamp [0.8h,1.2h]).
 * xyzh(4,np) inout: h is replaced by hfact*(m/rho)^(1/3) of active
 * particles.  vxyzu(maxvxyzu,np), fxyzu(>=3,np), fext(3,np) (NULL = zero),
 * Bevol(4,np) (MHD only, else NULL).
 * Outputs (float): divcurlv(lddivcurlv,np) component 1 = div v;
 * divcurlB(4,np) (MHD); alphaind(3,np) component 3 = d(div v)/dt;
 * gradh(ldgradh,np) component 1 = 1/Omega; dvdx(9,np).
 * The tree of the preceding phsph_maketree must still be current.
 */
int phsph_densityiterate(phsph_ctx *ctx, int32_t icall, int32_t np,
                         double *xyzh, int32_t ldxyzh,
                         const double *vxyzu, int32_t ldvxyzu,
                         const double *fxyzu, int32_t ldfxyzu,
                         const double *fext, const double *Bevol,
                         const int8_t *iphase,
                         float *divcurlv, int32_t lddivcurlv,
                         float *divcurlB, float *alphaind,
                         float *gradh, int32_t ldgradh, float *dvdx);

/*
 * cons2prim_everything (cons2prim.f90:274-456) for ieos 1 and 2: fills
 * eos_vars(ldeos,np) components 1..3 (P, c_s, T placeholder 0), the dust
 * fraction dustfrac(1,np) = s^2/(1+s^2) from dustevol(1,np) (one-fluid
 * dust only, else both NULL) and alphaind(2,:) = alpha_loc from
 * alphaind(3,:) and dvdx (Cullen-Dehnen, shock_capturing.f90:131-178).
 */
int phsph_cons2prim(phsph_ctx *ctx, int32_t np, const double *xyzh,
                    int32_t ldxyzh, const double *vxyzu, int32_t ldvxyzu,
                    const float *dvdx, float *alphaind,
                    const double *dustevol, double *dustfrac,
                    double *eos_vars, int32_t ldeos, const int8_t *iphase);

/*
 * force (force.F90:193-263) for hydro + Cullen-Dehnen viscosity +
 * conductivity, ideal MHD with div B cleaning and one-fluid dust.
 * Inputs: xyzh, vxyzu, gradh (1/Omega, float), alphaind (alpha in comp. 1),
 * eos_vars (P, c_s), divcurlB (MHD), Bevol (MHD), dustfrac (dust).
 * Outputs: fxyzu(ldfxyzu,np) (du/dt in component 4 when maxvxyzu = 4),
 * divcurlv(1,:) = div v from the force loop (float), dBevol(4,np) (MHD),
 * divBsymm(np) (MHD, float), ddustevol(1,np), deltav(3,np), tstop(np) (dust),
 * and the global timestep constraints dtcourant, dtforce (global-timestep
 * build, force.F90:3312-3327).  Arrays for disabled physics may be NULL.
 */
int phsph_force(phsph_ctx *ctx, int32_t icall, int32_t np,
                const double *xyzh, int32_t ldxyzh,
                const double *vxyzu, int32_t ldvxyzu,
                double *fxyzu, int32_t ldfxyzu,
                float *divcurlv, int32_t lddivcurlv,
                const float *divcurlB, const double *Bevol, double *dBevol,
                float *divBsymm,
                const double *dustfrac, double *ddustevol, double *deltav,
                double *tstop,
                const float *gradh, int32_t ldgradh, const float *alphaind,
                const double *eos_vars, int32_t ldeos, const int8_t *iphase,
                double *dtcourant, double *dtforce);

/*
 * Self-gravity potential energy output.  The reference accumulates
 * poten(i) = epoti in force, with epoti = 0.5*m*(phi_pairs + m*potensoft0/h)
 * + 0.5*m*phi_tree (force.F90:2912-2927); it is module state of part, so the array
 * (np floats, host or device) is registered once and written by every
 * phsph_force call while params.gravity is set.  Passing NULL stops the
 * output.  The array must stay valid until the next phsph_set_poten or
 * phsph_destroy.
 */
int phsph_set_poten(phsph_ctx *ctx, float *poten);

/*
 * Block (individual) timesteps.  The reference keeps the bins as module
 * state (part:ibin, ibin_old, ibin_wake; timestep_ind:nbinmax, istepfrac,
 * ibinnow); here they are registered on the context and used by every
 * phsph_force call while params.ind_timesteps is set.  The arrays must stay
 * valid until the next phsph_set_bins or phsph_destroy.  phsph_force then,
 * per active particle (force.F90:3272-3310): takes dti = min of the force,
 * cleaning and dust timesteps if below dtc and dtmax, places it with
 * get_newbin (utils_indtimesteps.f90:230-287, decreases allowed when
 * icall < 2), applies the Saitoh-Makino limiter ibin >= max(ibin_old of the
 * neighbours) - 1, and flags every neighbour j with
 * ibin_wake(j) = max(ibinnow-1, ibin_wake(j)) (force.F90:1346-1352).
 * dtcourant and dtforce are returned as 1e29, as in the reference.
 */
int phsph_set_bins(phsph_ctx *ctx, int8_t *ibin, const int8_t *ibin_old,
                   int8_t *ibin_wake, int32_t nbinmax, int32_t istepfrac,
                   int32_t ibinnow);

/* nbinmax after the last phsph_force with block timesteps: the largest
 * bin of its active particles (force.F90:792). */
int phsph_get_nbinmax(phsph_ctx *ctx, int32_t *nbinmax);

/*
 * set_active_particles (utils_indtimesteps.f90:114-185): a live particle
 * (h > 0) is active when mod(istepfrac, 2**(nbinmax - ibin)) == 0; iphase is
 * set to +1 (active) or -1 (inactive) for gas.  Returns the counts and
 * ibinnow, the largest active bin.  A bin above nbinmax is an error.
 */
int phsph_set_active_particles(phsph_ctx *ctx, int32_t np, const double *xyzh,
                               int32_t ldxyzh, int8_t *iphase,
                               const int8_t *ibin, int32_t nbinmax,
                               int32_t istepfrac, int32_t *nactive,
                               int32_t *nalive, int32_t *ibinnow);

/*
 * Leapfrog integrator.  The reference's step (step_leapfrog.f90:91-760) is
 * a sequence of particle loops around derivs; each loop is one entry point
 * here, so that a caller keeps the reference's control flow (derivs, the
 * velocity-error check and its iterations) and moves only the loops to the
 * device.  Arrays follow the other entry points: Fortran a(ld, np), host or
 * device pointers, NULL for physics that is off.  Only gas particles are
 * modelled: iphase (NULL = all active) marks active (>0) and inactive (<0)
 * gas, and a particle is dead when h < tiny(h) (part.F90:931-941).  Bevol,
 * dBevol and Bpred have 4 components (MHD), dustevol, ddustevol and dustpred
 * 1 (one-fluid dust); vpred has the layout of vxyzu.  Counts returned
 * through pointers may be NULL.  dtmax is params.dtmax.
 */

/*
 * init_step (step_leapfrog.f90:44-83) with block timesteps: when
 * time < tiny(time) every ibin is first set to nbinmax; then
 * twas(i) = time + 0.5*dtmax/2**ibin(i).  An error without
 * params.ind_timesteps.
 */
int phsph_init_step(phsph_ctx *ctx, int32_t np, double time, int8_t *ibin,
                    int32_t nbinmax, double *twas);

/*
 * The predictor loop of step (step_leapfrog.f90:150-237): for each live
 * particle v,u += hdti*f with hdti = dtsph/2 (global timestep) or
 * twas(i) - t (block timesteps, where active particles also set
 * ibin_old(i) = ibin(i)); u is floored to ufloor when ufloor > 0;
 * Bevol += hdti*dBevol; dustevol = |dustevol + hdti*ddustevol|.
 * ibin, ibin_old and twas are read only with block timesteps.
 */
int phsph_step_predictor(phsph_ctx *ctx, int32_t np, double t, double dtsph,
                         const double *xyzh, int32_t ldxyzh,
                         double *vxyzu, int32_t ldvxyzu,
                         const double *fxyzu, int32_t ldfxyzu,
                         double *Bevol, const double *dBevol,
                         double *dustevol, const double *ddustevol,
                         const int8_t *iphase, const int8_t *ibin,
                         int8_t *ibin_old, const double *twas, double ufloor,
                         int32_t *nvfloor);

/* substep_sph (substepping.F90:241-264): x += dt*v for live particles, the
 * drift of step when there are no external forces or sinks
 * (step_leapfrog.f90:279). */
int phsph_substep_sph(phsph_ctx *ctx, int32_t np, double dt, double *xyzh,
                      int32_t ldxyzh, const double *vxyzu, int32_t ldvxyzu);

/*
 * The predict_sph loop of step (step_leapfrog.f90:307-400) at
 * timei = t + dtsph: h -= dtsph*dhdrho*rho*div v; vpred = v + hdti*f with
 * hdti = dtsph/2 or timei - twas(i); vpred(4) floored to ufloor;
 * Bpred = Bevol + hdti*dBevol; dustpred = dustevol + hdti*ddustevol; and the
 * Cullen-Dehnen switch alphaind(1) = alpha_loc if alpha_loc is larger, else
 * relaxed towards it with decay time h/(0.1 c_s) (implicit step).
 * divcurlv(1,:) is div v, eos_vars(2,:) the sound speed, alphaind(2,:)
 * alpha_loc.
 */
int phsph_step_predict_sph(phsph_ctx *ctx, int32_t np, double timei,
                           double dtsph, double *xyzh, int32_t ldxyzh,
                           const double *vxyzu, int32_t ldvxyzu,
                           const double *fxyzu, int32_t ldfxyzu,
                           double *vpred, const double *Bevol,
                           const double *dBevol, double *Bpred,
                           const double *dustevol, const double *ddustevol,
                           double *dustpred, const float *divcurlv,
                           int32_t lddivcurlv, const double *eos_vars,
                           int32_t ldeos, float *alphaind,
                           const int8_t *iphase, const double *twas,
                           double ufloor, int32_t *nvfloor);

/*
 * One pass of the corrector loop of step (step_leapfrog.f90:482-625) at
 * timei = t + dtsph.  Global timestep: v,u += dtsph/2*f, Bevol and dustevol
 * likewise, and the velocity-error sums the reference passes to
 * check_velocity_error: errmax = max |v - vpred|^2, v2sum = sum |v|^2 and
 * npc, the number of particles summed.  Block timesteps: active particles
 * complete their step (dti = timei - twas + dt(ibin)/2, twas += dti,
 * ibin_wake = 0), every particle is synchronised to timei, u is floored, and
 * particles flagged for waking move to bin min(ibin_wake, nbinmax) with twas
 * at the half step of that bin (nwake counts them); errmax, v2sum and npc are
 * then 0.  ibin, ibin_wake and twas are used only with block timesteps.
 */
int phsph_step_corrector(phsph_ctx *ctx, int32_t np, double timei,
                         double dtsph, const double *xyzh, int32_t ldxyzh,
                         double *vxyzu, int32_t ldvxyzu,
                         const double *fxyzu, int32_t ldfxyzu,
                         const double *vpred, double *Bevol,
                         const double *dBevol, double *dustevol,
                         const double *ddustevol, const int8_t *iphase,
                         int8_t *ibin, int8_t *ibin_wake, int32_t nbinmax,
                         double *twas, double ufloor, double *errmax,
                         double *v2sum, int32_t *npc, int32_t *nwake,
                         int32_t *nvfloor);

/*
 * phsph_step_corrector without allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during the
 * call (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, results and errors as
 * phsph_step_corrector, except that every array must already be device
 * memory (PHSPH_ERR_INPUT otherwise: nothing is staged) and np is at most
 * 16,777,216 (PHSPH_ERR_CAPACITY above).  The reduction partials live in a
 * fixed 2 MiB device array of the library, one per GPU, held until the
 * process ends and looked up by phsph_create; the call makes no hipMalloc
 * or hipFree and records no kernel timings.  Calls that overlap on one GPU
 * (contexts on different streams) share that array and race.
 */
int phsph_step_corrector_noalloc(phsph_ctx *ctx, int32_t np, double timei,
                                 double dtsph, const double *xyzh, int32_t ldxyzh,
                                 double *vxyzu, int32_t ldvxyzu,
                                 const double *fxyzu, int32_t ldfxyzu,
                                 const double *vpred, double *Bevol,
                                 const double *dBevol, double *dustevol,
                                 const double *ddustevol, const int8_t *iphase,
                                 int8_t *ibin, int8_t *ibin_wake, int32_t nbinmax,
                                 double *twas, double ufloor, double *errmax,
                                 double *v2sum, int32_t *npc, int32_t *nwake,
                                 int32_t *nvfloor);

/*
 * The until_converged loop of step (step_leapfrog.f90:650-696), run when
 * the velocity error asks for another iteration: vpred = v, Bpred = Bevol,
 * dustpred = dustevol (block timesteps: active particles only) and, with a
 * global timestep, v, Bevol, dustevol -= dtsph/2 * their derivatives.  As
 * in the reference this loop also runs over dead particles.
 */
int phsph_step_reset(phsph_ctx *ctx, int32_t np, double dtsph,
                     double *vxyzu, int32_t ldvxyzu, const double *fxyzu, int32_t ldfxyzu,
                     double *vpred, double *Bevol, const double *dBevol,
                     double *Bpred, double *dustevol, const double *ddustevol,
                     double *dustpred, const int8_t *iphase);

/*
 * Whole steps on device-resident state.  phsph_step_graph runs `nsub`
 * steps of the reference's evolve loop with the control flow inside the
 * library (no host round trip per loop):
 *  - block timesteps (params.ind_timesteps): per substep evol_prestep's
 *    istepfrac and set_active_particles (evolve.f90:200-222,
 *    utils_indtimesteps.f90:114-185), step (step_leapfrog.f90:91-760) at
 *    dt = dtmax/2**nbinmax, then update_time_and_dt's time and
 *    change_nbinmax (evolve_utils.F90, utils_indtimesteps.f90:186-223);
 *  - global timestep: step at clock->dt, then dt = min(dtforce, dtcourant,
 *    dterr, dtmax + epsilon(dtmax)) (evolve.f90:105-160).
 * step is predictor, substep_sph drift, predict_sph, derivs (maketree,
 * densityiterate, cons2prim, force: deriv.f90:114-189), then the corrector
 * with check_velocity_error (step_leapfrog.f90:769-843) and, while it
 * fails, the reset and derivs(icall = 2), up to `maxits` iterations
 * (PHSPH_ERR_NOCONVERGE when they do not converge and maxits > 1).
 * With `graphs` = 1 the loops without data-dependent control (predictor,
 * drift and predict_sph with their reductions; the corrector with its
 * reduction) are captured as HIP graphs, replayed with the step's times
 * through hipGraphExecUpdate; with 0 the same kernels are launched
 * directly.  Both give bit-identical states, equal to those of the
 * individual entry points called in the same order.
 * Every array must be device memory (PHSPH_ERR_INPUT otherwise); arrays of
 * physics that is off may be NULL.  The state must hold the derivatives of
 * the current time (the first derivs and init_step of the reference's
 * initial.F90 are made through the individual entry points).  `clock` is
 * read and updated in place; `out` (room for nsub) receives one summary
 * per step and may be NULL.  No reference counterpart as one call: it is
 * the reference's evolve loop restricted to the SPH path (no sinks,
 * external forces, cooling or dumps).
 */
typedef struct phsph_step_state {
  double *xyzh, *vxyzu, *fxyzu, *vpred;   /* xyzh(4,:), vxyzu/fxyzu/vpred(ld,:) */
  double *Bevol, *dBevol, *Bpred;         /* (4,:), MHD */
  double *dustevol, *ddustevol, *dustpred, *dustfrac, *deltav, *tstop; /* one-fluid dust */
  float *divcurlv, *divcurlB, *divBsymm, *alphaind, *gradh, *dvdx;
  double *eos_vars;                       /* (3,:) */
  float *poten;                           /* self-gravity potential, may be NULL */
  int8_t *iphase, *ibin, *ibin_old, *ibin_wake;  /* block timesteps; iphase NULL: all active */
  double *twas;
  int32_t np, ldxyzh, ldvxyzu, ldfxyzu, lddivcurlv, ldgradh, ldeos, pad;
} phsph_step_state;

typedef struct phsph_step_clock {
  double time, tlast;        /* evolve.f90 time and tlast (block timesteps) */
  double dt;                 /* global timestep: dt of the next step */
  double dtcourant, dtforce; /* of the last force call (global timestep) */
  int32_t nbinmax, istepfrac, ibinnow, pad;
} phsph_step_clock;

typedef struct phsph_step_summary {
  double dt, dtnew, dterr;   /* step dt, derivs' dtnew, timestep:dterr */
  int32_t its, nactive, nwake, nvfloorp, nvfloorps, nvfloorc;
  int32_t ngraph, ninst;     /* graph launches of the step; of them, freshly instantiated */
} phsph_step_summary;

int phsph_step_graph(phsph_ctx *ctx, const phsph_step_state *state,
                     phsph_step_clock *clock, int32_t nsub, double ufloor,
                     double tolv, int32_t maxits, int32_t graphs,
                     phsph_step_summary *out);

/*
 * Configuration of the neighbour-list kernels, picked by the library's
 * autotuner (no counterpart in the reference): candidate leaves staged per
 * chunk `cl` (8 or 16), wavefronts per block `lw` (2 or 4), `r2` = 1 to
 * store each list entry's squared separation for the h iteration (0:
 * recomputed from the positions), and `q` = 1 to run the density list
 * kernel as persistent wavefronts drawing groups of active particles from
 * per-XCD work queues (0: one group per wavefront).  Every configuration
 * gives bit-identical results.  By default the first densityiterate of a
 * context with at least 32768 active particles times each configuration
 * once on the caller's own particles and keeps the fastest.
 * phsph_set_tuning fixes a configuration (cl = 0 asks for a new tuning on
 * the next large call); phsph_get_tuning returns the current one, whether
 * it was tuned (tuned = 1) and, if `ms`/`cfg` are given (room for 16), the
 * time of each configuration tried (cfg: cl, lw, r2, q per configuration)
 * and their count.  `rfac` receives the first search radius the next
 * densityiterate uses, in units of radkern*h at entry (1.03, 1.05 or 1.1):
 * it is adapted after every call, one level narrower while no particle
 * outgrows its list, one level wider (and not narrower again) when more
 * than 1/200 of them do.  Neighbour and iteration counts do not depend on
 * it; the sums agree to round-off (the candidate leaves come in another
 * order).
 */
int phsph_set_tuning(phsph_ctx *ctx, int32_t cl, int32_t lw, int32_t r2, int32_t q);
int phsph_get_tuning(phsph_ctx *ctx, int32_t *cl, int32_t *lw, int32_t *r2, int32_t *q,
                     int32_t *tuned, float *ms, int32_t *cfg, int32_t *ntried, double *rfac);

/* get_neighbour_stats (dens.F90:1137-1155) of the last densityiterate. */
int phsph_get_neighbour_stats(phsph_ctx *ctx, double *trialmean,
                              double *actualmean, int32_t *maxtrial,
                              int32_t *maxactual, int64_t *nrhocalc,
                              int64_t *nactualtot);

/*
 * Per-particle statistics of the last densityiterate, in the caller's
 * particle order: the actual neighbour count including the particle itself
 * (cell%nneigh, dens.F90:1267) and the number of the h iteration at which
 * the particle's own h converged (dens.F90:1425-1440).  Inactive and dead
 * particles get 0.  Either pointer may be NULL.
 */
int phsph_get_particle_stats(phsph_ctx *ctx, int32_t np, int32_t *nneigh,
                             int32_t *nits);

/*
 * Number of interacting pairs (j with q_i < R or q_j < R, force.F90:1287)
 * summed over the active particles of the last force call.
 */
int phsph_get_force_stats(phsph_ctx *ctx, int64_t *npairs);

/*
 * Adjoints with respect to positions.  Phantom has no adjoint: these are
 * the vector-Jacobian products of its forward sums, for gradient-based
 * uses of the hot path (relaxation of initial conditions to a target
 * density, sensitivities).  Hydrodynamics only (no MHD, dust or
 * self-gravity).  The h-rho constraint rho_i = m (hfact/h_i)^3 of
 * densityiterate (dens.F90:1382-1466) is differentiated implicitly, so h
 * must be converged (a tight tolh makes the gradient that of the discrete
 * solution).  Neighbour lists are recomputed from the current tree; no
 * pair data is stored between the forward and the adjoint.  Arrays follow
 * the other entry points (host or device, a(ld, np)); gradients dxyz are
 * (3, np) with leading dimension ldd.
 */

/*
 * Adjoint of densityiterate's density summation (get_density_sums,
 * dens.F90:578-857): dxyz_k = sum_i [rhobar_i drho_i/dx_k + hbar_i dh_i/dx_k]
 * with rho_i = m (hfact/h_i)^3 at the converged h.  xyzh: the positions and
 * h after phsph_maketree + phsph_densityiterate(icall 1) on them.  rhobar,
 * hbar: (np) adjoints of rho and h (either may be NULL = 0).
 */
int phsph_density_vjp(phsph_ctx *ctx, int32_t np, const double *xyzh, int32_t ldxyzh,
                      const double *rhobar, const double *hbar, double *dxyz, int32_t lddxyz);

/*
 * The pressure acceleration of force (force.F90:1595-1596 without
 * viscosity, conductivity or other physics: a_i = -sum_j m [P_i/(Omega_i
 * rho_i^2) dW_i + P_j/(Omega_j rho_j^2) dW_j] r_hat, with Omega in FP64)
 * for a barotropic closure P = polyk rho (ieos 1) or P = (gamma-1) u rho
 * with u held fixed (ieos 2, u from vxyzu), and its adjoint
 * dxyz = d/dx sum_i fbar_i . a_i through the separations, h(x) and
 * Omega(x).  xyzh as for phsph_density_vjp.  fxyz(ldfxyz, np) (may be
 * NULL) receives a; fbar (ldfbar >= 3, may be NULL: forward only).
 */
int phsph_force_vjp(phsph_ctx *ctx, int32_t np, const double *xyzh, int32_t ldxyzh,
                    const double *vxyzu, int32_t ldvxyzu, const double *fbar, int32_t ldfbar,
                    double *fxyz, int32_t ldfxyz, double *dxyz, int32_t lddxyz);

/*
 * One leapfrog step density -> force -> kick (step_leapfrog.f90:150-625
 * with the pressure acceleration above): v_half = v + dt/2 a(x),
 * x' = x + dt v_half, then maketree + densityiterate at x' and
 * v' = v_half + dt/2 a(x').  Forward: xyzh_new (positions and converged h
 * at x'), vxyzu_new (u copied), either may be NULL.  Adjoint, when xbar or
 * vbar is given ((3, np), ld ldxbar / ldvbar; NULL = 0): dxyz and dvxyz
 * (ld lddxyz) receive dL/dx and dL/dv of L = xbar.x' + vbar.v'.  The tree
 * and h at x and x' are rebuilt for the adjoint instead of stored.  xyzh's
 * h is the start of the density iteration at x.  The context's tree is
 * left at x (after an adjoint) or x' (forward only).
 */
int phsph_step_vjp(phsph_ctx *ctx, int32_t np, double dt, const double *xyzh, int32_t ldxyzh,
                   const double *vxyzu, int32_t ldvxyzu, double *xyzh_new, double *vxyzu_new,
                   const double *xbar, int32_t ldxbar, const double *vbar, int32_t ldvbar,
                   double *dxyz, double *dvxyz, int32_t lddxyz);

/*
 * Precision (PHSPH_PREC_*) the pair sums of the last phsph_densityiterate
 * and phsph_force actually ran in; -1 before the first call.  With
 * params.precision = PHSPH_PREC_MIXED the hydro and self-gravity sums run
 * in mixed precision (the multipole expansions stay FP64); the density sums
 * with MHD, and the force pair terms with MHD or dust, have FP64 kernels
 * only and report PHSPH_PREC_FP64.
 * No reference counterpart (the reference is FP64 throughout).
 */
int phsph_get_precision(phsph_ctx *ctx, int32_t *dens_prec, int32_t *force_prec);

/*
 * Kernel timing of the last call on the context's stream, measured with
 * HIP events around each launch: up to `cap` (name, milliseconds) entries
 * are written; returns the number of entries available.
 */
int phsph_get_timings(phsph_ctx *ctx, const char **names, double *ms,
                      int32_t cap);

#ifdef __cplusplus
}
#endif

#endif /* PHSPH_H */
