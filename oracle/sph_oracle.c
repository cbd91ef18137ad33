/*
 * oracle/sph_oracle.c -- CPU restatement (plain C, FP64) of Phantom's SPH
 * hot path: kd-tree build, tree walk, density iteration, cons2prim and the
 * force pair loop (hydro, Cullen-Dehnen viscosity, conductivity, ideal MHD
 * with div B cleaning, one-fluid dust).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker of the HIP
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it.  The product library never links or calls it.
 *
 * Semantics follow the reference executed with one OpenMP thread (the only
 * execution whose tree is independent of the thread count), cell by cell
 * in node order.  One deliberate difference: in the div B sum the density
 * rho_j = rhoh(h_j) of a neighbour uses the smoothing length it had when
 * densityiterate was entered.  The reference reads h_j while other cells
 * are overwriting it (an acknowledged race, config.F90:215-217,
 * dens.F90:810-811), so its value there depends on the cell schedule.
 *
 * Each routine cites the reference lines it restates.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../include/phsph.h"

#define EPS64 2.220446049250313e-16      /* epsilon(1d0)                  */
#define TINY64 2.2250738585072014e-308   /* tiny(1d0)                     */
#define HUGE64 1.7976931348623157e308    /* huge(1d0)                     */
#define BIGNUMBER 1.e29                  /* timestep.f90:40               */
#define PI_PH 3.1415926536               /* physcon.f90:23 (sic)          */
#define MAXDENSITS 100                   /* dens.F90:101                  */
#define ISIZECELLCACHE 1000              /* dens.F90:99, force.F90:56     */
#define MINPART 10                       /* config.F90:113                */
#define ISTACKSIZE 512                   /* kdtree.F90:131                */
#define MAXBINS 30                       /* utils_indtimesteps.f90:29     */

/* rhosum indices (dens.F90:51-93), 0-based */
enum { R_RHO = 0, R_GRADH, R_GRADSOFT, R_DIVV,
       R_DVXDX, R_DVXDY, R_DVXDZ, R_DVYDX, R_DVYDY, R_DVYDZ, R_DVZDX, R_DVZDY, R_DVZDZ,
       R_DAXDX, R_DAXDY, R_DAXDZ, R_DAYDX, R_DAYDY, R_DAYDZ, R_DAZDX, R_DAZDY, R_DAZDZ,
       R_RXX, R_RXY, R_RXZ, R_RYY, R_RYZ, R_RZZ,
       R_DIVB, R_DBXDX, R_DBXDY, R_DBXDZ, R_DBYDX, R_DBYDY, R_DBYDZ, R_DBZDX, R_DBZDY, R_DBZDZ,
       NRHOSUM };

typedef struct {
  phsph_params p;
  /* kernel constants (kernel_cubic.f90:23-30, kernel_quintic.f90:23-30) */
  double radkern, radkern2, cnormk, wab0, gradh0, dphidh0;
  double dxbound, dybound, dzbound, hdlx, hdly, hdlz;
  int32_t np;
  int64_t ncellsmax;
  /* tree (kdtree.F90:31-36, neigh_kdtree.f90:30-36), 1-based node index */
  phsph_kdnode *node;
  int32_t *leaf_is_active;   /* [ncellsmax+2] */
  int32_t *ir_lo, *ir_hi;    /* inoderange(1:2, n), 1-based positions   */
  int32_t *inodeparts;       /* [np+1], 1-based, signed                 */
  double *treecache;         /* 5 per position, position p at 5*(p-1)   */
  int64_t ncells;
  int32_t maxlevel, minlevel, maxlevel_indexed;
  int tree_built;
  /* neighbour statistics (dens.F90:106-108) */
  int64_t nneightry, nneighact, maxneightry, ncalc, nptot;
  int32_t maxneighact;
  /* longest neighbour list of the last force call (cells' getneigh lists) */
  int64_t maxforcetrial;
  /* block timesteps (part:ibin, ibin_old, ibin_wake; timestep_ind module) */
  int8_t *ibin, *ibin_wake;
  const int8_t *ibin_old;
  /* self-gravity: node second moments quads(6) about the centre of mass
   * (kdtree.F90:724-750), 6 per node slot; poten output (part:poten) */
  double *quads;
  float *poten;
  int32_t nbinmax, istepfrac, ibinnow, nbinmaxnew;
  char err[256];
} orc_ctx;

/* ------------------------------------------------------------------ */
/* kernels                                                             */
/* ------------------------------------------------------------------ */

/* get_kernel (kernel_cubic.f90:34-51, kernel_quintic.f90:34-56) */
static inline void get_kernel(int kind, double q2, double q, double *w, double *g) {
  if (kind == PHSPH_KERNEL_CUBIC) {
    if (q < 1.) {
      *w = 0.75 * q2 * q - 1.5 * q2 + 1.;
      *g = q * (2.25 * q - 3.);
    } else if (q < 2.) {
      double t = q - 2.;
      *w = -0.25 * (t * t * t);
      *g = -0.75 * (t * t);
    } else { *w = 0.; *g = 0.; }
  } else {
    if (q < 1.) {
      double q4 = q2 * q2;
      *w = -10. * q4 * q + 30. * q4 - 60. * q2 + 66.;
      *g = q * (-50. * q2 * q + 120. * q2 - 120.);
    } else if (q < 2.) {
      double a = q - 3., b = q - 2.;
      double a2 = a * a, b2 = b * b;
      *w = -(a2 * a2 * a) + 6. * (b2 * b2 * b);
      *g = -5. * (a2 * a2) + 30. * (b2 * b2);
    } else if (q < 3.) {
      double a = q - 3.;
      double a2 = a * a;
      *w = -(a2 * a2 * a);
      *g = -5. * (a2 * a2);
    } else { *w = 0.; *g = 0.; }
  }
}

/* grkern (kernel_cubic.f90:67-78, kernel_quintic.f90:75-88) */
static inline double grkern(int kind, double q2, double q) {
  double w, g;
  get_kernel(kind, q2, q, &w, &g);
  return g;
}

/* get_kernel_grav1 dphidh term (kernel_cubic.f90:80-101,
 * kernel_quintic.f90:90-122) */
static inline double dphidh_kern(int kind, double q2, double q) {
  double q4 = q2 * q2;
  if (kind == PHSPH_KERNEL_CUBIC) {
    if (q < 1.) return -0.6 * q4 * q + 1.5 * q4 - 2. * q2 + 1.4;
    if (q < 2.) return 0.2 * q4 * q - 1.5 * q4 + 4. * q2 * q - 4. * q2 + 1.6;
    return 0.;
  } else {
    double q6 = q4 * q2;
    if (q < 1.) return q6 * q / 21. - q6 / 6. + q4 / 2. - 11. * q2 / 10. + 239. / 210.;
    if (q < 2.) return -q6 * q / 42. + q6 / 4. - q4 * q + 7. * q4 / 4. - 5. * q2 * q / 6. - 17. * q2 / 20. + 473. / 420.;
    if (q < 3.) return q6 * q / 210. - q6 / 12. + 3. * q4 * q / 5. - 9. * q4 / 4. + 9. * q2 * q / 2. - 81. * q2 / 20. + 243. / 140.;
    return 0.;
  }
}

/* ------------------------------------------------------------------ */
/* particle helpers (part.F90:779-858)                                 */
/* ------------------------------------------------------------------ */
static inline double rhoh(const orc_ctx *o, double h, double m) {
  double t = o->p.hfact / fabs(h);
  return m * (t * t * t);
}
static inline double dhdrho(const orc_ctx *o, double h, double m) {
  return -h / (3. * rhoh(o, h, m));
}
static inline double hrho8(const orc_ctx *o, double rho, double m) {
  return o->p.hfact * pow(m / fabs(rho), 1.0 / 3.0);
}

/* ------------------------------------------------------------------ */
/* context                                                             */
/* ------------------------------------------------------------------ */
static void set_derived(orc_ctx *o) {
  if (o->p.kernel == PHSPH_KERNEL_CUBIC) {
    o->radkern = 2.; o->radkern2 = 4.; o->cnormk = 1. / PI_PH;
    o->wab0 = 1.; o->gradh0 = -3.; o->dphidh0 = 1.4;
  } else {
    o->radkern = 3.; o->radkern2 = 9.; o->cnormk = 1. / (120. * PI_PH);
    o->wab0 = 66.; o->gradh0 = -3. * 66.; o->dphidh0 = 239. / 210.;
  }
  /* set_boundary (boundary.f90:103-108) */
  o->dxbound = o->p.xmax - o->p.xmin;
  o->dybound = o->p.ymax - o->p.ymin;
  o->dzbound = o->p.zmax - o->p.zmin;
  o->hdlx = 0.5 * o->dxbound; o->hdly = 0.5 * o->dybound; o->hdlz = 0.5 * o->dzbound;
}

void orc_default_params(phsph_params *p, int32_t kernel) {
  memset(p, 0, sizeof(*p));
  p->kernel = kernel;
  p->periodic = 1;
  p->maxvxyzu = 4;
  p->ieos = 2;
  p->idrag = 2;
  p->ishock_heating = 1; p->ipdv_heating = 1; p->iresistive_heating = 1;
  p->xmin = -0.5; p->xmax = 0.5; p->ymin = -0.5; p->ymax = 0.5; p->zmin = -0.5; p->zmax = 0.5;
  p->hfact = (kernel == PHSPH_KERNEL_CUBIC) ? 1.2 : 1.0;
  p->tolh = 1.e-4;
  p->massoftype_gas = 1.;
  p->gamma = 5. / 3.; p->polyk = 1.;
  p->alpha = 0.; p->alphamax = 1.; p->alphau = 1.; p->alphaB = 1.; p->beta = 2.;
  p->C_cour = 0.3; p->C_force = 0.25; p->dtmax = 1.e29;
  p->psidecayfac = 1.; p->overcleanfac = 1.;
  p->K_code = 1.;
  p->tree_accuracy = 0.5;
  /* grainsizecgs, graindenscgs (dust.f90:38-39) with udist = umass = 1 and
   * seff of init_drag (dust.f90:94-97) in those units */
  p->grainsize = 0.1;
  p->graindens = 3.;
  p->seff = PI_PH / sqrt(2.) * 5. / 64. * ((2. * 1.67262158e-24) / 1.) / (2.367e-15 / (1. * 1.));
}

orc_ctx *orc_create(const phsph_params *p) {
  orc_ctx *o = (orc_ctx *)calloc(1, sizeof(orc_ctx));
  o->p = *p;
  set_derived(o);
  o->nptot = -1;
  return o;
}

static void free_tree(orc_ctx *o) {
  free(o->quads); o->quads = NULL;
  free(o->node); free(o->leaf_is_active); free(o->ir_lo); free(o->ir_hi);
  free(o->inodeparts); free(o->treecache);
  o->node = NULL; o->leaf_is_active = NULL; o->ir_lo = o->ir_hi = NULL;
  o->inodeparts = NULL; o->treecache = NULL;
  o->tree_built = 0;
}

void orc_destroy(orc_ctx *o) {
  if (!o) return;
  free_tree(o);
  free(o);
}

const char *orc_last_error(const orc_ctx *o) { return o->err; }

/* ------------------------------------------------------------------ */
/* tree build                                                          */
/* ------------------------------------------------------------------ */
typedef struct { int32_t node, parent, level, npnode; double xmin[3], xmax[3]; } bstack_t;

#define TC(o, pos, c) ((o)->treecache[5 * (size_t)((pos) - 1) + (c)])

/* sort_particles_in_cell (kdtree.F90:937-1001): two-pointer partition on
 * treecache(iaxis,:) <= xpivot, swapping inodeparts and treecache(1:5). */
static void sort_particles_in_cell(orc_ctx *o, int iaxis, int32_t imin, int32_t imax,
                                   int32_t *min_l, int32_t *max_l, int32_t *min_r, int32_t *max_r,
                                   int32_t *nl, int32_t *nr, double xpivot) {
  int32_t i = imin, j = imax;
  double xi_c = TC(o, i, iaxis), xj_c = TC(o, j, iaxis);
  int i_lt = xi_c <= xpivot, j_lt = xj_c <= xpivot;
  while (i < j) {
    if (i_lt) {
      i = i + 1;
      xi_c = TC(o, i, iaxis);
      i_lt = xi_c <= xpivot;
    } else {
      if (!j_lt) {
        j = j - 1;
        xj_c = TC(o, j, iaxis);
        j_lt = xj_c <= xpivot;
      } else {
        int32_t ps = o->inodeparts[i];
        double sw[5];
        memcpy(sw, &TC(o, i, 0), sizeof(sw));
        o->inodeparts[i] = o->inodeparts[j];
        memcpy(&TC(o, i, 0), &TC(o, j, 0), sizeof(sw));
        o->inodeparts[j] = ps;
        memcpy(&TC(o, j, 0), sw, sizeof(sw));
        i = i + 1;
        j = j - 1;
        xi_c = TC(o, i, iaxis);
        xj_c = TC(o, j, iaxis);
        i_lt = xi_c <= xpivot;
        j_lt = xj_c <= xpivot;
      }
    }
  }
  if (!i_lt) i = i - 1;
  if (j_lt) j = j + 1;
  *min_l = imin; *max_l = i; *min_r = j; *max_r = imax;
  *nl = *max_l - *min_l + 1;
  *nr = *max_r - *min_r + 1;
}

/* construct_node (kdtree.F90:531-929), non-MPI, non-APR branch. */
static int construct_node(orc_ctx *o, int32_t nnode, int32_t mymum, int32_t level,
                          const double xmini[3], const double xmaxi[3], int32_t npnode,
                          int32_t *il, int32_t *ir, int32_t *nl, int32_t *nr,
                          double xminl[3], double xmaxl[3], double xminr[3], double xmaxr[3],
                          int *wassplit) {
  phsph_kdnode *nd = &o->node[nnode];
  int nodeisactive = 0;
  int32_t npcounter = 0;
  if (o->ir_lo[nnode] > 0) {
    for (int32_t i = o->ir_lo[nnode]; i <= o->ir_hi[nnode]; i++)
      if (o->inodeparts[i] > 0) { nodeisactive = 1; break; }
    npcounter = o->ir_hi[nnode] - o->ir_lo[nnode] + 1;
  }
  if (npcounter != npnode) {
    snprintf(o->err, sizeof(o->err), "maketree: node %d has %d particles, expected %d", nnode, npcounter, npnode);
    return PHSPH_ERR_STATE;
  }
  *il = *ir = *nl = *nr = 0;
  *wassplit = 0;
  if (npnode < 1) return PHSPH_OK;

  double hmax = 0., xcofm = 0., ycofm = 0., zcofm = 0., totmass = 0.;
  double pmassi = o->p.massoftype_gas, dfac, fac;
  dfac = (pmassi > 0.) ? 1. / pmassi : 1.;
  int32_t i1 = o->ir_lo[nnode];
  for (int32_t i = i1; i <= i1 + npnode - 1; i++) {
    double xi = TC(o, i, 0), yi = TC(o, i, 1), zi = TC(o, i, 2), hi = TC(o, i, 3);
    pmassi = TC(o, i, 4);
    fac = pmassi * dfac;
    hmax = (hi > hmax) ? hi : hmax;
    totmass = totmass + pmassi;
    xcofm = xcofm + fac * xi;
    ycofm = ycofm + fac * yi;
    zcofm = zcofm + fac * zi;
  }
  double cofm[3] = {xcofm, ycofm, zcofm};
  if (totmass > 0.) {
    double d = totmass * dfac;
    cofm[0] = cofm[0] / d; cofm[1] = cofm[1] / d; cofm[2] = cofm[2] / d;
  }
  if (totmass <= 0.) {
    snprintf(o->err, sizeof(o->err), "maketree: totmass_node <= 0");
    return PHSPH_ERR_INPUT;
  }
  double r2max = 0.;
  double q[6] = {0., 0., 0., 0., 0., 0.};
  for (int32_t i = i1; i <= i1 + npnode - 1; i++) {
    double dx = TC(o, i, 0) - cofm[0];
    double dy = TC(o, i, 1) - cofm[1];
    double dz = TC(o, i, 2) - cofm[2];
    double dr2 = dx * dx + dy * dy + dz * dz;
    r2max = (dr2 > r2max) ? dr2 : r2max;
    if (o->quads) {
      double pm = TC(o, i, 4);
      q[0] = q[0] + pm * (dx * dx);
      q[1] = q[1] + pm * (dx * dy);
      q[2] = q[2] + pm * (dx * dz);
      q[3] = q[3] + pm * (dy * dy);
      q[4] = q[4] + pm * (dy * dz);
      q[5] = q[5] + pm * (dz * dz);
    }
  }
  if (o->quads)
    for (int k = 0; k < 6; k++) o->quads[6 * (size_t)nnode + k] = q[k];
  nd->xcen[0] = cofm[0]; nd->xcen[1] = cofm[1]; nd->xcen[2] = cofm[2];
  nd->size = sqrt(r2max) + EPS64;
  nd->hmax = hmax;
  nd->mass = totmass;
  nd->parent = mymum;
  nd->idum = 0;

  *wassplit = (npnode > MINPART);
  if (!*wassplit) {
    nd->leftchild = 0; nd->rightchild = 0;
    if (level > o->maxlevel) o->maxlevel = level;
    if (level < o->minlevel) o->minlevel = level;
    if (o->p.ind_timesteps) o->leaf_is_active[nnode] = nodeisactive ? 1 : -1;
    else o->leaf_is_active[nnode] = 1;
    return PHSPH_OK;
  }
  /* split along the longest axis at the centre of mass (:815-816) */
  int iaxis = 0;
  double ext = xmaxi[0] - xmini[0];
  for (int a = 1; a < 3; a++) {
    double e = xmaxi[a] - xmini[a];
    if (e > ext) { ext = e; iaxis = a; }
  }
  double xpivot = cofm[iaxis];
  if (level < o->maxlevel_indexed) {
    *il = 2 * nnode; *ir = *il + 1;
  } else {
    o->ncells += 2;
    *ir = (int32_t)o->ncells; *il = (int32_t)(o->ncells - 1);
    if (*ir > o->ncellsmax) {
      snprintf(o->err, sizeof(o->err), "maketree: number of nodes exceeds ncellsmax");
      return PHSPH_ERR_CAPACITY;
    }
  }
  nd->leftchild = *il; nd->rightchild = *ir;
  o->leaf_is_active[nnode] = 0;

  sort_particles_in_cell(o, iaxis, o->ir_lo[nnode], o->ir_hi[nnode], &o->ir_lo[*il], &o->ir_hi[*il],
                         &o->ir_lo[*ir], &o->ir_hi[*ir], nl, nr, xpivot);
  if (*nl == npnode || *nr == npnode) {
    *nl = npnode / 2;
    o->ir_lo[*il] = o->ir_lo[nnode];
    o->ir_hi[*il] = o->ir_lo[nnode] + *nl - 1;
    o->ir_lo[*ir] = o->ir_lo[nnode] + *nl;
    o->ir_hi[*ir] = o->ir_hi[nnode];
    *nr = npnode - *nl;
  }
  for (int c = 0; c < 3; c++) { xminl[c] = xmaxl[c] = TC(o, o->ir_lo[*il], c); }
  for (int32_t ip = o->ir_lo[*il] + 1; ip <= o->ir_hi[*il]; ip++)
    for (int c = 0; c < 3; c++) {
      double v = TC(o, ip, c);
      if (v < xminl[c]) xminl[c] = v;
      if (v > xmaxl[c]) xmaxl[c] = v;
    }
  for (int c = 0; c < 3; c++) { xminr[c] = xmaxr[c] = TC(o, o->ir_lo[*ir], c); }
  for (int32_t ip = o->ir_lo[*ir] + 1; ip <= o->ir_hi[*ir]; ip++)
    for (int c = 0; c < 3; c++) {
      double v = TC(o, ip, c);
      if (v < xminr[c]) xminr[c] = v;
      if (v > xmaxr[c]) xmaxr[c] = v;
    }
  return PHSPH_OK;
}

/* maketree (kdtree.F90:117-314) with construct_root_node (:347-488),
 * executed as with numthreads = 1: no queue phase, one depth-first stack. */
int orc_maketree(orc_ctx *o, double *xyzh, int32_t np, const int8_t *iphase) {
  free_tree(o);
  o->np = np;
  int64_t maxp = (o->p.maxp > np) ? o->p.maxp : np;
  o->ncellsmax = 2 * maxp;
  size_t nslots = (size_t)o->ncellsmax + 2;
  o->node = (phsph_kdnode *)calloc(nslots, sizeof(phsph_kdnode));
  if (o->p.gravity) o->quads = (double *)calloc((size_t)nslots * 6, sizeof(double));
  o->leaf_is_active = (int32_t *)calloc(nslots, sizeof(int32_t));
  o->ir_lo = (int32_t *)calloc(nslots, sizeof(int32_t));
  o->ir_hi = (int32_t *)calloc(nslots, sizeof(int32_t));
  o->inodeparts = (int32_t *)calloc((size_t)np + 2, sizeof(int32_t));
  o->treecache = (double *)calloc(5 * ((size_t)np + 1), sizeof(double));
  if (!o->node || !o->leaf_is_active || !o->ir_lo || !o->ir_hi || !o->inodeparts || !o->treecache) {
    snprintf(o->err, sizeof(o->err), "maketree: out of host memory");
    return PHSPH_ERR_CAPACITY;
  }
  /* construct_root_node: bounds start from particle 1 before any wrap */
  double xminp = xyzh[0], yminp = xyzh[1], zminp = xyzh[2];
  double xmaxp = xminp, ymaxp = yminp, zmaxp = zminp;
  for (int32_t i = 0; i < np; i++) {
    double *r = &xyzh[4 * (size_t)i];
    if (r[3] < TINY64) continue;
    if (o->p.periodic) {  /* cross_boundary (boundary.f90:123-160) */
      if (r[0] < o->p.xmin) r[0] = r[0] + o->dxbound;
      else if (r[0] > o->p.xmax) r[0] = r[0] - o->dxbound;
      if (r[1] < o->p.ymin) r[1] = r[1] + o->dybound;
      else if (r[1] > o->p.ymax) r[1] = r[1] - o->dybound;
      if (r[2] < o->p.zmin) r[2] = r[2] + o->dzbound;
      else if (r[2] > o->p.zmax) r[2] = r[2] - o->dzbound;
    }
    if (isnan(r[0]) || isnan(r[1]) || isnan(r[2])) {
      snprintf(o->err, sizeof(o->err), "maketree: NaN in particle position %d", i + 1);
      return PHSPH_ERR_NAN;
    }
    if (r[0] < xminp) xminp = r[0];
    if (r[1] < yminp) yminp = r[1];
    if (r[2] < zminp) zminp = r[2];
    if (r[0] > xmaxp) xmaxp = r[0];
    if (r[1] > ymaxp) ymaxp = r[1];
    if (r[2] > zmaxp) zmaxp = r[2];
  }
  int32_t nproot = 0;
  for (int32_t i = 0; i < np; i++) {
    const double *r = &xyzh[4 * (size_t)i];
    if (r[3] < TINY64) continue;
    nproot++;
    int active = (iphase == NULL) ? 1 : (iphase[i] > 0);
    o->inodeparts[nproot] = (o->p.ind_timesteps && !active) ? -(i + 1) : (i + 1);
    TC(o, nproot, 0) = r[0]; TC(o, nproot, 1) = r[1]; TC(o, nproot, 2) = r[2]; TC(o, nproot, 3) = r[3];
    TC(o, nproot, 4) = o->p.massoftype_gas;
  }
  if (nproot == 0) {
    snprintf(o->err, sizeof(o->err), "maketree: no particles or all particles dead/accreted");
    return PHSPH_ERR_NODATA;
  }
  o->ir_lo[1] = 1; o->ir_hi[1] = nproot;

  o->maxlevel = 0;
  o->minlevel = 32 - 1;
  o->maxlevel_indexed = (int32_t)(log((double)(o->ncellsmax + 1)) / log(2.)) - 1;
  o->ncells = ((int64_t)1 << (o->maxlevel_indexed + 1)) - 1;

  bstack_t *stack = (bstack_t *)malloc(ISTACKSIZE * sizeof(bstack_t));
  int istack = 1;
  stack[0].node = 1; stack[0].parent = 0; stack[0].level = 0; stack[0].npnode = nproot;
  stack[0].xmin[0] = xminp; stack[0].xmin[1] = yminp; stack[0].xmin[2] = zminp;
  stack[0].xmax[0] = xmaxp; stack[0].xmax[1] = ymaxp; stack[0].xmax[2] = zmaxp;
  int rc = PHSPH_OK;
  while (istack > 0) {
    bstack_t s = stack[istack - 1];
    istack--;
    int32_t il, ir, nl, nr;
    double xminl[3], xmaxl[3], xminr[3], xmaxr[3];
    int wassplit;
    rc = construct_node(o, s.node, s.parent, s.level, s.xmin, s.xmax, s.npnode, &il, &ir, &nl, &nr,
                        xminl, xmaxl, xminr, xmaxr, &wassplit);
    if (rc) break;
    if (wassplit) {
      if (istack + 2 > ISTACKSIZE) {
        snprintf(o->err, sizeof(o->err), "maketree: stack size exceeded");
        rc = PHSPH_ERR_CAPACITY; break;
      }
      bstack_t *a = &stack[istack++];
      a->node = il; a->parent = s.node; a->level = s.level + 1; a->npnode = nl;
      memcpy(a->xmin, xminl, sizeof(xminl)); memcpy(a->xmax, xmaxl, sizeof(xmaxl));
      bstack_t *b = &stack[istack++];
      b->node = ir; b->parent = s.node; b->level = s.level + 1; b->npnode = nr;
      memcpy(b->xmin, xminr, sizeof(xminr)); memcpy(b->xmax, xmaxr, sizeof(xmaxr));
    }
  }
  free(stack);
  if (rc) return rc;
  if (o->maxlevel < o->maxlevel_indexed) o->ncells = ((int64_t)1 << (o->maxlevel + 1)) - 1;
  o->tree_built = 1;
  return PHSPH_OK;
}

int orc_get_tree(const orc_ctx *o, phsph_kdnode *node, int32_t *inoderange, int32_t *inodeparts,
                 int32_t *leaf_is_active, int64_t cap_nodes, int64_t *ncells, int32_t *maxlevel) {
  if (!o->tree_built) return PHSPH_ERR_STATE;
  int64_t n = o->ncells < cap_nodes ? o->ncells : cap_nodes;
  for (int64_t k = 0; k < n; k++) {
    if (node) node[k] = o->node[k + 1];
    if (inoderange) { inoderange[2 * k] = o->ir_lo[k + 1]; inoderange[2 * k + 1] = o->ir_hi[k + 1]; }
    if (leaf_is_active) leaf_is_active[k] = o->leaf_is_active[k + 1];
  }
  if (inodeparts) for (int32_t p = 0; p < o->np; p++) inodeparts[p] = o->inodeparts[p + 1];
  if (ncells) *ncells = o->ncells;
  if (maxlevel) *maxlevel = o->maxlevel;
  return PHSPH_OK;
}

/* ------------------------------------------------------------------ */
/* tree walk                                                           */
/* ------------------------------------------------------------------ */

/* get_sep (kdtree.F90:1468-1501) */
static inline void get_sep(const orc_ctx *o, const double x1[3], const double x2[3], double *dx,
                           double *dy, double *dz, double *xoff, double *yoff, double *zoff, double *r2) {
  *xoff = *yoff = *zoff = 0.;
  *dx = x1[0] - x2[0]; *dy = x1[1] - x2[1]; *dz = x1[2] - x2[2];
  if (o->p.periodic) {
    if (fabs(*dx) > o->hdlx) { *xoff = copysign(o->dxbound, *dx); *dx = *dx - *xoff; }
    if (fabs(*dy) > o->hdly) { *yoff = copysign(o->dybound, *dy); *dy = *dy - *yoff; }
    if (fabs(*dz) > o->hdlz) { *zoff = copysign(o->dzbound, *dz); *dz = *dz - *zoff; }
  }
  *r2 = *dx * *dx + *dy * *dy + *dz * *dz;
}

typedef struct {
  int32_t *list;     /* 1-based particle indices                  */
  double *cache;     /* 4 per entry: x, y, z (offset), 1/h        */
  int64_t cap;
  int32_t n;
} nlist_t;

static int nlist_reserve(nlist_t *l, int64_t need) {
  if (need <= l->cap) return 0;
  int64_t nc = l->cap ? l->cap : 4096;
  while (nc < need) nc *= 2;
  int32_t *nl2 = (int32_t *)realloc(l->list, nc * sizeof(int32_t));
  double *nc2 = (double *)realloc(l->cache, nc * 4 * sizeof(double));
  if (!nl2 || !nc2) return -1;
  l->list = nl2; l->cache = nc2; l->cap = nc;
  return 0;
}

/* getneigh (kdtree.F90:1221-1347) + cache_neighbours (:1175-1213) for
 * the non-gravity walk.  get_hj selects rcut = max(rcuti, radkern*hmax_j). */
static int getneigh(const orc_ctx *o, const double xpos[3], double xsizei, double rcuti, int get_hj,
                    nlist_t *l) {
  int32_t nstack[64];
  int istack = 1;
  nstack[0] = 1;
  l->n = 0;
  while (istack != 0) {
    int32_t n = nstack[--istack];
    const phsph_kdnode *nd = &o->node[n];
    double dx, dy, dz, xo, yo, zo, r2;
    get_sep(o, xpos, nd->xcen, &dx, &dy, &dz, &xo, &yo, &zo, &r2);
    double rcut = rcuti;
    if (get_hj) {
      double rcutj = o->radkern * nd->hmax;
      rcut = (rcuti > rcutj) ? rcuti : rcutj;
    }
    double s = xsizei + nd->size + rcut;
    if (r2 < s * s) {
      if (o->leaf_is_active[n] != 0) {
        int32_t npnode = o->ir_hi[n] - o->ir_lo[n] + 1;
        if (nlist_reserve(l, (int64_t)l->n + npnode)) return PHSPH_ERR_CAPACITY;
        for (int32_t k = 0; k < npnode; k++) {
          int32_t pos = o->ir_lo[n] + k;
          int32_t idx = l->n + k;
          l->list[idx] = abs(o->inodeparts[pos]);
          l->cache[4 * idx + 0] = TC(o, pos, 0) + xo;
          l->cache[4 * idx + 1] = TC(o, pos, 1) + yo;
          l->cache[4 * idx + 2] = TC(o, pos, 2) + zo;
          l->cache[4 * idx + 3] = 1. / TC(o, pos, 3);
        }
        l->n += npnode;
      } else {
        if (istack + 2 > 64) return PHSPH_ERR_CAPACITY;
        if (nd->leftchild != 0) nstack[istack++] = nd->leftchild;
        if (nd->rightchild != 0) nstack[istack++] = nd->rightchild;
      }
    }
  }
  return PHSPH_OK;
}

static void kernel_softening(int kind, double q2, double q, double *pot, double *fs);

/* ------------------------------------------------------------------ */
/* self-gravity: dual tree walk and expansions                         */
/* ------------------------------------------------------------------ */
#define LENFGRAV 20
#define MAXDEPTH 32   /* kdtree.F90:42 */

/* compute_M2L (kdtree.F90:1702-1781): quadrupole source at separation
 * (dx,dy,dz), dr1 = 1/r, into the 20 coefficients of fnode */
static void compute_M2L(double dx, double dy, double dz, double dr1, double q0, const double *quads,
                        double *fnode) {
  double dr12 = dr1 * dr1;
  double dx2 = dx * dx, dx3 = dx * dx2, dy2 = dy * dy, dy3 = dy * dy2, dz2 = dz * dz, dz3 = dz * dz2;
  double g0 = -dr1;
  double g1 = 1. * dr12 * g0;
  double g2 = -3. * dr12 * g1;
  double g3 = -5. * dr12 * g2;
  double g2dx = g2 * dx, g2dy = g2 * dy, g2dz = g2 * dz;
  double D3[10], D2[6], D1[3];
  D3[0] = 3. * g2dx + g3 * dx3;
  D3[1] = g2dy + g3 * dx2 * dy;
  D3[2] = g2dz + g3 * dx2 * dz;
  D3[3] = g2dx + g3 * dy2 * dx;
  D3[4] = g3 * dx * dy * dz;
  D3[5] = g2dx + g3 * dz2 * dx;
  D3[6] = 3. * g2dy + g3 * dy3;
  D3[7] = g2dz + g3 * dy2 * dz;
  D3[8] = g2dy + g3 * dz2 * dy;
  D3[9] = 3. * g2dz + g3 * dz3;
  D2[0] = g1 + g2 * dx2;
  D2[1] = g2dx * dy;
  D2[2] = g2dx * dz;
  D2[3] = g1 + g2 * dy2;
  D2[4] = g2dy * dz;
  D2[5] = g1 + g2 * dz2;
  D1[0] = g1 * dx; D1[1] = g1 * dy; D1[2] = g1 * dz;
  double qxx = quads[0], qxy = quads[1], qxz = quads[2], qyy = quads[3], qyz = quads[4], qzz = quads[5];
  fnode[0] = fnode[0] + (D1[0] * q0 + 0.5 * (D3[0] * qxx + 2. * (D3[1] * qxy + D3[2] * qxz + D3[4] * qyz) +
                                             D3[3] * qyy + D3[5] * qzz));
  fnode[1] = fnode[1] + (D1[1] * q0 + 0.5 * (D3[1] * qxx + 2. * (D3[3] * qxy + D3[4] * qxz + D3[7] * qyz) +
                                             D3[6] * qyy + D3[8] * qzz));
  fnode[2] = fnode[2] + (D1[2] * q0 + 0.5 * (D3[2] * qxx + 2. * (D3[4] * qxy + D3[5] * qxz + D3[8] * qyz) +
                                             D3[7] * qyy + D3[9] * qzz));
  for (int k = 0; k < 6; k++) fnode[3 + k] = fnode[3 + k] + D2[k] * q0;
  for (int k = 0; k < 10; k++) fnode[9 + k] = fnode[9 + k] + D3[k] * q0;
  fnode[19] = fnode[19] + g0 * q0 -
              0.5 * (D2[0] * qxx + D2[3] * qyy + D2[5] * qzz + 2 * (D2[1] * qxy + D2[2] * qxz + D2[4] * qyz));
}

/* propagate_fnode_to_node (kdtree.F90:1527-1560): shift the expansion
 * fs about the parent centre by (dx,dy,dz) into f */
static void propagate_fnode_to_node(double *f, const double *fs, double dx, double dy, double dz) {
  f[0] = fs[0] + dx * (fs[3] + 0.5 * (dx * fs[9] + dy * fs[10] + dz * fs[11])) +
         dy * (fs[4] + 0.5 * (dx * fs[10] + dy * fs[12] + dz * fs[13])) +
         dz * (fs[5] + 0.5 * (dx * fs[11] + dy * fs[13] + dz * fs[14]));
  f[1] = fs[1] + dx * (fs[4] + 0.5 * (dx * fs[10] + dy * fs[12] + dz * fs[13])) +
         dy * (fs[6] + 0.5 * (dx * fs[12] + dy * fs[15] + dz * fs[16])) +
         dz * (fs[7] + 0.5 * (dx * fs[13] + dy * fs[16] + dz * fs[17]));
  f[2] = fs[2] + dx * (fs[5] + 0.5 * (dx * fs[11] + dy * fs[13] + dz * fs[14])) +
         dy * (fs[7] + 0.5 * (dx * fs[13] + dy * fs[16] + dz * fs[17])) +
         dz * (fs[8] + 0.5 * (dx * fs[14] + dy * fs[17] + dz * fs[18]));
  f[3] = fs[3] + dx * fs[9] + dy * fs[10] + dz * fs[11];
  f[4] = fs[4] + dx * fs[10] + dy * fs[12] + dz * fs[13];
  f[5] = fs[5] + dx * fs[11] + dy * fs[13] + dz * fs[14];
  f[6] = fs[6] + dx * fs[12] + dy * fs[15] + dz * fs[16];
  f[7] = fs[7] + dx * fs[13] + dy * fs[16] + dz * fs[17];
  f[8] = fs[8] + dx * fs[14] + dy * fs[17] + dz * fs[18];
  for (int k = 9; k < 19; k++) f[k] = fs[k];
  f[19] = fs[19] + dx * (fs[0] + 0.5 * (dx * fs[3] + dy * fs[4] + dz * fs[5])) +
          dy * (fs[1] + 0.5 * (dx * fs[4] + dy * fs[6] + dz * fs[7])) +
          dz * (fs[2] + 0.5 * (dx * fs[5] + dy * fs[7] + dz * fs[8]));
}

/* expand_fgrav_in_taylor_series (kdtree.F90:1799-1840), including the
 * reference's use of dfxy in the z term of the first potential bracket */
static void expand_fgrav(const double *fn, double dx, double dy, double dz, double *fx, double *fy, double *fz,
                         double *pot) {
  double fxi = fn[0], fyi = fn[1], fzi = fn[2];
  double dfxx = fn[3], dfxy = fn[4], dfxz = fn[5], dfyy = fn[6], dfyz = fn[7], dfzz = fn[8];
  double d2fxxx = fn[9], d2fxxy = fn[10], d2fxxz = fn[11], d2fxyy = fn[12], d2fxyz = fn[13], d2fxzz = fn[14];
  double d2fyyy = fn[15], d2fyyz = fn[16], d2fyzz = fn[17], d2fzzz = fn[18];
  double poti = fn[19];
  fxi = fxi + dx * (dfxx + 0.5 * (dx * d2fxxx + dy * d2fxxy + dz * d2fxxz)) +
        dy * (dfxy + 0.5 * (dx * d2fxxy + dy * d2fxyy + dz * d2fxyz)) +
        dz * (dfxz + 0.5 * (dx * d2fxxz + dy * d2fxyz + dz * d2fxzz));
  fyi = fyi + dx * (dfxy + 0.5 * (dx * d2fxxy + dy * d2fxyy + dz * d2fxyz)) +
        dy * (dfyy + 0.5 * (dx * d2fxyy + dy * d2fyyy + dz * d2fyyz)) +
        dz * (dfyz + 0.5 * (dx * d2fxyz + dy * d2fyyz + dz * d2fyzz));
  fzi = fzi + dx * (dfxz + 0.5 * (dx * d2fxxz + dy * d2fxyz + dz * d2fxzz)) +
        dy * (dfyz + 0.5 * (dx * d2fxyz + dy * d2fyyz + dz * d2fyzz)) +
        dz * (dfzz + 0.5 * (dx * d2fxzz + dy * d2fyzz + dz * d2fzzz));
  poti = poti - (dx * (fxi - 0.5 * (dx * dfxx + dy * dfxy + dz * dfxy)) +
                 dy * (fyi - 0.5 * (dx * dfxy + dy * dfyy + dz * dfyz)) +
                 dz * (fzi - 0.5 * (dx * dfxz + dy * dfyz + dz * dfzz)));
  *fx = fxi; *fy = fyi; *fz = fzi; *pot = poti;
}

/* cache_neighbours (kdtree.F90:1175-1213) of leaf n with offsets */
static int cache_leaf(const orc_ctx *o, int32_t n, double xo, double yo, double zo, nlist_t *l) {
  int32_t npnode = o->ir_hi[n] - o->ir_lo[n] + 1;
  if (nlist_reserve(l, (int64_t)l->n + npnode)) return PHSPH_ERR_CAPACITY;
  for (int32_t k = 0; k < npnode; k++) {
    int32_t pos = o->ir_lo[n] + k;
    int32_t idx = l->n + k;
    l->list[idx] = abs(o->inodeparts[pos]);
    l->cache[4 * idx + 0] = TC(o, pos, 0) + xo;
    l->cache[4 * idx + 1] = TC(o, pos, 1) + yo;
    l->cache[4 * idx + 2] = TC(o, pos, 2) + zo;
    l->cache[4 * idx + 3] = 1. / TC(o, pos, 3);
  }
  l->n += npnode;
  return PHSPH_OK;
}

/* getneigh_dual (kdtree.F90:1357-1461) with open_nodes (:1591-1645) and
 * node_interaction (:1653-1694).  The reference caches a parent's
 * per-level expansion (fnodecache) for the next leaf below it; the cached
 * values are the very ones recomputed here, so the result is the same. */
static int getneigh_dual(const orc_ctx *o, int32_t icell, nlist_t *l, double fnode[LENFGRAV]) {
  int32_t branch[MAXDEPTH + 1];
  int nparents = 1;
  branch[1] = icell;
  int32_t j = icell;
  while (o->node[j].parent != 0) {
    j = o->node[j].parent;
    if (nparents + 1 > MAXDEPTH) return PHSPH_ERR_CAPACITY;
    branch[++nparents] = j;
  }
  double fb[MAXDEPTH + 1][LENFGRAV];
  memset(fb, 0, sizeof(fb));
  const double tree_acc2 = o->p.tree_accuracy * o->p.tree_accuracy;
  int32_t st_dst[4 * MAXDEPTH], st_src[4 * MAXDEPTH], st_br[4 * MAXDEPTH];
  int istack = 1;
  st_dst[0] = 1; st_src[0] = 1; st_br[0] = nparents;
  l->n = 0;
  while (istack > 0) {
    istack--;
    int32_t idst = st_dst[istack], isrc = st_src[istack], idstbranch = st_br[istack];
    int stackit;
    double xo = 0., yo = 0., zo = 0.;
    if (idst == isrc) {
      stackit = 1;
    } else {
      const phsph_kdnode *nd = &o->node[idst], *ns = &o->node[isrc];
      double dx, dy, dz, r2;
      get_sep(o, nd->xcen, ns->xcen, &dx, &dy, &dz, &xo, &yo, &zo, &r2);
      double rcut_dst = nd->hmax * o->radkern, rcut_src = ns->hmax * o->radkern;
      double rcut = (rcut_dst > rcut_src) ? rcut_dst : rcut_src;
      double ss = nd->size + ns->size;
      double rcut2 = (ss + rcut) * (ss + rcut);
      int wellsep = (tree_acc2 * r2 > ss * ss) && (r2 > rcut2);
      if (wellsep) {
        double dr1 = 1. / sqrt(r2);
        compute_M2L(dx, dy, dz, dr1, ns->mass, &o->quads[6 * (size_t)isrc], fb[idstbranch]);
        stackit = 0;
      } else {
        stackit = 1;
      }
    }
    if (!stackit) continue;
    /* open_nodes */
    const phsph_kdnode *ns = &o->node[isrc];
    int ibranchnext, isdstleaf;
    if (idstbranch - 1 > 0) { ibranchnext = idstbranch - 1; isdstleaf = 0; }
    else { ibranchnext = idstbranch; isdstleaf = 1; }
    int32_t idstnext = branch[ibranchnext];
    if (o->leaf_is_active[isrc] != 0) {
      if (isdstleaf) {
        int rc = cache_leaf(o, isrc, xo, yo, zo, l);
        if (rc) return rc;
      } else {
        if (istack + 1 > 4 * MAXDEPTH) return PHSPH_ERR_CAPACITY;
        st_dst[istack] = idstnext; st_src[istack] = isrc; st_br[istack] = ibranchnext; istack++;
      }
    } else {
      if (istack + 2 > 4 * MAXDEPTH) return PHSPH_ERR_CAPACITY;
      if (ns->leftchild != 0) { st_dst[istack] = idstnext; st_src[istack] = ns->leftchild; st_br[istack] = ibranchnext; istack++; }
      if (ns->rightchild != 0) { st_dst[istack] = idstnext; st_src[istack] = ns->rightchild; st_br[istack] = ibranchnext; istack++; }
    }
  }
  /* downward pass: shift each level's expansion to the next centre */
  double facc[LENFGRAV], ftmp[LENFGRAV];
  memset(facc, 0, sizeof(facc));
  for (int i = nparents; i >= 2; i--) {
    int32_t iparent = branch[i];
    double dx, dy, dz, xo, yo, zo, r2;
    get_sep(o, o->node[branch[i - 1]].xcen, o->node[iparent].xcen, &dx, &dy, &dz, &xo, &yo, &zo, &r2);
    for (int k = 0; k < LENFGRAV; k++) ftmp[k] = facc[k] + fb[i][k];
    propagate_fnode_to_node(facc, ftmp, dx, dy, dz);
  }
  for (int k = 0; k < LENFGRAV; k++) fnode[k] = facc[k] + fb[1][k];
  return PHSPH_OK;
}

/* ------------------------------------------------------------------ */
/* density                                                             */
/* ------------------------------------------------------------------ */
typedef struct {
  int32_t icell, npcell, nits, nneightry;
  int32_t arr_index[MINPART + 1];
  int32_t nneigh[MINPART + 1];
  double h[MINPART + 1], h_old[MINPART + 1];
  double xpv[MINPART + 1][14];  /* x y z vx vy vz u Bx By Bz psi fx fy fz (dens.F90:33-47) */
  double rhosums[MINPART + 1][NRHOSUM];
  double xpos[3], xsizei, rcuti, hmax;
} celldens_t;

typedef struct {
  const double *xyzh, *vxyzu, *fxyzu, *fext, *Bevol;
  const double *h_in;  /* snapshot of h at entry (div B rho_j) */
  int32_t ldv, ldf;
} densin_t;

/* get_density_sums (dens.F90:578-857) for one particle. */
static void get_density_sums(const orc_ctx *o, int32_t i, const double *xpv, double hi, double hi1,
                             double hi21, const nlist_t *l, int32_t *nneighi, double *rs,
                             const densin_t *in) {
  const int kind = o->p.kernel;
  const double pm = o->p.massoftype_gas;
  const int getdB = o->p.mhd;
  memset(rs, 0, NRHOSUM * sizeof(double));
  *nneighi = 1;
  double fxi = xpv[11], fyi = xpv[12], fzi = xpv[13];
  for (int32_t n = 0; n < l->n; n++) {
    int32_t j = l->list[n];
    if (j == i) continue;
    double dx, dy, dz;
    if (n < ISIZECELLCACHE) {
      dx = xpv[0] - l->cache[4 * n + 0];
      dy = xpv[1] - l->cache[4 * n + 1];
      dz = xpv[2] - l->cache[4 * n + 2];
    } else {
      const double *rj = &in->xyzh[4 * (size_t)(j - 1)];
      dx = xpv[0] - rj[0]; dy = xpv[1] - rj[1]; dz = xpv[2] - rj[2];
    }
    if (o->p.periodic) {
      if (fabs(dx) > 0.5 * o->dxbound) dx = dx - copysign(o->dxbound, dx);
      if (fabs(dy) > 0.5 * o->dybound) dy = dy - copysign(o->dybound, dy);
      if (fabs(dz) > 0.5 * o->dzbound) dz = dz - copysign(o->dzbound, dz);
    }
    double rij2 = dx * dx + dy * dy + dz * dz;
    double q2i = rij2 * hi21;
    if (!(q2i < o->radkern2)) continue;
    double rij = sqrt(rij2);
    double qi = rij * hi1;
    double wabi, grkerni, dphidhi = 0.;
    get_kernel(kind, q2i, qi, &wabi, &grkerni);
    if (o->p.gravity) dphidhi = dphidh_kern(kind, q2i, qi);
    double dwdhi = (-qi * grkerni - 3. * wabi);
    rs[R_RHO] = rs[R_RHO] + wabi * pm;
    rs[R_GRADH] = rs[R_GRADH] + dwdhi * pm;
    rs[R_GRADSOFT] = rs[R_GRADSOFT] + dphidhi * pm;
    *nneighi += 1;
    double rij1 = 1. / (rij + EPS64);
    double rij1grkern = rij1 * grkerni;
    double runix = dx * rij1grkern * pm;
    double runiy = dy * rij1grkern * pm;
    double runiz = dz * rij1grkern * pm;
    const double *vj = &in->vxyzu[(size_t)in->ldv * (j - 1)];
    double dvx = xpv[3] - vj[0], dvy = xpv[4] - vj[1], dvz = xpv[5] - vj[2];
    double projv = dvx * runix + dvy * runiy + dvz * runiz;
    rs[R_DIVV] += projv;
    rs[R_DVXDX] += dvx * runix; rs[R_DVXDY] += dvx * runiy; rs[R_DVXDZ] += dvx * runiz;
    rs[R_DVYDX] += dvy * runix; rs[R_DVYDY] += dvy * runiy; rs[R_DVYDZ] += dvy * runiz;
    rs[R_DVZDX] += dvz * runix; rs[R_DVZDY] += dvz * runiy; rs[R_DVZDZ] += dvz * runiz;
    double fxj = 0., fyj = 0., fzj = 0.;
    if (in->fxyzu) {
      const double *fj = &in->fxyzu[(size_t)in->ldf * (j - 1)];
      fxj = fj[0]; fyj = fj[1]; fzj = fj[2];
    }
    if (in->fext) {
      const double *ej = &in->fext[3 * (size_t)(j - 1)];
      fxj = fxj + ej[0]; fyj = fyj + ej[1]; fzj = fzj + ej[2];
    }
    double dax = fxi - fxj, day = fyi - fyj, daz = fzi - fzj;
    rs[R_DAXDX] += dax * runix; rs[R_DAXDY] += dax * runiy; rs[R_DAXDZ] += dax * runiz;
    rs[R_DAYDX] += day * runix; rs[R_DAYDY] += day * runiy; rs[R_DAYDZ] += day * runiz;
    rs[R_DAZDX] += daz * runix; rs[R_DAZDY] += daz * runiy; rs[R_DAZDZ] += daz * runiz;
    rs[R_RXX] -= dx * runix; rs[R_RXY] -= dx * runiy; rs[R_RXZ] -= dx * runiz;
    rs[R_RYY] -= dy * runiy; rs[R_RYZ] -= dy * runiz; rs[R_RZZ] -= dz * runiz;
    if (getdB) {
      double rhoi = rhoh(o, hi, pm);
      double rhoj = rhoh(o, in->h_in[j - 1], pm);
      const double *Bj = &in->Bevol[4 * (size_t)(j - 1)];
      double dBx = xpv[7] * rhoi - Bj[0] * rhoj;
      double dBy = xpv[8] * rhoi - Bj[1] * rhoj;
      double dBz = xpv[9] * rhoi - Bj[2] * rhoj;
      double projdB = dBx * runix + dBy * runiy + dBz * runiz;
      rs[R_DIVB] += projdB;
      rs[R_DBXDX] += dBx * runix; rs[R_DBXDY] += dBx * runiy; rs[R_DBXDZ] += dBx * runiz;
      rs[R_DBYDX] += dBy * runix; rs[R_DBYDY] += dBy * runiy; rs[R_DBYDZ] += dBy * runiz;
      rs[R_DBZDX] += dBz * runix; rs[R_DBZDY] += dBz * runiy; rs[R_DBZDZ] += dBz * runiz;
    }
  }
}

/* finish_rhosum (dens.F90:1470-1507) */
static void finish_rhosum(const orc_ctx *o, const double *rs, double pm, double hi, int iterating,
                          double *rhoi, double *gradhi, double *rhohi, double *gradsofti,
                          double *dhdrhoi_out, double *omegai_out) {
  double hi1 = 1. / hi, hi21 = hi1 * hi1, hi31 = hi1 * hi21, hi41 = hi21 * hi21;
  *rhoi = o->cnormk * (rs[R_RHO] + o->wab0 * pm) * hi31;
  double gh = o->cnormk * (rs[R_GRADH] + o->gradh0 * pm) * hi41;
  double dhdrhoi = dhdrho(o, hi, pm);
  double omegai = 1. - dhdrhoi * gh;
  *gradhi = 1. / omegai;
  if (iterating) {
    *rhohi = rhoh(o, hi, pm);
    *dhdrhoi_out = dhdrhoi;
    *omegai_out = omegai;
  } else {
    double gs = (rs[R_GRADSOFT] + o->dphidh0 * pm) * hi21;
    *gradsofti = gs * dhdrhoi;
  }
}

/* exactlinear (dens.F90:1086-1101) */
static inline void exactlinear(double *gx, double *gy, double *gz, double dAx, double dAy, double dAz,
                               const double rm[6], double ddenom) {
  *gx = (dAx * rm[0] + dAy * rm[1] + dAz * rm[2]) * ddenom;
  *gy = (dAx * rm[1] + dAy * rm[3] + dAz * rm[4]) * ddenom;
  *gz = (dAx * rm[2] + dAy * rm[4] + dAz * rm[5]) * ddenom;
}

/* calculate_rmatrix_from_sums (dens.F90:866-891) */
static void rmatrix_from_sums(const double *rs, double *denom, double rm[6]) {
  double rxx = rs[R_RXX], rxy = rs[R_RXY], rxz = rs[R_RXZ], ryy = rs[R_RYY], ryz = rs[R_RYZ], rzz = rs[R_RZZ];
  *denom = rxx * ryy * rzz + 2. * rxy * rxz * ryz - rxx * ryz * ryz - ryy * rxz * rxz - rzz * rxy * rxy;
  rm[0] = ryy * rzz - ryz * ryz;
  rm[1] = rxz * ryz - rzz * rxy;
  rm[2] = rxy * ryz - rxz * ryy;
  rm[3] = rzz * rxx - rxz * rxz;
  rm[4] = rxy * rxz - rxx * ryz;
  rm[5] = rxx * ryy - rxy * rxy;
}

/* velocity gradient (exact-linear when det != 0), as used by both
 * calculate_divcurlv_from_sums (dens.F90:899-967) and
 * calculate_strain_from_sums (:997-1052) */
static void grad_from_sums(const double *rs, int i0, double termnorm, double denom, const double rm[6],
                           double g[9]) {
  if (fabs(denom) > TINY64) {
    double dd = 1. / denom, gx, gy, gz;
    for (int c = 0; c < 3; c++) {
      exactlinear(&gx, &gy, &gz, rs[i0 + 3 * c], rs[i0 + 3 * c + 1], rs[i0 + 3 * c + 2], rm, dd);
      g[3 * c] = -gx; g[3 * c + 1] = -gy; g[3 * c + 2] = -gz;
    }
  } else {
    for (int k = 0; k < 9; k++) g[k] = -rs[i0 + k] * termnorm;
  }
}

typedef struct {
  float *divcurlv; int32_t lddcv;
  float *divcurlB, *alphaind, *gradh; int32_t ldgradh;
  float *dvdx;
  double *xyzh;
} densout_t;

/* store_results (dens.F90:1511-1681) */
static void store_results(orc_ctx *o, const celldens_t *c, densout_t *out, int64_t *nneightry,
                          int64_t *nneighact, int64_t *maxneightry, int32_t *maxneighact, int64_t *npc,
                          int64_t *ncalc) {
  const double pm = o->p.massoftype_gas;
  for (int ip = 0; ip < c->npcell; ip++) {
    int32_t lli = o->inodeparts[c->arr_index[ip]];
    double hi = c->h[ip];
    const double *rs = c->rhosums[ip];
    double hi1 = 1. / hi, hi21 = hi1 * hi1, hi41 = hi21 * hi21;
    double rhoi, gradhi, gradsofti, dum1, dum2, dum3;
    finish_rhosum(o, rs, pm, hi, 0, &rhoi, &gradhi, &dum1, &gradsofti, &dum2, &dum3);
    double hnew = hrho8(o, rhoi, pm);
    out->xyzh[4 * (size_t)(lli - 1) + 3] = hnew;
    TC(o, c->arr_index[ip], 3) = hnew;
    float gh32 = (float)gradhi;
    out->gradh[(size_t)out->ldgradh * (lli - 1)] = gh32;
    if (o->p.gravity && out->ldgradh > 1) out->gradh[(size_t)out->ldgradh * (lli - 1) + 1] = (float)gradsofti;
    /* calculate_divcurlB branch (fast_divcurlB = .true.) */
    double gradh_r4 = (double)gh32;
    double rho1i = 1. / rhoi;
    double term = o->cnormk * gradh_r4 * rho1i * hi41;
    double denom, rm[6], g[9], ga[9];
    rmatrix_from_sums(rs, &denom, rm);
    out->divcurlv[(size_t)out->lddcv * (lli - 1)] = (float)(-rs[R_DIVV] * term);
    /* d(div v)/dt (dens.F90:927-964) */
    grad_from_sums(rs, R_DVXDX, term, denom, rm, g);
    double div_a;
    if (fabs(denom) > TINY64) {
      grad_from_sums(rs, R_DAXDX, term, denom, rm, ga);
      div_a = -(-ga[0] + -ga[4] + -ga[8]);
    } else {
      div_a = -term * (rs[R_DAXDX] + rs[R_DAYDY] + rs[R_DAZDZ]);
    }
    double d5 = div_a - (g[0] * g[0] + g[4] * g[4] + g[8] * g[8] +
                         2. * (g[1] * g[3] + g[2] * g[6] + g[5] * g[7]));
    if (out->alphaind) out->alphaind[3 * (size_t)(lli - 1) + 2] = (float)d5;
    if (o->p.mhd && out->divcurlB) {
      float *dB = &out->divcurlB[4 * (size_t)(lli - 1)];
      dB[0] = (float)(-rs[R_DIVB] * term);
      dB[1] = (float)(-(rs[R_DBZDY] - rs[R_DBYDZ]) * term);
      dB[2] = (float)(-(rs[R_DBXDZ] - rs[R_DBZDX]) * term);
      dB[3] = (float)(-(rs[R_DBYDX] - rs[R_DBXDY]) * term);
    }
    if (out->dvdx)
      for (int k = 0; k < 9; k++) out->dvdx[9 * (size_t)(lli - 1) + k] = (float)g[k];
    *nneightry += c->nneightry;
    *nneighact += c->nneigh[ip];
    if (c->nneightry > *maxneightry) *maxneightry = c->nneightry;
    if (c->nneigh[ip] > *maxneighact) *maxneighact = c->nneigh[ip];
  }
  *npc += c->npcell;
  *ncalc += (int64_t)c->npcell * c->nits;
}

/* densityiterate (dens.F90:117-568), non-MPI path.  nits_part (may be
 * NULL) receives per-particle the number of the iteration at which its own
 * h converged (1-based), for comparison with per-particle iteration. */
int orc_densityiterate(orc_ctx *o, int32_t icall, int32_t np, double *xyzh, const double *vxyzu,
                       int32_t ldv, const double *fxyzu, int32_t ldf, const double *fext,
                       const double *Bevol, const int8_t *iphase, float *divcurlv, int32_t lddcv,
                       float *divcurlB, float *alphaind, float *gradh, int32_t ldgradh, float *dvdx,
                       int32_t *nits_part, int32_t *nneigh_part) {
  if (!o->tree_built || np != o->np) {
    snprintf(o->err, sizeof(o->err), "densityiterate: tree not built for this np");
    return PHSPH_ERR_STATE;
  }
  double *h_in = (double *)malloc((size_t)np * sizeof(double));
  for (int32_t i = 0; i < np; i++) h_in[i] = xyzh[4 * (size_t)i + 3];
  densin_t in = {xyzh, vxyzu, fxyzu, fext, Bevol, h_in, ldv, ldf};
  densout_t out = {divcurlv, lddcv, divcurlB, alphaind, gradh, ldgradh, dvdx, xyzh};
  int64_t nneightry = 0, nneighact = 0, maxneightry = 0, ncalc = 0, npc = 0;
  int32_t maxneighact = 0;
  int rc_all = PHSPH_OK;
  const double pm = o->p.massoftype_gas;
  if (nits_part) for (int32_t i = 0; i < np; i++) nits_part[i] = 0;
  if (nneigh_part) for (int32_t i = 0; i < np; i++) nneigh_part[i] = 0;

#pragma omp parallel reduction(+ : nneightry, nneighact, ncalc, npc) reduction(max : maxneightry, maxneighact)
  {
    nlist_t l = {0};
    celldens_t c;
#pragma omp for schedule(dynamic, 64)
    for (int64_t icell = 1; icell <= o->ncells; icell++) {
      if (o->leaf_is_active[icell] <= 0) continue;
      if (rc_all) continue;
      const phsph_kdnode *nd = &o->node[icell];
      /* get_cell_location (neigh_kdtree.f90:354-365) */
      c.icell = (int32_t)icell;
      c.xpos[0] = nd->xcen[0]; c.xpos[1] = nd->xcen[1]; c.xpos[2] = nd->xcen[2];
      c.xsizei = nd->size;
      c.rcuti = o->radkern * nd->hmax;
      c.hmax = nd->hmax;
      c.nits = 0;
      int rc = getneigh(o, c.xpos, c.xsizei, c.rcuti, 0, &l);
      if (rc) { rc_all = rc; continue; }
      /* start_cell (dens.F90:1293-1378) */
      c.npcell = 0;
      for (int32_t ip = o->ir_lo[icell]; ip <= o->ir_hi[icell]; ip++) {
        int32_t i = o->inodeparts[ip];
        if (i < 0) continue;
        if (iphase && iphase[i - 1] <= 0) continue;
        int k = c.npcell++;
        c.arr_index[k] = ip;
        const double *r = &xyzh[4 * (size_t)(i - 1)];
        const double *v = &vxyzu[(size_t)ldv * (i - 1)];
        c.xpv[k][0] = r[0]; c.xpv[k][1] = r[1]; c.xpv[k][2] = r[2];
        c.h[k] = h_in[i - 1]; c.h_old[k] = h_in[i - 1];
        c.xpv[k][3] = v[0]; c.xpv[k][4] = v[1]; c.xpv[k][5] = v[2];
        c.xpv[k][6] = (o->p.maxvxyzu >= 4) ? v[3] : 0.;
        double fx = 0., fy = 0., fz = 0.;
        if (fxyzu) { const double *f = &fxyzu[(size_t)ldf * (i - 1)]; fx = f[0]; fy = f[1]; fz = f[2]; }
        if (fext) { const double *e = &fext[3 * (size_t)(i - 1)]; fx = fx + e[0]; fy = fy + e[1]; fz = fz + e[2]; }
        c.xpv[k][11] = fx; c.xpv[k][12] = fy; c.xpv[k][13] = fz;
        if (o->p.mhd) {
          const double *B = &Bevol[4 * (size_t)(i - 1)];
          c.xpv[k][7] = B[0]; c.xpv[k][8] = B[1]; c.xpv[k][9] = B[2]; c.xpv[k][10] = B[3];
        } else {
          c.xpv[k][7] = c.xpv[k][8] = c.xpv[k][9] = c.xpv[k][10] = 0.;
        }
      }
      int32_t conv_at[MINPART + 1];
      for (int k = 0; k < c.npcell; k++) conv_at[k] = 0;
      /* compute_cell (dens.F90:1203-1271) */
#define COMPUTE_CELL()                                                                          \
  for (int k = 0; k < c.npcell; k++) {                                                          \
    int32_t lli = o->inodeparts[c.arr_index[k]];                                                \
    double hi = c.h[k], hi1 = 1. / hi, hi21 = hi1 * hi1;                                        \
    get_density_sums(o, lli, c.xpv[k], hi, hi1, hi21, &l, &c.nneigh[k], c.rhosums[k], &in);      \
  }                                                                                             \
  c.nneightry = l.n;
      COMPUTE_CELL();
      int converged = 0;
      while (!converged) {
        /* finish_cell (dens.F90:1382-1466) */
        c.nits++;
        int cell_conv = 1;
        for (int k = 0; k < c.npcell; k++) {
          double hi = c.h[k], hi_old = c.h_old[k];
          double rhoi, gradhi, rhohi, gs, dhdrhoi, omegai;
          finish_rhosum(o, c.rhosums[k], pm, hi, 1, &rhoi, &gradhi, &rhohi, &gs, &dhdrhoi, &omegai);
          double func = rhohi - rhoi;
          double dfdh1 = (omegai > TINY64) ? dhdrhoi / omegai : dhdrhoi / fabs(omegai + EPS64);
          double hnew = hi - func * dfdh1;
          if (hnew > 1.2 * hi) hnew = 1.2 * hi;
          else if (hnew < 0.8 * hi) hnew = 0.8 * hi;
          int conv = ((fabs(hnew - hi) / hi_old) < o->p.tolh && omegai > 0. && hi > 0.);
          if (cell_conv) cell_conv = conv;
          if (!conv && c.nits >= MAXDENSITS) {
            snprintf(o->err, sizeof(o->err), "densityiterate: could not converge in density");
            rc_all = PHSPH_ERR_NOCONVERGE;
          }
          if (conv) { c.h[k] = hi; if (!conv_at[k]) conv_at[k] = c.nits; }
          else { c.h[k] = hnew; conv_at[k] = 0; }
        }
        converged = cell_conv;
        if (rc_all) break;
        /* compute_hmax (dens.F90:1275-1289) */
        int redo = 0;
        if (c.npcell > 0) {
          double hmx = c.h[0];
          for (int k = 1; k < c.npcell; k++) if (c.h[k] > hmx) hmx = c.h[k];
          double hmax = 1.01 * hmx;
          if (hmax > c.hmax) redo = 1;
          c.hmax = hmax;
          c.rcuti = o->radkern * hmax;
        }
        if (icall == 0) converged = 1;
        if (!converged) {
          if (redo) {
            /* set_hmaxcell (neigh_kdtree.f90:115-131) */
#pragma omp critical(orc_hmax)
            {
              int32_t nn = c.icell;
              o->node[nn].hmax = c.hmax;
              while (o->node[nn].parent != 0) {
                nn = o->node[nn].parent;
                if (c.hmax > o->node[nn].hmax) o->node[nn].hmax = c.hmax;
              }
            }
            rc = getneigh(o, c.xpos, c.xsizei, c.rcuti, 0, &l);
            if (rc) { rc_all = rc; break; }
          }
          COMPUTE_CELL();
        }
      }
#undef COMPUTE_CELL
      if (rc_all) continue;
      if (nits_part)
        for (int k = 0; k < c.npcell; k++)
          nits_part[o->inodeparts[c.arr_index[k]] - 1] = conv_at[k] ? conv_at[k] : c.nits;
      if (nneigh_part)
        for (int k = 0; k < c.npcell; k++) nneigh_part[o->inodeparts[c.arr_index[k]] - 1] = c.nneigh[k];
      store_results(o, &c, &out, &nneightry, &nneighact, &maxneightry, &maxneighact, &npc, &ncalc);
    }
    free(l.list); free(l.cache);
  }
  free(h_in);
  o->nneightry = nneightry; o->nneighact = nneighact; o->maxneightry = maxneightry;
  o->maxneighact = maxneighact; o->ncalc = ncalc; o->nptot = npc;
  return rc_all;
}

int orc_get_neighbour_stats(const orc_ctx *o, double *trialmean, double *actualmean, int32_t *maxtrial,
                            int32_t *maxactual, int64_t *nrhocalc, int64_t *nactualtot) {
  if (o->nptot > 0) {
    *trialmean = (double)o->nneightry / (double)o->nptot;
    *actualmean = (double)o->nneighact / (double)o->nptot;
    *maxtrial = (int32_t)o->maxneightry;
    *maxactual = o->maxneighact;
    *nrhocalc = o->ncalc;
    *nactualtot = o->nneighact;
  } else {
    *trialmean = *actualmean = -1; *maxtrial = *maxactual = -1; *nrhocalc = *nactualtot = -1;
  }
  return PHSPH_OK;
}

/* ------------------------------------------------------------------ */
/* cons2prim                                                           */
/* ------------------------------------------------------------------ */

/* xi_limiter (shock_capturing.f90:150-178) */
static double xi_limiter(const float *dv) {
  double divv = (double)dv[0] + (double)dv[4] + (double)dv[8];
  double cx = (double)dv[7] - (double)dv[5];
  double cy = (double)dv[2] - (double)dv[6];
  double cz = (double)dv[3] - (double)dv[1];
  double m = (-divv > 0.) ? -divv : 0.;
  double fac = m * m;
  double traceS = cx * cx + cy * cy + cz * cz;
  if (fac + traceS > EPS64) return fac / (fac + traceS);
  return 1.;
}

/* get_alphaloc (shock_capturing.f90:131-144) */
static double get_alphaloc(double divvdti, double cs, double hi, double xil, double amin, double amax) {
  double m = (-divvdti > 0.) ? -divvdti : 0.;
  double source = 10. * (hi * hi) * xil * m;
  double temp = cs * cs;
  if (temp > EPS64) {
    double a = source / temp;
    if (a > amax) a = amax;
    if (a < amin) a = amin;
    return a;
  }
  return amin;
}

/* equationofstate ieos 1 and 2 (eos.f90:184-240): returns P/rho_gas, c_s */
static void eos12(const orc_ctx *o, double rhogas, double u, double *ponrho, double *cs) {
  if (o->p.ieos == 1) {
    *ponrho = o->p.polyk;
    *cs = sqrt(*ponrho);
  } else {
    double g = o->p.gamma;
    if (o->p.maxvxyzu >= 4) {
      *ponrho = (g > 1.0001) ? (g - 1.) * u : 2. / 3. * u;
    } else {
      *ponrho = o->p.polyk * pow(rhogas, g - 1.);
    }
    *cs = sqrt(g * *ponrho);
  }
}

/* cons2prim_everything (cons2prim.f90:274-456) */
int orc_cons2prim(orc_ctx *o, int32_t np, const double *xyzh, const double *vxyzu, int32_t ldv,
                  const float *dvdx, float *alphaind, const double *dustevol, double *dustfrac,
                  double *eos_vars, int32_t ldeos, const int8_t *iphase) {
  (void)iphase;
  const double pm = o->p.massoftype_gas;
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < np; i++) {
    double hi = xyzh[4 * (size_t)i + 3];
    if (hi < TINY64) continue;
    double rhoi = rhoh(o, hi, pm);
    double rhogas = rhoi;
    if (o->p.use_dustfrac && o->p.ndustsmall > 0) {
      double s = dustevol[i];
      double eps = s * s / (1. + s * s);
      dustfrac[i] = eps;
      rhogas = rhoi * (1. - eps);
    }
    double u = (o->p.maxvxyzu >= 4) ? vxyzu[(size_t)ldv * i + 3] : 0.;
    double ponrho, cs;
    eos12(o, rhogas, u, &ponrho, &cs);
    eos_vars[(size_t)ldeos * i + 0] = ponrho * rhogas;
    eos_vars[(size_t)ldeos * i + 1] = cs;
    if (ldeos > 2) eos_vars[(size_t)ldeos * i + 2] = 0.;
    if (alphaind) {
      double xil = xi_limiter(&dvdx[9 * (size_t)i]);
      alphaind[3 * (size_t)i + 1] =
          (float)get_alphaloc((double)alphaind[3 * (size_t)i + 2], cs, hi, xil, o->p.alpha, o->p.alphamax);
    }
  }
  return PHSPH_OK;
}

/* ------------------------------------------------------------------ */
/* force                                                               */
/* ------------------------------------------------------------------ */

/* get_ts (dust.f90:161-276) with dv2 = 0 (the one-fluid calls,
 * force.F90:1793,2357) for idrag = 1 (Epstein/Stokes, dust.f90:188-245;
 * cste_mu and coeff_gei_1 of init_drag, dust.f90:78-79), 2 (constant K)
 * and 3 (constant stopping time) */
static double get_ts(const orc_ctx *o, double rhogas, double rhodust, double spsoundgas) {
  if (o->p.idrag == 1) {
    const double sgrain = o->p.grainsize, densgrain = o->p.graindens;
    const double cste_mu = sqrt(2. / (PI_PH * o->p.gamma)), coeff_gei_1 = sqrt(8. / (PI_PH * o->p.gamma));
    double rhosum = rhogas + rhodust;
    double lambda = o->p.seff / rhogas;
    double kn_eff = (sgrain > 0.) ? 9. * lambda / (4. * sgrain) : HUGE64;
    double dragcoeff = 0., f = 0., abs_dv = 0., dv2 = 0.;
    if (kn_eff >= 1.) {
      dragcoeff = (densgrain > TINY64) ? coeff_gei_1 * spsoundgas / (densgrain * sgrain) : HUGE64;
      f = 1.;  /* dv2 = 0: no Kwok correction */
      (void)dv2;
    } else {
      double viscmol_nu = cste_mu * lambda * spsoundgas;
      double re_dust = 2. * sgrain * abs_dv / viscmol_nu;
      if (re_dust <= 1.) {
        dragcoeff = 4.5 * viscmol_nu / (densgrain * sgrain * sgrain);
        f = 1.;
      } else if (re_dust <= 800.) {
        dragcoeff = 9. / (densgrain * sgrain * pow(re_dust, 0.6));
        f = abs_dv;
      } else {
        dragcoeff = 0.163075 / (densgrain * sgrain);
        f = abs_dv;
      }
    }
    double ts1 = (dragcoeff == HUGE64) ? HUGE64 : dragcoeff * f * rhosum;
    return (ts1 > 0.) ? 1. / ts1 : HUGE64;
  }
  if (o->p.idrag == 2) {
    double rhosum = rhogas + rhodust;
    if (o->p.K_code > 0.) return rhogas * rhodust / (o->p.K_code * rhosum);
    return HUGE64;
  }
  if (o->p.idrag == 3) return o->p.K_code;
  return 0.;
}

typedef struct {
  /* per active particle of the cell (force.F90:2172-2514) */
  double x, y, z, h, vx, vy, vz, u, rho, rhogas, pr, cs, pro2, vwave, alpha, gradh;
  double Bx_e, By_e, Bz_e, psi, divB;   /* Bevol and divcurlB(1) */
  double sxx, sxy, sxz, syy, syz, szz;  /* MHD stress / rho^2 terms */
  double eps, tstop;
  double gradsoft;                      /* gradh(2) (self-gravity) */
} fpart_t;

/* get_stress (force.F90:2068-2168) without physical viscosity */
static void get_stress(const orc_ctx *o, double pr, double cs, double rho, double rho1, double pm,
                       double Bx, double By, double Bz, double *pro2, double *vwave, double s[6]) {
  s[0] = s[1] = s[2] = s[3] = s[4] = s[5] = 0.;
  if (o->p.mhd) {
    double Brx = Bx * rho1, Bry = By * rho1, Brz = Bz * rho1;
    double Bro2 = Brx * Brx + Bry * Bry + Brz * Brz;
    double va2 = Bro2 * rho;
    *vwave = sqrt(cs * cs + va2);
    s[0] = s[0] - pm * Brx * Brx;
    s[1] = s[1] - pm * Brx * Bry;
    s[2] = s[2] - pm * Brx * Brz;
    s[3] = s[3] - pm * Bry * Bry;
    s[4] = s[4] - pm * Bry * Brz;
    s[5] = s[5] - pm * Brz * Brz;
    *pro2 = pr * rho1 * rho1 + 0. + 0.5 * Bro2;
  } else {
    *pro2 = pr * rho1 * rho1 + 0.;
    *vwave = cs;
  }
}

typedef struct {
  const double *xyzh, *vxyzu, *Bevol, *dustfrac, *eos_vars;
  const float *gradh, *alphaind, *divcurlB;
  int32_t ldv, ldgradh, ldeos;
} forcein_t;

enum { F_FX = 0, F_FY, F_FZ, F_POT, F_DRHODT, F_DUDTDISS, F_DENDTDISS, F_DIVBSYM, F_DBX, F_DBY,
       F_DBZ, F_DIVBDIFF, F_DDUSTEVOL, F_DUDTDUST, F_DVX, F_DVY, F_DVZ, NFSUM };

/* compute_forces (force.F90:914-2060) for one active gas particle. */
static void compute_forces(const orc_ctx *o, int32_t i, const fpart_t *pi, const nlist_t *l,
                           const forcein_t *in, double *fsum, double *vsigmax, int *ibin_neighi) {
  const int kind = o->p.kernel;
  const double pm = o->p.massoftype_gas;
  const double beta = o->p.beta, alphau = o->p.alphau, alphaB = o->p.alphaB;
  const int mhd = o->p.mhd, dust = o->p.use_dustfrac && o->p.ndustsmall > 0;
  const int thermal = o->p.maxvxyzu >= 4;
  const int grav = o->p.gravity != 0;
  double hi = pi->h, hi1 = 1. / hi, hi21 = hi1 * hi1, hi41 = hi21 * hi21;
  double gradhi = pi->gradh;
  double rhoi = pi->rho, rho1i = 1. / rhoi;
  double hfacgrkern = hi41 * o->cnormk * gradhi;
  double Bxi = 0., Byi = 0., Bzi = 0., psii = 0.;
  if (mhd) { Bxi = pi->Bx_e * rhoi; Byi = pi->By_e * rhoi; Bzi = pi->Bz_e * rhoi; psii = pi->psi; }
  double epsi = 0., epssumi = 0., tsi = 0., epstsi = 0., sqrtrhoepsi = 0.;
  if (dust) {
    epsi = pi->eps; epssumi = epsi;
    tsi = pi->tstop;
    if (o->p.ilimitdustflux) { double lim = hi / pi->cs; if (lim < tsi) tsi = lim; }
    epstsi = epsi * tsi;
    sqrtrhoepsi = sqrt(epsi / (1. - epsi));
  }
  double rho21i = rho1i * rho1i;
  double mrhoi5 = 0.5 * pm * rho1i;
  double auterm = mrhoi5 * alphau;
  double avBterm = mrhoi5 * alphaB * rho1i;
  const int useresistiveheat = o->p.iresistive_heating > 0;
  memset(fsum, 0, NFSUM * sizeof(double));
  *vsigmax = 0.;
  *ibin_neighi = 0;
  const int ind = o->p.ind_timesteps && o->ibin_old;
  const int8_t ibinnow_m1 = (int8_t)(o->ibinnow - 1);
  for (int32_t n = 0; n < l->n; n++) {
    int32_t j = l->list[n];
    if (j == i) continue;
    double xj, yj, zj, hj1;
    if (n < ISIZECELLCACHE) {
      xj = l->cache[4 * n]; yj = l->cache[4 * n + 1]; zj = l->cache[4 * n + 2]; hj1 = l->cache[4 * n + 3];
    } else {
      const double *r = &in->xyzh[4 * (size_t)(j - 1)];
      xj = r[0]; yj = r[1]; zj = r[2]; hj1 = 1. / r[3];
    }
    double dx = pi->x - xj, dy = pi->y - yj, dz = pi->z - zj;
    if (o->p.periodic) {
      if (fabs(dx) > 0.5 * o->dxbound) dx = dx - copysign(o->dxbound, dx);
      if (fabs(dy) > 0.5 * o->dybound) dy = dy - copysign(o->dybound, dy);
      if (fabs(dz) > 0.5 * o->dzbound) dz = dz - copysign(o->dzbound, dz);
    }
    double rij2 = dx * dx + dy * dy + dz * dz;
    double q2i = rij2 * hi21;
    double hj21 = hj1 * hj1;
    double q2j = rij2 * hj21;
    if (!(q2i < o->radkern2 || q2j < o->radkern2)) {
      if (grav) {
        /* trial neighbour outside both kernels: 1/r^2 (force.F90:1991-2053) */
        double rij1 = 1. / sqrt(rij2);
        double fgrav = rij1 * rij1 * rij1;
        double phii = -rij1;
        double fgravj = fgrav * pm;
        fsum[F_FX] = fsum[F_FX] - dx * fgravj;
        fsum[F_FY] = fsum[F_FY] - dy * fgravj;
        fsum[F_FZ] = fsum[F_FZ] - dz * fgravj;
        fsum[F_POT] = fsum[F_POT] + pm * phii;
      }
      continue;
    }
    double rij1, qi;
    if (rij2 > TINY64) { rij1 = 1. / sqrt(rij2); qi = (rij2 * rij1) * hi1; }
    else { rij1 = 0.; qi = 0.; }
    double grkerni = 0., grkernj = 0.;
    double phii = 0., fgravi = 0., fgravj_s = 0.;
    if (q2i < o->radkern2) {
      grkerni = grkern(kind, q2i, qi) * hfacgrkern;
      if (grav) {
        double fmi;
        kernel_softening(kind, q2i, qi, &phii, &fmi);
        phii = phii * hi1;
        fmi = fmi * hi21;
        double dsofti = pi->gradsoft * grkerni;
        fgravi = fmi + dsofti;
      }
    } else if (grav) {
      phii = -rij1;
      fgravi = rij1 * rij1;
    }
    int usej;
    const double gradhj = (double)in->gradh[(size_t)in->ldgradh * (j - 1)];
    if (q2j < o->radkern2) {
      double qj = (rij2 * rij1) * hj1;
      grkernj = grkern(kind, q2j, qj) * hj21 * hj21 * o->cnormk * gradhj;
      usej = 1;
      if (grav) {
        double phij, fmj;
        kernel_softening(kind, q2j, qj, &phij, &fmj);
        fmj = fmj * hj21;
        double dsoftj = (double)in->gradh[(size_t)in->ldgradh * (j - 1) + 1] * grkernj;
        fgravj_s = fmj + dsoftj;
      }
    } else {
      grkernj = 0.; usej = 0;
      if (grav) fgravj_s = rij1 * rij1;
    }
    double fgrav = grav ? 0.5 * (pm * fgravi + pm * fgravj_s) : 0.;
    if (mhd) usej = 1;
    if (dust) usej = 1;
    if (thermal && !o->p.gravity) usej = 1;
    if (ind) {
      /* force.F90:1346-1356: j is a neighbour of an active particle; flag
         it to be woken, and note the largest bin among the neighbours */
      if (o->ibin_wake[j - 1] < ibinnow_m1) {
#pragma omp atomic write
        o->ibin_wake[j - 1] = ibinnow_m1;
      }
      if (o->ibin_old[j - 1] > *ibin_neighi) *ibin_neighi = o->ibin_old[j - 1];
    }
    double runix = dx * rij1, runiy = dy * rij1, runiz = dz * rij1;
    const double *vj = &in->vxyzu[(size_t)in->ldv * (j - 1)];
    double dvx = pi->vx - vj[0], dvy = pi->vy - vj[1], dvz = pi->vz - vj[2];
    double projv = dvx * runix + dvy * runiy + dvz * runiz;
    double denij = 0.;
    if (thermal) denij = pi->u - vj[3];
    double vsigi = pi->vwave - beta * projv; if (vsigi < 0.) vsigi = 0.;
    double vsigavi = pi->alpha * pi->vwave - beta * projv; if (vsigavi < 0.) vsigavi = 0.;
    if (vsigi > *vsigmax) *vsigmax = vsigi;
    double Bxj = 0., Byj = 0., Bzj = 0., psij = 0., dBx = 0., dBy = 0., dBz = 0., projBi = 0., projBj = 0.,
           projdB = 0., dB2 = 0., divBdiffterm = 0.;
    if (mhd) {
      double hjm = in->xyzh[4 * (size_t)(j - 1) + 3];
      double rhojm = rhoh(o, hjm, pm);
      const double *Bj = &in->Bevol[4 * (size_t)(j - 1)];
      Bxj = Bj[0] * rhojm; Byj = Bj[1] * rhojm; Bzj = Bj[2] * rhojm; psij = Bj[3];
      dBx = Bxi - Bxj; dBy = Byi - Byj; dBz = Bzi - Bzj;
      projBi = Bxi * runix + Byi * runiy + Bzi * runiz;
      if (usej) projBj = Bxj * runix + Byj * runiy + Bzj * runiz;
      projdB = dBx * runix + dBy * runiy + dBz * runiz;
      dB2 = dBx * dBx + dBy * dBy + dBz * dBz;
      divBdiffterm = -pm * projdB * grkerni;
    }
    double rhoj = 0., rho1j = 0., rho21j = 0., mrhoj5 = 0., autermj = 0., avBtermj = 0., prj = 0., pro2j = 0.,
           vwavej = 0., vsigavj = 0., epsj = 0., sqrtrhoepsj = 0., rhogasj = 0., csj = 0.;
    double sj[6] = {0, 0, 0, 0, 0, 0};
    if (usej) {
      double hj = 1. / hj1;
      rhoj = rhoh(o, hj, pm);
      rho1j = 1. / rhoj;
      rho21j = rho1j * rho1j;
      if (dust) {
        epsj = in->dustfrac[j - 1];
        rhogasj = rhoj * (1. - epsj);
        sqrtrhoepsj = sqrt(epsj / (1. - epsj));
      } else {
        rhogasj = rhoj;
      }
      double alphaj = (double)in->alphaind[3 * (size_t)(j - 1)];
      prj = in->eos_vars[(size_t)in->ldeos * (j - 1)];
      csj = in->eos_vars[(size_t)in->ldeos * (j - 1) + 1];
      get_stress(o, prj, csj, rhoj, rho1j, pm, Bxj, Byj, Bzj, &pro2j, &vwavej, sj);
      mrhoj5 = 0.5 * pm * rho1j;
      autermj = mrhoj5 * alphau;
      avBtermj = mrhoj5 * alphaB * rho1j;
      double vsigj = vwavej - beta * projv; if (vsigj < 0.) vsigj = 0.;
      vsigavj = alphaj * vwavej - beta * projv; if (vsigavj < 0.) vsigavj = 0.;
      if (vsigj > *vsigmax) *vsigmax = vsigj;
    }
    /* gas-gas pair */
    double qrho2i = 0., qrho2j = 0.;
    if (projv < 0.) {
      qrho2i = -0.5 * rho1i * vsigavi * projv;
      if (usej) qrho2j = -0.5 * rho1j * vsigavj * projv;
    }
    double dudtdissi = pm * qrho2i * projv * grkerni;
    double gradpi = pm * (pi->pro2 + qrho2i) * grkerni;
    double gradpj = 0.;
    if (usej) gradpj = pm * (pro2j + qrho2j) * grkernj;
    double dendissterm = 0.;
    if (thermal) {
      double vsigu;
      if (o->p.gravity) vsigu = fabs(projv);
      else {
        double rhoav1 = 2. / (rhoi + rhoj);
        vsigu = sqrt(fabs(pi->pr - prj) * rhoav1);
      }
      dendissterm = vsigu * denij * (auterm * grkerni + autermj * grkernj);
    }
    double dBevolx = 0., dBevoly = 0., dBevolz = 0., divBsymmterm = 0., dudtresist = 0.;
    if (mhd) {
      double ax = dvx - projv * runix, ay = dvy - projv * runiy, az = dvz - projv * runiz;
      double vsigB = sqrt(ax * ax + ay * ay + az * az);
      double dBdissterm = (avBterm * grkerni + avBtermj * grkernj) * vsigB;
      if (useresistiveheat) dudtresist = -0.5 * dB2 * dBdissterm;
      double pmjrho21grkerni = pm * rho21i * grkerni;
      double pmjrho21grkernj = pm * rho21j * grkernj;
      double termi = pmjrho21grkerni * projBi;
      divBsymmterm = termi + pmjrho21grkernj * projBj;
      double dBrhoterm = -termi;
      double dpsiterm = o->p.overcleanfac * (pmjrho21grkerni * psii * pi->vwave + pmjrho21grkernj * psij * vwavej);
      dBevolx = dBrhoterm * dvx + dBdissterm * dBx - dpsiterm * runix;
      dBevoly = dBrhoterm * dvy + dBdissterm * dBy - dpsiterm * runiy;
      dBevolz = dBrhoterm * dvz + dBdissterm * dBz - dpsiterm * runiz;
    }
    double projsxi = (pi->sxx * runix + pi->sxy * runiy + pi->sxz * runiz) * grkerni;
    double projsyi = (pi->sxy * runix + pi->syy * runiy + pi->syz * runiz) * grkerni;
    double projszi = (pi->sxz * runix + pi->syz * runiy + pi->szz * runiz) * grkerni;
    double projsxj = 0., projsyj = 0., projszj = 0.;
    if (usej) {
      projsxj = (sj[0] * runix + sj[1] * runiy + sj[2] * runiz) * grkernj;
      projsyj = (sj[1] * runix + sj[3] * runiy + sj[4] * runiz) * grkernj;
      projszj = (sj[2] * runix + sj[4] * runiy + sj[5] * runiz) * grkernj;
    }
    double gradp = gradpi + gradpj;
    fsum[F_FX] = fsum[F_FX] - runix * (gradp + fgrav) - (projsxi + projsxj);
    fsum[F_FY] = fsum[F_FY] - runiy * (gradp + fgrav) - (projsyi + projsyj);
    fsum[F_FZ] = fsum[F_FZ] - runiz * (gradp + fgrav) - (projszi + projszj);
    if (grav) fsum[F_POT] = fsum[F_POT] + pm * phii;
    fsum[F_DRHODT] = fsum[F_DRHODT] + projv * grkerni;
    if (thermal) {
      fsum[F_DUDTDISS] = fsum[F_DUDTDISS] + dudtdissi + dudtresist;
      fsum[F_DENDTDISS] = fsum[F_DENDTDISS] + dendissterm;
    }
    if (mhd) {
      fsum[F_DIVBSYM] += divBsymmterm;
      fsum[F_DBX] += dBevolx; fsum[F_DBY] += dBevoly; fsum[F_DBZ] += dBevolz;
      fsum[F_DIVBDIFF] += divBdiffterm;
    }
    if (dust) {
      double tsj = get_ts(o, rhogasj, rhoj * epsj, csj);
      if (o->p.ilimitdustflux) { double lim = (1. / hj1) / csj; if (lim < tsj) tsj = lim; }
      double epstsj = epsj * tsj;
      if (epsi > 0. || epsj > 0.) {
        double grkernav = 0.5 * (grkerni + grkernj);
        double dft = pm * sqrtrhoepsj * rho1j *
                     ((tsi - epstsi) * (1. - epsi) / (1. - epssumi) + (tsj - epstsj) * (1. - epsj) / (1. - epsj)) *
                     (pi->pr - prj) * grkernav * rij1;
        fsum[F_DDUSTEVOL] = fsum[F_DDUSTEVOL] - dft;
        if (thermal) fsum[F_DUDTDUST] = fsum[F_DUDTDUST] - sqrtrhoepsi * dft * denij;
      }
      if (epsi < 1.) {
        double term = tsi / (1. - epssumi) * pm * (pi->pro2 * grkerni + pro2j * grkernj);
        fsum[F_DVX] += term * runix; fsum[F_DVY] += term * runiy; fsum[F_DVZ] += term * runiz;
      }
    }
  }
}

void orc_get_newbin(double dti, double dtmax, int8_t *ibini, int allow_decrease, int32_t nbinmax,
                    int32_t istepfrac);

/* force (force.F90:193-905) + start_cell (:2172-2514) +
 * finish_cell_and_store_results (:2649-3330), global timestep build. */
int orc_force(orc_ctx *o, int32_t icall, int32_t np, const double *xyzh, const double *vxyzu, int32_t ldv,
              double *fxyzu, int32_t ldf, float *divcurlv, int32_t lddcv, const float *divcurlB,
              const double *Bevol, double *dBevol, float *divBsymm, const double *dustfrac, double *ddustevol,
              double *deltav, double *tstop, const float *gradh, int32_t ldgradh, const float *alphaind,
              const double *eos_vars, int32_t ldeos, const int8_t *iphase, double *dtcourant_out,
              double *dtforce_out) {
  if (!o->tree_built || np != o->np) {
    snprintf(o->err, sizeof(o->err), "force: tree not built for this np");
    return PHSPH_ERR_STATE;
  }
  const double pm = o->p.massoftype_gas;
  const int mhd = o->p.mhd, dust = o->p.use_dustfrac && o->p.ndustsmall > 0;
  const int thermal = o->p.maxvxyzu >= 4;
  forcein_t in = {xyzh, vxyzu, Bevol, dustfrac, eos_vars, gradh, alphaind, divcurlB, ldv, ldgradh, ldeos};
  double dtcourant = BIGNUMBER, dtforce = BIGNUMBER;
  int nbinmaxnew = 0, ncheckbin = 0;
  const int ind = o->p.ind_timesteps != 0;
  if (ind && (!o->ibin || !o->ibin_old || !o->ibin_wake)) {
    snprintf(o->err, sizeof(o->err), "force: individual timesteps need orc_set_bins");
    return PHSPH_ERR_STATE;
  }
  int rc_all = PHSPH_OK;
  int64_t maxtry = 0;
#pragma omp parallel reduction(min : dtcourant, dtforce) reduction(max : nbinmaxnew, maxtry) reduction(+ : ncheckbin)
  {
    nlist_t l = {0};
#pragma omp for schedule(dynamic, 64)
    for (int64_t icell = 1; icell <= o->ncells; icell++) {
      if (o->leaf_is_active[icell] <= 0) continue;
      if (rc_all) continue;
      fpart_t parts[MINPART + 1];
      int32_t gidx[MINPART + 1];
      int npcell = 0;
      for (int32_t ip = o->ir_lo[icell]; ip <= o->ir_hi[icell]; ip++) {
        int32_t i = o->inodeparts[ip];
        if (i < 0) continue;
        if (iphase && iphase[i - 1] <= 0) continue;
        fpart_t *q = &parts[npcell];
        gidx[npcell] = i;
        npcell++;
        const double *r = &xyzh[4 * (size_t)(i - 1)];
        const double *v = &vxyzu[(size_t)ldv * (i - 1)];
        double hi = r[3];
        /* rhoanddhdrho (part.F90:805-816) */
        double hi1 = 1. / fabs(hi);
        double t = o->p.hfact * hi1;
        double rhoi = pm * (t * t * t);
        double rho1i = 1. / rhoi;
        q->x = r[0]; q->y = r[1]; q->z = r[2]; q->h = hi;
        q->vx = v[0]; q->vy = v[1]; q->vz = v[2]; q->u = thermal ? v[3] : 0.;
        q->rho = rhoi;
        q->eps = dust ? dustfrac[i - 1] : 0.;
        q->rhogas = rhoi * (1. - q->eps);
        q->pr = eos_vars[(size_t)ldeos * (i - 1)];
        q->cs = eos_vars[(size_t)ldeos * (i - 1) + 1];
        if (mhd) {
          const double *B = &Bevol[4 * (size_t)(i - 1)];
          q->Bx_e = B[0]; q->By_e = B[1]; q->Bz_e = B[2]; q->psi = B[3];
          q->divB = (double)divcurlB[4 * (size_t)(i - 1)];
        } else {
          q->Bx_e = q->By_e = q->Bz_e = q->psi = q->divB = 0.;
        }
        double s[6];
        get_stress(o, q->pr, q->cs, rhoi, rho1i, pm, q->Bx_e * rhoi, q->By_e * rhoi, q->Bz_e * rhoi, &q->pro2,
                   &q->vwave, s);
        q->sxx = s[0]; q->sxy = s[1]; q->sxz = s[2]; q->syy = s[3]; q->syz = s[4]; q->szz = s[5];
        q->tstop = dust ? get_ts(o, q->rhogas, rhoi * q->eps, q->cs) : 0.;
        q->alpha = (double)alphaind[3 * (size_t)(i - 1)];
        double g = (double)gradh[(size_t)ldgradh * (i - 1)];
        q->gradh = (g > 0.) ? g : 1.;
        q->gradsoft = (o->p.gravity && ldgradh > 1) ? (double)gradh[(size_t)ldgradh * (i - 1) + 1] : 0.;
      }
      if (npcell == 0) continue;
      const phsph_kdnode *nd = &o->node[icell];
      double fnode[LENFGRAV];
      int rc = o->p.gravity ? getneigh_dual(o, (int32_t)icell, &l, fnode)
                            : getneigh(o, nd->xcen, nd->size, o->radkern * nd->hmax, 1, &l);
      if (rc) { rc_all = rc; continue; }
      if ((int64_t)l.n > maxtry) maxtry = l.n;
      for (int k = 0; k < npcell; k++) {
        const fpart_t *q = &parts[k];
        int32_t i = gidx[k];
        double fsum[NFSUM], vsigmax;
        int ibin_neighi;
        compute_forces(o, i, q, &l, &in, fsum, &vsigmax, &ibin_neighi);
        /* finish_cell_and_store_results */
        double hi = q->h, hi1 = 1. / hi, rhoi = q->rho, rho1i = 1. / rhoi;
        double dtc = o->p.dtmax, dtf = BIGNUMBER, dtclean = BIGNUMBER, dtdusti = BIGNUMBER;
        if (o->p.gravity) {
          /* self-contribution and the far field (force.F90:2909-2927) */
          double potensoft0, dum;
          kernel_softening(o->p.kernel, 0., 0., &potensoft0, &dum);
          double epoti = 0.5 * pm * (fsum[F_POT] + pm * potensoft0 * hi1);
          double fxi, fyi, fzi, poti;
          expand_fgrav(fnode, q->x - nd->xcen[0], q->y - nd->xcen[1], q->z - nd->xcen[2], &fxi, &fyi, &fzi, &poti);
          fsum[F_FX] = fsum[F_FX] + fxi;
          fsum[F_FY] = fsum[F_FY] + fyi;
          fsum[F_FZ] = fsum[F_FZ] + fzi;
          epoti = epoti + 0.5 * pm * poti;
          if (o->poten) o->poten[i - 1] = (float)epoti;
        }
        if (mhd) {
          double Bx = q->Bx_e * rhoi, By = q->By_e * rhoi, Bz = q->Bz_e * rhoi;
          double B2i = Bx * Bx + By * By + Bz * Bz;
          double divBsymmi = fsum[F_DIVBSYM];
          double frac_divB = 0.;
          if (B2i > 0.) {
            double betai = 2. * q->pr / B2i;
            if (betai < 2.) frac_divB = 1.;
            else if (betai < 10.) frac_divB = (10. - betai) * 0.125;
          }
          fsum[F_FX] = fsum[F_FX] - Bx * divBsymmi * frac_divB;
          fsum[F_FY] = fsum[F_FY] - By * divBsymmi * frac_divB;
          fsum[F_FZ] = fsum[F_FZ] - Bz * divBsymmi * frac_divB;
          if (divBsymm) divBsymm[i - 1] = (float)(rhoi * divBsymmi);
        }
        double f2i = fsum[F_FX] * fsum[F_FX] + fsum[F_FY] * fsum[F_FY] + fsum[F_FZ] * fsum[F_FZ];
        double *fo = &fxyzu[(size_t)ldf * (i - 1)];
        fo[0] = fsum[F_FX]; fo[1] = fsum[F_FY]; fo[2] = fsum[F_FZ];
        double drhodti = pm * fsum[F_DRHODT];
        double divvi = -drhodti * rho1i;
        divcurlv[(size_t)lddcv * (i - 1)] = (float)divvi;
        if (thermal) {
          double fac = (q->rhogas > TINY64) ? rhoi / q->rhogas : 0.;
          double pdv_work = fac * q->pr * rho1i * rho1i * drhodti;
          double fxyz4 = 0.;
          if (o->p.ipdv_heating > 0) fxyz4 = fxyz4 + pdv_work;
          if (o->p.ishock_heating > 0) fxyz4 = fxyz4 + fac * fsum[F_DUDTDISS];
          fxyz4 = fxyz4 + fac * fsum[F_DENDTDISS];
          if (dust) fxyz4 = fxyz4 + 0.5 * fac * rho1i * fsum[F_DUDTDUST];
          fo[3] = fxyz4;
        }
        if (mhd) {
          double *dB = &dBevol[4 * (size_t)(i - 1)];
          dB[0] = fsum[F_DBX]; dB[1] = fsum[F_DBY]; dB[2] = fsum[F_DBZ];
          if (o->p.psidecayfac > 0.) {
            double vcleani = o->p.overcleanfac * q->vwave;
            double dtau = o->p.psidecayfac * vcleani * hi1;
            dB[3] = -vcleani * fsum[F_DIVBDIFF] * rho1i - q->psi * dtau - 0.5 * q->psi * divvi;
            dtclean = o->p.C_cour * hi / (vcleani + TINY64);
          }
        }
        if (dust) {
          double om = 1. - q->eps;
          ddustevol[i - 1] = 0.5 * (fsum[F_DDUSTEVOL] * rho1i / (om * om));
          double *dv = &deltav[3 * (size_t)(i - 1)];
          dv[0] = fsum[F_DVX]; dv[1] = fsum[F_DVY]; dv[2] = fsum[F_DVZ];
        }
        double vsigdtc = (vsigmax > q->vwave) ? vsigmax : q->vwave;
        if (vsigdtc > TINY64) {
          double am = (o->p.alpha > 1.) ? o->p.alpha : 1.;
          dtc = o->p.C_cour * hi / (vsigdtc * am);
        }
        if (thermal) {
          double eni = q->u;
          if (eni + dtc * fo[3] < EPS64 && eni > EPS64) fo[3] = fo[3] / (1. - dtc * fo[3] / eni);
        }
        if (fabs(f2i) > EPS64) dtf = o->p.C_force * sqrt(hi / sqrt(f2i));
        if (dust) {
          if (q->eps > 0. && q->cs > 0. && q->eps > EPS64) {
            double tseff = (1. - q->eps) / q->eps * (q->eps * q->tstop);
            double dtdustdenom = q->eps * tseff * q->cs * q->cs;
            if (dtdustdenom > TINY64) dtdusti = o->p.C_force * hi * hi / dtdustdenom;
          }
          if (tstop) tstop[i - 1] = q->tstop;
        }
        if (ind) {
          /* force.F90:3272-3310: per-particle timestep, new bin, limiter */
          double dti = dtc;
          double dtitmp = dtf;
          if (dtclean < dtitmp) dtitmp = dtclean;
          if (dtdusti < dtitmp) dtitmp = dtdusti;
          if (dtitmp < dti + TINY64 && dtitmp < o->p.dtmax) dti = dtitmp;
          int8_t b = o->ibin[i - 1];
          orc_get_newbin(dti, o->p.dtmax, &b, icall < 2, o->nbinmax, o->istepfrac);
          if (b < ibin_neighi - 1) b = (int8_t)(ibin_neighi - 1);
          o->ibin[i - 1] = b;
          if (b > nbinmaxnew) nbinmaxnew = b;
          ncheckbin++;
        } else {
          if (dtc < dtcourant) dtcourant = dtc;
          double m = dtf;
          if (dtdusti < m) m = dtdusti;
          if (dtclean < m) m = dtclean;
          if (m < dtforce) dtforce = m;
        }
      }
    }
    free(l.list); free(l.cache);
  }
  if (dtcourant_out) *dtcourant_out = dtcourant;
  if (dtforce_out) *dtforce_out = dtforce;
  /* no active particle: nbinmax is kept (force.F90:748-756) */
  o->nbinmaxnew = (ncheckbin == 0) ? o->nbinmax : nbinmaxnew;
  o->maxforcetrial = maxtry;
  return rc_all;
}

/* get_newbin (utils_indtimesteps.f90:230-287) */
void orc_get_newbin(double dti, double dtmax, int8_t *ibini, int allow_decrease, int32_t nbinmax,
                    int32_t istepfrac) {
  const double dlog2 = 1.4426950408889634; /* 1/log(2) */
  int8_t ibin_oldi = *ibini;
  int ibin_newi;
  if (dti > dtmax) ibin_newi = 0;
  else if (dti < TINY64) ibin_newi = MAXBINS;
  else {
    ibin_newi = (int)(log(2. * dtmax / dti) * dlog2 - EPS64);
    if (ibin_newi < 0) ibin_newi = 0;
  }
  if (ibin_newi > ibin_oldi) {
    *ibini = (int8_t)ibin_newi;
  } else if (ibin_newi < ibin_oldi && ibin_oldi <= nbinmax && allow_decrease) {
    /* timestep can only go up if bins are synchronised, one bin at a time */
    if (istepfrac % (1 << (nbinmax - (ibin_oldi - 1))) == 0) *ibini = (int8_t)(ibin_oldi - 1);
  }
}

void orc_set_bins(orc_ctx *o, int8_t *ibin, const int8_t *ibin_old, int8_t *ibin_wake, int32_t nbinmax,
                  int32_t istepfrac, int32_t ibinnow) {
  o->ibin = ibin;
  o->ibin_old = ibin_old;
  o->ibin_wake = ibin_wake;
  o->nbinmax = nbinmax;
  o->istepfrac = istepfrac;
  o->ibinnow = ibinnow;
}

int32_t orc_get_nbinmax(const orc_ctx *o) { return o->nbinmaxnew; }
/* the longest cell neighbour list of the last force call: at most 1000
 * (force.F90:58, maxcellcache) means every position came from the cache */
int64_t orc_get_max_force_trial(const orc_ctx *o) { return o->maxforcetrial; }

/* register the potential energy output of force (part:poten, real*4) */
void orc_set_poten(orc_ctx *o, float *poten) { o->poten = poten; }

/* node second moments (quads) of the tree, 6 per node, node ids 1..ncells */
int orc_get_quads(const orc_ctx *o, double *quads, int64_t cap) {
  if (!o->quads) return PHSPH_ERR_STATE;
  for (int64_t n = 1; n <= o->ncells && n - 1 < cap; n++)
    for (int k = 0; k < 6; k++) quads[6 * (n - 1) + k] = o->quads[6 * n + k];
  return PHSPH_OK;
}

/* set_active_particles (utils_indtimesteps.f90:114-185): a live particle is
 * active when mod(istepfrac, 2**(nbinmax-ibin)) == 0; iphase = +-1 (gas).
 * Returns ibinnow, or -1 if a bin exceeds nbinmax. */
int32_t orc_set_active_particles(int32_t np, const double *xyzh, int8_t *iphase, const int8_t *ibin,
                                 int32_t nbinmax, int32_t istepfrac, int32_t *nactive, int32_t *nalive) {
  int32_t na = 0, nl = 0;
  for (int32_t i = 0; i < np; i++) {
    if (!(xyzh[4 * (size_t)i + 3] > 0.)) continue;
    nl++;
    int b = ibin[i];
    if (b > nbinmax) return -1;
    int act = (istepfrac % (1 << (nbinmax - b))) == 0;
    if (act) na++;
    iphase[i] = act ? 1 : -1;
  }
  int32_t ibinnow = nbinmax;
  for (int32_t k = 0; ibinnow == nbinmax && k < nbinmax; k++)
    if (istepfrac % (1 << (nbinmax - k)) == 0) ibinnow = k;
  if (nactive) *nactive = na;
  if (nalive) *nalive = nl;
  return ibinnow;
}

/* kernel_softening (kernel_cubic.f90:103-121, kernel_quintic.f90:124-152):
 * softened potential and force of a unit mass at q = r/h */
static void kernel_softening(int kind, double q2, double q, double *pot, double *fs) {
  if (kind == PHSPH_KERNEL_CUBIC) {
    if (q < 1.) {
      double q4 = q2 * q2;
      *pot = q4 * q / 10. - 3. * q4 / 10. + 2. * q2 / 3. - 7. / 5.;
      *fs = q * (15. * q2 * q - 36. * q2 + 40.) / 30.;
    } else if (q < 2.) {
      double q4 = q2 * q2, q6 = q4 * q2;
      *pot = (q * (-q4 * q + 9. * q4 - 30. * q2 * q + 40. * q2 - 48.) + 2.) / (30. * q);
      *fs = (-5. * q6 + 36. * q4 * q - 90. * q4 + 80. * q2 * q - 2.) / (30. * q2);
    } else {
      *pot = -1. / q;
      *fs = 1. / q2;
    }
  } else {
    if (q < 1.) {
      double q4 = q2 * q2, q6 = q4 * q2;
      *pot = -q6 * q / 168. + q6 / 42. - q4 / 10. + 11. * q2 / 30. - 239. / 210.;
      *fs = q * (-35. * q4 * q + 120. * q4 - 336. * q2 + 616.) / 840.;
    } else if (q < 2.) {
      double q4 = q2 * q2, q6 = q4 * q2, q8 = q6 * q2;
      *pot = (q * (5. * q6 * q - 60. * q6 + 280. * q4 * q - 588. * q4 + 350. * q2 * q + 476. * q2 - 1892.) - 5.) /
             (1680. * q);
      *fs = (35. * q8 - 360. * q6 * q + 1400. * q6 - 2352. * q4 * q + 1050. * q4 + 952. * q2 * q + 5.) / (1680. * q2);
    } else if (q < 3.) {
      double q4 = q2 * q2, q6 = q4 * q2, q8 = q6 * q2;
      *pot = (q * (-q6 * q + 20. * q6 - 168. * q4 * q + 756. * q4 - 1890. * q2 * q + 2268. * q2 - 2916.) + 507.) /
             (1680. * q);
      *fs = (-7. * q8 + 120. * q6 * q - 840. * q6 + 3024. * q4 * q - 5670. * q4 + 4536. * q2 * q - 507.) /
            (1680. * q2);
    } else {
      *pot = -1. / q;
      *fs = 1. / q2;
    }
  }
}

void orc_kernel_softening(int32_t kind, double q, double *pot, double *fs) { kernel_softening(kind, q * q, q, pot, fs); }

/* Kernel value helpers exported for the kernel identity tests. */
void orc_kernel(int32_t kind, double q, double *w, double *g, double *dphidh) {
  get_kernel(kind, q * q, q, w, g);
  *dphidh = dphidh_kern(kind, q * q, q);
}

/* ------------------------------------------------------------------ */
/* leapfrog integrator loops (step_leapfrog.f90, substepping.F90)      */
/* Gas particles only: iphase NULL = all active, >0 active, <0         */
/* inactive; a particle is dead when h < tiny(h) (part.F90:931-941).   */
/* ------------------------------------------------------------------ */
#define AVDECAYCONST 0.1 /* shock_capturing.f90:32 */

static inline int orc_active(const int8_t *iphase, int32_t i) { return iphase == NULL || iphase[i] > 0; }

/* get_dt (utils_indtimesteps.f90:45-51): dtmax/2**ibin */
static inline double get_dt_bin(double dtmax, int ibin) {
  double p = 1.;
  for (int k = 0; k < ibin; k++) p = p * 2.;
  return dtmax / p;
}

/* init_step (step_leapfrog.f90:44-83), block timesteps */
int orc_init_step(orc_ctx *o, int32_t np, double time, int8_t *ibin, int32_t nbinmax, double *twas) {
  if (!o->p.ind_timesteps) {
    snprintf(o->err, sizeof(o->err), "init_step: needs individual timesteps");
    return PHSPH_ERR_INPUT;
  }
  if (time < TINY64)
    for (int32_t i = 0; i < np; i++) ibin[i] = (int8_t)nbinmax;
  for (int32_t i = 0; i < np; i++) twas[i] = time + 0.5 * get_dt_bin(o->p.dtmax, ibin[i]);
  return PHSPH_OK;
}

/* predictor loop of step (step_leapfrog.f90:150-237) */
int orc_step_predictor(orc_ctx *o, int32_t np, double t, double dtsph, const double *xyzh, double *vxyzu, int32_t ldv,
                       const double *fxyzu, int32_t ldf, double *Bevol, const double *dBevol, double *dustevol,
                       const double *ddustevol, const int8_t *iphase, const int8_t *ibin, int8_t *ibin_old,
                       const double *twas, double ufloor, int32_t *nvfloor) {
  const int ind = o->p.ind_timesteps, mhd = o->p.mhd, dust = o->p.use_dustfrac && o->p.ndustsmall > 0;
  const int nv = o->p.maxvxyzu;
  const double timei = t, hdtsph = 0.5 * dtsph;
  int32_t nfl = 0;
  for (int32_t i = 0; i < np; i++) {
    if (xyzh[4 * (size_t)i + 3] < TINY64) continue;
    double hdti;
    if (ind) {
      if (orc_active(iphase, i)) ibin_old[i] = ibin[i];
      hdti = twas[i] - timei;
    } else {
      hdti = hdtsph;
    }
    double *v = vxyzu + (size_t)ldv * i;
    const double *f = fxyzu + (size_t)ldf * i;
    for (int k = 0; k < nv; k++) v[k] = v[k] + hdti * f[k];
    if (ufloor > 0. && nv >= 4 && v[3] < ufloor) {
      v[3] = ufloor;
      nfl++;
    }
    if (mhd)
      for (int k = 0; k < 4; k++) Bevol[4 * (size_t)i + k] = Bevol[4 * (size_t)i + k] + hdti * dBevol[4 * (size_t)i + k];
    if (dust) dustevol[i] = fabs(dustevol[i] + hdti * ddustevol[i]);
  }
  if (nvfloor) *nvfloor = nfl;
  return PHSPH_OK;
}

/* substep_sph (substepping.F90:241-264) */
int orc_substep_sph(orc_ctx *o, int32_t np, double dt, double *xyzh, const double *vxyzu, int32_t ldv) {
  (void)o;
  for (int32_t i = 0; i < np; i++) {
    double *x = xyzh + 4 * (size_t)i;
    if (x[3] < TINY64) continue;
    const double *v = vxyzu + (size_t)ldv * i;
    x[0] = x[0] + dt * v[0];
    x[1] = x[1] + dt * v[1];
    x[2] = x[2] + dt * v[2];
  }
  return PHSPH_OK;
}

/* predict_sph loop of step (step_leapfrog.f90:307-400), gas, nalpha = 3
 * (Cullen-Dehnen switch, config.F90:161) */
int orc_step_predict_sph(orc_ctx *o, int32_t np, double timei, double dtsph, double *xyzh, const double *vxyzu,
                         int32_t ldv, const double *fxyzu, int32_t ldf, double *vpred, const double *Bevol,
                         const double *dBevol, double *Bpred, const double *dustevol, const double *ddustevol,
                         double *dustpred, const float *divcurlv, int32_t lddcv, const double *eos_vars, int32_t ldeos,
                         float *alphaind, const int8_t *iphase, const double *twas, double ufloor, int32_t *nvfloor) {
  (void)iphase;
  const int ind = o->p.ind_timesteps, mhd = o->p.mhd, dust = o->p.use_dustfrac && o->p.ndustsmall > 0;
  const int nv = o->p.maxvxyzu;
  const double pm = o->p.massoftype_gas, hdtsph = 0.5 * dtsph;
  int32_t nfl = 0;
  for (int32_t i = 0; i < np; i++) {
    double *x = xyzh + 4 * (size_t)i;
    if (x[3] < TINY64) continue;
    /* h prediction (:332) */
    x[3] = x[3] - dtsph * dhdrho(o, x[3], pm) * rhoh(o, x[3], pm) * (double)divcurlv[(size_t)lddcv * i];
    double hdti = ind ? timei - twas[i] : hdtsph;
    const double *v = vxyzu + (size_t)ldv * i;
    const double *f = fxyzu + (size_t)ldf * i;
    double *vp = vpred + (size_t)ldv * i;
    for (int k = 0; k < nv; k++) vp[k] = v[k] + hdti * f[k];
    if (ufloor > 0. && nv >= 4 && vp[3] < ufloor) {
      vp[3] = ufloor;
      nfl++;
    }
    if (mhd)
      for (int k = 0; k < 4; k++) Bpred[4 * (size_t)i + k] = Bevol[4 * (size_t)i + k] + hdti * dBevol[4 * (size_t)i + k];
    if (dust) dustpred[i] = dustevol[i] + hdti * ddustevol[i];
    /* viscosity switch (:377-396) */
    double hi = x[3];
    double spsoundi = eos_vars[(size_t)ldeos * i + 1];
    double tdecay1 = AVDECAYCONST * spsoundi / hi;
    double ddenom = 1. / (1. + dtsph * tdecay1);
    double alphaloci = (double)alphaind[3 * (size_t)i + 1];
    double a1 = (double)alphaind[3 * (size_t)i];
    if (a1 < alphaloci)
      alphaind[3 * (size_t)i] = (float)alphaloci;
    else
      alphaind[3 * (size_t)i] = (float)((a1 + dtsph * alphaloci * tdecay1) * ddenom);
  }
  if (nvfloor) *nvfloor = nfl;
  return PHSPH_OK;
}

/* one pass of the corrector loop of step (step_leapfrog.f90:482-625) */
int orc_step_corrector(orc_ctx *o, int32_t np, double timei, double dtsph, const double *xyzh, double *vxyzu,
                       int32_t ldv, const double *fxyzu, int32_t ldf, const double *vpred, double *Bevol,
                       const double *dBevol, double *dustevol, const double *ddustevol, const int8_t *iphase,
                       int8_t *ibin, int8_t *ibin_wake, int32_t nbinmax, double *twas, double ufloor, double *errmax,
                       double *v2sum, int32_t *npc, int32_t *nwake, int32_t *nvfloor) {
  const int ind = o->p.ind_timesteps, mhd = o->p.mhd, dust = o->p.use_dustfrac && o->p.ndustsmall > 0;
  const int nv = o->p.maxvxyzu;
  const double hdtsph = 0.5 * dtsph, dtmax = o->p.dtmax;
  /* ibin_dts(ithdt,:) and ibin_dts(ittwas,:) at time_now = timei
   * (step_leapfrog.f90:156-161, 73-80) */
  double thdt[MAXBINS + 1], ttwas[MAXBINS + 1];
  for (int b = 0; b <= MAXBINS; b++) {
    double dtb = get_dt_bin(dtmax, b);
    double dt1 = 1.0 / dtb;
    thdt[b] = 0.5 * dtb;
    ttwas[b] = ((double)(int64_t)(timei * dt1) + 0.5) * dtb;
  }
  double emax = 0., v2 = 0.;
  int32_t n = 0, nw = 0, nfl = 0;
  for (int32_t i = 0; i < np; i++) {
    if (xyzh[4 * (size_t)i + 3] < TINY64) continue;
    double *v = vxyzu + (size_t)ldv * i;
    const double *f = fxyzu + (size_t)ldf * i;
    double *B = mhd ? Bevol + 4 * (size_t)i : NULL;
    const double *dB = mhd ? dBevol + 4 * (size_t)i : NULL;
    if (ind) {
      if (orc_active(iphase, i)) {
        ibin_wake[i] = 0;
        double hdti = timei - twas[i];
        double dti = hdti + thdt[ibin[i]];
        for (int k = 0; k < nv; k++) v[k] = v[k] + dti * f[k];
        if (mhd)
          for (int k = 0; k < 4; k++) B[k] = B[k] + dti * dB[k];
        if (dust) dustevol[i] = dustevol[i] + dti * ddustevol[i];
        twas[i] = twas[i] + dti;
      }
      double hdti = timei - twas[i];
      for (int k = 0; k < nv; k++) v[k] = v[k] + hdti * f[k];
      if (ufloor > 0. && nv >= 4 && v[3] < ufloor) {
        v[3] = ufloor;
        nfl++;
      }
      if (mhd)
        for (int k = 0; k < 4; k++) B[k] = B[k] + hdti * dB[k];
      if (dust) dustevol[i] = dustevol[i] + hdti * ddustevol[i];
      if (ibin_wake[i] > ibin[i]) {
        int w = ibin_wake[i] < nbinmax ? ibin_wake[i] : nbinmax;
        nw++;
        twas[i] = ttwas[w];
        ibin[i] = (int8_t)w;
        ibin_wake[i] = 0;
      }
    } else {
      const double *vp = vpred + (size_t)ldv * i;
      double vxi = v[0] + hdtsph * f[0];
      double vyi = v[1] + hdtsph * f[1];
      double vzi = v[2] + hdtsph * f[2];
      double eni = 0.;
      if (nv >= 4) eni = v[3] + hdtsph * f[3];
      double erri = (vxi - vp[0]) * (vxi - vp[0]) + (vyi - vp[1]) * (vyi - vp[1]) + (vzi - vp[2]) * (vzi - vp[2]);
      if (erri > emax) emax = erri;
      v2 = v2 + (vxi * vxi + vyi * vyi + vzi * vzi);
      n++;
      v[0] = vxi;
      v[1] = vyi;
      v[2] = vzi;
      if (nv >= 4) v[3] = eni;
      if (mhd)
        for (int k = 0; k < 4; k++) B[k] = B[k] + hdtsph * dB[k];
      if (dust) dustevol[i] = dustevol[i] + hdtsph * ddustevol[i];
    }
  }
  if (errmax) *errmax = emax;
  if (v2sum) *v2sum = v2;
  if (npc) *npc = n;
  if (nwake) *nwake = nw;
  if (nvfloor) *nvfloor = nfl;
  return PHSPH_OK;
}

/* until_converged loop of step (step_leapfrog.f90:650-696) */
int orc_step_reset(orc_ctx *o, int32_t np, double dtsph, double *vxyzu, int32_t ldv,
                   const double *fxyzu, int32_t ldf, double *vpred, double *Bevol, const double *dBevol,
                   double *Bpred, double *dustevol, const double *ddustevol, double *dustpred,
                   const int8_t *iphase) {
  const int ind = o->p.ind_timesteps, mhd = o->p.mhd, dust = o->p.use_dustfrac && o->p.ndustsmall > 0;
  const int nv = o->p.maxvxyzu;
  const double hdtsph = 0.5 * dtsph;
  /* this loop has no dead-particle test in the reference */
  for (int32_t i = 0; i < np; i++) {
    double *v = vxyzu + (size_t)ldv * i;
    double *vp = vpred + (size_t)ldv * i;
    if (ind) {
      if (!orc_active(iphase, i)) continue;
      for (int k = 0; k < nv; k++) vp[k] = v[k];
      if (mhd)
        for (int k = 0; k < 4; k++) Bpred[4 * (size_t)i + k] = Bevol[4 * (size_t)i + k];
      if (dust) dustpred[i] = dustevol[i];
    } else {
      const double *f = fxyzu + (size_t)ldf * i;
      for (int k = 0; k < nv; k++) vp[k] = v[k];
      if (mhd)
        for (int k = 0; k < 4; k++) Bpred[4 * (size_t)i + k] = Bevol[4 * (size_t)i + k];
      if (dust) dustpred[i] = dustevol[i];
      for (int k = 0; k < nv; k++) v[k] = v[k] - hdtsph * f[k];
      if (mhd)
        for (int k = 0; k < 4; k++) Bevol[4 * (size_t)i + k] = Bevol[4 * (size_t)i + k] - hdtsph * dBevol[4 * (size_t)i + k];
      if (dust) dustevol[i] = dustevol[i] - hdtsph * ddustevol[i];
    }
  }
  return PHSPH_OK;
}
