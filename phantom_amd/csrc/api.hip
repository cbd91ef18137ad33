// api.hip -- the extern "C" entry points declared in include/phsph.h.
//
// Each entry point accepts host or device pointers.  Host arrays are staged
// through context-owned device buffers: inputs and in/out arrays are copied
// in, outputs are copied back after the kernels complete.  Device arrays
// are used in place, so a caller that keeps its particle arrays in HBM pays
// no transfer.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "ctx.h"

DevConst make_devconst(const phsph_params &p) {
  DevConst d;
  memset(&d, 0, sizeof(d));
  d.kernel = p.kernel;
  d.periodic = p.periodic;
  d.maxvxyzu = p.maxvxyzu;
  d.mhd = p.mhd;
  d.dust = (p.use_dustfrac && p.ndustsmall > 0) ? 1 : 0;
  d.gravity = p.gravity;
  d.ind_timesteps = p.ind_timesteps;
  d.ieos = p.ieos;
  d.ishock_heating = p.ishock_heating;
  d.ipdv_heating = p.ipdv_heating;
  d.iresistive_heating = p.iresistive_heating;
  d.ilimitdustflux = p.ilimitdustflux;
  d.idrag = p.idrag;
  // init_drag (dust.f90:78-79)
  d.cste_mu = sqrt(2. / (PH_PI * p.gamma));
  d.coeff_gei_1 = sqrt(8. / (PH_PI * p.gamma));
  d.grainsize = p.grainsize;
  d.graindens = p.graindens;
  d.seff = p.seff;
  if (p.kernel == PHSPH_KERNEL_CUBIC) {  // kernel_cubic.f90:23-30
    d.radkern = 2.; d.radkern2 = 4.; d.cnormk = 1. / PH_PI;
    d.wab0 = 1.; d.gradh0 = -3.; d.dphidh0 = 1.4;
  } else {  // kernel_quintic.f90:23-30
    d.radkern = 3.; d.radkern2 = 9.; d.cnormk = 1. / (120. * PH_PI);
    d.wab0 = 66.; d.gradh0 = -3. * 66.; d.dphidh0 = 239. / 210.;
  }
  d.xmin = p.xmin; d.xmax = p.xmax; d.ymin = p.ymin; d.ymax = p.ymax; d.zmin = p.zmin; d.zmax = p.zmax;
  d.dxbound = p.xmax - p.xmin; d.dybound = p.ymax - p.ymin; d.dzbound = p.zmax - p.zmin;
  d.hdlx = 0.5 * d.dxbound; d.hdly = 0.5 * d.dybound; d.hdlz = 0.5 * d.dzbound;
  d.hfact = p.hfact; d.tolh = p.tolh; d.pmass = p.massoftype_gas; d.gamma = p.gamma; d.polyk = p.polyk;
  d.alpha = p.alpha; d.alphamax = p.alphamax; d.alphau = p.alphau; d.alphaB = p.alphaB; d.beta = p.beta;
  d.C_cour = p.C_cour; d.C_force = p.C_force; d.dtmax = p.dtmax; d.psidecayfac = p.psidecayfac;
  d.overcleanfac = p.overcleanfac; d.K_code = p.K_code;
  d.tree_acc2 = p.tree_accuracy * p.tree_accuracy;
  return d;
}

void ctx_clear_timings(phsph_ctx *c) {
  for (auto &t : c->timings) {
    hipEventDestroy(t.t0);
    hipEventDestroy(t.t1);
  }
  c->timings.clear();
}

static const char *validate(const phsph_params *p) {
  if (p->kernel != PHSPH_KERNEL_CUBIC && p->kernel != PHSPH_KERNEL_QUINTIC) return "kernel must be cubic or quintic";
  if (p->ieos != 1 && p->ieos != 2) return "only ieos 1 (isothermal) and 2 (ideal gas) are supported";
  if (p->maxvxyzu != 3 && p->maxvxyzu != 4) return "maxvxyzu must be 3 or 4";
  if (!(p->massoftype_gas > 0.)) return "massoftype_gas must be > 0";
  if (!(p->hfact > 0.)) return "hfact must be > 0";
  if (!(p->tolh > 0.)) return "tolh must be > 0";
  if (!(p->xmax > p->xmin && p->ymax > p->ymin && p->zmax > p->zmin)) return "empty box";
  if (p->use_dustfrac && (p->ndustsmall < 0 || p->ndustsmall > 1)) return "ndustsmall must be 0 or 1";
  if (p->use_dustfrac && p->ndustsmall == 1 && p->idrag != 1 && p->idrag != 2 && p->idrag != 3)
    return "one-fluid dust supports idrag 1 (Epstein/Stokes), 2 (constant K) and 3 (constant stopping time)";
  if (p->gravity && (p->mhd || (p->use_dustfrac && p->ndustsmall > 0)))
    return "self-gravity is available for hydrodynamics (not with MHD or dust)";
  // read only with gravity, in 0..1 (neigh_kdtree.f90:345)
  if (p->gravity && !(p->tree_accuracy >= 0. && p->tree_accuracy <= 1.)) return "tree_accuracy must be in 0..1";
  if (p->precision != PHSPH_PREC_FP64 && p->precision != PHSPH_PREC_MIXED) return "unknown precision";
  return nullptr;
}

static bool is_device_ptr(const void *ptr) {
  if (!ptr) return false;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, ptr);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Staging of one caller array.  copy_in: the device copy starts with the
// caller's contents; copy_out: copied back by Stager::finish.
struct Stager {
  phsph_ctx *c;
  int slot = 0;
  struct Back {
    void *host;
    void *dev;
    size_t bytes;
  };
  std::vector<Back> backs;
  explicit Stager(phsph_ctx *ctx) : c(ctx) {}
  template <class T>
  int stage(T *ptr, size_t count, bool copy_in, bool copy_out, T **dev) {
    *dev = nullptr;
    if (!ptr) return PHSPH_OK;
    if (is_device_ptr(ptr)) {
      *dev = ptr;
      return PHSPH_OK;
    }
    if (slot >= 24) PH_RET(c, PHSPH_ERR_STATE, "internal: too many staged arrays");
    DBuf<char> &b = c->stage[slot++];
    const size_t bytes = count * sizeof(T);
    PH_HIPCHECK(c, b.ensure(bytes));
    if (copy_in) PH_HIPCHECK(c, hipMemcpyAsync(b.p, (const void *)ptr, bytes, hipMemcpyHostToDevice, c->stream));
    if (copy_out) backs.push_back({(void *)ptr, b.p, bytes});
    *dev = reinterpret_cast<T *>(b.p);
    return PHSPH_OK;
  }
  int finish() {
    for (auto &bk : backs)
      PH_HIPCHECK(c, hipMemcpyAsync(bk.host, bk.dev, bk.bytes, hipMemcpyDeviceToHost, c->stream));
    PH_HIPCHECK(c, hipStreamSynchronize(c->stream));
    return PHSPH_OK;
  }
};

#define STAGE(st, ptr, count, in, out, dev)                           \
  do {                                                                \
    int _rc = (st).stage((ptr), (size_t)(count), (in), (out), (dev)); \
    if (_rc) return _rc;                                              \
  } while (0)

struct SetWords {
  uint32_t w[32];
};
__global__ void k_set_words(uint32_t *dst, SetWords b, int nw) {
  const int i = threadIdx.x;
  if (i < nw) dst[i] = b.w[i];
}
hipError_t dev_set(phsph_ctx *c, void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return hipSuccess;
  if (bytes % 4 != 0 || bytes > sizeof(SetWords) || ((uintptr_t)dst & 3u) != 0)
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream);
  SetWords b;
  memcpy(b.w, src, bytes);
  hipLaunchKernelGGL(k_set_words, dim3(1), dim3(32), 0, c->stream, (uint32_t *)dst, b, (int)(bytes / 4));
  return hipGetLastError();
}

__global__ void k_zero_i32_api(int32_t *a, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// effective surface density of the gas mean free path in code units
// (init_drag, dust.f90:94-97)
static double seff_code_units(double umass, double udist) {
  const double cste_seff = PH_PI / sqrt(2.) * 5. / 64.;
  const double mass_mol_gas = (2. * 1.67262158e-24) / umass;   // physcon.f90:41
  const double cross_section_gas = 2.367e-15 / (udist * udist);  // physcon.f90:43
  return cste_seff * mass_mol_gas / cross_section_gas;
}

extern "C" {

void phsph_default_params(phsph_params *p, int32_t kernel) {
  memset(p, 0, sizeof(*p));
  p->kernel = kernel;
  p->periodic = 1;
  p->maxvxyzu = 4;
  p->ieos = 2;
  p->idrag = 2;
  p->ishock_heating = 1;
  p->ipdv_heating = 1;
  p->iresistive_heating = 1;
  p->xmin = -0.5; p->xmax = 0.5; p->ymin = -0.5; p->ymax = 0.5; p->zmin = -0.5; p->zmax = 0.5;
  p->hfact = (kernel == PHSPH_KERNEL_CUBIC) ? 1.2 : 1.0;
  p->tolh = 1.e-4;
  p->massoftype_gas = 1.;
  p->gamma = 5. / 3.;
  p->polyk = 1.;
  p->alpha = 0.; p->alphamax = 1.; p->alphau = 1.; p->alphaB = 1.; p->beta = 2.;
  p->C_cour = 0.3; p->C_force = 0.25; p->dtmax = 1.e29;
  p->psidecayfac = 1.; p->overcleanfac = 1.;
  p->K_code = 1.;
  p->tree_accuracy = 0.5;
  p->grainsize = 0.1;  // grainsizecgs, graindenscgs (dust.f90:38-39) with udist = umass = 1
  p->graindens = 3.;
  p->seff = seff_code_units(1., 1.);
}

int phsph_create(const phsph_params *p, int32_t device, phsph_ctx **out) {
  if (!out || !p) return PHSPH_ERR_INPUT;
  *out = nullptr;
  phsph_ctx *c = new phsph_ctx();
  c->device = device;
  const char *msg = validate(p);
  if (msg) {
    c->err = msg;
    *out = c;  // so that phsph_last_error can report; caller destroys
    return PHSPH_ERR_INPUT;
  }
  c->p = *p;
  c->dc = make_devconst(*p);
  // initial list capacities and the list budget (KiB) may be set from the
  // environment, e.g. PHSPH_CAPS="gcand=64,dlist=32,flist=32,p2p=16,gnear=8,budget_kib=1":
  // larger values avoid second search passes, smaller ones exercise them
  if (const char *env = getenv("PHSPH_CAPS")) {
    std::string spec(env);
    size_t pos = 0;
    while (pos < spec.size()) {
      size_t end = spec.find(',', pos);
      if (end == std::string::npos) end = spec.size();
      const std::string item = spec.substr(pos, end - pos);
      const size_t eq = item.find('=');
      if (eq != std::string::npos) {
        const std::string key = item.substr(0, eq);
        const int v = atoi(item.c_str() + eq + 1);
        if (v >= 1 && v <= (1 << 24)) {
          if (key == "gcand") c->caps.gcand = v;
          else if (key == "dlist") c->caps.dlist = v;
          else if (key == "flist") c->caps.flist = v;
          else if (key == "p2p") c->caps.p2p = v;
          else if (key == "gnear") c->caps.gnear = v;
          else if (key == "budget_kib") c->list_budget = (size_t)v << 10;
        }
      }
      pos = end + 1;
    }
  }
  hipError_t e = hipSetDevice(device);
  // a blocking stream: work the caller queued on the null stream (copies
  // into device arrays it passes) completes before the library's kernels
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamDefault);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) {
    c->err = std::string("HIP error ") + hipGetErrorString(e);
    *out = c;
    return PHSPH_ERR_HIP;
  }
  *out = c;
  // device address of the scratch array of phsph_step_corrector_noalloc
  // on this device (once per device, with the device current)
  return step_scratch_init(c);
}

int phsph_destroy(phsph_ctx *c) {
  if (!c) return PHSPH_OK;
  if (c->stream) {
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    ctx_clear_timings(c);
    for (hipGraphExec_t &x : c->step_graph)
      if (x) {
        hipGraphExecDestroy(x);
        x = nullptr;
      }
    hipStreamDestroy(c->stream);
  }
  delete c;  // device buffers are released by their destructors
  return PHSPH_OK;
}

int phsph_set_params(phsph_ctx *c, const phsph_params *p) {
  if (!c || !p) return PHSPH_ERR_INPUT;
  const char *msg = validate(p);
  if (msg) PH_RET(c, PHSPH_ERR_INPUT, msg);
  c->p = *p;
  c->dc = make_devconst(*p);
  c->dlist_reuse = false;
  return PHSPH_OK;
}

const char *phsph_last_error(const phsph_ctx *c) { return c ? c->err.c_str() : "null context"; }

int phsph_maketree(phsph_ctx *c, double *xyzh, int32_t ld, int32_t np, const int8_t *iphase,
                   int32_t *leaf_is_active, int64_t *ncells) {
  if (c) c->dlist_reuse = false;
  if (!c) return PHSPH_ERR_INPUT;
  if (!xyzh || np <= 0 || ld < 4) PH_RET(c, PHSPH_ERR_INPUT, "maketree: bad xyzh/np/ld");
  if ((int64_t)np >= (1 << 27)) PH_RET(c, PHSPH_ERR_CAPACITY, "maketree: at most 2^27-1 particles");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  ctx_clear_timings(c);
  Stager st(c);
  double *dx;
  int8_t *dph;
  STAGE(st, xyzh, (size_t)ld * np, true, true, &dx);
  STAGE(st, const_cast<int8_t *>(iphase), (size_t)np, true, false, &dph);
  int rc = tree_build(c, dx, ld, np, dph);
  int rc2 = st.finish();
  if (rc) return rc;
  if (rc2) return rc2;
  if (ncells) *ncells = c->ncells;
  if (leaf_is_active) {
    if (is_device_ptr(leaf_is_active)) {
      PH_HIPCHECK(c, hipMemcpyAsync(leaf_is_active, c->leaf_flag.p + 1, c->ncells * sizeof(int32_t),
                                    hipMemcpyDeviceToDevice, c->stream));
    } else {
      PH_HIPCHECK(c, hipMemcpyAsync(leaf_is_active, c->leaf_flag.p + 1, c->ncells * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, c->stream));
    }
    PH_HIPCHECK(c, hipStreamSynchronize(c->stream));
  }
  return PHSPH_OK;
}

int phsph_get_tree(phsph_ctx *c, phsph_kdnode *node, int32_t *inoderange, int32_t *inodeparts, int64_t cap_nodes,
                   int32_t *maxlevel) {
  if (!c) return PHSPH_ERR_INPUT;
  if (!c->tree_valid) PH_RET(c, PHSPH_ERR_STATE, "get_tree: no tree");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  const int64_t n = std::min<int64_t>(cap_nodes, c->ncells);
  if (node && n > 0) {
    PH_HIPCHECK(c, hipMemcpy(node, c->nd.p + 1, n * sizeof(phsph_kdnode), hipMemcpyDeviceToHost));
    if (c->p.gravity && c->ghmax.p) {
      // with self-gravity the node hmax follows the density relists
      // (set_hmaxcell), as the reference's node array does
      std::vector<unsigned long long> hm(n);
      PH_HIPCHECK(c, hipMemcpy(hm.data(), c->ghmax.p + 1, n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
      for (int64_t k = 0; k < n; k++) memcpy(&node[k].hmax, &hm[k], sizeof(double));
    }
  }
  if (inoderange && n > 0) {
    std::vector<int32_t> lo(n), hi(n), fl(n);
    PH_HIPCHECK(c, hipMemcpy(lo.data(), c->nd_lo.p + 1, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    PH_HIPCHECK(c, hipMemcpy(hi.data(), c->nd_hi.p + 1, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    PH_HIPCHECK(c, hipMemcpy(fl.data(), c->leaf_flag.p + 1, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<phsph_kdnode> nd(n);
    PH_HIPCHECK(c, hipMemcpy(nd.data(), c->nd.p + 1, n * sizeof(phsph_kdnode), hipMemcpyDeviceToHost));
    // slots below the deepest leaves are never constructed: report them
    // empty, as the reference leaves them untouched (kdtree.F90:161)
    std::vector<char> used(n + 1, 0);
    used[0] = 1;  // root (index 0 = node 1)
    for (int64_t k = 0; k < n; k++) {
      if (!used[k]) continue;
      if (fl[k] == 0) {
        int64_t l = (int64_t)nd[k].leftchild - 1, r = (int64_t)nd[k].rightchild - 1;
        if (l >= 0 && l < n) used[l] = 1;
        if (r >= 0 && r < n) used[r] = 1;
      }
    }
    for (int64_t k = 0; k < n; k++) {
      inoderange[2 * k] = used[k] ? lo[k] + 1 : 0;
      inoderange[2 * k + 1] = used[k] ? hi[k] + 1 : 0;
    }
  }
  if (inodeparts)
    PH_HIPCHECK(c, hipMemcpy(inodeparts, c->iparts.p, (size_t)c->nlive * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (maxlevel) *maxlevel = c->maxlevel;
  return PHSPH_OK;
}

int phsph_set_tuning(phsph_ctx *c, int32_t cl, int32_t lw, int32_t r2, int32_t q) {
  if (!c) return PHSPH_ERR_INPUT;
  if (cl == 0) {
    c->tune_mode = TUNE_PENDING;
    return PHSPH_OK;
  }
  if ((cl != 8 && cl != 16) || (lw != 2 && lw != 4) || (r2 != 0 && r2 != 1) || (q != 0 && q != 1))
    PH_RET(c, PHSPH_ERR_INPUT, "set_tuning: cl must be 8 or 16, lw 2 or 4, r2 and q 0 or 1");
  c->tune = ListTune{cl, lw, r2, q};
  c->tune_mode = TUNE_FIXED;
  return PHSPH_OK;
}

int phsph_get_tuning(phsph_ctx *c, int32_t *cl, int32_t *lw, int32_t *r2, int32_t *q, int32_t *tuned, float *ms,
                     int32_t *cfg, int32_t *ntried, double *rfac) {
  if (!c) return PHSPH_ERR_INPUT;
  if (rfac) *rfac = RFAC_LEVELS[c->rfac_level];
  if (cl) *cl = c->tune.cl;
  if (lw) *lw = c->tune.lw;
  if (r2) *r2 = c->tune.r2;
  if (q) *q = c->tune.q;
  if (tuned) *tuned = c->tune_mode == TUNE_DONE ? 1 : 0;
  if (ntried) *ntried = c->tune_ntried;
  for (int k = 0; k < c->tune_ntried; k++) {
    if (ms) ms[k] = c->tune_ms[k];
    if (cfg) {
      cfg[4 * k] = c->tune_tried[k].cl;
      cfg[4 * k + 1] = c->tune_tried[k].lw;
      cfg[4 * k + 2] = c->tune_tried[k].r2;
      cfg[4 * k + 3] = c->tune_tried[k].q;
    }
  }
  return PHSPH_OK;
}

int phsph_get_tree_grav(phsph_ctx *c, phsph_kdnode_grav *node, int64_t cap_nodes) {
  if (!c) return PHSPH_ERR_INPUT;
  if (!c->tree_valid) PH_RET(c, PHSPH_ERR_STATE, "get_tree_grav: no tree");
  if (!c->p.gravity || !c->quads.p) PH_RET(c, PHSPH_ERR_STATE, "get_tree_grav: self-gravity is off");
  if (!node) PH_RET(c, PHSPH_ERR_INPUT, "get_tree_grav: missing array");
  const int64_t n = std::min<int64_t>(cap_nodes, c->ncells);
  if (n <= 0) return PHSPH_OK;
  std::vector<phsph_kdnode> nd(n);
  std::vector<double> q((size_t)n * 6);
  int rc = phsph_get_tree(c, nd.data(), nullptr, nullptr, n, nullptr);
  if (rc) return rc;
  rc = phsph_get_quads(c, q.data(), n);
  if (rc) return rc;
  for (int64_t k = 0; k < n; k++) {
    phsph_kdnode_grav &g = node[k];
    for (int d = 0; d < 3; d++) g.xcen[d] = nd[k].xcen[d];
    g.size = nd[k].size;
    g.hmax = nd[k].hmax;
    g.mass = nd[k].mass;
    g.leftchild = nd[k].leftchild;
    g.rightchild = nd[k].rightchild;
    g.parent = nd[k].parent;
    g.idum = nd[k].idum;
    g.tobecached = 1;  // construct_node (kdtree.F90:788-789)
    g.cached = 0;
    for (int m = 0; m < 6; m++) g.quads[m] = q[(size_t)k * 6 + m];
  }
  return PHSPH_OK;
}

int phsph_get_quads(phsph_ctx *c, double *quads, int64_t cap_nodes) {
  if (!c) return PHSPH_ERR_INPUT;
  if (!c->tree_valid) PH_RET(c, PHSPH_ERR_STATE, "get_quads: no tree");
  if (!c->p.gravity || !c->quads.p) PH_RET(c, PHSPH_ERR_STATE, "get_quads: self-gravity is off");
  if (!quads) PH_RET(c, PHSPH_ERR_INPUT, "get_quads: missing array");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  const int64_t n = std::min<int64_t>(cap_nodes, c->ncells);
  if (n > 0)
    PH_HIPCHECK(c, hipMemcpy(quads, c->quads.p + 6, (size_t)n * 6 * sizeof(double), hipMemcpyDeviceToHost));
  return PHSPH_OK;
}

int phsph_densityiterate(phsph_ctx *c, int32_t icall, int32_t np, double *xyzh, int32_t ldx, const double *vxyzu,
                         int32_t ldv, const double *fxyzu, int32_t ldf, const double *fext, const double *Bevol,
                         const int8_t *iphase, float *divcurlv, int32_t lddcv, float *divcurlB, float *alphaind,
                         float *gradh, int32_t ldgradh, float *dvdx) {
  if (!c) return PHSPH_ERR_INPUT;
  if (!c->tree_valid || np != c->np) PH_RET(c, PHSPH_ERR_STATE, "densityiterate: call maketree first");
  if (icall != 0 && icall != 1) PH_RET(c, PHSPH_ERR_INPUT, "densityiterate: icall must be 0 or 1");
  if (!xyzh || !vxyzu || !divcurlv || !gradh || ldx < 4 || ldv < 3 || lddcv < 1 || ldgradh < 1)
    PH_RET(c, PHSPH_ERR_INPUT, "densityiterate: missing array");
  if (c->p.gravity && ldgradh < 2) PH_RET(c, PHSPH_ERR_INPUT, "densityiterate: self-gravity needs gradh(2,:)");
  if (c->p.mhd && !Bevol) PH_RET(c, PHSPH_ERR_INPUT, "densityiterate: MHD needs Bevol");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  ctx_clear_timings(c);
  Stager st(c);
  DensIOArgs a;
  memset(&a, 0, sizeof(a));
  STAGE(st, xyzh, (size_t)ldx * np, true, true, &a.xyzh);
  a.ldx = ldx;
  double *t;
  STAGE(st, const_cast<double *>(vxyzu), (size_t)ldv * np, true, false, &t);
  a.vxyzu = t; a.ldv = ldv;
  STAGE(st, const_cast<double *>(fxyzu), (size_t)ldf * np, true, false, &t);
  a.fxyzu = t; a.ldf = ldf;
  STAGE(st, const_cast<double *>(fext), (size_t)3 * np, true, false, &t);
  a.fext = t;
  STAGE(st, const_cast<double *>(c->p.mhd ? Bevol : nullptr), (size_t)4 * np, true, false, &t);
  a.Bevol = t;
  int8_t *ph;
  STAGE(st, const_cast<int8_t *>(iphase), (size_t)np, true, false, &ph);
  a.iphase = ph;
  STAGE(st, divcurlv, (size_t)lddcv * np, true, true, &a.divcurlv);
  a.lddcv = lddcv;
  STAGE(st, c->p.mhd ? divcurlB : nullptr, (size_t)4 * np, true, true, &a.divcurlB);
  STAGE(st, alphaind, (size_t)3 * np, true, true, &a.alphaind);
  STAGE(st, gradh, (size_t)ldgradh * np, true, true, &a.gradh);
  a.ldgradh = ldgradh;
  STAGE(st, dvdx, (size_t)9 * np, true, true, &a.dvdx);
  int rc = dens_run(c, icall, a);
  int rc2 = st.finish();
  return rc ? rc : rc2;
}

int phsph_cons2prim(phsph_ctx *c, int32_t np, const double *xyzh, int32_t ldx, const double *vxyzu, int32_t ldv,
                    const float *dvdx, float *alphaind, const double *dustevol, double *dustfrac, double *eos_vars,
                    int32_t ldeos, const int8_t *iphase) {
  (void)iphase;
  if (!c) return PHSPH_ERR_INPUT;
  if (!xyzh || !vxyzu || !eos_vars || ldeos < 2 || ldx < 4) PH_RET(c, PHSPH_ERR_INPUT, "cons2prim: missing array");
  const bool dust = c->p.use_dustfrac && c->p.ndustsmall > 0;
  if (dust && (!dustevol || !dustfrac)) PH_RET(c, PHSPH_ERR_INPUT, "cons2prim: dust needs dustevol and dustfrac");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  ctx_clear_timings(c);
  Stager st(c);
  double *dx, *dv, *de, *dd, *deos;
  float *ddv, *da;
  STAGE(st, const_cast<double *>(xyzh), (size_t)ldx * np, true, false, &dx);
  STAGE(st, const_cast<double *>(vxyzu), (size_t)ldv * np, true, false, &dv);
  STAGE(st, const_cast<float *>(dvdx), (size_t)9 * np, true, false, &ddv);
  STAGE(st, alphaind, (size_t)3 * np, true, true, &da);
  STAGE(st, const_cast<double *>(dust ? dustevol : nullptr), (size_t)np, true, false, &de);
  STAGE(st, dust ? dustfrac : nullptr, (size_t)np, true, true, &dd);
  STAGE(st, eos_vars, (size_t)ldeos * np, true, true, &deos);
  int rc = cons2prim_run(c, np, dx, ldx, dv, ldv, ddv, da, de, dd, deos, ldeos);
  int rc2 = st.finish();
  return rc ? rc : rc2;
}

int phsph_force(phsph_ctx *c, int32_t icall, int32_t np, const double *xyzh, int32_t ldx, const double *vxyzu,
                int32_t ldv, double *fxyzu, int32_t ldf, float *divcurlv, int32_t lddcv, const float *divcurlB,
                const double *Bevol, double *dBevol, float *divBsymm, const double *dustfrac, double *ddustevol,
                double *deltav, double *tstop, const float *gradh, int32_t ldgradh, const float *alphaind,
                const double *eos_vars, int32_t ldeos, const int8_t *iphase, double *dtcourant, double *dtforce) {
  if (!c) return PHSPH_ERR_INPUT;
  if (!c->tree_valid || np != c->np) PH_RET(c, PHSPH_ERR_STATE, "force: call maketree first");
  if (!xyzh || !vxyzu || !fxyzu || !divcurlv || !gradh || !alphaind || !eos_vars || ldx < 4 || ldv < 3 ||
      ldf < 3 || lddcv < 1 || ldgradh < 1 || ldeos < 2)
    PH_RET(c, PHSPH_ERR_INPUT, "force: missing array");
  if (c->p.maxvxyzu >= 4 && (ldv < 4 || ldf < 4)) PH_RET(c, PHSPH_ERR_INPUT, "force: vxyzu/fxyzu need 4 components");
  if (c->p.gravity && ldgradh < 2) PH_RET(c, PHSPH_ERR_INPUT, "force: self-gravity needs gradh(2,:)");
  const bool mhd = c->p.mhd, dust = c->p.use_dustfrac && c->p.ndustsmall > 0;
  if (mhd && (!Bevol || !dBevol || !divcurlB)) PH_RET(c, PHSPH_ERR_INPUT, "force: MHD needs Bevol, dBevol, divcurlB");
  if (dust && (!dustfrac || !ddustevol || !deltav)) PH_RET(c, PHSPH_ERR_INPUT, "force: dust needs dust arrays");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  ctx_clear_timings(c);
  Stager st(c);
  ForceIOArgs a;
  memset(&a, 0, sizeof(a));
  a.np = np;
  double *t;
  float *tf;
  int8_t *ph;
  STAGE(st, const_cast<double *>(xyzh), (size_t)ldx * np, true, false, &t);
  a.xyzh = t; a.ldx = ldx;
  STAGE(st, const_cast<double *>(vxyzu), (size_t)ldv * np, true, false, &t);
  a.vxyzu = t; a.ldv = ldv;
  STAGE(st, fxyzu, (size_t)ldf * np, true, true, &a.fxyzu);
  a.ldf = ldf;
  STAGE(st, divcurlv, (size_t)lddcv * np, true, true, &a.divcurlv);
  a.lddcv = lddcv;
  STAGE(st, const_cast<float *>(mhd ? divcurlB : nullptr), (size_t)4 * np, true, false, &tf);
  a.divcurlB = tf;
  STAGE(st, const_cast<double *>(mhd ? Bevol : nullptr), (size_t)4 * np, true, false, &t);
  a.Bevol = t;
  STAGE(st, mhd ? dBevol : nullptr, (size_t)4 * np, true, true, &a.dBevol);
  STAGE(st, mhd ? divBsymm : nullptr, (size_t)np, true, true, &a.divBsymm);
  STAGE(st, const_cast<double *>(dust ? dustfrac : nullptr), (size_t)np, true, false, &t);
  a.dustfrac = t;
  STAGE(st, dust ? ddustevol : nullptr, (size_t)np, true, true, &a.ddustevol);
  STAGE(st, dust ? deltav : nullptr, (size_t)3 * np, true, true, &a.deltav);
  STAGE(st, dust ? tstop : nullptr, (size_t)np, true, true, &a.tstop);
  STAGE(st, const_cast<float *>(gradh), (size_t)ldgradh * np, true, false, &tf);
  a.gradh = tf; a.ldgradh = ldgradh;
  STAGE(st, const_cast<float *>(alphaind), (size_t)3 * np, true, false, &tf);
  a.alphaind = tf;
  STAGE(st, const_cast<double *>(eos_vars), (size_t)ldeos * np, true, false, &t);
  a.eos_vars = t; a.ldeos = ldeos;
  STAGE(st, const_cast<int8_t *>(iphase), (size_t)np, true, false, &ph);
  a.iphase = ph;
  if (c->p.ind_timesteps) {
    if (!c->bins_set) PH_RET(c, PHSPH_ERR_STATE, "force: individual timesteps need phsph_set_bins");
    STAGE(st, c->bins_ibin, (size_t)np, true, true, &a.ibin);
    STAGE(st, const_cast<int8_t *>(c->bins_ibin_old), (size_t)np, true, false, &ph);
    a.ibin_old = ph;
    STAGE(st, c->bins_ibin_wake, (size_t)np, true, true, &a.ibin_wake);
    a.nbinmax = c->bins_nbinmax;
    a.istepfrac = c->bins_istepfrac;
    a.ibinnow = c->bins_ibinnow;
  }
  c->poten_out = nullptr;
  if (c->p.gravity && c->poten_user) STAGE(st, c->poten_user, (size_t)np, true, true, &c->poten_out);
  int rc = force_run(c, icall, a, dtcourant, dtforce);
  int rc2 = st.finish();
  return rc ? rc : rc2;
}

int phsph_set_poten(phsph_ctx *c, float *poten) {
  if (!c) return PHSPH_ERR_INPUT;
  c->poten_user = poten;
  return PHSPH_OK;
}

int phsph_set_bins(phsph_ctx *c, int8_t *ibin, const int8_t *ibin_old, int8_t *ibin_wake, int32_t nbinmax,
                   int32_t istepfrac, int32_t ibinnow) {
  if (!c) return PHSPH_ERR_INPUT;
  if (!ibin || !ibin_old || !ibin_wake) PH_RET(c, PHSPH_ERR_INPUT, "set_bins: missing array");
  if (nbinmax < 0 || nbinmax > 30) PH_RET(c, PHSPH_ERR_INPUT, "set_bins: nbinmax outside 0..30");
  c->bins_ibin = ibin;
  c->bins_ibin_old = ibin_old;
  c->bins_ibin_wake = ibin_wake;
  c->bins_nbinmax = nbinmax;
  c->bins_istepfrac = istepfrac;
  c->bins_ibinnow = ibinnow;
  c->bins_set = true;
  return PHSPH_OK;
}

int phsph_get_nbinmax(phsph_ctx *c, int32_t *nbinmax) {
  if (!c) return PHSPH_ERR_INPUT;
  if (c->nbinmax_new < 0) PH_RET(c, PHSPH_ERR_STATE, "get_nbinmax: no force call with block timesteps");
  if (nbinmax) *nbinmax = c->nbinmax_new;
  return PHSPH_OK;
}

int phsph_set_active_particles(phsph_ctx *c, int32_t np, const double *xyzh, int32_t ldx, int8_t *iphase,
                               const int8_t *ibin, int32_t nbinmax, int32_t istepfrac, int32_t *nactive,
                               int32_t *nalive, int32_t *ibinnow) {
  if (!c) return PHSPH_ERR_INPUT;
  if (np < 0 || !xyzh || !iphase || !ibin || ldx < 4) PH_RET(c, PHSPH_ERR_INPUT, "set_active_particles: missing array");
  if (nbinmax < 0 || nbinmax > 30) PH_RET(c, PHSPH_ERR_INPUT, "set_active_particles: nbinmax outside 0..30");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  ctx_clear_timings(c);
  Stager st(c);
  double *dx;
  int8_t *dph, *db;
  STAGE(st, const_cast<double *>(xyzh), (size_t)ldx * np, true, false, &dx);
  STAGE(st, iphase, (size_t)np, true, true, &dph);
  STAGE(st, const_cast<int8_t *>(ibin), (size_t)np, true, false, &db);
  int32_t na = 0, nl = 0;
  int rc = set_active_run(c, np, dx, ldx, dph, db, nbinmax, istepfrac, &na, &nl);
  int rc2 = st.finish();
  if (rc) return rc;
  if (rc2) return rc2;
  // the largest active bin (utils_indtimesteps.f90:170-176)
  int32_t bnow = nbinmax;
  for (int32_t k = 0; bnow == nbinmax && k < nbinmax; k++)
    if (istepfrac % (1 << (nbinmax - k)) == 0) bnow = k;
  if (nactive) *nactive = na;
  if (nalive) *nalive = nl;
  if (ibinnow) *ibinnow = bnow;
  return PHSPH_OK;
}

// ---- leapfrog integrator loops (step_leapfrog.f90, substepping.F90) ----

int phsph_init_step(phsph_ctx *c, int32_t np, double time, int8_t *ibin, int32_t nbinmax, double *twas) {
  if (!c) return PHSPH_ERR_INPUT;
  if (!c->p.ind_timesteps) PH_RET(c, PHSPH_ERR_STATE, "init_step: needs individual timesteps");
  if (np < 0 || (np > 0 && (!ibin || !twas))) PH_RET(c, PHSPH_ERR_INPUT, "init_step: missing array");
  if (nbinmax < 0 || nbinmax > 30) PH_RET(c, PHSPH_ERR_INPUT, "init_step: nbinmax outside 0..30");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  ctx_clear_timings(c);
  Stager st(c);
  int8_t *db;
  double *dt;
  STAGE(st, ibin, (size_t)np, true, true, &db);
  STAGE(st, twas, (size_t)np, false, true, &dt);
  int rc = step_init_run(c, np, time, db, nbinmax, dt);
  int rc2 = st.finish();
  return rc ? rc : rc2;
}

// common argument checks of the step loops: the arrays each physics option
// needs and the leading dimensions
static int step_check(phsph_ctx *c, const char *who, int32_t np, int ldx, int ldv, int ldf) {
  if (np < 0) PH_RET(c, PHSPH_ERR_INPUT, std::string(who) + ": np < 0");
  if (ldx < 4 || ldv < c->p.maxvxyzu || (ldf >= 0 && ldf < c->p.maxvxyzu))
    PH_RET(c, PHSPH_ERR_INPUT, std::string(who) + ": leading dimension below the component count");
  return PHSPH_OK;
}

int phsph_step_predictor(phsph_ctx *c, int32_t np, double t, double dtsph, const double *xyzh, int32_t ldx,
                         double *vxyzu, int32_t ldv, const double *fxyzu, int32_t ldf, double *Bevol,
                         const double *dBevol, double *dustevol, const double *ddustevol, const int8_t *iphase,
                         const int8_t *ibin, int8_t *ibin_old, const double *twas, double ufloor, int32_t *nvfloor) {
  if (!c) return PHSPH_ERR_INPUT;
  int rc = step_check(c, "step_predictor", np, ldx, ldv, ldf);
  if (rc) return rc;
  const bool mhd = c->p.mhd, dust = c->p.use_dustfrac && c->p.ndustsmall > 0, ind = c->p.ind_timesteps;
  if (np > 0 && (!xyzh || !vxyzu || !fxyzu || (mhd && (!Bevol || !dBevol)) || (dust && (!dustevol || !ddustevol)) ||
                 (ind && (!ibin || !ibin_old || !twas))))
    PH_RET(c, PHSPH_ERR_INPUT, "step_predictor: missing array");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  ctx_clear_timings(c);
  Stager st(c);
  StepArgs a;
  memset(&a, 0, sizeof(a));
  double *t8;
  STAGE(st, const_cast<double *>(xyzh), (size_t)ldx * np, true, false, &t8);
  a.xyzh = t8; a.ldx = ldx;
  STAGE(st, vxyzu, (size_t)ldv * np, true, true, &a.vxyzu);
  a.ldv = ldv;
  STAGE(st, const_cast<double *>(fxyzu), (size_t)ldf * np, true, false, &t8);
  a.fxyzu = t8; a.ldf = ldf;
  STAGE(st, mhd ? Bevol : nullptr, (size_t)4 * np, true, true, &a.Bevol);
  STAGE(st, const_cast<double *>(mhd ? dBevol : nullptr), (size_t)4 * np, true, false, &t8);
  a.dBevol = t8;
  STAGE(st, dust ? dustevol : nullptr, (size_t)np, true, true, &a.dustevol);
  STAGE(st, const_cast<double *>(dust ? ddustevol : nullptr), (size_t)np, true, false, &t8);
  a.ddustevol = t8;
  int8_t *t1;
  STAGE(st, const_cast<int8_t *>(iphase), (size_t)np, true, false, &t1);
  a.iphase = t1;
  if (ind) {
    STAGE(st, const_cast<int8_t *>(ibin), (size_t)np, true, false, &t1);
    a.ibin_in = t1;
    STAGE(st, ibin_old, (size_t)np, true, true, &a.ibin_old);
    STAGE(st, const_cast<double *>(twas), (size_t)np, true, false, &t8);
    a.twas = t8;
  }
  rc = step_predictor_run(c, np, t, dtsph, a, ufloor, nvfloor);
  int rc2 = st.finish();
  return rc ? rc : rc2;
}

int phsph_substep_sph(phsph_ctx *c, int32_t np, double dt, double *xyzh, int32_t ldx, const double *vxyzu,
                      int32_t ldv) {
  if (!c) return PHSPH_ERR_INPUT;
  if (np < 0 || ldx < 4 || ldv < 3) PH_RET(c, PHSPH_ERR_INPUT, "substep_sph: bad size");
  if (np > 0 && (!xyzh || !vxyzu)) PH_RET(c, PHSPH_ERR_INPUT, "substep_sph: missing array");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  ctx_clear_timings(c);
  Stager st(c);
  double *dx, *dv;
  STAGE(st, xyzh, (size_t)ldx * np, true, true, &dx);
  STAGE(st, const_cast<double *>(vxyzu), (size_t)ldv * np, true, false, &dv);
  // the particles move: the tree of the last maketree no longer describes them
  c->tree_valid = false;
  int rc = step_drift_run(c, np, dt, dx, ldx, dv, ldv);
  int rc2 = st.finish();
  return rc ? rc : rc2;
}

int phsph_step_predict_sph(phsph_ctx *c, int32_t np, double timei, double dtsph, double *xyzh, int32_t ldx,
                           const double *vxyzu, int32_t ldv, const double *fxyzu, int32_t ldf, double *vpred,
                           const double *Bevol, const double *dBevol, double *Bpred, const double *dustevol,
                           const double *ddustevol, double *dustpred, const float *divcurlv, int32_t lddcv,
                           const double *eos_vars, int32_t ldeos, float *alphaind, const int8_t *iphase,
                           const double *twas, double ufloor, int32_t *nvfloor) {
  if (!c) return PHSPH_ERR_INPUT;
  int rc = step_check(c, "step_predict_sph", np, ldx, ldv, ldf);
  if (rc) return rc;
  if (lddcv < 1 || ldeos < 2) PH_RET(c, PHSPH_ERR_INPUT, "step_predict_sph: bad leading dimension");
  const bool mhd = c->p.mhd, dust = c->p.use_dustfrac && c->p.ndustsmall > 0, ind = c->p.ind_timesteps;
  if (np > 0 && (!xyzh || !vxyzu || !fxyzu || !vpred || !divcurlv || !eos_vars || !alphaind ||
                 (mhd && (!Bevol || !dBevol || !Bpred)) || (dust && (!dustevol || !ddustevol || !dustpred)) ||
                 (ind && !twas)))
    PH_RET(c, PHSPH_ERR_INPUT, "step_predict_sph: missing array");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  ctx_clear_timings(c);
  Stager st(c);
  StepArgs a;
  memset(&a, 0, sizeof(a));
  double *t8;
  float *tf;
  int8_t *t1;
  STAGE(st, xyzh, (size_t)ldx * np, true, true, &a.xyzh);
  a.ldx = ldx;
  STAGE(st, const_cast<double *>(vxyzu), (size_t)ldv * np, true, false, &a.vxyzu);
  a.ldv = ldv;
  STAGE(st, const_cast<double *>(fxyzu), (size_t)ldf * np, true, false, &t8);
  a.fxyzu = t8; a.ldf = ldf;
  STAGE(st, vpred, (size_t)ldv * np, true, true, &a.vpred);
  STAGE(st, const_cast<double *>(mhd ? Bevol : nullptr), (size_t)4 * np, true, false, &a.Bevol);
  STAGE(st, const_cast<double *>(mhd ? dBevol : nullptr), (size_t)4 * np, true, false, &t8);
  a.dBevol = t8;
  STAGE(st, mhd ? Bpred : nullptr, (size_t)4 * np, true, true, &a.Bpred);
  STAGE(st, const_cast<double *>(dust ? dustevol : nullptr), (size_t)np, true, false, &a.dustevol);
  STAGE(st, const_cast<double *>(dust ? ddustevol : nullptr), (size_t)np, true, false, &t8);
  a.ddustevol = t8;
  STAGE(st, dust ? dustpred : nullptr, (size_t)np, true, true, &a.dustpred);
  STAGE(st, const_cast<float *>(divcurlv), (size_t)lddcv * np, true, false, &tf);
  a.divcurlv = tf; a.lddcv = lddcv;
  STAGE(st, const_cast<double *>(eos_vars), (size_t)ldeos * np, true, false, &t8);
  a.eos_vars = t8; a.ldeos = ldeos;
  STAGE(st, alphaind, (size_t)3 * np, true, true, &a.alphaind);
  STAGE(st, const_cast<int8_t *>(iphase), (size_t)np, true, false, &t1);
  a.iphase = t1;
  if (ind) STAGE(st, const_cast<double *>(twas), (size_t)np, true, false, &a.twas);
  rc = step_predict_sph_run(c, np, timei, dtsph, a, ufloor, nvfloor);
  int rc2 = st.finish();
  return rc ? rc : rc2;
}

int phsph_step_corrector(phsph_ctx *c, int32_t np, double timei, double dtsph, const double *xyzh, int32_t ldx,
                         double *vxyzu, int32_t ldv, const double *fxyzu, int32_t ldf, const double *vpred,
                         double *Bevol, const double *dBevol, double *dustevol, const double *ddustevol,
                         const int8_t *iphase, int8_t *ibin, int8_t *ibin_wake, int32_t nbinmax, double *twas,
                         double ufloor, double *errmax, double *v2sum, int32_t *npc, int32_t *nwake,
                         int32_t *nvfloor) {
  if (!c) return PHSPH_ERR_INPUT;
  int rc = step_check(c, "step_corrector", np, ldx, ldv, ldf);
  if (rc) return rc;
  const bool mhd = c->p.mhd, dust = c->p.use_dustfrac && c->p.ndustsmall > 0, ind = c->p.ind_timesteps;
  if (np > 0 && (!xyzh || !vxyzu || !fxyzu || (!ind && !vpred) || (mhd && (!Bevol || !dBevol)) ||
                 (dust && (!dustevol || !ddustevol)) || (ind && (!ibin || !ibin_wake || !twas))))
    PH_RET(c, PHSPH_ERR_INPUT, "step_corrector: missing array");
  if (ind && (nbinmax < 0 || nbinmax > 30)) PH_RET(c, PHSPH_ERR_INPUT, "step_corrector: nbinmax outside 0..30");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  ctx_clear_timings(c);
  Stager st(c);
  StepArgs a;
  memset(&a, 0, sizeof(a));
  double *t8;
  int8_t *t1;
  STAGE(st, const_cast<double *>(xyzh), (size_t)ldx * np, true, false, &a.xyzh);
  a.ldx = ldx;
  STAGE(st, vxyzu, (size_t)ldv * np, true, true, &a.vxyzu);
  a.ldv = ldv;
  STAGE(st, const_cast<double *>(fxyzu), (size_t)ldf * np, true, false, &t8);
  a.fxyzu = t8; a.ldf = ldf;
  STAGE(st, const_cast<double *>(ind ? nullptr : vpred), (size_t)ldv * np, true, false, &a.vpred);
  STAGE(st, mhd ? Bevol : nullptr, (size_t)4 * np, true, true, &a.Bevol);
  STAGE(st, const_cast<double *>(mhd ? dBevol : nullptr), (size_t)4 * np, true, false, &t8);
  a.dBevol = t8;
  STAGE(st, dust ? dustevol : nullptr, (size_t)np, true, true, &a.dustevol);
  STAGE(st, const_cast<double *>(dust ? ddustevol : nullptr), (size_t)np, true, false, &t8);
  a.ddustevol = t8;
  STAGE(st, const_cast<int8_t *>(iphase), (size_t)np, true, false, &t1);
  a.iphase = t1;
  if (ind) {
    STAGE(st, ibin, (size_t)np, true, true, &a.ibin);
    STAGE(st, ibin_wake, (size_t)np, true, true, &a.ibin_wake);
    STAGE(st, twas, (size_t)np, true, true, &a.twas);
  }
  rc = step_corrector_run(c, np, timei, dtsph, a, nbinmax, ufloor, errmax, v2sum, npc, nwake, nvfloor);
  int rc2 = st.finish();
  return rc ? rc : rc2;
}

int phsph_step_corrector_noalloc(phsph_ctx *c, int32_t np, double timei, double dtsph, const double *xyzh,
                                 int32_t ldx, double *vxyzu, int32_t ldv, const double *fxyzu, int32_t ldf,
                                 const double *vpred, double *Bevol, const double *dBevol, double *dustevol,
                                 const double *ddustevol, const int8_t *iphase, int8_t *ibin, int8_t *ibin_wake,
                                 int32_t nbinmax, double *twas, double ufloor, double *errmax, double *v2sum,
                                 int32_t *npc, int32_t *nwake, int32_t *nvfloor) {
  if (!c) return PHSPH_ERR_INPUT;
  int rc = step_check(c, "step_corrector_noalloc", np, ldx, ldv, ldf);
  if (rc) return rc;
  const bool mhd = c->p.mhd, dust = c->p.use_dustfrac && c->p.ndustsmall > 0, ind = c->p.ind_timesteps;
  if (np > 0 && (!xyzh || !vxyzu || !fxyzu || (!ind && !vpred) || (mhd && (!Bevol || !dBevol)) ||
                 (dust && (!dustevol || !ddustevol)) || (ind && (!ibin || !ibin_wake || !twas))))
    PH_RET(c, PHSPH_ERR_INPUT, "step_corrector_noalloc: missing array");
  if (ind && (nbinmax < 0 || nbinmax > 30))
    PH_RET(c, PHSPH_ERR_INPUT, "step_corrector_noalloc: nbinmax outside 0..30");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  StepArgs a;
  memset(&a, 0, sizeof(a));
  a.xyzh = const_cast<double *>(xyzh); a.ldx = ldx;
  a.vxyzu = vxyzu; a.ldv = ldv;
  a.fxyzu = fxyzu; a.ldf = ldf;
  a.vpred = ind ? nullptr : const_cast<double *>(vpred);
  a.Bevol = mhd ? Bevol : nullptr;
  a.dBevol = mhd ? dBevol : nullptr;
  a.dustevol = dust ? dustevol : nullptr;
  a.ddustevol = dust ? ddustevol : nullptr;
  a.iphase = iphase;
  if (ind) {
    a.ibin = ibin;
    a.ibin_wake = ibin_wake;
    a.twas = twas;
  }
  // no staging: every array the kernel reads or writes must be device memory
  if (np > 0) {
    const void *need[] = {a.xyzh, a.vxyzu, a.fxyzu, a.vpred, a.Bevol, a.dBevol, a.dustevol, a.ddustevol, a.iphase,
                          a.ibin, a.ibin_wake, a.twas};
    for (const void *q : need)
      if (q && !is_device_ptr(q))
        PH_RET(c, PHSPH_ERR_INPUT, "step_corrector_noalloc: every array must be device memory");
  }
  return step_corrector_noalloc_run(c, np, timei, dtsph, a, nbinmax, ufloor, errmax, v2sum, npc, nwake, nvfloor);
}

int phsph_step_graph(phsph_ctx *c, const phsph_step_state *st, phsph_step_clock *clock, int32_t nsub,
                     double ufloor, double tolv, int32_t maxits, int32_t graphs, phsph_step_summary *out) {
  if (!c) return PHSPH_ERR_INPUT;
  if (!st || !clock || nsub < 0 || maxits < 1) PH_RET(c, PHSPH_ERR_INPUT, "step_graph: missing state or clock");
  const phsph_step_state &S = *st;
  const int np = S.np;
  const bool mhd = c->p.mhd, dust = c->p.use_dustfrac && c->p.ndustsmall > 0, ind = c->p.ind_timesteps;
  const bool grav = c->p.gravity;
  if (np <= 0 || (int64_t)np >= (1 << 27)) PH_RET(c, PHSPH_ERR_INPUT, "step_graph: np outside 1..2^27-1");
  if (S.ldxyzh < 4 || S.ldvxyzu < c->p.maxvxyzu || S.ldfxyzu < c->p.maxvxyzu || S.ldvxyzu < 3 ||
      S.lddivcurlv < 1 || S.ldgradh < (grav ? 2 : 1) || S.ldeos < 2)
    PH_RET(c, PHSPH_ERR_INPUT, "step_graph: leading dimension below the component count");
  if (ind && (clock->nbinmax < 0 || clock->nbinmax > 30))
    PH_RET(c, PHSPH_ERR_INPUT, "step_graph: nbinmax outside 0..30");
  // every array the step reads or writes, device memory only
  const void *need[] = {S.xyzh, S.vxyzu, S.fxyzu, S.vpred, S.divcurlv, S.alphaind, S.gradh, S.dvdx, S.eos_vars};
  for (const void *q : need)
    if (!is_device_ptr(q)) PH_RET(c, PHSPH_ERR_INPUT, "step_graph: every array must be device memory");
  if (mhd) {
    const void *m[] = {S.Bevol, S.dBevol, S.Bpred, S.divcurlB, S.divBsymm};
    for (const void *q : m)
      if (!is_device_ptr(q)) PH_RET(c, PHSPH_ERR_INPUT, "step_graph: MHD needs device Bevol, dBevol, Bpred, divcurlB, divBsymm");
  }
  if (dust) {
    const void *d[] = {S.dustevol, S.ddustevol, S.dustpred, S.dustfrac, S.deltav, S.tstop};
    for (const void *q : d)
      if (!is_device_ptr(q)) PH_RET(c, PHSPH_ERR_INPUT, "step_graph: dust needs its device arrays");
  }
  if (ind) {
    const void *b[] = {S.iphase, S.ibin, S.ibin_old, S.ibin_wake, S.twas};
    for (const void *q : b)
      if (!is_device_ptr(q)) PH_RET(c, PHSPH_ERR_INPUT, "step_graph: block timesteps need device iphase and bins");
  } else if (S.iphase && !is_device_ptr(S.iphase)) {
    PH_RET(c, PHSPH_ERR_INPUT, "step_graph: every array must be device memory");
  }
  if (grav && S.poten && !is_device_ptr(S.poten)) PH_RET(c, PHSPH_ERR_INPUT, "step_graph: poten must be device memory");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  phsph_step_clock k = *clock;
  int rc = step_graph_run(c, S, k, nsub, ufloor, tolv, maxits, graphs != 0, out);
  *clock = k;
  return rc;
}

int phsph_step_reset(phsph_ctx *c, int32_t np, double dtsph, double *vxyzu, int32_t ldv, const double *fxyzu,
                     int32_t ldf, double *vpred, double *Bevol, const double *dBevol, double *Bpred,
                     double *dustevol, const double *ddustevol, double *dustpred, const int8_t *iphase) {
  if (!c) return PHSPH_ERR_INPUT;
  int rc = step_check(c, "step_reset", np, 4, ldv, ldf);
  if (rc) return rc;
  const bool mhd = c->p.mhd, dust = c->p.use_dustfrac && c->p.ndustsmall > 0;
  if (np > 0 && (!vxyzu || !fxyzu || !vpred || (mhd && (!Bevol || !dBevol || !Bpred)) ||
                 (dust && (!dustevol || !ddustevol || !dustpred))))
    PH_RET(c, PHSPH_ERR_INPUT, "step_reset: missing array");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  ctx_clear_timings(c);
  Stager st(c);
  StepArgs a;
  memset(&a, 0, sizeof(a));
  double *t8;
  int8_t *t1;
  STAGE(st, vxyzu, (size_t)ldv * np, true, true, &a.vxyzu);
  a.ldv = ldv;
  STAGE(st, const_cast<double *>(fxyzu), (size_t)ldf * np, true, false, &t8);
  a.fxyzu = t8; a.ldf = ldf;
  STAGE(st, vpred, (size_t)ldv * np, true, true, &a.vpred);
  STAGE(st, mhd ? Bevol : nullptr, (size_t)4 * np, true, true, &a.Bevol);
  STAGE(st, const_cast<double *>(mhd ? dBevol : nullptr), (size_t)4 * np, true, false, &t8);
  a.dBevol = t8;
  STAGE(st, mhd ? Bpred : nullptr, (size_t)4 * np, true, true, &a.Bpred);
  STAGE(st, dust ? dustevol : nullptr, (size_t)np, true, true, &a.dustevol);
  STAGE(st, const_cast<double *>(dust ? ddustevol : nullptr), (size_t)np, true, false, &t8);
  a.ddustevol = t8;
  STAGE(st, dust ? dustpred : nullptr, (size_t)np, true, true, &a.dustpred);
  STAGE(st, const_cast<int8_t *>(iphase), (size_t)np, true, false, &t1);
  a.iphase = t1;
  rc = step_reset_run(c, np, dtsph, a);
  int rc2 = st.finish();
  return rc ? rc : rc2;
}

int phsph_get_neighbour_stats(phsph_ctx *c, double *trialmean, double *actualmean, int32_t *maxtrial,
                              int32_t *maxactual, int64_t *nrhocalc, int64_t *nactualtot) {
  if (!c) return PHSPH_ERR_INPUT;
  if ((trialmean || maxtrial) && c->tree_valid && c->nrhocalc >= 0) {
    PH_HIPCHECK(c, hipSetDevice(c->device));
    int rc = dens_trial_stats(c);
    if (rc) return rc;
  }
  if (trialmean) *trialmean = c->trialmean;
  if (actualmean) *actualmean = c->actualmean;
  if (maxtrial) *maxtrial = c->maxtrial;
  if (maxactual) *maxactual = c->maxactual;
  if (nrhocalc) *nrhocalc = c->nrhocalc;
  if (nactualtot) *nactualtot = c->nactualtot;
  return PHSPH_OK;
}

int phsph_get_particle_stats(phsph_ctx *c, int32_t np, int32_t *nneigh, int32_t *nits) {
  if (!c) return PHSPH_ERR_INPUT;
  if (!c->tree_valid || np != c->np || c->nrhocalc < 0)
    PH_RET(c, PHSPH_ERR_STATE, "get_particle_stats: no densityiterate on this tree");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  Stager st(c);
  int32_t *dn, *di;
  if (nneigh) {
    STAGE(st, nneigh, (size_t)np, false, true, &dn);
    hipLaunchKernelGGL(k_zero_i32_api, dim3(nblocks(np, 256)), dim3(256), 0, c->stream, dn, np);
  } else dn = nullptr;
  if (nits) {
    STAGE(st, nits, (size_t)np, false, true, &di);
    hipLaunchKernelGGL(k_zero_i32_api, dim3(nblocks(np, 256)), dim3(256), 0, c->stream, di, np);
  } else di = nullptr;
  int rc = dens_particle_stats(c, dn, di);
  int rc2 = st.finish();
  return rc ? rc : rc2;
}

int phsph_get_force_stats(phsph_ctx *c, int64_t *npairs) {
  if (!c) return PHSPH_ERR_INPUT;
  if (npairs) *npairs = c->force_npairs;
  return PHSPH_OK;
}

int phsph_density_vjp(phsph_ctx *c, int32_t np, const double *xyzh, int32_t ldx, const double *rhobar,
                      const double *hbar, double *dxyz, int32_t ldd) {
  if (!c) return PHSPH_ERR_INPUT;
  if (!xyzh || !dxyz || np <= 0 || ldx < 4 || ldd < 3) PH_RET(c, PHSPH_ERR_INPUT, "density_vjp: bad arguments");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  ctx_clear_timings(c);
  Stager st(c);
  double *dX, *dR, *dH, *dD;
  STAGE(st, const_cast<double *>(xyzh), (size_t)ldx * np, true, false, &dX);
  STAGE(st, const_cast<double *>(rhobar), (size_t)np, true, false, &dR);
  STAGE(st, const_cast<double *>(hbar), (size_t)np, true, false, &dH);
  STAGE(st, dxyz, (size_t)ldd * np, false, true, &dD);
  int rc = vjp_density_run(c, np, dX, ldx, dR, dH, dD, ldd);
  int rc2 = st.finish();
  return rc ? rc : rc2;
}

int phsph_force_vjp(phsph_ctx *c, int32_t np, const double *xyzh, int32_t ldx, const double *vxyzu, int32_t ldv,
                    const double *fbar, int32_t ldfb, double *fxyz, int32_t ldf, double *dxyz, int32_t ldd) {
  if (!c) return PHSPH_ERR_INPUT;
  if (!xyzh || np <= 0 || ldx < 4 || (fxyz && ldf < 3) || (fbar && (ldfb < 3 || !dxyz || ldd < 3)))
    PH_RET(c, PHSPH_ERR_INPUT, "force_vjp: bad arguments");
  if (vxyzu && ldv < c->p.maxvxyzu) PH_RET(c, PHSPH_ERR_INPUT, "force_vjp: bad ldvxyzu");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  ctx_clear_timings(c);
  Stager st(c);
  double *dX, *dV, *dFb, *dF, *dD;
  STAGE(st, const_cast<double *>(xyzh), (size_t)ldx * np, true, false, &dX);
  STAGE(st, const_cast<double *>(vxyzu), (size_t)ldv * np, true, false, &dV);
  STAGE(st, const_cast<double *>(fbar), (size_t)ldfb * np, true, false, &dFb);
  STAGE(st, fxyz, (size_t)ldf * np, false, true, &dF);
  STAGE(st, fbar ? dxyz : nullptr, (size_t)ldd * np, false, true, &dD);
  int rc = vjp_force_run(c, np, dX, ldx, dV, ldv, dFb, ldfb, dF, ldf, dD, ldd);
  int rc2 = st.finish();
  return rc ? rc : rc2;
}

int phsph_step_vjp(phsph_ctx *c, int32_t np, double dt, const double *xyzh, int32_t ldx, const double *vxyzu,
                   int32_t ldv, double *xyzh_new, double *vxyzu_new, const double *xbar, int32_t ldxb,
                   const double *vbar, int32_t ldvb, double *dxyz, double *dvxyz, int32_t ldd) {
  if (!c) return PHSPH_ERR_INPUT;
  const bool back = xbar || vbar;
  if (!xyzh || !vxyzu || np <= 0 || ldx < 4 || ldv < c->p.maxvxyzu || (xbar && ldxb < 3) || (vbar && ldvb < 3) ||
      (back && (!dxyz || !dvxyz || ldd < 3)))
    PH_RET(c, PHSPH_ERR_INPUT, "step_vjp: bad arguments");
  if ((int64_t)np >= (1 << 27)) PH_RET(c, PHSPH_ERR_CAPACITY, "step_vjp: at most 2^27-1 particles");
  PH_HIPCHECK(c, hipSetDevice(c->device));
  ctx_clear_timings(c);
  Stager st(c);
  double *dX, *dV, *dXn, *dVn, *dXb, *dVb, *dDx, *dDv;
  STAGE(st, const_cast<double *>(xyzh), (size_t)ldx * np, true, false, &dX);
  STAGE(st, const_cast<double *>(vxyzu), (size_t)ldv * np, true, false, &dV);
  STAGE(st, xyzh_new, (size_t)ldx * np, true, true, &dXn);
  STAGE(st, vxyzu_new, (size_t)ldv * np, true, true, &dVn);
  STAGE(st, const_cast<double *>(xbar), (size_t)ldxb * np, true, false, &dXb);
  STAGE(st, const_cast<double *>(vbar), (size_t)ldvb * np, true, false, &dVb);
  STAGE(st, back ? dxyz : nullptr, (size_t)ldd * np, false, true, &dDx);
  STAGE(st, back ? dvxyz : nullptr, (size_t)ldd * np, false, true, &dDv);
  int rc = vjp_step_run(c, np, dt, dX, ldx, dV, ldv, dXn, ldx, dVn, ldv, dXb, ldxb, dVb, ldvb, dDx, ldd, dDv, ldd);
  int rc2 = st.finish();
  return rc ? rc : rc2;
}

int phsph_get_precision(phsph_ctx *c, int32_t *dens_prec, int32_t *force_prec) {
  if (!c) return PHSPH_ERR_INPUT;
  if (dens_prec) *dens_prec = c->dens_prec_used;
  if (force_prec) *force_prec = c->force_prec_used;
  return PHSPH_OK;
}

int phsph_get_timings(phsph_ctx *c, const char **names, double *ms, int32_t cap) {
  if (!c) return PHSPH_ERR_INPUT;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  const int n = (int)c->timings.size();
  c->timing_names.resize(n);
  for (int k = 0; k < n && k < cap; k++) {
    float t = 0.f;
    hipEventElapsedTime(&t, c->timings[k].t0, c->timings[k].t1);
    c->timing_names[k] = c->timings[k].name;
    if (names) names[k] = c->timing_names[k].c_str();
    if (ms) ms[k] = (double)t;
  }
  return n;
}

}  // extern "C"
