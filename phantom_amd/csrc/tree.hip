// tree.hip -- kd-tree build on the GPU with the reference's split rule
// (maketree, kdtree.F90:117-314; construct_node :531-929;
// sort_particles_in_cell :937-1001).
//
// The reference builds depth-first with one thread per subtree; with one
// OpenMP thread its result (leaf membership, node centres, in-leaf order)
// is a pure function of the input.  This build reproduces that function:
//
//  * Every node's centre of mass is the sequential FP64 sum over the node's
//    particles in their current order, divided by (total mass * 1/m_gas)
//    (:612-675).  Nodes with at most SUB_MAX particles are built inside one
//    workgroup, one node per thread, with exactly that sequential
//    arithmetic on LDS-resident particles.  Larger nodes (the top levels)
//    sum in parallel and bound the difference to the sequential sum
//    (|S_seq - S| <= 2 gamma_{n-1} sum|a_i|); if any particle lies within
//    that bound of the split plane the node is re-summed sequentially, so
//    the split is always the reference's.
//  * The partition is the reference's two-pointer swap partition.  For a
//    parallel node its permutation is computed directly: the k-th element
//    > pivot in the left part swaps with the k-th element <= pivot counted
//    from the right end of the right part, everything else stays.
//  * Children are numbered 2n, 2n+1 (Gafton & Rosswog indexing, :820-823).
#include "stream.h"

#define SUB_MAX TREE_SUB_MAX  // particles per LDS subtree
#define SUB_BS 256
#define SUB_QMAX 192  // >= 2*floor(SUB_MAX/(MINPART+1)) nodes per level

enum { CNT_NEXT = 0, CNT_ERR = 1, CNT_FALLBACK = 2, CNT_MAXLEVEL = 3, CNT_MINLEVEL = 4, CNT_NAN = 5,
       CNT_NLEAF = 6, CNT_LIVE = 7, CNT_NCHUNK = 8, CNT_NBIG = 9, CNT_SEQMAX = 10, CNT_N = 11 };
enum { TERR_NODEID = 1, TERR_QUEUE = 2, TERR_COUNT = 4 };

// ------------------------------------------------------------------------
// block helpers (256 threads = 4 waves)
// ------------------------------------------------------------------------
template <int BS>
__device__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0.;
  for (int k = 0; k < BS / 64; k++) r += sh[k];  // fixed order: deterministic
  __syncthreads();
  return r;
}
template <int BS>
__device__ double block_max(double v, double *sh) {
  v = wave_max(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = sh[0];
  for (int k = 1; k < BS / 64; k++) r = fmax(r, sh[k]);
  __syncthreads();
  return r;
}
template <int BS>
__device__ double block_min(double v, double *sh) {
  v = wave_min(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = sh[0];
  for (int k = 1; k < BS / 64; k++) r = fmin(r, sh[k]);
  __syncthreads();
  return r;
}
// exclusive scan of 0/1 flags over the block; returns exclusive prefix and
// writes the block total to *total
template <int BS>
__device__ int block_excl_scan(int b, int *shi, int *total) {
  int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned long long m = __ballot(b);
  int pre = __popcll(m & ((1ull << l) - 1ull));
  int wt = __popcll(m);
  __syncthreads();
  if (l == 0) shi[w] = wt;
  __syncthreads();
  int off = 0, tot = 0;
  for (int k = 0; k < BS / 64; k++) {
    if (k < w) off += shi[k];
    tot += shi[k];
  }
  __syncthreads();
  *total = tot;
  return off + pre;
}

// ------------------------------------------------------------------------
// root: periodic wrap, NaN check, bounds, live flags (construct_root_node)
// ------------------------------------------------------------------------
__global__ void k_root_wrap(DevConst dc, double *xyzh, int ld, int np, int32_t *live, double *bred,
                            int32_t *counters) {
  __shared__ double sh[8];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double mn[3] = {PH_HUGE64, PH_HUGE64, PH_HUGE64}, mx[3] = {-PH_HUGE64, -PH_HUGE64, -PH_HUGE64};
  int isl = 0;
  if (i < np) {
    double *r = xyzh + (size_t)i * ld;
    double h = r[3];
    if (h >= PH_TINY64) {
      isl = 1;
      double x = r[0], y = r[1], z = r[2];
      if (dc.periodic) {  // cross_boundary (boundary.f90:123-160)
        if (x < dc.xmin) x = xadd(x, dc.dxbound);
        else if (x > dc.xmax) x = xsub(x, dc.dxbound);
        if (y < dc.ymin) y = xadd(y, dc.dybound);
        else if (y > dc.ymax) y = xsub(y, dc.dybound);
        if (z < dc.zmin) z = xadd(z, dc.dzbound);
        else if (z > dc.zmax) z = xsub(z, dc.dzbound);
        r[0] = x; r[1] = y; r[2] = z;
      }
      if (isnan(x) || isnan(y) || isnan(z)) atomicAdd(&counters[CNT_NAN], 1);
      mn[0] = x; mn[1] = y; mn[2] = z;
      mx[0] = x; mx[1] = y; mx[2] = z;
    }
    live[i] = isl;
  }
  for (int c = 0; c < 3; c++) {
    double a = block_min<256>(mn[c], sh);
    double b = block_max<256>(mx[c], sh);
    if (threadIdx.x == 0) {
      bred[(size_t)blockIdx.x * 6 + c] = a;
      bred[(size_t)blockIdx.x * 6 + 3 + c] = b;
    }
  }
}

// final bounds: starts from particle 1 before any wrap (kdtree.F90:375-380)
__global__ void k_root_bounds(const double *bred, int nb, double x0, double y0, double z0, double *box) {
  __shared__ double sh[8];
  double mn[3] = {x0, y0, z0}, mx[3] = {x0, y0, z0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    for (int c = 0; c < 3; c++) {
      mn[c] = fmin(mn[c], bred[(size_t)b * 6 + c]);
      mx[c] = fmax(mx[c], bred[(size_t)b * 6 + 3 + c]);
    }
  for (int c = 0; c < 3; c++) {
    double a = block_min<256>(mn[c], sh);
    double b = block_max<256>(mx[c], sh);
    if (threadIdx.x == 0) { box[c] = a; box[3 + c] = b; }
  }
}

// stable compaction of live particles: per-block counts
__global__ void k_count_blocks(const int32_t *live, int np, int32_t *bcount) {
  __shared__ int shi[8];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int b = (i < np) ? live[i] : 0;
  int tot;
  block_excl_scan<256>(b, shi, &tot);
  if (threadIdx.x == 0) bcount[blockIdx.x] = tot;
}
// exclusive scan of block counts in place, total to *total: one block of
// 1024 threads, four consecutive counts per thread
__global__ void __launch_bounds__(1024) k_scan_blocks(int32_t *bcount, int nb, int32_t *total) {
  __shared__ int wsum[16];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int carry = 0;
  for (int base = 0; base < nb; base += 4096) {
    const int i0 = base + tid * 4;
    int v[4];
    for (int k = 0; k < 4; k++) v[k] = (i0 + k < nb) ? bcount[i0 + k] : 0;
    const int t = v[0] + v[1] + v[2] + v[3];
    int incl = t;
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(incl, o, 64);
      if (lane >= o) incl += u;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int woff = 0, btot = 0;
    for (int k = 0; k < 16; k++) {
      const int x = wsum[k];
      if (k < w) woff += x;
      btot += x;
    }
    int run = carry + woff + incl - t;
    for (int k = 0; k < 4; k++) {
      if (i0 + k < nb) bcount[i0 + k] = run;
      run += v[k];
    }
    carry += btot;
    __syncthreads();
  }
  if (tid == 0) *total = carry;
}
// scatter live particles into the tree cache (construct_root_node :429-456)
__global__ void k_fill_cache(DevConst dc, const double *xyzh, int ld, int np, const int32_t *live,
                             const int32_t *bcount, const int8_t *iphase, double *tx, double *ty, double *tz,
                             double *th, int32_t *iparts) {
  __shared__ int shi[8];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int b = (i < np) ? live[i] : 0;
  int tot;
  int pre = block_excl_scan<256>(b, shi, &tot);
  if (b) {
    int pos = bcount[blockIdx.x] + pre;
    const double *r = xyzh + (size_t)i * ld;
    tx[pos] = r[0]; ty[pos] = r[1]; tz[pos] = r[2]; th[pos] = r[3];
    int active = iphase ? (iphase[i] > 0) : 1;
    iparts[pos] = (dc.ind_timesteps && !active) ? -(i + 1) : (i + 1);
  }
}

__global__ void k_fill_i32(int32_t *a, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// ------------------------------------------------------------------------
// shared per-node finishing: write the node, decide the split
// ------------------------------------------------------------------------
__device__ __forceinline__ int pick_axis(const double *box) {
  // maxloc(xmaxi - xmini): first maximum (kdtree.F90:815)
  double e0 = xsub(box[3], box[0]), e1 = xsub(box[4], box[1]), e2 = xsub(box[5], box[2]);
  int a = 0;
  double e = e0;
  if (e1 > e) { e = e1; a = 1; }
  if (e2 > e) { a = 2; }
  return a;
}

// ------------------------------------------------------------------------
// large nodes (more than SUB_MAX particles).  Every large node of a level is
// cut into chunks of BN_CH consecutive tree positions, and each pass of
// construct_node runs over all chunks of all large nodes of the level at
// once, one workgroup per chunk.  One wavefront per node combines its chunk
// results in a fixed order.  Per level: plan -> sums -> centre -> ambiguity
// test -> sequential sums (ambiguous nodes only) -> size and left count ->
// node record -> partition (count, scan, scatter, swap) -> child boxes ->
// children.
// ------------------------------------------------------------------------
#define BN_CH 2048
#define BN_BS 256
#define BN_U (BN_CH / BN_BS)
#define BN_SEQ 1024     // positions per LDS stage of the sequential sums
#define BN_SEQ_BS 512   // wavefront 0 sums, the others stage
#define BN_RED 24       // doubles per chunk in ch_red: sums/size/quads, then
#define BN_BOX 12       // from BN_BOX the two children's bounding boxes

// exclusive scan of per-thread counts over the block (block total in *total)
template <int BS>
__device__ int block_excl_scan_int(int v, int *shi, int *total) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  int incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(incl, o, 64);
    if (l >= o) incl += u;
  }
  __syncthreads();
  if (l == 63) shi[w] = incl;
  __syncthreads();
  int off = 0, tot = 0;
  for (int k = 0; k < BS / 64; k++) {
    if (k < w) off += shi[k];
    tot += shi[k];
  }
  __syncthreads();
  *total = tot;
  return off + incl - v;
}

// K per-thread values reduced over the block in a fixed order (kind 0: sum,
// 1: max, 2: min); the result is valid in thread 0
template <int BS, int K>
__device__ void block_reduce(double (&v)[K], const int (&kind)[K], double *sh) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; k++) {
    double x = v[k];
    for (int o = 32; o > 0; o >>= 1) {
      const double y = __shfl_xor(x, o, 64);
      x = (kind[k] == 0) ? x + y : (kind[k] == 1) ? fmax(x, y) : fmin(x, y);
    }
    v[k] = x;
  }
  __syncthreads();
  if (l == 0)
    for (int k = 0; k < K; k++) sh[w * K + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < K; k++) {
      double x = sh[k];
      for (int ww = 1; ww < BS / 64; ww++) {
        const double y = sh[ww * K + k];
        x = (kind[k] == 0) ? x + y : (kind[k] == 1) ? fmax(x, y) : fmin(x, y);
      }
      v[k] = x;
    }
}

__device__ __forceinline__ const double *axis_arr(int a, const double *tx, const double *ty, const double *tz) {
  return (a == 0) ? tx : (a == 1) ? ty : tz;
}

// chunk c of the level: its node (frontier entry) and positions [p0, p1];
// false for a block past the level's chunk count
__device__ __forceinline__ bool bn_chunk(const BigNode *bn, const int32_t *ch_node, const int32_t *counters,
                                         int maxch, int c, int &f, int &p0, int &p1) {
  const int nch = counters[CNT_NCHUNK];
  if (nch > maxch || c >= nch) return false;
  f = ch_node[c];
  p0 = bn[f].lo + (c - bn[f].chunk0) * BN_CH;
  p1 = min(bn[f].hi, p0 + BN_CH - 1);
  return true;
}

// the level's large nodes, their chunks and split axes
__global__ void __launch_bounds__(1024) k_bn_plan(const int32_t *frontier, const double *fbox, int nfront,
                                                  const int32_t *nd_lo, const int32_t *nd_hi, BigNode *bn,
                                                  int32_t *ch_node, int maxch, int32_t *counters) {
  __shared__ int shi[16];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nfront; base += 1024) {
    const int f = base + threadIdx.x;
    int nch = 0, lo = 0, hi = -1, nnode = 0;
    if (f < nfront) {
      nnode = frontier[f];
      lo = nd_lo[nnode];
      hi = nd_hi[nnode];
      const int n = hi - lo + 1;
      nch = (n > SUB_MAX) ? (n + BN_CH - 1) / BN_CH : 0;
    }
    int tot;
    const int off = carry + block_excl_scan_int<1024>(nch, shi, &tot);
    if (f < nfront) {
      BigNode b;
      for (int k = 0; k < 3; k++) b.com[k] = 0.;
      b.tot = 0.; b.hmax = 0.; b.delta = 0.; b.pivot = 0.;
      b.lo = lo; b.hi = hi; b.nnode = nnode;
      b.iaxis = pick_axis(fbox + (size_t)f * 6);
      b.chunk0 = off; b.nchunk = nch;
      b.nl = 0; b.part = 0; b.nbad = 0; b.pad_ = 0;
      bn[f] = b;
      for (int k = 0; k < nch; k++)
        if (off + k < maxch) ch_node[off + k] = f;
    }
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    counters[CNT_NCHUNK] = carry;
    if (carry > maxch) atomicOr(&counters[CNT_ERR], TERR_QUEUE);
  }
}

// parallel sums of m*x over a chunk, their absolute values, and hmax
__global__ void __launch_bounds__(BN_BS) k_bn_sums(DevConst dc, const BigNode *bn, const int32_t *ch_node,
                                                   int maxch, const int32_t *counters, const double *tx,
                                                   const double *ty, const double *tz, const double *th,
                                                   double *ch_red) {
  __shared__ double sh[BN_BS / 64 * 7];
  const int c = blockIdx.x;
  int f, p0, p1;
  if (!bn_chunk(bn, ch_node, counters, maxch, c, f, p0, p1)) return;
  const double pm = dc.pmass;
  const double dfac = (pm > 0.) ? 1. / pm : 1.;
  const double fac = xmul(pm, dfac);
  double v[7] = {0., 0., 0., 0., 0., 0., 0.};
#pragma unroll
  for (int u = 0; u < BN_U; u++) {
    const int p = p0 + u * BN_BS + threadIdx.x;
    if (p <= p1) {
      const double vx = xmul(fac, tx[p]), vy = xmul(fac, ty[p]), vz = xmul(fac, tz[p]);
      v[0] += vx; v[1] += vy; v[2] += vz;
      v[3] += fabs(vx); v[4] += fabs(vy); v[5] += fabs(vz);
      v[6] = fmax(v[6], th[p]);
    }
  }
  const int kind[7] = {0, 0, 0, 0, 0, 0, 1};
  block_reduce<BN_BS, 7>(v, kind, sh);
  if (threadIdx.x == 0)
    for (int k = 0; k < 7; k++) ch_red[(size_t)c * BN_RED + k] = v[k];
}

// centre of mass from the parallel sums, and the bound on its distance from
// the sequential one on the split axis (|S_seq - S| <= 2 gamma_{n-1} sum|a_i|)
__global__ void k_bn_centre(DevConst dc, BigNode *bn, int maxch, const int32_t *counters, const double *ch_red) {
  const int f = blockIdx.x, lane = threadIdx.x;
  if (counters[CNT_NCHUNK] > maxch) return;
  BigNode *b = bn + f;
  const int nchunk = b->nchunk, chunk0 = b->chunk0;
  if (nchunk == 0) return;
  double s[3] = {0., 0., 0.}, a[3] = {0., 0., 0.}, hm = 0.;
  for (int k = lane; k < nchunk; k += 64) {
    const double *r = ch_red + (size_t)(chunk0 + k) * BN_RED;
    s[0] += r[0]; s[1] += r[1]; s[2] += r[2];
    a[0] += r[3]; a[1] += r[4]; a[2] += r[5];
    hm = fmax(hm, r[6]);
  }
  for (int k = 0; k < 3; k++) {
    s[k] = wave_sum(s[k]);
    a[k] = wave_sum(a[k]);
  }
  hm = wave_max(hm);
  if (lane == 0) {
    const int n = b->hi - b->lo + 1;
    const double pm = dc.pmass;
    const double dfac = (pm > 0.) ? 1. / pm : 1.;
    const double tot = xmul((double)n, pm);
    const double D = xmul(tot, dfac);
    double com[3];
    for (int k = 0; k < 3; k++) com[k] = s[k] / D;
    const double uu = 1.1102230246251565e-16;
    const double nu = (double)(n - 1) * uu;
    const double gam = nu / (1. - nu);
    const int ia = b->iaxis;
    const double a_ax = a[ia], c_ax = com[ia];
    for (int k = 0; k < 3; k++) b->com[k] = com[k];
    b->tot = tot;
    b->hmax = hm;
    b->delta = 2. * (2. * gam * a_ax / D + (2. * gam + 4. * uu) * fabs(c_ax)) + 1e-300;
    b->pivot = c_ax;
  }
}

// does any particle of the chunk lie within delta of the split plane?
__global__ void __launch_bounds__(BN_BS) k_bn_amb(const BigNode *bn, const int32_t *ch_node, int maxch,
                                                  const int32_t *counters, const double *tx, const double *ty,
                                                  const double *tz, int32_t *ch_i) {
  const int c = blockIdx.x;
  int f, p0, p1;
  if (!bn_chunk(bn, ch_node, counters, maxch, c, f, p0, p1)) return;
  const double *tax = axis_arr(bn[f].iaxis, tx, ty, tz);
  const double c_ax = bn[f].pivot, delta = bn[f].delta;
  int amb = 0;
#pragma unroll
  for (int u = 0; u < BN_U; u++) {
    const int p = p0 + u * BN_BS + threadIdx.x;
    if (p <= p1) amb |= (fabs(tax[p] - c_ax) <= delta);
  }
  amb = __syncthreads_or(amb);
  if (threadIdx.x == 0) ch_i[(size_t)c * 4 + 0] = amb;
}

// cnt additions of one running sum in order, from an LDS row, with the
// next 16 values loaded while the current 16 are added (two register sets
// in turn).  Reads up to 15 values past cnt: the rows are BN_SEQ_LD long.
__device__ __forceinline__ double seq_chain(double acc, const double *src, int cnt) {
  int k = 0;
  if (cnt >= 16) {
    double v[16], w[16];
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = src[u];
    while (true) {
#pragma unroll
      for (int u = 0; u < 16; u++) w[u] = src[k + 16 + u];
#pragma unroll
      for (int u = 0; u < 16; u++) acc = xadd(acc, v[u]);
      k += 16;
      if (k + 16 > cnt) break;
#pragma unroll
      for (int u = 0; u < 16; u++) v[u] = src[k + 16 + u];
#pragma unroll
      for (int u = 0; u < 16; u++) acc = xadd(acc, w[u]);
      k += 16;
      if (k + 16 > cnt) break;
    }
  }
  for (; k < cnt; k++) acc = xadd(acc, src[k]);
  return acc;
}

// exact sequential sums (kdtree.F90:655-665) of the nodes with a particle
// within delta of the plane.  The four running sums (x, y, z, mass) are four
// dependent chains in lanes 0-3 of wavefront 0; the other wavefronts stage
// the products fac*x of the next BN_SEQ positions into LDS meanwhile.  The
// rows are padded so that the four lanes read distinct LDS banks.
#define BN_SEQ_LD (BN_SEQ + 20)
__global__ void __launch_bounds__(BN_SEQ_BS) k_bn_seq(DevConst dc, BigNode *bn, int maxch, int32_t *counters,
                                                      const int32_t *ch_i, const double *tx, const double *ty,
                                                      const double *tz) {
  __shared__ double buf[2][3][BN_SEQ_LD];
  __shared__ double mrow[BN_SEQ_LD];
  __shared__ double seq[4];
  const int f = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
  if (counters[CNT_NCHUNK] > maxch) return;
  const int nchunk = bn[f].nchunk, chunk0 = bn[f].chunk0;
  if (nchunk == 0) return;
  int amb = 0;
  for (int k = tid; k < nchunk; k += BN_SEQ_BS) amb |= ch_i[(size_t)(chunk0 + k) * 4 + 0];
  amb = __syncthreads_or(amb);
  if (!amb) return;
  const int lo = bn[f].lo, hi = bn[f].hi, iaxis = bn[f].iaxis;
  const double pm = dc.pmass;
  const double dfac = (pm > 0.) ? 1. / pm : 1.;
  const double fac = xmul(pm, dfac);
  const int n = hi - lo + 1;
  const int nst = (n + BN_SEQ - 1) / BN_SEQ;
  auto fill = [&](int st) {
    const int base = lo + st * BN_SEQ;
    const int cnt = min(BN_SEQ, hi - base + 1);
    double(*dst)[BN_SEQ_LD] = buf[st & 1];
    for (int k = tid - 64; k < cnt; k += BN_SEQ_BS - 64) {
      dst[0][k] = xmul(fac, tx[base + k]);
      dst[1][k] = xmul(fac, ty[base + k]);
      dst[2][k] = xmul(fac, tz[base + k]);
    }
  };
  for (int k = tid; k < BN_SEQ_LD; k += BN_SEQ_BS) mrow[k] = pm;
  if (tid >= 64) fill(0);
  __syncthreads();
  double acc = 0.;
  for (int st = 0; st < nst; st++) {
    if (tid >= 64) {
      if (st + 1 < nst) fill(st + 1);
    } else if (lane < 4) {
      const int cnt = min(BN_SEQ, hi - (lo + st * BN_SEQ) + 1);
      acc = seq_chain(acc, lane < 3 ? buf[st & 1][lane] : mrow, cnt);
    }
    __syncthreads();
  }
  if (tid < 4) seq[tid] = acc;
  __syncthreads();
  if (tid == 0) {
    const double tot = seq[3];
    const double D = xmul(tot, dfac);
    BigNode *b = bn + f;
    for (int k = 0; k < 3; k++) b->com[k] = seq[k] / D;
    b->tot = tot;
    b->pivot = b->com[iaxis];
    atomicAdd(&counters[CNT_FALLBACK], 1);
    atomicMax(&counters[CNT_SEQMAX], n);
  }
}

// size (kdtree.F90:697-782) and the count of particles on the left of the
// pivot, per chunk; under self-gravity also the chunk's share of the second
// moments about the centre of mass (kdtree.F90:724-750)
__global__ void __launch_bounds__(BN_BS) k_bn_size(DevConst dc, const BigNode *bn, const int32_t *ch_node,
                                                   int maxch, const int32_t *counters, const double *tx,
                                                   const double *ty, const double *tz, double *ch_red,
                                                   int32_t *ch_i, int quads) {
  __shared__ double sh[BN_BS / 64 * 20];
  const int c = blockIdx.x;
  int f, p0, p1;
  if (!bn_chunk(bn, ch_node, counters, maxch, c, f, p0, p1)) return;
  const double c0 = bn[f].com[0], c1 = bn[f].com[1], c2 = bn[f].com[2], pivot = bn[f].pivot;
  const double *tax = axis_arr(bn[f].iaxis, tx, ty, tz);
  const double pm = dc.pmass;
  // [8..13] bounding box of the particles on the left of the pivot (the
  // left child after a split at the pivot), [14..19] of those on the right
  double v[20] = {0., 0., 0., 0., 0., 0., 0., 0.};
  for (int k = 0; k < 3; k++) {
    v[8 + k] = PH_HUGE64; v[11 + k] = -PH_HUGE64;
    v[14 + k] = PH_HUGE64; v[17 + k] = -PH_HUGE64;
  }
#pragma unroll
  for (int u = 0; u < BN_U; u++) {
    const int p = p0 + u * BN_BS + threadIdx.x;
    if (p <= p1) {
      const double r[3] = {tx[p], ty[p], tz[p]};
      v[0] = fmax(v[0], r2_exact(xsub(r[0], c0), xsub(r[1], c1), xsub(r[2], c2)));
      const bool left = tax[p] <= pivot;
      v[1] += left ? 1. : 0.;
      const int o = left ? 8 : 14;
      for (int k = 0; k < 3; k++) {
        v[o + k] = fmin(v[o + k], r[k]);
        v[o + 3 + k] = fmax(v[o + 3 + k], r[k]);
      }
      if (quads) {
        const double dx = r[0] - c0, dy = r[1] - c1, dz = r[2] - c2;
        v[2] += pm * (dx * dx); v[3] += pm * (dx * dy); v[4] += pm * (dx * dz);
        v[5] += pm * (dy * dy); v[6] += pm * (dy * dz); v[7] += pm * (dz * dz);
      }
    }
  }
  const int kind[20] = {1, 0, 0, 0, 0, 0, 0, 0, 2, 2, 2, 1, 1, 1, 2, 2, 2, 1, 1, 1};
  block_reduce<BN_BS, 20>(v, kind, sh);
  if (threadIdx.x == 0) {
    ch_red[(size_t)c * BN_RED + 0] = v[0];
    ch_i[(size_t)c * 4 + 1] = (int)v[1];
    if (quads)
      for (int k = 0; k < 6; k++) ch_red[(size_t)c * BN_RED + 1 + k] = v[2 + k];
    for (int k = 0; k < 12; k++) ch_red[(size_t)c * BN_RED + BN_BOX + k] = v[8 + k];
  }
}

// the node record, and the split: left count or, with everything on one
// side, half and half in place (kdtree.F90:856-865)
__global__ void k_bn_node(BigNode *bn, int maxch, int32_t *counters, const double *ch_red, const int32_t *ch_i,
                          phsph_kdnode *nd, int32_t *leaf_flag, int level, int maxlevel_indexed, int64_t nslots,
                          double *quads) {
  const int f = blockIdx.x, lane = threadIdx.x;
  if (counters[CNT_NCHUNK] > maxch) return;
  BigNode *b = bn + f;
  const int nchunk = b->nchunk, chunk0 = b->chunk0;
  if (nchunk == 0) return;
  double r2 = 0.;
  int cl = 0;
  for (int k = lane; k < nchunk; k += 64) {
    r2 = fmax(r2, ch_red[(size_t)(chunk0 + k) * BN_RED + 0]);
    cl += ch_i[(size_t)(chunk0 + k) * 4 + 1];
  }
  r2 = wave_max(r2);
  cl = wave_sum_i(cl);
  if (quads) {  // second moments: chunk shares in chunk order per lane, then across lanes
    double q[6] = {0., 0., 0., 0., 0., 0.};
    for (int k = lane; k < nchunk; k += 64)
      for (int j = 0; j < 6; j++) q[j] += ch_red[(size_t)(chunk0 + k) * BN_RED + 1 + j];
    for (int j = 0; j < 6; j++) q[j] = wave_sum(q[j]);
    if (lane == 0)
      for (int j = 0; j < 6; j++) quads[6 * (size_t)b->nnode + j] = q[j];
  }
  if (lane == 0) {
    const int nnode = b->nnode, il = 2 * nnode, ir = il + 1;
    const int n = b->hi - b->lo + 1;
    phsph_kdnode e;
    e.xcen[0] = b->com[0]; e.xcen[1] = b->com[1]; e.xcen[2] = b->com[2];
    e.size = xadd(sqrt(r2), PH_EPS64);
    e.hmax = b->hmax;
    e.mass = b->tot;
    e.leftchild = il;
    e.rightchild = ir;
    e.parent = nnode / 2;
    e.idum = 0;
    nd[nnode] = e;
    leaf_flag[nnode] = 0;
    b->nbad = 0;
    if (level >= maxlevel_indexed || (int64_t)ir >= nslots) {
      atomicOr(&counters[CNT_ERR], TERR_NODEID);
      b->nl = 0;
      b->part = 0;
    } else if (cl > 0 && cl < n) {
      b->nl = cl;
      b->part = 1;
    } else {
      b->nl = n / 2;
      b->part = 0;
    }
  }
}

// sort_particles_in_cell as a permutation: the k-th element > pivot in the
// left part [lo, lo+nl-1] swaps with the k-th element <= pivot counted from
// the right end of the right part.  Per chunk, the counts of both kinds.
__global__ void __launch_bounds__(BN_BS) k_bn_pcount(const BigNode *bn, const int32_t *ch_node, int maxch,
                                                     const int32_t *counters, const double *tx, const double *ty,
                                                     const double *tz, int32_t *ch_i) {
  __shared__ double sh[BN_BS / 64 * 2];
  const int c = blockIdx.x;
  int f, p0, p1;
  if (!bn_chunk(bn, ch_node, counters, maxch, c, f, p0, p1)) return;
  if (!bn[f].part) return;
  const int lend = bn[f].lo + bn[f].nl - 1;
  const double pivot = bn[f].pivot;
  const double *tax = axis_arr(bn[f].iaxis, tx, ty, tz);
  double v[2] = {0., 0.};
#pragma unroll
  for (int u = 0; u < BN_U; u++) {
    const int p = p0 + u * BN_BS + threadIdx.x;
    if (p <= p1) {
      const bool le = tax[p] <= pivot;
      if (p <= lend) v[0] += le ? 0. : 1.;
      else v[1] += le ? 1. : 0.;
    }
  }
  const int kind[2] = {0, 0};
  block_reduce<BN_BS, 2>(v, kind, sh);
  if (threadIdx.x == 0) {
    ch_i[(size_t)c * 4 + 2] = (int)v[0];
    ch_i[(size_t)c * 4 + 3] = (int)v[1];
  }
}

// per node: chunk offsets of both kinds in position order
__global__ void k_bn_pscan(BigNode *bn, int maxch, int32_t *counters, const int32_t *ch_i, int32_t *ch_off) {
  const int f = blockIdx.x, lane = threadIdx.x;
  if (counters[CNT_NCHUNK] > maxch) return;
  BigNode *b = bn + f;
  if (b->nchunk == 0 || !b->part) return;
  const int nchunk = b->nchunk, chunk0 = b->chunk0;
  int carry_l = 0, carry_r = 0;
  for (int base = 0; base < nchunk; base += 64) {
    const int k = base + lane;
    int vl = 0, vr = 0;
    if (k < nchunk) {
      vl = ch_i[(size_t)(chunk0 + k) * 4 + 2];
      vr = ch_i[(size_t)(chunk0 + k) * 4 + 3];
    }
    int il = vl, ir = vr;
    for (int o = 1; o < 64; o <<= 1) {
      const int ul = __shfl_up(il, o, 64), ur = __shfl_up(ir, o, 64);
      if (lane >= o) { il += ul; ir += ur; }
    }
    if (k < nchunk) {
      ch_off[(size_t)(chunk0 + k) * 2 + 0] = carry_l + il - vl;
      ch_off[(size_t)(chunk0 + k) * 2 + 1] = carry_r + ir - vr;
    }
    carry_l += __shfl(il, 63, 64);
    carry_r += __shfl(ir, 63, 64);
  }
  if (lane == 0) {
    b->nbad = min(carry_l, carry_r);
    if (carry_l != carry_r) atomicOr(&counters[CNT_ERR], TERR_COUNT);
  }
}

// ranks of the misplaced elements: the k-th left one goes to tmpA[lo+k],
// the k-th right one counted from the right end to tmpB[lo+k]
__global__ void __launch_bounds__(BN_BS) k_bn_pscatter(const BigNode *bn, const int32_t *ch_node, int maxch,
                                                       const int32_t *counters, const double *tx,
                                                       const double *ty, const double *tz, const int32_t *ch_off,
                                                       int32_t *tmpA, int32_t *tmpB) {
  __shared__ int shi[BN_BS / 64];
  const int c = blockIdx.x;
  int f, p0, p1;
  if (!bn_chunk(bn, ch_node, counters, maxch, c, f, p0, p1)) return;
  if (!bn[f].part) return;
  const int lo = bn[f].lo, lend = lo + bn[f].nl - 1, nbad = bn[f].nbad;
  const double pivot = bn[f].pivot;
  const double *tax = axis_arr(bn[f].iaxis, tx, ty, tz);
  const int q0 = p0 + threadIdx.x * BN_U;  // BN_U consecutive positions per thread: ranks follow positions
  int fl[BN_U], fr[BN_U], nl_ = 0, nr_ = 0;
#pragma unroll
  for (int u = 0; u < BN_U; u++) {
    const int p = q0 + u;
    fl[u] = 0;
    fr[u] = 0;
    if (p <= p1) {
      const bool le = tax[p] <= pivot;
      if (p <= lend) fl[u] = le ? 0 : 1;
      else fr[u] = le ? 1 : 0;
    }
    nl_ += fl[u];
    nr_ += fr[u];
  }
  int tl, tr;
  int rl = ch_off[(size_t)c * 2 + 0] + block_excl_scan_int<BN_BS>(nl_, shi, &tl);
  int rr = ch_off[(size_t)c * 2 + 1] + block_excl_scan_int<BN_BS>(nr_, shi, &tr);
#pragma unroll
  for (int u = 0; u < BN_U; u++) {
    if (fl[u]) {
      const int k = rl++;
      if (k < nbad) tmpA[lo + k] = q0 + u;
    }
    if (fr[u]) {
      const int k = nbad - 1 - rr++;
      if (k >= 0) tmpB[lo + k] = q0 + u;
    }
  }
}

// the swaps: chunk c of a node takes pairs k0 .. k0+BN_CH-1
__global__ void __launch_bounds__(BN_BS) k_bn_swap(const BigNode *bn, const int32_t *ch_node, int maxch,
                                                   const int32_t *counters, const int32_t *tmpA,
                                                   const int32_t *tmpB, double *tx, double *ty, double *tz,
                                                   double *th, int32_t *iparts) {
  const int c = blockIdx.x;
  int f, p0, p1;
  if (!bn_chunk(bn, ch_node, counters, maxch, c, f, p0, p1)) return;
  if (!bn[f].part) return;
  const int lo = bn[f].lo, hi = bn[f].hi, nbad = bn[f].nbad;
  const int k0 = p0 - lo;
#pragma unroll
  for (int u = 0; u < BN_U; u++) {
    const int k = k0 + u * BN_BS + threadIdx.x;
    if (k < nbad) {
      const int pa = tmpA[lo + k], pb = tmpB[lo + k];
      if (pa >= lo && pa <= hi && pb >= lo && pb <= hi) {
        const double t0 = tx[pa], t1 = ty[pa], t2 = tz[pa], t3 = th[pa];
        const int32_t ti = iparts[pa];
        tx[pa] = tx[pb]; ty[pa] = ty[pb]; tz[pa] = tz[pb]; th[pa] = th[pb]; iparts[pa] = iparts[pb];
        tx[pb] = t0; ty[pb] = t1; tz[pb] = t2; th[pb] = t3; iparts[pb] = ti;
      }
    }
  }
}

// bounding boxes of both children, per chunk
__global__ void __launch_bounds__(BN_BS) k_bn_box(const BigNode *bn, const int32_t *ch_node, int maxch,
                                                  const int32_t *counters, const double *tx, const double *ty,
                                                  const double *tz, double *ch_red) {
  __shared__ double sh[BN_BS / 64 * 12];
  const int c = blockIdx.x;
  int f, p0, p1;
  if (!bn_chunk(bn, ch_node, counters, maxch, c, f, p0, p1)) return;
  // a node split at the pivot has its children's boxes from k_bn_size
  if (bn[f].nl == 0 || bn[f].part) return;
  const int hA = bn[f].lo + bn[f].nl - 1;
  double v[12];
  for (int k = 0; k < 3; k++) {
    v[k] = PH_HUGE64; v[3 + k] = -PH_HUGE64;
    v[6 + k] = PH_HUGE64; v[9 + k] = -PH_HUGE64;
  }
#pragma unroll
  for (int u = 0; u < BN_U; u++) {
    const int p = p0 + u * BN_BS + threadIdx.x;
    if (p <= p1) {
      const double r[3] = {tx[p], ty[p], tz[p]};
      const int o = (p <= hA) ? 0 : 6;
      for (int k = 0; k < 3; k++) {
        v[o + k] = fmin(v[o + k], r[k]);
        v[o + 3 + k] = fmax(v[o + 3 + k], r[k]);
      }
    }
  }
  const int kind[12] = {2, 2, 2, 1, 1, 1, 2, 2, 2, 1, 1, 1};
  block_reduce<BN_BS, 12>(v, kind, sh);
  if (threadIdx.x == 0)
    for (int k = 0; k < 12; k++) ch_red[(size_t)c * BN_RED + BN_BOX + k] = v[k];
}

// children: ranges, boxes, next frontier
__global__ void k_bn_finish(const BigNode *bn, int maxch, int32_t *counters, const double *ch_red,
                            int32_t *nd_lo, int32_t *nd_hi, int32_t *next_front, double *next_box) {
  const int f = blockIdx.x, lane = threadIdx.x;
  if (counters[CNT_NCHUNK] > maxch) return;
  const BigNode *b = bn + f;
  const int nchunk = b->nchunk, chunk0 = b->chunk0, nl = b->nl;
  if (nchunk == 0 || nl == 0) return;
  double v[12];
  for (int k = 0; k < 3; k++) {
    v[k] = PH_HUGE64; v[3 + k] = -PH_HUGE64;
    v[6 + k] = PH_HUGE64; v[9 + k] = -PH_HUGE64;
  }
  for (int j = lane; j < nchunk; j += 64) {
    const double *r = ch_red + (size_t)(chunk0 + j) * BN_RED + BN_BOX;
    for (int k = 0; k < 3; k++) {
      v[k] = fmin(v[k], r[k]); v[3 + k] = fmax(v[3 + k], r[3 + k]);
      v[6 + k] = fmin(v[6 + k], r[6 + k]); v[9 + k] = fmax(v[9 + k], r[9 + k]);
    }
  }
  for (int k = 0; k < 3; k++) {
    v[k] = wave_min(v[k]); v[3 + k] = wave_max(v[3 + k]);
    v[6 + k] = wave_min(v[6 + k]); v[9 + k] = wave_max(v[9 + k]);
  }
  if (lane == 0) {
    const int il = 2 * b->nnode, ir = il + 1;
    const int lA = b->lo, hA = b->lo + nl - 1, lB = b->lo + nl, hB = b->hi;
    nd_lo[il] = lA; nd_hi[il] = hA; nd_lo[ir] = lB; nd_hi[ir] = hB;
    const int idx = atomicAdd(&counters[CNT_NEXT], 2);
    next_front[idx] = il;
    next_front[idx + 1] = ir;
    for (int k = 0; k < 6; k++) {
      next_box[(size_t)idx * 6 + k] = v[k];
      next_box[(size_t)(idx + 1) * 6 + k] = v[6 + k];
    }
    const int nbig = (hA - lA + 1 > SUB_MAX ? 1 : 0) + (hB - lB + 1 > SUB_MAX ? 1 : 0);
    if (nbig) atomicAdd(&counters[CNT_NBIG], nbig);
  }
}

// ------------------------------------------------------------------------
// small nodes: one workgroup builds the whole subtree in LDS, level by
// level, with the reference's sequential arithmetic.  A node of more than
// SUB_WAVE_MIN particles is built by one wavefront: its maxima, minima and
// the partition run across the lanes (the partition as the permutation of
// sort_particles_in_cell, see k_bn_pcount), and its four running sums are
// four dependent chains in lanes 0-3.  Smaller nodes take one thread each.
// ------------------------------------------------------------------------
#define SUB_WAVE_MIN 64

// the running sum of one node's row in position order: lanes 0-2 add
// fac*x, fac*y, fac*z, lane 3 adds the particle mass (kdtree.F90:655-665)
__device__ __forceinline__ double sub_chain(const double *row, int a0, int a1, bool ism, double pm, double fac) {
  double acc = 0.;
  int k = a0;
  for (; k + 16 <= a1 + 1; k += 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = row[k + u];
#pragma unroll
    for (int u = 0; u < 16; u++) acc = xadd(acc, ism ? pm : xmul(fac, v[u]));
  }
  for (; k <= a1; k++) acc = xadd(acc, ism ? pm : xmul(fac, row[k]));
  return acc;
}

__global__ void __launch_bounds__(SUB_BS) k_subtrees(DevConst dc, const int32_t *frontier, const double *fbox,
                                                     int nfront, double *tx, double *ty, double *tz, double *th,
                                                     int32_t *iparts, int32_t *nd_lo, int32_t *nd_hi,
                                                     phsph_kdnode *nd, int32_t *leaf_flag, int level0,
                                                     int maxlevel_indexed, int64_t nslots, int32_t *counters) {
  __shared__ double sx[SUB_MAX], sy[SUB_MAX], sz[SUB_MAX], shh[SUB_MAX];
  __shared__ int32_t sip[SUB_MAX];
  __shared__ int32_t sidx[SUB_MAX];  // partition pairs of the wavefront-built nodes
  __shared__ int32_t qn[2][SUB_QMAX], ql[2][SUB_QMAX], qh[2][SUB_QMAX];
  __shared__ double qb[2][SUB_QMAX][6];
  __shared__ int qcount[2];
  const int f = blockIdx.x;
  const int root = frontier[f];
  const int lo = nd_lo[root], hi = nd_hi[root];
  const int n = hi - lo + 1;
  if (n > SUB_MAX) return;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int k = tid; k < n; k += SUB_BS) {
    sx[k] = tx[lo + k]; sy[k] = ty[lo + k]; sz[k] = tz[lo + k]; shh[k] = th[lo + k]; sip[k] = iparts[lo + k];
  }
  if (tid == 0) {
    qn[0][0] = root; ql[0][0] = 0; qh[0][0] = n - 1;
    for (int c = 0; c < 6; c++) qb[0][0][c] = fbox[(size_t)f * 6 + c];
    qcount[0] = 1;
    qcount[1] = 0;
  }
  __syncthreads();
  const double pm = dc.pmass;
  const double dfac = (pm > 0.) ? 1. / pm : 1.;
  const double fac = xmul(pm, dfac);
  int cur = 0;
  int level = level0;
  int maxlev = 0, minlev = 1 << 30;
  while (qcount[cur] > 0) {
    const int nq = min(qcount[cur], SUB_QMAX);
    const int nxt = cur ^ 1;
    // nodes of more than SUB_WAVE_MIN particles: one wavefront each
    for (int t = wv; t < nq; t += SUB_BS / 64) {
      const int a0 = ql[cur][t], a1 = qh[cur][t];
      const int npn = a1 - a0 + 1;
      if (npn <= SUB_WAVE_MIN) continue;
      const int nnode = qn[cur][t];
      double hmax = 0.;
      for (int k = a0 + lane; k <= a1; k += 64) hmax = (shh[k] > hmax) ? shh[k] : hmax;
      hmax = wave_max(hmax);
      double acc = 0.;
      if (lane < 4) acc = sub_chain(lane == 0 ? sx : (lane == 1 ? sy : sz), a0, a1, lane == 3, pm, fac);
      const double xc = __shfl(acc, 0, 64), yc = __shfl(acc, 1, 64), zc = __shfl(acc, 2, 64);
      const double tot = __shfl(acc, 3, 64);
      const double D = xmul(tot, dfac);
      const double com0 = xc / D, com1 = yc / D, com2 = zc / D;
      double r2max = 0.;
      for (int k = a0 + lane; k <= a1; k += 64) {
        const double r2 = r2_exact(xsub(sx[k], com0), xsub(sy[k], com1), xsub(sz[k], com2));
        r2max = (r2 > r2max) ? r2 : r2max;
      }
      r2max = wave_max(r2max);
      const int il = 2 * nnode, ir = il + 1;
      const bool bad_id = (level >= maxlevel_indexed || (int64_t)ir >= nslots);
      if (lane == 0) {
        phsph_kdnode e;
        e.xcen[0] = com0; e.xcen[1] = com1; e.xcen[2] = com2;
        e.size = xadd(sqrt(r2max), PH_EPS64);
        e.hmax = hmax;
        e.mass = tot;
        e.leftchild = il;
        e.rightchild = ir;
        e.parent = nnode / 2;
        e.idum = 0;
        nd[nnode] = e;
        nd_lo[nnode] = lo + a0;
        nd_hi[nnode] = lo + a1;
        leaf_flag[nnode] = 0;
        if (bad_id) atomicOr(&counters[CNT_ERR], TERR_NODEID);
      }
      if (bad_id) continue;
      const double *box = qb[cur][t];
      const int iaxis = pick_axis(box);
      const double pivot = (iaxis == 0) ? com0 : (iaxis == 1) ? com1 : com2;
      const double *sa = (iaxis == 0) ? sx : (iaxis == 1) ? sy : sz;
      int cl = 0;
      for (int base = a0; base <= a1; base += 64) {
        const int k = base + lane;
        cl += __popcll(__ballot(k <= a1 && sa[k] <= pivot));
      }
      int nl = npn / 2;  // everything on one side: halves in place (kdtree.F90:856-865)
      if (cl > 0 && cl < npn) {
        // the k-th element > pivot of the left part [a0, a0+cl-1] swaps
        // with the k-th element <= pivot counted from the right end
        const int lend = a0 + cl - 1;
        const unsigned long long lt = lanemask_lt();
        int nbad = 0;
        for (int base = a0; base <= lend; base += 64) {
          const int k = base + lane;
          const bool b = k <= lend && !(sa[k] <= pivot);
          const unsigned long long m = __ballot(b);
          if (b) sidx[a0 + nbad + __popcll(m & lt)] = k;
          nbad += __popcll(m);
        }
        int nr = 0;
        for (int base = lend + 1; base <= a1; base += 64) {
          const int k = base + lane;
          const bool b = k <= a1 && sa[k] <= pivot;
          const unsigned long long m = __ballot(b);
          const int j = nbad - 1 - (nr + __popcll(m & lt));
          if (b && j >= 0) sidx[a0 + nbad + j] = k;
          nr += __popcll(m);
        }
        if (nr != nbad && lane == 0) atomicOr(&counters[CNT_ERR], TERR_COUNT);
        const int npair = min(nbad, nr);
        wave_lds_sync();
        for (int j = lane; j < npair; j += 64) {
          const int pa = sidx[a0 + j], pb = sidx[a0 + nbad + j];
          const double t0 = sx[pa], t1 = sy[pa], t2 = sz[pa], t3 = shh[pa];
          const int32_t ti = sip[pa];
          sx[pa] = sx[pb]; sy[pa] = sy[pb]; sz[pa] = sz[pb]; shh[pa] = shh[pb]; sip[pa] = sip[pb];
          sx[pb] = t0; sy[pb] = t1; sz[pb] = t2; shh[pb] = t3; sip[pb] = ti;
        }
        wave_lds_sync();
        nl = cl;
      }
      const int lA = a0, hA = a0 + nl - 1, lB = a0 + nl, hB = a1;
      double bA[6], bB[6];
      for (int c = 0; c < 3; c++) {
        bA[c] = PH_HUGE64; bA[3 + c] = -PH_HUGE64;
        bB[c] = PH_HUGE64; bB[3 + c] = -PH_HUGE64;
      }
      for (int k = lA + lane; k <= hA; k += 64) {
        bA[0] = fmin(bA[0], sx[k]); bA[1] = fmin(bA[1], sy[k]); bA[2] = fmin(bA[2], sz[k]);
        bA[3] = fmax(bA[3], sx[k]); bA[4] = fmax(bA[4], sy[k]); bA[5] = fmax(bA[5], sz[k]);
      }
      for (int k = lB + lane; k <= hB; k += 64) {
        bB[0] = fmin(bB[0], sx[k]); bB[1] = fmin(bB[1], sy[k]); bB[2] = fmin(bB[2], sz[k]);
        bB[3] = fmax(bB[3], sx[k]); bB[4] = fmax(bB[4], sy[k]); bB[5] = fmax(bB[5], sz[k]);
      }
      for (int c = 0; c < 3; c++) {
        bA[c] = wave_min(bA[c]); bA[3 + c] = wave_max(bA[3 + c]);
        bB[c] = wave_min(bB[c]); bB[3 + c] = wave_max(bB[3 + c]);
      }
      if (lane == 0) {
        const int slot = atomicAdd(&qcount[nxt], 2);
        if (slot + 1 >= SUB_QMAX) {
          atomicOr(&counters[CNT_ERR], TERR_QUEUE);
        } else {
          qn[nxt][slot] = il; ql[nxt][slot] = lA; qh[nxt][slot] = hA;
          qn[nxt][slot + 1] = ir; ql[nxt][slot + 1] = lB; qh[nxt][slot + 1] = hB;
          for (int c = 0; c < 6; c++) { qb[nxt][slot][c] = bA[c]; qb[nxt][slot + 1][c] = bB[c]; }
        }
      }
    }
    // nodes of at most SUB_WAVE_MIN particles: one thread each
    for (int t = tid; t < nq; t += SUB_BS) {
      const int nnode = qn[cur][t];
      const int a0 = ql[cur][t], a1 = qh[cur][t];
      const int npn = a1 - a0 + 1;
      if (npn > SUB_WAVE_MIN) continue;
      // construct_node: sequential sums in node order
      int nodeisactive = 0;
      for (int k = a0; k <= a1; k++)
        if (sip[k] > 0) { nodeisactive = 1; break; }
      double hmax = 0., xc = 0., yc = 0., zc = 0., tot = 0.;
      for (int k = a0; k <= a1; k++) {
        hmax = (shh[k] > hmax) ? shh[k] : hmax;
        tot = xadd(tot, pm);
        xc = xadd(xc, xmul(fac, sx[k]));
        yc = xadd(yc, xmul(fac, sy[k]));
        zc = xadd(zc, xmul(fac, sz[k]));
      }
      double D = xmul(tot, dfac);
      double com0 = xc / D, com1 = yc / D, com2 = zc / D;
      double r2max = 0.;
      for (int k = a0; k <= a1; k++) {
        double r2 = r2_exact(xsub(sx[k], com0), xsub(sy[k], com1), xsub(sz[k], com2));
        r2max = (r2 > r2max) ? r2 : r2max;
      }
      phsph_kdnode e;
      e.xcen[0] = com0; e.xcen[1] = com1; e.xcen[2] = com2;
      e.size = xadd(sqrt(r2max), PH_EPS64);
      e.hmax = hmax;
      e.mass = tot;
      e.parent = nnode / 2;
      e.idum = 0;
      nd_lo[nnode] = lo + a0;
      nd_hi[nnode] = lo + a1;
      if (npn <= PH_MINPART) {
        e.leftchild = 0; e.rightchild = 0;
        nd[nnode] = e;
        leaf_flag[nnode] = dc.ind_timesteps ? (nodeisactive ? 1 : -1) : 1;
        maxlev = max(maxlev, level);
        minlev = min(minlev, level);
        continue;
      }
      const int il = 2 * nnode, ir = il + 1;
      e.leftchild = il; e.rightchild = ir;
      nd[nnode] = e;
      leaf_flag[nnode] = 0;
      if (level >= maxlevel_indexed || (int64_t)ir >= nslots) {
        atomicOr(&counters[CNT_ERR], TERR_NODEID);
        continue;
      }
      const double *box = qb[cur][t];
      const int iaxis = pick_axis(box);
      const double pivot = (iaxis == 0) ? com0 : (iaxis == 1) ? com1 : com2;
      double *sa = (iaxis == 0) ? sx : (iaxis == 1) ? sy : sz;
      // sort_particles_in_cell (kdtree.F90:937-1001), verbatim two-pointer
      int i = a0, j = a1;
      bool ilt = sa[i] <= pivot, jlt = sa[j] <= pivot;
      while (i < j) {
        if (ilt) {
          i = i + 1;
          ilt = sa[i] <= pivot;
        } else if (!jlt) {
          j = j - 1;
          jlt = sa[j] <= pivot;
        } else {
          double t0 = sx[i], t1 = sy[i], t2 = sz[i], t3 = shh[i];
          int32_t ti = sip[i];
          sx[i] = sx[j]; sy[i] = sy[j]; sz[i] = sz[j]; shh[i] = shh[j]; sip[i] = sip[j];
          sx[j] = t0; sy[j] = t1; sz[j] = t2; shh[j] = t3; sip[j] = ti;
          i = i + 1;
          j = j - 1;
          ilt = sa[i] <= pivot;
          jlt = sa[j] <= pivot;
        }
      }
      if (!ilt) i = i - 1;
      int nl = i - a0 + 1;
      int nr = a1 - (jlt ? j + 1 : j) + 1;
      if (nl == npn || nr == npn) nl = npn / 2;
      const int lA = a0, hA = a0 + nl - 1, lB = a0 + nl, hB = a1;
      double bA[6] = {sx[lA], sy[lA], sz[lA], sx[lA], sy[lA], sz[lA]};
      for (int k = lA + 1; k <= hA; k++) {
        bA[0] = fmin(bA[0], sx[k]); bA[1] = fmin(bA[1], sy[k]); bA[2] = fmin(bA[2], sz[k]);
        bA[3] = fmax(bA[3], sx[k]); bA[4] = fmax(bA[4], sy[k]); bA[5] = fmax(bA[5], sz[k]);
      }
      double bB[6] = {sx[lB], sy[lB], sz[lB], sx[lB], sy[lB], sz[lB]};
      for (int k = lB + 1; k <= hB; k++) {
        bB[0] = fmin(bB[0], sx[k]); bB[1] = fmin(bB[1], sy[k]); bB[2] = fmin(bB[2], sz[k]);
        bB[3] = fmax(bB[3], sx[k]); bB[4] = fmax(bB[4], sy[k]); bB[5] = fmax(bB[5], sz[k]);
      }
      int slot = atomicAdd(&qcount[nxt], 2);
      if (slot + 1 >= SUB_QMAX) {
        atomicOr(&counters[CNT_ERR], TERR_QUEUE);
        continue;
      }
      qn[nxt][slot] = il; ql[nxt][slot] = lA; qh[nxt][slot] = hA;
      qn[nxt][slot + 1] = ir; ql[nxt][slot + 1] = lB; qh[nxt][slot + 1] = hB;
      for (int c = 0; c < 6; c++) { qb[nxt][slot][c] = bA[c]; qb[nxt][slot + 1][c] = bB[c]; }
    }
    __syncthreads();
    if (tid == 0) qcount[cur] = 0;
    cur ^= 1;
    level++;
    __syncthreads();
  }
  for (int k = tid; k < n; k += SUB_BS) {
    tx[lo + k] = sx[k]; ty[lo + k] = sy[k]; tz[lo + k] = sz[k]; th[lo + k] = shh[k]; iparts[lo + k] = sip[k];
  }
  // block max/min of levels
  if (maxlev > 0 || minlev < (1 << 30)) {
    atomicMax(&counters[CNT_MAXLEVEL], maxlev);
    atomicMin(&counters[CNT_MINLEVEL], minlev);
  }
}

// leaves in node-id order (stable compaction over node slots)
__global__ void k_leaf_count(const int32_t *leaf_flag, int64_t ncells, int32_t *bcount) {
  __shared__ int shi[8];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + 1;
  int b = (i <= ncells) ? (leaf_flag[i] != 0) : 0;
  int tot;
  block_excl_scan<256>(b, shi, &tot);
  if (threadIdx.x == 0) bcount[blockIdx.x] = tot;
}
__global__ void k_leaf_fill(const int32_t *leaf_flag, int64_t ncells, const int32_t *bcount, int32_t *leaves) {
  __shared__ int shi[8];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + 1;
  int b = (i <= ncells) ? (leaf_flag[i] != 0) : 0;
  int tot;
  int pre = block_excl_scan<256>(b, shi, &tot);
  if (b) leaves[bcount[blockIdx.x] + pre] = (int32_t)i;
}
__global__ void k_pleaf(const int32_t *leaves, int nleaves, const int32_t *nd_lo, const int32_t *nd_hi,
                        int32_t *pleaf) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nleaves) return;
  int c = leaves[l];
  for (int p = nd_lo[c]; p <= nd_hi[c]; p++) pleaf[p] = l;
}

// ------------------------------------------------------------------------
int tree_build(phsph_ctx *c, double *d_xyzh, int ld, int np, const int8_t *d_iphase) {
  hipStream_t s = c->stream;
  c->tree_valid = false;
  c->np = np;
  const int64_t maxp = (c->p.maxp > np) ? c->p.maxp : np;
  const int64_t ncellsmax = 2 * maxp;
  c->maxlevel_indexed = (int)(log((double)(ncellsmax + 1)) / log(2.)) - 1;
  const int64_t nidx = ((int64_t)1 << (c->maxlevel_indexed + 1));  // ids 1 .. 2^(L+1)-1
  c->nslots = nidx;
  const unsigned nb = nblocks(np, 256);
  PH_HIPCHECK(c, c->counters.ensure(CNT_N));
  PH_HIPCHECK(c, c->nd.ensure(nidx));
  PH_HIPCHECK(c, c->nd_lo.ensure(nidx));
  PH_HIPCHECK(c, c->nd_hi.ensure(nidx));
  PH_HIPCHECK(c, c->leaf_flag.ensure(nidx));
  PH_HIPCHECK(c, c->tx.ensure(np)); PH_HIPCHECK(c, c->ty.ensure(np));
  PH_HIPCHECK(c, c->tz.ensure(np)); PH_HIPCHECK(c, c->th.ensure(np));
  PH_HIPCHECK(c, c->iparts.ensure(np));
  PH_HIPCHECK(c, c->tmpi.ensure(np)); PH_HIPCHECK(c, c->tmpi2.ensure(np));
  PH_HIPCHECK(c, c->dred.ensure((size_t)nb * 6 + 16));
  PH_HIPCHECK(c, c->frontier_a.ensure((size_t)np + nb + 2));
  PH_HIPCHECK(c, c->frontier_b.ensure((size_t)np + nb + 2));
  PH_HIPCHECK(c, c->fbox_a.ensure((size_t)(np + 2) * 6));
  PH_HIPCHECK(c, c->fbox_b.ensure((size_t)(np + 2) * 6));
  // large nodes are disjoint sets of more than SUB_MAX particles: at most
  // np/(SUB_MAX+1) per level, twice that in a frontier, and their chunks at
  // most np/BN_CH + one per node
  const int bn_cap = 2 * (np / (SUB_MAX + 1)) + 4;
  const int maxch = np / BN_CH + np / (SUB_MAX + 1) + 2;
  PH_HIPCHECK(c, c->bnode.ensure(bn_cap));
  PH_HIPCHECK(c, c->ch_node.ensure(maxch));
  PH_HIPCHECK(c, c->ch_i.ensure((size_t)maxch * 4));
  PH_HIPCHECK(c, c->ch_off.ensure((size_t)maxch * 2));
  PH_HIPCHECK(c, c->ch_red.ensure((size_t)maxch * BN_RED));
  // under self-gravity the large nodes' second moments are formed here
  double *quads = nullptr;
  if (c->p.gravity) {
    PH_HIPCHECK(c, c->quads.ensure((size_t)c->nslots * 6));
    quads = c->quads.p;
  }
  int32_t *live = c->tmpi.p;
  int32_t *bcount = c->frontier_b.p;  // block counts of the compaction

  // the reference seeds the bounds with particle 1 before wrapping it
  double x0[3];
  PH_HIPCHECK(c, hipMemcpyAsync(x0, d_xyzh, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
  int32_t hc[CNT_N] = {0, 0, 0, 0, 1 << 30, 0, 0, 0, 0, 0, 0};
  PH_HIPCHECK(c, dev_set(c, c->counters.p, hc, sizeof(hc)));
  PH_HIPCHECK(c, hipStreamSynchronize(s));
  {
    LaunchTimer lt(c, "tree_root");
    hipLaunchKernelGGL(k_root_wrap, dim3(nb), dim3(256), 0, s, c->dc, d_xyzh, ld, np, live, c->dred.p,
                       c->counters.p);
    hipLaunchKernelGGL(k_count_blocks, dim3(nb), dim3(256), 0, s, live, np, bcount);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, bcount, (int)nb, c->counters.p + CNT_LIVE);
    hipLaunchKernelGGL(k_fill_cache, dim3(nb), dim3(256), 0, s, c->dc, d_xyzh, ld, np, live, bcount, d_iphase,
                       c->tx.p, c->ty.p, c->tz.p, c->th.p, c->iparts.p);
  }
  PH_HIPCHECK(c, hipGetLastError());
  int32_t cnt[CNT_N];
  PH_HIPCHECK(c, hipMemcpyAsync(cnt, c->counters.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
  PH_HIPCHECK(c, hipStreamSynchronize(s));
  if (cnt[CNT_NAN] > 0) PH_RET(c, PHSPH_ERR_NAN, "maketree: NaN in particle position");
  c->nlive = cnt[CNT_LIVE];
  if (c->nlive == 0) PH_RET(c, PHSPH_ERR_NODATA, "maketree: no particles or all particles dead/accreted");

  double *rootbox = c->dred.p + (size_t)nb * 6;
  hipLaunchKernelGGL(k_root_bounds, dim3(1), dim3(256), 0, s, c->dred.p, (int)nb, x0[0], x0[1], x0[2], rootbox);
  hipLaunchKernelGGL(k_fill_i32, dim3(nblocks(c->nslots, 256)), dim3(256), 0, s, c->leaf_flag.p, c->nslots, 0);
  int32_t root_init[2] = {0, c->nlive - 1};
  PH_HIPCHECK(c, dev_set(c, c->nd_lo.p + 1, &root_init[0], sizeof(int32_t)));
  PH_HIPCHECK(c, dev_set(c, c->nd_hi.p + 1, &root_init[1], sizeof(int32_t)));
  int32_t one = 1;
  PH_HIPCHECK(c, dev_set(c, c->frontier_a.p, &one, sizeof(int32_t)));
  PH_HIPCHECK(c, hipMemcpyAsync(c->fbox_a.p, rootbox, 6 * sizeof(double), hipMemcpyDeviceToDevice, s));
  const bool dbg = getenv("PHSPH_DEBUG") != nullptr;
  int nfront = 1;
  int nbig = (c->nlive > SUB_MAX) ? 1 : 0;  // large nodes in the frontier
  int level = 0;
  int32_t *fa = c->frontier_a.p, *fb = c->frontier_b.p;
  double *ba = c->fbox_a.p, *bb = c->fbox_b.p;
  while (nfront > 0) {
    int32_t zero = 0;
    PH_HIPCHECK(c, dev_set(c, c->counters.p + CNT_NEXT, &zero, sizeof(int32_t)));
    static_assert(CNT_SEQMAX == CNT_NBIG + 1, "adjacent level counters");
    const int32_t zero2[2] = {0, 0};
    PH_HIPCHECK(c, dev_set(c, c->counters.p + CNT_NBIG, zero2, sizeof(zero2)));
    if (nbig < nfront) {
      LaunchTimer lt(c, "tree_subtrees");
      hipLaunchKernelGGL(k_subtrees, dim3(nfront), dim3(SUB_BS), 0, s, c->dc, fa, ba, nfront, c->tx.p, c->ty.p,
                         c->tz.p, c->th.p, c->iparts.p, c->nd_lo.p, c->nd_hi.p, c->nd.p, c->leaf_flag.p, level,
                         c->maxlevel_indexed, c->nslots, c->counters.p);
    }
    if (nbig > 0) {
      if (nfront > bn_cap) PH_RET(c, PHSPH_ERR_STATE, "maketree: large-node frontier beyond its capacity");
      LaunchTimer lt(c, "tree_bignodes");
      // chunks <= sum over large nodes of ceil(n/BN_CH) <= nlive/BN_CH + nbig
      const int gch = std::min(maxch, (c->nlive + BN_CH - 1) / BN_CH + nbig);
      BigNode *bn = c->bnode.p;
      int32_t *cnt_d = c->counters.p;
      const int32_t *chn = c->ch_node.p;
      hipLaunchKernelGGL(k_bn_plan, dim3(1), dim3(1024), 0, s, fa, ba, nfront, c->nd_lo.p, c->nd_hi.p, bn,
                         c->ch_node.p, maxch, cnt_d);
      hipLaunchKernelGGL(k_bn_sums, dim3(gch), dim3(BN_BS), 0, s, c->dc, bn, chn, maxch, cnt_d, c->tx.p, c->ty.p,
                         c->tz.p, c->th.p, c->ch_red.p);
      hipLaunchKernelGGL(k_bn_centre, dim3(nfront), dim3(64), 0, s, c->dc, bn, maxch, cnt_d, c->ch_red.p);
      hipLaunchKernelGGL(k_bn_amb, dim3(gch), dim3(BN_BS), 0, s, bn, chn, maxch, cnt_d, c->tx.p, c->ty.p, c->tz.p,
                         c->ch_i.p);
      hipLaunchKernelGGL(k_bn_seq, dim3(nfront), dim3(BN_SEQ_BS), 0, s, c->dc, bn, maxch, cnt_d, c->ch_i.p,
                         c->tx.p, c->ty.p, c->tz.p);
      hipLaunchKernelGGL(k_bn_size, dim3(gch), dim3(BN_BS), 0, s, c->dc, bn, chn, maxch, cnt_d, c->tx.p, c->ty.p,
                         c->tz.p, c->ch_red.p, c->ch_i.p, quads ? 1 : 0);
      hipLaunchKernelGGL(k_bn_node, dim3(nfront), dim3(64), 0, s, bn, maxch, cnt_d, c->ch_red.p, c->ch_i.p,
                         c->nd.p, c->leaf_flag.p, level, c->maxlevel_indexed, c->nslots, quads);
      hipLaunchKernelGGL(k_bn_pcount, dim3(gch), dim3(BN_BS), 0, s, bn, chn, maxch, cnt_d, c->tx.p, c->ty.p,
                         c->tz.p, c->ch_i.p);
      hipLaunchKernelGGL(k_bn_pscan, dim3(nfront), dim3(64), 0, s, bn, maxch, cnt_d, c->ch_i.p, c->ch_off.p);
      hipLaunchKernelGGL(k_bn_pscatter, dim3(gch), dim3(BN_BS), 0, s, bn, chn, maxch, cnt_d, c->tx.p, c->ty.p,
                         c->tz.p, c->ch_off.p, c->tmpi.p, c->tmpi2.p);
      hipLaunchKernelGGL(k_bn_swap, dim3(gch), dim3(BN_BS), 0, s, bn, chn, maxch, cnt_d, c->tmpi.p, c->tmpi2.p,
                         c->tx.p, c->ty.p, c->tz.p, c->th.p, c->iparts.p);
      hipLaunchKernelGGL(k_bn_box, dim3(gch), dim3(BN_BS), 0, s, bn, chn, maxch, cnt_d, c->tx.p, c->ty.p, c->tz.p,
                         c->ch_red.p);
      hipLaunchKernelGGL(k_bn_finish, dim3(nfront), dim3(64), 0, s, bn, maxch, cnt_d, c->ch_red.p, c->nd_lo.p,
                         c->nd_hi.p, fb, bb);
    }
    PH_HIPCHECK(c, hipGetLastError());
    PH_HIPCHECK(c, hipMemcpyAsync(cnt, c->counters.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    if (dbg)
      fprintf(stderr, "maketree level %d: frontier %d, large %d, sequential sums so far %d, longest this level %d\n",
              level, nfront, nbig, cnt[CNT_FALLBACK], cnt[CNT_SEQMAX]);
    if (cnt[CNT_ERR]) break;
    nfront = cnt[CNT_NEXT];
    nbig = cnt[CNT_NBIG];
    std::swap(fa, fb);
    std::swap(ba, bb);
    level++;
    if (level > 64) PH_RET(c, PHSPH_ERR_CAPACITY, "maketree: more than 64 levels");
  }
  if (cnt[CNT_ERR] & TERR_NODEID)
    PH_RET(c, PHSPH_ERR_CAPACITY,
           "maketree: tree deeper than maxlevel_indexed; pass a larger maxp (ncellsmax = 2*maxp)");
  if (cnt[CNT_ERR]) PH_RET(c, PHSPH_ERR_STATE, "maketree: internal queue overflow");
  c->maxlevel = cnt[CNT_MAXLEVEL];
  c->minlevel = cnt[CNT_MINLEVEL];
  c->ncells = (c->maxlevel < c->maxlevel_indexed) ? (((int64_t)1 << (c->maxlevel + 1)) - 1)
                                                  : (((int64_t)1 << (c->maxlevel_indexed + 1)) - 1);
  c->tree_fallbacks = cnt[CNT_FALLBACK];
  // leaves in node-id order
  const unsigned nbl = nblocks(c->ncells, 256);
  PH_HIPCHECK(c, c->tmpi2.ensure((size_t)nbl + 1));
  PH_HIPCHECK(c, c->leaves.ensure(c->nlive + 1));
  PH_HIPCHECK(c, c->pleaf.ensure(c->nlive + 1));
  {
    LaunchTimer lt(c, "tree_leaves");
    hipLaunchKernelGGL(k_leaf_count, dim3(nbl), dim3(256), 0, s, c->leaf_flag.p, c->ncells, c->tmpi2.p);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, c->tmpi2.p, (int)nbl,
                       c->counters.p + CNT_NLEAF);
    hipLaunchKernelGGL(k_leaf_fill, dim3(nbl), dim3(256), 0, s, c->leaf_flag.p, c->ncells, c->tmpi2.p,
                       c->leaves.p);
  }
  PH_HIPCHECK(c, hipMemcpyAsync(cnt, c->counters.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
  PH_HIPCHECK(c, hipStreamSynchronize(s));
  c->nleaves = cnt[CNT_NLEAF];
  hipLaunchKernelGGL(k_pleaf, dim3(nblocks(c->nleaves, 256)), dim3(256), 0, s, c->leaves.p, c->nleaves,
                     c->nd_lo.p, c->nd_hi.p, c->pleaf.p);
  PH_HIPCHECK(c, hipGetLastError());
  c->tree_valid = true;
  // statistics of a density call belong to the tree it ran on
  c->nact_dens = 0;
  c->trial_stale = true;
  c->nrhocalc = -1;
  c->nactualtot = -1;
  if (c->p.gravity) return grav_tree(c);
  return PHSPH_OK;
}
