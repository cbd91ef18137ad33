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
#include "ctx.h"

#define SUB_MAX 1024  // particles per LDS subtree
#define SUB_BS 256
#define SUB_QMAX 192  // >= 2*floor(SUB_MAX/(MINPART+1)) nodes per level
#define BIG_BS 512

enum { CNT_NEXT = 0, CNT_ERR = 1, CNT_FALLBACK = 2, CNT_MAXLEVEL = 3, CNT_MINLEVEL = 4, CNT_NAN = 5,
       CNT_NLEAF = 6, CNT_LIVE = 7, CNT_N = 8 };
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
// exclusive scan of block counts (single block, sequential chunks)
__global__ void k_scan_blocks(int32_t *bcount, int nb, int32_t *total) {
  __shared__ int carry;
  __shared__ int tmp[256];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 256) {
    int i = base + threadIdx.x;
    int v = (i < nb) ? bcount[i] : 0;
    int tot;
    // counts exceed 1: plain Hillis-Steele scan
    tmp[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
      int t = (threadIdx.x >= o) ? tmp[threadIdx.x - o] : 0;
      __syncthreads();
      tmp[threadIdx.x] += t;
      __syncthreads();
    }
    int incl = tmp[threadIdx.x];
    tot = tmp[255];
    if (i < nb) bcount[i] = carry + incl - v;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
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
// large nodes: one workgroup of BIG_BS threads per node.  The streaming
// passes take BIG_U elements per thread per iteration with independent
// loads, so that a node of a million particles keeps enough loads in
// flight on its one CU.
// ------------------------------------------------------------------------
#define BIG_U 8
#define BIG_CHUNK 2048

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

__device__ __forceinline__ const double *axis_arr(int a, const double *tx, const double *ty, const double *tz) {
  return (a == 0) ? tx : (a == 1) ? ty : tz;
}

__global__ void __launch_bounds__(BIG_BS) k_big_nodes(DevConst dc, const int32_t *frontier, const double *fbox,
                                                      int nfront, double *tx, double *ty, double *tz, double *th,
                                                      int32_t *iparts, int32_t *nd_lo, int32_t *nd_hi,
                                                      phsph_kdnode *nd, int32_t *leaf_flag, int32_t *tmpA,
                                                      int32_t *tmpB, int level, int maxlevel_indexed,
                                                      int64_t nslots, int32_t *next_front, double *next_box,
                                                      int32_t *counters) {
  __shared__ double sh[BIG_BS / 64];
  __shared__ int shi[BIG_BS / 64];
  __shared__ double chunk[4][BIG_CHUNK];
  const int f = blockIdx.x;
  const int nnode = frontier[f];
  const int lo = nd_lo[nnode], hi = nd_hi[nnode];
  const int n = hi - lo + 1;
  if (n <= SUB_MAX) return;  // handled by the subtree kernel
  const int tid = threadIdx.x;
  const double pm = dc.pmass;
  const double dfac = (pm > 0.) ? 1. / pm : 1.;
  const double fac = xmul(pm, dfac);
  double box[6];
  for (int c = 0; c < 6; c++) box[c] = fbox[(size_t)f * 6 + c];
  constexpr int STEP = BIG_BS * BIG_U;

  // parallel sums + bound
  double s[3] = {0., 0., 0.}, a[3] = {0., 0., 0.}, hmax = 0.;
  for (int base = lo; base <= hi; base += STEP) {
#pragma unroll
    for (int u = 0; u < BIG_U; u++) {
      const int p = base + u * BIG_BS + tid;
      if (p <= hi) {
        const double vx = xmul(fac, tx[p]), vy = xmul(fac, ty[p]), vz = xmul(fac, tz[p]);
        s[0] += vx; s[1] += vy; s[2] += vz;
        a[0] += fabs(vx); a[1] += fabs(vy); a[2] += fabs(vz);
        hmax = fmax(hmax, th[p]);
      }
    }
  }
  for (int c = 0; c < 3; c++) {
    s[c] = block_sum<BIG_BS>(s[c], sh);
    a[c] = block_sum<BIG_BS>(a[c], sh);
  }
  hmax = block_max<BIG_BS>(hmax, sh);
  double tot = xmul((double)n, pm);
  double D = xmul(tot, dfac);
  double com[3];
  for (int c = 0; c < 3; c++) com[c] = s[c] / D;
  const double uu = 1.1102230246251565e-16;
  const double nu = (double)(n - 1) * uu;
  const double gam = nu / (1. - nu);
  const int iaxis = pick_axis(box);
  const double *tax = axis_arr(iaxis, tx, ty, tz);
  const double a_ax = (iaxis == 0) ? a[0] : (iaxis == 1) ? a[1] : a[2];
  const double c_ax = (iaxis == 0) ? com[0] : (iaxis == 1) ? com[1] : com[2];
  double delta = 2. * (2. * gam * a_ax / D + (2. * gam + 4. * uu) * fabs(c_ax)) + 1e-300;
  // any particle within delta of the plane?
  int amb = 0;
  for (int base = lo; base <= hi; base += STEP) {
#pragma unroll
    for (int u = 0; u < BIG_U; u++) {
      const int p = base + u * BIG_BS + tid;
      if (p <= hi) amb |= (fabs(tax[p] - c_ax) <= delta);
    }
  }
  int namb;
  block_excl_scan<BIG_BS>(amb, shi, &namb);
  if (namb > 0) {
    // exact sequential sums (kdtree.F90:655-665): the products fac*x are
    // independent and are formed by all threads into LDS; the four running
    // sums (x, y, z, mass) are four dependent chains, run by lanes 0-3 of
    // the first wavefront in lockstep, one addition per particle each
    __shared__ double seq[4];
    double acc = 0.;
    for (int base = lo; base <= hi; base += BIG_CHUNK) {
      const int cnt = min(BIG_CHUNK, hi - base + 1);
      __syncthreads();
      for (int k = tid; k < cnt; k += BIG_BS) {
        chunk[0][k] = xmul(fac, tx[base + k]);
        chunk[1][k] = xmul(fac, ty[base + k]);
        chunk[2][k] = xmul(fac, tz[base + k]);
        chunk[3][k] = pm;
      }
      __syncthreads();
      if (tid < 4) {
        // row tid: x, y, z products or the masses; LDS reads issued ahead
        // of the dependent additions, no branch per element
        const double *src = chunk[tid];
        int k = 0;
        for (; k + 8 <= cnt; k += 8) {
          double v[8];
#pragma unroll
          for (int u = 0; u < 8; u++) v[u] = src[k + u];
#pragma unroll
          for (int u = 0; u < 8; u++) acc = xadd(acc, v[u]);
        }
        for (; k < cnt; k++) acc = xadd(acc, src[k]);
      }
    }
    if (tid < 4) seq[tid] = acc;
    __syncthreads();
    tot = seq[3];
    D = xmul(tot, dfac);
    for (int c = 0; c < 3; c++) com[c] = seq[c] / D;
    if (tid == 0) atomicAdd(&counters[CNT_FALLBACK], 1);
  }
  // size: sqrt(max r^2) + epsilon (kdtree.F90:697-782), and the count of
  // particles on the left of the pivot
  const double pivot = (iaxis == 0) ? com[0] : (iaxis == 1) ? com[1] : com[2];
  double r2max = 0.;
  int cl = 0;
  for (int base = lo; base <= hi; base += STEP) {
#pragma unroll
    for (int u = 0; u < BIG_U; u++) {
      const int p = base + u * BIG_BS + tid;
      if (p <= hi) {
        const double r2 = r2_exact(xsub(tx[p], com[0]), xsub(ty[p], com[1]), xsub(tz[p], com[2]));
        r2max = fmax(r2max, r2);
        cl += (tax[p] <= pivot);
      }
    }
  }
  r2max = block_max<BIG_BS>(r2max, sh);
  const int il = 2 * nnode, ir = il + 1;
  if (tid == 0) {
    phsph_kdnode e;
    e.xcen[0] = com[0]; e.xcen[1] = com[1]; e.xcen[2] = com[2];
    e.size = xadd(sqrt(r2max), PH_EPS64);
    e.hmax = hmax;
    e.mass = tot;
    e.leftchild = il;
    e.rightchild = ir;
    e.parent = nnode / 2;
    e.idum = 0;
    nd[nnode] = e;
    leaf_flag[nnode] = 0;
    if (level >= maxlevel_indexed || (int64_t)ir >= nslots) atomicOr(&counters[CNT_ERR], TERR_NODEID);
  }
  if (level >= maxlevel_indexed || (int64_t)ir >= nslots) return;
  // partition (sort_particles_in_cell): the k-th element > pivot in the
  // left part swaps with the k-th element <= pivot counted from the right
  int nl;
  block_excl_scan_int<BIG_BS>(cl, shi, &nl);
  if (nl > 0 && nl < n) {
    const int lend = lo + nl - 1;  // last position of the left part
    int carry = 0;
    for (int base = lo; base <= lend; base += STEP) {
      // thread tid takes BIG_U consecutive positions, so ranks follow position
      const int p0 = base + tid * BIG_U;
      int b[BIG_U], c = 0;
#pragma unroll
      for (int u = 0; u < BIG_U; u++) {
        const int p = p0 + u;
        b[u] = (p <= lend) ? !(tax[p] <= pivot) : 0;
        c += b[u];
      }
      int t;
      int r = block_excl_scan_int<BIG_BS>(c, shi, &t);
#pragma unroll
      for (int u = 0; u < BIG_U; u++)
        if (b[u]) tmpA[lo + carry + r++] = p0 + u;
      carry += t;
    }
    const int nbad = carry;
    carry = 0;
    for (int base = 0; base < hi - lend; base += STEP) {
      const int q0 = base + tid * BIG_U;  // distance from the right end
      int b[BIG_U], c = 0;
#pragma unroll
      for (int u = 0; u < BIG_U; u++) {
        const int p = hi - (q0 + u);
        b[u] = (p > lend) ? (tax[p] <= pivot) : 0;
        c += b[u];
      }
      int t;
      int r = block_excl_scan_int<BIG_BS>(c, shi, &t);
#pragma unroll
      for (int u = 0; u < BIG_U; u++)
        if (b[u]) tmpB[lo + carry + r++] = hi - (q0 + u);
      carry += t;
    }
    if (carry != nbad && tid == 0) atomicOr(&counters[CNT_ERR], TERR_COUNT);
    __syncthreads();
    for (int k = tid; k < nbad; k += BIG_BS) {
      const int pa = tmpA[lo + k], pb = tmpB[lo + k];
      const double t0 = tx[pa], t1 = ty[pa], t2 = tz[pa], t3 = th[pa];
      const int32_t ti = iparts[pa];
      tx[pa] = tx[pb]; ty[pa] = ty[pb]; tz[pa] = tz[pb]; th[pa] = th[pb]; iparts[pa] = iparts[pb];
      tx[pb] = t0; ty[pb] = t1; tz[pb] = t2; th[pb] = t3; iparts[pb] = ti;
    }
    __syncthreads();
  } else {
    nl = n / 2;  // all on one side: split in half in place (kdtree.F90:856-865)
  }
  const int lA = lo, hA = lo + nl - 1, lB = lo + nl, hB = hi;
  double bmn[2][3], bmx[2][3];
  for (int side = 0; side < 2; side++) {
    const int l0 = side ? lB : lA, h0 = side ? hB : hA;
    double m0[3] = {PH_HUGE64, PH_HUGE64, PH_HUGE64}, m1[3] = {-PH_HUGE64, -PH_HUGE64, -PH_HUGE64};
    for (int base = l0; base <= h0; base += STEP) {
#pragma unroll
      for (int u = 0; u < BIG_U; u++) {
        const int p = base + u * BIG_BS + tid;
        if (p <= h0) {
          const double x = tx[p], y = ty[p], z = tz[p];
          m0[0] = fmin(m0[0], x); m0[1] = fmin(m0[1], y); m0[2] = fmin(m0[2], z);
          m1[0] = fmax(m1[0], x); m1[1] = fmax(m1[1], y); m1[2] = fmax(m1[2], z);
        }
      }
    }
    for (int c = 0; c < 3; c++) {
      bmn[side][c] = block_min<BIG_BS>(m0[c], sh);
      bmx[side][c] = block_max<BIG_BS>(m1[c], sh);
    }
  }
  if (tid == 0) {
    nd_lo[il] = lA; nd_hi[il] = hA; nd_lo[ir] = lB; nd_hi[ir] = hB;
    int idx = atomicAdd(&counters[CNT_NEXT], 2);
    next_front[idx] = il;
    next_front[idx + 1] = ir;
    for (int c = 0; c < 3; c++) {
      next_box[(size_t)idx * 6 + c] = bmn[0][c];
      next_box[(size_t)idx * 6 + 3 + c] = bmx[0][c];
      next_box[(size_t)(idx + 1) * 6 + c] = bmn[1][c];
      next_box[(size_t)(idx + 1) * 6 + 3 + c] = bmx[1][c];
    }
  }
}

// ------------------------------------------------------------------------
// small nodes: one workgroup builds the whole subtree in LDS, one node per
// thread per level, with the reference's sequential arithmetic
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(SUB_BS) k_subtrees(DevConst dc, const int32_t *frontier, const double *fbox,
                                                     int nfront, double *tx, double *ty, double *tz, double *th,
                                                     int32_t *iparts, int32_t *nd_lo, int32_t *nd_hi,
                                                     phsph_kdnode *nd, int32_t *leaf_flag, int level0,
                                                     int maxlevel_indexed, int64_t nslots, int32_t *counters) {
  __shared__ double sx[SUB_MAX], sy[SUB_MAX], sz[SUB_MAX], shh[SUB_MAX];
  __shared__ int32_t sip[SUB_MAX];
  __shared__ int32_t qn[2][SUB_QMAX], ql[2][SUB_QMAX], qh[2][SUB_QMAX];
  __shared__ double qb[2][SUB_QMAX][6];
  __shared__ int qcount[2];
  const int f = blockIdx.x;
  const int root = frontier[f];
  const int lo = nd_lo[root], hi = nd_hi[root];
  const int n = hi - lo + 1;
  if (n > SUB_MAX) return;
  const int tid = threadIdx.x;
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
  int cur = 0;
  int level = level0;
  int maxlev = 0, minlev = 1 << 30;
  while (qcount[cur] > 0) {
    const int nq = min(qcount[cur], SUB_QMAX);
    const int nxt = cur ^ 1;
    for (int t = tid; t < nq; t += SUB_BS) {
      const int nnode = qn[cur][t];
      const int a0 = ql[cur][t], a1 = qh[cur][t];
      const int npn = a1 - a0 + 1;
      // construct_node: sequential sums in node order
      int nodeisactive = 0;
      for (int k = a0; k <= a1; k++)
        if (sip[k] > 0) { nodeisactive = 1; break; }
      double hmax = 0., xc = 0., yc = 0., zc = 0., tot = 0.;
      for (int k = a0; k <= a1; k++) {
        double fac = xmul(pm, dfac);
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
  int32_t *live = c->tmpi.p;
  int32_t *bcount = c->frontier_b.p;  // block counts of the compaction

  // the reference seeds the bounds with particle 1 before wrapping it
  double x0[3];
  PH_HIPCHECK(c, hipMemcpyAsync(x0, d_xyzh, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
  int32_t hc[CNT_N] = {0, 0, 0, 0, 1 << 30, 0, 0, 0};
  PH_HIPCHECK(c, hipMemcpyAsync(c->counters.p, hc, sizeof(hc), hipMemcpyHostToDevice, s));
  PH_HIPCHECK(c, hipStreamSynchronize(s));
  {
    LaunchTimer lt(c, "tree_root");
    hipLaunchKernelGGL(k_root_wrap, dim3(nb), dim3(256), 0, s, c->dc, d_xyzh, ld, np, live, c->dred.p,
                       c->counters.p);
    hipLaunchKernelGGL(k_count_blocks, dim3(nb), dim3(256), 0, s, live, np, bcount);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(256), 0, s, bcount, (int)nb, c->counters.p + CNT_LIVE);
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
  PH_HIPCHECK(c, hipMemcpyAsync(c->nd_lo.p + 1, &root_init[0], sizeof(int32_t), hipMemcpyHostToDevice, s));
  PH_HIPCHECK(c, hipMemcpyAsync(c->nd_hi.p + 1, &root_init[1], sizeof(int32_t), hipMemcpyHostToDevice, s));
  int32_t one = 1;
  PH_HIPCHECK(c, hipMemcpyAsync(c->frontier_a.p, &one, sizeof(int32_t), hipMemcpyHostToDevice, s));
  PH_HIPCHECK(c, hipMemcpyAsync(c->fbox_a.p, rootbox, 6 * sizeof(double), hipMemcpyDeviceToDevice, s));
  int nfront = 1;
  int level = 0;
  int32_t *fa = c->frontier_a.p, *fb = c->frontier_b.p;
  double *ba = c->fbox_a.p, *bb = c->fbox_b.p;
  while (nfront > 0) {
    int32_t zero = 0;
    PH_HIPCHECK(c, hipMemcpyAsync(c->counters.p + CNT_NEXT, &zero, sizeof(int32_t), hipMemcpyHostToDevice, s));
    {
      LaunchTimer lt(c, "tree_subtrees");
      hipLaunchKernelGGL(k_subtrees, dim3(nfront), dim3(SUB_BS), 0, s, c->dc, fa, ba, nfront, c->tx.p, c->ty.p,
                         c->tz.p, c->th.p, c->iparts.p, c->nd_lo.p, c->nd_hi.p, c->nd.p, c->leaf_flag.p, level,
                         c->maxlevel_indexed, c->nslots, c->counters.p);
    }
    {
      LaunchTimer lt(c, "tree_bignodes");
      hipLaunchKernelGGL(k_big_nodes, dim3(nfront), dim3(BIG_BS), 0, s, c->dc, fa, ba, nfront, c->tx.p, c->ty.p,
                         c->tz.p, c->th.p, c->iparts.p, c->nd_lo.p, c->nd_hi.p, c->nd.p, c->leaf_flag.p,
                         c->tmpi.p, c->tmpi2.p, level, c->maxlevel_indexed, c->nslots, fb, bb, c->counters.p);
    }
    PH_HIPCHECK(c, hipGetLastError());
    PH_HIPCHECK(c, hipMemcpyAsync(cnt, c->counters.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
    PH_HIPCHECK(c, hipStreamSynchronize(s));
    if (cnt[CNT_ERR]) break;
    nfront = cnt[CNT_NEXT];
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
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(256), 0, s, c->tmpi2.p, (int)nbl,
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
