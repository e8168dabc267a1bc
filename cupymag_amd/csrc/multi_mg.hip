// Batched multigrid-preconditioned CG: many independent systems, each with
// its own multigrid hierarchy (any mesh sizes, any depths), advanced
// together with one launch per phase for the whole batch.
//
// As in multi.hip, every system owns contiguous workgroup ranges sized from
// its own rows only (its outer operator for the CG vectors, each level's
// operator for that level's V-cycle work), so its arithmetic does not depend
// on the batch.  The V-cycle runs level by level: at depth d every system
// that still has a coarser level smooths, forms its residual and restricts
// it; systems whose coarsest level is d + 1 solve it densely right after;
// on the way up every system prolongs and post-smooths at its own depths.
// The kernel count per iteration is that of one hierarchy of the deepest
// member, whatever the batch size.
#include <algorithm>
#include <cstring>
#include <vector>

#include "device.h"
#include "mg.h"
#include "mg_dev.h"
#include "multi_mg.h"
#include "sell.h"

namespace cpm {

struct OuterSeg {
  const int64_t* sptr;
  const int32_t* col;
  const double* v64;      // every slot (the FP64 fallback of the FP32 products)
  const double* v64h;     // the other FP64 products' values (sell_v64)
  const double* dg64;
  const float* v32;       // FP32 form (reliable-update systems)
  const double* dgc;
  GridArg ga;             // for v32 and v64
  GridArg ga64;           // for v64h
  int32_t nb, stencil;
  int64_t n;
  const double* b;
  double* x;
  double* r;
  void *p, *q;            // FP32 for the reliable-update systems (delta > 0), else FP64
  const void* z;
  CgState* st;
  int32_t blk0, nblk;    // vector phases: grid_for(n) workgroups
  int32_t sblk0, snblk;  // SpMV: the operator's SELL grid
  int32_t use_x0, maxiter;
  double tol;
  double delta;          // > 0: FP32 products with FP64 residual replacements (mg.hip)
  int32_t floating;      // floating hierarchy (mg.hip): F r into the V-cycle, F^T on its output
};

struct SmoothSeg {
  const int64_t* sptr;
  const int32_t* col;
  const void* val;
  const void* dg;     // diagonal blocks: sell_diag_t of the value type
  GridArg ga;
  int32_t nb, stencil;
  void *x, *r, *b, *d0, *d1;
  const void* dinv;
  const double* b64;  // depth 0: the system's FP64 outer residual
  double c0, c1[8], c2[8];
  Dims3 fd, cd;
  double vscale;      // FP16 values: their scale
  int32_t anchor;     // node 0 left out of the transfers
  void *bc, *xc;      // next level's b and x
  void *rc, *dc;      // next level's r and d0 (fused first step)
  const void* dinvc;  // next level's D^-1
  double c0c;         // next level's first Chebyshev coefficient
  int32_t fuse;       // the next level has a first step (not a dense coarsest level): fused with the restriction
  int32_t nsteps;     // Chebyshev steps of this level's smoothing (the coarsest of a smoothed hierarchy: its solve)
  const double* fpart;  // depth 0 of a floating hierarchy: partials of the sums of b64
  int64_t fstride;
  int32_t fn;
  const CgState* st;
  int32_t blk0, nblk;
};

struct DenseSeg {
  int64_t n;
  const void* cinv;
  const void* b;
  void* x;
  const CgState* st;
  int32_t blk0, nblk;
};

// ---------------------------------------------------------------- kernels
// K columns per system (interleaved as in mg.hip); a system's V-cycle work
// runs while any of its columns is live, its CG updates per live column.

template <int BS, int K, class T, bool ZERO, bool FROM64>
__global__ void __launch_bounds__(kBlock) kbm_first(const SmoothSeg* __restrict__ seg, const int32_t* __restrict__ map) {
  const SmoothSeg& d = seg[map[blockIdx.x]];
  if (all_done<K>(d.st)) return;
  mg_first_rows<BS, K, T, ZERO, FROM64>((int64_t)blockIdx.x - d.blk0, d.nblk, d.nb, d.b64, (T*)d.b, (T*)d.r,
                                        (T*)d.d0, (T*)d.x, (const T*)d.dinv, d.c0, FROM64 ? d.fpart : nullptr,
                                        d.fstride, d.fn);
}

// MODE 0 / 4: Chebyshev step k (d_in = d[(k-1)&1], d_out = d[k&1]; 4: also
// x += d_in); MODE 1: the residual after step `k` (d_in = d[k&1]); MODE 2:
// r = b - A x; MODE 3: r = b - A x and d[0] = c0 D^-1 r (mg_dev.h).
template <int BS, int K, class T, int MODE, class VT>
__global__ void __launch_bounds__(kBlock) kbm_sweep(const SmoothSeg* __restrict__ seg, const int32_t* __restrict__ map,
                                                    int k) {
  const SmoothSeg& d = seg[map[blockIdx.x]];
  if (k >= d.nsteps || all_done<K>(d.st)) return;
  T* dv[2] = {(T*)d.d0, (T*)d.d1};
  const T* din = (MODE == 0 || MODE == 4) ? dv[(k - 1) & 1] : dv[k & 1];
  T* dout = dv[k & 1];
  const double c1 = (MODE == 0 || MODE == 4) ? d.c1[k] : 0.0;
  const double c2 = (MODE == 0 || MODE == 4) ? d.c2[k] : (MODE == 3) ? d.c0 : 0.0;
  const int64_t lb = xcd_perm((int64_t)blockIdx.x - d.blk0, d.nblk, d.ga.xg);
  if (d.stencil)
    mg_sweep_rows<BS, K, T, MODE, true, VT>(lb, d.nblk, d.nb, d.sptr, d.col, (const VT*)d.val,
                                            (const sell_diag_t<VT>*)d.dg, d.ga, din, dout,
                                            (T*)d.x, (T*)d.r, (const T*)d.b, (const T*)d.dinv, c1, c2, d.vscale);
  else
    mg_sweep_rows<BS, K, T, MODE, false, VT>(lb, d.nblk, d.nb, d.sptr, d.col, (const VT*)d.val,
                                             (const sell_diag_t<VT>*)d.dg, d.ga, din, dout,
                                             (T*)d.x, (T*)d.r, (const T*)d.b, (const T*)d.dinv, c1, c2, d.vscale);
}

template <int BS, int K, class T>
__global__ void __launch_bounds__(kBlock) kbm_restrict(const SmoothSeg* __restrict__ seg,
                                                       const int32_t* __restrict__ map) {
  const SmoothSeg& d = seg[map[blockIdx.x]];
  if (all_done<K>(d.st)) return;
  const int64_t lb = (int64_t)blockIdx.x - d.blk0;
  if (d.fuse)
    mg_restrict_first_rows<BS, K, T>(lb, d.nblk, d.fd, d.cd, d.anchor != 0, (const T*)d.r, (T*)d.bc, (T*)d.rc,
                                     (T*)d.dc, (T*)d.xc, (const T*)d.dinvc, d.c0c);
  else
    mg_restrict_rows<BS * K, T>(lb, d.nblk, d.fd, d.cd, d.anchor != 0, (const T*)d.r, (T*)d.bc);
}

template <int W, int K, class T>
__global__ void __launch_bounds__(kBlock) kbm_prolong(const SmoothSeg* __restrict__ seg,
                                                      const int32_t* __restrict__ map) {
  const SmoothSeg& d = seg[map[blockIdx.x]];
  if (all_done<K>(d.st)) return;
  mg_prolong_rows<W, T>((int64_t)blockIdx.x - d.blk0, d.nblk, d.fd, d.cd, d.anchor != 0, (const T*)d.xc, (T*)d.x);
}

template <int K, class T>
__global__ void __launch_bounds__(kBlock) kbm_dense(const DenseSeg* __restrict__ seg, const int32_t* __restrict__ map) {
  const DenseSeg& d = seg[map[blockIdx.x]];
  if (all_done<K>(d.st)) return;
  mg_dense_rows<K, T>((int64_t)blockIdx.x - d.blk0, d.nblk, d.n, (const T*)d.cinv, (const T*)d.b, (T*)d.x);
}

// q = A p and partials of p_j.q_j (PLAIN: y = A x for the start, no partials)
template <int BS, int K, bool PLAIN>
__global__ void __launch_bounds__(kBlock) kbm_spmv(const OuterSeg* __restrict__ seg, const int32_t* __restrict__ map,
                                                   double* partials) {
  const OuterSeg& d = seg[map[blockIdx.x]];
  if (PLAIN ? !d.use_x0 : all_done<K>(d.st)) return;
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) acc[j] = 0.0;
  const int64_t lb = xcd_perm((int64_t)blockIdx.x - d.sblk0, d.snblk, d.ga.xg);
  if (!PLAIN && d.delta > 0.0) {
    // reliable-update systems: FP32 p and q, the FP32 form, or every FP64
    // slot two at a time once the system's products went to FP64 (as the
    // one-system kernel reads)
    const float* in = (const float*)d.p;
    float* out = (float*)d.q;
    if (!any_p64<K>(d.st)) {
      if (d.stencil)
        sell_rows<BS, K, float, true, false, true, float>(lb, d.snblk, d.nb, d.sptr, d.col, d.v32, d.dgc, in,
                                                          nullptr, out, d.ga, acc);
      else
        sell_rows<BS, K, float, true, false, false, float>(lb, d.snblk, d.nb, d.sptr, d.col, d.v32, d.dgc, in,
                                                           nullptr, out, d.ga, acc);
    } else {
      GridArg g64 = d.ga;
      g64.kuhn = 0;
      if (d.stencil)
        sell_rows<BS, K, float, true, false, true, double, 2>(lb, d.snblk, d.nb, d.sptr, d.col, d.v64, d.dg64, in,
                                                              nullptr, out, g64, acc);
      else
        sell_rows<BS, K, float, true, false, false, double, 2>(lb, d.snblk, d.nb, d.sptr, d.col, d.v64, d.dg64, in,
                                                               nullptr, out, g64, acc);
    }
  } else {
    // FP64 products with the half copy of the values
    const double* in = PLAIN ? d.x : (const double*)d.p;
    double* out = (double*)d.q;
    if (d.stencil)
      sell_rows<BS, K, double, !PLAIN, false, true>(lb, d.snblk, d.nb, d.sptr, d.col, d.v64h, d.dg64, in, nullptr,
                                                    out, d.ga64, acc);
    else
      sell_rows<BS, K, double, !PLAIN, false, false>(lb, d.snblk, d.nb, d.sptr, d.col, d.v64h, d.dg64, in, nullptr,
                                                     out, d.ga64, acc);
  }
  if (!PLAIN) block_partial<K>(acc, partials);
}

// reliable update of the systems that asked for one: r = b - A x (FP64
// form) on those columns, partials of their r.r (SpMV workgroup layout)
template <int BS, int K>
__global__ void __launch_bounds__(kBlock) kbm_true(const OuterSeg* __restrict__ seg, const int32_t* __restrict__ map,
                                                   double* partials) {
  const OuterSeg& d = seg[map[blockIdx.x]];
  if (!any_upd<K>(d.st)) return;
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) acc[j] = 0.0;
  const int64_t lb = xcd_perm((int64_t)blockIdx.x - d.sblk0, d.snblk, d.ga.xg);
  if (d.stencil)
    pcg_true_rows<BS, K, true>(lb, d.snblk, d.nb, d.sptr, d.col, d.v64h, d.dg64, d.ga64, d.x, d.b, d.r, d.st, acc);
  else
    pcg_true_rows<BS, K, false>(lb, d.snblk, d.nb, d.sptr, d.col, d.v64h, d.dg64, d.ga64, d.x, d.b, d.r, d.st, acc);
  block_partial<K>(acc, partials);
}

template <int K>
__global__ void __launch_bounds__(kBlock) kbm_init(const OuterSeg* __restrict__ seg, const int32_t* __restrict__ map,
                                                   double* partials) {
  const OuterSeg& d = seg[map[blockIdx.x]];
  double acc[2 * K];
#pragma unroll
  for (int j = 0; j < 2 * K; ++j) acc[j] = 0.0;
  pcg_init_rows<K>((int64_t)blockIdx.x - d.blk0, d.nblk, d.n, d.b, d.x, (const double*)d.q, d.r, d.use_x0, acc);
  block_partial<2 * K>(acc, partials);
}

template <int K>
__global__ void __launch_bounds__(kBlock) kbm_zero_b(const OuterSeg* __restrict__ seg,
                                                     const int32_t* __restrict__ map) {
  const OuterSeg& d = seg[map[blockIdx.x]];
  if (!d.use_x0) return;
  zero_b_rows<K>((int64_t)blockIdx.x - d.blk0, d.nblk, d.n, d.x, d.st);
}

template <int K>
__global__ void __launch_bounds__(kBlock) kbm_xr(const OuterSeg* __restrict__ seg, const int32_t* __restrict__ map,
                                                 double* partials) {
  const OuterSeg& d = seg[map[blockIdx.x]];
  if (all_done<K>(d.st)) return;
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) acc[j] = 0.0;
  const int64_t lb = (int64_t)blockIdx.x - d.blk0;
  if (d.delta > 0.0)
    pcg_xr_rows<K, float>(lb, d.nblk, d.n, d.st, d.x, d.r, (const float*)d.p, (const float*)d.q, acc);
  else
    pcg_xr_rows<K, double>(lb, d.nblk, d.n, d.st, d.x, d.r, (const double*)d.p, (const double*)d.q, acc);
  block_partial<K>(acc, partials);
}

// floating systems: partials of the per-component sums of r (read by the
// first Chebyshev step of depth 0)
template <int BS, int K>
__global__ void __launch_bounds__(kBlock) kbm_rsum(const OuterSeg* __restrict__ seg, const int32_t* __restrict__ map,
                                                   double* partials) {
  const OuterSeg& d = seg[map[blockIdx.x]];
  if (!d.floating || all_done<K>(d.st)) return;
  double acc[BS * K];
#pragma unroll
  for (int c = 0; c < BS * K; ++c) acc[c] = 0.0;
  pcg_rsum_rows<BS * K>((int64_t)blockIdx.x - d.blk0, d.nblk, d.n, d.r, acc);
  block_partial<BS * K>(acc, partials);
}

template <int BS, int K, class T>
__global__ void __launch_bounds__(kBlock) kbm_dot(const OuterSeg* __restrict__ seg, const int32_t* __restrict__ map,
                                                  double* partials) {
  const OuterSeg& d = seg[map[blockIdx.x]];
  if (all_done<K>(d.st)) return;
  double acc[2 * K];
#pragma unroll
  for (int j = 0; j < 2 * K; ++j) acc[j] = 0.0;
  const int64_t lb = (int64_t)blockIdx.x - d.blk0;
  const bool rel = d.delta > 0.0;
  bool done = false;
  if constexpr (K == 1) {
    if (d.floating) {
      if (rel)
        pcg_dot_rows<K, T, BS, float>(lb, d.nblk, d.n, d.st, d.r, (const float*)d.q, (const T*)d.z, acc);
      else
        pcg_dot_rows<K, T, BS, double>(lb, d.nblk, d.n, d.st, d.r, (const double*)d.q, (const T*)d.z, acc);
      done = true;
    }
  }
  if (!done) {
    if (rel)
      pcg_dot_rows<K, T, 0, float>(lb, d.nblk, d.n, d.st, d.r, (const float*)d.q, (const T*)d.z, acc);
    else
      pcg_dot_rows<K, T, 0, double>(lb, d.nblk, d.n, d.st, d.r, (const double*)d.q, (const T*)d.z, acc);
  }
  block_partial<2 * K>(acc, partials);
}

template <int BS, int K, class T>
__global__ void __launch_bounds__(kBlock) kbm_pupd(const OuterSeg* __restrict__ seg, const int32_t* __restrict__ map) {
  const OuterSeg& d = seg[map[blockIdx.x]];
  if (all_done<K>(d.st)) return;
  const int64_t lb = (int64_t)blockIdx.x - d.blk0;
  const bool rel = d.delta > 0.0;
  if constexpr (K == 1) {
    if (d.floating) {
      if (rel)
        pcg_pupd_rows<K, T, BS, float>(lb, d.nblk, d.n, d.st, (const T*)d.z, (float*)d.p, d.r);
      else
        pcg_pupd_rows<K, T, BS, double>(lb, d.nblk, d.n, d.st, (const T*)d.z, (double*)d.p, d.r);
      return;
    }
  }
  if (rel)
    pcg_pupd_rows<K, T, 0, float>(lb, d.nblk, d.n, d.st, (const T*)d.z, (float*)d.p, d.r);
  else
    pcg_pupd_rows<K, T, 0, double>(lb, d.nblk, d.n, d.st, (const T*)d.z, (double*)d.p, d.r);
}

enum : int { MRED_INIT = 0, MRED_ALPHA = 1, MRED_RR = 2, MRED_RZ = 3, MRED_TRUE = 4 };

// One workgroup per system: the system's partials in a fixed order (those
// of its SpMV workgroups for MRED_ALPHA and MRED_TRUE, of its vector
// workgroups otherwise),
// as the one-system reducers of mg.hip sum them.
template <int K, int MODE>
__global__ void __launch_bounds__(kRedBlock) kbm_red(const OuterSeg* __restrict__ seg, int G,
                                                     const double* __restrict__ partials) {
  constexpr int NV = (MODE == MRED_INIT || MODE == MRED_RZ) ? 2 * K : K;
  const OuterSeg& d = seg[blockIdx.x];
  if (MODE == MRED_TRUE ? !any_upd<K>(d.st) : (MODE != MRED_INIT && all_done<K>(d.st))) return;
  __shared__ double lds[NV * (kRedBlock / 64)];
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  constexpr bool SP = (MODE == MRED_ALPHA || MODE == MRED_TRUE);  // partials of SpMV workgroups
  const int32_t b0 = SP ? d.sblk0 : d.blk0, nbk = SP ? d.snblk : d.nblk;
  for (int i = threadIdx.x; i < nbk; i += kRedBlock)
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += partials[(int64_t)k * G + b0 + i];
  block_sum_t<NV, kRedBlock>(v, lds);
  if (threadIdx.x != 0) return;
  if constexpr (MODE == MRED_INIT) {
    if (d.delta > 0.0)
      cg_state_init<K, true>(d.st, v, d.tol, d.maxiter);
    else
      cg_state_init<K, false>(d.st, v, d.tol, d.maxiter);
  } else if constexpr (MODE == MRED_ALPHA) {
    if (d.delta > 0.0)
      cg_state_update<K, RED_ALPHA_PCG_REL>(d.st, v, 0.0);
    else
      cg_state_update<K, RED_ALPHA>(d.st, v, 0.0);
  } else if constexpr (MODE == MRED_RR) {
    if (d.delta > 0.0)
      cg_state_update<K, RED_RR_PCG_REL>(d.st, v, d.delta);
    else
      cg_state_update<K, RED_RR_PCG>(d.st, v, 0.0);
  } else if constexpr (MODE == MRED_TRUE) {
    cg_state_update<K, RED_RR_PCG_TRUE>(d.st, v, 0.0);
  } else {
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (!d.st[j].done) pcg_state_rz(d.st[j], v[j], v[K + j]);
  }
}

// ---------------------------------------------------------------- host

namespace {

template <class Seg>
struct Table {
  std::vector<Seg> h;
  std::vector<int32_t> map;
  Seg* d = nullptr;
  int32_t* dmap = nullptr;
  size_t cap = 0, capmap = 0;
  int G = 0;
};

template <class Seg>
int upload(Table<Seg>& t, hipStream_t s) {
  if (t.h.size() > t.cap) {
    if (t.d) CPM_HIP(hipFree(t.d));
    t.cap = std::max(t.h.size(), 2 * t.cap);
    CPM_HIP(hipMalloc(&t.d, sizeof(Seg) * t.cap));
  }
  if (t.map.size() > t.capmap) {
    if (t.dmap) CPM_HIP(hipFree(t.dmap));
    t.capmap = std::max(t.map.size(), t.capmap + t.capmap / 2);
    CPM_HIP(hipMalloc(&t.dmap, sizeof(int32_t) * t.capmap));
  }
  if (!t.h.empty()) CPM_HIP(hipMemcpyAsync(t.d, t.h.data(), sizeof(Seg) * t.h.size(), hipMemcpyHostToDevice, s));
  if (!t.map.empty())
    CPM_HIP(hipMemcpyAsync(t.dmap, t.map.data(), sizeof(int32_t) * t.map.size(), hipMemcpyHostToDevice, s));
  t.G = (int)t.map.size();
  return CPM_OK;
}

template <class Seg>
void free_table(Table<Seg>& t) {
  if (t.d) (void)hipFree(t.d);
  if (t.dmap) (void)hipFree(t.dmap);
  t.d = nullptr, t.dmap = nullptr, t.cap = t.capmap = 0;
}

}  // namespace

struct MultiMgTables {
  Table<OuterSeg> outer;               // map: vector workgroups
  std::vector<int32_t> smap;           // SpMV workgroups
  int32_t* dsmap = nullptr;
  size_t capsmap = 0;
  std::vector<Table<SmoothSeg>> smooth;  // per depth
  std::vector<Table<DenseSeg>> dense;    // per depth
  std::vector<Table<SmoothSeg>> csm;     // per depth: coarsest levels solved by Chebyshev steps
  CgState* st = nullptr;
  CgState* hst = nullptr;
  double* partials = nullptr;
  size_t cap_sys = 0, cap_part = 0;
  ~MultiMgTables() {
    free_table(outer);
    for (auto& t : smooth) free_table(t);
    for (auto& t : dense) free_table(t);
    for (auto& t : csm) free_table(t);
    if (st) (void)hipFree(st);
    if (hst) (void)hipHostFree(hst);
    if (partials) (void)hipFree(partials);
    if (dsmap) (void)hipFree(dsmap);
  }
};

MultiMgSolver::MultiMgSolver() : T(new MultiMgTables()) {}
MultiMgSolver::~MultiMgSolver() { delete T; }

template <int BS, int K, class T, class VT>
static int run(cpm_ctx* ctx, MultiMgTables& W, const std::vector<MultiMgSys>& sys, int hint, int32_t* iters,
               double* resid, int* failed) {
  hipStream_t s = ctx->stream;
  const int ns = (int)sys.size();
  // smoothing degree at depth d: the finest level's, or the coarser levels'
  // (the same for every system of the batch)
  const int m0 = sys[0].M->degree, m1 = sys[0].M->degree1 ? sys[0].M->degree1 : m0;
  auto mdeg = [&](int d) { return d == 0 ? m0 : m1; };
  int depth = 0;  // deepest level index over the batch
  for (const auto& y : sys) depth = std::max(depth, (int)y.M->L.size() - 1);
  if ((size_t)(ns * K) > W.cap_sys) {
    if (W.st) CPM_HIP(hipFree(W.st));
    if (W.hst) CPM_HIP(hipHostFree(W.hst));
    W.cap_sys = std::max((size_t)(ns * K), 2 * W.cap_sys);
    CPM_HIP(hipMalloc(&W.st, sizeof(CgState) * W.cap_sys));
    CPM_HIP(hipHostMalloc(&W.hst, sizeof(CgState) * W.cap_sys, hipHostMallocDefault));
  }
  // outer (fine level, FP64) segments
  auto& O = W.outer;
  O.h.clear();
  O.map.clear();
  W.smap.clear();
  for (int j = 0; j < ns; ++j) {
    const MultiMgSys& y = sys[j];
    cpm_mg* M = y.M;
    const Sell& S = M->L[0].A->sell;
    OuterSeg o{};
    o.sptr = S.sptr, o.col = S.col, o.v64 = S.v64, o.v64h = sell_v64(S), o.dg64 = S.dg64, o.v32 = S.v32;
    o.dgc = S.dgc;
    o.delta = y.delta;
    o.ga = CPM_GRID_ARG32(S);  // v64 holds every slot: the FP64 products read it in full (h64 = 0)
    o.ga64 = CPM_GRID_ARG64(S);
    o.nb = S.nbrows, o.stencil = S.stencil, o.n = M->L[0].n * K;
    o.b = y.b, o.x = y.x, o.r = M->r, o.p = M->p, o.q = M->q, o.z = M->L[0].x;
    o.st = W.st + j * K;
    o.blk0 = (int32_t)O.map.size();
    o.nblk = grid_for(o.n);
    o.sblk0 = (int32_t)W.smap.size();
    o.snblk = sell_grid(S);
    o.use_x0 = y.use_x0, o.maxiter = y.maxiter, o.tol = y.tol;
    o.floating = M->floating;
    for (int b = 0; b < o.nblk; ++b) O.map.push_back(j);
    for (int b = 0; b < o.snblk; ++b) W.smap.push_back(j);
    O.h.push_back(o);
  }
  // partials: 2K per vector workgroup (BS*K for the sums of a floating
  // system's residual), K per SpMV workgroup
  {
    const size_t need = (size_t)std::max(2 * K, BS * K) * std::max(O.map.size(), W.smap.size());
    if (need > W.cap_part) {
      if (W.partials) CPM_HIP(hipFree(W.partials));
      W.cap_part = std::max(need, W.cap_part + W.cap_part / 2);
      CPM_HIP(hipMalloc(&W.partials, sizeof(double) * W.cap_part));
    }
  }
  double* pt = W.partials;
  // V-cycle segments per depth
  if ((int)W.smooth.size() < depth + 1) W.smooth.resize(depth + 1);
  if ((int)W.dense.size() < depth + 1) W.dense.resize(depth + 1);
  if ((int)W.csm.size() < depth + 1) W.csm.resize(depth + 1);
  // the smoothing fields of level d of system j (operator values, diagonal
  // blocks, level vectors, Chebyshev coefficients)
  auto level_fields = [&](SmoothSeg& g, int j, int d) {
    cpm_mg* M = sys[j].M;
    const MgLevel& Lv = M->L[d];
    const Sell& S = Lv.A->sell;
    g.sptr = S.sptr, g.col = S.col;
    // smoothing values: FP16 form, or (recovered stiffness of a
    // floating level 0, else the level's operator) FP32 / FP64
    g.val = M->half            ? (const void*)Lv.v16
            : (M->prec == 32) ? (Lv.kv32 ? (const void*)Lv.kv32 : (const void*)S.v32)
                              : (Lv.kv64 ? (const void*)Lv.kv64 : (const void*)S.v64);
    g.dg = M->half            ? (const void*)Lv.dg32
           : (M->prec == 32) ? (const void*)(Lv.kdgc ? Lv.kdgc : S.dgc)
                             : (const void*)(Lv.kdg64 ? Lv.kdg64 : S.dg64);
    g.vscale = M->half ? Lv.vscale : 1.0;
    g.ga = CPM_GRID_ARG(S);
    // values keeping half of the stencil: the FP16 form (k_mg_make_half),
    // or the level's own FP32 form
    if (M->half ? Lv.khalf != 0 : (g.val == (const void*)S.v32 && S.half32)) g.ga.kuhn = 2;
    // the level's own FP64 form: its half copy when there is one
    if (g.val == (const void*)S.v64 && S.half64) {
      g.val = S.v64h;
      g.ga = CPM_GRID_ARG64(S);
    }
    g.nb = S.nbrows, g.stencil = S.stencil;
    g.x = Lv.x, g.r = Lv.r, g.b = Lv.b, g.d0 = Lv.d0, g.d1 = Lv.d1, g.dinv = Lv.dinv;
    g.c0 = M->c0[d];
    for (int k = 0; k < 8; ++k) g.c1[k] = M->c1[d * 8 + k], g.c2[k] = M->c2[d * 8 + k];
    g.st = W.st + j * K;
  };
  for (int d = 0; d <= depth; ++d) {
    auto& Sm = W.smooth[d];
    auto& Dn = W.dense[d];
    auto& Cs = W.csm[d];
    Sm.h.clear(), Sm.map.clear(), Dn.h.clear(), Dn.map.clear(), Cs.h.clear(), Cs.map.clear();
    for (int j = 0; j < ns; ++j) {
      cpm_mg* M = sys[j].M;
      const int last = (int)M->L.size() - 1;
      if (d < last) {
        const MgLevel& Lv = M->L[d];
        const MgLevel& Lc = M->L[d + 1];
        const Sell& S = Lv.A->sell;
        SmoothSeg g{};
        level_fields(g, j, d);
        g.nsteps = M->level_degree(d);
        g.b64 = (d == 0) ? M->r : nullptr;
        g.fd = Dims3{{Lv.dims[0], Lv.dims[1], Lv.dims[2]}};
        g.cd = Dims3{{Lc.dims[0], Lc.dims[1], Lc.dims[2]}};
        g.bc = Lc.b, g.xc = Lc.x;
        g.anchor = M->anchor;
        // the next level's first step with the restriction, unless it is a
        // dense coarsest level
        const bool fuse = d + 1 < last || M->csmooth;
        g.fuse = fuse ? 1 : 0;
        g.rc = Lc.r, g.dc = Lc.d0, g.dinvc = Lc.dinv;
        g.c0c = fuse ? M->c0[d + 1] : 0.0;
        if (d == 0 && M->floating) {
          // the partials kbm_rsum leaves for this system's vector workgroups
          g.fpart = pt + O.h[j].blk0;
          g.fstride = (int64_t)O.map.size();
          g.fn = O.h[j].nblk;
        }
        g.blk0 = (int32_t)Sm.map.size();
        g.nblk = sell_grid(S);
        for (int b = 0; b < g.nblk; ++b) Sm.map.push_back((int32_t)Sm.h.size());
        Sm.h.push_back(g);
      } else if (d == last && M->csmooth) {
        // coarsest level solved by Chebyshev steps 1 .. csmooth-1 (step 0
        // ran with the restriction)
        SmoothSeg g{};
        level_fields(g, j, d);
        g.nsteps = M->csmooth;
        g.blk0 = (int32_t)Cs.map.size();
        g.nblk = sell_grid(M->L[d].A->sell);
        for (int b = 0; b < g.nblk; ++b) Cs.map.push_back((int32_t)Cs.h.size());
        Cs.h.push_back(g);
      } else if (d == last) {
        DenseSeg e{};
        e.n = M->L[d].n, e.cinv = M->cinv, e.b = M->L[d].b, e.x = M->L[d].x, e.st = W.st + j * K;
        e.blk0 = (int32_t)Dn.map.size();
        e.nblk = (int32_t)std::max<int64_t>(1, std::min<int64_t>(kMaxGrid, (e.n + 3) / 4));
        for (int b = 0; b < e.nblk; ++b) Dn.map.push_back((int32_t)Dn.h.size());
        Dn.h.push_back(e);
      }
    }
    if (int rc = upload(Sm, s)) return rc;
    if (int rc = upload(Dn, s)) return rc;
    if (int rc = upload(Cs, s)) return rc;
  }
  int csteps = 0;  // most Chebyshev steps of a smoothed coarsest level in the batch
  for (const auto& y : sys) csteps = std::max(csteps, y.M->csmooth);
  if (int rc = upload(O, s)) return rc;
  if (W.smap.size() > W.capsmap) {
    if (W.dsmap) CPM_HIP(hipFree(W.dsmap));
    W.capsmap = std::max(W.smap.size(), W.capsmap + W.capsmap / 2);
    CPM_HIP(hipMalloc(&W.dsmap, sizeof(int32_t) * W.capsmap));
  }
  CPM_HIP(hipMemcpyAsync(W.dsmap, W.smap.data(), sizeof(int32_t) * W.smap.size(), hipMemcpyHostToDevice, s));
  const int G = O.G, Gs = (int)W.smap.size();
  bool any_x0 = false, any_rel = false;
  int maxit = 0;
  for (const auto& y : sys)
    any_x0 |= y.use_x0 != 0, any_rel |= y.delta > 0.0, maxit = std::max(maxit, (int)y.maxiter);

  bool any_fl = false;
  for (const auto& y : sys) any_fl |= y.M->floating != 0;
  auto vcycle = [&]() {
    if (any_fl) kbm_rsum<BS, K><<<G, kBlock, 0, s>>>(O.d, O.dmap, pt);
    for (int d = 0; d < depth; ++d) {
      auto& Sm = W.smooth[d];
      if (Sm.G) {
        // depth > 0: the first step ran fused with the restriction
        if (d == 0) kbm_first<BS, K, T, true, true><<<Sm.G, kBlock, 0, s>>>(Sm.d, Sm.dmap);
        const int m = mdeg(d);
        for (int k = 1; k < m; ++k) kbm_sweep<BS, K, T, 0, VT><<<Sm.G, kBlock, 0, s>>>(Sm.d, Sm.dmap, k);
        kbm_sweep<BS, K, T, 1, VT><<<Sm.G, kBlock, 0, s>>>(Sm.d, Sm.dmap, m - 1);
        kbm_restrict<BS, K, T><<<Sm.G, kBlock, 0, s>>>(Sm.d, Sm.dmap);
      }
      auto& Dn = W.dense[d + 1];
      if (Dn.G) kbm_dense<K, T><<<Dn.G, kBlock, 0, s>>>(Dn.d, Dn.dmap);
      auto& Cs = W.csm[d + 1];
      if (Cs.G)
        for (int k = 1; k < csteps; ++k) kbm_sweep<BS, K, T, 0, VT><<<Cs.G, kBlock, 0, s>>>(Cs.d, Cs.dmap, k);
    }
    for (int d = depth - 1; d >= 0; --d) {
      auto& Sm = W.smooth[d];
      if (!Sm.G) continue;
      kbm_prolong<BS * K, K, T><<<Sm.G, kBlock, 0, s>>>(Sm.d, Sm.dmap);
      const int m = mdeg(d);
      if (m >= 2) {
        // the residual with the first step's direction, then x += d_0 with the second step
        kbm_sweep<BS, K, T, 3, VT><<<Sm.G, kBlock, 0, s>>>(Sm.d, Sm.dmap, 0);
        kbm_sweep<BS, K, T, 4, VT><<<Sm.G, kBlock, 0, s>>>(Sm.d, Sm.dmap, 1);
        for (int k = 2; k < m; ++k) kbm_sweep<BS, K, T, 0, VT><<<Sm.G, kBlock, 0, s>>>(Sm.d, Sm.dmap, k);
      } else {
        kbm_sweep<BS, K, T, 2, VT><<<Sm.G, kBlock, 0, s>>>(Sm.d, Sm.dmap, 0);
        kbm_first<BS, K, T, false, false><<<Sm.G, kBlock, 0, s>>>(Sm.d, Sm.dmap);
      }
    }
  };

  // start
  if (any_x0) kbm_spmv<BS, K, true><<<Gs, kBlock, 0, s>>>(O.d, W.dsmap, pt);
  kbm_init<K><<<G, kBlock, 0, s>>>(O.d, O.dmap, pt);
  kbm_red<K, MRED_INIT><<<ns, kRedBlock, 0, s>>>(O.d, G, pt);
  if (any_x0) kbm_zero_b<K><<<G, kBlock, 0, s>>>(O.d, O.dmap);
  vcycle();
  kbm_dot<BS, K, T><<<G, kBlock, 0, s>>>(O.d, O.dmap, pt);
  kbm_red<K, MRED_RZ><<<ns, kRedBlock, 0, s>>>(O.d, G, pt);
  CPM_HIP(hipGetLastError());
  auto issue = [&](int todo) -> int {
    for (int k = 0; k < todo; ++k) {
      kbm_pupd<BS, K, T><<<G, kBlock, 0, s>>>(O.d, O.dmap);
      kbm_spmv<BS, K, false><<<Gs, kBlock, 0, s>>>(O.d, W.dsmap, pt);
      kbm_red<K, MRED_ALPHA><<<ns, kRedBlock, 0, s>>>(O.d, Gs, pt);
      kbm_xr<K><<<G, kBlock, 0, s>>>(O.d, O.dmap, pt);
      kbm_red<K, MRED_RR><<<ns, kRedBlock, 0, s>>>(O.d, G, pt);
      if (any_rel) {
        kbm_true<BS, K><<<Gs, kBlock, 0, s>>>(O.d, W.dsmap, pt);
        kbm_red<K, MRED_TRUE><<<ns, kRedBlock, 0, s>>>(O.d, Gs, pt);
      }
      vcycle();
      kbm_dot<BS, K, T><<<G, kBlock, 0, s>>>(O.d, O.dmap, pt);
      kbm_red<K, MRED_RZ><<<ns, kRedBlock, 0, s>>>(O.d, G, pt);
    }
    CPM_HIP(hipGetLastError());
    return CPM_OK;
  };
  const int nst = ns * K;
  auto read = [&]() -> int {
    CPM_HIP(hipMemcpyAsync(W.hst, W.st, sizeof(CgState) * nst, hipMemcpyDeviceToHost, s));
    CPM_HIP(hipStreamSynchronize(s));
    return CPM_OK;
  };
  auto all_live_done = [&]() {
    for (int j = 0; j < nst; ++j)
      if (!W.hst[j].done) return false;
    return true;
  };
  if (int rc = read()) return rc;
  // iterations per host check: the hint (the slowest member's previous
  // count) less one, then one at a time up to two past it, then a quarter of
  // the count so far (an iteration after the last live member has converged
  // costs a V-cycle's worth of launches, a check one round trip)
  int batch = hint > 2 ? hint - 1 : 4;
  int issued = 0;
  while (!all_live_done() && issued <= maxit) {
    const int todo = std::max(1, std::min(std::min(batch, 1024), maxit - issued + 1));
    if (int rc = issue(todo)) return rc;
    issued += todo;
    if (int rc = read()) return rc;
    int it_max = 0;
    for (int j = 0; j < nst; ++j)
      if (!W.hst[j].done) it_max = std::max(it_max, (int)W.hst[j].it);
    batch = (hint > 2 && it_max < hint + 2) ? 1 : std::max(2, std::min(16, it_max / 4 + 2));
  }
  for (int j = 0; j < ns; ++j) {
    failed[j] = 0;
    for (int c = 0; c < K; ++c) {
      const CgState& h = W.hst[j * K + c];
      iters[j * K + c] = h.it;
      resid[j * K + c] = h.resid;
      if (h.fail || !h.done) failed[j] = 1;
    }
  }
  return CPM_OK;
}

int multi_mg_pcg(cpm_ctx* ctx, MultiMgSolver& W, const std::vector<MultiMgSys>& sys, int K, int hint,
                 int32_t* iters, double* resid, int* failed) {
  if (sys.empty()) return CPM_OK;
  const cpm_mg* M0 = sys[0].M;
  for (const auto& y : sys) {
    if (!y.M || !y.b || !y.x) return fail(CPM_EINVAL, "batched multigrid CG: null system");
    if (y.M->ctx != ctx) return fail(CPM_EINVAL, "batched multigrid CG: a hierarchy uses another context");
    if (y.M->bs != M0->bs || y.M->prec != M0->prec || y.M->half != M0->half || y.M->degree != M0->degree ||
        y.M->degree1 != M0->degree1)
      return fail(CPM_EINVAL, "batched multigrid CG: hierarchies differ in block size, precision or degree");
    if (y.M->L.size() > 16) return fail(CPM_EINVAL, "batched multigrid CG: too many levels");
    if (!(y.delta >= 0.0 && y.delta < 1.0)) return fail(CPM_EINVAL, "batched multigrid CG: delta must be in [0, 1)");
  }
  if (K != 1 && !(K == 3 && M0->bs == 1))
    return fail(CPM_EINVAL, "batched multigrid CG: 1 right-hand side, or 3 for scalar operators");
  for (const auto& y : sys)
    if (y.M->floating && K != 1) return fail(CPM_EINVAL, "batched multigrid CG: a floating hierarchy solves one column");
  for (size_t i = 0; i < sys.size(); ++i)
    for (size_t j = i + 1; j < sys.size(); ++j)
      if (sys[i].M == sys[j].M)
        return fail(CPM_EINVAL, "batched multigrid CG: two systems share one hierarchy (and its workspace)");
  for (const auto& y : sys)
    if (int rc = mg_reserve(y.M, K)) return rc;
  MultiMgTables& T = *W.T;
  if (M0->half) {
    if (M0->bs == 3) return run<3, 1, float, _Float16>(ctx, T, sys, hint, iters, resid, failed);
    if (K == 3) return run<1, 3, float, _Float16>(ctx, T, sys, hint, iters, resid, failed);
    return run<1, 1, float, _Float16>(ctx, T, sys, hint, iters, resid, failed);
  }
  if (M0->bs == 3)
    return M0->prec == 32 ? run<3, 1, float, float>(ctx, T, sys, hint, iters, resid, failed)
                          : run<3, 1, double, double>(ctx, T, sys, hint, iters, resid, failed);
  if (K == 3)
    return M0->prec == 32 ? run<1, 3, float, float>(ctx, T, sys, hint, iters, resid, failed)
                          : run<1, 3, double, double>(ctx, T, sys, hint, iters, resid, failed);
  return M0->prec == 32 ? run<1, 1, float, float>(ctx, T, sys, hint, iters, resid, failed)
                        : run<1, 1, double, double>(ctx, T, sys, hint, iters, resid, failed);
}

}  // namespace cpm

using namespace cpm;

extern "C" int cpm_cg_mg_batch(cpm_ctx* ctx, int count, int nrhs, cpm_mg* const* mgs, const double* const* b,
                               double* const* x, const int32_t* use_x0, double tol, int maxiter, int32_t* iters,
                               double* resid, int32_t* failed) {
  if (!ctx || count < 0 || (count > 0 && (!mgs || !b || !x || !iters || !resid || !failed)))
    return fail(CPM_EINVAL, "cpm_cg_mg_batch: null argument");
  if (count == 0) return CPM_OK;
  CPM_HIP(hipSetDevice(ctx->device));
  std::vector<MultiMgSys> sys(count);
  for (int j = 0; j < count; ++j) {
    sys[j].M = mgs[j];
    sys[j].b = b[j];
    sys[j].x = x[j];
    sys[j].use_x0 = use_x0 ? use_x0[j] : 0;
    sys[j].tol = tol;
    sys[j].maxiter = maxiter;
  }
  MultiMgSolver W;
  std::vector<int> f(count);
  if (int rc = multi_mg_pcg(ctx, W, sys, nrhs, 0, iters, resid, f.data())) return rc;
  for (int j = 0; j < count; ++j) failed[j] = f[j];
  return CPM_OK;
}
