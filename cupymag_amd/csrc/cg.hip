// CSR SpMV and the FP64 conjugate-gradient solver with CuPy's stopping rule.
//
// Replaces cuSPARSE `A @ x` and cupyx.scipy.sparse.linalg.cg as called by
// solve_cg (cupymag/solvers/linear_solvers.py:59-62).  One CG iteration is
// three kernels on the context stream:
//   pupd : p = r (first iteration) or p = r + (rho/rho_old) p
//   spmv : q = A p, fused p.q partials; k_red sets alpha
//   xr   : x += alpha p, r -= alpha q, fused r.r partials; k_red advances
//          the iteration count and tests ||r|| <= atol
// The CG scalars live in device memory (CgState); every kernel returns
// immediately once `done` is set, so the host enqueues batches of iterations
// and reads the state back only between batches instead of once per
// iteration.  Both dot products are written as per-workgroup partials and
// added by a one-workgroup kernel (k_red) in fixed order: deterministic, and
// without a device-scope fence per workgroup.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "device.h"

namespace cpm {

// y = A x with G lanes per row.
template <int G>
__global__ void __launch_bounds__(kBlock) k_spmv(int32_t nrows, const int32_t* __restrict__ ptr,
                                                 const int32_t* __restrict__ idx, const double* __restrict__ val,
                                                 const double* __restrict__ x, double* __restrict__ y) {
  const int lane = threadIdx.x & (G - 1);
  const int64_t stride = (int64_t)gridDim.x * (kBlock / G);
  for (int64_t row = (int64_t)blockIdx.x * (kBlock / G) + threadIdx.x / G; row < nrows; row += stride) {
    const int32_t b = ptr[row], e = ptr[row + 1];
    double s = 0.0;
    for (int32_t j = b + lane; j < e; j += G) s += val[j] * x[idx[j]];
    s = group_sum<G>(s);
    if (lane == 0) y[row] = s;
  }
}

// q = A p and the partials of p.q (k_red then sets alpha = rho / (p.q)).
template <int G>
__global__ void __launch_bounds__(kBlock) k_cg_spmv(int32_t nrows, const int32_t* __restrict__ ptr,
                                                    const int32_t* __restrict__ idx, const double* __restrict__ val,
                                                    const double* __restrict__ p, double* __restrict__ q,
                                                    const CgState* st, double* partials) {
  if (st->done) return;
  const int lane = threadIdx.x & (G - 1);
  const int64_t stride = (int64_t)gridDim.x * (kBlock / G);
  double acc[1] = {0.0};
  for (int64_t row = (int64_t)blockIdx.x * (kBlock / G) + threadIdx.x / G; row < nrows; row += stride) {
    const int32_t b = ptr[row], e = ptr[row + 1];
    double s = 0.0;
    for (int32_t j = b + lane; j < e; j += G) s += val[j] * p[idx[j]];
    s = group_sum<G>(s);
    if (lane == 0) {
      q[row] = s;
      acc[0] += p[row] * s;
    }
  }
  block_partial<1>(acc, partials);
}

__global__ void __launch_bounds__(kBlock) k_cg_pupd(int64_t n, const double* __restrict__ r, double* __restrict__ p,
                                                    const CgState* st) {
  if (st->done) return;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  if (st->it == 0) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) p[i] = r[i];
  } else {
    const double beta = st->rho / st->rho_old;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) p[i] = r[i] + beta * p[i];
  }
}

__global__ void __launch_bounds__(kBlock) k_cg_xr(int64_t n, double* __restrict__ x, double* __restrict__ r,
                                                  const double* __restrict__ p, const double* __restrict__ q,
                                                  const CgState* st, double* partials) {
  if (st->done) return;
  const double a = st->alpha;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  double acc[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    x[i] = x[i] + a * p[i];
    const double ri = r[i] - a * q[i];
    r[i] = ri;
    acc[0] += ri * ri;
  }
  block_partial<1>(acc, partials);
}

// ---- CG without device allocation (cpm_cg_noalloc) ----------------------------
// For callers that cannot allocate device memory (code under graph capture,
// tight serving loops): the solve's vectors r, p, q, its reduction partials,
// grid-sum counter and CG state live in one file-scope device array, held by
// every GPU the library runs on until the process ends.  Calls on one GPU
// share it: overlapping calls race.
constexpr int64_t kNoAllocMaxRows = int64_t(1) << 20;
constexpr int64_t kScratchHead = 64;  // counter (64 unsigned = 32 doubles), then the CgState
constexpr int64_t kScratchDoubles = kScratchHead + kMaxPartials + 3 * kNoAllocMaxRows;
static_assert(sizeof(CgState) <= 32 * sizeof(double), "CgState must fit the scratch head");
static __device__ __align__(256) double scratch[kScratchDoubles];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[64] = {};

// the device address of `scratch` on `device` (the current device), looked
// up once per device
int noalloc_scratch(int device, double** out) {
  if (device < 0 || device >= 64) return fail(CPM_EINVAL, "no-allocation scratch: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(CPM_EHIP, std::string("no-allocation scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (double*)p;
  }
  *out = g_scratch_ptr[device];
  return CPM_OK;
}

__global__ void k_zero_counter(unsigned* c) {
  if (threadIdx.x < 64) c[threadIdx.x] = 0u;
}

// Preconditioned CG (cupyx cg with M, which solve_cg forwards,
// linear_solvers.py:59-62): z = M r with the partials of r.z (k_red then
// sets rho_old = rho, rho = r.z), p = z or z + (rho/rho_old) p.
template <int G>
__global__ void __launch_bounds__(kBlock) k_cg_prec(int32_t nrows, const int32_t* __restrict__ ptr,
                                                    const int32_t* __restrict__ idx, const double* __restrict__ val,
                                                    const double* __restrict__ r, double* __restrict__ z,
                                                    const CgState* st, double* partials) {
  if (st->done) return;
  const int lane = threadIdx.x & (G - 1);
  const int64_t stride = (int64_t)gridDim.x * (kBlock / G);
  double acc[1] = {0.0};
  for (int64_t row = (int64_t)blockIdx.x * (kBlock / G) + threadIdx.x / G; row < nrows; row += stride) {
    const int32_t b = ptr[row], e = ptr[row + 1];
    double s = 0.0;
    for (int32_t j = b + lane; j < e; j += G) s += val[j] * r[idx[j]];
    s = group_sum<G>(s);
    if (lane == 0) {
      z[row] = s;
      acc[0] += r[row] * s;
    }
  }
  block_partial<1>(acc, partials);
}

__global__ void __launch_bounds__(kBlock) k_cg_pupd_z(int64_t n, const double* __restrict__ z, double* __restrict__ p,
                                                      const CgState* st) {
  if (st->done) return;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  if (st->it == 0) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) p[i] = z[i];
  } else {
    const double beta = st->rho / st->rho_old;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) p[i] = z[i] + beta * p[i];
  }
}

// r = b - A x (use_x0, q holds A x) or x = 0, r = b; ||b||^2, r.r and atol.
__global__ void __launch_bounds__(kBlock) k_cg_init(int64_t n, const double* __restrict__ b, double* __restrict__ x,
                                                    double* __restrict__ r, const double* __restrict__ q, int use_x0,
                                                    double tol, int32_t maxiter, CgState* st, double* partials,
                                                    unsigned* counter) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  double acc[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double bi = b[i];
    double ri;
    if (use_x0) {
      ri = bi - q[i];
    } else {
      x[i] = 0.0;
      ri = bi;
    }
    r[i] = ri;
    acc[0] += bi * bi;
    acc[1] += ri * ri;
  }
  if (grid_sum<2>(acc, partials, counter) && threadIdx.x == 0) {
    st->bnorm2 = acc[0];
    st->rho = acc[1];
    st->rho_old = 0.0;
    st->tol = tol;
    st->atol = tol * sqrt(acc[0]);
    st->it = 0;
    st->fail = 0;
    st->maxiter = maxiter;
    st->resid = sqrt(acc[1]);
    // CuPy returns b itself when ||b|| == 0, and never iterates when maxiter <= 0
    st->done = (acc[0] == 0.0 || maxiter <= 0) ? 1 : 0;
    if (acc[0] != 0.0 && maxiter <= 0) st->fail = 0;
  }
}

// When ||b|| == 0 the result is b (all zeros), also when an x0 was given.
__global__ void __launch_bounds__(kBlock) k_cg_zero_b(int64_t n, double* __restrict__ x, const CgState* st) {
  if (st->bnorm2 != 0.0) return;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) x[i] = 0.0;
}

// workgroups of a CSR SpMV launch with G lanes per row, at most A->grid_cap
template <int G>
static int csr_grid(const cpm_mat* A) {
  const int64_t g = ((int64_t)A->nrows + kBlock / G - 1) / (kBlock / G);
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, A->grid_cap));
}

template <int G>
static void launch_spmv_g(const cpm_mat* A, const double* val, const double* x, double* y, hipStream_t s) {
  const int grid = csr_grid<G>(A);
  k_spmv<G><<<grid, kBlock, 0, s>>>(A->nrows, A->ptr, A->idx, val, x, y);
}

template <int G>
static void launch_cg_prec_g(cpm_ctx* ctx, const cpm_mat* Mp, const CgWork& w, int* grid_out) {
  const int grid = csr_grid<G>(Mp);
  k_cg_prec<G><<<grid, kBlock, 0, ctx->stream>>>(Mp->nrows, Mp->ptr, Mp->idx, Mp->val, w.r, w.z, w.st, w.partials);
  *grid_out = grid;
}

template <int G>
static void launch_cg_spmv_g(cpm_ctx* ctx, const cpm_mat* A, const CgWork& w, int* grid_out) {
  const int grid = csr_grid<G>(A);
  k_cg_spmv<G><<<grid, kBlock, 0, ctx->stream>>>(A->nrows, A->ptr, A->idx, A->val, w.p, w.q, w.st, w.partials);
  *grid_out = grid;
}

#define CPM_LANES_DISPATCH(L, F, ...) \
  switch (L) {                        \
    case 1: F<1>(__VA_ARGS__); break;   \
    case 2: F<2>(__VA_ARGS__); break;   \
    case 4: F<4>(__VA_ARGS__); break;   \
    case 8: F<8>(__VA_ARGS__); break;   \
    case 16: F<16>(__VA_ARGS__); break; \
    case 32: F<32>(__VA_ARGS__); break; \
    default: F<64>(__VA_ARGS__); break; \
  }

int launch_spmv(cpm_ctx* ctx, const cpm_mat* A, int valset, const double* x, double* y) {
  if (A->nrows == 0) return CPM_OK;
  if (valset == 0 && A->spmv_sell && A->sell.sptr && A->sell.bs == 1) return sell_spmv64(ctx, A, 1, x, y);
  const double* val = A->val + (int64_t)valset * A->nnz;
  CPM_LANES_DISPATCH(A->lanes, launch_spmv_g, A, val, x, y, ctx->stream);
  CPM_HIP(hipGetLastError());
  return CPM_OK;
}

// Mp: the preconditioner (z = Mp r each iteration, CuPy's recurrence with
// M), or nullptr for plain CG.
// Wk: the vectors, partials, grid-sum counter and CG state to use (the
// context's, grown on demand, when null).
int run_cg(cpm_ctx* ctx, const cpm_mat* A, const double* b, double* x, int use_x0, double tol, int maxiter, int hint,
           int32_t* iters, double* resid, const cpm_mat* Mp, const CgWork* Wk) {
  const int64_t n = A->nrows;
  if (A->ncols != A->nrows) return fail(CPM_EINVAL, "CG needs a square matrix");
  if (Mp && (Mp->nrows != n || Mp->ncols != n)) return fail(CPM_EINVAL, "CG: the preconditioner must be n x n like A");
  if (Wk && Mp) return fail(CPM_EINVAL, "CG: a given workspace has no room for the preconditioner's vector");
  CgWork w;
  if (Wk) {
    w = *Wk;
  } else {
    if (int rc = ensure_workspace(ctx, n)) return rc;
    w = CgWork{ctx->r, ctx->p, ctx->q, ctx->z, ctx->partials, ctx->counter, ctx->cg};
  }
  CgState* st = w.st;
  hipStream_t s = ctx->stream;
  const int gv = grid_for(n);
  if (use_x0)
    if (int rc = launch_spmv(ctx, A, 0, x, w.q)) return rc;
  k_cg_init<<<gv, kBlock, 0, s>>>(n, b, x, w.r, w.q, use_x0, tol, maxiter, st, w.partials, w.counter);
  k_cg_zero_b<<<gv, kBlock, 0, s>>>(n, x, st);
  CPM_HIP(hipGetLastError());
  CgState h{};
  int batch = hint > 2 ? hint - 1 : 8;
  const bool prof = prof_on(ctx, A);
  const int max_batch = prof ? 256 : 4096;
  int issued = 0;
  for (;;) {
    const int todo = std::max(1, std::min(std::min(batch, max_batch), maxiter - issued));
    const int it_before = h.it;
    for (int k = 0; k < todo; ++k) {
      if (Mp) {
        int gm = 0;
        CPM_LANES_DISPATCH(Mp->lanes, launch_cg_prec_g, ctx, Mp, w, &gm);
        k_red<1, RED_RHO><<<1, kRedBlock, 0, s>>>(gm, w.partials, st, 0.0);
        k_cg_pupd_z<<<gv, kBlock, 0, s>>>(n, w.z, w.p, st);
      } else {
        k_cg_pupd<<<gv, kBlock, 0, s>>>(n, w.r, w.p, st);
      }
      if (prof) CPM_HIP(hipEventRecord(ctx->ev[2 * k], s));
      int gsp = 0;
      CPM_LANES_DISPATCH(A->lanes, launch_cg_spmv_g, ctx, A, w, &gsp);
      if (prof) CPM_HIP(hipEventRecord(ctx->ev[2 * k + 1], s));
      k_red<1, RED_ALPHA><<<1, kRedBlock, 0, s>>>(gsp, w.partials, st, 0.0);
      k_cg_xr<<<gv, kBlock, 0, s>>>(n, x, w.r, w.p, w.q, st, w.partials);
      if (Mp)
        k_red<1, RED_RR_PCG><<<1, kRedBlock, 0, s>>>(gv, w.partials, st, 0.0);
      else
        k_red<1, RED_RR><<<1, kRedBlock, 0, s>>>(gv, w.partials, st, 0.0);
    }
    issued += todo;
    CPM_HIP(hipGetLastError());
    CPM_HIP(hipMemcpyAsync(ctx->pinned, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
    CPM_HIP(hipStreamSynchronize(s));
    std::memcpy(&h, ctx->pinned, sizeof(CgState));
    if (prof) {
      // only launches that did work (iterations it_before .. h.it-1 of this batch)
      cpm_mat* Am = const_cast<cpm_mat*>(A);
      for (int k = 0; k < h.it - it_before && k < todo; ++k) {
        float ms = 0.f;
        CPM_HIP(hipEventElapsedTime(&ms, ctx->ev[2 * k], ctx->ev[2 * k + 1]));
        Am->prof_ms += ms;
        Am->prof_launches += 1;
      }
    }
    if (h.done || issued >= maxiter) break;
    batch = std::max(2, std::min(64, h.it / 4 + 2));
  }
  if (iters) *iters = h.it;
  if (resid) *resid = h.resid;
  if (h.fail || (!h.done && h.it >= maxiter)) return h.it > 0 ? h.it : 1;
  return CPM_OK;
}

}  // namespace cpm

extern "C" int cpm_spmv(cpm_ctx* ctx, const cpm_mat* A, int valset, const double* x, double* y) {
  if (!ctx || !A || !x || !y) return cpm::fail(CPM_EINVAL, "cpm_spmv: null argument");
  if (valset < 0 || valset >= A->nvals) return cpm::fail(CPM_EINVAL, "cpm_spmv: valset out of range");
  return cpm::launch_spmv(ctx, A, valset, x, y);
}

extern "C" int cpm_cg(cpm_ctx* ctx, const cpm_mat* A, const double* b, double* x, int use_x0, double tol, int maxiter,
                      int32_t* iters, double* resid) {
  if (!ctx || !A || !b || !x) return cpm::fail(CPM_EINVAL, "cpm_cg: null argument");
  return cpm::run_cg(ctx, A, b, x, use_x0, tol, maxiter, 0, iters, resid, nullptr);
}

extern "C" int cpm_cg_noalloc(cpm_ctx* ctx, const cpm_mat* A, const double* b, double* x, int use_x0, double tol,
                              int maxiter, int32_t* iters, double* resid) {
  using namespace cpm;
  if (!ctx || !A || !b || !x) return fail(CPM_EINVAL, "cpm_cg_noalloc: null argument");
  if (A->nrows > kNoAllocMaxRows)
    return fail(CPM_EINVAL, "cpm_cg_noalloc: " + std::to_string(A->nrows) + " rows exceed the maximum of " +
                                std::to_string(kNoAllocMaxRows) + " (use cpm_cg)");
  if (!ctx->noalloc) return fail(CPM_EHIP, "cpm_cg_noalloc: the context has no scratch address");
  double* base = ctx->noalloc;
  CgWork w;
  w.counter = (unsigned*)base;
  w.st = (CgState*)(base + 32);
  w.partials = base + kScratchHead;
  w.r = w.partials + kMaxPartials;
  w.p = w.r + kNoAllocMaxRows;
  w.q = w.p + kNoAllocMaxRows;
  w.z = nullptr;
  // the grid-sum counter must start at zero (a fill kernel, also under capture)
  k_zero_counter<<<1, 64, 0, ctx->stream>>>(w.counter);
  CPM_HIP(hipGetLastError());
  return run_cg(ctx, A, b, x, use_x0, tol, maxiter, 0, iters, resid, nullptr, &w);
}

extern "C" int cpm_cg_precond(cpm_ctx* ctx, const cpm_mat* A, const cpm_mat* M, const double* b, double* x,
                              int use_x0, double tol, int maxiter, int32_t* iters, double* resid) {
  if (!ctx || !A || !M || !b || !x) return cpm::fail(CPM_EINVAL, "cpm_cg_precond: null argument");
  return cpm::run_cg(ctx, A, b, x, use_x0, tol, maxiter, 0, iters, resid, M);
}

extern "C" int cpm_prof_enable(cpm_ctx* ctx, int on) {
  if (!ctx) return cpm::fail(CPM_EINVAL, "cpm_prof_enable: null context");
  // timing-only events: no system-scope fence (no cache write-back) at
  // record, so a pair around one launch times the kernel, not the flush
  if (on && !ctx->ev[0])
    for (auto& e : ctx->ev) CPM_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
  ctx->prof = on ? 1 : 0;
  return CPM_OK;
}

extern "C" int cpm_prof_only(cpm_ctx* ctx, const cpm_mat* A) {
  if (!ctx) return cpm::fail(CPM_EINVAL, "cpm_prof_only: null context");
  if (A && A->ctx != ctx) return cpm::fail(CPM_EINVAL, "cpm_prof_only: matrix of another context");
  ctx->prof_mat = A;
  return CPM_OK;
}

extern "C" int cpm_mat_prof_outer(cpm_mat* A, double* total_ms, int64_t* launches, int reset) {
  if (!A) return cpm::fail(CPM_EINVAL, "cpm_mat_prof_outer: null matrix");
  if (total_ms) *total_ms = A->prof2_ms;
  if (launches) *launches = A->prof2_launches;
  if (reset) {
    A->prof2_ms = 0.0;
    A->prof2_launches = 0;
  }
  return CPM_OK;
}

extern "C" int cpm_mat_prof(cpm_mat* A, double* total_ms, int64_t* launches, int reset) {
  if (!A) return cpm::fail(CPM_EINVAL, "cpm_mat_prof: null matrix");
  if (total_ms) *total_ms = A->prof_ms;
  if (launches) *launches = A->prof_launches;
  if (reset) {
    A->prof_ms = 0.0;
    A->prof_launches = 0;
  }
  return CPM_OK;
}
