// Internal declarations shared by the HIP translation units of
// libcupymag_hip.so: context / operator structs, launch shapes, and the
// deterministic wave64 + LDS reductions used by every dot product.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "cupymag_hip.h"
#include "errors.h"
#include "sell.h"

#define CPM_HIP(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return cpm::fail(CPM_EHIP, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

namespace cpm {

constexpr int kBlock = 256;      // 4 wave64 per workgroup
constexpr int kMaxGrid = 2048;   // 8 workgroups per CU on 256 CUs; grid-stride beyond
constexpr int kMaxRed = 12;      // max values reduced together
constexpr int kSellMaxGrid = 8192;   // default SELL SpMV grid cap
constexpr int kMaxSpmvGrid = 32768;  // largest SpMV grid cap an operator can be tuned to
// doubles in ctx->partials: kMaxRed per workgroup of a vector kernel, or up
// to K = 3 per workgroup of a tuned SpMV
constexpr int kMaxPartials = (kMaxRed * kMaxGrid > 3 * kMaxSpmvGrid) ? kMaxRed * kMaxGrid : 3 * kMaxSpmvGrid;
constexpr int kRedBlock = 1024;  // threads of the one-workgroup reducers

// Per-solve CG scalars, device resident.
struct CgState {
  double rho, rho_old, alpha, atol, bnorm2, resid, pq, tol;
  int32_t it, done, fail, maxiter;
  double maxr;      // mixed precision: residual norm at the last FP64 update
  int32_t upd;      // mixed precision: FP64 residual update requested
  int32_t p64;      // mixed-precision PCG: the outer products have switched to FP64
  double beta;      // preconditioned CG: (r - r_old).z / rho_old
  int32_t it_upd;   // mixed-precision PCG: iteration of the last FP64 residual
  int32_t pad;
};

// mixed-precision PCG: longest run of iterations between two FP64
// residuals (a refinement cycle takes 6-8 when the FP32 products carry it)
constexpr int kRelCycle = 40;

}  // namespace cpm

struct cpm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  double* partials = nullptr;     // kMaxRed * kMaxGrid doubles
  unsigned* counter = nullptr;    // grid-reduction arrival counter
  double* scal = nullptr;         // device scalar scratch (256 doubles)
  double* pinned = nullptr;       // pinned host mirror (256 doubles)
  cpm::CgState* cg = nullptr;     // device CG state (16 slots)
  // CG workspace (grown on demand)
  double* r = nullptr;
  double* p = nullptr;
  double* q = nullptr;
  double* z = nullptr;  // preconditioned CG: z = M r
  double* noalloc = nullptr;  // device address of the no-allocation CG scratch (cg.hip)
  int64_t ws_len = 0;
  int ncu = 256;
  // optional timing of every CG SpMV launch with HIP events (cpm_prof_enable),
  // of one operator's launches only when prof_mat is set (cpm_prof_only)
  int prof = 0;
  const cpm_mat* prof_mat = nullptr;
  hipEvent_t ev[2 * 256] = {};
  // FP32 workspace of the mixed-precision solver (+ one FP64 scratch vector)
  float *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *x32 = nullptr, *b32 = nullptr;
  double* y64 = nullptr;
  int64_t ws32_len = 0;
  // second lane of solves (ensure_lane2): a non-blocking stream of default
  // priority, its partials (kMaxPartials doubles) and pinned state mirror,
  // and the events that order it against the context's stream
  hipStream_t stream2 = nullptr;
  double* partials2 = nullptr;
  double* pinned2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

struct cpm_mat {
  cpm_ctx* ctx = nullptr;
  int32_t nrows = 0, ncols = 0;
  int64_t nnz = 0;
  int nvals = 1;
  int32_t* ptr = nullptr;
  int32_t* idx = nullptr;
  double* val = nullptr;   // nvals * nnz
  int lanes = 8;           // threads per row of the vector-CSR kernels
  int grid_cap = cpm::kMaxGrid;  // workgroup cap of the CSR SpMV launches (grid-stride beyond)
  int spmv_sell = 0;       // cpm_spmv on value set 0 uses the FP64 SELL form (scalar operators)
  cpm::Sell sell;          // optional SELL-64 form (cpm_mat_add_sell)
  double prof_ms = 0.0;    // accumulated CG SpMV kernel time (profiling on)
  int64_t prof_launches = 0;
  double prof2_ms = 0.0;   // multigrid PCG: accumulated time of the outer FP32 product (profiling on)
  int64_t prof2_launches = 0;
};

namespace cpm {

// timing of the launches on operator A requested (cpm_prof_enable, cpm_prof_only)
inline bool prof_on(const cpm_ctx* ctx, const cpm_mat* A) {
  return ctx->prof && (!ctx->prof_mat || ctx->prof_mat == A);
}

inline int grid_for(int64_t work_items, int items_per_block = kBlock) {
  int64_t g = (work_items + items_per_block - 1) / items_per_block;
  if (g < 1) g = 1;
  if (g > kMaxGrid) g = kMaxGrid;
  return (int)g;
}

// Position w of a workgroup range of g workgroups, permuted so that the rows
// a workgroup shares data with run on the same XCD.  The dispatcher deals
// workgroups to the eight XCDs round-robin (workgroup w runs on XCD w % 8)
// and every XCD has an L2 of its own.  With G > 1, each XCD takes G
// consecutive positions of every run of 8 G (the tail of the range keeps
// the dispatch order): when 8 G workgroups cover one plane of a periodic
// grid, the previous grid line and the previous plane of a row (the
// mirrored blocks of a half-stored stencil, the gathered vector entries)
// were read by the same XCD a few workgroups earlier.  G < 0: each XCD one
// contiguous run of the range.  A bijection of [0, g).  A workgroup's
// partial sum stays at its dispatch position w, so launches that cover one
// system with g workgroups, alone or as a segment of a batched launch, add
// the same rows in the same order.
__host__ __device__ __forceinline__ int64_t xcd_perm(int64_t w, int64_t g, int32_t G) {
  constexpr int64_t kXcd = 8;
  if (G < 0) {
    const int64_t q = g / kXcd, rem = g % kXcd, x = w % kXcd;
    return x * q + (x < rem ? x : rem) + w / kXcd;
  }
  if (G <= 1) return w;
  const int64_t span = kXcd * G;
  if (w >= (g / span) * span) return w;
  return (w / span) * span + (w % kXcd) * G + (w / kXcd) % G;
}

int ensure_workspace(cpm_ctx* ctx, int64_t n);
// the context's second lane of solves (stream2, partials2, pinned2, the
// fork and join events), created on first use
int ensure_lane2(cpm_ctx* ctx);
int read_cg_state(cpm_ctx* ctx, int slot, CgState* out);

// y = A_valset x on the context stream (no sync).
int launch_spmv(cpm_ctx* ctx, const cpm_mat* A, int valset, const double* x, double* y);
// CG on the context stream; syncs to test convergence.  `hint` = expected
// iteration count (0 = unknown), used only to size launch batches.
// Device memory of one CG solve: vectors r, p, q (z: preconditioned CG),
// the reduction partials (kMaxPartials doubles), the grid-sum counter
// (zero between uses) and the CG state.
struct CgWork {
  double *r, *p, *q, *z;
  double* partials;
  unsigned* counter;
  CgState* st;
};
// Mp: preconditioner applied as z = Mp r (CuPy's cg with M), or nullptr.
// Wk: the workspace to use (nullptr: the context's, grown on demand).
int run_cg(cpm_ctx* ctx, const cpm_mat* A, const double* b, double* x, int use_x0, double tol, int maxiter,
           int hint, int32_t* iters, double* resid, const cpm_mat* Mp = nullptr, const CgWork* Wk = nullptr);
// Device address of the no-allocation CG scratch on the current device
// (looked up once per device; cg.hip).
int noalloc_scratch(int device, double** out);

// ---- device reductions (blockDim.x == kBlock) -------------------------------

template <int NV>
__device__ inline void wave_sum(double (&v)[NV]) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += __shfl_xor(v[k], o, 64);
}

// Sum over the workgroup; the result is valid in thread 0.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double* lds /* NV*4 */) {
  wave_sum<NV>(v);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[k * 4 + wid] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = ((lds[k * 4 + 0] + lds[k * 4 + 1]) + lds[k * 4 + 2]) + lds[k * 4 + 3];
  __syncthreads();
}

// Sum over a workgroup of NT threads; the result is valid in thread 0.  The
// waves' sums are added in wave order.
template <int NV, int NT>
__device__ inline void block_sum_t(double (&v)[NV], double* lds /* NV*NT/64 */) {
  constexpr int NW = NT / 64;
  wave_sum<NV>(v);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[k * NW + wid] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double t = lds[k * NW];
      for (int w = 1; w < NW; ++w) t += lds[k * NW + w];
      v[k] = t;
    }
  __syncthreads();
}

// Deterministic grid-wide sum: every workgroup stores its partial, the last
// one to arrive adds the partials in block order.  Returns true in the last
// workgroup, where thread 0 holds the totals in `v`.
template <int NV>
__device__ inline bool grid_sum(double (&v)[NV], double* partials, unsigned* counter) {
  __shared__ double lds[NV * 4];
  __shared__ int am_last;
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[k * gridDim.x + blockIdx.x] = v[k];
    __threadfence();
    unsigned t = atomicAdd(counter, 1u);
    am_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!am_last) return false;
  __threadfence();
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = 0.0;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += kBlock)
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] += __hip_atomic_load(&partials[k * gridDim.x + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  block_sum<NV>(s, lds);
  if (threadIdx.x == 0) {
    *counter = 0u;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = s[k];
  }
  return true;
}

// Workgroup partial sum stored at partials[k * gridDim.x + blockIdx.x] with a
// plain store.  A separate one-workgroup kernel adds the partials after the
// producing kernel has finished: this avoids a device-scope fence and an
// atomic per workgroup, which on a multi-XCD GPU cost an L2 write-back each.
template <int NV>
__device__ inline void block_partial(double (&v)[NV], double* partials) {
  __shared__ double lds[NV * 4];
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[k * gridDim.x + blockIdx.x] = v[k];
}

// Row sum of a CSR row with G lanes (G a power of two <= 64), lanes of a
// row contiguous within the wavefront; the sum is valid in every lane.
template <int G>
__device__ inline double group_sum(double s) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
  return s;
}

// CG scalars after a reduction (thread 0).  v[j]: the reduced value of
// column j; only the columns in use (live, or with an update requested for
// RED_RR_TRUE) are touched.
enum : int {
  RED_ALPHA = 0,
  RED_RR = 1,
  RED_RR_REL = 2,
  RED_RR_TRUE = 3,
  RED_RR_PCG = 4,
  RED_RR_PCG_REL = 5,   // preconditioned CG with reliable updates: residual norm, update request
  RED_RR_PCG_TRUE = 6,  // ... the FP64 residual replaced the recurrence residual
  RED_ALPHA_PCG_REL = 7, // alpha of the preconditioned CG whose products may be FP32
  RED_RHO = 8            // preconditioned CG (CuPy's recurrence): rho_old = rho, rho = r.z
};

template <int MODE>
__device__ inline bool cg_state_uses(const CgState& c) {
  return (MODE == RED_RR_TRUE || MODE == RED_RR_PCG_TRUE) ? c.upd != 0 : !c.done;
}

template <int K, int MODE>
__device__ inline void cg_state_update(CgState* st, const double (&v)[K], double delta) {
#pragma unroll
  for (int j = 0; j < K; ++j) {
    CgState& c = st[j];
    if (!cg_state_uses<MODE>(c)) continue;
    if (MODE == RED_ALPHA) {
      c.pq = v[j];
      c.alpha = c.rho / v[j];
    } else if (MODE == RED_RHO) {
      c.rho_old = c.rho;
      c.rho = v[j];
    } else if (MODE == RED_ALPHA_PCG_REL) {
      // p.q of an SPD operator is positive; an FP32 product that has lost
      // it (search direction in the operator's softest modes) gives no
      // step: x and r stay, the FP64 residual is taken and the remaining
      // products of the solve run in FP64
      c.pq = v[j];
      c.alpha = c.rho / v[j];
      if (!(v[j] > 0.0) || !(c.rho > 0.0) || !isfinite(c.alpha)) {
        c.alpha = 0.0;
        c.upd = 1;
        c.p64 = 1;
      }
    } else if (MODE == RED_RR) {
      // CuPy: rho = r.r; stop when ||r|| <= atol after the update
      c.rho_old = c.rho;
      c.rho = v[j];
      c.it += 1;
      c.resid = sqrt(v[j]);
      if (c.resid <= c.atol) {
        c.done = 1;
      } else if (c.it >= c.maxiter) {
        c.done = 1;
        c.fail = 1;
      }
    } else if (MODE == RED_RR_REL) {
      c.rho_old = c.rho;
      c.rho = v[j];
      c.it += 1;
      c.resid = sqrt(v[j]);
      if (c.resid < delta * c.maxr || c.resid <= c.atol || c.it >= c.maxiter) c.upd = 1;
    } else if (MODE == RED_RR_PCG) {
      // preconditioned CG: the residual norm only (rho = r.z comes later)
      c.it += 1;
      c.resid = sqrt(v[j]);
      if (c.resid <= c.atol) {
        c.done = 1;
      } else if (c.it >= c.maxiter || !isfinite(c.resid)) {
        c.done = 1;
        c.fail = 1;
      }
    } else if (MODE == RED_RR_PCG_REL) {
      // recurrence residual of a PCG whose operator products are inexact:
      // an FP64 residual is requested when it has dropped by delta since
      // the last one, or meets the tolerance, or at maxiter
      c.it += 1;
      c.resid = sqrt(v[j]);
      if (c.resid < delta * c.maxr || c.resid <= c.atol || c.it >= c.maxiter || c.it - c.it_upd >= kRelCycle ||
          !isfinite(c.resid))
        c.upd = 1;
    } else if (MODE == RED_RR_PCG_TRUE) {
      // c.resid still holds the recurrence residual and c.maxr the FP64
      // residual of the last replacement.  An FP64 residual more than 10x
      // above the recurrence's or not half the last one, or a cycle of
      // kRelCycle iterations that did not bring the recurrence residual
      // down by delta, means the FP32 products no longer carry the
      // refinement (seen on the anchored operators at tolerances near
      // 1e-13): the rest of the solve multiplies in FP64
      //
      // Once the products are FP64, a replacement that follows a recurrence
      // residual <= atol and finds the FP64 residual no lower than half the
      // last one has reached the attainable accuracy of FP64 products
      // (about eps64 ||A|| ||x||, 1e-13 of ||b|| on the anchored elasticity
      // operator at 64^3): the solve ends there, converged, as the
      // reference's test on the recurrence residual would.
      // (a replacement taken with another column's, upd = 2, comes before
      // this column's residual has dropped by delta: only the 10x test
      // applies to it)
      const double tr = sqrt(v[j]);
      const bool own = c.upd == 1;
      const bool stalled = own && tr > 0.5 * c.maxr;
      const bool floor = c.p64 && stalled && c.resid <= c.atol;
      if (tr > 10.0 * c.resid || stalled || (own && c.it - c.it_upd >= kRelCycle)) c.p64 = 1;
      c.it_upd = c.it;
      c.upd = 0;
      c.resid = tr;
      c.maxr = c.resid;
      if (c.resid <= c.atol || (floor && isfinite(tr))) {
        c.done = 1;
      } else if (c.it >= c.maxiter || !isfinite(c.resid)) {
        c.done = 1;
        c.fail = 1;
      }
    } else {  // RED_RR_TRUE: FP64 residual replaced the FP32 recurrence residual
      c.upd = 0;
      c.rho = v[j];
      c.resid = sqrt(v[j]);
      c.maxr = c.resid;
      if (c.resid <= c.atol) {
        c.done = 1;
      } else if (c.it >= c.maxiter) {
        c.done = 1;
        c.fail = 1;
      }
    }
  }
  if constexpr (K > 1 && (MODE == RED_RR_REL || MODE == RED_RR_PCG_REL)) {
    // one FP64 residual pass serves every column: when a live column asks
    // for one, the other live columns take it too (upd = 2), so that the
    // columns' requests do not each cost a pass of the FP64 operator
    bool any = false;
#pragma unroll
    for (int j = 0; j < K; ++j) any = any || (!st[j].done && st[j].upd == 1);
    if (any)
#pragma unroll
      for (int j = 0; j < K; ++j)
        if (!st[j].done && st[j].upd == 0) st[j].upd = 2;
  }
}

// One workgroup: add the partials of the preceding kernel (deterministic
// order) and advance the CG scalars.
// Launched with kRedBlock threads.
template <int K, int MODE>
__global__ void __launch_bounds__(kRedBlock) k_red(int nblk, const double* __restrict__ partials, CgState* st,
                                                   double delta) {
  bool any = false;
#pragma unroll
  for (int j = 0; j < K; ++j) any |= cg_state_uses<MODE>(st[j]);
  if (!any) return;
  __shared__ double lds[K * (kRedBlock / 64)];
  double v[K];
#pragma unroll
  for (int j = 0; j < K; ++j) v[j] = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kRedBlock)
#pragma unroll
    for (int j = 0; j < K; ++j) v[j] += partials[j * nblk + i];
  block_sum_t<K, kRedBlock>(v, lds);
  if (threadIdx.x == 0) cg_state_update<K, MODE>(st, v, delta);
}

}  // namespace cpm
