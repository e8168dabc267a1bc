// SpMV autotuner: times the launch schedules of an operator on the device
// and keeps the fastest.
//
// Candidates (each timed as the CG uses it: the SpMV with its fused dot
// partials followed by the one-workgroup reduction of those partials):
//   CSR  : lanes per row G in {2, 4, 8, 16, 32, 64} (= 64/G rows per
//          wavefront) x workgroup cap in {2048, 8192, 32768} (grid-stride
//          beyond the cap)
//   SELL : workgroup cap in {2048, 4096, 8192, 16384, 32768} for the FP32
//          and FP64 forms (64 rows per wavefront, one thread per row)
// The format of the standalone product y = A x (cpm_spmv on value set 0 of
// a scalar operator) is the faster of the tuned CSR and the tuned FP64 SELL
// schedules.  Results depend only on the operator, so a caller can tune one
// operator of a size class and apply the schedule to the others
// (cpm_mat_set_schedule).
#include <algorithm>
#include <vector>

#include "device.h"
#include "sell_dev.h"

namespace cpm {

template <int G>
__global__ void __launch_bounds__(kBlock) k_tune_csr(int32_t nrows, const int32_t* __restrict__ ptr,
                                                     const int32_t* __restrict__ idx, const double* __restrict__ val,
                                                     const double* __restrict__ p, double* __restrict__ q,
                                                     double* partials) {
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

template <int BS, class T, bool ST>
__global__ void __launch_bounds__(kBlock) k_tune_sell(int32_t nbrows, const int64_t* __restrict__ sptr,
                                                      const int32_t* __restrict__ col, const T* __restrict__ val,
                                                      const double* __restrict__ dg, const T* __restrict__ x,
                                                      T* __restrict__ y, double* partials, GridArg ga) {
  double acc[1] = {0.0};
  sell_rows<BS, 1, T, true, false, ST>(xcd_perm(blockIdx.x, gridDim.x, ga.xg), gridDim.x, nbrows, sptr, col, val, dg, x, nullptr, y, ga, acc);
  block_partial<1>(acc, partials);
}

__global__ void __launch_bounds__(kRedBlock) k_tune_red(int nblk, const double* __restrict__ partials,
                                                        double* out) {
  __shared__ double lds[kRedBlock / 64];
  double v[1] = {0.0};
  for (int i = threadIdx.x; i < nblk; i += kRedBlock) v[0] += partials[i];
  block_sum_t<1, kRedBlock>(v, lds);
  if (threadIdx.x == 0) out[0] = v[0];
}

namespace {

constexpr int kReps = 5;

// Average microseconds of `launch()` over kReps launches after one warm-up.
template <class F>
int time_us(cpm_ctx* ctx, hipEvent_t e0, hipEvent_t e1, F&& launch, float* us) {
  launch();
  CPM_HIP(hipGetLastError());
  CPM_HIP(hipEventRecord(e0, ctx->stream));
  for (int r = 0; r < kReps; ++r) launch();
  CPM_HIP(hipEventRecord(e1, ctx->stream));
  CPM_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  CPM_HIP(hipEventElapsedTime(&ms, e0, e1));
  *us = ms * 1e3f / kReps;
  return CPM_OK;
}

template <int G>
int grid_csr(const cpm_mat* A, int cap) {
  const int64_t g = ((int64_t)A->nrows + kBlock / G - 1) / (kBlock / G);
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

struct Buffers {
  void* x = nullptr;
  void* y = nullptr;
  ~Buffers() {
    if (x) (void)hipFree(x);
    if (y) (void)hipFree(y);
  }
};

}  // namespace

int tune_csr(cpm_ctx* ctx, cpm_mat* A, hipEvent_t e0, hipEvent_t e1, double* scratch, std::vector<cpm_tune_entry>& out) {
  Buffers b;
  CPM_HIP(hipMalloc(&b.x, sizeof(double) * std::max(1, A->ncols)));
  CPM_HIP(hipMalloc(&b.y, sizeof(double) * std::max(1, A->nrows)));
  CPM_HIP(hipMemsetAsync(b.x, 0, sizeof(double) * std::max(1, A->ncols), ctx->stream));
  const double* x = (const double*)b.x;
  double* y = (double*)b.y;
  float best = 1e30f;
  int best_g = A->lanes, best_cap = A->grid_cap;
  for (int g : {2, 4, 8, 16, 32, 64})
    for (int cap : {2048, 8192, 32768}) {
      int grid = 0;
      auto launch = [&]() {
        switch (g) {
#define CPM_TUNE_G(GG)                                                                                     \
  case GG:                                                                                                 \
    grid = grid_csr<GG>(A, cap);                                                                           \
    k_tune_csr<GG><<<grid, kBlock, 0, ctx->stream>>>(A->nrows, A->ptr, A->idx, A->val, x, y, ctx->partials); \
    break;
          CPM_TUNE_G(2)
          CPM_TUNE_G(4)
          CPM_TUNE_G(8)
          CPM_TUNE_G(16)
          CPM_TUNE_G(32)
          default:
            CPM_TUNE_G(64)
#undef CPM_TUNE_G
        }
        k_tune_red<<<1, kRedBlock, 0, ctx->stream>>>(grid, ctx->partials, scratch);
      };
      float us = 0.f;
      if (int rc = time_us(ctx, e0, e1, launch, &us)) return rc;
      out.push_back(cpm_tune_entry{CPM_FMT_CSR, g, cap, us});
      if (us < best) {
        best = us;
        best_g = g;
        best_cap = cap;
      }
    }
  A->lanes = best_g;
  A->grid_cap = best_cap;
  return CPM_OK;
}

template <int BS, class T>
static int tune_sell_t(cpm_ctx* ctx, cpm_mat* A, hipEvent_t e0, hipEvent_t e1, double* scratch,
                       std::vector<cpm_tune_entry>& out, float* best_us, int* best_cap) {
  Sell& S = A->sell;
  const int64_t n = (int64_t)S.nbrows * BS;
  Buffers b;
  CPM_HIP(hipMalloc(&b.x, sizeof(T) * n));
  CPM_HIP(hipMalloc(&b.y, sizeof(T) * n));
  CPM_HIP(hipMemsetAsync(b.x, 0, sizeof(T) * n, ctx->stream));
  const T* x = (const T*)b.x;
  T* y = (T*)b.y;
  const T* val;
  const double* dg;
  if constexpr (sizeof(T) == 8) {
    val = sell_v64(S);
    dg = S.dg64;
  } else {
    val = S.v32;
    dg = S.dgc;
  }
  const GridArg ga = sizeof(T) == 8 ? CPM_GRID_ARG64(S) : CPM_GRID_ARG32(S);
  const int saved = S.grid_cap;
  *best_us = 1e30f;
  *best_cap = saved;
  for (int cap : {2048, 4096, 8192, 16384, 32768}) {
    S.grid_cap = cap;
    const int grid = sell_grid(S);
    auto launch = [&]() {
      if (S.stencil)
        k_tune_sell<BS, T, true><<<grid, kBlock, 0, ctx->stream>>>(S.nbrows, S.sptr, S.col, val, dg, x, y,
                                                                   ctx->partials, ga);
      else
        k_tune_sell<BS, T, false><<<grid, kBlock, 0, ctx->stream>>>(S.nbrows, S.sptr, S.col, val, dg, x, y,
                                                                    ctx->partials, ga);
      k_tune_red<<<1, kRedBlock, 0, ctx->stream>>>(grid, ctx->partials, scratch);
    };
    float us = 0.f;
    int rc = time_us(ctx, e0, e1, launch, &us);
    if (rc) {
      S.grid_cap = saved;
      return rc;
    }
    out.push_back(cpm_tune_entry{sizeof(T) == 8 ? CPM_FMT_SELL64 : CPM_FMT_SELL32, 64, cap, us});
    if (us < *best_us) {
      *best_us = us;
      *best_cap = cap;
    }
  }
  S.grid_cap = saved;
  return CPM_OK;
}

}  // namespace cpm

using namespace cpm;

extern "C" int cpm_mat_autotune(cpm_mat* A, cpm_tune_entry* out, int max_out, int* n_out) {
  if (!A || !n_out) return fail(CPM_EINVAL, "cpm_mat_autotune: null argument");
  cpm_ctx* ctx = A->ctx;
  CPM_HIP(hipSetDevice(ctx->device));
  if (A->nrows == 0) {
    *n_out = 0;
    return CPM_OK;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  CPM_HIP(hipEventCreate(&e0));
  CPM_HIP(hipEventCreate(&e1));
  std::vector<cpm_tune_entry> res;
  double* scratch = ctx->scal + 200;
  int rc = tune_csr(ctx, A, e0, e1, scratch, res);
  float csr_best = 1e30f;
  for (const auto& r : res)
    if (r.lanes == A->lanes && r.grid_cap == A->grid_cap) csr_best = r.us;
  A->spmv_sell = 0;
  if (rc == CPM_OK && A->sell.sptr) {
    float us32 = 0.f, us64 = 0.f;
    int cap32 = 0, cap64 = 0;
    if (A->sell.bs == 3) {
      rc = tune_sell_t<3, float>(ctx, A, e0, e1, scratch, res, &us32, &cap32);
      if (rc == CPM_OK) rc = tune_sell_t<3, double>(ctx, A, e0, e1, scratch, res, &us64, &cap64);
    } else {
      rc = tune_sell_t<1, float>(ctx, A, e0, e1, scratch, res, &us32, &cap32);
      if (rc == CPM_OK) rc = tune_sell_t<1, double>(ctx, A, e0, e1, scratch, res, &us64, &cap64);
    }
    // one cap for both value types: the mixed solvers run the FP32 form
    // every iteration and the FP64 form only for residual updates
    if (rc == CPM_OK) A->sell.grid_cap = cap32;
    if (rc == CPM_OK && A->sell.bs == 1 && us64 < csr_best) A->spmv_sell = 1;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  *n_out = (int)res.size();
  if (out)
    for (int i = 0; i < std::min((int)res.size(), max_out); ++i) out[i] = res[i];
  return CPM_OK;
}

extern "C" int cpm_mat_set_schedule(cpm_mat* A, const cpm_schedule* s) {
  if (!A || !s) return fail(CPM_EINVAL, "cpm_mat_set_schedule: null argument");
  const int l = s->lanes;
  if (l < 1 || l > 64 || (l & (l - 1))) return fail(CPM_EINVAL, "cpm_mat_set_schedule: lanes must be 1..64, a power of 2");
  if (s->grid_cap < 1 || s->grid_cap > kMaxSpmvGrid || s->sell_grid_cap < 1 || s->sell_grid_cap > kMaxSpmvGrid)
    return fail(CPM_EINVAL, "cpm_mat_set_schedule: grid cap out of range");
  if (s->spmv_sell && !(A->sell.sptr && A->sell.bs == 1))
    return fail(CPM_EINVAL, "cpm_mat_set_schedule: the SELL product needs a scalar SELL form");
  A->lanes = l;
  A->grid_cap = s->grid_cap;
  A->sell.grid_cap = s->sell_grid_cap;
  A->spmv_sell = s->spmv_sell ? 1 : 0;
  return CPM_OK;
}

extern "C" int cpm_mat_get_schedule(const cpm_mat* A, cpm_schedule* s) {
  if (!A || !s) return fail(CPM_EINVAL, "cpm_mat_get_schedule: null argument");
  s->lanes = A->lanes;
  s->grid_cap = A->grid_cap;
  s->sell_grid_cap = A->sell.sptr ? A->sell.grid_cap : 0;
  s->spmv_sell = A->spmv_sell;
  return CPM_OK;
}
