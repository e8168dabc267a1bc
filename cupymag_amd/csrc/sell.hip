// Sliced-ELL operators and the multi-RHS / mixed-precision CG.
//
// SELL-64: rows are grouped in slices of 64 (one wavefront); within a slice
// every row owns `w` slots (w = longest row of the slice, short rows padded
// with zero values pointing at column 0), stored slot-major with the 64
// lanes contiguous, so each load of a wavefront is one 512 B (FP64) or
// 256 B (FP32) coalesced segment.  One thread per row, no shuffles.
//
// The diagonal block of every row is kept apart, in FP64, for both the FP64
// and the FP32 form.  In the FP32 form it is compensated so that the row sums
// of each displacement component (the action on rigid translations, resp.
// on constants for scalar operators) equal those of the FP64 operator: the
// rounding of the off-diagonal values to FP32 is added back to the diagonal.
// The operators are anchored at one node only, so their smallest eigenvalue
// (the near-translation mode) scales like 1/n^3; without the compensation
// the FP32 rounding of the row sums is of that size at 128^3 and the FP32
// solves no longer improve that mode.
//
// Block size BS = 3 stores the 3x3 blocks of the elasticity operator
// (assemble_elasticity.py:389-398) with one column index per block and the
// vectors interleaved node-major (3i + d) instead of the reference's
// [ux; uy; uz]; the solver permutes in and out around the solve.
//
// The CG here is the same recurrence as cg.hip (CuPy's cupyx cg: rho = r.r,
// p = r + (rho/rho_old) p, alpha = rho/(p.q), stop when ||r|| <= tol*||b||
// after the update) for K right-hand sides that share the operator: every
// column keeps its own scalars and stops independently, and the operator is
// streamed from HBM once per iteration for all of them (the three m
// components of Micromagnetics.py:221-223 and :281-290 share A1 / A2).
//
// The mixed-precision solver runs that CG with FP32 values and vectors (dot
// products accumulated in FP64) and refines in FP64 by reliable updates: the
// FP32 correction is periodically folded into the FP64 solution and the
// residual replaced by the FP64 b - A x, until that residual meets tol*||b||.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <hip/hip_ext.h>

#include "device.h"
#include "sell.h"
#include "sell_dev.h"

namespace cpm {

// ---------------------------------------------------------------- building

int sell_build(cpm_ctx* ctx, Sell& S, int bs, int32_t nbrows, const int64_t* indptr, const int32_t* indices,
               const double* values, int64_t n_csr_rows, const int32_t* dims) {
  // Block-row b covers CSR rows {d*nbrows + b : d < bs} (bs = 3: the SoA DOF
  // order of the elasticity operator).  Its block columns are the columns
  // < nbrows of CSR row b; entry (d, e) of block (b, c) sits at CSR row
  // d*nbrows + b, column e*nbrows + c.
  if (bs != 1 && bs != 3) return fail(CPM_EINVAL, "SELL block size must be 1 or 3");
  if ((int64_t)bs * nbrows != n_csr_rows) return fail(CPM_EINVAL, "SELL: rows do not match the block size");
  const int64_t nsl = (nbrows + 63) / 64;
  std::vector<int64_t> sptr(nsl + 1, 0);
  std::vector<int32_t> blen(nbrows);
  std::vector<int32_t> slots(nbrows);
  for (int32_t b = 0; b < nbrows; ++b) {
    int64_t len = 0, diag = 0;
    for (int64_t p = indptr[b]; p < indptr[b + 1]; ++p) {
      len += indices[p] < nbrows;
      diag += indices[p] == b;
    }
    blen[b] = (int32_t)len;
    slots[b] = (int32_t)(len - diag);
    if (bs == 3)
      for (int d = 0; d < 3; ++d)
        if (indptr[d * nbrows + b + 1] - indptr[d * nbrows + b] != 3 * len)
          return fail(CPM_EINVAL, "SELL: block row is not a full 3x3 block pattern");
  }
  // Index-free form: every off-diagonal block column is a neighbour of the
  // row on the periodic dims[0] x dims[1] x dims[2] grid at one of <= 32
  // fixed offsets in {-1,0,1}^3.  Otherwise (or without dims) keep indices.
  std::vector<int32_t> codes;  // packed offsets, sorted
  std::vector<int8_t> slot_of;  // per CSR entry of row segment 0: stencil slot (-1 diagonal / skipped)
  bool stencil = dims && dims[0] >= 3 && dims[1] >= 3 && dims[2] >= 3 &&
                 (int64_t)dims[0] * dims[1] * dims[2] == nbrows;
  auto code_of = [&](int32_t b, int32_t c, int32_t* code) {
    const int32_t nyz = dims[1] * dims[2];
    const int32_t bi = b / nyz, bj = (b / dims[2]) % dims[1], bk = b % dims[2];
    const int32_t ci = c / nyz, cj = (c / dims[2]) % dims[1], ck = c % dims[2];
    int32_t d[3] = {ci - bi, cj - bj, ck - bk};
    for (int a = 0; a < 3; ++a) {
      if (d[a] > dims[a] / 2) d[a] -= dims[a];
      if (d[a] < -(dims[a] / 2)) d[a] += dims[a];
      if (d[a] < -1 || d[a] > 1) return false;
    }
    *code = (d[0] + 1) | ((d[1] + 1) << 2) | ((d[2] + 1) << 4);
    return true;
  };
  if (stencil) {
    std::vector<char> seen(64, 0);
    for (int32_t b = 0; b < nbrows && stencil; ++b)
      for (int64_t p = indptr[b]; p < indptr[b + 1]; ++p) {
        const int32_t c = indices[p];
        if (c >= nbrows || c == b) continue;
        int32_t code;
        if (!code_of(b, c, &code)) {
          stencil = false;
          break;
        }
        seen[code] = 1;
      }
    for (int k = 0; k < 64 && stencil; ++k)
      if (seen[k]) codes.push_back(k);
    if (codes.size() > 32) stencil = false;
  }
  std::vector<int32_t> slot_index(64, -1);
  for (size_t k = 0; k < codes.size(); ++k) slot_index[codes[k]] = (int32_t)k;
  for (int64_t s = 0; s < nsl; ++s) {
    int32_t w = 0;
    for (int64_t b = s * 64; b < std::min<int64_t>(nbrows, s * 64 + 64); ++b) w = std::max(w, slots[b]);
    if (stencil) w = (int32_t)codes.size();
    sptr[s + 1] = sptr[s] + w;
  }
  const int64_t nslot = sptr[nsl];
  const int bb = bs * bs;
  std::vector<int32_t> col(nslot * 64, 0);
  std::vector<double> val(nslot * 64 * bb, 0.0);
  std::vector<double> dg(std::max<int64_t>(1, (int64_t)nbrows * bb), 0.0), dgc(dg.size(), 0.0);
  for (int32_t b = 0; b < nbrows; ++b) {
    const int64_t s = b / 64, l = b % 64;
    const int64_t len = blen[b];
    double comp[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // sum over off-diagonal blocks of (a - fp32(a))
    int64_t j = 0;
    for (int64_t p = indptr[b]; p < indptr[b + 1]; ++p) {
      const int32_t c = indices[p];
      if (c >= nbrows) continue;
      const int64_t k = p - indptr[b];  // position of c inside the row segment
      double blk[9];
      for (int d = 0; d < bs; ++d)
        for (int e = 0; e < bs; ++e) {
          const int64_t q = (bs == 1) ? p : indptr[d * nbrows + b] + e * len + k;
          if (indices[q] != e * nbrows + c) return fail(CPM_EINVAL, "SELL: unexpected 3x3 block layout");
          blk[d * bs + e] = values[q];
        }
      if (c == b) {
        for (int t = 0; t < bb; ++t) dg[(int64_t)b * bb + t] = blk[t];
        continue;
      }
      int64_t slot = j;
      if (stencil) {
        int32_t code = 0;
        code_of(b, c, &code);
        slot = slot_index[code];
      }
      col[(sptr[s] + slot) * 64 + l] = c;
      for (int t = 0; t < bb; ++t) {
        val[((sptr[s] + slot) * bb + t) * 64 + l] = blk[t];
        comp[t] += blk[t] - (double)(float)blk[t];
      }
      ++j;
    }
    for (int t = 0; t < bb; ++t) dgc[(int64_t)b * bb + t] = dg[(int64_t)b * bb + t] + comp[t];
  }
  S.bs = bs;
  S.nbrows = nbrows;
  S.stencil = stencil ? 1 : 0;
  S.nst = stencil ? (int32_t)codes.size() : 0;
  S.kuhn = 0;
  if (stencil && codes.size() == (size_t)kKuhnSlots) {
    S.kuhn = 1;
    for (int k = 0; k < kKuhnSlots; ++k) S.kuhn &= codes[k] == kKuhnCodes[k] ? 1 : 0;
  }
  if (stencil)
    for (int a = 0; a < 3; ++a) S.dims[a] = dims[a];
  S.xg = stencil ? sell_xcd_group(dims) : 0;
  auto shift = [](int32_t d) {
    int k = 0;
    while (k < 30 && (1 << k) < d) ++k;
    return (d > 0 && (1 << k) == d) ? k + 1 : 0;
  };
  S.shz = stencil ? shift(dims[2]) : 0;
  S.shy = stencil ? shift(dims[1]) : 0;
  S.nslices = (int32_t)nsl;
  S.nslot = nslot;
  // FP32 form: half of a slot-symmetric Kuhn stencil (every block once, the
  // mirrored neighbour reads it transposed), else every slot
  S.half32 = 0;
  std::vector<float> v32;
  if (S.kuhn) {  // a stencil form: dims given, each >= 3
    const GridArg ga{dims[0], dims[1], dims[2], kKuhnSlots, nullptr, 1};
    if (kuhn_slot_symmetric(bs, nbrows, ga, val.data())) {
      S.half32 = 1;
      v32.assign(nsl * 7 * bb * 64, 0.0f);
      kuhn_half32(bs, nbrows, nsl, ga, val.data(), dg.data(), v32.data(), dgc.data());
    }
  }
  // FP64 half form when the mirrored blocks agree to rounding (the
  // elasticity blocks K_ab and K_ba^T are sums of the same products in
  // different orders: they differ in the last bits; scalar operators are
  // symmetric exactly, and their half products equal the full ones)
  S.half64 = 0;
  std::vector<double> v64h;
  if (S.half32) {
    const GridArg ga{dims[0], dims[1], dims[2], kKuhnSlots, nullptr, 1};
    if (kuhn_slot_symmetric(bs, nbrows, ga, val.data(), 1e-14)) {
      S.half64 = 1;
      v64h.resize(nsl * 7 * bb * 64);
      for (int64_t s = 0; s < nsl; ++s)
        std::memcpy(&v64h[s * 7 * bb * 64], &val[(s * kKuhnSlots + 7) * bb * 64], sizeof(double) * 7 * bb * 64);
    }
  }
  if (!S.half32) {
    v32.resize(val.size());
    for (size_t i = 0; i < val.size(); ++i) v32[i] = (float)val[i];
  }
  hipStream_t st = ctx->stream;
  CPM_HIP(hipMalloc(&S.sptr, sizeof(int64_t) * (nsl + 1)));
  if (stencil) {
    CPM_HIP(hipMalloc(&S.soff, sizeof(int32_t) * std::max<size_t>(1, codes.size())));
    CPM_HIP(hipMemcpyAsync(S.soff, codes.data(), sizeof(int32_t) * codes.size(), hipMemcpyHostToDevice, st));
  }
  CPM_HIP(hipMalloc(&S.col, sizeof(int32_t) * (stencil ? 1 : std::max<int64_t>(1, col.size()))));
  CPM_HIP(hipMalloc(&S.v64, sizeof(double) * std::max<size_t>(1, val.size())));
  CPM_HIP(hipMalloc(&S.v32, sizeof(float) * std::max<size_t>(1, v32.size())));
  CPM_HIP(hipMalloc(&S.dg64, sizeof(double) * dg.size()));
  CPM_HIP(hipMalloc(&S.dgc, sizeof(double) * dgc.size()));
  CPM_HIP(hipMemcpyAsync(S.dg64, dg.data(), sizeof(double) * dg.size(), hipMemcpyHostToDevice, st));
  CPM_HIP(hipMemcpyAsync(S.dgc, dgc.data(), sizeof(double) * dgc.size(), hipMemcpyHostToDevice, st));
  CPM_HIP(hipMemcpyAsync(S.sptr, sptr.data(), sizeof(int64_t) * (nsl + 1), hipMemcpyHostToDevice, st));
  if (!stencil) CPM_HIP(hipMemcpyAsync(S.col, col.data(), sizeof(int32_t) * col.size(), hipMemcpyHostToDevice, st));
  CPM_HIP(hipMemcpyAsync(S.v64, val.data(), sizeof(double) * val.size(), hipMemcpyHostToDevice, st));
  CPM_HIP(hipMemcpyAsync(S.v32, v32.data(), sizeof(float) * v32.size(), hipMemcpyHostToDevice, st));
  if (S.half64) {
    CPM_HIP(hipMalloc(&S.v64h, sizeof(double) * v64h.size()));
    CPM_HIP(hipMemcpyAsync(S.v64h, v64h.data(), sizeof(double) * v64h.size(), hipMemcpyHostToDevice, st));
  }
  CPM_HIP(hipStreamSynchronize(st));
  return CPM_OK;
}

void sell_free(Sell& S) {
  for (void* p : {(void*)S.sptr, (void*)S.col, (void*)S.v64, (void*)S.v64h, (void*)S.v32, (void*)S.dg64, (void*)S.dgc,
                  (void*)S.soff})
    if (p) (void)hipFree(p);
  S = Sell{};
}

// ---------------------------------------------------------------- kernels

// One system per launch: the bodies in sell_dev.h over the whole grid.
template <int BS, int K, class T, bool CG, bool GATE = false, bool ADDX32 = false, bool ST = false>
__global__ void __launch_bounds__(kBlock) k_sell_spmv(int32_t nbrows, const int64_t* __restrict__ sptr,
                                                      const int32_t* __restrict__ col, const T* __restrict__ val,
                                                      const double* __restrict__ dg, const T* __restrict__ x,
                                                      const float* __restrict__ x32, T* __restrict__ y,
                                                      const CgState* st, double* partials, GridArg ga) {
  // y = A x (ADDX32: A (x + x32)); CG: per-column partials of x.y.  GATE:
  // run only when some column requested an FP64 residual update.
  bool any = false;
#pragma unroll
  for (int j = 0; j < K; ++j) any |= CG ? !st[j].done : (GATE ? st[j].upd != 0 : true);
  if (!any) return;
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) acc[j] = 0.0;
  sell_rows<BS, K, T, CG, ADDX32, ST>(xcd_perm(blockIdx.x, gridDim.x, ga.xg), gridDim.x, nbrows, sptr, col, val, dg, x, x32, y, ga, acc);
  if (CG) block_partial<K>(acc, partials);
}

template <int K, class T>
__global__ void __launch_bounds__(kBlock) k_mcg_pupd(int64_t n, const T* __restrict__ r, T* __restrict__ p,
                                                     const CgState* st) {
  pupd_rows<K, T>(blockIdx.x, gridDim.x, n, r, p, st);
}

template <int K, class T>
__global__ void __launch_bounds__(kBlock) k_mcg_xr(int64_t n, T* __restrict__ x, T* __restrict__ r,
                                                   const T* __restrict__ p, const T* __restrict__ q, const CgState* st,
                                                   double* partials) {
  xr_rows<K, T>(blockIdx.x, gridDim.x, n, x, r, p, q, st, partials);
}

// CG start (FP64 CG, or REL: the mixed solver's FP64 residual + FP32
// copies); the column states are set by the last workgroup.
template <int K, bool REL>
__global__ void __launch_bounds__(kBlock) k_mcg_init(int64_t n, const double* __restrict__ b, double* __restrict__ x,
                                                     const double* __restrict__ y, double* __restrict__ r,
                                                     float* __restrict__ r32, float* __restrict__ x32, int use_x0,
                                                     double tol, int32_t maxiter, CgState* st, double* partials,
                                                     unsigned* counter) {
  double acc[2 * K];
#pragma unroll
  for (int j = 0; j < 2 * K; ++j) acc[j] = 0.0;
  init_rows<K, REL>(blockIdx.x, gridDim.x, n, b, x, y, r, r32, x32, use_x0, acc);
  if (grid_sum<2 * K>(acc, partials, counter) && threadIdx.x == 0) cg_state_init<K, REL>(st, acc, tol, maxiter);
}

template <int K>
__global__ void __launch_bounds__(kBlock) k_mcg_zero_b(int64_t n, double* __restrict__ x, const CgState* st) {
  zero_b_rows<K>(blockIdx.x, gridDim.x, n, x, st);
}

template <int K>
__global__ void __launch_bounds__(kBlock) k_rel_r(int64_t n, const double* __restrict__ b, const double* __restrict__ y,
                                                  double* __restrict__ x, float* __restrict__ x32,
                                                  double* __restrict__ r64, float* __restrict__ r32,
                                                  const CgState* st, double* partials) {
  rel_r_rows<K>(blockIdx.x, gridDim.x, n, b, y, x, x32, r64, r32, st, partials);
}

// ---------------------------------------------------------------- dispatch

// one thread per (block) row, up to A.grid_cap (<= kMaxSpmvGrid) workgroups
// (3 partials each fit the partials buffer)
int sell_xcd_group(const int32_t dims[3]) {
  const int64_t plane = (int64_t)dims[1] * dims[2];
  if (plane % kBlock) return 1;
  int g = 1;
  while (g < 64 && (plane / kBlock) % (8LL * 2 * g) == 0) g *= 2;
  return g;
}

int sell_grid(const Sell& A) {
  int64_t g = ((((int64_t)A.nbrows + 63) & ~63) + kBlock - 1) / kBlock;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, A.grid_cap));
}
static_assert(3 * kMaxSpmvGrid <= kMaxPartials, "partials buffer too small for the SELL grid");

// ev (profiling, may be null): start and stop events taken by the dispatch
// itself (hipExtLaunchKernelGGL), so they time the kernel alone
template <int BS, int K, class T, bool CG, bool GATE = false, bool ADDX32 = false>
static int spmv_t(cpm_ctx* ctx, const Sell& A, const T* x, const float* x32, T* y, hipEvent_t* ev = nullptr) {
  const T* val;
  const double* dg;
  if constexpr (sizeof(T) == 8) {
    val = sell_v64(A);
    dg = A.dg64;
  } else {
    val = A.v32;
    dg = A.dgc;
  }
  const GridArg ga = sizeof(T) == 8 ? CPM_GRID_ARG64(A) : CPM_GRID_ARG32(A);
  const dim3 grid(sell_grid(A)), block(kBlock);
  if (ev) {
    if (A.stencil)
      hipExtLaunchKernelGGL(k_sell_spmv<BS, K, T, CG, GATE, ADDX32, true>, grid, block, 0, ctx->stream, ev[0], ev[1],
                            0, A.nbrows, (const int64_t*)A.sptr, (const int32_t*)A.col, val, dg, x, x32, y,
                            (CgState*)ctx->cg, ctx->partials, ga);
    else
      hipExtLaunchKernelGGL(k_sell_spmv<BS, K, T, CG, GATE, ADDX32, false>, grid, block, 0, ctx->stream, ev[0],
                            ev[1], 0, A.nbrows, (const int64_t*)A.sptr, (const int32_t*)A.col, val, dg, x, x32, y,
                            (CgState*)ctx->cg, ctx->partials, ga);
  } else if (A.stencil) {
    k_sell_spmv<BS, K, T, CG, GATE, ADDX32, true><<<grid, block, 0, ctx->stream>>>(
        A.nbrows, A.sptr, A.col, val, dg, x, x32, y, ctx->cg, ctx->partials, ga);
  } else {
    k_sell_spmv<BS, K, T, CG, GATE, ADDX32, false><<<grid, block, 0, ctx->stream>>>(
        A.nbrows, A.sptr, A.col, val, dg, x, x32, y, ctx->cg, ctx->partials, ga);
  }
  CPM_HIP(hipGetLastError());
  return CPM_OK;
}

template <class T>
static int launch_sell(cpm_ctx* ctx, const Sell& A, int K, const T* x, T* y) {
  if (A.bs == 1 && K == 1) return spmv_t<1, 1, T, false>(ctx, A, x, nullptr, y);
  if (A.bs == 1 && K == 3) return spmv_t<1, 3, T, false>(ctx, A, x, nullptr, y);
  if (A.bs == 3 && K == 1) return spmv_t<3, 1, T, false>(ctx, A, x, nullptr, y);
  return fail(CPM_EINVAL, "SELL: unsupported (block size, rhs count)");
}

template <int K, int MODE>
static void reduce_t(cpm_ctx* ctx, int nblk, double delta = 0.0) {
  k_red<K, MODE><<<1, kRedBlock, 0, ctx->stream>>>(nblk, ctx->partials, ctx->cg, delta);
}

static int read_states(cpm_ctx* ctx, int K, CgState* h) {
  CPM_HIP(hipMemcpyAsync(ctx->pinned, ctx->cg, K * sizeof(CgState), hipMemcpyDeviceToHost, ctx->stream));
  CPM_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(h, ctx->pinned, K * sizeof(CgState));
  return CPM_OK;
}

// Host side of a CG run: enqueue `issue(todo)` iteration batches, read the
// column states back between batches, time the SpMV launches when
// profiling is on.  Returns after every column is done or maxiter is spent.
template <int K, class Issue>
static int drive(cpm_ctx* ctx, cpm_mat* M, int maxiter, int hint, Issue&& issue, int32_t* iters, double* resid,
                 int* failed) {
  CgState h[K] = {};
  if (int rc = read_states(ctx, K, h)) return rc;
  int batch = hint > 2 ? hint - 1 : 8;
  const bool prof = M && prof_on(ctx, M);
  const int max_batch = prof ? 256 : 4096;
  int issued = 0;
  bool all_done = true;
  for (int j = 0; j < K; ++j) all_done &= h[j].done != 0;
  while (!all_done && issued <= maxiter) {
    const int todo = std::max(1, std::min(std::min(batch, max_batch), maxiter - issued + 1));
    int it_before = 0;
    for (int j = 0; j < K; ++j) it_before = std::max(it_before, h[j].it);
    if (int rc = issue(todo)) return rc;
    issued += todo;
    if (int rc = read_states(ctx, K, h)) return rc;
    int it_max = 0;
    all_done = true;
    for (int j = 0; j < K; ++j) {
      it_max = std::max(it_max, h[j].it);
      all_done &= h[j].done != 0;
    }
    if (prof) {
      for (int k = 0; k < it_max - it_before && k < todo; ++k) {
        float ms = 0.f;
        CPM_HIP(hipEventElapsedTime(&ms, ctx->ev[2 * k], ctx->ev[2 * k + 1]));
        M->prof_ms += ms;
        M->prof_launches += 1;
      }
    }
    // as the multigrid PCG: one iteration per check up to two past the
    // previous solve's count, then a quarter of the count so far
    batch = (hint > 2 && it_max < hint + 2) ? 1 : std::max(2, std::min(64, it_max / 4 + 2));
  }
  *failed = 0;
  for (int j = 0; j < K; ++j) {
    iters[j] = h[j].it;
    resid[j] = h[j].resid;
    if (h[j].fail || !h[j].done) *failed = 1;
  }
  return CPM_OK;
}

// FP64 CG on the SELL form, K interleaved columns (the cupyx cg recurrence).
template <int BS, int K>
static int cg64_loop(cpm_ctx* ctx, cpm_mat* M, const double* b, double* x, int use_x0, double tol, int maxiter,
                     int hint, int32_t* iters, double* resid, int* failed) {
  const Sell& A = M->sell;
  const int64_t n = (int64_t)A.nbrows * BS * K;
  const int g = grid_for(n), gs = sell_grid(A);
  hipStream_t s = ctx->stream;
  CgState* st = ctx->cg;
  if (use_x0)
    if (int rc = spmv_t<BS, K, double, false>(ctx, A, x, nullptr, ctx->q)) return rc;
  k_mcg_init<K, false><<<g, kBlock, 0, s>>>(n, b, x, ctx->q, ctx->r, nullptr, nullptr, use_x0, tol, maxiter, st,
                                            ctx->partials, ctx->counter);
  k_mcg_zero_b<K><<<g, kBlock, 0, s>>>(n, x, st);
  CPM_HIP(hipGetLastError());
  auto issue = [&](int todo) -> int {
    for (int k = 0; k < todo; ++k) {
      k_mcg_pupd<K, double><<<g, kBlock, 0, s>>>(n, ctx->r, ctx->p, st);
      if (int rc = spmv_t<BS, K, double, true>(ctx, A, ctx->p, nullptr, ctx->q,
                                               prof_on(ctx, M) ? &ctx->ev[2 * k] : nullptr))
        return rc;
      reduce_t<K, RED_ALPHA>(ctx, gs);
      k_mcg_xr<K, double><<<g, kBlock, 0, s>>>(n, x, ctx->r, ctx->p, ctx->q, st, ctx->partials);
      reduce_t<K, RED_RR>(ctx, g);
    }
    CPM_HIP(hipGetLastError());
    return CPM_OK;
  };
  return drive<K>(ctx, M, maxiter, hint, issue, iters, resid, failed);
}

int sell_cg64(cpm_ctx* ctx, cpm_mat* M, int K, const double* b, double* x, int use_x0, double tol, int maxiter,
              int hint, int32_t* iters, double* resid, int* failed) {
  const Sell& A = M->sell;
  const int64_t n = (int64_t)A.nbrows * A.bs * K;
  if (int rc = ensure_workspace(ctx, n)) return rc;
  if (A.bs == 1 && K == 1) return cg64_loop<1, 1>(ctx, M, b, x, use_x0, tol, maxiter, hint, iters, resid, failed);
  if (A.bs == 1 && K == 3) return cg64_loop<1, 3>(ctx, M, b, x, use_x0, tol, maxiter, hint, iters, resid, failed);
  if (A.bs == 3 && K == 1) return cg64_loop<3, 1>(ctx, M, b, x, use_x0, tol, maxiter, hint, iters, resid, failed);
  return fail(CPM_EINVAL, "SELL CG: unsupported (block size, rhs count)");
}

int ensure_workspace32(cpm_ctx* ctx, int64_t n) {
  if (n <= ctx->ws32_len) return CPM_OK;
  for (void** b : {(void**)&ctx->r32, (void**)&ctx->p32, (void**)&ctx->q32, (void**)&ctx->x32, (void**)&ctx->b32,
                   (void**)&ctx->y64})
    if (*b) {
      CPM_HIP(hipFree(*b));
      *b = nullptr;
    }
  CPM_HIP(hipMalloc(&ctx->r32, sizeof(float) * n));
  CPM_HIP(hipMalloc(&ctx->p32, sizeof(float) * n));
  CPM_HIP(hipMalloc(&ctx->q32, sizeof(float) * n));
  CPM_HIP(hipMalloc(&ctx->x32, sizeof(float) * n));
  CPM_HIP(hipMalloc(&ctx->b32, sizeof(float) * n));
  CPM_HIP(hipMalloc(&ctx->y64, sizeof(double) * n));
  ctx->ws32_len = n;
  return CPM_OK;
}

// ---- mixed precision with reliable updates ----------------------------------
//
// FP32 CG recurrence (x32, r32, p32) on the FP32 operator; whenever the
// recurrence residual has dropped by `delta` since the last update (or meets
// the tolerance, or the iteration budget ends) the FP32 correction is folded
// into the FP64 solution and the residual is replaced by the true FP64
// residual b - A x64.  The search direction is kept across updates, so the
// iteration count stays close to that of FP64 CG; convergence is declared
// on the FP64 residual only.

template <int BS, int K>
static int rel_loop(cpm_ctx* ctx, cpm_mat* M, const double* b, double* x, int use_x0, double tol, int maxiter,
                    double delta, int hint, int32_t* iters, double* resid, int* failed) {
  const Sell& A = M->sell;
  const int64_t n = (int64_t)A.nbrows * BS * K;
  const int g = grid_for(n), gs = sell_grid(A);
  hipStream_t s = ctx->stream;
  CgState* st = ctx->cg;
  double* y64 = ctx->y64;
  if (use_x0)
    if (int rc = spmv_t<BS, K, double, false>(ctx, A, x, nullptr, y64)) return rc;
  k_mcg_init<K, true><<<g, kBlock, 0, s>>>(n, b, x, y64, ctx->r, ctx->r32, ctx->x32, use_x0, tol, maxiter, st,
                                           ctx->partials, ctx->counter);
  k_mcg_zero_b<K><<<g, kBlock, 0, s>>>(n, x, st);
  CPM_HIP(hipGetLastError());
  auto issue = [&](int todo) -> int {
    for (int k = 0; k < todo; ++k) {
      k_mcg_pupd<K, float><<<g, kBlock, 0, s>>>(n, ctx->r32, ctx->p32, st);
      if (int rc = spmv_t<BS, K, float, true>(ctx, A, ctx->p32, nullptr, ctx->q32,
                                              prof_on(ctx, M) ? &ctx->ev[2 * k] : nullptr))
        return rc;
      reduce_t<K, RED_ALPHA>(ctx, gs);
      k_mcg_xr<K, float><<<g, kBlock, 0, s>>>(n, ctx->x32, ctx->r32, ctx->p32, ctx->q32, st, ctx->partials);
      reduce_t<K, RED_RR_REL>(ctx, g, delta);
      // reliable update (runs only when a column asked for it): y = A (x64 + x32)
      if (int rc = spmv_t<BS, K, double, false, true, true>(ctx, A, x, ctx->x32, y64)) return rc;
      k_rel_r<K><<<g, kBlock, 0, s>>>(n, b, y64, x, ctx->x32, ctx->r, ctx->r32, st, ctx->partials);
      reduce_t<K, RED_RR_TRUE>(ctx, g);
    }
    CPM_HIP(hipGetLastError());
    return CPM_OK;
  };
  return drive<K>(ctx, M, maxiter, hint, issue, iters, resid, failed);
}

int sell_cg_mixed(cpm_ctx* ctx, cpm_mat* M, int K, const double* b, double* x, int use_x0, double tol, int maxiter,
                  double delta, int hint, int32_t* iters, double* resid, int* failed, int32_t* sweeps) {
  const Sell& A = M->sell;
  const int64_t n = (int64_t)A.nbrows * A.bs * K;
  if (int rc = ensure_workspace(ctx, n)) return rc;
  if (int rc = ensure_workspace32(ctx, n)) return rc;
  if (sweeps) *sweeps = 0;
  if (A.bs == 1 && K == 1) return rel_loop<1, 1>(ctx, M, b, x, use_x0, tol, maxiter, delta, hint, iters, resid, failed);
  if (A.bs == 1 && K == 3) return rel_loop<1, 3>(ctx, M, b, x, use_x0, tol, maxiter, delta, hint, iters, resid, failed);
  if (A.bs == 3 && K == 1) return rel_loop<3, 1>(ctx, M, b, x, use_x0, tol, maxiter, delta, hint, iters, resid, failed);
  return fail(CPM_EINVAL, "mixed CG: unsupported (block size, rhs count)");
}

// interleave K columns (col-major blocks of n) <-> [i*K + j]; BS = 3 SoA <-> node-major
__global__ void __launch_bounds__(kBlock) k_interleave(int64_t n, int K, const double* __restrict__ src,
                                                       double* __restrict__ dst, int inverse) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n * K; i += stride) {
    const int64_t node = i / K, j = i % K;
    if (!inverse) dst[i] = src[j * n + node];
    else dst[j * n + node] = src[i];
  }
}

int interleave(cpm_ctx* ctx, int64_t n, int K, const double* src, double* dst, int inverse) {
  k_interleave<<<grid_for(n * K), kBlock, 0, ctx->stream>>>(n, K, src, dst, inverse);
  CPM_HIP(hipGetLastError());
  return CPM_OK;
}

int sell_spmv64(cpm_ctx* ctx, const cpm_mat* M, int K, const double* x, double* y) {
  return launch_sell<double>(ctx, M->sell, K, x, y);
}

}  // namespace cpm

// ---------------------------------------------------------------- C ABI

extern "C" int cpm_mat_add_sell(cpm_mat* A, int bs, const int64_t* indptr, const int32_t* indices,
                                const double* values, const int32_t* grid_dims) {
  if (!A || !indptr || !indices || !values) return cpm::fail(CPM_EINVAL, "cpm_mat_add_sell: null argument");
  if (A->nrows != A->ncols) return cpm::fail(CPM_EINVAL, "cpm_mat_add_sell: operator must be square");
  if (bs != 1 && bs != 3) return cpm::fail(CPM_EINVAL, "cpm_mat_add_sell: block size must be 1 or 3");
  if (A->nrows % bs) return cpm::fail(CPM_EINVAL, "cpm_mat_add_sell: rows not divisible by the block size");
  if (indptr[A->nrows] != A->nnz) return cpm::fail(CPM_EINVAL, "cpm_mat_add_sell: pattern differs from the operator");
  cpm::sell_free(A->sell);
  return cpm::sell_build(A->ctx, A->sell, bs, A->nrows / bs, indptr, indices, values, A->nrows, grid_dims);
}

extern "C" int cpm_mat_sell_info(const cpm_mat* A, int* bs, int* stencil, int64_t* slots) {
  if (!A) return cpm::fail(CPM_EINVAL, "cpm_mat_sell_info: null matrix");
  if (bs) *bs = A->sell.bs;
  if (stencil)
    *stencil = A->sell.stencil ? (A->sell.kuhn ? (A->sell.half64 ? 4 : A->sell.half32 ? 3 : 2) : 1) : 0;
  if (slots) *slots = A->sell.nslot;
  return CPM_OK;
}

extern "C" int cpm_cg_sell(cpm_ctx* ctx, cpm_mat* A, int nrhs, const double* b, double* x, int use_x0, double tol,
                           int maxiter, int precision, double delta, int32_t* iters, double* resid) {
  if (!ctx || !A || !b || !x || !iters || !resid) return cpm::fail(CPM_EINVAL, "cpm_cg_sell: null argument");
  if (!A->sell.sptr) return cpm::fail(CPM_EINVAL, "cpm_cg_sell: operator has no SELL form (cpm_mat_add_sell)");
  if (nrhs != 1 && !(nrhs == 3 && A->sell.bs == 1))
    return cpm::fail(CPM_EINVAL, "cpm_cg_sell: nrhs must be 1 (or 3 for scalar operators)");
  int failed = 0;
  int rc;
  if (precision == 64)
    rc = cpm::sell_cg64(ctx, A, nrhs, b, x, use_x0, tol, maxiter, 0, iters, resid, &failed);
  else if (precision == 32)
    rc = cpm::sell_cg_mixed(ctx, A, nrhs, b, x, use_x0, tol, maxiter, delta, 0, iters, resid, &failed, nullptr);
  else
    return cpm::fail(CPM_EINVAL, "cpm_cg_sell: precision must be 64 or 32");
  if (rc) return rc;
  if (failed) {
    int m = 1;
    for (int j = 0; j < nrhs; ++j) m = std::max(m, (int)iters[j]);
    return m;
  }
  return CPM_OK;
}

extern "C" int cpm_spmv_sell(cpm_ctx* ctx, const cpm_mat* A, int nrhs, const double* x, double* y) {
  if (!ctx || !A || !x || !y) return cpm::fail(CPM_EINVAL, "cpm_spmv_sell: null argument");
  if (!A->sell.sptr) return cpm::fail(CPM_EINVAL, "cpm_spmv_sell: operator has no SELL form");
  if (nrhs != 1 && !(nrhs == 3 && A->sell.bs == 1)) return cpm::fail(CPM_EINVAL, "cpm_spmv_sell: bad nrhs");
  return cpm::sell_spmv64(ctx, A, nrhs, x, y);
}
