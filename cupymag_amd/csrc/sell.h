// Sliced-ELL (SELL-64) operator form and its solvers; see sell.hip.
#pragma once
#include <cstdint>

struct cpm_ctx;
struct cpm_mat;

namespace cpm {

struct Sell {
  int bs = 0;               // block size (1 scalar, 3 for the elasticity operator)
  int32_t nbrows = 0;       // block rows
  int32_t nslices = 0;      // ceil(nbrows / 64)
  int64_t nslot = 0;        // total slots; slice s owns slots [sptr[s], sptr[s+1])
  int64_t* sptr = nullptr;  // device
  int32_t* col = nullptr;   // device, [slot][64]
  double* v64 = nullptr;    // device, [slot][bs*bs][64], off-diagonal blocks
  float* v32 = nullptr;     // device, FP32 copy of v64
  double* dg64 = nullptr;   // device, [brow][bs*bs] diagonal blocks
  double* dgc = nullptr;    // device, diagonal blocks of the FP32 form, compensated (see sell.hip)
  // index-free form for DOFs numbered i*ny*nz + j*nz + k on a periodic grid:
  // slot s of every row holds the neighbour at offset soff[s], col is unused
  int stencil = 0;
  int32_t nst = 0;           // offsets per row (14 Kuhn, 26 hex)
  int32_t dims[3] = {0, 0, 0};
  int32_t* soff = nullptr;   // device, packed (di+1) | (dj+1) << 2 | (dk+1) << 4
  int kuhn = 0;              // the offsets are the Kuhn stencil in kKuhnCodes order (sell_dev.h)
  // Kuhn stencil symmetric slot by slot (A(i, i-d) = A(i-d, i)^T to 1e-12 of
  // its largest entry, every dim >= 3): v32 keeps the forward slots 7..13
  // only, 7 planes per slice ([slice*7 + u][bs*bs][64]), read with
  // GridArg::kuhn = 2 (CPM_GRID_ARG32)
  int half32 = 0;
  // the same for the FP64 products of the solvers when the stencil is
  // symmetric to rounding (1e-14 of its largest entry): v64h holds the
  // forward slots of v64 ([slice*7 + u][bs*bs][64]), read with
  // GridArg::kuhn = 2 and h64 = 1 (CPM_GRID_ARG64); v64 keeps every slot
  // for the multigrid set-up and the FP64 fallback of the FP32 products
  double* v64h = nullptr;
  int half64 = 0;
  int grid_cap = 8192;       // workgroup cap of the SpMV launches (tunable, grid-stride beyond)
  int xg = 0;                // XCD group of the row launches of a stencil form (GridArg::xg)
  int shz = 0, shy = 0;      // log2(dims[2]) + 1, log2(dims[1]) + 1 for powers of two, else 0 (GridArg)
};

// the grid argument of the kernels for a stencil-form operator (FP64 form)
#define CPM_GRID_ARG(S) (GridArg{(S).dims[0], (S).dims[1], (S).dims[2], (S).nst, (S).soff, (S).kuhn, (S).xg, (S).shz, (S).shy})
// ... for its FP32 form v32 (kuhn = 2 when v32 keeps half of the stencil;
// the FP64 products of a kernel that reads both ignore the difference)
#define CPM_GRID_ARG32(S) \
  (GridArg{(S).dims[0], (S).dims[1], (S).dims[2], (S).nst, (S).soff, (S).kuhn ? ((S).half32 ? 2 : 1) : 0, (S).xg, (S).shz, (S).shy})

// ... for the FP64 products of the solvers (sell_v64: the half form when
// there is one)
#define CPM_GRID_ARG64(S)                                                                                           \
  (GridArg{(S).dims[0], (S).dims[1], (S).dims[2], (S).nst, (S).soff, (S).kuhn ? ((S).half64 ? 2 : 1) : 0, (S).xg, \
           (S).shz, (S).shy, (S).half64})
inline const double* sell_v64(const Sell& S) { return S.half64 ? S.v64h : S.v64; }

// XCD group (xcd_perm) of the row launches of a stencil form on a grid of
// these dims: the largest power of two G <= 64 for which 8 G workgroups of
// kBlock rows tile one plane (ny nz rows) exactly, else 1 (dispatch order).
int sell_xcd_group(const int32_t dims[3]);
void sell_free(Sell& S);
int sell_grid(const Sell& A);
int sell_spmv64(cpm_ctx* ctx, const cpm_mat* M, int K, const double* x, double* y);
int sell_cg64(cpm_ctx* ctx, cpm_mat* M, int K, const double* b, double* x, int use_x0, double tol, int maxiter,
              int hint, int32_t* iters, double* resid, int* failed);
int sell_cg_mixed(cpm_ctx* ctx, cpm_mat* M, int K, const double* b, double* x, int use_x0, double tol, int maxiter,
                  double delta, int hint, int32_t* iters, double* resid, int* failed, int32_t* sweeps);
int interleave(cpm_ctx* ctx, int64_t n, int K, const double* src, double* dst, int inverse);

}  // namespace cpm
