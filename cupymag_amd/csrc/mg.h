// Multigrid-preconditioned CG (see mg.hip).
#pragma once
#include <cstdint>
#include <vector>

#include <hip/hip_runtime.h>

struct cpm_ctx;
struct cpm_mat;

namespace cpm {

struct CgState;

// What one PCG in flight uses of its context: the launch stream, the CG
// states of its columns, the reduction partials (kMaxPartials doubles) and
// a pinned host mirror of the states.  A context has two lanes, so that two
// independent solves of one step run at once on two streams.
struct PcgLane {
  hipStream_t stream = nullptr;
  CgState* st = nullptr;
  double* partials = nullptr;
  void* pinned = nullptr;
  int prof = 0;  // time the level-0 kernels of the first iteration of a batch (HIP events of the context)
  // another solve runs beside this one while its first batch runs: the
  // timed iteration of that batch is its last, so that the timings are of
  // the kernels alone
  int shared = 0;
};

struct MgLevel {
  cpm_mat* A = nullptr;
  int32_t nb = 0;           // nodes (block rows)
  int32_t dims[3] = {0, 0, 0};
  int64_t n = 0;            // nb * bs
  void *x = nullptr, *r = nullptr, *b = nullptr, *d0 = nullptr, *d1 = nullptr;
  void* dinv = nullptr;     // bs*bs per node: inverse diagonal block
  double lmax = 0.0;        // power-iteration estimate of lambda_max(D^-1 A)
  // FP16 form (precision 16): off-diagonal values times vscale^-1 (a power
  // of two), in the SELL layout, and diagonal blocks compensated for their
  // rounding: dgh in FP64 (during set-up only), dg32 its FP32 rounding, which
  // the FP16 sweeps read (and invert per node)
  void* v16 = nullptr;
  int khalf = 0;            // v16 keeps half of a symmetric Kuhn stencil (GridArg::kuhn = 2)
  double* dgh = nullptr;
  float* dg32 = nullptr;
  double vscale = 1.0;
  // floating hierarchy, level 0: the smoothing operator is the periodic
  // stiffness recovered from the anchored operator (owned copies; FP64 for
  // precision 64, the compensated FP32 form for precision 32); null: the
  // level's own SELL form
  double *kv64 = nullptr, *kdg64 = nullptr, *kdgc = nullptr;
  float* kv32 = nullptr;
};

}  // namespace cpm

struct cpm_mg {
  cpm_ctx* ctx = nullptr;
  int bs = 1, prec = 32, degree = 2;
  // Chebyshev degree of the levels between the finest and the coarsest
  // (cpm_mg_set_coarse_degree); 0: `degree` on every level
  int degree1 = 0;
  int anchor = 1;           // row 0 of the fine operator is an identity row: node 0 left out of the transfers
  // anchored fine operator over singular periodic levels: the V-cycle
  // solves the periodic problem, with the anchor applied around it
  int floating = 0;
  int half = 0;             // precision 16: FP32 vectors, FP16 off-diagonal values
  double ratio = 20.0;
  std::vector<cpm::MgLevel> L;
  void* cinv = nullptr;     // coarsest dense inverse (T), ncoarse x ncoarse
  int64_t ncoarse = 0;
  // csmooth > 0: no dense inverse; the coarsest level is solved by csmooth
  // Chebyshev steps on D^-1 A over [lmax / cratio, 1.1 lmax] (its first
  // step fused with the restriction, as on the other levels)
  int csmooth = 0;
  double cratio = 0.0;
  // Chebyshev coefficients: c0[l] = 1/theta; for step k of level l,
  // c1[8l + k] = rho_k rho_(k-1) and c2[8l + k] = 2 rho_k / delta
  std::vector<double> c0, c1, c2;
  // smoothing degree of level l (the coarsest: its Chebyshev solve, or 0)
  int level_degree(size_t l) const {
    if (l + 1 == L.size()) return csmooth;
    return (l == 0 || !degree1) ? degree : degree1;
  }
  // PCG vectors of the fine level, 8 bytes per entry (p and q hold FP32
  // entries in the mixed-precision solves); level and PCG vectors hold kcap
  // columns
  double *r = nullptr, *p = nullptr, *q = nullptr;
  int kcap = 1;
  // stream the V-cycle kernels go to (null: the context's stream); the
  // lane's stream during a solve, the side stream while a chunk of
  // iterations is captured; lpart: the partials of the solve in progress
  hipStream_t ls = nullptr;
  double* lpart = nullptr;
  // captured chunks of PCG iterations (hipGraph), replayed on the context
  // stream; keyed by the vectors and states they were captured with
  int graphs = 1;
  hipStream_t side = nullptr;
  struct Chunk {
    uint64_t key = 0;
    hipGraphExec_t exec = nullptr;
    int iters = 0;
  };
  std::vector<Chunk> chunks;
  std::vector<void*> allocs;
};



namespace cpm {

// PCG on the hierarchy's fine operator for K interleaved columns (K = 1, or
// 3 for a scalar operator), on its context's stream; iters and resid hold K
// entries.
// delta > 0: FP32 operator products with FP64 residual replacements (the
// stopping test on the FP64 residual); 0: FP64 products.
int mg_pcg(cpm_mg* M, int K, const double* b, double* x, int use_x0, double tol, int maxiter, double delta,
           int hint, int32_t* iters, double* resid, int* failed);
// the same on a lane of the hierarchy's context (its stream, states,
// partials and pinned mirror)
int mg_pcg_lane(cpm_mg* M, const PcgLane& ln, int K, const double* b, double* x, int use_x0, double tol,
                int maxiter, double delta, int hint, int32_t* iters, double* resid, int* failed);
// the context's first lane (its stream, CG states 0.., partials, pinned
// mirror; profiling as the context has it) and its second (stream2, CG
// states 8.., partials2, pinned2, no profiling; created on first use)
PcgLane lane1(cpm_ctx* ctx);
int lane2(cpm_ctx* ctx, PcgLane* out);
// grow the level and PCG vectors of M to K columns (synchronises the
// context's stream, or `s` when given)
int mg_reserve(cpm_mg* M, int K, hipStream_t s = nullptr);
// the fine operator a hierarchy was built on
const cpm_mat* mg_operator(const cpm_mg* M);

}  // namespace cpm
