// Batched CG over many independent systems (different operators and sizes)
// in one sequence of launches; see multi.hip.
#pragma once
#include <cstdint>
#include <vector>

#include <hip/hip_runtime.h>

struct cpm_ctx;
struct cpm_mat;

namespace cpm {

struct CgState;
struct BSys;
struct Sell;

// One system of a batched solve.  Vectors are interleaved as the SELL
// solvers expect (length n = block rows * bs * K); the workspace vectors
// belong to the caller and hold at least n entries each (FP64 CG: r, p, q;
// mixed precision: r, y64, r32, p32, q32, x32).
struct MultiSys {
  const cpm_mat* A = nullptr;
  int64_t n = 0;
  const double* b = nullptr;
  double* x = nullptr;
  int use_x0 = 0;
  double tol = 1e-7, delta = 0.3;
  int32_t maxiter = 5000;
  double *r = nullptr, *p = nullptr, *q = nullptr, *y64 = nullptr;
  float *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *x32 = nullptr;
};

// Device tables of a batched solve, reused across calls with the same
// batch composition (and the captured iteration graph with them).
struct MultiSolver {
  int cap_sys = 0, cap_blk = 0;
  BSys* d_sys = nullptr;          // device descriptors
  BSys* h_sys = nullptr;          // pinned staging
  int32_t* d_b2s = nullptr;       // workgroup -> system
  int32_t* h_b2s = nullptr;
  double* partials = nullptr;     // 6 * cap_blk
  CgState* d_st = nullptr;        // 3 per system
  CgState* h_st = nullptr;        // pinned mirror
  // captured iteration chunk (graphs on): valid for `gkey`
  hipStream_t side = nullptr;
  hipGraphExec_t gexec = nullptr;
  uint64_t gkey = 0;
  int gchunk = 0;
  ~MultiSolver();
};

// Solve every system of `sys` (all with SELL block size bs and K
// right-hand sides, FP64 CG or mixed precision) with one launch per CG
// phase for the whole batch.  Each system keeps its own scalars, stops on
// its own tolerance and uses its own workgroups, so its arithmetic does
// not depend on the other members of the batch.  Outputs per system:
// iters[K], resid[K] (entries j*K + k), failed[j].
int multi_cg(cpm_ctx* ctx, MultiSolver& M, int bs, int K, int mixed, const std::vector<MultiSys>& sys, int hint,
             bool graph, int32_t* iters, double* resid, int* failed);

// workgroups a system is given in batched launches (its SELL grid)
int multi_blocks(const Sell& A);

}  // namespace cpm
