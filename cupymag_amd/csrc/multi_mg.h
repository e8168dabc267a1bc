// Batched multigrid-preconditioned CG over many independent systems, each
// with its own hierarchy; see multi_mg.hip.
#pragma once
#include <cstdint>
#include <vector>

struct cpm_ctx;
struct cpm_mg;

namespace cpm {

struct MultiMgSys {
  cpm_mg* M = nullptr;  // hierarchy of the system's operator (its workspace is used)
  const double* b = nullptr;
  double* x = nullptr;
  int use_x0 = 0;
  double tol = 1e-7;
  int32_t maxiter = 5000;
  double delta = 0.0;  // > 0: FP32 products with FP64 residual replacements (see mg_pcg)
};

struct MultiMgTables;

// Device tables of a batched solve, reused across calls.
struct MultiMgSolver {
  MultiMgTables* T;
  MultiMgSolver();
  ~MultiMgSolver();
  MultiMgSolver(const MultiMgSolver&) = delete;
  MultiMgSolver& operator=(const MultiMgSolver&) = delete;
};

// Solve every system of `sys` (hierarchies of one block size, V-cycle
// precision and smoother degree; each hierarchy at most once; K interleaved
// columns each, K = 1 or 3 for scalar operators) with one launch per phase
// for the whole batch.  Each system keeps its own scalars, stops on its own
// tolerance and uses its own workgroups and reduction order (those of
// mg_pcg), so its result equals mg_pcg's on the same input bit for bit,
// whatever the other members.  iters and resid: K entries per system.
int multi_mg_pcg(cpm_ctx* ctx, MultiMgSolver& W, const std::vector<MultiMgSys>& sys, int K, int hint,
                 int32_t* iters, double* resid, int* failed);

}  // namespace cpm
