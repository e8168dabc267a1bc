// Batched CG: many independent systems advanced together, one launch per CG
// phase for the whole batch.
//
// Every system owns a contiguous range of workgroups (sized from its own row
// count only) and a descriptor with its operator, vectors and CG scalars; a
// table maps each workgroup to its system.  The SpMV and vector kernels run
// the same bodies as the single-system solvers (sell_dev.h) over the
// system's range, store their partial sums by workgroup, and a reducer with
// one workgroup per system adds its range in a fixed order.  A system that
// has converged costs its workgroups one early return per launch.  The
// result of a system is therefore independent of the batch it runs in, and
// the launch count per iteration is that of one solve, whatever the batch
// size: the small meshes of a batch share the launches of the large ones.
//
// The host enqueues iterations in batches and reads the scalars of all
// systems back between batches (as the single-system drive does).  With
// graphs on, a fixed chunk of iterations is captured once per batch
// composition on a side stream and replayed on the context stream.
#include <algorithm>
#include <cstring>
#include <vector>

#include "device.h"
#include "multi.h"
#include "sell_dev.h"

namespace cpm {

struct BSys {
  const int64_t* sptr;
  const int32_t* col;
  const double* v64;  // the FP64 products' values (sell_v64)
  const float* v32;
  const double* dg64;
  const double* dgc;
  GridArg ga;    // for v32
  GridArg ga64;  // for v64
  int32_t nbrows, stencil;
  int64_t n;
  const double* b;
  double* x;
  double *r, *p, *q, *y64;
  float *r32, *p32, *q32, *x32;
  CgState* st;
  int32_t blk0, nblk;
  int32_t use_x0, maxiter;
  double tol, delta;
};

enum : int { OP_CG = 0, OP_X0 = 1, OP_REL = 2 };

// OP_CG: q = A p (T-typed, partials of p.q); OP_X0: y = A x0 (FP64, into
// y64 for the mixed solver, q otherwise); OP_REL: y64 = A (x + x32) on the
// systems that requested a reliable update.
template <int BS, int K, class T, int OP, bool MIXED>
__global__ void __launch_bounds__(kBlock) kb_spmv(const BSys* __restrict__ sys, const int32_t* __restrict__ b2s,
                                                  double* partials) {
  const BSys& d = sys[b2s[blockIdx.x]];
  bool any = false;
#pragma unroll
  for (int j = 0; j < K; ++j)
    any |= (OP == OP_CG) ? !d.st[j].done : (OP == OP_REL ? d.st[j].upd != 0 : d.use_x0 != 0);
  if (!any) return;
  const T* val;
  const double* dg;
  const T* x;
  T* y;
  const float* x32 = nullptr;
  if constexpr (sizeof(T) == 4) {
    val = d.v32;
    dg = d.dgc;
    x = d.p32;
    y = d.q32;
  } else {
    val = d.v64;
    dg = d.dg64;
    if (OP == OP_CG) {
      x = d.p;
      y = d.q;
    } else {
      x = d.x;
      y = (OP == OP_REL || MIXED) ? d.y64 : d.q;
      if (OP == OP_REL) x32 = d.x32;
    }
  }
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) acc[j] = 0.0;
  const int64_t lb = xcd_perm((int64_t)blockIdx.x - d.blk0, d.nblk, d.ga.xg);
  const GridArg& ga = sizeof(T) == 8 ? d.ga64 : d.ga;
  if (d.stencil)
    sell_rows<BS, K, T, OP == OP_CG, OP == OP_REL, true>(lb, d.nblk, d.nbrows, d.sptr, d.col, val, dg, x, x32, y,
                                                          ga, acc);
  else
    sell_rows<BS, K, T, OP == OP_CG, OP == OP_REL, false>(lb, d.nblk, d.nbrows, d.sptr, d.col, val, dg, x, x32, y,
                                                           ga, acc);
  if (OP == OP_CG) block_partial<K>(acc, partials);
}

template <int K, class T>
__global__ void __launch_bounds__(kBlock) kb_pupd(const BSys* __restrict__ sys, const int32_t* __restrict__ b2s) {
  const BSys& d = sys[b2s[blockIdx.x]];
  const int64_t lb = (int64_t)blockIdx.x - d.blk0;
  if constexpr (sizeof(T) == 4)
    pupd_rows<K, float>(lb, d.nblk, d.n, d.r32, d.p32, d.st);
  else
    pupd_rows<K, double>(lb, d.nblk, d.n, d.r, d.p, d.st);
}

template <int K, class T>
__global__ void __launch_bounds__(kBlock) kb_xr(const BSys* __restrict__ sys, const int32_t* __restrict__ b2s,
                                                double* partials) {
  const BSys& d = sys[b2s[blockIdx.x]];
  const int64_t lb = (int64_t)blockIdx.x - d.blk0;
  if constexpr (sizeof(T) == 4)
    xr_rows<K, float>(lb, d.nblk, d.n, d.x32, d.r32, d.p32, d.q32, d.st, partials);
  else
    xr_rows<K, double>(lb, d.nblk, d.n, d.x, d.r, d.p, d.q, d.st, partials);
}

template <int K>
__global__ void __launch_bounds__(kBlock) kb_rel_r(const BSys* __restrict__ sys, const int32_t* __restrict__ b2s,
                                                   double* partials) {
  const BSys& d = sys[b2s[blockIdx.x]];
  const int64_t lb = (int64_t)blockIdx.x - d.blk0;
  rel_r_rows<K>(lb, d.nblk, d.n, d.b, d.y64, d.x, d.x32, d.r, d.r32, d.st, partials);
}

template <int K, bool REL>
__global__ void __launch_bounds__(kBlock) kb_init(const BSys* __restrict__ sys, const int32_t* __restrict__ b2s,
                                                  double* partials) {
  const BSys& d = sys[b2s[blockIdx.x]];
  const int64_t lb = (int64_t)blockIdx.x - d.blk0;
  double acc[2 * K];
#pragma unroll
  for (int j = 0; j < 2 * K; ++j) acc[j] = 0.0;
  init_rows<K, REL>(lb, d.nblk, d.n, d.b, d.x, REL ? d.y64 : d.q, d.r, d.r32, d.x32, d.use_x0, acc);
  block_partial<2 * K>(acc, partials);
}

template <int K>
__global__ void __launch_bounds__(kBlock) kb_zero_b(const BSys* __restrict__ sys, const int32_t* __restrict__ b2s) {
  const BSys& d = sys[b2s[blockIdx.x]];
  if (!d.use_x0) return;
  zero_b_rows<K>((int64_t)blockIdx.x - d.blk0, d.nblk, d.n, d.x, d.st);
}

// One workgroup per system: add the system's partials (its own workgroup
// range of a launch of G workgroups, in order) and advance its scalars.
enum : int { RED_INIT64 = 10, RED_INIT_REL = 11 };

template <int K, int MODE>
__global__ void __launch_bounds__(kRedBlock) kb_red(const BSys* __restrict__ sys, int G,
                                                    const double* __restrict__ partials) {
  constexpr bool INIT = MODE == RED_INIT64 || MODE == RED_INIT_REL;
  constexpr int NV = INIT ? 2 * K : K;
  const BSys& d = sys[blockIdx.x];
  if constexpr (!INIT) {
    bool any = false;
#pragma unroll
    for (int j = 0; j < K; ++j) any |= cg_state_uses<MODE>(d.st[j]);
    if (!any) return;
  }
  __shared__ double lds[NV * (kRedBlock / 64)];
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (int i = threadIdx.x; i < d.nblk; i += kRedBlock)
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += partials[(int64_t)k * G + d.blk0 + i];
  block_sum_t<NV, kRedBlock>(v, lds);
  if (threadIdx.x != 0) return;
  if constexpr (MODE == RED_INIT64)
    cg_state_init<K, false>(d.st, v, d.tol, d.maxiter);
  else if constexpr (MODE == RED_INIT_REL)
    cg_state_init<K, true>(d.st, v, d.tol, d.maxiter);
  else
    cg_state_update<K, MODE>(d.st, v, d.delta);
}

// ---------------------------------------------------------------- host

int multi_blocks(const Sell& A) { return sell_grid(A); }

MultiSolver::~MultiSolver() {
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (side) (void)hipStreamDestroy(side);
  for (void* p : {(void*)d_sys, (void*)d_b2s, (void*)partials, (void*)d_st})
    if (p) (void)hipFree(p);
  for (void* p : {(void*)h_sys, (void*)h_b2s, (void*)h_st})
    if (p) (void)hipHostFree(p);
}

static int multi_reserve(MultiSolver& M, int nsys, int nblk) {
  if (nsys > M.cap_sys) {
    const int cap = std::max(nsys, 2 * M.cap_sys);
    for (void** p : {(void**)&M.d_sys, (void**)&M.d_st})
      if (*p) CPM_HIP(hipFree(*p));
    for (void** p : {(void**)&M.h_sys, (void**)&M.h_st})
      if (*p) CPM_HIP(hipHostFree(*p));
    M.d_sys = nullptr, M.d_st = nullptr, M.h_sys = nullptr, M.h_st = nullptr;
    CPM_HIP(hipMalloc(&M.d_sys, sizeof(BSys) * cap));
    CPM_HIP(hipMalloc(&M.d_st, sizeof(CgState) * 3 * cap));
    CPM_HIP(hipHostMalloc(&M.h_sys, sizeof(BSys) * cap, hipHostMallocDefault));
    CPM_HIP(hipHostMalloc(&M.h_st, sizeof(CgState) * 3 * cap, hipHostMallocDefault));
    M.cap_sys = cap;
  }
  if (nblk > M.cap_blk) {
    const int cap = std::max(nblk, M.cap_blk + M.cap_blk / 2);
    for (void** p : {(void**)&M.d_b2s, (void**)&M.partials})
      if (*p) CPM_HIP(hipFree(*p));
    if (M.h_b2s) CPM_HIP(hipHostFree(M.h_b2s));
    M.d_b2s = nullptr, M.partials = nullptr, M.h_b2s = nullptr;
    CPM_HIP(hipMalloc(&M.d_b2s, sizeof(int32_t) * cap));
    CPM_HIP(hipMalloc(&M.partials, sizeof(double) * 6 * (int64_t)cap));
    CPM_HIP(hipHostMalloc(&M.h_b2s, sizeof(int32_t) * cap, hipHostMallocDefault));
    M.cap_blk = cap;
  }
  return CPM_OK;
}

template <int BS, int K, bool MIXED>
static int issue_iters(const MultiSolver& M, int nsys, int G, int todo, hipStream_t s) {
  const BSys* sys = M.d_sys;
  const int32_t* b2s = M.d_b2s;
  double* pt = M.partials;
  for (int k = 0; k < todo; ++k) {
    if (MIXED) {
      kb_pupd<K, float><<<G, kBlock, 0, s>>>(sys, b2s);
      kb_spmv<BS, K, float, OP_CG, true><<<G, kBlock, 0, s>>>(sys, b2s, pt);
      kb_red<K, RED_ALPHA><<<nsys, kRedBlock, 0, s>>>(sys, G, pt);
      kb_xr<K, float><<<G, kBlock, 0, s>>>(sys, b2s, pt);
      kb_red<K, RED_RR_REL><<<nsys, kRedBlock, 0, s>>>(sys, G, pt);
      kb_spmv<BS, K, double, OP_REL, true><<<G, kBlock, 0, s>>>(sys, b2s, pt);
      kb_rel_r<K><<<G, kBlock, 0, s>>>(sys, b2s, pt);
      kb_red<K, RED_RR_TRUE><<<nsys, kRedBlock, 0, s>>>(sys, G, pt);
    } else {
      kb_pupd<K, double><<<G, kBlock, 0, s>>>(sys, b2s);
      kb_spmv<BS, K, double, OP_CG, false><<<G, kBlock, 0, s>>>(sys, b2s, pt);
      kb_red<K, RED_ALPHA><<<nsys, kRedBlock, 0, s>>>(sys, G, pt);
      kb_xr<K, double><<<G, kBlock, 0, s>>>(sys, b2s, pt);
      kb_red<K, RED_RR><<<nsys, kRedBlock, 0, s>>>(sys, G, pt);
    }
  }
  CPM_HIP(hipGetLastError());
  return CPM_OK;
}

template <int BS, int K, bool MIXED>
static int multi_loop(cpm_ctx* ctx, MultiSolver& M, const std::vector<MultiSys>& sys, int hint, bool graph,
                      int32_t* iters, double* resid, int* failed) {
  const int nsys = (int)sys.size();
  int G = 0;
  for (const MultiSys& m : sys) G += multi_blocks(m.A->sell);
  if (int rc = multi_reserve(M, nsys, G)) return rc;
  int maxit = 0;
  bool any_x0 = false;
  int blk = 0;
  for (int j = 0; j < nsys; ++j) {
    const MultiSys& m = sys[j];
    const Sell& A = m.A->sell;
    BSys& d = M.h_sys[j];
    d.sptr = A.sptr;
    d.col = A.col;
    d.v64 = sell_v64(A);
    d.v32 = A.v32;
    d.dg64 = A.dg64;
    d.dgc = A.dgc;
    d.ga = CPM_GRID_ARG32(A);
    d.ga64 = CPM_GRID_ARG64(A);
    d.nbrows = A.nbrows;
    d.stencil = A.stencil;
    d.n = m.n;
    d.b = m.b;
    d.x = m.x;
    d.r = m.r;
    d.p = m.p;
    d.q = m.q;
    d.y64 = m.y64;
    d.r32 = m.r32;
    d.p32 = m.p32;
    d.q32 = m.q32;
    d.x32 = m.x32;
    d.st = M.d_st + 3 * j;
    d.blk0 = blk;
    d.nblk = multi_blocks(A);
    d.use_x0 = m.use_x0;
    d.maxiter = m.maxiter;
    d.tol = m.tol;
    d.delta = m.delta;
    for (int b = 0; b < d.nblk; ++b) M.h_b2s[blk + b] = j;
    blk += d.nblk;
    maxit = std::max(maxit, (int)m.maxiter);
    any_x0 |= m.use_x0 != 0;
  }
  hipStream_t s = ctx->stream;
  CPM_HIP(hipMemcpyAsync(M.d_sys, M.h_sys, sizeof(BSys) * nsys, hipMemcpyHostToDevice, s));
  CPM_HIP(hipMemcpyAsync(M.d_b2s, M.h_b2s, sizeof(int32_t) * G, hipMemcpyHostToDevice, s));
  if (any_x0) kb_spmv<BS, K, double, OP_X0, MIXED><<<G, kBlock, 0, s>>>(M.d_sys, M.d_b2s, M.partials);
  kb_init<K, MIXED><<<G, kBlock, 0, s>>>(M.d_sys, M.d_b2s, M.partials);
  kb_red<K, MIXED ? RED_INIT_REL : RED_INIT64><<<nsys, kRedBlock, 0, s>>>(M.d_sys, G, M.partials);
  if (any_x0) kb_zero_b<K><<<G, kBlock, 0, s>>>(M.d_sys, M.d_b2s);
  CPM_HIP(hipGetLastError());

  // iteration graph: a fixed chunk captured for this composition
  const int chunk = 16;
  const uint64_t key = ((uint64_t)(uintptr_t)M.d_sys * 1315423911ull) ^ ((uint64_t)(uintptr_t)M.partials << 7) ^
                       ((uint64_t)G << 32) ^ ((uint64_t)nsys << 8) ^ (uint64_t)(BS * 16 + K * 2 + MIXED);
  if (graph && (M.gexec == nullptr || M.gkey != key)) {
    if (M.gexec) {
      CPM_HIP(hipGraphExecDestroy(M.gexec));
      M.gexec = nullptr;
    }
    if (!M.side) CPM_HIP(hipStreamCreateWithFlags(&M.side, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    CPM_HIP(hipStreamBeginCapture(M.side, hipStreamCaptureModeThreadLocal));
    int rc = issue_iters<BS, K, MIXED>(M, nsys, G, chunk, M.side);
    hipError_t e = hipStreamEndCapture(M.side, &g);
    if (rc) return rc;
    if (e != hipSuccess) return fail(CPM_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&M.gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      M.gexec = nullptr;
      return fail(CPM_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    M.gkey = key;
    M.gchunk = chunk;
  }
  auto read = [&]() -> int {
    CPM_HIP(hipMemcpyAsync(M.h_st, M.d_st, sizeof(CgState) * 3 * nsys, hipMemcpyDeviceToHost, s));
    CPM_HIP(hipStreamSynchronize(s));
    return CPM_OK;
  };
  auto state = [&](int j, int k) -> const CgState& { return M.h_st[3 * j + k]; };
  if (int rc = read()) return rc;
  auto all_done = [&]() {
    for (int j = 0; j < nsys; ++j)
      for (int k = 0; k < K; ++k)
        if (!state(j, k).done) return false;
    return true;
  };
  int batch = hint > 2 ? hint - 1 : 8;
  int issued = 0;
  while (!all_done() && issued <= maxit) {
    int todo = std::max(1, std::min(std::min(batch, 4096), maxit - issued + 1));
    if (graph) {
      const int reps = std::max(1, todo / M.gchunk);
      for (int r = 0; r < reps; ++r) CPM_HIP(hipGraphLaunch(M.gexec, s));
      todo = reps * M.gchunk;
    } else if (int rc = issue_iters<BS, K, MIXED>(M, nsys, G, todo, s)) {
      return rc;
    }
    issued += todo;
    if (int rc = read()) return rc;
    int it_max = 0;
    for (int j = 0; j < nsys; ++j)
      for (int k = 0; k < K; ++k)
        if (!state(j, k).done) it_max = std::max(it_max, (int)state(j, k).it);
    batch = std::max(2, std::min(64, it_max / 4 + 2));
  }
  for (int j = 0; j < nsys; ++j) {
    failed[j] = 0;
    for (int k = 0; k < K; ++k) {
      const CgState& c = state(j, k);
      iters[j * K + k] = c.it;
      resid[j * K + k] = c.resid;
      if (c.fail || !c.done) failed[j] = 1;
    }
  }
  return CPM_OK;
}

int multi_cg(cpm_ctx* ctx, MultiSolver& M, int bs, int K, int mixed, const std::vector<MultiSys>& sys, int hint,
             bool graph, int32_t* iters, double* resid, int* failed) {
  if (sys.empty()) return CPM_OK;
  for (const MultiSys& m : sys) {
    if (!m.A || !m.A->sell.sptr || m.A->sell.bs != bs)
      return fail(CPM_EINVAL, "batched CG: every operator needs a SELL form of the batch's block size");
    if (m.n != (int64_t)m.A->sell.nbrows * bs * K) return fail(CPM_EINVAL, "batched CG: vector length mismatch");
    if (!m.b || !m.x || !m.r || (mixed ? (!m.y64 || !m.r32 || !m.p32 || !m.q32 || !m.x32) : (!m.p || !m.q)))
      return fail(CPM_EINVAL, "batched CG: missing vector or workspace");
  }
#define CPM_MULTI(BS_, K_)                                                                                   \
  if (bs == BS_ && K == K_)                                                                                   \
    return mixed ? multi_loop<BS_, K_, true>(ctx, M, sys, hint, graph, iters, resid, failed)                  \
                 : multi_loop<BS_, K_, false>(ctx, M, sys, hint, graph, iters, resid, failed);
  CPM_MULTI(1, 1)
  CPM_MULTI(1, 3)
  CPM_MULTI(3, 1)
#undef CPM_MULTI
  return fail(CPM_EINVAL, "batched CG: unsupported (block size, rhs count)");
}

}  // namespace cpm
