// Geometric multigrid preconditioner and preconditioned CG for the
// stiffness-type operators of structured periodic Kuhn meshes: elasticity
// (bs = 3), demag (bs = 1) and the Gauss-Seidel operator M + alpha1 K
// (bs = 1, three right-hand-side columns at once).
//
// Levels: a level of n cells per dimension has a coarser one of ceil(n/2)
// cells, its operator assembled on that mesh by the host assembler
// (rediscretisation).  For even n the coarse mesh is the Kuhn mesh of the
// even nodes: the Kuhn (Freudenthal) triangulation of cell size h refines
// the one of cell size 2h, so continuous P1 functions of the coarse mesh are
// P1 functions of the fine mesh and the prolongation is their nodal
// interpolation: fine node 2I + d, d in {0,1}^3, takes the coarse value at I
// (d = 0) or the mean of the values at I and I + d (the coarse edge in
// direction d), periodic in every dimension.  For odd n the coarse mesh is
// the regular Kuhn mesh of (n+1)/2 cells of the same periodic box and the
// prolongation is the same P1 interpolation at the fine nodes' nominal
// positions (non-nested: barycentric weights in the coarse Kuhn
// tetrahedron that holds the point).  Restriction is its transpose.
//
// Anchored operators (row 0 an identity row: the anchor of A_demag and
// A_el) come in two kinds of hierarchy.  Floating (coarse levels given as
// the singular periodic stiffness K, coarsest solve its pseudo-inverse):
// the V-cycle V works on the periodic problem, where the anchor is no point
// constraint for the coarse levels to approximate, and the preconditioner
// is z = F^T V(F r) + e_0 r_0 with F = I - e_0 1^T per component -- node 0
// takes minus the sum of the other entries of r (a right-hand side in the
// range of K), the output is shifted by its node-0 value and node 0 gets
// r_0 (the identity row).  With an exact V this is the inverse of the
// anchored operator; it is symmetric, so CG keeps its theory.  Level 0
// smooths with K, recovered on the device from the anchored operator (K
// annihilates the translations: the block of node 0's column in row i is
// minus the sum of the row's other blocks; row 0 by symmetry).  Anchored
// coarse levels instead: node 0 is left out of both transfers, so
// corrections never touch it.
//
// Smoother: Chebyshev iteration of degree m on D^-1 A (D = diagonal blocks,
// inverted once), on the interval [lmax/ratio, lmax] with lmax = 1.1 x a
// power-iteration estimate; the same polynomial before and after the coarse
// correction, and an exact (dense inverse) solve on the coarsest level, so
// the V-cycle is a symmetric positive definite preconditioner.  It runs in
// FP64, FP32 (precision 32, on the FP32 SELL form), or FP32 with the level
// operators' off-diagonal values in FP16 (precision 16: scaled by a power of
// two, diagonal blocks compensated for the rounding in FP64 and then stored
// in FP32, so the stored row sums equal the FP64 ones to the FP32 rounding of
// the diagonal; every Chebyshev step uses the inverse of those FP32 blocks).
// The restriction
// into a level that is not the coarsest is fused with that level's first
// Chebyshev step.
//
// The outer solver is CG preconditioned with one V-cycle per iteration: x
// and r in FP64, p and q in FP64 (FP32 when delta > 0), z from the V-cycle;
// beta = (r - r_old).z / rho_old = -alpha q.z / rho_old (flexible CG, robust
// to a preconditioner applied in lower precision), and
// CuPy's stopping rule on the residual: ||r|| <= tol ||b|| after the update,
// at least one iteration, b = 0 gives 0.  The products A p use the FP64
// form, or (delta > 0) the FP32 form with an FP64 residual replacement
// r = b - A x whenever the recurrence residual has dropped by delta and
// before stopping, so that the stopping test is on the FP64 residual.
// When a replacement finds the FP64 residual more than 10x above the
// recurrence's, or p.q is not positive, the FP32 products can no longer
// carry the refinement (tight tolerances on the anchored operators, whose
// condition number times the FP32 rounding nears 1): the rest of the solve
// takes its products in FP64 (a flag in the CG state, read by the product
// kernel, so captured iterations follow it too).
// K columns (K = 3 for the Gauss-Seidel solve) share the products and the
// V-cycle; each keeps its own scalars and stops on its own.  Every kernel
// returns at once when the solve is done, so the host enqueues batches of
// iterations and reads the state between them; after the first iteration
// of a batch the iterations are replayed from captured chunks (hipGraph).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <hip/hip_ext.h>

#include "device.h"
#include "mg.h"
#include "mg_dev.h"

namespace cpm {

// ---------------------------------------------------------------- kernels
// V-cycle kernels run while any of the solve's K columns is live.

// fpart (floating hierarchy, level 0): partials of the per-component sums
// of b64, fn of them at stride fstride
// The elementwise level kernels (first step, prolongation, restriction
// with the coarse level's first step) run over the level's row grid in the
// XCD order of its sweeps (xg = the level's GridArg::xg), so that each XCD
// writes the rows its L2 then serves to the sweep that follows.
template <int BS, int K, class T, bool ZERO, bool FROM64>
__global__ void __launch_bounds__(kBlock) k_mg_first(int32_t nb, const double* __restrict__ b64, T* __restrict__ b,
                                                     T* __restrict__ r, T* __restrict__ d, T* __restrict__ x,
                                                     const T* __restrict__ dinv, double c0, const CgState* st,
                                                     int32_t xg, const double* __restrict__ fpart = nullptr,
                                                     int64_t fstride = 0, int32_t fn = 0) {
  if (all_done<K>(st)) return;
  mg_first_rows<BS, K, T, ZERO, FROM64>(xcd_perm(blockIdx.x, gridDim.x, xg), gridDim.x, nb, b64, b, r, d, x, dinv,
                                        c0, fpart, fstride, fn);
}

// Periodic stiffness K from the anchored operator A of a floating
// hierarchy (A: node 0's rows and columns replaced by the identity, every
// other entry that of K; K annihilates the bs translations of a periodic
// mesh).  Rows i != 0: the block in the column of node 0 is minus the sum
// of the row's other blocks (diagonal block included), per component.
// Stencil SELL form, FP64 values v and diagonal blocks dg (copies of A's,
// changed in place).
template <int BS>
__global__ void __launch_bounds__(kBlock) k_mg_float_rows(int32_t nb, int32_t nst, GridArg ga, double* __restrict__ v,
                                                          const double* __restrict__ dg) {
  constexpr int BB = BS * BS;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t row = 1 + (int64_t)blockIdx.x * kBlock + threadIdx.x; row < nb; row += stride) {
    const int64_t s0 = (row >> 6) * nst, l = row & 63;
    const int32_t nyz = ga.ny * ga.nz;
    const int32_t ri = (int32_t)row / nyz, rj = ((int32_t)row / ga.nz) % ga.ny, rk = (int32_t)row % ga.nz;
    int64_t zslot = -1;
    double sum[BB];
#pragma unroll
    for (int t = 0; t < BB; ++t) sum[t] = dg[row * BB + t];
    for (int32_t k = 0; k < nst; ++k) {
      const int32_t code = ga.soff[k];
      int32_t i2 = ri + (code & 3) - 1, j2 = rj + ((code >> 2) & 3) - 1, k2 = rk + ((code >> 4) & 3) - 1;
      i2 = (i2 + ga.nx) % ga.nx, j2 = (j2 + ga.ny) % ga.ny, k2 = (k2 + ga.nz) % ga.nz;
      if (i2 == 0 && j2 == 0 && k2 == 0) {
        zslot = s0 + k;
        continue;
      }
#pragma unroll
      for (int t = 0; t < BB; ++t) sum[t] += v[((s0 + k) * BB + t) * 64 + l];
    }
    if (zslot < 0) continue;
#pragma unroll
    for (int t = 0; t < BB; ++t) v[(zslot * BB + t) * 64 + l] = -sum[t];
  }
}

// Row 0 of K (one thread, after k_mg_float_rows): the block in the column
// of neighbour c is the transpose of K's block (c, 0) (K is symmetric), the
// diagonal block minus the sum of them.
template <int BS>
__global__ void k_mg_float_row0(int32_t nst, GridArg ga, double* __restrict__ v, double* __restrict__ dg) {
  constexpr int BB = BS * BS;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double dsum[BB];
#pragma unroll
  for (int t = 0; t < BB; ++t) dsum[t] = 0.0;
  for (int32_t k = 0; k < nst; ++k) {
    const int32_t code = ga.soff[k];
    const int32_t di = (code & 3) - 1, dj = ((code >> 2) & 3) - 1, dk = ((code >> 4) & 3) - 1;
    const int32_t i2 = (di + ga.nx) % ga.nx, j2 = (dj + ga.ny) % ga.ny, k2 = (dk + ga.nz) % ga.nz;
    const int64_t c = ((int64_t)i2 * ga.ny + j2) * ga.nz + k2;
    // the slot of row c that points back to node 0: the opposite offset
    const int32_t back = (1 - di) | ((1 - dj) << 2) | ((1 - dk) << 4);
    int64_t cs = -1;
    for (int32_t q = 0; q < nst; ++q)
      if (ga.soff[q] == back) cs = (c >> 6) * nst + q;
    if (cs < 0) continue;  // no opposite offset: not a Kuhn stencil, which cpm_mg_create rejects
    const int64_t lc = c & 63;
#pragma unroll
    for (int d = 0; d < BS; ++d)
#pragma unroll
      for (int e = 0; e < BS; ++e) {
        const double a = v[(cs * BB + e * BS + d) * 64 + lc];
        v[(k * BB + d * BS + e) * 64] = a;  // row 0: slice 0, lane 0
        dsum[d * BS + e] += a;
      }
  }
#pragma unroll
  for (int t = 0; t < BB; ++t) dg[t] = -dsum[t];
}

#ifndef CPM_SWEEP_WAVES
#define CPM_SWEEP_WAVES 1
#endif
#ifndef CPM_OUTER_WAVES
#define CPM_OUTER_WAVES 1
#endif
template <int BS, int K, class T, int MODE, bool ST, class VT>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(CPM_SWEEP_WAVES))) k_mg_sweep(int32_t nb, const int64_t* __restrict__ sptr,
                                                     const int32_t* __restrict__ col, const VT* __restrict__ val,
                                                     const sell_diag_t<VT>* __restrict__ dg, GridArg ga,
                                                     const T* __restrict__ din, T* __restrict__ dout,
                                                     T* __restrict__ x, T* __restrict__ r, const T* __restrict__ b,
                                                     const T* __restrict__ dinv, double c1, double c2,
                                                     double vscale, const CgState* st) {
  if (all_done<K>(st)) return;
  mg_sweep_rows<BS, K, T, MODE, ST, VT>(xcd_perm(blockIdx.x, gridDim.x, ga.xg), gridDim.x, nb, sptr, col, val, dg, ga, din, dout, x, r, b, dinv,
                                        c1, c2, vscale);
}

// Low-precision form of a level operator: VT = _Float16: v16 =
// fp16(v64 / vscale) in the SELL layout (padding lanes included; for 3x3
// blocks a lane's entries 0..7 contiguous, 16 bytes, so that one 128-bit
// load fetches them, entry 8 of the 64 lanes after them); VT = float: fp32(v64)
// (vscale 1) in the layout of v64.  dgh = dg64 + the rounding of the row's
// off-diagonal values, so that the row sums with dgh equal the FP64 ones
// (the FP16 forms keep fp32(dgh): equal to the FP32 rounding of the diagonal).
template <int BS, class VT>
__global__ void __launch_bounds__(kBlock) k_mg_make_low(int32_t nb, int64_t nrows_pad, const int64_t* __restrict__ sptr,
                                                        int32_t nst, int stencil, const double* __restrict__ v64,
                                                        const double* __restrict__ dg64, double vscale,
                                                        VT* __restrict__ v16, double* __restrict__ dgh) {
  constexpr int BB = BS * BS;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x; row < nrows_pad; row += stride) {
    const int64_t s = row >> 6, l = row & 63;
    const int64_t s0 = stencil ? s * nst : sptr[s];
    const int64_t s1 = stencil ? s0 + nst : sptr[s + 1];
    double comp[BB];
#pragma unroll
    for (int t = 0; t < BB; ++t) comp[t] = 0.0;
    for (int64_t slot = s0; slot < s1; ++slot)
#pragma unroll
      for (int t = 0; t < BB; ++t) {
        const int64_t i = (slot * BB + t) * 64 + l;
        const double a = v64[i];
        const VT h = (VT)(a / vscale);
        const int64_t o =
            (sizeof(VT) == 2 && BB == 9) ? slot * 576 + (t < 8 ? 8 * l + t : 512 + l) : i;
        v16[o] = h;
        comp[t] += a - (double)h * vscale;
      }
    if (row < nb)
#pragma unroll
      for (int t = 0; t < BB; ++t) dgh[row * BB + t] = dg64[row * BB + t] + comp[t];
  }
}

// Symmetry of a Kuhn stencil operator, slot by slot: amax[0] = max over the
// backward blocks of |A(i, i-d) - A(i-d, i)^T| (the mirrored forward block
// of the neighbour row), amax[1] = max |A(i, i-d)|, as FP64 bit patterns
// (non-negative values order like their bits).
template <int BS>
__global__ void __launch_bounds__(kBlock) k_mg_kuhn_asym(int32_t nb, GridArg ga, const double* __restrict__ v64,
                                                         unsigned long long* __restrict__ amax) {
  constexpr int BB = BS * BS;
  double dm = 0.0, am = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x; row < nb; row += stride) {
    const int64_t s = row >> 6, l = row & 63;
    for (int t = 0; t < 7; ++t) {
      const int64_t c = kuhn_neighbour(row, ga, t);
      const int64_t cs = (c >> 6) * kKuhnSlots + (kKuhnSlots - 1 - t), lc = c & 63;
#pragma unroll
      for (int d = 0; d < BS; ++d)
#pragma unroll
        for (int e = 0; e < BS; ++e) {
          const double a = v64[((s * kKuhnSlots + t) * BB + d * BS + e) * 64 + l];
          const double an = v64[(cs * BB + e * BS + d) * 64 + lc];
          dm = fmax(dm, fabs(a - an));
          am = fmax(am, fabs(a));
        }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    dm = fmax(dm, __shfl_xor(dm, o));
    am = fmax(am, __shfl_xor(am, o));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&amax[0], (unsigned long long)__double_as_longlong(dm));
    atomicMax(&amax[1], (unsigned long long)__double_as_longlong(am));
  }
}

// FP16 half form of a symmetric Kuhn stencil operator (make_half_row).
template <int BS>
__global__ void __launch_bounds__(kBlock) k_mg_make_half(int32_t nb, int64_t nrows_pad, GridArg ga,
                                                         const double* __restrict__ v64,
                                                         const double* __restrict__ dg64, double vscale,
                                                         _Float16* __restrict__ v16, double* __restrict__ dgh) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x; row < nrows_pad; row += stride)
    make_half_row<BS>(row, nb, ga, v64, dg64, vscale, v16, dgh);
}

template <int W, int K, class T>
__global__ void __launch_bounds__(kBlock) k_mg_restrict(Dims3 fine, Dims3 coarse, int anchor,
                                                        const T* __restrict__ rf, T* __restrict__ bc,
                                                        const CgState* st) {
  if (all_done<K>(st)) return;
  mg_restrict_rows<W, T>(blockIdx.x, gridDim.x, fine, coarse, anchor != 0, rf, bc);
}

template <int BS, int K, class T>
__global__ void __launch_bounds__(kBlock) k_mg_restrict_first(Dims3 fine, Dims3 coarse, int anchor,
                                                              const T* __restrict__ rf, T* __restrict__ bc,
                                                              T* __restrict__ rc, T* __restrict__ dc,
                                                              T* __restrict__ xc, const T* __restrict__ dinvc,
                                                              double c0, const CgState* st, int32_t xg) {
  if (all_done<K>(st)) return;
  mg_restrict_first_rows<BS, K, T>(xcd_perm(blockIdx.x, gridDim.x, xg), gridDim.x, fine, coarse, anchor != 0, rf, bc,
                                   rc, dc, xc, dinvc, c0);
}

template <int W, int K, class T>
__global__ void __launch_bounds__(kBlock) k_mg_prolong(Dims3 fine, Dims3 coarse, int anchor,
                                                       const T* __restrict__ ec, T* __restrict__ xf,
                                                       const CgState* st, int32_t xg) {
  if (all_done<K>(st)) return;
  mg_prolong_rows<W, T>(xcd_perm(blockIdx.x, gridDim.x, xg), gridDim.x, fine, coarse, anchor != 0, ec, xf);
}

template <int K, class T>
__global__ void __launch_bounds__(kBlock) k_mg_dense(int64_t n, const T* __restrict__ cinv, const T* __restrict__ b,
                                                     T* __restrict__ x, const CgState* st) {
  if (all_done<K>(st)) return;
  mg_dense_rows<K, T>(blockIdx.x, gridDim.x, n, cinv, b, x);
}

template <int BS, bool ST>
__global__ void __launch_bounds__(kBlock) k_mg_power(int32_t nb, const int64_t* __restrict__ sptr,
                                                     const int32_t* __restrict__ col, const double* __restrict__ val,
                                                     const double* __restrict__ dg, GridArg ga,
                                                     const double* __restrict__ x, double* __restrict__ y,
                                                     const double* __restrict__ dinv, double* partials) {
  double acc[1] = {0.0};
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x; row < (int64_t)((nb + 63) & ~63); row += stride) {
    double sum[BS], xo[BS], dv[BS];
    sell_row<BS, 1, double, false, ST>(row, nb, sptr, col, val, dg, x, nullptr, ga, sum, xo);
    if (row >= nb) continue;
    apply_dinv<BS, 1, double>(dinv, row, sum, dv);
#pragma unroll
    for (int c = 0; c < BS; ++c) {
      y[row * BS + c] = dv[c];
      acc[0] += dv[c] * dv[c];
    }
  }
  block_partial<1>(acc, partials);
}

__global__ void __launch_bounds__(kBlock) k_scale(int64_t n, const double* __restrict__ y, double* __restrict__ x,
                                                  double s) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) x[i] = s * y[i];
}

// lam = sqrt(sum of the partials), one workgroup (power iteration)
__global__ void __launch_bounds__(kRedBlock) k_power_norm(int nblk, const double* __restrict__ partials,
                                                          double* __restrict__ lam) {
  __shared__ double lds[kRedBlock / 64];
  double v[1] = {0.0};
  for (int i = threadIdx.x; i < nblk; i += kRedBlock) v[0] += partials[i];
  block_sum_t<1, kRedBlock>(v, lds);
  if (threadIdx.x == 0) *lam = sqrt(v[0]);
}

// x = y / lam (lam > 0; otherwise x is left alone and the host reports it)
__global__ void __launch_bounds__(kBlock) k_scale_inv(int64_t n, const double* __restrict__ y, double* __restrict__ x,
                                                      const double* __restrict__ lam) {
  const double l = *lam;
  if (!(l > 0.0)) return;
  const double s = 1.0 / l;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) x[i] = s * y[i];
}

// ---- outer preconditioned CG (FP64) ------------------------------------------
// K columns interleaved (entry i belongs to column i % K, state st[i % K]).

// q = A p (SELL form with VT values: FP64, or FP32 with compensated FP64
// diagonals) and partials of p_j.q_j; PLAIN: y = A x, no partials.
template <int BS, int K, bool ST, bool PLAIN, class VT = double>
__global__ void __launch_bounds__(kBlock) k_pcg_spmv(int32_t nb, const int64_t* __restrict__ sptr,
                                                     const int32_t* __restrict__ col, const VT* __restrict__ val,
                                                     const double* __restrict__ dg, GridArg ga,
                                                     const double* __restrict__ p, double* __restrict__ q,
                                                     const CgState* st, double* partials) {
  if (!PLAIN && all_done<K>(st)) return;
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) acc[j] = 0.0;
  sell_rows<BS, K, double, !PLAIN, false, ST, VT>(xcd_perm(blockIdx.x, gridDim.x, ga.xg), gridDim.x, nb, sptr, col, val, dg, p, nullptr, q, ga,
                                                  acc);
  if (!PLAIN) block_partial<K>(acc, partials);
}

// q = A p of the mixed-precision PCG: the FP32 form with compensated FP64
// diagonals, or the FP64 form once the solve has switched its products to
// FP64 (CgState::p64); partials of p_j.q_j.  p and q are FP32 (the sums in
// FP64, q rounded once; its p.q partials use the rounded q that r -= alpha q
// reads).  The FP64 products walk the
// stored stencil offsets two slots at a time (the same slots in the same
// order as the compile-time Kuhn walk, so the same sums): the kernel's
// registers, and so its occupancy, are those of the FP32 path.
template <int BS, int K, bool ST>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(CPM_OUTER_WAVES))) k_pcg_spmv_rel(int32_t nb, const int64_t* __restrict__ sptr,
                                                         const int32_t* __restrict__ col, const float* __restrict__ v32,
                                                         const double* __restrict__ dgc,
                                                         const double* __restrict__ v64,
                                                         const double* __restrict__ dg64, GridArg ga,
                                                         const float* __restrict__ p, float* __restrict__ q,
                                                         const CgState* st, double* partials) {
  if (all_done<K>(st)) return;
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) acc[j] = 0.0;
  if (any_p64<K>(st)) {
    GridArg g64 = ga;
    g64.kuhn = 0;
    sell_rows<BS, K, float, true, false, ST, double, 2>(xcd_perm(blockIdx.x, gridDim.x, ga.xg), gridDim.x, nb, sptr,
                                                        col, v64, dg64, p, nullptr, q, g64, acc);
  } else
    sell_rows<BS, K, float, true, false, ST, float>(xcd_perm(blockIdx.x, gridDim.x, ga.xg), gridDim.x, nb, sptr, col,
                                                    v32, dgc, p, nullptr, q, ga, acc);
  block_partial<K>(acc, partials);
}

// reliable update, run when a column asked for one: r = b - A x (FP64
// form) on those columns, partials of their r.r (SpMV workgroup layout)
template <int BS, int K, bool ST>
__global__ void __launch_bounds__(kBlock) k_pcg_true(int32_t nb, const int64_t* __restrict__ sptr,
                                                     const int32_t* __restrict__ col, const double* __restrict__ val,
                                                     const double* __restrict__ dg, GridArg ga,
                                                     const double* __restrict__ x, const double* __restrict__ b,
                                                     double* __restrict__ r, const CgState* st, double* partials) {
  if (!any_upd<K>(st)) return;
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) acc[j] = 0.0;
  pcg_true_rows<BS, K, ST>(xcd_perm(blockIdx.x, gridDim.x, ga.xg), gridDim.x, nb, sptr, col, val, dg, ga, x, b, r, st, acc);
  block_partial<K>(acc, partials);
}

// r = b - y (use_x0) or x = 0, r = b; partials of ||b_j||^2, ||r_j||^2.
template <int K>
__global__ void __launch_bounds__(kBlock) k_pcg_init(int64_t n, const double* __restrict__ b, double* __restrict__ x,
                                                     const double* __restrict__ y, double* __restrict__ r,
                                                     int use_x0, double* partials) {
  double acc[2 * K];
#pragma unroll
  for (int j = 0; j < 2 * K; ++j) acc[j] = 0.0;
  pcg_init_rows<K>(blockIdx.x, gridDim.x, n, b, x, y, r, use_x0, acc);
  block_partial<2 * K>(acc, partials);
}

template <int K, bool REL>
__global__ void __launch_bounds__(kRedBlock) k_pcg_red_init(int nblk, const double* __restrict__ partials,
                                                            CgState* st, double tol, int32_t maxiter) {
  __shared__ double lds[2 * K * (kRedBlock / 64)];
  double v[2 * K];
#pragma unroll
  for (int j = 0; j < 2 * K; ++j) v[j] = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kRedBlock)
#pragma unroll
    for (int j = 0; j < 2 * K; ++j) v[j] += partials[j * nblk + i];
  block_sum_t<2 * K, kRedBlock>(v, lds);
  if (threadIdx.x == 0) cg_state_init<K, REL>(st, v, tol, maxiter);
}

template <int K>
__global__ void __launch_bounds__(kBlock) k_pcg_zero_b(int64_t n, double* __restrict__ x, const CgState* st) {
  zero_b_rows<K>(blockIdx.x, gridDim.x, n, x, st);
}

// live columns: x += alpha p, r -= alpha q; partials of r_j.r_j.
template <int K, class PQ>
__global__ void __launch_bounds__(kBlock) k_pcg_xr(int64_t n, double* __restrict__ x, double* __restrict__ r,
                                                   const PQ* __restrict__ p, const PQ* __restrict__ q,
                                                   const CgState* st, double* partials) {
  if (all_done<K>(st)) return;
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) acc[j] = 0.0;
  pcg_xr_rows<K, PQ>(blockIdx.x, gridDim.x, n, st, x, r, p, q, acc);
  block_partial<K>(acc, partials);
}

// partials of r_j.z_j and q_j.z_j (FB: see pcg_z)
template <int K, class T, int FB, class PQ>
__global__ void __launch_bounds__(kBlock) k_pcg_dot(int64_t n, const double* __restrict__ r,
                                                    const PQ* __restrict__ q, const T* __restrict__ z,
                                                    const CgState* st, double* partials) {
  if (all_done<K>(st)) return;
  double acc[2 * K];
#pragma unroll
  for (int j = 0; j < 2 * K; ++j) acc[j] = 0.0;
  pcg_dot_rows<K, T, FB, PQ>(blockIdx.x, gridDim.x, n, st, r, q, z, acc);
  block_partial<2 * K>(acc, partials);
}

// partials of the per-component sums of r (floating hierarchy: read by the
// first Chebyshev step of level 0)
template <int NV, int K>
__global__ void __launch_bounds__(kBlock) k_pcg_rsum(int64_t n, const double* __restrict__ r, const CgState* st,
                                                     double* partials) {
  if (all_done<K>(st)) return;
  double acc[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) acc[c] = 0.0;
  pcg_rsum_rows<NV>(blockIdx.x, gridDim.x, n, r, acc);
  block_partial<NV>(acc, partials);
}

// z = F^T y + e_0 r_0 in FP64 (the preconditioner of a floating hierarchy,
// for cpm_mg_apply)
template <int BS, class T>
__global__ void __launch_bounds__(kBlock) k_mg_float_out(int64_t n, const T* __restrict__ y,
                                                         const double* __restrict__ r, double* __restrict__ z) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) z[i] = pcg_z<BS, T>(i, y, r);
}

// live columns: rho = r.z; beta = -alpha q.z / rho_old (flexible CG)
template <int K>
__global__ void __launch_bounds__(kRedBlock) k_pcg_red_rz(int nblk, const double* __restrict__ partials,
                                                          CgState* st) {
  if (all_done<K>(st)) return;
  __shared__ double lds[2 * K * (kRedBlock / 64)];
  double v[2 * K];
#pragma unroll
  for (int j = 0; j < 2 * K; ++j) v[j] = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kRedBlock)
#pragma unroll
    for (int j = 0; j < 2 * K; ++j) v[j] += partials[j * nblk + i];
  block_sum_t<2 * K, kRedBlock>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (!st[j].done) pcg_state_rz(st[j], v[j], v[K + j]);
}

template <int K, class T, int FB, class PQ>
__global__ void __launch_bounds__(kBlock) k_pcg_pupd(int64_t n, const T* __restrict__ z, PQ* __restrict__ p,
                                                     const double* __restrict__ r, const CgState* st) {
  if (all_done<K>(st)) return;
  pcg_pupd_rows<K, T, FB, PQ>(blockIdx.x, gridDim.x, n, st, z, p, r);
}

// ---------------------------------------------------------------- host

static GridArg grid_arg(const Sell& S) { return CPM_GRID_ARG(S); }


// smoothing values of a level (FP64 or FP32 form): the recovered periodic
// stiffness on level 0 of a floating hierarchy, else the level's operator
template <class T>
static const T* lvals(const MgLevel& Lv) {
  if constexpr (sizeof(T) == 8)
    return Lv.kv64 ? Lv.kv64 : Lv.A->sell.v64;
  else
    return Lv.kv32 ? Lv.kv32 : Lv.A->sell.v32;
}
template <class T>
static const double* ldiag(const MgLevel& Lv) {
  if constexpr (sizeof(T) == 8)
    return Lv.kdg64 ? Lv.kdg64 : Lv.A->sell.dg64;
  else
    return Lv.kdgc ? Lv.kdgc : Lv.A->sell.dgc;
}

template <int BS, int K, class T, int MODE, bool ST, class VT>
static void launch_sweep(hipStream_t s, int g, const Sell& S, const VT* val, const sell_diag_t<VT>* dg,
                         const MgLevel& Lv,
                         const void* din, void* dout, double c1, double c2, double vscale, const CgState* st,
                         hipEvent_t* ev) {
  // values keeping half of the stencil: the FP16 form (k_mg_make_half), or
  // the level's own FP32 form
  GridArg ga = grid_arg(S);
  if (sizeof(VT) == 2 ? Lv.khalf != 0 : ((const void*)val == (const void*)S.v32 && S.half32)) ga.kuhn = 2;
  // the level's own FP64 form: its half copy when there is one
  if constexpr (sizeof(VT) == 8)
    if (val == S.v64 && S.half64) {
      val = S.v64h;
      ga = CPM_GRID_ARG64(S);
    }
  if (ev)
    hipExtLaunchKernelGGL(k_mg_sweep<BS, K, T, MODE, ST, VT>, dim3(g), dim3(kBlock), 0, s, ev[0], ev[1], 0,
                          (int32_t)S.nbrows, (const int64_t*)S.sptr, (const int32_t*)S.col, val, dg, ga,
                          (const T*)din, (T*)dout, (T*)Lv.x, (T*)Lv.r, (const T*)Lv.b, (const T*)Lv.dinv, c1, c2,
                          vscale, st);
  else
    k_mg_sweep<BS, K, T, MODE, ST, VT><<<g, kBlock, 0, s>>>(S.nbrows, S.sptr, S.col, val, dg, ga,
                                                           (const T*)din, (T*)dout, (T*)Lv.x, (T*)Lv.r,
                                                           (const T*)Lv.b, (const T*)Lv.dinv, c1, c2, vscale, st);
}

// ev (profiling, may be null): start and stop events taken by the dispatch
// itself (hipExtLaunchKernelGGL), so they time the kernel alone
template <int BS, int K, class T, int MODE>
static void sweep(cpm_mg* M, const MgLevel& Lv, const void* din, void* dout, double c1, double c2,
                  const CgState* st, hipEvent_t* ev = nullptr) {
  const Sell& S = Lv.A->sell;
  hipStream_t s = M->ls ? M->ls : M->ctx->stream;
  const int g = sell_grid(S);
  if constexpr (sizeof(T) == 4) {
    if (M->half) {
      const _Float16* v16 = (const _Float16*)Lv.v16;
      if (S.stencil)
        launch_sweep<BS, K, T, MODE, true>(s, g, S, v16, Lv.dg32, Lv, din, dout, c1, c2, Lv.vscale, st, ev);
      else
        launch_sweep<BS, K, T, MODE, false>(s, g, S, v16, Lv.dg32, Lv, din, dout, c1, c2, Lv.vscale, st, ev);
      return;
    }
  }
  if (S.stencil)
    launch_sweep<BS, K, T, MODE, true>(s, g, S, lvals<T>(Lv), ldiag<T>(Lv), Lv, din, dout, c1, c2, 1.0, st, ev);
  else
    launch_sweep<BS, K, T, MODE, false>(s, g, S, lvals<T>(Lv), ldiag<T>(Lv), Lv, din, dout, c1, c2, 1.0, st, ev);
}

// One V-cycle from level l for K columns: b of level 0 read from the FP64
// vector b64.  ev (profiling, may be null): two event pairs, taken by the
// first Chebyshev step of level 0 in the pre-smoothing (ev[0], ev[1]) and in
// the post-smoothing (ev[2], ev[3]).
// first_done: the first Chebyshev step of level l already ran (fused with
// the restriction from level l-1).
template <int BS, int K, class T>
static void vcycle(cpm_mg* M, size_t l, const double* b64, const CgState* st, hipEvent_t* ev = nullptr,
                   bool first_done = false) {
  MgLevel& Lv = M->L[l];
  hipStream_t s = M->ls ? M->ls : M->ctx->stream;
  // the level's sweep grid and XCD order (the coarsest level has no sweeps)
  const int gn = l + 1 == M->L.size() ? grid_for(Lv.nb) : sell_grid(Lv.A->sell);
  const int32_t xg = Lv.A->sell.xg;
  const double* c1 = &M->c1[l * 8];
  const double* c2 = &M->c2[l * 8];
  void* d[2] = {Lv.d0, Lv.d1};
  if (l + 1 == M->L.size()) {
    if (M->csmooth) {
      // Chebyshev solve: the first step ran with the restriction
      for (int k = 1; k < M->csmooth; ++k)
        sweep<BS, K, T, 0>(M, Lv, d[(k - 1) & 1], d[k & 1], c1[k], c2[k], st);
      return;
    }
    const int g = (int)std::min<int64_t>(kMaxGrid, (Lv.n + 3) / 4);
    k_mg_dense<K, T><<<g, kBlock, 0, s>>>(Lv.n, (const T*)M->cinv, (const T*)Lv.b, (T*)Lv.x, st);
    return;
  }
  const int m = M->level_degree(l);
  const double c0 = M->c0[l];
  // pre-smoothing from x = 0 (floating: b64's component sums are in the
  // partials of k_pcg_rsum, launched just before over grid_for(n) workgroups)
  if (l == 0 && b64) {
    const int gf = grid_for(Lv.n * K);
    k_mg_first<BS, K, T, true, true><<<gn, kBlock, 0, s>>>(Lv.nb, b64, (T*)Lv.b, (T*)Lv.r, (T*)d[0], (T*)Lv.x,
                                                           (const T*)Lv.dinv, c0, st, xg,
                                                           M->floating ? M->lpart : nullptr, gf, gf);
  }
  else if (!first_done)
    k_mg_first<BS, K, T, true, false><<<gn, kBlock, 0, s>>>(Lv.nb, nullptr, (T*)Lv.b, (T*)Lv.r, (T*)d[0],
                                                            (T*)Lv.x, (const T*)Lv.dinv, c0, st, xg);
  for (int k = 1; k < m; ++k)
    sweep<BS, K, T, 0>(M, Lv, d[(k - 1) & 1], d[k & 1], c1[k], c2[k], st, (k == 1 && ev) ? ev : nullptr);
  sweep<BS, K, T, 1>(M, Lv, d[(m - 1) & 1], nullptr, 0.0, 0.0, st);
  // coarse-grid correction
  MgLevel& Lc = M->L[l + 1];
  const Dims3 fd{{Lv.dims[0], Lv.dims[1], Lv.dims[2]}}, cd{{Lc.dims[0], Lc.dims[1], Lc.dims[2]}};
  if (l + 2 < M->L.size() || M->csmooth) {
    // restriction fused with the coarse level's first step
    k_mg_restrict_first<BS, K, T><<<sell_grid(Lc.A->sell), kBlock, 0, s>>>(
        fd, cd, M->anchor, (const T*)Lv.r, (T*)Lc.b, (T*)Lc.r, (T*)Lc.d0, (T*)Lc.x, (const T*)Lc.dinv, M->c0[l + 1],
        st, Lc.A->sell.xg);
    vcycle<BS, K, T>(M, l + 1, nullptr, st, nullptr, true);
  } else {
    k_mg_restrict<BS * K, K, T><<<grid_for(Lc.nb), kBlock, 0, s>>>(fd, cd, M->anchor, (const T*)Lv.r, (T*)Lc.b,
                                                                  st);
    vcycle<BS, K, T>(M, l + 1, nullptr, st);
  }
  k_mg_prolong<BS * K, K, T><<<gn, kBlock, 0, s>>>(fd, cd, M->anchor, (const T*)Lc.x, (T*)Lv.x, st, xg);
  // post-smoothing: the same polynomial on the new residual (degree >= 2:
  // the residual and the first step's direction in one pass, its x update
  // in the second step's pass)
  if (m >= 2) {
    sweep<BS, K, T, 3>(M, Lv, nullptr, d[0], 0.0, c0, st);
    sweep<BS, K, T, 4>(M, Lv, d[0], d[1], c1[1], c2[1], st, ev ? ev + 2 : nullptr);
    for (int k = 2; k < m; ++k) sweep<BS, K, T, 0>(M, Lv, d[(k - 1) & 1], d[k & 1], c1[k], c2[k], st);
  } else {
    sweep<BS, K, T, 2>(M, Lv, nullptr, nullptr, 0.0, 0.0, st);
    k_mg_first<BS, K, T, false, false><<<gn, kBlock, 0, s>>>(Lv.nb, nullptr, (T*)Lv.b, (T*)Lv.r, (T*)d[0],
                                                             (T*)Lv.x, (const T*)Lv.dinv, c0, st, xg);
  }
}

static void vcycle_any(cpm_mg* M, const double* b64, const CgState* st) {
  if (M->bs == 3) {
    if (M->prec == 32) vcycle<3, 1, float>(M, 0, b64, st);
    else vcycle<3, 1, double>(M, 0, b64, st);
  } else {
    if (M->prec == 32) vcycle<1, 1, float>(M, 0, b64, st);
    else vcycle<1, 1, double>(M, 0, b64, st);
  }
}

template <int BS>
// 20 power iterations on D^-1 A without host round trips (the norm stays on
// the device, in lam_dev); one read at the end.  The partials buffer
// (sell_grid(S) doubles) belongs to the hierarchy being built, not to the
// context, so that hierarchies built from several host threads on one
// context (their kernels interleaved on its stream) do not mix their sums.
static int power_lmax(cpm_mg* M, MgLevel& Lv, double* x, double* y, const double* dinv64, double* lam_dev,
                      double* part, double* lmax) {
  cpm_ctx* ctx = M->ctx;
  hipStream_t s = ctx->stream;
  const Sell& S = Lv.A->sell;
  const int g = sell_grid(S);
  std::vector<double> h(Lv.n);
  uint64_t st = 0x9E3779B97F4A7C15ull;
  for (auto& v : h) {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    v = 0.5 + (double)(st >> 11) * (1.0 / 9007199254740992.0);
  }
  for (int c = 0; c < BS; ++c) h[c] = 0.0;  // node 0 (the anchor of anchored operators)
  CPM_HIP(hipMemcpyAsync(x, h.data(), sizeof(double) * Lv.n, hipMemcpyHostToDevice, s));
  double nrm = 0.0;
  for (double v : h) nrm += v * v;
  nrm = std::sqrt(nrm);
  k_scale<<<grid_for(Lv.n), kBlock, 0, s>>>(Lv.n, x, x, 1.0 / nrm);
  const double *v64 = lvals<double>(Lv), *dg64 = ldiag<double>(Lv);
  for (int it = 0; it < 20; ++it) {
    if (S.stencil)
      k_mg_power<BS, true><<<g, kBlock, 0, s>>>(S.nbrows, S.sptr, S.col, v64, dg64, grid_arg(S), x, y, dinv64,
                                                part);
    else
      k_mg_power<BS, false><<<g, kBlock, 0, s>>>(S.nbrows, S.sptr, S.col, v64, dg64, grid_arg(S), x, y, dinv64,
                                                 part);
    k_power_norm<<<1, kRedBlock, 0, s>>>(g, part, lam_dev);
    k_scale_inv<<<grid_for(Lv.n), kBlock, 0, s>>>(Lv.n, y, x, lam_dev);
  }
  CPM_HIP(hipGetLastError());
  double lam = 0.0;
  CPM_HIP(hipMemcpyAsync(&lam, lam_dev, sizeof(double), hipMemcpyDeviceToHost, s));
  CPM_HIP(hipStreamSynchronize(s));
  if (!(lam > 0.0)) return fail(CPM_EINVAL, "multigrid: operator annihilates the power-iteration vector");
  *lmax = lam;
  return CPM_OK;
}

// inverse of every diagonal block (host, FP64)
static bool invert_blocks(int bs, int64_t nb, const std::vector<double>& dg, std::vector<double>& out) {
  out.assign(dg.size(), 0.0);
  for (int64_t i = 0; i < nb; ++i) {
    const double* a = &dg[i * bs * bs];
    double* o = &out[i * bs * bs];
    if (bs == 1) {
      if (a[0] == 0.0) return false;
      o[0] = 1.0 / a[0];
      continue;
    }
    const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                       a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (det == 0.0 || !std::isfinite(det)) return false;
    const double id = 1.0 / det;
    o[0] = (a[4] * a[8] - a[5] * a[7]) * id;
    o[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    o[3] = (a[5] * a[6] - a[3] * a[8]) * id;
    o[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    o[6] = (a[3] * a[7] - a[4] * a[6]) * id;
    o[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
  }
  return true;
}

static int mg_alloc(cpm_mg* M, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 8));
  if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? CPM_ENOMEM : CPM_EHIP, hipGetErrorString(e));
  M->allocs.push_back(*p);
  return CPM_OK;
}

// Level and outer vectors for K columns (allocated for one column at
// creation, grown on the first solve with more).
int mg_reserve(cpm_mg* M, int K, hipStream_t s) {
  if (K <= M->kcap) return CPM_OK;
  if (M->kcap == 0) return fail(CPM_ENOMEM, "multigrid: hierarchy unusable after a failed allocation");
  cpm_ctx* ctx = M->ctx;
  if (!s) s = ctx->stream;
  CPM_HIP(hipStreamSynchronize(s));
  // captured chunks refer to the vectors replaced below
  for (auto& c : M->chunks) (void)hipGraphExecDestroy(c.exec);
  M->chunks.clear();
  const size_t ts = M->prec == 32 ? 4 : 8;
  auto regrow = [&](void** p, size_t bytes) -> int {
    auto it = std::find(M->allocs.begin(), M->allocs.end(), *p);
    if (it != M->allocs.end()) M->allocs.erase(it);
    if (*p) CPM_HIP(hipFree(*p));
    *p = nullptr;
    return mg_alloc(M, p, bytes);
  };
  // on a failure the vectors are left partly freed: the hierarchy refuses
  // further solves (kcap = 0)
  M->kcap = 0;
  for (MgLevel& Lv : M->L)
    for (void** p : {&Lv.x, &Lv.r, &Lv.b, &Lv.d0, &Lv.d1})
      if (int rc = regrow(p, ts * Lv.n * K)) return rc;
  const int64_t n0 = M->L[0].n;
  for (void** p : {(void**)&M->r, (void**)&M->p, (void**)&M->q})
    if (int rc = regrow(p, 8 * n0 * K)) return rc;
  M->kcap = K;
  return CPM_OK;
}

// Preconditioned CG on level 0 for K columns (see the file comment); column
// j of b and x is entry i*K + j, with its own state ln.st[j].
// delta > 0: the outer products use the FP32 form of the operator, with an
// FP64 residual replacement whenever the recurrence residual has dropped by
// delta since the last one (and before stopping: the stopping test is on
// the FP64 residual); delta = 0: FP64 products.
template <int BS, int K, class T>
static int pcg_run(cpm_mg* M, const PcgLane& ln, const double* b, double* x, int use_x0, double tol, int maxiter,
                   double delta, int hint, int32_t* iters, double* resid, int* failed) {
  cpm_ctx* ctx = M->ctx;
  hipStream_t s = ln.stream;  // launch stream of the lambdas below
  MgLevel& L0 = M->L[0];
  const Sell& S = L0.A->sell;
  const int64_t n = L0.n * K;
  const int gs = sell_grid(S), gv = grid_for(n);
  CgState* st = ln.st;
  double* const partials = ln.partials;
  // the V-cycle launches on the lane's stream and reads its partials
  M->ls = ln.stream;
  M->lpart = partials;
  struct Restore {
    cpm_mg* M;
    ~Restore() {
      M->ls = nullptr;
      M->lpart = nullptr;
    }
  } restore{M};
  const bool rel = delta > 0.0;
  // ev (profiling): the product is timed by its dispatch (mixed, stencil form)
  // p and q: FP32 in the mixed-precision solves (rel), FP64 otherwise; in
  // and out are FP64 for the start's product (plain)
  auto spmv = [&](const void* in, void* out, bool plain, hipEvent_t* ev = nullptr) {
    if (!plain && rel) {
      if (S.stencil && ev)
        hipExtLaunchKernelGGL(k_pcg_spmv_rel<BS, K, true>, dim3(gs), dim3(kBlock), 0, s, ev[0], ev[1], 0,
                              (int32_t)S.nbrows, (const int64_t*)S.sptr, (const int32_t*)S.col,
                              (const float*)S.v32, (const double*)S.dgc, (const double*)S.v64,
                              (const double*)S.dg64, CPM_GRID_ARG32(S), (const float*)in, (float*)out,
                              (const CgState*)st, partials);
      else if (S.stencil)
        k_pcg_spmv_rel<BS, K, true><<<gs, kBlock, 0, s>>>(S.nbrows, S.sptr, S.col, S.v32, S.dgc, S.v64, S.dg64,
                                                          CPM_GRID_ARG32(S), (const float*)in, (float*)out, st,
                                                          partials);
      else
        k_pcg_spmv_rel<BS, K, false><<<gs, kBlock, 0, s>>>(S.nbrows, S.sptr, S.col, S.v32, S.dgc, S.v64, S.dg64,
                                                           CPM_GRID_ARG32(S), (const float*)in, (float*)out, st,
                                                           partials);
      return;
    }
    if (plain) {
      if (S.stencil)
        k_pcg_spmv<BS, K, true, true><<<gs, kBlock, 0, s>>>(S.nbrows, S.sptr, S.col, sell_v64(S), S.dg64,
                                                            CPM_GRID_ARG64(S), (const double*)in, (double*)out, st, partials);
      else
        k_pcg_spmv<BS, K, false, true><<<gs, kBlock, 0, s>>>(S.nbrows, S.sptr, S.col, sell_v64(S), S.dg64,
                                                             CPM_GRID_ARG64(S), (const double*)in, (double*)out, st, partials);
    } else {
      if (S.stencil)
        k_pcg_spmv<BS, K, true, false><<<gs, kBlock, 0, s>>>(S.nbrows, S.sptr, S.col, sell_v64(S), S.dg64,
                                                             CPM_GRID_ARG64(S), (const double*)in, (double*)out, st, partials);
      else
        k_pcg_spmv<BS, K, false, false><<<gs, kBlock, 0, s>>>(S.nbrows, S.sptr, S.col, sell_v64(S), S.dg64,
                                                              CPM_GRID_ARG64(S), (const double*)in, (double*)out, st, partials);
    }
  };
  // floating hierarchy (one column): z = F^T V(F r) + e_0 r_0, the sums of
  // r for F r just before the V-cycle, the shift in the consumers of z
  const bool fl = M->floating != 0;
  if (fl && K != 1) return fail(CPM_EINVAL, "multigrid CG: a floating hierarchy solves one column");
  constexpr int FB = (K == 1) ? BS : 0;
  auto precond = [&](hipEvent_t* ev) {
    if (fl) k_pcg_rsum<BS * K, K><<<gv, kBlock, 0, s>>>(n, M->r, st, partials);
    vcycle<BS, K, T>(M, 0, M->r, st, ev);
    if (fl && rel)
      k_pcg_dot<K, T, FB, float><<<gv, kBlock, 0, s>>>(n, M->r, (const float*)M->q, (const T*)L0.x, st, partials);
    else if (fl)
      k_pcg_dot<K, T, FB, double><<<gv, kBlock, 0, s>>>(n, M->r, M->q, (const T*)L0.x, st, partials);
    else if (rel)
      k_pcg_dot<K, T, 0, float><<<gv, kBlock, 0, s>>>(n, M->r, (const float*)M->q, (const T*)L0.x, st, partials);
    else
      k_pcg_dot<K, T, 0, double><<<gv, kBlock, 0, s>>>(n, M->r, M->q, (const T*)L0.x, st, partials);
    k_pcg_red_rz<K><<<1, kRedBlock, 0, s>>>(gv, partials, st);
  };
  auto pupd = [&]() {
    if (fl && rel)
      k_pcg_pupd<K, T, FB, float><<<gv, kBlock, 0, s>>>(n, (const T*)L0.x, (float*)M->p, M->r, st);
    else if (fl)
      k_pcg_pupd<K, T, FB, double><<<gv, kBlock, 0, s>>>(n, (const T*)L0.x, M->p, M->r, st);
    else if (rel)
      k_pcg_pupd<K, T, 0, float><<<gv, kBlock, 0, s>>>(n, (const T*)L0.x, (float*)M->p, M->r, st);
    else
      k_pcg_pupd<K, T, 0, double><<<gv, kBlock, 0, s>>>(n, (const T*)L0.x, M->p, M->r, st);
  };
  // start: r = b - A x0 (or b), state, z = M r, rho = r.z
  if (use_x0) spmv(x, M->q, true);
  k_pcg_init<K><<<gv, kBlock, 0, s>>>(n, b, x, M->q, M->r, use_x0, partials);
  if (rel)
    k_pcg_red_init<K, true><<<1, kRedBlock, 0, s>>>(gv, partials, st, tol, maxiter);
  else
    k_pcg_red_init<K, false><<<1, kRedBlock, 0, s>>>(gv, partials, st, tol, maxiter);
  if (use_x0) k_pcg_zero_b<K><<<gv, kBlock, 0, s>>>(n, x, st);
  precond(nullptr);
  CPM_HIP(hipGetLastError());
  // profiling: the first Chebyshev step of level 0 in the pre- and in the
  // post-smoothing of one iteration of every batch (launched directly, the
  // rest of the batch from the captured chunks) is timed with HIP events and
  // accumulated on the level-0 operator
  const bool prof = ln.prof && prof_on(ctx, L0.A) && M->degree >= 2 && K == 1;
  auto iterate = [&](int todo, hipEvent_t* ev) -> int {
    for (int k = 0; k < todo; ++k) {
      pupd();
      spmv(M->p, M->q, false, (k == 0 && ev && rel && S.stencil) ? ev + 4 : nullptr);
      if (rel)
        k_red<K, RED_ALPHA_PCG_REL><<<1, kRedBlock, 0, s>>>(gs, partials, st, 0.0);
      else
        k_red<K, RED_ALPHA><<<1, kRedBlock, 0, s>>>(gs, partials, st, 0.0);
      if (rel)
        k_pcg_xr<K, float><<<gv, kBlock, 0, s>>>(n, x, M->r, (const float*)M->p, (const float*)M->q, st, partials);
      else
        k_pcg_xr<K, double><<<gv, kBlock, 0, s>>>(n, x, M->r, M->p, M->q, st, partials);
      if (rel) {
        k_red<K, RED_RR_PCG_REL><<<1, kRedBlock, 0, s>>>(gv, partials, st, delta);
        if (S.stencil)
          k_pcg_true<BS, K, true><<<gs, kBlock, 0, s>>>(S.nbrows, S.sptr, S.col, sell_v64(S), S.dg64,
                                                        CPM_GRID_ARG64(S), x, b, M->r, st, partials);
        else
          k_pcg_true<BS, K, false><<<gs, kBlock, 0, s>>>(S.nbrows, S.sptr, S.col, sell_v64(S), S.dg64,
                                                         CPM_GRID_ARG64(S), x, b, M->r, st, partials);
        k_red<K, RED_RR_PCG_TRUE><<<1, kRedBlock, 0, s>>>(gs, partials, st, 0.0);
      } else {
        k_red<K, RED_RR_PCG><<<1, kRedBlock, 0, s>>>(gv, partials, st, 0.0);
      }
      precond((k == 0) ? ev : nullptr);
    }
    CPM_HIP(hipGetLastError());
    return CPM_OK;
  };
  // the captured chunks: `ch` iterations (2, and 1 for the last iterations
  // of a solve) with this solve's vectors and states (the kernels read every
  // scalar from the states, so a replay is the same work as launching them)
  hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
  auto capture = [&](int ch) -> int {
    uint64_t key = 1469598103934665603ull;
    uint64_t dbits;
    std::memcpy(&dbits, &delta, sizeof dbits);
    for (uint64_t v : {(uint64_t)(uintptr_t)x, (uint64_t)(uintptr_t)b, (uint64_t)(uintptr_t)M->r,
                       (uint64_t)(uintptr_t)L0.x, (uint64_t)(uintptr_t)st, (uint64_t)(uintptr_t)partials,
                       (uint64_t)K, (uint64_t)sizeof(T), (uint64_t)ch, dbits})
      key = (key ^ v) * 1099511628211ull;
    for (auto& c : M->chunks)
      if (c.key == key) gexec[ch] = c.exec;
    if (gexec[ch]) return CPM_OK;
    if (!M->side) CPM_HIP(hipStreamCreateWithFlags(&M->side, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    CPM_HIP(hipStreamBeginCapture(M->side, hipStreamCaptureModeThreadLocal));
    s = M->side;
    M->ls = M->side;
    const int rc = iterate(ch, nullptr);
    s = ln.stream;
    M->ls = ln.stream;
    hipError_t e = hipStreamEndCapture(M->side, &g);
    if (rc) return rc;
    if (e != hipSuccess) return fail(CPM_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&gexec[ch], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(CPM_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    // a hierarchy serves up to three solves (the Gauss-Seidel columns, g1*,
    // g2*), each with chunks of 2 and 1 iterations
    if (M->chunks.size() >= 8) {
      (void)hipGraphExecDestroy(M->chunks.front().exec);
      M->chunks.erase(M->chunks.begin());
    }
    M->chunks.push_back({key, gexec[ch], ch});
    return CPM_OK;
  };
  if (M->graphs)
    if (int rc = capture(2)) return rc;
  // cnt iterations from the chunks (launched directly when there are none)
  auto chunked = [&](int cnt) -> int {
    if (cnt <= 0) return CPM_OK;
    if (!gexec[2]) return iterate(cnt, nullptr);
    int k = 0;
    for (; k + 2 <= cnt; k += 2) CPM_HIP(hipGraphLaunch(gexec[2], s));
    if (k < cnt) {
      if (!gexec[1])
        if (int rc = capture(1)) return rc;
      CPM_HIP(hipGraphLaunch(gexec[1], s));
    }
    return CPM_OK;
  };
  // todo iterations; with tk >= 0 the iteration tk of them is launched
  // directly and timed with events
  auto issue = [&](int todo, int tk) -> int {
    if (tk < 0) return chunked(todo);
    if (int rc = chunked(tk)) return rc;
    if (int rc = iterate(1, &ctx->ev[0])) return rc;
    return chunked(todo - tk - 1);
  };
  CgState h[K];
  auto read = [&]() -> int {
    CPM_HIP(hipMemcpyAsync(ln.pinned, st, sizeof(CgState) * K, hipMemcpyDeviceToHost, s));
    CPM_HIP(hipStreamSynchronize(s));
    std::memcpy(h, ln.pinned, sizeof(CgState) * K);
    return CPM_OK;
  };
  auto all = [&]() {
    for (int j = 0; j < K; ++j)
      if (!h[j].done) return false;
    return true;
  };
  auto it_max = [&]() {
    int m = 0;
    for (int j = 0; j < K; ++j) m = std::max(m, (int)h[j].it);
    return m;
  };
  if (int rc = read()) return rc;
  // iterations per host check: the previous solve's count less one, then
  // one at a time up to two past it (an iteration launched after the solve
  // has converged costs a V-cycle's worth of launches, a check one
  // round trip), then a quarter of the count so far
  int batch = hint > 2 ? hint - 1 : 4;
  int issued = 0;
  cpm_mat* A0 = L0.A;
  while (!all() && issued <= maxiter) {
    const int todo = std::max(1, std::min(std::min(batch, 1024), maxiter - issued + 1));
    const int it_before = h[0].it;
    // the timed iteration: the first of the batch, or the last of the
    // first batch when another solve runs beside this one (its kernels
    // then share the GPU with the other solve's early iterations)
    const int tk = !prof ? -1 : (ln.shared && issued == 0) ? todo - 1 : 0;
    if (int rc = issue(todo, tk)) return rc;
    issued += todo;
    if (int rc = read()) return rc;
    if (tk >= 0) {
      // the timed iteration did V-cycle work unless the solve converged
      // before or in it
      const int base = it_before + tk;
      const bool worked = h[0].it > base + 1 || (h[0].it == base + 1 && (!h[0].done || h[0].fail));
      if (worked)
        for (int h2 = 0; h2 < 2; ++h2) {
          float ms = 0.f;
          CPM_HIP(hipEventElapsedTime(&ms, ctx->ev[2 * h2], ctx->ev[2 * h2 + 1]));
          A0->prof_ms += ms;
          A0->prof_launches += 1;
        }
      // the outer product of that iteration ran whenever the iteration did
      if (rel && S.stencil && h[0].it > base) {
        float ms = 0.f;
        CPM_HIP(hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]));
        A0->prof2_ms += ms;
        A0->prof2_launches += 1;
      }
    }
    const int itm = it_max();
    batch = (hint > 2 && itm < hint + 2) ? 1 : std::max(2, std::min(16, itm / 4 + 2));
  }
  *failed = 0;
  for (int j = 0; j < K; ++j) {
    iters[j] = h[j].it;
    resid[j] = h[j].resid;
    if (h[j].fail || !h[j].done) *failed = 1;
  }
  return CPM_OK;
}

const cpm_mat* mg_operator(const cpm_mg* M) { return M && !M->L.empty() ? M->L[0].A : nullptr; }

PcgLane lane1(cpm_ctx* ctx) {
  PcgLane ln;
  ln.stream = ctx->stream;
  ln.st = ctx->cg;
  ln.partials = ctx->partials;
  ln.pinned = ctx->pinned;
  ln.prof = ctx->prof;
  return ln;
}

int lane2(cpm_ctx* ctx, PcgLane* out) {
  if (int rc = ensure_lane2(ctx)) return rc;
  out->stream = ctx->stream2;
  out->st = ctx->cg + 8;
  out->partials = ctx->partials2;
  out->pinned = ctx->pinned2;
  out->prof = 0;
  return CPM_OK;
}

int mg_pcg_lane(cpm_mg* M, const PcgLane& ln, int K, const double* b, double* x, int use_x0, double tol,
                int maxiter, double delta, int hint, int32_t* iters, double* resid, int* failed) {
  if (K != 1 && !(K == 3 && M->bs == 1))
    return fail(CPM_EINVAL, "multigrid CG: 1 right-hand side, or 3 for a scalar operator");
  if (!(delta >= 0.0 && delta < 1.0)) return fail(CPM_EINVAL, "multigrid CG: delta must be in [0, 1)");
  if (int rc = mg_reserve(M, K, ln.stream)) return rc;
  const double d = delta;
  if (M->bs == 3)
    return M->prec == 32 ? pcg_run<3, 1, float>(M, ln, b, x, use_x0, tol, maxiter, d, hint, iters, resid, failed)
                         : pcg_run<3, 1, double>(M, ln, b, x, use_x0, tol, maxiter, d, hint, iters, resid, failed);
  if (K == 3)
    return M->prec == 32 ? pcg_run<1, 3, float>(M, ln, b, x, use_x0, tol, maxiter, d, hint, iters, resid, failed)
                         : pcg_run<1, 3, double>(M, ln, b, x, use_x0, tol, maxiter, d, hint, iters, resid, failed);
  return M->prec == 32 ? pcg_run<1, 1, float>(M, ln, b, x, use_x0, tol, maxiter, d, hint, iters, resid, failed)
                       : pcg_run<1, 1, double>(M, ln, b, x, use_x0, tol, maxiter, d, hint, iters, resid, failed);
}

int mg_pcg(cpm_mg* M, int K, const double* b, double* x, int use_x0, double tol, int maxiter, double delta,
           int hint, int32_t* iters, double* resid, int* failed) {
  return mg_pcg_lane(M, lane1(M->ctx), K, b, x, use_x0, tol, maxiter, delta, hint, iters, resid, failed);
}

}  // namespace cpm

using namespace cpm;

extern "C" int cpm_mg_destroy(cpm_mg* M) {
  if (!M) return CPM_OK;
  if (M->ctx) {
    (void)hipSetDevice(M->ctx->device);
    (void)hipStreamSynchronize(M->ctx->stream);
  }
  for (auto& c : M->chunks) (void)hipGraphExecDestroy(c.exec);
  if (M->side) (void)hipStreamDestroy(M->side);
  for (void* p : M->allocs) (void)hipFree(p);
  delete M;
  return CPM_OK;
}

// Chebyshev coefficients of every level for its degree (level_degree) over
// [lmax / ratio, 1.1 lmax] (the smoothed coarsest level: over [lmax / cratio,
// 1.1 lmax])
static void chebyshev_coefficients(cpm_mg* M) {
  const size_t nl = M->L.size();
  M->c0.assign(nl, 0.0);
  M->c1.assign(nl * 8, 0.0);
  M->c2.assign(nl * 8, 0.0);
  for (size_t l = 0; l < nl; ++l) {
    const bool last = l + 1 == nl;
    if (last && !M->csmooth) break;
    const double hi = 1.1 * M->L[l].lmax, lo = hi / (last ? M->cratio : M->ratio);
    const double th = 0.5 * (hi + lo), de = 0.5 * (hi - lo), sg = th / de;
    M->c0[l] = 1.0 / th;
    double rho = 1.0 / sg;
    for (int k = 1; k < M->level_degree(l); ++k) {
      const double rn = 1.0 / (2.0 * sg - rho);
      M->c1[l * 8 + k] = rn * rho;
      M->c2[l * 8 + k] = 2.0 * rn / de;
      rho = rn;
    }
  }
}

// cdeg > 0: the coarsest level is solved by cdeg Chebyshev steps over
// [lmax / cratio, 1.1 lmax] instead of the dense inverse coarse_inv
static int mg_create(cpm_ctx* ctx, int nlevels, cpm_mat* const* levels, const double* coarse_inv, int precision,
                     int degree, double ratio, int cdeg, double cratio, cpm_mg** out) {
  if (!ctx || !levels || (!coarse_inv && !cdeg) || !out) return fail(CPM_EINVAL, "cpm_mg_create: null argument");
  if (cdeg && (cdeg < 1 || cdeg > 8 || !(cratio > 1.0)))
    return fail(CPM_EINVAL, "cpm_mg_create: coarsest Chebyshev degree must be 1..8 and its ratio > 1");
  if (nlevels < 2 || nlevels > 16) return fail(CPM_EINVAL, "cpm_mg_create: 2..16 levels");
  if (precision != 16 && precision != 32 && precision != 64)
    return fail(CPM_EINVAL, "cpm_mg_create: precision must be 16, 32 or 64");
  const bool half = precision == 16;  // FP32 vectors, FP16 off-diagonal values
  if (half) precision = 32;
  if (degree < 1 || degree > 8) return fail(CPM_EINVAL, "cpm_mg_create: Chebyshev degree must be 1..8");
  if (!(ratio > 1.0)) return fail(CPM_EINVAL, "cpm_mg_create: eigenvalue ratio must be > 1");
  const int bs = levels[0] ? levels[0]->sell.bs : 0;
  for (int l = 0; l < nlevels; ++l) {
    const cpm_mat* A = levels[l];
    if (!A || !A->sell.sptr || A->sell.bs != bs || (bs != 1 && bs != 3))
      return fail(CPM_EINVAL, "cpm_mg_create: every level needs a SELL form of one block size (1 or 3)");
    if (!A->sell.stencil) return fail(CPM_EINVAL, "cpm_mg_create: levels must be structured periodic grids");
    if (l > 0) {
      const Sell& F = levels[l - 1]->sell;
      const Sell& C = A->sell;
      for (int a = 0; a < 3; ++a)
        if (C.dims[a] != (F.dims[a] + 1) / 2)
          return fail(CPM_EINVAL, "cpm_mg_create: each level must halve the grid (rounding up) in every dimension");
    }
  }
  CPM_HIP(hipSetDevice(ctx->device));
  auto* M = new cpm_mg();
  M->ctx = ctx;
  M->bs = bs;
  M->prec = precision;
  M->half = half ? 1 : 0;
  M->degree = degree;
  M->ratio = ratio;
  {
    // anchored operator (row 0 an identity row, as A_demag and A_el): with
    // anchored coarse levels node 0 stays out of the transfers; with
    // singular periodic coarse levels the hierarchy is floating (level 0
    // smooths with the periodic stiffness recovered from the fine operator,
    // node 0 takes part in the transfers).  Otherwise (A1) node 0 takes part.
    auto anchored = [&](const Sell& F, bool* a) -> hipError_t {
      const int64_t cnt = (int64_t)F.nst * bs * bs * 64;
      std::vector<double> v0(std::max<int64_t>(cnt, 1), 0.0);
      if (cnt > 0) {
        const hipError_t e = hipMemcpy(v0.data(), F.v64, sizeof(double) * cnt, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return e;
      }
      bool off = false;
      for (int64_t q = 0; q < cnt; q += 64) off |= v0[q] != 0.0;  // row 0 = lane 0 of slice 0
      *a = !off;
      return hipSuccess;
    };
    bool a[16];
    for (int l = 0; l < nlevels; ++l) {
      const hipError_t e = anchored(levels[l]->sell, &a[l]);
      if (e != hipSuccess) {
        delete M;
        return fail(CPM_EHIP, hipGetErrorString(e));
      }
    }
    for (int l = 2; l < nlevels; ++l)
      if (a[l] != a[1]) {
        delete M;
        return fail(CPM_EINVAL, "cpm_mg_create: coarse levels must all be anchored or all periodic");
      }
    if (!a[0] && a[1]) {
      delete M;
      return fail(CPM_EINVAL, "cpm_mg_create: anchored coarse levels under a periodic fine operator");
    }
    M->floating = (a[0] && !a[1]) ? 1 : 0;
    M->anchor = (a[0] && a[1]) ? 1 : 0;
    if (cdeg && a[0]) {
      delete M;
      return fail(CPM_EINVAL, "cpm_mg_create: an anchored operator needs the dense coarsest solve");
    }
    M->csmooth = cdeg;
    M->cratio = cratio;
    const Sell& F = levels[0]->sell;
    if (M->floating && (F.dims[0] < 3 || F.dims[1] < 3 || F.dims[2] < 3 || !F.kuhn)) {
      delete M;
      return fail(CPM_EINVAL, "cpm_mg_create: a floating hierarchy needs a Kuhn stencil with >= 3 fine nodes per "
                              "dimension");
    }
  }
  const size_t ts = precision == 32 ? 4 : 8;
  int rc = CPM_OK;
  auto A = [&](void** p, size_t bytes) {
    if (rc == CPM_OK) rc = mg_alloc(M, p, bytes);
  };
  hipStream_t s = ctx->stream;
  double *px = nullptr, *py = nullptr, *dinv64 = nullptr;
  for (int l = 0; l < nlevels && rc == CPM_OK; ++l) {
    MgLevel Lv;
    Lv.A = levels[l];
    const Sell& S = Lv.A->sell;
    Lv.nb = S.nbrows;
    for (int a = 0; a < 3; ++a) Lv.dims[a] = S.dims[a];
    Lv.n = (int64_t)Lv.nb * bs;
    for (void** p : {&Lv.x, &Lv.r, &Lv.b, &Lv.d0, &Lv.d1}) A(p, ts * Lv.n);
    A(&Lv.dinv, ts * Lv.nb * bs * bs);
    if (rc != CPM_OK) break;
    const int64_t nbb = (int64_t)Lv.nb * bs * bs;
    const int64_t nval = S.nslot * bs * bs * 64;
    hipError_t e = hipSuccess;
    if (l == 0 && M->floating) {
      // the periodic stiffness: FP64 copies of the fine operator's values
      // and diagonal blocks, node 0's rows and columns recovered in place
      A((void**)&Lv.kv64, sizeof(double) * std::max<int64_t>(1, nval));
      A((void**)&Lv.kdg64, sizeof(double) * nbb);
      if (rc != CPM_OK) break;
      e = hipMemcpyAsync(Lv.kv64, S.v64, sizeof(double) * nval, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess) e = hipMemcpyAsync(Lv.kdg64, S.dg64, sizeof(double) * nbb, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess) {
        if (bs == 3) {
          k_mg_float_rows<3><<<grid_for(Lv.nb), kBlock, 0, s>>>(S.nbrows, S.nst, grid_arg(S), Lv.kv64, Lv.kdg64);
          k_mg_float_row0<3><<<1, 64, 0, s>>>(S.nst, grid_arg(S), Lv.kv64, Lv.kdg64);
        } else {
          k_mg_float_rows<1><<<grid_for(Lv.nb), kBlock, 0, s>>>(S.nbrows, S.nst, grid_arg(S), Lv.kv64, Lv.kdg64);
          k_mg_float_row0<1><<<1, 64, 0, s>>>(S.nst, grid_arg(S), Lv.kv64, Lv.kdg64);
        }
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) {
        rc = fail(CPM_EHIP, hipGetErrorString(e));
        break;
      }
    }
    const double* v64 = Lv.kv64 ? Lv.kv64 : S.v64;
    const double* dg64 = Lv.kdg64 ? Lv.kdg64 : S.dg64;
    // inverse diagonal blocks: of the FP64 form, and of the (compensated) FP32 form
    std::vector<double> dg(nbb), inv64, inv_t;
    e = hipMemcpy(dg.data(), dg64, sizeof(double) * nbb, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      rc = fail(CPM_EHIP, hipGetErrorString(e));
      break;
    }
    if (!invert_blocks(bs, Lv.nb, dg, inv64)) {
      rc = fail(CPM_EINVAL, "cpm_mg_create: singular diagonal block");
      break;
    }
    if (half) {
      // scale: a power of two that brings the largest diagonal entry (a
      // bound of every entry of the SPD operator) to <= 2^14
      double dmax = 0.0;
      for (double v : dg) dmax = std::max(dmax, std::fabs(v));
      Lv.vscale = std::ldexp(1.0, std::max(-60, (int)std::ceil(std::log2(std::max(dmax, 1e-300) / 16384.0))));
      // half storage when the Kuhn stencil is symmetric slot by slot (to
      // 1e-12 of its largest entry) and every offset reaches its own node
      if (S.kuhn && S.dims[0] >= 3 && S.dims[1] >= 3 && S.dims[2] >= 3) {
        void* am = nullptr;
        A(&am, 2 * sizeof(unsigned long long));
        if (rc != CPM_OK) break;
        unsigned long long hm[2] = {0, 0};
        e = hipMemcpy(am, hm, sizeof(hm), hipMemcpyHostToDevice);
        if (e == hipSuccess) {
          if (bs == 3)
            k_mg_kuhn_asym<3><<<grid_for(Lv.nb), kBlock, 0, s>>>(S.nbrows, grid_arg(S), v64,
                                                                 (unsigned long long*)am);
          else
            k_mg_kuhn_asym<1><<<grid_for(Lv.nb), kBlock, 0, s>>>(S.nbrows, grid_arg(S), v64,
                                                                 (unsigned long long*)am);
          e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(hm, am, sizeof(hm), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(am);
        M->allocs.erase(std::find(M->allocs.begin(), M->allocs.end(), am));
        if (e != hipSuccess) {
          rc = fail(CPM_EHIP, hipGetErrorString(e));
          break;
        }
        double asym, amx;
        std::memcpy(&asym, &hm[0], sizeof(double));
        std::memcpy(&amx, &hm[1], sizeof(double));
        Lv.khalf = (asym <= 1e-12 * amx) ? 1 : 0;
      }
      A(&Lv.v16, sizeof(_Float16) * std::max<int64_t>(1, (Lv.khalf ? S.nslot / 2 : S.nslot) * bs * bs * 64));
      A((void**)&Lv.dgh, sizeof(double) * nbb);
      if (rc != CPM_OK) break;
      const int64_t pad = (int64_t)S.nslices * 64;
      if (Lv.khalf) {
        if (bs == 3)
          k_mg_make_half<3><<<grid_for(pad), kBlock, 0, s>>>(S.nbrows, pad, grid_arg(S), v64, dg64, Lv.vscale,
                                                             (_Float16*)Lv.v16, Lv.dgh);
        else
          k_mg_make_half<1><<<grid_for(pad), kBlock, 0, s>>>(S.nbrows, pad, grid_arg(S), v64, dg64, Lv.vscale,
                                                             (_Float16*)Lv.v16, Lv.dgh);
      } else if (bs == 3)
        k_mg_make_low<3, _Float16><<<grid_for(pad), kBlock, 0, s>>>(S.nbrows, pad, S.sptr, S.nst, S.stencil, v64,
                                                                    dg64, Lv.vscale, (_Float16*)Lv.v16, Lv.dgh);
      else
        k_mg_make_low<1, _Float16><<<grid_for(pad), kBlock, 0, s>>>(S.nbrows, pad, S.sptr, S.nst, S.stencil, v64,
                                                                    dg64, Lv.vscale, (_Float16*)Lv.v16, Lv.dgh);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) {
        rc = fail(CPM_EHIP, hipGetErrorString(e));
        break;
      }
    } else if (precision == 32 && Lv.kv64) {
      // the compensated FP32 form of the recovered stiffness
      A((void**)&Lv.kv32, sizeof(float) * std::max<int64_t>(1, nval));
      A((void**)&Lv.kdgc, sizeof(double) * nbb);
      if (rc != CPM_OK) break;
      const int64_t pad = (int64_t)S.nslices * 64;
      if (bs == 3)
        k_mg_make_low<3, float><<<grid_for(pad), kBlock, 0, s>>>(S.nbrows, pad, S.sptr, S.nst, S.stencil, v64, dg64,
                                                                 1.0, Lv.kv32, Lv.kdgc);
      else
        k_mg_make_low<1, float><<<grid_for(pad), kBlock, 0, s>>>(S.nbrows, pad, S.sptr, S.nst, S.stencil, v64, dg64,
                                                                 1.0, Lv.kv32, Lv.kdgc);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) {
        rc = fail(CPM_EHIP, hipGetErrorString(e));
        break;
      }
    }
    if (precision == 32) {
      e = hipMemcpy(dg.data(), half ? Lv.dgh : ldiag<float>(Lv), sizeof(double) * nbb, hipMemcpyDeviceToHost);
      // FP16 forms: the sweeps invert the FP32-rounded diagonal blocks per
      // node, so the stored D^-1 of the first step is the inverse of the
      // same rounded blocks (one D for every step of the polynomial)
      if (half)
        for (double& v : dg) v = (double)(float)v;
      if (e != hipSuccess || !invert_blocks(bs, Lv.nb, dg, inv_t)) {
        rc = fail(CPM_EINVAL, "cpm_mg_create: singular diagonal block (FP32 form)");
        break;
      }
      if (half) {
        // the FP32 diagonal blocks of the FP16 sweeps
        A((void**)&Lv.dg32, sizeof(float) * nbb);
        if (rc != CPM_OK) break;
        std::vector<float> d32(dg.begin(), dg.end());
        e = hipMemcpy(Lv.dg32, d32.data(), sizeof(float) * nbb, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
          rc = fail(CPM_EHIP, hipGetErrorString(e));
          break;
        }
        // the FP64 compensated blocks served the FP32 copy (and, rounded, the inverse)
        (void)hipFree(Lv.dgh);
        M->allocs.erase(std::find(M->allocs.begin(), M->allocs.end(), (void*)Lv.dgh));
        Lv.dgh = nullptr;
      }
      std::vector<float> f(inv_t.begin(), inv_t.end());
      e = hipMemcpy(Lv.dinv, f.data(), sizeof(float) * nbb, hipMemcpyHostToDevice);
    } else {
      e = hipMemcpy(Lv.dinv, inv64.data(), sizeof(double) * nbb, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
      rc = fail(CPM_EHIP, hipGetErrorString(e));
      break;
    }
    if (l + 1 < nlevels || cdeg) {
      // lambda_max(D^-1 A) by power iteration on the FP64 form
      void *vx = nullptr, *vy = nullptr, *vd = nullptr;
      A(&vx, 8 * Lv.n);
      A(&vy, 8 * Lv.n);
      const int gp = sell_grid(S);
      A(&vd, 8 * (nbb + 1 + gp));  // D^-1 (FP64), the norm of the iterate, the partials
      if (rc != CPM_OK) break;
      px = (double*)vx, py = (double*)vy, dinv64 = (double*)vd;
      e = hipMemcpy(dinv64, inv64.data(), sizeof(double) * nbb, hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        rc = fail(CPM_EHIP, hipGetErrorString(e));
        break;
      }
      double* lam_dev = dinv64 + nbb;
      rc = bs == 3 ? power_lmax<3>(M, Lv, px, py, dinv64, lam_dev, lam_dev + 1, &Lv.lmax)
                   : power_lmax<1>(M, Lv, px, py, dinv64, lam_dev, lam_dev + 1, &Lv.lmax);
      // the scratch of the estimate is not needed afterwards
      for (void* p : {vx, vy, vd}) {
        (void)hipFree(p);
        M->allocs.erase(std::find(M->allocs.begin(), M->allocs.end(), p));
      }
    }
    if (Lv.kv64 && precision != 64) {
      // the FP64 recovered stiffness served the conversions and the estimate
      for (void* p : {(void*)Lv.kv64, (void*)Lv.kdg64}) {
        (void)hipFree(p);
        M->allocs.erase(std::find(M->allocs.begin(), M->allocs.end(), p));
      }
      Lv.kv64 = Lv.kdg64 = nullptr;
    }
    M->L.push_back(Lv);
  }
  if (rc == CPM_OK) {
    chebyshev_coefficients(M);
    // coarsest dense inverse
    const int64_t nc = cdeg ? 0 : M->L.back().n;
    M->ncoarse = nc;
    if (!cdeg) A(&M->cinv, ts * nc * nc);
    if (rc == CPM_OK && !cdeg) {
      hipError_t e;
      if (precision == 32) {
        std::vector<float> f(coarse_inv, coarse_inv + nc * nc);
        e = hipMemcpy(M->cinv, f.data(), sizeof(float) * nc * nc, hipMemcpyHostToDevice);
      } else {
        e = hipMemcpy(M->cinv, coarse_inv, sizeof(double) * nc * nc, hipMemcpyHostToDevice);
      }
      if (e != hipSuccess) rc = fail(CPM_EHIP, hipGetErrorString(e));
    }
    const int64_t n0 = M->L[0].n;
    for (void** p : {(void**)&M->r, (void**)&M->p, (void**)&M->q}) A(p, 8 * n0);
  }
  if (rc != CPM_OK) {
    cpm_mg_destroy(M);
    return rc;
  }
  *out = M;
  return CPM_OK;
}

extern "C" int cpm_mg_create(cpm_ctx* ctx, int nlevels, cpm_mat* const* levels, const double* coarse_inv,
                             int precision, int degree, double ratio, cpm_mg** out) {
  if (!coarse_inv) return fail(CPM_EINVAL, "cpm_mg_create: null argument");
  return mg_create(ctx, nlevels, levels, coarse_inv, precision, degree, ratio, 0, 0.0, out);
}

extern "C" int cpm_mg_create_smoothed(cpm_ctx* ctx, int nlevels, cpm_mat* const* levels, int precision, int degree,
                                      double ratio, int cdegree, double cratio, cpm_mg** out) {
  if (cdegree < 1) return fail(CPM_EINVAL, "cpm_mg_create_smoothed: coarsest Chebyshev degree must be 1..8");
  return mg_create(ctx, nlevels, levels, nullptr, precision, degree, ratio, cdegree, cratio, out);
}

extern "C" int cpm_mg_set_coarse_degree(cpm_mg* M, int degree) {
  if (!M) return fail(CPM_EINVAL, "cpm_mg_set_coarse_degree: null hierarchy");
  if (degree < 0 || degree > 8 || (degree == 1) != (M->degree == 1))
    return fail(CPM_EINVAL, "cpm_mg_set_coarse_degree: degree must be 0 (the finest level's) or 2..8 (1 only with "
                            "a finest degree of 1)");
  M->degree1 = degree;
  chebyshev_coefficients(M);
  // the captured chunks hold the old sequence of sweeps
  for (auto& c : M->chunks) (void)hipGraphExecDestroy(c.exec);
  M->chunks.clear();
  return CPM_OK;
}

extern "C" int cpm_mg_info(const cpm_mg* M, int* nlevels, double* lmax, int max_levels) {
  if (!M || !nlevels) return fail(CPM_EINVAL, "cpm_mg_info: null argument");
  *nlevels = (int)M->L.size();
  if (lmax)
    for (int l = 0; l < std::min<int>(max_levels, (int)M->L.size()); ++l) lmax[l] = M->L[l].lmax;
  return CPM_OK;
}

extern "C" int cpm_mg_forms(const cpm_mg* M, int* half_stencil, int max_levels) {
  if (!M || !half_stencil) return fail(CPM_EINVAL, "cpm_mg_forms: null argument");
  for (int l = 0; l < std::min<int>(max_levels, (int)M->L.size()); ++l)
    half_stencil[l] = (M->half && M->L[l].khalf) ? 1 : 0;
  return CPM_OK;
}

extern "C" int cpm_mg_apply(cpm_mg* M, const double* r, double* z) {
  if (!M || !r || !z) return fail(CPM_EINVAL, "cpm_mg_apply: null argument");
  if (M->kcap == 0) return fail(CPM_ENOMEM, "cpm_mg_apply: hierarchy unusable after a failed allocation");
  cpm_ctx* ctx = M->ctx;
  CPM_HIP(hipSetDevice(ctx->device));
  CgState* st = ctx->cg + 15;  // a live state so that the gated kernels run
  CgState h{};
  CPM_HIP(hipMemcpy(st, &h, sizeof(CgState), hipMemcpyHostToDevice));
  const MgLevel& L0 = M->L[0];
  const int gv = grid_for(L0.n);
  if (M->floating) {
    // the preconditioner of the floating hierarchy: F^T V(F r) + e_0 r_0
    if (M->bs == 3)
      k_pcg_rsum<3, 1><<<gv, kBlock, 0, ctx->stream>>>(L0.n, r, st, ctx->partials);
    else
      k_pcg_rsum<1, 1><<<gv, kBlock, 0, ctx->stream>>>(L0.n, r, st, ctx->partials);
  }
  M->lpart = ctx->partials;  // the partials k_pcg_rsum just wrote (floating hierarchies)
  vcycle_any(M, r, st);
  M->lpart = nullptr;
  CPM_HIP(hipGetLastError());
  if (M->floating) {
    if (M->bs == 3) {
      if (M->prec == 32)
        k_mg_float_out<3, float><<<gv, kBlock, 0, ctx->stream>>>(L0.n, (const float*)L0.x, r, z);
      else
        k_mg_float_out<3, double><<<gv, kBlock, 0, ctx->stream>>>(L0.n, (const double*)L0.x, r, z);
    } else {
      if (M->prec == 32)
        k_mg_float_out<1, float><<<gv, kBlock, 0, ctx->stream>>>(L0.n, (const float*)L0.x, r, z);
      else
        k_mg_float_out<1, double><<<gv, kBlock, 0, ctx->stream>>>(L0.n, (const double*)L0.x, r, z);
    }
    CPM_HIP(hipGetLastError());
  } else if (M->prec == 64) {
    CPM_HIP(hipMemcpyAsync(z, L0.x, 8 * L0.n, hipMemcpyDeviceToDevice, ctx->stream));
  } else {
    std::vector<float> f(L0.n);
    std::vector<double> d(L0.n);
    CPM_HIP(hipMemcpyAsync(f.data(), L0.x, 4 * L0.n, hipMemcpyDeviceToHost, ctx->stream));
    CPM_HIP(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < L0.n; ++i) d[i] = f[i];
    CPM_HIP(hipMemcpyAsync(z, d.data(), 8 * L0.n, hipMemcpyHostToDevice, ctx->stream));
  }
  CPM_HIP(hipStreamSynchronize(ctx->stream));
  return CPM_OK;
}

extern "C" int cpm_cg_mg(cpm_mg* M, int nrhs, const double* b, double* x, int use_x0, double tol, int maxiter,
                         double delta, int32_t* iters, double* resid) {
  if (!M || !b || !x || !iters || !resid) return fail(CPM_EINVAL, "cpm_cg_mg: null argument");
  CPM_HIP(hipSetDevice(M->ctx->device));
  int failed = 0;
  if (int rc = mg_pcg(M, nrhs, b, x, use_x0, tol, maxiter, delta, 0, iters, resid, &failed)) return rc;
  int worst = 0;
  for (int j = 0; j < nrhs; ++j) worst = std::max(worst, (int)iters[j]);
  return failed ? std::max(1, worst) : CPM_OK;
}

extern "C" int cpm_mg_set_graphs(cpm_mg* M, int on) {
  if (!M) return fail(CPM_EINVAL, "cpm_mg_set_graphs: null hierarchy");
  M->graphs = on ? 1 : 0;
  return CPM_OK;
}
