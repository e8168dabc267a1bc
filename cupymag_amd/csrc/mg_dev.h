// Device bodies of the multigrid V-cycle and of the preconditioned CG,
// written over a workgroup range (lb, lg) like sell_dev.h, so that one
// system per launch (mg.hip) and many systems per segmented launch
// (multi_mg.hip) share them.
#pragma once
#include "sell_dev.h"

namespace cpm {

// Vectors of a level hold K right-hand sides interleaved per node:
// entry (c*BS + e)*K + j for block row c, component e, column j (the layout
// of the K-column SELL solvers); NV = BS*K entries per node.

// out = D^-1 v at `node` (BS x BS block applied to each of the K columns)
template <int BS, int K, class T>
__device__ __forceinline__ void apply_dinv(const T* __restrict__ dinv, int64_t node, const double (&v)[BS * K],
                                           double (&out)[BS * K]) {
#pragma unroll
  for (int d = 0; d < BS; ++d)
#pragma unroll
    for (int j = 0; j < K; ++j) {
      double s = 0.0;
#pragma unroll
      for (int e = 0; e < BS; ++e) s += (double)dinv[node * BS * BS + d * BS + e] * v[e * K + j];
      out[d * K + j] = s;
    }
}

// D^-1 v for a node from its diagonal block (FP32 in the FP16 level forms,
// read back from the cache: the row's product has just loaded it), the
// inverse by cofactors in FP64 -- the host's rule for the stored D^-1.
template <int BS, int K, class DG>
__device__ __forceinline__ void solve_diag(const DG* __restrict__ dg, int64_t node, const double (&v)[BS * K],
                                           double (&out)[BS * K]) {
  const DG* a = dg + node * BS * BS;
  if constexpr (BS == 1) {
    const double id = 1.0 / (double)a[0];
#pragma unroll
    for (int j = 0; j < K; ++j) out[j] = v[j] * id;
  } else {
    static_assert(BS == 3, "1x1 or 3x3 blocks");
    double m[9], inv[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) m[t] = (double)a[t];
    const double c0 = m[4] * m[8] - m[5] * m[7], c1 = m[5] * m[6] - m[3] * m[8], c2 = m[3] * m[7] - m[4] * m[6];
    const double id = 1.0 / (m[0] * c0 + m[1] * c1 + m[2] * c2);
    inv[0] = c0 * id;
    inv[1] = (m[2] * m[7] - m[1] * m[8]) * id;
    inv[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    inv[3] = c1 * id;
    inv[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    inv[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    inv[6] = c2 * id;
    inv[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    inv[8] = (m[0] * m[4] - m[1] * m[3]) * id;
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int j = 0; j < K; ++j) {
        double s = 0.0;
#pragma unroll
        for (int e = 0; e < 3; ++e) s += inv[d * 3 + e] * v[e * K + j];
        out[d * K + j] = s;
      }
  }
}

// all K columns of a solve have converged
template <int K>
__device__ __forceinline__ bool all_done(const CgState* st) {
  bool d = true;
#pragma unroll
  for (int j = 0; j < K; ++j) d = d && st[j].done;
  return d;
}

// some column of a mixed-precision solve has switched its products to FP64
template <int K>
__device__ __forceinline__ bool any_p64(const CgState* st) {
  bool h = false;
#pragma unroll
  for (int j = 0; j < K; ++j) h = h || st[j].p64 != 0;
  return h;
}

// First Chebyshev step.  ZERO: x0 = 0, so r = b (FROM64: b is the FP64
// outer residual, also stored as this level's b), d = c0 D^-1 r, x = d.
// Otherwise r is the current residual: d = c0 D^-1 r, x += d.
// Floating hierarchy (fpart non-null, level 0 from b64): node 0 takes
// b64_0 - sum_i b64_i per component, so that the right-hand side of the
// periodic problem sums to zero (F b, F = I - e_0 1^T); the sums are the
// fn partials fpart[v * fstride + k] of pcg_rsum_rows, added up by the
// first workgroup of the range, which holds node 0.
template <int BS, int K, class T, bool ZERO, bool FROM64>
__device__ __forceinline__ void mg_first_rows(int64_t lb, int64_t lg, int32_t nb, const double* __restrict__ b64,
                                              T* __restrict__ b, T* __restrict__ r, T* __restrict__ d,
                                              T* __restrict__ x, const T* __restrict__ dinv, double c0,
                                              const double* __restrict__ fpart = nullptr, int64_t fstride = 0,
                                              int32_t fn = 0) {
  constexpr int NV = BS * K;
  double fsum[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) fsum[c] = 0.0;
  if (FROM64 && fpart && lb == 0) {
    __shared__ double lds[NV * (kBlock / 64)];
    for (int k = threadIdx.x; k < fn; k += kBlock)
#pragma unroll
      for (int c = 0; c < NV; ++c) fsum[c] += fpart[c * fstride + k];
    block_sum_t<NV, kBlock>(fsum, lds);  // totals in thread 0, the one that takes node 0
  }
  const int64_t stride = lg * kBlock;
  for (int64_t i = lb * kBlock + threadIdx.x; i < nb; i += stride) {
    double rv[NV], dv[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      if (ZERO) {
        double bv = FROM64 ? b64[i * NV + c] : (double)b[i * NV + c];
        if (FROM64 && fpart && i == 0) bv -= fsum[c];
        if (FROM64) b[i * NV + c] = (T)bv;
        r[i * NV + c] = (T)bv;
        rv[c] = (double)(T)bv;
      } else {
        rv[c] = (double)r[i * NV + c];
      }
    }
    apply_dinv<BS, K, T>(dinv, i, rv, dv);
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      const T dn = (T)(c0 * dv[c]);
      d[i * NV + c] = dn;
      x[i * NV + c] = ZERO ? dn : (T)((double)x[i * NV + c] + (double)dn);
    }
  }
}

// MODE 0: Chebyshev step r -= A d_in, d_out = c1 d_in + c2 D^-1 r, x += d_out.
// MODE 1: r -= A d_in only (residual after the last step).
// MODE 2: r = b - A x.
// MODE 3: r = b - A x and the first step's direction d_out = c2 D^-1 r
//         (c1 = 0), x left alone: the post-smoothing's residual and first
//         step in one pass (x can not change while neighbours read it).
// MODE 4: MODE 0 that also adds the first step's d_in to x, rounded as the
//         separate updates were: x = (x + d_in) + d_out.
// Off-diagonal values of type VT (T, or _Float16 scaled by 1/vscale; the
// FP16 forms keep FP32 diagonal blocks and take D^-1 r from them, so these
// sweeps do not read the stored D^-1).
template <int BS, int K, class T, int MODE, bool ST, class VT = T>
__device__ __forceinline__ void mg_sweep_rows(int64_t lb, int64_t lg, int32_t nb, const int64_t* __restrict__ sptr,
                                              const int32_t* __restrict__ col, const VT* __restrict__ val,
                                              const sell_diag_t<VT>* __restrict__ dg, const GridArg& ga,
                                              const T* __restrict__ din, T* __restrict__ dout, T* __restrict__ x,
                                              T* __restrict__ r, const T* __restrict__ b,
                                              const T* __restrict__ dinv, double c1, double c2,
                                              double vscale = 1.0) {
  constexpr int NV = BS * K;
  constexpr bool RES = (MODE == 2 || MODE == 3);  // residual of x (else r -= A d_in)
  const T* src = RES ? x : din;
  const int64_t stride = lg * kBlock;
  for (int64_t row = lb * kBlock + threadIdx.x; row < (int64_t)((nb + 63) & ~63); row += stride) {
    double sum[NV], xo[NV];
    sell_row<BS, K, T, false, ST, VT, 0, sizeof(T) == 4>(row, nb, sptr, col, val, dg, src, nullptr, ga, sum, xo,
                                                         vscale);
    if (row >= nb) continue;
    double rv[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      rv[c] = RES ? (double)b[row * NV + c] - sum[c] : (double)r[row * NV + c] - sum[c];
      r[row * NV + c] = (T)rv[c];
      rv[c] = (double)(T)rv[c];
    }
    if (MODE == 0 || MODE == 3 || MODE == 4) {
      double dv[NV];
      if constexpr (sizeof(VT) == 2)
        solve_diag<BS, K>(dg, row, rv, dv);
      else
        apply_dinv<BS, K, T>(dinv, row, rv, dv);
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        const T dn = (MODE == 3) ? (T)(c2 * dv[c]) : (T)(c1 * xo[c] + c2 * dv[c]);
        dout[row * NV + c] = dn;
        if (MODE == 0) {
          x[row * NV + c] = (T)((double)x[row * NV + c] + (double)dn);
        } else if (MODE == 4) {
          const T x1 = (T)((double)x[row * NV + c] + xo[c]);
          x[row * NV + c] = (T)((double)x1 + (double)dn);
        }
      }
    }
  }
}

// Row `row` (< nrows_pad, padding rows included) of the FP16 half form of a
// symmetric Kuhn stencil operator, read by sell_row with GridArg::kuhn = 2:
// plane s*7 + u = fp16(v64 slot s*14 + 7 + u / vscale), a lane's 3x3 entries
// 0..7 in one 16-byte word (8 halves at 8*l), entry 8 at 512 + l; for rows
// < nb, dgh = dg64 + the rounding of the fourteen blocks the sweeps read for
// the row (its forward blocks, and for its backward slots the transposed
// forward blocks of the neighbours), so that the row sums with dgh equal the
// FP64 ones.
template <int BS>
__host__ __device__ inline void make_half_row(int64_t row, int32_t nb, const GridArg& ga,
                                              const double* __restrict__ v64, const double* __restrict__ dg64,
                                              double vscale, _Float16* __restrict__ v16,
                                              double* __restrict__ dgh) {
  constexpr int BB = BS * BS;
  const int64_t s = row >> 6, l = row & 63;
  double comp[BB];
  for (int t = 0; t < BB; ++t) comp[t] = 0.0;
  for (int u = 0; u < 7; ++u) {
    const int64_t slot = s * kKuhnSlots + 7 + u, plane = s * 7 + u;
    for (int t = 0; t < BB; ++t) {
      const double a = v64[(slot * BB + t) * 64 + l];
      const _Float16 h = (_Float16)(a / vscale);
      v16[BB == 9 ? plane * 576 + (t < 8 ? 8 * l + t : 512 + l) : plane * 64 + l] = h;
      comp[t] += a - (double)h * vscale;
    }
  }
  if (row >= nb) return;
  for (int t = 0; t < 7; ++t) {
    const int64_t c = kuhn_neighbour(row, ga, t);
    const int64_t cs = (c >> 6) * kKuhnSlots + (kKuhnSlots - 1 - t), lc = c & 63;
    for (int d = 0; d < BS; ++d)
      for (int e = 0; e < BS; ++e) {
        const double a = v64[((s * kKuhnSlots + t) * BB + d * BS + e) * 64 + l];
        const _Float16 h = (_Float16)(v64[(cs * BB + e * BS + d) * 64 + lc] / vscale);
        comp[d * BS + e] += a - (double)h * vscale;
      }
  }
  for (int t = 0; t < BB; ++t) dgh[row * BB + t] = dg64[row * BB + t] + comp[t];
}

__host__ __device__ __forceinline__ int64_t wrap_idx(int32_t i, int32_t j, int32_t k, const int32_t* dims) {
  i += (i < 0) ? dims[0] : 0;
  i -= (i >= dims[0]) ? dims[0] : 0;
  j += (j < 0) ? dims[1] : 0;
  j -= (j >= dims[1]) ? dims[1] : 0;
  k += (k < 0) ? dims[2] : 0;
  k -= (k >= dims[2]) ? dims[2] : 0;
  return ((int64_t)i * dims[1] + j) * dims[2] + k;
}

struct Dims3 {
  int32_t d[3];
};

__host__ __device__ __forceinline__ bool nested_dims(const Dims3& fine, const Dims3& coarse) {
  return fine.d[0] == 2 * coarse.d[0] && fine.d[1] == 2 * coarse.d[1] && fine.d[2] == 2 * coarse.d[2];
}

// Axes by decreasing fraction (t[p0] >= t[p1] >= t[p2]; the comparison
// order fixes the ties, the same for the prolongation and the restriction).
__host__ __device__ __forceinline__ void kuhn_order(const double (&t)[3], int& p0, int& p1, int& p2) {
  int tmp;
  p0 = 0, p1 = 1, p2 = 2;
  if (t[p1] > t[p0]) tmp = p0, p0 = p1, p1 = tmp;
  if (t[p2] > t[p1]) tmp = p1, p1 = p2, p2 = tmp;
  if (t[p1] > t[p0]) tmp = p0, p0 = p1, p1 = tmp;
}

// P1 interpolation weights of fine node (i,j,k) on the coarse Kuhn mesh of
// the same periodic box (fine node i lies at coarse coordinate i*nc/nf per
// axis): cell c, fractions t sorted t_p0 >= t_p1 >= t_p2, vertices c,
// c + e_p0, c + e_p0 + e_p1, c + (1,1,1) with weights 1 - t_p0, t_p0 - t_p1,
// t_p1 - t_p2, t_p2 (barycentric coordinates in the Kuhn tetrahedron that
// holds the point).  For nf = 2 nc these are the nested weights (0.5 pairs).
__host__ __device__ __forceinline__ void kuhn_weights(int32_t i, int32_t j, int32_t k, const Dims3& fine,
                                             const Dims3& coarse, int64_t (&v)[4], double (&w)[4]) {
  const int32_t idx[3] = {i, j, k};
  int32_t cell[3];
  double t[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    // 32-bit: fine index x coarse size stays far below 2^31 for any grid
    // the operators fit
    const int32_t q = idx[a] * coarse.d[a];
    cell[a] = q / fine.d[a];
    t[a] = (double)(q - cell[a] * fine.d[a]) / (double)fine.d[a];
  }
  int p0, p1, p2;
  kuhn_order(t, p0, p1, p2);
  int32_t o[3] = {0, 0, 0};
  v[0] = wrap_idx(cell[0], cell[1], cell[2], coarse.d);
  w[0] = 1.0 - t[p0];
  o[p0] = 1;
  v[1] = wrap_idx(cell[0] + o[0], cell[1] + o[1], cell[2] + o[2], coarse.d);
  w[1] = t[p0] - t[p1];
  o[p1] = 1;
  v[2] = wrap_idx(cell[0] + o[0], cell[1] + o[1], cell[2] + o[2], coarse.d);
  w[2] = t[p1] - t[p2];
  v[3] = wrap_idx(cell[0] + 1, cell[1] + 1, cell[2] + 1, coarse.d);
  w[3] = t[p2];
}

// (R r_f) at coarse node I (R = P^T of the P1 interpolation, W values per
// node); anchor: node 0 (the operators' anchor) excluded on both grids.
// Nested grids (fine = 2 x coarse) use the fixed stencil; other size pairs
// gather over the fine nodes whose cell is the node's own or the one below
// it in each dimension.
template <int W, class T>
__host__ __device__ __forceinline__ void restrict_node(int64_t I, const Dims3& fine, const Dims3& coarse, bool nested,
                                              bool anchor, const T* __restrict__ rf, double (&acc)[W]) {
#pragma unroll
  for (int c = 0; c < W; ++c) acc[c] = 0.0;
  if (anchor && I == 0) return;
  const int32_t i = (int32_t)(I / ((int64_t)coarse.d[1] * coarse.d[2]));
  const int32_t j = (int32_t)((I / coarse.d[2]) % coarse.d[1]);
  const int32_t k = (int32_t)(I % coarse.d[2]);
  if (nested) {
    const int64_t f0 = wrap_idx(2 * i, 2 * j, 2 * k, fine.d);
#pragma unroll
    for (int c = 0; c < W; ++c) acc[c] = (anchor && f0 == 0) ? 0.0 : (double)rf[f0 * W + c];
#pragma unroll
    for (int dd = 1; dd < 8; ++dd) {
      const int32_t a = dd & 1, bb = (dd >> 1) & 1, cc = (dd >> 2) & 1;
      const int64_t fp = wrap_idx(2 * i + a, 2 * j + bb, 2 * k + cc, fine.d);
      const int64_t fm = wrap_idx(2 * i - a, 2 * j - bb, 2 * k - cc, fine.d);
#pragma unroll
      for (int c = 0; c < W; ++c) {
        if (!anchor || fp != 0) acc[c] += 0.5 * (double)rf[fp * W + c];
        if (!anchor || fm != 0) acc[c] += 0.5 * (double)rf[fm * W + c];
      }
    }
    return;
  }
  // per dimension: fine indices with cell in {C - 1, C}, i.e.
  // ceil((C-1) nf / nc) .. ceil((C+1) nf / nc) - 1 (unwrapped)
  const int32_t C3[3] = {i, j, k};
  int32_t lo[3], hi[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int64_t nf = fine.d[a], ncd = coarse.d[a];
    auto ceil_div = [](int64_t num, int64_t den) -> int64_t {
      return num >= 0 ? (num + den - 1) / den : -((-num) / den);
    };
    lo[a] = (int32_t)ceil_div((int64_t)(C3[a] - 1) * nf, ncd);
    hi[a] = (int32_t)ceil_div((int64_t)(C3[a] + 1) * nf, ncd) - 1;
    if (hi[a] - lo[a] + 1 > nf) hi[a] = lo[a] + (int32_t)nf - 1;
  }
  if (coarse.d[0] >= 2 && coarse.d[1] >= 2 && coarse.d[2] >= 2) {
    // Candidates walked per dimension with their cell (unwrapped) and
    // remainder kept up to date by addition (fine index f lies at coarse
    // coordinate f nc / nf = cell + rem / nf; nc < nf, so a step of one fine
    // index advances the cell by at most one).  With >= 2 coarse nodes per
    // dimension, vertex cell + o_v of the Kuhn tetrahedron that holds the
    // fine node is this node iff o_v = o = C - cell, so the weight is that
    // of vertex o: 1 - t_p0 (o = 0), t_p0 - t_p1 (o = e_p0), t_p1 - t_p2
    // (o = e_p0 + e_p1), t_p2 (o = (1,1,1)) -- the values kuhn_weights
    // gives, added in the same order as the generic loop below.
    int32_t cell0[3], rem0[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int32_t nf = fine.d[a], ncd = coarse.d[a];
      const int32_t q = lo[a] * ncd;
      const int32_t c = q >= 0 ? q / nf : -((-q + nf - 1) / nf);  // floor
      cell0[a] = c;
      rem0[a] = q - c * nf;
    }
    int32_t c0 = cell0[0], r0 = rem0[0];
    for (int32_t fi = lo[0]; fi <= hi[0]; ++fi) {
      const int32_t wi = fi < 0 ? fi + fine.d[0] : (fi >= fine.d[0] ? fi - fine.d[0] : fi);
      const double t0 = (double)r0 / (double)fine.d[0];
      const int32_t o0 = C3[0] - c0;
      int32_t c1 = cell0[1], r1 = rem0[1];
      for (int32_t fj = lo[1]; fj <= hi[1]; ++fj) {
        const int32_t wj = fj < 0 ? fj + fine.d[1] : (fj >= fine.d[1] ? fj - fine.d[1] : fj);
        const double t1 = (double)r1 / (double)fine.d[1];
        const int32_t o1 = C3[1] - c1;
        int32_t c2 = cell0[2], r2 = rem0[2];
        for (int32_t fk = lo[2]; fk <= hi[2]; ++fk) {
          const int32_t wk = fk < 0 ? fk + fine.d[2] : (fk >= fine.d[2] ? fk - fine.d[2] : fk);
          const double t[3] = {t0, t1, (double)r2 / (double)fine.d[2]};
          const int32_t o[3] = {o0, o1, C3[2] - c2};
          r2 += coarse.d[2];
          if (r2 >= fine.d[2]) r2 -= fine.d[2], ++c2;
          const int64_t f = ((int64_t)wi * fine.d[1] + wj) * fine.d[2] + wk;
          if (anchor && f == 0) continue;
          int p0, p1, p2;
          kuhn_order(t, p0, p1, p2);
          double wt;
          if (o[0] == 0 && o[1] == 0 && o[2] == 0)
            wt = 1.0 - t[p0];
          else if (o[0] == 1 && o[1] == 1 && o[2] == 1)
            wt = t[p2];
          else if (o[p0] == 1 && o[p1] == 0 && o[p2] == 0)
            wt = t[p0] - t[p1];
          else if (o[p0] == 1 && o[p1] == 1 && o[p2] == 0)
            wt = t[p1] - t[p2];
          else
            wt = 0.0;
          if (wt == 0.0) continue;
#pragma unroll
          for (int c = 0; c < W; ++c) acc[c] += wt * (double)rf[f * W + c];
        }
        r1 += coarse.d[1];
        if (r1 >= fine.d[1]) r1 -= fine.d[1], ++c1;
      }
      r0 += coarse.d[0];
      if (r0 >= fine.d[0]) r0 -= fine.d[0], ++c0;
    }
    return;
  }
  for (int32_t fi = lo[0]; fi <= hi[0]; ++fi)
    for (int32_t fj = lo[1]; fj <= hi[1]; ++fj)
      for (int32_t fk = lo[2]; fk <= hi[2]; ++fk) {
        const int32_t wi = fi < 0 ? fi + fine.d[0] : (fi >= fine.d[0] ? fi - fine.d[0] : fi);
        const int32_t wj = fj < 0 ? fj + fine.d[1] : (fj >= fine.d[1] ? fj - fine.d[1] : fj);
        const int32_t wk = fk < 0 ? fk + fine.d[2] : (fk >= fine.d[2] ? fk - fine.d[2] : fk);
        const int64_t f = ((int64_t)wi * fine.d[1] + wj) * fine.d[2] + wk;
        if (anchor && f == 0) continue;
        int64_t v[4];
        double w[4];
        kuhn_weights(wi, wj, wk, fine, coarse, v, w);
        double wt = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) wt += (v[q] == I) ? w[q] : 0.0;
        if (wt == 0.0) continue;
#pragma unroll
        for (int c = 0; c < W; ++c) acc[c] += wt * (double)rf[f * W + c];
      }
}

// b_c = R r_f over the coarse nodes of a workgroup range.
template <int BS, class T>
__device__ __forceinline__ void mg_restrict_rows(int64_t lb, int64_t lg, const Dims3& fine, const Dims3& coarse,
                                                 bool anchor, const T* __restrict__ rf, T* __restrict__ bc) {
  const int64_t nc = (int64_t)coarse.d[0] * coarse.d[1] * coarse.d[2];
  const int64_t stride = lg * kBlock;
  const bool nested = nested_dims(fine, coarse);
  for (int64_t I = lb * kBlock + threadIdx.x; I < nc; I += stride) {
    double acc[BS];
    restrict_node<BS, T>(I, fine, coarse, nested, anchor, rf, acc);
#pragma unroll
    for (int c = 0; c < BS; ++c) bc[I * BS + c] = (T)acc[c];
  }
}

// The restriction fused with the first (x = 0) Chebyshev step of the coarse
// level: b_c = R r_f, r_c = b_c, d_c = c0 D_c^-1 r_c, x_c = d_c -- the same
// arithmetic as mg_restrict_rows followed by mg_first_rows<ZERO>.
template <int BS, int K, class T>
__device__ __forceinline__ void mg_restrict_first_rows(int64_t lb, int64_t lg, const Dims3& fine,
                                                       const Dims3& coarse, bool anchor, const T* __restrict__ rf,
                                                       T* __restrict__ bc, T* __restrict__ rc, T* __restrict__ dc,
                                                       T* __restrict__ xc, const T* __restrict__ dinvc, double c0) {
  constexpr int NV = BS * K;
  const int64_t nc = (int64_t)coarse.d[0] * coarse.d[1] * coarse.d[2];
  const int64_t stride = lg * kBlock;
  const bool nested = nested_dims(fine, coarse);
  for (int64_t I = lb * kBlock + threadIdx.x; I < nc; I += stride) {
    double acc[NV], rv[NV], dv[NV];
    restrict_node<NV, T>(I, fine, coarse, nested, anchor, rf, acc);
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      const T bt = (T)acc[c];
      bc[I * NV + c] = bt;
      rc[I * NV + c] = bt;
      rv[c] = (double)bt;
    }
    apply_dinv<BS, K, T>(dinvc, I, rv, dv);
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      const T dn = (T)(c0 * dv[c]);
      dc[I * NV + c] = dn;
      xc[I * NV + c] = dn;
    }
  }
}

// (P e_c) at fine node f (W values per node); anchor: node 0 excluded on
// both grids.
template <int W, class T>
__host__ __device__ __forceinline__ void prolong_node(int64_t f, const Dims3& fine, const Dims3& coarse, bool nested,
                                                      bool anchor, const T* __restrict__ ec, double (&v)[W]) {
  const int32_t i = (int32_t)(f / ((int64_t)fine.d[1] * fine.d[2]));
  const int32_t j = (int32_t)((f / fine.d[2]) % fine.d[1]);
  const int32_t k = (int32_t)(f % fine.d[2]);
  if (nested) {
    const int32_t a = i & 1, bb = j & 1, cc = k & 1;
    const int32_t I = i >> 1, J = j >> 1, K = k >> 1;
    const int64_t c0 = wrap_idx(I, J, K, coarse.d);
    if ((a | bb | cc) == 0) {
#pragma unroll
      for (int c = 0; c < W; ++c) v[c] = (!anchor || c0 != 0) ? (double)ec[c0 * W + c] : 0.0;
    } else {
      const int64_t c1 = wrap_idx(I + a, J + bb, K + cc, coarse.d);
#pragma unroll
      for (int c = 0; c < W; ++c)
        v[c] = 0.5 * (((!anchor || c0 != 0) ? (double)ec[c0 * W + c] : 0.0) +
                      ((!anchor || c1 != 0) ? (double)ec[c1 * W + c] : 0.0));
    }
    return;
  }
  int64_t cv[4];
  double w[4];
  kuhn_weights(i, j, k, fine, coarse, cv, w);
#pragma unroll
  for (int c = 0; c < W; ++c) v[c] = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if ((anchor && cv[q] == 0) || w[q] == 0.0) continue;
#pragma unroll
    for (int c = 0; c < W; ++c) v[c] += w[q] * (double)ec[cv[q] * W + c];
  }
}

// x_f += P e_c (P1 interpolation); anchor: node 0 excluded on both grids.
template <int BS, class T>
__device__ __forceinline__ void mg_prolong_rows(int64_t lb, int64_t lg, const Dims3& fine, const Dims3& coarse,
                                                bool anchor, const T* __restrict__ ec, T* __restrict__ xf) {
  const int64_t nf = (int64_t)fine.d[0] * fine.d[1] * fine.d[2];
  const int64_t stride = lg * kBlock;
  const bool nested = nested_dims(fine, coarse);
  for (int64_t f = lb * kBlock + threadIdx.x; f < nf; f += stride) {
    if (anchor && f == 0) continue;
    double v[BS];
    prolong_node<BS, T>(f, fine, coarse, nested, anchor, ec, v);
#pragma unroll
    for (int c = 0; c < BS; ++c) xf[f * BS + c] = (T)((double)xf[f * BS + c] + v[c]);
  }
}

// x = Cinv b on the coarsest level (n rows of Cinv, K columns of b and x):
// one wavefront per row.
template <int K, class T>
__device__ __forceinline__ void mg_dense_rows(int64_t lb, int64_t lg, int64_t n, const T* __restrict__ cinv,
                                              const T* __restrict__ b, T* __restrict__ x) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = lg * (kBlock / 64);
  for (int64_t row = lb * (kBlock / 64) + threadIdx.x / 64; row < n; row += waves) {
    double s[K];
#pragma unroll
    for (int j = 0; j < K; ++j) s[j] = 0.0;
    if constexpr (sizeof(T) == 4) {
      if ((n & 3) == 0) {
        // FP32 rows of up to 16^3 (x3) entries: four consecutive entries per
        // lane in one 128-bit load (rows start 16-byte aligned), four such
        // loads in flight per lane
        typedef float f32x4 __attribute__((ext_vector_type(4)));
        const f32x4* a4 = reinterpret_cast<const f32x4*>(cinv + row * n);
        const int64_t n4 = n >> 2;
#pragma unroll 4
        for (int64_t c4 = lane; c4 < n4; c4 += 64) {
          const f32x4 a = a4[c4];
          const int64_t c = c4 << 2;
#pragma unroll
          for (int j = 0; j < K; ++j)
            s[j] += ((double)a.x * (double)b[c * K + j] + (double)a.y * (double)b[(c + 1) * K + j]) +
                    ((double)a.z * (double)b[(c + 2) * K + j] + (double)a.w * (double)b[(c + 3) * K + j]);
        }
#pragma unroll
        for (int j = 0; j < K; ++j) s[j] = group_sum<64>(s[j]);
        if (lane == 0)
#pragma unroll
          for (int j = 0; j < K; ++j) x[row * K + j] = (T)s[j];
        continue;
      }
    }
    for (int64_t c = lane; c < n; c += 64) {
      const double a = (double)cinv[row * n + c];
#pragma unroll
      for (int j = 0; j < K; ++j) s[j] += a * (double)b[c * K + j];
    }
#pragma unroll
    for (int j = 0; j < K; ++j) s[j] = group_sum<64>(s[j]);
    if (lane == 0)
#pragma unroll
      for (int j = 0; j < K; ++j) x[row * K + j] = (T)s[j];
  }
}

// ---- preconditioned CG vector bodies ------------------------------------------
// Vectors of length n = rows * K, column j = i % K; st[j] is column j's state.

// r = b - y (use_x0) or x = 0, r = b; acc[j] = ||b_j||^2, acc[K + j] = ||r_j||^2.
template <int K>
__device__ __forceinline__ void pcg_init_rows(int64_t lb, int64_t lg, int64_t n, const double* __restrict__ b,
                                              double* __restrict__ x, const double* __restrict__ y,
                                              double* __restrict__ r, int use_x0, double (&acc)[2 * K]) {
  const int64_t stride = lg * kBlock;
  for (int64_t i = lb * kBlock + threadIdx.x; i < n; i += stride) {
    const int j = (int)(i % K);
    const double bi = b[i];
    double ri;
    if (use_x0) {
      ri = bi - y[i];
    } else {
      x[i] = 0.0;
      ri = bi;
    }
    r[i] = ri;
    acc[j] += bi * bi;
    acc[K + j] += ri * ri;
  }
}

// live columns: x += alpha p, r -= alpha q; acc[j] += r_j.r_j.  p and q of
// type PQ: FP64, or FP32 in the mixed-precision solves (whose products are
// FP32-accurate; x and r stay FP64).
template <int K, class PQ = double>
__device__ __forceinline__ void pcg_xr_rows(int64_t lb, int64_t lg, int64_t n, const CgState* st,
                                            double* __restrict__ x, double* __restrict__ r,
                                            const PQ* __restrict__ p, const PQ* __restrict__ q, double (&acc)[K]) {
  double alpha[K];
  bool live[K];
#pragma unroll
  for (int j = 0; j < K; ++j) live[j] = !st[j].done, alpha[j] = st[j].alpha;
  const int64_t stride = lg * kBlock;
  for (int64_t i = lb * kBlock + threadIdx.x; i < n; i += stride) {
    const int j = (int)(i % K);
    if (!live[j]) continue;
    x[i] += alpha[j] * (double)p[i];
    const double rn = r[i] - alpha[j] * (double)q[i];
    r[i] = rn;
    acc[j] += rn * rn;
  }
}

// acc[NV] += per-component sums of r (vector entry i to component i % NV)
template <int NV>
__device__ __forceinline__ void pcg_rsum_rows(int64_t lb, int64_t lg, int64_t n, const double* __restrict__ r,
                                              double (&acc)[NV]) {
  const int64_t stride = lg * kBlock;
  for (int64_t i = lb * kBlock + threadIdx.x; i < n; i += stride) acc[i % NV] += r[i];
}

// Entry i of the preconditioned residual.  FB = 0: the V-cycle's output y.
// Floating hierarchy (FB = block size, one column): z = F^T y + e_0 r_0,
// i.e. y shifted by its node-0 value (so that node 0 stays 0) and r_0 at
// node 0 (the anchor's identity row).
template <int FB, class T>
__device__ __forceinline__ double pcg_z(int64_t i, const T* __restrict__ z, const double* __restrict__ r) {
  if constexpr (FB == 0) {
    return (double)z[i];
  } else {
    return (i < FB) ? r[i] : (double)z[i] - (double)z[i % FB];
  }
}

// acc[j] = r_j.z_j, acc[K + j] = q_j.z_j (after the first iteration; q = A p
// of the iteration, type PQ)
template <int K, class T, int FB = 0, class PQ = double>
__device__ __forceinline__ void pcg_dot_rows(int64_t lb, int64_t lg, int64_t n, const CgState* st,
                                             const double* __restrict__ r, const PQ* __restrict__ q,
                                             const T* __restrict__ z, double (&acc)[2 * K]) {
  static_assert(FB == 0 || K == 1, "floating hierarchies solve one column");
  bool first[K];
#pragma unroll
  for (int j = 0; j < K; ++j) first[j] = st[j].it == 0;
  const int64_t stride = lg * kBlock;
  for (int64_t i = lb * kBlock + threadIdx.x; i < n; i += stride) {
    const int j = (int)(i % K);
    const double zi = pcg_z<FB, T>(i, z, r);
    acc[j] += r[i] * zi;
    if (!first[j]) acc[K + j] += (double)q[i] * zi;
  }
}

// live columns: p = z (first iteration) or z + beta p, formed in FP64 and
// stored as PQ
template <int K, class T, int FB = 0, class PQ = double>
__device__ __forceinline__ void pcg_pupd_rows(int64_t lb, int64_t lg, int64_t n, const CgState* st,
                                              const T* __restrict__ z, PQ* __restrict__ p,
                                              const double* __restrict__ r) {
  static_assert(FB == 0 || K == 1, "floating hierarchies solve one column");
  double beta[K];
  bool live[K], first[K];
#pragma unroll
  for (int j = 0; j < K; ++j) live[j] = !st[j].done, first[j] = st[j].it == 0, beta[j] = st[j].beta;
  const int64_t stride = lg * kBlock;
  for (int64_t i = lb * kBlock + threadIdx.x; i < n; i += stride) {
    const int j = (int)(i % K);
    if (!live[j]) continue;
    const double zi = pcg_z<FB, T>(i, z, r);
    p[i] = (PQ)(first[j] ? zi : zi + beta[j] * (double)p[i]);
  }
}

// Reliable update fused with the FP64 product: for the rows of a
// workgroup range, r = b - A x (FP64 form) on the columns that asked for an
// update; acc[j] += r_j.r_j.
template <int BS, int K, bool ST>
__device__ __forceinline__ void pcg_true_rows(int64_t lb, int64_t lg, int32_t nb, const int64_t* __restrict__ sptr,
                                              const int32_t* __restrict__ col, const double* __restrict__ val,
                                              const double* __restrict__ dg, const GridArg& ga,
                                              const double* __restrict__ x, const double* __restrict__ b,
                                              double* __restrict__ r, const CgState* st, double (&acc)[K]) {
  constexpr int NV = BS * K;
  bool upd[K];
#pragma unroll
  for (int j = 0; j < K; ++j) upd[j] = st[j].upd != 0;
  const int64_t stride = lg * kBlock;
  for (int64_t row = lb * kBlock + threadIdx.x; row < (int64_t)((nb + 63) & ~63); row += stride) {
    double sum[NV], xo[NV];
    sell_row<BS, K, double, false, ST>(row, nb, sptr, col, val, dg, x, nullptr, ga, sum, xo);
    if (row >= nb) continue;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int j = v % K;
      if (!upd[j]) continue;
      const double ri = b[row * NV + v] - sum[v];
      r[row * NV + v] = ri;
      acc[j] += ri * ri;
    }
  }
}

// some column asked for a reliable update
template <int K>
__device__ __forceinline__ bool any_upd(const CgState* st) {
  bool u = false;
#pragma unroll
  for (int j = 0; j < K; ++j) u = u || st[j].upd != 0;
  return u;
}

// rho = r.z, beta = (r - r_old).z / rho_old = -alpha q.z / rho_old (flexible
// CG: r - r_old = -alpha q), from the sums r.z and q.z.
// r.z of an SPD preconditioner is positive; a preconditioner applied in
// reduced precision to a residual at its rounding level can lose that, and
// then the direction restarts (beta = 0) and the products go to FP64.
__device__ inline void pcg_state_rz(CgState& c, double rz, double qz) {
  c.rho_old = c.rho;
  c.rho = rz;
  c.beta = (c.it == 0) ? 0.0 : -c.alpha * qz / c.rho_old;
  if (!isfinite(c.beta) || !(rz > 0.0)) {
    c.beta = 0.0;
    c.p64 = 1;
  }
}

}  // namespace cpm
