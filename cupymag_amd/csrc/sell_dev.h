// Device bodies of the SELL-64 SpMV and of the CG vector updates, written
// over a workgroup range (lb = this workgroup's index inside the range,
// lg = workgroups in the range) so that one body serves both a launch that
// covers a single system (lb = blockIdx.x, lg = gridDim.x) and a segmented
// launch in which every system of a batch owns its own range of workgroups.
// Partial sums are stored at partials[k * gridDim.x + blockIdx.x] in both
// cases; the reducer of a system adds the entries of its own range.
#pragma once
#include <cmath>
#include <type_traits>

#include "device.h"

#ifndef SELL_HALF_UNROLL
#define SELL_HALF_UNROLL 7
#endif
#ifndef SELL_FLOAT_UNROLL
#define SELL_FLOAT_UNROLL 7
#endif
#ifndef SELL_DOUBLE_UNROLL
#define SELL_DOUBLE_UNROLL 7
#endif

namespace cpm {

// Periodic structured grid of an index-free (stencil) SELL operator.
// kuhn: the slots are the 14 neighbours of a Kuhn mesh in the order of
// kKuhnCodes (the host checks soff), so the kernels decode them at compile
// time; 2: besides, the FP32 or FP16 values passed with it keep half of the
// stencil (the forward slots 7..13, 7 planes per slice: Sell::half32,
// k_mg_make_half).  FP64 values hold every slot unless h64 is set as well
// (Sell::half64), so one grid argument serves a kernel that reads both an
// FP32 half form and the full FP64 form.
struct GridArg {
  int32_t nx, ny, nz, nst;
  const int32_t* soff;
  int32_t kuhn;
  int32_t xg;        // XCD group of the row launches (xcd_perm); 0 or 1: dispatch order
  int32_t shz, shy;  // log2(nz) + 1 and log2(ny) + 1 when nz and ny are powers of two, else 0
  int32_t h64;       // with kuhn = 2: the FP64 values passed with it keep half of the stencil too
};

// Slot codes (di+1) | (dj+1) << 2 | (dk+1) << 4 of the neighbours of a node
// of a Kuhn mesh (cube diagonal (0,0,0)-(1,1,1)): +-e1, +-e2, +-e3,
// +-(e1+e2), +-(e2+e3), +-(e1+e3), +-(e1+e2+e3), ascending (the order the
// SELL builder gives them).
constexpr int kKuhnSlots = 14;
constexpr int32_t kKuhnCodes[kKuhnSlots] = {0, 1, 4, 5, 16, 17, 20, 22, 25, 26, 37, 38, 41, 42};

// Neighbour row of `row` at Kuhn slot t of a periodic grid (dims >= 3).
__host__ __device__ __forceinline__ int64_t kuhn_neighbour(int64_t row, const GridArg& ga, int t) {
  const int32_t code = kKuhnCodes[t];
  const int32_t nyz = ga.ny * ga.nz;
  int32_t i = (int32_t)(row / nyz) + (code & 3) - 1;
  int32_t j = (int32_t)((row / ga.nz) % ga.ny) + ((code >> 2) & 3) - 1;
  int32_t k = (int32_t)(row % ga.nz) + ((code >> 4) & 3) - 1;
  i += (i < 0) ? ga.nx : 0;
  i -= (i >= ga.nx) ? ga.nx : 0;
  j += (j < 0) ? ga.ny : 0;
  j -= (j >= ga.ny) ? ga.ny : 0;
  k += (k < 0) ? ga.nz : 0;
  k -= (k >= ga.nz) ? ga.nz : 0;
  return ((int64_t)i * ga.ny + j) * ga.nz + k;
}

// Host, at build time, for a Kuhn stencil form with FP64 values val
// ([slot][bs*bs][64], slot = slice*14 + t): whether A(i, i-d) = A(i-d, i)^T
// for every row and backward slot, to 1e-12 of the largest entry.
inline bool kuhn_slot_symmetric(int bs, int32_t nbrows, const GridArg& ga, const double* val, double tol = 1e-12) {
  const int bb = bs * bs;
  double amax = 0.0, dmax = 0.0;
  for (int64_t b = 0; b < nbrows; ++b) {
    const int64_t s = b >> 6, l = b & 63;
    for (int t = 0; t < 7; ++t) {
      const int64_t c = kuhn_neighbour(b, ga, t);
      const int64_t cs = (c >> 6) * kKuhnSlots + (kKuhnSlots - 1 - t), lc = c & 63;
      for (int d = 0; d < bs; ++d)
        for (int e = 0; e < bs; ++e) {
          const double a = val[((s * kKuhnSlots + t) * bb + d * bs + e) * 64 + l];
          const double an = val[(cs * bb + e * bs + d) * 64 + lc];
          dmax = dmax > std::fabs(a - an) ? dmax : std::fabs(a - an);
          amax = amax > std::fabs(a) ? amax : std::fabs(a);
        }
    }
  }
  return dmax <= tol * amax;
}

// Host: the FP32 half form of a slot-symmetric Kuhn stencil operator,
// v32[((slice*7 + u)*bs*bs + t)*64 + l] = fp32 of slot 7 + u (nsl*7*bs*bs*64
// values, read with GridArg::kuhn = 2), and dgc = dg + the rounding of the
// fourteen blocks as the FP32 product reads them (the row's forward blocks,
// and for its backward slots the neighbours' forward blocks, transposed), so
// that the row sums of the FP32 form equal the FP64 ones.
inline void kuhn_half32(int bs, int32_t nbrows, int64_t nsl, const GridArg& ga, const double* val,
                        const double* dg, float* v32, double* dgc) {
  const int bb = bs * bs;
  for (int64_t s = 0; s < nsl; ++s)
    for (int u = 0; u < 7; ++u)
      for (int t = 0; t < bb; ++t)
        for (int l = 0; l < 64; ++l)
          v32[((s * 7 + u) * bb + t) * 64 + l] = (float)val[((s * kKuhnSlots + 7 + u) * bb + t) * 64 + l];
  for (int64_t b = 0; b < nbrows; ++b) {
    const int64_t s = b >> 6, l = b & 63;
    double comp[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int u = 0; u < 7; ++u)
      for (int t = 0; t < bb; ++t) {
        const double a = val[((s * kKuhnSlots + 7 + u) * bb + t) * 64 + l];
        comp[t] += a - (double)(float)a;
      }
    for (int t = 0; t < 7; ++t) {
      const int64_t c = kuhn_neighbour(b, ga, t);
      const int64_t cs = (c >> 6) * kKuhnSlots + (kKuhnSlots - 1 - t), lc = c & 63;
      for (int d = 0; d < bs; ++d)
        for (int e = 0; e < bs; ++e) {
          const double a = val[((s * kKuhnSlots + t) * bb + d * bs + e) * 64 + l];
          comp[d * bs + e] += a - (double)(float)val[(cs * bb + e * bs + d) * 64 + lc];
        }
    }
    for (int t = 0; t < bb; ++t) dgc[b * bb + t] = dg[b * bb + t] + comp[t];
  }
}

// Diagonal blocks of a SELL form with off-diagonal values of type VT: FP64,
// or FP32 for the FP16 multigrid level forms.
template <class VT>
using sell_diag_t = typename std::conditional<sizeof(VT) == 2, float, double>::type;

// sum = row `row` of A x (ADDX32: A (x + x32)) for
// BS x BS blocks and K interleaved right-hand sides x[(c*BS + e)*K + j];
// xo = the row's own entries of x (+ x32).  Padding rows (row >= nbrows)
// only stream their (zero) slots and leave xo unset.  Off-diagonal values
// of type VT (default T); for VT = _Float16 they are stored times a power of
// two and their sum is scaled back by vscale, and the diagonal blocks are
// FP32 (sell_diag_t).  UF > 0: UF slots per trip through the stored offsets
// of the stencil (GridArg::kuhn = 0) -- fewer loads in flight and fewer
// registers, for product paths that rarely run inside a kernel whose
// occupancy is set by its main path.  The off-diagonal products are added
// in FP64, or in FP32 with ACC32 (the FP32 V-cycle sweeps, whose vectors
// are FP32); the diagonal block's product in FP64.
template <int BS, int K, class T, bool ADDX32, bool ST, class VT = T, int UF = 0, bool ACC32 = false>
__host__ __device__ __forceinline__ void sell_row(int64_t row, int32_t nbrows, const int64_t* __restrict__ sptr,
                                         const int32_t* __restrict__ col, const VT* __restrict__ val,
                                         const sell_diag_t<VT>* __restrict__ dg, const T* __restrict__ x,
                                         const float* __restrict__ x32, const GridArg& ga, double (&sum)[BS * K],
                                         double (&xo)[BS * K], double vscale = 1.0) {
  constexpr int BB = BS * BS;
  constexpr int NV = BS * K;
  const int64_t s = row >> 6, l = row & 63;
  const int64_t s0 = ST ? s * ga.nst : sptr[s];
  const int64_t s1 = ST ? s0 + ga.nst : sptr[s + 1];
  int32_t ri = 0, rj = 0, rk = 0;
  if (ST && row < nbrows) {
    const int32_t r = (int32_t)row;
    if (ga.shz && ga.shy) {
      const int32_t lz = ga.shz - 1, ly = ga.shy - 1;
      ri = r >> (lz + ly);
      rj = (r >> lz) & ((1 << ly) - 1);
      rk = r & ((1 << lz) - 1);
    } else {
      const int32_t nyz = ga.ny * ga.nz;
      ri = r / nyz;
      rj = (r / ga.nz) % ga.ny;
      rk = r % ga.nz;
    }
  }
  using AT = typename std::conditional<ACC32, float, double>::type;
  AT acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = (AT)0;
  // matrix planes are read once per launch: streaming (non-temporal) loads
  // keep them from evicting the gathered vector entries; two slots per
  // trip keep more loads in flight
  auto slot_fma_at = [&](int64_t slot, int64_t c) {
    AT xv[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      xv[v] = (AT)x[c * NV + v];
      if (ADDX32) xv[v] += (AT)x32[c * NV + v];
    }
    if constexpr (sizeof(VT) == 2 && BB == 9) {
      // FP16 3x3 blocks: entries 0..7 of a lane in one 16-byte word (one
      // 128-bit load), entry 8 on its own (2 loads per slot instead of 9)
      AT av[9];
      typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
      const u32x4* w = reinterpret_cast<const u32x4*>(val + slot * 576);
      const u32x4 w4 = __builtin_nontemporal_load(&w[l]);
      const uint32_t pw4[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t pw = pw4[q];
        av[2 * q] = (AT)__builtin_bit_cast(_Float16, (uint16_t)(pw & 0xffffu));
        av[2 * q + 1] = (AT)__builtin_bit_cast(_Float16, (uint16_t)(pw >> 16));
      }
      av[8] = (AT)__builtin_nontemporal_load(&val[slot * 576 + 512 + l]);
#pragma unroll
      for (int d = 0; d < BS; ++d)
#pragma unroll
        for (int e = 0; e < BS; ++e)
#pragma unroll
          for (int j = 0; j < K; ++j) acc[d * K + j] += av[d * BS + e] * xv[e * K + j];
    } else {
#pragma unroll
      for (int d = 0; d < BS; ++d)
#pragma unroll
        for (int e = 0; e < BS; ++e) {
          const AT a = (AT)__builtin_nontemporal_load(&val[(slot * BB + d * BS + e) * 64 + l]);
#pragma unroll
          for (int j = 0; j < K; ++j) acc[d * K + j] += a * xv[e * K + j];
        }
    }
  };
  auto slot_fma = [&](int64_t slot) {
    int64_t c;
    if (ST) {
      // neighbour on the periodic grid (padding rows gather row 0)
      const int32_t code = ga.soff[slot - s0];
      int32_t i2 = ri + (code & 3) - 1, j2 = rj + ((code >> 2) & 3) - 1, k2 = rk + ((code >> 4) & 3) - 1;
      i2 += (i2 < 0) ? ga.nx : 0;
      i2 -= (i2 >= ga.nx) ? ga.nx : 0;
      j2 += (j2 < 0) ? ga.ny : 0;
      j2 -= (j2 >= ga.ny) ? ga.ny : 0;
      k2 += (k2 < 0) ? ga.nz : 0;
      k2 -= (k2 >= ga.nz) ? ga.nz : 0;
      c = (row < nbrows) ? ((int64_t)i2 * ga.ny + j2) * ga.nz + k2 : 0;
    } else {
      c = __builtin_nontemporal_load(&col[slot * 64 + l]);
    }
    slot_fma_at(slot, c);
  };
  // slots per trip: loads of U slots in flight before their FMAs (7 = half
  // the 14 neighbours of a Kuhn row; a sweep of 2..7 at 128^3 favoured 7 for
  // every value type, most for FP16)
  constexpr int U = UF > 0                ? UF
                    : (sizeof(VT) == 2) ? SELL_HALF_UNROLL
                    : (sizeof(VT) == 4) ? SELL_FLOAT_UNROLL
                                        : SELL_DOUBLE_UNROLL;
  if (ST && ga.kuhn) {
    // Kuhn stencil: slot k's offset is known at compile time, so its
    // neighbour is the row plus one wrap-aware step per dimension (computed
    // once per row), in 32-bit arithmetic, with no per-slot loads of codes
    const int32_t r = (row < nbrows) ? (int32_t)row : 0;  // padding rows gather row 0's neighbours
    const int32_t nyz = ga.ny * ga.nz;
    const int32_t stp[3][3] = {
        {ri == 0 ? (ga.nx - 1) * nyz : -nyz, 0, ri == ga.nx - 1 ? (1 - ga.nx) * nyz : nyz},
        {rj == 0 ? (ga.ny - 1) * ga.nz : -ga.nz, 0, rj == ga.ny - 1 ? (1 - ga.ny) * ga.nz : ga.nz},
        {rk == 0 ? ga.nz - 1 : -1, 0, rk == ga.nz - 1 ? 1 - ga.nz : 1}};
    auto kuhn_full = [&]() {
#pragma unroll
      for (int t = 0; t < kKuhnSlots; t += 7) {
#pragma unroll
        for (int u = 0; u < 7; ++u) {
          const int32_t code = kKuhnCodes[t + u];
          const int32_t c = r + stp[0][code & 3] + stp[1][(code >> 2) & 3] + stp[2][(code >> 4) & 3];
          slot_fma_at(s * kKuhnSlots + t + u, c);
        }
      }
    };
    {
      if (ga.kuhn == 2 && (sizeof(VT) <= 4 || ga.h64)) {
        // half storage (FP32 and FP16 forms of slot-symmetric operators, the
        // FP64 form of those symmetric to rounding):
        // slice s holds planes s*7 + u with the forward blocks of slots 7 + u
        // (offsets +d); the backward block of slot t < 7 (offset -d) is
        // A(i, i-d) = A(i-d, i)^T, the forward block of slot 13 - t (plane
        // 6 - t) of the neighbour row, read transposed.  Plain (cached)
        // loads: every block serves two rows.
#pragma unroll
        for (int t = 0; t < kKuhnSlots; t += 7) {
#pragma unroll
          for (int u = 0; u < 7; ++u) {
            const int32_t code = kKuhnCodes[t + u];
            const int32_t c = r + stp[0][code & 3] + stp[1][(code >> 2) & 3] + stp[2][(code >> 4) & 3];
            const bool back = t == 0;
            const int64_t plane = back ? (int64_t)(c >> 6) * 7 + (6 - u) : s * 7 + u;
            const int64_t lane = back ? (int64_t)(c & 63) : l;
            AT xv[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
              xv[v] = (AT)x[(int64_t)c * NV + v];
              if (ADDX32) xv[v] += (AT)x32[(int64_t)c * NV + v];
            }
            if constexpr (sizeof(VT) >= 4) {
              AT av[BB];
#pragma unroll
              for (int q = 0; q < BB; ++q) av[q] = (AT)val[(plane * BB + q) * 64 + lane];
#pragma unroll
              for (int d = 0; d < BS; ++d)
#pragma unroll
                for (int e = 0; e < BS; ++e) {
                  const AT a = back ? av[e * BS + d] : av[d * BS + e];
#pragma unroll
                  for (int j = 0; j < K; ++j) acc[d * K + j] += a * xv[e * K + j];
                }
            } else if constexpr (BB == 9) {
              AT av[9];
              typedef uint32_t u32x4h __attribute__((ext_vector_type(4)));
              const u32x4h w4 = reinterpret_cast<const u32x4h*>(val + plane * 576)[lane];
              const uint32_t pw4[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
              for (int q = 0; q < 4; ++q) {
                av[2 * q] = (AT)__builtin_bit_cast(_Float16, (uint16_t)(pw4[q] & 0xffffu));
                av[2 * q + 1] = (AT)__builtin_bit_cast(_Float16, (uint16_t)(pw4[q] >> 16));
              }
              av[8] = (AT)val[plane * 576 + 512 + lane];
#pragma unroll
              for (int d = 0; d < BS; ++d)
#pragma unroll
                for (int e = 0; e < BS; ++e) {
                  const AT a = back ? av[e * BS + d] : av[d * BS + e];
#pragma unroll
                  for (int j = 0; j < K; ++j) acc[d * K + j] += a * xv[e * K + j];
                }
            } else {
              static_assert(BB == 1, "1x1 or 3x3 blocks");
              const AT a = (AT)val[plane * 64 + lane];
#pragma unroll
              for (int j = 0; j < K; ++j) acc[j] += a * xv[j];
            }
          }
        }
      } else {
        kuhn_full();
      }
    }
  } else {
    int64_t slot = s0;
    for (; slot + U - 1 < s1; slot += U) {
#pragma unroll
      for (int u = 0; u < U; ++u) slot_fma(slot + u);
    }
    for (; slot < s1; ++slot) slot_fma(slot);
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) sum[v] = (sizeof(VT) == 2) ? (double)acc[v] * vscale : (double)acc[v];
  if (row < nbrows) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      xo[v] = (double)x[row * NV + v];
      if (ADDX32) xo[v] += (double)x32[row * NV + v];
    }
#pragma unroll
    for (int d = 0; d < BS; ++d)
#pragma unroll
      for (int e = 0; e < BS; ++e) {
        const double a = (double)dg[row * BB + d * BS + e];
#pragma unroll
        for (int j = 0; j < K; ++j) sum[d * K + j] += a * xo[e * K + j];
      }
  }
}

// y = A x (ADDX32: A (x + x32)) over the block rows of a workgroup range; CG:
// acc[j] += x.y of column j over the rows this thread owns.
template <int BS, int K, class T, bool CG, bool ADDX32, bool ST, class VT = T, int UF = 0>
__device__ __forceinline__ void sell_rows(int64_t lb, int64_t lg, int32_t nbrows, const int64_t* __restrict__ sptr,
                                          const int32_t* __restrict__ col, const VT* __restrict__ val,
                                          const sell_diag_t<VT>* __restrict__ dg, const T* __restrict__ x,
                                          const float* __restrict__ x32, T* __restrict__ y, const GridArg& ga,
                                          double (&acc)[K]) {
  constexpr int NV = BS * K;
  const int64_t stride = lg * kBlock;
  for (int64_t row = lb * kBlock + threadIdx.x; row < (int64_t)((nbrows + 63) & ~63); row += stride) {
    double sum[NV], xo[NV];
    sell_row<BS, K, T, ADDX32, ST, VT, UF>(row, nbrows, sptr, col, val, dg, x, x32, ga, sum, xo);
    if (row < nbrows) {
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const T yv = (T)sum[v];
        y[row * NV + v] = yv;
        if (CG) acc[v % K] += xo[v] * (double)yv;
      }
    }
  }
}

// p = r (first iteration) or r + beta p, on the live columns.
template <int K, class T>
__device__ __forceinline__ void pupd_rows(int64_t lb, int64_t lg, int64_t n, const T* __restrict__ r,
                                          T* __restrict__ p, const CgState* st) {
  double beta[K];
  bool live[K], first[K];
  bool any = false;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    live[j] = !st[j].done;
    first[j] = st[j].it == 0;
    beta[j] = (live[j] && !first[j]) ? st[j].rho / st[j].rho_old : 0.0;
    any |= live[j];
  }
  if (!any) return;
  const int64_t stride = lg * kBlock;
  for (int64_t i = lb * kBlock + threadIdx.x; i < n; i += stride) {
    const int j = (int)(i % K);
    if (!live[j]) continue;
    p[i] = first[j] ? r[i] : (T)(r[i] + (T)beta[j] * p[i]);
  }
}

// x += alpha p, r -= alpha q on the live columns; partials of r.r.  The
// whole workgroup returns when no column is live (no partials then: the
// reducer skips such systems too).
template <int K, class T>
__device__ __forceinline__ void xr_rows(int64_t lb, int64_t lg, int64_t n, T* __restrict__ x, T* __restrict__ r,
                                        const T* __restrict__ p, const T* __restrict__ q, const CgState* st,
                                        double* partials) {
  double a[K];
  bool live[K];
  bool any = false;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    live[j] = !st[j].done;
    a[j] = live[j] ? st[j].alpha : 0.0;
    any |= live[j];
  }
  if (!any) return;
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) acc[j] = 0.0;
  const int64_t stride = lg * kBlock;
  for (int64_t i = lb * kBlock + threadIdx.x; i < n; i += stride) {
    const int j = (int)(i % K);
    if (!live[j]) continue;
    const T aj = (T)a[j];
    x[i] = x[i] + aj * p[i];
    const T ri = r[i] - aj * q[i];
    r[i] = ri;
    acc[j] += (double)ri * (double)ri;
  }
  block_partial<K>(acc, partials);
}

// Mixed precision, on the columns that requested an update (y = A (x64 +
// x32) computed just before): x64 += x32, x32 = 0, r64 = b - y, r32 = r64,
// partials of r64.r64.
template <int K>
__device__ __forceinline__ void rel_r_rows(int64_t lb, int64_t lg, int64_t n, const double* __restrict__ b,
                                           const double* __restrict__ y, double* __restrict__ x,
                                           float* __restrict__ x32, double* __restrict__ r64,
                                           float* __restrict__ r32, const CgState* st, double* partials) {
  bool upd[K];
  bool any = false;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    upd[j] = st[j].upd != 0;
    any |= upd[j];
  }
  if (!any) return;
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) acc[j] = 0.0;
  const int64_t stride = lg * kBlock;
  for (int64_t i = lb * kBlock + threadIdx.x; i < n; i += stride) {
    const int j = (int)(i % K);
    if (!upd[j]) continue;
    x[i] += (double)x32[i];
    x32[i] = 0.0f;
    const double ri = b[i] - y[i];
    r64[i] = ri;
    r32[i] = (float)ri;
    acc[j] += ri * ri;
  }
  block_partial<K>(acc, partials);
}

// CG start: x = 0, r = b (or r = b - y with y = A x0); acc[j] = ||b_j||^2,
// acc[K + j] = r_j.r_j.  REL (mixed precision): also r32 = r, x32 = 0.
template <int K, bool REL>
__device__ __forceinline__ void init_rows(int64_t lb, int64_t lg, int64_t n, const double* __restrict__ b,
                                          double* __restrict__ x, const double* __restrict__ y,
                                          double* __restrict__ r, float* __restrict__ r32,
                                          float* __restrict__ x32, int use_x0, double (&acc)[2 * K]) {
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
    if (REL) {
      r32[i] = (float)ri;
      x32[i] = 0.0f;
    }
    acc[j] += bi * bi;
    acc[K + j] += ri * ri;
  }
}

// State of K columns from the start sums (thread 0).  FP64 CG keeps CuPy's
// rule (return b when ||b|| = 0, otherwise at least one iteration); the
// mixed solver also stops at once when r0 already meets the tolerance.
template <int K, bool REL>
__device__ inline void cg_state_init(CgState* st, const double (&acc)[2 * K], double tol, int32_t maxiter) {
#pragma unroll
  for (int j = 0; j < K; ++j) {
    CgState& s = st[j];
    s.bnorm2 = acc[j];
    s.rho = acc[K + j];
    s.rho_old = 0.0;
    s.tol = tol;
    s.atol = tol * sqrt(acc[j]);
    s.it = 0;
    s.fail = 0;
    s.maxiter = maxiter;
    s.resid = sqrt(acc[K + j]);
    s.maxr = s.resid;
    s.upd = 0;
    s.p64 = 0;
    s.it_upd = 0;
    if (REL)
      s.done = (acc[j] == 0.0 || s.resid <= s.atol || maxiter <= 0) ? 1 : 0;
    else
      s.done = (acc[j] == 0.0 || maxiter <= 0) ? 1 : 0;
  }
}

// x = 0 on the columns whose right-hand side is zero (CuPy returns b there).
template <int K>
__device__ __forceinline__ void zero_b_rows(int64_t lb, int64_t lg, int64_t n, double* __restrict__ x,
                                            const CgState* st) {
  const int64_t stride = lg * kBlock;
  for (int64_t i = lb * kBlock + threadIdx.x; i < n; i += stride)
    if (st[i % K].bnorm2 == 0.0) x[i] = 0.0;
}

}  // namespace cpm
