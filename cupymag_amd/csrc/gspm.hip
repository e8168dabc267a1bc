// Field operators and the device-resident GSPM LLG step (FP64).
//
// Restates cupymag/core/Micromagnetics.py:110-307 as a sequence of fused
// HIP kernels on one stream.  The only host synchronisations are the CG
// batch boundaries and the final report (sumsq, avg_m), which the reference
// also pulls to the host every step (:299-309).
//
// Kernel map (reference line -> kernel):
//   :112-121 m@R, build_E0_from_m, E0@M^T             -> k_e0
//   :123     F_el @ E0                                -> launch_spmv
//   :127-129 compute_E_from_u (6 or 9 SpMVs + mixing) -> k_strain (one pass
//            over the shared Fx/Fy/Fz pattern, 9 row sums)
//   :132,143 volume averages of E and m100            -> k_strain_avg
//   :135-162 E -= avg, E += Ebar, E @ M^-T            -> k_efinal
//   :166-168 -(Fx m~x + Fy m~y + Fz m~z), b[0] = 0    -> k_demag_rhs
//   :171     Htilde = -F U                            -> k_hd
//   :175-215 effective field, f*dt + m                -> k_field1
//   :217-219,227,233,280-288 F1 @ .                   -> launch_spmv
//   :225-227 / :231-233 m1*, m2* and their RHS        -> k_gs1 / k_gs2
//   :237-278 m3*, stage-2 field, f*dt2 + m*           -> k_gs3_field2
//   :292-306 normalise, defect -> 0, sumsq, avg_m     -> k_project
#include <algorithm>
#include <thread>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "device.h"
#include "mg.h"
#include "multi.h"
#include "multi_mg.h"
#include "sweep_ctl.h"

struct cpm_fel;
namespace cpm {
int fel_apply(const cpm_fel* F, const double* E0, double* b);
int64_t fel_rows(const cpm_fel* F);
int fder_strain_demag(const cpm_fel* F, const double* u, int rot, const double* dR, const double* m,
                      const double* davg, double* E, double* bd);
int fder_hd(const cpm_fel* F, const double* U, double* H);
}  // namespace cpm

namespace cpm {

struct FieldPrm {
  double K1, c1, c2, c3, l100, dt, dt2;
  double Nd[3], H[3];
  double R[9];
  int rot;
  const double* Hdev;  // the external field in device memory (sweep control), else H
};



__device__ inline void rot_apply(const double* R, double a0, double a1, double a2, double& b0, double& b1,
                                 double& b2) {
  // (a0,a1,a2) @ R  (row vector times R), in j order
  b0 = a0 * R[0] + a1 * R[3] + a2 * R[6];
  b1 = a0 * R[1] + a1 * R[4] + a2 * R[7];
  b2 = a0 * R[2] + a1 * R[5] + a2 * R[8];
}

// ---- E0 = build_E0_from_m(m or m@R) [@ Mv^T] ------------------------------
__global__ void __launch_bounds__(kBlock) k_e0(int64_t n, const double* __restrict__ m, double l100, double l111,
                                               int rot, const double* __restrict__ R /*9, device*/,
                                               const double* __restrict__ Mv /*36, device*/,
                                               double* __restrict__ E0) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    double mx = m[3 * i], my = m[3 * i + 1], mz = m[3 * i + 2];
    if (rot) {
      double a, b, c;
      rot_apply(R, mx, my, mz, a, b, c);
      mx = a, my = b, mz = c;
    }
    const double t = 1.0 / 3.0;
    double e[6];
    e[0] = (1.5 * l100) * (mx * mx - t);
    e[1] = (1.5 * l100) * (my * my - t);
    e[2] = (1.5 * l100) * (mz * mz - t);
    e[3] = 2.0 * ((1.5 * l111) * (mx * my));
    e[4] = 2.0 * ((1.5 * l111) * (my * mz));
    e[5] = 2.0 * ((1.5 * l111) * (mx * mz));
    if (rot) {
      double o[6];
      for (int r = 0; r < 6; ++r) {
        double s = 0.0;
        for (int j = 0; j < 6; ++j) s += e[j] * Mv[r * 6 + j];
        o[r] = s;
      }
      for (int r = 0; r < 6; ++r) e[r] = o[r];
    }
    for (int v = 0; v < 6; ++v) E0[6 * i + v] = e[v];
  }
}

// ---- strain: 9 weak-form derivatives of u on the Fx/Fy/Fz pattern ------------
template <int G>
__global__ void __launch_bounds__(kBlock) k_strain(int32_t n, const int32_t* __restrict__ ptr,
                                                   const int32_t* __restrict__ idx, const double* __restrict__ val,
                                                   int64_t nnz, const double* __restrict__ u, int rot,
                                                   const double* __restrict__ R, double* __restrict__ E) {
  const int lane = threadIdx.x & (G - 1);
  const int64_t stride = (int64_t)gridDim.x * (kBlock / G);
  const double* Fx = val;
  const double* Fy = val + nnz;
  const double* Fz = val + 2 * nnz;
  for (int64_t row = (int64_t)blockIdx.x * (kBlock / G) + threadIdx.x / G; row < n; row += stride) {
    double d[3][3] = {};  // d[k][j] = (D_k u_j)_row
    for (int32_t p = ptr[row] + lane; p < ptr[row + 1]; p += G) {
      const int32_t c = idx[p];
      const double uj[3] = {u[c], u[n + c], u[2 * (int64_t)n + c]};
      const double f[3] = {Fx[p], Fy[p], Fz[p]};
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) d[k][j] += f[k] * uj[j];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int j = 0; j < 3; ++j) d[k][j] = group_sum<G>(d[k][j]);
    if (lane == 0) {
      double e[6];
      if (!rot) {
        e[0] = d[0][0];
        e[1] = d[1][1];
        e[2] = d[2][2];
        e[3] = d[1][0] + d[0][1];
        e[4] = d[2][1] + d[1][2];
        e[5] = d[2][0] + d[0][2];
      } else {
        // compute_derivatives.py:94-123; d[k][j]: derivative k of component j
        auto mix = [&](int col, int j) { return R[0 * 3 + col] * d[0][j] + R[1 * 3 + col] * d[1][j] + R[2 * 3 + col] * d[2][j]; };
        e[0] = mix(0, 0);
        e[1] = mix(1, 1);
        e[2] = mix(2, 2);
        e[3] = mix(1, 0) + mix(0, 1);
        e[4] = mix(2, 1) + mix(1, 2);
        e[5] = mix(2, 0) + mix(0, 2);
      }
#pragma unroll
      for (int v = 0; v < 6; ++v) E[6 * row + v] = e[v];
    }
  }
}

// ---- volume averages of E (6) and m@R (3); Ebar from the averaged m --------
__global__ void __launch_bounds__(kBlock) k_strain_avg(int64_t n, const double* __restrict__ W, double vol,
                                                       const double* __restrict__ E, const double* __restrict__ m,
                                                       int rot, const double* __restrict__ R, double l100,
                                                       double l111, const double* __restrict__ Es /*6*/,
                                                       const double* __restrict__ Mv, double* dscal, double* partials,
                                                       unsigned* counter) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  double acc[9] = {};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double w = W[i];
#pragma unroll
    for (int v = 0; v < 6; ++v) acc[v] += w * E[6 * i + v];
    if (rot) {
      double a, b, c;
      rot_apply(R, m[3 * i], m[3 * i + 1], m[3 * i + 2], a, b, c);
      acc[6] += w * a;
      acc[7] += w * b;
      acc[8] += w * c;
    }
  }
  if (grid_sum<9>(acc, partials, counter) && threadIdx.x == 0) {
    for (int v = 0; v < 6; ++v) dscal[S_AVGE + v] = acc[v] / vol;
    double am[3];
    for (int k = 0; k < 3; ++k) {
      if (rot) dscal[S_AVGM100 + k] = acc[6 + k] / vol;
      am[k] = rot ? dscal[S_AVGM100 + k] : dscal[S_AVGM + k];
    }
    // Micromagnetics.py:135-151
    double eb[6];
    eb[0] = 1.5 * l100 * (am[0] * am[0] - 1.0 / 3.0) + Es[0];
    eb[1] = 1.5 * l100 * (am[1] * am[1] - 1.0 / 3.0) + Es[1];
    eb[2] = 1.5 * l100 * (am[2] * am[2] - 1.0 / 3.0) + Es[2];
    eb[3] = 3.0 * l111 * am[0] * am[1] + Es[3];
    eb[4] = 3.0 * l111 * am[1] * am[2] + Es[4];
    eb[5] = 3.0 * l111 * am[0] * am[2] + Es[5];
    if (rot) {
      double o[6];
      for (int r = 0; r < 6; ++r) {
        double s = 0.0;
        for (int j = 0; j < 6; ++j) s += Mv[r * 6 + j] * eb[j];
        o[r] = s;
      }
      for (int r = 0; r < 6; ++r) eb[r] = o[r];
    }
    for (int v = 0; v < 6; ++v) dscal[S_EBAR + v] = eb[v];
  }
}

// E = (E - avg(E)) + Ebar, then E @ M^-T when rot111 (Micromagnetics.py:132,154-162)
__global__ void __launch_bounds__(kBlock) k_efinal(int64_t n, double* __restrict__ E, int rot,
                                                   const double* __restrict__ MvinvT, const double* __restrict__ dscal) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  double av[6], eb[6];
  for (int v = 0; v < 6; ++v) {
    av[v] = dscal[S_AVGE + v];
    eb[v] = dscal[S_EBAR + v];
  }
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    double e[6];
#pragma unroll
    for (int v = 0; v < 6; ++v) e[v] = (E[6 * i + v] - av[v]) + eb[v];
    if (rot) {
      double o[6];
      for (int c = 0; c < 6; ++c) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += e[k] * MvinvT[k * 6 + c];
        o[c] = s;
      }
      for (int c = 0; c < 6; ++c) e[c] = o[c];
    }
#pragma unroll
    for (int v = 0; v < 6; ++v) E[6 * i + v] = e[v];
  }
}

// ---- demag RHS: b = -(Fx m~x + Fy m~y + Fz m~z), b[0] = 0 ------------------
template <int G>
__global__ void __launch_bounds__(kBlock) k_demag_rhs(int32_t n, const int32_t* __restrict__ ptr,
                                                      const int32_t* __restrict__ idx, const double* __restrict__ val,
                                                      int64_t nnz, const double* __restrict__ m,
                                                      const double* __restrict__ avg, double* __restrict__ b) {
  const int lane = threadIdx.x & (G - 1);
  const int64_t stride = (int64_t)gridDim.x * (kBlock / G);
  const double a0 = avg[0], a1 = avg[1], a2 = avg[2];
  for (int64_t row = (int64_t)blockIdx.x * (kBlock / G) + threadIdx.x / G; row < n; row += stride) {
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int32_t p = ptr[row] + lane; p < ptr[row + 1]; p += G) {
      const int64_t c = idx[p];
      sx += val[p] * (m[3 * c] - a0);
      sy += val[nnz + p] * (m[3 * c + 1] - a1);
      sz += val[2 * nnz + p] * (m[3 * c + 2] - a2);
    }
    sx = group_sum<G>(sx);
    sy = group_sum<G>(sy);
    sz = group_sum<G>(sz);
    if (lane == 0) b[row] = (row == 0) ? 0.0 : -((sx + sy) + sz);
  }
}

// ---- Htilde_k = -F_k U ------------------------------------------------------
template <int G>
__global__ void __launch_bounds__(kBlock) k_hd(int32_t n, const int32_t* __restrict__ ptr,
                                               const int32_t* __restrict__ idx, const double* __restrict__ val,
                                               int64_t nnz, const double* __restrict__ U, double* __restrict__ H) {
  const int lane = threadIdx.x & (G - 1);
  const int64_t stride = (int64_t)gridDim.x * (kBlock / G);
  for (int64_t row = (int64_t)blockIdx.x * (kBlock / G) + threadIdx.x / G; row < n; row += stride) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int32_t p = ptr[row] + lane; p < ptr[row + 1]; p += G) {
      const double uc = U[idx[p]];
      s0 += val[p] * uc;
      s1 += val[nnz + p] * uc;
      s2 += val[2 * nnz + p] * uc;
    }
    s0 = group_sum<G>(s0);
    s1 = group_sum<G>(s1);
    s2 = group_sum<G>(s2);
    if (lane == 0) {
      H[row] = -s0;
      H[n + row] = -s1;
      H[2 * (int64_t)n + row] = -s2;
    }
  }
}

// Effective field at one node (Micromagnetics.py:181-211 / :246-274), field
// terms included, in the reference's left-to-right order.
__device__ inline void field_at(const FieldPrm& P, double m1, double m2, double m3, const double* e /*6*/,
                                const double Ht[3], const double Hb[3], double f[3]) {
  if (!P.rot) {
    const double s1 = m1 * m1, s2 = m2 * m2, s3 = m3 * m3;
    f[0] = -2.0 * P.K1 * m1 * (s2 + s3) + 0.5 * Hb[0] + 0.5 * Ht[0] + P.H[0] - 2.0 * P.c1 * m1 * (s2 + s3) -
           2.0 * P.c2 * (e[0] - P.l100) * m1 - P.c3 * (e[3] * m2 + e[5] * m3);
    f[1] = -2.0 * P.K1 * m2 * (s3 + s1) + 0.5 * Hb[1] + 0.5 * Ht[1] + P.H[1] - 2.0 * P.c1 * m2 * (s3 + s1) -
           2.0 * P.c2 * (e[1] - P.l100) * m2 - P.c3 * (e[4] * m3 + e[3] * m1);
    f[2] = -2.0 * P.K1 * m3 * (s1 + s2) + 0.5 * Hb[2] + 0.5 * Ht[2] + P.H[2] - 2.0 * P.c1 * m3 * (s1 + s2) -
           2.0 * P.c2 * (e[2] - P.l100) * m3 - P.c3 * (e[5] * m1 + e[4] * m2);
  } else {
    double a1, a2, a3;
    rot_apply(P.R, m1, m2, m3, a1, a2, a3);
    const double s1 = a1 * a1, s2 = a2 * a2, s3 = a3 * a3;
    const double v1 = -2.0 * P.K1 * a1 * (s2 + s3) - 2.0 * P.c1 * a1 * (s2 + s3) -
                      2.0 * P.c2 * (e[0] - P.l100) * a1 - P.c3 * (e[3] * a2 + e[5] * a3);
    const double v2 = -2.0 * P.K1 * a2 * (s3 + s1) - 2.0 * P.c1 * a2 * (s3 + s1) -
                      2.0 * P.c2 * (e[1] - P.l100) * a2 - P.c3 * (e[4] * a3 + e[3] * a1);
    const double v3 = -2.0 * P.K1 * a3 * (s1 + s2) - 2.0 * P.c1 * a3 * (s1 + s2) -
                      2.0 * P.c2 * (e[2] - P.l100) * a3 - P.c3 * (e[5] * a1 + e[4] * a2);
    // v @ R^T: out_k = sum_j v_j R[k][j]
    const double* R = P.R;
    const double o1 = v1 * R[0] + v2 * R[1] + v3 * R[2];
    const double o2 = v1 * R[3] + v2 * R[4] + v3 * R[5];
    const double o3 = v1 * R[6] + v2 * R[7] + v3 * R[8];
    f[0] = o1 + 0.5 * Hb[0] + 0.5 * Ht[0] + P.H[0];
    f[1] = o2 + 0.5 * Hb[1] + 0.5 * Ht[1] + P.H[1];
    f[2] = o3 + 0.5 * Hb[2] + 0.5 * Ht[2] + P.H[2];
  }
}

__device__ inline void hbar_of(const FieldPrm& P, const double* avg, double Hb[3]) {
  Hb[0] = -P.Nd[0] * avg[0];
  Hb[1] = -P.Nd[1] * avg[1];
  Hb[2] = -P.Nd[2] * avg[2];
}

// stage 1: f (3,n) and ftmp = f*dt + m (3,n)
__global__ void __launch_bounds__(kBlock) k_field1(int64_t n, FieldPrm P0, const double* __restrict__ m,
                                                   const double* __restrict__ E, int me, const double* __restrict__ Ht,
                                                   const double* __restrict__ avg, double* __restrict__ f,
                                                   double* __restrict__ ftmp) {
  FieldPrm P = P0;
  if (P0.Hdev)
    for (int k = 0; k < 3; ++k) P.H[k] = P0.Hdev[k];
  double Hb[3];
  hbar_of(P, avg, Hb);
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    double e[6] = {0, 0, 0, 0, 0, 0};
    if (me)
      for (int v = 0; v < 6; ++v) e[v] = E[6 * i + v];
    const double ht[3] = {Ht[i], Ht[n + i], Ht[2 * n + i]};
    const double m1 = m[3 * i], m2 = m[3 * i + 1], m3 = m[3 * i + 2];
    double fv[3];
    field_at(P, m1, m2, m3, e, ht, Hb, fv);
    f[i] = fv[0];
    f[n + i] = fv[1];
    f[2 * n + i] = fv[2];
    ftmp[i] = fv[0] * P.dt + m1;
    ftmp[n + i] = fv[1] * P.dt + m2;
    ftmp[2 * n + i] = fv[2] * P.dt + m3;
  }
}

// m1* = m1 + g2n m3 - g3n m2 ; rhs = f1*dt + m1*
__global__ void __launch_bounds__(kBlock) k_gs1(int64_t n, double dt, const double* __restrict__ m,
                                                const double* __restrict__ gn, const double* __restrict__ f,
                                                double* __restrict__ mstar, double* __restrict__ rhs) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double m1s = m[3 * i] + gn[n + i] * m[3 * i + 2] - gn[2 * n + i] * m[3 * i + 1];
    mstar[i] = m1s;
    rhs[i] = f[i] * dt + m1s;
  }
}

// m2* = m2 + g3n m1* - g1* m3 ; rhs = f2*dt + m2*
__global__ void __launch_bounds__(kBlock) k_gs2(int64_t n, double dt, const double* __restrict__ m,
                                                const double* __restrict__ gn, const double* __restrict__ g1s,
                                                const double* __restrict__ f, double* __restrict__ mstar,
                                                double* __restrict__ rhs) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double m2s = m[3 * i + 1] + gn[2 * n + i] * mstar[i] - g1s[i] * m[3 * i + 2];
    mstar[n + i] = m2s;
    rhs[i] = f[n + i] * dt + m2s;
  }
}

// m3* = m3 + g1* m2* - g2* m1* ; stage-2 RHS f(m*)*dt2 + m* (3,n)
__global__ void __launch_bounds__(kBlock) k_gs3_field2(int64_t n, FieldPrm P0, const double* __restrict__ m,
                                                       const double* __restrict__ gs, double* __restrict__ mstar,
                                                       const double* __restrict__ E, int me,
                                                       const double* __restrict__ Ht, const double* __restrict__ avg,
                                                       double* __restrict__ ftmp) {
  FieldPrm P = P0;
  if (P0.Hdev)
    for (int k = 0; k < 3; ++k) P.H[k] = P0.Hdev[k];
  double Hb[3];
  hbar_of(P, avg, Hb);
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double m1s = mstar[i], m2s = mstar[n + i];
    const double m3s = m[3 * i + 2] + gs[i] * m2s - gs[n + i] * m1s;
    mstar[2 * n + i] = m3s;
    double e[6] = {0, 0, 0, 0, 0, 0};
    if (me)
      for (int v = 0; v < 6; ++v) e[v] = E[6 * i + v];
    const double ht[3] = {Ht[i], Ht[n + i], Ht[2 * n + i]};
    double fv[3];
    field_at(P, m1s, m2s, m3s, e, ht, Hb, fv);
    ftmp[i] = fv[0] * P.dt2 + m1s;
    ftmp[n + i] = fv[1] * P.dt2 + m2s;
    ftmp[2 * n + i] = fv[2] * P.dt2 + m3s;
  }
}

// normalise m**, zero defects, sumsq = sum (m - m_prev)^2, m_prev = m, avg_m
__global__ void __launch_bounds__(kBlock) k_project(int64_t n, const double* __restrict__ mss,
                                                    const uint8_t* __restrict__ defect, const double* __restrict__ W,
                                                    double vol, double* __restrict__ m, double* __restrict__ mprev,
                                                    double* dscal, double* partials, unsigned* counter) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double a = mss[i], b = mss[n + i], c = mss[2 * n + i];
    const double mag = sqrt(a * a + b * b + c * c);
    double o[3];
    if (fabs(mag) < 1e-14 || defect[i]) {
      o[0] = o[1] = o[2] = 0.0;
    } else {
      o[0] = a / mag;
      o[1] = b / mag;
      o[2] = c / mag;
    }
    const double w = W[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double d = o[k] - mprev[3 * i + k];
      acc[0] += d * d;
      m[3 * i + k] = o[k];
      mprev[3 * i + k] = o[k];
      acc[1 + k] += w * o[k];
    }
  }
  if (grid_sum<4>(acc, partials, counter) && threadIdx.x == 0) {
    dscal[S_SUMSQ] = acc[0];
    for (int k = 0; k < 3; ++k) dscal[S_AVGM + k] = acc[1 + k] / vol;
  }
}

// m[defect] = 0, m_prev = m, avg_m (initialize_m + Micromagnetics.py:72)
__global__ void __launch_bounds__(kBlock) k_setm(int64_t n, const double* __restrict__ src,
                                                 const uint8_t* __restrict__ defect, const double* __restrict__ W,
                                                 double vol, double* __restrict__ m, double* __restrict__ mprev,
                                                 double* dscal, double* partials, unsigned* counter) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double w = W[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double v = defect[i] ? 0.0 : src[3 * i + k];
      m[3 * i + k] = v;
      mprev[3 * i + k] = v;
      acc[k] += w * v;
    }
  }
  if (grid_sum<3>(acc, partials, counter) && threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) dscal[S_AVGM + k] = acc[k] / vol;
}

// out[c] = sum_i W[i] f[i*nc + c]
template <int NC>
__global__ void __launch_bounds__(kBlock) k_wsum(int64_t n, const double* __restrict__ W, const double* __restrict__ f,
                                                 double* out, double* partials, unsigned* counter) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  double acc[NC] = {};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double w = W[i];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] += w * f[i * NC + c];
  }
  if (grid_sum<NC>(acc, partials, counter) && threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = acc[c];
}

#define CPM_G_DISPATCH(L, KERNEL, GRIDROWS, S, ...)                                         \
  switch (L) {                                                                                \
    case 1: KERNEL<1><<<grid_for(GRIDROWS, kBlock / 1), kBlock, 0, S>>>(__VA_ARGS__); break;    \
    case 2: KERNEL<2><<<grid_for(GRIDROWS, kBlock / 2), kBlock, 0, S>>>(__VA_ARGS__); break;    \
    case 4: KERNEL<4><<<grid_for(GRIDROWS, kBlock / 4), kBlock, 0, S>>>(__VA_ARGS__); break;    \
    case 8: KERNEL<8><<<grid_for(GRIDROWS, kBlock / 8), kBlock, 0, S>>>(__VA_ARGS__); break;    \
    case 16: KERNEL<16><<<grid_for(GRIDROWS, kBlock / 16), kBlock, 0, S>>>(__VA_ARGS__); break; \
    case 32: KERNEL<32><<<grid_for(GRIDROWS, kBlock / 32), kBlock, 0, S>>>(__VA_ARGS__); break; \
    default: KERNEL<64><<<grid_for(GRIDROWS, kBlock / 64), kBlock, 0, S>>>(__VA_ARGS__); break; \
  }

static int check_f3(const cpm_mat* F3) {
  if (!F3 || F3->nvals != 3 || F3->nrows != F3->ncols)
    return fail(CPM_EINVAL, "derivative operator must be square with 3 value sets (Fx, Fy, Fz)");
  return CPM_OK;
}

}  // namespace cpm

using namespace cpm;

struct cpm_sim {
  cpm_ctx* ctx = nullptr;
  cpm_sim_ops ops{};
  cpm_sim_params prm{};
  double vol = 1.0;
  int64_t n = 0;
  double* W = nullptr;
  uint8_t* defect = nullptr;
  double* consts = nullptr;  // R(9) Mv(36) MvinvT(36) Ebar_sigma(6)
  double* dscal = nullptr;
  double* hscal = nullptr;  // pinned host mirror of dscal (this simulation's own: set_m and the
                            // step report of simulations sharing a context may run on several threads)
  double *m = nullptr, *mprev = nullptr, *f = nullptr, *ftmp = nullptr, *b = nullptr, *gn = nullptr, *gs = nullptr;
  double *mstar = nullptr, *mss = nullptr, *U = nullptr, *Ht = nullptr, *E = nullptr, *bd = nullptr;
  double *E0 = nullptr, *bel = nullptr, *u = nullptr;
  // fast modes: interleaved right-hand side and warm-start copies of the solutions
  double *bint = nullptr, *uint = nullptr, *gnint = nullptr, *mssint = nullptr;
  std::vector<void*> allocs;
  // CG workspace of the batched solver (allocated when added to a batch)
  double *wr = nullptr, *wp = nullptr, *wq = nullptr, *wy = nullptr;
  float *wr32 = nullptr, *wp32 = nullptr, *wq32 = nullptr, *wx32 = nullptr;
  cpm_batch* batch = nullptr;
  int hint[10] = {};
  bool stepped = false;  // the warm-start buffers hold a previous solution
  // warm_start 2: the solutions of the solve before the last, per solve
  // (12 n: elasticity 3n, demag n, Gauss-Seidel 3n + n + n, m** 3n), and the
  // solutions each solve has kept so far (0, 1, 2 or more)
  double* xhist = nullptr;
  int nwarm[10] = {};
  // hysteresis control (cpm_sim_control): on, host copy of the next step's use_init, field list
  int ctl_on = 0;
  int ctl_use_init = 0;
  double* ctl_fields = nullptr;
};

static int sim_alloc(cpm_sim* S, double** p, int64_t count) {
  hipError_t e = hipMalloc(p, sizeof(double) * std::max<int64_t>(count, 1));
  if (e != hipSuccess) return fail(e == hipErrorOutOfMemory ? CPM_ENOMEM : CPM_EHIP, hipGetErrorString(e));
  S->allocs.push_back(*p);
  e = hipMemsetAsync(*p, 0, sizeof(double) * std::max<int64_t>(count, 1), S->ctx->stream);
  if (e != hipSuccess) return fail(CPM_EHIP, hipGetErrorString(e));
  return CPM_OK;
}

extern "C" int cpm_sim_destroy(cpm_sim* S) {
  if (!S) return CPM_OK;
  if (S->batch) (void)cpm_batch_remove(S->batch, S);
  if (S->ctx) {
    (void)hipSetDevice(S->ctx->device);
    (void)hipStreamSynchronize(S->ctx->stream);
  }
  for (void* p : S->allocs) (void)hipFree(p);
  if (S->hscal) (void)hipHostFree(S->hscal);
  delete S;
  return CPM_OK;
}

extern "C" int cpm_sim_create(cpm_ctx* ctx, const cpm_sim_ops* ops, const cpm_sim_params* prm, const double* W,
                              double volume, const uint8_t* defect, cpm_sim** out) {
  if (!ctx || !ops || !prm || !W || !out) return fail(CPM_EINVAL, "cpm_sim_create: null argument");
  const int64_t n = prm->n;
  if (n <= 0) return fail(CPM_EINVAL, "cpm_sim_create: n must be positive");
  if (prm->mode != CPM_MODE_FP64 && prm->mode != CPM_MODE_FAST64 && prm->mode != CPM_MODE_MIXED)
    return fail(CPM_EINVAL, "cpm_sim_create: unknown mode");
  if (prm->warm_start < 0 || prm->warm_start > 2)
    return fail(CPM_EINVAL, "cpm_sim_create: warm_start must be 0, 1 or 2");
  if (prm->mode != CPM_MODE_FP64) {
    for (const cpm_mat* A : {ops->A_demag, ops->A1, ops->A2})
      if (!A->sell.sptr || A->sell.bs != 1)
        return fail(CPM_EINVAL, "cpm_sim_create: fast modes need SELL forms (bs 1) of A_demag, A1, A2");
    if (prm->me && (!ops->A_el->sell.sptr || ops->A_el->sell.bs != 3))
      return fail(CPM_EINVAL, "cpm_sim_create: fast modes need the bs-3 SELL form of A_el");
    if (prm->mode == CPM_MODE_MIXED && !(prm->rel_delta > 0.0 && prm->rel_delta < 1.0))
      return fail(CPM_EINVAL, "cpm_sim_create: rel_delta must be in (0, 1)");
    if (prm->mode == CPM_MODE_MIXED && !(prm->rel_delta_gs >= 0.0 && prm->rel_delta_gs < 1.0))
      return fail(CPM_EINVAL, "cpm_sim_create: rel_delta_gs must be in [0, 1)");
    if (prm->mode == CPM_MODE_MIXED && !(prm->rel_delta_mg >= 0.0 && prm->rel_delta_mg < 1.0))
      return fail(CPM_EINVAL, "cpm_sim_create: rel_delta_mg must be in [0, 1)");
    if (prm->precond & ~(CPM_PRECOND_EL_DEMAG | CPM_PRECOND_GS))
      return fail(CPM_EINVAL, "cpm_sim_create: unknown precond bits");
    const bool pe = (prm->precond & CPM_PRECOND_EL_DEMAG) != 0, pg = (prm->precond & CPM_PRECOND_GS) != 0;
    if (pe && !ops->mg_demag)
      return fail(CPM_EINVAL, "cpm_sim_create: precond needs the multigrid hierarchy of A_demag");
    if (pe && prm->me && !ops->mg_el)
      return fail(CPM_EINVAL, "cpm_sim_create: precond needs the multigrid hierarchy of A_el");
    if (pg && !ops->mg_a1) return fail(CPM_EINVAL, "cpm_sim_create: precond needs the multigrid hierarchy of A1");
    if ((pe && (mg_operator(ops->mg_demag) != ops->A_demag || (prm->me && mg_operator(ops->mg_el) != ops->A_el))) ||
        (pg && mg_operator(ops->mg_a1) != ops->A1))
      return fail(CPM_EINVAL, "cpm_sim_create: a multigrid hierarchy does not belong to its operator");
  } else if (prm->precond) {
    return fail(CPM_EINVAL, "cpm_sim_create: precond is a fast-mode option");
  }
  auto sq = [&](const cpm_mat* A, int64_t rows) { return A && A->nrows == rows && A->ncols == rows; };
  if (!sq(ops->A_demag, n) || !sq(ops->A1, n) || !sq(ops->A2, n) || !sq(ops->F1, n))
    return fail(CPM_EINVAL, "cpm_sim_create: scalar operators must be n x n");
  if (ops->fder) {
    if (fel_rows(ops->fder) != 3 * n) return fail(CPM_EINVAL, "cpm_sim_create: matrix-free Fx, Fy, Fz of another mesh");
  } else {
    if (int rc = check_f3(ops->F3)) return rc;
    if (ops->F3->nrows != n) return fail(CPM_EINVAL, "cpm_sim_create: F3 must be n x n");
  }
  if (prm->me) {
    if (!sq(ops->A_el, 3 * n)) return fail(CPM_EINVAL, "cpm_sim_create: A_el must be 3n x 3n");
    if (ops->fel) {
      if (fel_rows(ops->fel) != 3 * n) return fail(CPM_EINVAL, "cpm_sim_create: matrix-free F_el of another mesh");
    } else if (!ops->F_el || ops->F_el->nrows != 3 * n || ops->F_el->ncols != 6 * n) {
      return fail(CPM_EINVAL, "cpm_sim_create: F_el must be 3n x 6n");
    }
  }
  CPM_HIP(hipSetDevice(ctx->device));
  auto* S = new cpm_sim();
  S->ctx = ctx;
  S->ops = *ops;
  S->prm = *prm;
  S->vol = volume;
  S->n = n;
  int rc = CPM_OK;
  auto A = [&](double** p, int64_t cnt) {
    if (rc == CPM_OK) rc = sim_alloc(S, p, cnt);
  };
  A(&S->W, n);
  A(&S->consts, 96);
  A(&S->dscal, S_NSCAL);
  if (rc == CPM_OK) {
    hipError_t e = hipHostMalloc(&S->hscal, sizeof(double) * S_NSCAL, hipHostMallocDefault);
    if (e != hipSuccess) rc = fail(CPM_ENOMEM, "pinned scalar buffer");
  }
  A(&S->m, 3 * n);
  A(&S->mprev, 3 * n);
  A(&S->f, 3 * n);
  A(&S->ftmp, 3 * n);
  A(&S->b, 3 * n);
  A(&S->gn, 3 * n);
  A(&S->gs, 2 * n);
  A(&S->mstar, 3 * n);
  A(&S->mss, 3 * n);
  A(&S->U, n);
  A(&S->Ht, 3 * n);
  A(&S->E, 6 * n);
  A(&S->bd, n);
  if (prm->me) {
    A(&S->E0, 6 * n);
    A(&S->bel, 3 * n);
    A(&S->u, 3 * n);
  }
  if (prm->mode != CPM_MODE_FP64) {
    A(&S->bint, 3 * n);
    A(&S->gnint, 3 * n);
    A(&S->mssint, 3 * n);
    if (prm->me) A(&S->uint, 3 * n);
    if (prm->warm_start == 2) A(&S->xhist, 12 * n);
  }
  if (rc == CPM_OK) {
    hipError_t e = hipMalloc(&S->defect, std::max<int64_t>(n, 1));
    if (e != hipSuccess) rc = fail(CPM_ENOMEM, "defect mask allocation");
    else S->allocs.push_back(S->defect);
  }
  if (rc == CPM_OK) rc = ensure_workspace(ctx, prm->me ? 3 * n : n);
  if (rc != CPM_OK) {
    cpm_sim_destroy(S);
    return rc;
  }
  std::vector<uint8_t> dm(n, 0);
  if (defect) std::memcpy(dm.data(), defect, n);
  std::vector<double> cst(96, 0.0);
  std::memcpy(cst.data(), prm->R, 9 * sizeof(double));
  std::memcpy(cst.data() + 9, prm->Mv, 36 * sizeof(double));
  std::memcpy(cst.data() + 45, prm->MvinvT, 36 * sizeof(double));
  std::memcpy(cst.data() + 81, prm->Ebar_sigma, 6 * sizeof(double));
  hipError_t e = hipMemcpyAsync(S->defect, dm.data(), n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(S->consts, cst.data(), 96 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(S->W, W, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    cpm_sim_destroy(S);
    return fail(CPM_EHIP, std::string("cpm_sim_create upload: ") + hipGetErrorString(e));
  }
  *out = S;
  return CPM_OK;
}

extern "C" int cpm_sim_set_m(cpm_sim* S, const double* m, double* avg_m) {
  if (!S || !m) return fail(CPM_EINVAL, "cpm_sim_set_m: null argument");
  cpm_ctx* ctx = S->ctx;
  k_setm<<<grid_for(S->n), kBlock, 0, ctx->stream>>>(S->n, m, S->defect, S->W, S->vol, S->m, S->mprev, S->dscal,
                                                      ctx->partials, ctx->counter);
  std::fill(S->nwarm, S->nwarm + 10, 0);  // a new trajectory: no history to extrapolate
  CPM_HIP(hipGetLastError());
  CPM_HIP(hipMemcpyAsync(S->hscal, S->dscal + S_AVGM, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  CPM_HIP(hipStreamSynchronize(ctx->stream));
  if (avg_m) std::memcpy(avg_m, S->hscal, 3 * sizeof(double));
  return CPM_OK;
}

extern "C" int cpm_sim_control(cpm_sim* S, const cpm_sweep_cfg* cfg, const double* fields) {
  if (!S) return fail(CPM_EINVAL, "cpm_sim_control: null simulation");
  cpm_ctx* ctx = S->ctx;
  CPM_HIP(hipSetDevice(ctx->device));
  CPM_HIP(hipStreamSynchronize(ctx->stream));
  double c[S_NSCAL - C_ON];
  std::memset(c, 0, sizeof(c));
  if (!cfg) {  // detach: the caller's H and use_init drive the steps again
    CPM_HIP(hipMemcpy(S->dscal + C_ON, c, sizeof(c), hipMemcpyHostToDevice));
    S->ctl_on = 0;
    S->ctl_use_init = 0;
    return CPM_OK;
  }
  if (cfg->npoints < 0 || (cfg->npoints > 0 && !fields))
    return fail(CPM_EINVAL, "cpm_sim_control: npoints must be >= 0, with `fields` (npoints x 3) when > 0");
  if (cfg->npoints > 0) {
    if (S->ctl_fields) {
      (void)hipFree(S->ctl_fields);
      S->allocs.erase(std::find(S->allocs.begin(), S->allocs.end(), (void*)S->ctl_fields));
      S->ctl_fields = nullptr;
    }
    if (int rc = sim_alloc(S, &S->ctl_fields, 3 * (int64_t)cfg->npoints)) return rc;
    CPM_HIP(hipMemcpy(S->ctl_fields, fields, sizeof(double) * 3 * cfg->npoints, hipMemcpyHostToDevice));
  }
  auto at = [&](int slot) -> double& { return c[slot - C_ON]; };
  at(C_ON) = 1.0;
  for (int k = 0; k < 3; ++k) {
    at(C_H + k) = cfg->npoints > 0 ? fields[k] : cfg->H0[k];
    at(C_DH + k) = cfg->dH[k];
    at(C_STOPDIR + k) = cfg->stop_dir[k];
    at(C_PH + k) = at(C_H + k);
  }
  at(C_ACC) = cfg->llg_accuracy;
  at(C_INITF) = cfg->initial_factor;
  at(C_STOPVAL) = cfg->stop_value;
  at(C_MINSTEP) = cfg->llg_min_step;
  at(C_SPP) = cfg->steps_per_point;
  at(C_NPTS) = cfg->npoints;
  CPM_HIP(hipMemcpy(S->dscal + C_ON, c, sizeof(c), hipMemcpyHostToDevice));
  S->ctl_on = 1;
  S->ctl_use_init = 0;
  return CPM_OK;
}

extern "C" int cpm_sim_get(cpm_sim* S, int field, double* dst) {
  if (!S || !dst) return fail(CPM_EINVAL, "cpm_sim_get: null argument");
  const double* src = nullptr;
  int64_t cnt = 0;
  switch (field) {
    case CPM_FIELD_M: src = S->m, cnt = 3 * S->n; break;
    case CPM_FIELD_HTILDE: src = S->Ht, cnt = 3 * S->n; break;
    case CPM_FIELD_E: src = S->E, cnt = 6 * S->n; break;
    case CPM_FIELD_U: src = S->U, cnt = S->n; break;
    case CPM_FIELD_UEL:
      if (!S->u) return fail(CPM_EINVAL, "cpm_sim_get: no displacement without magnetoelastic coupling");
      src = S->u, cnt = 3 * S->n;
      break;
    default: return fail(CPM_EINVAL, "cpm_sim_get: unknown field");
  }
  CPM_HIP(hipMemcpyAsync(dst, src, cnt * sizeof(double), hipMemcpyDeviceToDevice, S->ctx->stream));
  CPM_HIP(hipStreamSynchronize(S->ctx->stream));
  return CPM_OK;
}

// ---- the step as stages -------------------------------------------------------
//
// A step is six stages, each a few field kernels followed by one solve
// (stage 0, the elasticity solve, only with magnetoelastic coupling), then
// the projection.  cpm_sim_step runs them for one simulation; cpm_batch_step
// runs every stage for all simulations of a batch and their solves as one
// batched CG.

namespace cpm {

struct SolveSpec {
  int slot = 0;                // first report slot of the solve
  int K = 0;                   // right-hand sides sharing the operator
  const cpm_mat* A = nullptr;  // nullptr: the stage has no solve
  const double* b = nullptr;   // right-hand sides in the reference layout (K blocks of n, or 3n)
  double* x = nullptr;         // solutions, same layout
  // fast modes: interleaved copies (bs 3 or K = 3), nullptr when b / x are used directly
  double* bint = nullptr;
  double* xint = nullptr;
};

// Hysteresis control after a step (one thread, sweep_ctl.h).
__global__ void k_sweep_ctl(double* __restrict__ c, const double* __restrict__ fields) {
  sweep_ctl_update(c, fields);
}

static FieldPrm field_prm(const cpm_sim* S, const double* H) {
  const cpm_sim_params& p = S->prm;
  FieldPrm P{};
  P.K1 = p.K1;
  P.c1 = p.elasC[0];
  P.c2 = p.elasC[1];
  P.c3 = p.elasC[2];
  P.l100 = p.lambda100;
  P.dt = p.dt;
  P.dt2 = p.dt * p.alpha_damp;
  for (int k = 0; k < 3; ++k) {
    P.Nd[k] = p.Nd[k];
    P.H[k] = H ? H[k] : 0.0;
  }
  std::memcpy(P.R, p.R, sizeof(P.R));
  P.rot = p.rot111;
  P.Hdev = S->ctl_on ? S->dscal + C_H : nullptr;
  return P;
}

static inline bool fast_mode(const cpm_sim* S) { return S->prm.mode != CPM_MODE_FP64; }

// Field kernels of stage `stage` (0..5) and the description of its solve.
// ovl (overlap_demag): the demag right-hand side is formed in stage 0,
// where the demag solve starts beside the elasticity solve, and stage 1
// forms the strain only
static int stage_pre(cpm_sim* S, int stage, const FieldPrm& P, SolveSpec& sp, bool ovl = false) {
  cpm_ctx* ctx = S->ctx;
  hipStream_t s = ctx->stream;
  const cpm_sim_params& p = S->prm;
  const int64_t n = S->n;
  const int gv = grid_for(n);
  const double* dR = S->consts;
  const double* dMv = S->consts + 9;
  const double* dMvinvT = S->consts + 45;
  const double* dEs = S->consts + 81;
  const cpm_mat* F3 = S->ops.F3;
  const bool fast = fast_mode(S);
  sp = SolveSpec{};
  int rc = CPM_OK;
  switch (stage) {
    case 0:  // Micromagnetics.py:112-125
      if (!p.me) return CPM_OK;
      k_e0<<<gv, kBlock, 0, s>>>(n, S->m, p.lambda100, p.lambda111, p.rot111, dR, dMv, S->E0);
      if (S->ops.fel) {
        if ((rc = fel_apply(S->ops.fel, S->E0, S->bel))) return rc;
      } else if ((rc = launch_spmv(ctx, S->ops.F_el, 0, S->E0, S->bel))) {
        return rc;
      }
      if (ovl) {  // :166 (b_demag depends on m only)
        if (S->ops.fder) {
          if ((rc = fder_strain_demag(S->ops.fder, nullptr, p.rot111, dR, S->m, S->dscal + S_AVGM, nullptr, S->bd)))
            return rc;
        } else {
          CPM_G_DISPATCH(F3->lanes, k_demag_rhs, n, s, (int32_t)n, F3->ptr, F3->idx, F3->val, F3->nnz, S->m,
                         S->dscal + S_AVGM, S->bd);
        }
      }
      sp = SolveSpec{0, 1, S->ops.A_el, S->bel, S->u, fast ? S->bint : nullptr, fast ? S->uint : nullptr};
      break;
    case 1:  // :127-169
      if (S->ops.fder) {
        // strain (when coupled) and the demag right-hand side in one walk
        if ((rc = fder_strain_demag(S->ops.fder, p.me ? S->u : nullptr, p.rot111, dR, ovl ? nullptr : S->m,
                                    S->dscal + S_AVGM, S->E, ovl ? nullptr : S->bd)))
          return rc;
      } else if (p.me) {
        CPM_G_DISPATCH(F3->lanes, k_strain, n, s, (int32_t)n, F3->ptr, F3->idx, F3->val, F3->nnz, S->u, p.rot111,
                       dR, S->E);
      }
      if (p.me) {
        k_strain_avg<<<gv, kBlock, 0, s>>>(n, S->W, S->vol, S->E, S->m, p.rot111, dR, p.lambda100, p.lambda111,
                                           dEs, dMv, S->dscal, ctx->partials, ctx->counter);
        k_efinal<<<gv, kBlock, 0, s>>>(n, S->E, p.rot111, dMvinvT, S->dscal);
      }
      if (ovl) break;  // the demag solve ran in stage 0
      if (!S->ops.fder)
        CPM_G_DISPATCH(F3->lanes, k_demag_rhs, n, s, (int32_t)n, F3->ptr, F3->idx, F3->val, F3->nnz, S->m,
                       S->dscal + S_AVGM, S->bd);
      sp = SolveSpec{1, 1, S->ops.A_demag, S->bd, S->U};
      break;
    case 2:  // :171-223
      if (S->ops.fder) {
        if ((rc = fder_hd(S->ops.fder, S->U, S->Ht))) return rc;
      } else {
        CPM_G_DISPATCH(F3->lanes, k_hd, n, s, (int32_t)n, F3->ptr, F3->idx, F3->val, F3->nnz, S->U, S->Ht);
      }
      k_field1<<<gv, kBlock, 0, s>>>(n, P, S->m, S->E, p.me, S->Ht, S->dscal + S_AVGM, S->f, S->ftmp);
      for (int c = 0; c < 3; ++c)
        if ((rc = launch_spmv(ctx, S->ops.F1, 0, S->ftmp + c * n, S->b + c * n))) return rc;
      sp = SolveSpec{2, 3, S->ops.A1, S->b, S->gn, fast ? S->bint : nullptr, fast ? S->gnint : nullptr};
      break;
    case 3:  // :225-229
      k_gs1<<<gv, kBlock, 0, s>>>(n, p.dt, S->m, S->gn, S->f, S->mstar, S->ftmp);
      if ((rc = launch_spmv(ctx, S->ops.F1, 0, S->ftmp, S->b))) return rc;
      sp = SolveSpec{5, 1, S->ops.A1, S->b, S->gs};
      break;
    case 4:  // :231-235
      k_gs2<<<gv, kBlock, 0, s>>>(n, p.dt, S->m, S->gn, S->gs, S->f, S->mstar, S->ftmp);
      if ((rc = launch_spmv(ctx, S->ops.F1, 0, S->ftmp, S->b))) return rc;
      sp = SolveSpec{6, 1, S->ops.A1, S->b, S->gs + n};
      break;
    case 5:  // :237-290
      k_gs3_field2<<<gv, kBlock, 0, s>>>(n, P, S->m, S->gs, S->mstar, S->E, p.me, S->Ht, S->dscal + S_AVGM,
                                         S->ftmp);
      for (int c = 0; c < 3; ++c)
        if ((rc = launch_spmv(ctx, S->ops.F1, 0, S->ftmp + c * n, S->b + c * n))) return rc;
      sp = SolveSpec{7, 3, S->ops.A2, S->b, S->mss, fast ? S->bint : nullptr, fast ? S->mssint : nullptr};
      break;
    default: return fail(CPM_EINVAL, "stage out of range");
  }
  CPM_HIP(hipGetLastError());
  // fast modes: K right-hand sides (K = 3, or the 3 components of one bs-3
  // solve) are gathered into one interleaved vector for the SELL solvers
  if (sp.bint) return interleave(ctx, n, 3, sp.b, sp.bint, 0);
  return CPM_OK;
}

// After the solve of a stage: back to the reference layout.  The
// interleaved solution stays in place as the next warm start.
static int stage_post(cpm_sim* S, const SolveSpec& sp) {
  if (sp.A && sp.xint) return interleave(S->ctx, S->n, 3, sp.xint, sp.x, 1);
  return CPM_OK;
}

// :292-307 normalise / project, then the step scalars to `dst` (pinned).
static int stage_final(cpm_sim* S, double* dst) {
  cpm_ctx* ctx = S->ctx;
  k_project<<<grid_for(S->n), kBlock, 0, ctx->stream>>>(S->n, S->mss, S->defect, S->W, S->vol, S->m, S->mprev,
                                                        S->dscal, ctx->partials, ctx->counter);
  if (S->ctl_on) k_sweep_ctl<<<1, 1, 0, ctx->stream>>>(S->dscal, S->ctl_fields);
  CPM_HIP(hipGetLastError());
  CPM_HIP(hipMemcpyAsync(dst, S->dscal, S_NSCAL * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  return CPM_OK;
}

static void report_scalars(cpm_sim* S, const double* scal, cpm_step_report* rep) {
  S->stepped = true;
  rep->sumsq = scal[S_SUMSQ];
  for (int k = 0; k < 3; ++k) rep->avg_m[k] = scal[S_AVGM + k];
  if (S->ctl_on) {
    S->ctl_use_init = scal[C_USEINIT] != 0.0 ? 1 : 0;
    rep->use_init = S->ctl_use_init;
    rep->point_done = scal[C_PDONE] != 0.0 ? 1 : 0;
    rep->state = (int32_t)scal[C_STATE];
    rep->point = (int32_t)scal[C_POINT];
    rep->nstep = (int32_t)scal[C_NSTEP];
    for (int k = 0; k < 3; ++k) {
      rep->H[k] = scal[C_PH + k];
      rep->H_next[k] = scal[C_H + k];
    }
  }
}

// the use_init flag of a step: the control's when one is attached
static inline int step_use_init(const cpm_sim* S, int given) { return S->ctl_on ? S->ctl_use_init : given; }

// reliable-update factor of a solve: slots 0 and 1 are the elasticity and
// demag solves, slots 2..9 the Gauss-Seidel solves
static inline double solve_delta(const cpm_sim_params& p, int slot) {
  return (slot >= 2 && p.rel_delta_gs > 0.0) ? p.rel_delta_gs : p.rel_delta;
}

static inline int use_x0_fast(const cpm_sim* S, int use_init) {
  return (use_init || (S->prm.warm_start && S->stepped)) ? 1 : 0;
}

// x = 2 x - xp and xp = x (the old x): the solutions of the last two steps
// extrapolated linearly in time; first: xp = x only (one solution so far).
__global__ void __launch_bounds__(kBlock) k_extrap(int64_t n, double* __restrict__ x, double* __restrict__ xp,
                                                   int first) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double xi = x[i];
    if (!first) x[i] = 2.0 * xi - xp[i];
    xp[i] = xi;
  }
}

// warm_start 2: the initial guess of a fast-mode solve, formed in place in
// its solution vector (which holds the last step's solution): the linear
// extrapolation of the last two solutions once two are kept, the last one
// before that.
static void warm_guess(cpm_sim* S, const SolveSpec& sp) {
  if (S->prm.warm_start != 2 || !S->xhist || !sp.A) return;
  const int64_t n = S->n;
  const int64_t len = (sp.K == 3 || sp.A == S->ops.A_el) ? 3 * n : n;
  // offset of a solve's history (first report slot -> multiple of n)
  static constexpr int64_t kOff[8] = {0, 3, 4, -1, -1, 7, 8, 9};
  double* xx = sp.xint ? sp.xint : sp.x;
  int& h = S->nwarm[sp.slot];
  if (h >= 1)
    k_extrap<<<grid_for(len), kBlock, 0, S->ctx->stream>>>(len, xx, S->xhist + kOff[sp.slot] * n, h == 1 ? 1 : 0);
  h = std::min(h + 1, 2);
}

// The multigrid hierarchy a solve runs on, or null (plain CG): elasticity
// and demag (CPM_PRECOND_EL_DEMAG), the Gauss-Seidel solves on A1
// (CPM_PRECOND_GS).
static cpm_mg* solve_mg(const cpm_sim* S, const SolveSpec& sp) {
  const int pc = S->prm.precond;
  if ((pc & CPM_PRECOND_EL_DEMAG) && sp.slot == 0) return S->ops.mg_el;
  if ((pc & CPM_PRECOND_EL_DEMAG) && sp.slot == 1) return S->ops.mg_demag;
  if ((pc & CPM_PRECOND_GS) && sp.A == S->ops.A1) return S->ops.mg_a1;
  return nullptr;
}

// One solve of a single simulation: CSR CG per column (plain FP64 path, the
// reference's sequence) or the SELL solvers (fast modes).
static int solve_one(cpm_sim* S, const SolveSpec& sp, int use_init, cpm_step_report* rep,
                     const PcgLane* ln = nullptr) {
  cpm_ctx* ctx = S->ctx;
  const cpm_sim_params& p = S->prm;
  const int64_t n = S->n;
  if (!fast_mode(S)) {
    const int cols = sp.K;  // K = 3: three scalar solves on one operator
    const int64_t len = sp.A->nrows;
    for (int c = 0; c < cols; ++c) {
      int32_t it = 0;
      double res = 0.0;
      const int slot = sp.slot + c;
      int rc = run_cg(ctx, sp.A, sp.b + c * len, sp.x + c * len, use_init, p.tol, p.maxiter, S->hint[slot], &it,
                      &res);
      rep->iters[slot] = it;
      S->hint[slot] = it;
      if (rc > 0) {
        rep->failed = slot + 1;
        rep->resid = res;
      }
      if (rc) return rc;
    }
    return CPM_OK;
  }
  cpm_mat* Am = const_cast<cpm_mat*>(sp.A);
  const double* bb = sp.bint ? sp.bint : sp.b;
  double* xx = sp.xint ? sp.xint : sp.x;
  int32_t it[3] = {0, 0, 0};
  double res[3] = {0, 0, 0};
  int failed = 0;
  int hint = 0;
  for (int j = 0; j < sp.K; ++j) hint = std::max(hint, S->hint[sp.slot + j]);
  const int x0 = use_x0_fast(S, use_init);
  cpm_mg* mg = solve_mg(S, sp);
  const double dmg = (p.mode == CPM_MODE_MIXED) ? p.rel_delta_mg : 0.0;
  int rc = mg ? (ln ? mg_pcg_lane(mg, *ln, sp.K, bb, xx, x0, p.tol, p.maxiter, dmg, hint, it, res, &failed)
                    : mg_pcg(mg, sp.K, bb, xx, x0, p.tol, p.maxiter, dmg, hint, it, res, &failed))
           : (p.mode == CPM_MODE_MIXED)
               ? sell_cg_mixed(ctx, Am, sp.K, bb, xx, x0, p.tol, p.maxiter, solve_delta(p, sp.slot), hint, it, res,
                               &failed,
                               nullptr)
               : sell_cg64(ctx, Am, sp.K, bb, xx, x0, p.tol, p.maxiter, hint, it, res, &failed);
  if (rc) return rc;
  (void)n;
  for (int j = 0; j < sp.K; ++j) {
    rep->iters[sp.slot + j] = it[j];
    S->hint[sp.slot + j] = it[j];
  }
  if (failed) {
    int jf = 0;
    for (int j = 1; j < sp.K; ++j)
      if (it[j] > it[jf]) jf = j;
    rep->failed = sp.slot + jf + 1;
    rep->resid = res[jf];
    return std::max(1, (int)it[jf]);
  }
  return CPM_OK;
}

}  // namespace cpm

// The demag solve runs beside the elasticity solve: the two are
// independent (Micromagnetics.py:112-169; the demag right-hand side depends
// on m only), so with multigrid on both the demag PCG takes the context's
// second lane, from a host thread of its own, while this thread drives the
// elasticity PCG on the first; the V-cycles of the scalar operator fill the
// gaps of the elasticity solve's launches.  Each solve's arithmetic is the
// same as when they run one after the other.
static bool overlap_demag(const cpm_sim* S) {
  const cpm_sim_params& p = S->prm;
  return p.overlap && fast_mode(S) && p.me && (p.precond & CPM_PRECOND_EL_DEMAG) && S->ops.mg_el &&
         S->ops.mg_demag;
}

static int solve_el_and_demag(cpm_sim* S, const SolveSpec& se, int use_init, cpm_step_report* rep) {
  cpm_ctx* ctx = S->ctx;
  PcgLane l2;
  if (int rc = lane2(ctx, &l2)) return rc;
  const SolveSpec sd{1, 1, S->ops.A_demag, S->bd, S->U};
  warm_guess(S, sd);
  CPM_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
  int drc = CPM_OK;
  std::string derr;
  cpm_step_report drep{};
  std::thread demag([&] {
    if (hipSetDevice(ctx->device) != hipSuccess || hipStreamWaitEvent(ctx->stream2, ctx->ev_fork, 0) != hipSuccess) {
      drc = fail(CPM_EHIP, "demag solve: second lane not ready");
      derr = cpm_last_error();
      return;
    }
    drc = solve_one(S, sd, use_init, &drep, &l2);
    if (drc < 0) derr = cpm_last_error();
    (void)hipEventRecord(ctx->ev_join, ctx->stream2);
  });
  PcgLane l1 = lane1(ctx);
  l1.shared = 1;
  const int rc = solve_one(S, se, use_init, rep, &l1);
  demag.join();
  CPM_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
  rep->iters[sd.slot] = drep.iters[sd.slot];
  if (rc) return rc;  // as one after the other: the elasticity solve's failure comes first
  if (drc < 0) return fail(drc, derr);
  if (drc > 0) {
    rep->failed = drep.failed;
    rep->resid = drep.resid;
  }
  return drc;
}

extern "C" int cpm_sim_step(cpm_sim* S, const double* H, int use_init, cpm_step_report* rep) {
  if (!S || !rep || (!H && !S->ctl_on)) return fail(CPM_EINVAL, "cpm_sim_step: null argument");
  if (S->batch) return fail(CPM_EINVAL, "cpm_sim_step: the simulation belongs to a batch (cpm_batch_step)");
  cpm_ctx* ctx = S->ctx;
  std::memset(rep, 0, sizeof(*rep));
  use_init = step_use_init(S, use_init);
  const FieldPrm P = field_prm(S, S->ctl_on ? nullptr : H);
  const bool ovl = overlap_demag(S);
  for (int stage = 0; stage < 6; ++stage) {
    SolveSpec sp;
    if (int rc = stage_pre(S, stage, P, sp, ovl)) return rc;
    if (!sp.A) continue;
    warm_guess(S, sp);
    if (int rc = (ovl && stage == 0) ? solve_el_and_demag(S, sp, use_init, rep) : solve_one(S, sp, use_init, rep))
      return rc;
    if (int rc = stage_post(S, sp)) return rc;
  }
  if (int rc = stage_final(S, S->hscal)) return rc;
  CPM_HIP(hipStreamSynchronize(ctx->stream));
  report_scalars(S, S->hscal, rep);
  return CPM_OK;
}

// ---- batches of simulations --------------------------------------------------

struct cpm_batch {
  cpm_ctx* ctx = nullptr;
  int flags = 0;
  int mode = -1;
  std::vector<cpm_sim*> sims;
  cpm::MultiSolver ms[6];  // one per stage
  // multigrid-preconditioned members of each stage, per (precision, degree)
  std::map<int, cpm::MultiMgSolver> mgs[6];
  double* scal = nullptr;  // pinned, S_NSCAL per simulation
  int cap = 0;
};

static int batch_workspace(cpm_sim* S) {
  // per-simulation CG workspace of the batched solver: the longest system
  // of a step has 3n entries (elasticity, or the 3-RHS solves)
  if (S->wr) return CPM_OK;
  const int64_t len = 3 * S->n;
  int rc = CPM_OK;
  auto D = [&](double** p) {
    if (rc == CPM_OK) rc = sim_alloc(S, p, len);
  };
  auto F = [&](float** p) {
    if (rc != CPM_OK) return;
    hipError_t e = hipMalloc(p, sizeof(float) * len);
    if (e != hipSuccess) {
      rc = fail(e == hipErrorOutOfMemory ? CPM_ENOMEM : CPM_EHIP, hipGetErrorString(e));
      return;
    }
    S->allocs.push_back(*p);
  };
  D(&S->wr);
  if (S->prm.mode == CPM_MODE_MIXED) {
    D(&S->wy);
    F(&S->wr32);
    F(&S->wp32);
    F(&S->wq32);
    F(&S->wx32);
  } else {
    D(&S->wp);
    D(&S->wq);
  }
  return rc;
}

extern "C" int cpm_batch_create(cpm_ctx* ctx, int flags, cpm_batch** out) {
  if (!ctx || !out) return fail(CPM_EINVAL, "cpm_batch_create: null argument");
  if (flags & ~CPM_BATCH_GRAPH) return fail(CPM_EINVAL, "cpm_batch_create: unknown flags");
  auto* B = new cpm_batch();
  B->ctx = ctx;
  B->flags = flags;
  *out = B;
  return CPM_OK;
}

extern "C" int cpm_batch_destroy(cpm_batch* B) {
  if (!B) return CPM_OK;
  (void)hipSetDevice(B->ctx->device);
  (void)hipStreamSynchronize(B->ctx->stream);
  for (cpm_sim* S : B->sims) S->batch = nullptr;
  if (B->scal) (void)hipHostFree(B->scal);
  delete B;
  return CPM_OK;
}

extern "C" int cpm_batch_add(cpm_batch* B, cpm_sim* S) {
  if (!B || !S) return fail(CPM_EINVAL, "cpm_batch_add: null argument");
  if (S->batch) return fail(CPM_EINVAL, "cpm_batch_add: the simulation already belongs to a batch");
  if (S->ctx != B->ctx) return fail(CPM_EINVAL, "cpm_batch_add: the simulation uses another context");
  if (S->prm.mode == CPM_MODE_FP64)
    return fail(CPM_EINVAL, "cpm_batch_add: batches run the SELL paths (CPM_MODE_FAST64 or CPM_MODE_MIXED)");
  if (!B->sims.empty() && S->prm.mode != B->mode)
    return fail(CPM_EINVAL, "cpm_batch_add: every simulation of a batch uses the same mode");
  CPM_HIP(hipSetDevice(B->ctx->device));
  if (int rc = batch_workspace(S)) return rc;
  if ((int)B->sims.size() + 1 > B->cap) {
    const int cap = std::max(8, 2 * B->cap);
    double* p = nullptr;
    CPM_HIP(hipHostMalloc(&p, sizeof(double) * S_NSCAL * cap, hipHostMallocDefault));
    if (B->scal) CPM_HIP(hipHostFree(B->scal));
    B->scal = p;
    B->cap = cap;
  }
  B->mode = S->prm.mode;
  B->sims.push_back(S);
  S->batch = B;
  return CPM_OK;
}

extern "C" int cpm_batch_remove(cpm_batch* B, cpm_sim* S) {
  if (!B || !S) return fail(CPM_EINVAL, "cpm_batch_remove: null argument");
  auto it = std::find(B->sims.begin(), B->sims.end(), S);
  if (it == B->sims.end()) return fail(CPM_EINVAL, "cpm_batch_remove: the simulation is not in this batch");
  (void)hipSetDevice(B->ctx->device);
  CPM_HIP(hipStreamSynchronize(B->ctx->stream));
  B->sims.erase(it);
  S->batch = nullptr;
  return CPM_OK;
}

extern "C" int cpm_batch_size(const cpm_batch* B, int32_t* n) {
  if (!B || !n) return fail(CPM_EINVAL, "cpm_batch_size: null argument");
  *n = (int32_t)B->sims.size();
  return CPM_OK;
}

extern "C" int cpm_batch_step(cpm_batch* B, const double* H, const int32_t* use_init_in, cpm_step_report* reps) {
  const int32_t* use_init = use_init_in;
  if (!B || !H || !use_init || !reps) return fail(CPM_EINVAL, "cpm_batch_step: null argument");
  const int nb = (int)B->sims.size();
  if (nb == 0) return CPM_OK;
  cpm_ctx* ctx = B->ctx;
  CPM_HIP(hipSetDevice(ctx->device));
  std::vector<FieldPrm> P(nb);
  std::vector<int32_t> ui(nb);
  for (int j = 0; j < nb; ++j) {
    std::memset(&reps[j], 0, sizeof(cpm_step_report));
    P[j] = field_prm(B->sims[j], B->sims[j]->ctl_on ? nullptr : H + 3 * j);
    ui[j] = step_use_init(B->sims[j], use_init[j]);
  }
  use_init = ui.data();
  const bool mixed = B->mode == CPM_MODE_MIXED;
  std::vector<SolveSpec> sp(nb);
  std::vector<MultiSys> sys;
  std::vector<int> who;
  std::vector<int32_t> iters;
  std::vector<double> resid;
  std::vector<int> failed;
  std::vector<std::pair<int, MultiMgSys>> mgq;  // (member, system) still to solve
  std::vector<MultiMgSys> mgsys;
  std::vector<int> mgwho;
  std::vector<int32_t> done_it(3 * nb);
  std::vector<double> done_res(3 * nb);
  std::vector<int> done_fail(nb), solved(nb);
  for (int stage = 0; stage < 6; ++stage) {
    sys.clear();
    who.clear();
    mgq.clear();
    int hint = 0, mghint = 0, bs = 1, K = 1, mgK = 1;
    for (int j = 0; j < nb; ++j) {
      cpm_sim* S = B->sims[j];
      solved[j] = 0;
      if (int rc = stage_pre(S, stage, P[j], sp[j])) return rc;
      if (!sp[j].A) continue;
      warm_guess(S, sp[j]);
      const cpm_sim_params& p = S->prm;
      cpm_mg* mg = solve_mg(S, sp[j]);
      if (mg) {
        MultiMgSys y;
        y.M = mg;
        y.b = sp[j].bint ? sp[j].bint : sp[j].b;
        y.x = sp[j].xint ? sp[j].xint : sp[j].x;
        y.use_x0 = use_x0_fast(S, use_init[j]);
        y.tol = p.tol;
        y.maxiter = p.maxiter;
        y.delta = (p.mode == CPM_MODE_MIXED) ? p.rel_delta_mg : 0.0;
        mgq.emplace_back(j, y);
        mgK = sp[j].K;
        for (int k = 0; k < mgK; ++k) mghint = std::max(mghint, S->hint[sp[j].slot + k]);
        continue;
      }
      MultiSys m;
      m.A = sp[j].A;
      bs = sp[j].A->sell.bs;
      K = sp[j].K;
      m.n = (int64_t)sp[j].A->sell.nbrows * bs * K;
      m.b = sp[j].bint ? sp[j].bint : sp[j].b;
      m.x = sp[j].xint ? sp[j].xint : sp[j].x;
      m.use_x0 = use_x0_fast(S, use_init[j]);
      m.tol = p.tol;
      m.delta = solve_delta(p, sp[j].slot);
      m.maxiter = p.maxiter;
      m.r = S->wr;
      m.p = S->wp;
      m.q = S->wq;
      m.y64 = S->wy;
      m.r32 = S->wr32;
      m.p32 = S->wp32;
      m.q32 = S->wq32;
      m.x32 = S->wx32;
      sys.push_back(m);
      who.push_back(j);
      for (int k = 0; k < K; ++k) hint = std::max(hint, S->hint[sp[j].slot + k]);
    }
    if (sys.empty() && mgq.empty()) continue;
    const int ns = (int)sys.size();
    iters.assign(ns * K, 0);
    resid.assign(ns * K, 0.0);
    failed.assign(ns, 0);
    if (ns)
      if (int rc = multi_cg(ctx, B->ms[stage], bs, K, mixed ? 1 : 0, sys, hint, (B->flags & CPM_BATCH_GRAPH) != 0,
                            iters.data(), resid.data(), failed.data()))
        return rc;
    // multigrid members: one batched solve per (precision, degree) with each
    // hierarchy at most once (members sharing operators go to later rounds)
    while (!mgq.empty()) {
      const cpm_mg* M0 = mgq[0].second.M;
      const int key = (M0->prec + M0->half) * 16 + M0->degree;
      mgsys.clear();
      mgwho.clear();
      std::vector<std::pair<int, MultiMgSys>> rest;
      for (auto& e : mgq) {
        const cpm_mg* M = e.second.M;
        bool dup = false;
        for (const auto& y : mgsys) dup |= (y.M == M);
        if (dup || (M->prec + M->half) * 16 + M->degree != key)
          rest.push_back(e);
        else
          mgsys.push_back(e.second), mgwho.push_back(e.first);
      }
      const int nm = (int)mgsys.size();
      std::vector<int32_t> mit(nm * mgK);
      std::vector<double> mres(nm * mgK);
      std::vector<int> mfail(nm);
      if (int rc = multi_mg_pcg(ctx, B->mgs[stage][key], mgsys, mgK, mghint, mit.data(), mres.data(), mfail.data()))
        return rc;
      for (int i = 0; i < nm; ++i) {
        const int j = mgwho[i];
        for (int k = 0; k < mgK; ++k) done_it[3 * j + k] = mit[i * mgK + k], done_res[3 * j + k] = mres[i * mgK + k];
        done_fail[j] = mfail[i], solved[j] = 1;
      }
      mgq.swap(rest);
    }
    for (int i = 0; i < ns; ++i) {
      const int j = who[i];
      cpm_sim* S = B->sims[j];
      const SolveSpec& s = sp[j];
      for (int k = 0; k < K; ++k) {
        reps[j].iters[s.slot + k] = iters[i * K + k];
        S->hint[s.slot + k] = iters[i * K + k];
      }
      if (failed[i] && !reps[j].failed) {
        int kf = 0;
        for (int k = 1; k < K; ++k)
          if (iters[i * K + k] > iters[i * K + kf]) kf = k;
        reps[j].failed = s.slot + kf + 1;
        reps[j].resid = resid[i * K + kf];
      }
      if (int rc = stage_post(S, s)) return rc;
    }
    for (int j = 0; j < nb; ++j) {
      if (!solved[j]) continue;
      cpm_sim* S = B->sims[j];
      const SolveSpec& s = sp[j];
      for (int k = 0; k < s.K; ++k) {
        reps[j].iters[s.slot + k] = done_it[3 * j + k];
        S->hint[s.slot + k] = done_it[3 * j + k];
      }
      if (done_fail[j] && !reps[j].failed) {
        int kf = 0;
        for (int k = 1; k < s.K; ++k)
          if (done_it[3 * j + k] > done_it[3 * j + kf]) kf = k;
        reps[j].failed = s.slot + kf + 1;
        reps[j].resid = done_res[3 * j + kf];
      }
      if (int rc = stage_post(S, s)) return rc;
    }
  }
  for (int j = 0; j < nb; ++j)
    if (int rc = stage_final(B->sims[j], B->scal + (int64_t)j * S_NSCAL)) return rc;
  CPM_HIP(hipStreamSynchronize(ctx->stream));
  for (int j = 0; j < nb; ++j) report_scalars(B->sims[j], B->scal + (int64_t)j * S_NSCAL, &reps[j]);
  return CPM_OK;
}

// ---- standalone field operators -------------------------------------------

extern "C" int cpm_hd_from_u(cpm_ctx* ctx, const cpm_mat* F3, const double* U, double* H) {
  if (!ctx || !U || !H) return fail(CPM_EINVAL, "cpm_hd_from_u: null argument");
  if (int rc = check_f3(F3)) return rc;
  const int64_t n = F3->nrows;
  CPM_G_DISPATCH(F3->lanes, k_hd, n, ctx->stream, (int32_t)n, F3->ptr, F3->idx, F3->val, F3->nnz, U, H);
  CPM_HIP(hipGetLastError());
  return CPM_OK;
}

extern "C" int cpm_strain_from_u(cpm_ctx* ctx, const cpm_mat* F3, const double* u, const double* R, double* E) {
  if (!ctx || !u || !E) return fail(CPM_EINVAL, "cpm_strain_from_u: null argument");
  if (int rc = check_f3(F3)) return rc;
  const int64_t n = F3->nrows;
  double* dR = ctx->scal;  // first 9 slots of the context scratch
  if (R) CPM_HIP(hipMemcpyAsync(dR, R, 9 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  CPM_G_DISPATCH(F3->lanes, k_strain, n, ctx->stream, (int32_t)n, F3->ptr, F3->idx, F3->val, F3->nnz, u, R ? 1 : 0,
                 dR, E);
  CPM_HIP(hipGetLastError());
  CPM_HIP(hipStreamSynchronize(ctx->stream));
  return CPM_OK;
}

extern "C" int cpm_build_e0(cpm_ctx* ctx, int32_t n, double l100, double l111, const double* m, const double* R,
                            const double* Mv, double* E0) {
  if (!ctx || !m || !E0 || n < 0) return fail(CPM_EINVAL, "cpm_build_e0: bad argument");
  if ((R == nullptr) != (Mv == nullptr)) return fail(CPM_EINVAL, "cpm_build_e0: give both R and Mv or neither");
  double* d = ctx->scal + 16;  // 9 + 36 doubles of context scratch
  if (R) {
    CPM_HIP(hipMemcpyAsync(d, R, 9 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    CPM_HIP(hipMemcpyAsync(d + 9, Mv, 36 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  }
  k_e0<<<grid_for(n), kBlock, 0, ctx->stream>>>(n, m, l100, l111, R ? 1 : 0, d, d + 9, E0);
  CPM_HIP(hipGetLastError());
  CPM_HIP(hipStreamSynchronize(ctx->stream));
  return CPM_OK;
}

extern "C" int cpm_weighted_sum(cpm_ctx* ctx, int32_t n, int ncomp, const double* W, const double* f, double* out) {
  if (!ctx || !W || !f || !out || n < 0) return fail(CPM_EINVAL, "cpm_weighted_sum: bad argument");
  double* d = ctx->scal + 128;
  const int g = grid_for(n);
  hipStream_t s = ctx->stream;
  switch (ncomp) {
    case 1: k_wsum<1><<<g, kBlock, 0, s>>>(n, W, f, d, ctx->partials, ctx->counter); break;
    case 2: k_wsum<2><<<g, kBlock, 0, s>>>(n, W, f, d, ctx->partials, ctx->counter); break;
    case 3: k_wsum<3><<<g, kBlock, 0, s>>>(n, W, f, d, ctx->partials, ctx->counter); break;
    case 4: k_wsum<4><<<g, kBlock, 0, s>>>(n, W, f, d, ctx->partials, ctx->counter); break;
    case 5: k_wsum<5><<<g, kBlock, 0, s>>>(n, W, f, d, ctx->partials, ctx->counter); break;
    case 6: k_wsum<6><<<g, kBlock, 0, s>>>(n, W, f, d, ctx->partials, ctx->counter); break;
    case 7: k_wsum<7><<<g, kBlock, 0, s>>>(n, W, f, d, ctx->partials, ctx->counter); break;
    case 8: k_wsum<8><<<g, kBlock, 0, s>>>(n, W, f, d, ctx->partials, ctx->counter); break;
    default: return fail(CPM_EINVAL, "cpm_weighted_sum: ncomp must be 1..8");
  }
  CPM_HIP(hipGetLastError());
  CPM_HIP(hipMemcpyAsync(ctx->pinned, d, ncomp * sizeof(double), hipMemcpyDeviceToHost, s));
  CPM_HIP(hipStreamSynchronize(s));
  std::memcpy(out, ctx->pinned, ncomp * sizeof(double));
  return CPM_OK;
}
