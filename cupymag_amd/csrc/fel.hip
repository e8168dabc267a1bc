// Matrix-free magnetoelastic right-hand side b = F_el E0 on structured
// periodic Kuhn meshes.
//
// F_el is the operator of assemble_elasticity.py:268-348,507-527: for every
// tetrahedron e and local nodes a, b, F_e[3a+d][6b+v] = sum_q (B_a^T C)[d][v]
// N_q[b] detJ w_q, rows gid_a + d*n, columns 6 gid_b + v, rows {0, n, 2n}
// zeroed.  Its CSR form holds 270 entries per row triple (6.8 GB at 128^3)
// and is read once per step; here each DOF thread walks the 24 tetrahedra
// around its vertex (8 cubes of the periodic grid, 6 Kuhn tetrahedra each),
// recomputes the element geometry from the node coordinates exactly as the
// host assembler does (tet Jacobian rows = edge vectors, signed detJ, zero
// element when |detJ| < 1e-14, the 4-point table N_q with weights 1/24), and
// accumulates its three rows.  No atomics: every output is written by one
// thread in a fixed order, so the result is deterministic; it differs from
// the CSR product only in summation order.
//
// Mesh: kuhn_mesh order, node (i,j,k) of the (n+1)^3 grid at index
// (i*(n+1) + j)*(n+1) + k; cube c has the tetrahedra of the 6 axis
// permutations pi: vertices c, c + e_pi0, c + e_pi0 + e_pi1, c + (1,1,1),
// the last two swapped when that order is negatively oriented; DOF of a
// vertex = its indices mod n.
#include <algorithm>
#include <cmath>
#include <vector>

#include "device.h"

struct cpm_fel {
  cpm_ctx* ctx = nullptr;
  int32_t n = 0;             // cells per dimension
  double* coords = nullptr;  // (n+1)^3 x 3
  double* dR = nullptr;      // 9 doubles: the rotation of cpm_fel_strain
  double C[36];
  // regular lattice (every node at origin + (i hx, j hy, k hz)): F_el E0 is
  // one constant 27-point stencil, S[offset][d][v] (27 x 3 x 6), and so are
  // Fx, Fy, Fz, FS[offset][k] (27 x 3); device copies
  double* S = nullptr;
  double* FS = nullptr;
};

namespace cpm {

struct FelArgs {
  int32_t n;
  const double* coords;
  double C[36];
};

// local vertex offsets (bit 0: x, bit 1: y, bit 2: z) of the 6 Kuhn
// tetrahedra of a cube, in kuhn_mesh order, orientation fixed
struct KuhnTets {
  int8_t v[6][4];
};

static KuhnTets kuhn_tets() {
  KuhnTets K{};
  const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  for (int p = 0; p < 6; ++p) {
    int off[4][3] = {{0, 0, 0}};
    for (int s = 0; s < 3; ++s) {
      for (int a = 0; a < 3; ++a) off[s + 1][a] = off[s][a];
      off[s + 1][perms[p][s]] += 1;
    }
    // orientation: det of the rows off[1..3] - off[0]
    const int(*o)[3] = off;
    const int det = o[1][0] * (o[2][1] * o[3][2] - o[2][2] * o[3][1]) - o[1][1] * (o[2][0] * o[3][2] - o[2][2] * o[3][0]) +
                    o[1][2] * (o[2][0] * o[3][1] - o[2][1] * o[3][0]);
    int order[4] = {0, 1, 2, 3};
    if (det < 0) order[2] = 3, order[3] = 2;
    for (int a = 0; a < 4; ++a) {
      const int* q = off[order[a]];
      K.v[p][a] = (int8_t)(q[0] | (q[1] << 1) | (q[2] << 2));
    }
  }
  return K;
}

__global__ void __launch_bounds__(kBlock) k_fel_mf(FelArgs fa, KuhnTets kt, const double* __restrict__ E0,
                                                   double* __restrict__ b) {
  const int32_t n = fa.n;
  const int64_t N = (int64_t)n * n * n;
  const int64_t n1 = n + 1;
  const double a7 = 0.5854102, b7 = 0.1381966;  // N_q[b] of the 4-point rule (7-digit table)
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t dof = (int64_t)blockIdx.x * kBlock + threadIdx.x; dof < N; dof += stride) {
    const int32_t i = (int32_t)(dof / ((int64_t)n * n)), j = (int32_t)((dof / n) % n), k = (int32_t)(dof % n);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int o = 0; o < 8; ++o) {
      // cube whose local vertex l is (i,j,k): l = 1 - o (per axis)
      const int lx = 1 - (o & 1), ly = 1 - ((o >> 1) & 1), lz = 1 - ((o >> 2) & 1);
      int32_t ci = i - lx, cj = j - ly, ck = k - lz;
      ci += (ci < 0) ? n : 0;
      cj += (cj < 0) ? n : 0;
      ck += (ck < 0) ? n : 0;
      const int lcode = lx | (ly << 1) | (lz << 2);
      for (int t = 0; t < 6; ++t) {
        int a = -1;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (kt.v[t][q] == lcode) a = q;
        if (a < 0) continue;
        double x[4][3];
        int64_t gdof[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int c = kt.v[t][q];
          const int64_t vi = ci + (c & 1), vj = cj + ((c >> 1) & 1), vk = ck + ((c >> 2) & 1);
          const int64_t node = (vi * n1 + vj) * n1 + vk;
#pragma unroll
          for (int d = 0; d < 3; ++d) x[q][d] = fa.coords[node * 3 + d];
          gdof[q] = ((vi % n) * n + (vj % n)) * n + (vk % n);
        }
        double J[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) J[r][c] = x[r + 1][c] - x[0][c];
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        if (fabs(det) < 1e-14) continue;
        const double id = 1.0 / det;
        const double inv[3][3] = {
            {c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id, (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id},
            {c01 * id, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id, (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id},
            {c02 * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id, (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id}};
        // reference gradient of local node a: (-1,-1,-1) for a = 0, e_(a-1) otherwise
        double dref[3] = {a == 0 ? -1.0 : (a == 1 ? 1.0 : 0.0), a == 0 ? -1.0 : (a == 2 ? 1.0 : 0.0),
                          a == 0 ? -1.0 : (a == 3 ? 1.0 : 0.0)};
        double g[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) g[r] = inv[r][0] * dref[0] + inv[r][1] * dref[1] + inv[r][2] * dref[2];
        // B_a^T C (3 x 6)
        const double dx = g[0], dy = g[1], dz = g[2];
        double BtC[3][6];
#pragma unroll
        for (int v = 0; v < 6; ++v) {
          const double* Cc = fa.C;
          BtC[0][v] = dx * Cc[0 * 6 + v] + dy * Cc[3 * 6 + v] + dz * Cc[5 * 6 + v];
          BtC[1][v] = dy * Cc[1 * 6 + v] + dx * Cc[3 * 6 + v] + dz * Cc[4 * 6 + v];
          BtC[2][v] = dz * Cc[2 * 6 + v] + dy * Cc[4 * 6 + v] + dx * Cc[5 * 6 + v];
        }
        // sum_q sum_b N_q[b] E0_b, then x detJ / 24
        double Eb[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
          for (int bb = 0; bb < 4; ++bb) {
            const double w = (q == bb) ? a7 : b7;
#pragma unroll
            for (int v = 0; v < 6; ++v) Eb[v] += w * E0[gdof[bb] * 6 + v];
          }
        const double dw = det * (1.0 / 24.0);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          double s = 0.0;
#pragma unroll
          for (int v = 0; v < 6; ++v) s += BtC[d][v] * Eb[v];
          acc[d] += s * dw;
        }
      }
    }
    for (int d = 0; d < 3; ++d) b[d * N + dof] = (dof == 0) ? 0.0 : acc[d];
  }
}

// b = F_el E0 on a regular lattice: every DOF row is the same 27-point
// stencil of 3 x 6 blocks (periodic), S[(o*3 + d)*6 + v] for the neighbour at
// offset o = (di+1)*9 + (dj+1)*3 + (dk+1); rows 0, n, 2n stay zero
__global__ void __launch_bounds__(kBlock) k_fel_stencil(int32_t n, const double* __restrict__ S,
                                                        const double* __restrict__ E0, double* __restrict__ b) {
  const int64_t N = (int64_t)n * n * n;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t dof = (int64_t)blockIdx.x * kBlock + threadIdx.x; dof < N; dof += stride) {
    const int32_t i = (int32_t)(dof / ((int64_t)n * n)), j = (int32_t)((dof / n) % n), k = (int32_t)(dof % n);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int di = -1; di <= 1; ++di) {
      int32_t ii = i + di;
      ii += (ii < 0) ? n : 0;
      ii -= (ii >= n) ? n : 0;
      for (int dj = -1; dj <= 1; ++dj) {
        int32_t jj = j + dj;
        jj += (jj < 0) ? n : 0;
        jj -= (jj >= n) ? n : 0;
#pragma unroll
        for (int dk = -1; dk <= 1; ++dk) {
          int32_t kk = k + dk;
          kk += (kk < 0) ? n : 0;
          kk -= (kk >= n) ? n : 0;
          const int o = (di + 1) * 9 + (dj + 1) * 3 + (dk + 1);
          const double* e = E0 + (((int64_t)ii * n + jj) * n + kk) * 6;
          double ev[6];
#pragma unroll
          for (int v = 0; v < 6; ++v) ev[v] = e[v];
#pragma unroll
          for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int v = 0; v < 6; ++v) acc[d] += S[(o * 3 + d) * 6 + v] * ev[v];
        }
      }
    }
    for (int d = 0; d < 3; ++d) b[d * N + dof] = (dof == 0) ? 0.0 : acc[d];
  }
}

// Host: the stencil of k_fel_stencil for a regular lattice of spacings h,
// from the element terms k_fel_mf forms per tetrahedron (B_a^T C times
// detJ / 24, and the weights of the 4-point rule summed per node).
static void fel_stencil(const double h[3], const double C[36], const KuhnTets& kt, std::vector<double>& S) {
  const double a7 = 0.5854102, b7 = 0.1381966;
  S.assign(27 * 3 * 6, 0.0);
  for (int o = 0; o < 8; ++o) {
    const int lx = 1 - (o & 1), ly = 1 - ((o >> 1) & 1), lz = 1 - ((o >> 2) & 1);
    const int lcode = lx | (ly << 1) | (lz << 2);
    for (int t = 0; t < 6; ++t) {
      int a = -1;
      for (int q = 0; q < 4; ++q)
        if (kt.v[t][q] == lcode) a = q;
      if (a < 0) continue;
      double x[4][3];
      for (int q = 0; q < 4; ++q) {
        const int c = kt.v[t][q];
        x[q][0] = (c & 1) * h[0], x[q][1] = ((c >> 1) & 1) * h[1], x[q][2] = ((c >> 2) & 1) * h[2];
      }
      double J[3][3];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) J[r][c] = x[r + 1][c] - x[0][c];
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      if (std::fabs(det) < 1e-14) continue;
      const double id = 1.0 / det;
      const double inv[3][3] = {
          {c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id, (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id},
          {c01 * id, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id, (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id},
          {c02 * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id, (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id}};
      const double dref[3] = {a == 0 ? -1.0 : (a == 1 ? 1.0 : 0.0), a == 0 ? -1.0 : (a == 2 ? 1.0 : 0.0),
                              a == 0 ? -1.0 : (a == 3 ? 1.0 : 0.0)};
      double g[3];
      for (int r = 0; r < 3; ++r) g[r] = inv[r][0] * dref[0] + inv[r][1] * dref[1] + inv[r][2] * dref[2];
      const double dx = g[0], dy = g[1], dz = g[2];
      double BtC[3][6];
      for (int v = 0; v < 6; ++v) {
        BtC[0][v] = dx * C[0 * 6 + v] + dy * C[3 * 6 + v] + dz * C[5 * 6 + v];
        BtC[1][v] = dy * C[1 * 6 + v] + dx * C[3 * 6 + v] + dz * C[4 * 6 + v];
        BtC[2][v] = dz * C[2 * 6 + v] + dy * C[4 * 6 + v] + dx * C[5 * 6 + v];
      }
      const double dw = det * (1.0 / 24.0);
      for (int bb = 0; bb < 4; ++bb) {
        double wsum = 0.0;
        for (int q = 0; q < 4; ++q) wsum += (q == bb) ? a7 : b7;
        // node bb of the tetrahedron, relative to the row's vertex
        const int c = kt.v[t][bb];
        const int oi = (c & 1) - lx, oj = ((c >> 1) & 1) - ly, ok = ((c >> 2) & 1) - lz;
        const int off = (oi + 1) * 9 + (oj + 1) * 3 + (ok + 1);
        for (int d = 0; d < 3; ++d)
          for (int v = 0; v < 6; ++v) S[(off * 3 + d) * 6 + v] += BtC[d][v] * dw * wsum;
      }
    }
  }
}

// Host: the stencil of k_fder_stencil for a regular lattice of spacings h:
// FS[o*3 + k] = sum over the tetrahedra at the row's vertex a and their
// nodes b at offset o of sum_q N_q[a] (dN_b/dx_k) detJ / 24 (the element
// terms of k_fder).
static void fder_stencil(const double h[3], const KuhnTets& kt, std::vector<double>& FS) {
  const double a7 = 0.5854102, b7 = 0.1381966;
  FS.assign(27 * 3, 0.0);
  for (int o = 0; o < 8; ++o) {
    const int lx = 1 - (o & 1), ly = 1 - ((o >> 1) & 1), lz = 1 - ((o >> 2) & 1);
    const int lcode = lx | (ly << 1) | (lz << 2);
    for (int t = 0; t < 6; ++t) {
      int a = -1;
      for (int q = 0; q < 4; ++q)
        if (kt.v[t][q] == lcode) a = q;
      if (a < 0) continue;
      double x[4][3];
      for (int q = 0; q < 4; ++q) {
        const int c = kt.v[t][q];
        x[q][0] = (c & 1) * h[0], x[q][1] = ((c >> 1) & 1) * h[1], x[q][2] = ((c >> 2) & 1) * h[2];
      }
      double J[3][3];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) J[r][c] = x[r + 1][c] - x[0][c];
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      if (std::fabs(det) < 1e-14) continue;
      const double id = 1.0 / det;
      const double inv[3][3] = {
          {c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id, (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id},
          {c01 * id, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id, (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id},
          {c02 * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id, (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id}};
      const double dw = det * (1.0 / 24.0);
      for (int b = 0; b < 4; ++b) {
        const double dref[3] = {b == 0 ? -1.0 : (b == 1 ? 1.0 : 0.0), b == 0 ? -1.0 : (b == 2 ? 1.0 : 0.0),
                                b == 0 ? -1.0 : (b == 3 ? 1.0 : 0.0)};
        double g[3];
        for (int r = 0; r < 3; ++r) g[r] = inv[r][0] * dref[0] + inv[r][1] * dref[1] + inv[r][2] * dref[2];
        double f[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < 4; ++q) {
          const double nq = (q == a) ? a7 : b7;
          for (int c = 0; c < 3; ++c) f[c] += nq * g[c] * dw;
        }
        const int cb = kt.v[t][b];
        const int oi = (cb & 1) - lx, oj = ((cb >> 1) & 1) - ly, ok = ((cb >> 2) & 1) - lz;
        const int off = (oi + 1) * 9 + (oj + 1) * 3 + (ok + 1);
        for (int c = 0; c < 3; ++c) FS[off * 3 + c] += f[c];
      }
    }
  }
}

// ---- matrix-free weak-form derivatives Fx, Fy, Fz -------------------------
// F_k[a][b] = sum_q N_q[a] (dN_b/dx_k) detJ w_q (assemble_demag.py:235-305,
// the host assembler's scalar_row with the same products in the same
// order): per element the values of the assembled CSR, summed over the 24
// tetrahedra of a vertex in a fixed order instead of along the CSR row.
// Geometry of tetrahedron t of the cube at (ci, cj, ck): corner DOFs,
// reference-to-physical gradients of the 4 shape functions and detJ; false
// for the zero element (|detJ| < 1e-14).
__device__ inline bool kuhn_tet_grads(const double* __restrict__ coords, int32_t n, const KuhnTets& kt, int t,
                                      int32_t ci, int32_t cj, int32_t ck, int64_t (&gdof)[4], double (&g)[4][3],
                                      double& det) {
  const int64_t n1 = n + 1;
  double x[4][3];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int c = kt.v[t][q];
    const int64_t vi = ci + (c & 1), vj = cj + ((c >> 1) & 1), vk = ck + ((c >> 2) & 1);
    const int64_t node = (vi * n1 + vj) * n1 + vk;
#pragma unroll
    for (int d = 0; d < 3; ++d) x[q][d] = coords[node * 3 + d];
    gdof[q] = ((vi % n) * n + (vj % n)) * n + (vk % n);
  }
  double J[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) J[r][c] = x[r + 1][c] - x[0][c];
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  if (fabs(det) < 1e-14) return false;
  const double id = 1.0 / det;
  const double inv[3][3] = {
      {c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id, (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id},
      {c01 * id, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id, (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id},
      {c02 * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id, (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id}};
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const double dref[3] = {b == 0 ? -1.0 : (b == 1 ? 1.0 : 0.0), b == 0 ? -1.0 : (b == 2 ? 1.0 : 0.0),
                            b == 0 ? -1.0 : (b == 3 ? 1.0 : 0.0)};
#pragma unroll
    for (int r = 0; r < 3; ++r) g[b][r] = inv[r][0] * dref[0] + inv[r][1] * dref[1] + inv[r][2] * dref[2];
  }
  return true;
}

struct FderIn {
  const double* u;    // strain: displacement [ux; uy; uz] (3n), or null
  const double* m;    // demag RHS: m (n, 3), or null
  const double* avg;  // demag RHS: device average magnetization (3)
  const double* U;    // Htilde: the demag potential (n), or null
  const double* x;    // one product F_k x (n), or null
  int32_t k;          // ... its component
  int32_t rot;        // strain in the rotated frame (R mixing)
  const double* R;    // device 3x3 (rot)
};
struct FderOut {
  double* E;   // (n, 6) strain
  double* bd;  // (n) demag RHS -(Fx m~x + Fy m~y + Fz m~z), entry 0 = 0
  double* H;   // (3, n) Htilde = -F U
  double* y;   // (n) F_k x
};

// The outputs of the derivative products at one DOF (k_fder, k_fder_stencil):
// D[k][j] = (F_k u_j), sm[k] = (F_k m~_k), sh[k] = (F_k U), sx = (F_k x).
__device__ inline void fder_finish(int64_t dof, int64_t N, const FderIn& in, const FderOut& out, double (&D)[3][3],
                                   const double (&sm)[3], const double (&sh)[3], double sx) {
  if (in.u) {
    double e[6];
    if (!in.rot) {
      e[0] = D[0][0];
      e[1] = D[1][1];
      e[2] = D[2][2];
      e[3] = D[1][0] + D[0][1];
      e[4] = D[2][1] + D[1][2];
      e[5] = D[2][0] + D[0][2];
    } else {
      // compute_derivatives.py:94-123; D[k][j]: derivative k of component j
      const double* R = in.R;
      auto mix = [&](int col, int jj) { return R[col] * D[0][jj] + R[3 + col] * D[1][jj] + R[6 + col] * D[2][jj]; };
      e[0] = mix(0, 0);
      e[1] = mix(1, 1);
      e[2] = mix(2, 2);
      e[3] = mix(1, 0) + mix(0, 1);
      e[4] = mix(2, 1) + mix(1, 2);
      e[5] = mix(2, 0) + mix(0, 2);
    }
#pragma unroll
    for (int v = 0; v < 6; ++v) out.E[6 * dof + v] = e[v];
  }
  if (in.m) out.bd[dof] = (dof == 0) ? 0.0 : -((sm[0] + sm[1]) + sm[2]);
  if (in.U) {
    out.H[dof] = -sh[0];
    out.H[N + dof] = -sh[1];
    out.H[2 * N + dof] = -sh[2];
  }
  if (in.x) out.y[dof] = sx;
}

// One thread per DOF walks the 24 tetrahedra of its vertex and forms the
// products requested in `in` (each output written once: deterministic).
__global__ void __launch_bounds__(kBlock) k_fder(int32_t n, const double* __restrict__ coords, KuhnTets kt,
                                                 FderIn in, FderOut out) {
  const int64_t N = (int64_t)n * n * n;
  const double a7 = 0.5854102, b7 = 0.1381966;  // N_q[a] of the 4-point rule (7-digit table)
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  const double a0 = in.m ? in.avg[0] : 0.0, a1 = in.m ? in.avg[1] : 0.0, a2 = in.m ? in.avg[2] : 0.0;
  for (int64_t dof = (int64_t)blockIdx.x * kBlock + threadIdx.x; dof < N; dof += stride) {
    const int32_t i = (int32_t)(dof / ((int64_t)n * n)), j = (int32_t)((dof / n) % n), k = (int32_t)(dof % n);
    double D[3][3] = {};       // D[k][j] = (F_k u_j)_dof
    double sm[3] = {0, 0, 0};  // (F_k m~_k)_dof
    double sh[3] = {0, 0, 0};  // (F_k U)_dof
    double sx = 0.0;           // (F_k x)_dof
    for (int o = 0; o < 8; ++o) {
      const int lx = 1 - (o & 1), ly = 1 - ((o >> 1) & 1), lz = 1 - ((o >> 2) & 1);
      int32_t ci = i - lx, cj = j - ly, ck = k - lz;
      ci += (ci < 0) ? n : 0;
      cj += (cj < 0) ? n : 0;
      ck += (ck < 0) ? n : 0;
      const int lcode = lx | (ly << 1) | (lz << 2);
      for (int t = 0; t < 6; ++t) {
        int a = -1;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (kt.v[t][q] == lcode) a = q;
        if (a < 0) continue;
        int64_t gd[4];
        double g[4][3], det;
        if (!kuhn_tet_grads(coords, n, kt, t, ci, cj, ck, gd, g, det)) continue;
        const double dw = det * (1.0 / 24.0);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          double f[3] = {0.0, 0.0, 0.0};
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double nq = (q == a) ? a7 : b7;
#pragma unroll
            for (int c = 0; c < 3; ++c) f[c] += nq * g[b][c] * dw;
          }
          const int64_t cb = gd[b];
          if (in.u) {
            const double uj[3] = {in.u[cb], in.u[N + cb], in.u[2 * N + cb]};
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
              for (int jj = 0; jj < 3; ++jj) D[c][jj] += f[c] * uj[jj];
          }
          if (in.m) {
            sm[0] += f[0] * (in.m[3 * cb] - a0);
            sm[1] += f[1] * (in.m[3 * cb + 1] - a1);
            sm[2] += f[2] * (in.m[3 * cb + 2] - a2);
          }
          if (in.U) {
            const double uc = in.U[cb];
#pragma unroll
            for (int c = 0; c < 3; ++c) sh[c] += f[c] * uc;
          }
          if (in.x) sx += f[in.k] * in.x[cb];
        }
      }
    }
    fder_finish(dof, N, in, out, D, sm, sh, sx);
  }
}

// b = F x for the products of k_fder on a regular lattice: every row is the
// same 27-point stencil, FS[o*3 + k] the coefficient of F_k at neighbour
// offset o = (di+1)*9 + (dj+1)*3 + (dk+1)
__global__ void __launch_bounds__(kBlock) k_fder_stencil(int32_t n, const double* __restrict__ FS, FderIn in,
                                                         FderOut out) {
  const int64_t N = (int64_t)n * n * n;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  const double a0 = in.m ? in.avg[0] : 0.0, a1 = in.m ? in.avg[1] : 0.0, a2 = in.m ? in.avg[2] : 0.0;
  for (int64_t dof = (int64_t)blockIdx.x * kBlock + threadIdx.x; dof < N; dof += stride) {
    const int32_t i = (int32_t)(dof / ((int64_t)n * n)), j = (int32_t)((dof / n) % n), k = (int32_t)(dof % n);
    double D[3][3] = {};
    double sm[3] = {0, 0, 0};
    double sh[3] = {0, 0, 0};
    double sx = 0.0;
    for (int di = -1; di <= 1; ++di) {
      int32_t ii = i + di;
      ii += (ii < 0) ? n : 0;
      ii -= (ii >= n) ? n : 0;
      for (int dj = -1; dj <= 1; ++dj) {
        int32_t jj = j + dj;
        jj += (jj < 0) ? n : 0;
        jj -= (jj >= n) ? n : 0;
#pragma unroll
        for (int dk = -1; dk <= 1; ++dk) {
          int32_t kk = k + dk;
          kk += (kk < 0) ? n : 0;
          kk -= (kk >= n) ? n : 0;
          const int o = (di + 1) * 9 + (dj + 1) * 3 + (dk + 1);
          const double f[3] = {FS[o * 3], FS[o * 3 + 1], FS[o * 3 + 2]};
          const int64_t cb = ((int64_t)ii * n + jj) * n + kk;
          if (in.u) {
            const double uj[3] = {in.u[cb], in.u[N + cb], in.u[2 * N + cb]};
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
              for (int q = 0; q < 3; ++q) D[c][q] += f[c] * uj[q];
          }
          if (in.m) {
            sm[0] += f[0] * (in.m[3 * cb] - a0);
            sm[1] += f[1] * (in.m[3 * cb + 1] - a1);
            sm[2] += f[2] * (in.m[3 * cb + 2] - a2);
          }
          if (in.U) {
            const double uc = in.U[cb];
#pragma unroll
            for (int c = 0; c < 3; ++c) sh[c] += f[c] * uc;
          }
          if (in.x) sx += f[in.k] * in.x[cb];
        }
      }
    }
    fder_finish(dof, N, in, out, D, sm, sh, sx);
  }
}

int fder_apply(const cpm_fel* F, const FderIn& in, const FderOut& out) {
  static const KuhnTets kt = kuhn_tets();
  const int64_t N = (int64_t)F->n * F->n * F->n;
  if (F->FS)
    k_fder_stencil<<<grid_for(N), kBlock, 0, F->ctx->stream>>>(F->n, F->FS, in, out);
  else
    k_fder<<<grid_for(N), kBlock, 0, F->ctx->stream>>>(F->n, F->coords, kt, in, out);
  CPM_HIP(hipGetLastError());
  return CPM_OK;
}

int fder_strain_demag(const cpm_fel* F, const double* u, int rot, const double* dR, const double* m,
                      const double* davg, double* E, double* bd) {
  FderIn in{};
  in.u = u;
  in.rot = rot;
  in.R = dR;
  in.m = m;
  in.avg = davg;
  FderOut out{};
  out.E = E;
  out.bd = bd;
  return fder_apply(F, in, out);
}

int fder_hd(const cpm_fel* F, const double* U, double* H) {
  FderIn in{};
  in.U = U;
  FderOut out{};
  out.H = H;
  return fder_apply(F, in, out);
}

int fel_apply(const cpm_fel* F, const double* E0, double* b) {
  const int64_t N = (int64_t)F->n * F->n * F->n;
  if (F->S) {
    k_fel_stencil<<<grid_for(N), kBlock, 0, F->ctx->stream>>>(F->n, F->S, E0, b);
    CPM_HIP(hipGetLastError());
    return CPM_OK;
  }
  FelArgs fa;
  fa.n = F->n;
  fa.coords = F->coords;
  std::copy(F->C, F->C + 36, fa.C);
  static const KuhnTets kt = kuhn_tets();
  k_fel_mf<<<grid_for(N), kBlock, 0, F->ctx->stream>>>(fa, kt, E0, b);
  CPM_HIP(hipGetLastError());
  return CPM_OK;
}

int64_t fel_rows(const cpm_fel* F) { return 3 * (int64_t)F->n * F->n * F->n; }

}  // namespace cpm

using namespace cpm;

extern "C" int cpm_fel_create(cpm_ctx* ctx, int32_t n, const double* coords, const double* C, cpm_fel** out) {
  if (!ctx || !coords || !out) return fail(CPM_EINVAL, "cpm_fel_create: null argument");
  if (n < 2) return fail(CPM_EINVAL, "cpm_fel_create: at least 2 cells per dimension");
  CPM_HIP(hipSetDevice(ctx->device));
  auto* F = new cpm_fel();
  F->ctx = ctx;
  F->n = n;
  if (C)
    std::copy(C, C + 36, F->C);
  else
    std::fill(F->C, F->C + 36, 0.0);
  const size_t bytes = sizeof(double) * 3 * (size_t)(n + 1) * (n + 1) * (n + 1);
  hipError_t e = hipMalloc(&F->coords, bytes);
  if (e == hipSuccess) e = hipMalloc(&F->dR, sizeof(double) * 9);
  if (e == hipSuccess) e = hipMemcpy(F->coords, coords, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    // a regular lattice (nodes at origin + (i hx, j hy, k hz) to 1e-12 of
    // the spacing): the F_el product and the derivatives as stencils
    const int64_t n1 = n + 1;
    double h[3];
    for (int a = 0; a < 3; ++a) h[a] = coords[((a == 0 ? n1 * n1 : a == 1 ? n1 : 1)) * 3 + a] - coords[a];
    bool regular = h[0] > 0 && h[1] > 0 && h[2] > 0;
    for (int64_t i = 0; i < n1 && regular; ++i)
      for (int64_t j = 0; j < n1 && regular; ++j)
        for (int64_t k = 0; k < n1 && regular; ++k) {
          const double* x = coords + ((i * n1 + j) * n1 + k) * 3;
          const int64_t idx[3] = {i, j, k};
          for (int a = 0; a < 3; ++a)
            if (std::fabs(x[a] - (coords[a] + idx[a] * h[a])) > 1e-12 * h[a]) regular = false;
        }
    if (regular && C) {
      std::vector<double> S;
      fel_stencil(h, F->C, kuhn_tets(), S);
      e = hipMalloc(&F->S, sizeof(double) * S.size());
      if (e == hipSuccess) e = hipMemcpy(F->S, S.data(), sizeof(double) * S.size(), hipMemcpyHostToDevice);
    }
    if (regular && e == hipSuccess) {
      std::vector<double> FS;
      fder_stencil(h, kuhn_tets(), FS);
      e = hipMalloc(&F->FS, sizeof(double) * FS.size());
      if (e == hipSuccess) e = hipMemcpy(F->FS, FS.data(), sizeof(double) * FS.size(), hipMemcpyHostToDevice);
    }
  }
  if (e != hipSuccess) {
    if (F->coords) (void)hipFree(F->coords);
    if (F->dR) (void)hipFree(F->dR);
    if (F->S) (void)hipFree(F->S);
    if (F->FS) (void)hipFree(F->FS);
    delete F;
    return fail(e == hipErrorOutOfMemory ? CPM_ENOMEM : CPM_EHIP, hipGetErrorString(e));
  }
  *out = F;
  return CPM_OK;
}

extern "C" int cpm_fel_destroy(cpm_fel* F) {
  if (!F) return CPM_OK;
  (void)hipSetDevice(F->ctx->device);
  (void)hipStreamSynchronize(F->ctx->stream);
  if (F->coords) (void)hipFree(F->coords);
  if (F->dR) (void)hipFree(F->dR);
  if (F->S) (void)hipFree(F->S);
  if (F->FS) (void)hipFree(F->FS);
  delete F;
  return CPM_OK;
}

extern "C" int cpm_fel_apply(cpm_fel* F, const double* E0, double* b) {
  if (!F || !E0 || !b) return fail(CPM_EINVAL, "cpm_fel_apply: null argument");
  CPM_HIP(hipSetDevice(F->ctx->device));
  return fel_apply(F, E0, b);
}

extern "C" int cpm_fel_strain(cpm_fel* F, const double* u, const double* R, double* E) {
  if (!F || !u || !E) return fail(CPM_EINVAL, "cpm_fel_strain: null argument");
  CPM_HIP(hipSetDevice(F->ctx->device));
  if (R) CPM_HIP(hipMemcpy(F->dR, R, 9 * sizeof(double), hipMemcpyHostToDevice));
  if (int rc = fder_strain_demag(F, u, R ? 1 : 0, F->dR, nullptr, nullptr, E, nullptr)) return rc;
  CPM_HIP(hipStreamSynchronize(F->ctx->stream));
  return CPM_OK;
}

extern "C" int cpm_fel_hd(cpm_fel* F, const double* U, double* H) {
  if (!F || !U || !H) return fail(CPM_EINVAL, "cpm_fel_hd: null argument");
  CPM_HIP(hipSetDevice(F->ctx->device));
  return fder_hd(F, U, H);
}

extern "C" int cpm_fel_dmul(cpm_fel* F, int k, const double* x, double* y) {
  if (!F || !x || !y) return fail(CPM_EINVAL, "cpm_fel_dmul: null argument");
  if (k < 0 || k > 2) return fail(CPM_EINVAL, "cpm_fel_dmul: k must be 0, 1 or 2 (Fx, Fy, Fz)");
  CPM_HIP(hipSetDevice(F->ctx->device));
  FderIn in{};
  in.x = x;
  in.k = k;
  FderOut out{};
  out.y = y;
  return fder_apply(F, in, out);
}
