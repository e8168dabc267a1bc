// Host-side FEM assembly straight into CSR (runs once per mesh, on the CPU).
//
// Replaces the Numba COO assemblers of the reference
//   cupymag/physics/assemble_demag.py:95-341        (K_e, F_e, anchor at DOF 0)
//   cupymag/physics/assemble_Gauss_Seidel.py:88-214  (K_e, M_e)
//   cupymag/physics/assemble_elasticity.py:194-527   (B^T C B, B^T C S, anchors)
// followed by the SciPy COO->CSR conversion and the defect-DOF surgery of
//   cupymag/utils/final_assembly.py:59-205
// and cupymag/core/assemble_simulation_operators.py:42-107.
//
// The reference materialises npe^2 (elasticity: 3npe x 6npe) COO triplets
// per element and lets SciPy sum duplicates; at a 128^3 Kuhn mesh that is
// ~3.6e9 triplets for F_el alone.  Here every DOF row is owned by one thread,
// which walks the elements incident to that DOF in ascending element order and
// accumulates the row's entries in place.  The sparsity pattern is the set of
// (row, col) pairs the reference's COO contains (explicit zeros included);
// the per-entry arithmetic is the reference's (signed detJ, J^-1 applied to
// the reference-space gradients, Gauss sums, alpha*K_e + M_e per element
// entry).  Only the order in which duplicate element contributions are
// summed differs (SciPy sorts them), a difference of a few ulp per entry.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "cupymag_hip.h"
#include "errors.h"

namespace {

struct Tables {
  int npe = 0, ngp = 0;
  double N[8][8];
  double dN[8][8][3];
  double w[8];
};

// Shape tables of the assemblers (get_N / get_dN).  Hex: trilinear shape
// functions and their reference gradients at the 2x2x2 points +-0.57735026919
// rounded to 12 decimals (cupymag/mesh/ShapeHex.py:254-393).  Tet: barycentric
// values at the 4-point rule rounded to 7 decimals, constant gradients,
// weights 1/24 (cupymag/mesh/ShapeTet.py:205-255).
Tables make_tables(int npe) {
  Tables t;
  t.npe = npe;
  if (npe == 8) {
    const double g = 0.57735026919;
    const double sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    const double sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    const double sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    auto r12 = [](double v) { return std::nearbyint(v * 1e12) / 1e12; };
    t.ngp = 8;
    int q = 0;
    for (int ix = 0; ix < 2; ++ix)
      for (int iy = 0; iy < 2; ++iy)
        for (int iz = 0; iz < 2; ++iz, ++q) {
          double r = ix ? g : -g, s = iy ? g : -g, u = iz ? g : -g;
          for (int a = 0; a < 8; ++a) {
            t.N[q][a] = r12((1 + sx[a] * r) * (1 + sy[a] * s) * (1 + sz[a] * u) / 8.0);
            t.dN[q][a][0] = r12(sx[a] * (1 + sy[a] * s) * (1 + sz[a] * u) / 8.0);
            t.dN[q][a][1] = r12(sy[a] * (1 + sx[a] * r) * (1 + sz[a] * u) / 8.0);
            t.dN[q][a][2] = r12(sz[a] * (1 + sx[a] * r) * (1 + sy[a] * s) / 8.0);
          }
          t.w[q] = 1.0;
        }
  } else {
    const double a7 = 0.5854102, b7 = 0.1381966;
    const double d[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    t.ngp = 4;
    for (int q = 0; q < 4; ++q) {
      for (int a = 0; a < 4; ++a) {
        t.N[q][a] = (a == q) ? a7 : b7;
        for (int k = 0; k < 3; ++k) t.dN[q][a][k] = d[a][k];
      }
      t.w[q] = 1.0 / 24.0;
    }
  }
  return t;
}

struct Geo {
  double det[8];
  double gr[8][8][3];  // [gp][node][x,y,z]
  bool zero;           // |detJ| < 1e-14 at some Gauss point -> zero element
};

// Jacobian conventions of the reference: hex J[j][k] = sum_i x_i[j] dN_i[k]
// (ShapeHex.element_jacobian), tet J[i][k] = x_{i+1}[k] - x_0[k]
// (ShapeTet.element_jacobian); gradients J^-1 dN in both cases.
void geometry(const Tables& t, const double x[8][3], Geo& g) {
  g.zero = false;
  const int ngp_geo = (t.npe == 8) ? t.ngp : 1;
  for (int q = 0; q < ngp_geo; ++q) {
    double J[3][3];
    if (t.npe == 8) {
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          double s = 0.0;
          for (int i = 0; i < 8; ++i) s += x[i][j] * t.dN[q][i][k];
          J[j][k] = s;
        }
    } else {
      for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) J[i][k] = x[i + 1][k] - x[0][k];
    }
    double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    if (std::fabs(det) < 1e-14) {
      g.zero = true;
      return;
    }
    double id = 1.0 / det;
    double inv[3][3] = {
        {c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id, (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id},
        {c01 * id, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id, (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id},
        {c02 * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id, (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id}};
    g.det[q] = det;
    for (int a = 0; a < t.npe; ++a)
      for (int j = 0; j < 3; ++j)
        g.gr[q][a][j] = inv[j][0] * t.dN[q][a][0] + inv[j][1] * t.dN[q][a][1] + inv[j][2] * t.dN[q][a][2];
  }
  if (t.npe != 8)
    for (int q = 1; q < t.ngp; ++q) {
      g.det[q] = g.det[0];
      std::memcpy(g.gr[q], g.gr[0], sizeof(g.gr[0]));
    }
}

// Row `a` of the scalar element matrices: K, M, Fx, Fy, Fz.
void scalar_row(const Tables& t, const Geo& g, int a, double K[8], double M[8], double F[3][8]) {
  for (int b = 0; b < t.npe; ++b) {
    K[b] = M[b] = F[0][b] = F[1][b] = F[2][b] = 0.0;
  }
  if (g.zero) return;
  for (int q = 0; q < t.ngp; ++q) {
    const double dw = g.det[q] * t.w[q];
    const double* ga = g.gr[q][a];
    for (int b = 0; b < t.npe; ++b) {
      const double* gb = g.gr[q][b];
      K[b] += (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]) * dw;
      M[b] += t.N[q][a] * t.N[q][b] * dw;
      for (int k = 0; k < 3; ++k) F[k][b] += t.N[q][a] * gb[k] * dw;
    }
  }
}

// B matrix column entries for node b: B[k][3b+d]
inline void bcol(const double* gr, double B[6][3]) {
  const double dx = gr[0], dy = gr[1], dz = gr[2];
  B[0][0] = dx; B[0][1] = 0;  B[0][2] = 0;
  B[1][0] = 0;  B[1][1] = dy; B[1][2] = 0;
  B[2][0] = 0;  B[2][1] = 0;  B[2][2] = dz;
  B[3][0] = dy; B[3][1] = dx; B[3][2] = 0;
  B[4][0] = 0;  B[4][1] = dz; B[4][2] = dy;
  B[5][0] = dz; B[5][1] = 0;  B[5][2] = dx;
}

// Rows 3a+da (da = 0..2) of the elastic element matrices:
//   Kel[da][3b+db] = sum_q sum_k (B^T C)[3a+da][k] B[k][3b+db] dw
//   Fel[da][6b+v]  = sum_q (B^T C)[3a+da][v] N_q[b] dw
void elastic_row(const Tables& t, const Geo& g, int a, const double C[6][6], double Kel[3][24],
                 double Fel[3][48]) {
  const int npe = t.npe;
  for (int d = 0; d < 3; ++d) {
    for (int j = 0; j < 3 * npe; ++j) Kel[d][j] = 0.0;
    for (int j = 0; j < 6 * npe; ++j) Fel[d][j] = 0.0;
  }
  if (g.zero) return;
  for (int q = 0; q < t.ngp; ++q) {
    const double dw = g.det[q] * t.w[q];
    double Ba[6][3];
    bcol(g.gr[q][a], Ba);
    double BtC[3][6];
    for (int d = 0; d < 3; ++d)
      for (int j = 0; j < 6; ++j) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += Ba[k][d] * C[k][j];
        BtC[d][j] = s;
      }
    for (int b = 0; b < npe; ++b) {
      double Bb[6][3];
      bcol(g.gr[q][b], Bb);
      for (int d = 0; d < 3; ++d) {
        for (int e = 0; e < 3; ++e) {
          double s = 0.0;
          for (int k = 0; k < 6; ++k) s += BtC[d][k] * Bb[k][e];
          Kel[d][3 * b + e] += s * dw;
        }
        const double nb = t.N[q][b] * dw;
        for (int v = 0; v < 6; ++v) Fel[d][6 * b + v] += BtC[d][v] * nb;
      }
    }
  }
}

template <class F>
void parallel_rows(int64_t n, int nthreads, F&& body) {
  if (nthreads <= 1 || n < 4096) {
    body(0, n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = 2048;
  std::atomic<int64_t> next{0};
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&]() {
      for (;;) {
        int64_t s = next.fetch_add(chunk);
        if (s >= n) break;
        body(s, std::min(n, s + chunk));
      }
    });
  for (auto& x : th) x.join();
}

}  // namespace

struct cpm_host_ops {
  int32_t ndof = 0;
  // base scalar pattern: A_demag, Fx, Fy, Fz share it
  std::vector<int64_t> ptr;
  std::vector<int32_t> idx;
  std::vector<double> vK, vFx, vFy, vFz;
  // Gauss-Seidel operators after defect surgery
  std::vector<int64_t> gs_ptr;
  std::vector<int32_t> gs_idx;
  std::vector<double> vA1, vA2;
  std::vector<int64_t> f1_ptr;
  std::vector<int32_t> f1_idx;
  std::vector<double> vF1;
  // elasticity
  std::vector<int64_t> el_ptr;
  std::vector<int32_t> el_idx;
  std::vector<double> vAel;
  std::vector<int64_t> fe_ptr;
  std::vector<int32_t> fe_idx;
  std::vector<double> vFel;
};

extern "C" int cpm_host_assemble(const double* coords, int64_t n_nodes, const int32_t* elems, int64_t n_elems,
                                 int npe, int elem_stride, const int32_t* gid, int32_t ndof,
                                 const uint8_t* defect_dof, const double* coef, int flags, int nthreads,
                                 cpm_host_ops** out) {
  const bool want_elastic = (flags & CPM_HA_ELASTIC) != 0;
  const bool anchors = (flags & CPM_HA_NO_ANCHORS) == 0;
  try {
    if (!coords || !elems || !gid || !coef || !out) return cpm::fail(CPM_EINVAL, "null argument");
    if (npe != 4 && npe != 8) return cpm::fail(CPM_EINVAL, "npe must be 4 (tet) or 8 (hex)");
    if (ndof <= 0) return cpm::fail(CPM_EINVAL, "ndof must be positive");
    for (int64_t i = 0; i < n_nodes; ++i)
      if (gid[i] < 0 || gid[i] >= ndof) return cpm::fail(CPM_EINVAL, "global_id out of range");
    for (int64_t e = 0; e < n_elems; ++e)
      for (int a = 0; a < npe; ++a) {
        int32_t v = elems[e * elem_stride + a];
        if (v < 0 || v >= n_nodes) return cpm::fail(CPM_EINVAL, "element node index out of range");
      }
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    const double alpha1 = coef[0], alpha2 = coef[1];
    double C[6][6];
    for (int i = 0; i < 36; ++i) C[i / 6][i % 6] = coef[2 + i];
    const Tables T = make_tables(npe);
    auto* H = new cpm_host_ops();
    H->ndof = ndof;
    const int64_t n = ndof;

    // DOF -> incident (element, local node), ascending element order
    std::vector<int64_t> iptr(n + 1, 0);
    for (int64_t e = 0; e < n_elems; ++e)
      for (int a = 0; a < npe; ++a) iptr[gid[elems[e * elem_stride + a]] + 1]++;
    for (int64_t d = 0; d < n; ++d) iptr[d + 1] += iptr[d];
    std::vector<int64_t> ilist(iptr[n]);
    {
      std::vector<int64_t> fill(iptr.begin(), iptr.end() - 1);
      for (int64_t e = 0; e < n_elems; ++e)
        for (int a = 0; a < npe; ++a) ilist[fill[gid[elems[e * elem_stride + a]]]++] = e * npe + a;
    }
    auto row_cols = [&](int64_t d, std::vector<int32_t>& cols) {
      cols.clear();
      for (int64_t k = iptr[d]; k < iptr[d + 1]; ++k) {
        int64_t e = ilist[k] / npe;
        for (int b = 0; b < npe; ++b) cols.push_back(gid[elems[e * elem_stride + b]]);
      }
      std::sort(cols.begin(), cols.end());
      cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    };
    // pattern
    H->ptr.assign(n + 1, 0);
    parallel_rows(n, nthreads, [&](int64_t s, int64_t t) {
      std::vector<int32_t> cols;
      for (int64_t d = s; d < t; ++d) {
        row_cols(d, cols);
        H->ptr[d + 1] = (int64_t)cols.size();
      }
    });
    for (int64_t d = 0; d < n; ++d) H->ptr[d + 1] += H->ptr[d];
    const int64_t nnz = H->ptr[n];
    H->idx.resize(nnz);
    H->vK.assign(nnz, 0.0);
    H->vFx.assign(nnz, 0.0);
    H->vFy.assign(nnz, 0.0);
    H->vFz.assign(nnz, 0.0);
    std::vector<double> vA1(nnz, 0.0), vA2(nnz, 0.0), vM(nnz, 0.0);
    if (want_elastic) {
      if (3 * nnz * 6 > INT32_MAX) return cpm::fail(CPM_EINVAL, "elasticity operator exceeds int32 indexing");
      H->vAel.assign((size_t)9 * nnz, 0.0);
      H->vFel.assign((size_t)18 * nnz, 0.0);
    }
    const int64_t nnzel = 3 * nnz;  // per displacement component block of rows
    parallel_rows(n, nthreads, [&](int64_t s, int64_t t) {
      std::vector<int32_t> cols;
      double x[8][3];
      Geo g;
      double K[8], M[8], F[3][8], Kel[3][24], Fel[3][48];
      for (int64_t d = s; d < t; ++d) {
        row_cols(d, cols);
        const int64_t r0 = H->ptr[d];
        const int64_t len = (int64_t)cols.size();
        std::copy(cols.begin(), cols.end(), H->idx.begin() + r0);
        for (int64_t k = iptr[d]; k < iptr[d + 1]; ++k) {
          const int64_t e = ilist[k] / npe;
          const int a = (int)(ilist[k] % npe);
          const int32_t* en = elems + e * elem_stride;
          for (int b = 0; b < npe; ++b)
            for (int j = 0; j < 3; ++j) x[b][j] = coords[(int64_t)en[b] * 3 + j];
          geometry(T, x, g);
          scalar_row(T, g, a, K, M, F);
          int64_t pos[8];
          for (int b = 0; b < npe; ++b)
            pos[b] = std::lower_bound(cols.begin(), cols.end(), gid[en[b]]) - cols.begin();
          for (int b = 0; b < npe; ++b) {
            const int64_t p = r0 + pos[b];
            H->vK[p] += K[b];
            vM[p] += M[b];
            H->vFx[p] += F[0][b];
            H->vFy[p] += F[1][b];
            H->vFz[p] += F[2][b];
            vA1[p] += K[b] * alpha1 + M[b];
            vA2[p] += K[b] * alpha2 + M[b];
          }
          if (want_elastic) {
            elastic_row(T, g, a, C, Kel, Fel);
            for (int da = 0; da < 3; ++da) {
              // A_el row da*n+d: [cols | cols+n | cols+2n]
              const int64_t ar = da * nnzel + 3 * r0;
              // F_el row da*n+d: 6*col+v
              const int64_t fr = da * 2 * nnzel + 6 * r0;
              for (int b = 0; b < npe; ++b) {
                for (int db = 0; db < 3; ++db) H->vAel[ar + db * len + pos[b]] += Kel[da][3 * b + db];
                for (int v = 0; v < 6; ++v) H->vFel[fr + 6 * pos[b] + v] += Fel[da][6 * b + v];
              }
            }
          }
        }
      }
    });
    // anchor of the demag potential at DOF 0 (assemble_demag.py:307-341)
    for (int64_t d = 0; d < n && anchors; ++d)
      for (int64_t p = H->ptr[d]; p < H->ptr[d + 1]; ++p)
        if (d == 0 || H->idx[p] == 0) H->vK[p] = (d == 0 && H->idx[p] == 0) ? 1.0 : 0.0;
    // Gauss-Seidel surgery: defect rows/cols -> identity (final_assembly.py:59-112),
    // F1 defect rows dropped (final_assembly.py:115-158)
    auto isdef = [&](int64_t d) { return defect_dof && defect_dof[d]; };
    H->gs_ptr.assign(n + 1, 0);
    H->f1_ptr.assign(n + 1, 0);
    for (int64_t d = 0; d < n; ++d) {
      int64_t c = 0, cf = 0;
      if (isdef(d)) {
        c = 1;
      } else {
        for (int64_t p = H->ptr[d]; p < H->ptr[d + 1]; ++p) c += !isdef(H->idx[p]);
        cf = H->ptr[d + 1] - H->ptr[d];
      }
      H->gs_ptr[d + 1] = H->gs_ptr[d] + c;
      H->f1_ptr[d + 1] = H->f1_ptr[d] + cf;
    }
    H->gs_idx.resize(H->gs_ptr[n]);
    H->vA1.resize(H->gs_ptr[n]);
    H->vA2.resize(H->gs_ptr[n]);
    H->f1_idx.resize(H->f1_ptr[n]);
    H->vF1.resize(H->f1_ptr[n]);
    for (int64_t d = 0; d < n; ++d) {
      int64_t q = H->gs_ptr[d], qf = H->f1_ptr[d];
      if (isdef(d)) {
        H->gs_idx[q] = (int32_t)d;
        H->vA1[q] = H->vA2[q] = 1.0;
        continue;
      }
      for (int64_t p = H->ptr[d]; p < H->ptr[d + 1]; ++p) {
        if (!isdef(H->idx[p])) {
          H->gs_idx[q] = H->idx[p];
          H->vA1[q] = vA1[p];
          H->vA2[q] = vA2[p];
          ++q;
        }
        H->f1_idx[qf] = H->idx[p];
        H->vF1[qf] = vM[p];
        ++qf;
      }
    }
    if (want_elastic) {
      // patterns of A_el (3n x 3n) and F_el (3n x 6n)
      H->el_ptr.assign(3 * n + 1, 0);
      H->fe_ptr.assign(3 * n + 1, 0);
      for (int da = 0; da < 3; ++da)
        for (int64_t d = 0; d <= n; ++d) {
          if (da > 0 && d == 0) continue;
          H->el_ptr[da * n + d] = da * nnzel + 3 * H->ptr[d];
          H->fe_ptr[da * n + d] = da * 2 * nnzel + 6 * H->ptr[d];
        }
      H->el_idx.resize((size_t)9 * nnz);
      H->fe_idx.resize((size_t)18 * nnz);
      parallel_rows(n, nthreads, [&](int64_t s, int64_t t) {
        for (int64_t d = s; d < t; ++d) {
          const int64_t r0 = H->ptr[d], len = H->ptr[d + 1] - r0;
          for (int da = 0; da < 3; ++da) {
            const int64_t ar = da * nnzel + 3 * r0, fr = da * 2 * nnzel + 6 * r0;
            for (int db = 0; db < 3; ++db)
              for (int64_t j = 0; j < len; ++j) H->el_idx[ar + db * len + j] = (int32_t)(db * n + H->idx[r0 + j]);
            for (int64_t j = 0; j < len; ++j)
              for (int v = 0; v < 6; ++v) H->fe_idx[fr + 6 * j + v] = 6 * H->idx[r0 + j] + v;
          }
        }
      });
      // anchors {0, n, 2n} (assemble_elasticity.py:461-527)
      const int64_t anc[3] = {0, n, 2 * n};
      for (int64_t r = 0; r < 3 * n && anchors; ++r)
        for (int64_t p = H->el_ptr[r]; p < H->el_ptr[r + 1]; ++p) {
          const int64_t c = H->el_idx[p];
          bool ra = (r == anc[0] || r == anc[1] || r == anc[2]);
          bool ca = (c == anc[0] || c == anc[1] || c == anc[2]);
          if (ra || ca) H->vAel[p] = (r == c) ? 1.0 : 0.0;
        }
      for (int a = 0; a < 3 && anchors; ++a)
        for (int64_t p = H->fe_ptr[anc[a]]; p < H->fe_ptr[anc[a] + 1]; ++p) H->vFel[p] = 0.0;
    }
    *out = H;
    return CPM_OK;
  } catch (const std::bad_alloc&) {
    return cpm::fail(CPM_ENOMEM, "host assembly: out of memory");
  } catch (const std::exception& e) {
    return cpm::fail(CPM_EINTERNAL, std::string("host assembly: ") + e.what());
  }
}

extern "C" int cpm_host_ops_array(cpm_host_ops* H, int which, void** ptr, int64_t* len) {
  if (!H || !ptr || !len) return cpm::fail(CPM_EINVAL, "null argument");
#define CPM_ARR(id, vec)            \
  case id:                          \
    *ptr = (void*)H->vec.data();    \
    *len = (int64_t)H->vec.size();  \
    return CPM_OK;
  switch (which) {
    CPM_ARR(CPM_HA_PTR, ptr)
    CPM_ARR(CPM_HA_IDX, idx)
    CPM_ARR(CPM_HA_ADEMAG, vK)
    CPM_ARR(CPM_HA_FX, vFx)
    CPM_ARR(CPM_HA_FY, vFy)
    CPM_ARR(CPM_HA_FZ, vFz)
    CPM_ARR(CPM_HA_GS_PTR, gs_ptr)
    CPM_ARR(CPM_HA_GS_IDX, gs_idx)
    CPM_ARR(CPM_HA_A1, vA1)
    CPM_ARR(CPM_HA_A2, vA2)
    CPM_ARR(CPM_HA_F1_PTR, f1_ptr)
    CPM_ARR(CPM_HA_F1_IDX, f1_idx)
    CPM_ARR(CPM_HA_F1, vF1)
    CPM_ARR(CPM_HA_EL_PTR, el_ptr)
    CPM_ARR(CPM_HA_EL_IDX, el_idx)
    CPM_ARR(CPM_HA_AEL, vAel)
    CPM_ARR(CPM_HA_FE_PTR, fe_ptr)
    CPM_ARR(CPM_HA_FE_IDX, fe_idx)
    CPM_ARR(CPM_HA_FEL, vFel)
    default:
      return cpm::fail(CPM_EINVAL, "unknown host array id");
  }
#undef CPM_ARR
}

extern "C" int cpm_host_ops_free(cpm_host_ops* H) {
  delete H;
  return CPM_OK;
}

extern "C" int cpm_host_element_matrices(const double* coords, int64_t n_nodes, const int32_t* elems,
                                         int64_t n_elems, int npe, int elem_stride, const double* C36,
                                         double* K, double* M, double* F, double* Kel, double* Fel) {
  if (!coords || !elems) return cpm::fail(CPM_EINVAL, "null argument");
  if (npe != 4 && npe != 8) return cpm::fail(CPM_EINVAL, "npe must be 4 or 8");
  const Tables T = make_tables(npe);
  double C[6][6] = {};
  if (C36)
    for (int i = 0; i < 36; ++i) C[i / 6][i % 6] = C36[i];
  const int nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
  for (int64_t e = 0; e < n_elems; ++e)
    for (int a = 0; a < npe; ++a)
      if (elems[e * elem_stride + a] < 0 || elems[e * elem_stride + a] >= n_nodes)
        return cpm::fail(CPM_EINVAL, "element node index out of range");
  parallel_rows(n_elems, nthreads, [&](int64_t s, int64_t t) {
    double x[8][3];
    Geo g;
    double k[8], m[8], f[3][8], kel[3][24], fel[3][48];
    for (int64_t e = s; e < t; ++e) {
      const int32_t* en = elems + e * elem_stride;
      for (int b = 0; b < npe; ++b)
        for (int j = 0; j < 3; ++j) x[b][j] = coords[(int64_t)en[b] * 3 + j];
      geometry(T, x, g);
      for (int a = 0; a < npe; ++a) {
        scalar_row(T, g, a, k, m, f);
        for (int b = 0; b < npe; ++b) {
          const int64_t o = (e * npe + a) * npe + b;
          if (K) K[o] = k[b];
          if (M) M[o] = m[b];
          if (F)
            for (int c = 0; c < 3; ++c) F[(int64_t)c * n_elems * npe * npe + o] = f[c][b];
        }
        if (Kel || Fel) {
          elastic_row(T, g, a, C, kel, fel);
          for (int da = 0; da < 3; ++da) {
            const int64_t row = e * 3 * npe + 3 * a + da;
            if (Kel)
              for (int j = 0; j < 3 * npe; ++j) Kel[row * 3 * npe + j] = kel[da][j];
            if (Fel)
              for (int j = 0; j < 6 * npe; ++j) Fel[row * 6 * npe + j] = fel[da][j];
          }
        }
      }
    }
  });
  return CPM_OK;
}
