/*
 * cupymag_hip.h -- C ABI of libcupymag_hip.so, the MI355X (gfx950) engine
 * behind the GSPM LLG time step of CuPyMag.
 *
 * The reference (CuPyMag v0.9.1) is pure Python on CuPy; it has no FFI.  Its
 * hot path is reached through a handful of Python call signatures, and each
 * entry point below replaces the CuPy/cuSPARSE work behind one of them
 * (citations are reference file:line).  The Python package `cupymag_amd`
 * keeps those signatures and calls this library through ctypes.
 *
 * Conventions
 *   - Every function returns int: CPM_OK (0) on success, a negative CPM_E*
 *     code on argument/HIP/memory errors, and for the CG solvers a positive
 *     value = CuPy's `info` (iterations done without convergence).
 *     cpm_last_error() returns the message of the last failure on the
 *     calling thread.
 *   - Vectors are device pointers to contiguous float64 arrays owned by the
 *     caller (PyTorch-ROCm tensors); the library never keeps a caller pointer
 *     past the call, except buffers handed to cpm_sim_* which belong to the
 *     simulation object.  Matrices and simulations are opaque handles owned
 *     by the library and must be destroyed explicitly.
 *   - All device work is enqueued on the context's stream (the caller's
 *     hipStream_t or the null stream).  Calls that
 *     return host scalars (CG iteration counts, sums, step reports)
 *     synchronise that stream, as the reference synchronises on them.
 *   - One context per host thread; a context is not re-entrant.
 *   - Layouts follow the reference: m is (n,3) row-major, u is [ux;uy;uz]
 *     (assemble_elasticity.py:391), E0 is 6n interleaved [6*i+v]
 *     (final_assembly.py:259-264), E is (n,6) row-major
 *     (compute_derivatives.py:125).
 */
#ifndef CUPYMAG_HIP_H
#define CUPYMAG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPM_ABI_VERSION 3

#define CPM_OK 0
#define CPM_EINVAL (-1)
#define CPM_ENOMEM (-2)
#define CPM_EHIP (-3)
#define CPM_EINTERNAL (-4)
#define CPM_ENODEV (-5)

/* ------------------------------------------------------------------ */
/* library / errors                                                    */
/* ------------------------------------------------------------------ */

/* Message of the last failed call on this thread ("" if none). */
const char* cpm_last_error(void);
/* CPM_ABI_VERSION of the built library. */
int cpm_abi_version(void);

/* ------------------------------------------------------------------ */
/* context                                                             */
/* ------------------------------------------------------------------ */
typedef struct cpm_ctx cpm_ctx;

/* Replaces CuPy's implicit current-device/default-stream state
 * (cupymag/__init__.py:21-24).  `stream` is the hipStream_t to enqueue on;
 * NULL is the device's null stream (PyTorch's default stream), so work
 * enqueued by the caller and by the library is ordered. */
int cpm_ctx_create(int device, void* stream, cpm_ctx** out);
int cpm_ctx_destroy(cpm_ctx* ctx);
int cpm_ctx_synchronize(cpm_ctx* ctx);

/* ------------------------------------------------------------------ */
/* sparse operators                                                    */
/* ------------------------------------------------------------------ */
typedef struct cpm_mat cpm_mat;

/* Upload a CSR operator from host arrays.  Replaces
 * cpx.csr_matrix(A_scipy, dtype=float_cp) in
 * cupymag/utils/final_assembly.py:176-177,200-201.  `values` holds
 * `nvalsets` arrays of length nnz back to back that share the pattern
 * (1, or 3 for the Fx/Fy/Fz derivative triple of
 * core/assemble_simulation_operators.py:56-63).  indptr is int64 on the
 * host; the device copy is int32 (nnz must fit). */
int cpm_mat_create_csr(cpm_ctx* ctx, int32_t nrows, int32_t ncols, int64_t nnz, const int64_t* indptr,
                       const int32_t* indices, const double* values, int nvalsets, cpm_mat** out);
int cpm_mat_destroy(cpm_mat* A);
int cpm_mat_shape(const cpm_mat* A, int32_t* nrows, int32_t* ncols, int64_t* nnz);

/* y = A_valset x.  Replaces `A @ x` on cupyx.scipy.sparse.csr_matrix
 * (cuSPARSE SpMV), e.g. cupymag/core/Micromagnetics.py:123,217-219. */
int cpm_spmv(cpm_ctx* ctx, const cpm_mat* A, int valset, const double* x, double* y);

/* Unpreconditioned CG with CuPy's stopping rule.  Replaces
 * cupyx.scipy.sparse.linalg.cg called from solve_cg,
 * cupymag/solvers/linear_solvers.py:59-62: x0 = 0 unless use_x0; returns b
 * (x = 0) when ||b|| = 0; atol = tol*||b||; at least one iteration; stop when
 * ||r|| <= atol after the update.  x is input (x0, read iff use_x0) and
 * output.  Returns 0 when converged, else CuPy's info (= iterations).
 * *iters and *resid (= ||r||, the recurrence residual) are always set. */
int cpm_cg(cpm_ctx* ctx, const cpm_mat* A, const double* b, double* x, int use_x0, double tol, int maxiter,
           int32_t* iters, double* resid);

/* Preconditioned CG with CuPy's recurrence for a given M.  Replaces
 * cupyx.scipy.sparse.linalg.cg(A, b, M=M, ...) as solve_cg forwards M,
 * cupymag/solvers/linear_solvers.py:59-62: each iteration z = M r,
 * rho = r.z, p = z (first) or z + (rho/rho_old) p, q = A p,
 * alpha = rho/(p.q), x += alpha p, r -= alpha q, stop when ||r|| <= atol.
 * M is an n x n operator uploaded like A (cpm_mat_create_csr).  Same
 * x0, zero-b, return-value and *iters / *resid conventions as cpm_cg. */
int cpm_cg_precond(cpm_ctx* ctx, const cpm_mat* A, const cpm_mat* M, const double* b, double* x, int use_x0,
                   double tol, int maxiter, int32_t* iters, double* resid);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  cpm_cg on the same arguments with the same kernels,
 * results and return values, but the solve's vectors, reduction partials and
 * CG state live in a fixed device array of the library (one per GPU, held
 * until the process ends; its address is looked up when a context is
 * created), so the call allocates nothing.  At most 1,048,576 rows (larger
 * A: CPM_EINVAL, use cpm_cg).  Calls that overlap on one GPU share that
 * array and race; each call starts from a cleared counter. */
int cpm_cg_noalloc(cpm_ctx* ctx, const cpm_mat* A, const double* b, double* x, int use_x0, double tol, int maxiter,
                   int32_t* iters, double* resid);

/* Kernel timing (no counterpart in the reference, which has only wall-clock
 * prints, Micromagnetics.py:49,106,348-351): with `on`, every CG SpMV launch
 * is bracketed by HIP events on the context stream and its duration is
 * accumulated on the matrix; cpm_mat_prof reads (and optionally resets) the
 * total milliseconds and the number of launches that did work. */
int cpm_prof_enable(cpm_ctx* ctx, int on);
int cpm_mat_prof(cpm_mat* A, double* total_ms, int64_t* launches, int reset);
/* No counterpart in the reference (measurement).  Restrict the timing of
 * cpm_prof_enable to the launches on operator A (NULL: every operator), so
 * that the other solves run as they do without profiling. */
int cpm_prof_only(cpm_ctx* ctx, const cpm_mat* A);
/* No counterpart in the reference (measurement).  With profiling on, the
 * outer FP32 product of the mixed-precision multigrid PCG on this (level-0)
 * operator is timed by its dispatch in the first iteration of every
 * iteration batch; read (and optionally reset) like cpm_mat_prof. */
int cpm_mat_prof_outer(cpm_mat* A, double* total_ms, int64_t* launches, int reset);

/* Attach a SELL-64 form (sliced ELL, slices of 64 rows, FP64 and FP32 values)
 * to a square operator from the same host CSR arrays it was created from.
 * bs = 3 groups rows/columns {i, n+i, 2n+i} into 3x3 blocks (the elasticity
 * operator, DOF order of assemble_elasticity.py:391) and its solver works on
 * node-major interleaved vectors.  grid_dims (3 ints, may be NULL): the DOF
 * grid of a structured periodic mesh; when the pattern is a fixed
 * neighbour stencil of that grid the column indices are not stored. */
int cpm_mat_add_sell(cpm_mat* A, int bs, const int64_t* indptr, const int32_t* indices, const double* values,
                     const int32_t* grid_dims);
/* SELL form in use: block size, stencil = 1 if index-free (grid_dims given
 * to cpm_mat_add_sell, DOF (i,j,k) numbered (i*ny + j)*nz + k on a periodic
 * grid, and every block column a neighbour at an offset in {-1,0,1}^3), 2
 * if those offsets are the 14 neighbours of a Kuhn mesh (decoded at compile
 * time by the kernels), 3 if besides the operator is symmetric slot by slot
 * and its FP32 form keeps half of the stencil (each off-diagonal block once,
 * read transposed for the mirrored neighbour), 4 if moreover it is
 * symmetric to rounding (1e-14 of its largest entry) and the FP64 products
 * of the solvers read a half copy of the FP64 form the same way, 0 if
 * indexed; and the number of 64-row slots (of the full FP64 form). */
int cpm_mat_sell_info(const cpm_mat* A, int* bs, int* stencil, int64_t* slots);
/* CG on the SELL form for nrhs interleaved right-hand sides (b[i*nrhs + j];
 * nrhs = 3 only for bs = 1; bs = 3 vectors are node-major interleaved).
 * precision 64: CuPy's CG recurrence per column (as cpm_cg).  precision 32:
 * FP32 CG (FP64 dot accumulation) refined in FP64 by reliable updates: when
 * the FP32 recurrence residual has dropped by `delta` the correction is added
 * to the FP64 solution and the residual replaced by the FP64 b - A x; stops
 * when that FP64 residual is <= tol*||b||.
 * iters/resid get nrhs entries.  Returns 0 or a positive count when a
 * column did not converge within maxiter. */
int cpm_cg_sell(cpm_ctx* ctx, cpm_mat* A, int nrhs, const double* b, double* x, int use_x0, double tol, int maxiter,
                int precision, double delta, int32_t* iters, double* resid);
/* y = A x on the SELL form (FP64), nrhs interleaved right-hand sides. */
int cpm_spmv_sell(cpm_ctx* ctx, const cpm_mat* A, int nrhs, const double* x, double* y);

/* ------------------------------------------------------------------ */
/* SpMV autotuning                                                     */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference (cuSPARSE picks its own kernel behind
 * `A @ x`, final_assembly.py:176-177).  The launch schedule of an operator:
 * CSR lanes per row (64/lanes rows per wavefront) and workgroup cap, SELL
 * workgroup cap, and the format of the standalone product y = A x of a
 * scalar operator (spmv_sell = 1: the FP64 SELL form; results equal to
 * rounding).  Every solver and product reads it. */
typedef struct cpm_schedule {
  int32_t lanes;          /* CSR: threads per row, power of 2 in [1, 64] */
  int32_t grid_cap;       /* CSR: at most this many workgroups per launch */
  int32_t sell_grid_cap;  /* SELL: at most this many workgroups per launch (0: no SELL form) */
  int32_t spmv_sell;      /* cpm_spmv on value set 0 runs the SELL form */
} cpm_schedule;

#define CPM_FMT_CSR 0
#define CPM_FMT_SELL64 1
#define CPM_FMT_SELL32 2

typedef struct cpm_tune_entry {
  int32_t format;   /* CPM_FMT_* */
  int32_t lanes;    /* CSR lanes per row (64 for SELL: one row per lane) */
  int32_t grid_cap; /* workgroup cap */
  float us;         /* average microseconds of the SpMV + its dot reduction */
} cpm_tune_entry;

/* Time every candidate schedule of A on its context's stream (SpMV with its
 * fused dot partials + the reduction of those partials, as one CG
 * iteration runs it) and keep the fastest: CSR lanes x cap, SELL cap (FP32
 * form, the mixed solvers' hot kernel), and the product format.  Writes up
 * to max_out timings to out (may be NULL) and their count to n_out.
 * Synchronises the stream. */
int cpm_mat_autotune(cpm_mat* A, cpm_tune_entry* out, int max_out, int* n_out);
int cpm_mat_set_schedule(cpm_mat* A, const cpm_schedule* s);
int cpm_mat_get_schedule(const cpm_mat* A, cpm_schedule* s);

/* ------------------------------------------------------------------ */
/* field operators                                                     */
/* ------------------------------------------------------------------ */

/* H[k*n + i] = -(F_k U)_i, k = x,y,z on the 3-valset operator F3.
 * Replaces ComputeDerivatives.compute_Hd_from_U,
 * cupymag/utils/compute_derivatives.py:36-58. */
int cpm_hd_from_u(cpm_ctx* ctx, const cpm_mat* F3, const double* U, double* H);

/* Strain (n,6) [E11,E22,E33,E12,E23,E13] from u = [ux;uy;uz] by the weak-form
 * derivative triple F3, optionally mixed by the 3x3 rotation R (host array,
 * row-major; NULL = no rotation).  Replaces
 * ComputeDerivatives.compute_E_from_u, compute_derivatives.py:60-125. */
int cpm_strain_from_u(cpm_ctx* ctx, const cpm_mat* F3, const double* u, const double* R, double* E);

/* Spontaneous strain E0 (6n, interleaved) of m (n,3).  With R and Mv (host,
 * row-major 3x3 and 6x6) non-NULL, m is first rotated m100 = m R and the
 * Voigt vector mapped E0 <- E0 Mv^T (Micromagnetics.py:112-119).  Replaces
 * build_E0_from_m, cupymag/utils/final_assembly.py:208-266. */
int cpm_build_e0(cpm_ctx* ctx, int32_t n, double lambda100, double lambda111, const double* m, const double* R,
                 const double* Mv, double* E0);

/* out[c] = sum_i W[i] f[i*ncomp + c] (host out, ncomp <= 8), FP64
 * accumulation.  With W the precomputed per-DOF quadrature weights, out/V
 * (V = total weight) is VolumeAverage.compute_average_field_gpu,
 * volume_average.py:155-216. */
int cpm_weighted_sum(cpm_ctx* ctx, int32_t n, int ncomp, const double* W, const double* f, double* out);

/* ------------------------------------------------------------------ */
/* multigrid-preconditioned CG (structured periodic Kuhn meshes)        */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference, whose solve_cg passes M=None
 * (linear_solvers.py:20,59-62): an optional preconditioner for the slowly
 * converging solves (elasticity, demag, the Gauss-Seidel operator A1).
 * levels[0] is the operator (its index-free SELL form), levels[l+1] the
 * same operator assembled on the coarse Kuhn mesh of level l: ceil(n/2)
 * cells per dimension (the even nodes for even n, the regular mesh of the
 * same box for odd n); the last level is solved with coarse_inv (dense
 * n x n, row-major, in the SELL vector order: node-major interleaved for
 * bs 3).  Transfers: P1 interpolation (nested for even n, at the nominal
 * node positions for odd n) and its transpose.  When row 0 of levels[0] is
 * an identity row (the anchor of A_demag and A_el), the coarse levels are
 * either all anchored the same way (node 0 is then left out of both
 * transfers) or all the singular periodic stiffness K, with coarse_inv its
 * pseudo-inverse: a floating hierarchy, whose V-cycle V solves the periodic
 * problem (level 0 smooths with K, recovered on the device from levels[0]
 * by K's annihilation of the translations) and whose preconditioner is
 * F^T V(F r) + e_0 r_0, F = I - e_0 1^T per component (one column only).
 * Smoother: Chebyshev of `degree` on D^-1 A over [lmax/ratio, lmax],
 * lmax = 1.1 x a power-iteration estimate.  precision: the arithmetic of
 * the V-cycle (the outer CG is FP64): 64, 32, or 16 = FP32 vectors with the
 * level operators' off-diagonal values stored in FP16 (scaled by a power of
 * two, diagonal blocks in FP64 compensated for the rounding).  The levels
 * are borrowed and must outlive the hierarchy. */
typedef struct cpm_mg cpm_mg;
int cpm_mg_create(cpm_ctx* ctx, int nlevels, cpm_mat* const* levels, const double* coarse_inv, int precision,
                  int degree, double ratio, cpm_mg** out);
/* No counterpart in the reference (as cpm_mg_create).  The same hierarchy
 * for an operator without an anchor (the Gauss-Seidel operator A1 = M +
 * alpha1 K, whose coarse levels are well conditioned): no dense inverse; the
 * last level is solved by `cdegree` (1..8) Chebyshev steps on D^-1 A over
 * [lmax/cratio, lmax], the first fused with the restriction.  An anchored
 * levels[0] is refused (CPM_EINVAL). */
int cpm_mg_create_smoothed(cpm_ctx* ctx, int nlevels, cpm_mat* const* levels, int precision, int degree,
                           double ratio, int cdegree, double cratio, cpm_mg** out);
/* No counterpart in the reference (as cpm_mg_create).  Chebyshev degree of
 * the levels between the finest and the coarsest (0: the finest level's
 * degree on every level; 2..8, or 1 when the finest level's is 1): the
 * coarse levels cost a fraction of the finest, and a higher degree there
 * can save outer iterations.  Recomputes the coefficients and drops the
 * captured iteration chunks; batched solves need the same degrees. */
int cpm_mg_set_coarse_degree(cpm_mg* mg, int degree);
int cpm_mg_destroy(cpm_mg* mg);
/* levels and the lambda_max estimate of each level (host, up to max_levels) */
int cpm_mg_info(const cpm_mg* mg, int* nlevels, double* lmax, int max_levels);
/* storage of each level's FP16 form (host, up to max_levels): 1 = half of a
 * slot-symmetric Kuhn stencil (forward offsets only, the backward blocks read
 * transposed from the neighbour rows), 0 = every slot, or no FP16 form.  No
 * counterpart in the reference (its levels are CSR matrices). */
int cpm_mg_forms(const cpm_mg* mg, int* half_stencil, int max_levels);
/* z = the preconditioner applied to r: one V-cycle (floating hierarchy:
 * F^T V(F r) + e_0 r_0); device pointers, SELL vector order; synchronises */
int cpm_mg_apply(cpm_mg* mg, const double* r, double* z);
/* CG on levels[0] preconditioned by one V-cycle per iteration, FP64, on
 * the context's stream: CuPy's stopping rule on the recurrence residual
 * (||r|| <= tol*||b|| after the update, at least one iteration, b = 0 ->
 * x = 0), flexible beta = (r - r_old).z / rho_old.  nrhs = 1, or 3 for a
 * scalar operator: b and x hold nrhs interleaved columns (entry i*nrhs + j),
 * each with its own scalars and stopping test; iters and resid hold nrhs
 * entries.  delta > 0: the outer products use the FP32 form of the
 * operator on FP32 search directions (x and r stay FP64), with an FP64
 * residual replacement whenever the recurrence
 * residual drops by delta and before stopping (the stopping test is then
 * on the FP64 residual); 0: FP64 products.  Returns 0, or a positive
 * iteration count when maxiter was reached. */
/* No counterpart in the reference.  Iterations of cpm_cg_mg (and of the
 * multigrid solves of cpm_sim_step) are replayed from captured chunks
 * (hipGraph, 2 iterations each, cached per hierarchy for the vectors they
 * were captured with) unless turned off here; the results are the same
 * bit for bit.  Default on. */
int cpm_mg_set_graphs(cpm_mg* mg, int on);
int cpm_cg_mg(cpm_mg* mg, int nrhs, const double* b, double* x, int use_x0, double tol, int maxiter,
              double delta, int32_t* iters, double* resid);
/* No counterpart in the reference.  cpm_cg_mg on `count` independent
 * systems at once (hierarchies of one block size, V-cycle precision and
 * smoother degree, each at most once, all on ctx) with nrhs columns each:
 * one launch per phase for the whole batch, V-cycles level by level
 * whatever the depths.  Each system's result, iteration counts and
 * residuals (iters, resid: nrhs entries per system) equal cpm_cg_mg's on
 * the same input bit for bit.  failed[j] = 1 when a column of system j
 * reached maxiter; returns 0 or an error < 0. */
int cpm_cg_mg_batch(cpm_ctx* ctx, int count, int nrhs, cpm_mg* const* mgs, const double* const* b,
                    double* const* x, const int32_t* use_x0, double tol, int maxiter, int32_t* iters,
                    double* resid, int32_t* failed);

/* ------------------------------------------------------------------ */
/* matrix-free magnetoelastic right-hand side                           */
/* ------------------------------------------------------------------ */
/* b = F_el E0 without the F_el matrix (assemble_elasticity.py:268-348,
 * 507-527, used at Micromagnetics.py:123), for an n^3 periodic Kuhn mesh
 * whose (n+1)^3 node coordinates are in kuhn_mesh order: every element's
 * geometry is recomputed from the coordinates with the host assembler's
 * formulas (signed detJ, 7-digit 4-point table, zero element when
 * |detJ| < 1e-14).  C: the 6x6 stiffness (rotated when rot111), row-major.
 * E0: 6n interleaved; b: 3n [bx; by; bz], rows 0, n, 2n zero.  Results equal
 * the CSR product up to summation order.  Saves the 270 entries per row
 * triple of F_el (6.8 GB at 128^3).  On a regular lattice (every node at
 * origin + (i hx, j hy, k hz)) the element terms are the same in every cube,
 * so the product is one 27-point stencil of 3 x 6 blocks, summed once on the
 * host at creation from the same element formulas. */
typedef struct cpm_fel cpm_fel;
int cpm_fel_create(cpm_ctx* ctx, int32_t n, const double* coords, const double* C, cpm_fel** out);
int cpm_fel_destroy(cpm_fel* fel);
int cpm_fel_apply(cpm_fel* fel, const double* E0, double* b);

/* The weak-form derivative operators Fx, Fy, Fz (assemble_demag.py:235-305)
 * of the same mesh without their CSR form (0.9 GB at 128^3), element values
 * formed as the host assembler forms them (C of cpm_fel_create may be NULL
 * when only these are used).  Replace, in that order,
 * ComputeDerivatives.compute_E_from_u (compute_derivatives.py:60-125: u 3n
 * [ux;uy;uz], R host 3x3 or NULL, E (n,6) [E11,E22,E33,E12,E23,E13]),
 * compute_Hd_from_U (:36-58: H (3,n) = -F U) and a product Fk @ x (k = 0,1,2
 * for Fx, Fy, Fz).  Results equal the CSR products up to summation order. */
int cpm_fel_strain(cpm_fel* fel, const double* u, const double* R, double* E);
int cpm_fel_hd(cpm_fel* fel, const double* U, double* H);
int cpm_fel_dmul(cpm_fel* fel, int k, const double* x, double* y);

/* ------------------------------------------------------------------ */
/* device-resident GSPM LLG simulation (the hot loop)                  */
/* ------------------------------------------------------------------ */
typedef struct cpm_sim cpm_sim;

typedef struct cpm_sim_ops {
  const cpm_mat* A_demag; /* n x n, anchored at DOF 0 (assemble_demag.py:307-341) */
  const cpm_mat* F3;      /* Fx,Fy,Fz on one pattern (assemble_demag.py:235-305) */
  const cpm_mat* A1;      /* M + alpha1 K, defect surgery (assemble_simulation_operators.py:81) */
  const cpm_mat* A2;      /* M + alpha2 K (:82) */
  const cpm_mat* F1;      /* M, defect rows removed (:83) */
  const cpm_mat* A_el;    /* 3n x 3n or NULL when magnetoelastic coupling is off */
  const cpm_mat* F_el;    /* 3n x 6n or NULL */
  cpm_mg* mg_el;          /* multigrid hierarchy of A_el (used when params.precond) or NULL */
  cpm_mg* mg_demag;       /* multigrid hierarchy of A_demag (used when params.precond) or NULL */
  cpm_fel* fel;           /* matrix-free F_el (cpm_fel_create); used instead of F_el when not NULL */
  cpm_mg* mg_a1;          /* multigrid hierarchy of A1 (used when params.precond has CPM_PRECOND_GS) or NULL */
  cpm_fel* fder;          /* matrix-free Fx, Fy, Fz (cpm_fel of this mesh): strain, Htilde and the demag right-hand
                             side without F3 (which may then be NULL), or NULL */
} cpm_sim_ops;

#define CPM_PRECOND_EL_DEMAG 1 /* elasticity and demag solves: multigrid-preconditioned CG */
#define CPM_PRECOND_GS 2       /* Gauss-Seidel solves on A1 (g1n..g3n, g1*, g2*): multigrid-preconditioned CG */

typedef struct cpm_sim_params {
  int32_t n;                  /* number of DOFs */
  int32_t me;                 /* magnetoelastic coupling on */
  int32_t rot111;             /* crystal x axis along [111] */
  int32_t maxiter;            /* CG max iterations */
  double tol;                 /* CG relative tolerance */
  double dt;                  /* reduced time step */
  double alpha_damp;          /* Gilbert damping */
  double K1;                  /* reduced cubic anisotropy */
  double elasC[3];            /* reduced magnetoelastic constants (zero when me == 0) */
  double lambda100, lambda111;
  double Nd[3];               /* demagnetisation factors */
  double Ebar_sigma[6];       /* strain of the external stress (sigma_matrices.py:25-45) */
  double R[9];                /* rot_111_matrices.get_R_matrix (row-major) */
  double Mv[36];              /* rot_111_matrices.get_M_matrix (row-major) */
  double MvinvT[36];          /* inv(Mv)^T (Micromagnetics.py:54) */
  int32_t mode;               /* CPM_MODE_* */
  int32_t warm_start;         /* fast modes: 1 = start every CG from the previous step's solution
                                 (the reference does so only when use_init is set); 2 = from the
                                 linear extrapolation in time of the last two steps' solutions
                                 (2 x_n - x_{n-1}; 12 n doubles more device memory) */
  double rel_delta;           /* CPM_MODE_MIXED: FP64 residual update when the FP32 recurrence
                                 residual drops by this factor (reliable updates) */
  int32_t precond;            /* fast modes: CPM_PRECOND_* bits; the solves named run multigrid-
                                 preconditioned CG (cpm_cg_mg) on the hierarchies in cpm_sim_ops */
  int32_t overlap;            /* fast modes with multigrid elasticity and demag solves: the demag solve
                                 runs beside the elasticity solve, on the context's second stream
                                 (same results bit for bit; single simulations, not batches) */
  int32_t reserved[2];
  double rel_delta_gs;        /* CPM_MODE_MIXED: rel_delta of the Gauss-Seidel solves (A1, A2);
                                 0 = rel_delta */
  double rel_delta_mg;        /* CPM_MODE_MIXED with precond: the multigrid solves' outer products use
                                 the FP32 operator, with an FP64 residual replacement whenever the
                                 recurrence residual drops by this factor and before stopping;
                                 0 = FP64 products */
} cpm_sim_params;

#define CPM_MODE_FP64 0   /* plain FP64 path: CSR operators, the reference's sequence of solves */
#define CPM_MODE_FAST64 1 /* FP64, SELL-64 operators, 3-RHS solves for g1n..g3n and m1**..m3** */
#define CPM_MODE_MIXED 2  /* as FAST64 with FP32 CG refined in FP64 (reliable updates) */

/* Report of one step (cpm_sim_step). */
typedef struct cpm_step_report {
  double sumsq;       /* sum (m - m_prev)^2, Micromagnetics.py:299 */
  double avg_m[3];    /* volume average of the new m, :306 */
  int32_t iters[10];  /* CG iterations: elasticity, demag, g1n, g2n, g3n, g1*, g2*, m1**, m2**, m3** */
  int32_t failed;     /* index+1 of the solve that did not converge, 0 if all converged */
  double resid;       /* recurrence residual of the failed solve */
  /* with a hysteresis control attached (cpm_sim_control), filled from its
   * device state in the step's one scalar transfer: */
  int32_t point_done; /* the step completed a field point (LLG convergence, or the per-point cap) */
  int32_t state;      /* 0 running, 1 stop condition met (Micromagnetics.py:109), 2 field list done */
  int32_t point;      /* index of the current field point (points completed) */
  int32_t nstep;      /* steps at the current point (0 right after a point) */
  int32_t use_init;   /* use_init of the next step (:300) */
  int32_t pad;
  double H[3];        /* field of this step (reduced units) */
  double H_next[3];   /* field of the next step */
} cpm_step_report;

/* Hysteresis control of a simulation (Micromagnetics.py:109,299-335), kept
 * on the device and evaluated by a kernel after every step: use_init,
 * LLG convergence (sumsq < llg_accuracy after more than llg_min_step steps
 * at a point), the field step (H -= dH, or the next of `npoints` listed
 * fields), the stop condition (avg_m . stop_dir <= stop_value) and an
 * optional cap of steps per point.  Field and use_init of the steps then
 * come from this state (the H and use_init arguments of cpm_sim_step /
 * cpm_batch_step are ignored for the simulation), and the host learns the
 * outcome from the step report without a further transfer. */
typedef struct cpm_sweep_cfg {
  double H0[3];          /* first field (npoints = 0) */
  double dH[3];          /* field step on convergence */
  double llg_accuracy;   /* nDOF * llg_accuracy_factor (:88) */
  double initial_factor; /* use_init when sumsq < initial_factor * llg_accuracy (:300) */
  double stop_dir[3];
  double stop_value;
  int32_t llg_min_step;  /* 10 with magnetoelastic coupling, else 100 (:83-86) */
  int32_t steps_per_point; /* 0: no cap */
  int32_t npoints;       /* > 0: fields (npoints x 3, reduced) instead of H0 and dH */
  int32_t pad;
} cpm_sweep_cfg;

/* No counterpart as a call in the reference (its loop body, :109,299-335).
 * Attach (cfg) or detach (cfg NULL) the control; the step counters start at 0. */
int cpm_sim_control(cpm_sim* sim, const cpm_sweep_cfg* cfg, const double* fields);

/* Create a simulation on `ops` (borrowed; they must outlive the sim).
 * W: device pointer to the n per-DOF quadrature weights and `volume` their
 * total, so that avg = (W . f) / volume restates
 * VolumeAverage.compute_average_field_gpu (volume_average.py:155-216) as one
 * dot product; W is copied.  defect: host array of n bytes, 1 on defect
 * DOFs (extract_defect_dofs, final_assembly.py:27-57). */
int cpm_sim_create(cpm_ctx* ctx, const cpm_sim_ops* ops, const cpm_sim_params* prm, const double* W,
                   double volume, const uint8_t* defect, cpm_sim** out);
int cpm_sim_destroy(cpm_sim* sim);

#define CPM_FIELD_M 0       /* (n,3) */
#define CPM_FIELD_HTILDE 1  /* (3,n) component-major: Htilde1, Htilde2, Htilde3 */
#define CPM_FIELD_E 2       /* (n,6) */
#define CPM_FIELD_U 3       /* demag potential (n) */
#define CPM_FIELD_UEL 4     /* displacement [ux;uy;uz] (3n) */

/* Set m (device pointer, (n,3)); zeroes defect DOFs, computes avg_m
 * (initialize_m + the first average, Micromagnetics.py:70-73) and returns
 * it in avg_m (host, 3). */
int cpm_sim_set_m(cpm_sim* sim, const double* m, double* avg_m);
/* Copy a field (CPM_FIELD_*) into a caller device buffer. */
int cpm_sim_get(cpm_sim* sim, int field, double* dst);
/* One GSPM LLG step at external field H (reduced units), with the
 * warm-start flag use_init (Micromagnetics.py:110-307).  Returns 0, or the
 * positive CG info of the first solve that did not converge (the state is
 * then that of the failing solve, as when the reference raises). */
int cpm_sim_step(cpm_sim* sim, const double* H, int use_init, cpm_step_report* rep);

/* ------------------------------------------------------------------ */
/* batches: many simulations (any mesh sizes) stepped together          */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference, which runs one simulation per process
 * (Micromagnetics.py:34).  A batch steps all of its simulations through
 * the same stages of the step; each of the step's solves runs as one
 * batched CG over every member (one launch per CG phase for the whole
 * batch), so small meshes share the launches of the large ones.  Members
 * must use the batch's context and one of the SELL modes (all the same).
 * Members created with params.precond run their elasticity and demag
 * solves as one batched multigrid-preconditioned CG (V-cycles of all
 * hierarchies level by level, any depths) and the others as batched plain
 * CG; members that share a hierarchy are solved in turn.
 * A member's arithmetic does not depend on the other members: its results
 * are bit-identical alone and in any batch.  Members can be added and
 * removed between steps (continuous batching); a member cannot be stepped
 * with cpm_sim_step while in a batch.  Destroying a member removes it. */
typedef struct cpm_batch cpm_batch;

#define CPM_BATCH_GRAPH 1 /* replay captured chunks of CG iterations (hipGraph) */

int cpm_batch_create(cpm_ctx* ctx, int flags, cpm_batch** out);
int cpm_batch_destroy(cpm_batch* batch);
/* Allocates the member's batched-CG workspace (3n FP64 + 4 x 3n FP32 for
 * the mixed mode, 3 x 3n FP64 for FAST64) on first use. */
int cpm_batch_add(cpm_batch* batch, cpm_sim* sim);
int cpm_batch_remove(cpm_batch* batch, cpm_sim* sim);
int cpm_batch_size(const cpm_batch* batch, int32_t* n);
/* One step of every member, in the order they were added: H (host, 3 per
 * member, reduced units), use_init (host, 1 per member), reps (host, 1 per
 * member).  A member whose solve did not converge has reps[j].failed set
 * and its step carried on; the call still returns 0 (errors < 0 only). */
int cpm_batch_step(cpm_batch* batch, const double* H, const int32_t* use_init, cpm_step_report* reps);

/* ------------------------------------------------------------------ */
/* host-side assembly (CPU, once per mesh)                             */
/* ------------------------------------------------------------------ */
typedef struct cpm_host_ops cpm_host_ops;

/* Direct-to-CSR FEM assembly on the host.  Replaces the Numba assemblers
 * AssembleDemag / AssembleGaussSeidel / AssembleElasticity
 * (cupymag/physics/assemble_*.py) plus the COO->CSR and defect surgery of
 * cupymag/utils/final_assembly.py:59-205 and
 * core/assemble_simulation_operators.py:42-107.
 *   elems: (n_elems, elem_stride) int32, first npe columns are node ids.
 *   defect_dof: n bytes (1 = defect DOF) or NULL.
 *   coef: [alpha1, alpha2, C_voigt (36, row-major, already rotated)].
 *   flags: CPM_HA_ELASTIC assembles A_el and F_el as well;
 *   CPM_HA_NO_ANCHORS leaves out the anchors (A_demag row/column 0 and
 *   A_el rows/columns {0, n, 2n} stay the periodic stiffness, F_el rows
 *   {0, n, 2n} are kept): the singular operators that the floating
 *   multigrid levels use (no counterpart in the reference).
 * Arrays are read back with cpm_host_ops_array. */
#define CPM_HA_ELASTIC 1
#define CPM_HA_NO_ANCHORS 2
int cpm_host_assemble(const double* coords, int64_t n_nodes, const int32_t* elems, int64_t n_elems, int npe,
                      int elem_stride, const int32_t* gid, int32_t ndof, const uint8_t* defect_dof,
                      const double* coef, int flags, int nthreads, cpm_host_ops** out);

#define CPM_HA_PTR 0      /* int64 [n+1] pattern of A_demag, Fx, Fy, Fz */
#define CPM_HA_IDX 1      /* int32 */
#define CPM_HA_ADEMAG 2   /* double */
#define CPM_HA_FX 3
#define CPM_HA_FY 4
#define CPM_HA_FZ 5
#define CPM_HA_GS_PTR 6   /* pattern of A1/A2 after defect surgery */
#define CPM_HA_GS_IDX 7
#define CPM_HA_A1 8
#define CPM_HA_A2 9
#define CPM_HA_F1_PTR 10  /* pattern of F1 (defect rows removed) */
#define CPM_HA_F1_IDX 11
#define CPM_HA_F1 12
#define CPM_HA_EL_PTR 13  /* A_el, 3n x 3n */
#define CPM_HA_EL_IDX 14
#define CPM_HA_AEL 15
#define CPM_HA_FE_PTR 16  /* F_el, 3n x 6n */
#define CPM_HA_FE_IDX 17
#define CPM_HA_FEL 18
int cpm_host_ops_array(cpm_host_ops* h, int which, void** ptr, int64_t* len);
int cpm_host_ops_free(cpm_host_ops* h);

/* Dense element matrices (any output may be NULL):
 *   K, M: (n_elems, npe, npe); F: (3, n_elems, npe, npe);
 *   Kel: (n_elems, 3npe, 3npe); Fel: (n_elems, 3npe, 6npe) using C36.
 * The per-element integrals of assemble_demag.py:95-184,
 * assemble_Gauss_Seidel.py:88-160, assemble_elasticity.py:194-348; the
 * building block of the reference's COO API (build_coo_matrix*_numba). */
int cpm_host_element_matrices(const double* coords, int64_t n_nodes, const int32_t* elems, int64_t n_elems, int npe,
                              int elem_stride, const double* C36, double* K, double* M, double* F, double* Kel,
                              double* Fel);

#ifdef __cplusplus
}
#endif
#endif /* CUPYMAG_HIP_H */
