/*
 * rbfx.h -- C ABI of the MI355X (gfx950) dense RBF-collocation library.
 *
 * Drop-in boundary for the hot path of Updes' pde_solver / pde_solver_jit
 * (updes/operators.py:559-683).  The reference has no FFI of its own: its
 * boundary is the Python API, and every arithmetic step runs inside JAX/XLA
 * or lineax.  Each entry point below replaces one of those call sites (cited
 * per function).  Conventions:
 *   - every pointer argument named d_* is DEVICE memory owned by the caller;
 *     the library allocates no device memory and keeps no pointer after return;
 *   - matrices are ROW-MAJOR with a leading dimension ld (elements);
 *   - every call is asynchronous on the caller's stream (hipStream_t passed as
 *     void*; NULL = default stream) unless documented as synchronising;
 *   - the return value is RBFX_OK (0) or a negative error code; the message of
 *     the last error of the calling thread is available from rbfx_last_error();
 *   - no C++ exception crosses the ABI.
 */
#ifndef RBFX_H
#define RBFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RBFX_ABI_VERSION 3

enum rbfx_status {
  RBFX_OK = 0,
  RBFX_ERR_ARG = -1,       /* invalid argument / shape                       */
  RBFX_ERR_HIP = -2,       /* a HIP runtime call failed                      */
  RBFX_ERR_SINGULAR = -3,  /* exact zero pivot met (factors still written)   */
  RBFX_ERR_TIMEOUT = -4,   /* a bounded device-side wait expired             */
  RBFX_ERR_STALL = -5,     /* iterative refinement stalled                   */
};

/* RBF kernels (updes/utils.py:30-69). */
enum rbfx_kernel {
  RBFX_POLYHARMONIC = 0,      /* r^(2a+1), param = a            */
  RBFX_GAUSSIAN = 1,          /* exp(-(eps r)^2), param = eps   */
  RBFX_THIN_PLATE = 2,        /* r^(2a) log r, param = a        */
  RBFX_MULTIQUADRIC = 3,      /* sqrt(1+(eps r)^2)              */
  RBFX_INV_MULTIQUADRIC = 4,  /* 1/sqrt(1+(eps r)^2)            */
};

/* Operator basis of a lowered nodal operator. */
enum rbfx_op { RBFX_VAL = 0, RBFX_DX, RBFX_DY, RBFX_LAP, RBFX_DXX, RBFX_DYY, RBFX_DXY, RBFX_NOPS };

/* Which matrix rbfx_assemble / rbfx_matvec describe. */
enum rbfx_matrix {
  RBFX_MAT_K = 0,  /* collocation matrix [[opPhi opP];[bdPhi bdP];[P^T 0]]  */
  RBFX_MAT_A = 1,  /* interpolation matrix [[Phi P];[P^T 0]]                */
};

/*
 * A collocation problem on N sorted nodes (updes/cloud.py:403-408 ordering:
 * internal | dirichlet | neumann | robin | periodic groups).
 * Row classes of K (updes/assembly.py:141-362, 366-401):
 *   [0, Ni)                 interior rows, operator descriptor op_rbf/op_mon
 *   [Ni, Ni+Nd)             Dirichlet rows      phi
 *   [.., +Nn)               Neumann rows        grad phi . n
 *   [.., +Nr)               Robin rows          beta phi + grad phi . n
 *   [.., +Npd)              periodic value rows phi(x_i1,.) - phi(x_i2,.)
 *   [.., +Npd)              periodic flux rows  grad phi(x_i1,.).n1 + grad phi(x_i2,.).n2
 *   [N, N+M)                P^T rows
 * The descriptor coefficients are op_*_const[t] + (op_*_row[t] ? op_*_row[t][i] : 0).
 *
 * Local supports (updes/cloud.py:83-112 with support_size < N; the scatters
 * of updes/assembly.py:31-32, 127-128, 171-258): with RBFX_FLAG_LOCAL_SUPPORT
 * in flags, the RBF columns of node row i are restricted to the S =
 * support_cols node ids in d_support[i*S .. i*S+S) (ascending, without i),
 * and periodic rows to the union of the supports of both nodes of the pair.
 * Monomial columns and the P^T rows stay full.  Without the flag every node
 * but the row's own is a column (support_size = "max").
 */
enum rbfx_flags { RBFX_FLAG_LOCAL_SUPPORT = 1 };

typedef struct rbfx_problem {
  int64_t N;            /* number of nodes                                   */
  int64_t M;            /* number of monomials, 0..15                        */
  int32_t kernel;       /* enum rbfx_kernel                                  */
  int32_t flags;        /* enum rbfx_flags bits, or 0                        */
  double param;         /* a (polyharmonic / thin plate) or eps              */
  const double* d_x;    /* (N) node x, sorted order                          */
  const double* d_y;    /* (N) node y                                        */
  int64_t Ni, Nd, Nn, Nr, Npd;  /* row-class sizes; Npd = sum(Np)/2          */
  const double* d_nx;   /* (Nn+Nr+2*Npd) outward normals, sorted node order  */
  const double* d_ny;
  const double* d_robin_normal_x; /* (Nr) normal used by the Robin phi rows   */
  const double* d_robin_normal_y; /*      (updes/assembly.py:206 indexing)   */
  const double* d_beta; /* (Nr) Robin coefficients, sorted node order        */
  const int64_t* d_pd_i1; /* (Npd) first node of each periodic pair          */
  const int64_t* d_pd_i2; /* (Npd) partner node                               */
  double op_rbf_const[RBFX_NOPS];
  double op_mon_const[RBFX_NOPS];
  const double* d_op_rbf_row[RBFX_NOPS]; /* (Ni) per-row coefficients or NULL */
  const double* d_op_mon_row[RBFX_NOPS];
  int64_t support_cols;      /* S: support entries per node (flag set)      */
  const int32_t* d_support;  /* (N*S) ascending node ids per node, or NULL  */
} rbfx_problem;

/* ---- library ---------------------------------------------------------- */
int rbfx_abi_version(void);
const char* rbfx_last_error(void);
/* Queries the current device (CU count, LDS size).  Called implicitly. */
int rbfx_init(void);
/* Releases the streams and events the library created (the LU's look-ahead
 * stream and the timing events), after waiting for their work.  Call it
 * before the HIP runtime shuts down (the Python binding does so at exit);
 * the library re-creates them on its next use.  No reference counterpart. */
int rbfx_shutdown(void);

/* ---- K1 assembly ---------------------------------------------------------
 * Writes rows [row0, row0+nrows) of K or A (n = N+M columns) in FP64 and/or
 * FP32 (either output may be NULL).  Row r of the matrix goes to out[(r-row0)*ld].
 * Replaces updes/assembly.py:10-36 (assemble_Phi), 39-59 (assemble_P),
 * 62-85 (assemble_A), 93-137 (assemble_op_Phi_P), 141-362
 * (assemble_bd_Phi_P) and the concatenation in 384-396 (assemble_B).       */
int rbfx_assemble(const rbfx_problem* p, int matrix, int64_t row0, int64_t nrows,
                  double* d_out64, int64_t ld64, float* d_out32, int64_t ld32,
                  void* stream);

/* Matrix-free product y = K x (trans=0) or y = K^T x (trans=1) over all
 * n = N+M rows, recomputing every entry from the node coordinates in FP64.
 * With local supports only the support entries are visited: always for
 * trans=0, and for trans=1 when the workspace holds the reverse support
 * table (the size rbfx_matvec_workspace returns); a workspace of the dense
 * size takes the column kernel instead.  No workspace is needed for trans=0.
 * Replaces the products with B / A implied by updes/assembly.py:399,410,430. */
size_t rbfx_matvec_workspace(const rbfx_problem* p, int trans);
int rbfx_matvec(const rbfx_problem* p, int matrix, int trans, const double* d_x,
                double* d_y, void* d_work, size_t work_bytes, void* stream);

/* ---- K6 field evaluation -------------------------------------------------
 * out[t*npts + i] = sum_j lambda_j op_t phi(p_i, c_j) + sum_k gamma_k op_t p_k(p_i)
 * for every t set in ops_mask (bit t = enum rbfx_op), coefficients
 * coeffs = [lambda (N); gamma (M)].  All N centres are summed, the evaluation
 * point included (phi(0) for values, 0 for derivatives at r = 0).
 * Replaces updes/operators.py:118-150 (value/value_vec), 156-187
 * (gradient/gradient_vec) and 319-351 (laplacian/laplacian_vec).          */
int rbfx_eval(int kernel, double param, int64_t N, int64_t M, const double* d_cx,
              const double* d_cy, const double* d_coeffs, int64_t npts,
              const double* d_px, const double* d_py, uint32_t ops_mask,
              double* d_out, void* stream);

/* ---- right-hand side q ------------------------------------------------------
 * One term of the interior right-hand side: weight * w_i * op(field)(x_i),
 * the field given by its N + M RBF and monomial coefficients (as the
 * coefficient solve returns them), w_i = d_weight_row[i] or 1.            */
typedef struct rbfx_rhs_term {
  int32_t op;                  /* RBFX_VAL .. RBFX_DXY                        */
  int32_t reserved;            /* 0                                           */
  double weight;
  const double* d_weight_row;  /* (Ni) per-row factor, or NULL               */
  const double* d_coeffs;      /* (N + M) coefficients of the field          */
} rbfx_rhs_term;

/* The right-hand side [q; 0] of the augmented system (N + M entries, the
 * node order of p): interior rows i < Ni get constant + d_const_row[i] (or
 * 0 when NULL) + the sum of the terms at x_i (the lowered rhs_operator with
 * its fields); boundary rows Ni <= i < N get d_bc[i - Ni] (the boundary
 * values in node order, a Robin facet's value the first element of its
 * tuple); the M monomial rows get 0.  As in the reference, the values of
 * periodic facets are taken as given: pde_solver zeroes them beforehand
 * (updes/operators.py:546-555).  p needs N, M, Ni..Npd, kernel, param, d_x
 * and d_y.  Workspace: rbfx_rhs_workspace(p) bytes (none without terms).
 * Replaces assemble_q (updes/assembly.py:434-485).                         */
size_t rbfx_rhs_workspace(const rbfx_problem* p);
int rbfx_rhs(const rbfx_problem* p, double constant, const double* d_const_row, int nterms,
             const rbfx_rhs_term* terms, const double* d_bc, double* d_q, void* d_work,
             size_t work_bytes, void* stream);

/* Transpose of rbfx_eval for one operator t (adjoint of the evaluation):
 * out[j] = sum_i w_i op_t phi(p_i, c_j)  (j < N),
 * out[N+k] = sum_i w_i op_t p_k(p_i)     (k < M).  No reference counterpart
 * (the reference differentiates value() with jax.grad).                   */
int rbfx_eval_t(int kernel, double param, int64_t N, int64_t M, const double* d_cx,
                const double* d_cy, int64_t npts, const double* d_px,
                const double* d_py, int op, const double* d_w, double* d_out,
                void* stream);

/* ---- K2/K3 dense LU with partial pivoting ---------------------------------
 * In-place P A = L U of an n x n row-major matrix (unit lower L), LAPACK
 * getrf semantics with 0-based ipiv (row i was swapped with ipiv[i]).
 * Replaces the factorisations inside jnp.linalg.inv (updes/assembly.py:90)
 * and lineax QR (updes/operators.py:612-613).  *info = 0 or 1+index of the
 * first exactly-zero pivot.  Workspace size from rbfx_getrf_workspace.     */
size_t rbfx_getrf_workspace(int64_t n, int dtype_bytes);
int rbfx_getrf_f64(int64_t n, double* d_A, int64_t lda, int64_t* d_ipiv,
                   void* d_work, size_t work_bytes, int32_t* d_info, void* stream);
int rbfx_getrf_f32(int64_t n, float* d_A, int64_t lda, int64_t* d_ipiv,
                   void* d_work, size_t work_bytes, int32_t* d_info, void* stream);

/* Prepares the solve data of a factorisation: d_perm (n, int64) the row
 * permutation (row i of P A is row perm[i] of A) and d_dinv
 * (rbfx_getrs_dinv_elems(n) doubles: the FP64 inverses of the 64 x 64
 * diagonal blocks of L and U, then those of the 512 x 512 diagonal blocks
 * that the solves apply, then a condition estimate of each 512 x 512 block:
 * the FP64 solves substitute over the 64-row sub-blocks of an
 * ill-conditioned block instead of applying its inverse).  Synchronises the
 * stream (copies ipiv to the host once).                                    */
int64_t rbfx_getrs_dinv_elems(int64_t n);
int rbfx_getrs_prepare_f64(int64_t n, const double* d_LU, int64_t lda,
                           const int64_t* d_ipiv, int64_t* d_perm, double* d_dinv,
                           void* d_work, size_t work_bytes, void* stream);
int rbfx_getrs_prepare_f32(int64_t n, const float* d_LU, int64_t lda,
                           const int64_t* d_ipiv, int64_t* d_perm, double* d_dinv,
                           void* d_work, size_t work_bytes, void* stream);

/* Solves A x = b (trans=0) or A^T x = b (trans=1) in place on d_b (FP64
 * vector) with FP64 or FP32 factors; the arithmetic is FP64 in both cases
 * (FP32 factors are applied exactly as stored).  Work size from
 * rbfx_getrs_workspace.  Replaces the triangular solves of lineax QR and the
 * inv(A) @ v products (updes/assembly.py:410,430; operators.py:612-616).   */
size_t rbfx_getrs_workspace(int64_t n);
int rbfx_getrs_f64(int trans, int64_t n, const double* d_LU, int64_t lda,
                   const int64_t* d_perm, const double* d_dinv, double* d_b,
                   void* d_work, size_t work_bytes, void* stream);
int rbfx_getrs_f32(int trans, int64_t n, const float* d_LU, int64_t lda,
                   const int64_t* d_perm, const double* d_dinv, double* d_b,
                   void* d_work, size_t work_bytes, void* stream);
/* A X = B for nrhs right-hand sides at once (B: n x nrhs row-major, ldb >=
 * nrhs, overwritten by X) with FP64 LU factors and their solve data: blocked
 * substitution whose updates run as MFMA GEMMs.  Replaces the dense
 * jnp.linalg.inv of updes/assembly.py:87-90 (B = identity) that assemble_B
 * multiplies by D (updes/assembly.py:399).                                  */
size_t rbfx_getrs_multi_workspace(int64_t n, int64_t nrhs);
int rbfx_getrs_multi_f64(int64_t n, const double* d_LU, int64_t lda, const int64_t* d_perm,
                         const double* d_dinv, int64_t nrhs, double* d_B, int64_t ldb, void* d_work,
                         size_t work_bytes, void* stream);
/* Status of the last rbfx_getrs_* queued with this workspace: synchronises
 * the stream and returns RBFX_ERR_TIMEOUT if one of its device-side waits
 * expired (each gives up after about 4 s of real time; the solve then wrote
 * NaN into d_b instead of a partial result), else RBFX_OK.               */
int rbfx_getrs_status(int64_t n, void* d_work, size_t work_bytes, void* stream);
/* Allocation-free rbfx_getrs_f64: the same solve with the same kernels, its
 * scratch (permuted vector, superblock values, control words) in a fixed
 * device array of the library instead of a workspace.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * n is at most RBFX_GETRS_NOALLOC_MAX_N (larger n: RBFX_ERR_ARG, nothing
 * queued); the array (about 2 MB) is held on every device the call runs on
 * until the process ends.  All calls on one device share it: calls that
 * overlap on one GPU race.  A device-side wait that expired leaves NaN in
 * d_b (there is no status query for this variant).                        */
#define RBFX_GETRS_NOALLOC_MAX_N 65536
int rbfx_getrs_noalloc_f64(int trans, int64_t n, const double* d_LU, int64_t lda,
                           const int64_t* d_perm, const double* d_dinv, double* d_b,
                           void* stream);

/* ---- K5 FP64 GEMV ---------------------------------------------------------
 * y = alpha op(A) x + beta y, op = A (trans=0) or A^T (trans=1); A m x n.   */
size_t rbfx_gemv_workspace(int trans, int64_t m, int64_t n);
int rbfx_gemv_f64(int trans, int64_t m, int64_t n, double alpha, const double* d_A,
                  int64_t lda, const double* d_x, double beta, double* d_y,
                  void* d_work, size_t work_bytes, void* stream);

/* ---- GMRES-based iterative refinement -------------------------------------
 * Solves op(K) x = b (trans 0/1) in FP64 accuracy using FP32 LU factors of K
 * as a right preconditioner and FP64 residuals computed either from a stored
 * FP64 K (d_K64 != NULL) or matrix-free from the problem (d_K64 == NULL).
 * x holds the result.  Returns RBFX_ERR_STALL (x = best iterate) when the
 * residual does not reach tol * ||b|| within the iteration limits; the caller
 * then falls back to the FP64 factorisation.  Synchronises the stream.
 * No reference counterpart (the reference solves in FP64 only).            */
typedef struct rbfx_ir_stats {
  int32_t outer_iters;
  int32_t inner_iters;
  double final_relres;   /* ||b - K x|| / ||b||  (FP64)                     */
  int32_t stalled;
  int32_t reserved;
} rbfx_ir_stats;
size_t rbfx_gmres_ir_workspace(int64_t n, int32_t restart);
int rbfx_gmres_ir(int trans, int64_t n, const rbfx_problem* p, const double* d_K64,
                  int64_t ld64, const float* d_LU32, int64_t ld32,
                  const int64_t* d_perm32, const double* d_dinv32,
                  const double* d_b, double* d_x, double tol, int32_t restart,
                  int32_t max_inner, void* d_work, size_t work_bytes,
                  rbfx_ir_stats* stats, void* stream);

/* ---- FP32 preconditioner in a local cardinal basis --------------------------
 * The FP32-MFMA leg of the mixed-precision solve.  No reference counterpart
 * (the reference solves in FP64 only, updes/operators.py:612-613).
 * Column j of the basis C is node j's local cardinal function: the RBF
 * interpolant on the k nearest nodes S_j (j included) with the problem's M
 * monomials, 1 at node j and 0 at the rest of S_j, with RBF coefficients
 * beta_j (P^T beta_j = 0 to rounding) and monomial coefficients alpha_j.
 * K~ = diag(rs) D C (D: the first N rows of K) is N x N and well conditioned
 * where K is not; its FP32 LU preconditions GMRES on K itself:
 *   z = C K~^-1 diag(rs) r_D + [P (P^T P)^-1 r_P; 0]  (and its transpose).
 * Requires every node in every support (no RBFX_FLAG_LOCAL_SUPPORT).      */
typedef struct rbfx_cardinal {
  int64_t N, M, k;           /* nodes, monomials (<= 15), stencil size (k <= 48, k + M <= 63) */
  const double* d_x;         /* (N) node coordinates, sorted order                          */
  const double* d_y;
  const int32_t* d_stencil;  /* (N*k) stencil node ids per node, node j included            */
  double* d_beta;            /* (N*k) out of rbfx_cardinal_coeffs                            */
  double* d_alpha;           /* (N*M) out of rbfx_cardinal_coeffs                            */
  double* d_rs;              /* (N) row scales, out of rbfx_cardinal_form                   */
  const int64_t* d_rev_ptr;  /* (N+1) occurrences of node l in the stencils: entries         */
  const int32_t* d_rev_idx;  /*   d_rev_idx[d_rev_ptr[l] .. d_rev_ptr[l+1]) = j*k + t with
                                  d_stencil[j*k+t] == l, j ascending (rbfx_gmres_ir_cardinal) */
  const double* d_ptp_inv;   /* (M*M) inverse of P^T P, row-major (rbfx_gmres_ir_cardinal)   */
  /* Optional (all three or NULL) union tables of rbfx_cardinal_form: for each
   * block cb of 256 consecutive nodes, the sorted distinct ids of its stencils
   * d_union_idx[d_union_ptr[cb] .. d_union_ptr[cb+1]), and for each stencil
   * entry j*k+t the position of its id in its block's list.  With them the K
   * rows are staged through LDS; without them gathered from global memory. */
  const int64_t* d_union_ptr; /* (ceil(N/256)+1)                                             */
  const int32_t* d_union_idx;
  const int16_t* d_union_pos; /* (N*k)                                                       */
} rbfx_cardinal;

/* Local cardinal coefficients (beta, alpha) of every node.  Synchronises;
 * RBFX_ERR_ARG for a stencil id outside [0, N) or a node missing from its own
 * stencil, RBFX_ERR_SINGULAR for a stencil not unisolvent for the monomials.
 * Workspace: 256 bytes.                                                    */
int rbfx_cardinal_coeffs(const rbfx_problem* p, const rbfx_cardinal* c, void* d_work, size_t work_bytes,
                         void* stream);
/* rs and K~ = diag(rs) D C in FP32 (N x N, leading dimension ld32) from the
 * stored FP64 K (n x n, ld64).  Synchronises.  Workspace: 256 bytes.       */
int rbfx_cardinal_form(const rbfx_cardinal* c, const double* d_K64, int64_t ld64, float* d_out32,
                       int64_t ld32, void* d_work, size_t work_bytes, void* stream);
/* Solves op(K) x = b (trans 0/1, n = N + M) to FP64 accuracy: flexible GMRES
 * with FP64 residuals from the stored K, preconditioned by the FP32 LU of K~
 * (rbfx_getrf_f32 + rbfx_getrs_prepare_f32 of the N x N matrix).  The
 * forward system is solved row-scaled, diag(s) K x = diag(s) b with
 * s = [rs; 1] (the scales of K~), so that GMRES sees the preconditioned
 * operator diag(rs) D C K~^-1 without a diagonal similarity; the transposed
 * one unscaled.  A cycle stops as converged at a relative residual <= tol,
 * or at a normwise backward error eta = ||r|| / (k_fro ||x|| + ||b||) <=
 * 2^-54, or once the residual no longer improves at eta <= 1e-14; k_fro is
 * the Frobenius norm of the system's matrix (rbfx_frobenius_f64 of K with
 * the row scales rs for trans = 0, without for trans = 1; 0 disables the
 * eta tests).  RBFX_ERR_STALL otherwise (x = best iterate).  Synchronises. */
size_t rbfx_gmres_ir_cardinal_workspace(int64_t N, int64_t M, int32_t restart);
int rbfx_gmres_ir_cardinal(int trans, const rbfx_cardinal* c, const double* d_K64, int64_t ld64,
                           double k_fro, const float* d_LU32, int64_t ld32, const int64_t* d_perm32,
                           const double* d_dinv32, const double* d_b, double* d_x, double tol,
                           int32_t restart, int32_t max_inner, void* d_work, size_t work_bytes,
                           rbfx_ir_stats* stats, void* stream);
/* Frobenius norm of an m x n FP64 matrix whose rows i < n_scaled are scaled
 * by d_row_scale[i] (NULL: no scaling), returned on the host (synchronises). */
size_t rbfx_frobenius_workspace(int64_t m);
int rbfx_frobenius_f64(int64_t m, int64_t n, const double* d_A, int64_t lda, const double* d_row_scale,
                       int64_t n_scaled, double* h_norm, void* d_work, size_t work_bytes, void* stream);
/* Both norms of rbfx_frobenius_f64 from one read of A: with the row scales
 * (h_norm_scaled) and without (h_norm); the same workspace.  No counterpart
 * in the reference (the norms of GMRES-IR's backward-error test). */
int rbfx_frobenius_pair_f64(int64_t m, int64_t n, const double* d_A, int64_t lda, const double* d_row_scale,
                            int64_t n_scaled, double* h_norm_scaled, double* h_norm, void* d_work,
                            size_t work_bytes, void* stream);

/* ---- per-launch timing ----------------------------------------------------
 * Between rbfx_profile_begin() and rbfx_profile_end() every launch of the
 * main kernels is bracketed by HIP events on its own stream; _end
 * synchronises on them and returns per-kernel totals with the algorithmic
 * flops / bytes of each launch.  No reference counterpart (the reference
 * times with time.time / process_time around whole calls).                 */
typedef struct rbfx_prof_entry {
  char name[48];
  int64_t launches;
  double total_ms;
  double flops;   /* algorithmic flops summed over the launches   */
  double bytes;   /* algorithmic HBM bytes summed over the launches */
} rbfx_prof_entry;
int rbfx_profile_begin(void);
int rbfx_profile_end(rbfx_prof_entry* out, int max_entries, int* n_entries);

/* ---- test hook --------------------------------------------------------------
 * C = beta C + alpha A B (row-major, beta 0 or 1) with the MFMA kernel that
 * performs the LU trailing updates.                                         */
int rbfx_gemm_f64(int64_t m, int64_t n, int64_t k, double alpha, const double* d_A,
                  int64_t lda, const double* d_B, int64_t ldb, double beta, double* d_C,
                  int64_t ldc, void* stream);
int rbfx_gemm_f32(int64_t m, int64_t n, int64_t k, float alpha, const float* d_A,
                  int64_t lda, const float* d_B, int64_t ldb, float beta, float* d_C,
                  int64_t ldc, void* stream);

/* ---- tuning / test hook -----------------------------------------------------
 * Caps the workgroups of one panel-leaf launch of rbfx_getrf_* (process-wide;
 * 0 = the default: the device's CU count for FP64, a quarter of it for FP32,
 * or as many as three rows per thread need).  A leaf keeps its rows in
 * registers, one to three rows per thread of 256-thread workgroups, so a cap
 * below ceil(m/256) moves it to two or three rows per thread; a cap that
 * would need more than three fails the factorisation with an error.  No
 * counterpart in the reference.                                            */
int rbfx_set_panel_groups(int max_groups);

/* ---- tuning / test hook -----------------------------------------------------
 * Panels of rbfx_getrf_* with at least `rows` rows use 16-wide leaves instead
 * of 64-wide ones (process-wide; 0 = the default: 28,000 rows for FP64, none
 * for FP32; a negative value disables them).  No counterpart in the
 * reference.                                                               */
int rbfx_set_narrow_leaf_rows(int64_t rows);

/* ---- test hook --------------------------------------------------------------
 * Real-time bound of every device-side wait launched after the call
 * (process-wide, seconds; a negative value restores the default, 4 s).  A
 * bound of 0 makes every wait that is not satisfied at once expire, so the
 * failure handling of the solves (NaN result, RBFX_ERR_TIMEOUT) can be
 * exercised from the host while every kernel still exits normally.  No
 * counterpart in the reference.                                            */
int rbfx_set_wait_seconds(double seconds);

#ifdef __cplusplus
}
#endif
#endif /* RBFX_H */
