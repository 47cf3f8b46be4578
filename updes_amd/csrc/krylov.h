// Internal interface of the GMRES-IR driver (krylov.hip) shared with the
// preconditioners (krylov.hip: FP32 LU of K; cardinal.hip: FP32 LU of K in a
// local cardinal basis).
#pragma once
#include "common.h"

namespace rbfx {

// One preconditioned operator step of flexible GMRES: z = M^-1 v and
// w = op(K) z, both of length n.  Implementations may use their own scratch.
struct PairOp {
  virtual ~PairOp() = default;
  virtual int apply(const double* v, double* z, double* w) = 0;
};

// w = op(K) x in FP64 (the residual operator of the refinement).
struct ResidualOp {
  virtual ~ResidualOp() = default;
  virtual int applyK(const double* x, double* w) = 0;
};

struct FgmresOpts {
  double tol = 1e-11;  // relative residual ||b - op(K) x|| / ||b||
  int restart = 30;
  int max_inner = 150;
  // Frobenius norm of K, or 0.  When positive, the normwise backward error
  // eta = ||r|| / (||K||_F ||x|| + ||b||) of every cycle's iterate is checked:
  // at most kStopEta (a backward-stable FP64 solve) stops as converged, and so
  // does at most kFloorEta once the true residual no longer improves while
  // the Arnoldi estimate met its target (the FP64 residual's attainable
  // accuracy is reached).
  double knorm = 0.0;
};
constexpr double kStopEta = 0x1p-54;
constexpr double kFloorEta = 1e-14;

size_t fgmres_ws_bytes(int64_t n, int restart);
// Flexible restarted GMRES on the correction equation with FP64 residuals.
// Returns RBFX_OK or RBFX_ERR_STALL (x = best iterate).  Synchronises.
int fgmres_ir(PairOp& pre, ResidualOp& res, int64_t n, const double* b, double* x, const FgmresOpts& o,
              void* work, size_t work_bytes, rbfx_ir_stats* stats, hipStream_t s);

// FP64 GEMV (krylov.hip).
size_t gemv_workspace(int trans, int64_t m, int64_t n);
int gemv(int trans, int64_t m, int64_t n, double alpha, const double* A, int64_t lda, const double* x,
         double beta, double* y, double* part, hipStream_t s);
// y = op(K) x for the n x n residual matrix, timed as "residual_gemv(_t)".
int residual_gemv(int trans, int64_t n, const double* K, int64_t ld, const double* x, double* y, double* part,
                  hipStream_t s);

}  // namespace rbfx
