// Host-side bookkeeping of the solvers in plain C++ (no HIP): workspace
// carving, the small dense least-squares problem and the stopping decisions of
// flexible GMRES, and the launch geometry of the LU's panel leaves.  The
// library compiles host_logic.cpp like any other source; a CPU test builds the
// same file with AddressSanitizer and UBSan and drives it directly.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace rbfx {

// Bump allocator over a caller-provided workspace (256-byte aligned slices).
// take() returns nullptr, and takes nothing, when the slice does not fit.
struct Workspace {
  char* base;
  size_t size;
  size_t used = 0;
  Workspace(void* p, size_t n) : base(static_cast<char*>(p)), size(n) {}
  template <typename T>
  T* take(size_t count) {
    const size_t off = (used + 255) & ~size_t(255);
    if (off > size || count > (size - off) / sizeof(T)) return nullptr;
    used = off + count * sizeof(T);
    return reinterpret_cast<T*>(base + off);
  }
};
inline size_t ws_round(size_t bytes) { return (bytes + 255) & ~size_t(255); }

// Least-squares problem of one GMRES(m) cycle: the (k+2) x (k+1) Hessenberg
// matrix is reduced to triangular form by Givens rotations as its columns
// arrive, so |g[k+1]| is the residual norm of min ||beta e1 - H y||.
class HessenbergLS {
 public:
  explicit HessenbergLS(int m);
  int capacity() const { return m_; }
  // new cycle with right-hand side beta e1
  void start(double beta);
  // column k of H: h[0..k+1] (k+2 entries, h[k+1] the new basis vector's
  // norm); returns the updated residual estimate |g[k+1]|
  double add_column(int k, const double* h);
  // y[0..kdone) = R^-1 g (zero where R has a zero diagonal); valid until the
  // next call
  const double* solve(int kdone);
  double estimate(int k) const { return g_[k + 1] < 0 ? -g_[k + 1] : g_[k + 1]; }

 private:
  int m_;
  std::vector<double> H_, cs_, sn_, g_, y_;
};

// Inner iterations that a geometric decrease from rn to est over k1
// iterations still needs to reach target (1e30 when it does not decrease).
double iterations_needed(double rn, double est, int k1, double target);

// Outer-loop decisions of GMRES-IR from each cycle's true residual:
// converged on the relative residual or the backward error, stopped when
// three cycles in a row fail to halve the best residual, and converged at
// the FP64 floor when the last cycle's Arnoldi estimate met its target but
// the true residual did not follow while the backward error is small.
struct IrMonitor {
  double tol, stop_eta, floor_eta;
  double best = 1e300;
  int noprog = 0;
  bool est_met = false;
  enum Decision { kContinue = 0, kConverged = 1, kStalled = 2 };
  // rel = ||r|| / ||b||; eta = backward error (or +inf when not known)
  Decision next(double rel, double eta);
};

// Launch geometry of one register-resident panel leaf of m rows: G
// workgroups of `threads` threads holding `rows` rows per thread, R rows per
// workgroup.  Returns false when m needs more than max_rows rows per thread
// with at most gmax workgroups.
struct LeafShape {
  int rows = 0;
  int G = 0;
  int64_t R = 0;
};
bool leaf_shape(int64_t m, int64_t gmax, int threads, int max_rows, LeafShape& out);

}  // namespace rbfx
