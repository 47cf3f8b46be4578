// Host-side bookkeeping of the solvers (plain C++, see host_logic.h).
#include "host_logic.h"

#include <math.h>

#include <algorithm>

namespace rbfx {

HessenbergLS::HessenbergLS(int m)
    : m_(m > 0 ? m : 1),
      H_((size_t)(m_ + 1) * m_),
      cs_(m_),
      sn_(m_),
      g_(m_ + 1),
      y_(m_) {}

void HessenbergLS::start(double beta) {
  std::fill(g_.begin(), g_.end(), 0.0);
  g_[0] = beta;
}

double HessenbergLS::add_column(int k, const double* h) {
  if (k < 0 || k >= m_) return INFINITY;
  double* col = &H_[(size_t)k * (m_ + 1)];
  for (int i = 0; i <= k + 1; ++i) col[i] = h[i];
  // the rotations of the earlier columns
  for (int i = 0; i < k; ++i) {
    const double t = cs_[i] * col[i] + sn_[i] * col[i + 1];
    col[i + 1] = -sn_[i] * col[i] + cs_[i] * col[i + 1];
    col[i] = t;
  }
  // this column's rotation zeroes h[k+1]
  const double den = hypot(col[k], col[k + 1]);
  cs_[k] = den > 0.0 ? col[k] / den : 1.0;
  sn_[k] = den > 0.0 ? col[k + 1] / den : 0.0;
  col[k] = den;
  col[k + 1] = 0.0;
  g_[k + 1] = -sn_[k] * g_[k];
  g_[k] = cs_[k] * g_[k];
  return fabs(g_[k + 1]);
}

const double* HessenbergLS::solve(int kdone) {
  kdone = std::max(0, std::min(kdone, m_));
  for (int i = kdone - 1; i >= 0; --i) {
    double t = g_[i];
    for (int j = i + 1; j < kdone; ++j) t -= H_[(size_t)j * (m_ + 1) + i] * y_[j];
    const double d = H_[(size_t)i * (m_ + 1) + i];
    y_[i] = d != 0.0 ? t / d : 0.0;
  }
  return y_.data();
}

double iterations_needed(double rn, double est, int k1, double target) {
  if (!(rn > 0.0) || !(est > 0.0) || k1 <= 0) return est <= target ? 0.0 : 1e30;
  if (est <= target) return 0.0;
  const double rate = pow(est / rn, 1.0 / k1);
  return (rate < 1.0 && rate > 0.0) ? log(target / est) / log(rate) : 1e30;
}

IrMonitor::Decision IrMonitor::next(double rel, double eta) {
  if (!(rel == rel)) return kStalled;  // NaN
  if (rel <= tol) return kConverged;
  if (eta <= stop_eta) return kConverged;
  if (rel <= 0.5 * best) {
    best = rel;
    noprog = 0;
    return kContinue;
  }
  ++noprog;
  if (est_met && eta <= floor_eta) return kConverged;
  return noprog >= 3 ? kStalled : kContinue;
}

bool leaf_shape(int64_t m, int64_t gmax, int threads, int max_rows, LeafShape& out) {
  if (m <= 0 || gmax <= 0 || threads <= 0 || max_rows <= 0) return false;
  // as few workgroups as one row per thread allows; more rows per thread
  // only when the workgroups run out
  const int64_t per_group = (int64_t)threads * gmax;
  const int64_t rows = (m + per_group - 1) / per_group;
  if (rows > max_rows) return false;
  const int64_t cap = (int64_t)threads * rows;
  int64_t G = (m + cap - 1) / cap;
  const int64_t R = (m + G - 1) / G;
  G = (m + R - 1) / R;
  out.rows = (int)rows;
  out.G = (int)G;
  out.R = R;
  return true;
}

}  // namespace rbfx
