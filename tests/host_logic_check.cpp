// CPU driver of updes_amd/csrc/host_logic.cpp, built by tests/test_host_logic.py
// with AddressSanitizer and UBSan.  Prints "ok" and exits 0 when every check
// holds; any sanitizer report aborts the run.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <random>
#include <vector>

#include "../updes_amd/csrc/host_logic.h"

using namespace rbfx;

static int failures = 0;
#define CHECK(c)                                                \
  do {                                                          \
    if (!(c)) {                                                 \
      fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      ++failures;                                               \
    }                                                           \
  } while (0)

static void check_workspace() {
  std::vector<char> buf(1000);
  Workspace ws(buf.data(), buf.size());
  double* a = ws.take<double>(10);
  CHECK(a == reinterpret_cast<double*>(buf.data()));
  char* b = ws.take<char>(1);
  CHECK(b == buf.data() + 256);
  const size_t used = ws.used;
  CHECK(ws.take<double>(1000) == nullptr);
  CHECK(ws.used == used);
  CHECK(ws.take<int>(SIZE_MAX / 2) == nullptr);  // no overflow of count * size
  CHECK(ws.take<char>(1000 - 512) != nullptr);   // exactly fills the rest
  CHECK(ws.take<char>(1) == nullptr);
  CHECK(ws_round(0) == 0 && ws_round(1) == 256 && ws_round(256) == 256 && ws_round(257) == 512);
}

// min || beta e1 - H y || over y for the Hessenberg columns given, directly:
// returns the residual norm of y from HessenbergLS and the norm of H^T r.
static void check_hessenberg(int m, unsigned seed) {
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  HessenbergLS ls(m);
  const double beta = 2.5;
  ls.start(beta);
  std::vector<std::vector<double>> cols;
  for (int k = 0; k < m; ++k) {
    std::vector<double> h(k + 2);
    for (auto& v : h) v = nd(rng);
    h[k + 1] = fabs(h[k + 1]) + 0.1;
    cols.push_back(h);
    const double est = ls.add_column(k, h.data());
    const double* y = ls.solve(k + 1);
    std::vector<double> r(k + 2, 0.0);
    r[0] = beta;
    for (int j = 0; j <= k; ++j)
      for (int i = 0; i <= j + 1; ++i) r[i] -= cols[j][i] * y[j];
    double rn = 0.0;
    for (double v : r) rn += v * v;
    rn = sqrt(rn);
    CHECK(fabs(rn - est) <= 1e-10 * beta);
    CHECK(fabs(ls.estimate(k) - est) == 0.0);
    // optimality: H^T r = 0
    for (int j = 0; j <= k; ++j) {
      double t = 0.0;
      for (int i = 0; i <= j + 1; ++i) t += cols[j][i] * r[i];
      CHECK(fabs(t) <= 1e-9 * beta);
    }
  }
  CHECK(isinf(ls.add_column(m, cols[0].data())));  // beyond the capacity: refused
  // a zero column (breakdown) solves to zeros there, no division by zero
  HessenbergLS z(3);
  z.start(1.0);
  const double zero[3] = {0.0, 0.0, 0.0};
  z.add_column(0, zero);
  const double* y = z.solve(1);
  CHECK(y[0] == 0.0);
}

static void check_iterations_needed() {
  CHECK(fabs(iterations_needed(1.0, 1e-3, 3, 1e-9) - 6.0) < 1e-9);
  CHECK(iterations_needed(1.0, 2.0, 3, 1e-9) == 1e30);
  CHECK(iterations_needed(1.0, 1e-12, 3, 1e-9) == 0.0);
  CHECK(iterations_needed(0.0, 1e-3, 3, 1e-9) == 1e30);
  CHECK(iterations_needed(1.0, 1e-3, 0, 1e-9) == 1e30);
}

static void check_monitor() {
  {
    IrMonitor m{1e-12, 0x1p-54, 1e-14};
    CHECK(m.next(1e-3, INFINITY) == IrMonitor::kContinue);
    CHECK(m.next(1e-6, INFINITY) == IrMonitor::kContinue);
    CHECK(m.next(1e-13, INFINITY) == IrMonitor::kConverged);
  }
  {
    IrMonitor m{1e-12, 0x1p-54, 1e-14};
    CHECK(m.next(1e-3, INFINITY) == IrMonitor::kContinue);
    CHECK(m.next(0.9e-3, INFINITY) == IrMonitor::kContinue);
    CHECK(m.next(0.8e-3, INFINITY) == IrMonitor::kContinue);
    CHECK(m.next(0.7e-3, INFINITY) == IrMonitor::kStalled);
  }
  {
    IrMonitor m{1e-12, 0x1p-54, 1e-14};
    CHECK(m.next(NAN, INFINITY) == IrMonitor::kStalled);
  }
  {
    IrMonitor m{1e-12, 0x1p-54, 1e-14};
    CHECK(m.next(1e-9, 1e-17) == IrMonitor::kConverged);  // backward-stable solve
  }
  {
    IrMonitor m{1e-12, 0x1p-54, 1e-14};
    CHECK(m.next(1e-9, 1e-15) == IrMonitor::kContinue);
    m.est_met = true;
    CHECK(m.next(0.9e-9, 1e-15) == IrMonitor::kConverged);  // FP64 floor reached
    IrMonitor m2{1e-12, 0x1p-54, 1e-14};
    CHECK(m2.next(1e-9, 1e-13) == IrMonitor::kContinue);
    m2.est_met = true;
    CHECK(m2.next(0.9e-9, 1e-13) == IrMonitor::kContinue);  // backward error too large
  }
}

static void check_leaf_shape() {
  const int threads = 256, max_rows = 3;
  const int64_t gmax = 256;
  const int64_t ms[] = {1, 2, 255, 256, 257, 4000, 40006, 65535, 65536, 65537, 131072, 150000, 196607, 196608};
  for (int64_t m : ms) {
    LeafShape s;
    CHECK(leaf_shape(m, gmax, threads, max_rows, s));
    CHECK(s.rows >= 1 && s.rows <= max_rows);
    CHECK(s.G >= 1 && s.G <= gmax);
    CHECK((int64_t)s.G * s.R >= m);
    CHECK((int64_t)(s.G - 1) * s.R < m);  // no empty workgroup
    CHECK(s.R <= (int64_t)threads * s.rows);
    // one row per thread while the workgroups suffice
    if (m <= threads * gmax) CHECK(s.rows == 1);
  }
  LeafShape s;
  CHECK(!leaf_shape(196609, gmax, threads, max_rows, s));
  CHECK(!leaf_shape(0, gmax, threads, max_rows, s));
  CHECK(!leaf_shape(100, 0, threads, max_rows, s));
}

int main() {
  check_workspace();
  for (unsigned seed = 1; seed <= 3; ++seed) check_hessenberg(40, seed);
  check_iterations_needed();
  check_monitor();
  check_leaf_shape();
  if (failures) {
    fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  printf("ok\n");
  return 0;
}
