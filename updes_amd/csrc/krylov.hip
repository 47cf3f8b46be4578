// K5 FP64 GEMV and the GMRES-based iterative refinement driver.
//
// The reference solves B u = q with lineax QR in FP64 only
// (updes/operators.py:612-613).  Here the collocation system K c = [q; 0] can
// be factored in FP32 (MFMA) and the FP64 solution recovered by GMRES-IR
// (Carson & Higham, SIAM J. Sci. Comput. 40(2), 2018): FP32 factors
// right-precondition restarted flexible GMRES on the correction equation,
// residuals b - K x are formed in FP64 either from the stored FP64 matrix
// (GEMV) or matrix-free from the node coordinates (rbfx_matvec), and
// refinement stops with RBFX_ERR_STALL when it fails to reach the tolerance
// so that the caller can fall back to an FP64 factorisation.  The
// preconditioner is either the FP32 LU of K itself (here) or of K in a local
// cardinal basis (cardinal.hip).
#include <math.h>
#include <stdlib.h>

#include <algorithm>
#include <string>
#include <vector>

#include "krylov.h"

namespace rbfx {

constexpr int kGvThreads = 256;
// Inner iterations observed before the convergence rate is extrapolated.
constexpr int kRateWindow = 24;

// y[i] = alpha * sum_j A[i][j] x[j] + beta * y[i]: one block per row.
__global__ __launch_bounds__(kGvThreads) void gemv_n_block_kernel(int64_t m, int64_t n, double alpha,
                                                                  const double* __restrict__ A, int64_t lda,
                                                                  const double* __restrict__ x, double beta,
                                                                  double* __restrict__ y) {
  const int64_t i = blockIdx.x;
  const double* row = A + i * lda;
  double s = 0.0;
  for (int64_t j = threadIdx.x; j < n; j += kGvThreads) s += row[j] * x[j];
  __shared__ double red[kGvThreads / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0] + red[1] + red[2] + red[3];
    y[i] = alpha * t + (beta != 0.0 ? beta * y[i] : 0.0);
  }
}

// Same, one wave per row (many rows): 16-byte non-temporal loads, four in
// flight per lane.
__global__ __launch_bounds__(kGvThreads) void gemv_n_wave_kernel(int64_t m, int64_t n, double alpha,
                                                                 const double* __restrict__ A, int64_t lda,
                                                                 const double* __restrict__ x, double beta,
                                                                 double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * (kGvThreads / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= m) return;
  const double* row = A + i * lda;
  double s = 0.0;
  const bool vec = ((((uintptr_t)row) | ((uintptr_t)x)) & 15) == 0;
  if (vec) {
    const int64_t n2 = n / 2;
    typedef double vd2 __attribute__((ext_vector_type(2)));
    const vd2* r2 = reinterpret_cast<const vd2*>(row);
    const vd2* x2 = reinterpret_cast<const vd2*>(x);
    // four 16-byte loads in flight per lane
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int64_t j = lane;
    for (; j + 192 < n2; j += 256) {
      const vd2 a0 = __builtin_nontemporal_load(r2 + j);
      const vd2 a1 = __builtin_nontemporal_load(r2 + j + 64);
      const vd2 a2 = __builtin_nontemporal_load(r2 + j + 128);
      const vd2 a3 = __builtin_nontemporal_load(r2 + j + 192);
      const vd2 b0 = x2[j], b1 = x2[j + 64], b2 = x2[j + 128], b3 = x2[j + 192];
      s = __builtin_fma(a0.x, b0.x, s);
      s = __builtin_fma(a0.y, b0.y, s);
      s1 = __builtin_fma(a1.x, b1.x, s1);
      s1 = __builtin_fma(a1.y, b1.y, s1);
      s2 = __builtin_fma(a2.x, b2.x, s2);
      s2 = __builtin_fma(a2.y, b2.y, s2);
      s3 = __builtin_fma(a3.x, b3.x, s3);
      s3 = __builtin_fma(a3.y, b3.y, s3);
    }
    for (; j < n2; j += 64) {
      const vd2 a = __builtin_nontemporal_load(r2 + j);
      const vd2 b = x2[j];
      s = __builtin_fma(a.x, b.x, s);
      s = __builtin_fma(a.y, b.y, s);
    }
    s += (s1 + s2) + s3;
    if ((n & 1) && lane == 0) s += row[n - 1] * x[n - 1];
  } else {
    for (int64_t j = lane; j < n; j += 64) s += row[j] * x[j];
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) y[i] = alpha * s + (beta != 0.0 ? beta * y[i] : 0.0);
}

// y[j] = alpha * sum_i A[i][j] x[i] + beta * y[j], partial sums over row
// chunks (reduced by gemv_t_finish_kernel).  Workgroup (column block of
// kGvTCols, row chunk): lane l of wave w owns columns 128 b + 2 l and
// 128 b + 2 l + 1 and the rows rs + w, rs + w + 4, ... of the chunk, with eight
// 16-byte row loads in flight; the four waves' sums meet in LDS.
constexpr int kGvTCols = 128;
__global__ __launch_bounds__(kGvThreads) void gemv_t_kernel(int64_t m, int64_t n, const double* __restrict__ A,
                                                            int64_t lda, const double* __restrict__ x,
                                                            int64_t rows_per_chunk, double* __restrict__ part) {
  typedef double vd2 __attribute__((ext_vector_type(2)));
  __shared__ double red[kGvThreads / 64][kGvTCols];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t c0 = (int64_t)blockIdx.x * kGvTCols + 2 * lane;
  const int64_t rs = (int64_t)blockIdx.y * rows_per_chunk;
  const int64_t re = rs + rows_per_chunk < m ? rs + rows_per_chunk : m;
  double s0 = 0.0, s1 = 0.0;
  const bool vec = ((lda & 1) == 0) && ((reinterpret_cast<uintptr_t>(A) & 15) == 0);
  if (vec && c0 + 1 < n) {
    const double* a = A + c0;
    int64_t i = rs + w;
    for (; i + 28 < re; i += 32) {
      vd2 av[8];
      double xv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        av[u] = __builtin_nontemporal_load(reinterpret_cast<const vd2*>(a + (i + 4 * u) * lda));
        xv[u] = x[i + 4 * u];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        s0 = __builtin_fma(av[u].x, xv[u], s0);
        s1 = __builtin_fma(av[u].y, xv[u], s1);
      }
    }
    for (; i < re; i += 4) {
      const vd2 a0 = __builtin_nontemporal_load(reinterpret_cast<const vd2*>(a + i * lda));
      s0 = __builtin_fma(a0.x, x[i], s0);
      s1 = __builtin_fma(a0.y, x[i], s1);
    }
  } else if (c0 < n) {
    const bool two = c0 + 1 < n;
    for (int64_t i = rs + w; i < re; i += 4) {
      s0 = __builtin_fma(A[i * lda + c0], x[i], s0);
      if (two) s1 = __builtin_fma(A[i * lda + c0 + 1], x[i], s1);
    }
  }
  red[w][2 * lane] = s0;
  red[w][2 * lane + 1] = s1;
  __syncthreads();
  if (threadIdx.x < kGvTCols) {
    const int64_t j = (int64_t)blockIdx.x * kGvTCols + threadIdx.x;
    if (j < n) {
      double t = red[0][threadIdx.x];
#pragma unroll
      for (int q = 1; q < kGvThreads / 64; ++q) t += red[q][threadIdx.x];
      part[(int64_t)blockIdx.y * n + j] = t;
    }
  }
}

__global__ void gemv_t_finish_kernel(int64_t n, int64_t chunks, const double* __restrict__ part,
                                     double alpha, double beta, double* __restrict__ y) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double s = 0.0;
  for (int64_t c = 0; c < chunks; ++c) s += part[c * n + j];
  y[j] = alpha * s + (beta != 0.0 ? beta * y[j] : 0.0);
}

// Row chunks of the transposed product: about eight workgroups per CU.
static int64_t gemv_t_chunks(int64_t m, int64_t n) {
  const int64_t colb = ceil_div(n, kGvTCols);
  int64_t want = ceil_div(8 * (int64_t)device_info().num_cus, colb);
  want = std::max<int64_t>(1, std::min<int64_t>(want, ceil_div(m, 64)));
  return want;
}

int gemv_launch(int trans, int64_t m, int64_t n, double alpha, const double* A, int64_t lda, const double* x,
                double beta, double* y, double* part, hipStream_t s);

size_t gemv_workspace(int trans, int64_t m, int64_t n) {
  if (!trans) return 0;
  return ws_round(sizeof(double) * (size_t)gemv_t_chunks(m, n) * (size_t)n);
}

int gemv(int trans, int64_t m, int64_t n, double alpha, const double* A, int64_t lda, const double* x,
         double beta, double* y, double* part, hipStream_t s) {
  if (m <= 0 || n <= 0) return RBFX_OK;
  return gemv_launch(trans, m, n, alpha, A, lda, x, beta, y, part, s);
}

int residual_gemv(int trans, int64_t n, const double* K, int64_t ld, const double* x, double* y, double* part,
                  hipStream_t s) {
  const int pid = prof_start(trans ? "residual_gemv_t" : "residual_gemv", s);
  const int st = gemv(trans, n, n, 1.0, K, ld, x, 0.0, y, part, s);
  prof_stop(pid, s, 2.0 * n * n, 8.0 * ((double)n * n + 2.0 * n));
  return st;
}

int gemv_launch(int trans, int64_t m, int64_t n, double alpha, const double* A, int64_t lda, const double* x,
                double beta, double* y, double* part, hipStream_t s) {
  if (!trans) {
    if (m < 2048) {
      hipLaunchKernelGGL(gemv_n_block_kernel, dim3((unsigned)m), dim3(kGvThreads), 0, s, m, n, alpha, A,
                         lda, x, beta, y);
    } else {
      hipLaunchKernelGGL(gemv_n_wave_kernel, dim3((unsigned)ceil_div(m, kGvThreads / 64)), dim3(kGvThreads),
                         0, s, m, n, alpha, A, lda, x, beta, y);
    }
    RBFX_CHECK_LAUNCH();
    return RBFX_OK;
  }
  const int64_t chunks = gemv_t_chunks(m, n);
  const int64_t rpc = ceil_div(m, chunks);
  const int64_t used = ceil_div(m, rpc);
  RBFX_REQUIRE(part != nullptr, "gemv^T needs a workspace");
  hipLaunchKernelGGL(gemv_t_kernel, dim3((unsigned)ceil_div(n, kGvTCols), (unsigned)used), dim3(kGvThreads),
                     0, s, m, n, A, lda, x, rpc, part);
  RBFX_CHECK_LAUNCH();
  hipLaunchKernelGGL(gemv_t_finish_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, n, used, part,
                     alpha, beta, y);
  RBFX_CHECK_LAUNCH();
  return RBFX_OK;
}

// ---- small vector kernels ------------------------------------------------
// Deterministic dot product: fixed block partials then one block.
constexpr int kDotBlocks = 256;
__global__ __launch_bounds__(256) void dot_partial_kernel(int64_t n, const double* __restrict__ a,
                                                          const double* __restrict__ b,
                                                          double* __restrict__ part) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    s += a[i] * b[i];
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
__global__ __launch_bounds__(256) void dot_final_kernel(int nparts, const double* __restrict__ part,
                                                        double* __restrict__ out) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) s += part[i];
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) *out = red[0] + red[1] + red[2] + red[3];
}

// y = a*x + b*y
__global__ void axpby_kernel(int64_t n, double a, const double* __restrict__ x, double b,
                             double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + (b != 0.0 ? b * y[i] : 0.0);
}
// y = a*x (scaled copy)
__global__ void scale_copy_kernel(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}

__global__ void fill_kernel(int64_t n, double v, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}

// Frobenius norm partials: sum of squares of rows [blockIdx.x*rows, ...),
// row i scaled by rs[i] when i < nrs.
// Partial sums of squares over row blocks: part[b] with rows i < nrs scaled
// by rs[i] (rs may be null), and, when part1 is given, part1[b] unscaled
// from the same loads.
__global__ __launch_bounds__(256) void fro_partial_kernel(int64_t m, int64_t n, const double* __restrict__ A,
                                                          int64_t lda, int64_t rows, const double* __restrict__ rs,
                                                          int64_t nrs, double* __restrict__ part,
                                                          double* __restrict__ part1) {
  double s = 0.0, s1 = 0.0;
  const int64_t r0 = (int64_t)blockIdx.x * rows;
  const int64_t r1 = r0 + rows < m ? r0 + rows : m;
  for (int64_t i = r0; i < r1; ++i) {
    const double f = (rs && i < nrs) ? rs[i] : 1.0;
    double t = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += 256) {
      const double v = A[i * lda + j];
      t += v * v;
    }
    s += f * f * t;
    s1 += t;
  }
  __shared__ double red[2][4];
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_down(s, off, 64);
    s1 += __shfl_down(s1, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = s;
    red[1][threadIdx.x >> 6] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    if (part1) part1[blockIdx.x] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

static int dot(int64_t n, const double* a, const double* b, double* part, double* out, hipStream_t s) {
  hipLaunchKernelGGL(dot_partial_kernel, dim3(kDotBlocks), dim3(256), 0, s, n, a, b, part);
  RBFX_CHECK_LAUNCH();
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(256), 0, s, kDotBlocks, part, out);
  RBFX_CHECK_LAUNCH();
  return RBFX_OK;
}

}  // namespace rbfx

using namespace rbfx;

extern "C" size_t rbfx_gemv_workspace(int trans, int64_t m, int64_t n) {
  if (m <= 0 || n <= 0) return 0;
  return gemv_workspace(trans, m, n);
}

extern "C" int rbfx_gemv_f64(int trans, int64_t m, int64_t n, double alpha, const double* d_A, int64_t lda,
                             const double* d_x, double beta, double* d_y, void* d_work, size_t work_bytes,
                             void* stream) {
  RBFX_REQUIRE(m >= 0 && n >= 0 && lda >= n, "bad gemv shape");
  RBFX_REQUIRE(d_A && d_x && d_y && d_x != d_y, "bad gemv pointers");
  RBFX_REQUIRE(work_bytes >= gemv_workspace(trans, m, n) && (!trans || d_work), "gemv workspace too small");
  return gemv(trans, m, n, alpha, d_A, lda, d_x, beta, d_y, static_cast<double*>(d_work), as_stream(stream));
}

static int64_t fro_blocks(int64_t m) { return std::max<int64_t>(1, std::min<int64_t>(m, 4096)); }

extern "C" size_t rbfx_frobenius_workspace(int64_t m) {
  return 2 * ws_round(sizeof(double) * (size_t)fro_blocks(m)) + 64;  // room for rbfx_frobenius_pair_f64
}

// Both norms (scaled, unscaled) from one read of A; norm1 may be null.
static int frobenius(int64_t m, int64_t n, const double* A, int64_t lda, const double* rs, int64_t nrs,
                     double* norm, double* norm1, void* work, size_t work_bytes, hipStream_t s) {
  RBFX_REQUIRE(m > 0 && n > 0 && lda >= n && A && norm, "bad frobenius args");
  RBFX_REQUIRE(work && work_bytes >= rbfx_frobenius_workspace(m), "frobenius workspace too small");
  RBFX_REQUIRE(!rs || (nrs >= 0 && nrs <= m), "bad row scale count");
  const int64_t rows = ceil_div(m, fro_blocks(m));
  const int64_t used = ceil_div(m, rows);
  Workspace ws(work, work_bytes);
  double* part = ws.take<double>(fro_blocks(m));
  double* part1 = ws.take<double>(fro_blocks(m));
  RBFX_REQUIRE(part && part1, "frobenius workspace carve failed");
  hipLaunchKernelGGL(fro_partial_kernel, dim3((unsigned)used), dim3(256), 0, s, m, n, A, lda, rows, rs, nrs, part,
                     norm1 ? part1 : nullptr);
  RBFX_CHECK_LAUNCH();
  std::vector<double> hp(2 * used);
  RBFX_HIP(hipMemcpyAsync(hp.data(), part, used * sizeof(double), hipMemcpyDeviceToHost, s));
  if (norm1) RBFX_HIP(hipMemcpyAsync(hp.data() + used, part1, used * sizeof(double), hipMemcpyDeviceToHost, s));
  RBFX_HIP(hipStreamSynchronize(s));
  double t = 0.0, t1 = 0.0;
  for (int64_t b = 0; b < used; ++b) {
    t += hp[b];
    t1 += hp[used + b];
  }
  *norm = sqrt(t);
  if (norm1) *norm1 = sqrt(t1);
  return RBFX_OK;
}

extern "C" int rbfx_frobenius_f64(int64_t m, int64_t n, const double* d_A, int64_t lda, const double* d_row_scale,
                                  int64_t n_scaled, double* h_norm, void* d_work, size_t work_bytes, void* stream) {
  return frobenius(m, n, d_A, lda, d_row_scale, n_scaled, h_norm, nullptr, d_work, work_bytes, as_stream(stream));
}

extern "C" int rbfx_frobenius_pair_f64(int64_t m, int64_t n, const double* d_A, int64_t lda,
                                       const double* d_row_scale, int64_t n_scaled, double* h_norm_scaled,
                                       double* h_norm, void* d_work, size_t work_bytes, void* stream) {
  RBFX_REQUIRE(h_norm, "bad frobenius args");
  return frobenius(m, n, d_A, lda, d_row_scale, n_scaled, h_norm_scaled, h_norm, d_work, work_bytes,
                   as_stream(stream));
}

namespace rbfx {

size_t fgmres_ws_bytes(int64_t n, int m) {
  size_t b = 0;
  b += ws_round(sizeof(double) * (size_t)(m + 1) * (size_t)n);  // V
  b += ws_round(sizeof(double) * (size_t)m * (size_t)n);        // Z = M^-1 V
  b += 3 * ws_round(sizeof(double) * (size_t)n);                // r, w, u
  b += ws_round(sizeof(double) * (size_t)(2 * m + 8));          // h, h2, norms
  b += ws_round(sizeof(double) * kDotBlocks);                   // dot partials
  b += ws_round(std::max(gemv_workspace(1, m + 1, n), gemv_workspace(1, m, n)));
  return b + 2048;
}

int fgmres_ir(PairOp& pre, ResidualOp& res, int64_t n, const double* d_b, double* d_x, const FgmresOpts& o,
              void* work, size_t work_bytes, rbfx_ir_stats* stats, hipStream_t s) {
  const int m = o.restart;
  RBFX_REQUIRE(n > 0 && m > 0 && o.max_inner > 0 && o.tol > 0.0, "bad GMRES-IR parameters");
  RBFX_REQUIRE(work_bytes >= fgmres_ws_bytes(n, m), "GMRES-IR workspace too small");
  Workspace ws(work, work_bytes);
  double* V = ws.take<double>((size_t)(m + 1) * n);
  double* Z = ws.take<double>((size_t)m * n);
  double* r = ws.take<double>(n);
  double* w = ws.take<double>(n);
  double* u = ws.take<double>(n);
  double* hd = ws.take<double>(2 * m + 8);
  double* dpart = ws.take<double>(kDotBlocks);
  const size_t gw = std::max(gemv_workspace(1, m + 1, n), gemv_workspace(1, m, n));
  double* part = ws.take<double>(gw / sizeof(double) + 1);
  RBFX_REQUIRE(part != nullptr, "GMRES-IR workspace carve failed");
  const unsigned g = (unsigned)ceil_div(n, 256);
  int st;
  rbfx_ir_stats loc{0, 0, 0.0, 0, 0};

  auto host_dot = [&](const double* a, const double* b, double& out) -> int {
    int e = dot(n, a, b, dpart, hd + 2 * m + 6, s);
    if (e) return e;
    RBFX_HIP(hipMemcpyAsync(&out, hd + 2 * m + 6, sizeof(double), hipMemcpyDeviceToHost, s));
    RBFX_HIP(hipStreamSynchronize(s));
    return RBFX_OK;
  };

  double bn2 = 0.0;
  if ((st = host_dot(d_b, d_b, bn2))) return st;
  const double bn = sqrt(bn2);
  hipLaunchKernelGGL(fill_kernel, dim3(g), dim3(256), 0, s, n, 0.0, d_x);
  RBFX_CHECK_LAUNCH();
  if (bn == 0.0) {
    RBFX_HIP(hipStreamSynchronize(s));
    if (stats) *stats = loc;
    return RBFX_OK;
  }
  hipLaunchKernelGGL(scale_copy_kernel, dim3(g), dim3(256), 0, s, n, 1.0, d_b, r);
  RBFX_CHECK_LAUNCH();

  HessenbergLS ls(m);
  std::vector<double> hh(2 * m + 8), col(m + 2);
  // RBFX_IR_TRACE=1: one stderr line per cycle with the Arnoldi estimates
  static const bool trace = getenv("RBFX_IR_TRACE") && getenv("RBFX_IR_TRACE")[0] == '1';
  std::string tline;
  IrMonitor mon{o.tol, kStopEta, kFloorEta};
  int result = RBFX_ERR_STALL;
  for (int cycle = 0;; ++cycle) {
    if (cycle > 0) {
      // r = b - K x
      if ((st = res.applyK(d_x, w))) return st;
      hipLaunchKernelGGL(scale_copy_kernel, dim3(g), dim3(256), 0, s, n, 1.0, d_b, r);
      hipLaunchKernelGGL(axpby_kernel, dim3(g), dim3(256), 0, s, n, -1.0, w, 1.0, r);
      RBFX_CHECK_LAUNCH();
    }
    double rn2 = 0.0;
    if ((st = host_dot(r, r, rn2))) return st;
    const double rn = sqrt(rn2);
    const double rel = rn / bn;
    loc.outer_iters = cycle;
    loc.final_relres = rel;
    if (trace) {
      if (!tline.empty()) fprintf(stderr, "%s\n", tline.c_str());
      char buf[96];
      snprintf(buf, sizeof(buf), "[rbfx ir] n=%lld cycle %d relres %.3e inner %d:", (long long)n, cycle, rel,
               loc.inner_iters);
      tline = buf;
    }
    double eta = INFINITY, xn = -1.0;
    if (rel == rel && rel > o.tol && cycle > 0 && o.knorm > 0.0) {
      double xn2 = 0.0;
      if ((st = host_dot(d_x, d_x, xn2))) return st;
      xn = sqrt(xn2);
      eta = rn / (o.knorm * xn + bn);
      if (trace) {
        char buf[40];
        snprintf(buf, sizeof(buf), " (eta %.1e)", eta);
        tline += buf;
      }
    }
    const IrMonitor::Decision dec = mon.next(rel, eta);
    if (dec == IrMonitor::kConverged) {
      result = RBFX_OK;
      break;
    }
    if (dec == IrMonitor::kStalled) break;
    if (loc.inner_iters >= o.max_inner) break;

    // GMRES(m) on K M^-1 y = r, flexible form: Z keeps M^-1 V
    hipLaunchKernelGGL(scale_copy_kernel, dim3(g), dim3(256), 0, s, n, 1.0 / rn, r, V);
    RBFX_CHECK_LAUNCH();
    ls.start(rn);
    // The cycle's target: a residual at which the next cycle's test passes
    // (relative residual tol, or, once ||x|| is known, a backward error of
    // kStopEta with this ||x||), with a margin; but no deeper than one cycle
    // reaches in FP64 (1e-10 of its starting residual).
    // The first cycle may go to 1e-12 of its residual: it often meets the
    // tolerance itself, and the backward-error target (from ||Z y|| of its
    // current iterate, checked from 1e-8 on) stops it where it cannot.
    double target = 0.1 * o.tol * bn;
    if (xn >= 0.0) target = std::max(target, 0.5 * kStopEta * (o.knorm * xn + bn));
    const double inner_tol = std::max(target, (cycle == 0 ? 1e-12 : 1e-10) * rn);
    int kdone = 0;
    bool early_stall = false;
    mon.est_met = false;
    for (int k = 0; k < m; ++k) {
      double* vk = V + (size_t)k * n;
      if ((st = pre.apply(vk, Z + (size_t)k * n, w))) return st;
      // classical Gram-Schmidt, two passes
      double* h1 = hd;
      double* h2 = hd + (m + 1);
      if ((st = gemv(0, k + 1, n, 1.0, V, n, w, 0.0, h1, nullptr, s))) return st;
      if ((st = gemv(1, k + 1, n, -1.0, V, n, h1, 1.0, w, part, s))) return st;
      if ((st = gemv(0, k + 1, n, 1.0, V, n, w, 0.0, h2, nullptr, s))) return st;
      if ((st = gemv(1, k + 1, n, -1.0, V, n, h2, 1.0, w, part, s))) return st;
      if ((st = dot(n, w, w, dpart, hd + 2 * m + 4, s))) return st;
      RBFX_HIP(hipMemcpyAsync(hh.data(), hd, (2 * m + 6) * sizeof(double), hipMemcpyDeviceToHost, s));
      RBFX_HIP(hipStreamSynchronize(s));
      const double hn = sqrt(hh[2 * m + 4]);
      for (int i = 0; i <= k; ++i) col[i] = hh[i] + hh[m + 1 + i];
      col[k + 1] = hn;
      ++loc.inner_iters;
      kdone = k + 1;
      if (hn > 0.0) {
        hipLaunchKernelGGL(scale_copy_kernel, dim3(g), dim3(256), 0, s, n, 1.0 / hn, w, V + (size_t)(k + 1) * n);
        RBFX_CHECK_LAUNCH();
      }
      const double est = ls.add_column(k, col.data());
      if (trace) {
        char buf[24];
        snprintf(buf, sizeof(buf), " %.1e", est / bn);
        tline += buf;
      }
      double cycle_tol = inner_tol;
      if (cycle == 0 && o.knorm > 0.0 && est > inner_tol && est <= 1e-8 * rn) {
        // ||x|| of the cycle's current iterate x = Z y (the first cycle starts
        // from x = 0), for the backward-error target
        const double* yk = ls.solve(k + 1);
        RBFX_HIP(hipMemcpyAsync(hd, yk, (k + 1) * sizeof(double), hipMemcpyHostToDevice, s));
        if ((st = gemv(1, k + 1, n, 1.0, Z, n, hd, 0.0, u, part, s))) return st;
        double xk2 = 0.0;
        if ((st = host_dot(u, u, xk2))) return st;  // synchronises: yk outlives the copy
        // the backward-error target, or the rounding floor of this cycle's
        // residual (2^-53 ||K||_F ||x||), below which more iterations lower
        // the estimate but not the true residual
        const double xk = sqrt(xk2);
        cycle_tol = std::max({inner_tol, 0.5 * kStopEta * (o.knorm * xk + bn), 0x1p-53 * o.knorm * xk});
      }
      if (est <= cycle_tol || hn == 0.0) {
        mon.est_met = true;
        break;
      }
      if (loc.inner_iters >= o.max_inner) break;
      // Early stall: extrapolate the observed geometric rate of the residual
      // estimate; give up when it cannot reach the tolerance in the budget.
      if (k + 1 >= kRateWindow && loc.inner_iters + iterations_needed(rn, est, k + 1, inner_tol) > (double)o.max_inner) {
        early_stall = true;
        break;
      }
    }
    // y = R^-1 g ; x += Z y
    const double* y = ls.solve(kdone);
    RBFX_HIP(hipMemcpyAsync(hd, y, kdone * sizeof(double), hipMemcpyHostToDevice, s));
    if ((st = gemv(1, kdone, n, 1.0, Z, n, hd, 0.0, u, part, s))) return st;
    hipLaunchKernelGGL(axpby_kernel, dim3(g), dim3(256), 0, s, n, 1.0, u, 1.0, d_x);
    RBFX_CHECK_LAUNCH();
    RBFX_HIP(hipStreamSynchronize(s));  // y (host) must outlive the copy
    if (early_stall) {
      // report the residual of the improved iterate, then stop
      if ((st = res.applyK(d_x, w))) return st;
      hipLaunchKernelGGL(scale_copy_kernel, dim3(g), dim3(256), 0, s, n, 1.0, d_b, r);
      hipLaunchKernelGGL(axpby_kernel, dim3(g), dim3(256), 0, s, n, -1.0, w, 1.0, r);
      RBFX_CHECK_LAUNCH();
      double e2 = 0.0;
      if ((st = host_dot(r, r, e2))) return st;
      loc.final_relres = sqrt(e2) / bn;
      loc.outer_iters = cycle + 1;
      if (loc.final_relres <= o.tol) result = RBFX_OK;
      break;
    }
  }
  if (trace && !tline.empty()) fprintf(stderr, "%s -> %s\n", tline.c_str(), result == RBFX_OK ? "ok" : "stall");
  loc.stalled = result == RBFX_OK ? 0 : 1;
  if (stats) *stats = loc;
  if (result != RBFX_OK)
    set_error("GMRES-IR stalled: relres %.3e after %d inner iterations", loc.final_relres, loc.inner_iters);
  return result;
}

}  // namespace rbfx

// ---- FP32 LU of K as the preconditioner -------------------------------------
namespace {
struct StoredOrMatrixFreeK : ResidualOp {
  int trans;
  int64_t n;
  const rbfx_problem* p;
  const double* K64;
  int64_t ld64;
  double* part;
  size_t part_bytes;
  hipStream_t s;
  int applyK(const double* v, double* w) override {
    if (K64) return residual_gemv(trans, n, K64, ld64, v, w, part, s);
    return rbfx_matvec(p, RBFX_MAT_K, trans, v, w, part, part_bytes, s);
  }
};

struct LuPair : PairOp {
  int trans;
  int64_t n;
  const float* LU32;
  int64_t ld32;
  const int64_t* perm;
  const double* dinv;
  void* ts_work;
  size_t ts_bytes;
  ResidualOp* K;
  hipStream_t s;
  int apply(const double* v, double* z, double* w) override {
    hipLaunchKernelGGL(scale_copy_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, n, 1.0, v, z);
    RBFX_CHECK_LAUNCH();
    int st = rbfx_getrs_f32(trans, n, LU32, ld32, perm, dinv, z, ts_work, ts_bytes, s);
    if (st) return st;
    if ((st = rbfx_getrs_status(n, ts_work, ts_bytes, s))) return st;
    return K->applyK(z, w);
  }
};

size_t residual_ws_bytes(int64_t n) {
  return std::max({gemv_workspace(1, n, n), matvec_t_ws_bytes(n)});
}
}  // namespace

extern "C" size_t rbfx_gmres_ir_workspace(int64_t n, int32_t restart) {
  if (n <= 0 || restart <= 0) return 0;
  return fgmres_ws_bytes(n, restart) + ws_round(rbfx_getrs_workspace(n)) + ws_round(residual_ws_bytes(n)) + 1024;
}

extern "C" int rbfx_gmres_ir(int trans, int64_t n, const rbfx_problem* p, const double* d_K64, int64_t ld64,
                             const float* d_LU32, int64_t ld32, const int64_t* d_perm32,
                             const double* d_dinv32, const double* d_b, double* d_x, double tol,
                             int32_t restart, int32_t max_inner, void* d_work, size_t work_bytes,
                             rbfx_ir_stats* stats, void* stream) {
  RBFX_REQUIRE(n > 0 && restart > 0 && max_inner > 0 && tol > 0.0, "bad gmres_ir parameters");
  RBFX_REQUIRE(d_LU32 && d_perm32 && d_dinv32 && d_b && d_x && d_b != d_x, "bad gmres_ir pointers");
  RBFX_REQUIRE(d_K64 || p, "need a stored K or a problem for the residual");
  RBFX_REQUIRE(!d_K64 || ld64 >= n, "ld64 < n");
  RBFX_REQUIRE(!p || p->N + p->M == n, "problem size != n");
  RBFX_REQUIRE(work_bytes >= rbfx_gmres_ir_workspace(n, restart), "gmres_ir workspace too small");
  hipStream_t s = as_stream(stream);
  Workspace ws(d_work, work_bytes);
  const size_t fw = fgmres_ws_bytes(n, restart);
  void* fwork = ws.take<char>(fw);
  const size_t tsb = rbfx_getrs_workspace(n);
  void* tsw = ws.take<char>(tsb);
  const size_t rb = residual_ws_bytes(n);
  double* part = ws.take<double>(rb / sizeof(double) + 1);
  RBFX_REQUIRE(fwork && tsw && part, "gmres_ir workspace carve failed");
  StoredOrMatrixFreeK K;
  K.trans = trans;
  K.n = n;
  K.p = p;
  K.K64 = d_K64;
  K.ld64 = ld64;
  K.part = part;
  K.part_bytes = rb;
  K.s = s;
  LuPair pre;
  pre.trans = trans;
  pre.n = n;
  pre.LU32 = d_LU32;
  pre.ld32 = ld32;
  pre.perm = d_perm32;
  pre.dinv = d_dinv32;
  pre.ts_work = tsw;
  pre.ts_bytes = tsb;
  pre.K = &K;
  pre.s = s;
  FgmresOpts o;
  o.tol = tol;
  o.restart = restart;
  o.max_inner = max_inner;
  return fgmres_ir(pre, K, n, d_b, d_x, o, fwork, fw, stats, s);
}
