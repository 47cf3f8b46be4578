// K4: triangular solves with the LU factors (FP64 and FP32), A x = b and
// A^T x = b, for a single right-hand side.
//
// Replaces the solve phase of lineax QR (updes/operators.py:612-613) and the
// inv(A) @ v products (updes/assembly.py:410,430).  Blocked, synchronisation-
// free sweep: one workgroup per 64-row block row, taken in dependency order
// from an atomic ticket; each workgroup folds in the off-diagonal tiles of the
// blocks already solved (a monotone completion counter tells it how far it
// may go), then applies the precomputed inverse of its diagonal block and
// publishes its slice.  Row-major storage makes the no-transpose tiles
// row-contiguous and the transposed tiles column-coalesced.
#include <algorithm>
#include <climits>
#include <vector>

#include "common.h"

namespace rbfx {

int gemm_f64(int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda, const double* B,
             int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t s);

constexpr int TB = 64;  // block size of the sweep and of the diagonal inverses
constexpr int kTsThreads = 256;

__host__ __device__ inline int64_t imin64(int64_t a, int64_t b) { return a < b ? a : b; }

// ---- diagonal-block inverses ----------------------------------------------
// dinv[b][0] = inv(unit lower L_bb), dinv[b][1] = inv(U_bb), TB x TB row-major,
// identity-padded for the last partial block.
template <typename TF>
__global__ __launch_bounds__(TB) void diag_inv_kernel(const TF* __restrict__ LU, int64_t lda, int64_t n,
                                                      double* __restrict__ dinv) {
  typedef double T;
  __shared__ T S[TB][TB + 1];
  __shared__ T X[TB][TB + 1];
  const int64_t b = blockIdx.x;
  const int64_t o = b * TB;
  const int sz = (int)imin64(TB, n - o);
  const int c = threadIdx.x;
  for (int i = 0; i < TB; ++i) S[i][c] = (i < sz && c < sz) ? (T)LU[(o + i) * lda + o + c] : (i == c ? T(1) : T(0));
  __syncthreads();
  // L^-1 column c: x_i = -sum_{c<=k<i} L_ik x_k
  for (int i = 0; i < TB; ++i) X[i][c] = (i == c) ? T(1) : T(0);
  for (int i = c + 1; i < TB; ++i) {
    T s = 0;
    for (int k = c; k < i; ++k) s += S[i][k] * X[k][c];
    X[i][c] = (i < sz) ? -s : T(0);
  }
  T* Li = dinv + (b * 2) * TB * TB;
  for (int i = 0; i < TB; ++i) Li[i * TB + c] = X[i][c];
  __syncthreads();
  // U^-1 column c: x_i = (delta_ic - sum_{k>i} U_ik x_k) / U_ii
  for (int i = TB - 1; i >= 0; --i) {
    if (i > c) {
      X[i][c] = T(0);
      continue;
    }
    T s = (i == c) ? T(1) : T(0);
    for (int k = i + 1; k <= c; ++k) s -= S[i][k] * X[k][c];
    X[i][c] = s / S[i][i];
  }
  T* Ui = dinv + (b * 2 + 1) * TB * TB;
  for (int i = 0; i < TB; ++i) Ui[i * TB + c] = X[i][c];
}

// ---- permutation staging ------------------------------------------------------
// ctl[0] = ticket, ctl[1] = completion counter of the current sweep,
// ctl[2] = status of the whole solve (nonzero: a wait expired).  The solve's
// first gather clears all three; the reset before each sweep only the first
// two, so an expired wait in the first sweep survives the second.
__global__ void gather_perm_kernel(const double* __restrict__ b, const int64_t* __restrict__ perm,
                                   int64_t n, double* __restrict__ y, unsigned* __restrict__ ctl,
                                   int clear_status) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    ctl[0] = 0u;
    ctl[1] = 0u;
    if (clear_status) ctl[2] = 0u;
  }
  if (i < n) y[i] = b[perm ? perm[i] : i];
}

// Writes the solution back; a solve whose status is set writes NaN instead,
// so that a timed-out solve never reads as a result.
__global__ void scatter_perm_kernel(const double* __restrict__ y, const int64_t* __restrict__ perm,
                                    int64_t n, double* __restrict__ b, const unsigned* __restrict__ ctl) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool failed = ctl[2] != 0u;
  if (i < n) b[perm ? perm[i] : i] = failed ? __builtin_nan("") : y[i];
}

// ---- the sweep ---------------------------------------------------------------
// FWD: block rows processed 0..nb-1, else nb-1..0.
// TRANS=false: the tile (I,J) of the factor itself is used (row-contiguous);
// TRANS=true : the tile (J,I) transposed.
// WHICH: which diagonal inverse (0 = L, 1 = U) is applied.
// Factors are read in their storage type TF; the arithmetic is FP64, so FP32
// factors are applied exactly as stored (the FP32 preconditioner of GMRES-IR).
template <typename TF, bool FWD, bool TRANS, int WHICH>
__global__ __launch_bounds__(kTsThreads) void sweep_kernel(const TF* __restrict__ F, int64_t lda, int64_t n,
                                                           const double* __restrict__ dinv,
                                                           double* __restrict__ y, unsigned* __restrict__ ctl,
                                                           uint64_t wait_ticks) {
  typedef double T;
  __shared__ T acc[TB];
  __shared__ T part[4][TB];
  __shared__ T Ds[TB][TB + 1];
  __shared__ T red[TRANS ? 1 : TB][TB + 1];  // no-transpose partial sums, row x column
  __shared__ int s_I;
  __shared__ unsigned s_avail;
  const int tid = threadIdx.x;
  const int64_t nb = ceil_div(n, TB);
  unsigned* ticket = ctl;
  unsigned* done = ctl + 1;
  int* status = reinterpret_cast<int*>(ctl + 2);
  if (tid == 0) {
    const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_I = (int)t;
    s_avail = 0;
    // an earlier sweep of this solve gave up: nothing to compute
    if (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) s_I = INT_MAX;
  }
  __syncthreads();
  const int seq = s_I;
  if (seq >= nb) return;
  const int64_t I = FWD ? seq : nb - 1 - seq;
  const int64_t oI = I * TB;
  const int szI = (int)imin64(TB, n - oI);

  // Thread ownership of a tile, kept for every off-diagonal tile: lanes run
  // along the stored rows (coalesced), a wave owns 16 of the other index.
  //   no-transpose: column rr = tid%64 of block J, rows q*16 .. q*16+15 of
  //                 block I (16 partial row sums per thread, summed over the
  //                 lanes at the end)
  //   transpose:    column rr = tid%64 of block I, rows q*16 .. q*16+15 of block J
  const int rr = tid & 63;
  const int q = tid >> 6;

  // prefetch the diagonal inverse and the tile of the last dependency (the
  // one on the critical path) before waiting
  const double* D = dinv + (I * 2 + WHICH) * TB * TB;
  for (int e = tid; e < TB * TB; e += kTsThreads) Ds[e / TB][e % TB] = D[e];
  T pre[16];
  if (seq > 0) {
    const int64_t Jl = FWD ? I - 1 : I + 1;
    const int64_t oJ = Jl * TB;
    const int szJ = (int)imin64(TB, n - oJ);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int kk = q * 16 + k;
      if (TRANS)
        pre[k] = (rr < szI && kk < szJ) ? (T)F[(oJ + kk) * lda + oI + rr] : T(0);
      else
        pre[k] = (kk < szI && rr < szJ) ? (T)F[(oI + kk) * lda + oJ + rr] : T(0);
    }
  }

  // fold in the dependencies in completion order d = 0 .. seq-1 as they
  // become available (private partial sums per thread, no barrier per tile)
  T s = 0;
  T ps[TRANS ? 1 : 16];
#pragma unroll
  for (int k = 0; k < (TRANS ? 1 : 16); ++k) ps[k] = 0;
  int d = 0;
  while (d < seq) {
    if ((unsigned)d >= s_avail) {
      __syncthreads();
      if (tid == 0) {
        unsigned v;
        uint32_t spins = 0;
        WaitClock clk(wait_ticks);
        while ((v = ld_coh(done)) <= (unsigned)d) {
          __builtin_amdgcn_s_sleep(1);
          ++spins;
          if (clk.expired()) {
            __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            v = 0xffffffffu;
            break;
          }
          if ((spins & 255u) == 0 &&
              __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            v = 0xffffffffu;
            break;
          }
        }
        s_avail = v;
      }
      __syncthreads();
      if (s_avail == 0xffffffffu) return;
    }
    const int lim = (unsigned)seq < s_avail ? seq : (int)s_avail;
    // two tiles at a time while neither is the last dependency: all
    // their loads are issued before any is used (folding one tile at a time
    // is bound by the load latency once the factor no longer fits in the
    // Infinity Cache).  Every product is an explicit FMA in the same order as
    // tile by tile, so the result does not depend on how the tiles group
    constexpr int FB = 2;
    for (; d + FB <= lim && d + FB <= seq - 1; d += FB) {
      T f[FB][16];
      if constexpr (!TRANS) {
        T yk[FB];
#pragma unroll
        for (int u = 0; u < FB; ++u) {
          const int64_t oJ = (FWD ? d + u : nb - 1 - (d + u)) * TB;
          const bool ok = rr < (int)imin64(TB, n - oJ);
          yk[u] = ok ? ld_coh(y + oJ + rr) : T(0);
#pragma unroll
          for (int k = 0; k < 16; ++k)
            f[u][k] = (ok && q * 16 + k < szI) ? (T)F[(oI + q * 16 + k) * lda + oJ + rr] : T(0);
        }
#pragma unroll
        for (int u = 0; u < FB; ++u)
#pragma unroll
          for (int k = 0; k < 16; ++k) ps[k] = __builtin_fma(f[u][k], yk[u], ps[k]);
      } else {
        T yk[FB][16];
#pragma unroll
        for (int u = 0; u < FB; ++u) {
          const int64_t oJ = (FWD ? d + u : nb - 1 - (d + u)) * TB;
          const int szJ = (int)imin64(TB, n - oJ);
#pragma unroll
          for (int k = 0; k < 16; ++k) {
            const bool ok = rr < szI && q * 16 + k < szJ;
            f[u][k] = ok ? (T)F[(oJ + q * 16 + k) * lda + oI + rr] : T(0);
            yk[u][k] = ok ? ld_coh(y + oJ + q * 16 + k) : T(0);
          }
        }
#pragma unroll
        for (int u = 0; u < FB; ++u)
#pragma unroll
          for (int k = 0; k < 16; ++k) s = __builtin_fma(f[u][k], yk[u][k], s);
      }
    }
    for (; d < lim; ++d) {
      const int64_t J = FWD ? d : nb - 1 - d;
      const int64_t oJ = J * TB;
      const int szJ = (int)imin64(TB, n - oJ);
      if (!TRANS) {
        // one y entry per thread (its column), 16 coalesced row segments
        if (rr >= szJ) continue;
        const T yk = ld_coh(y + oJ + rr);
        if (d == seq - 1) {
#pragma unroll
          for (int k = 0; k < 16; ++k) ps[k] = __builtin_fma(pre[k], yk, ps[k]);
        } else {
#pragma unroll
          for (int k = 0; k < 16; ++k)
            if (q * 16 + k < szI) ps[k] = __builtin_fma((T)F[(oI + q * 16 + k) * lda + oJ + rr], yk, ps[k]);
        }
        continue;
      }
      if (rr >= szI) continue;
      if (d == seq - 1) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          const int kk = q * 16 + k;
          if (kk < szJ) s = __builtin_fma(pre[k], ld_coh(y + oJ + kk), s);
        }
      } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          const int kk = q * 16 + k;
          if (kk < szJ) s = __builtin_fma((T)F[(oJ + kk) * lda + oI + rr], ld_coh(y + oJ + kk), s);
        }
      }
    }
  }

  // acc = b_I - (sum of the folded tiles)
  if (!TRANS) {
    // row sums over the 64 column lanes: through LDS, 4 threads per row
#pragma unroll
    for (int k = 0; k < 16; ++k) red[q * 16 + k][rr] = ps[k];
    __syncthreads();
    const int r = tid >> 2, qq = tid & 3;
    T t = 0;
#pragma unroll
    for (int c = 0; c < 16; ++c) t += red[r][qq * 16 + c];
    t += __shfl_xor(t, 1, 64);
    t += __shfl_xor(t, 2, 64);
    if (qq == 0) acc[r] = r < szI ? y[oI + r] - t : T(0);
  } else {
    part[q][rr] = s;
    __syncthreads();
    if (tid < TB) acc[tid] = tid < szI ? y[oI + tid] - (part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid]) : T(0);
  }
  __syncthreads();
  // diagonal block: y_I = D acc (D = inverse of the diagonal block, or its transpose)
  {
    const int r = tid >> 2, qq = tid & 3;
    T t = 0;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const int cc = qq * 16 + c;
      t += (TRANS ? Ds[cc][r] : Ds[r][cc]) * acc[cc];
    }
    t += __shfl_xor(t, 1, 64);
    t += __shfl_xor(t, 2, 64);
    if (qq == 0 && r < szI) st_coh(y + oI + r, t);
  }
  // publish: the slice's coherent stores are complete before the counter moves
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid == 0) __hip_atomic_fetch_add(done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
static int launch_sweep(bool fwd, bool trans, int which, const T* F, int64_t lda, int64_t n,
                        const double* dinv, double* y, unsigned* ctl, hipStream_t s) {
  const unsigned nb = (unsigned)ceil_div(n, TB);
  // reset ticket/done before each sweep (the status stays)
  hipLaunchKernelGGL(gather_perm_kernel, dim3(1), dim3(1), 0, s, (const double*)nullptr,
                     (const int64_t*)nullptr, (int64_t)0, (double*)nullptr, ctl, 0);
  RBFX_CHECK_LAUNCH();
  const uint64_t ticks = device_info().wait_ticks;
  const int pid = prof_start(sizeof(T) == 8 ? "sweep_f64" : "sweep_f32", s);
#define RBFX_SWEEP(F_, T_, W_)                                                               \
  hipLaunchKernelGGL((sweep_kernel<T, F_, T_, W_>), dim3(nb), dim3(kTsThreads), 0, s, F, lda, n, \
                     dinv, y, ctl, ticks)
  if (fwd && !trans && which == 0) RBFX_SWEEP(true, false, 0);
  else if (!fwd && !trans && which == 1) RBFX_SWEEP(false, false, 1);
  else if (fwd && trans && which == 1) RBFX_SWEEP(true, true, 1);
  else if (!fwd && trans && which == 0) RBFX_SWEEP(false, true, 0);
  else {
    set_error("invalid sweep combination");
    return RBFX_ERR_ARG;
  }
#undef RBFX_SWEEP
  prof_stop(pid, s, (double)n * n, (double)sizeof(T) * n * (n + 1) / 2.0);
  RBFX_CHECK_LAUNCH();
  return RBFX_OK;
}

// Workspace of one solve: the permuted vector, then the control words.
static bool carve_getrs(void* work, size_t work_bytes, int64_t n, double*& y, unsigned*& ctl) {
  Workspace ws(work, work_bytes);
  y = ws.take<double>(n);
  ctl = ws.take<unsigned>(64);
  return y && ctl;
}

template <typename T>
static int getrs(int trans, int64_t n, const T* LU, int64_t lda, const int64_t* perm, const double* dinv,
                 double* b, void* work, size_t work_bytes, hipStream_t s) {
  RBFX_REQUIRE(n > 0 && lda >= n && LU && perm && dinv && b, "bad getrs args");
  double* y;
  unsigned* ctl;
  RBFX_REQUIRE(carve_getrs(work, work_bytes, n, y, ctl), "getrs workspace too small");
  const unsigned g = (unsigned)ceil_div(n, 256);
  int st;
  if (!trans) {
    hipLaunchKernelGGL(gather_perm_kernel, dim3(g), dim3(256), 0, s, b, perm, n, y, ctl, 1);
    RBFX_CHECK_LAUNCH();
    if ((st = launch_sweep<T>(true, false, 0, LU, lda, n, dinv, y, ctl, s))) return st;
    if ((st = launch_sweep<T>(false, false, 1, LU, lda, n, dinv, y, ctl, s))) return st;
    hipLaunchKernelGGL(scatter_perm_kernel, dim3(g), dim3(256), 0, s, y, (const int64_t*)nullptr, n, b,
                       ctl);
    RBFX_CHECK_LAUNCH();
  } else {
    hipLaunchKernelGGL(gather_perm_kernel, dim3(g), dim3(256), 0, s, b, (const int64_t*)nullptr, n, y,
                       ctl, 1);
    RBFX_CHECK_LAUNCH();
    if ((st = launch_sweep<T>(true, true, 1, LU, lda, n, dinv, y, ctl, s))) return st;
    if ((st = launch_sweep<T>(false, true, 0, LU, lda, n, dinv, y, ctl, s))) return st;
    hipLaunchKernelGGL(scatter_perm_kernel, dim3(g), dim3(256), 0, s, y, perm, n, b, ctl);
    RBFX_CHECK_LAUNCH();
  }
  return RBFX_OK;
}

// ---- many right-hand sides -------------------------------------------------------
__global__ void gather_rows_kernel(const double* __restrict__ B, int64_t ldb, const int64_t* __restrict__ perm,
                                   int64_t n, int64_t nrhs, double* __restrict__ T, int64_t ldt) {
  const int64_t i = blockIdx.y;
  const int64_t src = perm[i];
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nrhs; c += (int64_t)gridDim.x * blockDim.x)
    T[i * ldt + c] = B[src * ldb + c];
}

__global__ void copy_rows_kernel(const double* __restrict__ S, int64_t lds, int64_t rows, int64_t cols,
                                 double* __restrict__ D, int64_t ldd) {
  const int64_t i = blockIdx.y;
  if (i >= rows) return;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < cols; c += (int64_t)gridDim.x * blockDim.x)
    D[i * ldd + c] = S[i * lds + c];
}

static size_t getrs_multi_ws(int64_t n, int64_t nrhs) {
  const int64_t ld = ceil_div(nrhs, 8) * 8;
  return ws_round(sizeof(double) * (size_t)n * ld) + ws_round(sizeof(double) * (size_t)TB * ld) + 256;
}

// A X = B for nrhs right-hand sides (B n x nrhs row-major, overwritten by X)
// with FP64 LU factors: rows gathered by the permutation, then blocked
// forward and backward substitution over the TB-row blocks, each block one
// MFMA GEMM against the blocks already solved and one with its precomputed
// diagonal inverse.
static int getrs_multi(int64_t n, const double* LU, int64_t lda, const int64_t* perm, const double* dinv,
                       int64_t nrhs, double* B, int64_t ldb, void* work, size_t work_bytes, hipStream_t s) {
  RBFX_REQUIRE(n > 0 && nrhs > 0 && lda >= n && ldb >= nrhs && LU && perm && dinv && B, "bad getrs_multi args");
  RBFX_REQUIRE(work && work_bytes >= getrs_multi_ws(n, nrhs), "getrs_multi workspace too small");
  const int64_t ldt = ceil_div(nrhs, 8) * 8;
  Workspace ws(work, work_bytes);
  double* T = ws.take<double>((size_t)n * ldt);
  double* W = ws.take<double>((size_t)TB * ldt);
  RBFX_REQUIRE(T && W, "getrs_multi workspace carve failed");
  const unsigned gx = (unsigned)std::min<int64_t>(ceil_div(nrhs, 256), 64);
  RBFX_REQUIRE(n < (int64_t)65536 * 65536, "getrs_multi: n too large");
  hipLaunchKernelGGL(gather_rows_kernel, dim3(gx, (unsigned)n), dim3(256), 0, s, B, ldb, perm, n, nrhs, T, ldt);
  RBFX_CHECK_LAUNCH();
  const int64_t nb = ceil_div(n, TB);
  int st;
  auto diag = [&](int64_t blk, int which) -> int {
    const int64_t o = blk * TB;
    const int64_t sz = imin64(TB, n - o);
    const double* D = dinv + (blk * 2 + which) * TB * TB;
    if ((st = gemm_f64(sz, nrhs, sz, 1.0, D, TB, T + o * ldt, ldt, 0.0, W, ldt, s))) return st;
    hipLaunchKernelGGL(copy_rows_kernel, dim3(gx, (unsigned)sz), dim3(256), 0, s, W, ldt, sz, nrhs, T + o * ldt, ldt);
    RBFX_CHECK_LAUNCH();
    return RBFX_OK;
  };
  for (int64_t blk = 0; blk < nb; ++blk) {  // L (unit lower)
    const int64_t o = blk * TB;
    const int64_t sz = imin64(TB, n - o);
    if (o > 0 && (st = gemm_f64(sz, nrhs, o, -1.0, LU + o * lda, lda, T, ldt, 1.0, T + o * ldt, ldt, s)))
      return st;
    if ((st = diag(blk, 0))) return st;
  }
  for (int64_t blk = nb - 1; blk >= 0; --blk) {  // U
    const int64_t o = blk * TB;
    const int64_t sz = imin64(TB, n - o);
    const int64_t e = o + sz;
    if (e < n && (st = gemm_f64(sz, nrhs, n - e, -1.0, LU + o * lda + e, lda, T + e * ldt, ldt, 1.0, T + o * ldt,
                                ldt, s)))
      return st;
    if ((st = diag(blk, 1))) return st;
  }
  hipLaunchKernelGGL(copy_rows_kernel, dim3(gx, (unsigned)n), dim3(256), 0, s, T, ldt, n, nrhs, B, ldb);
  RBFX_CHECK_LAUNCH();
  return RBFX_OK;
}

// Status of the last solve queued with this workspace (synchronises).
static int getrs_status(int64_t n, void* work, size_t work_bytes, hipStream_t s) {
  RBFX_REQUIRE(n > 0 && work, "bad getrs_status args");
  double* y;
  unsigned* ctl;
  RBFX_REQUIRE(carve_getrs(work, work_bytes, n, y, ctl), "getrs workspace too small");
  unsigned st = 0;
  RBFX_HIP(hipMemcpyAsync(&st, ctl + 2, sizeof(st), hipMemcpyDeviceToHost, s));
  RBFX_HIP(hipStreamSynchronize(s));
  if (st != 0u) {
    set_error("triangular solve: a device-side wait expired after %.0f s (solution set to NaN)", kWaitSeconds);
    return RBFX_ERR_TIMEOUT;
  }
  return RBFX_OK;
}

template <typename T>
static int prepare(int64_t n, const T* LU, int64_t lda, const int64_t* ipiv, int64_t* perm, double* dinv,
                   hipStream_t s) {
  RBFX_REQUIRE(n > 0 && lda >= n && LU && ipiv && perm && dinv, "bad getrs_prepare args");
  std::vector<int64_t> hp(n), pm(n);
  RBFX_HIP(hipMemcpyAsync(hp.data(), ipiv, n * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  RBFX_HIP(hipStreamSynchronize(s));
  for (int64_t i = 0; i < n; ++i) pm[i] = i;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t p = hp[i];
    RBFX_REQUIRE(p >= i && p < n, "invalid pivot ipiv[%lld] = %lld", (long long)i, (long long)p);
    std::swap(pm[i], pm[p]);
  }
  RBFX_HIP(hipMemcpyAsync(perm, pm.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL((diag_inv_kernel<T>), dim3((unsigned)ceil_div(n, TB)), dim3(TB), 0, s, LU, lda, n, dinv);
  RBFX_CHECK_LAUNCH();
  RBFX_HIP(hipStreamSynchronize(s));  // pm must outlive the copy
  return RBFX_OK;
}

}  // namespace rbfx

using namespace rbfx;

extern "C" size_t rbfx_getrs_multi_workspace(int64_t n, int64_t nrhs) {
  return (n <= 0 || nrhs <= 0) ? 0 : getrs_multi_ws(n, nrhs);
}

extern "C" int rbfx_getrs_multi_f64(int64_t n, const double* d_LU, int64_t lda, const int64_t* d_perm,
                                    const double* d_dinv, int64_t nrhs, double* d_B, int64_t ldb, void* d_work,
                                    size_t work_bytes, void* stream) {
  return getrs_multi(n, d_LU, lda, d_perm, d_dinv, nrhs, d_B, ldb, d_work, work_bytes, as_stream(stream));
}

extern "C" int64_t rbfx_getrs_dinv_elems(int64_t n) { return n <= 0 ? 0 : ceil_div(n, TB) * 2 * TB * TB; }

extern "C" size_t rbfx_getrs_workspace(int64_t n) {
  return ws_round((size_t)n * sizeof(double)) + ws_round(64 * sizeof(unsigned)) + 512;
}

extern "C" int rbfx_getrs_prepare_f64(int64_t n, const double* d_LU, int64_t lda, const int64_t* d_ipiv,
                                      int64_t* d_perm, double* d_dinv, void* d_work, size_t work_bytes,
                                      void* stream) {
  (void)d_work;
  (void)work_bytes;
  return prepare<double>(n, d_LU, lda, d_ipiv, d_perm, d_dinv, as_stream(stream));
}

extern "C" int rbfx_getrs_prepare_f32(int64_t n, const float* d_LU, int64_t lda, const int64_t* d_ipiv,
                                      int64_t* d_perm, double* d_dinv, void* d_work, size_t work_bytes,
                                      void* stream) {
  (void)d_work;
  (void)work_bytes;
  return prepare<float>(n, d_LU, lda, d_ipiv, d_perm, d_dinv, as_stream(stream));
}

extern "C" int rbfx_getrs_f64(int trans, int64_t n, const double* d_LU, int64_t lda, const int64_t* d_perm,
                              const double* d_dinv, double* d_b, void* d_work, size_t work_bytes,
                              void* stream) {
  return getrs<double>(trans, n, d_LU, lda, d_perm, d_dinv, d_b, d_work, work_bytes, as_stream(stream));
}

extern "C" int rbfx_getrs_status(int64_t n, void* d_work, size_t work_bytes, void* stream) {
  return getrs_status(n, d_work, work_bytes, as_stream(stream));
}

extern "C" int rbfx_getrs_f32(int trans, int64_t n, const float* d_LU, int64_t lda, const int64_t* d_perm,
                              const double* d_dinv, double* d_b, void* d_work, size_t work_bytes,
                              void* stream) {
  return getrs<float>(trans, n, d_LU, lda, d_perm, d_dinv, d_b, d_work, work_bytes, as_stream(stream));
}
