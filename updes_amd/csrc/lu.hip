// K2/K3: right-looking blocked LU with partial pivoting (FP64 and FP32) for
// row-major matrices on gfx950.
//
// Replaces the factorisations that the reference runs through XLA/LAPACK:
// jnp.linalg.inv (updes/assembly.py:90, getrf + getri) and lineax QR
// (updes/operators.py:612-613).  The augmented collocation matrix is factored
// once: P K = L U.
//
// Structure per outer panel of NB columns (k0 .. k0+NB):
//   1. recursive panel factorisation: halves down to 64-wide leaves; each
//      leaf runs as ONE launch of co-resident workgroups that keep their rows
//      of the leaf in registers and agree on every pivot through one exchange per
//      column (every workgroup publishes its best row and a tagged record,
//      all of them poll the records and pick the same winner);
//      the leaves apply their interchanges to the panel's other columns and
//      invert their diagonal blocks; the splits join those inverses;
//   2. row interchanges of the columns outside the panel through a
//      gather/scatter plan (row-major rows are contiguous, so a swap is a
//      coalesced copy);
//   3. U12 = inv(L11) A12 with the MFMA GEMM;
//   4. trailing update A22 -= L21 U12 with an MFMA GEMM
//      (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32, 128x128 tiles,
//      k-tiles loaded global -> LDS by DMA, XCD-aware tile order).
#include <algorithm>
#include <atomic>
#include <mutex>
#include <type_traits>
#include <vector>

#include "common.h"

namespace rbfx {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
struct Mma;
// MFMA 16x16x4 operand layout (both types): lane l holds A[l%16][l/16] and
// B[l/16][l%16].  Accumulator element v of lane l is D[row(l,v)][l%16] with
//   f32: row = 4*(l/16) + v        f64: row = 4*v + l/16
// (measured on gfx950 with one-hot operands).
template <>
struct Mma<double> {
  typedef v4d acc_t;
  static constexpr int BK = 8;
  __device__ static inline acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int v) { return 4 * v + (lane >> 4); }
};
template <>
struct Mma<float> {
  typedef v4f acc_t;
  static constexpr int BK = 16;
  __device__ static inline acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// ---------------------------------------------------------------------------
// GEMM: C = beta*C + alpha*A*B, row-major, beta in {0,1}.
// Block tile 128x128, 4 waves (2x2), wave tile 64x64 = 4x4 MFMA 16x16 tiles.
constexpr int GBM = 128, GBN = 128, GTHREADS = 256;

// Register-staged version: the k-tiles of A and B are loaded two tiles
// ahead into two register slots and stored to an LDS double buffer.  It
// serves the shapes and alpha values the LDS-DMA kernel below does not take.
// Two waves per SIMD (<= 256 registers each) so one wave's barrier or load
// wait overlaps the other's MFMAs.
// USE only separates the kernel symbols of the trailing update (0), of the
// triangular solve (1), of the diagonal-block inverse (2) and of the updates
// inside a recursive panel (3) in profiles; the code is the same.
template <typename T, int USE>
__global__ __launch_bounds__(GTHREADS) __attribute__((amdgpu_waves_per_eu(2, 2))) void gemm_kernel(
                                                        int64_t m, int64_t n, int64_t k, T alpha,
                                                        const T* __restrict__ A, int64_t lda,
                                                        const T* __restrict__ B, int64_t ldb,
                                                        T beta, T* __restrict__ C, int64_t ldc,
                                                        int64_t tiles_m, int64_t tiles_n) {
  constexpr int BK = Mma<T>::BK;
  constexpr int AV = GBM * BK / GTHREADS;  // A elements per thread per k-tile
  constexpr int BV = BK * GBN / GTHREADS;  // B elements per thread per k-tile
  // double-buffered k-tiles: the next tile is written while this one is read
  __shared__ T As[2][BK][GBM];
  __shared__ T Bs[2][BK][GBN];

  // Tile order.  Tiles are numbered in groups of 8 tile rows, column by
  // column inside a group, so the ~64 workgroups resident on one XCD at a
  // time cover an 8 x 8 patch and share its A and B tiles in that XCD's L2.
  // Consecutive workgroup ids land on different XCDs (round robin), so XCD x
  // takes the x-th contiguous eighth of that numbering.
  const int64_t T_ = tiles_m * tiles_n;
  const int64_t L = blockIdx.x;
  const int64_t per = (T_ + 7) / 8;
  const int64_t tile = (L % 8) * per + L / 8;
  if (tile >= T_) return;
  constexpr int64_t kGroupM = 8;
  const int64_t gwidth = kGroupM * tiles_n;
  const int64_t first_m = (tile / gwidth) * kGroupM;
  const int64_t gm = tiles_m - first_m < kGroupM ? tiles_m - first_m : kGroupM;
  const int64_t tm = first_m + (tile % gwidth) % gm;
  const int64_t tn = (tile % gwidth) / gm;
  const int64_t row0 = tm * GBM, col0 = tn * GBN;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  const int wm = wid >> 1, wn = wid & 1;

  // The accumulators start from beta*C (beta is 0 or 1) and alpha is folded
  // into the A tile, so the epilogue only stores (no C reads held live).
  typename Mma<T>::acc_t acc[4][4];
  const bool add = beta != T(0);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t c = col0 + wn * 64 + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t r = row0 + wm * 64 + i * 16 + Mma<T>::row(lane, v);
        acc[i][j][v] = (add && r < m && c < n) ? C[r * ldc + c] : T(0);
      }
    }

  // global -> register staging maps
  // A tile (GBM x BK): thread -> row ar = tid % GBM, k chunk ak = (tid / GBM) * AV
  const int ar = tid % GBM, ak = (tid / GBM) * AV;
  // B tile (BK x GBN): thread -> k row bk = tid / (GBN / BV), col chunk bc
  const int bk = tid / (GBN / BV), bc = (tid % (GBN / BV)) * BV;
  // Two register slots: the loads of k-tile t+2 are issued while tile t is
  // multiplied, so a load has two tiles of MFMA work to arrive.
  T ra0[AV], rb0[BV], ra1[AV], rb1[BV];

  auto load = [&](int64_t kt, T (&ra)[AV], T (&rb)[BV]) {
    const int64_t gr = row0 + ar;
    const int64_t gk = kt + ak;
    const T* pa = A + gr * lda + gk;
    if (gr < m && gk + AV <= k) {
#pragma unroll
      for (int q = 0; q < AV; ++q) ra[q] = pa[q];
    } else {
#pragma unroll
      for (int q = 0; q < AV; ++q) ra[q] = (gr < m && gk + q < k) ? pa[q] : T(0);
    }
    const int64_t gkb = kt + bk;
    const int64_t gc = col0 + bc;
    const T* pb = B + gkb * ldb + gc;
    if (gkb < k && gc + BV <= n) {
#pragma unroll
      for (int q = 0; q < BV; ++q) rb[q] = pb[q];
    } else {
#pragma unroll
      for (int q = 0; q < BV; ++q) rb[q] = (gkb < k && gc + q < n) ? pb[q] : T(0);
    }
  };
  auto store = [&](int buf, const T (&ra)[AV], const T (&rb)[BV]) {
#pragma unroll
    for (int q = 0; q < AV; ++q) As[buf][ak + q][ar] = alpha * ra[q];
#pragma unroll
    for (int q = 0; q < BV; ++q) Bs[buf][bk][bc + q] = rb[q];
  };
  auto compute = [&](int buf) {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = As[buf][kk + (lane >> 4)][wm * 64 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = Bs[buf][kk + (lane >> 4)][wn * 64 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Mma<T>::mma(af[i], bf[j], acc[i][j]);
    }
  };

  load(0, ra0, rb0);
  if (BK < k) load(BK, ra1, rb1);
  store(0, ra0, rb0);
  __syncthreads();
  // LDS buffer 0 holds the even k-tiles, buffer 1 the odd ones; slot 0 the
  // register copy of the next even tile, slot 1 of the next odd one.  A
  // buffer is rewritten only after the barrier that follows its last reads.
  for (int64_t kt = 0; kt < k; kt += 2 * BK) {
    if (kt + 2 * BK < k) load(kt + 2 * BK, ra0, rb0);
    compute(0);
    if (kt + BK < k) store(1, ra1, rb1);
    __syncthreads();
    if (kt + BK >= k) break;
    if (kt + 3 * BK < k) load(kt + 3 * BK, ra1, rb1);
    compute(1);
    if (kt + 2 * BK < k) store(0, ra0, rb0);
    __syncthreads();
  }

  // epilogue: plain stores
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t c = col0 + wn * 64 + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t r = row0 + wm * 64 + i * 16 + Mma<T>::row(lane, v);
        if (r < m && c < n) C[r * ldc + c] = acc[i][j][v];
      }
    }
  }
}

// GEMM with LDS-DMA operand loads (global_load_lds_dwordx4): the k-tiles go
// from global memory straight into a ring of kGemmStages LDS stages, so no
// registers stage them and a tile is requested kGemmStages-1 tiles before it
// is multiplied.  One barrier per k-tile.
//
// Layouts (E = 16 / sizeof(T) elements per 16-byte chunk):
//   A stage: 128 rows x BK, chunk (m, g) = A[m][E g .. E g + E) at position
//     4 m + (g ^ ((m >> 2) & 3))  (the xor keeps the 16-lane groups of the
//     operand reads on distinct banks);
//   B stage: BK rows x 128, chunk (k, q) = B[k][E q .. E q + E) at k*(128/E) + q.
// MFMA k-assignment: lane group g = lane/16 of MFMA s (0 <= s < E) holds
// k = E g + s, so one 16-byte read gives a lane its A values of all E MFMAs.
// For B one 16-byte read gives E adjacent columns, which feed E adjacent
// n-tiles: tile j's column for lane column c is 16E (j/E) + E c + j%E.
// Requirements (checked by the host): k % BK == 0, n % E == 0, lda, ldb and
// ldc multiples of E, 16-byte aligned A, B and C.
constexpr int kGemmStages = 4;

// BT x BT block tile (BT = 128, or 64 for small FP64 products), 2 x 2 waves
// of (BT/2) x (BT/2), each TW x TW MFMA 16x16 tiles.
template <typename T, int USE, int BT, int STAGES = kGemmStages>
__device__ __forceinline__ void gemm_dma_body(int64_t m, int64_t n, int64_t k, T alpha, const T* __restrict__ A,
                                              int64_t lda, const T* __restrict__ B, int64_t ldb, T beta,
                                              T* __restrict__ C, int64_t ldc, int64_t tiles_m, int64_t tiles_n) {
  constexpr int BK = Mma<T>::BK;
  constexpr int E = 16 / (int)sizeof(T);
  constexpr int TW = BT / 32;      // MFMA tiles per wave and dimension
  constexpr int HB = BT / 2;       // wave tile extent
  constexpr int NG = BT / E;       // chunks per B row
  constexpr int NCH = 4 * BT;      // chunks per operand and stage
  constexpr int RC = NCH / GTHREADS;  // chunks per thread and operand
  static_assert(4 * E == BK, "one 16-byte chunk per lane group and k-tile");
  static_assert(BT * (BK / E) == NCH && BK * NG == NCH, "chunk counts");
  static_assert(TW % E == 0 || E == 2, "E adjacent n-tiles per 16-byte B read");
  static_assert(TW >= E, "E adjacent n-tiles per 16-byte B read");
  typedef T vec_t __attribute__((ext_vector_type(E)));
  __shared__ __align__(16) T sm[STAGES][2][NCH * E];

  // tile order as in gemm_kernel (groups of 8 tile rows, XCD-contiguous)
  const int64_t T_ = tiles_m * tiles_n;
  const int64_t L = blockIdx.x;
  const int64_t per = (T_ + 7) / 8;
  const int64_t tile = (L % 8) * per + L / 8;
  if (tile >= T_) return;
  constexpr int64_t kGroupM = 8;
  const int64_t gwidth = kGroupM * tiles_n;
  const int64_t first_m = (tile / gwidth) * kGroupM;
  const int64_t gm = tiles_m - first_m < kGroupM ? tiles_m - first_m : kGroupM;
  const int64_t tm = first_m + (tile % gwidth) % gm;
  const int64_t tn = (tile % gwidth) / gm;
  const int64_t row0 = tm * BT, col0 = tn * BT;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  const int wm = wid >> 1, wn = wid & 1;
  const int lg = lane >> 4, lc = lane & 15;

  // DMA sources of this thread (RC A chunks, RC B chunks per k-tile); rows
  // past m are clamped (their results are never stored)
  const T* srcA[RC];
  const T* srcB[RC];
#pragma unroll
  for (int r = 0; r < RC; ++r) {
    const int p = (r * 4 + wid) * 64 + lane;
    const int mm = p >> 2, g = (p & 3) ^ ((mm >> 2) & 3);
    const int64_t row = row0 + mm < m ? row0 + mm : m - 1;
    srcA[r] = A + row * lda + g * E;
    const int kk = p / NG, q = p % NG;
    const int64_t col = col0 + (int64_t)q * E < n ? col0 + (int64_t)q * E : n - E;
    srcB[r] = B + (int64_t)kk * ldb + col;
  }
  auto issue = [&](int64_t kt, int st) {
#pragma unroll
    for (int r = 0; r < RC; ++r) {
      const int base = (r * 4 + wid) * 64 * E;
      __builtin_amdgcn_global_load_lds((const void*)(srcA[r] + kt),
                                       (__attribute__((address_space(3))) void*)(&sm[st][0][base]), 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void*)(srcB[r] + kt * ldb),
                                       (__attribute__((address_space(3))) void*)(&sm[st][1][base]), 16, 0, 0);
    }
  };

  typename Mma<T>::acc_t acc[TW][TW];
  const bool add = beta != T(0);
  // alpha is +-1 here: the sum runs on alpha*C and alpha is applied at the end
  const bool neg = alpha < T(0);
#pragma unroll
  for (int i = 0; i < TW; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t r = row0 + wm * HB + i * 16 + Mma<T>::row(lane, v);
#pragma unroll
      for (int jg = 0; jg < TW / E; ++jg) {
        const int64_t c = col0 + wn * HB + 16 * E * jg + E * lc;
        vec_t x = {};
        if (add && r < m && c < n) x = *reinterpret_cast<const vec_t*>(C + r * ldc + c);
#pragma unroll
        for (int e = 0; e < E; ++e) acc[i][jg * E + e][v] = neg ? -x[e] : x[e];
      }
    }

  const int64_t nt = k / BK;
#pragma unroll
  for (int t = 0; t < STAGES - 1; ++t)
    if (t < nt) issue((int64_t)t * BK, t);

  for (int64_t t = 0; t < nt; ++t) {
    // tile t has landed once at most (tiles issued after it) x 2 RC loads of
    // this thread are outstanding; the barrier makes every thread's part
    // visible and ends all reads of the stage refilled below
    const int64_t ahead = nt - 1 - t < STAGES - 2 ? nt - 1 - t : STAGES - 2;
    if (ahead >= 2)
      __builtin_amdgcn_s_waitcnt(0xF70 | (4 * RC));
    else if (ahead == 1)
      __builtin_amdgcn_s_waitcnt(0xF70 | (2 * RC));
    else
      __builtin_amdgcn_s_waitcnt(0xF70);
    // A plain barrier: __syncthreads() would also wait for the loads of the
    // tiles still in flight (its release fence covers the LDS-DMA).  Every
    // LDS read of this wave has completed (the MFMAs consumed them).
    __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    if (t + STAGES - 1 < nt) issue((t + STAGES - 1) * BK, (int)((t + STAGES - 1) % STAGES));
    const int st = (int)(t % STAGES);
    const T* As = sm[st][0];
    const T* Bs = sm[st][1];
    vec_t a[TW], b[TW / E][E];
#pragma unroll
    for (int i = 0; i < TW; ++i) {
      const int mm = wm * HB + i * 16 + lc;
      a[i] = *reinterpret_cast<const vec_t*>(As + (mm * 4 + (lg ^ ((mm >> 2) & 3))) * E);
    }
#pragma unroll
    for (int jg = 0; jg < TW / E; ++jg)
#pragma unroll
      for (int s2 = 0; s2 < E; ++s2)
        b[jg][s2] = *reinterpret_cast<const vec_t*>(Bs + ((E * lg + s2) * NG + wn * HB / E + 16 * jg + lc) * E);
#pragma unroll
    for (int s2 = 0; s2 < E; ++s2)
#pragma unroll
      for (int i = 0; i < TW; ++i)
#pragma unroll
        for (int j = 0; j < TW; ++j) acc[i][j] = Mma<T>::mma(a[i][s2], b[j / E][s2][j % E], acc[i][j]);
  }

#pragma unroll
  for (int i = 0; i < TW; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t r = row0 + wm * HB + i * 16 + Mma<T>::row(lane, v);
      if (r >= m) continue;
#pragma unroll
      for (int jg = 0; jg < TW / E; ++jg) {
        const int64_t c = col0 + wn * HB + 16 * E * jg + E * lc;
        if (c >= n) continue;
        vec_t x;
#pragma unroll
        for (int e = 0; e < E; ++e) x[e] = neg ? -acc[i][jg * E + e][v] : acc[i][jg * E + e][v];
        *reinterpret_cast<vec_t*>(C + r * ldc + c) = x;
      }
    }
}

template <typename T, int USE, int STAGES = kGemmStages>
__global__ __launch_bounds__(GTHREADS) __attribute__((amdgpu_waves_per_eu(2, 2))) void gemm_dma_kernel(
    int64_t m, int64_t n, int64_t k, T alpha, const T* __restrict__ A, int64_t lda, const T* __restrict__ B,
    int64_t ldb, T beta, T* __restrict__ C, int64_t ldc, int64_t tiles_m, int64_t tiles_n) {
  gemm_dma_body<T, USE, 128, STAGES>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tiles_m, tiles_n);
}

// 64 x 64 tiles for FP64 products too small to fill the GPU with 128 x 128
// tiles (the in-panel updates, inverses and triangular solves of late
// panels): four times the workgroups, a quarter of the k-loop work each.
template <int USE>
__global__ __launch_bounds__(GTHREADS) void gemm_dma64_kernel(int64_t m, int64_t n, int64_t k, double alpha,
                                                             const double* __restrict__ A, int64_t lda,
                                                             const double* __restrict__ B, int64_t ldb,
                                                             double beta, double* __restrict__ C, int64_t ldc,
                                                             int64_t tiles_m, int64_t tiles_n) {
  gemm_dma_body<double, USE, 64>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tiles_m, tiles_n);
}

// small_lds: an FP32 product of the panel stream, which runs beside the
// trailing update: a two-stage ring (32 KB of LDS), so that its workgroups
// fit beside two update workgroups on a CU instead of waiting for one to end.
template <typename T, int USE = 0>
static int gemm_launch(int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, const T* B,
                       int64_t ldb, T beta, T* C, int64_t ldc, hipStream_t s, bool small_lds = false) {
  if (m <= 0 || n <= 0) return RBFX_OK;
  const int64_t tm = ceil_div(m, GBM), tn = ceil_div(n, GBN);
  const int64_t tiles = tm * tn;
  const int64_t grid = ((tiles + 7) / 8) * 8;
  constexpr int E = 16 / (int)sizeof(T);
  auto al16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  const bool dma = (alpha == T(1) || alpha == T(-1)) && k > 0 && k % Mma<T>::BK == 0 && n >= E && n % E == 0 && lda % E == 0 && ldb % E == 0 &&
                   ldc % E == 0 && al16(A) && al16(B) && al16(C);
  if (dma && sizeof(T) == 8 && tiles < 2 * (int64_t)device_info().num_cus) {
    const int64_t tm64 = ceil_div(m, (int64_t)64), tn64 = ceil_div(n, (int64_t)64);
    const int64_t grid64 = ((tm64 * tn64 + 7) / 8) * 8;
    hipLaunchKernelGGL((gemm_dma64_kernel<USE>), dim3((unsigned)grid64), dim3(GTHREADS), 0, s, m, n, k,
                       (double)alpha, reinterpret_cast<const double*>(A), lda, reinterpret_cast<const double*>(B),
                       ldb, (double)beta, reinterpret_cast<double*>(C), ldc, tm64, tn64);
  } else if (dma && small_lds && sizeof(T) == 4)
    hipLaunchKernelGGL((gemm_dma_kernel<T, USE, 2>), dim3((unsigned)grid), dim3(GTHREADS), 0, s, m, n, k, alpha, A,
                       lda, B, ldb, beta, C, ldc, tm, tn);
  else if (dma)
    hipLaunchKernelGGL((gemm_dma_kernel<T, USE>), dim3((unsigned)grid), dim3(GTHREADS), 0, s, m, n, k, alpha, A,
                       lda, B, ldb, beta, C, ldc, tm, tn);
  else
    hipLaunchKernelGGL((gemm_kernel<T, USE>), dim3((unsigned)grid), dim3(GTHREADS), 0, s, m, n, k, alpha, A,
                       lda, B, ldb, beta, C, ldc, tm, tn);
  RBFX_CHECK_LAUNCH();
  return RBFX_OK;
}

template <typename T>
static int gemm(int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, const T* B,
                int64_t ldb, T beta, T* C, int64_t ldc, hipStream_t s, bool in_panel = false) {
  if (m <= 0 || n <= 0) return RBFX_OK;
  // profile names: the trailing update of the factorisation (and direct
  // calls) vs the updates inside a recursive panel
  const int pid = prof_start(in_panel ? (sizeof(T) == 8 ? "gemm_panel_f64" : "gemm_panel_f32")
                                      : (sizeof(T) == 8 ? "gemm_f64" : "gemm_f32"),
                             s);
  const int st = in_panel ? gemm_launch<T, 3>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, true)
                          : gemm_launch<T, 0>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
  prof_stop(pid, s, 2.0 * m * n * k, (double)sizeof(T) * (2.0 * m * n + (double)m * k + (double)k * n));
  return st;
}


template <typename T>
struct PanelArgs {
  T* A;           // panel top-left (absolute row r0, column r0)
  int64_t lda;
  int64_t m;      // rows in the panel (r0 .. r0+m)
  int w;          // panel width <= kPanelWMax
  int64_t R;      // rows per workgroup
  int64_t r0;     // absolute row of the panel's first row
  int64_t left, right;  // columns of the enclosing outer panel left / right of the leaf
  int64_t* ipiv;  // (w) absolute pivot rows
  double* cand;   // [2][G][kRecWords] tagged (|value|, row) records (first two words used)
  T* cand_row;    // [2][G][w]
  unsigned seq_base;  // step count before this launch
  int* status;
  int32_t* info;
  uint64_t wait_ticks;  // real-time bound of each wait (wall_clock64 ticks)
};

constexpr int kPanelThreads = 256;
constexpr int kPanelWMax = 64;
// Narrow leaves for the panels that run beside a long trailing update: a
// 16-wide FP64 leaf (67 VGPRs) shares a CU with two trailing-update
// workgroups (2 x 216 VGPRs per SIMD), where a 64-wide one (164) displaces
// one of them.  More leaves and splits per panel, hidden under the update
// while it is long: panels of at least kNarrowLeafRows rows (measured best
// at n = 40,006 among 20k..36k).  FP32 only uses them on request: a 64-wide
// FP32 leaf (95 VGPRs) already fits beside two update workgroups, and the
// FP32 update is too short to hide the narrow leaves' extra launches
// (40k mixed step: 615 ms without, 621 ms from 28,000 rows, 625 ms from
// 20,000).
constexpr int kPanelWNarrow = 16;
constexpr int64_t kNarrowLeafRows = 28000;
static std::atomic<int64_t> g_narrow_leaf_rows{0};
// A pivot record (two 64-bit words) owns a 128-byte line: every workgroup
// polls every record, and records sharing a line would serialise the polls.
constexpr int kRecWords = 16;

// (value, index) maximum with the smaller index winning ties.
__device__ inline void argmax_merge(double& v, int64_t& i, double ov, int64_t oi) {
  if (ov > v || (ov == v && oi < i)) {
    v = ov;
    i = oi;
  }
}

// Pivot keys.  A candidate (|value|, row) is compared as one 64-bit key:
// the bits of |value| with the low kKeyLow bits replaced by the position of
// the row, inverted so that a lower position wins a tie of the truncated
// values:  bits 11..19 = 511 - workgroup, bits 0..10 = 2047 - row within the
// workgroup.  |value| >= 0, so its bits order like the values; key 0 means
// "no candidate" (no active row, or NaN).  The choice is partial pivoting on
// values rounded down to 32 mantissa bits, ties to the lowest row.
constexpr int kKeyLow = 20;
constexpr unsigned long long kKeyValMask = ~((1ull << kKeyLow) - 1);
constexpr int kKeyMaxGroups = 512, kKeyMaxRows = 2048;

__device__ inline unsigned long long pivot_key(double absval, int q, int local) {
  if (!(absval >= 0.0)) return 0ull;
  return ((unsigned long long)__double_as_longlong(absval) & kKeyValMask) |
         ((unsigned long long)(kKeyMaxGroups - 1 - q) << 11) | (unsigned long long)(kKeyMaxRows - 1 - local);
}

// Maximum of a 64-bit key over the wave with DPP row shifts and broadcasts
// (no LDS traffic): an inclusive max-scan inside each row of 16 lanes, then
// row 15 -> rows 1, 3 and lane 31 -> rows 2, 3, so lane 63 holds the result.
template <int CTRL, int ROWMASK>
__device__ inline unsigned long long dpp_u64(unsigned long long v) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)v, CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(v >> 32), CTRL, ROWMASK, 0xf, false);
  return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}
__device__ inline unsigned long long umax64(unsigned long long a, unsigned long long b) { return a > b ? a : b; }
__device__ inline unsigned long long wave_max_key(unsigned long long v) {
  v = umax64(v, dpp_u64<0x111, 0xf>(v));  // row_shr:1
  v = umax64(v, dpp_u64<0x112, 0xf>(v));  // row_shr:2
  v = umax64(v, dpp_u64<0x114, 0xf>(v));  // row_shr:4
  v = umax64(v, dpp_u64<0x118, 0xf>(v));  // row_shr:8
  v = umax64(v, dpp_u64<0x142, 0xa>(v));  // row_bcast:15
  v = umax64(v, dpp_u64<0x143, 0xc>(v));  // row_bcast:31
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), 63);
  return ((unsigned long long)hi << 32) | lo;
}

// Block-wide key maximum; one barrier.  s has one slot per wave and is
// reused by a later call only after at least one more barrier.
__device__ inline unsigned long long block_max_key(unsigned long long key, unsigned long long* s) {
  const unsigned long long wm = wave_max_key(key);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = wm;
  __syncthreads();
  unsigned long long b = s[0];
#pragma unroll
  for (int q = 1; q < kPanelThreads / 64; ++q) b = umax64(b, s[q]);
  return b;
}

// Position bookkeeping of the panel's row interchanges, replayed from the
// winners (every workgroup knows them) by one wave.  Only moved rows are
// listed: pos(p) = p unless p is in the list.
struct MoveList {
  int64_t phys[2 * kPanelWMax];
  int64_t pos[2 * kPanelWMax];
  int cnt;
};

// Orders lane 0's LDS writes before the wave's following LDS reads.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline int ml_find(const int64_t* keys, int cnt, int64_t key) {
  const int lane = threadIdx.x & 63;
  for (int base = 0; base < cnt; base += 64) {
    const int q = base + lane;
    const unsigned long long bal = __ballot(q < cnt && keys[q] == key);
    if (bal) return base + __ffsll((long long)bal) - 1;
  }
  return -1;
}

// Staging of the interchanges applied to the panel's other columns.
constexpr int kSwapStage = 2048;

// One leaf of the panel: columns [0, w) (w <= kPanelWMax) of rows [0, m) at
// a.A, factored by G workgroups of kPanelThreads threads.  Thread t of
// workgroup g holds rows g*R + t + q*kPanelThreads (q < ROWS) in registers.
// The column loop runs over all kPanelWMax columns and rotates every row's
// registers left by one after each column, so the current column is always
// register 0 (no run-time register indexing) and after the loop the rows
// are back in column order.  Columns >= w are zero and only rotated.
template <typename T, int ROWS, int W>
__global__ __launch_bounds__(kPanelThreads) void panel_kernel(PanelArgs<T> a) {
  static_assert(W <= kPanelWMax, "leaf wider than the candidate rows");
  __shared__ T urow[2 * W];  // winner row of the current column; zeros from w on
  __shared__ unsigned long long s_k1[kPanelThreads / 64], s_k2[kPanelThreads / 64];
  __shared__ int64_t s_win[W];
  __shared__ int s_zero[W];
  __shared__ int s_bad;
  __shared__ MoveList ml;
  __shared__ T stage[kSwapStage];

  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
  const int w = a.w;
  const int64_t rb = (int64_t)g * a.R;
  const int64_t re = rb + a.R < a.m ? rb + a.R : a.m;
  const int nr = re > rb ? (int)(re - rb) : 0;

  T P[ROWS][W];
  bool act[ROWS];
#pragma unroll
  for (int q = 0; q < ROWS; ++q) {
    const int local = tid + q * kPanelThreads;
    act[q] = local < nr;
    const T* src = a.A + (rb + local) * a.lda;
#pragma unroll
    for (int c = 0; c < W; ++c) P[q][c] = (act[q] && c < w) ? src[c] : T(0);
  }
  for (int c = tid; c < 2 * W; c += kPanelThreads) urow[c] = T(0);

  // key of this thread's best active row in register 0 (the current column)
  auto thread_key = [&]() -> unsigned long long {
    unsigned long long k = 0;
#pragma unroll
    for (int q = 0; q < ROWS; ++q)
      if (act[q]) k = umax64(k, pivot_key((double)fabs(P[q][0]), 0, tid + q * kPanelThreads));
    return k;
  };
  unsigned long long* recs = reinterpret_cast<unsigned long long*>(a.cand);

  // rotate every row left by one: column j+1 into register 0, column j to
  // the top
  auto rotate = [&]() {
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
      const T t0 = P[q][0];
#pragma unroll
      for (int c = 0; c + 1 < W; ++c) P[q][c] = P[q][c + 1];
      P[q][W - 1] = t0;
    }
  };
  for (int j = 0; j < w; ++j) {
    const int par = j & 1;
    const unsigned long long tag = (unsigned long long)((a.seq_base + (unsigned)(j + 1)) & 0xFFFFu);
    // 1. best active row of this workgroup for column j
    if (tid == 0) s_bad = 0;
    const unsigned long long k1 = block_max_key(thread_key(), s_k1);
    const int blocal = kKeyMaxRows - 1 - (int)(k1 & (kKeyMaxRows - 1));
    // 2. publish it.  The owner of that row stores the row (absolute column
    //    order), waits until the stores have completed, then stores the
    //    tagged record
    //      w0 = bits of |value| with the 16 low mantissa bits set to the tag,
    //      w1 = tag << 48 | row.
    //    A reader takes a record only when both words carry the tag of
    //    column j, so the record needs no separate flag and its two stores
    //    no order.  With no candidate, thread 0 publishes value -1.
    const bool owner = k1 != 0 && (blocal % kPanelThreads) == tid;
    if (owner) {
      T* dst = a.cand_row + ((int64_t)par * G + g) * w;
      const int qo = blocal / kPanelThreads;
#pragma unroll
      for (int q = 0; q < ROWS; ++q)
        if (q == qo) {
#pragma unroll
          for (int c = 0; c < W; ++c)
            if (j + c < w) st_coh(dst + j + c, P[q][c]);
        }
      __builtin_amdgcn_s_waitcnt(0);
    }
    if (owner || (k1 == 0 && tid == 0)) {
      const double bv = k1 ? __longlong_as_double((long long)(k1 & kKeyValMask)) : -1.0;
      const int64_t bi = k1 ? rb + blocal : INT64_MAX;
      unsigned long long* rec = recs + kRecWords * ((int64_t)par * G + g);
      st_coh(rec, (unsigned long long)(__double_as_longlong(bv) & ~0xFFFFll) | tag);
      st_coh(rec + 1, (tag << 48) | ((unsigned long long)bi & 0xFFFFFFFFFFFFull));
    }
    // 3.-4. every workgroup's record of column j (one thread each) and the
    //    winner across workgroups
    unsigned long long k2 = 0;
    for (int q = tid; q < G; q += kPanelThreads) {
      const unsigned long long* o = recs + kRecWords * ((int64_t)par * G + q);
      unsigned long long w0 = ld_coh(o), w1 = ld_coh(o + 1);
      uint32_t spins = 0;
      WaitClock clk(a.wait_ticks);
      while ((w0 & 0xFFFFull) != tag || (w1 >> 48) != tag) {
        __builtin_amdgcn_s_sleep(1);
        ++spins;
        if ((spins & 255u) == 0 &&
            __hip_atomic_load(a.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          s_bad = 1;
          break;
        }
        if (clk.expired()) {
          __hip_atomic_store(a.status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          s_bad = 1;
          break;
        }
        w0 = ld_coh(o);
        w1 = ld_coh(o + 1);
      }
      const double ov = __longlong_as_double((long long)(w0 & ~0xFFFFull));
      if (ov >= 0.0)
        k2 = umax64(k2, pivot_key(ov, q, (int)((int64_t)(w1 & 0xFFFFFFFFFFFFull) - (int64_t)q * a.R)));
    }
    k2 = block_max_key(k2, s_k2);
    if (s_bad) return;
    const int wq = kKeyMaxGroups - 1 - (int)((k2 >> 11) & (kKeyMaxGroups - 1));
    const int64_t wi = (int64_t)wq * a.R + (kKeyMaxRows - 1 - (int)(k2 & (kKeyMaxRows - 1)));
    if (k2 != 0 && (wq >= G || wi >= a.m)) {
      // a key that names no row cannot come from the records: stop
      if (tid == 0) __hip_atomic_store(a.status, 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    if (k2 == 0) {
      // no finite candidate anywhere (NaN column): every workgroup sees the
      // same key, so all of them leave here together
      if (g == 0 && tid == 0) {
        *a.info = -1;
        __hip_atomic_store(a.status, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      return;
    }
    if (tid < w) urow[tid] = ld_coh(a.cand_row + ((int64_t)par * G + wq) * w + tid);
    if (tid == 0) s_win[j] = wi;
#pragma unroll
    for (int q = 0; q < ROWS; ++q)
      if (rb + tid + q * kPanelThreads == wi) act[q] = false;
    __syncthreads();
    // 5. eliminate column j from the active rows: register c holds column
    //    j + c (urow beyond w is zero, which also leaves the finished
    //    columns rotated to the top untouched)
    const T piv = urow[j];
    const T rpiv = piv != T(0) ? T(1) / piv : T(0);
    if (tid == 0) s_zero[j] = piv == T(0);
    // Branch-free: an inactive row takes multiplier 0 (a divergent branch
    // here makes the compiler keep a second copy of the row registers).
    // Eliminated and rotated in one pass: register c takes column j+1+c,
    // the top register the multiplier of column j.
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
      const T p0 = P[q][0];
      const T l = act[q] ? p0 * rpiv : T(0);
#pragma unroll
      for (int c = 0; c + 1 < W; ++c) P[q][c] = P[q][c + 1] - l * urow[j + 1 + c];
      P[q][W - 1] = act[q] ? l : p0;
    }
  }
  // the remaining rotations bring the columns back into register order
  for (int j = w; j < W; ++j) rotate();

  // Replay the interchanges (LAPACK order: at step j rows at positions j and
  // pos(winner) swap) to find every row's final position; wave 0 only.
  if (tid < 64) {
    if (tid == 0) ml.cnt = 0;
    wave_sync();
    for (int j = 0; j < w; ++j) {
      const int64_t p = s_win[j];
      const int cnt = ml.cnt;
      const int ip = ml_find(ml.phys, cnt, p);
      const int64_t pp = ip >= 0 ? ml.pos[ip] : p;         // position of the winner
      const int iq = ml_find(ml.pos, cnt, (int64_t)j);
      const int64_t q = iq >= 0 ? ml.phys[iq] : (int64_t)j; // row at position j
      if (tid == 0) {
        int k = ip;
        if (k < 0) {
          k = ml.cnt++;
          ml.phys[k] = p;
        }
        ml.pos[k] = j;
        if (q != p) {
          // no listed row at position j means its occupant is row j itself,
          // which is then not listed either
          int kq = iq;
          if (kq < 0) {
            kq = ml.cnt++;
            ml.phys[kq] = q;
          }
          ml.pos[kq] = pp;
        }
        if (g == 0) {
          a.ipiv[j] = a.r0 + pp;
          if (s_zero[j] && *a.info == 0) *a.info = (int32_t)(a.r0 + j + 1);
        }
      }
      wave_sync();
    }
  }
  __syncthreads();
  // every row of this workgroup to its final position
  const int cnt = ml.cnt;
#pragma unroll
  for (int q = 0; q < ROWS; ++q) {
    const int local = tid + q * kPanelThreads;
    if (local >= nr) continue;
    int64_t dest = rb + local;
    for (int k = 0; k < cnt; ++k)
      if (ml.phys[k] == rb + local) {
        dest = ml.pos[k];
        break;
      }
    T* dst = a.A + dest * a.lda;
#pragma unroll
    for (int c = 0; c < W; ++c)
      if (c < w) dst[c] = P[q][c];
  }
  // The same interchanges on the outer panel's other columns (left and right
  // of this leaf), so the recursion above needs no row swaps of its own.
  // Each workgroup takes a disjoint slice of those columns, reads the moved
  // rows of its slice into LDS, then writes them to their final positions.
  const int64_t nout = a.left + a.right;
  if (cnt == 0 || nout == 0) return;
  const int64_t per = (nout + G - 1) / G;
  const int64_t o0 = (int64_t)g * per;
  const int64_t o1 = o0 + per < nout ? o0 + per : nout;
  const int64_t chunk = kSwapStage / cnt;  // columns per pass that fit the stage
  for (int64_t c0 = o0; c0 < o1; c0 += chunk) {
    const int64_t nc = c0 + chunk < o1 ? chunk : o1 - c0;
    for (int64_t e = tid; e < (int64_t)cnt * nc; e += kPanelThreads) {
      const int k = (int)(e / nc);
      const int64_t oc = c0 + e % nc;
      const int64_t col = oc < a.left ? oc - a.left : oc - a.left + w;
      stage[e] = a.A[ml.phys[k] * a.lda + col];
    }
    __syncthreads();
    for (int64_t e = tid; e < (int64_t)cnt * nc; e += kPanelThreads) {
      const int k = (int)(e / nc);
      const int64_t oc = c0 + e % nc;
      const int64_t col = oc < a.left ? oc - a.left : oc - a.left + w;
      a.A[ml.pos[k] * a.lda + col] = stage[e];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Row interchange plan: the net permutation of the sequential swaps
// ipiv[s..e) (absolute rows, ipiv[i] >= i) as (dst, src) pairs.
constexpr int kSwapMax = 512;

__global__ __launch_bounds__(64) void swap_plan_kernel(const int64_t* __restrict__ ipiv, int64_t s,
                                                       int64_t e, int64_t* __restrict__ dst,
                                                       int64_t* __restrict__ src,
                                                       int* __restrict__ count) {
  __shared__ int64_t in_src[kSwapMax];
  __shared__ int64_t out_pos[kSwapMax];
  __shared__ int64_t out_src[kSwapMax];
  __shared__ int64_t piv[kSwapMax];
  __shared__ int nout;
  const int lane = threadIdx.x;
  const int k = (int)(e - s);
  for (int t = lane; t < k; t += 64) {
    in_src[t] = s + t;
    piv[t] = ipiv[s + t];
  }
  if (lane == 0) nout = 0;
  __syncthreads();
  for (int i = 0; i < k; ++i) {
    const int64_t p = piv[i];
    if (p < e) {
      if (lane == 0 && p != s + i) {
        const int64_t tmp = in_src[i];
        in_src[i] = in_src[p - s];
        in_src[p - s] = tmp;
      }
    } else {
      // find p among out_pos (wave-parallel)
      int found = -1;
      for (int base = 0; base < nout; base += 64) {
        const int q = base + lane;
        const bool hit = q < nout && out_pos[q] == p;
        const unsigned long long bal = __ballot(hit);
        if (bal) {
          found = base + __ffsll((long long)bal) - 1;
          break;
        }
      }
      if (lane == 0) {
        int q = found;
        if (q < 0) {
          q = nout++;
          out_pos[q] = p;
          out_src[q] = p;
        }
        const int64_t tmp = in_src[i];
        in_src[i] = out_src[q];
        out_src[q] = tmp;
      }
    }
    __syncthreads();
  }
  if (lane == 0) {
    int c = 0;
    for (int t = 0; t < k; ++t)
      if (in_src[t] != s + t) {
        dst[c] = s + t;
        src[c] = in_src[t];
        ++c;
      }
    for (int q = 0; q < nout; ++q)
      if (out_src[q] != out_pos[q]) {
        dst[c] = out_pos[q];
        src[c] = out_src[q];
        ++c;
      }
    *count = c;
  }
}

// This workgroup's share (blockIdx.x of gridDim.x) of a copy of nc
// entries, with 16-byte accesses when both ends are 16-byte aligned.
template <typename T>
__device__ inline void copy_row_part(const T* __restrict__ s, T* __restrict__ d, int64_t nc) {
  constexpr int E = 16 / sizeof(T);
  typedef T V __attribute__((ext_vector_type(E)));
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (((reinterpret_cast<uintptr_t>(s) | reinterpret_cast<uintptr_t>(d)) & 15) == 0) {
    const int64_t nv = nc / E;
    for (int64_t v = t0; v < nv; v += stride) reinterpret_cast<V*>(d)[v] = reinterpret_cast<const V*>(s)[v];
    for (int64_t c = nv * E + t0; c < nc; c += stride) d[c] = s[c];
  } else {
    for (int64_t c = t0; c < nc; c += stride) d[c] = s[c];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void swap_gather_kernel(const T* __restrict__ A, int64_t lda,
                                                          int64_t c0, int64_t nc,
                                                          const int64_t* __restrict__ src,
                                                          const int* __restrict__ count,
                                                          T* __restrict__ tmp) {
  const int q = blockIdx.y;
  if (q >= *count) return;
  copy_row_part<T>(A + src[q] * lda + c0, tmp + (int64_t)q * nc, nc);
}

template <typename T>
__global__ __launch_bounds__(256) void swap_scatter_kernel(T* __restrict__ A, int64_t lda, int64_t c0,
                                                           int64_t nc,
                                                           const int64_t* __restrict__ dst,
                                                           const int* __restrict__ count,
                                                           const T* __restrict__ tmp) {
  const int q = blockIdx.y;
  if (q >= *count) return;
  copy_row_part<T>(tmp + (int64_t)q * nc, A + dst[q] * lda + c0, nc);
}

// ---------------------------------------------------------------------------
// U12 = inv(L11) A12 as one MFMA GEMM.  inv(L11) (unit lower, W <= kNB) is
// built by the panel recursion itself: every leaf inverts its own diagonal
// block (trtri_diag_kernel, FP64 arithmetic rounded to T) and every split
// of the recursion joins the inverses of its halves with two GEMMs,
// X21 = -X22 (L21 X11).
constexpr int kTri = 64;
constexpr int kNB = 512;  // outer panel width of the blocked LU

// Inverse of the sz x sz (sz <= 64) unit-lower block at L into X[0:sz, 0:sz]
// (zeros above the diagonal included), in FP64 arithmetic.  Blocked by 16:
//   1. each wave inverts one 16 x 16 diagonal block by forward substitution,
//      the solved row broadcast across the wave by lane shuffles;
//   2. block rows I = 1..3 in turn: S = L[I, 0:I] X[0:I, 0:I], then
//      X[I, 0:I] = -X_II S (two barriers per block row).
// Rows and columns from sz to 64 are padded with the identity.
// NT = 16 for the 16-wide leaves: the 16 x 16 substitution alone, without
// the three empty block-row joins of the 64 x 64 case.
template <typename T, int NT>
__global__ __launch_bounds__(256) void trtri_diag_kernel(const T* __restrict__ L, int64_t ldl, int sz,
                                                        T* __restrict__ X, int64_t ldx) {
  __shared__ double Ls[NT * (NT - 1) / 2];  // strictly lower part, packed by rows
  __shared__ double Xs[NT][NT + 1];
  __shared__ double Ss[16][NT + 1];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  auto lidx = [](int i, int k) { return i * (i - 1) / 2 + k; };
  // whole rows per pass (coalesced), all 16 loads in flight
  T lv[NT * NT / 256];
#pragma unroll
  for (int q = 0; q < NT * NT / 256; ++q) {
    const int e = q * 256 + t, i = e / NT, k = e % NT;
    lv[q] = (k < i && i < sz) ? L[(int64_t)i * ldl + k] : T(0);
  }
#pragma unroll
  for (int q = 0; q < NT * NT / 256; ++q) {
    const int e = q * 256 + t, i = e / NT, k = e % NT;
    if (k < i) Ls[lidx(i, k)] = (double)lv[q];
    Xs[i][k] = 0.0;
  }
  __syncthreads();
  if (16 * wv < NT) {
    const int c = lane & 15, r4 = lane >> 4, o = 16 * wv;
    double y[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) y[q] = (4 * q + r4 == c) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < 15; ++k) {
      const double xk = __shfl(y[k >> 2], ((k & 3) << 4) | c, 64);  // X[k][c]
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = 4 * q + r4;
        if (i > k) y[q] -= Ls[lidx(o + i, o + k)] * xk;
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) Xs[o + 4 * q + r4][o + c] = y[q];
  }
  __syncthreads();
  for (int I = 1; I < NT / 16; ++I) {
    const int nI = 16 * I, oI = 16 * I;
    for (int e = t; e < 16 * nI; e += 256) {
      const int i = e / nI, j = e % nI;
      double acc = 0.0;
      for (int k = j; k < nI; ++k) acc += Ls[lidx(oI + i, k)] * Xs[k][j];  // X is lower: k >= j
      Ss[i][j] = acc;
    }
    __syncthreads();
    for (int e = t; e < 16 * nI; e += 256) {
      const int i = e / nI, j = e % nI;
      double acc = 0.0;
      for (int k = 0; k <= i; ++k) acc += Xs[oI + i][oI + k] * Ss[k][j];
      Xs[oI + i][j] = -acc;
    }
    __syncthreads();
  }
  for (int e = t; e < sz * sz; e += 256) {
    const int i = e / sz, c = e % sz;
    X[(int64_t)i * ldx + c] = (T)Xs[i][c];
  }
}

// dst (rows x cols, ld ldd) = src (ld lds)
template <typename T>
__global__ __launch_bounds__(256) void copy_block_kernel(const T* __restrict__ src, int64_t lds, int rows,
                                                        int64_t cols, T* __restrict__ dst, int64_t ldd) {
  const int r = blockIdx.y;
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < cols; c += (int64_t)gridDim.x * 256)
    dst[(int64_t)r * ldd + c] = src[(int64_t)r * lds + c];
}

__global__ void zero_u32_kernel(unsigned* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// A failed or timed-out panel launch leaves status != 0: report it as
// info = -1 (the factors are then invalid).
__global__ void status_to_info_kernel(const int* status, int32_t* info) {
  if (threadIdx.x == 0 && *status != 0) *info = -1;
}

// ---------------------------------------------------------------------------
// Driver.
// Scratch of one stream of the factorisation (row-swap plan and staging,
// inverse of the diagonal block).
template <typename T>
struct LuScratch {
  int64_t* plan_dst;
  int64_t* plan_src;
  int* plan_count;
  int64_t plan_s = -1, plan_e = -1;  // swap range whose plan is in plan_*
  T* swap_tmp;
  int64_t swap_tmp_elems;
  T* tri_inv;  // [kNB][kNB] inverse of the panel's diagonal block
};

template <typename T>
struct LuCtx {
  T* A;
  int64_t n, lda;
  int64_t* ipiv;
  int32_t* info;
  hipStream_t s;
  // workspace
  double* cand;     // [2][G_max] (|value|, row) records of the panel launches
  T* cand_row;      // [2][G_max][kPanelWMax]
  int* status;
  unsigned seq;     // steps taken by the panel launches so far
  LuScratch<T> sc;
  int G_max;
  int wleaf = kPanelWMax;  // leaf width of the recursion
  int64_t pc0 = 0;  // outer panel being factored: columns [pc0, pc0 + pw)
  int pw = 0;
  bool on_panel_stream = false;  // the panel stream's context (runs beside the trailing update)
};

// Upper bound on the workgroups of one leaf launch (0: the CU count).
static std::atomic<int> g_panel_groups{0};
constexpr int kPanelMaxRows = 3;  // rows per thread of the register-resident leaf

template <typename T>
static int leaf(LuCtx<T>& c, int64_t r0, int w) {
  const DeviceInfo& di = device_info();
  const int64_t m = c.n - r0;
  int cap = g_panel_groups.load(std::memory_order_relaxed);
  if (cap <= 0 && sizeof(T) == 4) {
    // FP32: at most a quarter of the CUs (more only where three rows per
    // thread do not reach m).  Fewer workgroups, each with up to three rows
    // per thread, exchange faster beside the trailing update: 40k mixed step
    // 616 ms with up to 256 workgroups, 601 / 596 / 593 ms with 96 / 80 / 64.
    const int64_t need = ceil_div(m, (int64_t)kPanelThreads * kPanelMaxRows);
    cap = (int)std::max<int64_t>(di.num_cus / 4, need);
  }
  const int64_t gmax =
      std::min<int64_t>(cap > 0 ? std::min(cap, di.num_cus) : di.num_cus, kKeyMaxGroups - 1);
  // As few workgroups as one row per thread allows (every column's exchange
  // polls all G records); more rows per thread only when the CUs run out.
  LeafShape shape;
  RBFX_REQUIRE(leaf_shape(m, gmax, kPanelThreads, kPanelMaxRows, shape),
               "matrix too large for the register-resident panel (m=%lld)", (long long)m);
  const int rows = shape.rows;
  const int G = shape.G;
  const int64_t R = shape.R;
  PanelArgs<T> a;
  a.A = c.A + r0 * c.lda + r0;
  a.lda = c.lda;
  a.m = m;
  a.w = w;
  a.R = R;
  a.r0 = r0;
  a.left = r0 - c.pc0;
  a.right = c.pc0 + c.pw - (r0 + w);
  RBFX_REQUIRE(a.left >= 0 && a.right >= 0, "leaf outside its outer panel");
  RBFX_REQUIRE(w >= 1 && w <= kPanelWMax, "leaf width %d out of range", w);
  a.ipiv = c.ipiv + r0;
  a.cand = c.cand;
  a.cand_row = c.cand_row;
  a.seq_base = c.seq;
  a.status = c.status;
  a.info = c.info;
  a.wait_ticks = di.wait_ticks;
  RBFX_REQUIRE(G <= c.G_max && G < kKeyMaxGroups && R <= (int64_t)kPanelThreads * rows && R < kKeyMaxRows,
               "too many panel workgroups or rows");
  const int pid = prof_start(sizeof(T) == 8 ? "panel_f64" : "panel_f32", c.s);
  // A plain launch on the panel stream.  The workgroups wait for each
  // other, which needs all G of them resident at once: G is at most the CU
  // count and a workgroup fits an empty CU, so the ones not yet resident
  // start as soon as the trailing-update workgroups on some CU retire;
  // every wait is bounded in any case.  (A cooperative launch guarantees
  // the same but goes through a separate device queue, which added ~60 us
  // of dispatch latency per leaf, measured.)
  if (w <= kPanelWNarrow) {
    if (rows == 1)
      hipLaunchKernelGGL((panel_kernel<T, 1, kPanelWNarrow>), dim3(G), dim3(kPanelThreads), 0, c.s, a);
    else if (rows == 2)
      hipLaunchKernelGGL((panel_kernel<T, 2, kPanelWNarrow>), dim3(G), dim3(kPanelThreads), 0, c.s, a);
    else
      hipLaunchKernelGGL((panel_kernel<T, 3, kPanelWNarrow>), dim3(G), dim3(kPanelThreads), 0, c.s, a);
  } else {
    if (rows == 1)
      hipLaunchKernelGGL((panel_kernel<T, 1, kPanelWMax>), dim3(G), dim3(kPanelThreads), 0, c.s, a);
    else if (rows == 2)
      hipLaunchKernelGGL((panel_kernel<T, 2, kPanelWMax>), dim3(G), dim3(kPanelThreads), 0, c.s, a);
    else
      hipLaunchKernelGGL((panel_kernel<T, 3, kPanelWMax>), dim3(G), dim3(kPanelThreads), 0, c.s, a);
  }
  RBFX_CHECK_LAUNCH();
  prof_stop(pid, c.s, (double)m * w * w, 2.0 * sizeof(T) * m * w);
  c.seq += (unsigned)w;
  return RBFX_OK;
}

// Net row-interchange plan of the swaps ipiv[s..e) into dst's plan buffers,
// enqueued on c's stream.
template <typename T>
static int make_plan(LuCtx<T>& c, LuScratch<T>& dst, int64_t s, int64_t e) {
  RBFX_REQUIRE(e - s <= kSwapMax, "too many swaps in one plan");
  hipLaunchKernelGGL(swap_plan_kernel, dim3(1), dim3(64), 0, c.s, c.ipiv, s, e, dst.plan_dst, dst.plan_src,
                     dst.plan_count);
  RBFX_CHECK_LAUNCH();
  dst.plan_s = s;
  dst.plan_e = e;
  return RBFX_OK;
}

// Inverse of a leaf's w x w unit-lower diagonal block (at (r0, r0)) into X,
// enqueued on c's stream.
template <typename T>
static int leaf_inverse(LuCtx<T>& c, int64_t r0, int w, T* X, int64_t ldx) {
  RBFX_REQUIRE(w <= kTri, "leaf wider than the diagonal inverse kernel");
  const int pid = prof_start(sizeof(T) == 8 ? "trtri_f64" : "trtri_f32", c.s);
  if (w <= 16)
    hipLaunchKernelGGL((trtri_diag_kernel<T, 16>), dim3(1), dim3(256), 0, c.s, c.A + r0 * c.lda + r0, c.lda, w,
                       X, ldx);
  else
    hipLaunchKernelGGL((trtri_diag_kernel<T, kTri>), dim3(1), dim3(256), 0, c.s, c.A + r0 * c.lda + r0, c.lda, w,
                       X, ldx);
  RBFX_CHECK_LAUNCH();
  prof_stop(pid, c.s, (double)w * w * w / 3.0, (double)sizeof(T) * w * w);
  return RBFX_OK;
}

// Apply the swaps ipiv[s..e) to columns [c0, c0+nc).
template <typename T>
static int apply_swaps(LuCtx<T>& c, int64_t s, int64_t e, int64_t c0, int64_t nc) {
  if (nc <= 0 || e <= s) return RBFX_OK;
  int st;
  if ((c.sc.plan_s != s || c.sc.plan_e != e) && (st = make_plan(c, c.sc, s, e))) return st;
  const int maxcnt = (int)(2 * (e - s));
  const int pid = prof_start("row_swaps", c.s);
  // process the columns in slices that fit the staging buffer
  const int64_t slice = std::max<int64_t>(256, c.sc.swap_tmp_elems / maxcnt);
  for (int64_t off = 0; off < nc; off += slice) {
    const int64_t w = std::min<int64_t>(slice, nc - off);
    // a workgroup moves at least 4 x 16 bytes per thread of a row
    dim3 grid((unsigned)std::min<int64_t>(ceil_div(w, 256 * 4 * (16 / (int64_t)sizeof(T))), 64), (unsigned)maxcnt);
    hipLaunchKernelGGL((swap_gather_kernel<T>), grid, dim3(256), 0, c.s, c.A, c.lda, c0 + off, w,
                       c.sc.plan_src, c.sc.plan_count, c.sc.swap_tmp);
    RBFX_CHECK_LAUNCH();
    hipLaunchKernelGGL((swap_scatter_kernel<T>), grid, dim3(256), 0, c.s, c.A, c.lda, c0 + off, w,
                       c.sc.plan_dst, c.sc.plan_count, c.sc.swap_tmp);
    RBFX_CHECK_LAUNCH();
  }
  prof_stop(pid, c.s, 0.0, 4.0 * sizeof(T) * (e - s) * nc);
  return RBFX_OK;
}

// U12 = X A12 with X = inv(L11) of the W x W diagonal block at r0 (ld ldx)
// and columns [col, col+ncols): staged copy of A12, then one GEMM.
template <typename T>
static int trsm(LuCtx<T>& c, int64_t r0, int W, const T* X, int64_t ldx, int64_t col, int64_t ncols) {
  if (ncols <= 0) return RBFX_OK;
  int st;
  const int pid = prof_start(sizeof(T) == 8 ? "trsm_f64" : "trsm_f32", c.s);
  T* B = c.A + r0 * c.lda + col;
  const int64_t slice = std::max<int64_t>((int64_t)GBN, c.sc.swap_tmp_elems / W);
  for (int64_t off = 0; off < ncols; off += slice) {
    const int64_t w = std::min<int64_t>(slice, ncols - off);
    hipLaunchKernelGGL((copy_block_kernel<T>), dim3((unsigned)std::min<int64_t>(ceil_div(w, 256), 64), W),
                       dim3(256), 0, c.s, B + off, c.lda, W, w, c.sc.swap_tmp, w);
    RBFX_CHECK_LAUNCH();
    if ((st = gemm_launch<T, 1>(W, w, W, T(1), X, ldx, c.sc.swap_tmp, w, T(0), B + off, c.lda, c.s,
                                c.on_panel_stream)))
      return st;
  }
  prof_stop(pid, c.s, 2.0 * W * W * ncols, 3.0 * sizeof(T) * W * ncols);
  return RBFX_OK;
}

// Recursive factorisation of the panel rows [r0, n) x columns [r0, r0+W).
// It also leaves X = inv(L11) of the panel's W x W unit-lower diagonal block
// (ld ldx) in X's lower triangle; X's strictly upper part must hold zeros
// (the buffer is cleared once per factorisation and only ever receives
// lower-triangular entries).
template <typename T>
static int panel(LuCtx<T>& c, int64_t r0, int W, T* X, int64_t ldx) {
  int st;
  if (W <= c.wleaf) {
    if ((st = leaf(c, r0, W))) return st;
    return leaf_inverse(c, r0, W, X, ldx);
  }
  const int W1 = std::max(8, (W / 2 + 7) / 8 * 8);
  const int W2 = W - W1;
  // (every leaf applies its row interchanges to all columns of the outer
  // panel itself, so the halves need no swaps here)
  if ((st = panel(c, r0, W1, X, ldx))) return st;
  if ((st = trsm(c, r0, W1, X, ldx, r0 + W1, W2))) return st;
  const int64_t m2 = c.n - r0 - W1;
  const T* L21 = c.A + (r0 + W1) * c.lda + r0;
  if ((st = gemm<T>(m2, W2, W1, T(-1), L21, c.lda, c.A + r0 * c.lda + r0 + W1, c.lda, T(1),
                    c.A + (r0 + W1) * c.lda + r0 + W1, c.lda, c.s, true)))
    return st;
  T* X22 = X + (int64_t)W1 * ldx + W1;
  if ((st = panel(c, r0 + W1, W2, X22, ldx))) return st;
  // L21 (rows r0+W1 .. r0+W) is final now: later pivots only move rows
  // >= r0+W.  X21 = -X22 (L21 X11), the product staged in the panel's
  // swap buffer.
  RBFX_REQUIRE((int64_t)W1 * W2 <= c.sc.swap_tmp_elems, "inverse staging too small");
  T* Tm = c.sc.swap_tmp;
  const int pid = prof_start(sizeof(T) == 8 ? "trtri_f64" : "trtri_f32", c.s);
  if ((st = gemm_launch<T, 2>(W2, W1, W1, T(1), L21, c.lda, X, ldx, T(0), Tm, W1, c.s, c.on_panel_stream)))
    return st;
  st = gemm_launch<T, 2>(W2, W1, W2, T(-1), X22, ldx, Tm, W1, T(0), X + (int64_t)W1 * ldx, ldx, c.s,
                         c.on_panel_stream);
  prof_stop(pid, c.s, 2.0 * W2 * W1 * (double)(W1 + W2), (double)sizeof(T) * 3.0 * W1 * W2);
  return st;
}

static size_t getrf_ws_bytes(int64_t n, int bytes) {
  const int G = 1024;  // upper bound on workgroups of a leaf
  size_t b = 0;
  b += ws_round((size_t)2 * G * kRecWords * 8);   // cand records
  b += ws_round((size_t)2 * G * kPanelWMax * bytes);  // cand_row
  b += ws_round(64);                              // status
  for (int q = 0; q < 3; ++q) {                   // two slots of the main stream, one of the panel stream
    b += ws_round((size_t)2 * kSwapMax * 8) * 2;  // plan dst / src
    b += ws_round(64);                            // plan count
    b += ws_round((size_t)kNB * kNB * bytes);     // inverse of the diagonal block
  }
  // main stream: staging of row swaps (2*NB rows x min(n, 8192) columns) and
  // of the triangular solve's right-hand side (NB x n)
  b += ws_round((size_t)std::max<int64_t>(2 * kNB * std::min<int64_t>(n, 8192), (int64_t)kNB * n) * bytes);
  // panel stream: swaps and solves inside one panel (at most 2*NB x NB)
  b += ws_round((size_t)2 * kNB * kNB * bytes);
  return b;
}

template <typename T>
static bool carve_scratch(Workspace& ws, LuScratch<T>& sc, int64_t tmp_elems) {
  sc.plan_dst = ws.take<int64_t>(2 * kSwapMax);
  sc.plan_src = ws.take<int64_t>(2 * kSwapMax);
  sc.plan_count = ws.take<int>(16);
  sc.tri_inv = ws.take<T>((size_t)kNB * kNB);
  sc.swap_tmp_elems = tmp_elems;
  sc.swap_tmp = ws.take<T>(tmp_elems);
  sc.plan_s = sc.plan_e = -1;
  return sc.plan_dst && sc.plan_src && sc.plan_count && sc.tri_inv && (tmp_elems == 0 || sc.swap_tmp);
}

// The panels run on a second stream of this library (default priority, one
// per device) so that panel k+1 overlaps the trailing update of step k
// (look-ahead).  Enqueue sequences of concurrent calls on one device are
// serialised by a mutex, so the shared stream and events pair up correctly.
struct PanelStream {
  hipStream_t s = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
};
static std::mutex g_ps_mu;
static PanelStream g_ps[64];

// Releases the panel streams and events of every device (rbfx_shutdown).
int release_panel_streams() {
  std::lock_guard<std::mutex> lock(g_ps_mu);
  int dev0 = 0;
  RBFX_HIP(hipGetDevice(&dev0));
  for (int d = 0; d < 64; ++d) {
    PanelStream& p = g_ps[d];
    if (!p.s) continue;
    RBFX_HIP(hipSetDevice(d));
    RBFX_HIP(hipStreamSynchronize(p.s));
    for (int i = 0; i < 2; ++i) {
      if (p.ev[i]) RBFX_HIP(hipEventDestroy(p.ev[i]));
      p.ev[i] = nullptr;
    }
    RBFX_HIP(hipStreamDestroy(p.s));
    p.s = nullptr;
  }
  RBFX_HIP(hipSetDevice(dev0));
  return RBFX_OK;
}

static int panel_stream(PanelStream*& out) {
  int dev = 0;
  RBFX_HIP(hipGetDevice(&dev));
  RBFX_REQUIRE(dev >= 0 && dev < 64, "device index %d out of range", dev);
  PanelStream& p = g_ps[dev];
  if (!p.s) {
    RBFX_HIP(hipStreamCreateWithFlags(&p.s, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) RBFX_HIP(hipEventCreateWithFlags(&p.ev[i], hipEventDisableTiming));
  }
  out = &p;
  return RBFX_OK;
}

template <typename T>
static int getrf(int64_t n, T* A, int64_t lda, int64_t* ipiv, void* work, size_t work_bytes,
                 int32_t* info, hipStream_t s) {
  RBFX_REQUIRE(n > 0 && lda >= n, "bad n / lda");
  RBFX_REQUIRE(A && ipiv && info, "null pointer");
  RBFX_REQUIRE(n < (int64_t)INT32_MAX, "n too large");
  RBFX_REQUIRE(work && work_bytes >= getrf_ws_bytes(n, sizeof(T)), "getrf workspace too small");
  const DeviceInfo& di = device_info();
  std::lock_guard<std::mutex> lock(g_ps_mu);
  PanelStream* ps = nullptr;
  int st;
  if ((st = panel_stream(ps))) return st;
  LuCtx<T> c;
  c.A = A;
  c.n = n;
  c.lda = lda;
  c.ipiv = ipiv;
  c.info = info;
  c.s = s;
  Workspace ws(work, work_bytes);
  const int G = 1024;
  c.cand = ws.take<double>((size_t)2 * G * kRecWords);
  c.cand_row = ws.take<T>((size_t)2 * G * kPanelWMax);
  unsigned* ctl = ws.take<unsigned>(16);
  c.status = reinterpret_cast<int*>(ctl + 1);
  // main stream: two slots of (plan, inverse) filled by the panel stream
  // for alternate steps, one staging buffer; panel stream: its own scratch
  LuScratch<T> slot[2], sc_panel;
  const int64_t tmp_main = std::max<int64_t>(2 * kNB * std::min<int64_t>(n, 8192), (int64_t)kNB * n);
  const bool ok = carve_scratch(ws, slot[0], tmp_main) && carve_scratch(ws, slot[1], 0) &&
                  carve_scratch(ws, sc_panel, (int64_t)2 * kNB * kNB);
  slot[1].swap_tmp = slot[0].swap_tmp;
  slot[1].swap_tmp_elems = slot[0].swap_tmp_elems;
  RBFX_REQUIRE(ok && c.cand && c.cand_row, "workspace carve failed");
  c.seq = 0;
  c.G_max = G;
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, ctl, 16);
  RBFX_CHECK_LAUNCH();
  // pivot records start with tag 0, which no column uses
  hipLaunchKernelGGL(zero_u32_kernel, dim3(4 * G * kRecWords / 64), dim3(64), 0, s,
                     reinterpret_cast<unsigned*>(c.cand), 4 * G * kRecWords);
  RBFX_CHECK_LAUNCH();
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned*>(info), 1);
  RBFX_CHECK_LAUNCH();
  // the diagonal-block inverses: strictly upper parts stay zero from here on
  for (int q = 0; q < 2; ++q) {
    const int words = (int)((size_t)kNB * kNB * sizeof(T) / 4);
    hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)ceil_div(words, 256)), dim3(256), 0, s,
                       reinterpret_cast<unsigned*>(slot[q].tri_inv), words);
    RBFX_CHECK_LAUNCH();
  }

  // panel context: same matrix and pivot records, own stream and scratch
  LuCtx<T> cp = c;
  cp.s = ps->s;
  cp.on_panel_stream = true;
  cp.sc = sc_panel;
  hipEvent_t ready = ps->ev[0], done = ps->ev[1];
  // Panel stream, per step: factor the panel (which also leaves the inverse
  // of its diagonal block in the main stream's slot), then the
  // row-interchange plan for that slot.
  // Narrow leaves only for panels that run beside a trailing update (not
  // for panel 0, which runs alone).
  auto panel_step = [&](int64_t r0, int W, LuScratch<T>& out, bool beside_update) -> int {
    int e;
    const int pid = prof_start(sizeof(T) == 8 ? "panel_step_f64" : "panel_step_f32", cp.s);
    cp.pc0 = r0;
    cp.pw = W;
    const int64_t nl = g_narrow_leaf_rows.load(std::memory_order_relaxed);
    const int64_t narrow_rows = nl != 0 ? nl : sizeof(T) == 8 ? kNarrowLeafRows : -1;
    cp.wleaf = beside_update && narrow_rows > 0 && n - r0 >= narrow_rows ? kPanelWNarrow : kPanelWMax;
    if ((e = panel(cp, r0, W, out.tri_inv, (int64_t)kNB))) return e;
    prof_stop(pid, cp.s, (double)(n - r0) * W * W, 0.0);
    return make_plan(cp, out, r0, r0 + W);
  };
  // panel 0 after everything queued so far on the caller's stream
  RBFX_HIP(hipEventRecord(ready, s));
  RBFX_HIP(hipStreamWaitEvent(cp.s, ready, 0));
  if ((st = panel_step(0, (int)std::min<int64_t>(kNB, n), slot[0], false))) return st;
  RBFX_HIP(hipEventRecord(done, cp.s));
  int par = 0;
  for (int64_t k0 = 0; k0 < n; k0 += kNB, par ^= 1) {
    const int W = (int)std::min<int64_t>(kNB, n - k0);
    const int64_t k1 = k0 + W;
    const int64_t rest = n - k1;
    c.sc = slot[par];
    RBFX_HIP(hipStreamWaitEvent(s, done, 0));  // panel k0, its plan and inverse are ready
    if (rest <= 0) {
      if ((st = apply_swaps(c, k0, k1, 0, k0))) return st;
      break;
    }
    const T* L21 = A + k1 * lda + k0;
    const T* U12 = A + k0 * lda + k1;
    T* A22 = A + k1 * lda + k1;
    // The next panel's columns first (interchanges, U12 block, update): the
    // panel stream waits for exactly these.  It then factors that panel
    // while this stream does the rest of the step (look-ahead): the
    // interchanges of the columns left of the panel, which nothing in the
    // factorisation reads again, and those, U12 and the update of the
    // remaining columns.  The rest of the step is enqueued before the panel,
    // whose many small launches take the host a few ms: so the device never
    // waits for the host to reach the long update.
    const int Wn = (int)std::min<int64_t>(kNB, rest);
    if ((st = apply_swaps(c, k0, k1, k1, Wn))) return st;
    if ((st = trsm(c, k0, W, c.sc.tri_inv, (int64_t)kNB, k1, Wn))) return st;
    if ((st = gemm<T>(rest, Wn, W, T(-1), L21, lda, U12, lda, T(1), A22, lda, s))) return st;
    RBFX_HIP(hipEventRecord(ready, s));
    if ((st = apply_swaps(c, k0, k1, 0, k0))) return st;
    if (rest > Wn) {
      if ((st = apply_swaps(c, k0, k1, k1 + Wn, rest - Wn))) return st;
      if ((st = trsm(c, k0, W, c.sc.tri_inv, (int64_t)kNB, k1 + Wn, rest - Wn))) return st;
      if ((st = gemm<T>(rest, rest - Wn, W, T(-1), L21, lda, U12 + Wn, lda, T(1), A22 + Wn, lda, s)))
        return st;
    }
    RBFX_HIP(hipStreamWaitEvent(cp.s, ready, 0));
    if ((st = panel_step(k1, Wn, slot[par ^ 1], rest > Wn))) return st;
    RBFX_HIP(hipEventRecord(done, cp.s));
  }
  hipLaunchKernelGGL(status_to_info_kernel, dim3(1), dim3(64), 0, s, c.status, info);
  RBFX_CHECK_LAUNCH();
  return RBFX_OK;
}

// Exposed for the blocked triangular solves and tests.
int gemm_f64(int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
             const double* B, int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t s) {
  return gemm<double>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
}
int gemm_f32(int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
             const float* B, int64_t ldb, float beta, float* C, int64_t ldc, hipStream_t s) {
  return gemm<float>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

}  // namespace rbfx

using namespace rbfx;

extern "C" int rbfx_set_panel_groups(int max_groups) {
  RBFX_REQUIRE(max_groups >= 0, "max_groups must be >= 0");
  g_panel_groups.store(max_groups, std::memory_order_relaxed);
  return RBFX_OK;
}

extern "C" int rbfx_set_narrow_leaf_rows(int64_t rows) {
  g_narrow_leaf_rows.store(rows, std::memory_order_relaxed);
  return RBFX_OK;
}

extern "C" size_t rbfx_getrf_workspace(int64_t n, int dtype_bytes) {
  if (n <= 0 || (dtype_bytes != 4 && dtype_bytes != 8)) return 0;
  return getrf_ws_bytes(n, dtype_bytes);
}

extern "C" int rbfx_getrf_f64(int64_t n, double* d_A, int64_t lda, int64_t* d_ipiv, void* d_work,
                              size_t work_bytes, int32_t* d_info, void* stream) {
  return getrf<double>(n, d_A, lda, d_ipiv, d_work, work_bytes, d_info, as_stream(stream));
}

extern "C" int rbfx_getrf_f32(int64_t n, float* d_A, int64_t lda, int64_t* d_ipiv, void* d_work,
                              size_t work_bytes, int32_t* d_info, void* stream) {
  return getrf<float>(n, d_A, lda, d_ipiv, d_work, work_bytes, d_info, as_stream(stream));
}

// Test hook: C = beta C + alpha A B with the MFMA kernel.
extern "C" int rbfx_gemm_f64(int64_t m, int64_t n, int64_t k, double alpha, const double* d_A,
                             int64_t lda, const double* d_B, int64_t ldb, double beta, double* d_C,
                             int64_t ldc, void* stream) {
  RBFX_REQUIRE(m >= 0 && n >= 0 && k >= 0 && d_A && d_B && d_C, "bad gemm args");
  RBFX_REQUIRE(beta == 0.0 || beta == 1.0, "beta must be 0 or 1");
  return gemm<double>(m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc, as_stream(stream));
}
extern "C" int rbfx_gemm_f32(int64_t m, int64_t n, int64_t k, float alpha, const float* d_A,
                             int64_t lda, const float* d_B, int64_t ldb, float beta, float* d_C,
                             int64_t ldc, void* stream) {
  RBFX_REQUIRE(m >= 0 && n >= 0 && k >= 0 && d_A && d_B && d_C, "bad gemm args");
  RBFX_REQUIRE(beta == 0.0f || beta == 1.0f, "beta must be 0 or 1");
  return gemm<float>(m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc, as_stream(stream));
}
