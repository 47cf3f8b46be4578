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

#include <mutex>

#include "common.h"

namespace rbfx {

int gemm_f64(int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda, const double* B,
             int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t s);

constexpr int TB = 64;  // block size of the sweep and of the diagonal inverses

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

// ---- superblock sweep ---------------------------------------------------------
// The chain of a blocked triangular solve is what bounds it at these sizes:
// every exchange between workgroups is a coherent store, a counter and a
// poll through the memory side of the chip (~2 us each).  The superblock
// sweep makes the chain n/SB steps instead of n/TB: the solution advances by
// SB = 512 rows at a time, each superblock's rows solved with its 512 x 512
// diagonal inverse (built once per factorisation by sbinv_kernel), by NCH
// workgroups of CH rows:
//   1. fold the tiles of the superblocks already solved, in solve order
//      (each superblock counts its solved chunks), into private partial sums;
//   2. publish the chunk's CH accumulated values b - sum, count them in;
//   3. once all NCH chunks of the superblock are in, apply rows of its
//      diagonal inverse to the superblock's SB values and publish the CH
//      solution entries.
// Tickets are taken in start order and each superblock has NCH consecutive
// tickets, so the peers a workgroup waits for in step 3 have all started
// before any workgroup of a later superblock: the waits always end.
constexpr int SB = 512;
constexpr int CH = 32;
constexpr int NCH = SB / CH;
constexpr int kCtlLine = 32;  // words per superblock control line

// 512 x 512 inverse of a diagonal superblock of L (unit lower, which = 0) or
// U (which = 1) in FP64, from the 64 x 64 inverses D_i of its sub-blocks:
//   lower:  X_ii = D_i,  X_ij = -D_i sum_{j<=k<i} L_ik X_kj   (i > j)
//   upper:  X_ii = D_i,  X_ij = -D_i sum_{i<k<=j} U_ik X_kj   (i < j)
// One workgroup per (superblock, triangle); 64 x 64 tile products through
// LDS, each thread a 4 x 4 block of the tile.  Rows and columns past n are
// identity.
template <typename TF>
__global__ __launch_bounds__(256) void sbinv_kernel(const TF* __restrict__ LU, int64_t lda, int64_t n,
                                                    const double* __restrict__ dinv64, double* __restrict__ dinv512) {
  __shared__ double As[TB][TB + 1];
  __shared__ double Bs[TB][TB + 1];
  const int64_t K = blockIdx.x >> 1;
  const int which = blockIdx.x & 1;
  const int64_t o = K * SB;
  const int sz = (int)imin64(SB, n - o);
  const int nb = (sz + TB - 1) / TB;
  constexpr int NSUB = SB / TB;
  double* X = dinv512 + (K * 2 + which) * (int64_t)SB * SB;
  const int tid = threadIdx.x;
  const int tr = (tid >> 4) * 4, tc = (tid & 15) * 4;
  // X = blockdiag(D_i), identity past the matrix
  for (int64_t e = tid; e < (int64_t)SB * SB; e += 256) {
    const int r = (int)(e / SB), c = (int)(e % SB);
    const int bi = r / TB, bj = c / TB;
    double v = 0.0;
    if (bi == bj) {
      if (bi < nb)
        v = dinv64[((o / TB + bi) * 2 + which) * TB * TB + (r % TB) * TB + (c % TB)];
      else
        v = r == c ? 1.0 : 0.0;
    }
    X[e] = v;
  }
  __syncthreads();
  double acc[4][4];
  auto tile_mma = [&](double sign) {
#pragma unroll
    for (int q = 0; q < TB; ++q) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = As[tr + u][q];
        b[u] = Bs[q][tc + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = __builtin_fma(sign * a[u], b[v], acc[u][v]);
    }
  };
  for (int j = 0; j < nb; ++j) {
    for (int step = 1; step < nb; ++step) {
      // lower: rows below j; upper: rows above j, nearest first
      const int i = which == 0 ? j + step : j - step;
      if (i < 0 || i >= nb) break;
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
      const int k0 = which == 0 ? j : i + 1, k1 = which == 0 ? i - 1 : j;
      for (int k = k0; k <= k1; ++k) {
        // As = factor tile (i, k) (strictly lower / upper: no diagonal), Bs = X tile (k, j)
        for (int e = tid; e < TB * TB; e += 256) {
          const int r = e / TB, c = e % TB;
          const int gr = i * TB + r, gc = k * TB + c;
          As[r][c] = (gr < sz && gc < sz) ? (double)LU[(o + gr) * lda + o + gc] : 0.0;
          Bs[r][c] = X[(int64_t)(k * TB + r) * SB + j * TB + c];
        }
        __syncthreads();
        tile_mma(1.0);
        __syncthreads();
      }
      // X_ij = -D_i T: As = D_i, Bs = T
      for (int e = tid; e < TB * TB; e += 256) {
        const int r = e / TB, c = e % TB;
        As[r][c] = X[(int64_t)(i * TB + r) * SB + i * TB + c];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) Bs[tr + u][tc + v] = acc[u][v];
      __syncthreads();
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
      tile_mma(-1.0);
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) X[(int64_t)(i * TB + tr + u) * SB + j * TB + tc + v] = acc[u][v];
      __syncthreads();
    }
  }
  (void)NSUB;
}

// Condition estimate of every diagonal superblock T_KK of L (which = 0, unit
// lower) and U (which = 1): max(||T||_inf ||X||_inf, ||T||_1 ||X||_1) with X
// its 512 x 512 inverse.  The FP64 sweeps apply X only where this is at most
// kCondRobust; above it they solve the block by substitution over its 64-row
// sub-blocks (one refinement step with X reduces the block residual by about
// kappa eps, so X serves a backward-stable solve only for kappa well below
// eps^-1/2).
constexpr double kCondRobust = 1e7;
template <typename TF>
__global__ __launch_bounds__(256) void sbcond_kernel(const TF* __restrict__ LU, int64_t lda, int64_t n,
                                                     const double* __restrict__ dinv512, double* __restrict__ cond) {
  __shared__ double red[4][256];
  const int I = blockIdx.x >> 1, which = blockIdx.x & 1;
  const int64_t o = (int64_t)I * SB;
  const int sz = (int)imin64(SB, n - o);
  const double* X = dinv512 + (int64_t)blockIdx.x * SB * SB;
  double tr = 0, tc = 0, xr = 0, xc = 0;  // max row / column sums of |T| and |X|
  for (int q = threadIdx.x; q < sz; q += 256) {
    double a = 0, b = 0, cx = 0, dx = 0;
    for (int k = 0; k < sz; ++k) {
      // row q and column q of the triangle (unit diagonal for L)
      const bool inr = which ? k >= q : k < q;
      const bool inc = which ? k <= q : k > q;
      if (inr) a += fabs((double)LU[(o + q) * lda + o + k]);
      if (inc) b += fabs((double)LU[(o + k) * lda + o + q]);
      cx += fabs(X[(int64_t)q * SB + k]);
      dx += fabs(X[(int64_t)k * SB + q]);
    }
    if (!which) {
      a += 1.0;
      b += 1.0;
    }
    tr = fmax(tr, a);
    tc = fmax(tc, b);
    xr = fmax(xr, cx);
    xc = fmax(xc, dx);
  }
  red[0][threadIdx.x] = tr;
  red[1][threadIdx.x] = tc;
  red[2][threadIdx.x] = xr;
  red[3][threadIdx.x] = xc;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int t = 1; t < 256; ++t)
      for (int e = 0; e < 4; ++e) red[e][0] = fmax(red[e][0], red[e][t]);
    cond[blockIdx.x] = fmax(red[0][0] * red[2][0], red[1][0] * red[3][0]);
  }
}

// ctl[0] ticket, ctl[1] superblocks solved, ctl[2] status of the solve; the
// s-th processed superblock owns the 128-byte line ctl[kCtlLine (s + 1) ..]:
// word e < 3 counts its chunks that have published stage e (0: accumulated
// values, 1: first solution, 2: its residual), word 3 its solved chunks.
// Separate lines keep the polls of one superblock's counters off the others'.
// Superblocks complete in order (each folds its predecessor), so ctl[1] >= k
// means superblocks 0 .. k-1 are solved.
//
// REFINE (FP64 factors): the inverse of an ill-conditioned 512 x 512
// diagonal block applied to the accumulated values loses digits that a
// substitution keeps (measured on the adjoint gradients, and as a backward
// error of 6e-9 against 8e-17 for substitution with the same factors of a
// numerically singular system).  In U blocks one step of refinement inside
// the block restores them for moderately conditioned blocks: x0 = D acc,
// r = acc - U_KK x0, x = x0 + D r, at two more exchanges per superblock.
// Blocks of L or U whose condition estimate exceeds kCondRobust are solved
// by substitution over their 64-row sub-blocks instead.  The FP32 factors of
// the preconditioner need neither: an application error there only perturbs
// the preconditioner.
constexpr int kSweepThreads = 512;
constexpr int kSweepWaves = kSweepThreads / 64;
template <typename TF, bool FWD, bool TRANS, int WHICH, bool REFINE>
__global__ __launch_bounds__(kSweepThreads) void sbsweep_kernel(const TF* __restrict__ F, int64_t lda, int64_t n,
                                                                const double* __restrict__ D512,
                                                                const double* __restrict__ dinv64,
                                                                const double* __restrict__ cond, double* __restrict__ y,
                                                                double* __restrict__ P, unsigned* __restrict__ ctl,
                                                                uint64_t wait_ticks) {
  typedef double T;
  constexpr int RW = CH / kSweepWaves;        // chunk rows per wave (no-transpose)
  constexpr int NG = 2 * kSweepWaves;         // half-wave groups (transpose)
  constexpr int NV = SB * CH / kSweepThreads;  // operand values per thread
  constexpr int NP = TRANS ? 1 : RW;          // partial sums per thread
  static_assert(RW * kSweepWaves == CH && NV * NG == SB && NV == 8 * RW, "sweep layout");
  __shared__ T xs[SB];             // folded solution / accumulated values of the superblock
  __shared__ T xs2[SB];            // first solution, then its residual (REFINE)
  __shared__ T red[NG][CH + 1];    // cross-group partial sums
  __shared__ int s_t;
  __shared__ unsigned s_avail, s_solved;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t nsb = ceil_div(n, SB);
  unsigned* ticket = ctl;
  unsigned* solved = ctl + 1;
  int* status = reinterpret_cast<int*>(ctl + 2);
  if (tid == 0) {
    s_t = (int)__hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_avail = 0;
    s_solved = 0;
    if (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) s_t = INT_MAX;
  }
  __syncthreads();
  const int t = s_t;
  if ((int64_t)t >= nsb * NCH) return;
  const int s = t / NCH, c = t % NCH;
  unsigned* parts = ctl + kCtlLine * (s + 1);
  const int64_t I = FWD ? s : nsb - 1 - s;
  const int64_t oI = I * SB;
  const int szI = (int)imin64(SB, n - oI);
  const int rc0 = c * CH;                                   // first row of the chunk in the superblock
  const int nr = szI - rc0 < 0 ? 0 : (szI - rc0 < CH ? szI - rc0 : CH);
  // Thread layout of every CH x SB operand, for the chunk's rows:
  //   no-transpose: wave w takes rows RW w .. RW w + RW-1, lane l the columns
  //     l, l+64, ..., l+448 (coalesced rows of the factor); reduced over the
  //     lanes;
  //   transpose: lane l & 31 takes chunk column l & 31, the NG half-waves
  //     g = 2w + (l >> 5) the rows NV g .. NV g + NV-1 of M[rows][chunk
  //     columns]; reduced over the half-waves.
  const int tcol = lane & 31, tgrp = 2 * w + (lane >> 5);

  // bounded wait of thread 0 until *ctr >= target; false if the solve failed.
  // far: not on the critical path, poll less often
  auto wait_count = [&](const unsigned* ctr, unsigned target, unsigned& v, bool far) -> bool {
    uint32_t spins = 0;
    WaitClock clk(wait_ticks);
    while ((v = ld_coh(ctr)) < target) {
      if (far)
        __builtin_amdgcn_s_sleep(8);
      else
        __builtin_amdgcn_s_sleep(1);
      ++spins;
      if (clk.expired()) {
        __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return false;
      }
      if ((spins & 255u) == 0 && __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
        return false;
    }
    return true;
  };
  // reduce the per-thread partial sums of a product into red; after the
  // barrier, thread r < CH reads its row's value with row_value(r)
  auto reduce_rows = [&](T* ps) {
    if constexpr (!TRANS) {
#pragma unroll
      for (int k = 0; k < RW; ++k) {
        T v = ps[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) red[0][RW * w + k] = v;
      }
    } else {
      if (tid < NG * 32) red[tgrp][tcol] = ps[0];
    }
  };
  auto row_value = [&](int r) -> T {
    if constexpr (!TRANS) {
      return red[0][r];
    } else {
      T a = 0;
#pragma unroll
      for (int g = 0; g < NG; ++g) a += red[g][r];
      return a;
    }
  };
  // Loads of the operands are issued before the wait that releases their
  // use, so the chain of the sweep does not wait on HBM behind the other
  // workgroups' folds.
  // the factor's tile of column block oJ (szJ valid); diag: the stored
  // triangle of the diagonal block only (U_KK with its diagonal, or the
  // strictly lower part of the unit-lower L_KK; transposed as the sweep is)
  // Addresses are a wave-uniform base (scalar registers) plus one 32-bit
  // per-lane offset, so the NV loads in flight hold no vector address each.
  const int wu = __builtin_amdgcn_readfirstlane(w);
  const uint32_t hoff = (uint32_t)(lane >> 5) * NV;  // row offset of the upper half-wave (transpose)
  // Full tiles (every superblock but the last, off the diagonal) load
  // unmasked; the masks of the others compare against per-call limits, so no
  // per-value index is kept across the fold loop.
  auto load_F = [&](TF(&v)[NV], int64_t oJ, int szJ, bool diag) {
    const bool full = szJ == SB && nr == CH && !diag;
    if constexpr (!TRANS) {
      const int clim = szJ - lane;  // columns lane + 64m valid for 64m < clim
#pragma unroll
      for (int k = 0; k < RW; ++k) {
        const int r = RW * wu + k;
        const TF* row = F + (oI + rc0 + r) * lda + oJ;
        const int dlim = rc0 + r - lane;  // the diagonal entry: 64m == dlim
        if (full) {
#pragma unroll
          for (int m = 0; m < 8; ++m) v[8 * k + m] = row[(uint32_t)(lane + 64 * m)];
        } else if (r < nr) {
#pragma unroll
          for (int m = 0; m < 8; ++m) {
            const bool tri = WHICH == 1 ? 64 * m >= dlim : 64 * m < dlim;
            v[8 * k + m] = (64 * m < clim && (!diag || tri)) ? row[(uint32_t)(lane + 64 * m)] : TF(0);
          }
        } else {
#pragma unroll
          for (int m = 0; m < 8; ++m) v[8 * k + m] = TF(0);
        }
      }
    } else {
      const TF* base = F + (oJ + (int64_t)2 * wu * NV) * lda + oI + rc0;
      const uint32_t off = hoff * (uint32_t)lda + (uint32_t)tcol;
      if (full) {
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] = (base + q * lda)[off];
      } else {
        const int g0 = tgrp * NV;
        const int qlim = tcol < nr ? szJ - g0 : 0;  // rows g0 + q valid for q < qlim
        const int dlim = rc0 + tcol - g0;           // the diagonal entry: q == dlim
#pragma unroll
        for (int q = 0; q < NV; ++q) {
          const bool tri = WHICH == 1 ? q <= dlim : q > dlim;
          v[q] = (q < qlim && (!diag || tri)) ? (base + q * lda)[off] : TF(0);
        }
      }
    }
  };
  // rows of the superblock's diagonal inverse (columns for the transpose)
  const T* D = D512 + (I * 2 + WHICH) * (int64_t)SB * SB;
  auto load_D = [&](T(&v)[NV]) {
    if constexpr (!TRANS) {
#pragma unroll
      for (int k = 0; k < RW; ++k) {
        const int r = RW * wu + k;
        const T* row = D + (int64_t)(rc0 + r) * SB;
        if (r < nr) {
#pragma unroll
          for (int m = 0; m < 8; ++m) v[8 * k + m] = row[(uint32_t)(lane + 64 * m)];
        } else {
#pragma unroll
          for (int m = 0; m < 8; ++m) v[8 * k + m] = T(0);
        }
      }
    } else {
      const T* base = D + (int64_t)2 * wu * NV * SB + rc0;
      const uint32_t off = hoff * SB + (uint32_t)tcol;
      if (nr == CH) {
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] = (base + q * SB)[off];
      } else {
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] = tcol < nr ? (base + q * SB)[off] : T(0);
      }
    }
  };
  // ps += the operand held in v times the SB-vector vec (LDS), per thread
  auto product = [&](const auto& v, const T* vec, T* ps) {
    if constexpr (!TRANS) {
#pragma unroll
      for (int k = 0; k < RW; ++k) {
        T a = 0;
#pragma unroll
        for (int m = 0; m < 8; ++m) a = __builtin_fma((T)v[8 * k + m], vec[lane + 64 * m], a);
        ps[k] += a;
      }
    } else {
      T a = 0;
#pragma unroll
      for (int q = 0; q < NV; ++q) a = __builtin_fma((T)v[q], vec[tgrp * NV + q], a);
      ps[0] += a;
    }
  };
  auto apply = [&](const T(&dv)[NV], const T* vec) {
    T ps[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) ps[k] = 0;
    product(dv, vec, ps);
    reduce_rows(ps);
  };
  // publish the chunk's CH values of stage e into buf, wait for the
  // superblock's other chunks, then stage the superblock's SB values in dst
  auto exchange = [&](T* buf, int e, T value, T* dst) -> bool {
    if (tid < nr) st_coh(buf + oI + rc0 + tid, value);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) {
      __hip_atomic_fetch_add(parts + e, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      unsigned v;
      s_avail = wait_count(parts + e, (unsigned)NCH, v, false) ? 0u : 0xffffffffu;
    }
    __syncthreads();
    if (s_avail == 0xffffffffu) return false;
    for (int q = tid; q < SB; q += kSweepThreads) dst[q] = q < szI ? ld_coh(buf + oI + q) : T(0);
    __syncthreads();
    return true;
  };

  // 1. fold the superblocks solved before this one, in solve order.  The
  // predecessor's completion is polled on its own line; earlier superblocks
  // through the solved count, so a workgroup that starts late folds them
  // without a poll each.  FP32 factors leave room to hold the diagonal
  // inverse's rows during the last wait as well; FP64 ones load them after
  // the fold.
  constexpr bool kEarlyD = sizeof(TF) == 4;
  T ps[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) ps[k] = 0;
  T dv[NV];
  if (kEarlyD && s == 0) load_D(dv);
  for (int d = 0; d < s; ++d) {
    const int64_t J = FWD ? d : nsb - 1 - d;
    const int64_t oJ = J * SB;
    const int szJ = (int)imin64(SB, n - oJ);
    TF tv[NV];
    load_F(tv, oJ, szJ, false);
    if (kEarlyD && d == s - 1) load_D(dv);
    __syncthreads();  // xs free
    if (tid == 0 && s_solved <= (unsigned)d) {
      unsigned v;
      bool ok;
      if (d + 1 < s) {
        ok = wait_count(solved, (unsigned)(d + 1), v, true);
      } else {
        ok = wait_count(ctl + kCtlLine * (d + 1) + 3, (unsigned)NCH, v, false);
        v = d + 1;
      }
      s_solved = v;
      s_avail = ok ? 0u : 0xffffffffu;
    }
    __syncthreads();
    if (s_avail == 0xffffffffu) return;
    for (int e = tid; e < SB; e += kSweepThreads) xs[e] = e < szJ ? ld_coh(y + oJ + e) : T(0);
    __syncthreads();
    product(tv, xs, ps);
  }
  if (!kEarlyD) load_D(dv);
  // 2. accumulated values b - (sum of the folded tiles); exchange them
  __syncthreads();
  reduce_rows(ps);
  __syncthreads();
  T acc_r = tid < nr ? y[oI + rc0 + tid] - row_value(tid) : T(0);
  if (!exchange(P, 0, acc_r, xs)) return;
  // 3. x = rows of D applied to the superblock's accumulated values
  T x;
  if constexpr (REFINE) {
    if (cond[I * 2 + WHICH] > kCondRobust) {
      // ill-conditioned block: every chunk solves the whole block by
      // substitution over its 64-row sub-blocks with their 64 x 64 inverses
      // (xs: values still to solve, xs2: solution), then keeps its rows
      const int nsub = (szI + TB - 1) / TB;
      for (int st = 0; st < nsub; ++st) {
        const int b = FWD ? st : nsub - 1 - st;
        const int ob = b * TB;
        const double* D64 = dinv64 + ((I * (SB / TB) + b) * 2 + WHICH) * (int64_t)(TB * TB);
        if (tid < TB) {
          T a = 0;
          for (int c2 = 0; c2 < TB; ++c2)
            a = __builtin_fma(TRANS ? D64[c2 * TB + tid] : D64[tid * TB + c2], xs[ob + c2], a);
          xs2[ob + tid] = ob + tid < szI ? a : T(0);
        }
        __syncthreads();
        const int cend = szI - ob < TB ? szI - ob : TB;
        for (int i = tid; i < szI; i += kSweepThreads) {
          if (FWD ? i >= ob + TB : i < ob) {
            T a = 0;
            for (int c2 = 0; c2 < cend; ++c2) {
              const int64_t j = ob + c2;
              const TF t = TRANS ? F[(oI + j) * lda + oI + i] : F[(oI + i) * lda + oI + j];
              a = __builtin_fma((T)t, xs2[j], a);
            }
            xs[i] -= a;
          }
        }
        __syncthreads();
      }
      x = tid < nr ? xs2[rc0 + tid] : T(0);
      goto publish;
    }
  }
  apply(dv, xs);
  __syncthreads();
  x = tid < nr ? row_value(tid) : T(0);
  // the refinement step in U blocks; L blocks (unit diagonal, multipliers
  // at most 1) measured no better with it
  if constexpr (REFINE && WHICH == 1) {
    // x0 exchanged; r = acc - T_KK x0 (or T_KK^T x0) on the chunk's rows;
    // exchanged; x += D r.  The diagonal tile loads during the first
    // exchange, D's rows again during the second.
    T* X0 = P + n;
    T* Rb = P + 2 * n;
    TF uv[NV];
    load_F(uv, oI, szI, true);
    if (!exchange(X0, 1, x, xs2)) return;
    T pr[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) pr[k] = 0;
    product(uv, xs2, pr);
    load_D(dv);
    __syncthreads();
    reduce_rows(pr);
    __syncthreads();
    // L_KK has a unit diagonal that is not stored: its x0 term
    const T r_r = tid < nr ? xs[rc0 + tid] - row_value(tid) - (WHICH == 0 ? xs2[rc0 + tid] : T(0)) : T(0);
    if (!exchange(Rb, 2, r_r, xs2)) return;
    apply(dv, xs2);
    __syncthreads();
    if (tid < nr) x += row_value(tid);
  }
publish:
  if (tid < nr) st_coh(y + oI + rc0 + tid, x);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid == 0) {
    // the superblock's last chunk to finish counts it solved
    if (__hip_atomic_fetch_add(parts + 3, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == (unsigned)NCH - 1)
      __hip_atomic_fetch_add(solved, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Clears the ticket and the superblocks' control lines; the solve's status
// word ctl[2] stays.
__global__ void reset_sweep_ctl_kernel(unsigned* ctl, int64_t nsb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2 || (i >= kCtlLine && i < kCtlLine * (nsb + 1))) ctl[i] = 0u;
}

// 64 x 64 inverses of L and U, then the 512 x 512 ones, then the condition
// estimates of the 512 x 512 blocks.
static int64_t dinv64_elems(int64_t n) { return ceil_div(n, TB) * 2 * TB * TB; }

template <typename T>
static int launch_sbsweep(bool fwd, bool trans, int which, const T* F, int64_t lda, int64_t n,
                          const double* dinv, double* y, double* P, unsigned* ctl, hipStream_t s) {
  const double* D512 = dinv + dinv64_elems(n);
  const double* cond = D512 + ceil_div(n, (int64_t)SB) * 2 * (int64_t)SB * SB;
  const int64_t nsb = ceil_div(n, (int64_t)SB);
  hipLaunchKernelGGL(reset_sweep_ctl_kernel, dim3((unsigned)ceil_div(kCtlLine * (nsb + 1), 256)), dim3(256), 0, s,
                     ctl, nsb);
  RBFX_CHECK_LAUNCH();
  const uint64_t ticks = device_info().wait_ticks;
  const unsigned grid = (unsigned)(nsb * NCH);
  const int pid = prof_start(sizeof(T) == 8 ? "sweep_f64" : "sweep_f32", s);
  // one refinement step in the diagonal blocks of FP64 factors
  constexpr bool R64 = sizeof(T) == 8;
#define RBFX_SBSWEEP(F_, T_, W_)                                                                              \
  hipLaunchKernelGGL((sbsweep_kernel<T, F_, T_, W_, R64>), dim3(grid), dim3(kSweepThreads), 0, s, F, \
                     lda, n, D512, dinv, cond, y, P, ctl, ticks)
  if (fwd && !trans && which == 0) RBFX_SBSWEEP(true, false, 0);
  else if (!fwd && !trans && which == 1) RBFX_SBSWEEP(false, false, 1);
  else if (fwd && trans && which == 1) RBFX_SBSWEEP(true, true, 1);
  else if (!fwd && trans && which == 0) RBFX_SBSWEEP(false, true, 0);
  else {
    set_error("invalid sweep combination");
    return RBFX_ERR_ARG;
  }
#undef RBFX_SBSWEEP
  prof_stop(pid, s, (double)n * n, (double)sizeof(T) * n * (n + 1) / 2.0);
  RBFX_CHECK_LAUNCH();
  return RBFX_OK;
}

// Workspace of one solve: the permuted vector, the superblocks' accumulated
// values, then the control words (3 + one per superblock).
constexpr int64_t getrs_ctl_words(int64_t n) { return kCtlLine * ((n + SB - 1) / SB + 1) + 8; }
// P holds three n-vectors: the accumulated values, and for the refined U
// sweeps the first solutions and their residuals.
static bool carve_getrs(void* work, size_t work_bytes, int64_t n, double*& y, double*& P, unsigned*& ctl) {
  Workspace ws(work, work_bytes);
  y = ws.take<double>(n);
  P = ws.take<double>(3 * n);
  ctl = ws.take<unsigned>(getrs_ctl_words(n));
  return y && P && ctl;
}

// The solve itself on carved scratch: y (n), P (3n), ctl (getrs_ctl_words).
// Every scratch word that must start at zero is cleared by a kernel of the
// solve (no memset), and nothing is allocated.
template <typename T>
static int getrs_run(int trans, int64_t n, const T* LU, int64_t lda, const int64_t* perm, const double* dinv,
                     double* b, double* y, double* P, unsigned* ctl, hipStream_t s) {
  const unsigned g = (unsigned)ceil_div(n, 256);
  int st;
  if (!trans) {
    hipLaunchKernelGGL(gather_perm_kernel, dim3(g), dim3(256), 0, s, b, perm, n, y, ctl, 1);
    RBFX_CHECK_LAUNCH();
    if ((st = launch_sbsweep<T>(true, false, 0, LU, lda, n, dinv, y, P, ctl, s))) return st;
    if ((st = launch_sbsweep<T>(false, false, 1, LU, lda, n, dinv, y, P, ctl, s))) return st;
    hipLaunchKernelGGL(scatter_perm_kernel, dim3(g), dim3(256), 0, s, y, (const int64_t*)nullptr, n, b,
                       ctl);
    RBFX_CHECK_LAUNCH();
  } else {
    hipLaunchKernelGGL(gather_perm_kernel, dim3(g), dim3(256), 0, s, b, (const int64_t*)nullptr, n, y,
                       ctl, 1);
    RBFX_CHECK_LAUNCH();
    if ((st = launch_sbsweep<T>(true, true, 1, LU, lda, n, dinv, y, P, ctl, s))) return st;
    if ((st = launch_sbsweep<T>(false, true, 0, LU, lda, n, dinv, y, P, ctl, s))) return st;
    hipLaunchKernelGGL(scatter_perm_kernel, dim3(g), dim3(256), 0, s, y, perm, n, b, ctl);
    RBFX_CHECK_LAUNCH();
  }
  return RBFX_OK;
}

template <typename T>
static int getrs(int trans, int64_t n, const T* LU, int64_t lda, const int64_t* perm, const double* dinv,
                 double* b, void* work, size_t work_bytes, hipStream_t s) {
  RBFX_REQUIRE(n > 0 && lda >= n && LU && perm && dinv && b, "bad getrs args");
  RBFX_REQUIRE(n < (int64_t)INT32_MAX / 2, "getrs: n too large");
  double *y, *P;
  unsigned* ctl;
  RBFX_REQUIRE(carve_getrs(work, work_bytes, n, y, P, ctl), "getrs workspace too small");
  return getrs_run<T>(trans, n, LU, lda, perm, dinv, b, y, P, ctl, s);
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
  double *y, *P;
  unsigned* ctl;
  RBFX_REQUIRE(carve_getrs(work, work_bytes, n, y, P, ctl), "getrs workspace too small");
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
  const int64_t nsb = ceil_div(n, (int64_t)SB);
  double* d512 = dinv + dinv64_elems(n);
  hipLaunchKernelGGL((sbinv_kernel<T>), dim3((unsigned)(2 * nsb)), dim3(256), 0, s, LU, lda, n, (const double*)dinv,
                     d512);
  RBFX_CHECK_LAUNCH();
  hipLaunchKernelGGL((sbcond_kernel<T>), dim3((unsigned)(2 * nsb)), dim3(256), 0, s, LU, lda, n, (const double*)d512,
                     d512 + nsb * 2 * (int64_t)SB * SB);
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

extern "C" int64_t rbfx_getrs_dinv_elems(int64_t n) {
  return n <= 0 ? 0 : dinv64_elems(n) + ceil_div(n, (int64_t)SB) * (2 * (int64_t)SB * SB + 2);
}

extern "C" size_t rbfx_getrs_workspace(int64_t n) {
  if (n <= 0) return 0;
  return ws_round((size_t)n * sizeof(double)) + ws_round((size_t)3 * n * sizeof(double)) +
         ws_round(getrs_ctl_words(n) * sizeof(unsigned)) + 512;
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

namespace rbfx {
// ---- allocation-free solve ------------------------------------------------------
// Scratch of rbfx_getrs_noalloc_f64 for n <= RBFX_GETRS_NOALLOC_MAX_N, carved
// like a getrs workspace: y (n doubles), P (3n doubles), control words.  One
// copy per device.  Host code never uses the symbol itself: its device
// address comes from hipGetSymbolAddress, once per device.
constexpr int64_t kNoAllocMaxN = RBFX_GETRS_NOALLOC_MAX_N;
constexpr size_t kNoAllocBytes = 4 * (((size_t)kNoAllocMaxN * 8 + 255) / 256 * 256) +
                                 ((size_t)getrs_ctl_words(kNoAllocMaxN) * 4 + 255) / 256 * 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];

static int scratch_address(void*& out) {
  int dev = 0;
  RBFX_HIP(hipGetDevice(&dev));
  RBFX_REQUIRE(dev >= 0 && dev < 64, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    RBFX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    RBFX_REQUIRE(p != nullptr, "no device address for the solve's scratch array");
    g_scratch_ptr[dev] = p;
  }
  out = g_scratch_ptr[dev];
  return RBFX_OK;
}
}  // namespace rbfx

extern "C" int rbfx_getrs_noalloc_f64(int trans, int64_t n, const double* d_LU, int64_t lda,
                                      const int64_t* d_perm, const double* d_dinv, double* d_b, void* stream) {
  RBFX_REQUIRE(n > 0 && lda >= n && d_LU && d_perm && d_dinv && d_b, "bad getrs args");
  RBFX_REQUIRE(n <= kNoAllocMaxN, "rbfx_getrs_noalloc_f64: n = %lld exceeds its fixed scratch (n <= %lld)",
               (long long)n, (long long)kNoAllocMaxN);
  void* base = nullptr;
  int st;
  if ((st = scratch_address(base))) return st;
  double *y, *P;
  unsigned* ctl;
  RBFX_REQUIRE(carve_getrs(base, kNoAllocBytes, n, y, P, ctl), "solve scratch too small");
  return getrs_run<double>(trans, n, d_LU, lda, d_perm, d_dinv, d_b, y, P, ctl, as_stream(stream));
}

extern "C" int rbfx_getrs_status(int64_t n, void* d_work, size_t work_bytes, void* stream) {
  return getrs_status(n, d_work, work_bytes, as_stream(stream));
}

extern "C" int rbfx_getrs_f32(int trans, int64_t n, const float* d_LU, int64_t lda, const int64_t* d_perm,
                              const double* d_dinv, double* d_b, void* d_work, size_t work_bytes,
                              void* stream) {
  return getrs<float>(trans, n, d_LU, lda, d_perm, d_dinv, d_b, d_work, work_bytes, as_stream(stream));
}
