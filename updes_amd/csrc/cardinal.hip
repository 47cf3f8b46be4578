// FP32 preconditioner of the collocation system in a local cardinal basis:
// the FP32-MFMA LU leg of the mixed-precision solve.  No reference
// counterpart (the reference solves in FP64 only, updes/operators.py:612-613).
//
// The FP32 LU of K = [[D], [P^T 0]] itself is a poor preconditioner: with r^3
// kernels kappa(K) grows like N^2.5 (~1e12 at N = 40,000), far beyond
// 1/u_fp32.  Most of that is the conditioning of the RBF basis, not of the
// PDE: with c = A^-1 [u; 0] the system is B u = q with kappa(B) ~ 1e5-1e6.
// So the FP32 factorisation is taken after a change of basis c = C y, where
// column j of C holds the coefficients of node j's local cardinal function
// psi_j: the RBF interpolant on the k nearest nodes S_j (node j included),
// with the problem's monomials, equal to 1 at node j and 0 at the other nodes
// of S_j, whose RBF coefficients beta_j satisfy the moment conditions
// P^T beta_j = 0 (so every column of C lies in the RBF space of the solve).
// Then D C = B Psi with Psi[i][j] = psi_j(x_i), the identity on each stencil,
// and the N x N matrix
//     K~ = diag(rs) D C          (rs: inverse row maxima)
// is well enough conditioned for FP32 LU to precondition GMRES on K itself.
// The refinement keeps its iterate in the original coefficients c and its
// residuals in FP64 with the stored K, so the basis change costs no
// accuracy; the preconditioner maps a residual [r_D; r_P] to
//     z = C K~^-1 diag(rs) r_D + [P (P^T P)^-1 r_P; 0],
// and its transpose handles K^T (the adjoint).
#include <math.h>

#include <algorithm>
#include <vector>

#include "krylov.h"
#include "rbf_device.h"

namespace rbfx {

constexpr int kCardMaxK = 48;
constexpr int kCardMaxM = 15;
constexpr int kFormRows = 16;   // rows of K~ per workgroup of the gather formation kernel
constexpr int kFormCols = 256;  // columns per block of the union tables (rbfx_cardinal)
constexpr int kFormTR = 2;      // rows staged in LDS per pass of the LDS formation kernel
constexpr int kFormUMax = 2560; // union entries of one block staged in LDS
constexpr int kFormSlots = 5;   // union entries per thread and chunk of the staging loads
constexpr int kFormRB = 512;    // rows per workgroup of the LDS formation kernels
// ring formation kernel (k <= 24, M <= 6): rows of K[:, U] streamed into an
// LDS ring by producer waves, kRingDepth rows ahead of the consumer waves
constexpr int kRingUMax = 1280;                    // union entries + monomial columns per ring slot
constexpr int kRingDepth = 6;                      // rows in flight
constexpr int kRingSlots = kRingDepth + 2;         // a slot is refilled two rows after its last read
constexpr int kRingLoads = 2 * kRingUMax / kFormCols;  // 4-byte loads per producer lane and row
static_assert(kRingLoads * kFormCols == 2 * kRingUMax && kRingLoads * kRingDepth <= 63, "ring loads");
static_assert((kRingSlots & (kRingSlots - 1)) == 0, "ring slots: a power of two");
constexpr int kRedBlocks = 128; // partial sums of the small reductions

__device__ inline double wave_sum_d(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline double wave_max_d(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ inline double mono_val(int id, double x, double y) {
  int a, b;
  mono_exp(id, a, b);
  return ipow(x, a) * ipow(y, b);
}
__device__ inline double binom(int n, int r) {
  double v = 1.0;
  for (int i = 1; i <= r; ++i) v = v * (n - r + i) / i;
  return v;
}
// A stencil entry outside [0, N) is replaced by the node itself and flagged.
__device__ inline int64_t checked_id(int32_t v, int64_t N, int64_t self, int* status) {
  if (v >= 0 && v < N) return v;
  atomicCAS(status, 0, 1);
  return self;
}

struct CardArgs {
  int64_t N, M, k;
  const double* x;
  const double* y;
  const int32_t* stencil;
  double* beta;
  double* alpha;
  Kern kern;
  int* status;  // 1: stencil id out of range; 2: node missing from its stencil; 3: singular local system
};

// One wave per node j: solve the local system
//   [[Phi_S, P_S], [P_S^T, 0]] [beta'; a'] = [e_j; 0]
// (Phi_S with a zero diagonal, as the global A; monomials in the coordinates
// (x - x_j)/s, s the stencil radius), project beta' onto the null space of
// P_S^T (two passes against an orthonormal basis of P_S, so that the moment
// conditions hold to rounding), and expand the local polynomial a' into the
// global monomials.  MS bounds the local system's size k + M (32 for the
// usual stencils of at most 26 entries with at most 6 monomials: 12 KB of
// LDS, so about three times as many nodes in flight per CU as with 64).
template <int MS>
__global__ __launch_bounds__(64) void cardinal_coeffs_kernel(CardArgs a) {
  constexpr int kQ = MS == 64 ? kCardMaxM : 6;  // monomials of the projection basis
  __shared__ double A[MS][MS + 1];
  __shared__ double rhs[64];
  __shared__ double sol[64];
  __shared__ double Qs[kQ][64];
  const int64_t j = blockIdx.x;
  const int lane = threadIdx.x;
  const int k = (int)a.k, M = (int)a.M, S = k + M;
  const int32_t* st = a.stencil + j * a.k;
  const int64_t id = lane < k ? checked_id(st[lane], a.N, j, a.status) : j;
  const double cx = a.x[j], cy = a.y[j];
  const double xl = a.x[id], yl = a.y[id];
  if (__ballot(lane < k && id == j) == 0ull) {
    if (lane == 0) atomicCAS(a.status, 0, 2);
  }
  double sc = wave_max_d(lane < k ? sqrt((xl - cx) * (xl - cx) + (yl - cy) * (yl - cy)) : 0.0);
  if (!(sc > 0.0)) sc = 1.0;
  const double ux = (xl - cx) / sc, uy = (yl - cy) / sc;
  // row `lane` of the local saddle-point matrix
  for (int b = 0; b < k; ++b) {
    const double xb = __shfl(xl, b, 64), yb = __shfl(yl, b, 64);
    const double ubx = __shfl(ux, b, 64), uby = __shfl(uy, b, 64);
    if (lane < k) {
      const double dx = xl - xb, dy = yl - yb;
      A[lane][b] = lane == b ? 0.0 : radial_value(a.kern, sqrt(dist2(dx, dy)));
    } else if (lane < S) {
      A[lane][b] = mono_val(lane - k, ubx, uby);
    }
  }
  for (int m = 0; m < M; ++m) {
    if (lane < k)
      A[lane][k + m] = mono_val(m, ux, uy);
    else if (lane < S)
      A[lane][k + m] = 0.0;
  }
  // the node's own position in its stencil (the first, if listed twice)
  const unsigned long long selfmask = __ballot(lane < k && id == j);
  const int self = selfmask ? __ffsll((long long)selfmask) - 1 : 0;
  rhs[lane] = lane == self ? 1.0 : 0.0;
  __syncthreads();

  // Gaussian elimination with partial pivoting, a row per lane
  bool singular = false;
  for (int c = 0; c < S; ++c) {
    double v = (lane >= c && lane < S) ? fabs(A[lane][c]) : -1.0;
    int p = lane;
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(v, off, 64);
      const int op = __shfl_xor(p, off, 64);
      if (ov > v || (ov == v && op < p)) {
        v = ov;
        p = op;
      }
    }
    if (!(v > 0.0)) {  // uniform across the wave
      singular = true;
      break;
    }
    if (p != c) {
      if (lane < S) {
        const double t = A[c][lane];
        A[c][lane] = A[p][lane];
        A[p][lane] = t;
      }
      if (lane == 0) {
        const double t = rhs[c];
        rhs[c] = rhs[p];
        rhs[p] = t;
      }
    }
    __syncthreads();
    const double piv = A[c][c];
    if (lane > c && lane < S) {
      const double f = A[lane][c] / piv;
      for (int t = c + 1; t < S; ++t) A[lane][t] -= f * A[c][t];
      rhs[lane] -= f * rhs[c];
    }
    __syncthreads();
  }
  if (singular) {
    if (lane == 0) atomicCAS(a.status, 0, 3);
    if (lane < k) a.beta[j * a.k + lane] = __builtin_nan("");
    if (lane < M) a.alpha[j * a.M + lane] = __builtin_nan("");
    return;
  }
  for (int c = S - 1; c >= 0; --c) {
    if (lane == c) sol[c] = rhs[c] / A[c][c];
    __syncthreads();
    const double xc = sol[c];
    if (lane < c) rhs[lane] -= A[lane][c] * xc;
    __syncthreads();
  }

  // moment conditions to rounding: beta' minus its projection on span(P_S)
  double bl = lane < k ? sol[lane] : 0.0;
  for (int m = 0; m < M; ++m) {
    double v = lane < k ? mono_val(m, ux, uy) : 0.0;
    for (int pass = 0; pass < 2; ++pass)
      for (int mm = 0; mm < m; ++mm) {
        const double h = wave_sum_d(Qs[mm][lane] * v);
        v -= h * Qs[mm][lane];
      }
    const double nv = sqrt(wave_sum_d(v * v));
    Qs[m][lane] = nv > 0.0 ? v / nv : 0.0;
  }
  for (int pass = 0; pass < 2; ++pass)
    for (int m = 0; m < M; ++m) {
      const double h = wave_sum_d(Qs[m][lane] * bl);
      bl -= h * Qs[m][lane];
    }
  if (lane < k) a.beta[j * a.k + lane] = bl;

  // global monomial coefficients of the local polynomial sum_m a'_m p_m((x - x_j)/s)
  if (lane < M) {
    int ti, tj;
    mono_exp(lane, ti, tj);
    double acc = 0.0;
    for (int m = 0; m < M; ++m) {
      int ea, eb;
      mono_exp(m, ea, eb);
      if (ea < ti || eb < tj) continue;
      acc += sol[k + m] * ipow(1.0 / sc, ea + eb) * binom(ea, ti) * ipow(-cx, ea - ti) * binom(eb, tj) *
             ipow(-cy, eb - tj);
    }
    a.alpha[j * a.M + lane] = acc;
  }
}

struct FormArgs {
  int64_t N, M, k;
  const double* K;  // FP64 K (n x ld), rows [0, N) are D
  int64_t ld;
  const int32_t* stencil;
  const double* beta;
  const double* alpha;
  double* rs;
  float* out;  // N x ld32
  int64_t ld32;
  int* status;
};

// K~[i][j] before scaling: (D C)[i][j] = sum_t K[i][S_j[t]] beta_j[t] + sum_m K[i][N+m] alpha_j[m].
__device__ inline double dc_entry(const FormArgs& f, const double* Ki, int64_t j) {
  const int32_t* sj = f.stencil + j * f.k;
  const double* bj = f.beta + j * f.k;
  double acc = 0.0;
  for (int t = 0; t < f.k; ++t) acc = __builtin_fma(Ki[checked_id(sj[t], f.N, j, f.status)], bj[t], acc);
  for (int m = 0; m < f.M; ++m) acc = __builtin_fma(Ki[f.N + m], f.alpha[j * f.M + m], acc);
  return acc;
}

// rs[i] = 1 / max_{j in S_i} |(D C)[i][j]|: the row maximum of D C sits on
// the cardinal functions centred next to x_i.  One wave per row.
__global__ __launch_bounds__(256) void cardinal_rowscale_kernel(FormArgs f) {
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= f.N) return;
  const double* Ki = f.K + i * f.ld;
  double v = 0.0;
  if (lane < f.k) v = fabs(dc_entry(f, Ki, checked_id(f.stencil[i * f.k + lane], f.N, i, f.status)));
  v = wave_max_d(v);
  if (lane == 0) f.rs[i] = v > 0.0 ? 1.0 / v : (v == 0.0 ? 1.0 : v);
}

// K~ = diag(rs) D C in FP32.  Thread per column j of a 256-column block
// (its stencil, coefficients in registers), kFormRows rows per workgroup;
// the K entries are gathered from L2 (stencils are sorted by node id, so a
// wave's gathers of one row fall on a few cache lines).  Linear block b runs
// on XCD b % 8: b -> row tile (b % 8) + 8 * (b / 8 / ncb), column block
// (b / 8) % ncb, so the workgroups of one XCD share the rows in its L2.
template <int KMAX>
__global__ __launch_bounds__(256) void cardinal_form_kernel(FormArgs f, int64_t ncb, int64_t nrt) {
  const int64_t b = blockIdx.x;
  const int64_t q = b >> 3;
  const int64_t rt = (q / ncb) * 8 + (b & 7);
  const int64_t cb = q % ncb;
  if (rt >= nrt) return;
  const int64_t j = cb * 256 + threadIdx.x;
  const bool act = j < f.N;
  int32_t sj[KMAX];
  double bj[KMAX];
  double aj[kCardMaxM];
#pragma unroll
  for (int t = 0; t < KMAX; ++t) {
    const bool ok = act && t < f.k;
    sj[t] = ok ? (int32_t)checked_id(f.stencil[j * f.k + t], f.N, j, f.status) : 0;
    bj[t] = ok ? f.beta[j * f.k + t] : 0.0;
  }
#pragma unroll
  for (int m = 0; m < kCardMaxM; ++m) aj[m] = (act && m < f.M) ? f.alpha[j * f.M + m] : 0.0;
  const int64_t r0 = rt * kFormRows;
  for (int rr = 0; rr < kFormRows; ++rr) {
    const int64_t i = r0 + rr;
    if (i >= f.N) break;
    const double* Ki = f.K + i * f.ld;
    double acc = 0.0;
#pragma unroll
    for (int t = 0; t < KMAX; ++t)
      if (t < f.k) acc = __builtin_fma(Ki[sj[t]], bj[t], acc);
#pragma unroll
    for (int m = 0; m < kCardMaxM; ++m)
      if (m < f.M) acc = __builtin_fma(Ki[f.N + m], aj[m], acc);
    if (act) f.out[i * f.ld32 + j] = (float)(f.rs[i] * acc);
  }
}

// K~ = diag(rs) D C in FP32 with the K entries staged in LDS.  Workgroup
// (rb, cb) owns the kFormCols columns of block cb and the kFormRB rows of
// range rb.  Each thread keeps its column's coefficients, the LDS offsets of
// its stencil entries within a staged row, and its share of the block's union
// U (sorted distinct stencil ids, mostly contiguous runs, so the staging
// loads coalesce) in registers.  Every pass stages K[rows, U] for kFormTR
// rows: all of a chunk's loads are issued before any of them is written to
// LDS, so a pass waits on memory once per chunk, not once per load.  A block
// whose union exceeds kFormUMax re-reads its stencil ids and gathers from
// global memory instead (slow, correct).  Linear block b runs on XCD b % 8:
// b -> row range (b % 8) + 8 * (b / 8 / ncb), column block (b / 8) % ncb, so
// the workgroups of one XCD share the K rows they read in its L2.
template <int KMAX, int MMAX>
__global__ __launch_bounds__(kFormCols) void cardinal_form_lds_kernel(FormArgs f, const int64_t* __restrict__ uptr,
                                                                const int32_t* __restrict__ uidx,
                                                                const int16_t* __restrict__ upos, int64_t ncb,
                                                                int64_t nrb) {
  __shared__ double Ks[kFormTR][kFormUMax];
  __shared__ double Kp[kFormTR][MMAX];
  constexpr int kSlots = kFormUMax / kFormCols;     // union entries per thread
  constexpr int kChunks = kSlots / kFormSlots;      // staged one chunk at a time
  constexpr int kChunk = kFormSlots * kFormCols;    // union entries per chunk
  static_assert(kSlots * kFormCols == kFormUMax && kChunks * kFormSlots == kSlots, "union staging");
  const int64_t b = blockIdx.x;
  const int64_t q = b >> 3;
  const int64_t rb = (q / ncb) * 8 + (b & 7);
  const int64_t cb = q % ncb;
  if (rb >= nrb) return;
  const int tid = threadIdx.x;
  const int64_t u0 = uptr[cb];
  const int64_t nu64 = uptr[cb + 1] - u0;
  const bool lds = nu64 >= 0 && nu64 <= kFormUMax;  // uniform over the workgroup
  const int nu = lds ? (int)nu64 : 0;
  const int64_t j = cb * kFormCols + tid;
  const bool act = j < f.N;
  const int k = (int)f.k, M = (int)f.M;
  // coefficients and LDS byte offsets of this column's stencil entries
  // (loads issued unconditionally from clamped indices, then checked)
  uint32_t la[KMAX];
  double bj[KMAX];
  double aj[MMAX];
#pragma unroll
  for (int t = 0; t < KMAX; ++t) {
    const bool ok = act && t < k;
    const int64_t e = ok ? j * f.k + t : 0;
    const int p = upos[e];
    bj[t] = f.beta[e];
    la[t] = (uint32_t)p;
    if (!ok) bj[t] = 0.0;
    if (!(ok && p >= 0 && p < nu)) la[t] = 0u;
  }
#pragma unroll
  for (int t = 0; t < KMAX; ++t) la[t] *= (uint32_t)sizeof(double);
#pragma unroll
  for (int m = 0; m < MMAX; ++m) {
    const bool ok = act && m < M;
    const double a = f.alpha[ok ? j * f.M + m : 0];
    aj[m] = ok ? a : 0.0;
  }
  // byte offsets within a K row of this thread's union entries
  uint32_t uo[kSlots];
#pragma unroll
  for (int e = 0; e < kSlots; ++e) {
    const int u = tid + e * kFormCols;
    uo[e] = (uint32_t)uidx[u < nu ? u0 + u : (nu > 0 ? u0 : 0)];
  }
#pragma unroll
  for (int e = 0; e < kSlots; ++e) {
    const int u = tid + e * kFormCols;
    const int32_t id = (int32_t)uo[e];
    uo[e] = (u < nu ? (uint32_t)checked_id(id, f.N, 0, f.status) : 0u) * (uint32_t)sizeof(double);
  }
  // this thread's entry of the staged polynomial columns
  const int prr = tid / MMAX, pm = tid - prr * MMAX;
  const bool pw = prr < kFormTR && pm < M;
  const int64_t pcol = M > 0 ? f.N + (pm < M ? pm : M - 1) : 0;
  const int64_t rbeg = rb * kFormRB;
  const int64_t rend = rbeg + kFormRB < f.N ? rbeg + kFormRB : f.N;
  for (int64_t r0 = rbeg; r0 < rend; r0 += kFormTR) {
    const int nr = (int)(rend - r0 < kFormTR ? rend - r0 : kFormTR);
    if (lds) {
      __syncthreads();  // the previous pass is done with Ks
      const double kp = f.K[(r0 + (prr < nr ? prr : nr - 1)) * f.ld + pcol];
#pragma unroll
      for (int c = 0; c < kChunks; ++c) {
        if (c * kChunk < nu) {
          double v[kFormTR][kFormSlots];
#pragma unroll
          for (int rr = 0; rr < kFormTR; ++rr) {
            const char* Ki = reinterpret_cast<const char*>(f.K + (r0 + (rr < nr ? rr : nr - 1)) * f.ld);
#pragma unroll
            for (int s = 0; s < kFormSlots; ++s)
              v[rr][s] = *reinterpret_cast<const double*>(Ki + uo[c * kFormSlots + s]);
          }
#pragma unroll
          for (int rr = 0; rr < kFormTR; ++rr)
#pragma unroll
            for (int s = 0; s < kFormSlots; ++s) {
              const int u = c * kChunk + s * kFormCols + tid;
              if (u < nu) Ks[rr][u] = v[rr][s];
            }
        }
      }
      if (pw) Kp[prr][pm] = kp;
      __syncthreads();
      for (int rr = 0; rr < nr; ++rr) {
        const char* row = reinterpret_cast<const char*>(&Ks[rr][0]);
        double acc = 0.0;
#pragma unroll
        for (int t = 0; t < KMAX; ++t)
          if (t < k) acc = __builtin_fma(*reinterpret_cast<const double*>(row + la[t]), bj[t], acc);
#pragma unroll
        for (int m = 0; m < MMAX; ++m)
          if (m < M) acc = __builtin_fma(Kp[rr][m], aj[m], acc);
        if (act) f.out[(r0 + rr) * f.ld32 + j] = (float)(f.rs[r0 + rr] * acc);
      }
    } else {
      for (int rr = 0; rr < nr; ++rr) {
        const double* Ki = f.K + (r0 + rr) * f.ld;
        double acc = 0.0;
#pragma unroll
        for (int t = 0; t < KMAX; ++t)
          if (t < k && act) {
            const int64_t id = checked_id(f.stencil[j * f.k + t], f.N, j, f.status);
            acc = __builtin_fma(Ki[id], bj[t], acc);
          }
#pragma unroll
        for (int m = 0; m < MMAX; ++m)
          if (m < M) acc = __builtin_fma(Ki[f.N + m], aj[m], acc);
        if (act) f.out[(r0 + rr) * f.ld32 + j] = (float)(f.rs[r0 + rr] * acc);
      }
    }
  }
}

// K~ = diag(rs) D C with the rows of K[:, U] (U: the block's union, then the
// M monomial columns) streamed through a ring of kRingSlots LDS slots.
// Workgroup (rb, cb) of 2 x kFormCols threads: the first kFormCols are
// consumers (thread = column of block cb: coefficients and LDS offsets of
// its stencil in registers), the others producers that only issue
// direct-to-LDS 4-byte loads (lane l of a producer wave loads half l & 1 of
// one union entry, so 64 lanes fill 32 consecutive entries of a slot).
// Producers run kRingDepth rows ahead and wait on their own load counter
// for the row the consumers read next; one barrier per row hands it over.
// Producers never store and consumers never issue the ring's loads, so each
// wave's counter covers exactly its own kind of work.  A block whose union
// does not fit a slot is formed by the consumers from global memory.
// XCD-aware block order as in cardinal_form_lds_kernel.
template <int KMAX, int MMAX>
__global__ __launch_bounds__(2 * kFormCols) __attribute__((amdgpu_waves_per_eu(4, 4))) void cardinal_form_ring_kernel(FormArgs f, const int64_t* __restrict__ uptr,
                                                                     const int32_t* __restrict__ uidx,
                                                                     const int16_t* __restrict__ upos, int64_t ncb,
                                                                     int64_t nrb, const double* __restrict__ rsv) {
  __shared__ __align__(16) double ring[kRingSlots][kRingUMax];
  const int64_t b = blockIdx.x;
  const int64_t q = b >> 3;
  const int64_t rb = (q / ncb) * 8 + (b & 7);
  const int64_t cb = q % ncb;
  if (rb >= nrb) return;
  const int tid = threadIdx.x;
  const int k = (int)f.k, M = (int)f.M;
  const int64_t u0 = uptr[cb];
  const int64_t nu64 = uptr[cb + 1] - u0;
  const bool fits = nu64 >= 0 && nu64 + M <= kRingUMax;  // uniform over the workgroup
  const int nu = fits ? (int)nu64 : 0;
  const int64_t rbeg = rb * kFormRB;
  const int64_t rend = rbeg + kFormRB < f.N ? rbeg + kFormRB : f.N;
  const int nrows = (int)(rend - rbeg);

  if (tid >= kFormCols) {
    if (!fits) return;
    // ---- producer ----
    const int p = tid - kFormCols, pw = p >> 6, lane = p & 63;
    constexpr int kPW = kFormCols / 64;  // producer waves
    uint32_t go[kRingLoads];  // byte offsets within a K row of this lane's 4-byte words
#pragma unroll
    for (int e = 0; e < kRingLoads; ++e) {
      const int u = (e * kPW + pw) * 32 + (lane >> 1);
      const int32_t raw = uidx[u < nu ? u0 + u : (nu > 0 ? u0 : 0)];
      int64_t id = 0;
      if (u < nu)
        id = checked_id(raw, f.N, 0, f.status);
      else if (u < nu + M)
        id = f.N + (u - nu);
      go[e] = (uint32_t)id * (uint32_t)sizeof(double) + 4u * (uint32_t)(lane & 1);
    }
    auto issue = [&](int i) {  // row rbeg + i (clamped: every row issues the same loads) into its slot
      const int64_t row = rbeg + (i < nrows ? i : nrows - 1);
      const char* Kr = reinterpret_cast<const char*>(f.K + row * f.ld);
      char* slot = reinterpret_cast<char*>(&ring[i & (kRingSlots - 1)][0]);
#pragma unroll
      for (int e = 0; e < kRingLoads; ++e)
        __builtin_amdgcn_global_load_lds((const void*)(Kr + go[e]),
                                         (__attribute__((address_space(3))) void*)(slot + (e * kPW + pw) * 256), 4,
                                         0, 0);
    };
    for (int d = 0; d < kRingDepth; ++d) issue(d);
    // s_waitcnt vmcnt(kRingLoads * kRingDepth): vmcnt in bits [3:0] and [15:14]
    constexpr int kVm = kRingLoads * kRingDepth;
    constexpr int kWait = (kVm & 15) | ((kVm >> 4) << 14) | (7 << 4) | (15 << 8);
    for (int i = 0; i < nrows; ++i) {
      issue(i + kRingDepth);
      __builtin_amdgcn_s_waitcnt(kWait);  // row i has landed (kRingDepth rows issued after it)
      __builtin_amdgcn_s_barrier();
    }
    __builtin_amdgcn_s_waitcnt(0);  // no load may land after the workgroup ends
    return;
  }

  // ---- consumer ----
  const int64_t j = cb * kFormCols + tid;
  const bool act = j < f.N;
  // this column's coefficients and stencil positions (a column past N reads
  // column 0's and discards them); the loads are issued before any is used
  const double* pb = f.beta + (act ? j * f.k : 0);
  const int16_t* pu = upos + (act ? j * f.k : 0);
  const double* pa = f.alpha + (act ? j * f.M : 0);
  uint32_t la[KMAX];
  double bj[KMAX];
  double aj[MMAX];
#pragma unroll
  for (int t = 0; t < KMAX; ++t) {
    bj[t] = 0.0;
    la[t] = 0u;
    if (t < k) {
      bj[t] = pb[t];
      la[t] = (uint32_t)(int32_t)pu[t];
    }
  }
#pragma unroll
  for (int m = 0; m < MMAX; ++m) aj[m] = m < M ? pa[m] : 0.0;
#pragma unroll
  for (int t = 0; t < KMAX; ++t) {
    const int32_t pp = (int32_t)la[t];
    la[t] = (act && t < k && pp >= 0 && pp < nu) ? (uint32_t)pp * (uint32_t)sizeof(double) : 0u;
    if (!act) bj[t] = 0.0;
  }
#pragma unroll
  for (int m = 0; m < MMAX; ++m)
    if (!act) aj[m] = 0.0;
  if (!fits) {
    for (int64_t r = rbeg; r < rend; ++r) {
      const double* Ki = f.K + r * f.ld;
      double acc = 0.0;
#pragma unroll
      for (int t = 0; t < KMAX; ++t)
        if (t < k && act) {
          const int64_t id = checked_id(f.stencil[j * f.k + t], f.N, j, f.status);
          acc = __builtin_fma(Ki[id], bj[t], acc);
        }
#pragma unroll
      for (int m = 0; m < MMAX; ++m)
        if (m < M) acc = __builtin_fma(Ki[f.N + m], aj[m], acc);
      if (act) f.out[r * f.ld32 + j] = (float)(f.rs[r] * acc);
    }
    return;
  }
  // the monomial columns' LDS offsets; unused stencil and monomial slots
  // read entry 0 (a K value, finite) with a zero coefficient, so the row's
  // LDS reads need no branches and are issued together
  uint32_t lm[MMAX];
#pragma unroll
  for (int m = 0; m < MMAX; ++m) lm[m] = (uint32_t)(m < M ? nu + m : 0) * (uint32_t)sizeof(double);
  // row scales through the scalar cache (read-only here), one row ahead, so
  // that the loop has no vector load to wait for and its stores never stall it
  double rs_next = rsv[rbeg];
  for (int i = 0; i < nrows; ++i) {
    const double rs_i = rs_next;
    rs_next = rsv[rbeg + (i + 1 < nrows ? i + 1 : i)];
    __builtin_amdgcn_s_barrier();  // the producers' loads of row i have landed
    const char* row = reinterpret_cast<const char*>(&ring[i & (kRingSlots - 1)][0]);
    double acc = 0.0;
#pragma unroll
    for (int t = 0; t < KMAX; ++t) acc = __builtin_fma(*reinterpret_cast<const double*>(row + la[t]), bj[t], acc);
#pragma unroll
    for (int m = 0; m < MMAX; ++m) acc = __builtin_fma(*reinterpret_cast<const double*>(row + lm[m]), aj[m], acc);
    if (act) f.out[(rbeg + i) * f.ld32 + j] = (float)(rs_i * acc);
  }
}

// ---- the preconditioner's vector kernels ----------------------------------
struct VecArgs {
  int64_t N, M, k;
  const double* x;
  const double* y;
  const int32_t* stencil;
  const double* beta;
  const double* alpha;
  const int64_t* rev_ptr;
  const int32_t* rev_idx;
  const double* rs;
};

// out[i] = rs[i] v[i] for i < N, v[i] for N <= i < n (in place allowed)
__global__ void card_scale_kernel(int64_t N, int64_t n, const double* __restrict__ rs, const double* v,
                                  double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N)
    out[i] = rs[i] * v[i];
  else if (i < n)
    out[i] = v[i];
}

__global__ void card_copy_kernel(int64_t n, const double* __restrict__ v, double* __restrict__ t) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) t[i] = v[i];
}

// z[i] = rs[i] y[i] (i < N), z[i] = 0 (N <= i < n)
__global__ void card_scale_pad_kernel(int64_t N, int64_t n, const double* __restrict__ rs,
                                      const double* __restrict__ yv, double* __restrict__ z) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N)
    z[i] = rs[i] * yv[i];
  else if (i < n)
    z[i] = 0.0;
}

// z[l] = (C_beta y)[l] + (P s)[l] for l < N, over the reverse stencil table
// (the occurrences of node l in the stencils, ascending j).  A thread per
// node; a node that occurs in more than kHeavyRev stencils (the special
// points occur in all of them) is summed by its whole workgroup afterwards,
// in a fixed order, so that no thread walks N entries alone.
constexpr int kHeavyRev = 256;
__global__ __launch_bounds__(256) void card_cmap_kernel(VecArgs a, const double* __restrict__ yv,
                                                        const double* __restrict__ sP, double* __restrict__ z) {
  __shared__ double red[4];
  __shared__ int s_heavy[256];
  __shared__ int s_nheavy;
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nk = a.N * a.k;
  const int kk = (int)a.k;
  auto range = [&](int64_t node, int64_t& e0, int64_t& e1) {
    e0 = a.rev_ptr[node];
    e1 = a.rev_ptr[node + 1];
    e0 = e0 < 0 ? 0 : (e0 > nk ? nk : e0);
    e1 = e1 < e0 ? e0 : (e1 > nk ? nk : e1);
  };
  auto term = [&](int64_t e) {
    int idx = a.rev_idx[e];
    idx = idx < 0 ? 0 : (idx >= nk ? (int)nk - 1 : idx);
    return a.beta[idx] * yv[idx / kk];
  };
  auto poly = [&](int64_t node) {
    const double xl = a.x[node], yl = a.y[node];
    double acc = 0.0;
    for (int m = 0; m < a.M; ++m) acc = __builtin_fma(mono_val(m, xl, yl), sP[m], acc);
    return acc;
  };
  if (threadIdx.x == 0) s_nheavy = 0;
  __syncthreads();
  if (l < a.N) {
    int64_t e0, e1;
    range(l, e0, e1);
    if (e1 - e0 > kHeavyRev) {
      s_heavy[atomicAdd(&s_nheavy, 1)] = threadIdx.x;
    } else {
      double acc = 0.0;
      for (int64_t e = e0; e < e1; ++e) acc += term(e);
      z[l] = acc + poly(l);
    }
  }
  __syncthreads();
  const int nh = s_nheavy;
  for (int h = 0; h < nh; ++h) {
    // the heavy nodes of this block in thread order (fixed summation order)
    int th = 256;
    for (int q = 0; q < nh; ++q) {
      const int c = s_heavy[q];
      int rank = 0;
      for (int r = 0; r < nh; ++r) rank += s_heavy[r] < c;
      if (rank == h) th = c;
    }
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + th;
    int64_t e0, e1;
    range(node, e0, e1);
    double acc = 0.0;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += 256) acc += term(e);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) z[node] = (red[0] + red[1]) + (red[2] + red[3]) + poly(node);
    __syncthreads();
  }
}

// t[j] = (C_beta^T v[:N])[j] + (C_alpha^T v[N:])[j]
__global__ void card_cmap_t_kernel(VecArgs a, const double* __restrict__ v, double* __restrict__ t) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.N) return;
  const int32_t* sj = a.stencil + j * a.k;
  const double* bj = a.beta + j * a.k;
  double acc = 0.0;
  for (int q = 0; q < a.k; ++q) {
    const int32_t id = sj[q];
    const int64_t c = (id >= 0 && id < a.N) ? id : j;
    acc = __builtin_fma(bj[q], v[c], acc);
  }
  for (int m = 0; m < a.M; ++m) acc = __builtin_fma(a.alpha[j * a.M + m], v[a.N + m], acc);
  t[j] = acc;
}

// w[l] += (P zP)[l] for l < N
__global__ void card_add_p_kernel(VecArgs a, const double* __restrict__ zP, double* __restrict__ w) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= a.N) return;
  const double xl = a.x[l], yl = a.y[l];
  double acc = w[l];
  for (int m = 0; m < a.M; ++m) acc = __builtin_fma(mono_val(m, xl, yl), zP[m], acc);
  w[l] = acc;
}

// Deterministic M-output reductions over the nodes, block partials:
//   mode 0: out[m] = sum_j alpha[j][m] y[j]          (C_alpha y)
//   mode 1: out[m] = sum_l p_m(x_l) (v[l] - w[l])     (P^T (v - w))
__global__ __launch_bounds__(256) void card_reduce_partial_kernel(VecArgs a, int mode,
                                                                  const double* __restrict__ v,
                                                                  const double* __restrict__ w,
                                                                  double* __restrict__ part) {
  double acc[kCardMaxM];
#pragma unroll
  for (int m = 0; m < kCardMaxM; ++m) acc[m] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.N; i += (int64_t)gridDim.x * 256) {
    if (mode == 0) {
      const double yi = v[i];
#pragma unroll
      for (int m = 0; m < kCardMaxM; ++m)
        if (m < a.M) acc[m] = __builtin_fma(a.alpha[i * a.M + m], yi, acc[m]);
    } else {
      const double d = v[i] - w[i];
      const double xi = a.x[i], yi = a.y[i];
#pragma unroll
      for (int m = 0; m < kCardMaxM; ++m)
        if (m < a.M) acc[m] = __builtin_fma(mono_val(m, xi, yi), d, acc[m]);
    }
  }
  __shared__ double red[4][kCardMaxM];
#pragma unroll
  for (int m = 0; m < kCardMaxM; ++m) {
    double s = acc[m];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][m] = s;
  }
  __syncthreads();
  if (threadIdx.x < a.M)
    part[blockIdx.x * a.M + threadIdx.x] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// out = G (sum over nblocks of part), G = ptp_inv (M x M) or the identity
__global__ void card_reduce_final_kernel(int64_t M, int64_t nblocks, const double* __restrict__ part,
                                         const double* __restrict__ G, double* __restrict__ out) {
  __shared__ double sums[kCardMaxM];
  const int t = threadIdx.x;
  if (t < M) {
    double s = 0.0;
    for (int64_t b = 0; b < nblocks; ++b) s += part[b * M + t];
    sums[t] = s;
  }
  __syncthreads();
  if (t < M) {
    if (G) {
      double r = 0.0;
      for (int m = 0; m < M; ++m) r += G[t * M + m] * sums[m];
      out[t] = r;
    } else {
      out[t] = sums[t];
    }
  }
}

static int check_card(const rbfx_cardinal* c, bool need_rs = true) {
  RBFX_REQUIRE(c, "null cardinal basis");
  RBFX_REQUIRE(c->N > 0 && c->M >= 0 && c->M <= kCardMaxM, "cardinal basis: bad N / M");
  RBFX_REQUIRE(c->k >= 1 && c->k <= kCardMaxK && c->k <= c->N && c->k + c->M <= 63,
               "cardinal basis: stencil size k=%lld outside [1, %d] (k + M <= 63, k <= N)", (long long)c->k,
               kCardMaxK);
  RBFX_REQUIRE(c->N * c->k < (int64_t)INT32_MAX, "cardinal basis: N * k too large");
  RBFX_REQUIRE(c->d_x && c->d_y && c->d_stencil && c->d_beta && c->d_alpha && (c->d_rs || !need_rs),
               "cardinal basis: null array");
  return RBFX_OK;
}

static int read_status(int* d_status, hipStream_t s, const char* what) {
  int st = 0;
  RBFX_HIP(hipMemcpyAsync(&st, d_status, sizeof(int), hipMemcpyDeviceToHost, s));
  RBFX_HIP(hipStreamSynchronize(s));
  if (st == 1) {
    set_error("%s: a stencil node id is outside [0, N)", what);
    return RBFX_ERR_ARG;
  }
  if (st == 2) {
    set_error("%s: a node is missing from its own stencil", what);
    return RBFX_ERR_ARG;
  }
  if (st == 3) {
    set_error("%s: singular local interpolation system (stencil not unisolvent for the monomials)", what);
    return RBFX_ERR_SINGULAR;
  }
  return RBFX_OK;
}

__global__ void zero_int_kernel(int* p) { *p = 0; }

// ---- the two preconditioner pairs -----------------------------------------------
// w = diag(s) op(K) v with s = [rs; 1] when rs is set (the row-scaled
// forward system), else w = op(K) v.
struct StoredK : ResidualOp {
  int trans;
  int64_t n;
  const double* K64;
  int64_t ld64;
  double* part;
  const double* rs = nullptr;
  int64_t nrs = 0;
  hipStream_t s;
  int applyK(const double* v, double* w) override {
    int st = residual_gemv(trans, n, K64, ld64, v, w, part, s);
    if (st || !rs) return st;
    hipLaunchKernelGGL(card_scale_kernel, dim3((unsigned)ceil_div(nrs, 256)), dim3(256), 0, s, nrs, nrs, rs,
                       (const double*)w, w);
    RBFX_CHECK_LAUNCH();
    return RBFX_OK;
  }
};

struct CardPair : PairOp {
  int trans;
  VecArgs a;
  int64_t n;
  const float* LU32;
  int64_t ld32;
  const int64_t* perm;
  const double* dinv;
  const double* ptp_inv;
  void* ts_work;
  size_t ts_bytes;
  double* t;     // N
  double* red;   // kRedBlocks * M partials, then M small results
  StoredK* K;
  hipStream_t s;

  int solve32(int tr) {
    int st = rbfx_getrs_f32(tr, a.N, LU32, ld32, perm, dinv, t, ts_work, ts_bytes, s);
    if (st) return st;
    return rbfx_getrs_status(a.N, ts_work, ts_bytes, s);
  }
  int apply(const double* v, double* z, double* w) override {
    const unsigned gN = (unsigned)ceil_div(a.N, 256);
    double* part = red;
    double* small = red + kRedBlocks * kCardMaxM;
    const int pid = prof_start("cardinal_precond", s);
    int st;
    if (!trans) {
      // v is a residual of the row-scaled system: diag(rs) v_D is already the
      // right-hand side of K~ (v[:N] itself)
      hipLaunchKernelGGL(card_copy_kernel, dim3(gN), dim3(256), 0, s, a.N, v, t);
      RBFX_CHECK_LAUNCH();
      if ((st = solve32(0))) return st;
      if (a.M > 0) {
        hipLaunchKernelGGL(card_reduce_final_kernel, dim3(1), dim3(64), 0, s, a.M, (int64_t)1, v + a.N, ptp_inv,
                           small);
        RBFX_CHECK_LAUNCH();
      }
      hipLaunchKernelGGL(card_cmap_kernel, dim3(gN), dim3(256), 0, s, a, (const double*)t, (const double*)small, z);
      RBFX_CHECK_LAUNCH();
      if (a.M > 0) {
        hipLaunchKernelGGL(card_reduce_partial_kernel, dim3(kRedBlocks), dim3(256), 0, s, a, 0, (const double*)t,
                           (const double*)nullptr, part);
        RBFX_CHECK_LAUNCH();
        hipLaunchKernelGGL(card_reduce_final_kernel, dim3(1), dim3(64), 0, s, a.M, (int64_t)kRedBlocks,
                           (const double*)part, (const double*)nullptr, z + a.N);
        RBFX_CHECK_LAUNCH();
      }
      prof_stop(pid, s, 0.0, 0.0);
      return K->applyK(z, w);
    }
    hipLaunchKernelGGL(card_cmap_t_kernel, dim3(gN), dim3(256), 0, s, a, v, t);
    RBFX_CHECK_LAUNCH();
    if ((st = solve32(1))) return st;
    hipLaunchKernelGGL(card_scale_pad_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, a.N, n, a.rs,
                       (const double*)t, z);
    RBFX_CHECK_LAUNCH();
    prof_stop(pid, s, 0.0, 0.0);
    if ((st = K->applyK(z, w))) return st;
    if (a.M > 0) {
      hipLaunchKernelGGL(card_reduce_partial_kernel, dim3(kRedBlocks), dim3(256), 0, s, a, 1, v,
                         (const double*)w, part);
      RBFX_CHECK_LAUNCH();
      hipLaunchKernelGGL(card_reduce_final_kernel, dim3(1), dim3(64), 0, s, a.M, (int64_t)kRedBlocks,
                         (const double*)part, ptp_inv, z + a.N);
      RBFX_CHECK_LAUNCH();
      hipLaunchKernelGGL(card_add_p_kernel, dim3(gN), dim3(256), 0, s, a, (const double*)(z + a.N), w);
      RBFX_CHECK_LAUNCH();
    }
    return RBFX_OK;
  }
};

}  // namespace rbfx

using namespace rbfx;

extern "C" int rbfx_cardinal_coeffs(const rbfx_problem* p, const rbfx_cardinal* c, void* d_work,
                                    size_t work_bytes, void* stream) {
  int st;
  if ((st = check_card(c, false))) return st;
  RBFX_REQUIRE(p && p->N == c->N && p->M == c->M, "cardinal basis does not match the problem");
  RBFX_REQUIRE(!(p->flags & RBFX_FLAG_LOCAL_SUPPORT), "the cardinal basis needs every node in every support");
  RBFX_REQUIRE(d_work && work_bytes >= 256, "cardinal workspace too small");
  hipStream_t s = as_stream(stream);
  int* status = static_cast<int*>(d_work);
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, s, status);
  RBFX_CHECK_LAUNCH();
  CardArgs a{c->N, c->M, c->k, c->d_x, c->d_y, c->d_stencil, c->d_beta, c->d_alpha, make_kern(p->kernel, p->param),
             status};
  const int pid = prof_start("cardinal_coeffs", s);
  if (c->k + c->M <= 32 && c->M <= 6)
    hipLaunchKernelGGL(cardinal_coeffs_kernel<32>, dim3((unsigned)c->N), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL(cardinal_coeffs_kernel<64>, dim3((unsigned)c->N), dim3(64), 0, s, a);
  RBFX_CHECK_LAUNCH();
  prof_stop(pid, s, 0.0, 0.0);
  return read_status(status, s, "cardinal coefficients");
}

extern "C" int rbfx_cardinal_form(const rbfx_cardinal* c, const double* d_K64, int64_t ld64, float* d_out32,
                                  int64_t ld32, void* d_work, size_t work_bytes, void* stream) {
  int st;
  if ((st = check_card(c))) return st;
  RBFX_REQUIRE(d_K64 && d_out32 && ld64 >= c->N + c->M && ld32 >= c->N, "bad cardinal_form matrices");
  RBFX_REQUIRE(d_work && work_bytes >= 256, "cardinal workspace too small");
  hipStream_t s = as_stream(stream);
  int* status = static_cast<int*>(d_work);
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, s, status);
  RBFX_CHECK_LAUNCH();
  FormArgs f{c->N, c->M, c->k, d_K64, ld64, c->d_stencil, c->d_beta, c->d_alpha, c->d_rs, d_out32, ld32, status};
  hipLaunchKernelGGL(cardinal_rowscale_kernel, dim3((unsigned)ceil_div(c->N, 4)), dim3(256), 0, s, f);
  RBFX_CHECK_LAUNCH();
  const int64_t ncb = ceil_div(c->N, 256);
  const int pid = prof_start("cardinal_form", s);
  if (c->d_union_ptr && c->d_union_idx && c->d_union_pos) {
    const int64_t nrb = ceil_div(c->N, kFormRB);
    const int64_t blocks = ceil_div(nrb, 8) * 8 * ncb;
    RBFX_REQUIRE(blocks < (int64_t)UINT32_MAX, "cardinal_form grid too large");
#define RBFX_FORM_M(KM, MM)                                                                                \
  hipLaunchKernelGGL((cardinal_form_lds_kernel<KM, MM>), dim3((unsigned)blocks), dim3(kFormCols), 0, s, f, \
                     c->d_union_ptr, c->d_union_idx, c->d_union_pos, ncb, nrb)
#define RBFX_FORM(KM)      \
  if (c->M <= 6)           \
    RBFX_FORM_M(KM, 6);    \
  else                     \
    RBFX_FORM_M(KM, kCardMaxM)
#define RBFX_RING(KM)                                                                                       \
  hipLaunchKernelGGL((cardinal_form_ring_kernel<KM, 6>), dim3((unsigned)blocks), dim3(2 * kFormCols), 0, s, f, \
                     c->d_union_ptr, c->d_union_idx, c->d_union_pos, ncb, nrb, (const double*)c->d_rs)
    if (c->k <= 16 && c->M <= 6)
      RBFX_RING(16);
    else if (c->k <= 24 && c->M <= 6)
      RBFX_RING(24);
    else if (c->k <= 16)
      RBFX_FORM(16);
    else if (c->k <= 24)
      RBFX_FORM(24);
    else if (c->k <= 32)
      RBFX_FORM(32);
    else
      RBFX_FORM(48);
#undef RBFX_RING
#undef RBFX_FORM
#undef RBFX_FORM_M
  } else {
    const int64_t nrt = ceil_div(c->N, kFormRows);
    const int64_t blocks = ceil_div(nrt, 8) * 8 * ncb;
    RBFX_REQUIRE(blocks < (int64_t)UINT32_MAX, "cardinal_form grid too large");
#define RBFX_FORM(KM) \
  hipLaunchKernelGGL((cardinal_form_kernel<KM>), dim3((unsigned)blocks), dim3(256), 0, s, f, ncb, nrt)
    if (c->k <= 16)
      RBFX_FORM(16);
    else if (c->k <= 24)
      RBFX_FORM(24);
    else if (c->k <= 32)
      RBFX_FORM(32);
    else
      RBFX_FORM(48);
#undef RBFX_FORM
  }
  RBFX_CHECK_LAUNCH();
  // algorithmic bytes: K read once (8 n N) and K~ written (4 N^2)
  prof_stop(pid, s, 2.0 * (double)(c->k + c->M) * c->N * c->N,
            8.0 * (double)(c->N + c->M) * c->N + 4.0 * (double)c->N * c->N);
  return read_status(status, s, "cardinal form");
}

extern "C" size_t rbfx_gmres_ir_cardinal_workspace(int64_t N, int64_t M, int32_t restart) {
  if (N <= 0 || M < 0 || restart <= 0) return 0;
  const int64_t n = N + M;
  size_t b = fgmres_ws_bytes(n, restart);
  b += ws_round(rbfx_getrs_workspace(N));
  b += 2 * ws_round(sizeof(double) * (size_t)n);                            // t, scaled b
  b += ws_round(sizeof(double) * (size_t)(kRedBlocks + 1) * kCardMaxM);     // reductions
  b += ws_round(gemv_workspace(1, n, n));                                   // residual K^T partials
  return b + 2048;
}

extern "C" int rbfx_gmres_ir_cardinal(int trans, const rbfx_cardinal* c, const double* d_K64, int64_t ld64,
                                      double k_fro, const float* d_LU32, int64_t ld32, const int64_t* d_perm32,
                                      const double* d_dinv32, const double* d_b, double* d_x, double tol,
                                      int32_t restart, int32_t max_inner, void* d_work, size_t work_bytes,
                                      rbfx_ir_stats* stats, void* stream) {
  int st;
  if ((st = check_card(c))) return st;
  RBFX_REQUIRE(c->d_rev_ptr && c->d_rev_idx && (c->M == 0 || c->d_ptp_inv), "cardinal basis: null table");
  RBFX_REQUIRE(d_K64 && ld64 >= c->N + c->M, "need the stored FP64 K for the residual");
  RBFX_REQUIRE(d_LU32 && d_perm32 && d_dinv32 && ld32 >= c->N, "bad FP32 factors");
  RBFX_REQUIRE(d_b && d_x && d_b != d_x, "bad gmres_ir_cardinal vectors");
  RBFX_REQUIRE(restart > 0 && max_inner > 0 && tol > 0.0, "bad gmres_ir_cardinal parameters");
  RBFX_REQUIRE(work_bytes >= rbfx_gmres_ir_cardinal_workspace(c->N, c->M, restart),
               "gmres_ir_cardinal workspace too small");
  hipStream_t s = as_stream(stream);
  const int64_t N = c->N, n = c->N + c->M;
  Workspace ws(d_work, work_bytes);
  const size_t fw = fgmres_ws_bytes(n, restart);
  void* fwork = ws.take<char>(fw);
  const size_t tsb = rbfx_getrs_workspace(N);
  void* tsw = ws.take<char>(tsb);
  double* t = ws.take<double>(n);
  double* bs = ws.take<double>(n);
  double* red = ws.take<double>((size_t)(kRedBlocks + 1) * kCardMaxM);
  double* part = ws.take<double>(gemv_workspace(1, n, n) / sizeof(double) + 1);
  RBFX_REQUIRE(fwork && tsw && t && bs && red && part, "gmres_ir_cardinal workspace carve failed");
  StoredK K;
  K.trans = trans;
  K.n = n;
  K.K64 = d_K64;
  K.ld64 = ld64;
  K.part = part;
  K.s = s;
  const double* rhs = d_b;
  if (!trans) {
    // row-scaled forward system diag(s) K x = diag(s) b, s = [rs; 1]
    K.rs = c->d_rs;
    K.nrs = N;
    hipLaunchKernelGGL(card_scale_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, N, n,
                       (const double*)c->d_rs, d_b, bs);
    RBFX_CHECK_LAUNCH();
    rhs = bs;
  }
  CardPair pre;
  pre.trans = trans;
  pre.a = VecArgs{c->N, c->M, c->k, c->d_x, c->d_y, c->d_stencil, c->d_beta, c->d_alpha, c->d_rev_ptr,
                  c->d_rev_idx, c->d_rs};
  pre.n = n;
  pre.LU32 = d_LU32;
  pre.ld32 = ld32;
  pre.perm = d_perm32;
  pre.dinv = d_dinv32;
  pre.ptp_inv = c->d_ptp_inv;
  pre.ts_work = tsw;
  pre.ts_bytes = tsb;
  pre.t = t;
  pre.red = red;
  pre.K = &K;
  pre.s = s;
  FgmresOpts o;
  o.tol = tol;
  o.restart = restart;
  o.max_inner = max_inner;
  o.knorm = k_fro;
  return fgmres_ir(pre, K, n, rhs, d_x, o, fwork, fw, stats, s);
}
