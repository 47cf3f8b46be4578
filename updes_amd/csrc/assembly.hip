// K1: dense collocation assembly and the matrix-free products built on it.
//
// Reference: updes/assembly.py:10-137 (Phi, P, A, op(Phi, P)), 141-362
// (boundary rows), 366-401 (D = [[opPhi opP];[bdPhi bdP]]).  The reference
// assembles row by row with fori_loop + scatter over a support table that
// omits the node itself (updes/cloud.py:110-111), evaluating each entry by
// autodiff.  Here every entry is a closed form (rbf_device.h) and one launch
// writes the whole row-major (N+M) x (N+M) matrix: each thread owns two
// adjacent columns (16-byte non-temporal stores), loops over a tile of rows,
// and keeps its two centres in registers; the row context (node, normal,
// operator coefficients) is wave-uniform.  The kernel is bound by HBM writes.
//
// Local supports (support_size < N, updes/cloud.py:83-112): the same kernels,
// instantiated with LOC, zero every RBF column outside the row's support.  A
// row's support is an ascending id list; a column outside its [first, last]
// range is rejected with two compares (almost all of a row), the rest by a
// binary search in the list, so the dense write stays the bound.
#include <stdarg.h>

#include <algorithm>

#include "common.h"
#include "rbf_device.h"

namespace rbfx {

// Device copy of rbfx_problem (by value in the kernel arguments).
struct ProbDev {
  int64_t N, M, n;
  int64_t Ni, Nd, Nn, Nr, Npd;
  Kern k;
  const double* x;
  const double* y;
  const double* nx;
  const double* ny;
  const double* rnx;
  const double* rny;
  const double* beta;
  const int64_t* pd1;
  const int64_t* pd2;
  double rc[RBFX_NOPS];
  double mc[RBFX_NOPS];
  const double* rrow[RBFX_NOPS];
  const double* mrow[RBFX_NOPS];
  int matrix;
  bool local;          // RBFX_FLAG_LOCAL_SUPPORT: RBF columns restricted to sup
  int64_t S;           // support entries per node
  const int32_t* sup;  // (N, S) ascending node ids
};

enum RowClass { RC_INT = 0, RC_D, RC_N, RC_R, RC_PD, RC_PN, RC_PT, RC_VAL };

struct RowCtx {
  int cls;
  int mono;       // RC_PT: monomial id
  int64_t i;      // node of the row (self column excluded for INT/D/N/R/VAL)
  double xi, yi;
  double x2, y2;  // periodic partner
  double nx, ny;  // normal used by the RBF part (Neumann / Robin(bug) / pn first)
  double mnx, mny;  // normal used by the monomial part
  double n2x, n2y;  // pn second normal
  double beta;
  OpCoef rc, mc;
  int64_t i2;               // periodic partner node
  const int32_t* s1;        // local supports: support row of node i ...
  const int32_t* s2;        // ... and of the partner (periodic rows)
  int64_t lo1, hi1, lo2, hi2;  // first / last id of each support row
};

__device__ inline OpCoef make_coef(const double* c, const double* const* rows, int64_t i) {
  OpCoef o;
  double v[RBFX_NOPS];
#pragma unroll
  for (int t = 0; t < RBFX_NOPS; ++t) v[t] = c[t] + (rows[t] ? rows[t][i] : 0.0);
  o.v = v[RBFX_VAL];
  o.dx = v[RBFX_DX];
  o.dy = v[RBFX_DY];
  o.lap = v[RBFX_LAP];
  o.dxx = v[RBFX_DXX];
  o.dyy = v[RBFX_DYY];
  o.dxy = v[RBFX_DXY];
  o.hess = (o.dxx != 0.0) || (o.dyy != 0.0) || (o.dxy != 0.0);
  return o;
}

__device__ inline RowCtx row_ctx_base(const ProbDev& P, int64_t r) {
  RowCtx c;
  c.i2 = -1;
  c.mono = 0;
  c.beta = 0.0;
  c.nx = c.ny = c.mnx = c.mny = c.n2x = c.n2y = 0.0;
  c.x2 = c.y2 = 0.0;
  if (r >= P.N) {
    c.cls = RC_PT;
    c.mono = (int)(r - P.N);
    c.i = -1;
    c.xi = c.yi = 0.0;
    return c;
  }
  c.i = r;
  c.xi = P.x[r];
  c.yi = P.y[r];
  if (P.matrix == RBFX_MAT_A) {
    c.cls = RC_VAL;
    return c;
  }
  int64_t t = r;
  if (t < P.Ni) {
    c.cls = RC_INT;
    c.rc = make_coef(P.rc, P.rrow, r);
    c.mc = make_coef(P.mc, P.mrow, r);
    return c;
  }
  t -= P.Ni;
  if (t < P.Nd) {
    c.cls = RC_D;
    return c;
  }
  t -= P.Nd;
  const int64_t nrm0 = r - P.Ni - P.Nd;  // sorted_outward_normals index of node r
  if (t < P.Nn) {
    c.cls = RC_N;
    c.nx = c.mnx = P.nx[nrm0];
    c.ny = c.mny = P.ny[nrm0];
    return c;
  }
  t -= P.Nn;
  if (t < P.Nr) {
    c.cls = RC_R;
    c.beta = P.beta[t];
    c.nx = P.rnx[t];
    c.ny = P.rny[t];
    c.mnx = P.nx[nrm0];
    c.mny = P.ny[nrm0];
    return c;
  }
  t -= P.Nr;
  const bool flux = t >= P.Npd;
  if (flux) t -= P.Npd;
  const int64_t i1 = P.pd1[t], i2 = P.pd2[t];
  c.cls = flux ? RC_PN : RC_PD;
  c.i = i1;
  c.i2 = i2;
  c.xi = P.x[i1];
  c.yi = P.y[i1];
  c.x2 = P.x[i2];
  c.y2 = P.y[i2];
  if (flux) {
    c.nx = P.nx[i1 - P.Ni - P.Nd];
    c.ny = P.ny[i1 - P.Ni - P.Nd];
    c.n2x = P.nx[i2 - P.Ni - P.Nd];
    c.n2y = P.ny[i2 - P.Ni - P.Nd];
  }
  return c;
}

// Local supports (updes/cloud.py:108-112): the support row of a node and its
// id range; an empty range when S = 0.  The dense kernels only compare the
// ids; the support-only product reads the nodes they name and skips ids
// outside [0, N), so a malformed table gives wrong values but no stray access.
__device__ inline void support_row(const ProbDev& P, int64_t node, const int32_t*& s, int64_t& lo,
                                   int64_t& hi) {
  s = P.sup + node * P.S;
  if (P.S > 0) {
    lo = s[0];
    hi = s[P.S - 1];
  } else {
    lo = 1;
    hi = 0;
  }
}

__device__ inline bool in_support(const int32_t* s, int64_t S, int64_t lo, int64_t hi, int64_t j) {
  if (j < lo || j > hi) return false;  // most columns of a row end here
  int64_t a = 0, b = S - 1;
  while (a < b) {
    const int64_t m = (a + b) >> 1;
    if (s[m] < j)
      a = m + 1;
    else
      b = m;
  }
  return s[a] == j;
}

template <bool LOC>
__device__ inline RowCtx row_ctx(const ProbDev& P, int64_t r) {
  RowCtx c = row_ctx_base(P, r);
  if (LOC && c.cls != RC_PT) {
    support_row(P, c.i, c.s1, c.lo1, c.hi1);
    if (c.cls == RC_PD || c.cls == RC_PN) support_row(P, c.i2, c.s2, c.lo2, c.hi2);
  }
  return c;
}

// Entry (row context, centre column j < N).  With local supports the column
// must be in the row node's support, or (periodic rows, the concatenated
// supports of updes/assembly.py:219-221, 242-244) in its partner's.
template <bool LOC>
__device__ inline double elem_rbf(const ProbDev& P, const RowCtx& c, int64_t j, double xj,
                                  double yj) {
  if (LOC && c.cls != RC_PT) {
    bool in = in_support(c.s1, P.S, c.lo1, c.hi1, j);
    if (!in && (c.cls == RC_PD || c.cls == RC_PN)) in = in_support(c.s2, P.S, c.lo2, c.hi2, j);
    if (!in) return 0.0;
  }
  switch (c.cls) {
    case RC_INT:
      return j == c.i ? 0.0 : apply_op(P.k, c.rc, c.xi - xj, c.yi - yj);
    case RC_D:
    case RC_VAL: {
      if (j == c.i) return 0.0;
      const double dx = c.xi - xj, dy = c.yi - yj;
      return radial_value(P.k, sqrt(dist2(dx, dy)));
    }
    case RC_N: {
      if (j == c.i) return 0.0;
      const double dx = c.xi - xj, dy = c.yi - yj;
      return radial_f1(P.k, sqrt(dist2(dx, dy))) * __builtin_fma(dx, c.nx, dy * c.ny);
    }
    case RC_R: {
      if (j == c.i) return 0.0;
      const double dx = c.xi - xj, dy = c.yi - yj;
      const double r = sqrt(dist2(dx, dy));
      if (r == 0.0) return c.beta * P.k.phi0;
      double phi, f1, g2;
      radial(P.k, r, phi, f1, g2);
      return __builtin_fma(c.beta, phi, f1 * __builtin_fma(dx, c.nx, dy * c.ny));
    }
    case RC_PD: {
      const double dx1 = c.xi - xj, dy1 = c.yi - yj;
      const double dx2 = c.x2 - xj, dy2 = c.y2 - yj;
      return radial_value(P.k, sqrt(dist2(dx1, dy1))) - radial_value(P.k, sqrt(dist2(dx2, dy2)));
    }
    case RC_PN: {
      const double dx1 = c.xi - xj, dy1 = c.yi - yj;
      const double dx2 = c.x2 - xj, dy2 = c.y2 - yj;
      return __builtin_fma(radial_f1(P.k, sqrt(dist2(dx1, dy1))), __builtin_fma(dx1, c.nx, dy1 * c.ny),
                           radial_f1(P.k, sqrt(dist2(dx2, dy2))) * __builtin_fma(dx2, c.n2x, dy2 * c.n2y));
    }
    default:  // RC_PT: row N+k holds p_k(x_j)
    {
      OpCoef v{1.0, 0, 0, 0, 0, 0, 0, false};
      return mono_apply(c.mono, xj, yj, v);
    }
  }
}

// Entry (row context, monomial column N+k).
__device__ inline double elem_mono(const RowCtx& c, int kid) {
  double o[RBFX_NOPS];
  switch (c.cls) {
    case RC_INT:
      return mono_apply(kid, c.xi, c.yi, c.mc);
    case RC_D:
    case RC_VAL:
      mono_ops(kid, c.xi, c.yi, o);
      return o[RBFX_VAL];
    case RC_N:
      mono_ops(kid, c.xi, c.yi, o);
      return o[RBFX_DX] * c.mnx + o[RBFX_DY] * c.mny;
    case RC_R:
      mono_ops(kid, c.xi, c.yi, o);
      return c.beta * o[RBFX_VAL] + o[RBFX_DX] * c.mnx + o[RBFX_DY] * c.mny;
    case RC_PD: {
      double o2[RBFX_NOPS];
      mono_ops(kid, c.xi, c.yi, o);
      mono_ops(kid, c.x2, c.y2, o2);
      return o[RBFX_VAL] - o2[RBFX_VAL];
    }
    case RC_PN: {
      double o2[RBFX_NOPS];
      mono_ops(kid, c.xi, c.yi, o);
      mono_ops(kid, c.x2, c.y2, o2);
      return (o[RBFX_DX] * c.nx + o[RBFX_DY] * c.ny) + (o2[RBFX_DX] * c.n2x + o2[RBFX_DY] * c.n2y);
    }
    default:
      return 0.0;  // the M x M block of [P^T 0]
  }
}

typedef double vd2 __attribute__((ext_vector_type(2)));
typedef float vf2 __attribute__((ext_vector_type(2)));

constexpr int kAsmThreads = 256;
// Each thread writes kAsmPairs adjacent column pairs of every row of its
// block (pairs 2 kAsmThreads columns apart, so every store instruction of a
// wave covers 1 KB contiguous).  More pairs per thread (amortising the
// wave-uniform row context) measured slower: 4 pairs 5.47 vs 5.16 ms for K
// at 40k (176 VGPRs, two waves per SIMD).
constexpr int kAsmPairs = 1;
constexpr int kAsmCols = 2 * kAsmThreads * kAsmPairs;  // columns per block
constexpr int kAsmRows = 16;                            // rows per block

// sqrt of x for 2^-900 <= x <= 2^900: the library's FP64 square-root
// sequence (rsq estimate, a Goldschmidt step, two Newton corrections)
// without its range scaling and its 0 / inf fixups, so the same result.
__device__ inline double sqrt_mid(double x) {
  const double g = __builtin_amdgcn_rsq(x);
  double s = x * g, h = 0.5 * g;
  const double r = __builtin_fma(-h, s, 0.5);
  s = __builtin_fma(s, r, s);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-s, s, x);
  s = __builtin_fma(d, h, s);
  d = __builtin_fma(-s, s, x);
  return __builtin_fma(d, h, s);
}

// Entry of an interior collocation row (the operator applied to
// phi(., x_j) at x_i) without branches: apply_op's arithmetic, r = 0 and
// the zero diagonal (j == i) as selects.  mid: every r^2 of the wave is in
// sqrt_mid's range (r^2 = 0 aside).
template <int KIND, bool FAST, bool HESS>
__device__ inline double interior_entry(const Kern& k, const OpCoef& c, double dx, double dy, bool diag,
                                        bool mid) {
  const double r2 = dist2(dx, dy);
  const bool zero = !(r2 > 0.0);
  const double rs2 = zero ? 1.0 : r2;
  const double r = mid ? sqrt_mid(rs2) : sqrt(rs2);
  const double v = apply_op_pos_t<HESS>(k, c, dx, dy, rs2, r);
  return diag ? 0.0 : (zero ? c.v * k.phi0 : v);
}

__device__ inline bool in_sqrt_mid(double r2) { return !(r2 > 0.0) || (r2 >= 0x1p-900 && r2 <= 0x1p900); }

template <int KIND, bool FAST, bool LOC, bool W64, bool W32>
__global__ __launch_bounds__(kAsmThreads) void assemble_kernel(ProbDev P, int64_t row0,
                                                               int64_t nrows, double* out64,
                                                               int64_t ld64, float* out32,
                                                               int64_t ld32) {
  pin_kind<KIND, FAST>(P.k);
  const int64_t jb = (int64_t)blockIdx.x * kAsmCols + 2 * threadIdx.x;
  const int64_t rb = (int64_t)blockIdx.y * kAsmRows;
  if (jb >= P.n) return;
  double xj[kAsmPairs][2], yj[kAsmPairs][2];
#pragma unroll
  for (int q = 0; q < kAsmPairs; ++q)
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int64_t j = jb + 2 * kAsmThreads * q + e;
      xj[q][e] = j < P.N ? P.x[j] : 0.0;
      yj[q][e] = j < P.N ? P.y[j] : 0.0;
    }
  const int64_t rend = rb + kAsmRows < nrows ? rb + kAsmRows : nrows;
  for (int64_t rr = rb; rr < rend; ++rr) {
    const int64_t r = row0 + rr;
    const RowCtx c = row_ctx<LOC>(P, r);
#pragma unroll
    for (int q = 0; q < kAsmPairs; ++q) {
      const int64_t j0 = jb + 2 * kAsmThreads * q;
      if (j0 >= P.n) break;
      const bool two = (j0 + 1) < P.n;
      const double v0 = j0 < P.N ? elem_rbf<LOC>(P, c, j0, xj[q][0], yj[q][0]) : elem_mono(c, (int)(j0 - P.N));
      double v1 = 0.0;
      if (two)
        v1 = (j0 + 1) < P.N ? elem_rbf<LOC>(P, c, j0 + 1, xj[q][1], yj[q][1]) : elem_mono(c, (int)(j0 + 1 - P.N));
      if (W64) {
        double* dst = out64 + rr * ld64 + j0;
        if (two && (((uintptr_t)dst & 15) == 0)) {
          vd2 v = {v0, v1};
          __builtin_nontemporal_store(v, reinterpret_cast<vd2*>(dst));
        } else {
          __builtin_nontemporal_store(v0, dst);
          if (two) __builtin_nontemporal_store(v1, dst + 1);
        }
      }
      if (W32) {
        float* dst = out32 + rr * ld32 + j0;
        if (two && (((uintptr_t)dst & 7) == 0)) {
          vf2 v = {(float)v0, (float)v1};
          __builtin_nontemporal_store(v, reinterpret_cast<vf2*>(dst));
        } else {
          __builtin_nontemporal_store((float)v0, dst);
          if (two) __builtin_nontemporal_store((float)v1, dst + 1);
        }
      }
    }
  }
}

// y = K x: one block per kMvRows rows; each thread strides over columns.
constexpr int kMvThreads = 256;
constexpr int kMvRows = 4;

template <int KIND, bool FAST, bool LOC>
__global__ __launch_bounds__(kMvThreads) void matvec_kernel(ProbDev P, const double* __restrict__ xv,
                                                            double* __restrict__ yv, int64_t row_base) {
  pin_kind<KIND, FAST>(P.k);
  const int64_t r0 = row_base + (int64_t)blockIdx.x * kMvRows;
  RowCtx c[kMvRows];
  double acc[kMvRows];
#pragma unroll
  for (int q = 0; q < kMvRows; ++q) {
    const int64_t r = r0 + q < P.n ? r0 + q : P.n - 1;
    c[q] = row_ctx<LOC>(P, r);
    acc[q] = 0.0;
  }
  for (int64_t j = threadIdx.x; j < P.n; j += kMvThreads) {
    const double w = xv[j];
    if (j < P.N) {
      const double xj = P.x[j], yj = P.y[j];
#pragma unroll
      for (int q = 0; q < kMvRows; ++q) acc[q] += elem_rbf<LOC>(P, c[q], j, xj, yj) * w;
    } else {
#pragma unroll
      for (int q = 0; q < kMvRows; ++q) acc[q] += elem_mono(c[q], (int)(j - P.N)) * w;
    }
  }
  __shared__ double red[kMvRows][kMvThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kMvRows; ++q) {
    double v = acc[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[q][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < kMvRows) {
    const int q = threadIdx.x;
    double s = 0.0;
    for (int w = 0; w < kMvThreads / 64; ++w) s += red[q][w];
    if (r0 + q < P.n) yv[r0 + q] = s;
  }
}

// y = K x (or A x) on local supports, visiting only the support entries:
// one thread per node row (its support columns, for a periodic row also the
// partner's columns that are not in its own support, then the monomial
// columns), and one workgroup per P^T row (a reduction over the N nodes).
constexpr int kLocThreads = 256;

template <int KIND, bool FAST>
__global__ __launch_bounds__(kLocThreads) void matvec_local_kernel(ProbDev P, const double* __restrict__ xv,
                                                                   double* __restrict__ yv) {
  pin_kind<KIND, FAST>(P.k);
  const int64_t nbn = (P.N + kLocThreads - 1) / kLocThreads;
  if ((int64_t)blockIdx.x < nbn) {
    const int64_t r = (int64_t)blockIdx.x * kLocThreads + threadIdx.x;
    if (r >= P.N) return;
    const RowCtx c = row_ctx<true>(P, r);
    double acc = 0.0;
    // ids outside [0, N) (a malformed table) are skipped, never dereferenced
    for (int64_t q = 0; q < P.S; ++q) {
      const int64_t j = c.s1[q];
      if ((uint64_t)j < (uint64_t)P.N) acc += elem_rbf<false>(P, c, j, P.x[j], P.y[j]) * xv[j];
    }
    if (c.cls == RC_PD || c.cls == RC_PN) {
      for (int64_t q = 0; q < P.S; ++q) {
        const int64_t j = c.s2[q];
        if ((uint64_t)j < (uint64_t)P.N && !in_support(c.s1, P.S, c.lo1, c.hi1, j))
          acc += elem_rbf<false>(P, c, j, P.x[j], P.y[j]) * xv[j];
      }
    }
    for (int k = 0; k < (int)P.M; ++k) acc += elem_mono(c, k) * xv[P.N + k];
    yv[r] = acc;
    return;
  }
  // P^T row N + k: sum_j p_k(x_j) x_j over the nodes
  const int64_t r = P.N + ((int64_t)blockIdx.x - nbn);
  const RowCtx c = row_ctx<false>(P, r);
  double acc = 0.0;
  for (int64_t j = threadIdx.x; j < P.N; j += kLocThreads) acc += elem_rbf<false>(P, c, j, P.x[j], P.y[j]) * xv[j];
  __shared__ double red[kLocThreads / 64];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kLocThreads / 64; ++w) t += red[w];
    yv[r] = t;
  }
}

// y = K^T x (or A^T x) on local supports over the support entries only.
// The rows that name each column are gathered first into a reverse table
// (count, scan, fill; each column's rows then sorted, so the sums run in a
// fixed order), then one thread per column sums over its rows and the P^T
// rows, and one workgroup per monomial column sums over the node rows.
// Column set of node-matrix row r: its node's support, for a periodic row
// also the partner's ids that are not in it.  Ids outside [0, N) are skipped.
template <typename F>
__device__ inline void for_row_cols(const ProbDev& P, const RowCtx& c, F f) {
  for (int64_t q = 0; q < P.S; ++q) {
    const int64_t j = c.s1[q];
    if ((uint64_t)j < (uint64_t)P.N) f(j);
  }
  if (c.cls == RC_PD || c.cls == RC_PN)
    for (int64_t q = 0; q < P.S; ++q) {
      const int64_t j = c.s2[q];
      if ((uint64_t)j < (uint64_t)P.N && !in_support(c.s1, P.S, c.lo1, c.hi1, j)) f(j);
    }
}

__global__ void zero_i32_kernel(int* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ __launch_bounds__(kLocThreads) void rev_count_kernel(ProbDev P, int* __restrict__ cnt) {
  const int64_t r = (int64_t)blockIdx.x * kLocThreads + threadIdx.x;
  if (r >= P.N) return;
  const RowCtx c = row_ctx<true>(P, r);
  for_row_cols(P, c, [&](int64_t j) { atomicAdd(cnt + j, 1); });
}

// Exclusive scan of cnt[0..N) into off[0..N] by one workgroup of 1024.
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void rev_scan_kernel(const int* __restrict__ cnt, int64_t N,
                                                                int* __restrict__ off) {
  __shared__ int part[kScanThreads];
  const int t = threadIdx.x;
  const int64_t per = (N + kScanThreads - 1) / kScanThreads;
  const int64_t b = t * per, e = b + per < N ? b + per : N;
  int sum = 0;
  for (int64_t i = b; i < e; ++i) sum += cnt[i];
  part[t] = sum;
  __syncthreads();
  for (int d = 1; d < kScanThreads; d <<= 1) {  // inclusive Hillis-Steele scan
    const int v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? part[t - 1] : 0;
  for (int64_t i = b; i < e; ++i) {
    off[i] = run;
    run += cnt[i];
  }
  if (t == kScanThreads - 1) off[N] = part[t];
}

__global__ __launch_bounds__(kLocThreads) void rev_fill_kernel(ProbDev P, const int* __restrict__ off,
                                                               int* __restrict__ fill, int* __restrict__ rows) {
  const int64_t r = (int64_t)blockIdx.x * kLocThreads + threadIdx.x;
  if (r >= P.N) return;
  const RowCtx c = row_ctx<true>(P, r);
  for_row_cols(P, c, [&](int64_t j) { rows[off[j] + atomicAdd(fill + j, 1)] = (int)r; });
}

template <int KIND, bool FAST>
__global__ __launch_bounds__(kLocThreads) void matvec_local_t_kernel(ProbDev P, const int* __restrict__ off,
                                                                     int* __restrict__ rows,
                                                                     const double* __restrict__ xv,
                                                                     double* __restrict__ yv) {
  pin_kind<KIND, FAST>(P.k);
  const int64_t nbn = (P.N + kLocThreads - 1) / kLocThreads;
  if ((int64_t)blockIdx.x < nbn) {
    const int64_t j = (int64_t)blockIdx.x * kLocThreads + threadIdx.x;
    if (j >= P.N) return;
    const int b = off[j], e = off[j + 1];
    for (int a = b + 1; a < e; ++a) {  // insertion sort of this column's rows
      const int v = rows[a];
      int q = a - 1;
      while (q >= b && rows[q] > v) {
        rows[q + 1] = rows[q];
        --q;
      }
      rows[q + 1] = v;
    }
    const double xj = P.x[j], yj = P.y[j];
    double acc = 0.0;
    for (int a = b; a < e; ++a) {
      const int64_t r = rows[a];
      acc += elem_rbf<false>(P, row_ctx<false>(P, r), j, xj, yj) * xv[r];
    }
    for (int k = 0; k < (int)P.M; ++k) acc += elem_rbf<false>(P, row_ctx<false>(P, P.N + k), j, xj, yj) * xv[P.N + k];
    yv[j] = acc;
    return;
  }
  // monomial column N + k: sum over the node rows
  const int k = (int)((int64_t)blockIdx.x - nbn);
  double acc = 0.0;
  for (int64_t r = threadIdx.x; r < P.N; r += kLocThreads) acc += elem_mono(row_ctx<false>(P, r), k) * xv[r];
  __shared__ double red[kLocThreads / 64];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kLocThreads / 64; ++w) t += red[w];
    yv[P.N + k] = t;
  }
}

// Workspace of the reverse table: cnt, fill (N each), off (N+1), rows (N*2S).
static size_t local_t_ws_bytes(int64_t N, int64_t S) {
  return 2 * ws_round(sizeof(int) * (size_t)N) + ws_round(sizeof(int) * (size_t)(N + 1)) +
         ws_round(sizeof(int) * (size_t)N * (size_t)(2 * S > 0 ? 2 * S : 1));
}

// Partial y^T = x^T K over a chunk of rows: part[chunk][j].
constexpr int kMtThreads = 256;

template <int KIND, bool FAST, bool LOC>
__global__ __launch_bounds__(kMtThreads) void matvec_t_kernel(ProbDev P, const double* __restrict__ xv,
                                                              int64_t rows_per_chunk,
                                                              double* __restrict__ part) {
  pin_kind<KIND, FAST>(P.k);
  const int64_t j = (int64_t)blockIdx.x * kMtThreads + threadIdx.x;
  const int64_t rs = (int64_t)blockIdx.y * rows_per_chunk;
  const int64_t re = rs + rows_per_chunk < P.n ? rs + rows_per_chunk : P.n;
  const bool valid = j < P.n;
  const int64_t jj = valid ? j : 0;
  const bool isrbf = jj < P.N;
  const double xj = isrbf ? P.x[jj] : 0.0, yj = isrbf ? P.y[jj] : 0.0;
  double acc = 0.0;
  for (int64_t r = rs; r < re; ++r) {
    const RowCtx c = row_ctx<LOC>(P, r);
    const double e = isrbf ? elem_rbf<LOC>(P, c, jj, xj, yj) : elem_mono(c, (int)(jj - P.N));
    acc += e * xv[r];
  }
  if (valid) part[blockIdx.y * P.n + j] = acc;
}

__global__ void sum_chunks_kernel(const double* __restrict__ part, int64_t chunks, int64_t n,
                                  double* __restrict__ y) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double s = 0.0;
  for (int64_t c = 0; c < chunks; ++c) s += part[c * n + j];
  y[j] = s;
}

// y = A x over the value rows of A (rows < N, every node in every support):
// kMvRows rows per workgroup, each thread strides over the centre columns,
// then the monomial columns.  Where every r^2 of the wave is in sqrt_mid's
// range r comes from sqrt_mid (the same bits as sqrt), so the entries are
// elem_rbf's for these rows; the rows of P^T go through matvec_kernel.
template <int KIND, bool FAST>
__global__ __launch_bounds__(kMvThreads) void matvec_value_kernel(ProbDev P, const double* __restrict__ xv,
                                                                  double* __restrict__ yv) {
  pin_kind<KIND, FAST>(P.k);
  const int64_t r0 = (int64_t)blockIdx.x * kMvRows;
  int64_t ri[kMvRows];
  double xi[kMvRows], yi[kMvRows], acc[kMvRows];
#pragma unroll
  for (int q = 0; q < kMvRows; ++q) {
    ri[q] = r0 + q < P.N ? r0 + q : P.N - 1;
    xi[q] = P.x[ri[q]];
    yi[q] = P.y[ri[q]];
    acc[q] = 0.0;
  }
  for (int64_t j = threadIdx.x; j < P.N; j += kMvThreads) {
    const double w = xv[j], xj = P.x[j], yj = P.y[j];
    double d2[kMvRows];
    bool in = true;
#pragma unroll
    for (int q = 0; q < kMvRows; ++q) {
      d2[q] = dist2(xi[q] - xj, yi[q] - yj);
      in = in && in_sqrt_mid(d2[q]);
    }
    if (__all(in)) {
#pragma unroll
      for (int q = 0; q < kMvRows; ++q) {
        const double r = d2[q] > 0.0 ? sqrt_mid(d2[q]) : 0.0;
        const double e = j == ri[q] ? 0.0 : radial_value(P.k, r);
        acc[q] = __builtin_fma(e, w, acc[q]);
      }
    } else {
#pragma unroll
      for (int q = 0; q < kMvRows; ++q) {
        const double e = j == ri[q] ? 0.0 : radial_value(P.k, sqrt(d2[q]));
        acc[q] = __builtin_fma(e, w, acc[q]);
      }
    }
  }
  if (threadIdx.x < P.M) {
    const int m = threadIdx.x;
    const double w = xv[P.N + m];
#pragma unroll
    for (int q = 0; q < kMvRows; ++q) {
      double o[RBFX_NOPS];
      mono_ops(m, xi[q], yi[q], o);
      acc[q] = __builtin_fma(o[RBFX_VAL], w, acc[q]);
    }
  }
  __shared__ double red[kMvRows][kMvThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kMvRows; ++q) {
    double v = acc[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[q][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < kMvRows) {
    const int q = threadIdx.x;
    double t = 0.0;
    for (int w = 0; w < kMvThreads / 64; ++w) t += red[q][w];
    if (r0 + q < P.N) yv[r0 + q] = t;
  }
}

static int make_prob(const rbfx_problem* p, int matrix, ProbDev& P) {
  RBFX_REQUIRE(p != nullptr, "null problem");
  RBFX_REQUIRE(p->N > 0 && p->M >= 0 && p->M <= 15, "bad N=%lld / M=%lld", (long long)p->N,
               (long long)p->M);
  RBFX_REQUIRE(p->kernel >= 0 && p->kernel <= 4, "unknown kernel %d", p->kernel);
  RBFX_REQUIRE(matrix == RBFX_MAT_K || matrix == RBFX_MAT_A, "unknown matrix %d", matrix);
  RBFX_REQUIRE(p->d_x && p->d_y, "node coordinates missing");
  P.N = p->N;
  P.M = p->M;
  P.n = p->N + p->M;
  P.Ni = p->Ni;
  P.Nd = p->Nd;
  P.Nn = p->Nn;
  P.Nr = p->Nr;
  P.Npd = p->Npd;
  P.k = make_kern(p->kernel, p->param);
  P.x = p->d_x;
  P.y = p->d_y;
  P.nx = p->d_nx;
  P.ny = p->d_ny;
  P.rnx = p->d_robin_normal_x;
  P.rny = p->d_robin_normal_y;
  P.beta = p->d_beta;
  P.pd1 = p->d_pd_i1;
  P.pd2 = p->d_pd_i2;
  for (int t = 0; t < RBFX_NOPS; ++t) {
    P.rc[t] = p->op_rbf_const[t];
    P.mc[t] = p->op_mon_const[t];
    P.rrow[t] = p->d_op_rbf_row[t];
    P.mrow[t] = p->d_op_mon_row[t];
  }
  P.matrix = matrix;
  RBFX_REQUIRE((p->flags & ~(int32_t)RBFX_FLAG_LOCAL_SUPPORT) == 0, "unknown flags 0x%x", p->flags);
  P.local = (p->flags & RBFX_FLAG_LOCAL_SUPPORT) != 0;
  P.S = 0;
  P.sup = nullptr;
  if (P.local) {
    RBFX_REQUIRE(p->support_cols >= 0 && p->support_cols < p->N,
                 "support_cols=%lld must be in [0, N-1=%lld]", (long long)p->support_cols,
                 (long long)(p->N - 1));
    RBFX_REQUIRE(p->support_cols == 0 || p->d_support, "support table missing");
    P.S = p->support_cols;
    P.sup = p->d_support;
  }
  if (matrix == RBFX_MAT_K) {
    RBFX_REQUIRE(p->Ni + p->Nd + p->Nn + p->Nr + 2 * p->Npd == p->N,
                 "row classes do not add up: Ni+Nd+Nn+Nr+2Npd=%lld != N=%lld",
                 (long long)(p->Ni + p->Nd + p->Nn + p->Nr + 2 * p->Npd), (long long)p->N);
    RBFX_REQUIRE(p->Ni >= 0 && p->Nd >= 0 && p->Nn >= 0 && p->Nr >= 0 && p->Npd >= 0,
                 "negative row-class size");
    RBFX_REQUIRE((p->Nn + p->Nr + p->Npd == 0) || (p->d_nx && p->d_ny), "normals missing");
    RBFX_REQUIRE(p->Nr == 0 || (p->d_beta && p->d_robin_normal_x && p->d_robin_normal_y),
                 "Robin data missing");
    RBFX_REQUIRE(p->Npd == 0 || (p->d_pd_i1 && p->d_pd_i2), "periodic pairs missing");
  }
  return RBFX_OK;
}

// The interior collocation rows of K (rows < Ni, every node in every
// support): a kernel of their own, so that their entries compile to
// straight-line code (no row-class switch, r = 0 and the zero diagonal as
// selects) with a register budget of their own.  Same arithmetic as
// elem_rbf / elem_mono for these rows.
// Column pairs per thread of the interior kernel; measured at 40k: 2 pairs
// 3.49 vs 3.45 ms, and plain instead of non-temporal stores 5.19 ms.
constexpr int kIntPairs = 1;

// A wave-uniform double as scalar registers (its value is the same in every
// lane), so that branches on it are scalar branches.
__device__ inline double wave_uniform(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ inline OpCoef uniform_coef(OpCoef c) {
  c.v = wave_uniform(c.v);
  c.dx = wave_uniform(c.dx);
  c.dy = wave_uniform(c.dy);
  c.lap = wave_uniform(c.lap);
  c.dxx = wave_uniform(c.dxx);
  c.dyy = wave_uniform(c.dyy);
  c.dxy = wave_uniform(c.dxy);
  c.hess = __builtin_amdgcn_readfirstlane((int)c.hess) != 0;
  return c;
}

// One row of the interior kernel: the thread's column pairs, stored.
template <int KIND, bool FAST, bool HESS, bool W64, bool W32>
__device__ inline void interior_row(const ProbDev& P, const OpCoef& oc, int64_t r, int64_t rr, int64_t jb,
                                    const double (&xj)[kIntPairs][2], const double (&yj)[kIntPairs][2],
                                    double* out64, int64_t ld64, float* out32, int64_t ld32) {
  const double xi = P.x[r], yi = P.y[r];
#pragma unroll
  for (int q = 0; q < kIntPairs; ++q) {
    const int64_t j0 = jb + 2 * kAsmThreads * q;
    if (j0 >= P.n) break;
    const bool two = (j0 + 1) < P.n;
    double v0, v1 = 0.0;
    if ((j0 + 1) < P.N) {  // both columns are centres
      const double dx0 = xi - xj[q][0], dy0 = yi - yj[q][0];
      const double dx1 = xi - xj[q][1], dy1 = yi - yj[q][1];
      const bool mid = __all(in_sqrt_mid(dist2(dx0, dy0)) && in_sqrt_mid(dist2(dx1, dy1)));
      v0 = interior_entry<KIND, FAST, HESS>(P.k, oc, dx0, dy0, j0 == r, mid);
      v1 = interior_entry<KIND, FAST, HESS>(P.k, oc, dx1, dy1, j0 + 1 == r, mid);
    } else {
      // the last centre and the monomial columns
      const OpCoef mc = make_coef(P.mc, P.mrow, r);
      v0 = j0 < P.N ? interior_entry<KIND, FAST, HESS>(P.k, oc, xi - xj[q][0], yi - yj[q][0], j0 == r, false)
                    : mono_apply((int)(j0 - P.N), xi, yi, mc);
      if (two) v1 = mono_apply((int)(j0 + 1 - P.N), xi, yi, mc);
    }
    if (W64) {
      double* dst = out64 + rr * ld64 + j0;
      if (two && (((uintptr_t)dst & 15) == 0)) {
        vd2 v = {v0, v1};
        __builtin_nontemporal_store(v, reinterpret_cast<vd2*>(dst));
      } else {
        __builtin_nontemporal_store(v0, dst);
        if (two) __builtin_nontemporal_store(v1, dst + 1);
      }
    }
    if (W32) {
      float* dst = out32 + rr * ld32 + j0;
      if (two && (((uintptr_t)dst & 7) == 0)) {
        vf2 v = {(float)v0, (float)v1};
        __builtin_nontemporal_store(v, reinterpret_cast<vf2*>(dst));
      } else {
        __builtin_nontemporal_store((float)v0, dst);
        if (two) __builtin_nontemporal_store((float)v1, dst + 1);
      }
    }
  }
}

// Rows [rb, rend) of the block: with per-row operator coefficients (PERROW)
// each row loads its own, else the constant ones in oc serve every row.
template <int KIND, bool FAST, bool HESS, bool PERROW, bool W64, bool W32>
__device__ inline void interior_rows(const ProbDev& P, OpCoef oc, int64_t row0, int64_t rb, int64_t rend,
                                     int64_t jb, const double (&xj)[kIntPairs][2],
                                     const double (&yj)[kIntPairs][2], double* out64, int64_t ld64, float* out32,
                                     int64_t ld32) {
  for (int64_t rr = rb; rr < rend; ++rr) {
    const int64_t r = row0 + rr;
    if (PERROW) oc = uniform_coef(make_coef(P.rc, P.rrow, r));
    interior_row<KIND, FAST, HESS, W64, W32>(P, oc, r, rr, jb, xj, yj, out64, ld64, out32, ld32);
  }
}

template <int KIND, bool FAST, bool W64, bool W32>
__global__ __launch_bounds__(kAsmThreads) void assemble_interior_kernel(ProbDev P, int64_t row0, int64_t nrows,
                                                                        double* out64, int64_t ld64, float* out32,
                                                                        int64_t ld32) {
  pin_kind<KIND, FAST>(P.k);
  const int64_t jb = (int64_t)blockIdx.x * (2 * kAsmThreads * kIntPairs) + 2 * threadIdx.x;
  const int64_t rb = (int64_t)blockIdx.y * kAsmRows;
  if (jb >= P.n) return;
  double xj[kIntPairs][2], yj[kIntPairs][2];
#pragma unroll
  for (int q = 0; q < kIntPairs; ++q)
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int64_t j = jb + 2 * kAsmThreads * q + e;
      xj[q][e] = j < P.N ? P.x[j] : 0.0;
      yj[q][e] = j < P.N ? P.y[j] : 0.0;
    }
  const int64_t rend = rb + kAsmRows < nrows ? rb + kAsmRows : nrows;
  bool perrow = false;
#pragma unroll
  for (int t = 0; t < RBFX_NOPS; ++t) perrow |= P.rrow[t] != nullptr;
  // the coefficients are the same in every lane (scalar registers); without
  // per-row arrays they are the same for every row, and rows without
  // second-derivative terms compile without their division
  const OpCoef oc = uniform_coef(make_coef(P.rc, P.rrow, row0 + rb));
  if (perrow)
    interior_rows<KIND, FAST, true, true, W64, W32>(P, oc, row0, rb, rend, jb, xj, yj, out64, ld64, out32, ld32);
  else if (oc.hess)
    interior_rows<KIND, FAST, true, false, W64, W32>(P, oc, row0, rb, rend, jb, xj, yj, out64, ld64, out32, ld32);
  else
    interior_rows<KIND, FAST, false, false, W64, W32>(P, oc, row0, rb, rend, jb, xj, yj, out64, ld64, out32, ld32);
}

}  // namespace rbfx

using namespace rbfx;

template <int KIND, bool FAST>
static void launch_interior(dim3 grid, hipStream_t s, const ProbDev& P, int64_t row0, int64_t nrows,
                            double* out64, int64_t ld64, float* out32, int64_t ld32) {
  if (out64 && out32)
    hipLaunchKernelGGL((assemble_interior_kernel<KIND, FAST, true, true>), grid, dim3(kAsmThreads), 0, s, P, row0,
                       nrows, out64, ld64, out32, ld32);
  else if (out64)
    hipLaunchKernelGGL((assemble_interior_kernel<KIND, FAST, true, false>), grid, dim3(kAsmThreads), 0, s, P, row0,
                       nrows, out64, ld64, out32, ld32);
  else
    hipLaunchKernelGGL((assemble_interior_kernel<KIND, FAST, false, true>), grid, dim3(kAsmThreads), 0, s, P, row0,
                       nrows, out64, ld64, out32, ld32);
}

template <int KIND, bool FAST, bool LOC>
static void launch_assemble(dim3 grid, hipStream_t s, const ProbDev& P, int64_t row0, int64_t nrows,
                            double* out64, int64_t ld64, float* out32, int64_t ld32) {
  if (out64 && out32)
    hipLaunchKernelGGL((assemble_kernel<KIND, FAST, LOC, true, true>), grid, dim3(kAsmThreads), 0, s, P, row0,
                       nrows, out64, ld64, out32, ld32);
  else if (out64)
    hipLaunchKernelGGL((assemble_kernel<KIND, FAST, LOC, true, false>), grid, dim3(kAsmThreads), 0, s, P, row0,
                       nrows, out64, ld64, out32, ld32);
  else
    hipLaunchKernelGGL((assemble_kernel<KIND, FAST, LOC, false, true>), grid, dim3(kAsmThreads), 0, s, P, row0,
                       nrows, out64, ld64, out32, ld32);
}

extern "C" int rbfx_assemble(const rbfx_problem* p, int matrix, int64_t row0, int64_t nrows,
                             double* d_out64, int64_t ld64, float* d_out32, int64_t ld32,
                             void* stream) {
  ProbDev P;
  int st = make_prob(p, matrix, P);
  if (st) return st;
  RBFX_REQUIRE(row0 >= 0 && nrows >= 0 && row0 + nrows <= P.n, "row range out of bounds");
  RBFX_REQUIRE(d_out64 || d_out32, "no output");
  RBFX_REQUIRE(!d_out64 || ld64 >= P.n, "ld64 < n");
  RBFX_REQUIRE(!d_out32 || ld32 >= P.n, "ld32 < n");
  if (nrows == 0) return RBFX_OK;
  RBFX_REQUIRE(ceil_div(nrows, kAsmRows) <= 65535ll * 1024ll, "too many rows");
  hipStream_t s = as_stream(stream);
  const int pid = prof_start(d_out64 ? (d_out32 ? "assemble_f64_f32" : "assemble_f64") : "assemble_f32", s);
  // rows [row0, row0 + ni) are interior collocation rows of K
  const int64_t ni = (matrix == RBFX_MAT_K && !P.local) ? std::max<int64_t>(0, std::min(row0 + nrows, P.Ni) - row0) : 0;
  if (ni > 0) {
    dim3 gi((unsigned)ceil_div(P.n, 2 * kAsmThreads * kIntPairs), (unsigned)ceil_div(ni, kAsmRows));
    double* o64 = d_out64;
    float* o32 = d_out32;
    RBFX_KIND_SWITCH(P.k, { launch_interior<KIND_, FAST_>(gi, s, P, row0, ni, o64, ld64, o32, ld32); });
  }
  const int64_t r1 = row0 + ni, nr = nrows - ni;
  if (nr > 0) {
    dim3 grid((unsigned)ceil_div(P.n, kAsmCols), (unsigned)ceil_div(nr, kAsmRows));
    double* o64 = d_out64 ? d_out64 + ni * ld64 : nullptr;
    float* o32 = d_out32 ? d_out32 + ni * ld32 : nullptr;
    RBFX_KIND_SWITCH(P.k, {
      if (P.local)
        launch_assemble<KIND_, FAST_, true>(grid, s, P, r1, nr, o64, ld64, o32, ld32);
      else
        launch_assemble<KIND_, FAST_, false>(grid, s, P, r1, nr, o64, ld64, o32, ld32);
    });
  }
  prof_stop(pid, s, 0.0, (double)nrows * P.n * ((d_out64 ? 8.0 : 0.0) + (d_out32 ? 4.0 : 0.0)));
  RBFX_CHECK_LAUNCH();
  return RBFX_OK;
}

static int64_t mt_chunks(int64_t n) {
  // enough blocks to fill the chip: ceil(n/256) column blocks x chunks >= 4 * CUs
  const int64_t colb = ceil_div(n, kMtThreads);
  int64_t want = ceil_div(4 * (int64_t)device_info().num_cus, colb);
  if (want < 1) want = 1;
  if (want > n) want = n;
  return want;
}

namespace rbfx {
size_t matvec_t_ws_bytes(int64_t n) { return ws_round(sizeof(double) * (size_t)mt_chunks(n) * (size_t)n); }
}  // namespace rbfx

extern "C" size_t rbfx_matvec_workspace(const rbfx_problem* p, int trans) {
  if (!p || !trans) return 0;
  const size_t dense = matvec_t_ws_bytes(p->N + p->M);
  if ((p->flags & RBFX_FLAG_LOCAL_SUPPORT) && p->support_cols >= 0 && p->N > 0)
    return std::max(dense, local_t_ws_bytes(p->N, p->support_cols));
  return dense;
}

extern "C" int rbfx_matvec(const rbfx_problem* p, int matrix, int trans, const double* d_x,
                           double* d_y, void* d_work, size_t work_bytes, void* stream) {
  ProbDev P;
  int st = make_prob(p, matrix, P);
  if (st) return st;
  RBFX_REQUIRE(d_x && d_y && d_x != d_y, "matvec needs distinct x and y");
  hipStream_t s = as_stream(stream);
  // A = [[Phi P];[P^T 0]] is symmetric entry for entry when every node is in
  // every support (phi depends on |x_i - x_j| only): A^T x takes the row kernel
  const bool symmetric = matrix == RBFX_MAT_A && !P.local;
  if (!trans || symmetric) {
    const int pid = prof_start("matvec", s);
    const dim3 grid((unsigned)ceil_div(P.n, kMvRows));
    const dim3 grid_loc((unsigned)(ceil_div(P.N, kLocThreads) + P.M));
    RBFX_KIND_SWITCH(P.k, {
      if (P.local && !trans)
        hipLaunchKernelGGL((matvec_local_kernel<KIND_, FAST_>), grid_loc, dim3(kLocThreads), 0, s, P, d_x, d_y);
      else if (P.local)
        hipLaunchKernelGGL((matvec_kernel<KIND_, FAST_, true>), grid, dim3(kMvThreads), 0, s, P, d_x, d_y,
                           (int64_t)0);
      else if (matrix == RBFX_MAT_A) {
        // the value rows, then the rows of P^T
        hipLaunchKernelGGL((matvec_value_kernel<KIND_, FAST_>), dim3((unsigned)ceil_div(P.N, kMvRows)),
                           dim3(kMvThreads), 0, s, P, d_x, d_y);
        if (P.M > 0)
          hipLaunchKernelGGL((matvec_kernel<KIND_, FAST_, false>), dim3((unsigned)ceil_div(P.M, kMvRows)),
                             dim3(kMvThreads), 0, s, P, d_x, d_y, P.N);
      } else
        hipLaunchKernelGGL((matvec_kernel<KIND_, FAST_, false>), grid, dim3(kMvThreads), 0, s, P, d_x, d_y,
                           (int64_t)0);
    });
    prof_stop(pid, s, 0.0, 16.0 * P.n);
    RBFX_CHECK_LAUNCH();
    return RBFX_OK;
  }
  // (the reverse table indexes rows and entries with 32-bit ints)
  if (P.local && d_work && work_bytes >= local_t_ws_bytes(P.N, P.S) && P.N * (2 * P.S + 1) < INT32_MAX) {
    // reverse table in the workspace, then the support-only transposed product
    Workspace ws(d_work, work_bytes);
    int* cnt = ws.take<int>(P.N);
    int* fill = ws.take<int>(P.N);
    int* off = ws.take<int>(P.N + 1);
    int* rows = ws.take<int>((size_t)P.N * (size_t)(2 * P.S > 0 ? 2 * P.S : 1));
    RBFX_REQUIRE(cnt && fill && off && rows, "matvec workspace carve failed");
    const unsigned gz = (unsigned)ceil_div(P.N, 256), gr = (unsigned)ceil_div(P.N, kLocThreads);
    hipLaunchKernelGGL(zero_i32_kernel, dim3(gz), dim3(256), 0, s, cnt, P.N);
    hipLaunchKernelGGL(zero_i32_kernel, dim3(gz), dim3(256), 0, s, fill, P.N);
    hipLaunchKernelGGL(rev_count_kernel, dim3(gr), dim3(kLocThreads), 0, s, P, cnt);
    hipLaunchKernelGGL(rev_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, cnt, P.N, off);
    hipLaunchKernelGGL(rev_fill_kernel, dim3(gr), dim3(kLocThreads), 0, s, P, off, fill, rows);
    RBFX_CHECK_LAUNCH();
    const int pid = prof_start("matvec_t", s);
    const dim3 grid_loc((unsigned)(ceil_div(P.N, kLocThreads) + P.M));
    RBFX_KIND_SWITCH(P.k, {
      hipLaunchKernelGGL((matvec_local_t_kernel<KIND_, FAST_>), grid_loc, dim3(kLocThreads), 0, s, P, off, rows,
                         d_x, d_y);
    });
    prof_stop(pid, s, 0.0, 16.0 * P.n);
    RBFX_CHECK_LAUNCH();
    return RBFX_OK;
  }
  const int64_t chunks = mt_chunks(P.n);
  RBFX_REQUIRE(d_work && work_bytes >= sizeof(double) * (size_t)chunks * (size_t)P.n,
               "matvec workspace too small (%zu < %zu)", work_bytes,
               sizeof(double) * (size_t)chunks * (size_t)P.n);
  const int64_t rpc = ceil_div(P.n, chunks);
  const int64_t used_chunks = ceil_div(P.n, rpc);
  double* part = static_cast<double*>(d_work);
  const dim3 grid((unsigned)ceil_div(P.n, kMtThreads), (unsigned)used_chunks);
  RBFX_KIND_SWITCH(P.k, {
    if (P.local)
      hipLaunchKernelGGL((matvec_t_kernel<KIND_, FAST_, true>), grid, dim3(kMtThreads), 0, s, P, d_x, rpc, part);
    else
      hipLaunchKernelGGL((matvec_t_kernel<KIND_, FAST_, false>), grid, dim3(kMtThreads), 0, s, P, d_x, rpc, part);
  });
  RBFX_CHECK_LAUNCH();
  hipLaunchKernelGGL(sum_chunks_kernel, dim3((unsigned)ceil_div(P.n, 256)), dim3(256), 0, s, part,
                     used_chunks, P.n, d_y);
  RBFX_CHECK_LAUNCH();
  return RBFX_OK;
}
