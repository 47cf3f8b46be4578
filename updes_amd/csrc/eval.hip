// K6: evaluation of a field given by its RBF + monomial coefficients, and its
// transpose.  Reference: updes/operators.py:118-150 (value), 156-187
// (gradient), 319-351 (laplacian): a sum over ALL N centres (the evaluation
// point itself included, phi(0) for the value and 0 for derivatives) plus the
// monomial tail.  One block owns kEvPts points, its threads stride over the
// centres, and a block reduction produces every requested operator at once.
#include "common.h"
#include "rbf_device.h"

namespace rbfx {

constexpr int kEvThreads = 256;
constexpr int kEvPts = 4;

__global__ __launch_bounds__(kEvThreads) void eval_kernel(Kern k, int64_t N, int64_t M,
                                                          const double* __restrict__ cx,
                                                          const double* __restrict__ cy,
                                                          const double* __restrict__ coeffs,
                                                          int64_t npts, const double* __restrict__ px,
                                                          const double* __restrict__ py, uint32_t mask,
                                                          double* __restrict__ out) {
  const int64_t p0 = (int64_t)blockIdx.x * kEvPts;
  double xp[kEvPts], yp[kEvPts];
  double acc[kEvPts][RBFX_NOPS];
#pragma unroll
  for (int q = 0; q < kEvPts; ++q) {
    const int64_t p = p0 + q < npts ? p0 + q : npts - 1;
    xp[q] = px[p];
    yp[q] = py[p];
#pragma unroll
    for (int t = 0; t < RBFX_NOPS; ++t) acc[q][t] = 0.0;
  }
  const bool only_val = mask == (1u << RBFX_VAL);
  for (int64_t j = threadIdx.x; j < N; j += kEvThreads) {
    const double lam = coeffs[j];
    const double xj = cx[j], yj = cy[j];
#pragma unroll
    for (int q = 0; q < kEvPts; ++q) {
      const double dx = xp[q] - xj, dy = yp[q] - yj;
      if (only_val) {
        acc[q][RBFX_VAL] += lam * radial_value(k, sqrt(dist2(dx, dy)));
      } else {
        double o[RBFX_NOPS];
        all_ops(k, dx, dy, o);
#pragma unroll
        for (int t = 0; t < RBFX_NOPS; ++t) acc[q][t] += lam * o[t];
      }
    }
  }
  __shared__ double red[kEvPts][RBFX_NOPS][kEvThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kEvPts; ++q) {
#pragma unroll
    for (int t = 0; t < RBFX_NOPS; ++t) {
      double v = acc[q][t];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (lane == 0) red[q][t][wid] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < kEvPts * RBFX_NOPS) {
    const int q = threadIdx.x / RBFX_NOPS, t = threadIdx.x % RBFX_NOPS;
    const int64_t p = p0 + q;
    if (p < npts && (mask & (1u << t))) {
      double s = 0.0;
      for (int w = 0; w < kEvThreads / 64; ++w) s += red[q][t][w];
      for (int64_t m = 0; m < M; ++m) {
        double o[RBFX_NOPS];
        mono_ops((int)m, xp[q], yp[q], o);
        s += coeffs[N + m] * o[t];
      }
      out[(int64_t)t * npts + p] = s;
    }
  }
}

__device__ inline double op_value(const Kern& k, int op, double dx, double dy) {
  if (op == RBFX_VAL) return radial_value(k, sqrt(dist2(dx, dy)));
  double o[RBFX_NOPS];
  all_ops(k, dx, dy, o);
  return o[op];
}

// out[j] = sum_i w_i op phi(p_i, c_j): 64 centres per block x 4 point chunks.
__global__ __launch_bounds__(256) void eval_t_kernel(Kern k, int64_t N, const double* __restrict__ cx,
                                                     const double* __restrict__ cy, int64_t npts,
                                                     const double* __restrict__ px,
                                                     const double* __restrict__ py, int op,
                                                     const double* __restrict__ w,
                                                     double* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int chunk = threadIdx.x >> 6;
  const bool valid = j < N;
  const double xj = valid ? cx[j] : 0.0, yj = valid ? cy[j] : 0.0;
  const int64_t per = (npts + 3) / 4;
  const int64_t s = chunk * per, e = s + per < npts ? s + per : npts;
  double acc = 0.0;
  for (int64_t i = s; i < e; ++i) acc += w[i] * op_value(k, op, px[i] - xj, py[i] - yj);
  __shared__ double red[4][64];
  red[chunk][threadIdx.x & 63] = acc;
  __syncthreads();
  if (chunk == 0 && valid)
    out[j] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// out[N+m] = sum_i w_i op p_m(p_i): one block per monomial.
__global__ __launch_bounds__(256) void eval_t_mono_kernel(int64_t N, int64_t npts,
                                                          const double* __restrict__ px,
                                                          const double* __restrict__ py, int op,
                                                          const double* __restrict__ w,
                                                          double* __restrict__ out) {
  const int m = blockIdx.x;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < npts; i += 256) {
    double o[RBFX_NOPS];
    mono_ops(m, px[i], py[i], o);
    acc += w[i] * o[op];
  }
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[N + m] = red[0] + red[1] + red[2] + red[3];
}

// q's rows without the terms: interior constant (+ row array), boundary
// values, zeros for the monomial rows.
__global__ __launch_bounds__(256) void rhs_init_kernel(int64_t N, int64_t M, int64_t Ni, double c0,
                                                      const double* __restrict__ crow,
                                                      const double* __restrict__ bc, double* __restrict__ q) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= N + M) return;
  double v = 0.0;
  if (i < Ni)
    v = c0 + (crow ? crow[i] : 0.0);
  else if (i < N)
    v = bc[i - Ni];
  q[i] = v;
}

// q[i] += w * row[i] * vals[i] over the interior rows.
__global__ __launch_bounds__(256) void rhs_accum_kernel(int64_t Ni, double w, const double* __restrict__ row,
                                                       const double* __restrict__ vals, double* __restrict__ q) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < Ni) q[i] += w * (row ? row[i] : 1.0) * vals[i];
}

}  // namespace rbfx

using namespace rbfx;

extern "C" int rbfx_eval(int kernel, double param, int64_t N, int64_t M, const double* d_cx,
                         const double* d_cy, const double* d_coeffs, int64_t npts,
                         const double* d_px, const double* d_py, uint32_t ops_mask, double* d_out,
                         void* stream) {
  RBFX_REQUIRE(kernel >= 0 && kernel <= 4, "unknown kernel %d", kernel);
  RBFX_REQUIRE(N > 0 && M >= 0 && M <= 15, "bad N / M");
  RBFX_REQUIRE(npts >= 0, "npts < 0");
  RBFX_REQUIRE(ops_mask != 0 && ops_mask < (1u << RBFX_NOPS), "bad ops_mask 0x%x", ops_mask);
  RBFX_REQUIRE(d_cx && d_cy && d_coeffs && d_out && (npts == 0 || (d_px && d_py)), "null pointer");
  if (npts == 0) return RBFX_OK;
  const Kern k = make_kern(kernel, param);
  hipLaunchKernelGGL(eval_kernel, dim3((unsigned)ceil_div(npts, kEvPts)), dim3(kEvThreads), 0,
                     as_stream(stream), k, N, M, d_cx, d_cy, d_coeffs, npts, d_px, d_py, ops_mask,
                     d_out);
  RBFX_CHECK_LAUNCH();
  return RBFX_OK;
}

extern "C" int rbfx_eval_t(int kernel, double param, int64_t N, int64_t M, const double* d_cx,
                           const double* d_cy, int64_t npts, const double* d_px,
                           const double* d_py, int op, const double* d_w, double* d_out,
                           void* stream) {
  RBFX_REQUIRE(kernel >= 0 && kernel <= 4, "unknown kernel %d", kernel);
  RBFX_REQUIRE(N > 0 && M >= 0 && M <= 15, "bad N / M");
  RBFX_REQUIRE(op >= 0 && op < RBFX_NOPS, "bad op %d", op);
  RBFX_REQUIRE(npts >= 0, "npts < 0");
  RBFX_REQUIRE(d_cx && d_cy && d_w && d_out && (npts == 0 || (d_px && d_py)), "null pointer");
  const Kern k = make_kern(kernel, param);
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(eval_t_kernel, dim3((unsigned)ceil_div(N, 64)), dim3(256), 0, s, k, N, d_cx,
                     d_cy, npts, d_px, d_py, op, d_w, d_out);
  RBFX_CHECK_LAUNCH();
  if (M > 0) {
    hipLaunchKernelGGL(eval_t_mono_kernel, dim3((unsigned)M), dim3(256), 0, s, N, npts, d_px, d_py,
                       op, d_w, d_out);
    RBFX_CHECK_LAUNCH();
  }
  return RBFX_OK;
}

extern "C" size_t rbfx_rhs_workspace(const rbfx_problem* p) {
  if (!p || p->Ni < 0) return 0;
  return ws_round(sizeof(double) * (size_t)RBFX_NOPS * (size_t)p->Ni);
}

extern "C" int rbfx_rhs(const rbfx_problem* p, double constant, const double* d_const_row, int nterms,
                        const rbfx_rhs_term* terms, const double* d_bc, double* d_q, void* d_work,
                        size_t work_bytes, void* stream) {
  RBFX_REQUIRE(p != nullptr && d_q != nullptr, "null pointer");
  RBFX_REQUIRE(p->N > 0 && p->M >= 0 && p->M <= 15, "bad N / M");
  RBFX_REQUIRE(p->kernel >= 0 && p->kernel <= 4, "unknown kernel %d", p->kernel);
  RBFX_REQUIRE(p->Ni >= 0 && p->Nd >= 0 && p->Nn >= 0 && p->Nr >= 0 && p->Npd >= 0, "negative row-class size");
  RBFX_REQUIRE(p->Ni + p->Nd + p->Nn + p->Nr + 2 * p->Npd == p->N, "row classes do not add up to N");
  RBFX_REQUIRE(d_bc || p->N == p->Ni, "boundary values missing");
  RBFX_REQUIRE(nterms >= 0 && (nterms == 0 || terms), "bad terms");
  hipStream_t s = as_stream(stream);
  const int64_t N = p->N, M = p->M, Ni = p->Ni;
  hipLaunchKernelGGL(rhs_init_kernel, dim3((unsigned)ceil_div(N + M, 256)), dim3(256), 0, s, N, M, Ni, constant,
                     d_const_row, d_bc, d_q);
  RBFX_CHECK_LAUNCH();
  if (nterms == 0 || Ni == 0) return RBFX_OK;
  RBFX_REQUIRE(p->d_x && p->d_y, "node coordinates missing");
  Workspace ws(d_work, work_bytes);
  double* vals = ws.take<double>((size_t)RBFX_NOPS * (size_t)Ni);
  RBFX_REQUIRE(vals != nullptr, "rhs workspace too small");
  const Kern k = make_kern(p->kernel, p->param);
  for (int t = 0; t < nterms; ++t) {
    const rbfx_rhs_term& tm = terms[t];
    RBFX_REQUIRE(tm.op >= 0 && tm.op < RBFX_NOPS, "term %d: unknown operator %d", t, tm.op);
    RBFX_REQUIRE(tm.d_coeffs != nullptr, "term %d: coefficients missing", t);
    // the interior nodes are the first Ni nodes: evaluate op at them
    hipLaunchKernelGGL(eval_kernel, dim3((unsigned)ceil_div(Ni, kEvPts)), dim3(kEvThreads), 0, s, k, N, M, p->d_x,
                       p->d_y, tm.d_coeffs, Ni, p->d_x, p->d_y, 1u << tm.op, vals);
    RBFX_CHECK_LAUNCH();
    hipLaunchKernelGGL(rhs_accum_kernel, dim3((unsigned)ceil_div(Ni, 256)), dim3(256), 0, s, Ni, tm.weight,
                       tm.d_weight_row, vals + (size_t)tm.op * Ni, d_q);
    RBFX_CHECK_LAUNCH();
  }
  return RBFX_OK;
}
