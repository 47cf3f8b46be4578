// Device-side closed forms of the RBF kernels, their derivatives and the
// monomial basis (FP64).  Reference semantics:
//   updes/utils.py:30-69     kernels (polyharmonic r^(2a+1), gaussian
//                            exp(-(eps r)^2), thin plate r^(2a) log r with 0 at
//                            r = 0, (inverse) multiquadric)
//   updes/utils.py:92-134    monomials id 0..14
//   updes/operators.py:42-111 autodiff gradient / Laplacian / div-grad, whose
//                            NaN at r = 0 is replaced by 0 (operators.py:58,83,109)
// For a radial kernel phi(r) and d = x - c:
//   f1 = phi'(r)/r,  g2 = phi'' - phi'/r,
//   grad = f1 d,  Hess = (g2/r^2) d d^T + f1 I,  Lap = g2 + 2 f1.
// At r == 0: value phi(0) (1 for gaussian / multiquadrics, 0 otherwise) and
// every derivative 0.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rbfx.h"

namespace rbfx {

struct Kern {
  int kind;
  double p;    // a or eps
  double m;    // polyharmonic exponent 2a+1 / thin-plate 2a
  double e2;   // eps^2
  double phi0; // value at r = 0
  bool fast;   // a == 1 closed forms without pow
};

__host__ __device__ inline Kern make_kern(int kind, double param) {
  Kern k;
  k.kind = kind;
  k.p = param;
  k.e2 = param * param;
  k.m = 0.0;
  k.phi0 = 0.0;
  k.fast = false;
  if (kind == RBFX_POLYHARMONIC) {
    k.m = 2.0 * param + 1.0;
    k.fast = (param == 1.0);
  } else if (kind == RBFX_THIN_PLATE) {
    k.m = 2.0 * param;
    k.fast = (param == 1.0);
  } else {
    k.phi0 = 1.0;  // gaussian, multiquadric, inverse multiquadric
  }
  return k;
}

// Kernels are instantiated per (kind, a == 1): pinning the kind in the local
// copy of Kern turns every switch below into straight-line code.
template <int KIND, bool FAST>
__device__ inline void pin_kind(Kern& k) {
  k.kind = KIND;
  k.fast = FAST;
}

#define RBFX_KIND_SWITCH(k, ...)                                        \
  do {                                                                  \
    switch ((k).kind) {                                                 \
      case RBFX_POLYHARMONIC:                                           \
        if ((k).fast) {                                                 \
          constexpr int KIND_ = RBFX_POLYHARMONIC;                      \
          constexpr bool FAST_ = true;                                  \
          __VA_ARGS__;                                                  \
        } else {                                                        \
          constexpr int KIND_ = RBFX_POLYHARMONIC;                      \
          constexpr bool FAST_ = false;                                 \
          __VA_ARGS__;                                                  \
        }                                                               \
        break;                                                          \
      case RBFX_GAUSSIAN: {                                             \
        constexpr int KIND_ = RBFX_GAUSSIAN;                            \
        constexpr bool FAST_ = false;                                   \
        __VA_ARGS__;                                                    \
      } break;                                                          \
      case RBFX_THIN_PLATE:                                             \
        if ((k).fast) {                                                 \
          constexpr int KIND_ = RBFX_THIN_PLATE;                        \
          constexpr bool FAST_ = true;                                  \
          __VA_ARGS__;                                                  \
        } else {                                                        \
          constexpr int KIND_ = RBFX_THIN_PLATE;                        \
          constexpr bool FAST_ = false;                                 \
          __VA_ARGS__;                                                  \
        }                                                               \
        break;                                                          \
      case RBFX_MULTIQUADRIC: {                                         \
        constexpr int KIND_ = RBFX_MULTIQUADRIC;                        \
        constexpr bool FAST_ = false;                                   \
        __VA_ARGS__;                                                    \
      } break;                                                          \
      default: {                                                        \
        constexpr int KIND_ = RBFX_INV_MULTIQUADRIC;                    \
        constexpr bool FAST_ = false;                                   \
        __VA_ARGS__;                                                    \
      } break;                                                          \
    }                                                                   \
  } while (0)

// phi, f1 = phi'/r and g2 = phi'' - phi'/r (= r^2 f2) for r > 0.  Using g2
// keeps the Laplacian division-free: Lap phi = g2 + 2 f1.  Every sum is an
// explicit FMA or add, so each kernel that forms an entry gets the same bits
// (the assembly, the matrix-free products and the residuals must agree).
__device__ inline void radial(const Kern& k, double r, double& phi, double& f1, double& g2) {
  switch (k.kind) {
    case RBFX_POLYHARMONIC:
      if (k.fast) {
        phi = r * r * r;
        f1 = 3.0 * r;
        g2 = 3.0 * r;
      } else {
        const double rm2 = pow(r, k.m - 2.0);
        phi = rm2 * r * r;
        f1 = k.m * rm2;
        g2 = k.m * (k.m - 2.0) * rm2;
      }
      break;
    case RBFX_GAUSSIAN: {
      const double r2 = r * r;
      const double g = exp(-k.e2 * r2);
      phi = g;
      f1 = -2.0 * k.e2 * g;
      g2 = 4.0 * k.e2 * k.e2 * r2 * g;
      break;
    }
    case RBFX_THIN_PLATE: {
      const double lr = log(r);
      if (k.fast) {
        phi = r * r * lr;
        f1 = __builtin_fma(2.0, lr, 1.0);
        g2 = 2.0;
      } else {
        const double a = k.p;
        const double rm2 = pow(r, k.m - 2.0);
        phi = rm2 * r * r * lr;
        f1 = rm2 * __builtin_fma(2.0 * a, lr, 1.0);
        g2 = rm2 * __builtin_fma(4.0 * a * a - 4.0 * a, lr, 4.0 * a - 2.0);
      }
      break;
    }
    case RBFX_MULTIQUADRIC: {
      const double s = sqrt(__builtin_fma(k.e2 * r, r, 1.0));
      phi = s;
      f1 = k.e2 / s;
      g2 = -k.e2 * k.e2 * r * r / (s * s * s);
      break;
    }
    default: {  // RBFX_INV_MULTIQUADRIC
      const double s = 1.0 / sqrt(__builtin_fma(k.e2 * r, r, 1.0));
      phi = s;
      f1 = -k.e2 * s * s * s;
      g2 = 3.0 * k.e2 * k.e2 * r * r * s * s * s * s * s;
      break;
    }
  }
}

// Squared distance, the same bits in every kernel.
__device__ inline double dist2(double dx, double dy) { return __builtin_fma(dx, dx, dy * dy); }

__device__ inline double radial_value(const Kern& k, double r) {
  if (r == 0.0) return k.phi0;
  switch (k.kind) {
    case RBFX_POLYHARMONIC:
      return k.fast ? r * r * r : pow(r, k.m);
    case RBFX_GAUSSIAN:
      return exp(-k.e2 * r * r);
    case RBFX_THIN_PLATE:
      return k.fast ? r * r * log(r) : pow(r, k.m) * log(r);
    case RBFX_MULTIQUADRIC:
      return sqrt(__builtin_fma(k.e2 * r, r, 1.0));
    default:
      return 1.0 / sqrt(__builtin_fma(k.e2 * r, r, 1.0));
  }
}

// f1 only (gradient factor); 0 at r == 0.
__device__ inline double radial_f1(const Kern& k, double r) {
  if (r == 0.0) return 0.0;
  switch (k.kind) {
    case RBFX_POLYHARMONIC:
      return k.fast ? 3.0 * r : k.m * pow(r, k.m - 2.0);
    case RBFX_GAUSSIAN:
      return -2.0 * k.e2 * exp(-k.e2 * r * r);
    case RBFX_THIN_PLATE:
      return k.fast ? __builtin_fma(2.0, log(r), 1.0) : pow(r, k.m - 2.0) * __builtin_fma(2.0 * k.p, log(r), 1.0);
    case RBFX_MULTIQUADRIC:
      return k.e2 / sqrt(__builtin_fma(k.e2 * r, r, 1.0));
    default: {
      const double s = 1.0 / sqrt(__builtin_fma(k.e2 * r, r, 1.0));
      return -k.e2 * s * s * s;
    }
  }
}

// Linear combination of the operator basis applied to phi(., c) at x:
//   v phi + dx dphi/dx + dy dphi/dy + lap Lap phi + dxx/dyy/dxy second derivatives.
struct OpCoef {
  double v, dx, dy, lap, dxx, dyy, dxy;
  bool hess;  // any of dxx/dyy/dxy non-zero
};

// apply_op for r2 = dist2(dx, dy) > 0 and r = sqrt(r2), explicit FMAs.
// HESS: the second-derivative terms (c.hess), a compile-time choice for
// callers that branch once per row.
template <bool HESS>
__device__ inline double apply_op_pos_t(const Kern& k, const OpCoef& c, double dx, double dy, double r2, double r) {
  double phi, f1, g2;
  radial(k, r, phi, f1, g2);
  const double a1 = __builtin_fma(c.dx, dx, __builtin_fma(c.dy, dy, __builtin_fma(2.0, c.lap, c.dxx + c.dyy)));
  double hv = c.lap;
  if (HESS) hv += __builtin_fma(c.dxx * dx, dx, __builtin_fma(c.dyy * dy, dy, c.dxy * dx * dy)) / r2;
  return __builtin_fma(c.v, phi, __builtin_fma(f1, a1, g2 * hv));
}
__device__ inline double apply_op_pos(const Kern& k, const OpCoef& c, double dx, double dy, double r2, double r) {
  return c.hess ? apply_op_pos_t<true>(k, c, dx, dy, r2, r) : apply_op_pos_t<false>(k, c, dx, dy, r2, r);
}

__device__ inline double apply_op(const Kern& k, const OpCoef& c, double dx, double dy) {
  const double r2 = dist2(dx, dy);
  const double r = sqrt(r2);
  if (r == 0.0) return c.v * k.phi0;
  return apply_op_pos(k, c, dx, dy, r2, r);
}

// All seven operator values of phi at d (evaluation kernels).
__device__ inline void all_ops(const Kern& k, double dx, double dy, double out[RBFX_NOPS]) {
  const double r2 = dist2(dx, dy);
  const double r = sqrt(r2);
  if (r == 0.0) {
    out[0] = k.phi0;
#pragma unroll
    for (int t = 1; t < RBFX_NOPS; ++t) out[t] = 0.0;
    return;
  }
  double phi, f1, g2;
  radial(k, r, phi, f1, g2);
  const double f2 = g2 / r2;
  out[RBFX_VAL] = phi;
  out[RBFX_DX] = f1 * dx;
  out[RBFX_DY] = f1 * dy;
  out[RBFX_LAP] = g2 + 2.0 * f1;
  out[RBFX_DXX] = f2 * dx * dx + f1;
  out[RBFX_DYY] = f2 * dy * dy + f1;
  out[RBFX_DXY] = f2 * dx * dy;
}

// Monomial exponents of ids 0..14 (updes/utils.py:103-132).
__host__ __device__ inline void mono_exp(int id, int& a, int& b) {
  // degree d block starts at d(d+1)/2; inside the block x^(d-s) y^s
  int d = 0;
  while ((d + 1) * (d + 2) / 2 <= id) ++d;
  const int s = id - d * (d + 1) / 2;
  a = d - s;
  b = s;
}

__device__ inline double ipow(double v, int e) {
  double r = 1.0;
  for (int i = 0; i < e; ++i) r *= v;
  return r;
}

// All operator values of monomial id at (x, y).
__device__ inline void mono_ops(int id, double x, double y, double out[RBFX_NOPS]) {
  int a, b;
  mono_exp(id, a, b);
  const double xa = ipow(x, a), yb = ipow(y, b);
  const double xa1 = a >= 1 ? ipow(x, a - 1) : 0.0;
  const double yb1 = b >= 1 ? ipow(y, b - 1) : 0.0;
  const double xa2 = a >= 2 ? ipow(x, a - 2) : 0.0;
  const double yb2 = b >= 2 ? ipow(y, b - 2) : 0.0;
  out[RBFX_VAL] = xa * yb;
  out[RBFX_DX] = a * xa1 * yb;
  out[RBFX_DY] = b * xa * yb1;
  out[RBFX_DXX] = a * (a - 1) * xa2 * yb;
  out[RBFX_DYY] = b * (b - 1) * xa * yb2;
  out[RBFX_DXY] = a * b * xa1 * yb1;
  out[RBFX_LAP] = out[RBFX_DXX] + out[RBFX_DYY];
}

__device__ inline double mono_apply(int id, double x, double y, const OpCoef& c) {
  double o[RBFX_NOPS];
  mono_ops(id, x, y, o);
  return c.v * o[RBFX_VAL] + c.dx * o[RBFX_DX] + c.dy * o[RBFX_DY] + c.lap * o[RBFX_LAP] +
         c.dxx * o[RBFX_DXX] + c.dyy * o[RBFX_DYY] + c.dxy * o[RBFX_DXY];
}

}  // namespace rbfx
