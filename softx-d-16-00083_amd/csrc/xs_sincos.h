// FP64 sin/cos of 2*pi*r for any double r.
// Reduction by whole and quarter turns is exact (r - rint(r), then t - n/4),
// leaving x = 2*pi*u with |u| <= 1/8, |x| <= pi/4; sin and cos of x come from
// their Taylor series through x^17 and x^18, whose truncation error on that
// interval is below 1e-17.  Coefficients are 1/k! evaluated at compile time.
#pragma once
#include <hip/hip_runtime.h>

namespace xs {

__host__ __device__ constexpr double inv_factorial(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= i;
  return 1.0 / f;
}

// a * b + c with the constant c taken as a scalar (SGPR) operand: the
// series' coefficients then need no vector registers and no copies (the
// compiler otherwise keeps them in VGPRs and copies one before each v_fmac).
__device__ __forceinline__ double fma_k(double a, double b, double c) {
  double r;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
  return r;
}

__device__ __forceinline__ void sincos_2pi(double r, double *s, double *c) {
  constexpr double kTwoPi = 6.283185307179586476925286766559;
  const double t = r - rint(r);            // [-1/2, 1/2]
  const double n = rint(4.0 * t);          // quarter turns, -2..2
  const double x = kTwoPi * (t - 0.25 * n);  // [-pi/4, pi/4]
  const double x2 = x * x;
  // sin x = x - x^3 * (1/3! - x^2/5! + x^4/7! - ... - x^14/17!)
  double ps = fma_k(x2, -inv_factorial(17), inv_factorial(15));
  ps = fma_k(ps, x2, -inv_factorial(13));
  ps = fma_k(ps, x2, inv_factorial(11));
  ps = fma_k(ps, x2, -inv_factorial(9));
  ps = fma_k(ps, x2, inv_factorial(7));
  ps = fma_k(ps, x2, -inv_factorial(5));
  ps = fma_k(ps, x2, inv_factorial(3));
  const double sx = fma(-x * x2, ps, x);
  // cos x = 1 - x^2 * (1/2! - x^2/4! + x^4/6! - ... + x^16/18!)
  double pc = fma_k(x2, inv_factorial(18), -inv_factorial(16));
  pc = fma_k(pc, x2, inv_factorial(14));
  pc = fma_k(pc, x2, -inv_factorial(12));
  pc = fma_k(pc, x2, inv_factorial(10));
  pc = fma_k(pc, x2, -inv_factorial(8));
  pc = fma_k(pc, x2, inv_factorial(6));
  pc = fma_k(pc, x2, -inv_factorial(4));
  pc = fma_k(pc, x2, inv_factorial(2));
  const double cx = fma(-x2, pc, 1.0);
  // sin/cos(x + q pi/2), selected without branches: odd q swaps sin and
  // cos; sin changes sign for q = 2, 3 and cos for q = 1, 2
  const int q = ((int)n) & 3;
  const bool odd = (q & 1) != 0;
  const double ss = odd ? cx : sx, cc = odd ? sx : cx;
  *s = (q & 2) ? -ss : ss;
  *c = ((q ^ (q >> 1)) & 1) ? -cc : cc;
}

}  // namespace xs
