// Fixed-point floor of the FP64 histogram kernel's distance estimate.  Plain
// C++ as well as device code, so that CPU tests drive the same arithmetic
// with reciprocal square roots of the worst documented accuracy.
//
// The fast path estimates sqrt(s)/bin from y ~ 1/sqrt(s) (v_rsq_f64, within
// 2^-23 relative) and one Newton step, g = s y, t = g y,
// estimate = g (1.5 - 0.5 t) / bin.  The band test of the kernel bounds the
// estimate's distance from the reference expression's IEEE quotient
// sqrt(s)/bin by E bins (E = 1/2 - band).
//
// With an integer e, e 2^-32 >= E + 2^-32, the last FMA of the estimate adds
// m = 2^20 + e 2^-32: T = RN(estimate + m) lies in [2^20, 2^21), where the
// ulp of a double is 2^-32, so T's high word is kFixHi + K with
// K = floor(estimate + e 2^-32) and its low word L the fraction in units of
// 2^-32.  If L >= 2e, the IEEE quotient t lies in (K, K + 1):
//   t >= K + (L - e) 2^-32 - 2^-33 - E >= K + e 2^-32 - 2^-33 - E > K,
//   t <= K + (L - e) 2^-32 + 2^-33 + E <= K + 1 - e 2^-32 + 2^-33 + E < K + 1
// (2^-33 bounds T's own rounding), and K is the bin.  One unsigned compare
// of L replaces fract, the band test and the conversion, and the high word
// addresses a histogram slot directly (kFixHi folds into the slot offset).
// When s is zero, negative (rounding), subnormal or not finite, the estimate
// is NaN or infinite, whose low word is zero, so the exact evaluation takes
// it.  T stays below 2^21 for bin grids of at most kNarrowMax bins, since an
// estimate exceeds the bin count by no more than its error bound.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define XS_FIX_HD __host__ __device__
#else
#define XS_FIX_HD
#endif

namespace xs {
namespace hist {

constexpr uint32_t kNarrowMax = 1u << 19;  // bins of a grid the fixed-point path takes
constexpr uint32_t kFixHi = 0x41300000u;   // high word of the double 2^20

struct FixBand {
  double m;      // 2^20 + e 2^-32
  uint32_t lm1;  // 2e - 1: the estimate is floored directly when L > lm1
};

// From the half-width of the band test: E = 1/2 - band is exact for band in
// [1/4, 1/2] (Sterbenz), and a wider error bound sends every pair to the
// exact evaluation (L > 0xffffffff never holds).
XS_FIX_HD inline FixBand fix_band(double band) {
  const double E = 0.5 - band;
  if (!(E >= 0.0 && E <= 0.25)) return FixBand{0x1p20, 0xffffffffu};
  const uint32_t e = (uint32_t)__builtin_ceil(E * 0x1p32) + 1u;
  return FixBand{0x1p20 + (double)e * 0x1p-32, 2u * e - 1u};
}

// T from s and y ~ 1/sqrt(s): returns T's high word (kFixHi + K) and stores
// its low word L in lo.
XS_FIX_HD inline uint32_t fix_from_rsq(double s, double y, double inv_bin, double m,
                                       uint32_t &lo) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  const double g = s * y;
  const double t = g * y;
  const double T = __builtin_fma(g, __builtin_fma(t, -0.5 * inv_bin, 1.5 * inv_bin), m);
  uint64_t w;
  __builtin_memcpy(&w, &T, sizeof w);
  lo = (uint32_t)w;
  return (uint32_t)(w >> 32);
}

}  // namespace hist
}  // namespace xs
