// CPU check of the fixed-point floor of the FP64 histogram kernel
// (softx-d-16-00083_amd/csrc/xs_fixbin.h), the arithmetic the kernel runs
// after v_rsq_f64.  Reciprocal square roots are formed in extended precision
// and perturbed to the documented worst accuracy of v_rsq_f64 (2^-23
// relative) and inside it; distances are drawn uniformly and at or within a
// few ulps of bin edges.  Wherever the low word passes the test, the high
// word must give the reference expression (unsigned)(sqrt(s) / bin) in IEEE
// double.  Two unsound variants (no margin; the low word tested against e
// instead of 2e) must fail on the same samples, so the check tells a sound
// band from an unsound one.
#include <cmath>
#include <cstdio>
#include <random>

#include "xs_fixbin.h"
#include "xs_hist_plan.h"

using namespace xs::hist;

static double rsq_with_error(double s, long double e) {
  return (double)((1.0L + e) / sqrtl((long double)s));
}

struct Tally {
  unsigned long long fast = 0, exact = 0, wrong = 0;
};

static void one(double s, double bin, double y, const FixBand &fb, Tally &t) {
  const uint32_t want = (uint32_t)(std::sqrt(s) / bin);
  uint32_t lo;
  const uint32_t hi = fix_from_rsq(s, y, 1.0 / bin, fb.m, lo);
  if (lo > fb.lm1) {
    if (hi - kFixHi == want)
      ++t.fast;
    else
      ++t.wrong;
  } else {
    ++t.exact;
  }
}

// Samples of one grid; uniform fractions in u, edge samples in edge.
static void grid(uint32_t nhist, double bin, const FixBand &fb, std::mt19937_64 &rng, Tally &u,
                 Tally &edge) {
  std::uniform_int_distribution<uint32_t> kd(0, nhist - 1);
  std::uniform_real_distribution<double> fd(0.0, 1.0);
  std::uniform_real_distribution<long double> ed(-0x1p-23L, 0x1p-23L);
  const long double errs[4] = {-0x1p-23L, 0x1p-23L, -0x1p-23L * (1 - 0x1p-10L), 0};
  for (int i = 0; i < 200000; ++i) {
    const double t = kd(rng) + fd(rng);
    const double s = (t * bin) * (t * bin);
    one(s, bin, rsq_with_error(s, ed(rng)), fb, u);
  }
  for (int i = 0; i < 20000; ++i) {
    const uint32_t k = kd(rng) + (i & 1);
    for (int p = 20; p <= 52; p += 4)
      for (int sg = -1; sg <= 1; ++sg) {
        const double t = k + sg * std::ldexp(1.0, -p) * (1 + (i % 3));
        double s = (t * bin) * (t * bin);
        for (int j = 0; j < 3; ++j, s = std::nextafter(s, INFINITY))
          for (long double e : errs) one(s, bin, rsq_with_error(s, e), fb, edge);
      }
  }
}

int main() {
  std::mt19937_64 rng(20261203);
  const uint32_t grids[4] = {64, 6357, 97251, kNarrowMax};
  const double bins[4] = {0.005, 0.0173, 0.1, 1.0 / 3.0};
  int bad = 0;
  for (uint32_t nh : grids)
    for (double bin : bins) {
      const FixBand fb = fix_band(band_for(nh));
      Tally u, edge;
      grid(nh, bin, fb, rng, u, edge);
      std::printf("nhist %u bin %.4g: uniform fast %llu exact %llu wrong %llu; edges fast %llu exact %llu wrong %llu\n",
                  nh, bin, u.fast, u.exact, u.wrong, edge.fast, edge.exact, edge.wrong);
      if (u.wrong || edge.wrong) bad = 1;
      // the fast path takes nearly every uniform sample, so the check is not vacuous
      if (u.exact * 1000 > u.fast) bad = 1;
      if (edge.exact == 0) bad = 1;
      // two unsound variants must be caught where the estimate's error
      // exceeds 2^-32 bins: no margin (e = 1, only T's own rounding
      // covered), and the low word tested against e instead of 2e
      if (nh >= 97251) {
        const FixBand none{0x1p20 + 0x1p-32, 1u};
        const FixBand half{fb.m, (fb.lm1 + 1) / 2 - 1};
        Tally u0, e0, u1, e1;
        grid(nh, bin, none, rng, u0, e0);
        grid(nh, bin, half, rng, u1, e1);
        if (e0.wrong == 0 || e1.wrong == 0) {
          std::printf("an unsound band passed nhist %u bin %g (%llu, %llu wrong)\n", nh, bin,
                      e0.wrong, e1.wrong);
          bad = 1;
        }
      }
    }
  // zero, negative (rounding) and not-finite squared distances: the low word
  // is zero, so the exact evaluation decides
  const FixBand fb = fix_band(band_for(97251));
  const double odd[4] = {0.0, -0x1p-60, INFINITY, NAN};
  for (double s : odd) {
    uint32_t lo;
    fix_from_rsq(s, 1.0 / std::sqrt(s), 200.0, fb.m, lo);
    if (lo > fb.lm1) {
      std::printf("s = %g passed the fast test\n", s);
      bad = 1;
    }
  }
  // a band too wide for the fixed-point form sends every pair to the exact path
  const FixBand wide = fix_band(0.2);
  if (wide.lm1 != 0xffffffffu) bad = 1;
  std::printf(bad ? "FAIL\n" : "OK\n");
  return bad;
}
