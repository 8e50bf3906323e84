// CPU checks of the batched matrix-core pair-sum plan (csrc/xs_mfma_plan.h),
// built with g++ by tests/test_hist_plan.py.
//
// 1. Every pair type is exactly one column, and every column lies in
//    exactly one group, whose columns are consecutive and inside it.
// 2. A group holds columns of one slice count only, that count covers each
//    column's bound (bound < 2^(7 nb)), and the group's k range covers every
//    column's bins (32 kt >= nhist).
// 3. Groups' count slices (nb x NT x kt x 64 units from b_off) do not
//    overlap and end at b_bytes; launches cover the groups in order.
// 4. Bounds of 2^35 and more are refused.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "xs_mfma_plan.h"

using namespace xs;

static int failures = 0;

#define CHECK(cond, ...)                 \
  do {                                   \
    if (!(cond)) {                       \
      ++failures;                        \
      std::printf("FAIL: " __VA_ARGS__); \
      std::printf("\n");                 \
    }                                    \
  } while (0)

static void check_case(unsigned ntypes, unsigned seed, unsigned nq) {
  std::mt19937_64 rng(seed);
  std::vector<unsigned long long> off(ntypes), bound(ntypes);
  std::vector<uint32_t> nh(ntypes);
  unsigned long long o = 0;
  for (unsigned t = 0; t < ntypes; ++t) {
    nh[t] = 1 + (uint32_t)(rng() % 40000);
    off[t] = o;
    o += nh[t];
    // pair counts spread over every slice count 1..5
    const int bits = (int)(rng() % 35);
    bound[t] = bits == 0 ? 0 : (rng() & ((1ull << bits) - 1)) | (1ull << (bits - 1));
  }
  MfmaBatchPlan p;
  CHECK(mfma_batch_plan(off, nh, bound, nq, &p), "plan refused (seed %u)", seed);
  CHECK(p.cols.size() == ntypes, "columns %zu != types %u", p.cols.size(), ntypes);
  std::vector<int> seen_type(ntypes, 0), seen_col(ntypes, 0);
  for (const MfmaCol &c : p.cols) {
    CHECK(c.type < ntypes, "type %u", c.type);
    if (c.type >= ntypes) continue;
    ++seen_type[c.type];
    CHECK(c.hist_off == off[c.type] && c.nhist == nh[c.type], "column of type %u", c.type);
  }
  for (unsigned t = 0; t < ntypes; ++t) CHECK(seen_type[t] == 1, "type %u seen %d times", t, seen_type[t]);
  unsigned long long next_b = 0;
  uint32_t tile_next = 0, max_nh = 0;
  for (const MfmaLaunch &L : p.launches) {
    CHECK(L.tile0 == tile_next, "launch starts at group %u, expected %u", L.tile0, tile_next);
    CHECK(L.nb >= 1 && L.nb <= (uint32_t)kMfmaMaxSlicesB, "nb %u", L.nb);
    CHECK(L.nt == (L.nb <= 2 ? 2u : 1u), "nt %u for nb %u", L.nt, L.nb);
    for (uint32_t g = L.tile0; g < L.tile0 + L.ntiles; ++g) {
      const MfmaTile &d = p.tiles[g];
      CHECK(d.ncol >= 1 && d.ncol <= 32 * L.nt, "group %u holds %u columns", g, d.ncol);
      CHECK(d.b_off == next_b, "group %u slices at %llu, expected %llu", g, d.b_off, next_b);
      next_b += (unsigned long long)L.nb * L.nt * d.kt * 64;
      for (uint32_t c = d.col0; c < d.col0 + d.ncol && c < ntypes; ++c) {
        ++seen_col[c];
        const MfmaCol &col = p.cols[c];
        CHECK((bound[col.type] >> (7 * L.nb)) == 0ull, "column %u bound beyond %u slices", c, L.nb);
        // the fewest slices that cover the bound (grouping never takes more)
        CHECK(L.nb == 1 || (bound[col.type] >> (7 * (L.nb - 1))) != 0ull,
              "column %u takes %u slices, fewer cover it", c, L.nb);
        CHECK(32ull * d.kt >= col.nhist, "column %u: %u bins beyond k range %u", c, col.nhist, d.kt);
        max_nh = std::max(max_nh, col.nhist);
      }
    }
    tile_next += L.ntiles;
  }
  CHECK(tile_next == p.tiles.size(), "launches cover %u of %zu groups", tile_next, p.tiles.size());
  for (unsigned c = 0; c < ntypes; ++c) CHECK(seen_col[c] == 1, "column %u in %d groups", c, seen_col[c]);
  CHECK(p.b_bytes == next_b * 16, "b_bytes %zu, slices end at %llu", p.b_bytes, next_b * 16);
  CHECK(32ull * p.kt_max >= max_nh, "kt_max %u below %u bins", p.kt_max, max_nh);
  const unsigned MT = (nq + 127) / 128 * 4;
  CHECK(p.a_bytes == (size_t)kMfmaSlicesA * MT * p.kt_max * 1024, "a_bytes %zu", p.a_bytes);
  std::printf("case types=%u seed=%u groups=%zu launches=%zu b_bytes=%zu\n", ntypes, seed,
              p.tiles.size(), p.launches.size(), p.b_bytes);
}

int main() {
  check_case(1, 1, 1);
  check_case(31, 2, 100);
  check_case(65, 3, 2048);
  for (unsigned s = 4; s < 24; ++s) check_case(200 + 97 * s, s, 2048);
  {
    MfmaBatchPlan p;
    const std::vector<unsigned long long> off{0, 10};
    const std::vector<uint32_t> nh{10, 10};
    CHECK(!mfma_batch_plan(off, nh, {5ull, 1ull << 35}, 64, &p), "bound 2^35 accepted");
    CHECK(mfma_batch_plan(off, nh, {5ull, (1ull << 35) - 1}, 64, &p), "bound 2^35 - 1 refused");
  }
  if (failures) {
    std::printf("%d failures\n", failures);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
