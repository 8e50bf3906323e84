// CPU checks of the histogram work-item tables (csrc/xs_hist_plan.h), built
// with g++ by tests/test_hist_plan.py.
//
// 1. Bound: every work item holds fewer than 2^32 pairs (I unit atoms x J run
//    atoms with full tiles of kTile atoms), because the kernels count a bin
//    in a uint32 LDS slot that may be flushed only at the end of the item.
// 2. Coverage: the items of a pair type visit every tile pair exactly once,
//    following the kernels' rules for the item's own I unit (FP64 binning:
//    J tile 2u pairs with I tile 2u only; J tile 2u + 1 with I tile 2u and
//    itself; FP32 binning: J tile u is the diagonal tile).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "xs_hist_plan.h"

using namespace xs::hist;

static int failures = 0;

#define CHECK(cond, ...)                 \
  do {                                   \
    if (!(cond)) {                       \
      ++failures;                        \
      std::printf("FAIL: " __VA_ARGS__); \
      std::printf("\n");                 \
    }                                    \
  } while (0)

static std::vector<TypeDesc> one_type(uint32_t a_n, uint32_t b_n, bool same) {
  TypeDesc d{};
  d.nhist = 1;
  d.a_base = 0;
  d.a_n = a_n;
  d.b_base = same ? 0 : a_n;
  d.b_n = same ? a_n : b_n;
  d.same = same;
  return {d};
}

// Bound on items of one pair type with full tiles.
static void check_bound(uint32_t a_n, uint32_t b_n, bool same, bool paired) {
  const auto types = one_type(a_n, b_n, same);
  std::vector<Item> items;
  add_items(types, 0, 1, items, paired);
  CHECK(!items.empty(), "no items (a_n %u)", a_n);
  uint64_t worst = 0;
  for (const Item &it : items) {
    CHECK(it.jcount >= 1 && it.jcount <= max_jrun(paired), "jcount %u", it.jcount);
    const uint64_t iatoms = (uint64_t)(paired ? 2 : 1) * kTile;
    const uint64_t pairs = iatoms * (uint64_t)it.jcount * kTile;
    if (pairs > worst) worst = pairs;
  }
  CHECK(worst < (1ull << 32), "item with %llu pairs (a_n %u b_n %u same %d paired %d)",
        (unsigned long long)worst, a_n, b_n, (int)same, (int)paired);
  std::printf("bound a_n=%u b_n=%u same=%d paired=%d items=%zu max_pairs=%llu\n", a_n, b_n,
              (int)same, (int)paired, items.size(), (unsigned long long)worst);
}

// Every (I tile, J tile) visit of the items, counted per unordered pair
// (like elements) or per ordered pair (cross elements).
static void check_coverage(uint32_t a_n, uint32_t b_n, bool same, bool paired, uint32_t split) {
  const auto types = one_type(a_n, b_n, same);
  std::vector<Item> items;
  add_items(types, 0, 1, items, paired, split);
  const uint32_t nb = same ? a_n : b_n;
  std::vector<uint32_t> seen((size_t)a_n * nb, 0u);
  auto visit = [&](uint32_t i, uint32_t j) {
    if (same && i > j) std::swap(i, j);
    ++seen[(size_t)i * nb + j];
  };
  for (const Item &it : items) {
    const uint32_t i0 = paired ? 2 * it.ti : it.ti;
    const uint32_t i1 = paired ? std::min(i0 + 2, a_n) : i0 + 1;
    for (uint32_t m = 0; m < it.jcount; ++m) {
      const uint32_t j = (it.jstart + m) % nb;
      if (same && j == i0) {
        visit(i0, i0);  // diagonal tile of the unit's first tile
      } else if (same && paired && j == i0 + 1 && i0 + 1 < a_n) {
        visit(i0, j);   // I tile 2u fully
        visit(j, j);    // I tile 2u + 1 diagonal
      } else {
        for (uint32_t i = i0; i < i1; ++i) visit(i, j);
      }
    }
  }
  size_t bad = 0;
  for (uint32_t i = 0; i < a_n; ++i)
    for (uint32_t j = same ? i : 0; j < nb; ++j)
      if (seen[(size_t)i * nb + j] != 1) ++bad;
  CHECK(bad == 0, "%zu tile pairs not visited exactly once (a_n %u b_n %u same %d paired %d split %u)",
        bad, a_n, b_n, (int)same, (int)paired, split);
}

int main() {
  // coverage on small and ragged tile counts, batch (1) and single-structure
  // (kSingleSplit) splits of the J runs
  for (uint32_t split : {1u, kSingleSplit})
    for (uint32_t n : {1u, 2u, 3u, 4u, 5u, 7u, 8u, 64u, 65u, 513u, 1000u})
      for (bool paired : {false, true}) {
        check_coverage(n, 0, true, paired, split);
        check_coverage(n, n / 3 + 1, false, paired, split);
      }
  // the bound where the J runs would be longest: elements of 2 * 10^7 atoms
  // (20,001 tiles) and more
  for (bool paired : {false, true}) {
    check_bound(20001, 0, true, paired);
    check_bound(20000, 0, true, paired);
    check_bound(500, 40000, false, paired);
    check_bound(40000, 40000, false, paired);
  }
  if (failures) {
    std::printf("%d failures\n", failures);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
