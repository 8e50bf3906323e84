// Host-side planning of the batched matrix-core pair sums (xs_batch_run with
// XS_PAIRSUM_MFMA): the column order, the column groups and their slice
// layout.  Plain C++ (no HIP), so that the same code builds into libxs_hip
// and into CPU tests of the tables.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace xs {

constexpr int kMfmaSlicesA = 5;      // int8 slices of the sinc matrix
constexpr int kMfmaMaxSlicesB = 5;   // 7-bit count slices (counts below 2^35)

// Columns are the pair types, sorted by count slices (most first) and then
// by bin count (largest first), so that every group of NT column tiles of 32
// columns shares one slice count and one k range (its largest bin count).
struct MfmaCol {
  unsigned long long hist_off;  // first bin of the column's histogram
  uint32_t nhist, type;         // its bin count; its row of S
};
struct MfmaTile {
  unsigned long long b_off;  // first count slice of the group, in 16-byte units
  uint32_t kt, col0, ncol;   // k tiles of 32 bins; first column (sorted order); columns
  uint32_t pad;
};
struct MfmaLaunch {
  uint32_t tile0, ntiles, nb, nt;  // groups [tile0, tile0 + ntiles), slices, tiles per group
};
struct MfmaBatchPlan {
  std::vector<MfmaCol> cols;
  std::vector<MfmaTile> tiles;
  std::vector<MfmaLaunch> launches;
  uint32_t kt_max = 0;  // k tiles of the sinc slices (largest bin count)
  uint32_t nq = 0;
  size_t a_bytes = 0, b_bytes = 0;
};

// Builds the plan from each pair type's histogram offset, bin count and an
// upper bound of its counts (its pair count).  False if a bound reaches 2^35
// (more count slices than the kernels take).  A group holds 32 NT columns of
// one slice count (NT = 2 for up to two slices, else 1: more accumulators
// would not fit), never columns of the next.  Its slices are laid out
// [nb][NT][kt][64] (16-byte units) at b_off; the sinc slices take
// kMfmaSlicesA x MT x kt_max KiB, MT = 4 ceil(nq / 128).
inline bool mfma_batch_plan(const std::vector<unsigned long long> &hist_off,
                            const std::vector<uint32_t> &nhist,
                            const std::vector<unsigned long long> &count_bound, unsigned nq,
                            MfmaBatchPlan *plan) {
  MfmaBatchPlan p;
  const size_t n = nhist.size();
  std::vector<uint32_t> nb(n);
  for (size_t t = 0; t < n; ++t) {
    uint32_t k = 1;
    while (k < (uint32_t)kMfmaMaxSlicesB && (count_bound[t] >> (7 * k)) != 0ull) ++k;
    if ((count_bound[t] >> (7 * k)) != 0ull) return false;
    nb[t] = k;
  }
  std::vector<uint32_t> order(n);
  for (size_t t = 0; t < n; ++t) order[t] = (uint32_t)t;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return nb[a] != nb[b] ? nb[a] > nb[b] : nhist[a] > nhist[b];
  });
  p.cols.resize(n);
  for (size_t c = 0; c < n; ++c) p.cols[c] = MfmaCol{hist_off[order[c]], nhist[order[c]], order[c]};
  unsigned long long b_off = 0;
  uint32_t max_nh = 0;
  for (size_t c0 = 0; c0 < n;) {
    const uint32_t k = nb[order[c0]];
    size_t c1 = c0;
    while (c1 < n && nb[order[c1]] == k) ++c1;
    const uint32_t NT = k <= 2 ? 2 : 1;
    MfmaLaunch L{(uint32_t)p.tiles.size(), 0u, k, NT};
    for (size_t g = c0; g < c1; g += 32 * NT) {
      const uint32_t kt = std::max<uint32_t>(1u, (nhist[order[g]] + 31) / 32);
      const uint32_t ncol = (uint32_t)std::min<size_t>(32 * NT, c1 - g);
      p.tiles.push_back(MfmaTile{b_off, kt, (uint32_t)g, ncol, 0u});
      b_off += (unsigned long long)k * NT * kt * 64;
      max_nh = std::max(max_nh, nhist[order[g]]);
      ++L.ntiles;
    }
    p.launches.push_back(L);
    c0 = c1;
  }
  const unsigned MT = (nq + 127) / 128 * 4;
  p.kt_max = std::max<uint32_t>(1u, (max_nh + 31) / 32);
  p.nq = nq;
  p.a_bytes = (size_t)kMfmaSlicesA * MT * p.kt_max * 1024;
  p.b_bytes = (size_t)b_off * 16;
  *plan = std::move(p);
  return true;
}

}  // namespace xs
