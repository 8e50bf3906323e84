// Host-side planning of the histogram kernels: tile, pair-type and work-item
// tables, the equal-count partition of an element, and the bounds they keep.
// Plain C++ (no HIP), so that the same code builds into libxs_hip and into
// CPU tests of the tables.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <utility>
#include <vector>

namespace xs {
namespace hist {

constexpr int kTile = 1024;      // atoms per tile == threads per work group
constexpr int kSortCap = 2048;   // J tiles sorted per work item
constexpr int kMaxEl = 16;
// J-run split of a single structure's work items (add_items `split`): the
// items of one structure are nearly equal, so the last round of work groups
// leaves CUs idle unless there are many; a batch plan has items enough.
constexpr uint32_t kSingleSplit = 8;

struct Tile {
  uint32_t off, cnt;  // atom range in the sorted arrays
};

// One pair type of one structure.
struct TypeDesc {
  unsigned long long hist_off;  // offset of its nhist bins in the output
  uint32_t nhist;
  uint32_t a_base, a_n;         // I tiles: global tile index base and count
  uint32_t b_base, b_n;         // J tiles
  uint32_t same;                // like-element pair type: j > i within a tile
  uint32_t ovf;                 // overflow counter index
  uint32_t pad;
  double bin, inv_bin, band;
};

struct Item {
  uint32_t type, ti, jstart, jcount;
};

// Half-width test of the FP64 fast path: an estimate whose fractional part
// lies within 1/2 - band of 1/2 is floored directly.  The estimate and the
// IEEE quotient differ by less than 2^-45.3 relative (see sqrt_over_bin in
// xs_hist.hip), i.e. nhist 2^-45.3 bins; the band keeps 16 times that (and
// 2^-50 for the floor of values near 0) away from every integer.
inline double band_for(uint32_t nhist) {
  return 0.5 - std::ldexp((double)nhist, -41) - std::ldexp(1.0, -50);
}

inline uint32_t wcap_for(uint32_t nhist) { return nhist <= 8192 ? 8192u : 32768u; }

// Longest J run of one work item.  The LDS window counts in uint32 and is
// flushed only when it slides or the item ends, so a slot can collect every
// pair of the item: the run is capped so that an item holds fewer than 2^32
// pairs (I unit atoms x kTile atoms per J tile x run).  Paired units hold up
// to 2 kTile atoms: 2047 tiles; single tiles: kSortCap.
constexpr uint32_t max_jrun(bool paired) {
  return (uint32_t)std::min<uint64_t>((uint64_t)kSortCap,
                                      0xffffffffull / ((paired ? 2ull : 1ull) * kTile * kTile));
}
static_assert((uint64_t)max_jrun(true) * 2 * kTile * kTile < (1ull << 32), "paired item pairs");
static_assert((uint64_t)max_jrun(false) * kTile * kTile < (1ull << 32), "item pairs");

// Work items of pair types [first, last) (appended).  The I side of an item
// is one tile, or with `paired` (FP64 binning) the tile pair {2 ti, 2 ti + 1}
// (an I unit).  Like-element pairs: I unit ti with itself and the next h
// units cyclically (h = (n-1)/2, or n/2 for the first half when n is even,
// n = units), so each unordered unit pair is visited once and items have
// equal length; the J run is the tiles of those units.  Cross-element pairs:
// each I unit against the J tiles in parts.  Paired items are split twice as
// finely, so that there are as many items as unpaired ones; `split` divides
// every J run further.
inline void add_items(const std::vector<TypeDesc> &types, uint32_t first, uint32_t last,
                      std::vector<Item> &items, bool paired, uint32_t split = 1) {
  for (uint32_t t = first; t < last; ++t) {
    const TypeDesc &d = types[t];
    if (!d.a_n || !d.b_n) continue;
    if (d.same) {
      const uint32_t nt = d.a_n;
      const uint32_t nu = paired ? (nt + 1) / 2 : nt;
      for (uint32_t u = 0; u < nu; ++u) {
        const uint32_t h = (nu & 1) ? (nu - 1) / 2 : (u < nu / 2 ? nu / 2 : nu / 2 - 1);
        uint32_t start = u, total = h + 1;
        if (paired) {
          start = 2 * u;
          total = 2 * (h + 1);
          // the last unit holds one tile when nt is odd
          if ((nt & 1) && (nu - 1 + nu - u) % nu <= h) --total;
        }
        const uint32_t parts = (paired ? 4 : 1) * split;
        const uint32_t chunk = std::min<uint32_t>((total + parts - 1) / parts, max_jrun(paired));
        for (uint32_t c = 0; c < total; c += chunk)
          items.push_back(Item{t, u, (start + c) % nt, std::min<uint32_t>(chunk, total - c)});
      }
    } else {
      const uint32_t nu = paired ? (d.a_n + 1) / 2 : d.a_n;
      const uint32_t parts = std::min<uint32_t>(d.b_n, (paired ? 8u : 2u) * split);
      const uint32_t per = std::min<uint32_t>((d.b_n + parts - 1) / parts, max_jrun(paired));
      for (uint32_t u = 0; u < nu; ++u)
        for (uint32_t j0 = 0; j0 < d.b_n; j0 += per)
          items.push_back(Item{t, u, j0, std::min(per, d.b_n - j0)});
    }
  }
}

// Longest first (estimated pairs: I unit atoms x atoms of its J run).
inline void order_items(std::vector<Item> &items, const std::vector<TypeDesc> &types,
                        const std::vector<Tile> &tiles, bool paired) {
  std::vector<std::pair<double, uint32_t>> cost(items.size());
  for (size_t i = 0; i < items.size(); ++i) {
    const Item &it = items[i];
    const TypeDesc &d = types[it.type];
    double jn = 0;
    for (uint32_t m = 0; m < it.jcount; ++m) jn += tiles[d.b_base + (it.jstart + m) % d.b_n].cnt;
    double in = 0;
    if (paired) {
      for (uint32_t a = 2 * it.ti; a < std::min(2 * it.ti + 2, d.a_n); ++a) in += tiles[d.a_base + a].cnt;
    } else {
      in = tiles[d.a_base + it.ti].cnt;
    }
    cost[i] = {-in * jn, (uint32_t)i};
  }
  std::stable_sort(cost.begin(), cost.end());
  std::vector<Item> out(items.size());
  for (size_t i = 0; i < items.size(); ++i) out[i] = items[cost[i].second];
  items.swap(out);
}

// Type descriptors of one structure whose element e has tiles [ebase[e], ebase[e]+en[e]).
inline void add_types(unsigned n_el, const uint32_t *ebase, const uint32_t *en, uint32_t nhist,
                      double bin, unsigned long long hist_off, uint32_t ovf,
                      std::vector<TypeDesc> &types) {
  unsigned long long off = hist_off;
  for (unsigned a = 0; a < n_el; ++a)
    for (unsigned b = a; b < n_el; ++b, off += nhist) {
      TypeDesc d{};
      d.hist_off = off;
      d.nhist = nhist;
      d.a_base = ebase[a];
      d.a_n = en[a];
      d.b_base = ebase[b];
      d.b_n = en[b];
      d.same = a == b;
      d.ovf = ovf;
      d.bin = bin;
      d.inv_bin = 1.0 / bin;
      d.band = band_for(nhist);
      types.push_back(d);
    }
}

// Equal-count partition of one element for the device path: gx slabs along
// x and gy columns along y in each slab, all of equal count, so every column
// holds ne / (gx gy) atoms (+-1) whatever the shape; columns are cut along z
// into ceil(count / kTile) (at least nz) equal tiles.  gx, gy make tiles
// about cubic and as full as possible; nx, ny, nz make slabs, columns and
// tiles of a uniform-density structure narrower than 0.9 of the window cap
// (the cap itself is checked on the tile boxes).
inline void choose_equal_count(uint32_t ne, const double bbox[6], double cap, bool paired,
                               uint32_t &gx, uint32_t &gy, uint32_t &nz) {
  double ext[3];
  for (int a = 0; a < 3; ++a) ext[a] = std::max(bbox[3 + a] - bbox[a], 1e-6);
  auto need = [&](double e) -> uint32_t {
    return std::isfinite(cap) ? (uint32_t)std::max(1.0, std::ceil(e / (0.9 * cap))) : 1u;
  };
  const uint32_t nx = need(ext[0]), ny = need(ext[1]);
  nz = need(ext[2]);
  if (paired) nz += nz & 1u;
  const double T = std::max(1.0, std::ceil(ne / (double)kTile));
  const double L = std::cbrt(ext[0] * ext[1] * ext[2] / T);
  const long cx = std::lround(ext[0] / L), cy = std::lround(ext[1] / L);
  double best = -INFINITY;
  gx = std::max(nx, 1u);
  gy = std::max(ny, 1u);
  for (long a = cx - 2; a <= cx + 2; ++a)
    for (long b = cy - 2; b <= cy + 2; ++b) {
      const uint64_t ga = std::max<uint64_t>(std::max<long>(a, 1), nx);
      const uint64_t gb = std::max<uint64_t>(std::max<long>(b, 1), ny);
      if (ga * gb > std::max<uint64_t>(ne, 1) || ga * gb > 65536) continue;
      const double col = (double)ne / (double)(ga * gb);
      double tpc = std::max(std::ceil(col / kTile), (double)nz);
      if (paired) tpc = 2.0 * std::ceil(tpc / 2.0);
      const double fill = ne / ((double)(ga * gb) * tpc * kTile);
      const double shape = std::fabs(std::log(ext[0] / ga / L)) + std::fabs(std::log(ext[1] / gb / L)) +
                           std::fabs(std::log(ext[2] / tpc / L));
      const double score = fill - 0.01 * shape;
      if (score > best) {
        best = score;
        gx = (uint32_t)ga;
        gy = (uint32_t)gb;
      }
    }
}

// Tiles of the equal-count partition (positions from `pos`), column by column.
// paired: an even number of runs per column, so that tiles 2u and 2u + 1 are
// neighbouring runs of one column (the I tile pairs of the FP64 kernel).
inline void equal_count_tiles(uint32_t pos, uint32_t ne, uint32_t gx, uint32_t gy, uint32_t nz,
                              bool paired, std::vector<Tile> &tiles) {
  for (uint64_t sl = 0; sl < gx; ++sl) {
    const uint64_t S = (sl + 1) * ne / gx - sl * ne / gx;
    for (uint64_t c = 0; c < gy; ++c) {
      const uint32_t C = (uint32_t)((c + 1) * S / gy - c * S / gy);
      if (!C) continue;
      uint32_t runs = std::max((C + kTile - 1) / (uint32_t)kTile, nz);
      if (paired) runs += runs & 1u;
      const uint32_t nt = std::min(C, runs);
      for (uint32_t t = 0, o = 0; t < nt; ++t) {
        const uint32_t len = (C - o) / (nt - t);
        tiles.push_back(Tile{pos + o, len});
        o += len;
      }
      pos += C;
    }
  }
}

}  // namespace hist
}  // namespace xs
