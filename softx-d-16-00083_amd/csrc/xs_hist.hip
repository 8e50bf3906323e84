// H1: pair-distance histogram on gfx950.
//
// Pipeline on the context stream:
//   1. bounding box of all atoms (hist_bbox_partial, hist_grid_setup), read back
//   2. a uniform grid of cubic cells sized for ~800 atoms of the most numerous
//      element (and small enough that two cell diagonals fit the LDS window);
//      per-element counting sort by cell (hist_cell_count, read back, host
//      prefix sums, hist_scatter).  A tile is one cell, or a 1024-atom chunk of
//      a fuller cell, so its atoms lie in a cube.  Counts are order
//      independent: the permutation does not change the histogram.
//   3. per-tile bounding boxes (hist_tile_boxes)
//   4. histogram kernel, one work item per (I tile, run of J tiles):
//        - the I tile's 1024 atoms sit one per thread in registers
//        - the J tiles are sorted by the lowest bin their pairs with the I tile
//          can reach (from the two boxes), and swept in that order
//        - counts go to a circular LDS window of WCAP uint32 bins that slides
//          up the histogram; bins that the remaining J tiles cannot reach are
//          flushed to the global uint64 histogram with one atomic per non-zero
//          bin, so flush traffic is O(nhist) per item, not O(pairs)
//        - a pair whose bin is outside the window (never, unless a tile's
//          distance span exceeds WCAP bins) takes a global atomic directly.
//
// The bin expression is the reference's (unsigned)(r / bin) with
// r = sqrt((dx*dx + dy*dy) + dz*dz): CalcFunctions.cpp:784-788 (double) and
// CalcFunctionsCUDA.cu:832-833 (float).  It is evaluated with correctly
// rounded sqrt and division and without contraction, so the counts are the
// oracle's bit for bit.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "xs_internal.h"

namespace xs {
namespace hist {

constexpr int kTile = 1024;     // atoms per tile == threads per work group
constexpr int kStage = 256;     // J atoms staged in LDS per step
constexpr int kSortCap = 2048;  // J tiles sorted per work item
constexpr int kMaxEl = 16;
constexpr int kMaxTypes = kMaxEl * (kMaxEl + 1) / 2;
constexpr int kBoxBlocks = 512; // partial bounding boxes

struct Item {
  uint32_t p, ti, jstart, jcount;
};

struct Geom {
  uint32_t off[kMaxEl], cnt[kMaxEl], tile_base[kMaxEl], ntiles[kMaxEl];
};

struct Tile {
  uint32_t off, cnt;  // global atom range in sorted order
};

struct Grid {
  double lo[3];
  double inv_cell;
  uint32_t g[3];
};

template <class T>
struct Args {
  const T *x, *y, *z;           // sorted coordinates
  const Tile *tiles;
  const double *boxes;          // 6 doubles per tile: lo xyz, hi xyz
  const Item *items;
  unsigned long long *hist;     // ntypes * nhist
  unsigned long long *overflow;
  T bin;
  double bin_d, inv_bin_d, band;
  uint32_t nhist;
  Geom g;
  uint8_t ta[kMaxTypes], tb[kMaxTypes];
};

// ---------------------------------------------------------------- sorting

template <class T>
__global__ void hist_bbox_partial(const T *__restrict__ x, const T *__restrict__ y,
                                  const T *__restrict__ z, size_t n, double *__restrict__ part) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double v[3] = {(double)x[i], (double)y[i], (double)z[i]};
    for (int a = 0; a < 3; ++a) {
      lo[a] = fmin(lo[a], v[a]);
      hi[a] = fmax(hi[a], v[a]);
    }
  }
  __shared__ double s[6][256];
  for (int a = 0; a < 3; ++a) {
    s[a][threadIdx.x] = lo[a];
    s[3 + a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = fmin(s[a][threadIdx.x], s[a][threadIdx.x + w]);
        s[3 + a][threadIdx.x] = fmax(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// bbox[0..2] = lo, bbox[3..5] = hi of all atoms
__global__ void hist_bbox_reduce(const double *__restrict__ part, int nparts,
                                 double *__restrict__ bbox) {
  __shared__ double s[6][256];
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int i = threadIdx.x; i < nparts; i += blockDim.x)
    for (int a = 0; a < 3; ++a) {
      v[a] = fmin(v[a], part[i * 6 + a]);
      v[3 + a] = fmax(v[3 + a], part[i * 6 + 3 + a]);
    }
  for (int a = 0; a < 6; ++a) s[a][threadIdx.x] = v[a];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = fmin(s[a][threadIdx.x], s[a][threadIdx.x + w]);
        s[3 + a][threadIdx.x] = fmax(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) bbox[threadIdx.x] = s[threadIdx.x][0];
}

template <class T>
__global__ void hist_cell_count(const T *__restrict__ x, const T *__restrict__ y,
                                const T *__restrict__ z, Geom g, Grid grid,
                                uint32_t *__restrict__ keys, uint32_t *__restrict__ count) {
  const int e = blockIdx.y;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.cnt[e]) return;
  const size_t gi = (size_t)g.off[e] + i;
  const double v[3] = {(double)x[gi], (double)y[gi], (double)z[gi]};
  uint32_t c[3];
  for (int a = 0; a < 3; ++a) {
    const double f = (v[a] - grid.lo[a]) * grid.inv_cell;
    const int ci = isfinite(f) ? (int)f : 0;
    c[a] = (uint32_t)min(max(ci, 0), (int)grid.g[a] - 1);
  }
  const uint32_t m = grid.g[0] * grid.g[1] * grid.g[2];
  const uint32_t key = (c[0] * grid.g[1] + c[1]) * grid.g[2] + c[2];
  keys[gi] = key;
  atomicAdd(&count[(size_t)e * m + key], 1u);
}

template <class T>
__global__ void hist_scatter(const T *__restrict__ x, const T *__restrict__ y,
                             const T *__restrict__ z, Geom g, uint32_t m,
                             const uint32_t *__restrict__ keys, uint32_t *__restrict__ start,
                             T *__restrict__ sx, T *__restrict__ sy, T *__restrict__ sz) {
  const int e = blockIdx.y;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.cnt[e]) return;
  const size_t gi = (size_t)g.off[e] + i;
  const uint32_t pos = atomicAdd(&start[(size_t)e * m + keys[gi]], 1u);
  sx[pos] = x[gi];
  sy[pos] = y[gi];
  sz[pos] = z[gi];
}

template <class T>
__global__ void hist_tile_boxes(const T *__restrict__ x, const T *__restrict__ y,
                                const T *__restrict__ z, const Tile *__restrict__ tiles,
                                double *__restrict__ boxes) {
  const uint32_t tg = blockIdx.x;
  const Tile t = tiles[tg];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = threadIdx.x; i < t.cnt; i += blockDim.x) {
    size_t gi = (size_t)t.off + i;
    double v[3] = {(double)x[gi], (double)y[gi], (double)z[gi]};
    for (int a = 0; a < 3; ++a) {
      lo[a] = fmin(lo[a], v[a]);
      hi[a] = fmax(hi[a], v[a]);
    }
  }
  __shared__ double s[6][256];
  for (int a = 0; a < 3; ++a) {
    s[a][threadIdx.x] = lo[a];
    s[3 + a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = fmin(s[a][threadIdx.x], s[a][threadIdx.x + w]);
        s[3 + a][threadIdx.x] = fmax(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) boxes[(size_t)tg * 6 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void fill_u64(unsigned long long *__restrict__ p, size_t n, unsigned long long v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------- histogram

// Conservative bin range [kl, ku] of all pairs between two boxes.
__device__ __forceinline__ void box_bins(const double *__restrict__ bi, const double *__restrict__ bj,
                                         double bin, uint32_t nhist, uint32_t &kl, uint32_t &ku) {
  double lo2 = 0.0, hi2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    double gap = fmax(fmax(bj[a] - bi[3 + a], bi[a] - bj[3 + a]), 0.0);
    double ext = fmax(bj[3 + a] - bi[a], bi[3 + a] - bj[a]);
    lo2 += gap * gap;
    hi2 += ext * ext;
  }
  const double top = (double)(nhist - 1);
  double fl = floor(sqrt(lo2) / bin) - 1.0;
  double fh = floor(sqrt(hi2) / bin) + 1.0;
  fl = fl < 0.0 ? 0.0 : (fl > top ? top : fl);
  fh = fh > top ? top : (fh < fl ? fl : fh);
  kl = (uint32_t)fl;
  ku = (uint32_t)fh;
}

template <int WCAP>
__device__ __forceinline__ void flush_window(uint32_t *win, uint32_t lo, uint32_t hi,
                                             unsigned long long *__restrict__ hp) {
  for (uint32_t k = lo + threadIdx.x; k < hi; k += kTile) {
    const uint32_t slot = k & (WCAP - 1);
    const uint32_t v = win[slot];
    if (v) {
      atomicAdd(&hp[k], (unsigned long long)v);
      win[slot] = 0u;
    }
  }
}

// Bin index of a squared distance s.
// double: the reference CPU expression (unsigned)(sqrt(s)/bin) in IEEE double.
//   Fast path: sqrt(s)/bin estimated with v_rsq_f64 and one Newton step times
//   1/bin.  The estimate and the IEEE value both lie within 2^-32 relative of
//   the real quotient (v_rsq_f64 is accurate to far better than 2^-16, which
//   the Newton step squares), so when the estimate's fractional part is more
//   than nhist*2^-30 away from an integer, its floor is the IEEE result.
//   Otherwise -- about nhist*2^-29 of the pairs, and every NaN -- the IEEE
//   sqrt and division are evaluated.  `band` = 1/2 - nhist*2^-30 - 2^-40.
// float: the IEEE single-precision expression of the reference GPU kernel,
//   evaluated directly (correctly rounded sqrtf and division).
template <class T>
__device__ __forceinline__ uint32_t bin_index(T s, T bin, double inv_bin_d, double band,
                                              uint32_t nhist);

template <>
__device__ __forceinline__ uint32_t bin_index<double>(double s, double bin, double inv_bin_d,
                                                      double band, uint32_t nhist) {
  const double y = __builtin_amdgcn_rsq(s);
  double g = s * y;
  g = fma(g, fma(-g, 0.5 * y, 0.5), g);
  const double kd = g * inv_bin_d;
  // kd < 2^31 here: the host rejects boxes whose diagonal exceeds 2^31 bins
  if (__builtin_expect(fabs(__builtin_amdgcn_fract(kd) - 0.5) < band, 1)) return (uint32_t)kd;
  const double t = __builtin_sqrt(s) / bin;
  return t < (double)nhist ? (uint32_t)t : nhist;
}

template <>
__device__ __forceinline__ uint32_t bin_index<float>(float s, float bin, double inv_bin_d,
                                                     double band, uint32_t nhist) {
  (void)inv_bin_d;
  (void)band;
  const float t = __builtin_sqrtf(s) / bin;
  return t < (float)nhist ? (uint32_t)t : nhist;
}

template <class T, int WCAP, bool DIAG>
__device__ __forceinline__ void pair_block(T xi, T yi, T zi, uint32_t il, const T *sx,
                                           const T *sy, const T *sz, uint32_t cn,
                                           uint32_t jbase, T bin, double inv_bin_d,
                                           double band, uint32_t nhist, uint32_t B, uint32_t wlim,
                                           uint32_t *win, unsigned long long *__restrict__ hp,
                                           unsigned long long *__restrict__ ovf) {
#pragma clang fp contract(off)
#pragma unroll 4
  for (uint32_t j = 0; j < cn; ++j) {
    if (DIAG && jbase + j <= il) continue;
    const T dx = xi - sx[j];
    const T dy = yi - sy[j];
    const T dz = zi - sz[j];
    T s = dx * dx + dy * dy;
    s = s + dz * dz;
    const uint32_t k = bin_index<T>(s, bin, inv_bin_d, band, nhist);
    if (__builtin_expect(k - B < wlim, 1)) {
      atomicAdd(&win[k & (WCAP - 1)], 1u);
    } else if (k < nhist) {
      atomicAdd(&hp[k], 1ull);
    } else {
      atomicAdd(ovf, 1ull);
    }
  }
}

template <class T, int WCAP>
__global__ __launch_bounds__(kTile) void hist_kernel(Args<T> A) {
  __shared__ uint32_t win[WCAP];
  __shared__ unsigned long long keys[kSortCap];
  __shared__ T sx[kStage], sy[kStage], sz[kStage];

  const Item it = A.items[blockIdx.x];
  const uint32_t a = A.ta[it.p], b = A.tb[it.p];
  const bool same = a == b;
  const uint32_t tid = threadIdx.x;
  const uint32_t tiI = A.g.tile_base[a] + it.ti;
  const Tile I = A.tiles[tiI];
  const bool ivalid = tid < I.cnt;
  T xi = 0, yi = 0, zi = 0;
  if (ivalid) {
    const size_t gi = (size_t)I.off + tid;
    xi = A.x[gi];
    yi = A.y[gi];
    zi = A.z[gi];
  }
  const double *bi = A.boxes + 6 * (size_t)tiI;
  const uint32_t ntb = A.g.ntiles[b];
  const uint32_t n = it.jcount;
  uint32_t npow = 1;
  while (npow < n) npow <<= 1;
  unsigned long long *hp = A.hist + (size_t)it.p * A.nhist;

  for (uint32_t m = tid; m < npow; m += kTile) {
    unsigned long long key = ~0ull;
    if (m < n) {
      const uint32_t tj = (it.jstart + m) % ntb;
      uint32_t kl, ku;
      box_bins(bi, A.boxes + 6 * (size_t)(A.g.tile_base[b] + tj), A.bin_d, A.nhist, kl, ku);
      key = ((unsigned long long)kl << 32) | m;
    }
    keys[m] = key;
  }
  for (uint32_t w = tid; w < WCAP; w += kTile) win[w] = 0u;
  __syncthreads();
  for (uint32_t k = 2; k <= npow; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = tid; i < npow; i += kTile) {
        const uint32_t ixj = i ^ j;
        if (ixj > i) {
          const unsigned long long u = keys[i], v = keys[ixj];
          const bool up = (i & k) == 0;
          if ((u > v) == up) {
            keys[i] = v;
            keys[ixj] = u;
          }
        }
      }
      __syncthreads();
    }
  }

  uint32_t B = (uint32_t)(keys[0] >> 32);
  for (uint32_t s = 0; s < n; ++s) {
    const uint32_t m = (uint32_t)(keys[s] & 0xffffffffu);
    const uint32_t tj = (it.jstart + m) % ntb;
    const uint32_t tjg = A.g.tile_base[b] + tj;
    uint32_t kl, ku;
    box_bins(bi, A.boxes + 6 * (size_t)tjg, A.bin_d, A.nhist, kl, ku);
    uint32_t wlim = 0;
    if (ku - kl + 1 <= (uint32_t)WCAP) {
      const uint32_t nb = ku + 1 > (uint32_t)WCAP ? ku + 1 - WCAP : 0u;
      if (nb > B) {
        flush_window<WCAP>(win, B, min(min(nb, B + (uint32_t)WCAP), A.nhist), hp);
        B = nb;
        __syncthreads();
      }
      wlim = min((uint32_t)WCAP, A.nhist - B);
    }
    const Tile J = A.tiles[tjg];
    const bool diag = same && tj == it.ti;
    for (uint32_t s0 = 0; s0 < J.cnt; s0 += kStage) {
      const uint32_t cn = min((uint32_t)kStage, J.cnt - s0);
      if (tid < cn) {
        sx[tid] = A.x[(size_t)J.off + s0 + tid];
        sy[tid] = A.y[(size_t)J.off + s0 + tid];
        sz[tid] = A.z[(size_t)J.off + s0 + tid];
      }
      __syncthreads();
      if (ivalid) {
        if (diag)
          pair_block<T, WCAP, true>(xi, yi, zi, tid, sx, sy, sz, cn, s0, A.bin, A.inv_bin_d,
                                    A.band, A.nhist, B, wlim, win, hp, A.overflow);
        else
          pair_block<T, WCAP, false>(xi, yi, zi, tid, sx, sy, sz, cn, s0, A.bin, A.inv_bin_d,
                                     A.band, A.nhist, B, wlim, win, hp, A.overflow);
      }
      __syncthreads();
    }
  }
  flush_window<WCAP>(win, B, min(B + (uint32_t)WCAP, A.nhist), hp);
}

// ---------------------------------------------------------------- host side

// Work items over the element tile lists.  Like-element pairs: tile ti pairs
// with itself and the next h tiles cyclically (h = (n-1)/2, or n/2 for the
// first half when n is even), so every unordered tile pair is visited once and
// items have equal length.  Cross-element pairs: each I tile against the J
// tiles in two halves.  Longest items first.
static void build_items(const Geom &g, unsigned n_el, std::vector<Item> &items) {
  items.clear();
  unsigned p = 0;
  for (unsigned a = 0; a < n_el; ++a) {
    for (unsigned b = a; b < n_el; ++b, ++p) {
      const uint32_t nta = g.ntiles[a], ntb = g.ntiles[b];
      if (!nta || !ntb) continue;
      if (a == b) {
        for (uint32_t ti = 0; ti < nta; ++ti) {
          uint32_t h = (nta & 1) ? (nta - 1) / 2 : (ti < nta / 2 ? nta / 2 : nta / 2 - 1);
          uint32_t total = h + 1;
          for (uint32_t c = 0; c < total; c += kSortCap)
            items.push_back(Item{p, ti, (ti + c) % ntb, std::min<uint32_t>(kSortCap, total - c)});
        }
      } else {
        const uint32_t parts = ntb >= 2 ? 2 : 1;
        const uint32_t per = std::min<uint32_t>((ntb + parts - 1) / parts, kSortCap);
        for (uint32_t ti = 0; ti < nta; ++ti)
          for (uint32_t j0 = 0; j0 < ntb; j0 += per)
            items.push_back(Item{p, ti, j0, std::min(per, ntb - j0)});
      }
    }
  }
  std::stable_sort(items.begin(), items.end(),
                   [](const Item &u, const Item &v) { return u.jcount > v.jcount; });
}

template <class T, int WCAP>
static int launch_hist(xs_ctx *ctx, const Args<T> &A, size_t nitems) {
  hipEvent_t ev;
  XS_TRY(prof_begin(ctx, "hist", &ev));
  hipLaunchKernelGGL((hist_kernel<T, WCAP>), dim3((unsigned)nitems), dim3(kTile), 0, ctx->stream, A);
  XS_HIP_CHECK(hipGetLastError());
  XS_TRY(prof_end(ctx, "hist", ev));
  return XS_OK;
}

// Cell edge: cubes holding about kFill atoms of the most numerous element at
// its mean density over the bounding box, no wider than two diagonals of the
// window allow when the histogram is longer than the window.
constexpr double kFill = 800.0;
constexpr uint32_t kMaxCellsPerAxis = 256;

static Grid make_grid(const double bbox[6], unsigned max_count, double bin, uint32_t nhist,
                      uint32_t wcap) {
  Grid gr{};
  double ext[3], vol = 1.0;
  for (int a = 0; a < 3; ++a) {
    gr.lo[a] = bbox[a];
    ext[a] = std::max(bbox[3 + a] - bbox[a], 0.0);
  }
  // volume of the box, with degenerate axes widened to the typical spacing
  double span = std::max({ext[0], ext[1], ext[2], 1e-6});
  double spacing = span / std::max(1.0, std::cbrt((double)std::max(max_count, 1u)));
  for (int a = 0; a < 3; ++a) vol *= std::max(ext[a], spacing);
  double cell = std::cbrt(kFill * vol / std::max(max_count, 1u));
  if (nhist > wcap) cell = std::min(cell, (wcap - 8) * bin / (2.0 * std::sqrt(3.0)));
  cell = std::max(cell, span / kMaxCellsPerAxis);
  cell = std::max(cell, 1e-9);
  gr.inv_cell = 1.0 / cell;
  for (int a = 0; a < 3; ++a)
    gr.g[a] = (uint32_t)std::min<double>(kMaxCellsPerAxis, std::max(1.0, std::ceil(ext[a] / cell + 1e-9)));
  return gr;
}

template <class T>
static int run(xs_ctx *ctx, const T *xyz_in, size_t n, const unsigned *el_count, unsigned n_el,
               double bin, unsigned nhist, uint64_t *hist_out, uint64_t *ovf_out, bool device) {
  const unsigned ntypes = ntypes_of(n_el);
  const size_t hsize = (size_t)ntypes * nhist;
  const uint32_t wcap = nhist <= 8192 ? 8192u : 32768u;
  Geom g{};
  uint32_t off = 0;
  unsigned max_count = 0;
  for (unsigned e = 0; e < n_el; ++e) {
    g.off[e] = off;
    g.cnt[e] = el_count[e];
    off += el_count[e];
    max_count = std::max(max_count, el_count[e]);
  }
  XS_TRY(ensure_pinned(ctx, 4096));

  // 1. bounding box: on the host for host input; on the device (small
  //    dedicated workspace, read back) for device input.  Nothing else is
  //    enqueued before the workspace is sized below.
  double bbox[6] = {0, 0, 0, 0, 0, 0};
  if (n > 0) {
    if (!device) {
      for (int a = 0; a < 3; ++a) {
        bbox[a] = INFINITY;
        bbox[3 + a] = -INFINITY;
      }
      for (int a = 0; a < 3; ++a)
        for (size_t i = 0; i < n; ++i) {
          const double v = (double)xyz_in[a * n + i];
          bbox[a] = std::min(bbox[a], v);
          bbox[3 + a] = std::max(bbox[3 + a], v);
        }
    } else {
      XS_TRY(ensure_workspace(ctx, Carver::need({kBoxBlocks * 6 * sizeof(double), 8 * sizeof(double)})));
      Carver cb(ctx->ws);
      double *part = cb.take<double>(kBoxBlocks * 6);
      double *bbox_d = cb.take<double>(8);
      const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, kBoxBlocks);
      hipLaunchKernelGGL((hist_bbox_partial<T>), dim3(nb), dim3(256), 0, ctx->stream, xyz_in,
                         xyz_in + n, xyz_in + 2 * n, n, part);
      hipLaunchKernelGGL(hist_bbox_reduce, dim3(1), dim3(256), 0, ctx->stream, part, (int)nb, bbox_d);
      XS_HIP_CHECK(hipGetLastError());
      XS_HIP_CHECK(hipMemcpyAsync(ctx->pin, bbox_d, 6 * sizeof(double), hipMemcpyDeviceToHost,
                                  ctx->stream));
      XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
      std::memcpy(bbox, ctx->pin, sizeof(bbox));
    }
    for (int a = 0; a < 6; ++a)
      if (!std::isfinite(bbox[a])) return fail(XS_ERR_ARG, "xs_hist: non-finite coordinates");
    const double diag = std::sqrt((bbox[3] - bbox[0]) * (bbox[3] - bbox[0]) +
                                  (bbox[4] - bbox[1]) * (bbox[4] - bbox[1]) +
                                  (bbox[5] - bbox[2]) * (bbox[5] - bbox[2]));
    if (diag / bin >= 2147483648.0)
      return fail(XS_ERR_RANGE, "xs_hist: the structure spans more than 2^31 bins");
  }
  const Grid grid = make_grid(bbox, max_count, bin, nhist, wcap);
  const uint32_t m = grid.g[0] * grid.g[1] * grid.g[2];

  // 2. one workspace sizing for the whole call (upper bounds for tiles/items)
  const size_t max_tiles = n / kTile + (size_t)n_el * m + n_el;
  const size_t max_items = max_tiles * (2 + max_tiles / kSortCap) + (size_t)n_el * n_el;
  const size_t need = Carver::need({device ? 0 : 3 * n * sizeof(T), 3 * n * sizeof(T),
                                    n * sizeof(uint32_t), (size_t)n_el * m * sizeof(uint32_t),
                                    (size_t)n_el * m * sizeof(uint32_t), max_tiles * sizeof(Tile),
                                    max_tiles * 6 * sizeof(double), max_items * sizeof(Item),
                                    device ? 0 : hsize * sizeof(uint64_t), sizeof(uint64_t)});
  XS_TRY(ensure_workspace(ctx, need));
  XS_TRY(ensure_pinned(ctx, Carver::need({(size_t)n_el * m * sizeof(uint32_t),
                                          (size_t)n_el * m * sizeof(uint32_t),
                                          max_tiles * sizeof(Tile), max_items * sizeof(Item)})));
  Carver cv(ctx->ws);
  T *stage = device ? nullptr : cv.take<T>(3 * n);
  T *sorted = cv.take<T>(3 * n);
  uint32_t *keys = cv.take<uint32_t>(n);
  uint32_t *count = cv.take<uint32_t>((size_t)n_el * m);
  uint32_t *start_d = cv.take<uint32_t>((size_t)n_el * m);
  Tile *tiles_d = cv.take<Tile>(max_tiles);
  double *boxes = cv.take<double>(max_tiles * 6);
  Item *items_d = cv.take<Item>(max_items);
  unsigned long long *hist_d =
      device ? reinterpret_cast<unsigned long long *>(hist_out) : cv.take<unsigned long long>(hsize);
  unsigned long long *ovf_d = (device && ovf_out) ? reinterpret_cast<unsigned long long *>(ovf_out)
                                                  : cv.take<unsigned long long>(1);
  const T *xyz = xyz_in;
  if (!device && n) {
    XS_HIP_CHECK(hipMemcpyAsync(stage, xyz_in, 3 * n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    xyz = stage;
  }
  const T *x = xyz, *y = xyz + n, *z = xyz + 2 * n;
  T *sxp = sorted, *syp = sorted + n, *szp = sorted + 2 * n;

  hipEvent_t ev;
  XS_TRY(prof_begin(ctx, "hist_prep", &ev));
  {
    const unsigned zb = (unsigned)std::min<size_t>((hsize + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_u64, dim3(std::max(zb, 1u)), dim3(256), 0, ctx->stream, hist_d, hsize, 0ull);
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, ctx->stream, ovf_d, (size_t)1, 0ull);
  }
  std::vector<Tile> tiles;
  std::vector<Item> items;
  if (n > 0 && ntypes > 0) {
    // 3. cell counts, read back; prefix sums, tiles and work items on the host
    const size_t cbytes = (size_t)n_el * m * sizeof(uint32_t);
    XS_HIP_CHECK(hipMemsetAsync(count, 0, cbytes, ctx->stream));
    const dim3 eg((max_count + 255) / 256, n_el);
    if (max_count)
      hipLaunchKernelGGL((hist_cell_count<T>), eg, dim3(256), 0, ctx->stream, x, y, z, g, grid, keys,
                         count);
    XS_HIP_CHECK(hipGetLastError());
    Carver pc(ctx->pin);
    uint32_t *count_h = pc.take<uint32_t>((size_t)n_el * m);
    uint32_t *start_h = pc.take<uint32_t>((size_t)n_el * m);
    Tile *tiles_h = pc.take<Tile>(max_tiles);
    Item *items_h = pc.take<Item>(max_items);
    XS_HIP_CHECK(hipMemcpyAsync(count_h, count, cbytes, hipMemcpyDeviceToHost, ctx->stream));
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (unsigned e = 0; e < n_el; ++e) {
      g.tile_base[e] = (uint32_t)tiles.size();
      uint32_t pos = g.off[e];
      for (uint32_t c = 0; c < m; ++c) {
        const uint32_t cnt = count_h[(size_t)e * m + c];
        start_h[(size_t)e * m + c] = pos;
        for (uint32_t t = 0; t < cnt; t += kTile)
          tiles.push_back(Tile{pos + t, std::min<uint32_t>(kTile, cnt - t)});
        pos += cnt;
      }
      g.ntiles[e] = (uint32_t)tiles.size() - g.tile_base[e];
    }
    build_items(g, n_el, items);
    if (tiles.size() > max_tiles || items.size() > max_items)
      return fail(XS_ERR_RANGE, "xs_hist: tile bound exceeded");
    std::memcpy(tiles_h, tiles.data(), tiles.size() * sizeof(Tile));
    std::memcpy(items_h, items.data(), items.size() * sizeof(Item));
    XS_HIP_CHECK(hipMemcpyAsync(start_d, start_h, cbytes, hipMemcpyHostToDevice, ctx->stream));
    if (!tiles.empty())
      XS_HIP_CHECK(hipMemcpyAsync(tiles_d, tiles_h, tiles.size() * sizeof(Tile),
                                  hipMemcpyHostToDevice, ctx->stream));
    if (!items.empty())
      XS_HIP_CHECK(hipMemcpyAsync(items_d, items_h, items.size() * sizeof(Item),
                                  hipMemcpyHostToDevice, ctx->stream));
    if (max_count)
      hipLaunchKernelGGL((hist_scatter<T>), eg, dim3(256), 0, ctx->stream, x, y, z, g, m, keys,
                         start_d, sxp, syp, szp);
    if (!tiles.empty())
      hipLaunchKernelGGL((hist_tile_boxes<T>), dim3((unsigned)tiles.size()), dim3(256), 0,
                         ctx->stream, sxp, syp, szp, tiles_d, boxes);
    XS_HIP_CHECK(hipGetLastError());
  }
  XS_TRY(prof_end(ctx, "hist_prep", ev));

  // 4. histogram
  if (!items.empty()) {
    Args<T> A{};
    A.x = sxp;
    A.y = syp;
    A.z = szp;
    A.tiles = tiles_d;
    A.boxes = boxes;
    A.items = items_d;
    A.hist = hist_d;
    A.overflow = ovf_d;
    A.bin = (T)bin;
    A.bin_d = bin;
    A.inv_bin_d = 1.0 / bin;
    A.band = 0.5 - std::ldexp((double)nhist, -30) - std::ldexp(1.0, -40);
    A.nhist = nhist;
    A.g = g;
    unsigned p = 0;
    for (unsigned a = 0; a < n_el; ++a)
      for (unsigned b = a; b < n_el; ++b, ++p) {
        A.ta[p] = (uint8_t)a;
        A.tb[p] = (uint8_t)b;
      }
    if (wcap == 8192)
      XS_TRY((launch_hist<T, 8192>(ctx, A, items.size())));
    else
      XS_TRY((launch_hist<T, 32768>(ctx, A, items.size())));
  }

  if (!device) {
    XS_HIP_CHECK(hipMemcpyAsync(hist_out, hist_d, hsize * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                ctx->stream));
    XS_HIP_CHECK(hipMemcpyAsync(ctx->pin, ovf_d, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (ovf_out) std::memcpy(ovf_out, ctx->pin, sizeof(uint64_t));
  } else {
    // the pinned staging feeds the uploads above; keep it untouched until they ran
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
  return XS_OK;
}

}  // namespace hist
}  // namespace xs

using namespace xs;

extern "C" int xs_hist(xs_ctx *ctx, const void *xyz, size_t n_atoms, const unsigned *el_count,
                       unsigned n_el, double bin, unsigned nhist, uint64_t *hist,
                       uint64_t *overflow, unsigned flags) {
  XS_TRY(activate(ctx));
  if (!el_count || (!xyz && n_atoms) || !hist) return fail(XS_ERR_ARG, "xs_hist: null pointer");
  if (n_el == 0 || n_el > (unsigned)hist::kMaxEl)
    return fail(XS_ERR_RANGE, "xs_hist: n_el must be in [1, 16]");
  size_t tot = 0;
  for (unsigned e = 0; e < n_el; ++e) tot += el_count[e];
  if (tot != n_atoms) return fail(XS_ERR_ARG, "xs_hist: sum(el_count) != n_atoms");
  if (n_atoms >= (size_t(1) << 31)) return fail(XS_ERR_RANGE, "xs_hist: too many atoms");
  if (!(bin > 0.0) || !std::isfinite(bin)) return fail(XS_ERR_ARG, "xs_hist: bin must be > 0");
  if (nhist == 0) return fail(XS_ERR_ARG, "xs_hist: nhist must be > 0");
  const bool f64 = flags & XS_F64_BINNING;
  if (!f64 && nhist > (1u << 24))
    return fail(XS_ERR_RANGE, "xs_hist: nhist > 2^24 needs XS_F64_BINNING");
  const bool device = flags & XS_DEVICE_PTRS;
  if (f64)
    return hist::run<double>(ctx, static_cast<const double *>(xyz), n_atoms, el_count, n_el, bin,
                             nhist, hist, overflow, device);
  return hist::run<float>(ctx, static_cast<const float *>(xyz), n_atoms, el_count, n_el, bin, nhist,
                          hist, overflow, device);
}
