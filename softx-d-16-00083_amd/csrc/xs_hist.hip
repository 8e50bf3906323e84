// H1: pair-distance histogram on gfx950.
//
// Pipeline (all on the context stream, no host synchronisation in device mode):
//   1. bounding box of all atoms              (hist_bbox_partial, hist_grid_setup)
//   2. per-element counting sort by Morton cell (hist_morton_count, hist_scan, hist_scatter)
//      -> atoms of each element reordered so that 1024 consecutive atoms form a
//         compact tile; counts are order independent, so the histogram is not
//         affected by the permutation
//   3. per-tile bounding boxes                  (hist_tile_boxes)
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
constexpr int kMaxBits = 6;     // Morton bits per axis
constexpr int kBoxBlocks = 512; // partial bounding boxes

struct Item {
  uint32_t p, ti, jstart, jcount;
};

struct Geom {
  uint32_t off[kMaxEl], cnt[kMaxEl], tile_base[kMaxEl], ntiles[kMaxEl];
};

template <class T>
struct Args {
  const T *x, *y, *z;           // sorted coordinates
  const double *boxes;          // 6 doubles per tile: lo xyz, hi xyz
  const Item *items;
  unsigned long long *hist;     // ntypes * nhist
  unsigned long long *overflow;
  T bin;
  double bin_d;
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

// grid[0..2] = lo, grid[3..5] = cells per unit length
__global__ void hist_grid_setup(const double *__restrict__ part, int nparts, int bits,
                                double *__restrict__ grid) {
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
  if (threadIdx.x == 0) {
    for (int a = 0; a < 3; ++a) {
      double lo = s[a][0], hi = s[3 + a][0];
      double ext = hi - lo;
      grid[a] = lo;
      grid[3 + a] = (ext > 0 && isfinite(ext)) ? (double)(1 << bits) / ext : 0.0;
    }
  }
}

__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x3ff;
  v = (v | (v << 16)) & 0x030000FF;
  v = (v | (v << 8)) & 0x0300F00F;
  v = (v | (v << 4)) & 0x030C30C3;
  v = (v | (v << 2)) & 0x09249249;
  return v;
}

template <class T>
__global__ void hist_morton_count(const T *__restrict__ x, const T *__restrict__ y,
                                  const T *__restrict__ z, Geom g, const double *__restrict__ grid,
                                  int bits, uint32_t *__restrict__ keys,
                                  uint32_t *__restrict__ count) {
  const int e = blockIdx.y;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.cnt[e]) return;
  const size_t gi = (size_t)g.off[e] + i;
  const int cmax = (1 << bits) - 1;
  double v[3] = {(double)x[gi], (double)y[gi], (double)z[gi]};
  uint32_t c[3];
  for (int a = 0; a < 3; ++a) {
    double f = (v[a] - grid[a]) * grid[3 + a];
    int ci = isfinite(f) ? (int)f : 0;
    c[a] = (uint32_t)min(max(ci, 0), cmax);
  }
  const uint32_t key = spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
  keys[gi] = key;
  atomicAdd(&count[(size_t)e * (1u << (3 * bits)) + key], 1u);
}

// Exclusive scan of one element's cell counts (one block of 1024 threads per
// element), turned into absolute start positions.
__global__ __launch_bounds__(1024) void hist_scan(uint32_t *__restrict__ count, Geom g,
                                                  uint32_t m) {
  const int e = blockIdx.x;
  uint32_t *c = count + (size_t)e * m;
  __shared__ uint32_t wave_sum[16];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = g.off[e];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t base = 0; base < m; base += 1024) {
    const uint32_t idx = base + threadIdx.x;
    uint32_t v = idx < m ? c[idx] : 0u;
    uint32_t incl = v;
    for (int d = 1; d < 64; d <<= 1) {
      uint32_t t = __shfl_up(incl, d, 64);
      if (lane >= d) incl += t;
    }
    if (lane == 63) wave_sum[wave] = incl;
    __syncthreads();
    if (threadIdx.x < 16) {
      uint32_t w = wave_sum[threadIdx.x], wi = w;
      for (int d = 1; d < 16; d <<= 1) {
        uint32_t t = __shfl_up(wi, d, 16);
        if ((int)threadIdx.x >= d) wi += t;
      }
      wave_sum[threadIdx.x] = wi - w;  // exclusive prefix of wave totals
    }
    __syncthreads();
    const uint32_t excl = carry + wave_sum[wave] + incl - v;
    if (idx < m) c[idx] = excl;
    __syncthreads();
    if (threadIdx.x == 1023) carry = excl + v;
    __syncthreads();
  }
}

template <class T>
__global__ void hist_scatter(const T *__restrict__ x, const T *__restrict__ y,
                             const T *__restrict__ z, Geom g, int bits,
                             const uint32_t *__restrict__ keys, uint32_t *__restrict__ start,
                             T *__restrict__ sx, T *__restrict__ sy, T *__restrict__ sz) {
  const int e = blockIdx.y;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.cnt[e]) return;
  const size_t gi = (size_t)g.off[e] + i;
  const uint32_t pos = atomicAdd(&start[(size_t)e * (1u << (3 * bits)) + keys[gi]], 1u);
  sx[pos] = x[gi];
  sy[pos] = y[gi];
  sz[pos] = z[gi];
}

template <class T>
__global__ void hist_tile_boxes(const T *__restrict__ x, const T *__restrict__ y,
                                const T *__restrict__ z, Geom g, int n_el,
                                double *__restrict__ boxes) {
  const uint32_t tg = blockIdx.x;
  int e = 0;
  while (e + 1 < n_el && tg >= g.tile_base[e + 1]) ++e;
  const uint32_t t = tg - g.tile_base[e];
  const uint32_t i0 = t * kTile, i1 = min(g.cnt[e], i0 + kTile);
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    size_t gi = (size_t)g.off[e] + i;
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

template <class T, int WCAP, bool DIAG>
__device__ __forceinline__ void pair_block(T xi, T yi, T zi, uint32_t il, const T *sx,
                                           const T *sy, const T *sz, uint32_t cn,
                                           uint32_t jbase, T bin, uint32_t nhist, uint32_t B,
                                           uint32_t wlim, uint32_t *win,
                                           unsigned long long *__restrict__ hp,
                                           unsigned long long *__restrict__ ovf) {
#pragma clang fp contract(off)
  const T nh = (T)nhist;
#pragma unroll 4
  for (uint32_t j = 0; j < cn; ++j) {
    if (DIAG && jbase + j <= il) continue;
    const T dx = xi - sx[j];
    const T dy = yi - sy[j];
    const T dz = zi - sz[j];
    T s = dx * dx + dy * dy;
    s = s + dz * dz;
    const T r = __builtin_sqrt(s);
    const T t = r / bin;
    const uint32_t k = t < nh ? (uint32_t)t : nhist;
    if (k - B < wlim) {
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
  const uint32_t il = it.ti * kTile + tid;
  const bool ivalid = il < A.g.cnt[a];
  T xi = 0, yi = 0, zi = 0;
  if (ivalid) {
    const size_t gi = (size_t)A.g.off[a] + il;
    xi = A.x[gi];
    yi = A.y[gi];
    zi = A.z[gi];
  }
  const double *bi = A.boxes + 6 * (size_t)(A.g.tile_base[a] + it.ti);
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
    uint32_t kl, ku;
    box_bins(bi, A.boxes + 6 * (size_t)(A.g.tile_base[b] + tj), A.bin_d, A.nhist, kl, ku);
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
    const uint32_t jl0 = tj * kTile;
    const uint32_t jn = min((uint32_t)kTile, A.g.cnt[b] - jl0);
    const bool diag = same && tj == it.ti;
    const size_t jg0 = (size_t)A.g.off[b] + jl0;
    for (uint32_t s0 = 0; s0 < jn; s0 += kStage) {
      const uint32_t cn = min((uint32_t)kStage, jn - s0);
      if (tid < cn) {
        sx[tid] = A.x[jg0 + s0 + tid];
        sy[tid] = A.y[jg0 + s0 + tid];
        sz[tid] = A.z[jg0 + s0 + tid];
      }
      __syncthreads();
      if (ivalid) {
        if (diag)
          pair_block<T, WCAP, true>(xi, yi, zi, il, sx, sy, sz, cn, jl0 + s0, A.bin, A.nhist, B,
                                    wlim, win, hp, A.overflow);
        else
          pair_block<T, WCAP, false>(xi, yi, zi, il, sx, sy, sz, cn, jl0 + s0, A.bin, A.nhist, B,
                                     wlim, win, hp, A.overflow);
      }
      __syncthreads();
    }
  }
  flush_window<WCAP>(win, B, min(B + (uint32_t)WCAP, A.nhist), hp);
}

// ---------------------------------------------------------------- host side

static int morton_bits(const unsigned *el_count, unsigned n_el) {
  unsigned mx = 0;
  for (unsigned e = 0; e < n_el; ++e) mx = std::max(mx, el_count[e]);
  double cells = std::max(1.0, mx / 8.0);
  int bits = (int)std::ceil(std::log2(std::cbrt(cells)));
  return std::min(std::max(bits, 1), kMaxBits);
}

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

static int upload_items(xs_ctx *ctx, const Geom &g, unsigned n_el, const Item **dev,
                        size_t *count) {
  std::vector<uint32_t> key(g.cnt, g.cnt + n_el);
  key.push_back(n_el);
  key.push_back(kTile);
  static thread_local std::vector<Item> items;
  if (key != ctx->hist_items_key || !ctx->hist_items_dev) {
    build_items(g, n_el, items);
    const size_t bytes = std::max<size_t>(items.size(), 1) * sizeof(Item);
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (bytes > ctx->hist_items_cap) {
      if (ctx->hist_items_dev) XS_HIP_CHECK(hipFree(ctx->hist_items_dev));
      ctx->hist_items_dev = nullptr;
      ctx->hist_items_cap = 0;
      if (hipMalloc(&ctx->hist_items_dev, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(XS_ERR_NOMEM, "hipMalloc of the histogram work list failed");
      }
      ctx->hist_items_cap = bytes;
    }
    XS_TRY(ensure_pinned(ctx, bytes));
    std::memcpy(ctx->pin, items.data(), items.size() * sizeof(Item));
    XS_HIP_CHECK(hipMemcpyAsync(ctx->hist_items_dev, ctx->pin, items.size() * sizeof(Item),
                                hipMemcpyHostToDevice, ctx->stream));
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->hist_items_key = key;
    ctx->hist_items_n = items.size();
  }
  *dev = static_cast<const Item *>(ctx->hist_items_dev);
  *count = ctx->hist_items_n;
  return XS_OK;
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

template <class T>
static int run(xs_ctx *ctx, const T *xyz_in, size_t n, const unsigned *el_count, unsigned n_el,
               double bin, unsigned nhist, uint64_t *hist_out, uint64_t *ovf_out, bool device) {
  const unsigned ntypes = ntypes_of(n_el);
  const size_t hsize = (size_t)ntypes * nhist;
  Geom g{};
  uint32_t off = 0, tb = 0;
  for (unsigned e = 0; e < n_el; ++e) {
    g.off[e] = off;
    g.cnt[e] = el_count[e];
    g.tile_base[e] = tb;
    g.ntiles[e] = (el_count[e] + kTile - 1) / kTile;
    off += el_count[e];
    tb += g.ntiles[e];
  }
  const uint32_t ntiles = tb;
  const int bits = morton_bits(el_count, n_el);
  const uint32_t m = 1u << (3 * bits);

  const size_t need = Carver::need({
      device ? 0 : 3 * n * sizeof(T),       // staged input
      3 * n * sizeof(T),                    // sorted coordinates
      n * sizeof(uint32_t),                 // cell keys
      (size_t)n_el * m * sizeof(uint32_t),  // cell counts / starts
      kBoxBlocks * 6 * sizeof(double),      // partial boxes
      8 * sizeof(double),                   // grid
      (size_t)std::max<uint32_t>(ntiles, 1) * 6 * sizeof(double),
      device ? 0 : hsize * sizeof(uint64_t),
      sizeof(uint64_t)});
  XS_TRY(ensure_workspace(ctx, need));
  Carver cv(ctx->ws);
  T *stage = device ? nullptr : cv.take<T>(3 * n);
  T *sorted = cv.take<T>(3 * n);
  uint32_t *keys = cv.take<uint32_t>(n);
  uint32_t *count = cv.take<uint32_t>((size_t)n_el * m);
  double *part = cv.take<double>(kBoxBlocks * 6);
  double *grid = cv.take<double>(8);
  double *boxes = cv.take<double>((size_t)std::max<uint32_t>(ntiles, 1) * 6);
  unsigned long long *hist_d =
      device ? reinterpret_cast<unsigned long long *>(hist_out) : cv.take<unsigned long long>(hsize);
  unsigned long long *ovf_d = (device && ovf_out) ? reinterpret_cast<unsigned long long *>(ovf_out)
                                                  : cv.take<unsigned long long>(1);

  const T *xyz = xyz_in;
  if (!device) {
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
    XS_HIP_CHECK(hipMemsetAsync(count, 0, (size_t)n_el * m * sizeof(uint32_t), ctx->stream));
  }
  if (n > 0 && ntypes > 0) {
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, kBoxBlocks);
    hipLaunchKernelGGL((hist_bbox_partial<T>), dim3(nb), dim3(256), 0, ctx->stream, x, y, z, n, part);
    hipLaunchKernelGGL(hist_grid_setup, dim3(1), dim3(256), 0, ctx->stream, part, (int)nb, bits, grid);
    unsigned maxc = 0;
    for (unsigned e = 0; e < n_el; ++e) maxc = std::max(maxc, el_count[e]);
    dim3 eg((maxc + 255) / 256, n_el);
    if (maxc) {
      hipLaunchKernelGGL((hist_morton_count<T>), eg, dim3(256), 0, ctx->stream, x, y, z, g, grid,
                         bits, keys, count);
      hipLaunchKernelGGL(hist_scan, dim3(n_el), dim3(1024), 0, ctx->stream, count, g, m);
      hipLaunchKernelGGL((hist_scatter<T>), eg, dim3(256), 0, ctx->stream, x, y, z, g, bits, keys,
                         count, sxp, syp, szp);
    }
    if (ntiles)
      hipLaunchKernelGGL((hist_tile_boxes<T>), dim3(ntiles), dim3(256), 0, ctx->stream, sxp, syp,
                         szp, g, (int)n_el, boxes);
    XS_HIP_CHECK(hipGetLastError());
  }
  XS_TRY(prof_end(ctx, "hist_prep", ev));

  if (n > 0 && ntypes > 0) {
    const Item *items = nullptr;
    size_t nitems = 0;
    XS_TRY(upload_items(ctx, g, n_el, &items, &nitems));
    if (nitems) {
      Args<T> A{};
      A.x = sxp;
      A.y = syp;
      A.z = szp;
      A.boxes = boxes;
      A.items = items;
      A.hist = hist_d;
      A.overflow = ovf_d;
      A.bin = (T)bin;
      A.bin_d = bin;
      A.nhist = nhist;
      A.g = g;
      unsigned p = 0;
      for (unsigned a = 0; a < n_el; ++a)
        for (unsigned b = a; b < n_el; ++b, ++p) {
          A.ta[p] = (uint8_t)a;
          A.tb[p] = (uint8_t)b;
        }
      if (nhist <= 8192)
        XS_TRY((launch_hist<T, 8192>(ctx, A, nitems)));
      else
        XS_TRY((launch_hist<T, 32768>(ctx, A, nitems)));
    }
  }

  if (!device) {
    XS_HIP_CHECK(hipMemcpyAsync(hist_out, hist_d, hsize * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                ctx->stream));
    uint64_t ov = 0;
    XS_HIP_CHECK(hipMemcpyAsync(&ov, ovf_d, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (ovf_out) *ovf_out = ov;
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
