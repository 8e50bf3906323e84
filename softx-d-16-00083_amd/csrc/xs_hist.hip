// H1: pair-distance histograms on gfx950, for one structure or an ensemble.
//
// Tiling.  Each element's atoms are split into equal-count tiles of kTile
// atoms (all full but one), compact in space and no wider than two tile
// diagonals fit the LDS window: k-d splits on the host (kd_tiles), x-slab /
// y-column / z-run radix sorts on the device (eq_tiles_device), or, when
// those tiles are too wide for the window, the cells of a uniform grid.  A
// tile carries its bounding box.  Counts do not depend on atom order, so the
// permutation leaves the histogram unchanged.
//
// Kernel.  One work item = (pair type of a structure, I unit, run of J
// tiles); an I unit is one tile (FP32 binning) or two neighbouring tiles
// (FP64 binning, two I atoms per thread, so each J atom read from LDS serves
// two pairs).  The I atoms sit in registers; J tiles are sorted by the
// lowest bin their pairs with the I unit can reach (from the boxes) and
// swept in that order; J atoms are staged 256 at a time in LDS and read by
// every lane at the same address.  Counts go to a circular LDS window of
// WCAP uint32 bins that only slides up; bins the remaining J tiles cannot
// reach are flushed to the uint64 histogram with one atomic per non-zero bin,
// so flush traffic per item is at most nhist + WCAP whatever the pair count.
// The window is rotated (rarely) so that a J tile's bins never straddle the
// end of the slot array; the tile's pairs then count at slot k + const.
// A tile pair whose span exceeds the window takes global atomics per pair.
//
// Binning.  The reference's (unsigned)(r / bin), r = sqrt((dx*dx+dy*dy)+dz*dz):
//   double (CalcFunctions.cpp:784-788, the CPU path) with a fast exact path,
//   float  (CalcFunctionsCUDA.cu:832-833, the GPU path) in IEEE single precision.
// Counts equal the oracle's bit for bit in both modes.
//
// Pipelines.
//   xs_hist, device input: the bbox is computed on the GPU and read back; the
//     equal-count sorts run on the GPU and the tiles follow from the counts
//     (one more read-back of the tile boxes when the window caps their width;
//     the cell-grid fallback reads back per-cell counts).
//   xs_hist, host input, and xs_batch_plan: the sort, tiles, boxes and work
//     items are built on the host and uploaded.
//   xs_batch_run (xs_batch.hip): zero fills, histograms, pair sums and PDFs
//     are launches only, optionally captured once into a hipGraph.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "xs_fixbin.h"
#include "xs_hist_plan.h"
#include "xs_internal.h"

namespace xs {
namespace hist {

constexpr int kStage = 256;      // J atoms staged in LDS per step
constexpr int kBoxBlocks = 512;  // partial bounding boxes
constexpr int kChunk = 256;      // FP64 binning: J atoms between barriers (8 KB of records)
constexpr double kFill = 0.94 * kTile;  // target atoms in the fullest cell
constexpr uint32_t kMaxCellsPerAxis = 256;

struct Grid {
  double lo[3];
  double inv_cell[3];  // per axis (0 for a flat axis)
  uint32_t g[3];
};

template <class T>
struct Args {
  const T *x, *y, *z;
  const Tile *tiles;
  const double *boxes;  // 6 per tile: lo xyz, hi xyz
  const TypeDesc *types;
  const Item *items;
  unsigned long long *hist;
  unsigned long long *overflow;
  const double *rec;  // FP64 binning: J records, 4 per atom (hist_tile_records)
  uint32_t quot_f32;  // FP32 binning: the bin passed quot_f32_check_kernel
  uint32_t narrow;    // FP64 binning: every pair type has at most kNarrowMax bins
};

// ---------------------------------------------------------------- prep kernels

// Bounding boxes.  fmin / fmax drop a NaN operand, so a NaN coordinate is
// tracked by its own flag and turns the whole box into NaN, which the host's
// finiteness check (check_span) rejects before any pair kernel runs.
template <class T>
__global__ void hist_bbox_partial(const T *__restrict__ x, const T *__restrict__ y,
                                  const T *__restrict__ z, size_t n, double *__restrict__ part) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  int nan = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double v[3] = {(double)x[i], (double)y[i], (double)z[i]};
    for (int a = 0; a < 3; ++a) {
      nan |= v[a] != v[a];
      lo[a] = fmin(lo[a], v[a]);
      hi[a] = fmax(hi[a], v[a]);
    }
  }
  if (__syncthreads_or(nan)) {
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = NAN;
    return;
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

__global__ void hist_bbox_reduce(const double *__restrict__ part, int nparts,
                                 double *__restrict__ bbox) {
  __shared__ double s[6][256];
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  int nan = 0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x)
    for (int a = 0; a < 3; ++a) {
      const double l = part[i * 6 + a], h = part[i * 6 + 3 + a];
      nan |= (l != l) | (h != h);
      v[a] = fmin(v[a], l);
      v[3 + a] = fmax(v[3 + a], h);
    }
  if (__syncthreads_or(nan)) {
    if (threadIdx.x < 6) bbox[threadIdx.x] = NAN;
    return;
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

// Cell of every atom of one element (keys) and atoms per cell (count).
template <class T>
__global__ void hist_cell_count(const T *__restrict__ x, const T *__restrict__ y,
                                const T *__restrict__ z, uint32_t n, Grid grid,
                                uint32_t *__restrict__ keys, uint32_t *__restrict__ count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v[3] = {(double)x[i], (double)y[i], (double)z[i]};
  uint32_t c[3];
  for (int a = 0; a < 3; ++a) {
    const double f = (v[a] - grid.lo[a]) * grid.inv_cell[a];
    const int ci = isfinite(f) ? (int)f : 0;
    c[a] = (uint32_t)min(max(ci, 0), (int)grid.g[a] - 1);
  }
  const uint32_t key = (c[0] * grid.g[1] + c[1]) * grid.g[2] + c[2];
  keys[i] = key;
  atomicAdd(&count[key], 1u);
}

// Counting-sort scatter of one element: start[cell] is the next free slot.
template <class T>
__global__ void hist_scatter(const T *__restrict__ x, const T *__restrict__ y,
                             const T *__restrict__ z, uint32_t n, const uint32_t *__restrict__ keys,
                             uint32_t *__restrict__ start, T *__restrict__ sx, T *__restrict__ sy,
                             T *__restrict__ sz) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t pos = atomicAdd(&start[keys[i]], 1u);
  sx[pos] = x[i];
  sy[pos] = y[i];
  sz[pos] = z[i];
}

template <class T>
__global__ void hist_tile_boxes(const T *__restrict__ x, const T *__restrict__ y,
                                const T *__restrict__ z, const Tile *__restrict__ tiles,
                                double *__restrict__ boxes) {
  const Tile t = tiles[blockIdx.x];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = threadIdx.x; i < t.cnt; i += blockDim.x) {
    const size_t gi = (size_t)t.off + i;
    const double v[3] = {(double)x[gi], (double)y[gi], (double)z[gi]};
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
  if (threadIdx.x < 6) boxes[(size_t)blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void fill_u64(unsigned long long *__restrict__ p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// ---- equal-count tiles on the device (see eq_tiles_device) ----

__global__ void iota_u32(uint32_t *p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

// Part of rank r when n ranks are split into g equal-count parts, part s
// starting at floor(s n / g).
__device__ inline uint32_t equal_part(uint64_t r, uint64_t n, uint64_t g) {
  uint64_t s = r * g / n;
  while (s + 1 < g && (s + 1) * n / g <= r) ++s;
  while (s > 0 && s * n / g > r) --s;
  return (uint32_t)s;
}

// slab[ord[r]] = equal part of rank r (ord: atoms in x order)
__global__ void eq_slabs(const uint32_t *ord, uint32_t n, uint32_t gx, uint32_t *slab) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) slab[ord[r]] = equal_part(r, n, gx);
}

__global__ void gather_u32(const uint32_t *src, const uint32_t *ord, uint32_t n, uint32_t *dst) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) dst[r] = src[ord[r]];
}

// col[ord[r]] = s gy + equal part of r within slab s (ord: slab-major, y order)
__global__ void eq_columns(const uint32_t *sslab, const uint32_t *ord, uint32_t n, uint32_t gx,
                           uint32_t gy, uint32_t *col) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint64_t s = sslab[r];
  const uint64_t st = s * n / gx, en = (s + 1) * n / gx;
  col[ord[r]] = (uint32_t)s * gy + equal_part(r - st, en - st, gy);
}

template <class T>
__global__ void gather_xyz(const T *x, const T *y, const T *z, const uint32_t *ord, uint32_t n,
                           T *sx, T *sy, T *sz) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint32_t i = ord[r];
  sx[r] = x[i];
  sy[r] = y[i];
  sz[r] = z[i];
}

// ---------------------------------------------------------------- histogram kernel

// Conservative bin range [kl, ku] of all pairs between two boxes: the
// distance bounds scaled by 1 -+ 2^-20 (the FP32 binning's distance errs by
// less than 5 2^-24 relative) and one bin of margin each side.  Every pair's
// bin lies in [kl, ku], which the unchecked window slots rely on.
// fits: no pair can reach nhist (ku was not clamped).
__device__ __forceinline__ void box_bins(const double *__restrict__ bi, const double *__restrict__ bj,
                                         double bin, uint32_t nhist, uint32_t &kl, uint32_t &ku,
                                         bool &fits) {
  double lo2 = 0.0, hi2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double gap = fmax(fmax(bj[a] - bi[3 + a], bi[a] - bj[3 + a]), 0.0);
    const double ext = fmax(bj[3 + a] - bi[a], bi[3 + a] - bj[a]);
    lo2 += gap * gap;
    hi2 += ext * ext;
  }
  const double top = (double)(nhist - 1);
  double fl = floor(sqrt(lo2) / bin * (1.0 - 0x1p-20)) - 1.0;
  double fh = floor(sqrt(hi2) / bin * (1.0 + 0x1p-20)) + 1.0;
  fits = fh <= top;
  fl = fl < 0.0 ? 0.0 : (fl > top ? top : fl);
  fh = fh > top ? top : (fh < fl ? fl : fh);
  kl = (uint32_t)fl;
  ku = (uint32_t)fh;
}

// The window holds bins [B, B + WCAP); bin k sits in slot (k + D) mod WCAP.
// D changes only when the window is rotated (rotate_window).
template <int WCAP>
__device__ __forceinline__ void flush_window(uint32_t *win, uint32_t lo, uint32_t hi, uint32_t D,
                                             unsigned long long *__restrict__ hp) {
  for (uint32_t k = lo + threadIdx.x; k < hi; k += kTile) {
    const uint32_t slot = (k + D) & (WCAP - 1);
    const uint32_t v = win[slot];
    if (v) {
      atomicAdd(&hp[k], (unsigned long long)v);
      win[slot] = 0u;
    }
  }
}

// Rotates the window so that bin B sits in slot 0 (the new D is -B).  Called
// by the whole work group when a J tile's bins would straddle the end of the
// slot array, so that the tile's slots are k + const without a wrap.
template <int WCAP>
__device__ __forceinline__ void rotate_window(uint32_t *win, uint32_t B, uint32_t D) {
  constexpr int kPer = WCAP / kTile;
  uint32_t v[kPer];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < kPer; ++c) v[c] = win[(B + D + threadIdx.x + c * kTile) & (WCAP - 1)];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < kPer; ++c) win[threadIdx.x + c * kTile] = v[c];
  __syncthreads();
}

// Bin index of a squared distance s.
// double: the reference CPU expression (unsigned)(sqrt(s)/bin) in IEEE double.
//   Fast path: sqrt(s)/bin estimated with v_rsq_f64 and one Newton step times
//   1/bin.  The ISA reference guides of CDNA3 and CDNA4 document V_RSQ_F64 to
//   (2**29) ULP, i.e. y = (1 + e)/sqrt(s) with |e| <= 2^29 2^-52 = 2^-23
//   (relative to y's own ulp, which is at most 2^-52 y).  One Newton step
//   g (3 - t)/2, g = s y, t = g y, leaves sqrt(s)(1 + e)(1 - e - e^2/2)
//   = sqrt(s)(1 - 3e^2/2 - e^3/2): relative error below 1.5 2^-46 = 2^-45.42.
//   Its four FP64 roundings (g, t, the FMA, the product) add below 2^-50.4,
//   the IEEE sqrt and division of the reference below 2^-52, so estimate and
//   IEEE quotient differ by less than 2^-45.3 relative, nhist 2^-45.3 bins.
//   (Measured on gfx950 over 2^24 arguments: v_rsq_f64 within 2^-24.2.)
//   When the estimate's fractional part is more than nhist 2^-41 (over 16x
//   that) away from an integer, its floor is the IEEE result.  Otherwise --
//   about nhist 2^-40 of the pairs, and every NaN -- the IEEE sqrt and
//   division are evaluated.  `band` = 1/2 - nhist 2^-41 - 2^-50 (band_for).
//   The estimate stays below 2^31: the host rejects structures whose
//   bounding-box diagonal spans 2^31 bins.
// float: the IEEE single-precision expression of the reference GPU kernel,
//   evaluated directly (correctly rounded sqrtf and division).
// sqrt(s)/bin from v_rsq_f64 and one Newton step with 1/bin folded in:
// y = rsq(s), g = s y, t = g y, estimate = g (1.5 - 0.5 t) / bin
// = g * fma(t, -0.5/bin, 1.5/bin) (four FP64 operations after the rsq).
__device__ __forceinline__ double sqrt_over_bin(double s, double inv_bin) {
  const double y = __builtin_amdgcn_rsq(s);
  const double g = s * y;
  const double t = g * y;
  return g * fma(t, -0.5 * inv_bin, 1.5 * inv_bin);
}

template <class T>
__device__ __forceinline__ uint32_t bin_index(T s, T bin, double inv_bin, double band,
                                              uint32_t nhist);

template <>
__device__ __forceinline__ uint32_t bin_index<double>(double s, double bin, double inv_bin,
                                                      double band, uint32_t nhist) {
  const double kd = sqrt_over_bin(s, inv_bin);
  if (__builtin_expect(fabs(__builtin_amdgcn_fract(kd) - 0.5) < band, 1)) return (uint32_t)kd;
  const double t = __builtin_sqrt(s) / bin;
  return t < (double)nhist ? (uint32_t)t : nhist;
}

// float: (unsigned)(sqrtf(s) / bin) of the reference GPU kernel in IEEE
// single precision.  sqrtf is the IEEE one; the division is evaluated as
// (float)(r * inv_bin) in double, inv_bin = RN(1 / (double)bin).  That
// product lies within 2^-51.9 (relative) of r / bin, while a quotient of two
// floats lies at least 2^-49 from any midpoint m between floats: r - m bin
// is a non-zero multiple of ulp(m) ulp(bin) (m has 25 significant bits and
// an odd last one, so m bin has more than 24), i.e. |r/bin - m| >=
// ulp(m) ulp(bin) / bin >= 2^-49 m.  Rounding the product to float
// therefore gives the IEEE quotient.  inv_bin must be 1 / (double)(float)bin.
// IEEE sqrtf(s) for s >= 2^-96: v_sqrt_f32 is within one ulp there, and
// Tuckerman's rounding test picks the correctly rounded root among y and its
// two float neighbours y-, y+: RN(sqrt s) = y- if s <= y y-, y+ if s > y y+,
// else y (the fma residuals' signs are exact).  The general sqrtf adds a
// rescaling of s below 2^-96 and a zero/inf/NaN select; the s here are
// finite sums of squares, and below 2^-96 the root (< 2^-47) lies far below
// one bin, so the bin index is 0 either way.
__device__ __forceinline__ float sqrt_rn_f32(float s) {
  const float y = __builtin_amdgcn_sqrtf(s);
  const float dn = __int_as_float(__float_as_int(y) - 1);
  const float up = __int_as_float(__float_as_int(y) + 1);
  float r = fmaf(-dn, y, s) <= 0.0f ? dn : y;
  r = fmaf(-up, y, s) > 0.0f ? up : r;
  return r;
}

// sqrtf(s) / bin in IEEE single precision (before the conversion to a bin).
__device__ __forceinline__ float dist_over_bin_f32(float s, double inv_bin) {
  return (float)((double)sqrt_rn_f32(s) * inv_bin);
}

// RN(1 / bin) in single precision (IEEE division).
__device__ __forceinline__ float recip_f32(float bin) { return 1.0f / bin; }

// r / bin in FP32 operations only: q0 = r yb with yb = RN(1 / bin), the
// residual r - q0 bin by one FMA, and Markstein's correction q0 + e yb by
// another.  Three FP32 instructions against the FP64 multiply and two
// conversions of dist_over_bin_f32, which issue at the FP64 rate.  The
// correction is proven to give the IEEE quotient when q0 lies within one ulp
// of r / bin, but q0 can lie up to 1.5 ulp away, so the kernels use this form
// only for bins that quot_f32_check_kernel has passed.
__device__ __forceinline__ float quot_f32(float r, float bin, float yb) {
#pragma clang fp contract(off)
  const float q0 = r * yb;
  return fmaf(fmaf(-q0, bin, r), yb, q0);
}

// Counts the r in [1, 2) (all 2^23 significands, one per thread) for which
// quot_f32 differs from the double-product quotient, which equals the IEEE one
// for every r.  For 2^-60 <= bin <= 2^60 no count means quot_f32 is the IEEE
// quotient for every r the histogram kernels form:
// - q0, the residual and the correction scale exactly with the exponent of r
//   while they stay normal.  That holds wherever r / bin >= 1/4: then q0 is
//   normal, r >= 2^-62, and a non-zero residual, a multiple of
//   ulp(q0) ulp(bin), is at least 2^-109.  So one binade decides them all.
// - Where r / bin < 1/4, q0 stays below 1/4 (1 + 2^-23) and the correction
//   below 2^-24, so both quotients are below 1 and both bins are 0.
// The quotient stays below 2^31 (the host rejects wider spans), far from
// overflow.
__global__ void quot_f32_check_kernel(float bin, double inv_bin, unsigned long long *bad) {
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;  // grid of exactly 2^23 threads
  const float r = __uint_as_float(0x3f800000u | u);
  if (quot_f32(r, bin, recip_f32(bin)) != (float)((double)r * inv_bin)) atomicAdd(bad, 1ull);
}

// Exhaustive check of sqrt_rn_f32 against the IEEE root, one pass over every
// positive finite float s.  y is the IEEE root iff m-^2 < s < m+^2 with m-,
// m+ the midpoints between y and its float neighbours (a float s is never a
// midpoint's square: that square has more than 24 significant bits).  m+-
// have 25 significant bits and y >= 2^-75, so m+-^2 is exact in double and
// the test is exact.  bad[0] counts wrong roots for s >= 2^-96; bad[1] counts
// s < 2^-96 whose result is not in [0, 2^-47); bad[2] wrong roots for
// s < 2^-96.  The FP32 binning relies on bad[0] = bad[1] = 0 and, unless
// bad[2] = 0 too, on bins of at least 2^-40 (then every s < 2^-96 bins to 0
// with the IEEE root and with sqrt_rn_f32 alike).
__global__ void sqrt_f32_check_kernel(unsigned long long *bad) {
  unsigned long long b0 = 0, b1 = 0, b2 = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t u = 1u + blockIdx.x * blockDim.x + threadIdx.x; u < 0x7f800000u; u += stride) {
    const float s = __uint_as_float(u);
    const float y = sqrt_rn_f32(s);
    const uint32_t yb = __float_as_uint(y);
    bool ok = false;
    if (y > 0.0f && y < INFINITY) {
      const double yd = y;
      const double lo = 0.5 * (yd + (double)__uint_as_float(yb - 1u));
      const double hi = 0.5 * (yd + (double)__uint_as_float(yb + 1u));
      ok = lo * lo < (double)s && (double)s < hi * hi;
    }
    if (u >= 0x0f800000u) {  // s >= 2^-96
      b0 += !ok;
    } else {
      b1 += !(y >= 0.0f && y < 0x1p-47f);
      b2 += !ok;
    }
  }
  if (b0) atomicAdd(&bad[0], b0);
  if (b1) atomicAdd(&bad[1], b1);
  if (b2) atomicAdd(&bad[2], b2);
}

template <>
__device__ __forceinline__ uint32_t bin_index<float>(float s, float bin, double inv_bin,
                                                     double band, uint32_t nhist) {
  (void)bin;
  (void)band;
  const float t = dist_over_bin_f32(s, inv_bin);
  return t < (float)nhist ? (uint32_t)t : nhist;
}

// Fast part of the FP64 bin_index: the estimate's floor and whether it may
// differ from the IEEE result (then exact_bin decides).
__device__ __forceinline__ uint32_t fast_bin(double s, double inv_bin, double band, bool &bad) {
  const double kd = sqrt_over_bin(s, inv_bin);
  bad = !(fabs(__builtin_amdgcn_fract(kd) - 0.5) < band);
  return (uint32_t)kd;
}

// The same fast path in fixed point (xs_fixbin.h) for grids of at most
// kNarrowMax bins: T's high word kFixHi + bin, its low word L; the estimate
// is floored directly when L > FixBand::lm1.
__device__ __forceinline__ uint32_t fix_bin(double s, double inv_bin, double m, uint32_t &lo) {
  return fix_from_rsq(s, __builtin_amdgcn_rsq(s), inv_bin, m, lo);
}

template <class T>
__device__ __forceinline__ uint32_t exact_bin(T s, T bin, uint32_t nhist) {
  const T t = (T)__builtin_sqrt((double)s) / bin;
  return t < (T)nhist ? (uint32_t)t : nhist;
}

// An atom index the compiler cannot see through, so that coordinates read
// again at a later point are not merged with an earlier read (which would
// keep them in registers in between).
__device__ __forceinline__ size_t opaque_index(size_t g) {
  asm volatile("" : "+v"(g));
  return g;
}

// FP64 fast path of one I atom: its coordinates relative to the centre c of
// a J tile's box times -2, and its squared distance from c.  With the J
// atoms' coordinates relative to c (rj') and their squared norms (nj) in the
// J records, |ri - rj|^2 = n + nj + (-2 ri').rj' costs one add and three
// FMAs.
struct IAtom {
  double ax, ay, az, n;
};

// Centre c of a J tile's box (bi, 6 doubles) and the half-width of the FP64
// fast path's band for pairs formed relative to c.  The bound below is
// symmetric in the two atoms, so it holds with the J box as the one of
// radius rho.
// Relative to IEEE sqrt(s)/bin the estimate errs by (a) below nhist 2^-45.3
// bins from v_rsq_f64, the Newton step and the roundings of the reference
// expression, and (b) below u (64 R^2 + 36 R + 10 T) bins from forming s as
// |ri'|^2 + |rj'|^2 - 2 ri'.rj' with shifted, rounded coordinates, where
// u = 2^-53, R bounds |ri - c| in bins and T = distance in bins <= nhist + 1
// where it matters: the rounding error of s is below
// u (32 rho^2 + 36 rho d + 10 d^2) for |ri - c| <= rho, |rj - c| <= rho + d,
// and divides by d to give the error of the root for d >= bin / 2 (below
// that, both floors are 0).  Both bounds are taken 16-fold.
__device__ __forceinline__ double tile_band(const double *bi, double band0, double inv_bin,
                                            uint32_t nhist, double c[3]) {
  double h2 = 0.0, cm = 0.0;
  for (int a = 0; a < 3; ++a) {
    c[a] = 0.5 * (bi[a] + bi[3 + a]);
    const double e = bi[3 + a] - bi[a];
    h2 = fma(e, e, h2);
    cm = fmax(cm, fabs(c[a]));
  }
  const double R = (0.5 * sqrt(h2) * (1.0 + 0x1p-20) + 0x1p-40 * cm) * inv_bin;
  return band0 - 0x1p-49 * (fma(64.0 * R, R, 36.0 * R) + 10.0 * ((double)nhist + 1.0));
}

// FP32 groups of 8 J atoms (not the like-element diagonal tile): the
// reference GPU expression for 8 independent pairs, no per-pair branch.
// FQ: the quotient by bin as quot_f32 (yb = recip_f32(bin)), else as the
// rounded double product.
// SAFE: the J tile's bin range lies inside the window and below nhist
// (kl >= B, ku < B + WCAP, ku < nhist from the tile boxes) and maps to slots
// without a wrap, so every pair goes to slot k + off unchecked; otherwise
// each pair is tested.
template <int WCAP, bool SAFE, bool FQ>
__device__ __forceinline__ void pair_group8_f32(float xi, float yi, float zi, const float *sx,
                                                const float *sy, const float *sz, uint32_t ng,
                                                float bin, double inv_bin, float yb, uint32_t nhist,
                                                uint32_t B, uint32_t wlim, uint32_t D, uint32_t off,
                                                uint32_t *win, unsigned long long *__restrict__ hp,
                                                unsigned long long *__restrict__ ovf) {
#pragma clang fp contract(off)
  for (uint32_t j0 = 0; j0 < ng; j0 += 8) {
    uint32_t k[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const float dx = xi - sx[j0 + p];
      const float dy = yi - sy[j0 + p];
      const float dz = zi - sz[j0 + p];
      float s = dx * dx + dy * dy;
      s = s + dz * dz;
      // no clamp: a bin at or past nhist (t >= nhist, converted with
      // saturation) goes to the overflow counter below either way
      if constexpr (FQ)
        k[p] = (uint32_t)quot_f32(sqrt_rn_f32(s), bin, yb);
      else
        k[p] = (uint32_t)dist_over_bin_f32(s, inv_bin);
    }
    if constexpr (SAFE) {
#pragma unroll
      for (int p = 0; p < 8; ++p) atomicAdd(&win[k[p] + off], 1u);
    } else {
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        if (k[p] - B < wlim)
          atomicAdd(&win[(k[p] + D) & (WCAP - 1)], 1u);
        else if (k[p] < nhist)
          atomicAdd(&hp[k[p]], 1ull);
        else
          atomicAdd(ovf, 1ull);
      }
    }
  }
}

template <class T, int WCAP, bool DIAG>
__device__ __forceinline__ void pair_block(T xi, T yi, T zi, uint32_t il, const T *sx,
                                           const T *sy, const T *sz, uint32_t cn,
                                           uint32_t jbase, T bin, double inv_bin, double band,
                                           uint32_t nhist, uint32_t B, uint32_t wlim, uint32_t D,
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
    const uint32_t k = bin_index<T>(s, bin, inv_bin, band, nhist);
    if (__builtin_expect(k - B < wlim, 1)) {
      atomicAdd(&win[(k + D) & (WCAP - 1)], 1u);
    } else if (k < nhist) {
      atomicAdd(&hp[k], 1ull);
    } else {
      atomicAdd(ovf, 1ull);
    }
  }
}

// FP32 binning: one work item = (pair type, I tile, run of J tiles), one I
// atom per thread.
template <int WCAP, bool FQ>
__global__ __launch_bounds__(kTile) void hist_kernel_f32(Args<float> A) {
  using T = float;
  __shared__ uint32_t win[WCAP];
  __shared__ unsigned long long keys[kSortCap];
  __shared__ T sx[kStage], sy[kStage], sz[kStage];

  const Item it = A.items[blockIdx.x];
  const TypeDesc td = A.types[it.type];
  const uint32_t tid = threadIdx.x;
  const uint32_t tiI = td.a_base + it.ti;
  const Tile I = A.tiles[tiI];
  const bool ivalid = tid < I.cnt;
  T xi = 0, yi = 0, zi = 0;
  if (ivalid) {
    const size_t gi = (size_t)I.off + tid;
    xi = A.x[gi];
    yi = A.y[gi];
    zi = A.z[gi];
  }
  const T bin = (T)td.bin;
  const double inv_bin32 = 1.0 / (double)bin;  // for bin_index<float>
  const float yb = recip_f32(bin);              // for quot_f32
  const uint32_t nhist = td.nhist;
  const double *bi = A.boxes + 6 * (size_t)tiI;
  const uint32_t n = it.jcount;
  uint32_t npow = 1;
  while (npow < n) npow <<= 1;
  unsigned long long *hp = A.hist + td.hist_off;
  unsigned long long *ovf = A.overflow + td.ovf;

  for (uint32_t m = tid; m < npow; m += kTile) {
    unsigned long long key = ~0ull;
    if (m < n) {
      const uint32_t tj = (it.jstart + m) % td.b_n;
      uint32_t kl, ku;
      bool fits;
      box_bins(bi, A.boxes + 6 * (size_t)(td.b_base + tj), td.bin, nhist, kl, ku, fits);
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
  uint32_t D = 0u;  // bin k is counted in slot (k + D) mod WCAP
  for (uint32_t s = 0; s < n; ++s) {
    const uint32_t m = (uint32_t)(keys[s] & 0xffffffffu);
    const uint32_t tj = (it.jstart + m) % td.b_n;
    const uint32_t tjg = td.b_base + tj;
    uint32_t kl, ku;
    bool fits;
    box_bins(bi, A.boxes + 6 * (size_t)tjg, td.bin, nhist, kl, ku, fits);
    uint32_t wlim = 0;
    const bool span_ok = ku - kl + 1 <= (uint32_t)WCAP;
    if (span_ok) {
      const uint32_t nb = ku + 1 > (uint32_t)WCAP ? ku + 1 - WCAP : 0u;
      if (nb > B) {
        flush_window<WCAP>(win, B, min(min(nb, B + (uint32_t)WCAP), nhist), D, hp);
        B = nb;
        __syncthreads();
      }
      wlim = min((uint32_t)WCAP, nhist - B);
    }
    // unchecked slots k + off for the tile's bins [kl, ku] (inside [B, B + WCAP))
    const bool safe = span_ok && fits;
    uint32_t off = 0u;
    if (safe) {
      if (((kl + D) ^ (ku + D)) >= (uint32_t)WCAP) {
        rotate_window<WCAP>(win, B, D);
        D = 0u - B;
      }
      off = D - ((kl + D) & ~(uint32_t)(WCAP - 1));
    }
    const Tile J = A.tiles[tjg];
    const bool diag = td.same && tj == it.ti;
    for (uint32_t s0 = 0; s0 < J.cnt; s0 += kStage) {
      const uint32_t cn = min((uint32_t)kStage, J.cnt - s0);
      if (tid < cn) {
        const size_t g = (size_t)J.off + s0 + tid;
        sx[tid] = A.x[g];
        sy[tid] = A.y[g];
        sz[tid] = A.z[g];
      }
      __syncthreads();
      if (ivalid) {
        if (diag) {
          pair_block<T, WCAP, true>(xi, yi, zi, tid, sx, sy, sz, cn, s0, bin, inv_bin32, td.band,
                                    nhist, B, wlim, D, win, hp, ovf);
        } else {
          const uint32_t ng = cn & ~7u;
          if (safe)
            pair_group8_f32<WCAP, true, FQ>(xi, yi, zi, sx, sy, sz, ng, bin, inv_bin32, yb, nhist, B,
                                            wlim, D, off, win, hp, ovf);
          else
            pair_group8_f32<WCAP, false, FQ>(xi, yi, zi, sx, sy, sz, ng, bin, inv_bin32, yb, nhist, B,
                                             wlim, D, off, win, hp, ovf);
          pair_block<T, WCAP, false>(xi, yi, zi, tid, sx + ng, sy + ng, sz + ng, cn - ng, s0 + ng, bin,
                                     inv_bin32, td.band, nhist, B, wlim, D, win, hp, ovf);
        }
      }
      __syncthreads();
    }
  }
  flush_window<WCAP>(win, B, min(B + (uint32_t)WCAP, nhist), D, hp);
}

// Loads through the constant address space: with a uniform address they
// become scalar loads into SGPRs (the J records below), read-only data.
typedef const double __attribute__((address_space(4))) cdouble;

// FP64 groups of 4 J atoms against the thread's two I atoms a and b (8
// independent chains; each J record serves two pairs).  The J records (x',
// y', z', |r'|^2, relative to the centre of the J tile's box; rec) are read
// with scalar loads: the address is the same for every lane, so they take
// SGPRs and no LDS.  Counts of an I atom that does not exist (va / vb false)
// are dropped.  The exact evaluation reads the coordinates again (xg, yg, zg
// at ga, gb for the I atoms, jg + j for the J atoms).
// NARROW: the fixed-point fast path (fix_bin; nhist <= kNarrowMax), k[] then
// holding kFixHi + bin.
template <int WCAP, bool SAFE, bool NARROW>
__device__ __forceinline__ void pair_group4x2(const IAtom &ia, const IAtom &ib, bool va, bool vb,
                                              const double *xg, const double *yg,
                                              const double *zg, size_t ga, size_t gb, size_t jg,
                                              const double *rec, uint32_t ng, double bin,
                                              double inv_bin, double band, const FixBand &fb,
                                              uint32_t nhist, uint32_t B, uint32_t wlim,
                                              uint32_t D, uint32_t off, uint32_t *win,
                                              unsigned long long *__restrict__ hp,
                                              unsigned long long *__restrict__ ovf) {
#pragma clang fp contract(off)
  constexpr uint32_t kb = NARROW ? kFixHi : 0u;  // offset of k[] from the bin
  // (opaque, so that the slot address stays one shift-add of k[])
  uint32_t offk = off - kb;
  asm volatile("" : "+v"(offk));
  // SAFE: every lane increments for every pair, by 1 or, for an I atom that
  // does not exist (a stand-in whose bins lie in the J tile's range), by 0
  const uint32_t inca = va, incb = vb;
  cdouble *jr = (cdouble *)rec;
  for (uint32_t j0 = 0; j0 < ng; j0 += 4) {
    uint32_t k[8];
    bool bad[8], any = false;
    uint32_t lo[8];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const double rx = jr[4 * (j0 + p)], ry = jr[4 * (j0 + p) + 1], rz = jr[4 * (j0 + p) + 2],
                   rn = jr[4 * (j0 + p) + 3];
      double s = ia.n + rn;
      s = __builtin_fma(ia.az, rz, s);
      s = __builtin_fma(ia.ay, ry, s);
      s = __builtin_fma(ia.ax, rx, s);
      double t = ib.n + rn;
      t = __builtin_fma(ib.az, rz, t);
      t = __builtin_fma(ib.ay, ry, t);
      t = __builtin_fma(ib.ax, rx, t);
      if constexpr (NARROW) {
        k[p] = fix_bin(s, inv_bin, fb.m, lo[p]);
        k[4 + p] = fix_bin(t, inv_bin, fb.m, lo[4 + p]);
      } else {
        k[p] = fast_bin(s, inv_bin, band, bad[p]);
        k[4 + p] = fast_bin(t, inv_bin, band, bad[4 + p]);
        any |= bad[p] | bad[4 + p];
      }
    }
    if constexpr (NARROW) {
      // one test of the smallest low word; then all 8 pairs take the exact
      // path (its result equals the fast one wherever that one holds)
      const uint32_t m = min(min(min(lo[0], lo[1]), min(lo[2], lo[3])),
                             min(min(lo[4], lo[5]), min(lo[6], lo[7])));
      any = m <= fb.lm1;
#pragma unroll
      for (int p = 0; p < 8; ++p) bad[p] = true;
    }
    if (__builtin_expect(any, 0)) {
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        if (bad[p]) {
          const size_t g = opaque_index(p < 4 ? ga : gb), h = jg + j0 + (p & 3);
          const double dx = xg[g] - xg[h];
          const double dy = yg[g] - yg[h];
          const double dz = zg[g] - zg[h];
          double s = dx * dx + dy * dy;
          s = s + dz * dz;
          k[p] = exact_bin<double>(s, bin, nhist) + kb;
        }
      }
    }
    if constexpr (SAFE) {
      // slot k - kb + off, a window index (32-bit wrap-around)
#pragma unroll
      for (int p = 0; p < 8; ++p) atomicAdd(&win[k[p] + offk], p < 4 ? inca : incb);
    } else {
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        if (!(p < 4 ? va : vb)) continue;
        const uint32_t kp = k[p] - kb;
        if (kp - B < wlim)
          atomicAdd(&win[(kp + D) & (WCAP - 1)], 1u);
        else if (kp < nhist)
          atomicAdd(&hp[kp], 1ull);
        else
          atomicAdd(ovf, 1ull);
      }
    }
  }
}

// I atom relative to a centre c, for the fast path.
__device__ __forceinline__ IAtom centred(double x, double y, double z, const double c[3]) {
  const double u = x - c[0], v = y - c[1], w = z - c[2];
  return IAtom{-2.0 * u, -2.0 * v, -2.0 * w, fma(w, w, fma(v, v, u * u))};
}

// Records of the FP64 kernel's J atoms: each atom relative to the centre c
// of its tile's box (the centre tile_band takes), (x - c, y - c, z - c,
// |r - c|^2), four doubles per atom in the sorted order.
__global__ void hist_tile_records(const double *__restrict__ x, const double *__restrict__ y,
                                  const double *__restrict__ z, const Tile *__restrict__ tiles,
                                  const double *__restrict__ boxes, double *__restrict__ rec) {
  const Tile t = tiles[blockIdx.x];
  const double *b = boxes + 6 * (size_t)blockIdx.x;
  double c[3];
  for (int a = 0; a < 3; ++a) c[a] = 0.5 * (b[a] + b[3 + a]);
  for (uint32_t i = threadIdx.x; i < t.cnt; i += blockDim.x) {
    const size_t g = (size_t)t.off + i;
    const double u = x[g] - c[0], v = y[g] - c[1], w = z[g] - c[2];
    double *r = rec + 4 * g;
    r[0] = u;
    r[1] = v;
    r[2] = w;
    r[3] = fma(w, w, fma(v, v, u * u));
  }
}

// FP64 binning: one work item = (pair type, I tile pair {2 ti, 2 ti + 1},
// run of J tiles), two I atoms per thread (atom tid of each tile).  The J
// tiles are sorted and the window managed for the union box of the two I
// tiles.  Distances are formed relative to the centre of each J tile's box:
// the J records hold the J atoms relative to it (A.rec), and the I atoms are
// taken relative to it once per J tile.  Like-element types visit tile pairs
// by a half rotation over I tile pairs (add_items), so within the item's own
// tile pair the J tile 2 ti pairs with I tile 2 ti only (j > i) and J tile
// 2 ti + 1 with I tile 2 ti fully and with 2 ti + 1 for j > i; those two J
// tiles take the per-pair loop.  No J atoms are staged in LDS, so the work
// group synchronises only where the window moves.
template <int WCAP, bool NARROW>
__global__ __launch_bounds__(kTile) void hist_kernel_f64(Args<double> A) {
  __shared__ uint32_t win[WCAP];
  __shared__ unsigned long long keys[kSortCap];

  const Item it = A.items[blockIdx.x];
  const TypeDesc td = A.types[it.type];
  const uint32_t tid = threadIdx.x;
  const uint32_t tA = 2 * it.ti, tB = tA + 1;
  const bool hasB = tB < td.a_n;
  const Tile IA = A.tiles[td.a_base + tA];
  const Tile IB = hasB ? A.tiles[td.a_base + tB] : Tile{0u, 0u};
  if (IA.cnt == 0) return;
  const bool va = tid < IA.cnt, vb = tid < IB.cnt;
  // an I atom that does not exist stands in as the first atom of tile tA:
  // its counts are dropped, and its bins lie inside every J tile's range
  const size_t ga = (size_t)IA.off + (va ? tid : 0u);
  const size_t gb = vb ? (size_t)IB.off + tid : (size_t)IA.off;
  const double xa = A.x[ga], ya = A.y[ga], za = A.z[ga];
  const double xb = A.x[gb], yb = A.y[gb], zb = A.z[gb];
  const uint32_t nhist = td.nhist;
  const double bin = td.bin;
  // union box of the two I tiles
  double bu[6];
  {
    const double *b0 = A.boxes + 6 * (size_t)(td.a_base + tA);
    for (int a = 0; a < 6; ++a) bu[a] = b0[a];
    if (hasB) {
      const double *b1 = A.boxes + 6 * (size_t)(td.a_base + tB);
      for (int a = 0; a < 3; ++a) {
        bu[a] = fmin(bu[a], b1[a]);
        bu[3 + a] = fmax(bu[3 + a], b1[3 + a]);
      }
    }
  }
  const uint32_t n = it.jcount;
  uint32_t npow = 1;
  while (npow < n) npow <<= 1;
  unsigned long long *hp = A.hist + td.hist_off;
  unsigned long long *ovf = A.overflow + td.ovf;

  for (uint32_t m = tid; m < npow; m += kTile) {
    unsigned long long key = ~0ull;
    if (m < n) {
      const uint32_t tj = (it.jstart + m) % td.b_n;
      uint32_t kl, ku;
      bool fits;
      box_bins(bu, A.boxes + 6 * (size_t)(td.b_base + tj), bin, nhist, kl, ku, fits);
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
  uint32_t D = 0u;  // bin k is counted in slot (k + D) mod WCAP
  for (uint32_t s = 0; s < n; ++s) {
    const uint32_t m = (uint32_t)(keys[s] & 0xffffffffu);
    const uint32_t tj = (it.jstart + m) % td.b_n;
    const uint32_t tjg = td.b_base + tj;
    const double *bj = A.boxes + 6 * (size_t)tjg;
    uint32_t kl, ku;
    bool fits;
    box_bins(bu, bj, bin, nhist, kl, ku, fits);
    uint32_t wlim = 0;
    const bool span_ok = ku - kl + 1 <= (uint32_t)WCAP;
    if (span_ok) {
      const uint32_t nb = ku + 1 > (uint32_t)WCAP ? ku + 1 - WCAP : 0u;
      if (nb > B) {
        __syncthreads();
        flush_window<WCAP>(win, B, min(min(nb, B + (uint32_t)WCAP), nhist), D, hp);
        B = nb;
        __syncthreads();
      }
      wlim = min((uint32_t)WCAP, nhist - B);
    }
    const bool safe = span_ok && fits;
    uint32_t off = 0u;
    if (safe) {
      if (((kl + D) ^ (ku + D)) >= (uint32_t)WCAP) {
        rotate_window<WCAP>(win, B, D);
        D = 0u - B;
      }
      off = D - ((kl + D) & ~(uint32_t)(WCAP - 1));
    }
    const Tile J = A.tiles[tjg];
    const double *jx = A.x + J.off, *jy = A.y + J.off, *jz = A.z + J.off;
    // like-element types: J tiles of the item's own tile pair.  The I
    // coordinates are read again where the per-pair loop needs them, so that
    // they take no registers in the grouped loop.
    const bool jA = td.same && tj == tA, jB = td.same && hasB && tj == tB;
    if (jA) {
      const size_t g = opaque_index(ga);
      if (va)
        pair_block<double, WCAP, true>(A.x[g], A.y[g], A.z[g], tid, jx, jy, jz, J.cnt, 0u, bin,
                                       td.inv_bin, td.band, nhist, B, wlim, D, win, hp, ovf);
    } else if (jB) {
      const size_t g = opaque_index(ga), h = opaque_index(gb);
      if (va)
        pair_block<double, WCAP, false>(A.x[g], A.y[g], A.z[g], tid, jx, jy, jz, J.cnt, 0u, bin,
                                        td.inv_bin, td.band, nhist, B, wlim, D, win, hp, ovf);
      if (vb)
        pair_block<double, WCAP, true>(A.x[h], A.y[h], A.z[h], tid, jx, jy, jz, J.cnt, 0u, bin,
                                       td.inv_bin, td.band, nhist, B, wlim, D, win, hp, ovf);
    } else {
      double c[3];
      const double band = tile_band(bj, td.band, td.inv_bin, nhist, c);
      const FixBand fb = fix_band(band);
      const IAtom ia = centred(xa, ya, za, c), ib = centred(xb, yb, zb, c);
      const uint32_t ng = J.cnt & ~3u;
      // chunks of kChunk J atoms with a barrier between them: the work
      // group's waves read the same records, which the scalar cache then
      // serves once rather than once per wave
      for (uint32_t c0 = 0; c0 < ng; c0 += kChunk) {
        const uint32_t cn = min((uint32_t)kChunk, ng - c0);
        const size_t jg = (size_t)J.off + c0;
        const double *rec = A.rec + 4 * jg;
        if (va || vb) {
          if (safe)
            pair_group4x2<WCAP, true, NARROW>(ia, ib, va, vb, A.x, A.y, A.z, ga, gb, jg, rec, cn,
                                              bin, td.inv_bin, band, fb, nhist, B, wlim, D, off,
                                              win, hp, ovf);
          else
            pair_group4x2<WCAP, false, NARROW>(ia, ib, va, vb, A.x, A.y, A.z, ga, gb, jg, rec, cn,
                                               bin, td.inv_bin, band, fb, nhist, B, wlim, D, off,
                                               win, hp, ovf);
        }
        if (c0 + kChunk < ng) __syncthreads();
      }
      const size_t g = opaque_index(ga), h = opaque_index(gb);
      if (va)
        pair_block<double, WCAP, false>(A.x[g], A.y[g], A.z[g], tid, jx + ng, jy + ng, jz + ng,
                                        J.cnt - ng, ng, bin, td.inv_bin, td.band, nhist, B, wlim,
                                        D, win, hp, ovf);
      if (vb)
        pair_block<double, WCAP, false>(A.x[h], A.y[h], A.z[h], tid, jx + ng, jy + ng, jz + ng,
                                        J.cnt - ng, ng, bin, td.inv_bin, td.band, nhist, B, wlim,
                                        D, win, hp, ovf);
    }
    __syncthreads();
  }
  flush_window<WCAP>(win, B, min(B + (uint32_t)WCAP, nhist), D, hp);
}

// ---------------------------------------------------------------- host planning

// Cell-grid fallback (device input whose equal-count tiles would be wider
// than the window allows).  Cell sizing: lanes of a work group hold the
// atoms of one I tile, so a tile should be nearly full: cells are sized for
// kFill atoms in the fullest cell.  The first guess assumes uniform density over the bounding box; one
// refinement rescales the edge by cbrt(kFill / fullest count) after counting
// (spheres are ~2x denser than their box).  When the histogram exceeds the
// window, the edge is capped at (WCAP - 8) bin / (2 sqrt 3) so that the bin
// span between any two cells fits the window.
static Grid grid_with_cell(const double bbox[6], double cell, double cap) {
  Grid gr{};
  double ext[3];
  for (int a = 0; a < 3; ++a) {
    gr.lo[a] = bbox[a];
    ext[a] = std::max(bbox[3 + a] - bbox[a], 0.0);
  }
  const double span = std::max({ext[0], ext[1], ext[2], 1e-6});
  cell = std::min(cell, cap);
  cell = std::max(cell, span / kMaxCellsPerAxis);
  cell = std::max(cell, 1e-9);
  for (int a = 0; a < 3; ++a) {
    gr.g[a] = (uint32_t)std::min<double>(kMaxCellsPerAxis,
                                         std::max(1.0, std::ceil(ext[a] / cell - 1e-9)));
    gr.inv_cell[a] = ext[a] > 0.0 ? gr.g[a] / ext[a] : 0.0;
  }
  return gr;
}

static double cell_cap(double bin, uint32_t nhist, uint32_t wcap) {
  return nhist > wcap ? (wcap - 8) * bin / (2.0 * std::sqrt(3.0)) : INFINITY;
}

static double first_cell(const double bbox[6], unsigned max_count, double cap) {
  double ext[3], vol = 1.0;
  for (int a = 0; a < 3; ++a) ext[a] = std::max(bbox[3 + a] - bbox[a], 0.0);
  const double span = std::max({ext[0], ext[1], ext[2], 1e-6});
  const double spacing = span / std::max(1.0, std::cbrt((double)std::max(max_count, 1u)));
  for (int a = 0; a < 3; ++a) vol *= std::max(ext[a], spacing);
  return std::min(std::cbrt(kFill * vol / std::max(max_count, 1u)), cap);
}

// New edge from the fullest cell's count, or 0 when the current one is good.
static double refined_cell(double cell, uint32_t fullest, double cap) {
  if (fullest == 0) return 0.0;
  const bool over = fullest > (uint32_t)kTile;
  const bool under = fullest < 0.75 * kFill && cell < cap;
  if (!over && !under) return 0.0;
  // aim the fullest cell just below a full tile when shrinking, at kFill when growing
  const double target = over ? 0.99 * kTile : kFill;
  return std::min(cell * std::cbrt(target / fullest), cap);
}

// Tiles of one cell: ceil(c / kTile) chunks of near-equal size.
static void cell_tiles(uint32_t pos, uint32_t c, std::vector<Tile> &tiles) {
  if (!c) return;
  const uint32_t nt = (c + kTile - 1) / kTile;
  for (uint32_t t = 0, o = 0; t < nt; ++t) {
    const uint32_t len = (c - o) / (nt - t);
    tiles.push_back(Tile{pos + o, len});
    o += len;
  }
}

// Bounding box of n atoms (SoA xyz) on the host.  std::min / std::max drop a
// NaN, so NaN is tested for explicitly and makes the whole box NaN, which
// check_span rejects.
template <class T>
static void host_bbox(const T *xyz, size_t n, double bbox[6]) {
  bool nan = false;
  for (int a = 0; a < 3; ++a) {
    double lo = INFINITY, hi = -INFINITY;
    for (size_t i = 0; i < n; ++i) {
      const double v = (double)xyz[a * n + i];
      nan |= v != v;
      lo = std::min(lo, v);
      hi = std::max(hi, v);
    }
    bbox[a] = lo;
    bbox[3 + a] = hi;
  }
  if (nan)
    for (int a = 0; a < 6; ++a) bbox[a] = NAN;
}

static int check_span(const double bbox[6], double bin) {
  for (int a = 0; a < 6; ++a)
    if (!std::isfinite(bbox[a])) return fail(XS_ERR_ARG, "histogram: non-finite coordinates");
  double d2 = 0;
  for (int a = 0; a < 3; ++a) d2 += (bbox[3 + a] - bbox[a]) * (bbox[3 + a] - bbox[a]);
  if (std::sqrt(d2) / bin >= 2147483648.0)
    return fail(XS_ERR_RANGE, "histogram: the structure spans more than 2^31 bins");
  return XS_OK;
}

// Equal-count k-d tiles of one element (indices idx[0, ne) into x, y, z): the
// longest axis of a part's bounding box is split at a multiple of `unit`
// atoms (nth_element) until a part holds at most `unit` atoms and spans at
// most `cap` on every axis.  Every part of an element but one is full, so the
// lanes of the I tiles are busy whatever the structure's shape (grid cells
// over a sphere leave about half of them idle), and parts are compact boxes.
// unit = kTile: the parts are the tiles.  unit = 2 kTile (paired, for the
// FP64 kernel's I tile pairs): each part of two or more atoms is halved along
// its longest axis into tiles 2u and 2u + 1.
// Leaves come out depth first, left part first (neighbouring tiles are close).
template <class T>
static void kd_tiles(const T *x, const T *y, const T *z, uint32_t *idx, uint32_t ne, double cap,
                     bool paired, std::vector<std::pair<uint32_t, uint32_t>> &leaves) {
  const T *c[3] = {x, y, z};
  const uint32_t unit = paired ? 2u * kTile : (uint32_t)kTile;
  auto longest = [&](uint32_t lo, uint32_t hi) {
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = lo; i < hi; ++i)
      for (int a = 0; a < 3; ++a) {
        const double v = (double)c[a][idx[i]];
        mn[a] = std::min(mn[a], v);
        mx[a] = std::max(mx[a], v);
      }
    int ax = 0;
    for (int a = 1; a < 3; ++a)
      if (mx[a] - mn[a] > mx[ax] - mn[ax]) ax = a;
    return std::make_pair(ax, mx[ax] - mn[ax]);
  };
  auto split = [&](uint32_t lo, uint32_t hi, uint32_t left, int ax) {
    const T *ca = c[ax];
    std::nth_element(idx + lo, idx + lo + left, idx + hi,
                     [ca](uint32_t a, uint32_t b) { return ca[a] < ca[b]; });
  };
  std::vector<std::pair<uint32_t, uint32_t>> stack{{0u, ne}};
  while (!stack.empty()) {
    const auto part = stack.back();
    stack.pop_back();
    const uint32_t lo = part.first, n = part.second - part.first;
    if (!n) continue;
    const auto ax = longest(lo, part.second);
    if (n == 1 || (n <= unit && !(ax.second > cap))) {
      if (paired && n >= 2) {
        const uint32_t h = n - n / 2;
        split(lo, part.second, h, ax.first);
        leaves.push_back({lo, h});
        leaves.push_back({lo + h, n - h});
      } else {
        leaves.push_back({lo, n});
      }
      continue;
    }
    const uint32_t nl = std::max<uint32_t>(2u, (n + unit - 1) / unit);
    uint32_t left = (nl / 2) * unit;
    if (left >= n) left = n / 2;
    split(lo, part.second, left, ax.first);
    stack.push_back({lo + left, part.second});
    stack.push_back({lo, lo + left});
  }
}

// Host-side sort of one structure: its atoms in tile order into sx, sy, sz
// (n each), its tiles (equal-count k-d tiles per element, atom offsets from
// atom_base) and their boxes appended to tiles / boxes; element e's tiles
// are [ebase[e], ebase[e] + en[e]) counted from the first tile appended.
// xyz: SoA of the structure, element-major.
template <class T>
static void sort_structure_host(const T *xyz, size_t n, const unsigned *el_count, unsigned n_el,
                                double bin, uint32_t nhist, T *sx, T *sy, T *sz, uint32_t atom_base,
                                std::vector<Tile> &tiles, std::vector<double> &boxes, uint32_t *ebase,
                                uint32_t *en) {
  const double cap = cell_cap(bin, nhist, wcap_for(nhist));
  std::vector<uint32_t> idx;
  std::vector<std::pair<uint32_t, uint32_t>> leaves;
  const size_t tile0 = tiles.size();
  size_t off = 0;
  for (unsigned e = 0; e < n_el; ++e) {
    const uint32_t ne = el_count[e];
    idx.resize(ne);
    for (uint32_t i = 0; i < ne; ++i) idx[i] = i;
    leaves.clear();
    kd_tiles<T>(xyz + off, xyz + n + off, xyz + 2 * n + off, idx.data(), ne, cap, sizeof(T) == 8,
                leaves);
    ebase[e] = (uint32_t)(tiles.size() - tile0);
    for (const auto &l : leaves) tiles.push_back(Tile{(uint32_t)(atom_base + off + l.first), l.second});
    en[e] = (uint32_t)(tiles.size() - tile0) - ebase[e];
    for (uint32_t d = 0; d < ne; ++d) {
      const uint32_t i = idx[d];
      sx[off + d] = xyz[off + i];
      sy[off + d] = xyz[n + off + i];
      sz[off + d] = xyz[2 * n + off + i];
    }
    off += ne;
  }
  for (size_t t = tile0; t < tiles.size(); ++t) {
    double b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = tiles[t].off; i < tiles[t].off + tiles[t].cnt; ++i) {
      const uint32_t li = i - atom_base;
      const double v[3] = {(double)sx[li], (double)sy[li], (double)sz[li]};
      for (int a = 0; a < 3; ++a) {
        b[a] = std::min(b[a], v[a]);
        b[3 + a] = std::max(b[3 + a], v[a]);
      }
    }
    boxes.insert(boxes.end(), b, b + 6);
  }
}

// J records of the FP64 kernel (no-op for FP32 binning).
template <class T>
static void launch_records(hipStream_t st, const T *x, const T *y, const T *z, const Tile *tiles,
                           const double *boxes, size_t ntiles, double *rec) {
  if constexpr (sizeof(T) == 8) {
    if (ntiles)
      hipLaunchKernelGGL(hist_tile_records, dim3((unsigned)ntiles), dim3(256), 0, st, x, y, z,
                         tiles, boxes, rec);
  }
}

template <class T>
static int launch_hist(xs_ctx *ctx, const Args<T> &A, size_t nitems, uint32_t wcap, bool prof) {
  if (!nitems) return XS_OK;
  hipEvent_t ev = nullptr;
  if (prof) XS_TRY(prof_begin(ctx, "hist", &ev));
  const dim3 grid((unsigned)nitems), block(kTile);
  if constexpr (sizeof(T) == 8) {
    if (wcap == 8192) {
      if (A.narrow)
        hipLaunchKernelGGL((hist_kernel_f64<8192, true>), grid, block, 0, ctx->stream, A);
      else
        hipLaunchKernelGGL((hist_kernel_f64<8192, false>), grid, block, 0, ctx->stream, A);
    } else {
      if (A.narrow)
        hipLaunchKernelGGL((hist_kernel_f64<32768, true>), grid, block, 0, ctx->stream, A);
      else
        hipLaunchKernelGGL((hist_kernel_f64<32768, false>), grid, block, 0, ctx->stream, A);
    }
  } else {
    if (wcap == 8192) {
      if (A.quot_f32)
        hipLaunchKernelGGL((hist_kernel_f32<8192, true>), grid, block, 0, ctx->stream, A);
      else
        hipLaunchKernelGGL((hist_kernel_f32<8192, false>), grid, block, 0, ctx->stream, A);
    } else {
      if (A.quot_f32)
        hipLaunchKernelGGL((hist_kernel_f32<32768, true>), grid, block, 0, ctx->stream, A);
      else
        hipLaunchKernelGGL((hist_kernel_f32<32768, false>), grid, block, 0, ctx->stream, A);
    }
  }
  XS_HIP_CHECK(hipGetLastError());
  if (prof) XS_TRY(prof_end(ctx, "hist", ev));
  return XS_OK;
}

static int check_counters(xs_ctx *ctx) {
  if (!ctx->quot_check) XS_HIP_CHECK(hipMalloc((void **)&ctx->quot_check, 32));
  return ensure_pinned(ctx, 64);
}

// FP32 binning computes the IEEE sqrtf with sqrt_rn_f32, whose correction
// step needs v_sqrt_f32 within one ulp of the root's own binade.  Once per
// context, sqrt_f32_check_kernel compares it with the IEEE root on every
// positive float (about 2 ms, synchronises); binning in FP32 is refused if
// any root from 2^-96 up differs, and for bins below 2^-40 unless the check
// also held below 2^-96.
static int sqrt_f32_checked(xs_ctx *ctx, double bin) {
  if (ctx->sqrt_f32_state < 0) {
    XS_TRY(check_counters(ctx));
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, ctx->stream, ctx->quot_check, (size_t)4);
    hipLaunchKernelGGL(sqrt_f32_check_kernel, dim3(8192), dim3(256), 0, ctx->stream, ctx->quot_check);
    XS_HIP_CHECK(hipGetLastError());
    XS_HIP_CHECK(hipMemcpyAsync(ctx->pin, ctx->quot_check, 24, hipMemcpyDeviceToHost, ctx->stream));
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const auto *bad = static_cast<const unsigned long long *>(ctx->pin);
    ctx->sqrt_f32_state = (bad[0] || bad[1]) ? 0 : (bad[2] ? 2 : 1);
  }
  if (ctx->sqrt_f32_state == 0)
    return fail(XS_ERR_DEVICE, "FP32 binning: this device's v_sqrt_f32 fails the exhaustive check "
                               "of the correctly rounded square root; use XS_F64_BINNING");
  if (ctx->sqrt_f32_state == 2 && !((float)bin >= 0x1p-40f))
    return fail(XS_ERR_RANGE, "FP32 binning: bins below 2^-40 need XS_F64_BINNING on this device");
  return XS_OK;
}

// Whether FP32 binning may take quot_f32 for this bin: quot_f32_check_kernel
// over all 2^23 significands, once per bin and context (synchronises).
static int quot_f32_exact(xs_ctx *ctx, double bin_d, bool *ok) {
  *ok = false;
  const float bin = (float)bin_d;
  if (!(bin >= 0x1p-60f && bin <= 0x1p60f)) return XS_OK;
  uint32_t key;
  std::memcpy(&key, &bin, sizeof key);
  const auto it = ctx->quot_f32_ok.find(key);
  if (it != ctx->quot_f32_ok.end()) {
    *ok = it->second;
    return XS_OK;
  }
  XS_TRY(check_counters(ctx));
  hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, ctx->stream, ctx->quot_check, (size_t)1);
  hipLaunchKernelGGL(quot_f32_check_kernel, dim3(1u << 15), dim3(256), 0, ctx->stream, bin,
                     1.0 / (double)bin, ctx->quot_check);
  XS_HIP_CHECK(hipGetLastError());
  XS_HIP_CHECK(hipMemcpyAsync(ctx->pin, ctx->quot_check, 8, hipMemcpyDeviceToHost, ctx->stream));
  XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  *ok = *static_cast<const unsigned long long *>(ctx->pin) == 0;
  ctx->quot_f32_ok[key] = *ok;
  return XS_OK;
}

// ---------------------------------------------------------------- xs_hist

// Scratch of eq_tiles_device for elements of up to maxne atoms.
template <class T>
static size_t eq_scratch_bytes(uint32_t maxne, size_t *sort_bytes) {
  size_t b1 = 0, b2 = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, b1, (const T *)nullptr, (T *)nullptr,
                                           (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                           (int)maxne);
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, b2, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                           (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                           (int)maxne);
  *sort_bytes = align_up(std::max(b1, b2), 256);
  return Carver::need({maxne * sizeof(T), 6 * (size_t)maxne * sizeof(uint32_t), *sort_bytes});
}

// Equal-count tiles of every element on the device: stable radix sorts by
// x (slabs), by y then slab (columns), by z then column (tiles); the sorted
// coordinates go to sxp, syp, szp (element-major, as the input).
template <class T>
static int eq_tiles_device(xs_ctx *ctx, const T *xyz, size_t n, const unsigned *el_count,
                           unsigned n_el, const uint32_t *eoff, const double bbox[6], double cap,
                           void *scratch, uint32_t maxne, T *sxp, T *syp, T *szp,
                           std::vector<Tile> &tiles, uint32_t *ebase, uint32_t *en) {
  size_t sb = 0;
  (void)eq_scratch_bytes<T>(maxne, &sb);
  Carver cv(scratch);
  T *kout = cv.take<T>(maxne);
  uint32_t *iota = cv.take<uint32_t>(maxne), *ordA = cv.take<uint32_t>(maxne),
           *ordB = cv.take<uint32_t>(maxne), *lab = cv.take<uint32_t>(maxne),
           *k32a = cv.take<uint32_t>(maxne), *k32b = cv.take<uint32_t>(maxne);
  void *tmp = cv.take<char>(sb);
  hipStream_t st = ctx->stream;
  for (unsigned e = 0; e < n_el; ++e) {
    const uint32_t ne = el_count[e];
    ebase[e] = (uint32_t)tiles.size();
    en[e] = 0;
    if (!ne) continue;
    const T *ex = xyz + eoff[e], *ey = xyz + n + eoff[e], *ez = xyz + 2 * n + eoff[e];
    uint32_t gx, gy, nz;
    choose_equal_count(ne, bbox, cap, sizeof(T) == 8, gx, gy, nz);
    const unsigned nb = (ne + 255) / 256;
    size_t b = sb;
    hipLaunchKernelGGL(iota_u32, dim3(nb), dim3(256), 0, st, iota, ne);
    XS_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, b, ex, kout, iota, ordA, (int)ne, 0,
                                                    (int)(sizeof(T) * 8), st));
    hipLaunchKernelGGL(eq_slabs, dim3(nb), dim3(256), 0, st, ordA, ne, gx, lab);
    b = sb;
    XS_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, b, ey, kout, iota, ordA, (int)ne, 0,
                                                    (int)(sizeof(T) * 8), st));
    hipLaunchKernelGGL(gather_u32, dim3(nb), dim3(256), 0, st, lab, ordA, ne, k32a);
    b = sb;
    XS_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, b, k32a, k32b, ordA, ordB, (int)ne, 0,
                                                    bits_for(gx), st));
    hipLaunchKernelGGL(eq_columns, dim3(nb), dim3(256), 0, st, k32b, ordB, ne, gx, gy, lab);
    b = sb;
    XS_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, b, ez, kout, iota, ordA, (int)ne, 0,
                                                    (int)(sizeof(T) * 8), st));
    hipLaunchKernelGGL(gather_u32, dim3(nb), dim3(256), 0, st, lab, ordA, ne, k32a);
    b = sb;
    XS_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, b, k32a, k32b, ordA, ordB, (int)ne, 0,
                                                    bits_for((uint64_t)gx * gy), st));
    hipLaunchKernelGGL((gather_xyz<T>), dim3(nb), dim3(256), 0, st, ex, ey, ez, ordB, ne,
                       sxp + eoff[e], syp + eoff[e], szp + eoff[e]);
    XS_HIP_CHECK(hipGetLastError());
    equal_count_tiles(eoff[e], ne, gx, gy, nz, sizeof(T) == 8, tiles);
    en[e] = (uint32_t)tiles.size() - ebase[e];
  }
  return XS_OK;
}


template <class T>
static int run_device(xs_ctx *ctx, const T *xyz, size_t n, const unsigned *el_count,
                      unsigned n_el, double bin, unsigned nhist, unsigned long long *hist_d,
                      unsigned long long *ovf_user, bool quot, bool equal_count = true) {
  const unsigned ntypes = ntypes_of(n_el);
  const size_t hsize = (size_t)ntypes * nhist;
  const uint32_t wcap = wcap_for(nhist);
  uint32_t eoff[kMaxEl];
  for (unsigned e = 0, o = 0; e < n_el; ++e) {
    eoff[e] = o;
    o += el_count[e];
  }
  // 1. bounding box on the GPU, read back
  XS_TRY(ensure_pinned(ctx, 4096));
  XS_TRY(ensure_workspace(ctx, Carver::need({kBoxBlocks * 6 * sizeof(double), 8 * sizeof(double)})));
  double bbox[6];
  {
    Carver cb(ctx->ws);
    double *part = cb.take<double>(kBoxBlocks * 6);
    double *bbox_d = cb.take<double>(8);
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, kBoxBlocks);
    hipLaunchKernelGGL((hist_bbox_partial<T>), dim3(nb), dim3(256), 0, ctx->stream, xyz, xyz + n,
                       xyz + 2 * n, n, part);
    hipLaunchKernelGGL(hist_bbox_reduce, dim3(1), dim3(256), 0, ctx->stream, part, (int)nb, bbox_d);
    XS_HIP_CHECK(hipGetLastError());
    XS_HIP_CHECK(hipMemcpyAsync(ctx->pin, bbox_d, 6 * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    std::memcpy(bbox, ctx->pin, sizeof(bbox));
  }
  XS_TRY(check_span(bbox, bin));

  // 2. workspace: sorted coordinates, cell keys, overflow counter, scratch
  //    of the equal-count sorts
  uint32_t maxne = 0;
  for (unsigned e = 0; e < n_el; ++e) maxne = std::max<uint32_t>(maxne, el_count[e]);
  size_t sortb = 0;
  const size_t eqb = equal_count ? eq_scratch_bytes<T>(std::max(maxne, 1u), &sortb) : 0;
  const size_t nrec = sizeof(T) == 8 ? 4 * n : 0;  // FP64 J records
  XS_TRY(ensure_workspace(ctx, Carver::need({3 * n * sizeof(T), n * sizeof(uint32_t),
                                             sizeof(unsigned long long), eqb,
                                             nrec * sizeof(double)})));
  Carver cv(ctx->ws);
  T *sorted = cv.take<T>(3 * n);
  uint32_t *keys = cv.take<uint32_t>(n);
  unsigned long long *ovf_own = cv.take<unsigned long long>(1);
  void *eq_scratch = cv.take<char>(eqb);
  double *rec = cv.take<double>(nrec);
  unsigned long long *ovf_d = ovf_user ? ovf_user : ovf_own;
  T *sxp = sorted, *syp = sorted + n, *szp = sorted + 2 * n;

  // 3. tiles.  Equal-count tiles (stable sorts on the device; the tiles
  //    follow from the counts alone) or, as the fallback, a cell grid per
  //    element sized for full tiles (counts read back, tiles on the host;
  //    the counts go through the table buffer, which holds nothing else yet).
  hipEvent_t ev;
  XS_TRY(prof_begin(ctx, "hist_prep", &ev));
  const double cap = cell_cap(bin, nhist, wcap);
  std::vector<Grid> grids(n_el);
  std::vector<uint32_t> m_of(n_el), soff(n_el + 1, 0u), start_h;
  std::vector<Tile> tiles;
  uint32_t ebase[kMaxEl], en[kMaxEl];
  if (equal_count)
    XS_TRY(eq_tiles_device<T>(ctx, xyz, n, el_count, n_el, eoff, bbox, cap, eq_scratch,
                              std::max(maxne, 1u), sxp, syp, szp, tiles, ebase, en));
  for (unsigned e = 0; e < n_el && !equal_count; ++e) {
    const uint32_t ne = el_count[e];
    double cell = first_cell(bbox, ne, cap);
    for (int pass = 0;; ++pass) {
      grids[e] = grid_with_cell(bbox, cell, cap);
      const uint32_t m = grids[e].g[0] * grids[e].g[1] * grids[e].g[2];
      m_of[e] = m;
      XS_TRY(ensure_aux(ctx, (size_t)m * sizeof(uint32_t)));
      XS_TRY(ensure_pinned(ctx, (size_t)m * sizeof(uint32_t)));
      uint32_t *count = static_cast<uint32_t *>(ctx->aux);
      XS_HIP_CHECK(hipMemsetAsync(count, 0, (size_t)m * sizeof(uint32_t), ctx->stream));
      if (ne)
        hipLaunchKernelGGL((hist_cell_count<T>), dim3((ne + 255) / 256), dim3(256), 0, ctx->stream,
                           xyz + eoff[e], xyz + n + eoff[e], xyz + 2 * n + eoff[e], ne, grids[e],
                           keys + eoff[e], count);
      XS_HIP_CHECK(hipGetLastError());
      XS_HIP_CHECK(hipMemcpyAsync(ctx->pin, count, (size_t)m * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, ctx->stream));
      XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
      const uint32_t *c_h = static_cast<const uint32_t *>(ctx->pin);
      const uint32_t fullest = ne ? *std::max_element(c_h, c_h + m) : 0u;
      const double nc = pass == 0 ? refined_cell(cell, fullest, cap) : 0.0;
      if (nc == 0.0) {
        soff[e + 1] = soff[e] + m;
        start_h.resize(soff[e + 1]);
        ebase[e] = (uint32_t)tiles.size();
        uint32_t pos = eoff[e];
        for (uint32_t c = 0; c < m; ++c) {
          start_h[soff[e] + c] = pos;
          cell_tiles(pos, c_h[c], tiles);
          pos += c_h[c];
        }
        en[e] = (uint32_t)tiles.size() - ebase[e];
        break;
      }
      cell = nc;
    }
  }
  std::vector<TypeDesc> types;
  add_types(n_el, ebase, en, nhist, bin, 0ull, 0u, types);
  std::vector<Item> items;
  add_items(types, 0, (uint32_t)types.size(), items, sizeof(T) == 8, kSingleSplit);
  order_items(items, types, tiles, sizeof(T) == 8);

  // 4. tables (exact sizes) in the table buffer; the stream is idle here
  const size_t sb = start_h.size() * sizeof(uint32_t), tb = tiles.size() * sizeof(Tile),
               yb = types.size() * sizeof(TypeDesc), ib = items.size() * sizeof(Item),
               bb = tiles.size() * 6 * sizeof(double);
  XS_TRY(ensure_aux(ctx, Carver::need({sb, tb, yb, ib, bb})));
  Carver ca(ctx->aux);
  uint32_t *start_d = ca.take<uint32_t>(start_h.size());
  Tile *tiles_d = ca.take<Tile>(tiles.size());
  TypeDesc *types_d = ca.take<TypeDesc>(types.size());
  Item *items_d = ca.take<Item>(items.size());
  double *boxes = ca.take<double>(tiles.size() * 6);
  if (sb) XS_HIP_CHECK(hipMemcpy(start_d, start_h.data(), sb, hipMemcpyHostToDevice));
  if (tb) XS_HIP_CHECK(hipMemcpy(tiles_d, tiles.data(), tb, hipMemcpyHostToDevice));
  XS_HIP_CHECK(hipMemcpy(types_d, types.data(), yb, hipMemcpyHostToDevice));
  if (ib) XS_HIP_CHECK(hipMemcpy(items_d, items.data(), ib, hipMemcpyHostToDevice));
  const unsigned zb = (unsigned)std::min<size_t>((hsize + 255) / 256, 4096);
  hipLaunchKernelGGL(fill_u64, dim3(std::max(zb, 1u)), dim3(256), 0, ctx->stream, hist_d, hsize);
  hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, ctx->stream, ovf_d, (size_t)1);
  for (unsigned e = 0; e < n_el && !equal_count; ++e)
    if (el_count[e])
      hipLaunchKernelGGL((hist_scatter<T>), dim3((el_count[e] + 255) / 256), dim3(256), 0, ctx->stream,
                         xyz + eoff[e], xyz + n + eoff[e], xyz + 2 * n + eoff[e], el_count[e],
                         keys + eoff[e], start_d + soff[e], sxp, syp, szp);
  if (!tiles.empty())
    hipLaunchKernelGGL((hist_tile_boxes<T>), dim3((unsigned)tiles.size()), dim3(256), 0, ctx->stream,
                       sxp, syp, szp, tiles_d, boxes);
  launch_records<T>(ctx->stream, sxp, syp, szp, tiles_d, boxes, tiles.size(), rec);
  XS_HIP_CHECK(hipGetLastError());
  if (equal_count && std::isfinite(cap) && !tiles.empty()) {
    // equal-count tiles of a structure far from uniform density can be wider
    // than the window allows: then the cell grid instead
    std::vector<double> bx(tiles.size() * 6);
    XS_HIP_CHECK(hipMemcpyAsync(bx.data(), boxes, bb, hipMemcpyDeviceToHost, ctx->stream));
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    bool wide = false;
    for (size_t t = 0; t < tiles.size() && !wide; ++t)
      for (int a = 0; a < 3; ++a)
        if (bx[6 * t + 3 + a] - bx[6 * t + a] > cap) wide = true;
    if (wide) {
      XS_TRY(prof_end(ctx, "hist_prep", ev));
      return run_device<T>(ctx, xyz, n, el_count, n_el, bin, nhist, hist_d, ovf_user, quot, false);
    }
  }
  XS_TRY(prof_end(ctx, "hist_prep", ev));
  Args<T> A{sxp, syp, szp, tiles_d, boxes, types_d, items_d, hist_d, ovf_d, rec};
  A.quot_f32 = quot;
  A.narrow = nhist <= kNarrowMax;
  return launch_hist<T>(ctx, A, items.size(), wcap, true);
}

// Host coordinates: one synchronous copy to the device (a buffer of the
// context apart from the workspace), the device path for everything else
// (bounding box, equal-count tiles, J records, kernel), and the counts back
// over the same stream.  The host only checks the span first, so a bad input
// is rejected before anything is copied.
template <class T>
static int run_host(xs_ctx *ctx, const T *xyz, size_t n, const unsigned *el_count, unsigned n_el,
                    double bin, unsigned nhist, uint64_t *hist_out, uint64_t *ovf_out, bool quot) {
  const size_t hsize = (size_t)ntypes_of(n_el) * nhist;
  {
    double bbox[6];
    host_bbox<T>(xyz, n, bbox);
    XS_TRY(check_span(bbox, bin));
  }
  XS_TRY(ensure_io(ctx, Carver::need({3 * n * sizeof(T), hsize * 8, 8})));
  Carver cv(ctx->io);
  T *xyz_d = cv.take<T>(3 * n);
  unsigned long long *hist_d = cv.take<unsigned long long>(hsize);
  unsigned long long *ovf_d = cv.take<unsigned long long>(1);
  // the stream may still read the buffer from an earlier call
  XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  XS_HIP_CHECK(hipMemcpy(xyz_d, xyz, 3 * n * sizeof(T), hipMemcpyHostToDevice));
  XS_TRY(run_device<T>(ctx, xyz_d, n, el_count, n_el, bin, nhist, hist_d, ovf_d, quot));
  XS_HIP_CHECK(hipMemcpyAsync(hist_out, hist_d, hsize * 8, hipMemcpyDeviceToHost, ctx->stream));
  XS_HIP_CHECK(hipMemcpyAsync(ctx->pin, ovf_d, 8, hipMemcpyDeviceToHost, ctx->stream));
  XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (ovf_out) std::memcpy(ovf_out, ctx->pin, 8);
  return XS_OK;
}

}  // namespace hist
}  // namespace xs

// ---------------------------------------------------------------- batch plans

namespace xs {
namespace hist {

static void free_plan(xs_batch *b) {
  if (!b) return;
  if (b->ctx) b->ctx->plans.erase(b);
  for (void *p : {b->coords, b->tiles, b->boxes, b->types, b->items, (void *)b->ovf, b->recs})
    if (p) (void)hipFree(p);
  release_batch_run(b->run);
  delete b;
}

// The tail of a plan, host-built or device-built: work items (the
// 8192-bin window class, then the 32768 class, each longest first), the
// device tables and overflow counters, and, for a host-built plan, the
// upload of its sorted coordinates (host_sorted, 3 b->natoms values; a
// device-built plan has filled b->coords already).  Frees b on failure.
template <class T>
static int plan_finish(xs_ctx *ctx, xs_batch *b, const std::vector<Tile> &tiles,
                       const std::vector<double> &boxes, const std::vector<TypeDesc> &types,
                       unsigned long long hoff, xs_batch **out, const T *host_sorted) {
  b->hist_len = hoff;
  b->ntypes_total = types.size();
  std::vector<Item> all, small, large;
  add_items(types, 0, (uint32_t)types.size(), all, sizeof(T) == 8);
  for (const Item &it : all) (types[it.type].nhist <= 8192 ? small : large).push_back(it);
  order_items(small, types, tiles, sizeof(T) == 8);
  order_items(large, types, tiles, sizeof(T) == 8);
  b->items8192 = small.size();
  b->items32768 = large.size();
  small.insert(small.end(), large.begin(), large.end());
  const size_t nc = b->natoms;
  const bool ok =
      (b->coords || hipMalloc(&b->coords, 3 * nc * sizeof(T)) == hipSuccess) &&
      hipMalloc(&b->tiles, std::max<size_t>(tiles.size(), 1) * sizeof(Tile)) == hipSuccess &&
      hipMalloc(&b->boxes, std::max<size_t>(boxes.size(), 6) * 8) == hipSuccess &&
      hipMalloc(&b->types, std::max<size_t>(types.size(), 1) * sizeof(TypeDesc)) == hipSuccess &&
      hipMalloc(&b->items, std::max<size_t>(small.size(), 1) * sizeof(Item)) == hipSuccess &&
      hipMalloc((void **)&b->ovf, std::max<size_t>(b->nstruct, 1) * 8) == hipSuccess &&
      (sizeof(T) == 4 || hipMalloc(&b->recs, std::max<size_t>(nc, 1) * 32) == hipSuccess);
  if (!ok) {
    (void)hipGetLastError();
    free_plan(b);
    return fail(XS_ERR_NOMEM, "xs_batch_plan: device allocation failed");
  }
  T *c = static_cast<T *>(b->coords);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess && host_sorted)
    e = hipMemcpy(c, host_sorted, 3 * nc * sizeof(T), hipMemcpyHostToDevice);
  if (e == hipSuccess && !tiles.empty())
    e = hipMemcpy(b->tiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice);
  if (e == hipSuccess && !boxes.empty())
    e = hipMemcpy(b->boxes, boxes.data(), boxes.size() * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess && !types.empty())
    e = hipMemcpy(b->types, types.data(), types.size() * sizeof(TypeDesc), hipMemcpyHostToDevice);
  if (e == hipSuccess && !small.empty())
    e = hipMemcpy(b->items, small.data(), small.size() * sizeof(Item), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch_records<T>(ctx->stream, c, c + nc, c + 2 * nc, static_cast<const Tile *>(b->tiles),
                      static_cast<const double *>(b->boxes), tiles.size(),
                      static_cast<double *>(b->recs));
    e = hipGetLastError();
  }
  if (e != hipSuccess) {
    free_plan(b);
    return fail(XS_ERR_HIP, std::string("xs_batch_plan: upload failed: ") + hipGetErrorString(e));
  }
  *out = b;
  ctx->plans.insert(b);
  return XS_OK;
}

// Per-structure part of a batch plan, built on a worker thread.
struct StructPart {
  int rc = XS_OK;
  std::string err;
  uint32_t nh = 0;
  std::vector<Tile> tiles;  // atom offsets already global
  std::vector<double> boxes;
  uint32_t ebase[kMaxEl], en[kMaxEl];
};

// Threads of the host planning: the machine's, at most 16.
static unsigned plan_threads() {
  const unsigned h = std::thread::hardware_concurrency();
  return std::max(1u, std::min(h ? h : 1u, 16u));
}

template <class T>
static int plan_build(xs_ctx *ctx, unsigned nstruct, const unsigned *n_el, const unsigned *el_count,
                      const double *const *xyz, double bin, const unsigned *nhist_in, bool f64,
                      bool quot, xs_batch **out) {
  // 1. sizes and validation (serial, cheap)
  std::vector<size_t> nat(nstruct), ecf(nstruct), abase(nstruct);
  size_t ec = 0, ntot_all = 0;
  for (unsigned s = 0; s < nstruct; ++s) {
    const unsigned ne = n_el[s];
    if (ne == 0 || ne > (unsigned)kMaxEl)
      return fail(XS_ERR_RANGE, "xs_batch_plan: n_el must be in [1, 16]");
    size_t n = 0;
    for (unsigned e = 0; e < ne; ++e) n += el_count[ec + e];
    if (n == 0 || n >= (size_t(1) << 31))
      return fail(XS_ERR_RANGE, "xs_batch_plan: a structure must hold 1 to 2^31-1 atoms");
    nat[s] = n;
    ecf[s] = ec;
    abase[s] = ntot_all;
    ntot_all += n;
    ec += ne;
  }
  if (ntot_all >= (size_t(1) << 32))
    return fail(XS_ERR_RANGE, "xs_batch_plan: an ensemble must hold fewer than 2^32 atoms");
  const size_t nc = std::max<size_t>(ntot_all, 1);
  // every element is written by the workers below: no value-initialisation
  // (a serial pass over 24 bytes per atom before the parallel part)
  std::unique_ptr<T[]> sorted_buf(new T[3 * nc]);
  T *const sorted = sorted_buf.get();

  // 2. per structure, on worker threads (largest first): bounding box, bin
  //    count, k-d tiles; sorted coordinates straight into their place
  std::vector<StructPart> parts(nstruct);
  std::vector<unsigned> order(nstruct);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](unsigned a, unsigned c) { return nat[a] > nat[c]; });
  std::atomic<unsigned> next{0};
  auto worker = [&]() {
    std::vector<T> tmp;
    for (unsigned k; (k = next.fetch_add(1)) < nstruct;) {
      const unsigned s = order[k];
      StructPart &P = parts[s];
      const size_t n = nat[s];
      const double *X = xyz[s];
      double bbox[6];
      host_bbox<double>(X, n, bbox);
      P.rc = check_span(bbox, bin);
      if (P.rc != XS_OK) {
        P.err = xs_last_error();
        continue;
      }
      // bin count as the reference's GetHistSize (CalcFunctions.cpp:224-243) unless given
      uint32_t nh = nhist_in ? nhist_in[s] : 0u;
      if (!nh) {
        const double dx = bbox[3] - bbox[0], dy = bbox[4] - bbox[1], dz = bbox[5] - bbox[2];
        nh = (uint32_t)(std::sqrt(dx * dx + dy * dy + dz * dz) / bin) + 1;
      }
      if (!f64 && nh > (1u << 24)) {
        P.rc = XS_ERR_RANGE;
        P.err = "xs_batch_plan: nhist > 2^24 needs XS_F64_BINNING";
        continue;
      }
      P.nh = nh;
      const T *src;
      if constexpr (std::is_same<T, double>::value) {
        src = X;
      } else {
        tmp.resize(3 * n);
        for (size_t i = 0; i < 3 * n; ++i) tmp[i] = (T)X[i];
        src = tmp.data();
      }
      if (sizeof(T) < sizeof(double)) {
        // coordinates beyond the float range become infinite in FP32 binning
        double fb[6];
        host_bbox<T>(src, n, fb);
        P.rc = check_span(fb, bin);
        if (P.rc != XS_OK) {
          P.err = xs_last_error();
          continue;
        }
      }
      const size_t a0 = abase[s];
      sort_structure_host<T>(src, n, el_count + ecf[s], n_el[s], bin, nh, sorted + a0,
                             sorted + nc + a0, sorted + 2 * nc + a0, (uint32_t)a0, P.tiles, P.boxes,
                             P.ebase, P.en);
    }
  };
  {
    const unsigned nt = std::min(plan_threads(), nstruct);
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto &th : pool) th.join();
  }
  for (unsigned s = 0; s < nstruct; ++s)
    if (parts[s].rc != XS_OK) return fail(parts[s].rc, parts[s].err);

  // 3. tables in structure order (serial)
  auto *b = new xs_batch();
  b->ctx = ctx;
  b->f64 = f64;
  b->quot_f32 = quot;
  b->bin = bin;
  b->nstruct = nstruct;
  std::vector<Tile> tiles;
  std::vector<double> boxes;
  std::vector<TypeDesc> types;
  unsigned long long hoff = 0;
  for (unsigned s = 0; s < nstruct; ++s) {
    StructPart &P = parts[s];
    const unsigned ne = n_el[s];
    const uint32_t nh = P.nh;
    const uint32_t tbase = (uint32_t)tiles.size();
    tiles.insert(tiles.end(), P.tiles.begin(), P.tiles.end());
    boxes.insert(boxes.end(), P.boxes.begin(), P.boxes.end());
    std::vector<Tile>().swap(P.tiles);
    std::vector<double>().swap(P.boxes);
    uint32_t ebase[kMaxEl];
    for (unsigned e = 0; e < ne; ++e) ebase[e] = tbase + P.ebase[e];
    const size_t t0 = types.size();
    add_types(ne, ebase, P.en, nh, bin, hoff, s, types);
    b->type_first.push_back(t0);
    for (size_t t = t0; t < types.size(); ++t) {
      b->type_hist_off.push_back(types[t].hist_off);
      b->type_nhist.push_back(nh);
    }
    b->pdf_off.push_back(b->pdf_len);
    b->pdf_len += (unsigned long long)(1 + ntypes_of(ne)) * nh;
    b->n_el.push_back(ne);
    b->el_first.push_back(ecf[s]);
    b->nhist.push_back(nh);
    b->ntot.push_back((unsigned)nat[s]);
    b->hist_off.push_back(hoff);
    for (unsigned e = 0; e < ne; ++e) b->el_count.push_back(el_count[ecf[s] + e]);
    hoff += (unsigned long long)ntypes_of(ne) * nh;
  }
  b->natoms = nc;
  return plan_finish<T>(ctx, b, tiles, boxes, types, hoff, out, sorted);
}

// ---- batch plans from coordinates already on the device ----
//
// The same plan as plan_build, tiled on the GPU: every structure's
// coordinates are gathered into one SoA array, and three stable radix sorts
// over all atoms of all structures at once order them by (element segment,
// x), then (slab, y), then (column, z), with 64-bit keys whose high word is
// the segment / slab / column number and whose low word the order-keeping
// bits of the coordinate as float.  Slabs and columns hold equal counts, so
// the tiles follow from the counts alone (equal_count_tiles, on the host).
// A structure whose equal-count tiles come out wider than its window allows
// (an element that forms a hollow shell) is re-tiled on the host with the
// k-d split, from its own coordinates, as plan_build does.

struct SegDev {
  uint32_t start, n;  // atoms [start, start + n) of the concatenated arrays
  uint32_t gx, slab0;  // its slabs: global numbers slab0 .. slab0 + gx - 1
};
struct SlabDev {
  uint32_t start, n;  // positions of the slab after the x sort
  uint32_t gy, col0;  // its columns: global numbers col0 .. col0 + gy - 1
};

// Bits of (float)v whose unsigned order is the order of v.
__device__ __forceinline__ uint32_t ord_bits(double v) {
  const uint32_t u = __float_as_uint((float)v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// raw[a], raw[N + a], raw[2N + a] = atom a of the concatenation of the
// structures' own SoA arrays (structure s holds atoms abase[s] .. abase[s+1]).
__global__ void batch_gather(const double *const *__restrict__ src,
                             const unsigned long long *__restrict__ abase, unsigned nstruct,
                             size_t ntot, double *__restrict__ raw) {
  for (size_t a = blockIdx.x * (size_t)blockDim.x + threadIdx.x; a < ntot;
       a += (size_t)gridDim.x * blockDim.x) {
    unsigned lo = 0, hi = nstruct - 1;
    while (lo < hi) {
      const unsigned mid = (lo + hi + 1) / 2;
      if (abase[mid] <= a) lo = mid;
      else hi = mid - 1;
    }
    const size_t n = abase[lo + 1] - abase[lo], i = a - abase[lo];
    const double *p = src[lo];
    raw[a] = p[i];
    raw[ntot + a] = p[n + i];
    raw[2 * ntot + a] = p[2 * n + i];
  }
}

// Bounding box of every segment (one work group each), NaN if one is NaN.
__global__ void seg_boxes(const double *__restrict__ x, const double *__restrict__ y,
                          const double *__restrict__ z, const SegDev *__restrict__ segs,
                          double *__restrict__ box) {
  const SegDev sg = segs[blockIdx.x];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  int nan = 0;
  for (uint32_t i = sg.start + threadIdx.x; i < sg.start + sg.n; i += blockDim.x) {
    const double v[3] = {x[i], y[i], z[i]};
    for (int a = 0; a < 3; ++a) {
      nan |= v[a] != v[a];
      lo[a] = fmin(lo[a], v[a]);
      hi[a] = fmax(hi[a], v[a]);
    }
  }
  __shared__ double sh[6][256];
  for (int a = 0; a < 3; ++a) {
    sh[a][threadIdx.x] = lo[a];
    sh[3 + a][threadIdx.x] = hi[a];
  }
  const int any_nan = __syncthreads_or(nan);
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 3; ++a) {
        sh[a][threadIdx.x] = fmin(sh[a][threadIdx.x], sh[a][threadIdx.x + w]);
        sh[3 + a][threadIdx.x] = fmax(sh[3 + a][threadIdx.x], sh[3 + a][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) box[(size_t)blockIdx.x * 6 + threadIdx.x] = any_nan ? NAN : sh[threadIdx.x][0];
}

// Keys of the first sort: (segment, x), values the atom numbers.
__global__ void seg_keys(const double *__restrict__ x, const SegDev *__restrict__ segs,
                         unsigned long long *__restrict__ keys, uint32_t *__restrict__ vals) {
  const SegDev sg = segs[blockIdx.x];
  for (uint32_t i = sg.start + threadIdx.x; i < sg.start + sg.n; i += blockDim.x) {
    keys[i] = ((unsigned long long)blockIdx.x << 32) | ord_bits(x[i]);
    vals[i] = i;
  }
}

// After the (segment, x) sort: the slab of each atom.
__global__ void slab_labels(const unsigned long long *__restrict__ keys,
                            const uint32_t *__restrict__ vals, const SegDev *__restrict__ segs,
                            size_t ntot, uint32_t *__restrict__ lab) {
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < ntot;
       p += (size_t)gridDim.x * blockDim.x) {
    const SegDev sg = segs[keys[p] >> 32];
    lab[vals[p]] = sg.slab0 + equal_part(p - sg.start, sg.n, sg.gx);
  }
}

// After the (slab, y) sort: the column of each atom.
__global__ void col_labels(const unsigned long long *__restrict__ keys,
                           const uint32_t *__restrict__ vals, const SlabDev *__restrict__ slabs,
                           size_t ntot, uint32_t *__restrict__ lab) {
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < ntot;
       p += (size_t)gridDim.x * blockDim.x) {
    const SlabDev sb = slabs[keys[p] >> 32];
    lab[vals[p]] = sb.col0 + equal_part(p - sb.start, sb.n, sb.gy);
  }
}

// Keys (label, coordinate) of the second and third sorts.
__global__ void label_keys(const double *__restrict__ c, const uint32_t *__restrict__ lab,
                           size_t ntot, unsigned long long *__restrict__ keys,
                           uint32_t *__restrict__ vals) {
  for (size_t a = blockIdx.x * (size_t)blockDim.x + threadIdx.x; a < ntot;
       a += (size_t)gridDim.x * blockDim.x) {
    keys[a] = ((unsigned long long)lab[a] << 32) | ord_bits(c[a]);
    vals[a] = (uint32_t)a;
  }
}

// The plan's coordinates in the final order (as T).
template <class T>
__global__ void gather_sorted(const double *__restrict__ raw, const uint32_t *__restrict__ ord,
                              size_t ntot, size_t nc, T *__restrict__ out) {
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < ntot;
       p += (size_t)gridDim.x * blockDim.x) {
    const uint32_t a = ord[p];
    out[p] = (T)raw[a];
    out[nc + p] = (T)raw[ntot + a];
    out[2 * nc + p] = (T)raw[2 * ntot + a];
  }
}

static unsigned grid_for(size_t n) {
  return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 16384));
}

// One stable sort of (key, value) pairs over all atoms on the low `bits`
// key bits (scratch from tmp, sized by the caller).
static int sort_u64(xs_ctx *ctx, void *tmp, size_t tmp_bytes, const unsigned long long *kin,
                    unsigned long long *kout, const uint32_t *vin, uint32_t *vout, size_t n,
                    int bits) {
  size_t b = tmp_bytes;
  XS_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, b, kin, kout, vin, vout, (int)n, 0, bits,
                                                  ctx->stream));
  return XS_OK;
}

template <class T>
static int plan_build_device(xs_ctx *ctx, unsigned nstruct, const unsigned *n_el,
                             const unsigned *el_count, const double *const *xyz, double bin,
                             const unsigned *nhist_in, bool f64, bool quot, xs_batch **out) {
  // 1. sizes (host)
  std::vector<size_t> nat(nstruct), ecf(nstruct);
  std::vector<unsigned long long> abase(nstruct + 1, 0ull);
  size_t ec = 0, nseg = 0;
  for (unsigned s = 0; s < nstruct; ++s) {
    const unsigned ne = n_el[s];
    if (ne == 0 || ne > (unsigned)kMaxEl)
      return fail(XS_ERR_RANGE, "xs_batch_plan: n_el must be in [1, 16]");
    size_t n = 0;
    for (unsigned e = 0; e < ne; ++e) n += el_count[ec + e];
    if (n == 0 || n >= (size_t(1) << 31))
      return fail(XS_ERR_RANGE, "xs_batch_plan: a structure must hold 1 to 2^31-1 atoms");
    nat[s] = n;
    ecf[s] = ec;
    abase[s + 1] = abase[s] + n;
    ec += ne;
    nseg += ne;
  }
  const size_t ntot = abase[nstruct];
  if (ntot >= (size_t(1) << 31))
    return fail(XS_ERR_RANGE, "xs_batch_plan: device tiling takes fewer than 2^31 atoms in all");
  std::vector<SegDev> segs(nseg);
  for (unsigned s = 0, g = 0; s < nstruct; ++s) {
    uint32_t o = (uint32_t)abase[s];
    for (unsigned e = 0; e < n_el[s]; ++e, ++g) {
      segs[g] = SegDev{o, el_count[ecf[s] + e], 1u, 0u};
      o += el_count[ecf[s] + e];
    }
  }

  // 2. gather and per-segment boxes (one read-back)
  hipStream_t st = ctx->stream;
  DevMem raw, ptrs, segd, boxd;
  XS_TRY(raw.alloc(3 * ntot * sizeof(double), "batch coordinates"));
  XS_TRY(ptrs.alloc(nstruct * sizeof(void *) + (nstruct + 1) * 8, "batch pointers"));
  XS_TRY(segd.alloc(nseg * sizeof(SegDev), "batch segments"));
  XS_TRY(boxd.alloc(nseg * 6 * sizeof(double), "batch boxes"));
  XS_HIP_CHECK(hipStreamSynchronize(st));
  XS_HIP_CHECK(hipMemcpy(ptrs.p, xyz, nstruct * sizeof(void *), hipMemcpyHostToDevice));
  auto *abase_d = reinterpret_cast<unsigned long long *>(static_cast<char *>(ptrs.p) +
                                                         nstruct * sizeof(void *));
  XS_HIP_CHECK(hipMemcpy(abase_d, abase.data(), (nstruct + 1) * 8, hipMemcpyHostToDevice));
  XS_HIP_CHECK(hipMemcpy(segd.p, segs.data(), nseg * sizeof(SegDev), hipMemcpyHostToDevice));
  double *rx = raw.at<double>(), *ry = rx + ntot, *rz = rx + 2 * ntot;
  hipLaunchKernelGGL(batch_gather, dim3(grid_for(ntot)), dim3(256), 0, st,
                     static_cast<const double *const *>(ptrs.p), abase_d, nstruct, ntot, rx);
  hipLaunchKernelGGL(seg_boxes, dim3((unsigned)nseg), dim3(256), 0, st, rx, ry, rz,
                     segd.at<SegDev>(), boxd.at<double>());
  XS_HIP_CHECK(hipGetLastError());
  std::vector<double> sbox(nseg * 6);
  XS_HIP_CHECK(hipMemcpyAsync(sbox.data(), boxd.p, nseg * 6 * sizeof(double), hipMemcpyDeviceToHost, st));
  XS_HIP_CHECK(hipStreamSynchronize(st));

  // 3. bin counts, window caps, slabs and columns (host)
  std::vector<uint32_t> nh_of(nstruct);
  std::vector<double> cap_of(nstruct);
  std::vector<SlabDev> slabs;
  std::vector<uint32_t> seg_gy(nseg), seg_nz(nseg);
  uint32_t ncols = 0;
  for (unsigned s = 0, g = 0; s < nstruct; ++s) {
    double bb[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    bool nan = false;
    for (unsigned e = 0; e < n_el[s]; ++e) {
      const double *b = &sbox[(g + e) * 6];
      for (int a = 0; a < 3; ++a) {
        nan |= !(b[a] == b[a]) || !(b[3 + a] == b[3 + a]);
        bb[a] = std::min(bb[a], b[a]);
        bb[3 + a] = std::max(bb[3 + a], b[3 + a]);
      }
    }
    if (nan)
      for (int a = 0; a < 6; ++a) bb[a] = NAN;
    XS_TRY(check_span(bb, bin));
    uint32_t nh = nhist_in ? nhist_in[s] : 0u;
    if (!nh) {
      // GetHistSize (CalcFunctions.cpp:224-243) over the union of the element boxes
      const double dx = bb[3] - bb[0], dy = bb[4] - bb[1], dz = bb[5] - bb[2];
      nh = (uint32_t)(std::sqrt(dx * dx + dy * dy + dz * dz) / bin) + 1;
    }
    if (!f64) {
      if (nh > (1u << 24)) return fail(XS_ERR_RANGE, "xs_batch_plan: nhist > 2^24 needs XS_F64_BINNING");
      double fb[6];
      for (int a = 0; a < 6; ++a) fb[a] = (double)(float)bb[a];
      XS_TRY(check_span(fb, bin));  // coordinates beyond the float range
    }
    nh_of[s] = nh;
    cap_of[s] = cell_cap(bin, nh, wcap_for(nh));
    for (unsigned e = 0; e < n_el[s]; ++e, ++g) {
      uint32_t gx, gy, nz;
      choose_equal_count(segs[g].n, &sbox[g * 6], cap_of[s], sizeof(T) == 8, gx, gy, nz);
      segs[g].gx = gx;
      segs[g].slab0 = (uint32_t)slabs.size();
      seg_gy[g] = gy;
      seg_nz[g] = nz;
      for (uint64_t sl = 0; sl < gx; ++sl) {
        const uint32_t a0 = (uint32_t)(sl * segs[g].n / gx), a1 = (uint32_t)((sl + 1) * segs[g].n / gx);
        slabs.push_back(SlabDev{segs[g].start + a0, a1 - a0, gy, ncols});
        ncols += gy;
      }
    }
  }

  // 4. the three sorts (device)
  DevMem slabd, keys, vals, lab, tmp;
  size_t tb = 0;
  XS_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (const unsigned long long *)nullptr,
                                                  (unsigned long long *)nullptr,
                                                  (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                  (int)ntot, 0, 64, st));
  XS_TRY(slabd.alloc(std::max<size_t>(slabs.size(), 1) * sizeof(SlabDev), "batch slabs"));
  XS_TRY(keys.alloc(2 * ntot * 8, "batch sort keys"));
  XS_TRY(vals.alloc(2 * ntot * 4, "batch sort values"));
  XS_TRY(lab.alloc(ntot * 4, "batch labels"));
  XS_TRY(tmp.alloc(tb, "batch sort scratch"));
  XS_HIP_CHECK(hipMemcpy(segd.p, segs.data(), nseg * sizeof(SegDev), hipMemcpyHostToDevice));
  XS_HIP_CHECK(hipMemcpy(slabd.p, slabs.data(), slabs.size() * sizeof(SlabDev), hipMemcpyHostToDevice));
  unsigned long long *k0 = keys.at<unsigned long long>(), *k1 = k0 + ntot;
  uint32_t *v0 = vals.at<uint32_t>(), *v1 = v0 + ntot, *lb = lab.at<uint32_t>();
  hipLaunchKernelGGL(seg_keys, dim3((unsigned)nseg), dim3(256), 0, st, rx, segd.at<SegDev>(), k0, v0);
  XS_TRY(sort_u64(ctx, tmp.p, tb, k0, k1, v0, v1, ntot, 32 + bits_for(nseg)));
  hipLaunchKernelGGL(slab_labels, dim3(grid_for(ntot)), dim3(256), 0, st, k1, v1, segd.at<SegDev>(),
                     ntot, lb);
  hipLaunchKernelGGL(label_keys, dim3(grid_for(ntot)), dim3(256), 0, st, ry, lb, ntot, k0, v0);
  XS_TRY(sort_u64(ctx, tmp.p, tb, k0, k1, v0, v1, ntot, 32 + bits_for(slabs.size())));
  hipLaunchKernelGGL(col_labels, dim3(grid_for(ntot)), dim3(256), 0, st, k1, v1, slabd.at<SlabDev>(),
                     ntot, lb);
  hipLaunchKernelGGL(label_keys, dim3(grid_for(ntot)), dim3(256), 0, st, rz, lb, ntot, k0, v0);
  XS_TRY(sort_u64(ctx, tmp.p, tb, k0, k1, v0, v1, ntot, 32 + bits_for(std::max(ncols, 1u))));
  XS_HIP_CHECK(hipGetLastError());

  // 5. the plan: sorted coordinates, equal-count tiles and their boxes
  auto *b = new xs_batch();
  b->ctx = ctx;
  b->f64 = f64;
  b->quot_f32 = quot;
  b->bin = bin;
  b->nstruct = nstruct;
  const size_t nc = ntot;
  b->natoms = nc;
  std::vector<Tile> tiles;
  std::vector<uint32_t> tile0(nstruct + 1), ebase_all(nseg), en_all(nseg);
  for (unsigned s = 0, g = 0; s < nstruct; ++s) {
    tile0[s] = (uint32_t)tiles.size();
    for (unsigned e = 0; e < n_el[s]; ++e, ++g) {
      ebase_all[g] = (uint32_t)tiles.size();
      equal_count_tiles(segs[g].start, segs[g].n, segs[g].gx, seg_gy[g], seg_nz[g], sizeof(T) == 8,
                        tiles);
      en_all[g] = (uint32_t)tiles.size() - ebase_all[g];
    }
  }
  tile0[nstruct] = (uint32_t)tiles.size();
  DevMem tiled, boxes_d;
  if (hipMalloc(&b->coords, 3 * nc * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    free_plan(b);
    return fail(XS_ERR_NOMEM, "xs_batch_plan: device allocation failed");
  }
  T *c = static_cast<T *>(b->coords);
  hipLaunchKernelGGL((gather_sorted<T>), dim3(grid_for(ntot)), dim3(256), 0, st, rx, v1, ntot, nc, c);
  int rc = tiled.alloc(std::max<size_t>(tiles.size(), 1) * sizeof(Tile), "batch tiles");
  if (rc == XS_OK) rc = boxes_d.alloc(std::max<size_t>(tiles.size(), 1) * 48, "batch tile boxes");
  if (rc != XS_OK) {
    free_plan(b);
    return rc;
  }
  std::vector<double> boxes(tiles.size() * 6);
  hipError_t e = hipMemcpy(tiled.p, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL((hist_tile_boxes<T>), dim3((unsigned)tiles.size()), dim3(256), 0, st, c, c + nc,
                       c + 2 * nc, tiled.at<Tile>(), boxes_d.at<double>());
    e = hipMemcpyAsync(boxes.data(), boxes_d.p, boxes.size() * 8, hipMemcpyDeviceToHost, st);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    free_plan(b);
    return fail(XS_ERR_HIP, std::string("xs_batch_plan: device tiling failed: ") + hipGetErrorString(e));
  }

  // 6. structures whose tiles exceed their window cap: k-d tiles on the host
  std::vector<std::vector<Tile>> kd_t(nstruct);
  std::vector<std::vector<double>> kd_b(nstruct);
  std::vector<std::array<uint32_t, kMaxEl>> kd_eb(nstruct), kd_en(nstruct);
  std::vector<char> redo(nstruct, 0);
  for (unsigned s = 0; s < nstruct; ++s) {
    if (!std::isfinite(cap_of[s])) continue;
    for (uint32_t t = tile0[s]; t < tile0[s + 1] && !redo[s]; ++t)
      for (int a = 0; a < 3; ++a)
        if (boxes[6 * t + 3 + a] - boxes[6 * t + a] > cap_of[s]) redo[s] = 1;
  }
  for (unsigned s = 0; s < nstruct; ++s) {
    if (!redo[s]) continue;
    const size_t n = nat[s];
    std::vector<double> xs(3 * n);
    std::vector<T> src(3 * n), sx(n), sy(n), sz(n);
    e = hipMemcpy(xs.data(), xyz[s], 3 * n * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      free_plan(b);
      return fail(XS_ERR_HIP, std::string("xs_batch_plan: read-back failed: ") + hipGetErrorString(e));
    }
    for (size_t i = 0; i < 3 * n; ++i) src[i] = (T)xs[i];
    sort_structure_host<T>(src.data(), n, el_count + ecf[s], n_el[s], bin, nh_of[s], sx.data(),
                           sy.data(), sz.data(), (uint32_t)abase[s], kd_t[s], kd_b[s],
                           kd_eb[s].data(), kd_en[s].data());
    e = hipMemcpy(c + abase[s], sx.data(), n * sizeof(T), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c + nc + abase[s], sy.data(), n * sizeof(T), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(c + 2 * nc + abase[s], sz.data(), n * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      free_plan(b);
      return fail(XS_ERR_HIP, std::string("xs_batch_plan: upload failed: ") + hipGetErrorString(e));
    }
  }

  // 7. tables in structure order, as plan_build
  std::vector<Tile> ftiles;
  std::vector<double> fboxes;
  std::vector<TypeDesc> types;
  unsigned long long hoff = 0;
  for (unsigned s = 0, g = 0; s < nstruct; ++s) {
    const unsigned ne = n_el[s];
    const uint32_t nh = nh_of[s];
    const uint32_t tbase = (uint32_t)ftiles.size();
    uint32_t ebase[kMaxEl], en[kMaxEl];
    if (redo[s]) {
      ftiles.insert(ftiles.end(), kd_t[s].begin(), kd_t[s].end());
      fboxes.insert(fboxes.end(), kd_b[s].begin(), kd_b[s].end());
      for (unsigned el = 0; el < ne; ++el) {
        ebase[el] = tbase + kd_eb[s][el];
        en[el] = kd_en[s][el];
      }
    } else {
      ftiles.insert(ftiles.end(), tiles.begin() + tile0[s], tiles.begin() + tile0[s + 1]);
      fboxes.insert(fboxes.end(), boxes.begin() + 6 * (size_t)tile0[s], boxes.begin() + 6 * (size_t)tile0[s + 1]);
      for (unsigned el = 0; el < ne; ++el) {
        ebase[el] = tbase + (ebase_all[g + el] - tile0[s]);
        en[el] = en_all[g + el];
      }
    }
    g += ne;
    const size_t t0 = types.size();
    add_types(ne, ebase, en, nh, bin, hoff, s, types);
    b->type_first.push_back(t0);
    for (size_t t = t0; t < types.size(); ++t) {
      b->type_hist_off.push_back(types[t].hist_off);
      b->type_nhist.push_back(nh);
    }
    b->pdf_off.push_back(b->pdf_len);
    b->pdf_len += (unsigned long long)(1 + ntypes_of(ne)) * nh;
    b->n_el.push_back(ne);
    b->el_first.push_back(ecf[s]);
    b->nhist.push_back(nh);
    b->ntot.push_back((unsigned)nat[s]);
    b->hist_off.push_back(hoff);
    for (unsigned el = 0; el < ne; ++el) b->el_count.push_back(el_count[ecf[s] + el]);
    hoff += (unsigned long long)ntypes_of(ne) * nh;
  }
  return plan_finish<T>(ctx, b, ftiles, fboxes, types, hoff, out, nullptr);
}

// Histogram launches of a plan (zero fills + kernels): no host work, capturable.
template <class T>
static int plan_hist(xs_batch *b, unsigned long long *hist_d, bool prof) {
  xs_ctx *ctx = b->ctx;
  const unsigned zb = (unsigned)std::min<unsigned long long>((b->hist_len + 255) / 256, 8192);
  hipLaunchKernelGGL(fill_u64, dim3(std::max(zb, 1u)), dim3(256), 0, ctx->stream, hist_d,
                     (size_t)b->hist_len);
  hipLaunchKernelGGL(fill_u64, dim3(1), dim3(256), 0, ctx->stream, b->ovf, (size_t)b->nstruct);
  XS_HIP_CHECK(hipGetLastError());
  const T *c = static_cast<const T *>(b->coords);
  const Item *items = static_cast<const Item *>(b->items);
  Args<T> A{c,
            c + b->natoms,
            c + 2 * b->natoms,
            static_cast<const Tile *>(b->tiles),
            static_cast<const double *>(b->boxes),
            static_cast<const TypeDesc *>(b->types),
            items,
            hist_d,
            b->ovf,
            static_cast<const double *>(b->recs)};
  A.quot_f32 = b->quot_f32;
  A.narrow = std::all_of(b->nhist.begin(), b->nhist.end(), [](unsigned h) { return h <= kNarrowMax; });
  XS_TRY(launch_hist<T>(ctx, A, b->items8192, 8192, prof));
  A.items = items + b->items8192;
  XS_TRY(launch_hist<T>(ctx, A, b->items32768, 32768, prof));
  return XS_OK;
}

}  // namespace hist

void destroy_batch(xs_batch *b) { hist::free_plan(b); }

int batch_hist_launch(xs_batch *b, unsigned long long *hist_d, bool prof) {
  return b->f64 ? hist::plan_hist<double>(b, hist_d, prof) : hist::plan_hist<float>(b, hist_d, prof);
}

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
  if (n_atoms == 0) {
    const size_t hsize = (size_t)ntypes_of(n_el) * nhist;
    if (device) {
      XS_HIP_CHECK(hipMemsetAsync(hist, 0, hsize * 8, ctx->stream));
      if (overflow) XS_HIP_CHECK(hipMemsetAsync(overflow, 0, 8, ctx->stream));
    } else {
      std::memset(hist, 0, hsize * 8);
      if (overflow) *overflow = 0;
    }
    return XS_OK;
  }
  auto *h = reinterpret_cast<unsigned long long *>(hist);
  auto *o = reinterpret_cast<unsigned long long *>(overflow);
  bool quot = false;
  if (!f64) XS_TRY(hist::sqrt_f32_checked(ctx, bin));
  if (!f64 && !(flags & XS_F32_DOUBLE_QUOTIENT)) XS_TRY(hist::quot_f32_exact(ctx, bin, &quot));
  if (device) {
    if (f64)
      return hist::run_device<double>(ctx, static_cast<const double *>(xyz), n_atoms, el_count, n_el,
                                      bin, nhist, h, o, false);
    return hist::run_device<float>(ctx, static_cast<const float *>(xyz), n_atoms, el_count, n_el, bin,
                                   nhist, h, o, quot);
  }
  if (f64)
    return hist::run_host<double>(ctx, static_cast<const double *>(xyz), n_atoms, el_count, n_el, bin,
                                  nhist, hist, overflow, false);
  return hist::run_host<float>(ctx, static_cast<const float *>(xyz), n_atoms, el_count, n_el, bin,
                               nhist, hist, overflow, quot);
}

extern "C" int xs_batch_plan(xs_ctx *ctx, unsigned nstruct, const unsigned *n_el,
                             const unsigned *el_count, const double *const *xyz, double bin,
                             const unsigned *nhist, unsigned flags, xs_batch **plan) {
  XS_TRY(activate(ctx));
  if (!plan || !n_el || !el_count || !xyz) return fail(XS_ERR_ARG, "xs_batch_plan: null pointer");
  *plan = nullptr;
  if (nstruct == 0) return fail(XS_ERR_ARG, "xs_batch_plan: no structures");
  if (!(bin > 0.0) || !std::isfinite(bin)) return fail(XS_ERR_ARG, "xs_batch_plan: bin must be > 0");
  for (unsigned s = 0; s < nstruct; ++s)
    if (!xyz[s]) return fail(XS_ERR_ARG, "xs_batch_plan: null structure");
  const bool device = flags & XS_DEVICE_PTRS;
  if (flags & XS_F64_BINNING)
    return device ? hist::plan_build_device<double>(ctx, nstruct, n_el, el_count, xyz, bin, nhist, true,
                                                    false, plan)
                  : hist::plan_build<double>(ctx, nstruct, n_el, el_count, xyz, bin, nhist, true, false,
                                             plan);
  bool quot = false;
  XS_TRY(hist::sqrt_f32_checked(ctx, bin));
  if (!(flags & XS_F32_DOUBLE_QUOTIENT)) XS_TRY(hist::quot_f32_exact(ctx, bin, &quot));
  return device ? hist::plan_build_device<float>(ctx, nstruct, n_el, el_count, xyz, bin, nhist, false,
                                                 quot, plan)
                : hist::plan_build<float>(ctx, nstruct, n_el, el_count, xyz, bin, nhist, false, quot,
                                          plan);
}

extern "C" int xs_batch_free(xs_batch *plan) {
  if (!plan) return XS_OK;
  (void)hipSetDevice(plan->ctx->device);
  (void)hipStreamSynchronize(plan->ctx->stream);
  hist::free_plan(plan);
  return XS_OK;
}

extern "C" int xs_batch_info(const xs_batch *plan, unsigned *nhist, unsigned long long *hist_off,
                             unsigned long long *pdf_off, unsigned long long *type_first,
                             unsigned long long *totals) {
  if (!plan) return fail(XS_ERR_ARG, "xs_batch_info: null plan");
  for (unsigned s = 0; s < plan->nstruct; ++s) {
    if (nhist) nhist[s] = plan->nhist[s];
    if (hist_off) hist_off[s] = plan->hist_off[s];
    if (pdf_off) pdf_off[s] = plan->pdf_off[s];
    if (type_first) type_first[s] = plan->type_first[s];
  }
  if (totals) {
    totals[0] = plan->hist_len;
    totals[1] = plan->pdf_len;
    totals[2] = plan->ntypes_total;
  }
  return XS_OK;
}
