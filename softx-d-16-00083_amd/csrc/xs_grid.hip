// Shared device helpers: owned allocations, synchronous copies, hipcub radix
// sort of uint32 pairs, and the uniform cell grid that the ensemble builder
// (image merging) and RearrangementInt (cut-off pairs) search.
//
// Grid sizing: a pair closer than d lies in the same or adjacent cells when
// the cell edge h >= d; the edge is d * 1.01 (the margin covers the rounding
// of the cell coordinates floor((v - lo) / h), computed as (v - lo) * (1/h)
// for both points of a pair), grown by 1.25x until the grid has at most
// max(2^16, 8 n) and 2^24 cells.  The cell of a point is clamped to the grid;
// points lie inside the bounding box, so the clamp never moves them.
#include "xs_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace xs {

namespace {

constexpr int kThreads = 256;
constexpr int kBboxBlocks = 512;

inline unsigned nblk(size_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

__global__ __launch_bounds__(kThreads) void bbox_partial(const double *x, const double *y,
                                                         const double *z, size_t n,
                                                         double *part) {
  __shared__ double red[6][kThreads];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kThreads) {
    const double v[3] = {x[i], y[i], z[i]};
    for (int k = 0; k < 3; ++k) {
      lo[k] = fmin(lo[k], v[k]);
      hi[k] = fmax(hi[k], v[k]);
    }
  }
  for (int k = 0; k < 3; ++k) {
    red[k][threadIdx.x] = lo[k];
    red[3 + k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + w]);
        red[3 + k][threadIdx.x] = fmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void cell_keys(const double *x, const double *y, const double *z, size_t n, Grid G,
                          uint32_t *key, uint32_t *idx) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int cx = min(max(grid_coord(x[i], G.lo[0], G.inv, G.g[0]), 0), G.g[0] - 1);
  const int cy = min(max(grid_coord(y[i], G.lo[1], G.inv, G.g[1]), 0), G.g[1] - 1);
  const int cz = min(max(grid_coord(z[i], G.lo[2], G.inv, G.g[2]), 0), G.g[2] - 1);
  key[i] = grid_cell(G, cx, cy, cz);
  idx[i] = (uint32_t)i;
}

__global__ void fill_zero_u32(uint32_t *p, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void cell_ranges(const uint32_t *skey, size_t n, uint32_t *beg, uint32_t *end) {
  const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t k = skey[p];
  if (p == 0 || skey[p - 1] != k) beg[k] = (uint32_t)p;
  if (p + 1 == n || skey[p + 1] != k) end[k] = (uint32_t)(p + 1);
}

__global__ void gather_points(const uint32_t *order, size_t n, const double *x, const double *y,
                              const double *z, double *X, double *Y, double *Z) {
  const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t i = order[p];
  X[p] = x[i];
  Y[p] = y[i];
  Z[p] = z[i];
}

}  // namespace

int DevMem::alloc(size_t bytes, const char *what) {
  if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    return fail(XS_ERR_NOMEM,
                std::string("hipMalloc of ") + std::to_string(bytes) + " bytes (" + what + ") failed");
  }
  return XS_OK;
}

int upload(xs_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!bytes) return XS_OK;
  XS_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return XS_OK;
}

int download(xs_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!bytes) return XS_OK;
  XS_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return XS_OK;
}

int sort_pairs_u32(xs_ctx *ctx, const uint32_t *k_in, uint32_t *k_out, const uint32_t *v_in,
                   uint32_t *v_out, size_t n, int bits) {
  if (n > (size_t)INT32_MAX) return fail(XS_ERR_RANGE, "sort of more than 2^31 keys");
  size_t bytes = 0;
  XS_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k_in, k_out, v_in, v_out, (int)n,
                                                  0, bits, ctx->stream));
  DevMem tmp;
  XS_TRY(tmp.alloc(bytes, "radix sort"));
  XS_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, k_in, k_out, v_in, v_out, (int)n,
                                                  0, bits, ctx->stream));
  XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return XS_OK;
}

int bits_for(uint64_t v) {
  int b = 1;
  while (b < 32 && (v >> b)) ++b;
  return b;
}

int grid_sort(xs_ctx *ctx, const double *x, const double *y, const double *z, size_t n,
              double min_cell, SortedPoints &out) {
  if (!n) return fail(XS_ERR_ARG, "grid_sort: no points");
  if (n >= (size_t)INT32_MAX / 4) return fail(XS_ERR_RANGE, "grid_sort: too many points");
  if (!(min_cell > 0) || !std::isfinite(min_cell))
    return fail(XS_ERR_ARG, "grid_sort: cell edge must be positive and finite");
  out.n = n;
  // bounding box
  const unsigned nbb = std::min<unsigned>(kBboxBlocks, nblk(n));
  DevMem part;
  XS_TRY(part.alloc(6 * nbb * sizeof(double), "bbox"));
  bbox_partial<<<nbb, kThreads, 0, ctx->stream>>>(x, y, z, n, part.at<double>());
  XS_HIP_CHECK(hipGetLastError());
  std::vector<double> pb(6 * nbb);
  XS_TRY(download(ctx, pb.data(), part.p, pb.size() * sizeof(double)));
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (unsigned b = 0; b < nbb; ++b)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], pb[6 * b + k]);
      hi[k] = std::max(hi[k], pb[6 * b + 3 + k]);
    }
  double ext[3];
  for (int k = 0; k < 3; ++k) {
    ext[k] = hi[k] - lo[k];
    if (!std::isfinite(ext[k])) return fail(XS_ERR_RANGE, "grid_sort: coordinates are not finite");
  }
  const double cap = (double)std::min<size_t>(size_t(1) << 24, std::max<size_t>(size_t(1) << 16, 8 * n));
  double h = min_cell * 1.01;
  auto cells = [&](double inv) {
    double c = 1;
    for (int k = 0; k < 3; ++k) c *= std::floor(ext[k] * inv) + 1;
    return c;
  };
  while (cells(1.0 / h) > cap) h *= 1.25;
  Grid &G = out.G;
  G.inv = 1.0 / h;
  for (int k = 0; k < 3; ++k) {
    G.lo[k] = lo[k];
    G.g[k] = (int)std::floor(ext[k] * G.inv) + 1;
  }
  const size_t nc = (size_t)G.g[0] * G.g[1] * G.g[2];
  out.ncells = nc;
  // layout: X, Y, Z (doubles), then key, idx, skey, order, beg, end (uint32)
  const size_t o_u = align_up(3 * n * sizeof(double), 256);
  XS_TRY(out.mem.alloc(o_u + (4 * n + 2 * nc) * sizeof(uint32_t), "grid"));
  out.X = out.mem.at<double>();
  out.Y = out.X + n;
  out.Z = out.Y + n;
  uint32_t *key = out.mem.at<uint32_t>(o_u), *idx = key + n, *skey = idx + n;
  out.order = skey + n;
  out.beg = out.order + n;
  out.end = out.beg + nc;
  cell_keys<<<nblk(n), kThreads, 0, ctx->stream>>>(x, y, z, n, G, key, idx);
  XS_HIP_CHECK(hipGetLastError());
  XS_TRY(sort_pairs_u32(ctx, key, skey, idx, out.order, n, bits_for(nc)));
  fill_zero_u32<<<nblk(2 * nc), kThreads, 0, ctx->stream>>>(out.beg, 2 * nc);
  XS_HIP_CHECK(hipGetLastError());
  cell_ranges<<<nblk(n), kThreads, 0, ctx->stream>>>(skey, n, out.beg, out.end);
  gather_points<<<nblk(n), kThreads, 0, ctx->stream>>>(out.order, n, x, y, z, out.X, out.Y, out.Z);
  XS_HIP_CHECK(hipGetLastError());
  return XS_OK;
}

}  // namespace xs
