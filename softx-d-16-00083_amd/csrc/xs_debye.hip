// H5/H6: direct Debye pair sums on gfx950.
//
// S_p(q) = sum_{pairs in p} sin(q r)/(q r + 1e-8)   (CalcFunctions.cpp:577-583)
//
// Work item = (I tile, J tile) of up to 32 x 32 atoms; tiles never cross an
// element or a structural block, so each item feeds exactly one pair type and
// one block pair (the partial sums of calcIntPartialDebye,
// CalcFunctions.cpp:662-703; the plain sum is the one-block case).
// A work group of 4 waves first computes the item's pair distances in FP64
// into LDS (r/2pi, 1/r and, on a uniform q mesh, sin/cos of the phase step
// dq r), then each lane owns 32 consecutive q and each half-wave every other
// pair of its wave's quarter: sin(q r) is seeded once per pair and lane from
// a 64-point table of the turn and a short series (sincos_2pi_tab), advanced
// one q by the stored rotation and then
// by the three-term recurrence s_{k+1} = 2 cos(dq r) s_k - s_{k-1} (one FMA;
// a rounding error made at step j grows by |sin((k-j) dq r)/sin(dq r)| <=
// k-j, so over 32 q it stays near 1e-14), so a pair-q term costs 2 FP64
// instructions plus 1/32 of the seed.  FP64 because the intensity at low q
// is a strongly cancelling difference (see xs_iq.hip).  The 1e-8 of the
// denominator is applied to first order through the pair's weight
// 1/r - (1e-8/q_c)/r^2, one weight for every 4 neighbouring q of a lane.
// The half-waves' sums are added by a cross-lane shuffle, the four waves'
// in LDS, and the result to S with FP64 atomics; a last pass divides S by q.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "xs_internal.h"
#include "xs_sincos.h"

namespace xs {
namespace debye {

constexpr int kTT = 32;       // atoms per tile
constexpr int kQPT = 32;      // q per lane
constexpr int kSub = 4;       // q per denominator weight (uniform mesh)
constexpr int kWaves = 4;
constexpr int kBlock = 64 * kWaves;
constexpr int kQBlock = 32 * kQPT;  // q per work group (a half-wave covers them)
constexpr double kTwoPi = 6.283185307179586476925286766559;
static_assert((kWaves - 1) * kQPT * 32 <= 4 * kTT * kTT, "wave sums must fit the pair buffer");

constexpr int kMaxElem = 16;
constexpr int kMaxTypes = kMaxElem * (kMaxElem + 1) / 2;

struct Tile {
  uint32_t off, cnt, el, blk;
};
struct Item {
  uint32_t ti, tj, p, part;
};

// The q mesh as found on the device (allocation-free entry point): its step
// and whether it is uniform (the test of uniform_mesh below).
struct MeshInfo {
  double dq;
  uint32_t uniform, pad;
};

// sin and cos of 2 pi r for any double r, from a table of the 64 points
// k/64 of a turn (tab[k] = (sin, cos) of 2 pi k/64, within an ulp) and a
// short series: t = r - rint(r) and u = t - n/64 with n = rint(64 t) are
// exact (u is a multiple of ulp(t), |u| <= |t|), |u| <= 1/128, so
// x = 2 pi u lies within 0.0491 rad, where sin through x^9 and cos through
// x^8 leave truncation errors below 1e-19 relative; then
// sin(a + x) = sin a cos x + cos a sin x and likewise for cos.  24 VALU
// instructions and one LDS read against 37 for the quarter-turn series of
// sincos_2pi, with the same few-ulp result.
__device__ __forceinline__ void sincos_2pi_tab(double r, const double2 *tab, double *s, double *c) {
  const double t = r - rint(r);
  const double n = rint(64.0 * t);
  const double x = kTwoPi * fma(n, -1.0 / 64.0, t);
  const double x2 = x * x;
  double ps = fma_k(x2, inv_factorial(9), -inv_factorial(7));
  ps = fma_k(ps, x2, inv_factorial(5));
  ps = fma_k(ps, x2, -inv_factorial(3));
  const double sx = fma(x * x2, ps, x);
  double pc = fma_k(x2, inv_factorial(8), -inv_factorial(6));
  pc = fma_k(pc, x2, inv_factorial(4));
  pc = fma_k(pc, x2, -inv_factorial(2));
  const double cx = fma(x2, pc, 1.0);
  const double2 a = tab[(int)n & 63];
  *s = fma(a.x, cx, a.y * sx);
  *c = fma(a.y, cx, -(a.x * sx));
}

// S[p][k] /= q[k] (0 where q[k] = 0): debye_kernel adds up
// q sin(q r)/(q r + 1e-8) and leaves the division by q to this pass.
__global__ void divide_by_q_kernel(double *__restrict__ S, const double *__restrict__ q, unsigned nq,
                                   size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double qk = q[i % nq];
    S[i] = qk == 0.0 ? 0.0 : S[i] / qk;
  }
}

// mesh == nullptr: the host decided UNIFORM and passed dq.  Otherwise both
// variants are launched and the one that does not match mesh->uniform
// returns at once (the flag is the same for the whole grid).
template <bool UNIFORM>
__global__ __launch_bounds__(kBlock) void debye_kernel(const double *__restrict__ x,
                                                       const double *__restrict__ y,
                                                       const double *__restrict__ z,
                                                       const Tile *__restrict__ tiles,
                                                       const Item *__restrict__ items,
                                                       const double *__restrict__ q, unsigned nq,
                                                       double dq, unsigned nparts,
                                                       double *__restrict__ S,
                                                       const MeshInfo *__restrict__ mesh) {
  if (mesh) {
    if ((mesh->uniform != 0) != UNIFORM) return;
    dq = mesh->dq;
  }
  __shared__ double4 pr[kTT * kTT];
  __shared__ double2 tab[64];
  __shared__ double ecs[(kQPT / kSub) * 32];  // e / q_c per (group, lane & 31)
  if (threadIdx.x < 64) {
    double s, c;
    sincos_2pi((double)threadIdx.x / 64.0, &s, &c);
    tab[threadIdx.x] = make_double2(s, c);
  }
  const Item it = items[blockIdx.x];
  const Tile A = tiles[it.ti], B = tiles[it.tj];
  const bool diag = it.ti == it.tj;
  for (unsigned m = threadIdx.x; m < kTT * kTT; m += kBlock) {
    const unsigned il = m / kTT, jl = m % kTT;
    double4 v = make_double4(0.0, 0.0, 1.0, 1.0);
    if (il < A.cnt && jl < B.cnt && (!diag || jl > il)) {
      const size_t i = A.off + il, j = B.off + jl;
      const double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r > 0.0) {
        v.x = r / kTwoPi;
        v.y = 1.0 / r;
        if (UNIFORM) {
          double s, c;
          sincos_2pi(dq * v.x, &s, &c);
          v.z = s;
          v.w = c;
        }
      }
    }
    pr[m] = v;
  }
  __syncthreads();

  // lane: 32 consecutive q of the block (lane & 31) for every other pair of
  // its wave's quarter (lane >> 5), so that the seed of the recurrence is
  // spread over 32 q.
  // The denominator: sin(q r)/(q r + e) = (sin(q r)/q) (1/r - (e/q)/r^2)
  // to first order (e = 1e-8; the second order is below 1e-16 relative for
  // q r > 0.1).  Each term is weighted by w = 1/r - (e/q)/r^2, the pair's
  // weight for its q.  On a uniform mesh one weight serves kSub neighbouring
  // q, taken at their centre q_c: (e/q)/r^2 is then off by a factor within
  // 1 +- 1.5 dq/q, so the 1e-8 correction (of order 1e-6 of I(q) where the
  // low-q intensity cancels, less elsewhere) moves by a few percent of itself
  // at the lowest q of a mesh and by less above.
  // One FMA per q for the weights of kSub q, two per term (sum, recurrence).
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6, half = lane >> 5;
  const unsigned iq0 = blockIdx.y * kQBlock + (lane & 31) * kQPT;
  double acc[kQPT];
#pragma unroll
  for (int m = 0; m < kQPT; ++m) acc[m] = 0.0;
  const double q0 = iq0 < nq ? q[iq0] : 0.0;
  double qv[UNIFORM ? 1 : kQPT], eq[UNIFORM ? 1 : kQPT];  // q and e / q (any mesh)
  if (UNIFORM) {
    // ecs[g][l]: e / q_c of group g of lane l's q
    for (unsigned m = threadIdx.x; m < (kQPT / kSub) * 32; m += kBlock) {
      const unsigned g = m / 32, l = m % 32;
      const unsigned i0 = blockIdx.y * kQBlock + l * kQPT;
      const double qc = fma((double)(g * kSub) + 0.5 * (kSub - 1), dq, i0 < nq ? q[i0] : 0.0);
      ecs[m] = qc > 0.0 ? 1e-8 / qc : 0.0;
    }
    __syncthreads();
  } else {
#pragma unroll
    for (int k = 0; k < kQPT; ++k) {
      qv[k] = iq0 + k < nq ? q[iq0 + k] : 0.0;
      eq[k] = qv[k] > 0.0 ? 1e-8 / qv[k] : 0.0;
    }
  }
  const unsigned pstart = wave * (kTT * kTT / kWaves), pend = pstart + kTT * kTT / kWaves;
  if (UNIFORM) {
    // two pairs per step, so that their recurrences (each a chain of
    // dependent FMAs) interleave; an empty pair (v.y = 0) has weight 0
    for (unsigned m = pstart + half; m < pend; m += 4) {
      const double4 va = pr[m], vb = pr[m + 2];
      const double ira = va.y, ira2 = va.y * va.y, irb = vb.y, irb2 = vb.y * vb.y;
      // e / q_c of the lane's groups from LDS; each group's two weights are
      // formed just before its kSub q, so that only two are live (126 VGPRs,
      // 4 waves per SIMD, against 164 with all 16 weights in registers)
      double ev[kQPT / kSub];
#pragma unroll
      for (int g = 0; g < kQPT / kSub; ++g) ev[g] = ecs[g * 32 + (lane & 31)];
      // s_0 = sin(q0 r), s_1 by one rotation, then the three-term
      // recurrence s_{k+1} = 2 cos(dq r) s_k - s_{k-1} (one FMA a q)
      double sa0, ca0, sb0, cb0;
      sincos_2pi_tab(q0 * va.x, tab, &sa0, &ca0);
      sincos_2pi_tab(q0 * vb.x, tab, &sb0, &cb0);
      double spa = sa0, sa = fma(sa0, va.w, ca0 * va.z);
      double spb = sb0, sb = fma(sb0, vb.w, cb0 * vb.z);
      const double c2a = 2.0 * va.w, c2b = 2.0 * vb.w;
      double wa = fma(-ev[0], ira2, ira), wb = fma(-ev[0], irb2, irb);
      acc[0] = fma(spa, wa, acc[0]);
      acc[0] = fma(spb, wb, acc[0]);
#pragma unroll
      for (int k = 1; k < kQPT; ++k) {
        if (k % kSub == 0) {
          wa = fma(-ev[k / kSub], ira2, ira);
          wb = fma(-ev[k / kSub], irb2, irb);
        }
        acc[k] = fma(sa, wa, acc[k]);
        acc[k] = fma(sb, wb, acc[k]);
        if (k + 1 < kQPT) {
          const double sna = fma(c2a, sa, -spa), snb = fma(c2b, sb, -spb);
          spa = sa;
          sa = sna;
          spb = sb;
          sb = snb;
        }
      }
    }
  } else {
    for (unsigned m = pstart + half; m < pend; m += 2) {
      const double4 v = pr[m];
      if (v.y == 0.0) continue;
      const double ir = v.y, ir2 = v.y * v.y;
      // any mesh: the weight of each q itself
#pragma unroll
      for (int k = 0; k < kQPT; ++k) {
        double s, c;
        sincos_2pi_tab(qv[k] * v.x, tab, &s, &c);
        acc[k] = fma(s, fma(-eq[k], ir2, ir), acc[k]);
      }
    }
  }
  // sum over the half-waves (lane l + 32 onto lane l, a cross-lane shuffle)
  // and the waves: waves 1..3 park their 32 x 32 sums of q sin(q r)/(q r +
  // 1e-8) in LDS (3 x 32 x 32 doubles, 24 KiB); the division by q is left to
  // divide_by_q_kernel, once per output element
  double res[kQPT];
#pragma unroll
  for (int k = 0; k < kQPT; ++k) res[k] = acc[k] + __shfl_down(acc[k], 32, 64);
  __syncthreads();
  double *park = reinterpret_cast<double *>(pr);
  if (wave > 0 && half == 0) {
#pragma unroll
    for (int k = 0; k < kQPT; ++k) park[((wave - 1) * kQPT + k) * 32 + lane] = res[k];
  }
  __syncthreads();
  if (wave == 0 && half == 0) {
#pragma unroll
    for (int k = 0; k < kQPT; ++k) {
      double t = res[k];
      for (int w = 0; w < kWaves - 1; ++w) t += park[(w * kQPT + k) * 32 + lane];
      if (iq0 + k < nq && t != 0.0)
        atomicAdd(&S[((size_t)it.p * nparts + it.part) * nq + iq0 + k], t);
    }
  }
}

static bool uniform_mesh(const std::vector<double> &q, double *dq) {
  const size_t n = q.size();
  *dq = n > 1 ? (q[n - 1] - q[0]) / (double)(n - 1) : 0.0;
  double scale = 0.0;
  for (double v : q) scale = std::max(scale, std::fabs(v));
  for (size_t k = 0; k < n; ++k)
    if (std::fabs(q[k] - (q[0] + (double)k * *dq)) > 1e-13 * scale) return false;
  return true;
}

static int run(xs_ctx *ctx, const double *xyz_in, size_t n, const unsigned *el_count, unsigned n_el,
               const unsigned *blk_count, unsigned nblocks, const double *q_in, unsigned nq,
               double *S_out, bool device) {
  const unsigned ntypes = ntypes_of(n_el);
  const unsigned nparts = nblocks * (nblocks + 1) / 2;
  const size_t ssize = (size_t)ntypes * nparts * nq;
  // tiles that stay inside one (element, block) segment
  std::vector<Tile> tiles;
  std::vector<uint32_t> el_first(n_el + 1);
  uint32_t off = 0;
  for (unsigned e = 0; e < n_el; ++e) {
    el_first[e] = (uint32_t)tiles.size();
    for (unsigned b = 0; b < nblocks; ++b) {
      const uint32_t cnt = blk_count ? blk_count[e * nblocks + b] : el_count[e];
      for (uint32_t t = 0; t < cnt; t += kTT)
        tiles.push_back(Tile{off + t, std::min<uint32_t>(kTT, cnt - t), e, b});
      off += cnt;
    }
  }
  el_first[n_el] = (uint32_t)tiles.size();
  std::vector<Item> items;
  unsigned p = 0;
  for (unsigned a = 0; a < n_el; ++a)
    for (unsigned b = a; b < n_el; ++b, ++p)
      for (uint32_t ti = el_first[a]; ti < el_first[a + 1]; ++ti)
        for (uint32_t tj = (a == b ? ti : el_first[b]); tj < el_first[b + 1]; ++tj) {
          const uint32_t bi = tiles[ti].blk, bj = tiles[tj].blk;
          const uint32_t lo = std::min(bi, bj), hi = std::max(bi, bj);
          items.push_back(Item{ti, tj, p, nblocks * lo - (lo * (lo + 1)) / 2 + hi});
        }
  std::vector<double> qh(nq);
  if (device) {
    XS_HIP_CHECK(hipMemcpyAsync(qh.data(), q_in, (size_t)nq * 8, hipMemcpyDeviceToHost, ctx->stream));
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  } else {
    std::memcpy(qh.data(), q_in, (size_t)nq * 8);
  }
  double dq = 0.0;
  const bool uniform = uniform_mesh(qh, &dq);

  const size_t tb = tiles.size() * sizeof(Tile), ib = items.size() * sizeof(Item);
  XS_TRY(ensure_workspace(ctx, Carver::need({device ? 0 : 3 * n * 8, device ? 0 : (size_t)nq * 8,
                                             device ? 0 : ssize * 8, tb, ib})));
  Carver cv(ctx->ws);
  const double *xyz = xyz_in, *qd = q_in;
  double *Sd = S_out;
  if (!device) {
    double *x2 = cv.take<double>(3 * n);
    double *q2 = cv.take<double>(nq);
    Sd = cv.take<double>(ssize);
    XS_HIP_CHECK(hipMemcpyAsync(x2, xyz_in, 3 * n * 8, hipMemcpyHostToDevice, ctx->stream));
    XS_HIP_CHECK(hipMemcpyAsync(q2, q_in, (size_t)nq * 8, hipMemcpyHostToDevice, ctx->stream));
    xyz = x2;
    qd = q2;
  }
  Tile *td = cv.take<Tile>(tiles.size());
  Item *idv = cv.take<Item>(items.size());
  XS_TRY(ensure_pinned(ctx, tb + ib + 256));
  // the pinned buffer may still feed an earlier upload: wait for the stream
  XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  std::memcpy(ctx->pin, tiles.data(), tb);
  std::memcpy(static_cast<char *>(ctx->pin) + align_up(tb, 256), items.data(), ib);
  if (tb) XS_HIP_CHECK(hipMemcpyAsync(td, ctx->pin, tb, hipMemcpyHostToDevice, ctx->stream));
  if (ib)
    XS_HIP_CHECK(hipMemcpyAsync(idv, static_cast<char *>(ctx->pin) + align_up(tb, 256), ib,
                                hipMemcpyHostToDevice, ctx->stream));
  XS_HIP_CHECK(hipMemsetAsync(Sd, 0, ssize * 8, ctx->stream));
  if (!items.empty()) {
    hipEvent_t ev;
    XS_TRY(prof_begin(ctx, "debye", &ev));
    dim3 grid((unsigned)items.size(), (nq + kQBlock - 1) / kQBlock);
    if (uniform)
      hipLaunchKernelGGL(debye_kernel<true>, grid, dim3(kBlock), 0, ctx->stream, xyz, xyz + n,
                         xyz + 2 * n, td, idv, qd, nq, dq, nparts, Sd, nullptr);
    else
      hipLaunchKernelGGL(debye_kernel<false>, grid, dim3(kBlock), 0, ctx->stream, xyz, xyz + n,
                         xyz + 2 * n, td, idv, qd, nq, dq, nparts, Sd, nullptr);
    hipLaunchKernelGGL(divide_by_q_kernel, dim3((unsigned)std::min<size_t>((ssize + 255) / 256, 4096)),
                       dim3(256), 0, ctx->stream, Sd, qd, nq, ssize);
    XS_HIP_CHECK(hipGetLastError());
    XS_TRY(prof_end(ctx, "debye", ev));
  }
  if (!device) {
    XS_HIP_CHECK(hipMemcpyAsync(S_out, Sd, ssize * 8, hipMemcpyDeviceToHost, ctx->stream));
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  } else {
    // keep the pinned staging alive until the uploads have been consumed
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
  return XS_OK;
}

// ---------------------------------------------------------------- allocation-free entry point
//
// xs_debye_pairsum_noalloc runs the same kernels with their tile and item
// tables built on the device in `scratch`, a file-scope device array, so it
// allocates nothing, copies nothing to or from the host and never waits:
// mesh_info_kernel finds dq and whether the mesh is uniform, table_kernel
// writes the tables, fill_kernel clears S, and both debye_kernel variants are
// launched (the one that does not match the mesh returns at once).  Host code
// must not pass `scratch` itself to a kernel (that is a host-side stand-in,
// not a device address): scratch_ptr() gives the device address, taken once
// per device with hipGetSymbolAddress.  Every call on a device shares its
// scratch, so calls that overlap on one device race.

constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// The device address of `scratch` on the current device (which must be
// `device`), looked up on the first call for that device.
static int scratch_ptr(int device, void **out) {
  if (device < 0 || device >= kMaxDevices)
    return fail(XS_ERR_DEVICE, "xs_debye_pairsum_noalloc: device id beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(XS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return XS_OK;
}

// Tables of one structure (one structural block), in the order run() builds
// them on the host: tiles of 32 atoms per element; items per pair type
// p = (a <= b), I tile major, J tiles from the I tile on (a == b) or all.
struct TableArgs {
  uint32_t n_el, ntiles;
  unsigned long long nitems;
  uint32_t el_off[kMaxElem + 1];     // first atom of element e
  uint32_t tile_base[kMaxElem + 1];  // first tile of element e
  unsigned long long item_off[kMaxTypes + 1];  // first item of pair type p
};

__global__ void mesh_info_kernel(const double *__restrict__ q, unsigned nq, MeshInfo *mesh) {
  __shared__ double smax[256];
  __shared__ int sbad[256];
  const double dq = nq > 1 ? (q[nq - 1] - q[0]) / (double)(nq - 1) : 0.0;
  double m = 0.0;
  for (unsigned k = threadIdx.x; k < nq; k += blockDim.x) m = fmax(m, fabs(q[k]));
  smax[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
    __syncthreads();
  }
  const double scale = smax[0];
  int bad = 0;
  for (unsigned k = threadIdx.x; k < nq; k += blockDim.x)
    bad |= fabs(q[k] - (q[0] + (double)k * dq)) > 1e-13 * scale;
  sbad[threadIdx.x] = bad;
  __syncthreads();
  if (threadIdx.x == 0) {
    int any = 0;
    for (int t = 0; t < 256; ++t) any |= sbad[t];
    mesh->dq = dq;
    mesh->uniform = any ? 0u : 1u;
    mesh->pad = 0u;
  }
}

__global__ void table_kernel(TableArgs A, Tile *__restrict__ tiles, Item *__restrict__ items) {
  const unsigned long long k = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
  if (k < A.ntiles) {
    uint32_t e = 0;
    while (e + 1 < A.n_el && A.tile_base[e + 1] <= k) ++e;
    const uint32_t t = (uint32_t)k - A.tile_base[e];
    const uint32_t cnt = A.el_off[e + 1] - A.el_off[e];
    tiles[k] = Tile{A.el_off[e] + t * kTT, min((uint32_t)kTT, cnt - t * kTT), e, 0u};
  }
  if (k >= A.nitems) return;
  uint32_t p = 0, a = 0, b = 0;
  // pair type p = (a, b) of item k: types in the order (0,0),(0,1),...,(1,1),...
  for (;;) {
    if (A.item_off[p + 1] > k) break;
    ++p;
    if (++b == A.n_el) b = ++a;
  }
  const unsigned long long r = k - A.item_off[p];
  const uint64_t Ta = A.tile_base[a + 1] - A.tile_base[a];
  uint64_t i, j;
  if (a == b) {
    // row i holds Ta - i items (J tiles i..Ta-1); rows before i hold
    // c(i) = i Ta - i (i - 1) / 2
    const double T2 = 2.0 * (double)Ta + 1.0;
    double g = floor(0.5 * (T2 - sqrt(fmax(T2 * T2 - 8.0 * (double)r, 0.0))));
    i = g < 0.0 ? 0 : (uint64_t)g;
    if (i >= Ta) i = Ta - 1;
    auto c = [Ta](uint64_t v) { return v * Ta - v * (v - 1) / 2; };
    while (i > 0 && c(i) > r) --i;
    while (i + 1 < Ta && c(i + 1) <= r) ++i;
    j = i + (r - c(i));
  } else {
    const uint64_t Tb = A.tile_base[b + 1] - A.tile_base[b];
    i = r / Tb;
    j = r % Tb;
  }
  items[k] = Item{A.tile_base[a] + (uint32_t)i, A.tile_base[b] + (uint32_t)j, p, 0u};
}

__global__ void fill_kernel(double *__restrict__ p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

static int check_args(const double *xyz, size_t n, const unsigned *el_count, unsigned n_el,
                      const double *q, unsigned nq, double *S) {
  if (!el_count || !q || !S || (!xyz && n)) return fail(XS_ERR_ARG, "debye: null pointer");
  if (n_el == 0 || n_el > 16) return fail(XS_ERR_RANGE, "debye: n_el must be in [1, 16]");
  size_t tot = 0;
  for (unsigned e = 0; e < n_el; ++e) tot += el_count[e];
  if (tot != n) return fail(XS_ERR_ARG, "debye: sum(el_count) != n_atoms");
  if (nq == 0) return fail(XS_ERR_ARG, "debye: nq must be > 0");
  return XS_OK;
}

// Layout of the scratch for one call: MeshInfo, tiles, items.
static size_t scratch_need(size_t ntiles, unsigned long long nitems, size_t *tiles_at,
                           size_t *items_at) {
  *tiles_at = 256;
  *items_at = align_up(*tiles_at + ntiles * sizeof(Tile), 256);
  return *items_at + (size_t)nitems * sizeof(Item);
}

static int run_noalloc(xs_ctx *ctx, const double *xyz, size_t n, const unsigned *el_count,
                       unsigned n_el, const double *q, unsigned nq, double *S) {
  const unsigned ntypes = ntypes_of(n_el);
  TableArgs A{};
  A.n_el = n_el;
  uint64_t tiles = 0, off = 0;
  for (unsigned e = 0; e < n_el; ++e) {
    A.el_off[e] = (uint32_t)off;
    A.tile_base[e] = (uint32_t)tiles;
    off += el_count[e];
    tiles += (el_count[e] + kTT - 1) / kTT;
  }
  A.el_off[n_el] = (uint32_t)off;
  A.tile_base[n_el] = (uint32_t)tiles;
  A.ntiles = (uint32_t)tiles;
  unsigned long long items = 0;
  unsigned p = 0;
  for (unsigned a = 0; a < n_el; ++a)
    for (unsigned b = a; b < n_el; ++b, ++p) {
      const uint64_t ta = A.tile_base[a + 1] - A.tile_base[a];
      const uint64_t tb = A.tile_base[b + 1] - A.tile_base[b];
      A.item_off[p] = items;
      items += a == b ? ta * (ta + 1) / 2 : ta * tb;
    }
  A.item_off[ntypes] = items;
  A.nitems = items;
  size_t tiles_at, items_at;
  const size_t need = scratch_need((size_t)tiles, items, &tiles_at, &items_at);
  if (need > kScratchBytes || items >= (1ull << 31))
    return fail(XS_ERR_RANGE, "xs_debye_pairsum_noalloc: " + std::to_string(tiles) + " tiles and " +
                                  std::to_string(items) + " tile pairs need " + std::to_string(need) +
                                  " bytes of tables, beyond the " + std::to_string(kScratchBytes) +
                                  "-byte scratch (about 65,000 atoms); use xs_debye_pairsum");
  void *base = nullptr;
  XS_TRY(scratch_ptr(ctx->device, &base));
  char *sp = static_cast<char *>(base);
  auto *mesh = reinterpret_cast<MeshInfo *>(sp);
  auto *td = reinterpret_cast<Tile *>(sp + tiles_at);
  auto *idv = reinterpret_cast<Item *>(sp + items_at);
  hipStream_t st = ctx->stream;
  const size_t ssize = (size_t)ntypes * nq;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)std::min<size_t>((ssize + 255) / 256, 4096)),
                     dim3(256), 0, st, S, ssize);
  hipLaunchKernelGGL(mesh_info_kernel, dim3(1), dim3(256), 0, st, q, nq, mesh);
  const unsigned long long nt = std::max<unsigned long long>(items, tiles);
  if (nt) {
    hipLaunchKernelGGL(table_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, A, td,
                       idv);
  }
  if (items) {
    dim3 grid((unsigned)items, (nq + kQBlock - 1) / kQBlock);
    hipLaunchKernelGGL(debye_kernel<true>, grid, dim3(kBlock), 0, st, xyz, xyz + n, xyz + 2 * n, td,
                       idv, q, nq, 0.0, 1u, S, mesh);
    hipLaunchKernelGGL(debye_kernel<false>, grid, dim3(kBlock), 0, st, xyz, xyz + n, xyz + 2 * n, td,
                       idv, q, nq, 0.0, 1u, S, mesh);
    hipLaunchKernelGGL(divide_by_q_kernel, dim3((unsigned)std::min<size_t>((ssize + 255) / 256, 4096)),
                       dim3(256), 0, st, S, q, nq, ssize);
  }
  XS_HIP_CHECK(hipGetLastError());
  return XS_OK;
}

}  // namespace debye

// Called by xs_init with the context's device current.
int debye_scratch_init(int device) {
  void *p = nullptr;
  return debye::scratch_ptr(device, &p);
}

}  // namespace xs

using namespace xs;

extern "C" int xs_debye_pairsum(xs_ctx *ctx, const double *xyz, size_t n_atoms,
                                const unsigned *el_count, unsigned n_el, const double *q,
                                unsigned nq, double *S, unsigned flags) {
  XS_TRY(activate(ctx));
  XS_TRY(debye::check_args(xyz, n_atoms, el_count, n_el, q, nq, S));
  return debye::run(ctx, xyz, n_atoms, el_count, n_el, nullptr, 1, q, nq, S,
                    flags & XS_DEVICE_PTRS);
}

extern "C" int xs_partial_debye_pairsum(xs_ctx *ctx, const double *xyz, size_t n_atoms,
                                        const unsigned *el_count, unsigned n_el,
                                        const unsigned *blk_count, unsigned nblocks,
                                        const double *q, unsigned nq, double *S,
                                        unsigned flags) {
  XS_TRY(activate(ctx));
  XS_TRY(debye::check_args(xyz, n_atoms, el_count, n_el, q, nq, S));
  if (!blk_count || nblocks == 0) return fail(XS_ERR_ARG, "partial debye: no blocks");
  for (unsigned e = 0; e < n_el; ++e) {
    size_t t = 0;
    for (unsigned b = 0; b < nblocks; ++b) t += blk_count[e * nblocks + b];
    if (t != el_count[e]) return fail(XS_ERR_ARG, "partial debye: blk_count does not sum to el_count");
  }
  return debye::run(ctx, xyz, n_atoms, el_count, n_el, blk_count, nblocks, q, nq, S,
                    flags & XS_DEVICE_PTRS);
}

extern "C" int xs_debye_pairsum_noalloc(xs_ctx *ctx, const double *xyz, size_t n_atoms,
                                        const unsigned *el_count, unsigned n_el, const double *q,
                                        unsigned nq, double *S, unsigned flags) {
  XS_TRY(activate(ctx));
  XS_TRY(debye::check_args(xyz, n_atoms, el_count, n_el, q, nq, S));
  if (!(flags & XS_DEVICE_PTRS))
    return fail(XS_ERR_ARG, "xs_debye_pairsum_noalloc: needs device pointers (XS_DEVICE_PTRS)");
  if (nq > 65535u * (unsigned)debye::kQBlock)
    return fail(XS_ERR_RANGE, "xs_debye_pairsum_noalloc: nq too large");
  return debye::run_noalloc(ctx, xyz, n_atoms, el_count, n_el, q, nq, S);
}
