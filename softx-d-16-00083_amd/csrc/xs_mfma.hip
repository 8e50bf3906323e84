// H3 as a matrix product on the int8 matrix cores (v_mfma_i32_32x32x32_i8).
//
// S[c][q] = sum_k A[q][k] H[c][k],  A[q][k] = sin(q r_k) / (q r_k + 1e-8),
// r_k = (k + 1/2) bin, H = histogram counts (one column c per pair type).
//
// Exact-integer splitting.  A is evaluated in FP64 and split into NA signed
// int8 slices, A = sum_i a_i 2^(-6-7i) + e with |a_0| <= 64, |a_i| <= 64 and
// |e| <= 2^(-7 NA - 0) (the rounding residual of the last slice); counts are
// split exactly into NB 7-bit slices, H = sum_j b_j 2^(7j), b_j in [0, 127].
// Every slice product sum_k a_i b_j is computed exactly by the int8 MFMA in
// int32.  Products with the same d = j - i share the scale 2^(7d-6) and one
// int32 accumulator; accumulators are added into S (FP64) every 32,768 bins,
// before any can overflow (|acc| <= 5 * 64 * 127 * 32768 < 2^31).  All
// NA * NB products are kept: dropping a_4 b_0 would leave an error of
// 2^-28 per count on bins with small counts.  With NA = 5 the result differs
// from the FP64 sum by at most 2^-35 sum_k H[c][k] per entry, against the
// 2^-24-per-term of an FP32 (or split-BF16, FP32 accumulated) product; the
// cancellation of I(q) at its minima amplifies this by up to ~10^4, which
// FP32 accumulation cannot absorb.
//
// Layout.  Slices are stored pre-tiled in the MFMA operand layout: for a
// 32 x 32 (rows x k) tile, lane l holds row l % 32 and k = 16 (l / 32) + j in
// byte j of a 16-byte register group, so a wave loads a tile as one
// contiguous 1 KiB read.  A and B use the same k assignment, which is all the
// product needs.  Accumulator element 4i + jj of lane l is row
// 8i + 4 (l / 32) + jj, column l % 32.
//
// Work groups: 4 waves stacked along q (128 q) x NT tiles of 32 columns;
// tiles are ordered so that the work groups that share a q block run on the
// same XCD (consecutive work-group ids go to different XCDs).
//
// Ensembles (pairsum_mfma_batch): the columns are the pair types of many
// structures, each with its own bin count.  They come sorted and grouped by
// xs_mfma_plan.h; each group of NT column tiles has its own k range and
// count slices, one launch covers the groups of one slice count, and the
// columns' results go to their pair types' rows of S.  The slice count of a
// type comes from its pair count, so nothing is read back.
#include <algorithm>
#include <cmath>

#include "xs_internal.h"
#include "xs_sincos.h"

namespace xs {
namespace mm {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kNA = kMfmaSlicesA;        // sinc slices
constexpr int kMaxNB = kMfmaMaxSlicesB;  // count slices (counts < 2^35)
constexpr unsigned kFlush = 1024;  // k tiles (of 32 bins) per int32 accumulation
constexpr double kPi = 3.14159265358979323846;

__global__ void max_u64(const unsigned long long *__restrict__ h, size_t n,
                        unsigned long long *__restrict__ out) {
  unsigned long long m = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    m = max(m, h[i]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned long long)__shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0 && m) atomicMax(out, m);
}

__global__ void zero_u64(unsigned long long *p) { *p = 0ull; }

// One thread per (q tile, k tile, lane): 16 sinc values, NA slices of 16 bytes.
__global__ void slice_sinc(const double *__restrict__ q, unsigned nq, unsigned nhist, double bin,
                           unsigned MT, unsigned KT, v4i *__restrict__ As) {
  const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t total = (size_t)MT * KT * 64;
  if (tid >= total) return;
  const unsigned lane = tid & 63;
  const size_t tile = tid >> 6;
  const unsigned kt = tile % KT;
  const unsigned mt = tile / KT;
  const unsigned row = mt * 32 + (lane & 31);
  const unsigned k0 = kt * 32 + 16 * (lane >> 5);
  int8_t sl[kNA][16];
  const double qv = row < nq ? q[row] : 0.0;
  double s = 0, c = 1, ds = 0, dc = 1;
  if (qv != 0.0) {
    const double turns = qv * bin / (2.0 * kPi);
    sincos_2pi(((double)k0 + 0.5) * turns, &s, &c);
    sincos_2pi(turns, &ds, &dc);
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const unsigned k = k0 + j;
    double v = 0.0;
    if (qv != 0.0 && k < nhist) v = s / (qv * (((double)k + 0.5) * bin) + 1e-8);
    const double sn = fma(s, dc, c * ds);
    c = fma(c, dc, -(s * ds));
    s = sn;
    v *= 64.0;
#pragma unroll
    for (int i = 0; i < kNA; ++i) {
      const double a = rint(v);
      sl[i][j] = (int8_t)a;
      v = (v - a) * 128.0;
    }
  }
#pragma unroll
  for (int i = 0; i < kNA; ++i) {
    v4i w;
    __builtin_memcpy(&w, sl[i], 16);
    As[(size_t)i * total + tid] = w;
  }
}

// One thread per (column tile, k tile, lane): 16 counts, NB 7-bit slices.
__global__ void slice_counts(const unsigned long long *__restrict__ hist, unsigned ncols,
                             unsigned nhist, unsigned NTn, unsigned KT, int nb,
                             v4i *__restrict__ Bs) {
  const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t total = (size_t)NTn * KT * 64;
  if (tid >= total) return;
  const unsigned lane = tid & 63;
  const size_t tile = tid >> 6;
  const unsigned kt = tile % KT;
  const unsigned nt = tile / KT;
  const unsigned col = nt * 32 + (lane & 31);
  const unsigned k0 = kt * 32 + 16 * (lane >> 5);
  unsigned long long cnt[16];
#pragma unroll
  for (int j = 0; j < 16; ++j)
    cnt[j] = (col < ncols && k0 + j < nhist) ? hist[(size_t)col * nhist + k0 + j] : 0ull;
  for (int s = 0; s < nb; ++s) {
    int8_t b[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) b[j] = (int8_t)((cnt[j] >> (7 * s)) & 127ull);
    v4i w;
    __builtin_memcpy(&w, b, 16);
    Bs[(size_t)s * total + tid] = w;
  }
}

// Counts of an ensemble's pair types: one thread per (group, column tile t
// of the group, k tile below kt_max, lane); threads past their group's k
// range return.  Group g's slices go to b_off in the layout [nb][NT][kt][64];
// columns past ncols and bins past a column's nhist are zero.
__global__ void slice_counts_batch(const unsigned long long *__restrict__ hist,
                                   const MfmaCol *__restrict__ cols, unsigned ncols,
                                   const MfmaTile *__restrict__ tiles, unsigned tile0,
                                   unsigned ntiles, unsigned kt_max, int nb, int NT,
                                   v4i *__restrict__ Bs) {
  const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t per_t = (size_t)kt_max * 64, per_group = per_t * NT;
  if (tid >= per_group * ntiles) return;
  const unsigned g = (unsigned)(tid / per_group);
  const size_t r = tid % per_group;
  const unsigned t = (unsigned)(r / per_t);
  const unsigned kt = (unsigned)((r % per_t) >> 6);
  const unsigned lane = tid & 63;
  const MfmaTile d = tiles[tile0 + g];
  if (kt >= d.kt) return;
  const unsigned cg = t * 32 + (lane & 31);  // column within the group
  const unsigned k0 = kt * 32 + 16 * (lane >> 5);
  unsigned long long cnt[16];
  MfmaCol c{0ull, 0u, 0u};
  if (cg < d.ncol && d.col0 + cg < ncols) c = cols[d.col0 + cg];
#pragma unroll
  for (int j = 0; j < 16; ++j) cnt[j] = k0 + j < c.nhist ? hist[c.hist_off + k0 + j] : 0ull;
  for (int s = 0; s < nb; ++s) {
    int8_t b[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) b[j] = (int8_t)((cnt[j] >> (7 * s)) & 127ull);
    v4i w;
    __builtin_memcpy(&w, b, 16);
    Bs[d.b_off + ((size_t)(s * NT + t) * d.kt + kt) * 64 + lane] = w;
  }
}

// The product.  A work group is 4 waves stacked along q (4 x 32 q) over NT
// column tiles.  Per k tile it needs 4 NA A tiles (each wave its own) and
// NB NT B tiles (the same for all 4 waves), 1 KiB each in the pre-tiled
// layout, so a global_load_lds_dwordx4 of one wave writes exactly one tile
// (lane l's 16 bytes at 16 l).  The tiles go through three LDS stages: those
// of k tile kt + 2 are fetched while the products of kt run, and the
// products read their operands from LDS (ds_read_b128 at 16 l, no bank
// conflicts).  Each B tile is fetched once per work group, tile b by wave
// b mod 4, which costs one barrier per k tile: before it, each wave waits
// for its own loads of stage kt (vmcnt) and for its own reads of stage
// kt - 1 (lgkmcnt); after it, stage kt is complete in LDS and stage kt - 1,
// which is refilled next with k tile kt + 2, is read by no wave.
// The int32 accumulators go to S (FP64) every kFlush k tiles; a wave owns
// its output elements, so the first flush stores and later ones add.
//
// Measured on the 768-histogram workload (768 x 16,384 bins x 8,192 q):
// operands loaded straight into registers kept the MFMA pipe busy 35% of
// the time, waiting on load latency (two waves per SIMD at 250 VGPRs, no
// registers for a prefetch); per-wave LDS stages one k tile ahead, with
// every wave fetching its own copy of the B tiles (9 KiB per wave per k
// tile), 50%.
//
// BATCH: the column tiles come in groups of NT with their own k range and
// slices (tiles[tile0 + group]: the count slices at b_off, laid out
// [NB][NT][kt][64]), A is laid out with KT = the largest k range, and the
// columns go to the rows of S that cols[] names.  Otherwise all columns
// share KT and B is laid out [NB][NTn][KT][64].
template <int NA, int NB, int NT, bool BATCH>
__global__ __launch_bounds__(256, 2) void pairsum_mfma_lds_kernel(const v4i *__restrict__ As,
                                                                const v4i *__restrict__ Bs,
                                                                unsigned MT, unsigned NTn, unsigned KT,
                                                                unsigned nq, unsigned ncols,
                                                                double *__restrict__ S,
                                                                const MfmaTile *__restrict__ tiles,
                                                                const MfmaCol *__restrict__ cols,
                                                                unsigned tile0) {
  constexpr int DMIN = -(NA - 1);
  constexpr int DMAX = NB - 1;
  constexpr int NG = DMAX - DMIN + 1;
  constexpr int NBT = NB * NT;         // B tiles per k tile, shared by the waves
  constexpr int TPS = 4 * NA + NBT;    // tiles per stage
  constexpr int NS = 3;                // stages
  // loads per stage that every wave issues at least (waves b < NBT mod 4
  // issue one more; waiting for this count still covers their older stage)
  constexpr int VM = NA + NBT / 4;
  static_assert(VM <= 15, "vmcnt field");
  __shared__ v4i stage[NS][TPS][64];
  const unsigned lane = threadIdx.x & 63;
  const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned ngrp = NTn / NT;
  unsigned tile = blockIdx.x;
  if ((gridDim.x & 7u) == 0) tile = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  const unsigned mg = tile / ngrp, ng = tile % ngrp;
  const unsigned mt = mg * 4 + wave;
  const size_t sA = (size_t)MT * KT * 64;
  const v4i *ap = As + (size_t)mt * KT * 64 + lane;
  // this group's k tiles, B slice stride, B column-tile stride, B base
  unsigned KTg = KT;
  size_t sB = (size_t)NTn * KT * 64, tB = (size_t)KT * 64;
  const v4i *bp = Bs + (size_t)ng * NT * KT * 64 + lane;
  unsigned col0 = ng * NT * 32, colend = ncols;
  if constexpr (BATCH) {
    const MfmaTile d = tiles[tile0 + ng];
    KTg = d.kt;
    sB = (size_t)NT * d.kt * 64;
    tB = (size_t)d.kt * 64;
    bp = Bs + d.b_off + lane;
    col0 = d.col0;
    colend = min(ncols, d.col0 + d.ncol);
  }
  typedef __attribute__((address_space(3))) void lds_t;

  // this wave's loads of k tile kt into stage st: its NA A tiles, and the
  // B tiles b = wave, wave + 4, ... (b = j NT + t: slice j, column tile t)
  auto fetch = [&](unsigned kt, int st) {
#pragma unroll
    for (int i = 0; i < NA; ++i)
      __builtin_amdgcn_global_load_lds((const void *)(ap + i * sA + (size_t)kt * 64),
                                       (lds_t *)&stage[st][wave * NA + i][0], 16, 0, 0);
#pragma unroll
    for (int b0 = 0; b0 < NBT; b0 += 4) {
      const int b = b0 + (int)wave;
      if (b < NBT) {
        const int j = b / NT, t = b % NT;
        __builtin_amdgcn_global_load_lds((const void *)(bp + j * sB + t * tB + (size_t)kt * 64),
                                         (lds_t *)&stage[st][4 * NA + b][0], 16, 0, 0);
      }
    }
  };
  // all waves: own LDS reads done, then the barrier; the empty asm keeps the
  // compiler from moving LDS accesses across it
  auto sync_lds = [&]() {
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
  };

  for (unsigned kb = 0; kb < KTg; kb += kFlush) {
    const unsigned ke = min(KTg, kb + kFlush);
    v16i acc[NG][NT];
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[g][t] = v16i{0};
    // no wave still reads a stage of the previous block
    sync_lds();
    fetch(kb, 0);
    if (kb + 1 < ke) fetch(kb + 1, 1);
    int st = 0;
    for (unsigned kt = kb; kt < ke; ++kt) {
      if (kt + 1 < ke)
        __builtin_amdgcn_s_waitcnt(0x0f70 | VM);  // vmcnt(VM): own loads of stage kt landed
      else
        __builtin_amdgcn_s_waitcnt(0x0f70);       // vmcnt(0)
      sync_lds();
      if (kt + 2 < ke) fetch(kt + 2, st == 0 ? 2 : st - 1);
      v4i a[NA], b[NB][NT];
#pragma unroll
      for (int i = 0; i < NA; ++i) a[i] = stage[st][wave * NA + i][lane];
#pragma unroll
      for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t) b[j][t] = stage[st][4 * NA + j * NT + t][lane];
#pragma unroll
      for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
          for (int t = 0; t < NT; ++t)
            acc[j - i - DMIN][t] =
                __builtin_amdgcn_mfma_i32_32x32x32_i8(a[i], b[j][t], acc[j - i - DMIN][t], 0, 0, 0);
      st = st == 2 ? 0 : st + 1;
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const unsigned col = col0 + t * 32 + (lane & 31);
      if (col >= colend) continue;
      size_t srow = col;
      if constexpr (BATCH) srow = cols[col].type;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const unsigned row = mt * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        if (row >= nq) continue;
        double v = 0.0;
#pragma unroll
        for (int g = 0; g < NG; ++g) v = fma((double)acc[g][t][e], ldexp(1.0, 7 * (g + DMIN) - 6), v);
        double *o = S + srow * nq + row;
        *o = kb == 0 ? v : *o + v;
      }
    }
  }
}

template <int NT, bool BATCH>
static void launch_gemm(int nb, dim3 grid, hipStream_t st, const v4i *As, const v4i *Bs, unsigned MT,
                        unsigned NTn, unsigned KT, unsigned nq, unsigned ncols, double *S,
                        const MfmaTile *tiles = nullptr, const MfmaCol *cols = nullptr,
                        unsigned tile0 = 0) {
#define XS_MM_LAUNCH(NBV, NTV)                                                                       \
  hipLaunchKernelGGL((pairsum_mfma_lds_kernel<kNA, NBV, NTV, BATCH>), grid, dim3(256), 0, st, As, Bs, \
                     MT, NTn, KT, nq, ncols, S, tiles, cols, tile0)
  switch (nb) {
    case 1:
      XS_MM_LAUNCH(1, NT);
      break;
    case 2:
      XS_MM_LAUNCH(2, NT);
      break;
    case 3:
      if constexpr (NT == 1) XS_MM_LAUNCH(3, 1);  // three count slices with two column tiles would spill
      break;
    default:
      if constexpr (NT == 1) {
        if (nb == 4)
          XS_MM_LAUNCH(4, 1);
        else
          XS_MM_LAUNCH(5, 1);
      }
      break;
  }
#undef XS_MM_LAUNCH
}

}  // namespace mm

int pairsum_mfma(xs_ctx *ctx, const unsigned long long *hist, unsigned ncols, unsigned nhist,
                 double bin, const double *q, unsigned nq, double *S) {
  using namespace mm;
  const unsigned MT = (nq + 127) / 128 * 4;
  const unsigned KT = (nhist + 31) / 32;
  const unsigned NTmax = (ncols + 63) / 64 * 2;  // column tiles for either NT
  const size_t aBytes = (size_t)kNA * MT * KT * 1024, bBytes = (size_t)kMaxNB * NTmax * KT * 1024;
  XS_TRY(ensure_matrix(ctx, Carver::need({aBytes, bBytes, 8})));
  XS_TRY(ensure_pinned(ctx, 64));
  Carver cv(ctx->mat);
  v4i *As = cv.take<v4i>(aBytes / 16);
  v4i *Bs = cv.take<v4i>(bBytes / 16);
  unsigned long long *mx = cv.take<unsigned long long>(1);
  // largest count -> number of count slices (one read back)
  const size_t hn = (size_t)ncols * nhist;
  hipLaunchKernelGGL(zero_u64, dim3(1), dim3(1), 0, ctx->stream, mx);
  hipLaunchKernelGGL(max_u64, dim3((unsigned)std::min<size_t>((hn + 255) / 256, 2048)), dim3(256), 0,
                     ctx->stream, hist, hn, mx);
  XS_HIP_CHECK(hipGetLastError());
  XS_HIP_CHECK(hipMemcpyAsync(ctx->pin, mx, 8, hipMemcpyDeviceToHost, ctx->stream));
  XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  const unsigned long long cmax = *static_cast<unsigned long long *>(ctx->pin);
  int nb = 1;
  while (nb < kMaxNB && (cmax >> (7 * nb)) != 0ull) ++nb;
  if ((cmax >> (7 * nb)) != 0ull)
    return fail(XS_ERR_RANGE, "xs_hist_pairsum: a bin count exceeds 2^35 (matrix-core path)");
  // two column tiles per wave unless the accumulators would not fit
  const int NT = (ncols > 32 && nb <= 2) ? 2 : 1;  // nb = 3 with NT = 2 would spill
  const unsigned NTn = (ncols + 32 * NT - 1) / (32 * NT) * NT;

  hipEvent_t ev;
  XS_TRY(prof_begin(ctx, "pairsum_slices", &ev));
  const size_t ta = (size_t)MT * KT * 64, tb = (size_t)NTn * KT * 64;
  hipLaunchKernelGGL(slice_sinc, dim3((unsigned)((ta + 255) / 256)), dim3(256), 0, ctx->stream, q, nq,
                     nhist, bin, MT, KT, As);
  hipLaunchKernelGGL(slice_counts, dim3((unsigned)((tb + 255) / 256)), dim3(256), 0, ctx->stream, hist,
                     ncols, nhist, NTn, KT, nb, Bs);
  XS_HIP_CHECK(hipGetLastError());
  XS_TRY(prof_end(ctx, "pairsum_slices", ev));
  XS_TRY(prof_begin(ctx, "pairsum_mfma", &ev));
  const dim3 grid((MT / 4) * (NTn / NT));
  if (NT == 2)
    launch_gemm<2, false>(nb, grid, ctx->stream, As, Bs, MT, NTn, KT, nq, ncols, S);
  else
    launch_gemm<1, false>(nb, grid, ctx->stream, As, Bs, MT, NTn, KT, nq, ncols, S);
  XS_HIP_CHECK(hipGetLastError());
  XS_TRY(prof_end(ctx, "pairsum_mfma", ev));
  return XS_OK;
}

int pairsum_mfma_batch(xs_ctx *ctx, const MfmaBatchPlan &plan, const unsigned long long *hist,
                       const MfmaCol *cols_d, const MfmaTile *tiles_d, void *a_buf, void *b_buf,
                       double bin, const double *q, unsigned nq, double *S, bool prof) {
  using namespace mm;
  if (nq != plan.nq) return fail(XS_ERR_ARG, "pairsum_mfma_batch: plan made for another nq");
  const unsigned MT = (nq + 127) / 128 * 4;
  const unsigned KT = plan.kt_max;
  const unsigned ncols = (unsigned)plan.cols.size();
  v4i *As = static_cast<v4i *>(a_buf), *Bs = static_cast<v4i *>(b_buf);
  hipEvent_t ev{};
  if (prof) XS_TRY(prof_begin(ctx, "batch_pairsum", &ev));
  const size_t ta = (size_t)MT * KT * 64;
  hipLaunchKernelGGL(slice_sinc, dim3((unsigned)((ta + 255) / 256)), dim3(256), 0, ctx->stream, q, nq,
                     KT * 32, bin, MT, KT, As);
  for (const MfmaLaunch &L : plan.launches) {
    const size_t tb = (size_t)L.ntiles * L.nt * KT * 64;
    hipLaunchKernelGGL(slice_counts_batch, dim3((unsigned)((tb + 255) / 256)), dim3(256), 0,
                       ctx->stream, hist, cols_d, ncols, tiles_d, L.tile0, L.ntiles, KT, (int)L.nb,
                       (int)L.nt, Bs);
    const dim3 grid((MT / 4) * L.ntiles);
    if (L.nt == 2)
      launch_gemm<2, true>((int)L.nb, grid, ctx->stream, As, Bs, MT, L.ntiles * 2, KT, nq, ncols, S,
                           tiles_d, cols_d, L.tile0);
    else
      launch_gemm<1, true>((int)L.nb, grid, ctx->stream, As, Bs, MT, L.ntiles, KT, nq, ncols, S,
                           tiles_d, cols_d, L.tile0);
  }
  XS_HIP_CHECK(hipGetLastError());
  if (prof) XS_TRY(prof_end(ctx, "batch_pairsum", ev));
  return XS_OK;
}

}  // namespace xs
