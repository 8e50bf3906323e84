// H3 histogram -> pair sums and H4 RDF/PDF/rPDF on gfx950.
//
// H3: S_p(q) = sum_k h_p[k] sin(q r_k)/(q r_k), r_k = (k+1/2) bin.
//   One thread per q, one work group per (256 q, 2048-bin chunk, pair type).
//   The chunk's counts (as double, exact below 2^53) and 1/(k+1/2) are staged
//   in LDS; every lane reads the same LDS word (broadcast) and empty bins are
//   skipped uniformly.  The terms are FP64: at low q the intensity is a
//   difference of the pair sum and the self term that cancels by up to four
//   orders of magnitude, which FP32 terms (6e-8 relative each) cannot resolve
//   to 1e-5.  A bin costs four FP64 instructions per q: sin((k+1/2) q bin)
//   is advanced from bin to bin by Reinsch's form of the three-term
//   recurrence (an add and an FMA), seeded once per chunk from the series in
//   xs_sincos.h, and the term's weight and sum take one FMA each.
//   Chunk sums are scaled by 1/(q bin) and added into the result with FP64
//   atomics.  The reference's denominator is q r + 1e-8 (CalcFunctions.cpp:861);
//   1/(x + e) = (1/x)(1 - e/x) + O((e/x)^2) is applied per term, because at the
//   cancelling low-q points that 1e-8 shifts I(q) by ~1e-6 relative.
//   q = 0 gives S = 0 as the reference does (sin 0 = 0).
// H4: the reference's PDF formulas (CalcFunctions.cpp:919-957) in FP64, one
//   thread per bin, same operation order and no contraction, so the result
//   equals the FP64 CPU path bit for bit for the same histogram.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "xs_internal.h"
#include "xs_sincos.h"

namespace xs {
namespace iq {

constexpr int kQBlock = 256;
constexpr int kChunk = (int)kPairsumChunk;
constexpr double kPi = 3.14159265358979323846;

__global__ void fill_f64(double *__restrict__ p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// Sum over one staged chunk of bins [k0, k0 + kn) for one q (qv != 0).
// bins[b] = (h y, h y^2) with y = 1/(k0 + b + 1/2), so that the term
// h y (1 - eps_q y) sin((k + 1/2) q bin) is one FMA for its weight and one
// to add it.  The sine s_k = sin((k + 1/2) t), t = q bin, is advanced by
// Reinsch's form of the three-term recurrence (Stoer and Bulirsch,
// Introduction to Numerical Analysis, sec. 2.3.3): with the difference
// d_k = s_{k+1} - s_k and lam = -4 sin^2(t/2) = 2 cos t - 2,
//   s_{k+1} = s_k + d_k,   d_{k+1} = d_k + lam s_{k+1}
// (one add, one FMA), seeded with d_{k0} = 2 sin(t/2) cos((k0 + 1) t).  The
// plain recurrence s_{k+1} = 2 cos t s_k - s_{k-1} rounds 2 cos t, whose
// error shifts the frequency by ~eps cot t: over a 2048-bin chunk its sines
// drifted by up to 2e-11 at t = 1e-4..1e-3 (an FP64 simulation against
// extended precision), against 4e-16..2e-15 for this form and 2e-14 for a
// rotation of (sin, cos) by t (four FMAs).  Two accumulators (even and odd
// bins) halve the dependent chain of the sums.
__device__ __forceinline__ double pairsum_chunk(const double2 *bins, unsigned k0, unsigned kn,
                                                double qv, double bin) {
  const double turns = qv * bin / (2.0 * kPi);  // phase step per bin, in turns
  double s, c, h, g;
  sincos_2pi(((double)k0 + 0.5) * turns, &s, &c);
  sincos_2pi(0.5 * turns, &h, &g);
  const double lam = -4.0 * h * h;
  double d = 2.0 * h * fma(c, g, -(s * h));  // 2 sin(t/2) cos((k0 + 1) t)
  const double eps_q = 1e-8 / (qv * bin);
  double acc0 = 0.0, acc1 = 0.0;
  unsigned b = 0;
  for (; b + 2 <= kn; b += 2) {
    const double2 e0 = bins[b], e1 = bins[b + 1];
    acc0 = fma(fma(-eps_q, e0.y, e0.x), s, acc0);
    s += d;
    d = fma(lam, s, d);
    acc1 = fma(fma(-eps_q, e1.y, e1.x), s, acc1);
    s += d;
    d = fma(lam, s, d);
  }
  if (b < kn) {
    const double2 e0 = bins[b];
    acc0 = fma(fma(-eps_q, e0.y, e0.x), s, acc0);
  }
  return acc0 + acc1;
}

__device__ __forceinline__ void stage_chunk(double2 *bins, const unsigned long long *__restrict__ h,
                                            unsigned k0, unsigned kn) {
  for (unsigned b = threadIdx.x; b < kn; b += kQBlock) {
    const double y = 1.0 / ((double)(k0 + b) + 0.5), hy = (double)h[b] * y;
    bins[b] = make_double2(hy, hy * y);
  }
}

__global__ __launch_bounds__(kQBlock) void pairsum_kernel(const unsigned long long *__restrict__ hist,
                                                          unsigned nhist, const double *__restrict__ q,
                                                          unsigned nq, double bin,
                                                          double *__restrict__ S) {
  __shared__ double2 bins[kChunk];
  const unsigned p = blockIdx.z;
  const unsigned k0 = blockIdx.y * kChunk;
  const unsigned kn = min((unsigned)kChunk, nhist - k0);
  stage_chunk(bins, hist + (size_t)p * nhist + k0, k0, kn);
  __syncthreads();
  const unsigned iq = blockIdx.x * kQBlock + threadIdx.x;
  if (iq >= nq) return;
  const double qv = q[iq];
  if (qv == 0.0) return;
  const double acc = pairsum_chunk(bins, k0, kn, qv, bin);
  if (acc != 0.0) atomicAdd(&S[(size_t)p * nq + iq], acc / (qv * bin));
}

// The plain FP32 form (XS_PAIRSUM_F32), for comparison: every term
// h sin(q r)/(q r + 1e-8) evaluated in FP32 with the IEEE-accurate sinf, as
// the reference GPU kernels evaluate theirs in FP32 (CalcFunctionsCUDA.cu:
// 923-985; they take __sinf and 1e-6f, here sinf and the CPU path's 1e-8 so
// that the difference to FP64 is the FP32 arithmetic alone), summed in FP32
// over each 2048-bin chunk; chunk sums go to S with atomics.  Counts are
// rounded to float (exact below 2^24).
__global__ __launch_bounds__(kQBlock) void pairsum_f32_kernel(const unsigned long long *__restrict__ hist,
                                                              unsigned nhist, const double *__restrict__ q,
                                                              unsigned nq, double bin,
                                                              double *__restrict__ S) {
  __shared__ float2 bins[kChunk];
  const unsigned p = blockIdx.z;
  const unsigned k0 = blockIdx.y * kChunk;
  const unsigned kn = min((unsigned)kChunk, nhist - k0);
  const unsigned long long *h = hist + (size_t)p * nhist + k0;
  const float binf = (float)bin;
  for (unsigned b = threadIdx.x; b < kn; b += kQBlock)
    bins[b] = make_float2((float)h[b], ((float)(k0 + b) + 0.5f) * binf);
  __syncthreads();
  const unsigned iq = blockIdx.x * kQBlock + threadIdx.x;
  if (iq >= nq) return;
  const float qv = (float)q[iq];
  if (qv == 0.0f) return;
  float acc = 0.0f;
  for (unsigned b = 0; b < kn; ++b) {
    const float2 e = bins[b];
    if (e.x != 0.0f) {
      const float x = qv * e.y;
      acc += e.x * sinf(x) / (x + 1e-8f);
    }
  }
  if (acc != 0.0f) atomicAdd(&S[(size_t)p * nq + iq], (double)acc);
}

// Many histograms: blockIdx.x = chunk, blockIdx.y = block of 256 q.
__global__ __launch_bounds__(kQBlock) void pairsum_batch_kernel(
    const unsigned long long *__restrict__ hist, const PairsumChunk *__restrict__ chunks,
    const double *__restrict__ q, unsigned nq, double bin, double *__restrict__ S) {
  __shared__ double2 bins[kChunk];
  const PairsumChunk ch = chunks[blockIdx.x];
  stage_chunk(bins, hist + ch.hist_off, ch.k0, ch.kn);
  __syncthreads();
  const unsigned iq = blockIdx.y * kQBlock + threadIdx.x;
  if (iq >= nq) return;
  const double qv = q[iq];
  if (qv == 0.0) return;
  const double acc = pairsum_chunk(bins, ch.k0, ch.kn, qv, bin);
  if (acc != 0.0) atomicAdd(&S[(size_t)ch.type * nq + iq], acc / (qv * bin));
}

struct PdfArgs {
  const unsigned long long *hist;
  double *pdf;
  unsigned nhist, n_el, ntot;
  int type;
  int cpu_radius;
  double bin, rho0;
  unsigned el_count[16];
  double sl[16];
};

// All PDF values of bin k of one structure: partials at pdf[nhist + p*nhist + k],
// total at pdf[k].  Operation order of CalcFunctions.cpp:919-957.
__device__ __forceinline__ void pdf_point(const unsigned long long *__restrict__ hist,
                                          double *__restrict__ pdf, unsigned k, unsigned nhist,
                                          unsigned n_el, unsigned ntot, const unsigned *el_count,
                                          const double *sl, double rho0, int type, int cpu_radius,
                                          double bin) {
#pragma clang fp contract(off)
  double fav2 = 0;
  for (unsigned e = 0; e < n_el; ++e) fav2 += sl[e] * el_count[e];
  fav2 /= ntot;
  fav2 *= fav2;
  const double nt2 = (double)ntot * (double)ntot;
  double total = 0.0;
  size_t start = 0;
  for (unsigned a = 0; a < n_el; ++a) {
    for (unsigned b = a; b < n_el; ++b, start += nhist) {
      const double wab = sl[a] * sl[b] / fav2;
      const size_t i = start + k;
      const double cnt = (double)hist[i];
      const double r = ((cpu_radius ? (double)i : (double)k) + 0.5) * bin;
      double v;
      if (type == XS_PDF_RDF) {
        v = cnt * (2. / (bin * ntot));
      } else if (type == XS_PDF_PDF) {
        const double mult = 0.5 / (kPi * bin * rho0 * ntot);
        v = cnt * mult / (r * r);
      } else {
        const double mult = 2. / (bin * ntot);
        const double sub = (b > a) ? 8. * kPi * rho0 * (double)el_count[a] * (double)el_count[b] / nt2
                                   : 4. * kPi * rho0 * ((double)el_count[a] * (double)el_count[a]) / nt2;
        v = cnt * mult / r - sub * r;
      }
      pdf[nhist + i] = v;
      total += v * wab;
    }
  }
  pdf[k] = total;
}

__global__ void pdf_kernel(PdfArgs A) {
  const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= A.nhist) return;
  pdf_point(A.hist, A.pdf, k, A.nhist, A.n_el, A.ntot, A.el_count, A.sl, A.rho0, A.type,
            A.cpu_radius, A.bin);
}

// Many structures: 256-bin blocks laid end to end, structure found by binary
// search over the descriptors' first block.
__global__ void pdf_batch_kernel(const unsigned long long *__restrict__ hist, double *__restrict__ pdf,
                                 const PdfDesc *__restrict__ desc, unsigned nstruct, int type,
                                 int cpu_radius, double bin) {
  unsigned lo = 0, hi = nstruct - 1;
  while (lo < hi) {
    const unsigned mid = (lo + hi + 1) / 2;
    if (desc[mid].blk0 <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const PdfDesc &d = desc[lo];
  const unsigned k = (blockIdx.x - d.blk0) * blockDim.x + threadIdx.x;
  if (k >= d.nhist) return;
  pdf_point(hist + d.hist_off, pdf + d.pdf_off, k, d.nhist, d.n_el, d.ntot, d.el_count, d.sl, d.rho0,
            type, cpu_radius, bin);
}

}  // namespace iq

int batch_pairsum_launch(xs_ctx *ctx, const unsigned long long *hist, const PairsumChunk *chunks,
                         size_t nchunks, size_t ntypes, double bin, const double *q, unsigned nq,
                         double *S, bool prof) {
  const size_t ssize = ntypes * nq;
  const unsigned zb = (unsigned)std::min<size_t>((ssize + 255) / 256, 8192);
  hipLaunchKernelGGL(iq::fill_f64, dim3(std::max(zb, 1u)), dim3(256), 0, ctx->stream, S, ssize);
  if (!nchunks) return XS_OK;
  hipEvent_t ev = nullptr;
  if (prof) XS_TRY(prof_begin(ctx, "batch_pairsum", &ev));
  hipLaunchKernelGGL(iq::pairsum_batch_kernel, dim3((unsigned)nchunks, (nq + iq::kQBlock - 1) / iq::kQBlock),
                     dim3(iq::kQBlock), 0, ctx->stream, hist, chunks, q, nq, bin, S);
  XS_HIP_CHECK(hipGetLastError());
  if (prof) XS_TRY(prof_end(ctx, "batch_pairsum", ev));
  return XS_OK;
}

int batch_pdf_launch(xs_ctx *ctx, const unsigned long long *hist, const PdfDesc *desc,
                     unsigned nstruct, unsigned nblocks, int pdf_type, bool cpu_radius, double bin,
                     double *pdf, bool prof) {
  if (!nblocks) return XS_OK;
  hipEvent_t ev = nullptr;
  if (prof) XS_TRY(prof_begin(ctx, "batch_pdf", &ev));
  hipLaunchKernelGGL(iq::pdf_batch_kernel, dim3(nblocks), dim3(256), 0, ctx->stream, hist, pdf, desc,
                     nstruct, pdf_type, cpu_radius ? 1 : 0, bin);
  XS_HIP_CHECK(hipGetLastError());
  if (prof) XS_TRY(prof_end(ctx, "batch_pdf", ev));
  return XS_OK;
}

}  // namespace xs

using namespace xs;

extern "C" int xs_hist_pairsum(xs_ctx *ctx, const uint64_t *hist, unsigned ntypes, unsigned nhist,
                               double bin, const double *q, unsigned nq, double *S,
                               unsigned flags) {
  XS_TRY(activate(ctx));
  if (!hist || !q || !S) return fail(XS_ERR_ARG, "xs_hist_pairsum: null pointer");
  if (!(bin > 0.0) || !std::isfinite(bin)) return fail(XS_ERR_ARG, "xs_hist_pairsum: bin must be > 0");
  if (nhist > (1u << 24)) return fail(XS_ERR_RANGE, "xs_hist_pairsum: nhist > 2^24");
  if (ntypes == 0 || nq == 0 || nhist == 0) return XS_OK;
  const bool device = flags & XS_DEVICE_PTRS;
  const size_t hsize = (size_t)ntypes * nhist, ssize = (size_t)ntypes * nq;
  XS_TRY(ensure_workspace(ctx, Carver::need({device ? 0 : hsize * 8, device ? 0 : ssize * 8,
                                             device ? 0 : (size_t)nq * 8})));
  Carver cv(ctx->ws);
  const unsigned long long *hd = reinterpret_cast<const unsigned long long *>(hist);
  double *Sd = S;
  const double *qd = q;
  if (!device) {
    auto *h2 = cv.take<unsigned long long>(hsize);
    Sd = cv.take<double>(ssize);
    auto *q2 = cv.take<double>(nq);
    XS_HIP_CHECK(hipMemcpyAsync(h2, hist, hsize * 8, hipMemcpyHostToDevice, ctx->stream));
    XS_HIP_CHECK(hipMemcpyAsync(q2, q, (size_t)nq * 8, hipMemcpyHostToDevice, ctx->stream));
    hd = h2;
    qd = q2;
  }
  if ((flags & XS_PAIRSUM_MFMA) && (flags & XS_PAIRSUM_F32))
    return fail(XS_ERR_ARG, "xs_hist_pairsum: XS_PAIRSUM_MFMA and XS_PAIRSUM_F32 exclude each other");
  if (flags & XS_PAIRSUM_MFMA) {
    XS_TRY(pairsum_mfma(ctx, hd, ntypes, nhist, bin, qd, nq, Sd));
  } else {
    const unsigned zb = (unsigned)std::min<size_t>((ssize + 255) / 256, 4096);
    hipLaunchKernelGGL(iq::fill_f64, dim3(zb), dim3(256), 0, ctx->stream, Sd, ssize);
    hipEvent_t ev;
    XS_TRY(prof_begin(ctx, "hist_pairsum", &ev));
    dim3 grid((nq + iq::kQBlock - 1) / iq::kQBlock, (nhist + iq::kChunk - 1) / iq::kChunk, ntypes);
    if (flags & XS_PAIRSUM_F32)
      hipLaunchKernelGGL(iq::pairsum_f32_kernel, grid, dim3(iq::kQBlock), 0, ctx->stream, hd, nhist,
                         qd, nq, bin, Sd);
    else
      hipLaunchKernelGGL(iq::pairsum_kernel, grid, dim3(iq::kQBlock), 0, ctx->stream, hd, nhist, qd,
                         nq, bin, Sd);
    XS_HIP_CHECK(hipGetLastError());
    XS_TRY(prof_end(ctx, "hist_pairsum", ev));
  }
  if (!device) {
    XS_HIP_CHECK(hipMemcpyAsync(S, Sd, ssize * 8, hipMemcpyDeviceToHost, ctx->stream));
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
  return XS_OK;
}

extern "C" int xs_pdf(xs_ctx *ctx, const uint64_t *hist, unsigned n_el, const unsigned *el_count,
                      unsigned nhist, double bin, int pdf_type, double rho0, const double *sl,
                      double *pdf, unsigned flags) {
  XS_TRY(activate(ctx));
  if (!hist || !el_count || !sl || !pdf) return fail(XS_ERR_ARG, "xs_pdf: null pointer");
  if (n_el == 0 || n_el > 16) return fail(XS_ERR_RANGE, "xs_pdf: n_el must be in [1, 16]");
  if (pdf_type < XS_PDF_RDF || pdf_type > XS_PDF_RPDF) return fail(XS_ERR_ARG, "xs_pdf: bad type");
  if (!(bin > 0.0)) return fail(XS_ERR_ARG, "xs_pdf: bin must be > 0");
  if (nhist == 0) return XS_OK;
  iq::PdfArgs A{};
  unsigned ntot = 0;
  for (unsigned e = 0; e < n_el; ++e) {
    A.el_count[e] = el_count[e];
    A.sl[e] = sl[e];
    ntot += el_count[e];
  }
  if (ntot == 0) return fail(XS_ERR_ARG, "xs_pdf: no atoms");
  const unsigned ntypes = ntypes_of(n_el);
  const bool device = flags & XS_DEVICE_PTRS;
  const size_t hsize = (size_t)ntypes * nhist, psize = (size_t)(1 + ntypes) * nhist;
  Carver cv(nullptr);
  if (!device) {
    XS_TRY(ensure_workspace(ctx, Carver::need({hsize * 8, psize * 8})));
    cv = Carver(ctx->ws);
  }
  A.hist = reinterpret_cast<const unsigned long long *>(hist);
  A.pdf = pdf;
  if (!device) {
    auto *h2 = cv.take<unsigned long long>(hsize);
    A.pdf = cv.take<double>(psize);
    XS_HIP_CHECK(hipMemcpyAsync(h2, hist, hsize * 8, hipMemcpyHostToDevice, ctx->stream));
    A.hist = h2;
  }
  A.nhist = nhist;
  A.n_el = n_el;
  A.ntot = ntot;
  A.type = pdf_type;
  A.cpu_radius = (flags & XS_PDF_CPU_RADIUS) ? 1 : 0;
  A.bin = bin;
  A.rho0 = rho0;
  hipEvent_t ev;
  XS_TRY(prof_begin(ctx, "pdf", &ev));
  hipLaunchKernelGGL(iq::pdf_kernel, dim3((nhist + 255) / 256), dim3(256), 0, ctx->stream, A);
  XS_HIP_CHECK(hipGetLastError());
  XS_TRY(prof_end(ctx, "pdf", ev));
  if (!device) {
    XS_HIP_CHECK(hipMemcpyAsync(pdf, A.pdf, psize * 8, hipMemcpyDeviceToHost, ctx->stream));
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
  return XS_OK;
}
