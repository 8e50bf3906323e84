// Ensemble runs: histograms, pair sums and PDFs of every structure of a plan.
//
// The reference computes one structure per program run (main.cpp:151-233);
// an ensemble of particles there is a sequence of runs, each paying its own
// launches and host synchronisations.  A plan (xs_batch_plan) holds the
// sorted coordinates, tiles and work items of all structures on the device,
// so a run is a fixed sequence of launches with no host work:
//   zero fills -> histogram kernels (8192- then 32768-bin windows)
//   -> zero fill + pair-sum kernel over all (pair type, 2048-bin chunk)
//   -> PDF kernel over all structures' bins.
// With XS_BATCH_GRAPH the sequence is captured once into a hipGraph on the
// context's stream and replayed while the output pointers, nq, the PDF type
// and the flags stay the same.  The graph holds kernel nodes only: buffers
// are cleared by fill kernels, never by memset nodes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "xs_internal.h"

namespace xs {

void release_batch_run(BatchRunState &r) {
  if (r.exec) (void)hipGraphExecDestroy(r.exec);
  if (r.graph) (void)hipGraphDestroy(r.graph);
  if (r.buf) (void)hipFree(r.buf);
  r = BatchRunState();
}

namespace {

struct RunLayout {
  size_t chunks, pdfdesc, q, hist, S, pdf, mm_cols, mm_tiles, mm_a, mm_b, total;
};

// Byte offsets inside the run buffer; host-mode staging only when needed;
// the matrix-core tables and slices when mm (a plan with mm.nq != 0) is given.
RunLayout layout(const xs_batch *b, size_t nchunks, unsigned nq, bool host, bool want_S,
                 bool want_pdf, const MfmaBatchPlan *mm) {
  RunLayout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    off = align_up(off, 256);
    const size_t at = off;
    off += bytes;
    return at;
  };
  L.chunks = take(std::max<size_t>(nchunks, 1) * sizeof(PairsumChunk));
  L.pdfdesc = take((size_t)b->nstruct * sizeof(PdfDesc));
  L.q = take(host && want_S ? (size_t)nq * 8 : 0);
  L.hist = take(host ? b->hist_len * 8 : 0);
  L.S = take(host && want_S ? b->ntypes_total * nq * 8 : 0);
  L.pdf = take(host && want_pdf ? b->pdf_len * 8 : 0);
  L.mm_cols = take(mm ? mm->cols.size() * sizeof(MfmaCol) : 0);
  L.mm_tiles = take(mm ? mm->tiles.size() * sizeof(MfmaTile) : 0);
  L.mm_a = take(mm ? mm->a_bytes : 0);
  L.mm_b = take(mm ? mm->b_bytes : 0);
  L.total = align_up(off, 256);
  return L;
}

std::vector<PairsumChunk> make_chunks(const xs_batch *b) {
  std::vector<PairsumChunk> c;
  for (size_t t = 0; t < b->ntypes_total; ++t)
    for (uint32_t k0 = 0; k0 < b->type_nhist[t]; k0 += kPairsumChunk)
      c.push_back(PairsumChunk{b->type_hist_off[t] + k0, (uint32_t)t, k0,
                               std::min<uint32_t>(kPairsumChunk, b->type_nhist[t] - k0), 0u});
  return c;
}

std::vector<PdfDesc> make_pdf_desc(const xs_batch *b, const double *rho0, const double *sl,
                                   unsigned *nblocks) {
  std::vector<PdfDesc> d(b->nstruct);
  unsigned blk = 0;
  for (unsigned s = 0; s < b->nstruct; ++s) {
    PdfDesc &x = d[s];
    std::memset(&x, 0, sizeof(x));
    x.hist_off = b->hist_off[s];
    x.pdf_off = b->pdf_off[s];
    x.nhist = b->nhist[s];
    x.n_el = b->n_el[s];
    x.ntot = b->ntot[s];
    x.blk0 = blk;
    x.rho0 = rho0 ? rho0[s] : 0.0;
    for (unsigned e = 0; e < x.n_el; ++e) {
      x.el_count[e] = b->el_count[b->el_first[s] + e];
      x.sl[e] = sl ? sl[b->el_first[s] + e] : 0.0;
    }
    blk += (x.nhist + 255) / 256;
  }
  *nblocks = blk;
  return d;
}

struct RunPtrs {
  unsigned long long *hist;
  const double *q;
  double *S;
  double *pdf;
};

// An upper bound of every count of each pair type: its number of pairs.
std::vector<unsigned long long> pair_counts(const xs_batch *b) {
  std::vector<unsigned long long> c(b->ntypes_total);
  for (unsigned s = 0; s < b->nstruct; ++s) {
    const unsigned *ec = &b->el_count[b->el_first[s]];
    size_t t = b->type_first[s];
    for (unsigned i = 0; i < b->n_el[s]; ++i)
      for (unsigned j = i; j < b->n_el[s]; ++j, ++t)
        c[t] = i == j ? (unsigned long long)ec[i] * (ec[i] - (ec[i] > 0)) / 2
                      : (unsigned long long)ec[i] * ec[j];
  }
  return c;
}

int launch_all(xs_batch *b, const RunPtrs &P, const RunLayout &L, size_t nchunks, unsigned nq,
               int pdf_type, unsigned pdf_blocks, bool cpu_radius, bool mfma, bool prof) {
  char *buf = static_cast<char *>(b->run.buf);
  XS_TRY(batch_hist_launch(b, P.hist, prof));
  if (P.S && mfma)
    XS_TRY(pairsum_mfma_batch(b->ctx, b->run.mm, P.hist, reinterpret_cast<const MfmaCol *>(buf + L.mm_cols),
                              reinterpret_cast<const MfmaTile *>(buf + L.mm_tiles), buf + L.mm_a,
                              buf + L.mm_b, b->bin, P.q, nq, P.S, prof));
  else if (P.S)
    XS_TRY(batch_pairsum_launch(b->ctx, P.hist, reinterpret_cast<const PairsumChunk *>(buf + L.chunks),
                                nchunks, b->ntypes_total, b->bin, P.q, nq, P.S, prof));
  if (P.pdf)
    XS_TRY(batch_pdf_launch(b->ctx, P.hist, reinterpret_cast<const PdfDesc *>(buf + L.pdfdesc),
                            b->nstruct, pdf_blocks, pdf_type, cpu_radius, b->bin, P.pdf, prof));
  return XS_OK;
}

}  // namespace
}  // namespace xs

using namespace xs;

extern "C" int xs_batch_run(xs_batch *plan, uint64_t *hist, const double *q, unsigned nq, double *S,
                            int pdf_type, const double *rho0, const double *sl, double *pdf,
                            unsigned flags) {
  if (!plan) return fail(XS_ERR_ARG, "xs_batch_run: null plan");
  xs_ctx *ctx = plan->ctx;
  XS_TRY(activate(ctx));
  const bool device = flags & XS_DEVICE_PTRS;
  const bool graph = flags & XS_BATCH_GRAPH;
  const bool want_S = S != nullptr && nq > 0;
  const bool want_pdf = pdf != nullptr;
  if (!hist) return fail(XS_ERR_ARG, "xs_batch_run: null hist");
  if (want_S && !q) return fail(XS_ERR_ARG, "xs_batch_run: null q");
  if (want_pdf) {
    if (pdf_type < XS_PDF_RDF || pdf_type > XS_PDF_RPDF)
      return fail(XS_ERR_ARG, "xs_batch_run: bad PDF type");
    if (!sl || (pdf_type != XS_PDF_RDF && !rho0))
      return fail(XS_ERR_ARG, "xs_batch_run: PDF needs sl (and rho0 unless RDF)");
  }
  if (want_S && nq > 65535u * 256u) return fail(XS_ERR_RANGE, "xs_batch_run: nq too large");
  if (graph && ctx->stream == nullptr)
    return fail(XS_ERR_ARG, "xs_batch_run: graph capture needs a non-default stream");

  BatchRunState &R = plan->run;
  const bool mfma = want_S && (flags & XS_PAIRSUM_MFMA);
  if (mfma && R.mm.nq != nq) {
    // column order, groups and slice counts for this nq (the count bounds
    // are the pair counts, so no read-back: the run stays capturable)
    if (!mfma_batch_plan(plan->type_hist_off, plan->type_nhist, pair_counts(plan), nq, &R.mm))
      return fail(XS_ERR_RANGE, "xs_batch_run: a pair type has 2^35 pairs or more, beyond the "
                                "matrix-core path; run without XS_PAIRSUM_MFMA");
    R.mm_uploaded = false;
  }
  const std::vector<PairsumChunk> chunks = make_chunks(plan);
  unsigned pdf_blocks = 0;
  const std::vector<PdfDesc> desc = make_pdf_desc(plan, want_pdf ? rho0 : nullptr, want_pdf ? sl : nullptr,
                                                  &pdf_blocks);
  const RunLayout L = layout(plan, chunks.size(), nq, !device, want_S, want_pdf, mfma ? &R.mm : nullptr);

  // (re)allocate the run buffer and upload the descriptors when they change
  bool fresh = false;
  if (L.total > R.buf_cap) {
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (R.exec) (void)hipGraphExecDestroy(R.exec);
    if (R.graph) (void)hipGraphDestroy(R.graph);
    R.exec = nullptr;
    R.graph = nullptr;
    if (R.buf) XS_HIP_CHECK(hipFree(R.buf));
    R.buf = nullptr;
    R.buf_cap = 0;
    if (hipMalloc(&R.buf, L.total) != hipSuccess) {
      R.buf = nullptr;
      (void)hipGetLastError();
      return fail(XS_ERR_NOMEM, "xs_batch_run: device allocation failed");
    }
    R.buf_cap = L.total;
    fresh = true;
  }
  char *buf = static_cast<char *>(R.buf);
  std::vector<double> key_rho0, key_sl;
  if (want_pdf) {
    key_rho0.assign(rho0 ? rho0 : nullptr, rho0 ? rho0 + plan->nstruct : nullptr);
    key_sl.assign(sl, sl + plan->el_count.size());
  }
  const bool desc_changed = fresh || key_rho0 != R.key_rho0 || key_sl != R.key_sl;
  if (fresh || desc_changed) {
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (!chunks.empty())
      XS_HIP_CHECK(hipMemcpy(buf + L.chunks, chunks.data(), chunks.size() * sizeof(PairsumChunk),
                             hipMemcpyHostToDevice));
    XS_HIP_CHECK(hipMemcpy(buf + L.pdfdesc, desc.data(), desc.size() * sizeof(PdfDesc),
                           hipMemcpyHostToDevice));
    R.key_rho0 = key_rho0;
    R.key_sl = key_sl;
  }
  if (mfma && (fresh || !R.mm_uploaded || R.mm_at != L.mm_cols)) {
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    XS_HIP_CHECK(hipMemcpy(buf + L.mm_cols, R.mm.cols.data(), R.mm.cols.size() * sizeof(MfmaCol),
                           hipMemcpyHostToDevice));
    XS_HIP_CHECK(hipMemcpy(buf + L.mm_tiles, R.mm.tiles.data(), R.mm.tiles.size() * sizeof(MfmaTile),
                           hipMemcpyHostToDevice));
    R.mm_uploaded = true;
    R.mm_at = L.mm_cols;
  }

  RunPtrs P{};
  if (device) {
    P.hist = reinterpret_cast<unsigned long long *>(hist);
    P.q = want_S ? q : nullptr;
    P.S = want_S ? S : nullptr;
    P.pdf = want_pdf ? pdf : nullptr;
  } else {
    P.hist = reinterpret_cast<unsigned long long *>(buf + L.hist);
    P.q = want_S ? reinterpret_cast<const double *>(buf + L.q) : nullptr;
    P.S = want_S ? reinterpret_cast<double *>(buf + L.S) : nullptr;
    P.pdf = want_pdf ? reinterpret_cast<double *>(buf + L.pdf) : nullptr;
    if (want_S)
      XS_HIP_CHECK(hipMemcpyAsync(buf + L.q, q, (size_t)nq * 8, hipMemcpyHostToDevice, ctx->stream));
  }
  const bool cpu_radius = flags & XS_PDF_CPU_RADIUS;

  if (!graph) {
    XS_TRY(launch_all(plan, P, L, chunks.size(), nq, pdf_type, pdf_blocks, cpu_radius, mfma, true));
  } else {
    const unsigned kflags = flags & (XS_DEVICE_PTRS | XS_PDF_CPU_RADIUS | (mfma ? XS_PAIRSUM_MFMA : 0u));
    const bool same = R.exec && R.key_hist == P.hist && R.key_S == P.S && R.key_pdf == P.pdf &&
                      R.key_q == P.q && R.key_nq == (want_S ? nq : 0u) &&
                      R.key_pdf_type == (want_pdf ? pdf_type : -1) && R.key_flags == kflags;
    if (!same) {
      XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
      if (R.exec) (void)hipGraphExecDestroy(R.exec);
      if (R.graph) (void)hipGraphDestroy(R.graph);
      R.exec = nullptr;
      R.graph = nullptr;
      XS_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      const int rc =
          launch_all(plan, P, L, chunks.size(), nq, pdf_type, pdf_blocks, cpu_radius, mfma, false);
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
      if (rc != XS_OK || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (rc != XS_OK) return rc;
        return fail(XS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
      }
      R.graph = g;
      const hipError_t ei = hipGraphInstantiate(&R.exec, g, nullptr, nullptr, 0);
      if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        R.graph = nullptr;
        R.exec = nullptr;
        return fail(XS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
      }
      R.key_hist = P.hist;
      R.key_S = P.S;
      R.key_pdf = P.pdf;
      R.key_q = P.q;
      R.key_nq = want_S ? nq : 0u;
      R.key_pdf_type = want_pdf ? pdf_type : -1;
      R.key_flags = kflags;
    }
    hipEvent_t ev;
    XS_TRY(prof_begin(ctx, "batch_graph", &ev));
    XS_HIP_CHECK(hipGraphLaunch(R.exec, ctx->stream));
    XS_TRY(prof_end(ctx, "batch_graph", ev));
  }

  if (!device) {
    XS_HIP_CHECK(hipMemcpyAsync(hist, P.hist, plan->hist_len * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (want_S)
      XS_HIP_CHECK(hipMemcpyAsync(S, P.S, plan->ntypes_total * nq * 8, hipMemcpyDeviceToHost,
                                  ctx->stream));
    if (want_pdf)
      XS_HIP_CHECK(hipMemcpyAsync(pdf, P.pdf, plan->pdf_len * 8, hipMemcpyDeviceToHost, ctx->stream));
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
  return XS_OK;
}

extern "C" int xs_batch_overflow(xs_batch *plan, uint64_t *overflow) {
  if (!plan || !overflow) return fail(XS_ERR_ARG, "xs_batch_overflow: null pointer");
  XS_TRY(activate(plan->ctx));
  XS_HIP_CHECK(hipMemcpyAsync(overflow, plan->ovf, (size_t)plan->nstruct * 8, hipMemcpyDeviceToHost,
                              plan->ctx->stream));
  XS_HIP_CHECK(hipStreamSynchronize(plan->ctx->stream));
  return XS_OK;
}
