// Pair sums of RearrangementInt (CalcFunctions.cpp:469-545).
//
// Cut-off sums: every ordered pair (i, j) of a block's cell atoms with
// 1e-6 < r_ij^2 < rmax2[i] (rmax2 = the largest squared neighbour distance
// of atom i plus Rcorr^2) contributes sin(q r_ij)/(q r_ij + 1e-8) to its
// element pair (CalcFunctions.cpp:485-499).  The reference tests all N^2
// pairs on one core and evaluates Nq sines per accepted pair.  Here the
// atoms are sorted into a grid of cells no narrower than the largest
// cut-off and each atom tests the atoms of the 27 cells around its own
// (FP64, the reference's expression, no contraction); a count pass sizes one
// segment of distances per element pair and a write pass fills them.  A
// second kernel sums the sines of each segment for all q (one thread per q,
// distances staged in LDS, FP64 atomics across chunks).  The neighbour terms
// (506-529) are a given list of distances and use the second kernel alone.
#include "xs_internal.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#pragma clang fp contract(off)

namespace xs {
namespace rearr {

constexpr int kThreads = 256;
constexpr unsigned kChunk = 1024;  // distances per sum work group

struct CutoffArgs {
  const double *x, *y, *z;           // by atom
  const double *X, *Y, *Z;           // by grid position
  const uint32_t *order, *beg, *end;  // grid order and cell ranges
  const uint32_t *el;
  const double *rmax2;
  Grid G;
  uint32_t n, n_el;
  unsigned long long *count, *slot;
  const unsigned long long *seg_off;
  double *r_out;
};

// Visit the accepted pairs of atom i: every atom j in the 27 grid cells
// around i's cell (cell edge >= sqrt(max rmax2), so no accepted pair is
// missed), f(type, r^2) for each.
template <class F>
__device__ inline void for_pairs(const CutoffArgs &A, uint32_t i, F f) {
  const double xi = A.x[i], yi = A.y[i], zi = A.z[i], ri2 = A.rmax2[i];
  const uint32_t ei = A.el[i] * A.n_el;
  const int cx = grid_coord(xi, A.G.lo[0], A.G.inv, A.G.g[0]);
  const int cy = grid_coord(yi, A.G.lo[1], A.G.inv, A.G.g[1]);
  const int cz = grid_coord(zi, A.G.lo[2], A.G.inv, A.G.g[2]);
  for (int gz = max(cz - 1, 0); gz <= min(cz + 1, A.G.g[2] - 1); ++gz)
    for (int gy = max(cy - 1, 0); gy <= min(cy + 1, A.G.g[1] - 1); ++gy)
      for (int gx = max(cx - 1, 0); gx <= min(cx + 1, A.G.g[0] - 1); ++gx) {
        const uint32_t k = grid_cell(A.G, gx, gy, gz);
        const uint32_t p1 = A.end[k];
        for (uint32_t p = A.beg[k]; p < p1; ++p) {
          // (rAtom[i] - rAtom[j]).sqr(), CalcFunctions.cpp:491
          const double dx = xi - A.X[p], dy = yi - A.Y[p], dz = zi - A.Z[p];
          const double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 < ri2 && r2 > 0.000001) f(ei + A.el[A.order[p]], r2);
        }
      }
}

// Count (WRITE = false) or write (WRITE = true) the accepted pairs, one
// thread per atom.  Counts are gathered per work group in LDS.  The write
// pass counts again, reserves each element pair's range in its segment with
// one global atomic per work group, and writes through LDS slots.
template <bool WRITE>
__global__ __launch_bounds__(kThreads) void cutoff_pairs(CutoffArgs A) {
  extern __shared__ unsigned long long lds[];
  const uint32_t ntypes = A.n_el * A.n_el;
  unsigned *lcount = reinterpret_cast<unsigned *>(lds + (WRITE ? ntypes : 0));
  unsigned long long *base = lds;  // [ntypes], write pass only
  for (uint32_t t = threadIdx.x; t < ntypes; t += kThreads) lcount[t] = 0;
  __syncthreads();
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  const bool vi = i < A.n;
  if (vi) for_pairs(A, i, [&](uint32_t t, double) { atomicAdd(&lcount[t], 1u); });
  __syncthreads();
  if (!WRITE) {
    for (uint32_t t = threadIdx.x; t < ntypes; t += kThreads)
      if (lcount[t]) atomicAdd(&A.count[t], (unsigned long long)lcount[t]);
    return;
  }
  for (uint32_t t = threadIdx.x; t < ntypes; t += kThreads) {
    base[t] = lcount[t] ? A.seg_off[t] + atomicAdd(&A.slot[t], (unsigned long long)lcount[t]) : 0;
    lcount[t] = 0;
  }
  __syncthreads();
  if (vi)
    for_pairs(A, i, [&](uint32_t t, double r2) {
      const unsigned s = atomicAdd(&lcount[t], 1u);
      A.r_out[base[t] + s] = sqrt(r2);
    });
}

struct Chunk {
  unsigned long long start;  // first distance
  uint32_t len, type;
};

// S[type * nq + iq] += sum over the chunk of sin(q r)/(q r + 1e-8).
__global__ __launch_bounds__(kThreads) void sinc_sums(const double *r, const Chunk *chunks,
                                                      const double *q, unsigned nq, double *S) {
  __shared__ double sr[kChunk];
  const Chunk c = chunks[blockIdx.x];
  for (uint32_t k = threadIdx.x; k < c.len; k += kThreads) sr[k] = r[c.start + k];
  __syncthreads();
  const unsigned iq = blockIdx.y * kThreads + threadIdx.x;
  if (iq >= nq) return;
  const double qq = q[iq];
  double acc = 0.0;
  for (uint32_t k = 0; k < c.len; ++k) {
    const double qr = sr[k] * qq;  // rij * q[iq], CalcFunctions.cpp:495
    acc += sin(qr) / (qr + 0.00000001);
  }
  atomicAdd(&S[(size_t)c.type * nq + iq], acc);
}

__global__ void zero_f64(double *p, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

inline unsigned nblk(size_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

// Chunks of at most kChunk distances inside each type segment.
std::vector<Chunk> make_chunks(const std::vector<unsigned long long> &seg_off,
                               const std::vector<unsigned long long> &seg_len) {
  std::vector<Chunk> ch;
  for (size_t t = 0; t < seg_len.size(); ++t)
    for (unsigned long long s = 0; s < seg_len[t]; s += kChunk)
      ch.push_back(Chunk{seg_off[t] + s, (uint32_t)std::min<unsigned long long>(kChunk, seg_len[t] - s),
                         (uint32_t)t});
  return ch;
}

// Sum the sines of distances r_d (device, grouped in segments) into S (host).
int segment_sums(xs_ctx *ctx, const double *r_d, const std::vector<unsigned long long> &seg_off,
                 const std::vector<unsigned long long> &seg_len, const double *q, unsigned nq,
                 double *S_host, const char *prof_name) {
  const size_t ntypes = seg_len.size();
  const std::vector<Chunk> ch = make_chunks(seg_off, seg_len);
  DevMem b;
  const size_t o_q = 0, o_S = align_up(nq * sizeof(double), 256),
               o_c = o_S + align_up(ntypes * nq * sizeof(double), 256);
  XS_TRY(b.alloc(o_c + ch.size() * sizeof(Chunk), "pair sums"));
  XS_HIP_CHECK(hipMemcpyAsync(b.at<double>(o_q), q, nq * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
  if (!ch.empty())
    XS_HIP_CHECK(hipMemcpyAsync(b.at<Chunk>(o_c), ch.data(), ch.size() * sizeof(Chunk),
                                hipMemcpyHostToDevice, ctx->stream));
  double *S = b.at<double>(o_S);
  zero_f64<<<nblk(ntypes * nq), kThreads, 0, ctx->stream>>>(S, ntypes * nq);
  XS_HIP_CHECK(hipGetLastError());
  hipEvent_t ev;
  XS_TRY(prof_begin(ctx, prof_name, &ev));
  if (!ch.empty()) {
    sinc_sums<<<dim3((unsigned)ch.size(), nblk(nq)), kThreads, 0, ctx->stream>>>(
        r_d, b.at<Chunk>(o_c), b.at<double>(o_q), nq, S);
    XS_HIP_CHECK(hipGetLastError());
  }
  XS_TRY(prof_end(ctx, prof_name, ev));
  XS_HIP_CHECK(hipMemcpyAsync(S_host, S, ntypes * nq * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return XS_OK;
}

}  // namespace rearr
}  // namespace xs

using namespace xs;
using namespace xs::rearr;

extern "C" int xs_cutoff_pairsum(xs_ctx *ctx, const double *xyz, size_t n_atoms, const unsigned *el,
                                 unsigned n_el, const double *rmax2, const double *q, unsigned nq,
                                 double *S, uint64_t *npairs, unsigned flags) {
  if (!ctx) return fail(XS_ERR_ARG, "xs_cutoff_pairsum: null xs_ctx");
  if (flags & XS_DEVICE_PTRS) return fail(XS_ERR_ARG, "xs_cutoff_pairsum: host arrays only");
  if (!n_el || n_el > 64) return fail(XS_ERR_ARG, "xs_cutoff_pairsum: n_el must be 1..64");
  if (!S || !npairs || (nq && !q) || (n_atoms && (!xyz || !el || !rmax2)))
    return fail(XS_ERR_ARG, "xs_cutoff_pairsum: null array");
  if (n_atoms >= (size_t)1 << 31) return fail(XS_ERR_RANGE, "xs_cutoff_pairsum: too many atoms");
  for (size_t i = 0; i < n_atoms; ++i)
    if (el[i] >= n_el) return fail(XS_ERR_ARG, "xs_cutoff_pairsum: element index >= n_el");
  XS_TRY(activate(ctx));
  const size_t ntypes = (size_t)n_el * n_el, n = n_atoms;
  std::fill(S, S + ntypes * nq, 0.0);
  std::fill(npairs, npairs + ntypes, 0ull);
  if (n < 2) return XS_OK;
  const double maxr2 = *std::max_element(rmax2, rmax2 + n);
  if (!(maxr2 > 0.000001)) return XS_OK;  // no pair can pass 1e-6 < r^2 < rmax2
  if (!std::isfinite(maxr2)) return fail(XS_ERR_ARG, "xs_cutoff_pairsum: rmax2 not finite");
  DevMem in;
  const size_t o_x = 0, o_r = align_up(3 * n * sizeof(double), 256),
               o_e = o_r + align_up(n * sizeof(double), 256),
               o_c = o_e + align_up(n * sizeof(uint32_t), 256);
  XS_TRY(in.alloc(o_c + 3 * ntypes * sizeof(unsigned long long), "cut-off inputs"));
  XS_HIP_CHECK(hipMemcpyAsync(in.at<double>(o_x), xyz, 3 * n * sizeof(double),
                              hipMemcpyHostToDevice, ctx->stream));
  XS_HIP_CHECK(hipMemcpyAsync(in.at<double>(o_r), rmax2, n * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
  XS_HIP_CHECK(hipMemcpyAsync(in.at<uint32_t>(o_e), el, n * sizeof(uint32_t),
                              hipMemcpyHostToDevice, ctx->stream));
  unsigned long long *count = in.at<unsigned long long>(o_c), *slot = count + ntypes,
                     *seg = slot + ntypes;
  XS_HIP_CHECK(hipMemsetAsync(count, 0, 2 * ntypes * sizeof(unsigned long long), ctx->stream));
  const double *x = in.at<double>(o_x), *y = x + n, *z = y + n;
  hipEvent_t ev;
  XS_TRY(prof_begin(ctx, "cutoff_pairs", &ev));
  SortedPoints sp;
  XS_TRY(grid_sort(ctx, x, y, z, n, std::sqrt(maxr2), sp));
  CutoffArgs A{x, y, z, sp.X, sp.Y, sp.Z, sp.order, sp.beg, sp.end, in.at<uint32_t>(o_e),
               in.at<double>(o_r), sp.G, (uint32_t)n, n_el, count, slot, seg, nullptr};
  const size_t lds_count = align_up(ntypes * sizeof(unsigned), 8);
  const size_t lds_write = ntypes * sizeof(unsigned long long) + lds_count;
  cutoff_pairs<false><<<nblk(n), kThreads, lds_count, ctx->stream>>>(A);
  XS_HIP_CHECK(hipGetLastError());
  std::vector<unsigned long long> len(ntypes), off(ntypes);
  XS_TRY(download(ctx, len.data(), count, ntypes * sizeof(unsigned long long)));
  unsigned long long total = 0;
  for (size_t t = 0; t < ntypes; ++t) {
    off[t] = total;
    total += len[t];
  }
  if (total > (1ull << 31))
    return fail(XS_ERR_RANGE, "xs_cutoff_pairsum: more than 2^31 pairs within the cut-offs");
  DevMem rd;
  XS_TRY(rd.alloc(total * sizeof(double), "pair distances"));
  if (total) {
    XS_HIP_CHECK(hipMemcpyAsync(seg, off.data(), ntypes * sizeof(unsigned long long),
                                hipMemcpyHostToDevice, ctx->stream));
    A.r_out = rd.at<double>();
    cutoff_pairs<true><<<nblk(n), kThreads, lds_write, ctx->stream>>>(A);
    XS_HIP_CHECK(hipGetLastError());
  }
  XS_TRY(prof_end(ctx, "cutoff_pairs", ev));
  for (size_t t = 0; t < ntypes; ++t) npairs[t] = len[t];
  if (!nq || !total) {
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return XS_OK;
  }
  return segment_sums(ctx, rd.at<double>(), off, len, q, nq, S, "pairlist_sums");
}

extern "C" int xs_pairlist_pairsum(xs_ctx *ctx, const double *r, const unsigned *type,
                                   size_t npairs, unsigned ntypes, const double *q, unsigned nq,
                                   double *S, unsigned flags) {
  if (!ctx) return fail(XS_ERR_ARG, "xs_pairlist_pairsum: null xs_ctx");
  if (flags & XS_DEVICE_PTRS) return fail(XS_ERR_ARG, "xs_pairlist_pairsum: host arrays only");
  if (!ntypes || !S || (nq && !q) || (npairs && (!r || !type)))
    return fail(XS_ERR_ARG, "xs_pairlist_pairsum: null array or no types");
  for (size_t p = 0; p < npairs; ++p)
    if (type[p] >= ntypes) return fail(XS_ERR_ARG, "xs_pairlist_pairsum: type >= ntypes");
  XS_TRY(activate(ctx));
  std::fill(S, S + (size_t)ntypes * nq, 0.0);
  if (!npairs || !nq) return XS_OK;
  // group by type (stable counting sort on the host)
  std::vector<unsigned long long> len(ntypes, 0), off(ntypes);
  for (size_t p = 0; p < npairs; ++p) ++len[type[p]];
  unsigned long long o = 0;
  for (unsigned t = 0; t < ntypes; ++t) {
    off[t] = o;
    o += len[t];
  }
  std::vector<double> sorted(npairs);
  std::vector<unsigned long long> pos(off);
  for (size_t p = 0; p < npairs; ++p) sorted[pos[type[p]]++] = r[p];
  DevMem rd;
  XS_TRY(rd.alloc(npairs * sizeof(double), "pair distances"));
  XS_HIP_CHECK(hipMemcpyAsync(rd.p, sorted.data(), npairs * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
  return segment_sums(ctx, rd.at<double>(), off, len, q, nq, S, "pairlist_sums");
}
