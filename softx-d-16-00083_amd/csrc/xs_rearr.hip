// Pair sums of RearrangementInt (CalcFunctions.cpp:469-545).
//
// Cut-off sums: every ordered pair (i, j) of a block's cell atoms with
// 1e-6 < r_ij^2 < rmax2[i] (rmax2 = the largest squared neighbour distance
// of atom i plus Rcorr^2) contributes sin(q r_ij)/(q r_ij + 1e-8) to its
// element pair (CalcFunctions.cpp:485-499).  The reference tests all N^2
// pairs on one core and evaluates Nq sines per accepted pair.  Here a tiled
// kernel tests the pairs (FP64, the reference's expression, no
// contraction), a count pass sizes one segment of distances per element
// pair and a write pass fills them; a second kernel sums the sines of each
// segment for all q (one thread per q, distances staged in LDS, FP64
// atomics across chunks).  The neighbour terms (506-529) are a given list
// of distances and use the second kernel alone.
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

// Count (WRITE = false) or write (WRITE = true) the accepted pairs.
// count[t] (t = el_i * n_el + el_j): pairs per element pair; in the write
// pass slot[t] hands out positions inside segment t (seg_off[t] ...).
template <bool WRITE>
__global__ __launch_bounds__(kThreads) void cutoff_pairs(
    const double *x, const double *y, const double *z, const uint32_t *el, const double *rmax2,
    uint32_t n, uint32_t n_el, unsigned long long *count, unsigned long long *slot,
    const unsigned long long *seg_off, double *r_out) {
  extern __shared__ unsigned char smem[];
  double *sx = reinterpret_cast<double *>(smem), *sy = sx + kThreads, *sz = sy + kThreads;
  uint32_t *se = reinterpret_cast<uint32_t *>(sz + kThreads);
  unsigned *lcount = se + kThreads;  // [n_el * n_el], count pass only
  const uint32_t ntypes = n_el * n_el;
  if (!WRITE)
    for (uint32_t t = threadIdx.x; t < ntypes; t += kThreads) lcount[t] = 0;
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  const bool vi = i < n;
  const double xi = vi ? x[i] : 0.0, yi = vi ? y[i] : 0.0, zi = vi ? z[i] : 0.0;
  const double ri2 = vi ? rmax2[i] : -1.0;
  const uint32_t ei = vi ? el[i] * n_el : 0;
  for (uint32_t j0 = 0; j0 < n; j0 += kThreads) {
    __syncthreads();
    const uint32_t jl = j0 + threadIdx.x;
    if (jl < n) {
      sx[threadIdx.x] = x[jl];
      sy[threadIdx.x] = y[jl];
      sz[threadIdx.x] = z[jl];
      se[threadIdx.x] = el[jl];
    }
    __syncthreads();
    const uint32_t m = min((uint32_t)kThreads, n - j0);
    if (!vi) continue;
    for (uint32_t jj = 0; jj < m; ++jj) {
      // (rAtom[i] - rAtom[j]).sqr(), CalcFunctions.cpp:491
      const double dx = xi - sx[jj], dy = yi - sy[jj], dz = zi - sz[jj];
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 < ri2 && r2 > 0.000001) {
        const uint32_t t = ei + se[jj];
        if (WRITE) {
          const unsigned long long s = atomicAdd(&slot[t], 1ull);
          r_out[seg_off[t] + s] = sqrt(r2);
        } else {
          atomicAdd(&lcount[t], 1u);
        }
      }
    }
  }
  if (!WRITE) {
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < ntypes; t += kThreads)
      if (lcount[t]) atomicAdd(&count[t], (unsigned long long)lcount[t]);
  }
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

// Device buffer released on scope exit.
struct Buf {
  void *p = nullptr;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    if (hipMalloc(&p, std::max<size_t>(bytes, 64)) != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      return fail(XS_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    return XS_OK;
  }
  template <class T>
  T *as(size_t byte_off = 0) const {
    return reinterpret_cast<T *>(static_cast<char *>(p) + byte_off);
  }
};

// Sum the sines of distances r_d (device, grouped in segments) into S (host).
int segment_sums(xs_ctx *ctx, const double *r_d, const std::vector<unsigned long long> &seg_off,
                 const std::vector<unsigned long long> &seg_len, const double *q, unsigned nq,
                 double *S_host, const char *prof_name) {
  const size_t ntypes = seg_len.size();
  const std::vector<Chunk> ch = make_chunks(seg_off, seg_len);
  Buf b;
  const size_t o_q = 0, o_S = align_up(nq * sizeof(double), 256),
               o_c = o_S + align_up(ntypes * nq * sizeof(double), 256);
  XS_TRY(b.alloc(o_c + ch.size() * sizeof(Chunk)));
  XS_HIP_CHECK(hipMemcpyAsync(b.as<double>(o_q), q, nq * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
  if (!ch.empty())
    XS_HIP_CHECK(hipMemcpyAsync(b.as<Chunk>(o_c), ch.data(), ch.size() * sizeof(Chunk),
                                hipMemcpyHostToDevice, ctx->stream));
  double *S = b.as<double>(o_S);
  zero_f64<<<nblk(ntypes * nq), kThreads, 0, ctx->stream>>>(S, ntypes * nq);
  XS_HIP_CHECK(hipGetLastError());
  hipEvent_t ev;
  XS_TRY(prof_begin(ctx, prof_name, &ev));
  if (!ch.empty()) {
    sinc_sums<<<dim3((unsigned)ch.size(), nblk(nq)), kThreads, 0, ctx->stream>>>(
        r_d, b.as<Chunk>(o_c), b.as<double>(o_q), nq, S);
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
  Buf in;
  const size_t o_x = 0, o_r = align_up(3 * n * sizeof(double), 256),
               o_e = o_r + align_up(n * sizeof(double), 256),
               o_c = o_e + align_up(n * sizeof(uint32_t), 256);
  XS_TRY(in.alloc(o_c + 3 * ntypes * sizeof(unsigned long long)));
  XS_HIP_CHECK(hipMemcpyAsync(in.as<double>(o_x), xyz, 3 * n * sizeof(double),
                              hipMemcpyHostToDevice, ctx->stream));
  XS_HIP_CHECK(hipMemcpyAsync(in.as<double>(o_r), rmax2, n * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
  XS_HIP_CHECK(hipMemcpyAsync(in.as<uint32_t>(o_e), el, n * sizeof(uint32_t),
                              hipMemcpyHostToDevice, ctx->stream));
  unsigned long long *count = in.as<unsigned long long>(o_c), *slot = count + ntypes,
                     *seg = slot + ntypes;
  XS_HIP_CHECK(hipMemsetAsync(count, 0, 2 * ntypes * sizeof(unsigned long long), ctx->stream));
  const double *x = in.as<double>(o_x), *y = x + n, *z = y + n;
  const size_t lds = 3 * kThreads * sizeof(double) + kThreads * sizeof(uint32_t) +
                     ntypes * sizeof(unsigned);
  hipEvent_t ev;
  XS_TRY(prof_begin(ctx, "cutoff_pairs", &ev));
  cutoff_pairs<false><<<nblk(n), kThreads, lds, ctx->stream>>>(
      x, y, z, in.as<uint32_t>(o_e), in.as<double>(o_r), (uint32_t)n, n_el, count, nullptr,
      nullptr, nullptr);
  XS_HIP_CHECK(hipGetLastError());
  std::vector<unsigned long long> len(ntypes), off(ntypes);
  XS_HIP_CHECK(hipMemcpyAsync(len.data(), count, ntypes * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, ctx->stream));
  XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  unsigned long long total = 0;
  for (size_t t = 0; t < ntypes; ++t) {
    off[t] = total;
    total += len[t];
  }
  if (total > (1ull << 31))
    return fail(XS_ERR_RANGE, "xs_cutoff_pairsum: more than 2^31 pairs within the cut-offs");
  Buf rd;
  XS_TRY(rd.alloc(total * sizeof(double)));
  if (total) {
    XS_HIP_CHECK(hipMemcpyAsync(seg, off.data(), ntypes * sizeof(unsigned long long),
                                hipMemcpyHostToDevice, ctx->stream));
    cutoff_pairs<true><<<nblk(n), kThreads, lds, ctx->stream>>>(
        x, y, z, in.as<uint32_t>(o_e), in.as<double>(o_r), (uint32_t)n, n_el, nullptr, slot, seg,
        rd.as<double>());
    XS_HIP_CHECK(hipGetLastError());
  }
  XS_TRY(prof_end(ctx, "cutoff_pairs", ev));
  for (size_t t = 0; t < ntypes; ++t) npairs[t] = len[t];
  if (!nq || !total) {
    XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return XS_OK;
  }
  return segment_sums(ctx, rd.as<double>(), off, len, q, nq, S, "pairlist_sums");
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
  Buf rd;
  XS_TRY(rd.alloc(npairs * sizeof(double)));
  XS_HIP_CHECK(hipMemcpyAsync(rd.p, sorted.data(), npairs * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
  return segment_sums(ctx, rd.as<double>(), off, len, q, nq, S, "pairlist_sums");
}
