// Internal declarations shared by the HIP translation units of libxs_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "xs_mfma_plan.h"

#include "xs_hip.h"

namespace xs {

// Thread-local error text behind xs_last_error().
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define XS_HIP_CHECK(call)                                                     \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return ::xs::fail(XS_ERR_HIP, std::string(#call) + ": " +               \
                                        hipGetErrorString(e_));                \
  } while (0)

#define XS_TRY(call)                                                           \
  do {                                                                         \
    int r_ = (call);                                                           \
    if (r_ != XS_OK) return r_;                                                \
  } while (0)

inline unsigned ntypes_of(unsigned n_el) { return n_el * (n_el + 1) / 2; }

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Event pair around one kernel launch, resolved lazily by xs_profile_read.
struct ProfRecord {
  std::string name;
  hipEvent_t start, stop;
};

// Per-plan state of xs_batch_run: the captured graph and its key, and the
// device buffers the run owns.
struct BatchRunState {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void *key_hist = nullptr, *key_S = nullptr, *key_pdf = nullptr, *key_q = nullptr;
  unsigned key_nq = 0, key_flags = 0;
  int key_pdf_type = -1;
  std::vector<double> key_rho0, key_sl;
  void *buf = nullptr;   // pair-sum and PDF descriptors, host-mode staging
  size_t buf_cap = 0;
  MfmaBatchPlan mm;      // XS_PAIRSUM_MFMA: columns, groups, buffer sizes (mm.nq = 0: none)
  bool mm_uploaded = false;
  size_t mm_at = 0;      // offset in buf of the uploaded matrix-core tables
};

}  // namespace xs

// A histogram plan over an ensemble of structures (built by xs_batch_plan).
struct xs_batch {
  xs_ctx *ctx = nullptr;
  bool f64 = true;
  bool quot_f32 = false;  // FP32 binning with the checked FP32 quotient form
  double bin = 0;
  unsigned nstruct = 0;
  std::vector<unsigned> n_el, el_count, nhist, ntot;
  std::vector<size_t> el_first;                   // first el_count index of each structure
  std::vector<unsigned long long> hist_off;       // per structure, in bins
  std::vector<unsigned long long> pdf_off;        // per structure, in doubles
  std::vector<unsigned long long> type_first;     // first pair type of each structure
  std::vector<unsigned long long> type_hist_off;  // per pair type
  std::vector<uint32_t> type_nhist;
  unsigned long long hist_len = 0, pdf_len = 0;
  size_t ntypes_total = 0;
  size_t items8192 = 0, items32768 = 0;
  size_t natoms = 0;      // per axis in the sorted coordinate arrays
  void *coords = nullptr;  // sorted SoA (float or double)
  void *tiles = nullptr, *boxes = nullptr, *types = nullptr, *items = nullptr;
  void *recs = nullptr;  // FP64 binning: J records, 4 doubles per atom
  unsigned long long *ovf = nullptr;  // one overflow counter per structure
  xs::BatchRunState run;
};

// Atoms built by xs_atoms_build: element-major SoA on the device.
struct xs_atoms {
  xs_ctx *ctx = nullptr;
  unsigned n_el = 0, nblocks = 0;
  std::vector<unsigned> el_count;        // [n_el]
  std::vector<unsigned> block_el_count;  // [nblocks * n_el]
  size_t n = 0;
  double *xyz = nullptr;  // x[n], y[n], z[n]
};

struct xs_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  size_t max_lds = 0;

  // Device workspace carved per call (grown on demand, never shrunk).
  void *ws = nullptr;
  size_t ws_size = 0;
  // Device buffer for per-call tables (tiles, work items), grown on demand.
  void *aux = nullptr;
  size_t aux_size = 0;
  // Device buffer of the matrix-core pair sums (operand slices), grown on demand.
  void *mat = nullptr;
  size_t mat_size = 0;
  // Device copies of host inputs and outputs (coordinates in, counts out) of
  // calls that take host pointers, apart from the workspace, grown on demand.
  void *io = nullptr;
  size_t io_size = 0;
  // Pinned host staging for small uploads (item lists, parameter tables).
  void *pin = nullptr;
  size_t pin_size = 0;

  // Batch plans made on this context (freed by xs_free if still alive).
  std::set<xs_batch *> plans;
  // Built ensembles on this context (freed by xs_free if still alive).
  std::set<xs_atoms *> atom_sets;

  // FP32 binning: bins (float bits) checked for the FP32 quotient form; the
  // outcome of the exhaustive check of the FP32 square root (-1 not run yet,
  // 0 failed, 1 IEEE for every positive float, 2 IEEE from 2^-96 up and
  // below 2^-47 under it); the device counters of both checks (4 x 8 bytes,
  // allocated on first use).
  std::map<uint32_t, bool> quot_f32_ok;
  int sqrt_f32_state = -1;
  unsigned long long *quot_check = nullptr;

  bool profiling = false;
  std::vector<xs::ProfRecord> prof_pending;
  std::map<std::string, std::pair<double, unsigned long long>> prof_done;
  std::vector<hipEvent_t> event_pool;
};

namespace xs {

// Make the context's device current for this thread.
int activate(xs_ctx *ctx);
// Ensure ctx->ws holds at least `bytes` (synchronises the stream if it grows).
int ensure_workspace(xs_ctx *ctx, size_t bytes);
// Ensure ctx->aux holds at least `bytes` (synchronises the stream if it grows).
int ensure_aux(xs_ctx *ctx, size_t bytes);
// Ensure ctx->mat holds at least `bytes` (synchronises the stream if it grows).
int ensure_matrix(xs_ctx *ctx, size_t bytes);
// Ensure ctx->io holds at least `bytes` (synchronises the stream if it grows).
int ensure_io(xs_ctx *ctx, size_t bytes);
// Ensure ctx->pin holds at least `bytes` (synchronises the stream if it grows).
int ensure_pinned(xs_ctx *ctx, size_t bytes);

// Profiling: wrap a launch with events when collection is on.
int prof_begin(xs_ctx *ctx, const char *name, hipEvent_t *start);
int prof_end(xs_ctx *ctx, const char *name, hipEvent_t start);

void release_batch_run(BatchRunState &r);
// Free a plan's device memory and unregister it from its context.
void destroy_batch(xs_batch *b);
// Launches of a batch plan: zero fills and kernels only, no host work or
// synchronisation, so they can be captured into a graph.  `prof` wraps them
// in profiling events (not while capturing).
int batch_hist_launch(xs_batch *b, unsigned long long *hist_d, bool prof);

// Free a built ensemble's device memory and unregister it from its context.
void destroy_atoms(xs_atoms *a);

// Device address of the allocation-free Debye entry point's scratch array,
// taken once per device (the device must be current).
int debye_scratch_init(int device);

// Device allocation owned by one call (freed on scope exit).
struct DevMem {
  void *p = nullptr;
  DevMem() = default;
  DevMem(const DevMem &) = delete;
  DevMem &operator=(const DevMem &) = delete;
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes, const char *what);
  template <class T>
  T *at(size_t byte_off = 0) const {
    return reinterpret_cast<T *>(static_cast<char *>(p) + byte_off);
  }
};

// Synchronous copies on the context stream.
int upload(xs_ctx *ctx, void *dst, const void *src, size_t bytes);
int download(xs_ctx *ctx, void *dst, const void *src, size_t bytes);

// Stable radix sort of (key, value) uint32 pairs on the low `bits` key bits
// (hipcub), synchronous on the context stream.
int sort_pairs_u32(xs_ctx *ctx, const uint32_t *k_in, uint32_t *k_out, const uint32_t *v_in,
                   uint32_t *v_out, size_t n, int bits);
int bits_for(uint64_t v);

// Uniform grid of cells over n points (device SoA x, y, z) for searches of
// pairs closer than a cut-off: cell edge >= min_cell (grown by 1.25x until
// the grid has at most max(2^16, 8 n) and 2^24 cells), points sorted by cell.
// Points i, j closer than min_cell / 1.01 lie in the same or adjacent cells,
// computed with grid_coord from either point's coordinates.
struct Grid {
  double lo[3];
  double inv;
  int g[3];
};

__device__ inline int grid_coord(double v, double lo, double inv, int g) {
  double f = floor((v - lo) * inv);
  f = fmin(fmax(f, -2.0), (double)g + 1.0);  // also maps NaN into range
  return (int)f;
}

__device__ inline uint32_t grid_cell(const Grid &G, int x, int y, int z) {
  return ((uint32_t)z * (uint32_t)G.g[1] + (uint32_t)y) * (uint32_t)G.g[0] + (uint32_t)x;
}

struct SortedPoints {
  Grid G;
  size_t n = 0, ncells = 0;
  uint32_t *order = nullptr;              // order[p]: point at sorted position p
  uint32_t *beg = nullptr, *end = nullptr;  // [beg[c], end[c]): positions of cell c
  double *X = nullptr, *Y = nullptr, *Z = nullptr;  // coordinates by sorted position
  DevMem mem;
};
int grid_sort(xs_ctx *ctx, const double *x, const double *y, const double *z, size_t n,
              double min_cell, SortedPoints &out);

// Pair sums of many histograms: one chunk = up to 2048 bins of one pair type.
struct PairsumChunk {
  unsigned long long hist_off;  // first bin of the chunk in the histogram array
  uint32_t type;                // row of S (S[type * nq + iq])
  uint32_t k0, kn, pad;         // bin index of the chunk start within its histogram, length
};
constexpr unsigned kPairsumChunk = 2048;
// S[ntypes * nq] is zeroed and accumulated.
int batch_pairsum_launch(xs_ctx *ctx, const unsigned long long *hist, const PairsumChunk *chunks,
                         size_t nchunks, size_t ntypes, double bin, const double *q, unsigned nq,
                         double *S, bool prof);

// H3 as an int8 matrix-core product (device pointers, S[ncols * nq] overwritten).
int pairsum_mfma(xs_ctx *ctx, const unsigned long long *hist, unsigned ncols, unsigned nhist,
                 double bin, const double *q, unsigned nq, double *S);

// Launches only (no allocation, copy or wait; graph-capturable): slices of
// the sinc matrix into a_buf, of the counts into b_buf, then the products;
// S[type * nq + iq] is overwritten for every type.
int pairsum_mfma_batch(xs_ctx *ctx, const MfmaBatchPlan &plan, const unsigned long long *hist,
                       const MfmaCol *cols_d, const MfmaTile *tiles_d, void *a_buf, void *b_buf,
                       double bin, const double *q, unsigned nq, double *S, bool prof);

// PDFs of many structures; blk0 = first 256-bin block of the structure.
struct PdfDesc {
  unsigned long long hist_off, pdf_off;
  uint32_t nhist, n_el, ntot, blk0;
  double rho0;
  uint32_t el_count[16];
  double sl[16];
};
int batch_pdf_launch(xs_ctx *ctx, const unsigned long long *hist, const PdfDesc *desc,
                     unsigned nstruct, unsigned nblocks, int pdf_type, bool cpu_radius, double bin,
                     double *pdf, bool prof);

// Bump allocator over the workspace for one call.
struct Carver {
  char *base;
  size_t used = 0;
  explicit Carver(void *b) : base(static_cast<char *>(b)) {}
  template <class T>
  T *take(size_t count) {
    used = align_up(used, 256);
    T *p = reinterpret_cast<T *>(base + used);
    used += count * sizeof(T);
    return p;
  }
  static size_t need(std::initializer_list<size_t> sizes) {
    size_t s = 0;
    for (size_t b : sizes) s = align_up(s, 256) + b;
    return align_up(s, 256);
  }
};

}  // namespace xs
