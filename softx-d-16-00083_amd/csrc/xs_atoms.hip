// Ensemble builder: block::calcAtoms (CalcFunctions.cpp:251-397) for all
// blocks of an ensemble on the GPU.
//
// Site merging (CalcFunctions.cpp:276-319).  The reference compares every
// pair of the Natom*Nsymm symmetry images of the cell atoms under up to 125
// cell translations -- O((Natom Nsymm)^2) on one core, hours for a
// finished 10^6-atom structure given as one cell.  Here the images are
// sorted into a uniform grid of cells at least 0.505 A wide and each
// (image, translation) thread tests only the images in the 27 grid cells
// around the translated position.  The reference accepts a pair only below
// 0.5 A, so every pair it can accept is among those; each is tested with
// its own FP64 expression, so the accepted (image, image, translation)
// triples are exactly the reference's.  The host sorts the few triples into
// the reference's list order and derives the covering images and their
// cumulative occupancies.
//
// Placement (CalcFunctions.cpp:320-392).  One thread per (copy, cell,
// covering image): occupancy draw, molecule rotation and displacement,
// cut-off, copy transform, atom displacement, copies' cut-off, all in FP64
// with the reference's operation order (no contraction).  Random draws come
// from Philox4x64-10 counters, so they do not depend on the launch order.
// Each thread writes a key (element, block) and a position; a stable radix
// sort by key yields the order of the reference's ra[el] lists, and a
// gather writes the element-major SoA result.
#include "xs_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#pragma clang fp contract(off)

namespace xs {
namespace atoms {

constexpr double kMergeR2 = 0.25;  // squared merge distance, CalcFunctions.cpp:301
constexpr double kMergeR = 0.5;
constexpr double kPi = 3.14159265358979323846;
constexpr int kThreads = 256;

// One entry of a covering image's list: the cell atom and symmetry image
// placed if the occupancy draw is <= occ.
struct Entry {
  double occ;
  uint32_t atom, symm, el, pad;
};

// ---- random draws ----------------------------------------------------------

struct P4 {
  uint64_t v[4];
};

// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter (c0, c1, c2, 0).
__device__ inline P4 philox(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t k0, uint64_t k1) {
  uint64_t c3 = 0;
  for (int round = 0; round < 10; ++round) {
    if (round) {
      k0 += 0x9E3779B97F4A7C15ull;
      k1 += 0xBB67AE8584CAA73Bull;
    }
    const uint64_t lo0 = 0xD2E7470EE14C6C93ull * c0, hi0 = __umul64hi(0xD2E7470EE14C6C93ull, c0);
    const uint64_t lo1 = 0xCA5A826395121157ull * c2, hi1 = __umul64hi(0xCA5A826395121157ull, c2);
    const uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  return P4{{c0, c1, c2, c3}};
}

__device__ inline double unit(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }

// |N(0, sd)| from two uniforms (Box-Muller).
__device__ inline double half_normal(double sd, uint64_t ur, uint64_t ua) {
  const double z = sqrt(-2.0 * log(1.0 - unit(ur))) * cos(2.0 * kPi * unit(ua));
  const double v = sd * z;
  return v < 0 ? -v : v;
}

// ---- site merging kernels --------------------------------------------------

__global__ void site_positions(const double *r_atom, const double *p_symm, const double *r_symm,
                               unsigned natom, unsigned nsymm, double *sx, double *sy,
                               double *sz) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)natom * nsymm) return;
  const size_t a = i / nsymm, s = i - a * nsymm;
  const double rx = r_atom[3 * a], ry = r_atom[3 * a + 1], rz = r_atom[3 * a + 2];
  const double *P = p_symm + 9 * s;
  // pSymm[row][s].dot(rAtom[a]) + rSymm[s], CalcFunctions.cpp:270-271
  sx[i] = (P[0] * rx + P[1] * ry + P[2] * rz) + r_symm[3 * s];
  sy[i] = (P[3] * rx + P[4] * ry + P[5] * rz) + r_symm[3 * s + 1];
  sz[i] = (P[6] * rx + P[7] * ry + P[8] * rz) + r_symm[3 * s + 2];
}

__global__ void fill_u32(uint32_t *p, size_t n, uint32_t v) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void iota_u32(uint32_t *p, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

// [beg[k], end[k]) = positions of key k in the sorted key array.
__global__ void key_ranges(const uint32_t *skey, size_t n, uint32_t *beg, uint32_t *end) {
  const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t k = skey[p];
  if (p == 0 || skey[p - 1] != k) beg[k] = (uint32_t)p;
  if (p + 1 == n || skey[p + 1] != k) end[k] = (uint32_t)(p + 1);
}

struct MergeArgs {
  const double *sx, *sy, *sz;  // by image index
  const double *SX, *SY, *SZ;  // by sorted position (grid order)
  const uint32_t *order, *beg, *end;
  const double *trans;
  unsigned ntrans;
  unsigned M;
  Grid G;
  unsigned long long *count;
  uint32_t *out;  // (i, j, t) triples
  unsigned long long cap;
};

// One thread per (image i, translation t): every later image j with
// |(r_i - r_j) + T_t|^2 < 0.25 (CalcFunctions.cpp:300-301) is appended.
__global__ __launch_bounds__(kThreads) void merge_images(MergeArgs A) {
  const size_t gid = blockIdx.x * (size_t)kThreads + threadIdx.x;
  if (gid >= (size_t)A.M * A.ntrans) return;
  const unsigned i = (unsigned)(gid / A.ntrans), t = (unsigned)(gid - (size_t)i * A.ntrans);
  const double xi = A.sx[i], yi = A.sy[i], zi = A.sz[i];
  const double tx = A.trans[3 * t], ty = A.trans[3 * t + 1], tz = A.trans[3 * t + 2];
  const int cx = grid_coord(xi + tx, A.G.lo[0], A.G.inv, A.G.g[0]);
  const int cy = grid_coord(yi + ty, A.G.lo[1], A.G.inv, A.G.g[1]);
  const int cz = grid_coord(zi + tz, A.G.lo[2], A.G.inv, A.G.g[2]);
  for (int z = max(cz - 1, 0); z <= min(cz + 1, A.G.g[2] - 1); ++z)
    for (int y = max(cy - 1, 0); y <= min(cy + 1, A.G.g[1] - 1); ++y)
      for (int x = max(cx - 1, 0); x <= min(cx + 1, A.G.g[0] - 1); ++x) {
        const uint32_t k = grid_cell(A.G, x, y, z);
        const uint32_t p1 = A.end[k];
        for (uint32_t p = A.beg[k]; p < p1; ++p) {
          const uint32_t j = A.order[p];
          if (j <= i) continue;
          const double dx = xi - A.SX[p] + tx;
          const double dy = yi - A.SY[p] + ty;
          const double dz = zi - A.SZ[p] + tz;
          if (dx * dx + dy * dy + dz * dz < kMergeR2) {
            const unsigned long long slot = atomicAdd(A.count, 1ull);
            if (slot < A.cap) {
              A.out[3 * slot] = i;
              A.out[3 * slot + 1] = j;
              A.out[3 * slot + 2] = t;
            }
          }
        }
      }
}

// ---- placement -------------------------------------------------------------

struct PlaceArgs {
  const uint32_t *rep_off;
  const Entry *ent;
  const double *r_atom, *dev, *p_symm, *r_symm, *r_copy, *cf;
  uint32_t *key;
  double *px, *py, *pz;
  unsigned long long ncand;
  double e[9], r_mean[3], r_mean_c[3];
  double rcut2, rcc2, dev_mol;
  unsigned nrep, ncell[3];
  unsigned key_stride, key_block, key_none;
  int cutoff, cutoff_c, mol_dev, mol_rot;
  unsigned long long seed, block;
};

__global__ __launch_bounds__(kThreads) void place_atoms(PlaceArgs A) {
  // candidate counts stay below 2^31 (checked by the host), so the index
  // arithmetic is 32-bit
  const uint32_t L = blockIdx.x * (uint32_t)kThreads + threadIdx.x;
  if (L >= A.ncand) return;
  const uint32_t C = L / A.nrep;  // (copy, cell) index
  const unsigned r = L - C * A.nrep;
  uint32_t t = C;
  const unsigned Z = t % A.ncell[2];
  t /= A.ncell[2];
  const unsigned Y = t % A.ncell[1];
  t /= A.ncell[1];
  const unsigned X = t % A.ncell[0];
  const uint32_t copy = t / A.ncell[0];
  uint32_t key = A.key_none;
  const unsigned e1 = A.rep_off[r + 1];
  unsigned k = A.rep_off[r];
  // The occupancy draw is in [0, 1): an entry with occ >= 1 at the head of
  // the list is taken whatever it is, so the draw is only made otherwise
  // (counter-based, so drawing later gives the same numbers).
  P4 d;
  bool drawn = false;
  if (!(A.ent[k].occ >= 1.0)) {
    d = philox(L, 0, 0, A.seed, A.block);
    drawn = true;
    const double randval = unit(d.v[0]);
    while (k < e1 && !(randval <= A.ent[k].occ)) ++k;  // CalcFunctions.cpp:346
  }
  if (k < e1) {
    const Entry E = A.ent[k];
    const double *ra = A.r_atom + 3 * (size_t)E.atom;
    const double *P = A.p_symm + 9 * (size_t)E.symm;
    const double rx = ra[0], ry = ra[1], rz = ra[2];
    const double x0 = P[0] * rx + P[1] * ry + P[2] * rz;
    const double y0 = P[3] * rx + P[4] * ry + P[5] * rz;
    const double z0 = P[6] * rx + P[7] * ry + P[8] * rz;
    double x1 = x0, y1 = y0, z1 = z0;
    if (A.mol_rot) {  // CalcFunctions.cpp:331-339, 351
      const P4 w = philox(C, 4, E.symm, A.seed, A.block);
      const double al = 2.0 * kPi * unit(w.v[0]), ga = kPi * unit(w.v[1]),
                   be = 2.0 * kPi * unit(w.v[2]);
      const double ca = cos(al), sa = sin(al), cb = cos(be), sb = sin(be), cg = cos(ga),
                   sg = sin(ga);
      // rows of Rz(al) Rx(be) Rz(ga)
      const double R0 = ca * cg - sa * cb * sg, R1 = -ca * sg - sa * cb * cg, R2 = sa * sb;
      const double R3 = sa * cg + ca * cb * sg, R4 = -sa * sg + ca * cb * cg, R5 = -ca * sb;
      const double R6 = sb * sg, R7 = sb * cg, R8 = cb;
      x1 = x0 * R0 + y0 * R1 + z0 * R2;
      y1 = x0 * R3 + y0 * R4 + z0 * R5;
      z1 = x0 * R6 + y0 * R7 + z0 * R8;
    }
    double mx = 0.0, my = 0.0, mz = 0.0;
    if (A.mol_dev) {  // CalcFunctions.cpp:322-329
      const P4 w = philox(C, 2, E.symm, A.seed, A.block);
      const P4 v = philox(C, 3, E.symm, A.seed, A.block);
      const double R = half_normal(A.dev_mol, w.v[2], v.v[0]);
      const double fi = 2.0 * kPi * unit(w.v[0]), th = kPi * unit(w.v[1]);
      mx = R * sin(th) * cos(fi);
      my = R * sin(th) * sin(fi);
      mz = R * cos(th);
    }
    // celltrans, CalcFunctions.cpp:321
    const double cx = A.e[0] * (double)X + A.e[3] * (double)Y + A.e[6] * (double)Z;
    const double cy = A.e[1] * (double)X + A.e[4] * (double)Y + A.e[7] * (double)Z;
    const double cz = A.e[2] * (double)X + A.e[5] * (double)Y + A.e[8] * (double)Z;
    const double *rs = A.r_symm + 3 * (size_t)E.symm;
    x1 += rs[0] + cx + mx;  // CalcFunctions.cpp:353
    y1 += rs[1] + cy + my;
    z1 += rs[2] + cz + mz;
    const double ux = x1 - A.r_mean[0], uy = y1 - A.r_mean[1], uz = z1 - A.r_mean[2];
    if (!A.cutoff || ux * ux + uy * uy + uz * uz <= A.rcut2) {  // CalcFunctions.cpp:354
      const double *F = A.cf + 9 * copy, *rc = A.r_copy + 3 * copy;
      double x2 = rc[0] + (ux * F[0] + uy * F[1] + uz * F[2]);  // 355-356
      double y2 = rc[1] + (ux * F[3] + uy * F[4] + uz * F[5]);
      double z2 = rc[2] + (ux * F[6] + uy * F[7] + uz * F[8]);
      const double sd = A.dev[E.atom];
      if (sd >= 1.e-7) {  // CalcFunctions.cpp:357-363
        if (!drawn) d = philox(L, 0, 0, A.seed, A.block);
        const P4 v = philox(L, 1, 0, A.seed, A.block);
        const double R = half_normal(sd, d.v[3], v.v[0]);
        const double fi = 2.0 * kPi * unit(d.v[1]), th = kPi * unit(d.v[2]);
        x2 += R * sin(th) * cos(fi);
        y2 += R * sin(th) * sin(fi);
        z2 += R * cos(th);
      }
      const double wx = x2 - A.r_mean_c[0], wy = y2 - A.r_mean_c[1], wz = z2 - A.r_mean_c[2];
      if (!A.cutoff_c || wx * wx + wy * wy + wz * wz <= A.rcc2) {  // CalcFunctions.cpp:364
        key = E.el * A.key_stride + A.key_block;
        A.px[L] = x2;
        A.py[L] = y2;
        A.pz[L] = z2;
      }
    }
  }
  A.key[L] = key;
}

__global__ void gather_atoms(const uint32_t *order, size_t n, const double *px, const double *py,
                             const double *pz, double *out) {
  const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t g = order[p];
  out[p] = px[g];
  out[n + p] = py[g];
  out[2 * n + p] = pz[g];
}

// ---- host side -------------------------------------------------------------

inline unsigned nblk(size_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

// Host staging of several arrays for one upload.
struct Pack {
  std::vector<char> host;
  template <class T>
  size_t add(const T *src, size_t n) {
    const size_t off = align_up(host.size(), 256);
    host.resize(off + std::max<size_t>(n * sizeof(T), 1));
    if (n) std::memcpy(host.data() + off, src, n * sizeof(T));
    return off;
  }
};

// Covering images of one block and their entry lists (CSR).
struct Covering {
  std::vector<uint32_t> off;
  std::vector<Entry> ent;
};

// Translations of the merge test, CalcFunctions.cpp:289-299: -2..2 cells
// along every axis with more than one cell, X outermost.
std::vector<double> merge_translations(const xs_block_desc &B) {
  int lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = B.ncell[k] == 1 ? 0 : -2;
    hi[k] = B.ncell[k] == 1 ? 1 : 3;
  }
  const double *e = B.cell;
  std::vector<double> T;
  for (int X = lo[0]; X < hi[0]; ++X)
    for (int Y = lo[1]; Y < hi[1]; ++Y)
      for (int Z = lo[2]; Z < hi[2]; ++Z)
        for (int k = 0; k < 3; ++k)
          T.push_back(e[k] * (double)X + e[3 + k] * (double)Y + e[6 + k] * (double)Z);
  return T;
}

// Accepted (i, j, t) triples of the merge test for one block.
int merge_triples(xs_ctx *ctx, const xs_block_desc &B, const std::vector<double> &trans,
                  std::vector<uint32_t> &triples) {
  const size_t M = (size_t)B.natom * B.nsymm;
  const unsigned nT = (unsigned)(trans.size() / 3);
  triples.clear();
  if (M < 2) return XS_OK;
  hipEvent_t ev;
  XS_TRY(prof_begin(ctx, "atoms_merge", &ev));
  Pack in;
  const size_t o_ra = in.add(B.r_atom, 3 * (size_t)B.natom);
  const size_t o_ps = in.add(B.p_symm, 9 * (size_t)B.nsymm);
  const size_t o_rs = in.add(B.r_symm, 3 * (size_t)B.nsymm);
  const size_t o_tr = in.add(trans.data(), trans.size());
  DevMem din, pos;
  XS_TRY(din.alloc(in.host.size(), "merge inputs"));
  XS_TRY(upload(ctx, din.p, in.host.data(), in.host.size()));
  XS_TRY(pos.alloc(3 * M * sizeof(double), "image positions"));
  double *sx = pos.at<double>(), *sy = sx + M, *sz = sy + M;
  site_positions<<<nblk(M), kThreads, 0, ctx->stream>>>(din.at<double>(o_ra), din.at<double>(o_ps),
                                                       din.at<double>(o_rs), B.natom, B.nsymm, sx,
                                                       sy, sz);
  XS_HIP_CHECK(hipGetLastError());
  // images sorted into a grid of >= 0.5 A cells
  SortedPoints sp;
  XS_TRY(grid_sort(ctx, sx, sy, sz, M, kMergeR, sp));
  // count pass, then write pass
  DevMem cnt, out;
  XS_TRY(cnt.alloc(sizeof(unsigned long long), "merge count"));
  MergeArgs A{sx, sy, sz, sp.X, sp.Y, sp.Z, sp.order, sp.beg, sp.end, din.at<double>(o_tr), nT,
              (unsigned)M, sp.G, cnt.at<unsigned long long>(), nullptr, 0};
  XS_HIP_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), ctx->stream));
  const size_t nthr = M * nT;
  merge_images<<<nblk(nthr), kThreads, 0, ctx->stream>>>(A);
  XS_HIP_CHECK(hipGetLastError());
  unsigned long long count = 0;
  XS_TRY(download(ctx, &count, cnt.p, sizeof(count)));
  if (count > (1ull << 28))
    return fail(XS_ERR_RANGE, "xs_atoms_build: more than 2^28 coinciding image pairs");
  if (count) {
    XS_TRY(out.alloc(3 * count * sizeof(uint32_t), "merge triples"));
    A.out = out.at<uint32_t>();
    A.cap = count;
    XS_HIP_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), ctx->stream));
    merge_images<<<nblk(nthr), kThreads, 0, ctx->stream>>>(A);
    XS_HIP_CHECK(hipGetLastError());
    unsigned long long again = 0;
    XS_TRY(download(ctx, &again, cnt.p, sizeof(again)));
    if (again != count)
      return fail(XS_ERR_HIP, "xs_atoms_build: merge passes disagree (" + std::to_string(count) +
                                  " vs " + std::to_string(again) + ")");
    triples.resize(3 * count);
    XS_TRY(download(ctx, triples.data(), out.p, triples.size() * sizeof(uint32_t)));
  }
  XS_TRY(prof_end(ctx, "atoms_merge", ev));
  return XS_OK;
}

// Entry lists in the reference's order (self, then later images by index
// and translation) and the covering images, CalcFunctions.cpp:280-319, 340-345.
void covering(const xs_block_desc &B, std::vector<uint32_t> &triples, Covering &cov) {
  const size_t M = (size_t)B.natom * B.nsymm, n = triples.size() / 3;
  std::vector<size_t> ord(n);
  std::iota(ord.begin(), ord.end(), size_t(0));
  std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
    const uint32_t *x = &triples[3 * a], *y = &triples[3 * b];
    if (x[0] != y[0]) return x[0] < y[0];
    if (x[1] != y[1]) return x[1] < y[1];
    return x[2] < y[2];
  });
  std::vector<size_t> off(M + 1);
  std::vector<Entry> ent;
  ent.reserve(M + n);
  size_t m = 0;
  for (size_t i = 0; i < M; ++i) {
    off[i] = ent.size();
    const uint32_t a = (uint32_t)(i / B.nsymm);
    ent.push_back(Entry{B.occ[a], a, (uint32_t)(i % B.nsymm), B.el[a], 0});
    for (; m < n && triples[3 * ord[m]] == i; ++m) {
      const uint32_t j = triples[3 * ord[m] + 1], aj = j / B.nsymm;
      const Entry &last = ent.back();
      const double occ = B.el[aj] != last.el ? last.occ + B.occ[aj] : last.occ;
      ent.push_back(Entry{occ, aj, j % B.nsymm, B.el[aj], 0});
    }
  }
  off[M] = ent.size();
  std::vector<char> covered(M, 0);
  cov.off.assign(1, 0);
  cov.ent.clear();
  for (size_t i = 0; i < M; ++i) {
    if (covered[i]) continue;
    for (size_t k = off[i]; k < off[i + 1]; ++k) {
      covered[(size_t)ent[k].atom * B.nsymm + ent[k].symm] = 1;
      cov.ent.push_back(ent[k]);
    }
    cov.off.push_back((uint32_t)cov.ent.size());
  }
}

bool finite_all(const double *p, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(p[i])) return false;
  return true;
}

int check_block(const xs_block_desc &B, unsigned b, unsigned n_el) {
  const std::string w = "xs_atoms_build: block " + std::to_string(b) + ": ";
  if (!B.natom || !B.nsymm || !B.ncopy || !B.ncell[0] || !B.ncell[1] || !B.ncell[2])
    return fail(XS_ERR_ARG, w + "natom, nsymm, ncopy and ncell must be >= 1");
  if (!B.el || !B.r_atom || !B.occ || !B.dev || !B.p_symm || !B.r_symm || !B.cell || !B.r_copy ||
      !B.cf)
    return fail(XS_ERR_ARG, w + "null array");
  for (unsigned a = 0; a < B.natom; ++a)
    if (B.el[a] >= n_el) return fail(XS_ERR_ARG, w + "element index >= n_el");
  if ((size_t)B.natom * B.nsymm >= (size_t)INT_MAX / 4)
    return fail(XS_ERR_RANGE, w + "too many symmetry images");
  if (!finite_all(B.r_atom, 3 * (size_t)B.natom) || !finite_all(B.p_symm, 9 * (size_t)B.nsymm) ||
      !finite_all(B.r_symm, 3 * (size_t)B.nsymm) || !finite_all(B.cell, 9) ||
      !finite_all(B.r_copy, 3 * (size_t)B.ncopy) || !finite_all(B.cf, 9 * (size_t)B.ncopy) ||
      !finite_all(B.occ, B.natom) || !finite_all(B.dev, B.natom))
    return fail(XS_ERR_ARG, w + "non-finite input");
  return XS_OK;
}

}  // namespace atoms

void destroy_atoms(xs_atoms *a) {
  if (!a) return;
  if (a->ctx) {
    (void)hipSetDevice(a->ctx->device);
    (void)hipStreamSynchronize(a->ctx->stream);
    a->ctx->atom_sets.erase(a);
  }
  if (a->xyz) (void)hipFree(a->xyz);
  delete a;
}

}  // namespace xs

using namespace xs;
using namespace xs::atoms;

extern "C" int xs_atoms_build(xs_ctx *ctx, const xs_block_desc *blocks, unsigned nblocks,
                              unsigned n_el, uint64_t seed, xs_atoms **atoms) {
  if (!atoms) return fail(XS_ERR_ARG, "xs_atoms_build: null output handle");
  *atoms = nullptr;
  if (!ctx) return fail(XS_ERR_ARG, "xs_atoms_build: null xs_ctx");
  if (!blocks || !nblocks || !n_el) return fail(XS_ERR_ARG, "xs_atoms_build: no blocks or elements");
  if ((uint64_t)n_el * nblocks >= (uint64_t)INT_MAX)
    return fail(XS_ERR_RANGE, "xs_atoms_build: n_el * nblocks too large");
  for (unsigned b = 0; b < nblocks; ++b) XS_TRY(check_block(blocks[b], b, n_el));
  XS_TRY(activate(ctx));

  // site merging, block by block
  std::vector<Covering> cov(nblocks);
  std::vector<unsigned long long> base(nblocks + 1, 0);
  for (unsigned b = 0; b < nblocks; ++b) {
    const xs_block_desc &B = blocks[b];
    const std::vector<double> trans = merge_translations(B);
    std::vector<uint32_t> triples;
    XS_TRY(merge_triples(ctx, B, trans, triples));
    covering(B, triples, cov[b]);
    const double nc = (double)B.ncopy * B.ncell[0] * B.ncell[1] * B.ncell[2] *
                      (double)(cov[b].off.size() - 1);
    if (nc + (double)base[b] >= (double)INT_MAX)
      return fail(XS_ERR_RANGE, "xs_atoms_build: more than 2^31 candidate sites");
    base[b + 1] = base[b] + (unsigned long long)nc;
  }
  const size_t total = base[nblocks];
  const uint32_t key_none = n_el * nblocks;
  const size_t nkeys = (size_t)key_none + 1;

  // placement
  DevMem cand, ranges;
  XS_TRY(cand.alloc(total * (4 * sizeof(uint32_t) + 3 * sizeof(double)), "candidates"));
  double *px = cand.at<double>(), *py = px + total, *pz = py + total;
  uint32_t *key = reinterpret_cast<uint32_t *>(pz + total), *val = key + total,
           *skey = val + total, *sval = skey + total;
  std::vector<std::unique_ptr<DevMem>> tables;
  hipEvent_t ev;
  XS_TRY(prof_begin(ctx, "atoms_place", &ev));
  for (unsigned b = 0; b < nblocks; ++b) {
    const xs_block_desc &B = blocks[b];
    Pack in;
    const size_t o_off = in.add(cov[b].off.data(), cov[b].off.size());
    const size_t o_ent = in.add(cov[b].ent.data(), cov[b].ent.size());
    const size_t o_ra = in.add(B.r_atom, 3 * (size_t)B.natom);
    const size_t o_dev = in.add(B.dev, B.natom);
    const size_t o_ps = in.add(B.p_symm, 9 * (size_t)B.nsymm);
    const size_t o_rs = in.add(B.r_symm, 3 * (size_t)B.nsymm);
    const size_t o_rc = in.add(B.r_copy, 3 * (size_t)B.ncopy);
    const size_t o_cf = in.add(B.cf, 9 * (size_t)B.ncopy);
    tables.emplace_back(new DevMem);
    DevMem &T = *tables.back();
    XS_TRY(T.alloc(in.host.size(), "block tables"));
    XS_TRY(upload(ctx, T.p, in.host.data(), in.host.size()));
    PlaceArgs A{};
    A.rep_off = T.at<uint32_t>(o_off);
    A.ent = T.at<Entry>(o_ent);
    A.r_atom = T.at<double>(o_ra);
    A.dev = T.at<double>(o_dev);
    A.p_symm = T.at<double>(o_ps);
    A.r_symm = T.at<double>(o_rs);
    A.r_copy = T.at<double>(o_rc);
    A.cf = T.at<double>(o_cf);
    A.key = key + base[b];
    A.px = px + base[b];
    A.py = py + base[b];
    A.pz = pz + base[b];
    A.ncand = base[b + 1] - base[b];
    for (int k = 0; k < 9; ++k) A.e[k] = B.cell[k];
    double rmc[3] = {0, 0, 0};  // mean copy position, CalcFunctions.cpp:263-266
    if (B.rcut_copies > 0) {
      for (unsigned c = 0; c < B.ncopy; ++c)
        for (int k = 0; k < 3; ++k) rmc[k] += B.r_copy[3 * c + k];
      for (int k = 0; k < 3; ++k) rmc[k] = rmc[k] / (double)B.ncopy;
    }
    for (int k = 0; k < 3; ++k) {
      A.r_mean[k] = B.r_mean[k];
      A.r_mean_c[k] = rmc[k];
      A.ncell[k] = B.ncell[k];
    }
    A.rcut2 = B.rcut * B.rcut;
    A.rcc2 = B.rcut_copies * B.rcut_copies;
    A.dev_mol = B.dev_mol;
    A.nrep = (unsigned)(cov[b].off.size() - 1);
    A.key_stride = nblocks;
    A.key_block = b;
    A.key_none = key_none;
    A.cutoff = B.rcut > 0;
    A.cutoff_c = B.rcut_copies > 0;
    A.mol_dev = B.dev_mol >= 1.e-7;
    A.mol_rot = B.mol_rotation != 0;
    A.seed = seed;
    A.block = b;
    if (A.ncand) {
      place_atoms<<<nblk(A.ncand), kThreads, 0, ctx->stream>>>(A);
      XS_HIP_CHECK(hipGetLastError());
    }
  }
  XS_TRY(prof_end(ctx, "atoms_place", ev));

  // element-major order: stable sort by (element, block), placed atoms first
  XS_TRY(prof_begin(ctx, "atoms_order", &ev));
  iota_u32<<<nblk(total), kThreads, 0, ctx->stream>>>(val, total);
  XS_HIP_CHECK(hipGetLastError());
  XS_TRY(sort_pairs_u32(ctx, key, skey, val, sval, total, bits_for(key_none)));
  XS_TRY(ranges.alloc(2 * nkeys * sizeof(uint32_t), "key ranges"));
  uint32_t *kb = ranges.at<uint32_t>(), *ke = kb + nkeys;
  fill_u32<<<nblk(2 * nkeys), kThreads, 0, ctx->stream>>>(kb, 2 * nkeys, 0u);
  key_ranges<<<nblk(total), kThreads, 0, ctx->stream>>>(skey, total, kb, ke);
  XS_HIP_CHECK(hipGetLastError());
  std::vector<uint32_t> kr(2 * nkeys);
  XS_TRY(download(ctx, kr.data(), ranges.p, kr.size() * sizeof(uint32_t)));

  std::unique_ptr<xs_atoms> res(new xs_atoms);
  res->ctx = ctx;
  res->n_el = n_el;
  res->nblocks = nblocks;
  res->el_count.assign(n_el, 0);
  res->block_el_count.assign((size_t)nblocks * n_el, 0);
  size_t n = 0;
  for (unsigned el = 0; el < n_el; ++el)
    for (unsigned b = 0; b < nblocks; ++b) {
      const size_t k = (size_t)el * nblocks + b;
      const unsigned c = kr[nkeys + k] - kr[k];
      res->el_count[el] += c;
      res->block_el_count[(size_t)b * n_el + el] = c;
      n += c;
    }
  res->n = n;
  if (n) {
    if (hipMalloc(&res->xyz, 3 * n * sizeof(double)) != hipSuccess) {
      res->xyz = nullptr;
      (void)hipGetLastError();
      return fail(XS_ERR_NOMEM, "xs_atoms_build: hipMalloc of the atoms failed");
    }
    gather_atoms<<<nblk(n), kThreads, 0, ctx->stream>>>(sval, n, px, py, pz, res->xyz);
    XS_HIP_CHECK(hipGetLastError());
  }
  XS_TRY(prof_end(ctx, "atoms_order", ev));
  XS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  ctx->atom_sets.insert(res.get());
  *atoms = res.release();
  return XS_OK;
}

extern "C" int xs_atoms_info(const xs_atoms *atoms, unsigned *el_count, unsigned *block_el_count,
                             uint64_t *n_atoms) {
  if (!atoms) return fail(XS_ERR_ARG, "xs_atoms_info: null atoms");
  if (el_count) std::copy(atoms->el_count.begin(), atoms->el_count.end(), el_count);
  if (block_el_count)
    std::copy(atoms->block_el_count.begin(), atoms->block_el_count.end(), block_el_count);
  if (n_atoms) *n_atoms = atoms->n;
  return XS_OK;
}

extern "C" int xs_atoms_copy(xs_atoms *atoms, double *xyz, unsigned flags) {
  if (!atoms) return fail(XS_ERR_ARG, "xs_atoms_copy: null atoms");
  if (!atoms->n) return XS_OK;
  if (!xyz) return fail(XS_ERR_ARG, "xs_atoms_copy: null xyz");
  xs_ctx *ctx = atoms->ctx;
  XS_TRY(activate(ctx));
  const size_t bytes = 3 * atoms->n * sizeof(double);
  if (flags & XS_DEVICE_PTRS) {
    XS_HIP_CHECK(hipMemcpyAsync(xyz, atoms->xyz, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return XS_OK;
  }
  return download(ctx, xyz, atoms->xyz, bytes);
}

extern "C" int xs_atoms_free(xs_atoms *atoms) {
  destroy_atoms(atoms);
  return XS_OK;
}
