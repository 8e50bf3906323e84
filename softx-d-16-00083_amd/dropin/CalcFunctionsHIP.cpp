// HIP drop-in for the XaNSoNS CPU calculators.
//
// Defines the four calculators with the CPU signatures that main.cpp declares
// and calls (main.cpp:65-68, 151-233), and the ensemble builder
// CalcAndPrintAtoms (main.cpp:38, 130), and computes them on an MI355X
// through the C ABI of libxs_hip (include/xs_hip.h).  Compiled against the
// reference's own typedefs.h (config, block, vect3d); linked in place of the
// CPU calculator definitions (CalcFunctions.cpp:546-1092) and of
// CalcAndPrintAtoms (CalcFunctions.cpp:398-450).
//
// What stays on the host, in FP64, exactly as the CPU path does it:
//   the form-factor weighting x 2 f_a f_b, the self term N_el f_el^2, the
//   division by Ntot and the polarisation factor (CalcFunctions.cpp:592-626,
//   704-751, 872-906), the Euler grid and rotation matrices, the polar
//   (q, phi) mesh with cos(theta) = 1 - sin^2(theta), the 2D normalisation and
//   the phi average (CalcFunctions.cpp:999-1075).
// What runs on the GPU: the pair sums (direct and histogram), the histogram
//   itself (FP64 binning, so the counts equal calcHist's) and the PDFs.
//
// Ownership and output conventions follow the reference: outputs are
// allocated here with new[] and released by main (main.cpp:239-245).
// Errors: a failing device call prints the library message and exits with a
// non-zero status (the reference printed and continued on garbage,
// CalcFunctionsCUDA.cu:878-879).
//
// Environment:
//   XANSONS_HIP_DEVICE   device index (default 0)
//   XANSONS_PDF_RADIUS   "cpu" (default): PDF/rPDF radius of the CPU loop,
//                        which offsets pair types p >= 1 by p*Nhist bins
//                        (CalcFunctions.cpp:942, 950); "local": the radius of
//                        the CUDA/OpenCL kernels (CalcFunctionsCUDA.cu:1057)
//   XANSONS_HIP_SEED     seed of the ensemble builder's random draws
//                        (default: the clock, as the reference seeds its own)
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "typedefs.h"
#include "xs_hip.h"

extern int myid, numprocs;

namespace {

xs_ctx *g_ctx = nullptr;

void release_ctx() {
  if (g_ctx) xs_free(g_ctx);
  g_ctx = nullptr;
}

void check(int rc, const char *where) {
  if (rc != XS_OK) {
    std::fprintf(stderr, "Error in %s: %s (code %d)\n", where, xs_last_error(), rc);
    std::exit(EXIT_FAILURE);
  }
}

xs_ctx *ctx() {
  if (!g_ctx) {
    const char *dev = std::getenv("XANSONS_HIP_DEVICE");
    check(xs_init(dev ? std::atoi(dev) : 0, &g_ctx), "xs_init");
    std::atexit(release_ctx);
  }
  return g_ctx;
}

// Element-major structure-of-arrays copy of the ensemble (CalcFunctions.cpp:159-196
// order; the CUDA path flattens the same way, CalcFunctionsCUDA.cu:1548-1552).
std::vector<double> flatten(const vector<vect3d<double> > *ra, const unsigned *NatomEl,
                            unsigned Nel, size_t *n_out) {
  size_t n = 0;
  for (unsigned e = 0; e < Nel; ++e) n += NatomEl[e];
  std::vector<double> xyz(3 * n);
  size_t o = 0;
  for (unsigned e = 0; e < Nel; ++e)
    for (unsigned i = 0; i < NatomEl[e]; ++i, ++o) {
      xyz[o] = ra[e][i].x;
      xyz[n + o] = ra[e][i].y;
      xyz[2 * n + o] = ra[e][i].z;
    }
  *n_out = n;
  return xyz;
}

double weight(const config *cfg, const vector<double *> &FF, const vector<double> &SL,
              unsigned el, unsigned iq) {
  return cfg->source == xray ? FF[el][iq] : SL[el];
}

void polar_1d(double *I, const double *q, double lambda, unsigned Nq) {
  for (unsigned iq = 0; iq < Nq; ++iq) {
    const double st = q[iq] * (lambda * 0.25 / PI);
    const double c2t = 1. - 2. * SQR(st);
    I[iq] *= 0.5 * (1. + SQR(c2t));
  }
}

// I(q) from pair sums S[p*Nq + iq]: the combination of CalcFunctions.cpp:592-626.
void combine_1d(double *I, const double *S, const config *cfg, const unsigned *NatomEl,
                const vector<double *> &FF, const vector<double> &SL, const double *q,
                unsigned Ntot) {
  const unsigned Nq = cfg->q.N;
  for (unsigned iq = 0; iq < Nq; ++iq) I[iq] = 0;
  unsigned p = 0;
  for (unsigned a = 0; a < cfg->Nel; ++a)
    for (unsigned b = a; b < cfg->Nel; ++b, ++p)
      for (unsigned iq = 0; iq < Nq; ++iq)
        I[iq] += S[(size_t)p * Nq + iq] * weight(cfg, FF, SL, a, iq) * weight(cfg, FF, SL, b, iq) * 2.;
  for (unsigned e = 0; e < cfg->Nel; ++e)
    for (unsigned iq = 0; iq < Nq; ++iq) I[iq] += NatomEl[e] * SQR(weight(cfg, FF, SL, e, iq));
  for (unsigned iq = 0; iq < Nq; ++iq) I[iq] /= Ntot;
  if (cfg->PolarFactor) polar_1d(I, q, cfg->lambda, Nq);
}

float seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration_cast<std::chrono::duration<float> >(std::chrono::steady_clock::now() - t0)
      .count();
}

// Rotation matrix rows for intrinsic Euler angles, R = R_a1(e.x) R_a2(e.y) R_a3(e.z),
// axes per convention as numbered in typedefs.h:59-70.
void rotation(const vect3d<double> &e, unsigned convention, double R[9]) {
  static const int axes[12][3] = {{0, 2, 0}, {0, 1, 0}, {1, 0, 1}, {1, 2, 1}, {2, 1, 2}, {2, 0, 2},
                                  {0, 2, 1}, {0, 1, 2}, {1, 0, 2}, {1, 2, 0}, {2, 1, 0}, {2, 0, 1}};
  const double ang[3] = {e.x, e.y, e.z};
  double acc[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int s = 0; s < 3; ++s) {
    const int ax = axes[convention][s], u = (ax + 1) % 3, v = (ax + 2) % 3;
    const double c = std::cos(ang[s]), sn = std::sin(ang[s]);
    double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    M[ax * 3 + ax] = 1;
    M[u * 3 + u] = c;
    M[u * 3 + v] = -sn;
    M[v * 3 + u] = sn;
    M[v * 3 + v] = c;
    double T[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        T[i * 3 + j] = acc[i * 3] * M[j] + acc[i * 3 + 1] * M[3 + j] + acc[i * 3 + 2] * M[6 + j];
    std::memcpy(acc, T, sizeof(T));
  }
  std::memcpy(R, acc, sizeof(acc));
}

}  // namespace

void calcIntDebye(double **I, const config *cfg, const unsigned int *NatomEl,
                  const vector<vect3d<double> > *ra, vector<double *> FF, vector<double> SL,
                  const double *q, unsigned int Ntot, int NumOMPthreads) {
  (void)NumOMPthreads;
  const auto t0 = std::chrono::steady_clock::now();
  const unsigned Nq = cfg->q.N, P = cfg->Nel * (cfg->Nel + 1) / 2;
  size_t n = 0;
  std::vector<double> xyz = flatten(ra, NatomEl, cfg->Nel, &n);
  std::vector<double> S((size_t)P * Nq, 0.0);
  check(xs_debye_pairsum(ctx(), xyz.data(), n, NatomEl, cfg->Nel, q, Nq, S.data(), 0),
        "calcIntDebye");
  *I = new double[Nq];
  combine_1d(*I, S.data(), cfg, NatomEl, FF, SL, q, Ntot);
  if (!myid) cout << "1D pattern calculation time: " << seconds_since(t0) << " s" << endl;
}

void calcIntPartialDebye(double **I, const config *cfg, const unsigned int *NatomEl,
                         const vector<vect3d<double> > *ra, vector<double *> FF, vector<double> SL,
                         const double *q, const block *Block, unsigned int Ntot,
                         int NumOMPthreads) {
  (void)NumOMPthreads;
  const auto t0 = std::chrono::steady_clock::now();
  const unsigned Nq = cfg->q.N, Nb = cfg->Nblocks, Nel = cfg->Nel;
  const unsigned Nparts = Nb * (Nb + 1) / 2, P = Nel * (Nel + 1) / 2;
  // atoms of element e in block b (CalcFunctions.cpp:650-661)
  std::vector<unsigned> blk((size_t)Nel * Nb, 0);
  for (unsigned e = 0; e < Nel; ++e)
    for (unsigned b = 0; b < Nb; ++b)
      for (unsigned t = 0; t < Block[b].Nid; ++t)
        if (Block[b].id[t] == e) {
          blk[(size_t)e * Nb + b] = Block[b].NatomElAll[t];
          break;
        }
  size_t n = 0;
  std::vector<double> xyz = flatten(ra, NatomEl, Nel, &n);
  std::vector<double> S((size_t)P * Nparts * Nq, 0.0);
  check(xs_partial_debye_pairsum(ctx(), xyz.data(), n, NatomEl, Nel, blk.data(), Nb, q, Nq,
                                 S.data(), 0),
        "calcIntPartialDebye");
  const size_t Isize = (size_t)Nparts * Nq;
  *I = new double[Isize + Nq];
  double *out = *I;
  for (size_t i = 0; i < Isize + Nq; ++i) out[i] = 0;
  unsigned p = 0;
  for (unsigned a = 0; a < Nel; ++a)
    for (unsigned b = a; b < Nel; ++b, ++p)
      for (unsigned part = 0; part < Nparts; ++part)
        for (unsigned iq = 0; iq < Nq; ++iq)
          out[(size_t)(part + 1) * Nq + iq] += S[((size_t)p * Nparts + part) * Nq + iq] *
                                               weight(cfg, FF, SL, a, iq) *
                                               weight(cfg, FF, SL, b, iq) * 2.;
  for (unsigned e = 0; e < Nel; ++e)
    for (unsigned b = 0; b < Nb; ++b) {
      const size_t start = (size_t)Nq * (1 + (Nb * b - (b * (b + 1)) / 2 + b));
      for (unsigned iq = 0; iq < Nq; ++iq)
        out[start + iq] += blk[(size_t)e * Nb + b] * SQR(weight(cfg, FF, SL, e, iq));
    }
  for (size_t i = Nq; i < Isize + Nq; ++i) out[i] /= Ntot;
  if (cfg->PolarFactor)
    for (unsigned part = 1; part < Nparts + 1; ++part) polar_1d(out + (size_t)part * Nq, q, cfg->lambda, Nq);
  for (unsigned part = 1; part < Nparts + 1; ++part)
    for (unsigned iq = 0; iq < Nq; ++iq) out[iq] += out[(size_t)part * Nq + iq];
  if (!myid) cout << "1D pattern calculation time: " << seconds_since(t0) << " s" << endl;
}

void calcPDFandDebye(double **I, double **PDF, const config *cfg, const unsigned int *NatomEl,
                     const vector<vect3d<double> > *ra, vector<double *> FF, vector<double> SL,
                     const double *q, unsigned int Ntot, int NumOMPthreads) {
  (void)NumOMPthreads;
  auto t0 = std::chrono::steady_clock::now();
  const unsigned Nel = cfg->Nel, P = Nel * (Nel + 1) / 2, Nhist = cfg->Nhist;
  size_t n = 0;
  std::vector<double> xyz = flatten(ra, NatomEl, Nel, &n);
  std::vector<uint64_t> hist((size_t)P * Nhist, 0);
  uint64_t overflow = 0;
  check(xs_hist(ctx(), xyz.data(), n, NatomEl, Nel, cfg->hist_bin, Nhist, hist.data(), &overflow,
                XS_F64_BINNING),
        "calcPDFandDebye (histogram)");
  if (overflow && !myid)
    cerr << "Warning: " << overflow << " pair distances fell beyond the last histogram bin." << endl;
  if (!myid) cout << "Histogram calculation time: " << seconds_since(t0) << " s" << endl;
  if (cfg->scenario > Debye_hist) {
    t0 = std::chrono::steady_clock::now();
    const char *mode = std::getenv("XANSONS_PDF_RADIUS");
    const unsigned flags = (mode && std::strcmp(mode, "local") == 0) ? 0u : XS_PDF_CPU_RADIUS;
    *PDF = new double[(size_t)(1 + P) * Nhist];
    check(xs_pdf(ctx(), hist.data(), Nel, NatomEl, Nhist, cfg->hist_bin, (int)cfg->PDFtype, cfg->p0,
                 SL.data(), *PDF, flags),
          "calcPDFandDebye (PDF)");
    if (!myid) cout << "PDF calculation time: " << seconds_since(t0) << " s" << endl;
  }
  if (cfg->scenario == Debye_hist || cfg->scenario == DebyePDF) {
    t0 = std::chrono::steady_clock::now();
    const unsigned Nq = cfg->q.N;
    std::vector<double> S((size_t)P * Nq, 0.0);
    check(xs_hist_pairsum(ctx(), hist.data(), P, Nhist, cfg->hist_bin, q, Nq, S.data(), 0),
          "calcPDFandDebye (intensity)");
    *I = new double[Nq];
    double *out = *I;
    for (unsigned iq = 0; iq < Nq; ++iq) out[iq] = 0;
    unsigned p = 0;
    for (unsigned a = 0; a < Nel; ++a)
      for (unsigned b = a; b < Nel; ++b, ++p)
        for (unsigned iq = 0; iq < Nq; ++iq)
          out[iq] += 2. * S[(size_t)p * Nq + iq] * weight(cfg, FF, SL, a, iq) * weight(cfg, FF, SL, b, iq);
    for (unsigned e = 0; e < Nel; ++e)
      for (unsigned iq = 0; iq < Nq; ++iq) out[iq] += NatomEl[e] * SQR(weight(cfg, FF, SL, e, iq));
    for (unsigned iq = 0; iq < Nq; ++iq) out[iq] /= Ntot;
    if (cfg->PolarFactor) polar_1d(out, q, cfg->lambda, Nq);
    if (!myid) cout << "1D pattern calculation time: " << seconds_since(t0) << " s" << endl;
  }
}

void calcInt2D(double ***I2D, double **I, const config *cfg, const unsigned int *NatomEl,
               const vector<vect3d<double> > *ra, vector<double *> FF, vector<double> SL,
               const double *q, unsigned int Ntot, int NumOMPthreads) {
  (void)NumOMPthreads;
  const auto t0 = std::chrono::steady_clock::now();
  const unsigned Nq = cfg->q.N, Nfi = cfg->Nfi, Nel = cfg->Nel;
  const size_t N2D = (size_t)Nq * Nfi;
  // polar mesh: q (cos(theta) cos(phi), cos(theta) sin(phi), -sin(theta)) with
  // sin(theta) = q lambda/4pi and cos(theta) = 1 - sin^2(theta) (CalcFunctions.cpp:999-1005, 1031-1032)
  std::vector<double> qvec(3 * N2D), w((size_t)Nel * N2D);
  const double dfi = 2. * PI / Nfi;
  for (unsigned iq = 0; iq < Nq; ++iq) {
    const double st = q[iq] * (cfg->lambda * 0.25 / PI);
    const double ct = 1. - SQR(st);
    for (unsigned ifi = 0; ifi < Nfi; ++ifi) {
      const size_t k = (size_t)iq * Nfi + ifi;
      qvec[3 * k] = ct * std::cos(ifi * dfi) * q[iq];
      qvec[3 * k + 1] = ct * std::sin(ifi * dfi) * q[iq];
      qvec[3 * k + 2] = -st * q[iq];
      for (unsigned e = 0; e < Nel; ++e) w[(size_t)e * N2D + k] = weight(cfg, FF, SL, e, iq);
    }
  }
  // Euler midpoint grid (CalcFunctions.cpp:1007-1017)
  const mesh3d &E = cfg->Euler;
  const double da = E.N.x < 2 ? 0 : (E.max.x - E.min.x) / E.N.x;
  const double db = E.N.y < 2 ? 0 : (E.max.y - E.min.y) / E.N.y;
  const double dg = E.N.z < 2 ? 0 : (E.max.z - E.min.z) / E.N.z;
  const unsigned Nrot = E.N.x * E.N.y * E.N.z;
  std::vector<double> rot((size_t)Nrot * 9);
  unsigned r = 0;
  for (unsigned ia = 0; ia < E.N.x; ++ia)
    for (unsigned ib = 0; ib < E.N.y; ++ib)
      for (unsigned ig = 0; ig < E.N.z; ++ig, ++r) {
        vect3d<double> e(E.min.x + (ia + 0.5) * da, E.min.y + (ib + 0.5) * db, E.min.z + (ig + 0.5) * dg);
        rotation(e, cfg->EulerConvention, &rot[(size_t)r * 9]);
      }
  size_t n = 0;
  std::vector<double> xyz = flatten(ra, NatomEl, Nel, &n);
  std::vector<double> acc(N2D, 0.0);
  check(xs_amp2d(ctx(), xyz.data(), n, NatomEl, Nel, qvec.data(), N2D, rot.data(), Nrot, w.data(),
                 acc.data(), 0),
        "calcInt2D");
  *I = new double[Nq];
  *I2D = new double *[Nq];
  const double norm = 1. / (Ntot * E.N.x * E.N.y * E.N.z);
  for (unsigned iq = 0; iq < Nq; ++iq) {
    (*I2D)[iq] = new double[Nfi];
    double f = 1.0;
    if (cfg->PolarFactor) {
      const double st = q[iq] * (cfg->lambda * 0.25 / PI);
      const double c2t = 1. - 2. * SQR(st);
      f = 0.5 * (1. + SQR(c2t));
    }
    double s = 0;
    for (unsigned ifi = 0; ifi < Nfi; ++ifi) {
      double v = acc[(size_t)iq * Nfi + ifi] * norm;
      if (cfg->PolarFactor) v *= f;
      (*I2D)[iq][ifi] = v;
      s += v;
    }
    (*I)[iq] = s / Nfi;
  }
  if (!myid) cout << "2D pattern calculation time: " << seconds_since(t0) << " s" << endl;
}

// ---------------------------------------------------------------------------
// Ensemble builder.  CalcAndPrintAtoms (CalcFunctions.cpp:398-450) with the
// signature main.cpp declares (main.cpp:38): the atoms of every block are
// built on the GPU by xs_atoms_build (block::calcAtoms, 251-397), copied into
// the reference's ra[el] lists, and the reference's own helpers then compute
// the bin count and the density and print the atoms as it does.
// The random draws come from Philox counters keyed by XANSONS_HIP_SEED
// (default: the clock, like the reference's generator, 259-260).

double GetAtomicDensity(const vector<vect3d<double> > *r, unsigned int Nel, unsigned int Ntot);
unsigned int GetHistSize(const vector<vect3d<double> > *r, unsigned int Nel, double hist_bin);
int printAtoms(const vector<vect3d<double> > *ra, string name, map<string, unsigned int> ID,
               unsigned int Ntot);

static_assert(sizeof(vect3d<double>) == 3 * sizeof(double), "vect3d<double> must be 3 doubles");

unsigned int CalcAndPrintAtoms(config *cfg, block *Block, vector<vect3d<double> > **ra,
                               unsigned int **NatomEl, map<string, unsigned int> ID) {
  unsigned int Ntot = 0;
  *ra = new vector<vect3d<double> >[cfg->Nel];
  *NatomEl = new unsigned int[cfg->Nel];
  if (myid) return Ntot;
  const unsigned nb = cfg->Nblocks, nel = cfg->Nel;
  std::vector<xs_block_desc> desc(nb);
  std::vector<std::vector<unsigned> > el(nb);
  std::vector<std::vector<double> > ps(nb), cf(nb);
  for (unsigned b = 0; b < nb; ++b) {
    const block &B = Block[b];
    for (unsigned t = 0; t < B.Nid; ++t) el[b].insert(el[b].end(), B.NatomEl[t], B.id[t]);
    for (unsigned s = 0; s < B.Nsymm; ++s)
      for (int row = 0; row < 3; ++row) {
        const vect3d<double> &v = B.pSymm[row][s];
        ps[b].insert(ps[b].end(), {v.x, v.y, v.z});
      }
    for (unsigned c = 0; c < B.Ncopy; ++c)
      for (int row = 0; row < 3; ++row) {
        const vect3d<double> &v = B.cf[row][c];
        cf[b].insert(cf[b].end(), {v.x, v.y, v.z});
      }
    xs_block_desc &d = desc[b];
    d.natom = B.Natom;
    d.nsymm = B.Nsymm;
    d.ncopy = B.Ncopy;
    d.ncell[0] = B.Ncell.x;
    d.ncell[1] = B.Ncell.y;
    d.ncell[2] = B.Ncell.z;
    d.el = el[b].data();
    d.r_atom = &B.rAtom[0].x;
    d.occ = B.occ;
    d.dev = B.dev;
    d.p_symm = ps[b].data();
    d.r_symm = &B.rSymm[0].x;
    d.cell = &B.e[0].x;
    d.r_copy = &B.rCopy[0].x;
    d.cf = cf[b].data();
    d.r_mean[0] = B.rMean.x;
    d.r_mean[1] = B.rMean.y;
    d.r_mean[2] = B.rMean.z;
    d.rcut = B.cutoff ? B.Rcut : 0.0;
    d.rcut_copies = B.cutoffcopies ? B.RcutCopies : 0.0;
    d.dev_mol = B.dev_mol;
    d.mol_rotation = B.ro_mol ? 1 : 0;
  }
  const char *s = std::getenv("XANSONS_HIP_SEED");
  const uint64_t seed =
      s ? std::strtoull(s, nullptr, 10)
        : (uint64_t)std::chrono::system_clock::now().time_since_epoch().count();
  xs_atoms *A = nullptr;
  check(xs_atoms_build(ctx(), desc.data(), nb, nel, seed, &A), "xs_atoms_build");
  std::vector<unsigned> el_count(nel), block_el_count((size_t)nb * nel);
  uint64_t n = 0;
  check(xs_atoms_info(A, el_count.data(), block_el_count.data(), &n), "xs_atoms_info");
  std::vector<double> xyz(3 * n);
  check(xs_atoms_copy(A, xyz.data(), 0), "xs_atoms_copy");
  xs_atoms_free(A);
  size_t o = 0;
  for (unsigned e = 0; e < nel; ++e) {
    (*ra)[e].resize(el_count[e]);
    for (unsigned i = 0; i < el_count[e]; ++i, ++o)
      (*ra)[e][i].assign(xyz[o], xyz[n + o], xyz[2 * n + o]);
    (*NatomEl)[e] = el_count[e];
    Ntot += el_count[e];
  }
  for (unsigned b = 0; b < nb; ++b) {  // NatomElAll and N of each block (CalcFunctions.cpp:320, 394)
    block &B = Block[b];
    B.N = 0;
    for (unsigned t = 0; t < B.Nid; ++t) {
      B.NatomElAll[t] = block_el_count[(size_t)b * nel + B.id[t]];
      B.N += B.NatomElAll[t];
    }
  }
  if (cfg->scenario > Debye) cfg->Nhist = GetHistSize(*ra, cfg->Nel, cfg->hist_bin);
  if (cfg->PrintAtoms) printAtoms(*ra, cfg->name, ID, Ntot);
  if ((cfg->scenario > Debye_hist) && (!cfg->p0)) {
    cfg->p0 = GetAtomicDensity(*ra, cfg->Nel, Ntot);
    cout << "Approximate atomic density of the sample is " << cfg->p0 << " A^(-3).\n" << endl;
  }
  return Ntot;
}

// ---------------------------------------------------------------------------
// RearrangementInt (CalcFunctions.cpp:469-545), signature of main.cpp:44.
// The cut-off pair sums and the neighbour sums run on the GPU
// (xs_cutoff_pairsum, xs_pairlist_pairsum); the weights, the pair count N,
// the partial-intensity slots and the polarisation factor are computed here
// as the reference does (its Rmax2 and PolarFactor1D are called as they are).

double Rmax2(vect3d<double> rAtom, vector<vect3d<double> > rAtomNeighb);
void PolarFactor1D(double *I, const double *q, double lambda, unsigned int Nq);

void RearrangementInt(double *I0, const double *q, const config *cfg, const block *Block,
                      vector<double *> FF, vector<double> SL, unsigned int Ntot) {
  const unsigned Nq = cfg->q.N, nel = cfg->Nel, nb = cfg->Nblocks;
  const bool xr = cfg->source == xray, partial = cfg->calcPartialIntensity;
  long long N = SQR(Ntot);  // formed in unsigned int, as the reference does
  std::vector<double> I(partial ? (size_t)(nb + 1) * Nq : Nq, 0.0);
  auto weight = [&](unsigned coeff, unsigned a, unsigned b, unsigned iq) {
    return xr ? coeff * FF[a][iq] * FF[b][iq] : coeff * SL[a] * SL[b];
  };
  for (unsigned iB = 0; iB < nb; ++iB) {
    const unsigned iqST = partial ? Nq : 0;  // one slot for every block (CalcFunctions.cpp:478-479)
    const block &B = Block[iB];
    if (!B.rearrangement) continue;
    const unsigned coeff = B.Ncopy * B.Ncell.x * B.Ncell.y * B.Ncell.z * B.Nsymm;
    const size_t n = B.Natom;
    std::vector<unsigned> el;
    for (unsigned t = 0; t < B.Nid; ++t) el.insert(el.end(), B.NatomEl[t], B.id[t]);
    std::vector<double> xyz(3 * n), r2(n);
    for (size_t i = 0; i < n; ++i) {
      xyz[i] = B.rAtom[i].x;
      xyz[n + i] = B.rAtom[i].y;
      xyz[2 * n + i] = B.rAtom[i].z;
      r2[i] = Rmax2(B.rAtom[i], B.rAtomNeighb[i]) + SQR(B.Rcorr);
    }
    const size_t nt = (size_t)nel * nel;
    std::vector<double> S_cut(nt * Nq), S_nb(nt * Nq);
    std::vector<uint64_t> npairs(nt);
    check(xs_cutoff_pairsum(ctx(), xyz.data(), n, el.data(), nel, r2.data(), q, Nq, S_cut.data(),
                            npairs.data(), 0),
          "xs_cutoff_pairsum");
    std::vector<double> rn;
    std::vector<unsigned> tn;
    std::vector<unsigned long long> per_type(nt, 0), per_el(nel, 0);
    for (size_t i = 0; i < n; ++i)
      for (size_t k = 0; k < B.rAtomNeighb[i].size(); ++k) {
        rn.push_back((B.rAtom[i] - B.rAtomNeighb[i][k]).mag());
        tn.push_back(el[i] * nel + B.idNeighb[i][k]);
        ++per_type[tn.back()];
        ++per_el[el[i]];
      }
    check(xs_pairlist_pairsum(ctx(), rn.data(), tn.data(), rn.size(), (unsigned)nt, q, Nq,
                              S_nb.data(), 0),
          "xs_pairlist_pairsum");
    for (unsigned ia = 0; ia < B.Nid; ++ia) {
      const unsigned id = B.id[ia];
      for (unsigned ja = 0; ja < B.Nid; ++ja) {  // CalcFunctions.cpp:484-503
        const unsigned jd = B.id[ja];
        const size_t t = (size_t)id * nel + jd;
        N -= (long long)coeff * (long long)npairs[t];
        for (unsigned iq = 0; iq < Nq; ++iq)
          I[iqST + iq] -= weight(coeff, id, jd, iq) * S_cut[t * Nq + iq];
      }
      for (unsigned b = 0; b < nel; ++b) {  // CalcFunctions.cpp:506-529
        const size_t t = (size_t)id * nel + b;
        if (!per_type[t]) continue;
        for (unsigned iq = 0; iq < Nq; ++iq)
          I[iqST + iq] += weight(coeff, id, b, iq) * S_nb[t * Nq + iq];
      }
      N += (long long)coeff * (long long)per_el[id];
    }
  }
  if (partial) {  // CalcFunctions.cpp:532-541
    unsigned iqST = Nq;
    for (unsigned iB = 0; iB < nb; ++iB, iqST += Nq) {
      const unsigned i0qST = Nq * (1 + iB * (2 * nb - iB + 1) / 2);
      if (cfg->PolarFactor) PolarFactor1D(I.data() + iqST, q, cfg->lambda, Nq);
      for (unsigned iq = 0; iq < Nq; ++iq) {
        I[iq] += I[iqST + iq];
        I0[i0qST + iq] = (I0[i0qST + iq] * Ntot + I[iqST + iq]) / sqrt(double(N));
      }
    }
  } else if (cfg->PolarFactor) {
    PolarFactor1D(I.data(), q, cfg->lambda, Nq);
  }
  for (unsigned iq = 0; iq < Nq; ++iq) I0[iq] = (I0[iq] * Ntot + I[iq]) / sqrt(double(N));
}
