/*
 * xs_oracle.c -- CPU oracle for the XaNSoNS scattering calculators.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product path): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Built with -O2 -fno-fast-math -ffp-contract=off so that every expression
 * below is evaluated with IEEE rounding in the written order, as the
 * reference's g++ -O3 x86-64 build does (no FMA in the baseline ISA).
 * Function-by-function reference citations are in xs_oracle.h.
 */
#include "xs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define XS_PI 3.14159265358979323846

static int clamp_threads(int nthreads) { return nthreads < 1 ? 1 : nthreads; }

static int thread_id(void) {
#ifdef _OPENMP
  return omp_get_thread_num();
#else
  return 0;
#endif
}

unsigned or_pair_index(unsigned a, unsigned b, unsigned n_el) {
  /* rows (0,0..n-1), (1,1..n-1), ...: row a starts after sum_{t<a}(n-t). */
  return a * n_el - a * (a - 1) / 2 + (b - a);
}

/* The k-th row visited by team member t of T in the reference's zig-zag
 * interleave (CalcFunctions.cpp:573-575): t, 2T-1-t, 2T+t, 4T-1-t, ... */
static size_t zigzag_row(size_t k, int t, int T) {
  size_t base = k * (size_t)T;
  return base + ((k & 1) ? (size_t)(T - 1 - t) : (size_t)t);
}

unsigned or_hist_size(const double *xyz, size_t n, double bin) {
  const double *x = xyz, *y = xyz + n, *z = xyz + 2 * n;
  if (n == 0) return 1;
  double lo[3] = {x[0], y[0], z[0]}, hi[3] = {x[0], y[0], z[0]};
  for (size_t i = 0; i < n; i++) {
    if (hi[0] < x[i]) hi[0] = x[i];
    if (hi[1] < y[i]) hi[1] = y[i];
    if (hi[2] < z[i]) hi[2] = z[i];
    if (lo[0] > x[i]) lo[0] = x[i];
    if (lo[1] > y[i]) lo[1] = y[i];
    if (lo[2] > z[i]) lo[2] = z[i];
  }
  double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  double diag = sqrt(dx * dx + dy * dy + dz * dz);
  return (unsigned)(diag / bin) + 1;
}

double or_atomic_density(const double *xyz, size_t n) {
  const double *x = xyz, *y = xyz + n, *z = xyz + 2 * n;
  double c[3] = {0, 0, 0};
  for (size_t i = 0; i < n; i++) {
    c[0] += x[i];
    c[1] += y[i];
    c[2] += z[i];
  }
  c[0] /= (double)n;
  c[1] /= (double)n;
  c[2] /= (double)n;
  double r2max = 0;
  for (size_t i = 0; i < n; i++) {
    double dx = c[0] - x[i], dy = c[1] - y[i], dz = c[2] - z[i];
    double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > r2max) r2max = d2;
  }
  double size = 2. * sqrt(r2max);
  double rs = 0.33333 * size;
  double rs2 = rs * rs;
  unsigned count = 0;
  for (size_t i = 0; i < n; i++) {
    double dx = c[0] - x[i], dy = c[1] - y[i], dz = c[2] - z[i];
    if (dx * dx + dy * dy + dz * dz < rs2) count++;
  }
  double R = sqrt(rs2);
  return count / (4. / 3. * XS_PI * R * R * R);
}

/* Shared driver for both histogram precisions: each team member owns a
 * private histogram copy and visits rows in the zig-zag order; the copies
 * are summed afterwards (CalcFunctions.cpp:765-804). */
#define HIST_BODY(T, SQRT)                                                     \
  size_t ntot = 0;                                                             \
  for (unsigned e = 0; e < n_el; e++) ntot += el_count[e];                     \
  const unsigned ntypes = n_el * (n_el + 1) / 2;                               \
  const size_t hsize = (size_t)ntypes * nhist;                                 \
  const T *x = xyz, *y = xyz + ntot, *z = xyz + 2 * ntot;                      \
  int nt = clamp_threads(nthreads);                                            \
  uint64_t *priv = (uint64_t *)calloc((size_t)nt * (hsize + 1), sizeof(uint64_t)); \
  if (!priv) return -1;                                                        \
  _Pragma("omp parallel num_threads(nt)") {                                    \
    int t = thread_id();                                                       \
    uint64_t *h = priv + (size_t)t * (hsize + 1);                              \
    uint64_t *ovf = h + hsize;                                                 \
    size_t oa = 0;                                                             \
    unsigned p = 0;                                                            \
    for (unsigned a = 0; a < n_el; oa += el_count[a], a++) {                   \
      size_t ob = oa;                                                          \
      for (unsigned b = a; b < n_el; ob += el_count[b], b++, p++) {            \
        uint64_t *hp = h + (size_t)p * nhist;                                  \
        for (size_t k = 0;; k++) {                                             \
          size_t i = zigzag_row(k, t, nt);                                     \
          if (i >= el_count[a]) break;                                         \
          T xi = x[oa + i], yi = y[oa + i], zi = z[oa + i];                    \
          size_t j0 = (a == b) ? i + 1 : 0;                                    \
          for (size_t j = j0; j < el_count[b]; j++) {                          \
            T dx = xi - x[ob + j], dy = yi - y[ob + j], dz = zi - z[ob + j];   \
            T s = dx * dx + dy * dy;                                           \
            s = s + dz * dz;                                                   \
            T r = SQRT(s);                                                     \
            unsigned idx = (unsigned)(r / bin);                                \
            if (idx < nhist)                                                   \
              hp[idx] += 1;                                                    \
            else                                                               \
              *ovf += 1;                                                       \
          }                                                                    \
        }                                                                      \
      }                                                                        \
    }                                                                          \
  }                                                                            \
  memset(hist, 0, hsize * sizeof(uint64_t));                                   \
  uint64_t ov = 0;                                                             \
  for (int t = 0; t < nt; t++) {                                               \
    const uint64_t *h = priv + (size_t)t * (hsize + 1);                        \
    for (size_t i = 0; i < hsize; i++) hist[i] += h[i];                        \
    ov += h[hsize];                                                            \
  }                                                                            \
  if (overflow) *overflow = ov;                                                \
  free(priv);                                                                  \
  return 0;

int or_hist_f64(const double *xyz, const unsigned *el_count, unsigned n_el,
                double bin, unsigned nhist, uint64_t *hist, uint64_t *overflow,
                int nthreads) {
  HIST_BODY(double, sqrt)
}

int or_hist_f32(const float *xyz, const unsigned *el_count, unsigned n_el,
                float bin, unsigned nhist, uint64_t *hist, uint64_t *overflow,
                int nthreads) {
  HIST_BODY(float, sqrtf)
}

void or_debye_pairsum(const double *xyz, const unsigned *el_count, unsigned n_el,
                      const double *q, unsigned nq, double *S, int nthreads) {
  size_t ntot = 0;
  for (unsigned e = 0; e < n_el; e++) ntot += el_count[e];
  const double *x = xyz, *y = xyz + ntot, *z = xyz + 2 * ntot;
  const unsigned ntypes = n_el * (n_el + 1) / 2;
  int nt = clamp_threads(nthreads);
  double *acc = (double *)calloc((size_t)nt * nq, sizeof(double));
  memset(S, 0, (size_t)ntypes * nq * sizeof(double));
  size_t oa = 0;
  unsigned p = 0;
  for (unsigned a = 0; a < n_el; oa += el_count[a], a++) {
    size_t ob = oa;
    for (unsigned b = a; b < n_el; ob += el_count[b], b++, p++) {
#pragma omp parallel num_threads(nt)
      {
        int t = thread_id();
        double *my = acc + (size_t)t * nq;
        memset(my, 0, nq * sizeof(double));
        for (size_t k = 0;; k++) {
          size_t i = zigzag_row(k, t, nt);
          if (i >= el_count[a]) break;
          size_t j0 = (a == b) ? i + 1 : 0;
          for (size_t j = j0; j < el_count[b]; j++) {
            double dx = x[oa + i] - x[ob + j], dy = y[oa + i] - y[ob + j],
                   dz = z[oa + i] - z[ob + j];
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            for (unsigned iq = 0; iq < nq; iq++) {
              double qr = r * q[iq];
              my[iq] += sin(qr) / (qr + 0.00000001);
            }
          }
        }
      }
      for (int t = 0; t < nt; t++)
        for (unsigned iq = 0; iq < nq; iq++) S[(size_t)p * nq + iq] += acc[(size_t)t * nq + iq];
    }
  }
  free(acc);
}

void or_partial_debye_pairsum(const double *xyz, const unsigned *el_count,
                              unsigned n_el, const unsigned *blk_count,
                              unsigned nblocks, const double *q, unsigned nq,
                              double *S, int nthreads) {
  size_t ntot = 0;
  for (unsigned e = 0; e < n_el; e++) ntot += el_count[e];
  const double *x = xyz, *y = xyz + ntot, *z = xyz + 2 * ntot;
  const unsigned ntypes = n_el * (n_el + 1) / 2;
  const unsigned nparts = nblocks * (nblocks + 1) / 2;
  const size_t psize = (size_t)nparts * nq;
  int nt = clamp_threads(nthreads);
  double *acc = (double *)calloc((size_t)nt * psize, sizeof(double));
  memset(S, 0, (size_t)ntypes * psize * sizeof(double));
  /* block id of every atom, element by element (block-major inside each
   * element, CalcFunctions.cpp:650-661) */
  unsigned *blk = (unsigned *)malloc((ntot ? ntot : 1) * sizeof(unsigned));
  {
    size_t o = 0;
    for (unsigned e = 0; e < n_el; e++)
      for (unsigned b = 0; b < nblocks; b++)
        for (unsigned k = 0; k < blk_count[e * nblocks + b]; k++) blk[o++] = b;
  }
  size_t oa = 0;
  unsigned p = 0;
  for (unsigned a = 0; a < n_el; oa += el_count[a], a++) {
    size_t ob = oa;
    for (unsigned b = a; b < n_el; ob += el_count[b], b++, p++) {
#pragma omp parallel num_threads(nt)
      {
        int t = thread_id();
        double *my = acc + (size_t)t * psize;
        memset(my, 0, psize * sizeof(double));
        for (size_t k = 0;; k++) {
          size_t i = zigzag_row(k, t, nt);
          if (i >= el_count[a]) break;
          size_t j0 = (a == b) ? i + 1 : 0;
          unsigned bi = blk[oa + i];
          for (size_t j = j0; j < el_count[b]; j++) {
            unsigned bj = blk[ob + j];
            unsigned lo = bi < bj ? bi : bj, hi = bi < bj ? bj : bi;
            unsigned part = nblocks * lo - (lo * (lo + 1)) / 2 + hi;
            double dx = x[oa + i] - x[ob + j], dy = y[oa + i] - y[ob + j],
                   dz = z[oa + i] - z[ob + j];
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            double *dst = my + (size_t)part * nq;
            for (unsigned iq = 0; iq < nq; iq++) {
              double qr = r * q[iq];
              dst[iq] += sin(qr) / (qr + 0.00000001);
            }
          }
        }
      }
      for (int t = 0; t < nt; t++)
        for (size_t i = 0; i < psize; i++) S[(size_t)p * psize + i] += acc[(size_t)t * psize + i];
    }
  }
  free(blk);
  free(acc);
}

void or_hist_pairsum(const uint64_t *hist, unsigned ntypes, unsigned nhist,
                     double bin, const double *q, unsigned nq, double *S,
                     int nthreads) {
  int nt = clamp_threads(nthreads);
  double *acc = (double *)calloc((size_t)nt * nq, sizeof(double));
  for (unsigned p = 0; p < ntypes; p++) {
    const uint64_t *h = hist + (size_t)p * nhist;
#pragma omp parallel num_threads(nt)
    {
      int t = thread_id();
      double *my = acc + (size_t)t * nq;
      memset(my, 0, nq * sizeof(double));
      for (unsigned k = (unsigned)t; k < nhist; k += (unsigned)nt) {
        if (!h[k]) continue;
        double r = ((double)k + 0.5) * bin;
        for (unsigned iq = 0; iq < nq; iq++) {
          double qr = r * q[iq];
          my[iq] += h[k] * sin(qr) / (qr + 0.00000001);
        }
      }
    }
    for (unsigned iq = 0; iq < nq; iq++) {
      double s = 0;
      for (int t = 0; t < nt; t++) s += acc[(size_t)t * nq + iq];
      S[(size_t)p * nq + iq] = s;
    }
  }
  free(acc);
}

void or_polar_1d(double *I, const double *q, unsigned nq, double lambda) {
  for (unsigned iq = 0; iq < nq; iq++) {
    double st = q[iq] * (lambda * 0.25 / XS_PI);
    double c2t = 1. - 2. * (st * st);
    I[iq] *= 0.5 * (1. + c2t * c2t);
  }
}

static double weight(const double *ff, const double *sl, unsigned el, unsigned iq,
                     unsigned nq) {
  return ff ? ff[(size_t)el * nq + iq] : sl[el];
}

void or_intensity_1d(const double *S, unsigned n_el, const unsigned *el_count,
                     const double *ff, const double *sl, const double *q,
                     unsigned nq, int polar, double lambda, double *I) {
  unsigned ntot = 0;
  for (unsigned e = 0; e < n_el; e++) ntot += el_count[e];
  for (unsigned iq = 0; iq < nq; iq++) I[iq] = 0;
  unsigned p = 0;
  for (unsigned a = 0; a < n_el; a++)
    for (unsigned b = a; b < n_el; b++, p++)
      for (unsigned iq = 0; iq < nq; iq++)
        I[iq] += S[(size_t)p * nq + iq] * weight(ff, sl, a, iq, nq) *
                 weight(ff, sl, b, iq, nq) * 2.;
  for (unsigned e = 0; e < n_el; e++)
    for (unsigned iq = 0; iq < nq; iq++) {
      double w = weight(ff, sl, e, iq, nq);
      I[iq] += el_count[e] * (w * w);
    }
  for (unsigned iq = 0; iq < nq; iq++) I[iq] /= ntot;
  if (polar) or_polar_1d(I, q, nq, lambda);
}

void or_partial_intensity_1d(const double *S, unsigned n_el,
                             const unsigned *blk_count, unsigned nblocks,
                             const double *ff, const double *sl,
                             const double *q, unsigned nq, int polar,
                             double lambda, unsigned ntot, double *I) {
  const unsigned nparts = nblocks * (nblocks + 1) / 2;
  const size_t psize = (size_t)nparts * nq;
  for (size_t i = 0; i < psize + nq; i++) I[i] = 0;
  unsigned p = 0;
  for (unsigned a = 0; a < n_el; a++)
    for (unsigned b = a; b < n_el; b++, p++)
      for (unsigned part = 0; part < nparts; part++)
        for (unsigned iq = 0; iq < nq; iq++)
          I[(size_t)(part + 1) * nq + iq] +=
              S[((size_t)p * nparts + part) * nq + iq] * weight(ff, sl, a, iq, nq) *
              weight(ff, sl, b, iq, nq) * 2.;
  for (unsigned e = 0; e < n_el; e++)
    for (unsigned bk = 0; bk < nblocks; bk++) {
      size_t start = (size_t)nq * (1 + (nblocks * bk - (bk * (bk + 1)) / 2 + bk));
      for (unsigned iq = 0; iq < nq; iq++) {
        double w = weight(ff, sl, e, iq, nq);
        I[start + iq] += blk_count[e * nblocks + bk] * (w * w);
      }
    }
  for (size_t i = nq; i < psize + nq; i++) I[i] /= ntot;
  if (polar)
    for (unsigned part = 1; part < nparts + 1; part++) or_polar_1d(I + (size_t)part * nq, q, nq, lambda);
  for (unsigned part = 1; part < nparts + 1; part++)
    for (unsigned iq = 0; iq < nq; iq++) I[iq] += I[(size_t)part * nq + iq];
}

void or_pdf(const uint64_t *hist, unsigned n_el, const unsigned *el_count,
            unsigned nhist, double bin, int type, double rho0, const double *sl,
            int cpu_radius_offset, double *pdf) {
  unsigned ntot = 0;
  for (unsigned e = 0; e < n_el; e++) ntot += el_count[e];
  const unsigned ntypes = n_el * (n_el + 1) / 2;
  for (size_t i = 0; i < (size_t)(1 + ntypes) * nhist; i++) pdf[i] = 0;
  double fav2 = 0;
  for (unsigned e = 0; e < n_el; e++) fav2 += sl[e] * el_count[e];
  fav2 /= ntot;
  fav2 *= fav2;
  size_t start = 0;
  for (unsigned a = 0; a < n_el; a++) {
    for (unsigned b = a; b < n_el; b++, start += nhist) {
      double wab = sl[a] * sl[b] / fav2;
      double *part = pdf + nhist;
      if (type == 0) {
        double mult = 2. / (bin * ntot);
        for (size_t i = start; i < start + nhist; i++) part[i] = hist[i] * mult;
      } else if (type == 1) {
        double mult = 0.5 / (XS_PI * bin * rho0 * ntot);
        for (size_t i = start; i < start + nhist; i++) {
          double r = ((cpu_radius_offset ? i : i - start) + 0.5) * bin;
          part[i] = hist[i] * mult / (r * r);
        }
      } else {
        double mult = 2. / (bin * ntot);
        double sub = (b > a) ? 8. * XS_PI * rho0 * (double)el_count[a] *
                                   (double)el_count[b] / ((double)ntot * (double)ntot)
                             : 4. * XS_PI * rho0 * ((double)el_count[a] * (double)el_count[a]) /
                                   ((double)ntot * (double)ntot);
        for (size_t i = start; i < start + nhist; i++) {
          double r = ((cpu_radius_offset ? i : i - start) + 0.5) * bin;
          part[i] = hist[i] * mult / r - sub * r;
        }
      }
      for (unsigned i = 0; i < nhist; i++) pdf[i] += part[start + i] * wab;
    }
  }
}

/* Elementary right-handed rotations about axis 0 (x), 1 (y), 2 (z). */
static void elementary(int axis, double ang, double M[9]) {
  double c = cos(ang), s = sin(ang);
  for (int i = 0; i < 9; i++) M[i] = 0;
  int u = (axis + 1) % 3, v = (axis + 2) % 3;
  M[axis * 3 + axis] = 1;
  M[u * 3 + u] = c;
  M[u * 3 + v] = -s;
  M[v * 3 + u] = s;
  M[v * 3 + v] = c;
}

static void matmul3(const double A[9], const double B[9], double C[9]) {
  double T[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      T[i * 3 + j] = A[i * 3 + 0] * B[0 * 3 + j] + A[i * 3 + 1] * B[1 * 3 + j] +
                     A[i * 3 + 2] * B[2 * 3 + j];
  memcpy(C, T, sizeof(T));
}

void or_rot_matrix(const double euler[3], int convention, double R[9]) {
  /* axis sequences of the 12 conventions, numbered as typedefs.h:59-70 */
  static const int seq[12][3] = {{0, 2, 0}, {0, 1, 0}, {1, 0, 1}, {1, 2, 1},
                                 {2, 1, 2}, {2, 0, 2}, {0, 2, 1}, {0, 1, 2},
                                 {1, 0, 2}, {1, 2, 0}, {2, 1, 0}, {2, 0, 1}};
  double A[9], B[9], C[9];
  elementary(seq[convention][0], euler[0], A);
  elementary(seq[convention][1], euler[1], B);
  elementary(seq[convention][2], euler[2], C);
  matmul3(A, B, R);
  matmul3(R, C, R);
}

void or_amp2d(const double *xyz, const unsigned *el_count, unsigned n_el,
              const double *qvec, size_t nqp, const double *rot, unsigned nrot,
              const double *ff, double *I, int nthreads) {
  size_t ntot = 0;
  for (unsigned e = 0; e < n_el; e++) ntot += el_count[e];
  const double *x = xyz, *y = xyz + ntot, *z = xyz + 2 * ntot;
  int nt = clamp_threads(nthreads);
  for (size_t k = 0; k < nqp; k++) I[k] = 0;
  for (unsigned r = 0; r < nrot; r++) {
    const double *R = rot + (size_t)r * 9;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long long k = 0; k < (long long)nqp; k++) {
      const double *qv = qvec + (size_t)k * 3;
      /* q' = R^T q (CalcFunctions.cpp:1020-1022, 1033) */
      double qx = qv[0] * R[0] + qv[1] * R[3] + qv[2] * R[6];
      double qy = qv[0] * R[1] + qv[1] * R[4] + qv[2] * R[7];
      double qz = qv[0] * R[2] + qv[1] * R[5] + qv[2] * R[8];
      double Ar = 0, Ai = 0;
      size_t o = 0;
      for (unsigned e = 0; e < n_el; o += el_count[e], e++) {
        double ar = 0, ai = 0;
        for (size_t j = o; j < o + el_count[e]; j++) {
          double ph = qx * x[j] + qy * y[j] + qz * z[j];
          ar += cos(ph);
          ai += sin(ph);
        }
        double w = ff[(size_t)e * nqp + k];
        Ar += w * ar;
        Ai += w * ai;
      }
      I[k] += Ar * Ar + Ai * Ai;
    }
  }
}

void or_calc_int2d(const double *xyz, const unsigned *el_count, unsigned n_el,
                   const double *q, unsigned nq, unsigned nfi,
                   const double euler_min[3], const double euler_max[3],
                   const unsigned euler_n[3], int convention, double lambda,
                   int polar, const double *ff, const double *sl, double *I2D,
                   double *I1D, int nthreads) {
  unsigned ntot = 0;
  for (unsigned e = 0; e < n_el; e++) ntot += el_count[e];
  const size_t nqp = (size_t)nq * nfi;
  double *qvec = (double *)malloc(nqp * 3 * sizeof(double));
  double *w = (double *)malloc((size_t)n_el * nqp * sizeof(double));
  const double dfi = 2. * XS_PI / nfi;
  for (unsigned iq = 0; iq < nq; iq++) {
    double st = q[iq] * (lambda * 0.25 / XS_PI);
    double ct = 1. - st * st; /* CalcFunctions.cpp:1001 */
    for (unsigned ifi = 0; ifi < nfi; ifi++) {
      size_t k = (size_t)iq * nfi + ifi;
      qvec[3 * k + 0] = ct * cos(ifi * dfi) * q[iq];
      qvec[3 * k + 1] = ct * sin(ifi * dfi) * q[iq];
      qvec[3 * k + 2] = -st * q[iq];
      for (unsigned e = 0; e < n_el; e++)
        w[(size_t)e * nqp + k] = ff ? ff[(size_t)e * nq + iq] : sl[e];
    }
  }
  double step[3];
  for (int a = 0; a < 3; a++)
    step[a] = euler_n[a] < 2 ? 0 : (euler_max[a] - euler_min[a]) / euler_n[a];
  unsigned nrot = euler_n[0] * euler_n[1] * euler_n[2];
  double *rot = (double *)malloc((size_t)(nrot ? nrot : 1) * 9 * sizeof(double));
  unsigned r = 0;
  for (unsigned ia = 0; ia < euler_n[0]; ia++)
    for (unsigned ib = 0; ib < euler_n[1]; ib++)
      for (unsigned ig = 0; ig < euler_n[2]; ig++, r++) {
        double e3[3] = {euler_min[0] + (ia + 0.5) * step[0],
                        euler_min[1] + (ib + 0.5) * step[1],
                        euler_min[2] + (ig + 0.5) * step[2]};
        or_rot_matrix(e3, convention, rot + (size_t)r * 9);
      }
  or_amp2d(xyz, el_count, n_el, qvec, nqp, rot, nrot, w, I2D, nthreads);
  double norm = 1. / (ntot * euler_n[0] * euler_n[1] * euler_n[2]);
  for (size_t k = 0; k < nqp; k++) I2D[k] *= norm;
  if (polar) {
    for (unsigned iq = 0; iq < nq; iq++) {
      double st = q[iq] * (lambda * 0.25 / XS_PI);
      double c2t = 1. - 2. * (st * st);
      double f = 0.5 * (1. + c2t * c2t);
      for (unsigned ifi = 0; ifi < nfi; ifi++) I2D[(size_t)iq * nfi + ifi] *= f;
    }
  }
  for (unsigned iq = 0; iq < nq; iq++) {
    double s = 0;
    for (unsigned ifi = 0; ifi < nfi; ifi++) s += I2D[(size_t)iq * nfi + ifi];
    I1D[iq] = s / nfi;
  }
  free(rot);
  free(w);
  free(qvec);
}

/* ---- Ensemble builder --------------------------------------------------- */

void or_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
  uint64_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint64_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    if (round) {
      k0 += 0x9E3779B97F4A7C15ULL;
      k1 += 0xBB67AE8584CAA73BULL;
    }
    const unsigned __int128 p0 = (unsigned __int128)0xD2E7470EE14C6C93ULL * c0;
    const unsigned __int128 p1 = (unsigned __int128)0xCA5A826395121157ULL * c2;
    const uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
    const uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double unit_draw(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }

static void draws(uint64_t a, uint64_t b, uint64_t c, const uint64_t key[2], uint64_t out[4]) {
  const uint64_t ctr[4] = {a, b, c, 0};
  or_philox4x64(ctr, key, out);
}

/* |N(0, sd)| from two uniforms (Box-Muller). */
static double half_normal(double sd, uint64_t ur, uint64_t ua) {
  const double z = sqrt(-2.0 * log(1.0 - unit_draw(ur))) * cos(2.0 * XS_PI * unit_draw(ua));
  const double v = sd * z;
  return v < 0 ? -v : v;
}

/* (x, y, z) = R (sin t cos f, sin t sin f, cos t), CalcFunctions.cpp:334, 375-378 */
static void spherical(double R, double fi, double theta, double v[3]) {
  v[0] = R * sin(theta) * cos(fi);
  v[1] = R * sin(theta) * sin(fi);
  v[2] = R * cos(theta);
}

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

typedef struct {
  double occ;
  unsigned atom, symm, el;
} or_site_entry;

long long or_calc_atoms(const or_block *blocks, unsigned nblocks, uint64_t seed, long long cap,
                        unsigned *el_out, unsigned *block_out, double *xyz_out) {
  long long n = 0;
  for (unsigned b = 0; b < nblocks; ++b) {
    const or_block *B = &blocks[b];
    const uint64_t key[2] = {seed, b};
    const unsigned M = B->natom * B->nsymm;
    const double *e = B->cell;
    /* sites rAtomCell (CalcFunctions.cpp:268-274) */
    double *rc = (double *)malloc(sizeof(double) * 3 * (M ? M : 1));
    for (unsigned a = 0; a < B->natom; ++a)
      for (unsigned s = 0; s < B->nsymm; ++s) {
        const double *P = B->p_symm + 9 * s, *r = B->r_atom + 3 * a;
        double *o = rc + 3 * ((size_t)a * B->nsymm + s);
        for (int k = 0; k < 3; ++k) o[k] = dot3(P + 3 * k, r) + B->r_symm[3 * s + k];
      }
    /* occupancy lists occNum (276-319): site i, then every later site j
     * within 0.5 A under a translation of up to +-2 cells along each axis
     * with more than one cell, occupancy summed when the element changes */
    size_t *off = (size_t *)calloc((size_t)M + 1, sizeof(size_t));
    size_t cap_e = (size_t)M + 16, ne = 0;
    or_site_entry *ent = (or_site_entry *)malloc(sizeof(or_site_entry) * cap_e);
    int lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = B->ncell[k] == 1 ? 0 : -2;
      hi[k] = B->ncell[k] == 1 ? 1 : 3;
    }
    for (unsigned i = 0; i < M; ++i) {
      off[i] = ne;
      const unsigned ai = i / B->nsymm;
      if (ne == cap_e) { cap_e *= 2; ent = (or_site_entry *)realloc(ent, sizeof(or_site_entry) * cap_e); }
      ent[ne++] = (or_site_entry){B->occ[ai], ai, i % B->nsymm, B->el[ai]};
      for (unsigned j = i + 1; j < M; ++j) {
        const unsigned aj = j / B->nsymm;
        for (int X = lo[0]; X < hi[0]; ++X)
          for (int Y = lo[1]; Y < hi[1]; ++Y)
            for (int Z = lo[2]; Z < hi[2]; ++Z) {
              double d[3];
              for (int k = 0; k < 3; ++k) {
                const double ct = e[k] * (double)X + e[3 + k] * (double)Y + e[6 + k] * (double)Z;
                d[k] = rc[3 * i + k] - rc[3 * j + k] + ct;
              }
              if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < 0.25) {
                const or_site_entry last = ent[ne - 1];
                if (ne == cap_e) { cap_e *= 2; ent = (or_site_entry *)realloc(ent, sizeof(or_site_entry) * cap_e); }
                const double occ = B->el[aj] != last.el ? last.occ + B->occ[aj] : last.occ;
                ent[ne++] = (or_site_entry){occ, aj, j % B->nsymm, B->el[aj]};
              }
            }
      }
    }
    off[M] = ne;
    /* copies' centre for RcutCopies (263-266) */
    double rmc[3] = {0, 0, 0};
    if (B->rcut_copies > 0) {
      for (unsigned c = 0; c < B->ncopy; ++c)
        for (int k = 0; k < 3; ++k) rmc[k] += B->r_copy[3 * c + k];
      for (int k = 0; k < 3; ++k) rmc[k] = rmc[k] / (double)B->ncopy;
    }
    /* covering sites: those not listed by an earlier covering site (342-344),
     * the same in every cell */
    unsigned char *covered = (unsigned char *)calloc(M ? M : 1, 1);
    unsigned R = 0;
    for (unsigned i = 0; i < M; ++i)
      if (!covered[i]) {
        ++R;
        for (size_t k = off[i]; k < off[i + 1]; ++k) covered[(size_t)ent[k].atom * B->nsymm + ent[k].symm] = 1;
      }
    const double rcut2 = B->rcut * B->rcut, rcc2 = B->rcut_copies * B->rcut_copies;
    uint64_t C = 0;
    for (unsigned c = 0; c < B->ncopy; ++c)
      for (unsigned X = 0; X < B->ncell[0]; ++X)
        for (unsigned Y = 0; Y < B->ncell[1]; ++Y)
          for (unsigned Z = 0; Z < B->ncell[2]; ++Z, ++C) {
            double ct[3];
            for (int k = 0; k < 3; ++k) ct[k] = e[k] * (double)X + e[3 + k] * (double)Y + e[6 + k] * (double)Z;
            memset(covered, 0, M ? M : 1);
            unsigned r = 0;
            for (unsigned i = 0; i < M; ++i) {
              if (covered[i]) continue;
              const uint64_t L = C * R + r++;
              uint64_t u[4];
              draws(L, 0, 0, key, u);
              const double randval = unit_draw(u[0]);
              for (size_t k = off[i]; k < off[i + 1]; ++k) covered[(size_t)ent[k].atom * B->nsymm + ent[k].symm] = 1;
              for (size_t k = off[i]; k < off[i + 1]; ++k) {
                if (!(randval <= ent[k].occ)) continue;
                const unsigned ia = ent[k].atom, is = ent[k].symm;
                const double *P = B->p_symm + 9 * is, *ra = B->r_atom + 3 * ia;
                double ra0[3], ra1[3], md[3] = {0, 0, 0};
                for (int q = 0; q < 3; ++q) ra0[q] = dot3(P + 3 * q, ra);
                if (B->mol_rotation) { /* 331-339, 352 */
                  uint64_t w[4];
                  draws(C, 4, is, key, w);
                  const double al = 2.0 * XS_PI * unit_draw(w[0]), ga = XS_PI * unit_draw(w[1]),
                               be = 2.0 * XS_PI * unit_draw(w[2]);
                  const double ca = cos(al), sa = sin(al), cb = cos(be), sb = sin(be), cg = cos(ga), sg = sin(ga);
                  /* rows of Rz(al) Rx(be) Rz(ga) */
                  const double Rm[9] = {ca * cg - sa * cb * sg, -ca * sg - sa * cb * cg, sa * sb,
                                        sa * cg + ca * cb * sg, -sa * sg + ca * cb * cg, -ca * sb,
                                        sb * sg, sb * cg, cb};
                  for (int q = 0; q < 3; ++q) ra1[q] = dot3(ra0, Rm + 3 * q);
                } else {
                  for (int q = 0; q < 3; ++q) ra1[q] = ra0[q];
                }
                if (B->dev_mol >= 1.e-7) { /* 322-329 */
                  uint64_t w[4], v[4];
                  draws(C, 2, is, key, w);
                  draws(C, 3, is, key, v);
                  spherical(half_normal(B->dev_mol, w[2], v[0]), 2.0 * XS_PI * unit_draw(w[0]),
                            XS_PI * unit_draw(w[1]), md);
                }
                for (int q = 0; q < 3; ++q) ra1[q] += B->r_symm[3 * is + q] + ct[q] + md[q];
                double u0[3];
                for (int q = 0; q < 3; ++q) u0[q] = ra1[q] - B->r_mean[q];
                if (!(B->rcut > 0) || dot3(u0, u0) <= rcut2) { /* 356 */
                  for (int q = 0; q < 3; ++q) ra1[q] = B->r_copy[3 * c + q] + dot3(u0, B->cf + 9 * c + 3 * q);
                  if (B->dev[ia] >= 1.e-7) { /* 359-365 */
                    uint64_t v[4];
                    double rd[3];
                    draws(L, 1, 0, key, v);
                    spherical(half_normal(B->dev[ia], u[3], v[0]), 2.0 * XS_PI * unit_draw(u[1]),
                              XS_PI * unit_draw(u[2]), rd);
                    for (int q = 0; q < 3; ++q) ra1[q] += rd[q];
                  }
                  double w0[3];
                  for (int q = 0; q < 3; ++q) w0[q] = ra1[q] - rmc[q];
                  if (!(B->rcut_copies > 0) || dot3(w0, w0) <= rcc2) { /* 366-369 */
                    if (n >= cap) { n = -1; goto done; }
                    el_out[n] = ent[k].el;
                    block_out[n] = b;
                    for (int q = 0; q < 3; ++q) xyz_out[3 * n + q] = ra1[q];
                    ++n;
                  }
                }
                break;
              }
            }
          }
  done:
    free(covered);
    free(ent);
    free(off);
    free(rc);
    if (n < 0) return -1;
  }
  return n;
}

/* ---- RearrangementInt pair sums ----------------------------------------- */

void or_cutoff_pairsum(const double *xyz, size_t n, const unsigned *el, unsigned n_el,
                       const double *rmax2, const double *q, unsigned nq, double *S,
                       uint64_t *npairs) {
  const double *x = xyz, *y = xyz + n, *z = xyz + 2 * n;
  memset(S, 0, sizeof(double) * (size_t)n_el * n_el * nq);
  memset(npairs, 0, sizeof(uint64_t) * (size_t)n_el * n_el);
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) {
      const double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 < rmax2[i] && r2 > 0.000001) {
        const size_t t = (size_t)el[i] * n_el + el[j];
        const double r = sqrt(r2);
        ++npairs[t];
        for (unsigned k = 0; k < nq; ++k) {
          const double qr = r * q[k];
          S[t * nq + k] += sin(qr) / (qr + 0.00000001);
        }
      }
    }
}

void or_pairlist_pairsum(const double *r, const unsigned *type, size_t np, unsigned ntypes,
                         const double *q, unsigned nq, double *S) {
  memset(S, 0, sizeof(double) * (size_t)ntypes * nq);
  for (size_t p = 0; p < np; ++p)
    for (unsigned k = 0; k < nq; ++k) {
      const double qr = r[p] * q[k];
      S[(size_t)type[p] * nq + k] += sin(qr) / (qr + 0.00000001);
    }
}
