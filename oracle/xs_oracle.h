/*
 * xs_oracle.h -- CPU oracle for the XaNSoNS scattering calculators.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product library never
 * links or calls it.
 *
 * Each function restates, in plain C (FP64 unless stated), one piece of the
 * reference's CPU path (CalcFunctions.cpp of XaNSoNS) with the same
 * expressions, constants and loop order.  The reference file:line that a
 * function follows is given in its comment.  How each function is pinned
 * (golden vectors from the reference executable, closed forms, invariants)
 * is listed in the test files tests/test_oracle_*.py; a function without
 * such a pin says "parity unpinned" here.
 *
 * Atom layout used throughout: structure-of-arrays, xyz[0..n) = x,
 * xyz[n..2n) = y, xyz[2n..3n) = z, with atoms grouped element-major
 * (el_count[0] atoms of element 0 first, then element 1, ...), which is the
 * order CalcFunctions.cpp:159-196 (block::sortAtoms) produces.
 *
 * Pair types p = (a, b), a <= b, are ordered (0,0),(0,1),...,(0,n-1),(1,1),...
 * as in CalcFunctions.cpp:777-778.
 */
#ifndef XS_ORACLE_H
#define XS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Pair-type index of (a,b), a<=b, for n_el elements. */
unsigned or_pair_index(unsigned a, unsigned b, unsigned n_el);

/* Number of histogram bins from the bounding-box diagonal of all atoms.
 * Follows GetHistSize, CalcFunctions.cpp:224-243. */
unsigned or_hist_size(const double *xyz, size_t n, double bin);

/* Approximate atomic density: atoms inside D/3 of the centroid divided by
 * that sphere's volume, D = 2*max distance from the centroid.
 * Follows GetCenter/GetEnsembleSize/GetAtomicDensity,
 * CalcFunctions.cpp:197-223. */
double or_atomic_density(const double *xyz, size_t n);

/* Pair-distance histogram with double-precision binning,
 * bin index = (unsigned)(sqrt((dx*dx+dy*dy)+dz*dz)/bin).
 * Follows calcHist, CalcFunctions.cpp:757-809 (zig-zag row partition
 * 779-781, j>i for like elements 782).  hist has ntypes*nhist entries and
 * is overwritten.  Pairs whose index is >= nhist (the reference would write
 * past its array) are counted in *overflow instead.  Returns 0. */
int or_hist_f64(const double *xyz, const unsigned *el_count, unsigned n_el,
                double bin, unsigned nhist, uint64_t *hist, uint64_t *overflow,
                int nthreads);

/* Same histogram with the reference GPU kernel's binning expression
 * evaluated in IEEE single precision with no contraction:
 * (unsigned)(sqrtf((dx*dx+dy*dy)+dz*dz)/bin), all operands float.
 * Follows calcHistKernel, CalcFunctionsCUDA.cu:832-833 and
 * kernelsPDF.cl:217-218 (whose builds use fast-math; this is the
 * correctly rounded form of the same expression). */
int or_hist_f32(const float *xyz, const unsigned *el_count, unsigned n_el,
                float bin, unsigned nhist, uint64_t *hist, uint64_t *overflow,
                int nthreads);

/* Direct Debye pair sums S[p*nq+iq] = sum_{pairs in p} sin(q r)/(q r + 1e-8).
 * Follows calcIntDebye, CalcFunctions.cpp:561-591. */
void or_debye_pairsum(const double *xyz, const unsigned *el_count, unsigned n_el,
                      const double *q, unsigned nq, double *S, int nthreads);

/* Partial (block-pair) Debye pair sums.  Atoms of each element are further
 * grouped block-major: blk_count[el*nblocks+b] atoms of element el belong
 * to block b.  S[(p*nparts + part)*nq + iq], part(iB<=jB) =
 * nblocks*iB - iB*(iB+1)/2 + jB.
 * Follows calcIntPartialDebye, CalcFunctions.cpp:662-703. */
void or_partial_debye_pairsum(const double *xyz, const unsigned *el_count,
                              unsigned n_el, const unsigned *blk_count,
                              unsigned nblocks, const double *q, unsigned nq,
                              double *S, int nthreads);

/* Histogram pair sums S[p*nq+iq] = sum_k h[k] sin(q r_k)/(q r_k + 1e-8),
 * r_k = (k+0.5)*bin, empty bins skipped.
 * Follows calcInt1DHist, CalcFunctions.cpp:844-871. */
void or_hist_pairsum(const uint64_t *hist, unsigned ntypes, unsigned nhist,
                     double bin, const double *q, unsigned nq, double *S,
                     int nthreads);

/* Powder intensity from pair sums: I = [sum_p 2 f_a f_b S_p + sum_el N_el f_el^2]/Ntot,
 * then the optional polarisation factor.  ff is n_el*nq (x-ray, per-q form
 * factors) or NULL, in which case sl[n_el] (neutron scattering lengths) is used.
 * Follows CalcFunctions.cpp:592-626 (direct Debye) and 872-906 (histogram). */
void or_intensity_1d(const double *S, unsigned n_el, const unsigned *el_count,
                     const double *ff, const double *sl, const double *q,
                     unsigned nq, int polar, double lambda, double *I);

/* Partial (block-pair) intensity; I has (1+nparts)*nq entries, total first.
 * Follows CalcFunctions.cpp:704-751. */
void or_partial_intensity_1d(const double *S, unsigned n_el,
                             const unsigned *blk_count, unsigned nblocks,
                             const double *ff, const double *sl,
                             const double *q, unsigned nq, int polar,
                             double lambda, unsigned ntot, double *I);

/* Polarisation factor 0.5*(1+cos^2(2 theta)), sin(theta) = q*lambda/(4 pi).
 * Follows PolarFactor1D, CalcFunctions.cpp:460-468. */
void or_polar_1d(double *I, const double *q, unsigned nq, double lambda);

/* PDF of the requested type (0 RDF, 1 PDF, 2 rPDF) from the histogram.
 * pdf has (1+ntypes)*nhist entries: total first, then partials.
 * Follows calcPDFandDebye, CalcFunctions.cpp:919-957.  The CPU loop there
 * evaluates the bin radius from the running index i = p*nhist + k
 * (CalcFunctions.cpp:942, 950), so for pair types p >= 1 the PDF/rPDF radius
 * is offset by p*nhist*bin; the CUDA and OpenCL kernels use the local index k
 * (CalcFunctionsCUDA.cu:1057, 1066; kernelsPDF.cl:152, 171).
 * cpu_radius_offset = 1 reproduces the CPU path, 0 the GPU kernels. */
void or_pdf(const uint64_t *hist, unsigned n_el, const unsigned *el_count,
            unsigned nhist, double bin, int type, double rho0, const double *sl,
            int cpu_radius_offset, double *pdf);

/* Rotation matrix (row-major R[9]) for intrinsic Euler angles in one of the
 * 12 conventions (0 XZX, 1 XYX, 2 YXY, 3 YZY, 4 ZYZ, 5 ZXZ, 6 XZY, 7 XYZ,
 * 8 YXZ, 9 YZX, 10 ZYX, 11 ZXY), R = R_axis1(e0) R_axis2(e1) R_axis3(e2).
 * Same result as calcRotMatrix, CalcFunctions.cpp:46-115, built here as a
 * product of elementary rotations. */
void or_rot_matrix(const double euler[3], int convention, double R[9]);

/* Orientation-summed squared amplitude for an explicit q-vector list:
 * I[k] = sum_r | sum_el ff[el*nqp+k] * sum_{j in el} exp(i (R_r^T q_k) . x_j) |^2
 * with rot[r*9..] = R_r row-major.  Follows calcInt2D's inner loops,
 * CalcFunctions.cpp:1026-1062 (q rotated by R^T, 1033). */
void or_amp2d(const double *xyz, const unsigned *el_count, unsigned n_el,
              const double *qvec, size_t nqp, const double *rot, unsigned nrot,
              const double *ff, double *I, int nthreads);

/* The reference's 2D calculator end to end: polar (q, phi) mesh with the
 * c = 1 - s^2 quirk, Euler midpoint grid, normalisation 1/(Ntot*Nrot),
 * optional polarisation factor, phi average.  I2D is nq*nfi row-major,
 * I1D is nq.  ff is n_el*nq or NULL (then sl is used).
 * Follows calcInt2D, CalcFunctions.cpp:969-1091. */
void or_calc_int2d(const double *xyz, const unsigned *el_count, unsigned n_el,
                   const double *q, unsigned nq, unsigned nfi,
                   const double euler_min[3], const double euler_max[3],
                   const unsigned euler_n[3], int convention, double lambda,
                   int polar, const double *ff, const double *sl, double *I2D,
                   double *I1D, int nthreads);

/* ---- Ensemble builder --------------------------------------------------- */

/* Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11): ten rounds of the
 * 4x64 bijection keyed by key[2], out = philox(ctr).  The random draws of
 * or_calc_atoms (and of the GPU builder) are functions of counters, so a
 * seed reproduces the same ensemble whatever the launch order. */
void or_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);

/* One structural block as block::calcAtoms sees it after ReadConfig
 * (ReadXML.cpp:254-323) and block::sortAtoms (CalcFunctions.cpp:158-196).
 * Row-major 3-vectors; p_symm[s*9 + row*3 + col] (rows R1, R2, R3 of
 * SymmEqPos s), cf[c*9 + row*3 + col] (rotation rows of copy c), cell = a, b, c.
 * rcut > 0 / rcut_copies > 0 switch the two cut-offs on; dev and dev_mol are
 * standard deviations (sqrt of Uiso). */
typedef struct {
  unsigned natom, nsymm, ncopy;
  unsigned ncell[3];
  const unsigned *el;
  const double *r_atom, *occ, *dev, *p_symm, *r_symm, *cell, *r_copy, *cf;
  double r_mean[3];
  double rcut, rcut_copies, dev_mol;
  int mol_rotation;
} or_block;

/* block::calcAtoms (CalcFunctions.cpp:251-397) for each block in turn:
 * sites = symmetry images of the cell atoms (268-274); sites closer than
 * 0.5 A under translations of up to +-2 cells share a position, their
 * occupancies accumulated across elements (276-319); then for every copy and
 * cell, one occupancy draw per not-yet-covered site picks at most one atom
 * (325-392), with the molecule and atom displacements, molecule rotation and
 * both cut-offs of the reference.  The reference's clock-seeded
 * default_random_engine is replaced by Philox4x64-10 counters:
 *   key = {seed, block}; candidate L = (((copy*Nx + X)*Ny + Y)*Nz + Z)*R + r
 *   (r = ordinal of the covering site in the cell, R = their number),
 *   C = L / R;
 *   philox({L,0,0,0}): occupancy draw, phi, theta, radius uniform of the atom
 *   displacement; philox({L,1,0,0})[0]: its angle uniform;
 *   philox({C,2,s,0}): phi, theta, radius uniform of molecule s's
 *   displacement; philox({C,3,s,0})[0]: its angle uniform;
 *   philox({C,4,s,0}): molecule s's Euler angles.
 * u = (x >> 11) 2^-53; |N(0, sd)| = sd |sqrt(-2 ln(1 - u)) cos(2 pi v)|.
 * Writes atom k's element, block and position in placement order; returns
 * the number of atoms, or -1 if more than cap. */
long long or_calc_atoms(const or_block *blocks, unsigned nblocks, uint64_t seed, long long cap,
                        unsigned *el_out, unsigned *block_out, double *xyz_out);

/* ---- RearrangementInt pair sums ----------------------------------------- */

/* S[(el[i]*n_el + el[j])*nq + iq] = sum over ordered pairs (i, j) with
 * 1e-6 < r_ij^2 < rmax2[i] of sin(q r_ij)/(q r_ij + 1e-8), r_ij^2 =
 * (dx^2 + dy^2) + dz^2; npairs[el[i]*n_el + el[j]] counts them.  xyz is SoA
 * (n atoms).  i and j ascend, so for element-major atoms each sum is
 * accumulated in the order of RearrangementInt's Itemp
 * (CalcFunctions.cpp:485-499). */
void or_cutoff_pairsum(const double *xyz, size_t n, const unsigned *el, unsigned n_el,
                       const double *rmax2, const double *q, unsigned nq, double *S,
                       uint64_t *npairs);

/* S[type[p]*nq + iq] = sum_p sin(q r_p)/(q r_p + 1e-8), p ascending: the
 * neighbour terms of RearrangementInt (CalcFunctions.cpp:506-529). */
void or_pairlist_pairsum(const double *r, const unsigned *type, size_t np, unsigned ntypes,
                         const double *q, unsigned nq, double *S);

#ifdef __cplusplus
}
#endif
#endif
