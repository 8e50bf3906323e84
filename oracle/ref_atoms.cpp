// Golden-vector and timing harness for the reference's ensemble builder.
//
// Test infrastructure, built by `make -C oracle refatoms` from the reference
// sources (ReadXML.cpp, IO.cpp, CalcFunctions.cpp, tinyxml2.cpp; main.cpp is
// replaced by this file).  It reads a start XML with the reference's own
// ReadConfig (ReadXML.cpp:180-323), prints every structural block as the
// builder sees it after parsing (positions made absolute by redefAtoms /
// redefSymm, sorted by sortAtoms, copies' rotation rows from calcRotMatrix,
// CalcFunctions.cpp:116-196), runs the reference's CalcAndPrintAtoms
// (CalcFunctions.cpp:398-450, which calls block::calcAtoms, 251-397) under a
// steady clock, and prints the atoms it placed.  Doubles are printed as C99
// hex floats, so the output is exact.
//
//   ref_atoms start.xml out.json [--no-atoms] [--rearrangement]
//
// With --rearrangement it also runs the reference's RearrangementInt
// (CalcFunctions.cpp:469-545) on I0[k] = 1 + k/64 (length (1 + Nparts) Nq
// with PartialIntensity, else Nq), timed, and adds "q", "sl", "ff",
// "rearr_seconds", "rearr_I" and per block "rcorr" and "neighbours"
// ([[el, x, y, z], ...] per cell atom, in the block's atom order).
//
// Output (one JSON object in out.json; ReadConfig's messages go to stdout): {"nel": .., "blocks": [..], "seconds": ..,
// "el_count": [..], "block_el_count": [[..]..], "atoms": [[x, y, z, ...] per
// element]}.
#include "typedefs.h"

#include <chrono>
#include <cstdio>
#include <cstring>

int myid = 0, numprocs = 1;

int ReadConfig(config *cfg, double **q, block **Block, const char *FileName, string elements[],
               map<string, unsigned int> *ID, vector<double> *SL, vector<double *> *FF);
void RearrangementInt(double *I, const double *q, const config *cfg, const block *Block,
                      vector<double *> FF, vector<double> SL, unsigned int Ntot);
unsigned int CalcAndPrintAtoms(config *cfg, block *Block, vector<vect3d<double> > **ra,
                               unsigned int **NatomEl, map<string, unsigned int> ID);

static void hex(FILE *f, double v) { fprintf(f, "\"%a\"", v); }

static void hex_vec(FILE *f, const vect3d<double> &v) {
  hex(f, v.x); fputc(',', f); hex(f, v.y); fputc(',', f); hex(f, v.z);
}

static void dump_block(FILE *f, const block &B) {
  fprintf(f, "{\"natom\":%u,\"nsymm\":%u,\"ncopy\":%u,\"ncell\":[%u,%u,%u],", B.Natom, B.Nsymm,
          B.Ncopy, B.Ncell.x, B.Ncell.y, B.Ncell.z);
  // element of every cell atom: id[] of its block-local type (sorted order)
  fprintf(f, "\"el\":[");
  unsigned a = 0;
  for (unsigned t = 0; t < B.Nid; ++t)
    for (unsigned k = 0; k < B.NatomEl[t]; ++k, ++a) fprintf(f, "%s%u", a ? "," : "", B.id[t]);
  fprintf(f, "],\"r_atom\":[");
  for (unsigned i = 0; i < B.Natom; ++i) { if (i) fputc(',', f); hex_vec(f, B.rAtom[i]); }
  fprintf(f, "],\"occ\":[");
  for (unsigned i = 0; i < B.Natom; ++i) { if (i) fputc(',', f); hex(f, B.occ[i]); }
  fprintf(f, "],\"dev\":[");
  for (unsigned i = 0; i < B.Natom; ++i) { if (i) fputc(',', f); hex(f, B.dev[i]); }
  fprintf(f, "],\"p_symm\":[");
  for (unsigned s = 0; s < B.Nsymm; ++s)
    for (int r = 0; r < 3; ++r) { if (s || r) fputc(',', f); hex_vec(f, B.pSymm[r][s]); }
  fprintf(f, "],\"r_symm\":[");
  for (unsigned s = 0; s < B.Nsymm; ++s) { if (s) fputc(',', f); hex_vec(f, B.rSymm[s]); }
  fprintf(f, "],\"cell\":[");
  for (int r = 0; r < 3; ++r) { if (r) fputc(',', f); hex_vec(f, B.e[r]); }
  fprintf(f, "],\"r_copy\":[");
  for (unsigned c = 0; c < B.Ncopy; ++c) { if (c) fputc(',', f); hex_vec(f, B.rCopy[c]); }
  fprintf(f, "],\"cf\":[");
  for (unsigned c = 0; c < B.Ncopy; ++c)
    for (int r = 0; r < 3; ++r) { if (c || r) fputc(',', f); hex_vec(f, B.cf[r][c]); }
  fprintf(f, "],\"r_mean\":[");
  hex_vec(f, B.rMean);
  fprintf(f, "],\"rcut\":");
  hex(f, B.cutoff ? B.Rcut : 0.0);
  fprintf(f, ",\"rcut_copies\":");
  hex(f, B.cutoffcopies ? B.RcutCopies : 0.0);
  fprintf(f, ",\"dev_mol\":");
  hex(f, B.dev_mol);
  fprintf(f, ",\"mol_rotation\":%d", B.ro_mol ? 1 : 0);
  if (B.rearrangement) {
    fprintf(f, ",\"rcorr\":");
    hex(f, B.Rcorr);
    fprintf(f, ",\"neighbours\":[");
    for (unsigned i = 0; i < B.Natom; ++i) {
      fprintf(f, "%s[", i ? "," : "");
      for (size_t k = 0; k < B.rAtomNeighb[i].size(); ++k) {
        fprintf(f, "%s[%u,", k ? "," : "", B.idNeighb[i][k]);
        hex_vec(f, B.rAtomNeighb[i][k]);
        fputc(']', f);
      }
      fputc(']', f);
    }
    fputc(']', f);
  }
  fputc('}', f);
}

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s start.xml out.json [--no-atoms]\n", argv[0]);
    return 2;
  }
  bool atoms = true, rearr = false;
  for (int a = 3; a < argc; ++a) {
    if (!strcmp(argv[a], "--no-atoms")) atoms = false;
    if (!strcmp(argv[a], "--rearrangement")) rearr = true;
  }
  config cfg;
  block *Block = NULL;
  double *q = NULL;
  string elements[128];
  map<string, unsigned int> ID;
  vector<double> SL;
  vector<double *> FF;
  if (ReadConfig(&cfg, &q, &Block, argv[1], elements, &ID, &SL, &FF)) return 1;
  // the blocks as calcAtoms receives them
  FILE *f = fopen(argv[2], "w");
  if (!f) {
    fprintf(stderr, "cannot write %s\n", argv[2]);
    return 1;
  }
  fprintf(f, "{\"nel\":%u,\"blocks\":[", cfg.Nel);
  for (unsigned b = 0; b < cfg.Nblocks; ++b) { if (b) fputc(',', f); dump_block(f, Block[b]); }
  fprintf(f, "],");
  fflush(f);
  vector<vect3d<double> > *ra = NULL;
  unsigned int *NatomEl = NULL;
  const auto t0 = std::chrono::steady_clock::now();
  CalcAndPrintAtoms(&cfg, Block, &ra, &NatomEl, ID);
  const double sec =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fprintf(f, "\"seconds\":%.6f,\"el_count\":[", sec);
  for (unsigned e = 0; e < cfg.Nel; ++e) fprintf(f, "%s%u", e ? "," : "", NatomEl[e]);
  fprintf(f, "],\"block_el_count\":[");
  for (unsigned b = 0; b < cfg.Nblocks; ++b) {
    fprintf(f, "%s[", b ? "," : "");
    for (unsigned e = 0; e < cfg.Nel; ++e) {
      unsigned n = 0;
      for (unsigned t = 0; t < Block[b].Nid; ++t)
        if (Block[b].id[t] == e) n = Block[b].NatomElAll[t];
      fprintf(f, "%s%u", e ? "," : "", n);
    }
    fprintf(f, "]");
  }
  fprintf(f, "],\"atoms\":[");
  for (unsigned e = 0; e < cfg.Nel; ++e) {
    fprintf(f, "%s[", e ? "," : "");
    if (atoms)
      for (size_t i = 0; i < ra[e].size(); ++i) { if (i) fputc(',', f); hex_vec(f, ra[e][i]); }
    fprintf(f, "]");
  }
  fprintf(f, "]");
  if (rearr) {
    const unsigned Nq = cfg.q.N;
    const unsigned nparts = cfg.Nblocks * (cfg.Nblocks + 1) / 2;
    const size_t len = cfg.calcPartialIntensity ? (size_t)(1 + nparts) * Nq : Nq;
    vector<double> I(len);
    for (size_t k = 0; k < len; ++k) I[k] = 1.0 + (double)k / 64.0;
    unsigned Ntot = 0;
    for (unsigned e = 0; e < cfg.Nel; ++e) Ntot += NatomEl[e];
    const auto t1 = std::chrono::steady_clock::now();
    RearrangementInt(I.data(), q, &cfg, Block, FF, SL, Ntot);
    const double rs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
    fprintf(f, ",\"source\":%u,\"partial\":%d,\"polar\":%d,\"lambda\":", cfg.source,
            cfg.calcPartialIntensity ? 1 : 0, cfg.PolarFactor ? 1 : 0);
    hex(f, cfg.lambda);
    fprintf(f, ",\"q\":[");
    for (unsigned k = 0; k < Nq; ++k) { if (k) fputc(',', f); hex(f, q[k]); }
    fprintf(f, "],\"sl\":[");
    for (size_t e = 0; e < SL.size(); ++e) { if (e) fputc(',', f); hex(f, SL[e]); }
    fprintf(f, "],\"ff\":[");
    for (size_t e = 0; e < FF.size(); ++e) {
      fprintf(f, "%s[", e ? "," : "");
      for (unsigned k = 0; k < Nq; ++k) { if (k) fputc(',', f); hex(f, FF[e][k]); }
      fputc(']', f);
    }
    fprintf(f, "],\"rearr_seconds\":%.6f,\"rearr_I\":[", rs);
    for (size_t k = 0; k < len; ++k) { if (k) fputc(',', f); hex(f, I[k]); }
    fputc(']', f);
  }
  fprintf(f, "}\n");
  fclose(f);
  return 0;
}
