/* CPU oracle for the OLCAO integral -> H(k)/S(k) path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (UMKC-CPG/olcao, src/uolcao) used to check the HIP product
 * path and to time the reference algorithm on host cores.  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it; the
 * product (olcao_amd/) never does.
 *
 * Conventions mirror the reference's Fortran arrays:
 *   - matrices are column-major; valeVale(n,n,nk) complex is stored as
 *     interleaved (re,im) doubles, element (r,c,k) at 2*(r + c*n + k*n*n);
 *   - sh(16,16) blocks are column-major: sh(i,j) at i + 16*j (0-based);
 *   - a lattice vector a_i is vec[3*i + xyz].
 */
#ifndef OLCAO_ORACLE_H
#define OLCAO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_OL = 0, ORC_KE = 1, ORC_NP = 2, ORC_EP = 3, ORC_HAM = 4, ORC_MV = 5, ORC_DM = 6, ORC_MM = 7,
       ORC_KO = 8, ORC_FORCE = 9 };
enum { ORC_OK = 0, ORC_ERR_CELLS = 1, ORC_ERR_INPUT = 2 };

typedef struct {
    /* atoms */
    int nAtoms;
    const double *pos;        /* [nAtoms][3], bohr */
    const int *atomType;      /* [nAtoms], 0-based type index */
    /* atom types */
    int nTypes, maxAlphas, maxStates;
    const int *nAlpha;        /* [nTypes] */
    const double *alpha;      /* [nTypes][maxAlphas], ascending */
    const int *lmAlphaIndex;  /* [nTypes][maxAlphas], each in {1,4,9,16} */
    const int *nVale;         /* [nTypes] m-expanded valence states */
    const int *nCore;         /* [nTypes] m-expanded core states */
    const int *lmIndex;       /* [nTypes][maxStates], 1..16; valence then core */
    const double *basis;      /* [nTypes][maxStates][maxAlphas], incl. angNorm */
    const int *element;       /* [nTypes] element id used by screening tables */
    /* potential */
    int nPotTypes, maxPotAlphas;
    const int *potNAlpha;     /* [nPotTypes] */
    const double *potAlpha;   /* [nPotTypes][maxPotAlphas], ascending */
    const double *potZ;       /* [nPotTypes] nuclear charge */
    const double *potNucAlpha;/* [nPotTypes] */
    int nPotSites;
    const double *potPos;     /* [nPotSites][3] */
    const int *potSiteType;   /* [nPotSites] 0-based pot type */
    /* lattice */
    double vec[9];            /* a_i = vec[3*i + xyz], bohr */
    int nCells;
    const double *cellDims;   /* [nCells][3], sorted by |L|^2 */
    const double *cellSizes;  /* [nCells] |L|^2 */
    /* k-points, Cartesian bohr^-1 */
    int nk;
    const double *kpt;        /* [nk][3] */
    double logThresh;         /* tau = -ln(cutoff) */
    int gamma;                /* 1: real Gamma-only build (no phases) */
    int rel;                  /* 1: HAM subtracts the mass-velocity term */
} orc_sys;

/* Primitive integrals (gaussIntegrals.f90).  op: 0 overlap, 1 kinetic,
 * 2 three-centre overlap (electron3CIntg), 3 nuclear (nuclear3CIntg),
 * 4 mass velocity (massvel2CIntg). */
void orc_boys(double T, double F[7]);
void orc_prim(int op, double a1, double a2, double a3, const double *A,
              const double *B, const double *C, int l1l2switch, double *sh);
void orc_prim_batch(int op, int n, const double *a, const double *A,
                    const double *B, const double *C, const int *sw,
                    double *sh);

/* Lattice (lattice.f90). */
/* Three-component primitives: op 5 dipole about C (dipole3CIntg), op 6
 * gradient on the ket (momentum2CIntg); sh: [3][16][16] column-major per
 * component; a: [n][2], A/B/C: [n][3]. */
void orc_prim3(int op, double a1, double a2, const double *A, const double *B, const double *C,
               int sw, double *sh);
void orc_prim3_batch(int op, int n, const double *a, const double *A, const double *B,
                     const double *C, const int *sw, double *sh);
/* KOverlap2CIntg (gaussIntegrals.f90:129408): real and imaginary parts of the
 * (16,16) column-major block of exp(-i dK . r) between the two Gaussians. */
void orc_primk(double a1, double a2, const double *A, const double *B, const double *dK, int sw,
               double *shr, double *shi);
void orc_primk_batch(int n, const double *a, const double *A, const double *B, const double *dK,
                     const int *sw, double *shr, double *shi);
/* Force-derivative primitives (computeForceIntg): op 10 d/dB kinetic
 * (dkinetic2CIntg), 11 d/dB nuclear (dnuclear3CIntgCB), 12 d/dB
 * three-Gaussian overlap (delectron3CIntgCB); sh: [3][16][16] column-major
 * per component of B; a: [n][3].  Nonzero return: a nuclear block with
 * la + lb = 6 (not available). */
int orc_primd(int op, double a1, double a2, double a3, const double *A, const double *B,
              const double *C, int sw, double *sh);
int orc_primd_batch(int op, int n, const double *a, const double *A, const double *B,
                    const double *C, const int *sw, double *sh);
double orc_neglig_limit(const double *vec, double logThresh,
                        double minAtomicAlpha, double minPotAlpha);
int orc_make_lattice(const double *vec, double negligLimit, int maxCells,
                     double *cellDims, double *cellSizes);
void orc_nearest_lattice_vector(const double *vec, const double *v,
                                double *out);

/* Integral drivers (integrals.F90 + intgSaving.F90).  vv is valeVale
 * (complex, or real when s->gamma), cv coreVale, cc coreCore; all zeroed
 * by the call.  term: global potential term index for ORC_EP.  coef:
 * potential coefficients [potDim] of one spin for ORC_HAM.  ORC_DM, ORC_MM:
 * Cartesian component `term` (0..2) of the dipole about coef[0..2] and of
 * the momentum (-i times the gradient).  ORC_MV is the
 * negated mass-velocity sweep (gaussOverlapMV).  ORC_FORCE: component term
 * of computeForceIntg's matrices for one spin, coef = [that spin's
 * potential coefficients (potDim), totalEnergy]; complex builds save the
 * full matrix (saveCurrentPair saveFull = 1).  nthreads <= 0: default. */
int orc_integral(const orc_sys *s, int op, int term, const double *coef,
                 double *vv, double *cv, double *cc, int nthreads);

/* Restrict the next sweeps to atom pairs with (pairIndex - offset) %
 * stride == 0 (bounded CPU-baseline samples); stride 1 restores all. */
void orc_set_pair_sample(int stride, int offset);
/* test-only: scale the three-centre image range (1 = the reference's) */
void orc_set_three_center_range_scale(double f);

/* Work counters of the last orc_integral call (pair-cells, primitive
 * pairs, three-centre triples) for reporting. */
void orc_last_counts(long long *pairCells, long long *primPairs,
                     long long *triples);

#ifdef __cplusplus
}
#endif
#endif
