// Internal shared definitions of the OLCAO MI355X runtime (host + device).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace olcao {

constexpr double kSmallThresh = 1.0e-8;   // src/constants.f90 smallThresh

// Three-centre sweep kinds.
enum ThreeKind { K3_HAM = 0, K3_EP = 1, K3_NP = 2 };
// Two-centre operators.
enum TwoOp { OP_OL = 0, OP_KE = 1 };

// Device view of the system and the screened work lists.  Passed by value
// to every kernel; all pointers are device memory.
struct DevCtx {
    // atoms
    const double *pos;          // [nAtoms][3]
    const int *atomType;        // [nAtoms]
    const int *cumulVale;       // [nAtoms]
    const int *cumulCore;       // [nAtoms]
    int valeDim, coreDim;
    // atom types
    int maxA, maxS;
    const double *alpha;        // [nT][maxA] integral exponents
    const double *elemAlpha;    // [nT][maxA] screening exponents (per element)
    const double *basis;        // [nT][maxS][maxA]
    const int *lmIndex;         // [nT][maxS] 1..16
    const int *nVale, *nCore;   // [nT]
    // potential
    int nPotTypes, maxPA, nPotSites;
    const int *potNA;           // [nPT]
    const double *potAlpha;     // [nPT][maxPA]
    const double *potZ;         // [nPT]
    const double *potNuc;       // [nPT]
    const int *potCumul;        // [nPT] first global term index
    const double *potPos;       // [nPS][3]
    const int *potSiteType;     // [nPS]
    // lattice
    double vec[9], inv[9];      // a_i = vec[3i+x]; fractional f_i = inv[3i+.] . v
    int nCorr;                  // nearest-image candidate corrections
    double corr[26 * 3];
    int nCells;
    const double *cellDims;     // [nCells][3] sorted by |L|^2
    const double *cellSizes;    // [nCells]
    int nk;
    const double *kcart;        // [nk][3]
    const double *phase;        // [nCells][nk][2] exp(i k.L)
    double tau;
    // atom pairs i <= j with at least one surviving cell
    int nPairs;
    const int *pairI, *pairJ;
    const double *pairLv;       // [nPairs][3] nearest lattice vector
    const int *pairPc;          // [nPairs+1] pair-cell range
    // pair-cells
    int nPc;
    const int *pcPair, *pcCell;
    const double *pcB;          // [nPc][3] shifted atom-2 position
    const double *pcS2;         // [nPc] squared separation
    const int *pcPp;            // [nPc+1] primitive-pair range
    // primitive pairs
    long long nPp;
    const int *ppPc;
    const int *ppAB;            // a1 | a2 << 8 | la << 16 | lb << 20
    const long long *ppOff;     // block offset (doubles)
    // optional work counters [16 classes][2]: three-Gaussian overlap and
    // nuclear evaluations of the last three-centre sweep
    unsigned long long *counters;
};

inline __host__ __device__ int pp_a1(int ab) { return ab & 255; }
inline __host__ __device__ int pp_a2(int ab) { return (ab >> 8) & 255; }
inline __host__ __device__ int pp_la(int ab) { return (ab >> 16) & 15; }
inline __host__ __device__ int pp_lb(int ab) { return (ab >> 20) & 15; }

// Kernel launchers (kernels.hip).  All asynchronous on `s`.
struct ClassLists {
    const int *idx[16];         // device list of primitive pairs per class
    int count[16];
};

void launch_two_center(const DevCtx &c, const ClassLists &cl, int op,
                       double *blocks, hipStream_t s);
// heads: 32 zeroed ints, the work counters of the persistent kernels.
void launch_three_center(const DevCtx &c, const ClassLists &cl, int *heads, int kind,
                         int termType, int termIdx, const double *coef,
                         int accumulate, double *blocks, hipStream_t s);
void launch_contract(const DevCtx &c, const double *blocks, double *X,
                     hipStream_t s);
void launch_bloch_scatter(const DevCtx &c, const double *X, double *vv,
                          double *cv, double *cc, int gamma, hipStream_t s);
void launch_fill(double *p, long long n, double v, hipStream_t s);
void launch_assemble(long long n2nk, int n, const double *NP, const double *KE,
                     const double *const *V, const double *coef, int nTerms,
                     double *H, int cplx, hipStream_t s);
// classIdx[k]: device list of input indices of class k = 4 la + lb.
void launch_prim_batch(int op, int n, const double *a, const double *A,
                       const double *B, const double *C, const int *const *classIdx,
                       const int *classCount, double *sh, hipStream_t s);
void launch_boys_batch(int n, const double *T, double *F, hipStream_t s);
void launch_extract_evecs(const double *A, int n, int nk, int numStates,
                          int cplx, double *out, hipStream_t s);

}  // namespace olcao
