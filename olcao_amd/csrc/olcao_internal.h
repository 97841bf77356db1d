// Internal shared definitions of the OLCAO MI355X runtime (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>
#include <stdint.h>

namespace olcao {

constexpr double kSmallThresh = 1.0e-8;   // src/constants.f90 smallThresh
constexpr int kMaxPotTypes = 8;           // potential types staged in LDS
constexpr int kMaxPotAlphas = 24;         // potential exponents per type
constexpr int kMaxPotSites = 1024;        // sites staged in LDS (28 B each)

// Three-centre sweep kinds.
enum ThreeKind { K3_HAM = 0, K3_EP = 1, K3_NP = 2 };
// Two-centre operators.
// OP_DM + c, OP_MM + c: Cartesian component c of the dipole about
// DevCtx::dipC and of the gradient (the momentum matrix's blocks).
// OP_KO + p: the k-overlap block for DevCtx::kq, its real part (p = 0) or
// minus its imaginary part (p = 1).
enum TwoOp { OP_OL = 0, OP_KE = 1, OP_MV = 2, OP_DM = 3, OP_MM = 6, OP_KO = 9 };

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
    const long long *ppOff;     // [nPp + 1] block offsets (doubles), store end appended
    // optional work counters [16 classes][2]: three-Gaussian overlap and
    // nuclear evaluations of the last three-centre sweep
    unsigned long long *counters;
    int anyNuc;                 // some potential type has Z >= smallThresh
    double dipC[3];             // dipole centre of the OP_DM sweeps
    double kq[3];               // deltaK of the OP_KO sweeps
    double fE2;                 // 2 totalEnergy of the force sweeps
};

inline __host__ __device__ int pp_a1(int ab) { return ab & 255; }
inline __host__ __device__ int pp_a2(int ab) { return (ab >> 8) & 255; }
inline __host__ __device__ int pp_la(int ab) { return (ab >> 16) & 15; }
inline __host__ __device__ int pp_lb(int ab) { return (ab >> 20) & 15; }

// One primitive pair as the three-centre kernels read it (one load).
struct PPRec {
    double A[3], B[3];          // atom-1 position, shifted atom-2 position
    double a1, a2;              // integral exponents
    double ea1, ea2;            // screening exponents
    double s2;                  // |A - B|^2
    long long off;              // block offset (doubles)
};

// Image entry of the per-pair three-centre lists (64 bits): the potential
// image C = site + nx a1 + ny a2 + nz a3 and how many terms each sweep kind
// takes from it, i.e. the reference's range tests evaluated once per
// geometry:
//   site [0,12) | nx+512 [12,22) | ny+512 [22,32) | nz+512 [32,42) |
//   nHam [42,48) | nEp [48,54) | nucHam [54] | nucNp [55]
// nHam / nEp: leading potential terms of the site's type inside range with
// the fused-sweep / separate-sweep expression; nucHam / nucNp: the nucleus.
constexpr int kImgLatOff = 512;
inline __host__ __device__ unsigned long long img_pack(int site, int nx, int ny, int nz, int nHam,
                                                       int nEp, int nucHam, int nucNp)
{
    return (unsigned long long)site | ((unsigned long long)(nx + kImgLatOff) << 12) |
           ((unsigned long long)(ny + kImgLatOff) << 22) |
           ((unsigned long long)(nz + kImgLatOff) << 32) | ((unsigned long long)nHam << 42) |
           ((unsigned long long)nEp << 48) | ((unsigned long long)nucHam << 54) |
           ((unsigned long long)nucNp << 55);
}
inline __host__ __device__ int img_site(unsigned long long e) { return (int)(e & 4095); }
inline __host__ __device__ int img_n(unsigned long long e, int k)   // k = 0, 1, 2 -> nx, ny, nz
{
    return (int)((e >> (12 + 10 * k)) & 1023) - kImgLatOff;
}
inline __host__ __device__ int img_nham(unsigned long long e) { return (int)((e >> 42) & 63); }
inline __host__ __device__ int img_nep(unsigned long long e) { return (int)((e >> 48) & 63); }
inline __host__ __device__ int img_nucham(unsigned long long e) { return (int)((e >> 54) & 1); }
inline __host__ __device__ int img_nucnp(unsigned long long e) { return (int)((e >> 55) & 1); }

// The message olcao_last_error(NULL) returns to the calling thread.
void set_noctx_error(const std::string &msg);

// Kernel launchers (kernels.hip).  All asynchronous on `s`.
// The image list of a record is cut into chunks of about kChunkItems
// fused-sweep items; each chunk is one lane of the three-centre kernels, so
// that a diffuse pair with thousands of images does not hold a wave alone.
constexpr int kChunkItems = 192;
struct ClassLists {
    const int *idx[16];         // device list of primitive pairs per class
    const PPRec *rec[16];       // the class's pairs as packed records, heaviest first
    const long long *ioff[16];  // [count + 1] image-list offsets per record
    const unsigned long long *img[16];   // image entries
    double *W[16];              // Hermite sums [nherm(la+lb)][nChunk] per class
    double *Wd[16];             // the force sweep's, one order up [nherm(la+lb+1)][nChunk]
    int count[16];
    // chunks, heaviest first: record, entry range
    const int *chRec[16];
    const long long *chBeg[16], *chEnd[16];
    int nChunk[16];
    // per record: its chunks' positions, in entry order
    const int *rcOff[16];       // [count + 1]
    const int *rcList[16];      // [nChunk]
    // records in the order of their first chunk, so that the finish kernel's
    // lanes read neighbouring W columns
    const int *finOrder[16];    // [count]
    // the entries of each chunk that carry nuclear work (for either sweep
    // kind), compacted: chunk u's are imgN[nOff[u] .. nOff[u + 1])
    const unsigned long long *imgN[16];
    const long long *nOff[16];  // [nChunk + 1]
};

// The classes k = 4 la + lb whose bit k is set in mask (independent, as for
// launch_three_center below).
void launch_two_center(const DevCtx &c, const ClassLists &cl, int op, int accumulate,
                       double *blocks, unsigned mask, hipStream_t s);
// accumulate: bit 0 adds the blocks onto the store; with kAccAddKE set the
// classes whose blocks are staged in LDS (nsh(la) nsh(lb) <= 81: all but
// d-f, f-d and f-f) instead add the kinetic block of each primitive pair
// themselves and write the sum (the fused PSCF sweep's kinetic term,
// integrals.F90:3444-3490), the other classes adding onto the store as with
// bit 0.  kFusedKEClasses marks the classes k = 4 la + lb that do so.
constexpr int kAccAddKE = 2;
constexpr unsigned kFusedKEClasses = 0x37ffu;
// mixed: three-Gaussian exponentials by the mixed-precision evaluation.
// Launches the classes k = 4 la + lb whose bit k is set in mask; different
// classes touch disjoint memory (their own W slice, their own primitive-pair
// blocks, their own work counters), so they may run on different streams.
void launch_three_center(const DevCtx &c, const ClassLists &cl, int kind, int termType,
                         int termIdx, const double *coef, int accumulate, double *blocks,
                         int mixed, unsigned mask, hipStream_t s);
// The force integrals' two-centre part along comp (classes la, lb <= 2 in
// mask): minus the ket derivative of the kinetic block plus c.fE2 times that
// of the overlap block, written to the primitive-pair store.
void launch_two_center_fd(const DevCtx &c, const ClassLists &cl, int comp, double *blocks,
                          unsigned mask, hipStream_t s);
// The force sweep's three-centre part (classes la, lb <= 2 in mask): comp < 0
// the Hermite sums of coef (one spin's coefficients) and the nuclei, one
// order up, into cl.Wd; comp = 0..2 minus the ket derivative along comp
// added onto the primitive-pair store.
void launch_three_center_force(const DevCtx &c, const ClassLists &cl, const double *coef, int comp,
                               double *blocks, unsigned mask, hipStream_t s);
// Dynamic LDS of the three-centre kernels: site positions and types, then
// (the tabulated form, used when the table fits kTableLdsMax) the per-lane
// term table of 1 / (zeta + a3), per 64-lane wave.
inline size_t k3c_site_bytes(const DevCtx &c)
{
    return (size_t)(((3 * c.nPotSites + (c.nPotSites + 1) / 2 + 1) & ~1)) * sizeof(double);
}
inline size_t k3c_table_bytes(const DevCtx &c)
{
    return (size_t)c.nPotTypes * c.maxPA * 64 * sizeof(double);
}
constexpr size_t kTableLdsMax = 48 * 1024;

// The fused Hamiltonian sweep (K3_HAM) of many cells, one launch per class
// and kernel form over all of them.  Cell3c: one cell's lists and buffers
// for one class, as launch_three_center takes them from its context; the
// finish flags as launch_three_center derives them from `accumulate`.
struct Cell3c {
    int ctx;                    // the cell's entry in Multi3c::ctxs
    int nChunk, nRec;
    const PPRec *recs;
    const int *chRec;
    const long long *chBeg, *chEnd, *nOff;
    const unsigned long long *img, *imgN;
    const double *coef;         // the spin's potential coefficients
    double *W;
    const int *finOrder, *rcOff, *rcList;
    double *blocks;
    int finAcc, finKE;          // finish: add onto the store / add the kinetic block
};
// One wave of a batched launch: 64 chunks (or finish records) of one cell
// from `first` on.
struct Task3c {
    int cell, first;
};
// The launches of class k over a batch (device pointers).  e3c[v]: the
// three-Gaussian pass in kernel form v (0 tabulated + mixed-precision
// exponential, 1 tabulated, 2 untabulated); nuc: the nuclear pass of the
// cells with nuclear terms; fin: the finish.  LDS: the largest of the
// cells' in each launch.
struct Multi3c {
    int k;
    const DevCtx *ctxs;
    const Cell3c *cells;
    const Task3c *e3c[3];
    int nE3c[3];
    const Task3c *nuc;
    int nNuc;
    const Task3c *fin;
    int nFin;
    size_t ldsE3c[3], ldsNuc;
};
void launch_three_center_multi(const Multi3c &m, hipStream_t s);
// Image lists: pass 1 counts entries and fused-sweep items per record, pass 2
// writes the entries at off[i]; *overflow is set if a lattice index does not
// fit the entry.
void launch_image_count(const DevCtx &c, const PPRec *recs, int n, int *nEnt, int *nItems,
                        hipStream_t s);
void launch_image_fill(const DevCtx &c, const PPRec *recs, int n, const long long *off,
                       unsigned long long *ent, int *overflow, hipStream_t s);
// Nuclear entries of each chunk [beg[u], end[u]) of img: counted into
// cnt[u] (outN null), or copied to outN from nOff[u] on (outN set).
void launch_nuc_compact(const unsigned long long *img, const long long *beg, const long long *end,
                        int nc, int *cnt, const long long *nOff, unsigned long long *outN,
                        hipStream_t s);
// Contraction work: segments of at most kSegPP (the default; a context may
// take 1..64, a wavefront's lanes) primitive pairs, and the
// pair-cells that have several segments {pair-cell, first partial, segment
// count}.
constexpr int kSegPP = 32;
// One contraction segment with everything its wave needs before the first
// block load: the atom types and state counts of its pair and the store
// offset of its first block.
struct ContractSeg {
    int pc, p0, np, part;       // pair-cell, first primitive pair, count, partial or -1
    int ti, tj, nt1, nt2;       // atom types, states (valence + core) per side
    long long base;             // ppOff[p0]
};
struct ContractPlan {
    const ContractSeg *segs;
    int nSeg;
    const int4 *multi;
    int nMulti;
    double *partial;            // [parts][maxS][maxS]
};
void launch_contract(const DevCtx &c, const ContractPlan &cp, const double *blocks, double *X,
                     hipStream_t s);
// mode bit 0: multiply the Bloch sum by -i (the momentum matrix,
// applyPhaseFactors runCode 1, intgSaving.F90:178-187; -X at Gamma,
// :727-731); bit 1: the full save (saveCurrentPair saveFull = 1, the force
// matrices), see bloch_store.
void launch_bloch_scatter(const DevCtx &c, const double *X, double *vv,
                          double *cv, double *cc, int gamma, int mode, hipStream_t s);
// Bloch sum, origin shift and scatter on the matrix cores (kbloch.hip):
// FP64 MFMA, or (split) exact BF16 slice products with FP32 accumulation
// over kBlochSlices fixed-point slices per operand.  pslice: the phase
// slices [nCells][nk (1 at Gamma)][re, im][8] as BF16 bit patterns, from
// phase_slices_host.
constexpr int kBlochSlices = 6;
void launch_bloch_mfma(const DevCtx &c, const double *X, const unsigned short *pslice, double *vv,
                       double *cv, double *cc, int gamma, int mode, bool split, hipStream_t s);
void phase_slices_host(const std::vector<double> &phase, int nCells, int nk, int gamma,
                       std::vector<unsigned short> &out);
void launch_fill(double *p, long long n, double v, hipStream_t s);
void launch_add(double *y, const double *x, long long n, hipStream_t s);   // y += x
// Zeroes `words` 4-byte words by a kernel: clears counters and status
// arrays in sequences that may be captured in a graph, where memset nodes
// are not used.
void launch_zero_words(void *p, long long words, hipStream_t s);
void launch_assemble(long long n2nk, int n, const double *NP, const double *KE,
                     const double *MV, const double *const *V, const double *coef, int nTerms,
                     double *H, int cplx, hipStream_t s);
void launch_ortho_finish(double *M, int n, int nb, int cplx, hipStream_t s);
// classIdx[k]: device list of input indices of class k = 4 la + lb.
void launch_prim_batch(int op, int n, const double *a, const double *A,
                       const double *B, const double *C, const int *const *classIdx,
                       const int *classCount, double *sh, hipStream_t s);
void launch_boys_batch(int n, const double *T, double *F, hipStream_t s);
void launch_herm_fill(double *A, int n, int nb, int cplx, hipStream_t s);
void launch_identity(double *M, int n, int nb, int cplx, hipStream_t s);
// Batched small-matrix generalised eigensolve in LDS (ksolve.hip): A (H upper
// in, eigenvectors out), B (S upper in, U out), w eigenvalues, info per k.
bool small_hegv_fits(int n, bool cplx);
void launch_small_hegv(bool cplx, int n, int nb, double *A, double *B, double *w, int *info,
                       hipStream_t s);
void launch_extract_evecs(const double *A, int n, int nk, int numStates,
                          int cplx, double *out, hipStream_t s);
// Valence density: Y(:, j, k) = f(j, k) X(:, j, k) for j < ns.
void launch_scale_states(const double *X, const double *f, int n, int ns, int nk, int cplx,
                         double *Y, hipStream_t s);
// out[m] = Tr[M_m (R0 + sgn R1)] summed over the nk upper triangles
// (M: device array of nMat matrix pointers; R1 may be null); partial holds
// 512 doubles per matrix.
void launch_traces(const double *const *M, int nMat, const double *R0, const double *R1,
                   double sgn, int n, int nk, int cplx, double *partial, double *out,
                   hipStream_t s);

// Cooperative tridiagonal reduction with the matrix resident in LDS across a
// group of workgroups (ktrd.hip).  A: full Hermitian (lower triangle read);
// V: explicit reflectors, column i = v_i; d, e: [nb][n]; tau: [nb][n]
// (complex for cplx); info[k] = -2 when a group barrier timed out.
constexpr int kTrdBlock = 64;   // reflectors per back-transform block
struct TrdPlan {
    int G = 0;            // workgroups per matrix
    int nGroups = 0;      // matrices in flight
    size_t lds = 0;       // dynamic LDS bytes per workgroup
    size_t pubDoubles = 0;
    size_t counters = 0;  // unsigned barrier counters (zeroed before launch)
};
bool trd_plan(int n, int nb, bool cplx, TrdPlan &p);
hipError_t launch_trd(bool cplx, const TrdPlan &p, int n, int nb, const double *A, double *V,
                      double *d, double *e, double *tau, double *pub, unsigned *cnt, int *info,
                      hipStream_t s);
// U_b = V_b T_b per block of kTrdBlock reflectors (rows j0+1..n-1 of columns
// j0..j0+kTrdBlock-1 of U); Tw: work of trd_tfactor_work doubles (T, then
// up to kTrdGramSplit partial Gram matrices per block and matrix: as many as
// bring the Gram workgroups to about kTrdGramWg)
constexpr int kTrdGramSplit = 8;
constexpr int kTrdGramWg = 576;
size_t trd_tfactor_work(int n, int nb, bool cplx);
void launch_trd_tfactor(bool cplx, int n, int nb, const double *V, const double *tau, double *Tw,
                        double *U, hipStream_t s);
void launch_real_to_field(const double *Z, double *X, long long count, bool cplx, hipStream_t s);

// Blocked reduction with the matrix in global memory (ktrdb.hip), for the
// sizes the LDS-resident reduction does not take: panels of kPanelB columns
// (start kernel, cooperative panel kernel), each followed by the caller's
// update of the trailing matrix A -= [W V] [V W]^H (one GEMM, k = 2 pb).
// Same outputs as launch_trd; W: n x 3 kPanelB per matrix, the panel's W, V
// and W again in columns [0, pb), [pb, 2 pb), [2 pb, 3 pb); pub: pubLen
// doubles per matrix; cnt: 64 words per matrix; matrices in flight: nGroups.
// C = alpha sum_p P_p + beta C (P: parts of m x ncols, contiguous; C
// column-major with leading dimension ldc): the sum of a split-K GEMM.
void launch_splitk_sum(const double *P, int parts, int m, int ncols, bool cplx, double alpha,
                       double beta, double *C, int ldc, hipStream_t s);
constexpr int kPanelB = 32;
constexpr int kTrdbMaxReal = 7168;     // = kDcBigMaxN (asserted in kdc.hip)
constexpr int kTrdbMaxComplex = 4096;
struct TrdbPlan {
    int G = 0;            // workgroups per matrix (<= 256)
    int nGroups = 0;      // matrices in flight
    size_t lds = 0;       // dynamic LDS bytes per workgroup
    size_t pubLen = 0;    // publication doubles per matrix
};
bool trdb_plan(int n, int nb, bool cplx, TrdbPlan &p);
hipError_t launch_trdb_start(bool cplx, const TrdbPlan &pl, const double *A, double *V, double *W,
                             double *d, double *e, double *tau, double *pub, unsigned *cnt,
                             int *info, int n, int nMat, int j0, hipStream_t s);
hipError_t launch_trdb_panel(bool cplx, const TrdbPlan &pl, const double *A, double *V, double *W,
                             double *d, double *e, double *tau, double *pub, unsigned *cnt,
                             int *info, int n, int nMat, int j0, int pb, hipStream_t s);

// Blocked batched Cholesky, diagonal step (kchol.hip): factors the updated
// diagonal block [j0, j0 + jb) of B (upper, column-major) into U and writes
// its inverse into Dinv (rows j0.., columns 0..jb-1; n x n per matrix);
// info takes the first non-positive pivot column (1-based).
constexpr int kCholBlock = 64;
void launch_chol_diag(bool cplx, const double *B, double *U, double *Dinv, int n, int nb, int j0,
                      int jb, int *info, hipStream_t s);

// Batched divide and conquer eigensolver for real symmetric tridiagonal
// matrices (kdc.hip): D (diagonal in, eigenvalues out, ascending), E
// (couplings), Q (eigenvectors out, column-major n x n per matrix); Qt, W:
// n x n work per matrix, ord: n ints per matrix.
constexpr int kDcMaxN = 64 * 14;
// merges past kDcMaxN keep their state in global memory (sorted poles, z and
// deflation flags in 20 B of LDS per entry): the tridiagonal solver's limit
constexpr int kDcBigMaxN = 7168;
struct DcLevel {
    int nodeBase = 0, nNodes = 0, tileBase = 0, nTiles = 0, maxM = 0;
    int rootBase = 0, nRoot = 0, colBase = 0, nCol = 0;
};
// Blocks of at most kDcLeaf entries are the tree's leaves, solved directly
// (k_dc_leaf) instead of merged up from 1 x 1 pieces.
constexpr int kDcLeaf = 12;
struct DcPlan {
    int n = 0;
    std::vector<int2> leaves;      // (offset, size) of the directly solved blocks
    std::vector<int4> nodes;       // (offset, first child size, size, -) by height
    std::vector<int4> tiles;       // GEMM tiles: (node, row tile, column tile, -)
    std::vector<int2> rootTiles;   // (node within level, root chunk)
    std::vector<int2> colTiles;    // (node within level, column chunk)
    std::vector<DcLevel> levels;
};
// device copies of the plan's lists and the state of split merges
struct DcDev {
    const int4 *nodes, *tiles;
    const int2 *rootTiles, *colTiles;
    double *stD;      // dc_state_doubles
    int *stI;         // dc_state_ints
    int4 *cnt;        // nodes x nb
    double *rho;      // nodes x nb
    const int2 *leaves;
};
// false (nothing planned or launched) when n is outside 1..kDcBigMaxN
bool dc_plan(int n, DcPlan &p);
size_t dc_state_doubles(int n, int nb);
size_t dc_state_ints(int n, int nb);
bool launch_dc(const DcPlan &p, const DcDev &dv, int nb, double *D, const double *E, double *Q,
               double *Qt, double *W, int *ord, hipStream_t s);

}  // namespace olcao
