// Host runtime of the OLCAO MI355X path: derived setup, screened work
// lists, device residency, kernel orchestration and the batched FP64
// generalised eigensolve.  Exposes the C-ABI declared in include/olcao_gpu.h.
//
// Derived setup restated from the reference:
//   basis / lmIndex / lmAlphaIndex   src/uolcao/basis.f90:55-243
//   reciprocal vectors               src/uolcao/lattice.f90:174-237
//   cell list (radius, border cells, sort by |L|^2)
//                                    src/uolcao/lattice.f90:310-443, 469-781
//   Cartesian k-points, phases       src/uolcao/kpoints.f90:424-489
//   screening radii                  src/uolcao/gaussRelations.f90:64-135
//   pair / cell / alpha-pair walk    src/uolcao/integrals.F90:323-570
// With ascending exponents the reference's alpha walk visits exactly the
// pairs with alphaDist(a1,a2) >= s^2, which is the list built here.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <thread>
#include <vector>

#include "../../include/olcao_gpu.h"
#include "olcao_internal.h"

using namespace olcao;

namespace {

constexpr double kPiHost = 3.141592653589793;

struct Fail {
    olcao_status st;
    std::string msg;
};

[[noreturn]] void fail(olcao_status st, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Fail{st, buf};
}

#define HIPCHK(x)                                                                     \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) fail(OLCAO_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

// Counts device allocations and frees of the library's buffers: a captured
// iteration graph holds device addresses and is replayed only while its
// context's count is the one it was captured at.  Inside an entry point the
// count is the context's (set by guard for the calling thread); elsewhere
// the process-wide one.
std::atomic<unsigned long long> g_allocEpoch{1};
thread_local std::atomic<unsigned long long> *tl_allocEpoch = nullptr;
inline void bump_alloc_epoch() { (tl_allocEpoch ? *tl_allocEpoch : g_allocEpoch).fetch_add(1); }

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    // one owner per allocation: copies would free it twice
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n)
    {
        o.p = nullptr;
        o.n = 0;
    }
    DBuf &operator=(DBuf &&) = delete;
    // at least count elements, kept when already as large (workspaces whose
    // size varies from call to call: a captured sequence allocates nothing
    // once an eager run has reached the largest)
    void reserve(size_t count)
    {
        if (p && count <= n) return;
        alloc(count);
    }
    void alloc(size_t count)
    {
        if (count == n && p) return;
        const hipError_t fe = release();
        if (fe != hipSuccess) fail(OLCAO_ERR_HIP, "hipFree: %s", hipGetErrorString(fe));
        if (count == 0) return;
        bump_alloc_epoch();
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            fail(OLCAO_ERR_ALLOC, "hipMalloc of %zu bytes failed", count * sizeof(T));
        }
        n = count;
    }
    void upload(const std::vector<T> &v, hipStream_t s)
    {
        alloc(v.size());
        if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T),
                                              hipMemcpyHostToDevice, s));
    }
    hipError_t release()
    {
        if (p) bump_alloc_epoch();
        const hipError_t e = p ? hipFree(p) : hipSuccess;
        p = nullptr;
        n = 0;
        return e;
    }
    ~DBuf()
    {
        // a destructor cannot report through olcao_status: a failed free
        // (which leaves a sticky device error the next call reports) is
        // written to stderr
        const hipError_t e = release();
        if (e != hipSuccess) std::fprintf(stderr, "olcao: hipFree failed: %s\n", hipGetErrorString(e));
    }
};

// Buffers of the tridiagonal eigensolver path, kept between calls.
struct EigWork {
    DBuf<double> V, U, W, Tw, tau, pub;    // reflectors, block factors, GEMM work
    DBuf<double> Z, Zt, Wd;                 // divide and conquer: vectors, work
    DBuf<double> Uc, Dinv, Y;               // Cholesky factor, inverted blocks, solve work
    DBuf<double> Ui;                        // U^-1 (upper triangular), when formed ahead
    DBuf<unsigned> cnt;
    DBuf<int> ord;
    DcPlan dc;
    DBuf<int4> dcNodes, dcTiles, dcCnt;
    DBuf<int2> dcRoot, dcCol, dcLeaf;
    DBuf<double> dcStD, dcRho;
    DBuf<int> dcStI;
    int lastPath = -1;     // path of the last gen_eig: 0 LDS solver, 1 tridiagonal, 2 library
    bool forceBlocked = false;   // the global-memory reduction also where the LDS one fits
    DBuf<double> Wp;       // blocked reduction: the panel's W (n x kPanelB per matrix)
    DBuf<double> Psk;      // split-K GEMM partial products
    // events around the tridiagonal reduction (set by a context; the
    // reduction is timed when they exist and the solve is not captured)
    hipEvent_t trdEv[2] = {nullptr, nullptr};
    bool trdTimed = false;
};

double angnorm(int lm)   // lm = 1..16, basis.f90:86-101
{
    const double pi = kPiHost;
    switch (lm) {
    case 1: return std::sqrt(1.0 / (4 * pi));
    case 2: case 3: case 4: return std::sqrt(3.0 / (4 * pi));
    case 5: case 6: case 7: return std::sqrt(15.0 / (4 * pi));
    case 8: return std::sqrt(15.0 / (16 * pi));
    case 9: return std::sqrt(5.0 / (16 * pi));
    case 10: return std::sqrt(105.0 / (4 * pi));
    case 11: return std::sqrt(105.0 / (16 * pi));
    case 12: case 13: return std::sqrt(35.0 / (32 * pi));
    case 14: return std::sqrt(7.0 / (16 * pi));
    default: return std::sqrt(21.0 / (32 * pi));
    }
}

int lclass(int lmAlphaIndex) { return lmAlphaIndex == 1 ? 0 : lmAlphaIndex == 4 ? 1 : lmAlphaIndex == 9 ? 2 : 3; }
int nshh(int l) { return (l + 1) * (l + 1); }

}  // namespace

struct olcao_ctx {
    std::string err;
    int precision = 0;
    // the Mode-B iteration as a captured graph: replayed while no buffer of
    // the library was (re)allocated since the capture
    bool useGraph = false, graphFailed = false, lastReplay = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    std::atomic<unsigned long long> allocEpoch{1};
    unsigned long long graphEpoch = 0;
    std::string graphErr;
    bool blochMfma = true;        // matrix-core Bloch sum (else the vector-ALU one)
    bool blochSplit = true;       // in the mixed form: split BF16 products (else FP64 MFMA)
    int chunkItems = kChunkItems; // three-centre items per image-list chunk
    int segPairs = kSegPP;        // primitive pairs per contraction segment
    std::vector<unsigned short> hPslice;   // phase slices of the split Bloch sum
    DBuf<unsigned short> dPslice;
    bool gamma = false;
    int rel = 0;                 // 1: scalar-relativistic (mass-velocity term)
    int spin = 1;                // 1 or 2 potential-coefficient sets
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    hipEvent_t ev[12] = {};
    // side streams: the three-centre classes fork onto them and join back
    static constexpr int kSide = 3;
    hipStream_t side[kSide] = {};
    hipEvent_t evFork = nullptr, evJoin[kSide] = {};
    // the solve's Cholesky of S, started on the last side stream beside the
    // Hamiltonian sweep (one Mode-B iteration): its rocBLAS handle (created
    // on first use: each handle holds device workspace, and contexts that
    // only take part in batched calls never need it), the events around it
    // and its status per k
    rocblas_handle blasSide = nullptr;
    hipEvent_t evS = nullptr, evChol = nullptr;
    DBuf<int> dCholInfo;
    bool preInv = false;   // that Cholesky also formed U^-1 (EigWork::Ui)
    // streams the per-class forks use: the context stream and the first
    // forkStreams - 1 side streams (the last one is left to that Cholesky
    // while it runs)
    int forkStreams = 1 + kSide;

    // host-side derived setup
    int nAtoms = 0, nT = 0, maxA = 0, maxS = 0, nPT = 0, maxPA = 0, nPS = 0;
    int valeDim = 0, coreDim = 0, potDim = 0, nk = 0;
    double tau = 0.0, vec[9] = {}, inv[9] = {}, recip[9] = {};
    std::vector<double> hPos, hAlpha, hElemAlpha, hBasis, hPotAlpha, hPotZ, hPotNuc, hPotPos;
    std::vector<int> hAtomType, hLmIndex, hLmAlpha, hNVale, hNCore, hCumV, hCumC, hPotNA,
        hPotCumul, hPotSiteType, hNAlpha;
    std::vector<double> hCellDims, hCellSizes, hKcart, hPhase, hCorr;
    // work lists
    std::vector<int> hPairI, hPairJ, hPairPc, hPcPair, hPcCell, hPcPp, hPpPc, hPpAB;
    std::vector<double> hPairLv, hPcB, hPcS2;
    std::vector<long long> hPpOff;
    long long blockDoubles = 0;
    std::vector<std::vector<int>> classIdx;
    // three-centre records per class: primitive pairs, heaviest first
    std::vector<std::vector<int>> recOrder;

    // device
    DBuf<double> dPos, dAlpha, dElemAlpha, dBasis, dPotAlpha, dPotZ, dPotNuc, dPotPos, dCellDims,
        dCellSizes, dKcart, dPhase, dPairLv, dPcB, dPcS2;
    DBuf<int> dAtomType, dLmIndex, dNVale, dNCore, dCumV, dCumC, dPotNA, dPotCumul, dPotSiteType,
        dPairI, dPairJ, dPairPc, dPcPair, dPcCell, dPcPp, dPpPc, dPpAB;
    DBuf<long long> dPpOff;
    DBuf<int> dClass[16];
    DBuf<PPRec> dRec[16];
    DBuf<long long> dIoff[16];
    DBuf<int> dChRec[16], dRcOff[16], dRcList[16], dFinOrder[16];
    DBuf<long long> dChBeg[16], dChEnd[16];
    DBuf<unsigned long long> dImg[16], dImgN[16];
    DBuf<long long> dNOff[16];
    DBuf<double> dW, dWd;
    long long imageEntries = 0;
    double listMs = 0.0;
    DBuf<double> dBlocks, dX, dCoef;
    DBuf<double> S, Aw, Bw, CV_S, CC_S, Tmp;
    // per spin: eigenvalues, tridiagonal work, solver and reduction status
    DBuf<double> D[2], E[2];
    DBuf<int> dInfo[2], tInfo[2];
    // eigenvectors of spin 2 (spin 1's stay in Aw) and every spin's
    // eigenvalues [nk][n] from the last solve, for the valence density
    DBuf<double> Aw2;
    std::vector<double> hEv[2];
    bool vecsValid[2] = {false, false};
    // valence density work: rho per spin, scaled eigenvectors, occupancies,
    // trace partial sums, results and the term-matrix pointer list
    DBuf<double> Rho[2], Ysc, dOcc, dTrPart, dTrOut;
    DBuf<const double *> dTrPtr;
    // the batched call's grouped solves: several cells' matrices in one batch
    DBuf<double> gA, gB, gD, gE;
    DBuf<int> gInfo, gTInfo;
    // the batched call's three-centre launch arguments (DevCtx, Cell3c and
    // Task3c arrays of all cells), held by the batch's first cell
    DBuf<unsigned char> dBatch;
    DBuf<double> H[2], CV_H[2], CC_H[2];   // per spin
    DBuf<double> Kvv, Kcv, Kcc;            // the k-overlap's imaginary-part sweep
    DBuf<unsigned long long> dCounters;
    // Mode A term matrices (orthogonalised when coreDim > 0)
    DBuf<double> KEm, NPm, MVm;
    std::vector<DBuf<double>> Vm;
    DBuf<double *> dVptr;
    bool haveS = false, haveH = false, haveScf = false;
    double timings[8] = {};
    double trdMs = 0.0;           // tridiagonal reductions of the current iteration

    DevCtx dev{};
    ClassLists cls{};
    bool eigLibrary = false;     // always rocSOLVER, also where the LDS solver fits
    EigWork eig;
    ContractPlan cplan{};
    DBuf<ContractSeg> dSegs;
    DBuf<int4> dMulti;
    DBuf<double> dPartial;

    size_t matDoubles() const { return (size_t)(gamma ? 1 : 2) * valeDim * valeDim * (gamma ? 1 : nk); }
    size_t cvDoubles() const { return (size_t)(gamma ? 1 : 2) * coreDim * valeDim * (gamma ? 1 : nk); }
    size_t ccDoubles() const { return (size_t)(gamma ? 1 : 2) * coreDim * coreDim * (gamma ? 1 : nk); }
};

namespace {

void inverse3(const double *v, double *inv)
{
    const double *a = v, *b = v + 3, *c = v + 6;
    double bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
    double ca[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
    double ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    double vol = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
    for (int x = 0; x < 3; ++x) {
        inv[x] = bc[x] / vol;
        inv[3 + x] = ca[x] / vol;
        inv[6 + x] = ab[x] / vol;
    }
}

double sq3(const double *v) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }

void nearest_host(const olcao_ctx &c, const double *v, double *out)
{
    double base[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        double f = v[0] * c.inv[3 * i] + v[1] * c.inv[3 * i + 1] + v[2] * c.inv[3 * i + 2];
        double r = std::trunc(std::fabs(f) + 0.5);
        if (f < 0.0) r = -r;
        for (int x = 0; x < 3; ++x) base[x] += c.vec[3 * i + x] * r;
    }
    double rem[3] = {v[0] - base[0], v[1] - base[1], v[2] - base[2]};
    double best = sq3(rem);
    int bi = -1;
    int nc = (int)c.hCorr.size() / 3;
    for (int q = 0; q < nc; ++q) {
        double r[3] = {rem[0] - c.hCorr[3 * q], rem[1] - c.hCorr[3 * q + 1], rem[2] - c.hCorr[3 * q + 2]};
        double s = sq3(r);
        if (s < best * (1.0 - 1e-14)) {
            best = s;
            bi = q;
        }
    }
    for (int x = 0; x < 3; ++x) out[x] = base[x] + (bi >= 0 ? c.hCorr[3 * bi + x] : 0.0);
}

// ---------------------------------------------------------------------------
void setup_system(olcao_ctx &c, const olcao_system &s)
{
    if (s.nAtoms <= 0 || s.nTypes <= 0 || s.nPotTypes <= 0 || s.maxAlphas <= 0 ||
        s.maxAlphas > 255)
        fail(OLCAO_ERR_INPUT, "invalid counts (nAtoms=%d nTypes=%d nPotTypes=%d maxAlphas=%d)",
             s.nAtoms, s.nTypes, s.nPotTypes, s.maxAlphas);
    if (!(s.basisCutoff > 0.0 && s.basisCutoff < 1.0))
        fail(OLCAO_ERR_INPUT, "basisCutoff must lie in (0,1)");
    c.gamma = s.gamma != 0;
    if (s.rel != 0 && s.rel != 1) fail(OLCAO_ERR_INPUT, "rel must be 0 or 1");
    if (s.spin != 1 && s.spin != 2) fail(OLCAO_ERR_INPUT, "spin must be 1 or 2");
    c.rel = s.rel;
    c.spin = s.spin;
    c.nAtoms = s.nAtoms;
    c.nT = s.nTypes;
    c.maxA = s.maxAlphas;
    c.tau = -std::log(s.basisCutoff);   // setCutoffThresh, lattice.f90:293
    std::memcpy(c.vec, s.cellVectors, sizeof(c.vec));
    inverse3(c.vec, c.inv);
    // reciprocal vectors b_i = 2 pi inv_i
    for (int i = 0; i < 9; ++i) c.recip[i] = 2.0 * kPiHost * c.inv[i];
    c.hPos.assign(s.atomPos, s.atomPos + 3 * s.nAtoms);
    c.hAtomType.assign(s.atomType, s.atomType + s.nAtoms);
    for (int a : c.hAtomType)
        if (a < 0 || a >= s.nTypes) fail(OLCAO_ERR_INPUT, "atom type out of range");

    // --- basis (basis.f90:55-243) ---
    int maxS = 1;
    for (int t = 0; t < s.nTypes; ++t) {
        int ns = 0;
        for (int r = 0; r < s.typeNumValeRadial[t]; ++r) ns += 2 * s.typeValeL[t * s.maxRadial + r] + 1;
        for (int r = 0; r < s.typeNumCoreRadial[t]; ++r) ns += 2 * s.typeCoreL[t * s.maxRadial + r] + 1;
        maxS = std::max(maxS, ns);
    }
    c.maxS = maxS;
    const int mA = c.maxA;
    c.hAlpha.assign((size_t)s.nTypes * mA, 0.0);
    c.hElemAlpha.assign((size_t)s.nTypes * mA, 0.0);
    c.hLmAlpha.assign((size_t)s.nTypes * mA, 0);
    c.hLmIndex.assign((size_t)s.nTypes * maxS, 0);
    c.hBasis.assign((size_t)s.nTypes * maxS * mA, 0.0);
    c.hNVale.assign(s.nTypes, 0);
    c.hNCore.assign(s.nTypes, 0);
    c.hNAlpha.assign(s.typeNumAlphas, s.typeNumAlphas + s.nTypes);
    for (int t = 0; t < s.nTypes; ++t) {
        const int na = s.typeNumAlphas[t];
        if (na <= 0 || na > mA) fail(OLCAO_ERR_INPUT, "type %d: bad alpha count", t);
        const double *al = s.typeAlphas + (size_t)t * mA;
        for (int a = 0; a < na; ++a) {
            if (!(al[a] > 0.0) || (a > 0 && !(al[a] > al[a - 1])))
                fail(OLCAO_ERR_INPUT, "type %d: alphas must be positive and strictly ascending", t);
            c.hAlpha[t * mA + a] = al[a];
        }
        const int *no = s.typeNumOrbAlphas + 4 * t;
        for (int a = 0; a < na; ++a)
            c.hLmAlpha[t * mA + a] = a < no[3] ? 16 : a < no[2] ? 9 : a < no[1] ? 4 : 1;
        int st = 0;
        auto add = [&](int l, const double *coef) {
            if (l < 0 || l > 3) fail(OLCAO_ERR_INPUT, "type %d: l=%d unsupported", t, l);
            const int nal = no[l];
            if (nal > na) fail(OLCAO_ERR_INPUT, "type %d: numOrbAlphas exceeds alphas", t);
            for (int k = 1; k <= 2 * l + 1; ++k) {
                const int lm = l * l + k;
                c.hLmIndex[t * maxS + st] = lm;
                for (int a = 0; a < nal; ++a) c.hBasis[((size_t)t * maxS + st) * mA + a] = coef[a] * angnorm(lm);
                ++st;
            }
        };
        for (int r = 0; r < s.typeNumValeRadial[t]; ++r)
            add(s.typeValeL[t * s.maxRadial + r], s.typeValeCoef + ((size_t)t * s.maxRadial + r) * mA);
        c.hNVale[t] = st;
        for (int r = 0; r < s.typeNumCoreRadial[t]; ++r)
            add(s.typeCoreL[t * s.maxRadial + r], s.typeCoreCoef + ((size_t)t * s.maxRadial + r) * mA);
        c.hNCore[t] = st - c.hNVale[t];
    }
    // screening exponents: those of the first type of each element
    // (makeAlphaDist, gaussRelations.f90:89-99)
    for (int t = 0; t < s.nTypes; ++t) {
        int first = t;
        for (int u = 0; u < t; ++u)
            if (s.typeElement[u] == s.typeElement[t]) { first = u; break; }
        for (int a = 0; a < mA; ++a) c.hElemAlpha[t * mA + a] = c.hAlpha[first * mA + a];
        if (first != t && s.typeNumAlphas[first] < s.typeNumAlphas[t])
            fail(OLCAO_ERR_INPUT, "types of one element must share their alphas");
    }
    c.hCumV.assign(s.nAtoms, 0);
    c.hCumC.assign(s.nAtoms, 0);
    for (int a = 0; a < s.nAtoms; ++a) {
        c.hCumV[a] = c.valeDim;
        c.hCumC[a] = c.coreDim;
        c.valeDim += c.hNVale[c.hAtomType[a]];
        c.coreDim += c.hNCore[c.hAtomType[a]];
    }

    // --- potential ---
    if (s.nPotTypes > kMaxPotTypes || s.maxPotAlphas > kMaxPotAlphas)
        fail(OLCAO_ERR_UNSUPPORTED, "at most %d potential types of at most %d exponents are supported",
             kMaxPotTypes, kMaxPotAlphas);
    if (s.nPotSites > kMaxPotSites)
        fail(OLCAO_ERR_UNSUPPORTED, "at most %d potential sites are supported", kMaxPotSites);
    c.nPT = s.nPotTypes;
    c.maxPA = s.maxPotAlphas;
    c.nPS = s.nPotSites;
    c.hPotNA.assign(s.potNumAlphas, s.potNumAlphas + s.nPotTypes);
    c.hPotAlpha.assign(s.potAlphas, s.potAlphas + (size_t)s.nPotTypes * s.maxPotAlphas);
    c.hPotZ.assign(s.potZ, s.potZ + s.nPotTypes);
    c.hPotNuc.assign(s.potNucAlpha, s.potNucAlpha + s.nPotTypes);
    c.hPotCumul.assign(s.nPotTypes, 0);
    for (int t = 0; t < s.nPotTypes; ++t) {
        c.hPotCumul[t] = c.potDim;
        c.potDim += s.potNumAlphas[t];
        for (int a = 0; a < s.potNumAlphas[t]; ++a) {
            double v = s.potAlphas[(size_t)t * s.maxPotAlphas + a];
            if (!(v > 0.0) || (a > 0 && !(v > s.potAlphas[(size_t)t * s.maxPotAlphas + a - 1])))
                fail(OLCAO_ERR_INPUT, "pot type %d: alphas must be positive, ascending", t);
        }
        if (!(s.potNucAlpha[t] > 0.0)) fail(OLCAO_ERR_INPUT, "pot type %d: nuclear alpha", t);
    }
    c.hPotPos.assign(s.potPos, s.potPos + 3 * (size_t)s.nPotSites);
    c.hPotSiteType.assign(s.potSiteType, s.potSiteType + s.nPotSites);
    for (int i = 0; i < s.nPotSites; ++i) {
        int t = c.hPotSiteType[i];
        if (t < 0 || t >= s.nPotTypes) fail(OLCAO_ERR_INPUT, "pot site type out of range");
        if (i > 0 && t < c.hPotSiteType[i - 1]) fail(OLCAO_ERR_INPUT, "pot sites must be grouped by type");
    }

    // --- k-points (kpoints.f90:424-454) ---
    c.nk = c.gamma ? 1 : s.nk;
    if (!c.gamma && s.nk <= 0) fail(OLCAO_ERR_INPUT, "nk must be positive");
    c.hKcart.assign(3 * (size_t)c.nk, 0.0);
    if (!c.gamma)
        for (int k = 0; k < s.nk; ++k)
            for (int x = 0; x < 3; ++x)
                c.hKcart[3 * k + x] = s.kpoints[3 * k] * c.recip[x] + s.kpoints[3 * k + 1] * c.recip[3 + x] +
                                      s.kpoints[3 * k + 2] * c.recip[6 + x];

    // --- cell list (lattice.f90:310-443, 469-781) ---
    double minA = 1e300, minP = 1e300;
    for (int t = 0; t < s.nTypes; ++t)
        for (int a = 0; a < s.typeNumAlphas[t]; ++a) minA = std::min(minA, c.hAlpha[t * mA + a]);
    for (int t = 0; t < s.nPotTypes; ++t)
        for (int a = 0; a < s.potNumAlphas[t]; ++a)
            minP = std::min(minP, s.potAlphas[(size_t)t * s.maxPotAlphas + a]);
    double nl = c.tau * std::max(2.0 / minA, 1.0 / minP + 0.5 / minA);
    double maxSq = 0.0;
    for (int i = 0; i < 3; ++i) maxSq = std::max(maxSq, sq3(c.vec + 3 * i));
    nl = nl + 0.75 * maxSq + std::sqrt(3.0 * maxSq * nl);
    nl = std::max(nl, 2.0 * c.tau / minA);
    int reps[3];
    for (int i = 0; i < 3; ++i) reps[i] = 2 * (int)std::ceil(std::sqrt(nl) / std::sqrt(sq3(c.vec + 3 * i))) + 1;
    double diag[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i)
        for (int x = 0; x < 3; ++x) diag[x] += c.vec[3 * i + x];
    struct Cell { double size; int idx; double L[3]; };
    std::vector<Cell> cells;
    int gen = 0;
    for (int i = 1; i <= reps[0]; ++i)
        for (int j = 1; j <= reps[1]; ++j)
            for (int k = 1; k <= reps[2]; ++k) {
                int rep[3] = {i - (1 + reps[0]) / 2, j - (1 + reps[1]) / 2, k - (1 + reps[2]) / 2};
                double L[3] = {0, 0, 0};
                for (int l = 0; l < 3; ++l)
                    for (int m = 0; m < 3; ++m) L[m] += c.vec[3 * l + m] * rep[l];
                double ss = sq3(L);
                if (ss > nl) {
                    // seven box corners reached through the x, y, z parts of a1+a2+a3
                    double t[3] = {L[0], L[1], L[2]};
                    bool in = false;
                    const int seq[7][2] = {{0, 1}, {1, 1}, {0, -1}, {2, 1}, {0, 1}, {1, -1}, {0, -1}};
                    for (int q = 0; q < 7 && !in; ++q) {
                        t[seq[q][0]] += seq[q][1] * diag[seq[q][0]];
                        if (sq3(t) < nl) in = true;
                    }
                    if (!in) continue;
                }
                Cell cc{ss, gen++, {L[0], L[1], L[2]}};
                cells.push_back(cc);
            }
    std::stable_sort(cells.begin(), cells.end(), [](const Cell &a, const Cell &b) { return a.size < b.size; });
    c.hCellDims.resize(3 * cells.size());
    c.hCellSizes.resize(cells.size());
    for (size_t q = 0; q < cells.size(); ++q) {
        c.hCellSizes[q] = cells[q].size;
        for (int x = 0; x < 3; ++x) c.hCellDims[3 * q + x] = cells[q].L[x];
    }
    const int nCells = (int)cells.size();

    // --- phases exp(i k.L) (kpoints.f90:459-489) ---
    if (!c.gamma) {
        c.hPhase.resize(2 * (size_t)nCells * c.nk);
        for (int q = 0; q < nCells; ++q)
            for (int k = 0; k < c.nk; ++k) {
                const double *kp = &c.hKcart[3 * k];
                const double *L = &c.hCellDims[3 * q];
                double dot = kp[0] * L[0] + kp[1] * L[1] + kp[2] * L[2];
                c.hPhase[2 * ((size_t)q * c.nk + k)] = std::cos(dot);
                c.hPhase[2 * ((size_t)q * c.nk + k) + 1] = std::sin(dot);
            }
    } else {
        c.hPhase.assign(2, 0.0);
    }

    // --- nearest-image corrections: lattice vectors v (|n_i| <= 1) that can be
    // nearer than the rounded point for some residual in the cell:
    // sum_i |a_i . v| > |v|^2 ---
    c.hCorr.clear();
    for (int i = -1; i <= 1; ++i)
        for (int j = -1; j <= 1; ++j)
            for (int k = -1; k <= 1; ++k) {
                if (!i && !j && !k) continue;
                double v[3];
                for (int x = 0; x < 3; ++x) v[x] = i * c.vec[x] + j * c.vec[3 + x] + k * c.vec[6 + x];
                double s = 0.0;
                for (int a = 0; a < 3; ++a)
                    s += std::fabs(c.vec[3 * a] * v[0] + c.vec[3 * a + 1] * v[1] + c.vec[3 * a + 2] * v[2]);
                if (s > sq3(v) * (1.0 + 1e-12))
                    for (int x = 0; x < 3; ++x) c.hCorr.push_back(v[x]);
            }
}

// Screened work lists (integrals.F90:323-570).
void build_lists(olcao_ctx &c)
{
    const int mA = c.maxA;
    const int nCells = (int)c.hCellSizes.size();
    const double maxCell = c.hCellSizes.empty() ? 0.0 : c.hCellSizes.back();
    c.hPairI.clear(); c.hPairJ.clear(); c.hPairLv.clear(); c.hPairPc.assign(1, 0);
    c.hPcPair.clear(); c.hPcCell.clear(); c.hPcB.clear(); c.hPcS2.clear(); c.hPcPp.assign(1, 0);
    c.hPpPc.clear(); c.hPpAB.clear(); c.hPpOff.clear();
    c.classIdx.assign(16, {});
    long long off = 0;
    for (int i = 0; i < c.nAtoms; ++i) {
        const int ti = c.hAtomType[i];
        const double *A = &c.hPos[3 * i];
        for (int j = i; j < c.nAtoms; ++j) {
            const int tj = c.hAtomType[j];
            const double *Bj = &c.hPos[3 * j];
            double dv[3] = {A[0] - Bj[0], A[1] - Bj[1], A[2] - Bj[2]};
            double lv[3];
            nearest_host(c, dv, lv);
            double r0[3] = {A[0] - Bj[0] - lv[0], A[1] - Bj[1] - lv[1], A[2] - Bj[2] - lv[2]};
            const double sep = sq3(r0);
            const double *ea1 = &c.hElemAlpha[ti * mA], *ea2 = &c.hElemAlpha[tj * mA];
            const double lim = c.tau / ea2[0] / (ea1[0] / (ea1[0] + ea2[0]));
            if (sep > lim) continue;
            const double maxLR = sep + lim + 2.0 * std::sqrt(sep * lim);
            if (maxLR > maxCell)
                fail(OLCAO_ERR_CELLS, "more lattice points needed for atom pair (%d,%d): "
                     "maxLatticeRadius=%g available=%g", i, j, maxLR, maxCell);
            const int pairIdx = (int)c.hPairI.size();
            bool anyPc = false;
            for (int k = 0; k < nCells; ++k) {
                if (c.hCellSizes[k] > maxLR) break;
                const double *L = &c.hCellDims[3 * k];
                double B[3] = {Bj[0] + lv[0] + L[0], Bj[1] + lv[1] + L[1], Bj[2] + lv[2] + L[2]};
                double d[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
                const double s2 = sq3(d);
                if (s2 > lim) continue;
                const int pc = (int)c.hPcPair.size();
                c.hPcPair.push_back(pairIdx);
                c.hPcCell.push_back(k);
                c.hPcB.insert(c.hPcB.end(), B, B + 3);
                c.hPcS2.push_back(s2);
                const int na1 = c.hNAlpha[ti], na2 = c.hNAlpha[tj];
                for (int a1 = 0; a1 < na1; ++a1) {
                    if (c.tau / ea2[0] / (ea1[a1] / (ea1[a1] + ea2[0])) < s2) break;
                    for (int a2 = 0; a2 < na2; ++a2) {
                        const double ad = c.tau / ea2[a2] / (ea1[a1] / (ea1[a1] + ea2[a2]));
                        if (ad < s2) break;
                        const int la = lclass(c.hLmAlpha[ti * mA + a1]);
                        const int lb = lclass(c.hLmAlpha[tj * mA + a2]);
                        const int pp = (int)c.hPpPc.size();
                        c.hPpPc.push_back(pc);
                        c.hPpAB.push_back(a1 | (a2 << 8) | (la << 16) | (lb << 20));
                        c.hPpOff.push_back(off);
                        off += (long long)nshh(la) * nshh(lb);
                        c.classIdx[la * 4 + lb].push_back(pp);
                    }
                }
                c.hPcPp.push_back((int)c.hPpPc.size());
                anyPc = true;
            }
            if (anyPc) {
                c.hPairI.push_back(i);
                c.hPairJ.push_back(j);
                c.hPairLv.insert(c.hPairLv.end(), lv, lv + 3);
                c.hPairPc.push_back((int)c.hPcPair.size());
            }
        }
    }
    c.blockDoubles = off;
    // heaviest primitive pairs first: estimated cost ~ volume of the widest
    // potential screening sphere, (budget (1/zeta + 1/a3min))^(3/2)
    double a3min = 1e300;
    for (int t = 0; t < c.nPT; ++t)
        if (c.hPotNA[t] > 0) a3min = std::min(a3min, c.hPotAlpha[(size_t)t * c.maxPA]);
        else a3min = std::min(a3min, c.hPotNuc[t]);
    std::vector<double> cost(c.hPpPc.size());
    for (size_t pp = 0; pp < c.hPpPc.size(); ++pp) {
        const int pc = c.hPpPc[pp], ab = c.hPpAB[pp];
        const int pr = c.hPcPair[pc];
        const int ti = c.hAtomType[c.hPairI[pr]], tj = c.hAtomType[c.hPairJ[pr]];
        const double e1 = c.hElemAlpha[ti * mA + pp_a1(ab)], e2 = c.hElemAlpha[tj * mA + pp_a2(ab)];
        const double budget = std::max(0.0, c.tau * (1.0 - c.hPcS2[pc] / (c.tau / e2 / (e1 / (e1 + e2)))));
        const double z = c.hAlpha[ti * mA + pp_a1(ab)] + c.hAlpha[tj * mA + pp_a2(ab)];
        cost[pp] = std::pow(budget * (1.0 / z + 1.0 / a3min), 1.5);
    }
    for (auto &lst : c.classIdx)
        std::stable_sort(lst.begin(), lst.end(), [&](int x, int y) { return cost[x] > cost[y]; });
}

PPRec make_rec(const olcao_ctx &c, int pp)
{
    const int pc = c.hPpPc[pp], ab = c.hPpAB[pp];
    const int pr = c.hPcPair[pc];
    const int i = c.hPairI[pr], j = c.hPairJ[pr];
    const int ti = c.hAtomType[i], tj = c.hAtomType[j];
    PPRec r;
    for (int x = 0; x < 3; ++x) {
        r.A[x] = c.hPos[3 * i + x];
        r.B[x] = c.hPcB[3 * pc + x];
    }
    r.a1 = c.hAlpha[ti * c.maxA + pp_a1(ab)];
    r.a2 = c.hAlpha[tj * c.maxA + pp_a2(ab)];
    r.ea1 = c.hElemAlpha[ti * c.maxA + pp_a1(ab)];
    r.ea2 = c.hElemAlpha[tj * c.maxA + pp_a2(ab)];
    r.s2 = c.hPcS2[pc];
    r.off = c.hPpOff[pp];
    return r;
}

void upload(olcao_ctx &c)
{
    hipStream_t s = c.stream;
    c.dPos.upload(c.hPos, s); c.dAtomType.upload(c.hAtomType, s);
    c.dCumV.upload(c.hCumV, s); c.dCumC.upload(c.hCumC, s);
    c.dAlpha.upload(c.hAlpha, s); c.dElemAlpha.upload(c.hElemAlpha, s);
    c.dBasis.upload(c.hBasis, s); c.dLmIndex.upload(c.hLmIndex, s);
    c.dNVale.upload(c.hNVale, s); c.dNCore.upload(c.hNCore, s);
    c.dPotNA.upload(c.hPotNA, s); c.dPotAlpha.upload(c.hPotAlpha, s);
    c.dPotZ.upload(c.hPotZ, s); c.dPotNuc.upload(c.hPotNuc, s);
    c.dPotCumul.upload(c.hPotCumul, s); c.dPotPos.upload(c.hPotPos, s);
    c.dPotSiteType.upload(c.hPotSiteType, s);
    c.dCellDims.upload(c.hCellDims, s); c.dCellSizes.upload(c.hCellSizes, s);
    c.dKcart.upload(c.hKcart, s); c.dPhase.upload(c.hPhase, s);
    phase_slices_host(c.hPhase, (int)c.hCellSizes.size(), c.nk, c.gamma, c.hPslice);
    c.dPslice.upload(c.hPslice, s);
    c.dPairI.upload(c.hPairI, s); c.dPairJ.upload(c.hPairJ, s);
    c.dPairLv.upload(c.hPairLv, s); c.dPairPc.upload(c.hPairPc, s);
    c.dPcPair.upload(c.hPcPair, s); c.dPcCell.upload(c.hPcCell, s);
    c.dPcB.upload(c.hPcB, s); c.dPcS2.upload(c.hPcS2, s); c.dPcPp.upload(c.hPcPp, s);
    c.dPpPc.upload(c.hPpPc, s); c.dPpAB.upload(c.hPpAB, s);
    {   // block offsets with the store's end appended (ppOff[nPp] = total)
        std::vector<long long> off(c.hPpOff);
        off.push_back(c.blockDoubles);
        c.dPpOff.upload(off, s);
        HIPCHK(hipStreamSynchronize(s));
    }
    for (int k = 0; k < 16; ++k) {
        c.dClass[k].upload(c.classIdx[k], s);
        c.cls.idx[k] = c.dClass[k].p;
        c.cls.count[k] = (int)c.classIdx[k].size();
        std::vector<PPRec> recs(c.classIdx[k].size());
        for (size_t q = 0; q < recs.size(); ++q) recs[q] = make_rec(c, c.classIdx[k][q]);
        c.dRec[k].upload(recs, s);
        c.cls.rec[k] = c.dRec[k].p;
    }
    c.dBlocks.alloc(std::max<long long>(c.blockDoubles, 1));
    c.dX.alloc(std::max<size_t>(c.hPcPair.size() * (size_t)c.maxS * c.maxS, 1));
    {   // contraction segments (at most segPairs primitive pairs each)
        if (2 * (size_t)c.maxS * c.maxA * sizeof(double) > 64 * 1024)
            fail(OLCAO_ERR_INPUT, "basis too large for the contraction kernel (states x exponents)");
        std::vector<ContractSeg> segs;
        std::vector<int4> multi;
        std::vector<std::pair<int, int>> cut;
        int parts = 0;
        for (size_t pc = 0; pc < c.hPcPair.size(); ++pc) {
            const int p0 = c.hPcPp[pc], p1 = c.hPcPp[pc + 1];
            cut.clear();
            for (int q = p0; q < p1; q += c.segPairs) cut.emplace_back(q, std::min(c.segPairs, p1 - q));
            const int pr = c.hPcPair[pc];
            const int ti = c.hAtomType[c.hPairI[pr]], tj = c.hAtomType[c.hPairJ[pr]];
            auto seg = [&](int q0, int nq, int part) {
                ContractSeg g;
                g.pc = (int)pc; g.p0 = q0; g.np = nq; g.part = part;
                g.ti = ti; g.tj = tj;
                g.nt1 = c.hNVale[ti] + c.hNCore[ti]; g.nt2 = c.hNVale[tj] + c.hNCore[tj];
                g.base = q0 < (int)c.hPpOff.size() ? c.hPpOff[q0] : c.blockDoubles;
                segs.push_back(g);
            };
            if (cut.size() <= 1) {
                seg(p0, p1 - p0, -1);
                continue;
            }
            multi.push_back(make_int4((int)pc, parts, (int)cut.size(), 0));
            for (size_t q = 0; q < cut.size(); ++q) seg(cut[q].first, cut[q].second, parts + (int)q);
            parts += (int)cut.size();
        }
        c.dSegs.upload(segs, s);
        c.dMulti.upload(multi, s);
        c.dPartial.alloc(std::max<size_t>((size_t)parts * c.maxS * c.maxS, 1));
        HIPCHK(hipStreamSynchronize(s));
        c.cplan.segs = c.dSegs.p;
        c.cplan.nSeg = (int)segs.size();
        c.cplan.multi = c.dMulti.p;
        c.cplan.nMulti = (int)multi.size();
        c.cplan.partial = c.dPartial.p;
    }
    c.dCoef.alloc(std::max(c.potDim * c.spin, 1));
    c.dCounters.alloc(32);

    DevCtx &d = c.dev;
    d.pos = c.dPos.p; d.atomType = c.dAtomType.p; d.cumulVale = c.dCumV.p; d.cumulCore = c.dCumC.p;
    d.valeDim = c.valeDim; d.coreDim = c.coreDim; d.maxA = c.maxA; d.maxS = c.maxS;
    d.alpha = c.dAlpha.p; d.elemAlpha = c.dElemAlpha.p; d.basis = c.dBasis.p;
    d.lmIndex = c.dLmIndex.p; d.nVale = c.dNVale.p; d.nCore = c.dNCore.p;
    d.nPotTypes = c.nPT; d.maxPA = c.maxPA; d.nPotSites = c.nPS;
    d.potNA = c.dPotNA.p; d.potAlpha = c.dPotAlpha.p; d.potZ = c.dPotZ.p; d.potNuc = c.dPotNuc.p;
    d.potCumul = c.dPotCumul.p; d.potPos = c.dPotPos.p; d.potSiteType = c.dPotSiteType.p;
    std::memcpy(d.vec, c.vec, sizeof(d.vec));
    std::memcpy(d.inv, c.inv, sizeof(d.inv));
    d.nCorr = (int)c.hCorr.size() / 3;
    for (size_t q = 0; q < c.hCorr.size(); ++q) d.corr[q] = c.hCorr[q];
    d.nCells = (int)c.hCellSizes.size();
    d.cellDims = c.dCellDims.p; d.cellSizes = c.dCellSizes.p;
    d.nk = c.nk; d.kcart = c.dKcart.p; d.phase = c.dPhase.p; d.tau = c.tau;
    d.nPairs = (int)c.hPairI.size();
    d.pairI = c.dPairI.p; d.pairJ = c.dPairJ.p; d.pairLv = c.dPairLv.p; d.pairPc = c.dPairPc.p;
    d.nPc = (int)c.hPcPair.size();
    d.pcPair = c.dPcPair.p; d.pcCell = c.dPcCell.p; d.pcB = c.dPcB.p; d.pcS2 = c.dPcS2.p;
    d.pcPp = c.dPcPp.p;
    d.nPp = (long long)c.hPpPc.size();
    d.ppPc = c.dPpPc.p; d.ppAB = c.dPpAB.p; d.ppOff = c.dPpOff.p;
    d.counters = c.dCounters.p;
    d.anyNuc = 0;
    for (int t = 0; t < c.nPT; ++t)
        if (c.hPotZ[t] >= kSmallThresh) d.anyNuc = 1;
    HIPCHK(hipStreamSynchronize(s));
}

// Per-pair image lists of the three-centre sweeps.  The screening depends on
// the geometry and the exponents only, so it runs once per context: count
// the entries and the fused-sweep items of every primitive pair, order each
// class's records by item count (heaviest first, so that the 64 lanes of a
// wave carry nearly equal work), then write the entries in that order.
void build_image_lists(olcao_ctx &c)
{
    hipStream_t s = c.stream;
    const auto t0 = std::chrono::steady_clock::now();
    DBuf<int> dCnt, dItems, dOverflow;
    dOverflow.alloc(1);
    HIPCHK(hipMemsetAsync(dOverflow.p, 0, sizeof(int), s));
    c.recOrder.assign(16, {});
    c.imageEntries = 0;
    for (int k = 0; k < 16; ++k) {
        const int n = c.cls.count[k];
        if (n == 0) continue;
        dCnt.alloc(n);
        dItems.alloc(n);
        launch_image_count(c.dev, c.cls.rec[k], n, dCnt.p, dItems.p, s);
        std::vector<int> cnt(n), items(n);
        HIPCHK(hipMemcpyAsync(cnt.data(), dCnt.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(items.data(), dItems.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        std::vector<int> order(n);
        for (int q = 0; q < n; ++q) order[q] = q;
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return items[x] > items[y]; });
        std::vector<PPRec> recs(n);
        std::vector<long long> off(n + 1, 0);
        c.recOrder[k].resize(n);
        for (int q = 0; q < n; ++q) {
            c.recOrder[k][q] = c.classIdx[k][order[q]];
            recs[q] = make_rec(c, c.recOrder[k][q]);
            off[q + 1] = off[q] + cnt[order[q]];
        }
        c.dRec[k].upload(recs, s);
        c.cls.rec[k] = c.dRec[k].p;
        c.dIoff[k].upload(off, s);
        c.cls.ioff[k] = c.dIoff[k].p;
        c.dImg[k].alloc(std::max<long long>(off[n], 1));
        c.cls.img[k] = c.dImg[k].p;
        c.imageEntries += off[n];
        launch_image_fill(c.dev, c.cls.rec[k], n, c.cls.ioff[k], c.dImg[k].p, dOverflow.p, s);
        // chunks of about chunkItems items, entries split evenly within a record
        struct Chunk { int rec; long long beg, end; double est; };
        std::vector<Chunk> ch;
        ch.reserve(n);
        for (int q = 0; q < n; ++q) {
            const int o = order[q];
            const long long ne = cnt[o];
            const int nch = (int)std::max<long long>(1, std::min<long long>(ne, (items[o] + c.chunkItems - 1) / c.chunkItems));
            for (int u = 0; u < nch; ++u) {
                const long long b = off[q] + ne * u / nch, e = off[q] + ne * (u + 1) / nch;
                ch.push_back({q, b, e, ne ? (double)items[o] * (double)(e - b) / (double)ne : 0.0});
            }
        }
        std::vector<int> corder(ch.size());
        for (size_t u = 0; u < ch.size(); ++u) corder[u] = (int)u;
        std::stable_sort(corder.begin(), corder.end(), [&](int x, int y) { return ch[x].est > ch[y].est; });
        const int nc = (int)ch.size();
        std::vector<int> chRec(nc), rcOff(n + 1, 0), rcList(nc), pos(nc);
        std::vector<long long> chBeg(nc), chEnd(nc);
        for (int u = 0; u < nc; ++u) {
            const Chunk &x = ch[corder[u]];
            chRec[u] = x.rec;
            chBeg[u] = x.beg;
            chEnd[u] = x.end;
            pos[corder[u]] = u;
        }
        // chunks were generated record by record in entry order
        for (int u = 0; u < nc; ++u) rcOff[ch[u].rec + 1]++;
        for (int q = 0; q < n; ++q) rcOff[q + 1] += rcOff[q];
        for (int u = 0; u < nc; ++u) rcList[u] = pos[u];
        std::vector<int> finOrder;
        finOrder.reserve(n);
        {
            std::vector<char> seen(n, 0);
            for (int u = 0; u < nc; ++u)
                if (!seen[chRec[u]]) {
                    seen[chRec[u]] = 1;
                    finOrder.push_back(chRec[u]);
                }
        }
        c.dFinOrder[k].upload(finOrder, s);
        c.dChRec[k].upload(chRec, s);
        c.dChBeg[k].upload(chBeg, s);
        c.dChEnd[k].upload(chEnd, s);
        c.dRcOff[k].upload(rcOff, s);
        c.dRcList[k].upload(rcList, s);
        HIPCHK(hipStreamSynchronize(s));
        c.cls.chRec[k] = c.dChRec[k].p;
        c.cls.chBeg[k] = c.dChBeg[k].p;
        c.cls.chEnd[k] = c.dChEnd[k].p;
        c.cls.rcOff[k] = c.dRcOff[k].p;
        c.cls.rcList[k] = c.dRcList[k].p;
        c.cls.finOrder[k] = c.dFinOrder[k].p;
        c.cls.nChunk[k] = nc;
        {   // compacted nuclear entries per chunk
            DBuf<int> dNc;
            dNc.alloc(std::max(nc, 1));
            launch_nuc_compact(c.dImg[k].p, c.dChBeg[k].p, c.dChEnd[k].p, nc, dNc.p, nullptr, nullptr, s);
            std::vector<int> ncnt(nc);
            HIPCHK(hipMemcpyAsync(ncnt.data(), dNc.p, nc * sizeof(int), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            std::vector<long long> noff(nc + 1, 0);
            for (int u = 0; u < nc; ++u) noff[u + 1] = noff[u] + ncnt[u];
            c.dNOff[k].upload(noff, s);
            c.dImgN[k].alloc(std::max<long long>(noff[nc], 1));
            launch_nuc_compact(c.dImg[k].p, c.dChBeg[k].p, c.dChEnd[k].p, nc, nullptr, c.dNOff[k].p,
                               c.dImgN[k].p, s);
            HIPCHK(hipStreamSynchronize(s));
            c.cls.imgN[k] = c.dImgN[k].p;
            c.cls.nOff[k] = c.dNOff[k].p;
        }
    }
    // Hermite-sum scratch of the three-centre sweeps, one slice per class
    size_t wTotal = 0;
    size_t wOff[16];
    for (int k = 0; k < 16; ++k) {
        const int L = k / 4 + k % 4;
        wOff[k] = wTotal;
        wTotal += (size_t)(L + 1) * (L + 2) * (L + 3) / 6 * (size_t)c.cls.nChunk[k];
    }
    c.dW.alloc(std::max<size_t>(wTotal, 1));
    for (int k = 0; k < 16; ++k) c.cls.W[k] = c.dW.p + wOff[k];
    int overflow = 0;
    HIPCHK(hipMemcpyAsync(&overflow, dOverflow.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    if (overflow & 1)
        fail(OLCAO_ERR_UNSUPPORTED, "a potential image lies more than %d cells away", kImgLatOff);
    if (overflow & 2) fail(OLCAO_ERR_HIP, "image-list passes disagree");
    c.listMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Per-class work spread over the context stream and the side streams and
// joined back onto the context stream.  The 16 angular classes are
// independent (each has its own primitive-pair blocks, W slice and
// counters).  Most of them are small launches whose time is the serial
// chain of one lane, not the chip's throughput; beside the largest class
// they cost little.  Longest first onto the least-loaded stream.
template <class F>
void fork_classes(olcao_ctx &c, const double (&w)[16], F launch)
{
    constexpr int nS = 1 + olcao_ctx::kSide;
    const int use = std::max(1, std::min(nS, c.forkStreams));
    hipStream_t st[nS] = {c.stream};
    for (int i = 1; i < nS; ++i) st[i] = c.side[i - 1];
    int order[16];
    for (int k = 0; k < 16; ++k) order[k] = k;
    std::stable_sort(order, order + 16, [&](int a, int b) { return w[a] > w[b]; });
    double load[nS] = {};
    unsigned mask[nS] = {};
    for (int k : order) {
        if (w[k] == 0.0) continue;
        const int i = (int)(std::min_element(load, load + use) - load);
        mask[i] |= 1u << k;
        load[i] += w[k];
    }
    HIPCHK(hipEventRecord(c.evFork, c.stream));
    for (int i = 1; i < nS; ++i)
        if (mask[i]) HIPCHK(hipStreamWaitEvent(st[i], c.evFork, 0));
    for (int i = 0; i < nS; ++i)
        if (mask[i]) launch(mask[i], st[i]);
    HIPCHK(hipGetLastError());
    for (int i = 1; i < nS; ++i)
        if (mask[i]) {
            HIPCHK(hipEventRecord(c.evJoin[i - 1], st[i]));
            HIPCHK(hipStreamWaitEvent(c.stream, c.evJoin[i - 1], 0));
        }
}

// Two-centre blocks of op (accumulated onto the store when acc) for the
// classes in classMask, weighted by pairs x block size.
void two_center_forked(olcao_ctx &c, int op, int acc, unsigned classMask = 0xffffu)
{
    double w[16];
    for (int k = 0; k < 16; ++k)
        w[k] = (classMask >> k & 1u) ? (double)c.cls.count[k] * nshh(k / 4) * nshh(k % 4) : 0.0;
    fork_classes(c, w, [&](unsigned mask, hipStream_t s) {
        launch_two_center(c.dev, c.cls, op, acc, c.dBlocks.p, mask, s);
    });
}

// The three-centre sweep, classes weighted by chunks x Hermite sums per item.
void three_center_forked(olcao_ctx &c, int kind, int tt, int ti, const double *coef,
                         int accumulate)
{
    double w[16];
    for (int k = 0; k < 16; ++k) {
        const int L = k / 4 + k % 4;
        w[k] = c.cls.count[k] ? (double)c.cls.nChunk[k] * ((L + 1) * (L + 2) * (L + 3) / 6) : 0.0;
    }
    const int mixed = c.precision == OLCAO_PREC_MIXED;
    fork_classes(c, w, [&](unsigned mask, hipStream_t s) {
        launch_three_center(c.dev, c.cls, kind, tt, ti, coef, accumulate, c.dBlocks.p, mixed, mask,
                            s);
    });
}

// Three-centre sweep with its work counters cleared first (a fill kernel,
// so the sequence stays capturable in a graph).
void three_center(olcao_ctx &c, int kind, int tt, int ti, const double *coef, int accumulate)
{
    launch_fill(reinterpret_cast<double *>(c.dCounters.p), 32, 0.0, c.stream);
    three_center_forked(c, kind, tt, ti, coef, accumulate);
}

// The force sweep's three-centre part (classes with la, lb <= 2), weighted
// like the three-centre sweep one Hermite order up.
void three_center_force_forked(olcao_ctx &c, const double *coef, int comp)
{
    double w[16];
    for (int k = 0; k < 16; ++k) {
        const int la = k / 4, lb = k % 4, L = la + lb + 1;
        w[k] = (la <= 2 && lb <= 2 && c.cls.count[k])
                   ? (double)c.cls.nChunk[k] * ((L + 1) * (L + 2) * (L + 3) / 6)
                   : 0.0;
    }
    fork_classes(c, w, [&](unsigned mask, hipStream_t s) {
        launch_three_center_force(c.dev, c.cls, coef, comp, c.dBlocks.p, mask, s);
    });
}

// ---------------------------------------------------------------------------
// One integral sweep into (vv, cv, cc) device matrices.
enum Sweep { SW_OL, SW_KE, SW_NP, SW_EP, SW_HAM, SW_MV, SW_DM, SW_MM, SW_KO };

// Primitive-pair blocks of one sweep.  SW_MV stores the negated
// mass-velocity blocks, the term as it enters H (gaussOverlapMV accumulates
// with a minus sign, integrals.F90:1434-1447).  SW_HAM is the fused PSCF
// block of one spin: kinetic, minus mass-velocity when rel = 1, plus the
// potential terms with that spin's coefficients and the nuclear attraction
// (integrals.F90:3211-3490).
void sweep_blocks(olcao_ctx &c, Sweep kind, int term, int spinIdx)
{
    hipStream_t s = c.stream;
    const DevCtx &d = c.dev;
    switch (kind) {
    case SW_OL: two_center_forked(c, OP_OL, 0); break;
    case SW_KE: two_center_forked(c, OP_KE, 0); break;
    case SW_MV: two_center_forked(c, OP_MV, 0); break;
    case SW_DM: two_center_forked(c, OP_DM + term, 0); break;   // centre in c.dev.dipC
    case SW_MM: two_center_forked(c, OP_MM + term, 0); break;
    case SW_KO: two_center_forked(c, OP_KO + term, 0); break;   // deltaK in c.dev.kq
    case SW_NP: three_center(c, K3_NP, 0, 0, nullptr, 0); break;
    case SW_EP: {
        int t = 0;
        while (t + 1 < c.nPT && c.hPotCumul[t + 1] <= term) ++t;
        three_center(c, K3_EP, t, term - c.hPotCumul[t], nullptr, 0);
        break;
    }
    case SW_HAM:
        if (c.rel) {
            two_center_forked(c, OP_KE, 0);
            two_center_forked(c, OP_MV, 1);
            three_center(c, K3_HAM, 0, 0, c.dCoef.p + (size_t)spinIdx * c.potDim, 1);
        } else {   // the finish kernels add the kinetic blocks of their classes
            two_center_forked(c, OP_KE, 0, ~kFusedKEClasses & 0xffffu);
            three_center(c, K3_HAM, 0, 0, c.dCoef.p + (size_t)spinIdx * c.potDim, 1 | kAccAddKE);
        }
        break;
    }
}

// Contraction, Bloch sum and scatter of the current blocks; mode bit 0:
// the Bloch sum times -i (the momentum matrix); bit 1: the full save.
void contract_scatter(olcao_ctx &c, double *vv, double *cv, double *cc, int mode = 0)
{
    hipStream_t s = c.stream;
    launch_contract(c.dev, c.cplan, c.dBlocks.p, c.dX.p, s);
    launch_fill(vv, (long long)c.matDoubles(), 0.0, s);
    if (c.coreDim) {
        launch_fill(cv, (long long)c.cvDoubles(), 0.0, s);
        launch_fill(cc, (long long)c.ccDoubles(), 0.0, s);
    }
    if (c.blochMfma)
        launch_bloch_mfma(c.dev, c.dX.p, c.dPslice.p, vv, cv, cc, c.gamma ? 1 : 0, mode,
                          c.precision == OLCAO_PREC_MIXED && c.blochSplit, s);
    else
        launch_bloch_scatter(c.dev, c.dX.p, vv, cv, cc, c.gamma ? 1 : 0, mode, s);
    HIPCHK(hipGetLastError());
}

void sweep(olcao_ctx &c, Sweep kind, int term, int spinIdx, double *vv, double *cv, double *cc)
{
    sweep_blocks(c, kind, term, spinIdx);
    // the momentum matrix, and the k-overlap's imaginary part, take -i
    contract_scatter(c, vv, cv, cc, kind == SW_MM || (kind == SW_KO && term == 1) ? 1 : 0);
}

// Orthogonalisation of valence against core (intgOrtho.F90, integrals.F90
// orthoOL :3608, ortho :3747, orthoHamPSCF :3913), per k with batched GEMMs:
//   overlap   S' = S_vv - 2 S_cv^H S_cv + S_cv^H S_cc S_cv
//   operator  O' = O_vv - S_cv^H O_cv - O_cv^H S_cv + S_cv^H O_cc S_cv
// then the upper triangle is kept and components below 1e-10 are zeroed, as
// makeValeVale packs them (intgOrtho.F90:142-190).  Nothing happens when
// coreDim = 0 (the reference packs without zeroing then).
void gemm_b(olcao_ctx &c, bool conjA, int m, int n, int k, double alpha, const double *A, int lda,
            long long sA, const double *B, int ldb, long long sB, double beta, double *Cm,
            int ldc, long long sC)
{
    const int nb = c.gamma ? 1 : c.nk;
    rocblas_status r;
    if (c.gamma) {
        r = rocblas_dgemm_strided_batched(c.blas, conjA ? rocblas_operation_transpose
                                                        : rocblas_operation_none,
                                          rocblas_operation_none, m, n, k, &alpha, A, lda, sA, B,
                                          ldb, sB, &beta, Cm, ldc, sC, nb);
    } else {
        const rocblas_double_complex al = {alpha, 0.0}, be = {beta, 0.0};
        r = rocblas_zgemm_strided_batched(
            c.blas, conjA ? rocblas_operation_conjugate_transpose : rocblas_operation_none,
            rocblas_operation_none, m, n, k, &al, (const rocblas_double_complex *)A, lda, sA,
            (const rocblas_double_complex *)B, ldb, sB, &be, (rocblas_double_complex *)Cm, ldc,
            sC, nb);
    }
    if (r != rocblas_status_success) fail(OLCAO_ERR_SOLVER, "gemm status %d", (int)r);
}

void ortho(olcao_ctx &c, double *Ovv, const double *Ocv, const double *Occ, bool overlap,
           bool zeroSmall = true)
{
    if (!c.coreDim) return;
    const int n = c.valeDim, nc = c.coreDim;
    const long long svv = (long long)n * n, scv = (long long)nc * n, scc = (long long)nc * nc;
    const double *Scv = c.CV_S.p;
    c.Tmp.alloc(c.cvDoubles());
    if (overlap) {
        gemm_b(c, true, n, n, nc, -2.0, Scv, nc, scv, Scv, nc, scv, 1.0, Ovv, n, svv);
    } else {
        gemm_b(c, true, n, n, nc, -1.0, Scv, nc, scv, Ocv, nc, scv, 1.0, Ovv, n, svv);
        gemm_b(c, true, n, n, nc, -1.0, Ocv, nc, scv, Scv, nc, scv, 1.0, Ovv, n, svv);
    }
    // T = O_cc S_cv, then O_vv += S_cv^H T
    gemm_b(c, false, nc, n, nc, 1.0, Occ, nc, scc, Scv, nc, scv, 0.0, c.Tmp.p, nc, scv);
    gemm_b(c, true, n, n, nc, 1.0, Scv, nc, scv, c.Tmp.p, nc, scv, 1.0, Ovv, n, svv);
    if (zeroSmall) {
        launch_ortho_finish(Ovv, n, c.gamma ? 1 : c.nk, c.gamma ? 1 : 2, c.stream);
        HIPCHK(hipGetLastError());
    }
}

// The overlap sweep with S_cv and S_cc kept for the later orthogonalisations.
void overlap_sweep(olcao_ctx &c)
{
    sweep(c, SW_OL, 0, 0, c.S.p, c.CV_S.p, c.CC_S.p);
    ortho(c, c.S.p, c.CV_S.p, c.CC_S.p, true);
    c.haveS = true;
}

// An operator sweep into vv, orthogonalised with its own core blocks (which
// are left in cv, cc).  The overlap sweep runs first if it has not.
void operator_sweep(olcao_ctx &c, Sweep kind, int term, int spinIdx, double *vv, double *cv,
                    double *cc)
{
    if (c.coreDim && !c.haveS) overlap_sweep(c);
    sweep(c, kind, term, spinIdx, vv, cv, cc);
    ortho(c, vv, cv, cc, false);
}

void copy_out(olcao_ctx &c, const double *src, double *dst, size_t n)
{
    if (!dst || n == 0) return;
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
}

// vv (and, when coreDim > 0, the raw cv, cc) of spin slot q to host buffers
// laid out spin-major.
void copy_out_spin(olcao_ctx &c, int q, double *vv, double *cv, double *cc)
{
    copy_out(c, c.H[q].p, vv ? vv + (size_t)q * c.matDoubles() : nullptr, c.matDoubles());
    if (c.coreDim) {
        copy_out(c, c.CV_H[q].p, cv ? cv + (size_t)q * c.cvDoubles() : nullptr, c.cvDoubles());
        copy_out(c, c.CC_H[q].p, cc ? cc + (size_t)q * c.ccDoubles() : nullptr, c.ccDoubles());
    }
}

void ensure_mats(olcao_ctx &c)
{
    c.S.alloc(c.matDoubles());
    for (int q = 0; q < c.spin; ++q) c.H[q].alloc(c.matDoubles());
    if (c.coreDim) {
        c.CV_S.alloc(c.cvDoubles()); c.CC_S.alloc(c.ccDoubles());
        for (int q = 0; q < c.spin; ++q) {
            c.CV_H[q].alloc(c.cvDoubles());
            c.CC_H[q].alloc(c.ccDoubles());
        }
    }
}

// C = Y diag(w) Y^H for the Hermitian C in A (full): the cooperative
// tridiagonal reduction C = Q T Q^H (ktrd.hip), the batched divide and
// conquer T = Z diag(w) Z^T (kdc.hip), and Y = Q Z by blocks of reflectors,
// X <- X - U_b (V_b^H X), with batched GEMMs.  D receives w, A receives Y,
// dinfo the reduction's status.
void gemm_one(EigWork &ew, rocblas_handle h, hipStream_t s, bool cplx, bool ta, int m, int nc, int k,
              double alpha, const double *A, int lda, const double *B, int ldb, double beta, double *C,
              int ldc);

// Trailing update C -= X Y^H (m x m, k columns; X = [W V], Y = [V W] from
// the panel buffer), batched with the strides given: W V^H + V W^H in one
// GEMM.  Both triangles are kept current.
void panel_update(rocblas_handle h, bool cplx, int m, int k, const double *X, const double *Y,
                  rocblas_stride sxy, double *C, rocblas_stride sc, int ld, int nb)
{
    rocblas_status r;
    if (cplx) {
        const rocblas_double_complex alpha = {-1.0, 0.0}, beta = {1.0, 0.0};
        r = rocblas_zgemm_strided_batched(h, rocblas_operation_none,
                                          rocblas_operation_conjugate_transpose, m, m, k, &alpha,
                                          (const rocblas_double_complex *)X, ld, sxy,
                                          (const rocblas_double_complex *)Y, ld, sxy, &beta,
                                          (rocblas_double_complex *)C, ld, sc, nb);
    } else {
        const double alpha = -1.0, beta = 1.0;
        r = rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_transpose, m, m,
                                          k, &alpha, X, ld, sxy, Y, ld, sxy, &beta, C, ld, sc, nb);
    }
    if (r != rocblas_status_success) fail(OLCAO_ERR_SOLVER, "panel update gemm status %d", (int)r);
}

// The blocked reduction (ktrdb.hip) of the batch in A (lower triangle used,
// overwritten by the trailing updates), nGroups matrices at a time: per
// panel of kPanelB columns the start kernel, the cooperative panel kernel
// and the GEMM update of the trailing matrix; then d(n-1).  Outputs as
// launch_trd: V, tau (in ew), D, E, dinfo (-2: a group barrier timed out).
void trdb_reduce(EigWork &ew, const TrdbPlan &pl, rocblas_handle h, hipStream_t s, bool cplx, int n,
                 int nb, double *A, double *D, double *E, int *dinfo)
{
    const int el = cplx ? 2 : 1;
    const size_t per = (size_t)n * n, perW = (size_t)n * 3 * kPanelB;
    ew.Wp.alloc(perW * el * nb);
    ew.pub.alloc(pl.pubLen * nb);
    ew.cnt.alloc((size_t)64 * nb);
    launch_fill(ew.V.p, (long long)(per * el * nb), 0.0, s);
    launch_zero_words(dinfo, nb, s);
    for (int m0 = 0; m0 < nb; m0 += pl.nGroups) {
        const int nm = std::min(pl.nGroups, nb - m0);
        const double *Am = A + el * per * m0;
        double *Aw = A + el * per * m0;
        double *Vm = ew.V.p + el * per * m0, *Wm = ew.Wp.p + el * perW * m0;
        double *Dm = D + (size_t)n * m0, *Em = E + (size_t)n * m0;
        double *taum = ew.tau.p + (size_t)el * n * m0, *pubm = ew.pub.p + pl.pubLen * m0;
        unsigned *cntm = ew.cnt.p + (size_t)64 * m0;
        int *im = dinfo + m0;
        for (int j0 = 0; j0 < n - 1; j0 += kPanelB) {
            const int pb = std::min(kPanelB, n - 1 - j0);
            HIPCHK(launch_trdb_start(cplx, pl, Am, Vm, Wm, Dm, Em, taum, pubm, cntm, im, n, nm, j0, s));
            HIPCHK(launch_trdb_panel(cplx, pl, Am, Vm, Wm, Dm, Em, taum, pubm, cntm, im, n, nm, j0, pb,
                                     s));
            const int t0 = j0 + pb;
            if (t0 < n)
                panel_update(h, cplx, n - t0, 2 * pb, Wm + el * (size_t)t0,
                             Wm + el * ((size_t)t0 + (size_t)n * pb), (rocblas_stride)perW,
                             Aw + el * ((size_t)t0 + (size_t)n * t0), (rocblas_stride)per, n, nm);
        }
        HIPCHK(launch_trdb_start(cplx, pl, Am, Vm, Wm, Dm, Em, taum, pubm, cntm, im, n, nm, n - 1, s));
    }
}

void trd_eig(EigWork &ew, const TrdPlan *plan, const TrdbPlan *bplan, rocblas_handle h,
             hipStream_t s, bool cplx, int n, int nb, double *A, double *D, double *E, int *dinfo,
             bool defer)
{
    const int el = cplx ? 2 : 1;
    const size_t per = (size_t)n * n;
    ew.V.alloc(per * el * nb);
    ew.U.alloc(per * el * nb);
    ew.W.alloc((size_t)kTrdBlock * n * el * nb);
    ew.tau.alloc((size_t)n * el * nb);
    if (plan) {
        ew.pub.alloc(plan->pubDoubles);
        ew.cnt.alloc(plan->counters);
    }
    ew.Z.alloc(per * nb);
    ew.Zt.alloc(per * nb);
    ew.Wd.alloc(per * nb);
    ew.ord.alloc((size_t)2 * n * nb);   // per matrix: sort order, then its inverse
    if (ew.dc.n != n) {
        if (!dc_plan(n, ew.dc))
            fail(OLCAO_ERR_UNSUPPORTED, "divide and conquer takes n <= %d, not %d", kDcBigMaxN, n);
        ew.dcNodes.upload(ew.dc.nodes, s);
        ew.dcTiles.upload(ew.dc.tiles, s);
        ew.dcRoot.upload(ew.dc.rootTiles, s);
        ew.dcCol.upload(ew.dc.colTiles, s);
        ew.dcLeaf.upload(ew.dc.leaves, s);
    }
    if (plan) launch_zero_words(ew.cnt.p, (long long)plan->counters, s);
    const bool timed = ew.trdEv[0] && !defer;
    if (timed) HIPCHK(hipEventRecord(ew.trdEv[0], s));
    if (plan)
        HIPCHK(launch_trd(cplx, *plan, n, nb, A, ew.V.p, D, E, ew.tau.p, ew.pub.p, ew.cnt.p, dinfo, s));
    else
        trdb_reduce(ew, *bplan, h, s, cplx, n, nb, A, D, E, dinfo);
    if (timed) {
        HIPCHK(hipEventRecord(ew.trdEv[1], s));
        ew.trdTimed = true;
    }
    ew.dcStD.alloc(dc_state_doubles(n, nb));
    ew.dcStI.alloc(dc_state_ints(n, nb));
    ew.dcCnt.alloc(ew.dc.nodes.size() * nb);
    ew.dcRho.alloc(ew.dc.nodes.size() * nb);
    const DcDev dv{ew.dcNodes.p, ew.dcTiles.p, ew.dcRoot.p, ew.dcCol.p, ew.dcStD.p, ew.dcStI.p,
                   ew.dcCnt.p, ew.dcRho.p, ew.dcLeaf.p};
    if (!launch_dc(ew.dc, dv, nb, D, E, ew.Z.p, ew.Zt.p, ew.Wd.p, ew.ord.p, s))
        fail(OLCAO_ERR_UNSUPPORTED, "divide and conquer takes n <= %d, not %d", kDcBigMaxN, n);
    HIPCHK(hipGetLastError());
    ew.Tw.alloc(trd_tfactor_work(n, nb, cplx));
    launch_trd_tfactor(cplx, n, nb, ew.V.p, ew.tau.p, ew.Tw.p, ew.U.p, s);
    HIPCHK(hipGetLastError());
    launch_real_to_field(ew.Z.p, A, (long long)per * nb, cplx, s);
    HIPCHK(hipGetLastError());
    const rocblas_stride st = (rocblas_stride)per, stW = (rocblas_stride)kTrdBlock * n;
    const int nref = n - 1;
    for (int j0 = ((nref - 1) / kTrdBlock) * kTrdBlock; j0 >= 0; j0 -= kTrdBlock) {
        const int mb = std::min(kTrdBlock, nref - j0), m = n - 1 - j0;
        const size_t offV = (size_t)(j0 + 1) + (size_t)j0 * n, offX = (size_t)(j0 + 1);
        rocblas_status r;
        if (nb == 1) {
            // W = V_b^H X (split-K over the block's rows), X -= U_b W
            gemm_one(ew, h, s, cplx, true, mb, n, m, 1.0, ew.V.p + (size_t)el * offV, n,
                     A + (size_t)el * offX, n, 0.0, ew.W.p, kTrdBlock);
            gemm_one(ew, h, s, cplx, false, m, n, mb, -1.0, ew.U.p + (size_t)el * offV, n, ew.W.p,
                     kTrdBlock, 1.0, A + (size_t)el * offX, n);
            continue;
        }
        if (cplx) {
            const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0}, mone = {-1.0, 0.0};
            auto *Vz = (const rocblas_double_complex *)ew.V.p, *Uz = (const rocblas_double_complex *)ew.U.p;
            auto *Xz = (rocblas_double_complex *)A, *Wz = (rocblas_double_complex *)ew.W.p;
            r = rocblas_zgemm_strided_batched(h, rocblas_operation_conjugate_transpose,
                                              rocblas_operation_none, mb, n, m, &one, Vz + offV, n,
                                              st, Xz + offX, n, st, &zero, Wz, kTrdBlock, stW, nb);
            if (r == rocblas_status_success)
                r = rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, m,
                                                  n, mb, &mone, Uz + offV, n, st, Wz, kTrdBlock, stW,
                                                  &one, Xz + offX, n, st, nb);
        } else {
            const double one = 1.0, zero = 0.0, mone = -1.0;
            r = rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none,
                                              mb, n, m, &one, ew.V.p + offV, n, st, A + offX, n, st,
                                              &zero, ew.W.p, kTrdBlock, stW, nb);
            if (r == rocblas_status_success)
                r = rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, m,
                                                  n, mb, &mone, ew.U.p + offV, n, st, ew.W.p, kTrdBlock,
                                                  stW, &one, A + offX, n, st, nb);
        }
        if (r != rocblas_status_success) fail(OLCAO_ERR_SOLVER, "back-transform gemm status %d", (int)r);
    }
}

// Batched GEMM on column-major n x n matrices (stride n^2), offsets in
// elements: C = alpha op(A) op(B) + beta C, op = N or C (T for real).
void gemm_nn(rocblas_handle h, bool cplx, bool ta, bool tb, int m, int nc, int k, double alpha,
             const double *A, const double *B, double beta, double *C, int n, int nb)
{
    const rocblas_stride st = (rocblas_stride)n * n;
    rocblas_status r;
    if (cplx) {
        const rocblas_double_complex al = {alpha, 0.0}, be = {beta, 0.0};
        r = rocblas_zgemm_strided_batched(
            h, ta ? rocblas_operation_conjugate_transpose : rocblas_operation_none,
            tb ? rocblas_operation_conjugate_transpose : rocblas_operation_none, m, nc, k, &al,
            (const rocblas_double_complex *)A, n, st, (const rocblas_double_complex *)B, n, st, &be,
            (rocblas_double_complex *)C, n, st, nb);
    } else {
        r = rocblas_dgemm_strided_batched(h, ta ? rocblas_operation_transpose : rocblas_operation_none,
                                          tb ? rocblas_operation_transpose : rocblas_operation_none, m,
                                          nc, k, &alpha, A, n, st, B, n, st, &beta, C, n, st, nb);
    }
    if (r != rocblas_status_success) fail(OLCAO_ERR_SOLVER, "gemm status %d", (int)r);
}

// C = alpha op(A) op(B) + beta C for one matrix (op(B) = B), with explicit
// leading dimensions.  A product with a short side (at most 64 rows or
// columns of C) and a long inner dimension runs split-K: chunks of 512 of the
// inner dimension as one strided batch of GEMMs into ew.Psk, then summed
// into C (launch_splitk_sum), so that the chunks fill the device.
void gemm_one(EigWork &ew, rocblas_handle h, hipStream_t s, bool cplx, bool ta, int m, int nc, int k,
              double alpha, const double *A, int lda, const double *B, int ldb, double beta, double *C,
              int ldc)
{
    const int el = cplx ? 2 : 1;
    constexpr int kc = 512;
    auto call = [&](int kk, const double *Ap, rocblas_stride sa, const double *Bp, rocblas_stride sb,
                    double al, double be, double *Cp, int ldcp, rocblas_stride scp, int batch) {
        rocblas_status r;
        if (cplx) {
            const rocblas_double_complex a = {al, 0.0}, b = {be, 0.0};
            r = rocblas_zgemm_strided_batched(
                h, ta ? rocblas_operation_conjugate_transpose : rocblas_operation_none,
                rocblas_operation_none, m, nc, kk, &a, (const rocblas_double_complex *)Ap, lda, sa,
                (const rocblas_double_complex *)Bp, ldb, sb, &b, (rocblas_double_complex *)Cp, ldcp, scp,
                batch);
        } else {
            r = rocblas_dgemm_strided_batched(h, ta ? rocblas_operation_transpose : rocblas_operation_none,
                                              rocblas_operation_none, m, nc, kk, &al, Ap, lda, sa, Bp, ldb,
                                              sb, &be, Cp, ldcp, scp, batch);
        }
        if (r != rocblas_status_success) fail(OLCAO_ERR_SOLVER, "gemm status %d", (int)r);
    };
    if (k < 4 * kc || (m > 64 && nc > 64)) {
        call(k, A, 0, B, 0, alpha, beta, C, ldc, 0, 1);
        return;
    }
    const int full = k / kc, rem = k % kc, parts = full + (rem > 0);
    const size_t part = (size_t)m * nc * el;
    ew.Psk.reserve(part * parts);
    // chunk c of the inner dimension: rows c kc.. of the stored A (op = ^H)
    // or its columns c kc.. (op = N); rows c kc.. of B
    const rocblas_stride sa = ta ? (rocblas_stride)kc : (rocblas_stride)kc * lda;
    call(kc, A, sa, B, (rocblas_stride)kc, 1.0, 0.0, ew.Psk.p, m, (rocblas_stride)(m * nc), full);
    if (rem > 0)
        call(rem, A + el * (size_t)sa * full, 0, B + el * (size_t)kc * full, 0, 1.0, 0.0,
             ew.Psk.p + part * full, m, 0, 1);
    launch_splitk_sum(ew.Psk.p, parts, m, nc, cplx, alpha, beta, C, ldc, s);
    HIPCHK(hipGetLastError());
}

// gemm_nn, or gemm_one for a single matrix
void gemm_big(EigWork &ew, rocblas_handle h, hipStream_t s, bool cplx, bool ta, int m, int nc, int k,
              double alpha, const double *A, const double *B, double beta, double *C, int n, int nb)
{
    if (nb == 1)
        gemm_one(ew, h, s, cplx, ta, m, nc, k, alpha, A, n, B, n, beta, C, n);
    else
        gemm_nn(h, cplx, ta, false, m, nc, k, alpha, A, B, beta, C, n, nb);
}

// S = U^H U for the batch in B (upper triangle read; B is overwritten), the
// factor into U (upper triangle; the rest is not touched), left-looking by
// 64-wide block rows: a GEMM applies the finished rows to block row J, the
// diagonal kernel factors and inverts the diagonal block (Dinv), a GEMM
// forms the row panel U(J, >J) = Dinv_J^H S(J, >J).  dinfo: first failing
// column (1-based).
void chol_blocked(EigWork &ew, rocblas_handle h, hipStream_t s, bool cplx, int n, int nb, double *B,
                  double *U, double *Dinv, int *dinfo)
{
    const int el = cplx ? 2 : 1;
    launch_zero_words(dinfo, nb, s);
    for (int j0 = 0; j0 < n; j0 += kCholBlock) {
        const int jb = std::min(kCholBlock, n - j0), rest = n - j0 - jb;
        if (j0 > 0)
            gemm_big(ew, h, s, cplx, true, jb, n - j0, j0, -1.0, U + el * (size_t)n * j0,
                     U + el * (size_t)n * j0, 1.0, B + el * ((size_t)j0 + (size_t)n * j0), n, nb);
        launch_chol_diag(cplx, B, U, Dinv, n, nb, j0, jb, dinfo, s);
        HIPCHK(hipGetLastError());
        if (rest > 0)
            gemm_nn(h, cplx, true, false, jb, rest, jb, 1.0, Dinv + el * (size_t)j0,
                    B + el * ((size_t)j0 + (size_t)n * (j0 + jb)), 0.0,
                    U + el * ((size_t)j0 + (size_t)n * (j0 + jb)), n, nb);
    }
}

// The three triangular solves by 64-wide blocks with the inverted diagonal
// blocks (GEMMs only).  Dinv holds block J's inverse in rows [j0, j0 + jb),
// columns [0, jb) of an n x n matrix per k.  Block J = rows/columns
// [j0, j0 + jb).
//   left:  Y = U^-H A      Y_J = Dinv_J^H (A_J - U_{<J,J}^H Y_{<J})  (A destroyed)
//   right: C = Y U^-1      C_J = (Y_J - C_{<J} U_{<J,J}) Dinv_J      (Y destroyed)
//   back:  X = U^-1 Z      X_J = Dinv_J (Z_J - U_{J,>J} X_{>J})       (Z destroyed)
void solve_left_h(EigWork &ew, rocblas_handle h, hipStream_t s, bool cplx, int n, int nb,
                  const double *U, const double *Dinv, double *A, double *Y)
{
    const int el = cplx ? 2 : 1;
    for (int j0 = 0; j0 < n; j0 += kCholBlock) {
        const int jb = std::min(kCholBlock, n - j0);
        if (j0 > 0)
            gemm_big(ew, h, s, cplx, true, jb, n, j0, -1.0, U + el * (size_t)n * j0, Y, 1.0,
                     A + el * (size_t)j0, n, nb);
        gemm_nn(h, cplx, true, false, jb, n, jb, 1.0, Dinv + el * (size_t)j0, A + el * (size_t)j0,
                0.0, Y + el * (size_t)j0, n, nb);
    }
}
void solve_right(EigWork &ew, rocblas_handle h, hipStream_t s, bool cplx, int n, int nb,
                 const double *U, const double *Dinv, double *Y, double *C)
{
    const int el = cplx ? 2 : 1;
    for (int j0 = 0; j0 < n; j0 += kCholBlock) {
        const int jb = std::min(kCholBlock, n - j0);
        if (j0 > 0)
            gemm_big(ew, h, s, cplx, false, n, jb, j0, -1.0, C, U + el * (size_t)n * j0, 1.0,
                     Y + el * (size_t)n * j0, n, nb);
        gemm_nn(h, cplx, false, false, n, jb, jb, 1.0, Y + el * (size_t)n * j0, Dinv + el * (size_t)j0,
                0.0, C + el * (size_t)n * j0, n, nb);
    }
}
void solve_back(EigWork &ew, rocblas_handle h, hipStream_t s, bool cplx, int n, int nb,
                const double *U, const double *Dinv, double *Z, double *X)
{
    const int el = cplx ? 2 : 1;
    const int last = ((n - 1) / kCholBlock) * kCholBlock;
    for (int j0 = last; j0 >= 0; j0 -= kCholBlock) {
        const int jb = std::min(kCholBlock, n - j0), rest = n - j0 - jb;
        if (rest > 0)
            gemm_big(ew, h, s, cplx, false, jb, n, rest, -1.0,
                     U + el * ((size_t)j0 + (size_t)n * (j0 + jb)), X + el * (size_t)(j0 + jb), 1.0,
                     Z + el * (size_t)j0, n, nb);
        gemm_nn(h, cplx, false, false, jb, n, jb, 1.0, Dinv + el * (size_t)j0, Z + el * (size_t)j0, 0.0,
                X + el * (size_t)j0, n, nb);
    }
}

// Whether the tridiagonal path takes (n, nb), and its launch plan.
bool trd_fits(int n, int nb, bool cplx, TrdPlan &plan) { return trd_plan(n, nb, cplx, plan); }

// The path gen_eig takes for (n, nb): 0 the LDS solver, 1 the tridiagonal
// device path (blocked Cholesky first), 2 rocSOLVER.
int eig_path(const EigWork &ew, int n, int nb, bool cplx, bool library)
{
    if (library) return 2;
    if (small_hegv_fits(n, cplx)) return 0;
    TrdPlan plan;
    TrdbPlan bplan;
    if ((!ew.forceBlocked && trd_fits(n, nb, cplx, plan)) || trdb_plan(n, nb, cplx, bplan)) return 1;
    return 2;
}

// zhegv-style status of the last gen_eig from its device info arrays
// (dinfo: the LDS solver's or the Cholesky's or rocSOLVER's; tinfo: the
// tridiagonal reduction's): 0 ok, > n not positive definite, 1..n no
// convergence.
void read_info(const EigWork &ew, const int *dinfo, const int *tinfo, int n, int nb,
               hipStream_t s, std::vector<int> &info)
{
    std::vector<int> a(nb, 0), b(nb, 0);
    HIPCHK(hipMemcpyAsync(a.data(), dinfo, nb * sizeof(int), hipMemcpyDeviceToHost, s));
    if (ew.lastPath == 1)
        HIPCHK(hipMemcpyAsync(b.data(), tinfo, nb * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    info.assign(nb, 0);
    for (int k = 0; k < nb; ++k) {
        if (ew.lastPath == 0)
            info[k] = a[k] < 0 ? 1 : a[k];   // Jacobi did not converge: reported in 1..n
        else if (ew.lastPath == 1)
            info[k] = a[k] != 0 ? a[k] + n : (b[k] == -2 ? 1 : 0);   // -2: a group barrier timed out
        else
            info[k] = a[k];
    }
}

// Batched generalised eigensolve A x = w B x from the upper triangles
// (what zhegv/dsygv with itype 1, 'V', 'U' compute, interfaces.F90:73-200):
// B = U^H U, C = U^-H A U^-1 (triangular solves on the Hermitian-filled A),
// C = Y diag(w) Y^H, X = U^-1 Y.  A is replaced by the eigenvectors (all n
// columns); B is overwritten.  info[k] > n flags a B(k) that is not positive
// definite, as zhegv does.  Three paths: the one-workgroup LDS solver for
// small n (everything in ksolve.hip); the device path where the tridiagonal
// reduction fits (blocked Cholesky, GEMM triangular solves, trd_eig); else,
// and always when library is set, rocSOLVER (potrf, trsm, heevd).
// defer: enqueue only, with no host synchronisation (graph capture); the
// caller reads the status with read_info once the work has run.  The
// library path synchronises inside and is refused then.
// preInfo: the tridiagonal path's Cholesky already ran (chol_blocked into
// ew.Uc and ew.Dinv, ordered before this call on s) and left its status
// there; B is not read.  preInv (with preInfo): U^-1 was formed too, and the
// transformations C = U^-H A U^-1 and X = U^-1 Y are products with it
// instead of blocked triangular solves.
void gen_eig(EigWork &ew, rocblas_handle h, hipStream_t s, bool cplx, int n, int nb, double *A,
             double *B, double *D, double *E, int *dinfo, int *tinfo, std::vector<int> &info,
             bool library, bool defer = false, const int *preInfo = nullptr,
             const double *preInv = nullptr)
{
    const rocblas_stride st = (rocblas_stride)n * n;
    info.assign(nb, 0);
    if (preInfo && eig_path(ew, n, nb, cplx, library) != 1)
        fail(OLCAO_ERR_STATE, "a Cholesky factor given to a solve that does not take one");
    if (!library && small_hegv_fits(n, cplx)) {
        // one workgroup per matrix, everything in LDS (ksolve.hip)
        ew.lastPath = 0;
        launch_small_hegv(cplx, n, nb, A, B, D, dinfo, s);
        HIPCHK(hipGetLastError());
        if (!defer) read_info(ew, dinfo, tinfo, n, nb, s, info);
        return;
    }
    TrdPlan plan;
    TrdbPlan bplan;
    const bool ldsTrd = !library && !ew.forceBlocked && trd_fits(n, nb, cplx, plan);
    if (ldsTrd || (!library && trdb_plan(n, nb, cplx, bplan))) {
        // blocked Cholesky, GEMM triangular solves, cooperative reduction,
        // divide and conquer, blocked back-transforms (all batched)
        ew.lastPath = 1;
        const int el = cplx ? 2 : 1;
        ew.Dinv.alloc((size_t)n * n * el * nb);
        ew.Y.alloc((size_t)n * n * el * nb);
        ew.Uc.alloc((size_t)n * n * el * nb);
        if (preInfo)
            HIPCHK(hipMemcpyAsync(dinfo, preInfo, nb * sizeof(int), hipMemcpyDeviceToDevice, s));
        else
            chol_blocked(ew, h, s, cplx, n, nb, B, ew.Uc.p, ew.Dinv.p, dinfo);
        if (!defer) {
            // stop before the reduction when some S(k) is not positive
            // definite, as zhegv does
            HIPCHK(hipMemcpyAsync(info.data(), dinfo, nb * sizeof(int), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            for (int k = 0; k < nb; ++k)
                if (info[k] != 0) {
                    for (int q = 0; q < nb; ++q) info[q] = info[q] != 0 ? info[q] + n : 0;
                    return;
                }
        }
        launch_herm_fill(A, n, nb, el, s);
        // with U^-1 = [[Ui_LL, Ui_LR], [0, Ui_RR]] (L the first m1 rows and
        // columns), each product is two GEMMs that skip the zero block:
        //   T = A U^-1:    T(:, L) = A(:, L) Ui_LL,  T(:, R) = A Ui(:, R)
        //   C = U^-H T:    C(L, :) = Ui_LL^H T(L, :), C(R, :) = Ui(:, R)^H T
        //   X = U^-1 Y:    X(L, :) = Ui(L, :) Y,      X(R, :) = Ui_RR Y(R, :)
        const int m1 = n / 2, m2 = n - m1;
        auto at = [&](const double *M, int r, int col) {
            return const_cast<double *>(M) + el * ((size_t)r + (size_t)n * col);
        };
        if (preInv) {
            double *Tm = ew.Y.p;
            gemm_nn(h, cplx, false, false, n, m1, m1, 1.0, A, preInv, 0.0, Tm, n, nb);
            gemm_nn(h, cplx, false, false, n, m2, n, 1.0, A, at(preInv, 0, m1), 0.0, at(Tm, 0, m1), n,
                    nb);
            gemm_nn(h, cplx, true, false, m1, n, m1, 1.0, preInv, Tm, 0.0, A, n, nb);
            gemm_nn(h, cplx, true, false, m2, n, n, 1.0, at(preInv, 0, m1), Tm, 0.0, at(A, m1, 0), n,
                    nb);
        } else {
            solve_left_h(ew, h, s, cplx, n, nb, ew.Uc.p, ew.Dinv.p, A, ew.Y.p);
            solve_right(ew, h, s, cplx, n, nb, ew.Uc.p, ew.Dinv.p, ew.Y.p, A);
        }
        trd_eig(ew, ldsTrd ? &plan : nullptr, ldsTrd ? nullptr : &bplan, h, s, cplx, n, nb, A, D, E,
                tinfo, defer);
        if (preInv) {
            gemm_nn(h, cplx, false, false, m1, n, n, 1.0, preInv, A, 0.0, ew.Y.p, n, nb);
            gemm_nn(h, cplx, false, false, m2, n, m2, 1.0, at(preInv, m1, m1), at(A, m1, 0), 0.0,
                    at(ew.Y.p, m1, 0), n, nb);
        } else {
            solve_back(ew, h, s, cplx, n, nb, ew.Uc.p, ew.Dinv.p, A, ew.Y.p);
        }
        HIPCHK(hipMemcpyAsync(A, ew.Y.p, (size_t)n * n * el * nb * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
        if (!defer) read_info(ew, dinfo, tinfo, n, nb, s, info);
        return;
    }
    if (defer)
        fail(OLCAO_ERR_UNSUPPORTED, "the library eigensolver (n = %d) synchronises inside and is "
                                    "not captured in a graph", n);
    ew.lastPath = 2;
    rocblas_status r;
    if (cplx)
        r = rocsolver_zpotrf_strided_batched(h, rocblas_fill_upper, n, (rocblas_double_complex *)B, n,
                                             st, dinfo, nb);
    else
        r = rocsolver_dpotrf_strided_batched(h, rocblas_fill_upper, n, B, n, st, dinfo, nb);
    if (r != rocblas_status_success) fail(OLCAO_ERR_SOLVER, "potrf status %d", (int)r);
    HIPCHK(hipMemcpyAsync(info.data(), dinfo, nb * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int k = 0; k < nb; ++k)
        if (info[k] != 0) {
            info[k] += n;   // zhegv's convention: B not positive definite
            return;
        }
    launch_herm_fill(A, n, nb, cplx ? 2 : 1, s);
    if (cplx) {
        const rocblas_double_complex one = {1.0, 0.0};
        auto *Az = (rocblas_double_complex *)A;
        auto *Bz = (const rocblas_double_complex *)B;
        r = rocblas_ztrsm_strided_batched(h, rocblas_side_left, rocblas_fill_upper,
                                          rocblas_operation_conjugate_transpose,
                                          rocblas_diagonal_non_unit, n, n, &one, Bz, n, st, Az, n,
                                          st, nb);
        if (r == rocblas_status_success)
            r = rocblas_ztrsm_strided_batched(h, rocblas_side_right, rocblas_fill_upper,
                                              rocblas_operation_none, rocblas_diagonal_non_unit, n,
                                              n, &one, Bz, n, st, Az, n, st, nb);
        if (r != rocblas_status_success) fail(OLCAO_ERR_SOLVER, "trsm status %d", (int)r);
        r = rocsolver_zheevd_strided_batched(h, rocblas_evect_original, rocblas_fill_upper, n, Az, n,
                                             st, D, n, E, n, dinfo, nb);
        if (r != rocblas_status_success) fail(OLCAO_ERR_SOLVER, "heevd status %d", (int)r);
        r = rocblas_ztrsm_strided_batched(h, rocblas_side_left, rocblas_fill_upper,
                                          rocblas_operation_none, rocblas_diagonal_non_unit, n, n,
                                          &one, Bz, n, st, Az, n, st, nb);
    } else {
        const double one = 1.0;
        r = rocblas_dtrsm_strided_batched(h, rocblas_side_left, rocblas_fill_upper,
                                          rocblas_operation_transpose, rocblas_diagonal_non_unit, n,
                                          n, &one, B, n, st, A, n, st, nb);
        if (r == rocblas_status_success)
            r = rocblas_dtrsm_strided_batched(h, rocblas_side_right, rocblas_fill_upper,
                                              rocblas_operation_none, rocblas_diagonal_non_unit, n,
                                              n, &one, B, n, st, A, n, st, nb);
        if (r != rocblas_status_success) fail(OLCAO_ERR_SOLVER, "trsm status %d", (int)r);
        r = rocsolver_dsyevd_strided_batched(h, rocblas_evect_original, rocblas_fill_upper, n, A, n,
                                             st, D, n, E, n, dinfo, nb);
        if (r != rocblas_status_success) fail(OLCAO_ERR_SOLVER, "syevd status %d", (int)r);
        r = rocblas_dtrsm_strided_batched(h, rocblas_side_left, rocblas_fill_upper,
                                          rocblas_operation_none, rocblas_diagonal_non_unit, n, n,
                                          &one, B, n, st, A, n, st, nb);
    }
    if (r != rocblas_status_success) fail(OLCAO_ERR_SOLVER, "trsm status %d", (int)r);
    HIPCHK(hipMemcpyAsync(info.data(), dinfo, nb * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
}

// Batched generalised eigensolve of H, S (upper) on the device, spin q:
// solve_enqueue queues the copies and the solve (defer: with no host
// synchronisation, for graph capture; info is then read by the caller with
// read_info), solve_finish checks the status and returns the eigenpairs.
// pre: prefactor_s ran after S was last written, and the solve uses its
// factor (the stream waits for it) instead of factoring S again.
void solve_enqueue(olcao_ctx &c, int q, const double *Hsrc, const double *Ssrc, bool defer,
                   std::vector<int> &info, bool pre = false)
{
    hipStream_t s = c.stream;
    const int n = c.valeDim, nk = c.nk;
    // spin q's eigenvectors stay in Aw (q = 0) or Aw2 (q = 1) for the density
    DBuf<double> &Aq = q ? c.Aw2 : c.Aw;
    c.vecsValid[q] = false;
    Aq.alloc(c.matDoubles());
    c.Bw.alloc(c.matDoubles());
    c.D[q].alloc((size_t)n * nk);
    c.E[q].alloc((size_t)n * nk);
    c.dInfo[q].alloc(nk);
    c.tInfo[q].alloc(nk);
    HIPCHK(hipMemcpyAsync(Aq.p, Hsrc, c.matDoubles() * 8, hipMemcpyDeviceToDevice, s));
    if (pre)
        HIPCHK(hipStreamWaitEvent(s, c.evChol, 0));
    else
        HIPCHK(hipMemcpyAsync(c.Bw.p, Ssrc, c.matDoubles() * 8, hipMemcpyDeviceToDevice, s));
    gen_eig(c.eig, c.blas, s, !c.gamma, n, nk, Aq.p, c.Bw.p, c.D[q].p, c.E[q].p, c.dInfo[q].p,
            c.tInfo[q].p, info, c.eigLibrary, defer, pre ? c.dCholInfo.p : nullptr,
            pre && c.preInv ? c.eig.Ui.p : nullptr);
}

// Where the solve takes the tridiagonal path, its Cholesky of S (which does
// not depend on H) is started on the last side stream as soon as S is
// complete, so that it runs beside the Hamiltonian sweep; every spin's
// solve then uses the one factor (solve_enqueue with pre).  The side
// stream's own rocBLAS handle keeps its GEMMs off the context handle.  The
// factor is the same computation on the same data as inside gen_eig, so the
// results are unchanged.  Returns whether it was started.
bool prefactor_s(olcao_ctx &c)
{
    const int n = c.valeDim, nk = c.nk;
    const bool cplx = !c.gamma;
    if (eig_path(c.eig, n, nk, cplx, c.eigLibrary) != 1) return false;
    const size_t m = (size_t)n * n * (cplx ? 2 : 1) * nk;
    if (!c.blasSide) {
        if (rocblas_create_handle(&c.blasSide) != rocblas_status_success) {
            c.blasSide = nullptr;
            fail(OLCAO_ERR_HIP, "rocblas_create_handle failed");
        }
        rocblas_set_stream(c.blasSide, c.side[olcao_ctx::kSide - 1]);
        rocblas_set_atomics_mode(c.blasSide, rocblas_atomics_not_allowed);
    }
    c.Bw.alloc(c.matDoubles());
    c.eig.Dinv.alloc(m);
    c.eig.Y.alloc(m);
    c.eig.Uc.alloc(m);
    c.dCholInfo.alloc(nk);
    hipStream_t t = c.side[olcao_ctx::kSide - 1];
    HIPCHK(hipEventRecord(c.evS, c.stream));
    HIPCHK(hipStreamWaitEvent(t, c.evS, 0));
    HIPCHK(hipMemcpyAsync(c.Bw.p, c.S.p, c.matDoubles() * 8, hipMemcpyDeviceToDevice, t));
    chol_blocked(c.eig, c.blasSide, t, cplx, n, nk, c.Bw.p, c.eig.Uc.p, c.eig.Dinv.p, c.dCholInfo.p);
    // where the reduction is the LDS-resident one (n <= 896), also U^-1 =
    // U^-1 I by the blocked back substitution (Bw, free now, holds I), so
    // that the solve's three triangular transformations become GEMMs
    TrdPlan plan;
    c.preInv = !c.eig.forceBlocked && trd_fits(n, nk, cplx, plan);
    if (c.preInv) {
        c.eig.Ui.alloc(m);
        launch_identity(c.Bw.p, n, nk, cplx ? 2 : 1, t);
        HIPCHK(hipGetLastError());
        solve_back(c.eig, c.blasSide, t, cplx, n, nk, c.eig.Uc.p, c.eig.Dinv.p, c.Bw.p, c.eig.Ui.p);
    }
    HIPCHK(hipEventRecord(c.evChol, t));
    return true;
}

void solve_finish(olcao_ctx &c, int q, int numStates, const std::vector<int> &info,
                  double *evals, double *evecs)
{
    hipStream_t s = c.stream;
    const int n = c.valeDim, nk = c.nk;
    DBuf<double> &Aq = q ? c.Aw2 : c.Aw;
    std::vector<double> w((size_t)n * nk);
    HIPCHK(hipMemcpyAsync(w.data(), c.D[q].p, w.size() * 8, hipMemcpyDeviceToHost, s));
    if (evecs) {
        DBuf<double> tmp;
        const int cp = c.gamma ? 1 : 2;
        tmp.alloc((size_t)nk * numStates * n * cp);
        launch_extract_evecs(Aq.p, n, nk, numStates, cp, tmp.p, s);
        HIPCHK(hipMemcpyAsync(evecs, tmp.p, tmp.n * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    HIPCHK(hipStreamSynchronize(s));
    if (c.eig.trdTimed) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c.eig.trdEv[0], c.eig.trdEv[1]));
        c.trdMs += ms;
        c.eig.trdTimed = false;
    }
    for (int k = 0; k < nk; ++k) {
        if (info[k] > n) fail(OLCAO_ERR_NOT_SPD, "S(k) not positive definite at k=%d (info=%d)", k, info[k]);
        if (info[k] != 0) fail(OLCAO_ERR_SOLVER, "eigensolver did not converge at k=%d (info=%d)", k, info[k]);
    }
    if (evals)
        for (int k = 0; k < nk; ++k)
            std::memcpy(evals + (size_t)k * numStates, w.data() + (size_t)k * n, numStates * sizeof(double));
    c.hEv[q].swap(w);
    c.vecsValid[q] = true;
}

void solve_device(olcao_ctx &c, int q, int numStates, const double *Hsrc, const double *Ssrc,
                  double *evals, double *evecs, bool pre = false)
{
    if (numStates <= 0 || numStates > c.valeDim)
        fail(OLCAO_ERR_INPUT, "numStates must lie in 1..%d", c.valeDim);
    std::vector<int> info;
    solve_enqueue(c, q, Hsrc, Ssrc, false, info, pre);
    solve_finish(c, q, numStates, info, evals, evecs);
}

thread_local std::string g_noctx_err;   // message of context-free calls
std::atomic<int> g_streamRotation{0};   // contexts created so far (stream placement)

template <class F>
olcao_status guard(olcao_ctx *c, F &&f)
{
    std::string &msg = c ? c->err : g_noctx_err;
    // allocations inside this entry point count against c's graph
    struct EpochScope {
        std::atomic<unsigned long long> *prev;
        explicit EpochScope(olcao_ctx *c) : prev(tl_allocEpoch) { tl_allocEpoch = c ? &c->allocEpoch : nullptr; }
        ~EpochScope() { tl_allocEpoch = prev; }
    } scope(c);
    try {
        f();
        msg.clear();
        return OLCAO_OK;
    } catch (const Fail &e) {
        msg = e.msg;
        return e.st;
    } catch (const std::exception &e) {
        msg = e.what();
        return OLCAO_ERR_ALLOC;
    }
}

}  // namespace

void olcao::set_noctx_error(const std::string &msg) { g_noctx_err = msg; }

// ===========================================================================
extern "C" {

olcao_status olcao_ctx_create(const olcao_system *sys, const olcao_options *opt, olcao_ctx **out)
{
    if (!sys || !out) return OLCAO_ERR_INPUT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_noctx_err = "no GPU visible";
        return OLCAO_ERR_NO_DEVICE;
    }
    olcao_ctx *c = new olcao_ctx();
    olcao_status st = guard(c, [&] {
        c->precision = opt ? opt->precision : 0;
        c->eigLibrary = opt && opt->eigensolver == OLCAO_EIG_LIBRARY;
        c->eig.forceBlocked = opt && opt->eigensolver == OLCAO_EIG_BLOCKED;
        c->blochMfma = !(opt && opt->bloch == OLCAO_BLOCH_VALU);
        c->blochSplit = !(opt && opt->bloch == OLCAO_BLOCH_MFMA64);
        c->useGraph = opt && opt->graph != 0;
        if (opt && opt->chunkItems) {
            if (opt->chunkItems < 16 || opt->chunkItems > (1 << 20))
                fail(OLCAO_ERR_INPUT, "chunkItems must lie in 16..1048576 (0: default)");
            c->chunkItems = opt->chunkItems;
        }
        if (opt && opt->segPairs) {
            if (opt->segPairs < 1 || opt->segPairs > 64)
                fail(OLCAO_ERR_INPUT, "segPairs must lie in 1..64 (0: default)");
            c->segPairs = opt->segPairs;
        }
        if (opt && (opt->bloch < OLCAO_BLOCH_AUTO || opt->bloch > OLCAO_BLOCH_MFMA64))
            fail(OLCAO_ERR_INPUT, "bloch must be OLCAO_BLOCH_AUTO, _VALU, _MFMA or _MFMA64");
        if (c->precision != OLCAO_PREC_FP64 && c->precision != OLCAO_PREC_MIXED)
            fail(OLCAO_ERR_INPUT, "precision must be OLCAO_PREC_FP64 or OLCAO_PREC_MIXED");
        // The runtime spreads a process's streams over a few hardware
        // queues in creation order, so the main stream is created after a
        // rotating number of side streams: the main streams of successive
        // contexts land on different hardware queues and their work can run
        // at the same time (with four streams per context they all shared
        // one queue, and a batch of cells ran no faster than one by one).
        const int rot = g_streamRotation.fetch_add(1) % (olcao_ctx::kSide + 1);
        for (int q = 0; q < rot; ++q) HIPCHK(hipStreamCreate(&c->side[q]));
        HIPCHK(hipStreamCreate(&c->stream));
        for (int q = rot; q < olcao_ctx::kSide; ++q) HIPCHK(hipStreamCreate(&c->side[q]));
        for (auto &e : c->ev) HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventCreateWithFlags(&c->evFork, hipEventDisableTiming));
        for (auto &e : c->eig.trdEv) HIPCHK(hipEventCreate(&e));
        for (auto &e : c->evJoin) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        if (rocblas_create_handle(&c->blas) != rocblas_status_success)
            fail(OLCAO_ERR_HIP, "rocblas_create_handle failed");
        rocblas_set_stream(c->blas, c->stream);
        // no atomics in the library's GEMMs and in rocSOLVER's BLAS calls:
        // the same inputs give the same bits whatever else shares the device
        rocblas_set_atomics_mode(c->blas, rocblas_atomics_not_allowed);
        HIPCHK(hipEventCreateWithFlags(&c->evS, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c->evChol, hipEventDisableTiming));
        setup_system(*c, *sys);
        build_lists(*c);
        upload(*c);
        build_image_lists(*c);
    });
    if (st != OLCAO_OK) {
        g_noctx_err = c->err;
        olcao_ctx_destroy(c);
        return st;
    }
    *out = c;
    return OLCAO_OK;
}

void olcao_ctx_destroy(olcao_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto q : c->side)
        if (q) (void)hipStreamSynchronize(q);
    if (c->graphExec) (void)hipGraphExecDestroy(c->graphExec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->blas) rocblas_destroy_handle(c->blas);
    if (c->blasSide) rocblas_destroy_handle(c->blasSide);
    if (c->evS) (void)hipEventDestroy(c->evS);
    if (c->evChol) (void)hipEventDestroy(c->evChol);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->evFork) (void)hipEventDestroy(c->evFork);
    for (auto e : c->eig.trdEv)
        if (e) (void)hipEventDestroy(e);
    for (auto e : c->evJoin)
        if (e) (void)hipEventDestroy(e);
    for (auto q : c->side)
        if (q) (void)hipStreamDestroy(q);
    hipStream_t s = c->stream;
    delete c;
    if (s) (void)hipStreamDestroy(s);
}

const char *olcao_last_error(const olcao_ctx *c)
{
    return c ? c->err.c_str() : g_noctx_err.c_str();
}

olcao_status olcao_get_dims(const olcao_ctx *c, long long dims[8])
{
    if (!c || !dims) return OLCAO_ERR_INPUT;
    dims[0] = c->valeDim; dims[1] = c->coreDim; dims[2] = c->nk; dims[3] = c->potDim;
    dims[4] = (long long)c->hCellSizes.size(); dims[5] = (long long)c->hPcPair.size();
    dims[6] = (long long)c->hPpPc.size(); dims[7] = (long long)c->hPairI.size();
    return OLCAO_OK;
}

olcao_status olcao_get_setup(const olcao_ctx *c, double *cellDims, double *cellSizes, double *kCart)
{
    if (!c) return OLCAO_ERR_INPUT;
    if (cellDims) std::memcpy(cellDims, c->hCellDims.data(), c->hCellDims.size() * 8);
    if (cellSizes) std::memcpy(cellSizes, c->hCellSizes.data(), c->hCellSizes.size() * 8);
    if (kCart) std::memcpy(kCart, c->hKcart.data(), c->hKcart.size() * 8);
    return OLCAO_OK;
}

olcao_status olcao_gauss_overlap(olcao_ctx *c, int op, double *vv, double *cv, double *cc)
{
    if (!c) return OLCAO_ERR_INPUT;
    return guard(c, [&] {
        if (op < 0 || op > 2) fail(OLCAO_ERR_INPUT, "op must be 0 (overlap), 1 (kinetic) or 2 (mass velocity)");
        if (op == 2 && !c->rel) fail(OLCAO_ERR_STATE, "the mass-velocity term needs rel = 1");
        ensure_mats(*c);
        if (op == 0) {
            overlap_sweep(*c);
            copy_out(*c, c->S.p, vv, c->matDoubles());
            if (c->coreDim) {
                copy_out(*c, c->CV_S.p, cv, c->cvDoubles());
                copy_out(*c, c->CC_S.p, cc, c->ccDoubles());
            }
        } else {
            operator_sweep(*c, op == 1 ? SW_KE : SW_MV, 0, 0, c->H[0].p, c->CV_H[0].p, c->CC_H[0].p);
            c->haveH = false;
            copy_out_spin(*c, 0, vv, cv, cc);
        }
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

olcao_status olcao_gauss_overlap_np(olcao_ctx *c, double *vv, double *cv, double *cc)
{
    if (!c) return OLCAO_ERR_INPUT;
    return guard(c, [&] {
        ensure_mats(*c);
        operator_sweep(*c, SW_NP, 0, 0, c->H[0].p, c->CV_H[0].p, c->CC_H[0].p);
        c->haveH = false;
        copy_out_spin(*c, 0, vv, cv, cc);
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

// The three Cartesian components of the dipole (about `center`) or
// momentum matrices, each as an operator sweep (orthogonalised against the
// core with its own core blocks, ortho3Terms integrals3Terms.F90:1388),
// copied out component-major.
static void vector_sweeps(olcao_ctx &c, Sweep kind, const double *center, double *vv, double *cv,
                          double *cc)
{
    ensure_mats(c);
    if (center)
        for (int x = 0; x < 3; ++x) c.dev.dipC[x] = center[x];
    for (int comp = 0; comp < 3; ++comp) {
        operator_sweep(c, kind, comp, 0, c.H[0].p, c.CV_H[0].p, c.CC_H[0].p);
        copy_out(c, c.H[0].p, vv ? vv + (size_t)comp * c.matDoubles() : nullptr, c.matDoubles());
        if (c.coreDim) {
            copy_out(c, c.CV_H[0].p, cv ? cv + (size_t)comp * c.cvDoubles() : nullptr, c.cvDoubles());
            copy_out(c, c.CC_H[0].p, cc ? cc + (size_t)comp * c.ccDoubles() : nullptr, c.ccDoubles());
        }
        HIPCHK(hipStreamSynchronize(c.stream));
    }
    c.haveH = false;
}

olcao_status olcao_gauss_overlap_dm(olcao_ctx *c, const double *center, double *vv, double *cv,
                                    double *cc)
{
    if (!c || !center) return OLCAO_ERR_INPUT;
    return guard(c, [&] { vector_sweeps(*c, SW_DM, center, vv, cv, cc); });
}

olcao_status olcao_gauss_overlap_mm(olcao_ctx *c, double *vv, double *cv, double *cc)
{
    if (!c) return OLCAO_ERR_INPUT;
    return guard(c, [&] { vector_sweeps(*c, SW_MM, nullptr, vv, cv, cc); });
}

// gaussKOverlap (integrals3Terms.F90:956-1384) for the three deltaK =
// recipVectors(:, axis) / numAxialKPoints(axis): per axis the sweep of the
// blocks' real parts plus the sweep of minus their imaginary parts under the
// Bloch sum's -i factor (the complex block's Bloch sum, scatter and core
// blocks), then ortho3Terms' full-matrix orthogonalisation without zeroing.
olcao_status olcao_gauss_k_overlap(olcao_ctx *c, const int numAxial[3], double *vv, double *cv,
                                   double *cc)
{
    return guard(c, [&] {
        if (!numAxial || numAxial[0] < 1 || numAxial[1] < 1 || numAxial[2] < 1)
            fail(OLCAO_ERR_INPUT, "numAxialKPoints must be positive");
        if (c->gamma)
            fail(OLCAO_ERR_UNSUPPORTED, "the k-overlap is complex; a Gamma context has no such sweep");
        ensure_mats(*c);
        if (c->coreDim && !c->haveS) overlap_sweep(*c);
        c->Kvv.alloc(c->matDoubles());
        if (c->coreDim) {
            c->Kcv.alloc(c->cvDoubles());
            c->Kcc.alloc(c->ccDoubles());
        }
        // b_axis = 2 pi (a_j x a_k) / V, the rows of the inverse lattice matrix times 2 pi
        for (int axis = 0; axis < 3; ++axis) {
            for (int x = 0; x < 3; ++x)
                c->dev.kq[x] = 2.0 * kPiHost * c->dev.inv[3 * axis + x] / numAxial[axis];
            sweep(*c, SW_KO, 0, 0, c->H[0].p, c->CV_H[0].p, c->CC_H[0].p);
            sweep(*c, SW_KO, 1, 0, c->Kvv.p, c->Kcv.p, c->Kcc.p);
            launch_add(c->H[0].p, c->Kvv.p, (long long)c->matDoubles(), c->stream);
            if (c->coreDim) {
                launch_add(c->CV_H[0].p, c->Kcv.p, (long long)c->cvDoubles(), c->stream);
                launch_add(c->CC_H[0].p, c->Kcc.p, (long long)c->ccDoubles(), c->stream);
            }
            HIPCHK(hipGetLastError());
            ortho(*c, c->H[0].p, c->CV_H[0].p, c->CC_H[0].p, false, false);
            copy_out(*c, c->H[0].p, vv ? vv + (size_t)axis * c->matDoubles() : nullptr, c->matDoubles());
            if (c->coreDim) {
                copy_out(*c, c->CV_H[0].p, cv ? cv + (size_t)axis * c->cvDoubles() : nullptr,
                         c->cvDoubles());
                copy_out(*c, c->CC_H[0].p, cc ? cc + (size_t)axis * c->ccDoubles() : nullptr,
                         c->ccDoubles());
            }
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        c->haveH = false;
    });
}

// computeForceIntg (forces.F90:60-898) for the coefficients coef
// [spin][potDim] and totalEnergy: per spin q and Cartesian component x
// (output slot 3 q + x), the Bloch sums of the ket derivatives d/dB_x of
// minus the potential terms' three-Gaussian blocks times the coefficients,
// plus the nuclear blocks times Z, minus the kinetic block, and
// -2 totalEnergy times the momentum block (+2 totalEnergy d/dB_x of the
// overlap), over the fused sweep's screened images; saved in full in a
// complex build (saveCurrentPair saveFull = 1: the block and its dagger, a
// diagonal atom block as its own dagger), as the upper triangle at Gamma
// (saveCurrentPairGamma); no orthogonalisation.  Bases with f orbitals are
// refused: the f-f nuclear derivative needs F_7, which the reference's Boys
// function does not give (its dnuclear3CIntgCB is 1e-2 off there).
olcao_status olcao_force_integrals(olcao_ctx *c, const double *coef, double totalEnergy,
                                   double *vv, double *cv, double *cc)
{
    if (!c || !coef) return OLCAO_ERR_INPUT;
    return guard(c, [&] {
        for (int k = 0; k < 16; ++k)
            if ((k / 4 == 3 || k % 4 == 3) && c->cls.count[k] > 0)
                fail(OLCAO_ERR_UNSUPPORTED,
                     "force integrals of f orbitals (the f-f nuclear derivative needs F_7)");
        ensure_mats(*c);
        HIPCHK(hipMemcpyAsync(c->dCoef.p, coef, (size_t)c->potDim * c->spin * 8,
                              hipMemcpyHostToDevice, c->stream));
        size_t tot = 0, off[16] = {};
        for (int k = 0; k < 16; ++k) {
            const int la = k / 4, lb = k % 4, L = la + lb + 1;
            off[k] = tot;
            if (la <= 2 && lb <= 2)
                tot += (size_t)((L + 1) * (L + 2) * (L + 3) / 6) * c->cls.nChunk[k];
        }
        c->dWd.reserve(std::max<size_t>(tot, 1));
        for (int k = 0; k < 16; ++k) c->cls.Wd[k] = c->dWd.p + off[k];
        c->dev.fE2 = 2.0 * totalEnergy;
        const int mode = c->gamma ? 0 : 2;
        for (int q = 0; q < c->spin; ++q) {
            three_center_force_forked(*c, c->dCoef.p + (size_t)q * c->potDim, -1);
            for (int x = 0; x < 3; ++x) {
                launch_two_center_fd(c->dev, c->cls, x, c->dBlocks.p, 0x0777u, c->stream);
                three_center_force_forked(*c, nullptr, x);
                contract_scatter(*c, c->H[0].p, c->CV_H[0].p, c->CC_H[0].p, mode);
                const size_t slot = (size_t)q * 3 + x;
                copy_out(*c, c->H[0].p, vv ? vv + slot * c->matDoubles() : nullptr, c->matDoubles());
                if (c->coreDim) {
                    copy_out(*c, c->CV_H[0].p, cv ? cv + slot * c->cvDoubles() : nullptr,
                             c->cvDoubles());
                    copy_out(*c, c->CC_H[0].p, cc ? cc + slot * c->ccDoubles() : nullptr,
                             c->ccDoubles());
                }
                HIPCHK(hipStreamSynchronize(c->stream));
            }
        }
        c->haveH = false;
    });
}

olcao_status olcao_elec_pot_gauss_overlap(olcao_ctx *c, int term, double *vv, double *cv, double *cc)
{
    if (!c) return OLCAO_ERR_INPUT;
    return guard(c, [&] {
        if (term < 0 || term >= c->potDim) fail(OLCAO_ERR_INPUT, "term out of range");
        ensure_mats(*c);
        operator_sweep(*c, SW_EP, term, 0, c->H[0].p, c->CV_H[0].p, c->CC_H[0].p);
        c->haveH = false;
        copy_out_spin(*c, 0, vv, cv, cc);
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

olcao_status olcao_gauss_overlap_ham_pscf(olcao_ctx *c, const double *coef, double *vv, double *cv, double *cc)
{
    if (!c || !coef) return OLCAO_ERR_INPUT;
    return guard(c, [&] {
        ensure_mats(*c);
        HIPCHK(hipMemcpyAsync(c->dCoef.p, coef, (size_t)c->potDim * c->spin * 8,
                              hipMemcpyHostToDevice, c->stream));
        for (int q = 0; q < c->spin; ++q) {
            operator_sweep(*c, SW_HAM, 0, q, c->H[q].p, c->CV_H[q].p, c->CC_H[q].p);
            copy_out_spin(*c, q, vv, cv, cc);
        }
        c->haveH = true;
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

// Solve every spin's H against S: evals [spin][nk][numStates], evecs
// [spin][nk][numStates][valeDim].  pre: S's factor from prefactor_s.
static void solve_spins(olcao_ctx *c, int numStates, double *evals, double *evecs,
                        bool pre = false)
{
    const size_t ev = (size_t)c->nk * numStates;
    const size_t vec = ev * c->valeDim * (c->gamma ? 1 : 2);
    for (int q = 0; q < c->spin; ++q)
        solve_device(*c, q, numStates, c->H[q].p, c->S.p, evals ? evals + q * ev : nullptr,
                     evecs ? evecs + q * vec : nullptr, pre);
}

olcao_status olcao_secular_eqn_pscf(olcao_ctx *c, int numStates, double *evals, double *evecs)
{
    if (!c) return OLCAO_ERR_INPUT;
    return guard(c, [&] {
        if (!c->haveS || !c->haveH)
            fail(OLCAO_ERR_STATE, "build S (olcao_gauss_overlap op 0) and H (olcao_gauss_overlap_ham_pscf) first");
        solve_spins(c, numStates, evals, evecs);
    });
}

olcao_status olcao_setup_scf(olcao_ctx *c)
{
    if (!c) return OLCAO_ERR_INPUT;
    return guard(c, [&] {
        ensure_mats(*c);
        const size_t m = c->matDoubles();
        overlap_sweep(*c);
        double *cv = c->CV_H[0].p, *cc = c->CC_H[0].p;   // scratch for the core blocks
        c->KEm.alloc(m);
        operator_sweep(*c, SW_KE, 0, 0, c->KEm.p, cv, cc);
        c->NPm.alloc(m);
        operator_sweep(*c, SW_NP, 0, 0, c->NPm.p, cv, cc);
        if (c->rel) {
            c->MVm.alloc(m);
            operator_sweep(*c, SW_MV, 0, 0, c->MVm.p, cv, cc);
        }
        c->Vm.resize(c->potDim);
        std::vector<double *> ptrs(c->potDim);
        for (int j = 0; j < c->potDim; ++j) {
            c->Vm[j].alloc(m);
            operator_sweep(*c, SW_EP, j, 0, c->Vm[j].p, cv, cc);
            ptrs[j] = c->Vm[j].p;
        }
        c->dVptr.alloc(std::max(c->potDim, 1));
        HIPCHK(hipMemcpyAsync(c->dVptr.p, ptrs.data(), ptrs.size() * sizeof(double *),
                              hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        c->haveH = false;
        c->haveScf = true;
    });
}

olcao_status olcao_get_eigen(const olcao_ctx *c, int spin, int numStates, double *evals,
                             double *evecs)
{
    if (!c) return OLCAO_ERR_INPUT;
    olcao_ctx *m = const_cast<olcao_ctx *>(c);
    return guard(m, [&] {
        const int n = c->valeDim, nk = c->nk, cp = c->gamma ? 1 : 2;
        if (spin < 0 || spin >= c->spin) fail(OLCAO_ERR_INPUT, "spin must lie in 0..%d", c->spin - 1);
        if (numStates <= 0 || numStates > n) fail(OLCAO_ERR_INPUT, "numStates must lie in 1..%d", n);
        if (!c->vecsValid[spin]) fail(OLCAO_ERR_STATE, "solve first");
        if (evals)
            for (int k = 0; k < nk; ++k)
                std::memcpy(evals + (size_t)k * numStates, c->hEv[spin].data() + (size_t)k * n,
                            numStates * sizeof(double));
        if (evecs) {
            DBuf<double> tmp;
            tmp.alloc((size_t)nk * numStates * n * cp);
            launch_extract_evecs((spin ? c->Aw2 : c->Aw).p, n, nk, numStates, cp, tmp.p, c->stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(evecs, tmp.p, tmp.n * 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
    });
}

olcao_status olcao_get_flags(const olcao_ctx *c, int flags[4])
{
    if (!c || !flags) return OLCAO_ERR_INPUT;
    flags[0] = c->gamma ? 1 : 0;
    flags[1] = c->rel;
    flags[2] = c->spin;
    flags[3] = c->precision;
    return OLCAO_OK;
}

olcao_status olcao_get_scf_term(const olcao_ctx *c, int which, int term, double *out)
{
    if (!c || !out) return OLCAO_ERR_INPUT;
    olcao_ctx *m = const_cast<olcao_ctx *>(c);
    return guard(m, [&] {
        if (!c->haveScf) fail(OLCAO_ERR_STATE, "call olcao_setup_scf first");
        const double *src = nullptr;
        size_t count = c->matDoubles();
        switch (which) {
        case 0: src = c->S.p; break;
        case 1: src = c->KEm.p; break;
        case 2: src = c->rel ? c->MVm.p : nullptr; break;
        case 3: src = c->NPm.p; break;
        case 4:
            if (term < 0 || term >= c->potDim) fail(OLCAO_ERR_INPUT, "term must lie in 0..%d", c->potDim - 1);
            src = c->Vm[term].p;
            break;
        case 5: src = c->CV_S.p; count = c->cvDoubles(); break;
        default: fail(OLCAO_ERR_INPUT, "which must lie in 0..5");
        }
        if (!src || count == 0) fail(OLCAO_ERR_STATE, "that matrix is not part of this system");
        HIPCHK(hipMemcpyAsync(out, src, count * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

olcao_status olcao_secular_eqn_scf(olcao_ctx *c, const double *coef, int numStates, double *evals, double *evecs)
{
    if (!c || !coef) return OLCAO_ERR_INPUT;
    return guard(c, [&] {
        if (!c->haveScf) fail(OLCAO_ERR_STATE, "call olcao_setup_scf first");
        ensure_mats(*c);
        HIPCHK(hipMemcpyAsync(c->dCoef.p, coef, (size_t)c->potDim * c->spin * 8,
                              hipMemcpyHostToDevice, c->stream));
        for (int q = 0; q < c->spin; ++q)
            launch_assemble((long long)c->matDoubles(), c->valeDim, c->NPm.p, c->KEm.p,
                            c->rel ? c->MVm.p : nullptr, c->dVptr.p,
                            c->dCoef.p + (size_t)q * c->potDim, c->potDim, c->H[q].p,
                            c->gamma ? 1 : 2, c->stream);
        HIPCHK(hipGetLastError());
        c->haveH = true;
        solve_spins(c, numStates, evals, evecs);
    });
}

olcao_status olcao_vale_charge(olcao_ctx *c, int numStates, const double *occupancy,
                               double *electronEnergy, double *traces, double *potRho)
{
    if (!c || !occupancy) return OLCAO_ERR_INPUT;
    return guard(c, [&] {
        const int n = c->valeDim, nk = c->nk, sp = c->spin, cp = c->gamma ? 1 : 2;
        if (!c->haveScf) fail(OLCAO_ERR_STATE, "call olcao_setup_scf first (the term matrices)");
        for (int q = 0; q < sp; ++q)
            if (!c->vecsValid[q]) fail(OLCAO_ERR_STATE, "solve for the eigenvectors first");
        if (numStates <= 0 || numStates > n) fail(OLCAO_ERR_INPUT, "numStates must lie in 1..%d", n);
        hipStream_t s = c->stream;
        // populations arrive in electronPopulation order [nk][spin][state]
        // (valeCharge.F90:160-169); a state whose populations sum below
        // smallThresh in magnitude is skipped (:214-216)
        const double smallThresh = 1e-8;
        std::vector<double> f((size_t)sp * nk * numStates, 0.0);   // [spin][nk][state]
        double ee[2] = {0.0, 0.0};
        for (int k = 0; k < nk; ++k)
            for (int j = 0; j < numStates; ++j) {
                double sum = 0.0;
                for (int q = 0; q < sp; ++q) sum += std::fabs(occupancy[((size_t)k * sp + q) * numStates + j]);
                if (sum < smallThresh) continue;
                for (int q = 0; q < sp; ++q) {
                    const double fq = occupancy[((size_t)k * sp + q) * numStates + j];
                    f[((size_t)q * nk + k) * numStates + j] = fq;
                    ee[q] += fq * c->hEv[q][(size_t)k * n + j];
                }
            }
        // rho(k) = X diag(f) X^H per spin (the zher / dsyr updates, :228-232, :273-277)
        c->dOcc.upload(f, s);
        c->Ysc.alloc(c->matDoubles());
        const int nb = c->gamma ? 1 : nk;
        for (int q = 0; q < sp; ++q) {
            const double *X = q ? c->Aw2.p : c->Aw.p;
            c->Rho[q].alloc(c->matDoubles());
            launch_scale_states(X, c->dOcc.p + (size_t)q * nk * numStates, n, numStates, nb, cp,
                                c->Ysc.p, s);
            HIPCHK(hipGetLastError());
            gemm_nn(c->blas, !c->gamma, false, true, n, n, numStates, 1.0, c->Ysc.p, X, 0.0,
                    c->Rho[q].p, n, nb);
        }
        // traces against S, NP, KE, MV and every V_j (:330-418), for the
        // total and, at spin 2, the difference density (:306-314)
        std::vector<const double *> mats = {c->S.p, c->NPm.p, c->KEm.p, c->rel ? c->MVm.p : c->KEm.p};
        for (int j = 0; j < c->potDim; ++j) mats.push_back(c->Vm[j].p);
        const int nMat = (int)mats.size();
        c->dTrPtr.upload(mats, s);
        c->dTrPart.alloc((size_t)512 * nMat);
        c->dTrOut.alloc((size_t)sp * nMat);
        for (int t = 0; t < sp; ++t) {
            launch_traces(c->dTrPtr.p, nMat, c->Rho[0].p, sp == 2 ? c->Rho[1].p : nullptr,
                          t == 0 ? 1.0 : -1.0, n, nb, cp, c->dTrPart.p, c->dTrOut.p + (size_t)t * nMat, s);
            HIPCHK(hipGetLastError());
        }
        std::vector<double> out((size_t)sp * nMat);
        HIPCHK(hipMemcpyAsync(out.data(), c->dTrOut.p, out.size() * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (int t = 0; t < sp; ++t) {
            const double *o = out.data() + (size_t)t * nMat;
            if (electronEnergy) electronEnergy[t] = ee[t];
            if (traces) {
                traces[4 * t + 0] = o[0];
                traces[4 * t + 1] = o[1];
                traces[4 * t + 2] = o[2];
                traces[4 * t + 3] = c->rel ? o[3] : 0.0;
            }
            if (potRho)
                for (int j = 0; j < c->potDim; ++j) potRho[(size_t)t * c->potDim + j] = o[4 + j];
        }
    });
}

// First half of a Mode-B iteration: potential coefficients in, overlap
// sweep, fused Hamiltonian sweep per spin, contraction, Bloch sums and
// orthogonalisation, all queued on the context's streams.  (Starting S(k)'s
// Cholesky factor on a stream of its own beside the Hamiltonian sweep was
// measured: the sweep slowed by as much as the solve gained.)
// The device work of the build, everything after the coefficient upload;
// events: record the stage events (not inside a graph capture).
// prefactor: start the solve's Cholesky of S beside the Hamiltonian sweep
// (prefactor_s); returns whether it was started.
static bool iteration_device(olcao_ctx &c, bool events, bool prefactor)
{
    hipStream_t s = c.stream;
    auto mark = [&](int e) {
        if (events) HIPCHK(hipEventRecord(c.ev[e], s));
    };
    overlap_sweep(c);
    mark(1);
    const bool pre = prefactor && prefactor_s(c);
    struct Restore {
        olcao_ctx &c;
        ~Restore() { c.forkStreams = 1 + olcao_ctx::kSide; }
    } restore{c};
    if (pre) c.forkStreams = olcao_ctx::kSide;
    // events: 0 start, 1 overlap done; per spin q: 2+3q before the
    // three-centre sweep, 3+3q after it, 4+3q after contraction, Bloch
    // sum and orthogonalisation; 8 after the solve
    for (int q = 0; q < c.spin; ++q) {
        // kinetic blocks: here for every class when rel = 1 (the
        // mass-velocity term follows them); otherwise only for the
        // classes whose finish kernels do not add them
        two_center_forked(c, OP_KE, 0, c.rel ? 0xffffu : ~kFusedKEClasses & 0xffffu);
        if (c.rel) two_center_forked(c, OP_MV, 1);
        launch_fill(reinterpret_cast<double *>(c.dCounters.p), 32, 0.0, s);
        mark(2 + 3 * q);
        three_center_forked(c, K3_HAM, 0, 0, c.dCoef.p + (size_t)q * c.potDim, c.rel ? 1 : 1 | kAccAddKE);
        mark(3 + 3 * q);
        contract_scatter(c, c.H[q].p, c.CV_H[q].p, c.CC_H[q].p);
        ortho(c, c.H[q].p, c.CV_H[q].p, c.CC_H[q].p, false);
        mark(4 + 3 * q);
    }
    c.haveH = true;
    return pre;
}

static void upload_coef(olcao_ctx &c, const double *coef)
{
    ensure_mats(c);
    HIPCHK(hipMemcpyAsync(c.dCoef.p, coef, (size_t)c.potDim * c.spin * 8, hipMemcpyHostToDevice,
                          c.stream));
}

static bool iteration_build(olcao_ctx &c, const double *coef, bool prefactor = false)
{
    c.trdMs = 0.0;
    upload_coef(c, coef);
    HIPCHK(hipEventRecord(c.ev[0], c.stream));
    return iteration_device(c, true, prefactor);
}

// Second half: the solve for all k and spins, then the stage times.
// Stage times from the iteration's events; solveMs < 0: the solve ran on
// this context's stream and ev[8] marks its end.
static void record_timings(olcao_ctx &c, double solveMs)
{
    auto el = [&](int a, int b) {
        float v = 0.f;
        HIPCHK(hipEventElapsedTime(&v, c.ev[a], c.ev[b]));
        return (double)v;
    };
    double ham = 0.0, t3c = 0.0, cb = 0.0;
    for (int q = 0; q < c.spin; ++q) {
        ham += el(q == 0 ? 1 : 1 + 3 * q, 3 + 3 * q);
        t3c += el(2 + 3 * q, 3 + 3 * q);
        cb += el(3 + 3 * q, 4 + 3 * q);
    }
    c.timings[0] = el(0, 1);
    c.timings[1] = ham;
    c.timings[2] = cb;
    c.timings[3] = solveMs < 0 ? el(1 + 3 * c.spin, 8) : solveMs;
    c.timings[4] = solveMs < 0 ? el(0, 8) : el(0, 1 + 3 * c.spin) + solveMs;
    c.timings[5] = t3c;
    c.timings[6] = 0.0;
    c.timings[7] = c.trdMs;
}

static void iteration_solve(olcao_ctx &c, int numStates, double *evals, bool pre = false)
{
    hipStream_t s = c.stream;
    solve_spins(&c, numStates, evals, nullptr, pre);
    HIPCHK(hipEventRecord(c.ev[8], s));
    HIPCHK(hipEventSynchronize(c.ev[8]));
    record_timings(c, -1.0);
    c.lastReplay = false;
}

// Capture the device work of one Mode-B iteration (build and the solve of
// every spin, with no host synchronisation) into a graph, right after an
// eager iteration allocated every buffer it uses.  The coefficients are
// uploaded before each replay and the status and eigenvalues read after
// it.  A graph with a memset node is not kept (such nodes have given wrong
// results on replays); a capture that fails leaves the context eager and
// the reason in graphErr.
static void capture_iteration(olcao_ctx &c)
{
    hipStream_t s = c.stream;
    hipGraph_t g = nullptr;
    try {
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            const bool pre = iteration_device(c, false, true);
            std::vector<int> info;
            for (int q = 0; q < c.spin; ++q) solve_enqueue(c, q, c.H[q].p, c.S.p, true, info, pre);
        } catch (...) {
            hipGraph_t junk = nullptr;
            (void)hipStreamEndCapture(s, &junk);
            if (junk) (void)hipGraphDestroy(junk);
            throw;
        }
        HIPCHK(hipStreamEndCapture(s, &g));
        size_t nn = 0;
        HIPCHK(hipGraphGetNodes(g, nullptr, &nn));
        std::vector<hipGraphNode_t> nodes(nn);
        HIPCHK(hipGraphGetNodes(g, nodes.data(), &nn));
        for (hipGraphNode_t nd : nodes) {
            hipGraphNodeType t;
            HIPCHK(hipGraphNodeGetType(nd, &t));
            if (t == hipGraphNodeTypeMemset) fail(OLCAO_ERR_UNSUPPORTED, "the captured iteration holds a memset node");
        }
        hipGraphExec_t ex = nullptr;
        HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
        if (c.graphExec) (void)hipGraphExecDestroy(c.graphExec);
        if (c.graph) (void)hipGraphDestroy(c.graph);
        c.graph = g;
        c.graphExec = ex;
        c.graphEpoch = c.allocEpoch.load();
    } catch (const Fail &e) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        c.graphFailed = true;
        c.graphErr = e.msg;
    }
}

// One Mode-B iteration by replaying the captured graph.
static void iteration_replay(olcao_ctx &c, const double *coef, int numStates, double *evals)
{
    hipStream_t s = c.stream;
    upload_coef(c, coef);
    HIPCHK(hipEventRecord(c.ev[0], s));
    HIPCHK(hipGraphLaunch(c.graphExec, s));
    HIPCHK(hipEventRecord(c.ev[8], s));
    const size_t ev = (size_t)c.nk * numStates;
    for (int q = 0; q < c.spin; ++q) {
        std::vector<int> info;
        read_info(c.eig, c.dInfo[q].p, c.tInfo[q].p, c.valeDim, c.nk, s, info);
        solve_finish(c, q, numStates, info, evals ? evals + q * ev : nullptr, nullptr);
    }
    HIPCHK(hipEventSynchronize(c.ev[8]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c.ev[0], c.ev[8]));
    for (double &t : c.timings) t = 0.0;
    c.timings[4] = ms;
    c.timings[6] = 1.0;
    c.haveS = c.haveH = true;
    c.lastReplay = true;
}

// One generalised solve over several cells' matrices of the same size and
// kind (the batched call): every member's H(k) and S(k) are copied into one
// strided batch on the first member's stream, solved in one gen_eig call
// (the batched device path or rocSOLVER's strided-batched one), and the
// eigenvalues, eigenvectors (kept for the valence density) and stage times
// go back to each member.  A member is a (cell, spin).
// ---------------------------------------------------------------------------
// The batched call's build.  Per spin q, three phases:
//   A  on each cell's streams, from the worker threads: the coefficients,
//      the overlap sweep (q = 0), the kinetic (and mass-velocity) blocks;
//   B  the fused Hamiltonian three-centre work of every cell, one launch per
//      class and kernel form over all cells (batch_three_center);
//   C  again per cell: contraction, Bloch sums, orthogonalisation.
// Each cell's blocks, W and matrices are its own, computed by the same
// kernels in the same order as in its single-context iteration.
static void batch_phase_a(olcao_ctx &c, const double *coef, int q)
{
    hipStream_t s = c.stream;
    if (q == 0) {
        c.trdMs = 0.0;
        upload_coef(c, coef);
        HIPCHK(hipEventRecord(c.ev[0], s));
        overlap_sweep(c);
        HIPCHK(hipEventRecord(c.ev[1], s));
    }
    two_center_forked(c, OP_KE, 0, c.rel ? 0xffffu : ~kFusedKEClasses & 0xffffu);
    if (c.rel) two_center_forked(c, OP_MV, 1);
    launch_fill(reinterpret_cast<double *>(c.dCounters.p), 32, 0.0, s);
    HIPCHK(hipEventRecord(c.ev[2 + 3 * q], s));
}

static void batch_phase_c(olcao_ctx &c, int q, hipEvent_t built)
{
    hipStream_t s = c.stream;
    HIPCHK(hipStreamWaitEvent(s, built, 0));
    HIPCHK(hipEventRecord(c.ev[3 + 3 * q], s));
    contract_scatter(c, c.H[q].p, c.CV_H[q].p, c.CC_H[q].p);
    ortho(c, c.H[q].p, c.CV_H[q].p, c.CC_H[q].p, false);
    HIPCHK(hipEventRecord(c.ev[4 + 3 * q], s));
    if (q == c.spin - 1) c.haveH = true;
}

// Phase B of spin q over `cells` (every cell's phase A complete): the
// launch arguments go to the device in one copy, then the launches run on
// own's streams.  A launch unit is one class over one group of cells: the
// three-Gaussian pass per kernel form, the nuclear pass, the finish, in
// that order on one stream; units are spread over the streams by work,
// heaviest first.  Cells are grouped by the dynamic LDS their kernels take
// (a launch reserves the largest of its cells') and each such set is split
// in two, alternating down the weight order, so that the heavy classes have
// independent launches for the other streams.  Waves take 64 chunks of one
// cell; the cells' chunk lists are interleaved task by task (each list runs
// heaviest first).  Leaves own.ev[11] recorded after all of it.
static void batch_three_center(olcao_ctx *const *ctxs, const std::vector<int> &cells, int q,
                               olcao_ctx &own)
{
    const int nc = (int)cells.size();
    std::vector<DevCtx> dcs(nc);
    for (int j = 0; j < nc; ++j) dcs[j] = ctxs[cells[j]]->dev;
    auto tabbed = [](const olcao_ctx &c) { return k3c_table_bytes(c.dev) <= kTableLdsMax; };
    constexpr int kSplit = 2;
    std::vector<int> group(nc);
    int nGroups = 0;
    {
        std::map<size_t, std::vector<int>> byLds;
        for (int j = 0; j < nc; ++j) {
            const olcao_ctx &c = *ctxs[cells[j]];
            byLds[k3c_site_bytes(c.dev) + (tabbed(c) ? k3c_table_bytes(c.dev) : 0)].push_back(j);
        }
        for (auto &kv : byLds) {
            for (size_t t = 0; t < kv.second.size(); ++t) group[kv.second[t]] = nGroups + (int)(t % kSplit);
            nGroups += (int)std::min<size_t>(kSplit, kv.second.size());
        }
    }
    std::vector<Cell3c> args;
    struct Unit {
        int k = 0;
        std::vector<Task3c> lists[5];   // e3c forms 0..2, nuclear, finish
        size_t ldsE3c[3] = {0, 0, 0}, ldsNuc = 0;
        double weight = 0.0;
        size_t off[5] = {};
    };
    std::vector<Unit> units;
    for (int k = 0; k < 16; ++k) {
        const int L = k / 4 + k % 4;
        const bool fuse = (kFusedKEClasses >> k) & 1u;
        for (int g = 0; g < nGroups; ++g) {
            Unit U;
            U.k = k;
            std::vector<int> members, form;
            for (int j = 0; j < nc; ++j) {
                const olcao_ctx &c = *ctxs[cells[j]];
                if (group[j] != g || c.cls.count[k] == 0) continue;
                Cell3c a{};
                a.ctx = j;
                a.nChunk = c.cls.nChunk[k];
                a.nRec = c.cls.count[k];
                a.recs = c.cls.rec[k];
                a.chRec = c.cls.chRec[k];
                a.chBeg = c.cls.chBeg[k];
                a.chEnd = c.cls.chEnd[k];
                a.nOff = c.cls.nOff[k];
                a.img = c.cls.img[k];
                a.imgN = c.cls.imgN[k];
                a.coef = c.dCoef.p + (size_t)q * c.potDim;
                a.W = c.cls.W[k];
                a.finOrder = c.cls.finOrder[k];
                a.rcOff = c.cls.rcOff[k];
                a.rcList = c.cls.rcList[k];
                a.blocks = c.dBlocks.p;
                // as launch_three_center with accumulate = 1 (| kAccAddKE when rel = 0)
                const bool addKE = fuse && !c.rel;
                a.finAcc = addKE ? 0 : 1;
                a.finKE = addKE ? 1 : 0;
                const bool tab = tabbed(c);
                const int v = tab ? (c.precision == OLCAO_PREC_MIXED ? 0 : 1) : 2;
                U.ldsE3c[v] = std::max(U.ldsE3c[v], k3c_site_bytes(c.dev) +
                                                        (tab ? k3c_table_bytes(c.dev) : 0) + 16);
                if (c.dev.anyNuc) U.ldsNuc = std::max(U.ldsNuc, k3c_site_bytes(c.dev) + 16);
                U.weight += (double)a.nChunk * ((L + 1) * (L + 2) * (L + 3) / 6);
                members.push_back((int)args.size());
                form.push_back(v);
                args.push_back(a);
            }
            if (members.empty()) continue;
            // tasks interleaved across the cells: every cell's t-th wave, then t + 1
            for (int t = 0;; ++t) {
                bool any = false;
                for (size_t m = 0; m < members.size(); ++m) {
                    const Cell3c &a = args[members[m]];
                    if (64 * t < a.nChunk) {
                        U.lists[form[m]].push_back({members[m], 64 * t});
                        if (dcs[a.ctx].anyNuc) U.lists[3].push_back({members[m], 64 * t});
                        any = true;
                    }
                    if (64 * t < a.nRec) {
                        U.lists[4].push_back({members[m], 64 * t});
                        any = true;
                    }
                }
                if (!any) break;
            }
            units.push_back(std::move(U));
        }
    }
    // one buffer: [DevCtx x nc][Cell3c x args][every unit's task lists]
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t off = up(sizeof(DevCtx) * nc);
    const size_t offCells = off;
    off = up(off + sizeof(Cell3c) * std::max<size_t>(1, args.size()));
    for (Unit &U : units)
        for (int l = 0; l < 5; ++l) {
            U.off[l] = off;
            off = up(off + sizeof(Task3c) * U.lists[l].size());
        }
    std::vector<unsigned char> host(off);
    std::memcpy(host.data(), dcs.data(), sizeof(DevCtx) * nc);
    if (!args.empty()) std::memcpy(host.data() + offCells, args.data(), sizeof(Cell3c) * args.size());
    for (const Unit &U : units)
        for (int l = 0; l < 5; ++l)
            if (!U.lists[l].empty())
                std::memcpy(host.data() + U.off[l], U.lists[l].data(), sizeof(Task3c) * U.lists[l].size());
    own.dBatch.alloc(off);
    HIPCHK(hipMemcpy(own.dBatch.p, host.data(), off, hipMemcpyHostToDevice));
    unsigned char *base = own.dBatch.p;
    // units over own's streams, heaviest first onto the least-loaded one
    constexpr int nS = 1 + olcao_ctx::kSide;
    hipStream_t st[nS] = {own.stream};
    for (int i = 1; i < nS; ++i) st[i] = own.side[i - 1];
    std::vector<int> order(units.size());
    for (size_t u = 0; u < units.size(); ++u) order[u] = (int)u;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return units[a].weight > units[b].weight; });
    double load[nS] = {};
    bool used[nS] = {};
    HIPCHK(hipEventRecord(own.evFork, own.stream));
    for (int u : order) {
        const Unit &U = units[u];
        const int i = (int)(std::min_element(load, load + nS) - load);
        load[i] += U.weight;
        if (i > 0 && !used[i]) HIPCHK(hipStreamWaitEvent(st[i], own.evFork, 0));
        used[i] = true;
        Multi3c m{};
        m.k = U.k;
        m.ctxs = reinterpret_cast<const DevCtx *>(base);
        m.cells = reinterpret_cast<const Cell3c *>(base + offCells);
        for (int v = 0; v < 3; ++v) {
            m.e3c[v] = reinterpret_cast<const Task3c *>(base + U.off[v]);
            m.nE3c[v] = (int)U.lists[v].size();
            m.ldsE3c[v] = U.ldsE3c[v];
        }
        m.nuc = reinterpret_cast<const Task3c *>(base + U.off[3]);
        m.nNuc = (int)U.lists[3].size();
        m.ldsNuc = U.ldsNuc;
        m.fin = reinterpret_cast<const Task3c *>(base + U.off[4]);
        m.nFin = (int)U.lists[4].size();
        launch_three_center_multi(m, st[i]);
    }
    HIPCHK(hipGetLastError());
    for (int i = 1; i < nS; ++i)
        if (used[i]) {
            HIPCHK(hipEventRecord(own.evJoin[i - 1], st[i]));
            HIPCHK(hipStreamWaitEvent(own.stream, own.evJoin[i - 1], 0));
        }
    HIPCHK(hipEventRecord(own.ev[11], own.stream));
}

struct SolveMember {
    int cell, spin;
};

static void grouped_solve(const std::vector<SolveMember> &mem, olcao_ctx *const *ctxs,
                          const int *numStates, double *const *evals, std::vector<olcao_status> &st)
{
    olcao_ctx &h = *ctxs[mem[0].cell];
    const int n = h.valeDim;
    const bool cplx = !h.gamma;
    const size_t per = (size_t)n * n * (cplx ? 2 : 1);
    int nbTot = 0;
    for (const auto &m : mem) nbTot += ctxs[m.cell]->nk;
    hipStream_t s = h.stream;
    h.gA.alloc(per * nbTot);
    h.gB.alloc(per * nbTot);
    h.gD.alloc((size_t)n * nbTot);
    h.gE.alloc((size_t)n * nbTot);
    h.gInfo.alloc(nbTot);
    h.gTInfo.alloc(nbTot);
    size_t off = 0;
    for (const auto &m : mem) {
        const olcao_ctx &c = *ctxs[m.cell];
        HIPCHK(hipMemcpyAsync(h.gA.p + off * per, c.H[m.spin].p, per * c.nk * 8, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(h.gB.p + off * per, c.S.p, per * c.nk * 8, hipMemcpyDeviceToDevice, s));
        off += c.nk;
    }
    HIPCHK(hipEventRecord(h.ev[9], s));
    std::vector<int> info;
    gen_eig(h.eig, h.blas, s, cplx, n, nbTot, h.gA.p, h.gB.p, h.gD.p, h.gE.p, h.gInfo.p, h.gTInfo.p, info,
            h.eigLibrary);
    HIPCHK(hipEventRecord(h.ev[10], s));
    std::vector<double> w((size_t)n * nbTot);
    HIPCHK(hipMemcpyAsync(w.data(), h.gD.p, w.size() * 8, hipMemcpyDeviceToHost, s));
    off = 0;
    for (const auto &m : mem) {
        olcao_ctx &c = *ctxs[m.cell];
        DBuf<double> &X = m.spin ? c.Aw2 : c.Aw;
        X.alloc(c.matDoubles());
        HIPCHK(hipMemcpyAsync(X.p, h.gA.p + off * per, per * c.nk * 8, hipMemcpyDeviceToDevice, s));
        off += c.nk;
    }
    HIPCHK(hipStreamSynchronize(s));
    h.eig.trdTimed = false;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h.ev[9], h.ev[10]));
    off = 0;
    for (const auto &m : mem) {
        olcao_ctx &c = *ctxs[m.cell];
        const int ns = numStates[m.cell];
        olcao_status bad = OLCAO_OK;
        for (int k = 0; k < c.nk && bad == OLCAO_OK; ++k) {
            const int v = info[off + k];
            if (v > n) {
                c.err = "S(k) not positive definite at k=" + std::to_string(k) + " (info=" + std::to_string(v) + ")";
                bad = OLCAO_ERR_NOT_SPD;
            } else if (v != 0) {
                c.err = "eigensolver did not converge at k=" + std::to_string(k) + " (info=" + std::to_string(v) + ")";
                bad = OLCAO_ERR_SOLVER;
            }
        }
        if (bad == OLCAO_OK && (ns <= 0 || ns > n)) {
            c.err = "numStates must lie in 1..";
            c.err += std::to_string(n);
            bad = OLCAO_ERR_INPUT;
        }
        if (bad != OLCAO_OK) {
            st[m.cell] = bad;
            off += c.nk;
            continue;
        }
        double *ev = evals[m.cell] + (size_t)m.spin * c.nk * ns;
        for (int k = 0; k < c.nk; ++k)
            std::memcpy(ev + (size_t)k * ns, w.data() + (off + k) * n, ns * sizeof(double));
        c.hEv[m.spin].assign(w.begin() + off * n, w.begin() + (off + c.nk) * n);
        c.vecsValid[m.spin] = true;
        record_timings(c, ms);
        off += c.nk;
    }
}

olcao_status olcao_scf_iteration(olcao_ctx *c, const double *coef, int numStates, double *evals)
{
    if (!c || !coef) return OLCAO_ERR_INPUT;
    return guard(c, [&] {
        if (numStates <= 0 || numStates > c->valeDim)
            fail(OLCAO_ERR_INPUT, "numStates must lie in 1..%d", c->valeDim);
        if (c->useGraph && c->graphExec && c->graphEpoch == c->allocEpoch.load()) {
            iteration_replay(*c, coef, numStates, evals);
            return;
        }
        const bool pre = iteration_build(*c, coef, true);
        iteration_solve(*c, numStates, evals, pre);
        if (c->useGraph && !c->graphFailed) capture_iteration(*c);
    });
}

// Whether the context's solve is the one-workgroup LDS solver (one kernel
// of this library per solve), the only solve the batched call runs
// alongside other cells' work.
static bool concurrent_solve(const olcao_ctx &c)
{
    return !c.eigLibrary && small_hegv_fits(c.valeDim, !c.gamma);
}

olcao_status olcao_scf_iteration_batch(int nCells, olcao_ctx *const *ctxs,
                                       const double *const *potCoeffs, const int *numStates,
                                       double *const *evals, int nWorkers, double *cellMs,
                                       int *failedCell)
{
    return olcao_scf_iteration_batch_mode(nCells, ctxs, potCoeffs, numStates, evals, nWorkers,
                                          OLCAO_BATCH_CELLS, cellMs, failedCell);
}

olcao_status olcao_scf_iteration_batch_mode(int nCells, olcao_ctx *const *ctxs,
                                            const double *const *potCoeffs, const int *numStates,
                                            double *const *evals, int nWorkers, int mode,
                                            double *cellMs, int *failedCell)
{
    if (failedCell) *failedCell = -1;
    if (nCells <= 0 || !ctxs || !potCoeffs || !numStates || !evals) return OLCAO_ERR_INPUT;
    if (mode != OLCAO_BATCH_CELLS && mode != OLCAO_BATCH_SHARED) return OLCAO_ERR_INPUT;
    for (int i = 0; i < nCells; ++i)
        if (!ctxs[i] || !potCoeffs[i] || !evals[i]) {
            if (failedCell) *failedCell = i;
            return OLCAO_ERR_INPUT;
        }
    nWorkers = std::max(1, std::min(nWorkers, 16));
    // the grouped solves below allocate on behalf of several contexts: a
    // captured iteration graph of any of them is captured again later
    for (int i = 0; i < nCells; ++i) ctxs[i]->graphEpoch = 0;
    // heaviest cells first: three-centre image entries plus the solve's n^3
    std::vector<int> order(nCells);
    std::vector<double> weight(nCells);
    for (int i = 0; i < nCells; ++i) {
        const olcao_ctx &c = *ctxs[i];
        weight[i] = (double)c.imageEntries +
                    (double)c.valeDim * c.valeDim * c.valeDim * c.nk * c.spin / 16.0;
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return weight[a] > weight[b]; });
    std::vector<olcao_status> st(nCells, OLCAO_OK);
    std::atomic<bool> failed{false};
    // one pass over `list`: nWorkers host threads take cells in order, each
    // cell on its own streams, so the device runs several cells at once
    auto run = [&](const std::vector<int> &list, int workers, auto &&step) {
        std::atomic<int> next{0};
        auto work = [&] {
            for (int t; (t = next.fetch_add(1)) < (int)list.size();) {
                if (failed.load()) return;
                const int i = list[t];
                st[i] = guard(ctxs[i], [&] { step(i); });
                if (st[i] != OLCAO_OK) failed.store(true);
            }
        };
        std::vector<std::thread> pool;
        for (int w = 1; w < std::min(workers, (int)list.size()); ++w) pool.emplace_back(work);
        work();
        for (auto &th : pool) th.join();
    };
    // 1. every cell's integral build, each cell followed by its solve when
    // that is the LDS solver.  OLCAO_BATCH_CELLS: each cell's whole build on
    // its own streams from the worker threads, concurrently.
    // OLCAO_BATCH_SHARED: per spin, the cells' own stages from the worker
    // threads and their three-centre work in shared launches over all cells.
    std::vector<char> alone(nCells);
    std::vector<int> serial;
    int maxSpin = 1;
    for (int i : order) {
        if ((alone[i] = !concurrent_solve(*ctxs[i]))) serial.push_back(i);
        maxSpin = std::max(maxSpin, ctxs[i]->spin);
    }
    if (mode == OLCAO_BATCH_CELLS)
        run(order, nWorkers, [&](int i) {
            iteration_build(*ctxs[i], potCoeffs[i]);
            if (!alone[i]) iteration_solve(*ctxs[i], numStates[i], evals[i]);
        });
    for (int q = 0; q < maxSpin && mode == OLCAO_BATCH_SHARED && !failed.load(); ++q) {
        std::vector<int> list;
        for (int i : order)
            if (ctxs[i]->spin > q) list.push_back(i);
        run(list, nWorkers, [&](int i) { batch_phase_a(*ctxs[i], potCoeffs[i], q); });
        if (failed.load()) break;
        olcao_ctx &own = *ctxs[list[0]];
        const olcao_status r = guard(&own, [&] {
            HIPCHK(hipDeviceSynchronize());
            batch_three_center(ctxs, list, q, own);
        });
        if (r != OLCAO_OK) {
            st[list[0]] = r;
            failed.store(true);
            break;
        }
        run(list, nWorkers, [&](int i) {
            olcao_ctx &c = *ctxs[i];
            batch_phase_c(c, q, own.ev[11]);
            if (q == c.spin - 1 && !alone[i]) iteration_solve(c, numStates[i], evals[i]);
        });
    }
    // 2. the other solves one at a time with nothing else on the device.
    // The tridiagonal path's reduction is a cooperative launch that needs
    // all of its workgroups resident.  rocSOLVER's solves gave wrong
    // eigenvalues for some cells (and once an S(k) it took for indefinite)
    // when they overlapped other cells' kernels from other host threads,
    // while the same cells solved alone were right every time and no
    // library buffer was written outside its bounds (guard pads around
    // every allocation stayed intact); the cause was not isolated.
    // Cells whose matrices have the same size and kind are solved together
    // in one batched call (grouped_solve).
    if (!failed.load() && !serial.empty()) {
        const olcao_status sync = guard(nullptr, [&] { HIPCHK(hipDeviceSynchronize()); });
        if (sync != OLCAO_OK) return sync;
        std::map<std::tuple<int, bool, bool>, std::vector<SolveMember>> groups;
        for (int i : serial)
            for (int q = 0; q < ctxs[i]->spin; ++q)
                groups[{ctxs[i]->valeDim, ctxs[i]->gamma, ctxs[i]->eigLibrary}].push_back({i, q});
        for (auto &g : groups) {
            const int first = g.second[0].cell;
            const olcao_status r = guard(ctxs[first], [&] {
                grouped_solve(g.second, ctxs, numStates, evals, st);
            });
            if (r != OLCAO_OK)
                for (const auto &m : g.second) {
                    st[m.cell] = r;
                    if (m.cell != first) ctxs[m.cell]->err = ctxs[first]->err;
                }
        }
    }
    for (int t = 0; t < nCells; ++t) {
        const int i = order[t];
        if (st[i] != OLCAO_OK) {
            if (failedCell) *failedCell = i;
            g_noctx_err = ctxs[i]->err;
            return st[i];
        }
    }
    if (cellMs)
        for (int i = 0; i < nCells; ++i) cellMs[i] = ctxs[i]->timings[4];
    return OLCAO_OK;
}

olcao_status olcao_get_work_counts(olcao_ctx *c, unsigned long long counts[32])
{
    if (!c || !counts) return OLCAO_ERR_INPUT;
    return guard(c, [&] {
        HIPCHK(hipMemcpyAsync(counts, c->dCounters.p, 32 * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

olcao_status olcao_get_graph_state(const olcao_ctx *c, int *state, const char **reason)
{
    if (!c || !state) return OLCAO_ERR_INPUT;
    if (!c->useGraph)
        *state = 0;
    else if (c->graphFailed)
        *state = 3;
    else
        *state = c->graphExec && c->graphEpoch == c->allocEpoch.load() ? 1 : 2;
    if (reason) *reason = c->graphErr.c_str();
    return OLCAO_OK;
}

olcao_status olcao_get_class_pairs(const olcao_ctx *c, long long counts[16])
{
    if (!c || !counts) return OLCAO_ERR_INPUT;
    for (int k = 0; k < 16; ++k) counts[k] = (long long)c->classIdx[k].size();
    return OLCAO_OK;
}

olcao_status olcao_get_timings(const olcao_ctx *c, double t[8])
{
    if (!c || !t) return OLCAO_ERR_INPUT;
    std::memcpy(t, c->timings, sizeof(c->timings));
    return OLCAO_OK;
}

static olcao_status solve_host(int n, int nb, double *A, double *B, double *w, bool cplx)
{
    return guard(nullptr, [&] {
        if (n <= 0 || nb <= 0) fail(OLCAO_ERR_INPUT, "n and nbatch must be positive");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) fail(OLCAO_ERR_NO_DEVICE, "no GPU");
        // the stream and the handle are released on every path, the
        // failures thrown below included
        struct Res {
            hipStream_t s = nullptr;
            rocblas_handle h = nullptr;
            ~Res()
            {
                if (h) rocblas_destroy_handle(h);
                if (s) (void)hipStreamDestroy(s);
            }
        } res;
        HIPCHK(hipStreamCreate(&res.s));
        if (rocblas_create_handle(&res.h) != rocblas_status_success) {
            res.h = nullptr;
            fail(OLCAO_ERR_HIP, "rocblas_create_handle failed");
        }
        hipStream_t s = res.s;
        rocblas_handle h = res.h;
        if (rocblas_set_stream(h, s) != rocblas_status_success ||
            rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed) != rocblas_status_success)
            fail(OLCAO_ERR_HIP, "rocblas handle set-up failed");
        const size_t per = (size_t)n * n * (cplx ? 2 : 1);
        DBuf<double> dA, dB, dD, dE;
        DBuf<int> dI, dTI;
        dA.alloc(per * nb); dB.alloc(per * nb); dD.alloc((size_t)n * nb); dE.alloc((size_t)n * nb);
        dI.alloc(nb);
        dTI.alloc(nb);
        HIPCHK(hipMemcpyAsync(dA.p, A, per * nb * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dB.p, B, per * nb * 8, hipMemcpyHostToDevice, s));
        std::vector<int> info;
        EigWork ew;
        // OLCAO_EIGENSOLVER: auto (default), library or blocked, as the
        // contexts' olcao_options.eigensolver
        const char *path = std::getenv("OLCAO_EIGENSOLVER");
        const bool lib = path && std::strcmp(path, "library") == 0;
        ew.forceBlocked = path && std::strcmp(path, "blocked") == 0;
        gen_eig(ew, h, s, cplx, n, nb, dA.p, dB.p, dD.p, dE.p, dI.p, dTI.p, info, lib);
        HIPCHK(hipMemcpyAsync(A, dA.p, per * nb * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(w, dD.p, (size_t)n * nb * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (int k = 0; k < nb; ++k) {
            if (info[k] > n) fail(OLCAO_ERR_NOT_SPD, "B not positive definite in batch %d", k);
            if (info[k] != 0) fail(OLCAO_ERR_SOLVER, "no convergence in batch %d", k);
        }
    });
}

olcao_status olcao_solve_zhegv(int n, int nb, double *A, double *B, double *w)
{
    return solve_host(n, nb, A, B, w, true);
}

olcao_status olcao_solve_dsygv(int n, int nb, double *A, double *B, double *w)
{
    return solve_host(n, nb, A, B, w, false);
}

olcao_status olcao_prim_batch(int op, int n, const double *a, const double *A, const double *B,
                              const double *C, const int *sw, double *sh)
{
    return guard(nullptr, [&] {
        if (op < 0 || (op > 4 && op != 7 && op != 8) || n < 0) fail(OLCAO_ERR_INPUT, "bad op or n");
        if (n == 0) return;
        for (int i = 0; i < n; ++i) {
            int lo = sw[i] & 15, hi = (sw[i] >> 4) & 15;
            auto ok = [](int v) { return v == 1 || v == 2 || v == 4 || v == 8; };
            if (!ok(lo) || !ok(hi) || (sw[i] >> 8)) fail(OLCAO_ERR_INPUT, "bad l1l2switch %d", sw[i]);
        }
        DBuf<double> da, dA, dB, dC, dsh;
        std::vector<double> va(a, a + 3 * (size_t)n), vA(A, A + 3 * (size_t)n), vB(B, B + 3 * (size_t)n),
            vC(C, C + 3 * (size_t)n);
        da.upload(va, 0); dA.upload(vA, 0); dB.upload(vB, 0); dC.upload(vC, 0);
        std::vector<std::vector<int>> byClass(16);
        auto lof = [](int nib) { return nib == 1 ? 0 : nib == 2 ? 1 : nib == 4 ? 2 : 3; };
        for (int i = 0; i < n; ++i) byClass[lof(sw[i] & 15) * 4 + lof((sw[i] >> 4) & 15)].push_back(i);
        DBuf<int> dIdx[16];
        const int *idx[16];
        int cnt[16];
        for (int k = 0; k < 16; ++k) {
            dIdx[k].upload(byClass[k], 0);
            idx[k] = dIdx[k].p;
            cnt[k] = (int)byClass[k].size();
        }
        dsh.alloc(256 * (size_t)n);
        launch_prim_batch(op, n, da.p, dA.p, dB.p, dC.p, idx, cnt, dsh.p, 0);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(sh, dsh.p, 256 * (size_t)n * 8, hipMemcpyDeviceToHost));
    });
}

olcao_status olcao_boys_batch(int n, const double *T, double *F)
{
    return guard(nullptr, [&] {
        if (n <= 0) return;
        DBuf<double> dT, dF;
        std::vector<double> vT(T, T + n);
        dT.upload(vT, 0);
        dF.alloc(7 * (size_t)n);
        launch_boys_batch(n, dT.p, dF.p, 0);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(F, dF.p, 7 * (size_t)n * 8, hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
