// Three-centre lattice-sum kernels (electronicPE / nuclearPE / the fused
// PSCF sweep, src/uolcao/integrals.F90:2533-2856, :3211-3430), one
// translation unit per bra angular momentum (-DOLCAO_LA=0..3).
//
// Work decomposition (one lane per primitive pair):
//   * the screening is geometry-only, so it runs once per system: every
//     primitive pair has a list of the potential images it sums, each entry
//     carrying how many leading potential terms (and whether the nucleus)
//     each sweep kind takes from that image (k_image_list);
//   * the pairs of a class are sorted by their number of items (image x
//     term) so that the 64 lanes of a wave carry nearly equal work, heaviest
//     waves first;
//   * lane l walks its own list: every loop trip evaluates one item, and a
//     lane whose image is used up reads its next entry (prefetched one
//     ahead), so lanes stay busy until their lists end;
//   * each lane keeps its pair's McMurchie-Davidson sums W_tuv = sum_items
//     w R_tuv (t+u+v <= la+lb, 35 numbers for d-d) in registers: no search,
//     no range test and no cross-lane reduction in the sweep itself;
//   * the three-Gaussian terms and the nuclear terms are separate passes
//     (NUC = false / true) so that each keeps its own register budget;
//   * k_herm_finish turns W into the Cartesian block (Hermite expansion
//     coefficients about P), transforms it to solid harmonics and stores it
//     in the primitive-pair store, one thread per primitive pair.
#include "device_common.hpp"

#ifndef OLCAO_LA
#error "compile with -DOLCAO_LA=<0..3>"
#endif

namespace olcao {

// TAB: the per-(pair, term) constants mu and pref0 of the three-Gaussian
// items are computed once per lane into an LDS table (they depend on the
// pair and the potential exponent only), which takes the division and the
// square root out of the item loop.  MIX: mixed-precision exponential.
// One lane per chunk of a record's image list (chRec, chBeg, chEnd; n
// chunks); the chunk's Hermite sums go to W[e][chunk].  The workgroup of NT
// threads (64 or 256) takes the chunks i0 .. i0 + NT - 1; its waves share
// the staged sites and tables, the term table has a column per thread.
// DL = 1: the Hermite sums one order up (the force sweep's ket derivatives).
template <int LA, int LB, bool NUC, bool TAB, bool MIX, int NT = 64, int DL = 0>
__device__ __forceinline__ void three_center_wave(const DevCtx &c, const PPRec *recs, int n, int i0,
                                                  const int *chRec, const long long *chBeg,
                                                  const long long *chEnd,
                                                  const unsigned long long *img, int kind,
                                                  int termType, int termIdx, const double *coef,
                                                  int addW, double *W)
{
    constexpr int L = LA + LB + DL, NH = nherm(L);
    __shared__ double tAlpha[kMaxPotTypes * kMaxPotAlphas];
    __shared__ double tCoef[kMaxPotTypes * kMaxPotAlphas];
    __shared__ double tZ[kMaxPotTypes], tNuc[kMaxPotTypes];
    __shared__ double sExp2[256];   // 2^(j/256) for the table exponential
    // potential sites: dynamic LDS, positions then types
    extern __shared__ double dyn[];
    double *sitePos = dyn;
    int *siteType = reinterpret_cast<int *>(dyn + 3 * c.nPotSites);

    const int tid = threadIdx.x, lane = tid & 63;
    for (int q = tid; q < c.nPotSites; q += NT) {
        sitePos[3 * q] = c.potPos[3 * q];
        sitePos[3 * q + 1] = c.potPos[3 * q + 1];
        sitePos[3 * q + 2] = c.potPos[3 * q + 2];
        siteType[q] = c.potSiteType[q];
    }
    for (int q = tid; q < c.nPotTypes * kMaxPotAlphas; q += NT) {
        const int t = q / kMaxPotAlphas, a = q - t * kMaxPotAlphas;
        const bool ok = a < c.potNA[t];
        tAlpha[q] = ok ? c.potAlpha[t * c.maxPA + a] : 0.0;
        tCoef[q] = (ok && coef) ? coef[c.potCumul[t] + a] : 1.0;
    }
    for (int t = tid; t < c.nPotTypes; t += NT) {
        tZ[t] = c.potZ[t];
        tNuc[t] = c.potNuc[t];
    }
    for (int j = tid; j < 256; j += NT) sExp2[j] = exp2(j * (1.0 / 256.0));
    __syncthreads();

    const int i = i0 + tid;
    const bool active = i < n;
    const int ri = active ? i : 0;
    const PPRec rec = recs[chRec[ri]];
    const PairConst q = make_pair(rec.a1, rec.a2, rec.A, rec.B);
    // term table [type * maxPA + alpha][thread] of 1 / (zeta + a3), after
    // the sites: the item's mu and prefactor follow from it with
    // multiplications and one square root, in the same operations as when
    // both were tabulated (one table instead of two: more waves per CU)
    double *tabInv = dyn + ((3 * c.nPotSites + (c.nPotSites + 1) / 2 + 1) & ~1);
    if (TAB && !NUC) {
        for (int t = 0; t < c.nPotTypes; ++t)
            for (int a = 0; a < c.maxPA; ++a) {
                const double a3 = tAlpha[t * kMaxPotAlphas + a];
                tabInv[(t * c.maxPA + a) * NT + tid] = a3 > 0.0 ? 1.0 / (q.zeta + a3) : 0.0;
            }
    }
    double acc[NH];
#pragma unroll
    for (int e = 0; e < NH; ++e) acc[e] = 0.0;

    long long p = active ? chBeg[ri] : 0;
    const long long pe = active ? chEnd[ri] : 0;
    unsigned long long nxt = p < pe ? img[p] : 0ull;
    int nt = 0, nn = 0, t = 0;
    double C[3] = {0.0, 0.0, 0.0};
    unsigned items = 0;
    while (true) {
        // next entry with work for this pass (the following one is prefetched)
        while (nn >= nt && p < pe) {
            const unsigned long long e = nxt;
            ++p;
            if (p < pe) nxt = img[p];
            const int site = img_site(e);
            const int st = siteType[site];
            int m;
            if (NUC)
                m = kind == K3_HAM ? img_nucham(e) : img_nucnp(e);
            else if (kind == K3_HAM)
                m = img_nham(e);
            else
                m = (st == termType && img_nep(e) > termIdx) ? 1 : 0;
            nt = m;
            nn = 0;
            if (m > 0) {
                t = st;
                const int nx = img_n(e, 0), ny = img_n(e, 1), nz = img_n(e, 2);
#pragma unroll
                for (int x = 0; x < 3; ++x)
                    C[x] = sitePos[3 * site + x] +
                           (nx * c.vec[x] + ny * c.vec[3 + x] + nz * c.vec[6 + x]);
            }
        }
        const bool has = nn < nt;
        if (__ballot(has) == 0ull) break;
        if (has) {
            if (NUC) {
                acc_nuc_herm<L>(q, tNuc[t], C, -tZ[t], acc);
            } else {
                const int term = kind == K3_HAM ? nn : termIdx;
                if (TAB) {
                    const int tq = t * kMaxPotAlphas + term;
                    const double inv = tabInv[(t * c.maxPA + term) * NT + tid];
                    const double wgt = kind == K3_HAM ? tCoef[tq] : 1.0;
                    const double tt = kPi * inv;
                    const double mu = q.zeta * tAlpha[tq] * inv;
                    const double pref0 = wgt * tt * sqrt(tt);
                    acc_e3c_herm_tab<L, MIX>(q, mu, pref0, C, sExp2, acc);
                } else {
                    const double wgt = kind == K3_HAM ? tCoef[t * kMaxPotAlphas + term] : 1.0;
                    acc_e3c_herm<L>(q, tAlpha[t * kMaxPotAlphas + term], C, wgt, acc);
                }
            }
            ++nn;
            ++items;
        }
    }
    if (active) {
        // all NH loads before the first store: the compiler cannot tell that
        // the NH words differ, and would otherwise wait for each in turn
        double *pw = W + i;
        if (NUC && addW) {   // only the nuclear pass adds onto the three-Gaussian pass's W
            double old[NH];
#pragma unroll
            for (int e = 0; e < NH; ++e) old[e] = pw[(long long)e * n];
#pragma unroll
            for (int e = 0; e < NH; ++e) pw[(long long)e * n] = old[e] + acc[e];
        } else {
#pragma unroll
            for (int e = 0; e < NH; ++e) pw[(long long)e * n] = acc[e];
        }
    }
    if (c.counters) {
        const double tot = wave_sum((double)items);
        if (lane == 0)
            atomicAdd(c.counters + 2 * (LA * 4 + LB) + (NUC ? 1 : 0), (unsigned long long)tot);
    }
}

template <int LA, int LB, bool NUC, bool TAB, bool MIX, int NT = 64>
__global__ __launch_bounds__(NT) void k_three_center(DevCtx c, const PPRec *recs, int n,
                                                     const int *chRec, const long long *chBeg,
                                                     const long long *chEnd,
                                                     const unsigned long long *img, int kind,
                                                     int termType, int termIdx,
                                                     const double *coef, int addW, double *W)
{
    three_center_wave<LA, LB, NUC, TAB, MIX, NT>(c, recs, n, (int)blockIdx.x * NT, chRec, chBeg,
                                                 chEnd, img, kind, termType, termIdx, coef, addW, W);
}

// The fused Hamiltonian sweep of many cells in one launch: wave b takes the
// 64 chunks of task b (a cell and its first chunk) with that cell's system,
// lists and coefficients.  NUC: the nuclear pass over the compacted
// entries, added onto the three-Gaussian pass's W.
template <int LA, int LB, bool NUC, bool TAB, bool MIX>
__global__ __launch_bounds__(64) void k_three_center_multi(const DevCtx *ctxs, const Cell3c *cells,
                                                           const Task3c *tasks)
{
    const Task3c tk = tasks[blockIdx.x];
    const Cell3c a = cells[tk.cell];
    const DevCtx c = ctxs[a.ctx];
    if (NUC)
        three_center_wave<LA, LB, true, false, false>(c, a.recs, a.nChunk, tk.first, a.chRec,
                                                      a.nOff, a.nOff + 1, a.imgN, K3_HAM, 0, 0,
                                                      a.coef, 1, a.W);
    else
        three_center_wave<LA, LB, false, TAB, MIX>(c, a.recs, a.nChunk, tk.first, a.chRec, a.chBeg,
                                                   a.chEnd, a.img, K3_HAM, 0, 0, a.coef, 0, a.W);
}

// W -> Cartesian block -> solid harmonics, one thread per primitive pair:
//   <a|O|b> = K_AB sum_tuv Ex^{ax bx}_t Ey^{ay by}_u Ez^{az bz}_v W_tuv,
// then the reference's Cartesian -> solid-harmonic map, one ket harmonic
// (column) at a time so that only one Cartesian column is live.  All loops
// are compile-time (static_for), so W and the E tables stay in registers.
// Blocks of up to kFinishStage doubles are staged in LDS and written (or
// accumulated) by the whole wave with store_staged_blocks, so that the
// stores to the primitive-pair store are contiguous runs instead of one
// 8-byte word per lane per element.
constexpr int kFinishStage = 81;

// W of record i: the sum of its chunks' W in entry order (rcOff, rcList;
// nChunk columns).  The wave takes the records finOrder[li0 .. li0 + 63].
template <int LA, int LB>
__device__ __forceinline__ void herm_finish_wave(const PPRec *recs, const double *W, int nrec,
                                                 int li0, const int *finOrder, const int *rcOff,
                                                 const int *rcList, int nChunk, int accumulate,
                                                 int addKE, double *blocks)
{
    constexpr int L = LA + LB, NH = nherm(L);
    constexpr int NA = ncart(LA), SA = nsh(LA), SB = nsh(LB);
    constexpr bool STAGE = SA * SB <= kFinishStage;
    __shared__ double sOut[STAGE ? 64 * SA * SB + 1 : 1];
    __shared__ long long sOff[64];
    const int lane = threadIdx.x;
    const int li = li0 + lane;
    const bool active = li < nrec;
    const int i = finOrder[active ? li : nrec - 1];
    const PPRec rec = recs[i];
    const PairConst q = make_pair(rec.a1, rec.a2, rec.A, rec.B);
    double Wl[NH];
#pragma unroll
    for (int e = 0; e < NH; ++e) Wl[e] = 0.0;
    if (active)
        for (int q = rcOff[i]; q < rcOff[i + 1]; ++q) {
            const int col = rcList[q];
#pragma unroll
            for (int e = 0; e < NH; ++e) Wl[e] += W[(long long)e * nChunk + col];
        }
    double E[3][LA + 1][LB + 1][L + 1];
#pragma unroll
    for (int x = 0; x < 3; ++x) herm_e<LA, LB>(q.P[x] - q.A[x], q.P[x] - q.B[x], 0.5 / q.zeta, E[x]);
    const double kab = exp(-q.xid2);
    double *o = blocks + rec.off;
    static_for<SB>([&](auto j_) __attribute__((always_inline)) {
        constexpr int j = decltype(j_)::value;
        double col[NA];
        static_for<NA>([&](auto a_) __attribute__((always_inline)) {
            constexpr int a = decltype(a_)::value;
            double v = 0.0;
            static_for<sh_nterm(j)>([&](auto t_) __attribute__((always_inline)) {
                constexpr int tj = decltype(t_)::value;
                constexpr int b = sh_cart(j, tj);
                v += sh_weight(j, tj) *
                     herm_contract_ct<LA, LB, cexp(a, 0), cexp(a, 1), cexp(a, 2), cexp(b, 0),
                                      cexp(b, 1), cexp(b, 2)>(E, Wl);
            });
            col[a] = v;
        });
#pragma unroll
        for (int ii = 0; ii < SA; ++ii) {
            double s = 0.0;
#pragma unroll
            for (int ti = 0; ti < sh_nterm(ii); ++ti) s += sh_weight(ii, ti) * col[sh_cart(ii, ti)];
            s *= kab;
            if (STAGE) {
                sOut[lane * SA * SB + ii * SB + j] = s;
            } else if (active) {
                double *p = o + ii * SB + j;
                *p = accumulate ? *p + s : s;
            }
        }
    });
    if (STAGE) {
        if (addKE) {   // + the kinetic block of the same primitive pair
            double pk[NA][ncart(LB)];
            two_center_cart<LA, LB, true>(q, pk);
            double sk[SA][SB];
            cart_to_sh<LA, LB>(pk, sk);
#pragma unroll
            for (int ii = 0; ii < SA; ++ii)
#pragma unroll
                for (int j = 0; j < SB; ++j) sOut[lane * SA * SB + ii * SB + j] += sk[ii][j];
        }
        sOff[lane] = active ? rec.off : -1;
        __syncthreads();
        store_staged_blocks<SA * SB>(sOut, sOff, min(64, nrec - li0), accumulate, blocks);
    }
}

template <int LA, int LB>
__global__ __launch_bounds__(64) void k_herm_finish(const PPRec *recs, const double *W, int nrec,
                                                    const int *finOrder, const int *rcOff,
                                                    const int *rcList, int nChunk,
                                                    int accumulate, int addKE, double *blocks)
{
    herm_finish_wave<LA, LB>(recs, W, nrec, (int)blockIdx.x * 64, finOrder, rcOff, rcList, nChunk,
                             accumulate, addKE, blocks);
}

// The finish of many cells in one launch: wave b takes 64 records of task b.
template <int LA, int LB>
__global__ __launch_bounds__(64) void k_herm_finish_multi(const Cell3c *cells, const Task3c *tasks)
{
    const Task3c tk = tasks[blockIdx.x];
    const Cell3c a = cells[tk.cell];
    herm_finish_wave<LA, LB>(a.recs, a.W, a.nRec, tk.first, a.finOrder, a.rcOff, a.rcList,
                             a.nChunk, a.finAcc, a.finKE, a.blocks);
}

// ---------------------------------------------------------------------------
// Force integrals (computeForceIntg, forces.F90:60-898): the ket derivative
// of the fused sweep's three-centre terms.  The Hermite sums of the spin's
// coefficients and the nuclei, one order up, then per Cartesian component X
// K_AB sum_tuv D E E W (the derivative table along X, herm_dket), negated
// (the force takes minus the potential terms' and plus the nuclear terms'
// derivatives, the fused sweep's signs reversed) and added onto the store,
// which holds the two-centre part (k_two_center_fd).
template <int LA, int LB, bool NUC, bool TAB>
__global__ __launch_bounds__(64) void k_three_center_d(DevCtx c, const PPRec *recs, int n,
                                                       const int *chRec, const long long *chBeg,
                                                       const long long *chEnd,
                                                       const unsigned long long *img,
                                                       const double *coef, int addW, double *W)
{
    three_center_wave<LA, LB, NUC, TAB, false, 64, 1>(c, recs, n, (int)blockIdx.x * 64, chRec,
                                                      chBeg, chEnd, img, K3_HAM, 0, 0, coef, addW,
                                                      W);
}

template <int LA, int LB, int X>
__global__ __launch_bounds__(64) void k_herm_dfinish(const PPRec *recs, const double *W, int nrec,
                                                     const int *finOrder, const int *rcOff,
                                                     const int *rcList, int nChunk, double *blocks)
{
    constexpr int L = LA + LB, NH = nherm(L + 1);
    constexpr int NA = ncart(LA), SA = nsh(LA), SB = nsh(LB);
    constexpr bool STAGE = SA * SB <= kFinishStage;
    __shared__ double sOut[STAGE ? 64 * SA * SB + 1 : 1];
    __shared__ long long sOff[64];
    const int lane = threadIdx.x;
    const int li = (int)blockIdx.x * 64 + lane;
    const bool active = li < nrec;
    const int i = finOrder[active ? li : nrec - 1];
    const PPRec rec = recs[i];
    const PairConst q = make_pair(rec.a1, rec.a2, rec.A, rec.B);
    double Wl[NH];
#pragma unroll
    for (int e = 0; e < NH; ++e) Wl[e] = 0.0;
    if (active)
        for (int r = rcOff[i]; r < rcOff[i + 1]; ++r) {
            const int col = rcList[r];
#pragma unroll
            for (int e = 0; e < NH; ++e) Wl[e] += W[(long long)e * nChunk + col];
        }
    double E[3][LA + 1][LB + 1][L + 1], D[LA + 1][LB + 1][L + 2];
#pragma unroll
    for (int x = 0; x < 3; ++x) herm_e<LA, LB>(q.P[x] - q.A[x], q.P[x] - q.B[x], 0.5 / q.zeta, E[x]);
    herm_dket<LA, LB>(E[X], q.P[X] - q.B[X], q.a1, q.a2, q.zeta, D);
    const double kab = -exp(-q.xid2);
    double *o = blocks + rec.off;
    static_for<SB>([&](auto j_) __attribute__((always_inline)) {
        constexpr int j = decltype(j_)::value;
        double col[NA];
        static_for<NA>([&](auto a_) __attribute__((always_inline)) {
            constexpr int a = decltype(a_)::value;
            double v = 0.0;
            static_for<sh_nterm(j)>([&](auto t_) __attribute__((always_inline)) {
                constexpr int tj = decltype(t_)::value;
                constexpr int b = sh_cart(j, tj);
                v += sh_weight(j, tj) *
                     herm_dcontract_ct<LA, LB, X, cexp(a, 0), cexp(a, 1), cexp(a, 2), cexp(b, 0),
                                       cexp(b, 1), cexp(b, 2)>(E, D, Wl);
            });
            col[a] = v;
        });
#pragma unroll
        for (int ii = 0; ii < SA; ++ii) {
            double s = 0.0;
#pragma unroll
            for (int ti = 0; ti < sh_nterm(ii); ++ti) s += sh_weight(ii, ti) * col[sh_cart(ii, ti)];
            s *= kab;
            if (STAGE)
                sOut[lane * SA * SB + ii * SB + j] = s;
            else if (active)
                o[ii * SB + j] += s;
        }
    });
    if (STAGE) {
        sOff[lane] = active ? rec.off : -1;
        __syncthreads();
        store_staged_blocks<SA * SB>(sOut, sOff, min(64, nrec - (int)blockIdx.x * 64), 1, blocks);
    }
}

static size_t site_bytes(const DevCtx &c) { return k3c_site_bytes(c); }
static size_t table_bytes(const DevCtx &c) { return k3c_table_bytes(c); }

// Threads per workgroup of the tabulated three-Gaussian pass: 256 when four
// waves sharing one copy of the sites and constants fit more waves into a
// CU's LDS than single waves each holding their own (the term table, a
// column per thread, dominates: 8 KB per wave for 16 terms).
constexpr size_t kCuLds = 160 * 1024;
constexpr size_t kStaticLds3c = (2 * kMaxPotTypes * kMaxPotAlphas + 2 * kMaxPotTypes + 256) * 8;
static int tab_threads(const DevCtx &c)
{
    const size_t shared = site_bytes(c) + 16 + kStaticLds3c, perWave = table_bytes(c);
    const size_t w1 = kCuLds / (shared + perWave), w4 = 4 * (kCuLds / (shared + 4 * perWave));
    return w4 > w1 ? 256 : 64;
}

template <int LA, int LB>
static void three_one(const DevCtx &c, const ClassLists &cl, int kind, int tt, int ti,
                      const double *coef, int accumulate, double *blocks, int mixed,
                      hipStream_t s)
{
    const int k = LA * 4 + LB;
    const int n = cl.count[k];
    if (n == 0) return;
    constexpr int NH = nherm(LA + LB);
    double *W = cl.W[k];
    const int nch = cl.nChunk[k];
    const dim3 grid((nch + 63) / 64), gridRec((n + 63) / 64);
    const size_t dyn = site_bytes(c) + 16;
    const bool tab = table_bytes(c) <= kTableLdsMax;
    const size_t dynTab = site_bytes(c) + table_bytes(c) + 16;
    const bool e3c = kind != K3_NP;
    const bool nuc = kind != K3_EP && c.anyNuc;
    const bool wide = tab && tab_threads(c) == 256;
    const dim3 grid4((nch + 255) / 256);
    const size_t dynTab4 = site_bytes(c) + 4 * table_bytes(c) + 16;
    if (e3c) {
        if (wide && mixed)
            hipLaunchKernelGGL((k_three_center<LA, LB, false, true, true, 256>), grid4, dim3(256),
                               dynTab4, s, c, cl.rec[k], nch, cl.chRec[k], cl.chBeg[k], cl.chEnd[k],
                               cl.img[k], kind, tt, ti, coef, 0, W);
        else if (wide)
            hipLaunchKernelGGL((k_three_center<LA, LB, false, true, false, 256>), grid4, dim3(256),
                               dynTab4, s, c, cl.rec[k], nch, cl.chRec[k], cl.chBeg[k], cl.chEnd[k],
                               cl.img[k], kind, tt, ti, coef, 0, W);
        else if (tab && mixed)
            hipLaunchKernelGGL((k_three_center<LA, LB, false, true, true>), grid, dim3(64), dynTab,
                               s, c, cl.rec[k], nch, cl.chRec[k], cl.chBeg[k], cl.chEnd[k], cl.img[k], kind,
                               tt, ti, coef, 0, W);
        else if (tab)
            hipLaunchKernelGGL((k_three_center<LA, LB, false, true, false>), grid, dim3(64), dynTab,
                               s, c, cl.rec[k], nch, cl.chRec[k], cl.chBeg[k], cl.chEnd[k], cl.img[k], kind,
                               tt, ti, coef, 0, W);
        else
            hipLaunchKernelGGL((k_three_center<LA, LB, false, false, false>), grid, dim3(64), dyn,
                               s, c, cl.rec[k], nch, cl.chRec[k], cl.chBeg[k], cl.chEnd[k], cl.img[k], kind,
                               tt, ti, coef, 0, W);
    }
    if (nuc)   // the compacted nuclear entries: chunk u spans nOff[u] .. nOff[u + 1]
        hipLaunchKernelGGL((k_three_center<LA, LB, true, false, false>), grid, dim3(64), dyn, s, c,
                           cl.rec[k], nch, cl.chRec[k], cl.nOff[k], cl.nOff[k] + 1, cl.imgN[k], kind,
                           tt, ti, coef, e3c ? 1 : 0, W);
    if (!e3c && !nuc) launch_fill(W, (long long)NH * nch, 0.0, s);
    // classes up to d-d stage their blocks and can take the kinetic term here
    constexpr bool kFuse = nsh(LA) * nsh(LB) <= kFinishStage;
    static_assert(kFuse == ((kFusedKEClasses >> (4 * LA + LB)) & 1u), "kinetic-fused class set");
    const bool addKE = kFuse && (accumulate & kAccAddKE);
    hipLaunchKernelGGL((k_herm_finish<LA, LB>), gridRec, dim3(64), 0, s, cl.rec[k], W, n,
                       cl.finOrder[k], cl.rcOff[k], cl.rcList[k], nch,
                       addKE ? 0 : (accumulate & 1), addKE ? 1 : 0, blocks);
}

// Force sweep of class (LA, LB): comp < 0 the Hermite sums of coef and the
// nuclei into cl.Wd (one order up); comp = X the finish of component X
// added onto the store.
template <int LA, int LB>
static void three_force(const DevCtx &c, const ClassLists &cl, const double *coef, int comp,
                        double *blocks, hipStream_t s)
{
    const int k = LA * 4 + LB;
    const int n = cl.count[k];
    if (n == 0) return;
    double *W = cl.Wd[k];
    const int nch = cl.nChunk[k];
    const dim3 grid((nch + 63) / 64), gridRec((n + 63) / 64);
    if (comp < 0) {
        const size_t dyn = site_bytes(c) + 16;
        const bool tab = table_bytes(c) <= kTableLdsMax;
        const size_t dynTab = site_bytes(c) + table_bytes(c) + 16;
        if (tab)
            hipLaunchKernelGGL((k_three_center_d<LA, LB, false, true>), grid, dim3(64), dynTab, s,
                               c, cl.rec[k], nch, cl.chRec[k], cl.chBeg[k], cl.chEnd[k], cl.img[k],
                               coef, 0, W);
        else
            hipLaunchKernelGGL((k_three_center_d<LA, LB, false, false>), grid, dim3(64), dyn, s, c,
                               cl.rec[k], nch, cl.chRec[k], cl.chBeg[k], cl.chEnd[k], cl.img[k],
                               coef, 0, W);
        if (c.anyNuc)
            hipLaunchKernelGGL((k_three_center_d<LA, LB, true, false>), grid, dim3(64), dyn, s, c,
                               cl.rec[k], nch, cl.chRec[k], cl.nOff[k], cl.nOff[k] + 1, cl.imgN[k],
                               coef, 1, W);
        return;
    }
    if (comp == 0)
        hipLaunchKernelGGL((k_herm_dfinish<LA, LB, 0>), gridRec, dim3(64), 0, s, cl.rec[k], W, n,
                           cl.finOrder[k], cl.rcOff[k], cl.rcList[k], nch, blocks);
    else if (comp == 1)
        hipLaunchKernelGGL((k_herm_dfinish<LA, LB, 1>), gridRec, dim3(64), 0, s, cl.rec[k], W, n,
                           cl.finOrder[k], cl.rcOff[k], cl.rcList[k], nch, blocks);
    else
        hipLaunchKernelGGL((k_herm_dfinish<LA, LB, 2>), gridRec, dim3(64), 0, s, cl.rec[k], W, n,
                           cl.finOrder[k], cl.rcOff[k], cl.rcList[k], nch, blocks);
}

// The fused Hamiltonian sweep of class (LA, LB) over the cells of a batch:
// the three-Gaussian pass per kernel form, the nuclear pass, the finish.
template <int LA, int LB>
static void three_multi(const Multi3c &m, hipStream_t s)
{
    for (int v = 0; v < 3; ++v) {
        if (m.nE3c[v] == 0) continue;
        const dim3 g(m.nE3c[v]);
        if (v == 0)
            hipLaunchKernelGGL((k_three_center_multi<LA, LB, false, true, true>), g, dim3(64),
                               m.ldsE3c[v], s, m.ctxs, m.cells, m.e3c[v]);
        else if (v == 1)
            hipLaunchKernelGGL((k_three_center_multi<LA, LB, false, true, false>), g, dim3(64),
                               m.ldsE3c[v], s, m.ctxs, m.cells, m.e3c[v]);
        else
            hipLaunchKernelGGL((k_three_center_multi<LA, LB, false, false, false>), g, dim3(64),
                               m.ldsE3c[v], s, m.ctxs, m.cells, m.e3c[v]);
    }
    if (m.nNuc)
        hipLaunchKernelGGL((k_three_center_multi<LA, LB, true, false, false>), dim3(m.nNuc),
                           dim3(64), m.ldsNuc, s, m.ctxs, m.cells, m.nuc);
    if (m.nFin)
        hipLaunchKernelGGL((k_herm_finish_multi<LA, LB>), dim3(m.nFin), dim3(64), 0, s, m.cells,
                           m.fin);
}

#define OLCAO_CAT2(a, b) a##b
#define OLCAO_CAT(a, b) OLCAO_CAT2(a, b)
// The classes (OLCAO_LA, lb) whose bit 4 OLCAO_LA + lb is set in mask.
void OLCAO_CAT(launch_3c_la, OLCAO_LA)(const DevCtx &c, const ClassLists &cl, int kind,
                                        int tt, int ti, const double *coef, int accumulate,
                                        double *blocks, int mixed, unsigned mask, hipStream_t s)
{
    constexpr int k0 = 4 * OLCAO_LA;
    if (mask >> (k0 + 0) & 1u) three_one<OLCAO_LA, 0>(c, cl, kind, tt, ti, coef, accumulate, blocks, mixed, s);
    if (mask >> (k0 + 1) & 1u) three_one<OLCAO_LA, 1>(c, cl, kind, tt, ti, coef, accumulate, blocks, mixed, s);
    if (mask >> (k0 + 2) & 1u) three_one<OLCAO_LA, 2>(c, cl, kind, tt, ti, coef, accumulate, blocks, mixed, s);
    if (mask >> (k0 + 3) & 1u) three_one<OLCAO_LA, 3>(c, cl, kind, tt, ti, coef, accumulate, blocks, mixed, s);
}

// The force sweep's classes (OLCAO_LA, lb <= 2) in mask; bases with f
// orbitals are refused before this is reached.
void OLCAO_CAT(launch_3c_force_la, OLCAO_LA)(const DevCtx &c, const ClassLists &cl,
                                              const double *coef, int comp, double *blocks,
                                              unsigned mask, hipStream_t s)
{
    constexpr int k0 = 4 * OLCAO_LA;
    if constexpr (OLCAO_LA <= 2) {
        if (mask >> (k0 + 0) & 1u) three_force<OLCAO_LA, 0>(c, cl, coef, comp, blocks, s);
        if (mask >> (k0 + 1) & 1u) three_force<OLCAO_LA, 1>(c, cl, coef, comp, blocks, s);
        if (mask >> (k0 + 2) & 1u) three_force<OLCAO_LA, 2>(c, cl, coef, comp, blocks, s);
    } else {
        (void)c; (void)cl; (void)coef; (void)comp; (void)blocks; (void)mask; (void)s;
    }
}

// The batched class m.k = 4 OLCAO_LA + lb.
void OLCAO_CAT(launch_3c_multi_la, OLCAO_LA)(const Multi3c &m, hipStream_t s)
{
    switch (m.k - 4 * OLCAO_LA) {
    case 0: three_multi<OLCAO_LA, 0>(m, s); break;
    case 1: three_multi<OLCAO_LA, 1>(m, s); break;
    case 2: three_multi<OLCAO_LA, 2>(m, s); break;
    case 3: three_multi<OLCAO_LA, 3>(m, s); break;
    default: break;
    }
}

}  // namespace olcao
