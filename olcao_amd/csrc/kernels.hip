// HIP kernels of the OLCAO integral -> H(k)/S(k) path for gfx950.
//
//   k_two_center    one lane per primitive pair: overlap or kinetic block
//                   (overlap2CIntg / kinetic2CIntg, gaussIntegrals.f90:26,
//                   :4422) -> primitive-pair block store
//   k_three_center  one wavefront per primitive pair, lanes over potential
//                   sites: screened lattice sum over potential images and
//                   terms (electronicPE / nuclearPE / the fused PSCF sweep,
//                   integrals.F90:2533-2856, :3211-3430), accumulated in
//                   Cartesian registers, wave-reduced, transformed once
//   k_contract      one wavefront per pair-cell: primitive -> orbital
//                   contraction (multWithBasisFn1, intgSaving.F90:56-89)
//   k_bloch         one workgroup per atom pair: Bloch sum over the pair's
//                   cells, origin-shift phase and scatter into the global
//                   upper-triangular matrices (applyPhaseFactors,
//                   kPointLatticeOriginShift, saveCurrentPair,
//                   intgSaving.F90:132-655)
//   k_assemble      H = NP + KE + sum_j c_j V_j over full matrices
//                   (secularEqn.F90:147-170)
#include <algorithm>

#include "device_common.hpp"

namespace olcao {

// ---------------------------------------------------------------------------
// two-centre blocks, one lane per primitive pair; blocks of up to 81 doubles
// are staged in LDS and stored block by block with the whole wave, so the
// stores are contiguous runs.
// op OP_MV stores the negated mass-velocity block (the sign with which the
// term enters H); accumulate adds to the stored block instead of writing it.
// KIND 0: the overlap, 1: the kinetic block, 2: the other operators by op.
// One kernel per kind, so that the overlap and kinetic sweeps carry only
// their own registers (with all operators behind a run-time switch the d-d
// instance took 512 registers and spilled, at one wave per SIMD).
template <int LA, int LB, int KIND>
__global__ __launch_bounds__(64) void k_two_center(DevCtx c, const int *list, int n,
                                                   int op, int accumulate, double *blocks)
{
    constexpr int SA = nsh(LA), SB = nsh(LB);
    constexpr bool STAGE = SA * SB <= 81;
    __shared__ long long sOff[64];
    const int lane = threadIdx.x;
    const int t = blockIdx.x * 64 + lane;
    const bool active = t < n;
    const int pp = list[active ? t : n - 1];
    const PPInfo p = load_pp(c, pp);
    const PairConst q = make_pair(c.alpha[p.ti * c.maxA + p.ia],
                                  c.alpha[p.tj * c.maxA + p.ib], p.A, p.B);
    const long long off = c.ppOff[pp];
    const int nr = min(64, n - (int)blockIdx.x * 64);
    if constexpr (KIND <= 1 && SA * SB > 45 && STAGE) {
        // d-d overlap and kinetic straight in the harmonics, staged by halves
        // of the harmonic rows: no Cartesian block in registers and half the
        // LDS of a whole block, so two waves per SIMD instead of one
        TwoCenterTables<LA, LB, KIND == 1> tb;
        two_center_tables<LA, LB, KIND == 1>(q, tb);
        constexpr int RA = (SA + 1) / 2, NE1 = RA * SB, NE2 = (SA - RA) * SB;
        __shared__ double sHalf[64 * NE1 + 1];
        two_center_sh_rows<LA, LB, KIND == 1, 0, RA>(
            tb, [&](int i, int j, double v) { sHalf[lane * NE1 + i * SB + j] = v; });
        sOff[lane] = off;
        __syncthreads();
        store_staged_blocks<NE1>(sHalf, sOff, nr, accumulate, blocks);
        __syncthreads();
        two_center_sh_rows<LA, LB, KIND == 1, RA, SA>(
            tb, [&](int i, int j, double v) { sHalf[lane * NE2 + (i - RA) * SB + j] = v; });
        sOff[lane] = off + NE1;
        __syncthreads();
        store_staged_blocks<NE2>(sHalf, sOff, nr, accumulate, blocks);
    } else {
        __shared__ double sOut[STAGE ? 64 * SA * SB + 1 : 1];
        double pc[ncart(LA)][ncart(LB)];
        if constexpr (KIND == 0) {
            two_center_cart<LA, LB, false>(q, pc);
        } else if constexpr (KIND == 1) {
            two_center_cart<LA, LB, true>(q, pc);
        } else {
            if (op >= OP_KO)
                kov_cart<LA, LB>(q, c.kq, op - OP_KO, pc);
            else if (op == OP_MV)
                massvel_cart<LA, LB>(q, pc);
            else if (op >= OP_MM)
                vec_cart<LA, LB, false>(q, op - OP_MM, c.dipC, pc);
            else if (op >= OP_DM)
                vec_cart<LA, LB, true>(q, op - OP_DM, c.dipC, pc);
            else
                two_center_cart<LA, LB, false>(q, pc);   // not reached: OL and KE have their kinds
        }
        double sh[SA][SB];
        cart_to_sh<LA, LB>(pc, sh);
        const double sg = op == OP_MV ? -1.0 : 1.0;
        if (STAGE) {
#pragma unroll
            for (int i = 0; i < SA; ++i)
#pragma unroll
                for (int j = 0; j < SB; ++j) sOut[lane * SA * SB + i * SB + j] = sg * sh[i][j];
            sOff[lane] = off;
            __syncthreads();
            store_staged_blocks<SA * SB>(sOut, sOff, nr, accumulate, blocks);
        } else if (active) {
            double *o = blocks + off;
#pragma unroll
            for (int i = 0; i < SA; ++i)
#pragma unroll
                for (int j = 0; j < SB; ++j)
                    o[i * SB + j] = accumulate ? o[i * SB + j] + sg * sh[i][j] : sg * sh[i][j];
        }
    }
}

// The force integrals' two-centre part along component comp (classes up to
// d-d): minus the ket derivative of the kinetic block plus c.fE2 (2
// totalEnergy) times that of the overlap block, i.e. -2 totalEnergy times
// the momentum block (computeForceIntg, forces.F90:648-730).  Its own kernel
// so that k_two_center, on every iteration's path, keeps its registers.
template <int LA, int LB>
__global__ __launch_bounds__(64) void k_two_center_fd(DevCtx c, const int *list, int n, int comp,
                                                      double *blocks)
{
    constexpr int SA = nsh(LA), SB = nsh(LB);
    __shared__ double sOut[64 * SA * SB + 1];
    __shared__ long long sOff[64];
    const int lane = threadIdx.x;
    const int t = blockIdx.x * 64 + lane;
    const bool active = t < n;
    const int pp = list[active ? t : n - 1];
    const PPInfo p = load_pp(c, pp);
    const PairConst q = make_pair(c.alpha[p.ti * c.maxA + p.ia],
                                  c.alpha[p.tj * c.maxA + p.ib], p.A, p.B);
    double pc[ncart(LA)][ncart(LB)];
    dket_cart<LA, LB>(q, comp, -1.0, c.fE2, pc);
    double sh[SA][SB];
    cart_to_sh<LA, LB>(pc, sh);
#pragma unroll
    for (int i = 0; i < SA; ++i)
#pragma unroll
        for (int j = 0; j < SB; ++j) sOut[lane * SA * SB + i * SB + j] = sh[i][j];
    sOff[lane] = c.ppOff[pp];
    __syncthreads();
    store_staged_blocks<SA * SB>(sOut, sOff, min(64, n - (int)blockIdx.x * 64), 0, blocks);
}

// ---------------------------------------------------------------------------
// contraction: X(s1,s2) = sum over the pair-cell's primitive pairs of
//   C1(a1,s1) * block(lm(s1), lm(s2)) * C2(a2,s2)
// The primitive pairs of a pair-cell are cut into segments of at most
// kSegPP pairs (an on-site pair-cell has several hundred, a distant one a
// few), one wavefront per segment, so that no pair-cell serialises the
// launch.  Lane q holds the q-th pair's record; each lane sums up to
// kContractOut outputs at a time, walking the pairs in order in batches of
// kQBatch whose block loads are all issued before the first is used (the
// lanes of one pair read neighbouring words of its block).  The two atom
// types' coefficient tables C(a, s) are staged in LDS.  A single-segment
// pair-cell writes X directly; otherwise each segment writes a partial block
// and k_contract_sum adds the partials in segment order (deterministic).
constexpr int kContractOut = 2;
constexpr int kQBatch = 16;

__global__ __launch_bounds__(64) void k_contract_seg(DevCtx c, const ContractSeg *segs, int nSeg,
                                                     const double *blocks, double *X,
                                                     double *partial)
{
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const int w = blockIdx.x;
    if (w >= nSeg) return;
    const ContractSeg sg = segs[w];
    const int pc = sg.pc, p0 = sg.p0, np = sg.np, part = sg.part;
    const int ti = sg.ti, tj = sg.tj, nt1 = sg.nt1, nt2 = sg.nt2;
    const int ms = c.maxS, ma = c.maxA;
    double *sC1 = lds, *sC2 = lds + ms * ma;
    const double *bseg = blocks + sg.base;
    const int myAB = lane < np ? c.ppAB[p0 + lane] : 0;
    const int myRel = lane < np ? (int)(c.ppOff[p0 + lane] - sg.base) : 0;
    for (int e = lane; e < nt1 * ma; e += 64) sC1[e] = c.basis[(long long)ti * ms * ma + e];
    for (int e = lane; e < nt2 * ma; e += 64) sC2[e] = c.basis[(long long)tj * ms * ma + e];
    __syncthreads();
    double *dst = (part < 0 ? X + (long long)pc * ms * ms : partial + (long long)part * ms * ms);
    const int nout = nt1 * nt2;
    for (int g = 0; g < nout; g += 64 * kContractOut) {
        double acc[kContractOut];
        int lm1v[kContractOut], lm2v[kContractOut], dv[kContractOut];
        const double *c1v[kContractOut], *c2v[kContractOut];
#pragma unroll
        for (int r = 0; r < kContractOut; ++r) {
            acc[r] = 0.0;
            const int o = g + r * 64 + lane;
            const bool ok = o < nout;
            const int s1 = ok ? o / nt2 : 0, s2 = ok ? o - s1 * nt2 : 0;
            lm1v[r] = ok ? c.lmIndex[ti * ms + s1] - 1 : 99;
            lm2v[r] = ok ? c.lmIndex[tj * ms + s2] - 1 : 99;
            dv[r] = ok ? s1 * ms + s2 : -1;
            c1v[r] = sC1 + s1 * ma;
            c2v[r] = sC2 + s2 * ma;
        }
        for (int q0 = 0; q0 < np; q0 += kQBatch) {
            double v[kQBatch][kContractOut];
#pragma unroll
            for (int u = 0; u < kQBatch; ++u) {
                const int ab = __builtin_amdgcn_readlane(myAB, q0 + u);
                const int rel = __builtin_amdgcn_readlane(myRel, q0 + u);
                const int n1 = nsh(pp_la(ab)), n2 = nsh(pp_lb(ab));
#pragma unroll
                for (int r = 0; r < kContractOut; ++r) {
                    const bool ok = q0 + u < np && lm1v[r] < n1 && lm2v[r] < n2;
                    v[u][r] = ok ? bseg[rel + lm1v[r] * n2 + lm2v[r]] : 0.0;
                }
            }
#pragma unroll
            for (int u = 0; u < kQBatch; ++u) {
                if (q0 + u >= np) break;
                const int ab = __builtin_amdgcn_readlane(myAB, q0 + u);
                const int n1 = nsh(pp_la(ab)), n2 = nsh(pp_lb(ab));
                const int a1 = pp_a1(ab), a2 = pp_a2(ab);
#pragma unroll
                for (int r = 0; r < kContractOut; ++r)
                    if (lm1v[r] < n1 && lm2v[r] < n2) acc[r] += c1v[r][a1] * v[u][r] * c2v[r][a2];
            }
        }
#pragma unroll
        for (int r = 0; r < kContractOut; ++r)
            if (dv[r] >= 0) dst[dv[r]] = acc[r];
    }
}

// X of the multi-segment pair-cells: sum of the partials in segment order.
__global__ __launch_bounds__(64) void k_contract_sum(DevCtx c, const int4 *multi, int nMulti,
                                                     const double *partial, double *X)
{
    const int m = blockIdx.x;
    if (m >= nMulti) return;
    const int4 e = multi[m];   // pc, first partial, segment count
    const int pc = e.x;
    const int pr = c.pcPair[pc];
    const int ti = c.atomType[c.pairI[pr]], tj = c.atomType[c.pairJ[pr]];
    const int nt1 = c.nVale[ti] + c.nCore[ti], nt2 = c.nVale[tj] + c.nCore[tj];
    const int ms = c.maxS;
    for (int o = threadIdx.x; o < nt1 * nt2; o += 64) {
        const int s1 = o / nt2, s2 = o - s1 * nt2;
        const long long d = (long long)s1 * ms + s2;
        double v = 0.0;
        for (int q = 0; q < e.z; ++q) v += partial[(long long)(e.y + q) * ms * ms + d];
        X[(long long)pc * ms * ms + d] = v;
    }
}

// ---------------------------------------------------------------------------
// Bloch sum + origin shift + scatter.  One wavefront per (atom pair, tile of
// KT k-points): the pair's contracted blocks X(cell) and the phases
// exp(i k.L_cell) of the tile are staged through LDS kBlochCells cells at a
// time, each lane accumulating up to kBlochOut outputs (k, s1, s2) with s1
// fastest, so that the scatter into the column-major matrices is
// contiguous.  Cells are summed in the pair's cell order.
constexpr int kBlochCells = 16;
constexpr int kBlochOut = 8;

__global__ __launch_bounds__(64) void k_bloch(DevCtx c, const double *X, double *vv,
                                              double *cv, double *cc, int gamma, int mode,
                                              int KT, int nTiles)
{
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const int pr = blockIdx.x / nTiles, tile = blockIdx.x - pr * nTiles;
    if (pr >= c.nPairs) return;
    const int i = c.pairI[pr], j = c.pairJ[pr];
    const int ti = c.atomType[i], tj = c.atomType[j];
    const int nv1 = c.nVale[ti], nv2 = c.nVale[tj];
    const int nt1 = nv1 + c.nCore[ti], nt2 = nv2 + c.nCore[tj];
    const int v1 = c.cumulVale[i], v2 = c.cumulVale[j];
    const int c1 = c.cumulCore[i], c2 = c.cumulCore[j];
    const long long n = c.valeDim, nc = c.coreDim;
    const int p0 = c.pairPc[pr], p1 = c.pairPc[pr + 1];
    const int ms = c.maxS;
    const int nk = gamma ? 1 : c.nk;
    const int k0 = tile * KT, nkt = min(nk, k0 + KT) - k0;
    if (nkt <= 0) return;
    const int per = nt1 * nt2;
    const int nout = nkt * per;
    double *sX = lds;                          // [kBlochCells][per]
    double *sP = sX + kBlochCells * per;       // [kBlochCells][nkt][2]
    const double lv[3] = {c.pairLv[3 * pr], c.pairLv[3 * pr + 1], c.pairLv[3 * pr + 2]};
    for (int g = 0; g < nout; g += 64 * kBlochOut) {
        double re[kBlochOut], im[kBlochOut];
#pragma unroll
        for (int r = 0; r < kBlochOut; ++r) re[r] = im[r] = 0.0;
        for (int q0 = p0; q0 < p1; q0 += kBlochCells) {
            const int ncl = min(kBlochCells, p1 - q0);
            for (int q = lane; q < ncl * per; q += 64) {
                const int cl = q / per, e = q - cl * per;
                const int s1 = e / nt2, s2 = e - s1 * nt2;
                sX[q] = X[(long long)(q0 + cl) * ms * ms + s1 * ms + s2];
            }
            if (!gamma)
                for (int q = lane; q < ncl * nkt; q += 64) {
                    const int cl = q / nkt, kk = q - cl * nkt;
                    const double *ph = c.phase + 2 * ((long long)c.pcCell[q0 + cl] * c.nk + k0 + kk);
                    sP[2 * q] = ph[0];
                    sP[2 * q + 1] = ph[1];
                }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kBlochOut; ++r) {
                const int o = g + r * 64 + lane;
                if (o >= nout) continue;
                const int kk = o / per, e = o - kk * per;
                const int s2 = e / nt1, s1 = e - s2 * nt1;
                const double *xs = sX + s1 * nt2 + s2;
                if (gamma) {
                    for (int cl = 0; cl < ncl; ++cl) re[r] += xs[cl * per];
                } else {
                    const double *ps = sP + 2 * kk;
                    for (int cl = 0; cl < ncl; ++cl) {
                        const double x = xs[cl * per];
                        re[r] += ps[2 * cl * nkt] * x;
                        im[r] += ps[2 * cl * nkt + 1] * x;
                    }
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < kBlochOut; ++r) {
            const int o = g + r * 64 + lane;
            if (o >= nout) continue;
            const int kk = o / per, e = o - kk * per;
            const int s2 = e / nt1, s1 = e - s2 * nt1;
            bloch_store(c, i, j, k0 + kk, s1, s2, nv1, nv2, v1, v2, c1, c2, lv, re[r], im[r], gamma,
                        mode, vv, cv, cc);
        }
    }
}

__global__ void k_zero_words(unsigned *p, long long n)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        p[t] = 0u;
}

__global__ void k_fill(double *p, long long n, double v)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        p[t] = v;
}

__global__ void k_add(double *y, const double *x, long long n)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        y[t] += x[t];
}

// H = NP + KE (+ MV) + sum_j c_j V_j on the upper triangle of every k
// matrix, in the reference's order; MV (the negated mass-velocity term,
// present when rel = 1) and a c_j == 0 are added unscaled
// (readPackedMatrixAccum, matrixSubs.F90:369-373; secularEqn.F90:147-170).
__global__ void k_assemble(long long total, int n, const double *NP, const double *KE,
                           const double *MV, const double *const *V, const double *coef,
                           int nTerms, double *H, int cplx)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long e = t / cplx;
        const long long inK = e % ((long long)n * n);
        const int r = (int)(inK % n), col = (int)(inK / n);
        if (r > col) {
            H[t] = 0.0;
            continue;
        }
        double v = NP[t] + KE[t];
        if (MV) v = v + MV[t];
        for (int j = 0; j < nTerms; ++j) {
            const double cj = coef[j];
            v = v + (cj == 0.0 ? V[j][t] : V[j][t] * cj);
        }
        H[t] = v;
    }
}

// Lower triangle from the upper one: A(r,c) = conj(A(c,r)) for r > c.
__global__ void k_herm_fill(double *A, int n, int nb, int cplx)
{
    const long long per = (long long)n * n;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < per * nb;
         t += (long long)gridDim.x * blockDim.x) {
        const long long k = t / per, e = t - k * per;
        const int r = (int)(e % n), col = (int)(e / n);
        if (r <= col) continue;
        const long long src = k * per + (long long)col + (long long)r * n;
        if (cplx == 2) {
            A[2 * t] = A[2 * src];
            A[2 * t + 1] = -A[2 * src + 1];
        } else {
            A[t] = A[src];
        }
    }
}

// first numStates eigenvector columns of each k, compacted
__global__ void k_extract(const double *A, int n, int nk, int ns, int cplx, double *out)
{
    const long long total = (long long)nk * ns * n * cplx;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long per = (long long)ns * n * cplx;
        const long long k = t / per, rem = t - k * per;
        out[t] = A[k * (long long)n * n * cplx + rem];
    }
}

// ---------------------------------------------------------------------------
// launchers
template <int LA, int LB>
static void two_one(const DevCtx &c, const ClassLists &cl, int op, int acc, double *blocks,
                    hipStream_t s)
{
    const int k = LA * 4 + LB;
    const int n = cl.count[k];
    if (n == 0) return;
    const dim3 grid((n + 63) / 64);
    if (op == OP_OL)
        hipLaunchKernelGGL((k_two_center<LA, LB, 0>), grid, dim3(64), 0, s, c, cl.idx[k], n, op, acc,
                           blocks);
    else if (op == OP_KE)
        hipLaunchKernelGGL((k_two_center<LA, LB, 1>), grid, dim3(64), 0, s, c, cl.idx[k], n, op, acc,
                           blocks);
    else
        hipLaunchKernelGGL((k_two_center<LA, LB, 2>), grid, dim3(64), 0, s, c, cl.idx[k], n, op, acc,
                           blocks);
}

void launch_two_center(const DevCtx &c, const ClassLists &cl, int op, int acc, double *blocks,
                       unsigned mask, hipStream_t s)
{
#define OLCAO_TWO(la, lb) \
    if (mask >> (4 * la + lb) & 1u) two_one<la, lb>(c, cl, op, acc, blocks, s);
    OLCAO_TWO(0, 0) OLCAO_TWO(0, 1) OLCAO_TWO(0, 2) OLCAO_TWO(0, 3)
    OLCAO_TWO(1, 0) OLCAO_TWO(1, 1) OLCAO_TWO(1, 2) OLCAO_TWO(1, 3)
    OLCAO_TWO(2, 0) OLCAO_TWO(2, 1) OLCAO_TWO(2, 2) OLCAO_TWO(2, 3)
    OLCAO_TWO(3, 0) OLCAO_TWO(3, 1) OLCAO_TWO(3, 2) OLCAO_TWO(3, 3)
#undef OLCAO_TWO
}

void launch_two_center_fd(const DevCtx &c, const ClassLists &cl, int comp, double *blocks,
                          unsigned mask, hipStream_t s)
{
#define OLCAO_FD(la, lb)                                                                       \
    if ((mask >> (4 * la + lb) & 1u) && cl.count[4 * la + lb] > 0)                             \
        hipLaunchKernelGGL((k_two_center_fd<la, lb>), dim3((cl.count[4 * la + lb] + 63) / 64), \
                           dim3(64), 0, s, c, cl.idx[4 * la + lb], cl.count[4 * la + lb], comp,  \
                           blocks);
    OLCAO_FD(0, 0) OLCAO_FD(0, 1) OLCAO_FD(0, 2)
    OLCAO_FD(1, 0) OLCAO_FD(1, 1) OLCAO_FD(1, 2)
    OLCAO_FD(2, 0) OLCAO_FD(2, 1) OLCAO_FD(2, 2)
#undef OLCAO_FD
}

void launch_3c_la0(const DevCtx &, const ClassLists &, int, int, int, const double *, int,
                   double *, int, unsigned, hipStream_t);
void launch_3c_la1(const DevCtx &, const ClassLists &, int, int, int, const double *, int,
                   double *, int, unsigned, hipStream_t);
void launch_3c_la2(const DevCtx &, const ClassLists &, int, int, int, const double *, int,
                   double *, int, unsigned, hipStream_t);
void launch_3c_la3(const DevCtx &, const ClassLists &, int, int, int, const double *, int,
                   double *, int, unsigned, hipStream_t);

void launch_3c_force_la0(const DevCtx &, const ClassLists &, const double *, int, double *,
                         unsigned, hipStream_t);
void launch_3c_force_la1(const DevCtx &, const ClassLists &, const double *, int, double *,
                         unsigned, hipStream_t);
void launch_3c_force_la2(const DevCtx &, const ClassLists &, const double *, int, double *,
                         unsigned, hipStream_t);

void launch_three_center_force(const DevCtx &c, const ClassLists &cl, const double *coef, int comp,
                               double *blocks, unsigned mask, hipStream_t s)
{
    if (mask & 0x000fu) launch_3c_force_la0(c, cl, coef, comp, blocks, mask, s);
    if (mask & 0x00f0u) launch_3c_force_la1(c, cl, coef, comp, blocks, mask, s);
    if (mask & 0x0f00u) launch_3c_force_la2(c, cl, coef, comp, blocks, mask, s);
}

void launch_3c_multi_la0(const Multi3c &, hipStream_t);
void launch_3c_multi_la1(const Multi3c &, hipStream_t);
void launch_3c_multi_la2(const Multi3c &, hipStream_t);
void launch_3c_multi_la3(const Multi3c &, hipStream_t);

void launch_three_center_multi(const Multi3c &m, hipStream_t s)
{
    switch (m.k / 4) {
    case 0: launch_3c_multi_la0(m, s); break;
    case 1: launch_3c_multi_la1(m, s); break;
    case 2: launch_3c_multi_la2(m, s); break;
    default: launch_3c_multi_la3(m, s); break;
    }
}

void launch_three_center(const DevCtx &c, const ClassLists &cl, int kind, int tt, int ti,
                         const double *coef, int accumulate, double *blocks, int mixed,
                         unsigned mask, hipStream_t s)
{
    if (mask & 0x000fu) launch_3c_la0(c, cl, kind, tt, ti, coef, accumulate, blocks, mixed, mask, s);
    if (mask & 0x00f0u) launch_3c_la1(c, cl, kind, tt, ti, coef, accumulate, blocks, mixed, mask, s);
    if (mask & 0x0f00u) launch_3c_la2(c, cl, kind, tt, ti, coef, accumulate, blocks, mixed, mask, s);
    if (mask & 0xf000u) launch_3c_la3(c, cl, kind, tt, ti, coef, accumulate, blocks, mixed, mask, s);
}

// ---------------------------------------------------------------------------
// Per-pair image lists of the three-centre sweeps, built once per geometry.
//
// For primitive pair i (one thread each) every image C = site + L of every
// potential site inside the widest range any sweep kind can use is visited
// (a box in fractional coordinates about the screening centre P, then the
// exact distance), and the reference's range tests are applied term by term
// exactly as the sweeps apply them (integrals.F90:2625-2663 electronic and
// nuclear sweeps, :3211-3430 the fused sweep; the term loop stops at the
// first term out of range because the exponents ascend).  Pass 1 counts the
// entries and the fused-sweep items, pass 2 writes the entries (one kernel
// with a run-time switch, so both passes run the same arithmetic; pass 2
// still never writes past its record's segment and flags any mismatch).
__global__ __launch_bounds__(64) void k_image_list(DevCtx c, const PPRec *recs, int n, int fill,
                                                   int *nEnt, int *nItems, const long long *off,
                                                   unsigned long long *ent, int *overflow)
{
    __shared__ double tAlpha[kMaxPotTypes * kMaxPotAlphas];
    __shared__ double tZ[kMaxPotTypes], tNuc[kMaxPotTypes];
    __shared__ int tNA[kMaxPotTypes];
    extern __shared__ double dyn[];
    double *sitePos = dyn;
    int *siteType = reinterpret_cast<int *>(dyn + 3 * c.nPotSites);
    const int lane = threadIdx.x;
    for (int q = lane; q < c.nPotSites; q += 64) {
        sitePos[3 * q] = c.potPos[3 * q];
        sitePos[3 * q + 1] = c.potPos[3 * q + 1];
        sitePos[3 * q + 2] = c.potPos[3 * q + 2];
        siteType[q] = c.potSiteType[q];
    }
    for (int q = lane; q < c.nPotTypes * kMaxPotAlphas; q += 64) {
        const int t = q / kMaxPotAlphas, a = q - t * kMaxPotAlphas;
        tAlpha[q] = a < c.potNA[t] ? c.potAlpha[t * c.maxPA + a] : 0.0;
    }
    for (int t = lane; t < c.nPotTypes; t += 64) {
        tZ[t] = c.potZ[t];
        tNuc[t] = c.potNuc[t];
        tNA[t] = c.potNA[t];
    }
    __syncthreads();
    const int i = blockIdx.x * 64 + lane;
    if (i >= n) return;
    const PPRec rec = recs[i];
    // screening centre and ranges (element-level exponents, gaussRelations.f90)
    const double acen = rec.ea1 / (rec.ea1 + rec.ea2);
    const double adist = c.tau / rec.ea2 / acen;
    double P[3];
    for (int x = 0; x < 3; ++x) P[x] = acen * (rec.A[x] - rec.B[x]) + rec.B[x];
    const double shrink = 1.0 - rec.s2 / adist;
    const double budget = c.tau * shrink;
    const double a12 = rec.a1 + rec.a2, e12 = rec.ea1 + rec.ea2, tau = c.tau;
    // fused sweep: tau (1 - s^2/AD) (a1 + a2 + a3) / (a1 + a2) / a3 (integral exponents);
    // separate sweeps: (1 - s^2/AD) tau (e1 + e2 + a3) / (e1 + e2) / a3
    auto tadHam = [&](double a3) { return budget * (a12 + a3) / a12 / a3; };
    auto tadEp = [&](double a3) { return shrink * (tau * (e12 + a3) / e12 / a3); };
    double invn[3];
    for (int q = 0; q < 3; ++q)
        invn[q] = sqrt(c.inv[3 * q] * c.inv[3 * q] + c.inv[3 * q + 1] * c.inv[3 * q + 1] +
                       c.inv[3 * q + 2] * c.inv[3 * q + 2]);
    long long w = fill ? off[i] : 0;
    const long long wend = fill ? off[i + 1] : 0;
    int cnt = 0, items = 0, bad = 0;
    for (int st = 0; st < c.nPotSites; ++st) {
        const int t = siteType[st];
        const int na = tNA[t];
        const bool zok = tZ[t] >= kSmallThresh;
        double R2 = -1.0;
        if (na > 0) R2 = fmax(tadHam(tAlpha[t * kMaxPotAlphas]), tadEp(tAlpha[t * kMaxPotAlphas]));
        if (zok) R2 = fmax(R2, fmax(tadHam(tNuc[t]), tadEp(tNuc[t])));
        if (!(R2 >= 0.0)) continue;
        const double r = sqrt(R2) * (1.0 + 1e-12) + 1e-12;
        double cov[3];
        for (int x = 0; x < 3; ++x) cov[x] = P[x] - sitePos[3 * st + x];
        int lo[3], hi[3];
        for (int q = 0; q < 3; ++q) {
            const double f = c.inv[3 * q] * cov[0] + c.inv[3 * q + 1] * cov[1] +
                             c.inv[3 * q + 2] * cov[2];
            lo[q] = (int)ceil(f - r * invn[q]);
            hi[q] = (int)floor(f + r * invn[q]);
        }
        for (int nz = lo[2]; nz <= hi[2]; ++nz)
            for (int ny = lo[1]; ny <= hi[1]; ++ny)
                for (int nx = lo[0]; nx <= hi[0]; ++nx) {
                    double Lv[3];
                    for (int x = 0; x < 3; ++x)
                        Lv[x] = nx * c.vec[x] + ny * c.vec[3 + x] + nz * c.vec[6 + x];
                    const double r0 = cov[0] - Lv[0], r1 = cov[1] - Lv[1], r2 = cov[2] - Lv[2];
                    const double d2 = r0 * r0 + r1 * r1 + r2 * r2;
                    if (d2 > R2) continue;
                    int nHam = 0, nEp = 0;
                    while (nHam < na && d2 <= tadHam(tAlpha[t * kMaxPotAlphas + nHam])) ++nHam;
                    while (nEp < na && d2 <= tadEp(tAlpha[t * kMaxPotAlphas + nEp])) ++nEp;
                    const int nucHam = zok && d2 <= tadHam(tNuc[t]) ? 1 : 0;
                    const int nucNp = zok && d2 <= tadEp(tNuc[t]) ? 1 : 0;
                    if (nHam + nEp + nucHam + nucNp == 0) continue;
                    ++cnt;
                    items += nHam;
                    if (fill) {
                        if (nx < -kImgLatOff || nx >= kImgLatOff || ny < -kImgLatOff ||
                            ny >= kImgLatOff || nz < -kImgLatOff || nz >= kImgLatOff)
                            bad |= 1;
                        if (w < wend)
                            ent[w++] = img_pack(st, nx, ny, nz, nHam, nEp, nucHam, nucNp);
                        else
                            bad |= 2;
                    }
                }
    }
    if (!fill) {
        nEnt[i] = cnt;
        nItems[i] = items;
    } else {
        if (w != wend) bad |= 2;
        if (bad) atomicOr(overflow, bad);
    }
}

static size_t site_lds(const DevCtx &c)
{
    return (size_t)c.nPotSites * (3 * sizeof(double) + sizeof(int)) + 16;
}

void launch_image_count(const DevCtx &c, const PPRec *recs, int n, int *nEnt, int *nItems,
                        hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_image_list, dim3((n + 63) / 64), dim3(64), site_lds(c), s, c, recs, n,
                       0, nEnt, nItems, nullptr, nullptr, nullptr);
}

void launch_image_fill(const DevCtx &c, const PPRec *recs, int n, const long long *off,
                       unsigned long long *ent, int *overflow, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_image_list, dim3((n + 63) / 64), dim3(64), site_lds(c), s, c, recs, n,
                       1, nullptr, nullptr, off, ent, overflow);
}

// One thread per chunk: the chunk's entries with nuclear work, in entry
// order (the nuclear passes read only these; most image entries lie outside
// the nuclear range and carry three-Gaussian terms only).
__global__ __launch_bounds__(256) void k_nuc_compact(const unsigned long long *img,
                                                     const long long *beg, const long long *end,
                                                     int nc, int *cnt, const long long *nOff,
                                                     unsigned long long *outN)
{
    const int u = blockIdx.x * 256 + threadIdx.x;
    if (u >= nc) return;
    long long w = outN ? nOff[u] : 0;
    int k = 0;
    for (long long p = beg[u]; p < end[u]; ++p) {
        const unsigned long long e = img[p];
        if (img_nucham(e) | img_nucnp(e)) {
            if (outN) outN[w + k] = e;
            ++k;
        }
    }
    if (!outN) cnt[u] = k;
}

void launch_nuc_compact(const unsigned long long *img, const long long *beg, const long long *end,
                        int nc, int *cnt, const long long *nOff, unsigned long long *outN,
                        hipStream_t s)
{
    if (nc <= 0) return;
    hipLaunchKernelGGL(k_nuc_compact, dim3((nc + 255) / 256), dim3(256), 0, s, img, beg, end, nc,
                       cnt, nOff, outN);
}

void launch_contract(const DevCtx &c, const ContractPlan &cp, const double *blocks, double *X,
                     hipStream_t s)
{
    if (cp.nSeg == 0) return;
    const size_t lds = (size_t)(2 * c.maxS * c.maxA) * sizeof(double);
    hipLaunchKernelGGL(k_contract_seg, dim3(cp.nSeg), dim3(64), lds, s, c, cp.segs, cp.nSeg,
                       blocks, X, cp.partial);
    if (cp.nMulti > 0)
        hipLaunchKernelGGL(k_contract_sum, dim3(cp.nMulti), dim3(64), 0, s, c, cp.multi,
                           cp.nMulti, cp.partial, X);
}

void launch_bloch_scatter(const DevCtx &c, const double *X, double *vv, double *cv,
                          double *cc, int gamma, int mode, hipStream_t s)
{
    if (c.nPairs == 0) return;
    // k tile: about 512 outputs per wavefront, and at least ~1024 wavefronts
    const int nk = gamma ? 1 : c.nk;
    const int per = c.maxS * c.maxS;
    int KT = std::max(1, std::min(nk, 64 * kBlochOut / std::max(per, 1)));
    while (KT > 1 && (long long)c.nPairs * ((nk + KT - 1) / KT) < 1024) KT = (KT + 1) / 2;
    const int nTiles = (nk + KT - 1) / KT;
    const size_t lds = ((size_t)kBlochCells * per + 2 * (size_t)kBlochCells * KT) * sizeof(double);
    hipLaunchKernelGGL(k_bloch, dim3(c.nPairs * nTiles), dim3(64), lds, s, c, X, vv, cv, cc,
                       gamma, mode, KT, nTiles);
}

void launch_zero_words(void *p, long long words, hipStream_t s)
{
    if (words <= 0) return;
    long long blocks = (words + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)blocks), dim3(256), 0, s,
                       static_cast<unsigned *>(p), words);
}

void launch_fill(double *p, long long n, double v, hipStream_t s)
{
    if (n <= 0) return;
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

void launch_add(double *y, const double *x, long long n, hipStream_t s)
{
    if (n <= 0) return;
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_add, dim3((unsigned)blocks), dim3(256), 0, s, y, x, n);
}

void launch_assemble(long long total, int n, const double *NP, const double *KE,
                     const double *MV, const double *const *V, const double *coef, int nTerms,
                     double *H, int cplx, hipStream_t s)
{
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_assemble, dim3((unsigned)blocks), dim3(256), 0, s, total, n, NP, KE,
                       MV, V, coef, nTerms, H, cplx);
}

// After orthogonalisation: keep the upper triangle, zero the rest, and zero
// components below 1e-10 in magnitude as makeValeVale packs them
// (intgOrtho.F90:142-190).
__global__ void k_ortho_finish(double *M, int n, int nb, int cplx)
{
    const long long per = (long long)n * n;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < per * nb * cplx;
         t += (long long)gridDim.x * blockDim.x) {
        const long long e = (t / cplx) % per;
        const int r = (int)(e % n), col = (int)(e / n);
        if (r > col || fabs(M[t]) < 1.0e-10) M[t] = 0.0;
    }
}

void launch_ortho_finish(double *M, int n, int nb, int cplx, hipStream_t s)
{
    long long total = (long long)n * n * nb * cplx;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_ortho_finish, dim3((unsigned)blocks), dim3(256), 0, s, M, n, nb, cplx);
}

// nb identity matrices of order n (cplx = 2: complex entries)
__global__ void k_identity(double *M, int n, int nb, int cplx)
{
    const long long per = (long long)n * n;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < per * nb * cplx;
         t += (long long)gridDim.x * blockDim.x) {
        const long long e = (t / cplx) % per;
        const bool re = t % cplx == 0;
        M[t] = (re && e % n == e / n) ? 1.0 : 0.0;
    }
}

void launch_identity(double *M, int n, int nb, int cplx, hipStream_t s)
{
    long long total = (long long)n * n * nb * cplx;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_identity, dim3((unsigned)blocks), dim3(256), 0, s, M, n, nb, cplx);
}

void launch_herm_fill(double *A, int n, int nb, int cplx, hipStream_t s)
{
    long long total = (long long)n * n * nb;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_herm_fill, dim3((unsigned)blocks), dim3(256), 0, s, A, n, nb, cplx);
}

void launch_extract_evecs(const double *A, int n, int nk, int ns, int cplx, double *out,
                          hipStream_t s)
{
    long long total = (long long)nk * ns * n * cplx;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_extract, dim3((unsigned)blocks), dim3(256), 0, s, A, n, nk, ns, cplx,
                       out);
}

// Y(:, j, k) = f(j, k) X(:, j, k) for the first ns eigenvectors of each
// k-point; X and Y are n x n column-major per k (cplx = 2 interleaved).
__global__ void k_scale_states(const double *X, const double *f, int n, int ns, int nk, int cplx,
                               double *Y)
{
    const long long total = (long long)nk * ns * n * cplx;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long e = t / cplx;
        const int r = (int)(e % n);
        const long long jk = e / n;
        const int j = (int)(jk % ns), k = (int)(jk / ns);
        const size_t off = ((size_t)k * n * n + (size_t)j * n + r) * cplx + t % cplx;
        Y[off] = f[(size_t)k * ns + j] * X[off];
    }
}

void launch_scale_states(const double *X, const double *f, int n, int ns, int nk, int cplx,
                         double *Y, hipStream_t s)
{
    long long total = (long long)nk * ns * n * cplx;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_scale_states, dim3((unsigned)blocks), dim3(256), 0, s, X, f, n, ns, nk,
                       cplx, Y);
}

// Partial traces Tr[M_m rho] over the upper triangles of nk matrices, rho =
// R0 + sgn R1 (R1 may be null): off-diagonal elements count twice with
// re*re + im*im, diagonal elements once with their real parts, as
// matrixElementMult sums a packed pair (matrixSubs.F90:66-112).  Block b of
// matrix m writes partial[m * gridDim.x + b]; k_trace_sum adds them in order.
__global__ __launch_bounds__(256) void k_trace_part(const double *const *M, const double *R0,
                                                    const double *R1, double sgn, int n, int nk,
                                                    int cplx, double *partial)
{
    const int m = blockIdx.y;
    const double *A = M[m];
    const long long total = (long long)nk * n * n;
    double acc = 0.0;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long rc = e % ((long long)n * n);
        const int r = (int)(rc % n), c = (int)(rc / n);
        if (r > c) continue;
        const size_t o = (size_t)e * cplx;
        double rr = R0[o], ri = cplx == 2 ? R0[o + 1] : 0.0;
        if (R1) {
            rr += sgn * R1[o];
            if (cplx == 2) ri += sgn * R1[o + 1];
        }
        if (r == c)
            acc += A[o] * rr;
        else
            acc += 2.0 * (A[o] * rr + (cplx == 2 ? A[o + 1] * ri : 0.0));
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)m * gridDim.x + blockIdx.x] = red[0];
}

__global__ void k_trace_sum(const double *partial, int parts, int nMat, double *out)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nMat) return;
    double acc = 0.0;
    for (int b = 0; b < parts; ++b) acc += partial[(size_t)m * parts + b];
    out[m] = acc;
}

constexpr int kTraceParts = 512;

void launch_traces(const double *const *M, int nMat, const double *R0, const double *R1,
                   double sgn, int n, int nk, int cplx, double *partial, double *out,
                   hipStream_t s)
{
    hipLaunchKernelGGL(k_trace_part, dim3(kTraceParts, nMat), dim3(256), 0, s, M, R0, R1, sgn, n,
                       nk, cplx, partial);
    hipLaunchKernelGGL(k_trace_sum, dim3((nMat + 63) / 64), dim3(64), 0, s, partial, kTraceParts,
                       nMat, out);
}

}  // namespace olcao
