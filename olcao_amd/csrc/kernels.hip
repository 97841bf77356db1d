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
#include "device_common.hpp"

namespace olcao {

// ---------------------------------------------------------------------------
// two-centre blocks
template <int LA, int LB>
__global__ __launch_bounds__(256) void k_two_center(DevCtx c, const int *list, int n,
                                                    int op, double *blocks)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int pp = list[t];
    const PPInfo p = load_pp(c, pp);
    const PairConst q = make_pair(c.alpha[p.ti * c.maxA + p.ia],
                                  c.alpha[p.tj * c.maxA + p.ib], p.A, p.B);
    double pc[ncart(LA)][ncart(LB)];
    if (op == OP_KE)
        two_center_cart<LA, LB, true>(q, pc);
    else
        two_center_cart<LA, LB, false>(q, pc);
    double sh[nsh(LA)][nsh(LB)];
    cart_to_sh<LA, LB>(pc, sh);
    double *o = blocks + c.ppOff[pp];
#pragma unroll
    for (int i = 0; i < nsh(LA); ++i)
#pragma unroll
        for (int j = 0; j < nsh(LB); ++j) o[i * nsh(LB) + j] = sh[i][j];
}

// ---------------------------------------------------------------------------
// contraction: X(s1,s2) = sum over the pair-cell's primitive pairs of
//   C1(a1,s1) * block(lm(s1), lm(s2)) * C2(a2,s2)
__global__ __launch_bounds__(256) void k_contract(DevCtx c, const double *blocks, double *X)
{
    const int lane = threadIdx.x & 63;
    const int pc = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (pc >= c.nPc) return;
    const int pr = c.pcPair[pc];
    const int ti = c.atomType[c.pairI[pr]], tj = c.atomType[c.pairJ[pr]];
    const int nt1 = c.nVale[ti] + c.nCore[ti], nt2 = c.nVale[tj] + c.nCore[tj];
    const int p0 = c.pcPp[pc], p1 = c.pcPp[pc + 1];
    const int ms = c.maxS, ma = c.maxA;
    double *xo = X + (long long)pc * ms * ms;
    for (int o = lane; o < nt1 * nt2; o += 64) {
        const int s1 = o / nt2, s2 = o - s1 * nt2;
        const int lm1 = c.lmIndex[ti * ms + s1] - 1, lm2 = c.lmIndex[tj * ms + s2] - 1;
        const double *c1 = c.basis + (long long)(ti * ms + s1) * ma;
        const double *c2 = c.basis + (long long)(tj * ms + s2) * ma;
        double sum = 0.0;
        for (int pp = p0; pp < p1; ++pp) {
            const int ab = c.ppAB[pp];
            const int n1 = nsh(pp_la(ab)), n2 = nsh(pp_lb(ab));
            if (lm1 < n1 && lm2 < n2)
                sum += c1[pp_a1(ab)] * blocks[c.ppOff[pp] + lm1 * n2 + lm2] * c2[pp_a2(ab)];
        }
        xo[s1 * ms + s2] = sum;
    }
}

// ---------------------------------------------------------------------------
// Bloch sum + origin shift + scatter, one workgroup per atom pair
__global__ __launch_bounds__(256) void k_bloch(DevCtx c, const double *X, double *vv,
                                               double *cv, double *cc, int gamma)
{
    const int pr = blockIdx.x;
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
    const double lv[3] = {c.pairLv[3 * pr], c.pairLv[3 * pr + 1], c.pairLv[3 * pr + 2]};
    const int per = nt1 * nt2;
    for (int o = threadIdx.x; o < nk * per; o += blockDim.x) {
        const int k = o / per;
        const int rem = o - k * per;
        const int s1 = rem / nt2, s2 = rem - s1 * nt2;
        double re = 0.0, im = 0.0;
        if (gamma) {
            for (int pc = p0; pc < p1; ++pc) re += X[(long long)pc * ms * ms + s1 * ms + s2];
        } else {
            for (int pc = p0; pc < p1; ++pc) {
                const double x = X[(long long)pc * ms * ms + s1 * ms + s2];
                const double *ph = c.phase + 2 * ((long long)c.pcCell[pc] * c.nk + k);
                re += ph[0] * x;
                im += ph[1] * x;
            }
            const double *kp = c.kcart + 3 * k;
            const double dot = kp[0] * lv[0] + kp[1] * lv[1] + kp[2] * lv[2];
            if (fabs(dot) >= kSmallThresh) {
                const double cr = cos(dot), ci = sin(dot);
                const double r2 = cr * re - ci * im;
                im = cr * im + ci * re;
                re = r2;
            }
        }
        const int cp = gamma ? 1 : 2;
        const bool a_core = s1 >= nv1, b_core = s2 >= nv2;
        if (!a_core && !b_core) {
            const long long idx = (v1 + s1) + (long long)(v2 + s2) * n + (long long)k * n * n;
            const bool zero = (i == j) && (s1 > s2);
            vv[idx * cp] = zero ? 0.0 : re;
            if (!gamma) vv[idx * cp + 1] = zero ? 0.0 : im;
        } else if (a_core && !b_core) {
            if (i != j) {   // for i == j the dagger assignment below wins
                const long long idx = (c1 + s1 - nv1) + (long long)(v2 + s2) * nc +
                                      (long long)k * nc * n;
                cv[idx * cp] = re;
                if (!gamma) cv[idx * cp + 1] = im;
            }
        } else if (!a_core && b_core) {
            const long long idx = (c2 + s2 - nv2) + (long long)(v1 + s1) * nc +
                                  (long long)k * nc * n;
            cv[idx * cp] = re;
            if (!gamma) cv[idx * cp + 1] = -im;
        } else {
            const long long idx = (c1 + s1 - nv1) + (long long)(c2 + s2 - nv2) * nc +
                                  (long long)k * nc * nc;
            cc[idx * cp] = re;
            if (!gamma) cc[idx * cp + 1] = im;
            if (i != j) {
                const long long id2 = (c2 + s2 - nv2) + (long long)(c1 + s1 - nv1) * nc +
                                      (long long)k * nc * nc;
                cc[id2 * cp] = re;
                if (!gamma) cc[id2 * cp + 1] = -im;
            }
        }
    }
}

__global__ void k_fill(double *p, long long n, double v)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        p[t] = v;
}

// H = NP + KE + sum_j c_j V_j on the upper triangle of every k matrix;
// c_j == 0 adds V_j unscaled (readPackedMatrixAccum, matrixSubs.F90:369-373).
__global__ void k_assemble(long long total, int n, const double *NP, const double *KE,
                           const double *const *V, const double *coef, int nTerms,
                           double *H, int cplx)
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
        for (int j = 0; j < nTerms; ++j) {
            const double cj = coef[j];
            v = v + (cj == 0.0 ? V[j][t] : V[j][t] * cj);
        }
        H[t] = v;
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
static void two_one(const DevCtx &c, const ClassLists &cl, int op, double *blocks,
                    hipStream_t s)
{
    const int k = LA * 4 + LB;
    const int n = cl.count[k];
    if (n == 0) return;
    hipLaunchKernelGGL((k_two_center<LA, LB>), dim3((n + 255) / 256), dim3(256), 0, s, c,
                       cl.idx[k], n, op, blocks);
}

void launch_two_center(const DevCtx &c, const ClassLists &cl, int op, double *blocks,
                       hipStream_t s)
{
    two_one<0, 0>(c, cl, op, blocks, s); two_one<0, 1>(c, cl, op, blocks, s);
    two_one<0, 2>(c, cl, op, blocks, s); two_one<0, 3>(c, cl, op, blocks, s);
    two_one<1, 0>(c, cl, op, blocks, s); two_one<1, 1>(c, cl, op, blocks, s);
    two_one<1, 2>(c, cl, op, blocks, s); two_one<1, 3>(c, cl, op, blocks, s);
    two_one<2, 0>(c, cl, op, blocks, s); two_one<2, 1>(c, cl, op, blocks, s);
    two_one<2, 2>(c, cl, op, blocks, s); two_one<2, 3>(c, cl, op, blocks, s);
    two_one<3, 0>(c, cl, op, blocks, s); two_one<3, 1>(c, cl, op, blocks, s);
    two_one<3, 2>(c, cl, op, blocks, s); two_one<3, 3>(c, cl, op, blocks, s);
}

void launch_3c_la0(const DevCtx &, const ClassLists &, int *, int, int, int, const double *,
                   int, double *, hipStream_t);
void launch_3c_la1(const DevCtx &, const ClassLists &, int *, int, int, int, const double *,
                   int, double *, hipStream_t);
void launch_3c_la2(const DevCtx &, const ClassLists &, int *, int, int, int, const double *,
                   int, double *, hipStream_t);
void launch_3c_la3(const DevCtx &, const ClassLists &, int *, int, int, int, const double *,
                   int, double *, hipStream_t);

void launch_three_center(const DevCtx &c, const ClassLists &cl, int *heads, int kind, int tt,
                         int ti, const double *coef, int accumulate, double *blocks, hipStream_t s)
{
    launch_3c_la0(c, cl, heads, kind, tt, ti, coef, accumulate, blocks, s);
    launch_3c_la1(c, cl, heads, kind, tt, ti, coef, accumulate, blocks, s);
    launch_3c_la2(c, cl, heads, kind, tt, ti, coef, accumulate, blocks, s);
    launch_3c_la3(c, cl, heads, kind, tt, ti, coef, accumulate, blocks, s);
}

void launch_contract(const DevCtx &c, const double *blocks, double *X, hipStream_t s)
{
    if (c.nPc == 0) return;
    hipLaunchKernelGGL(k_contract, dim3((c.nPc + 3) / 4), dim3(256), 0, s, c, blocks, X);
}

void launch_bloch_scatter(const DevCtx &c, const double *X, double *vv, double *cv,
                          double *cc, int gamma, hipStream_t s)
{
    if (c.nPairs == 0) return;
    hipLaunchKernelGGL(k_bloch, dim3(c.nPairs), dim3(256), 0, s, c, X, vv, cv, cc, gamma);
}

void launch_fill(double *p, long long n, double v, hipStream_t s)
{
    if (n <= 0) return;
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

void launch_assemble(long long total, int n, const double *NP, const double *KE,
                     const double *const *V, const double *coef, int nTerms, double *H,
                     int cplx, hipStream_t s)
{
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_assemble, dim3((unsigned)blocks), dim3(256), 0, s, total, n, NP, KE,
                       V, coef, nTerms, H, cplx);
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

}  // namespace olcao
