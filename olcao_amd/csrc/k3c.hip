// Three-centre lattice-sum kernels, one translation unit per bra angular
// momentum (compiled with -DOLCAO_LA=0..3 so the build parallelises).
#include "device_common.hpp"

#ifndef OLCAO_LA
#error "compile with -DOLCAO_LA=<0..3>"
#endif

namespace olcao {

// ---------------------------------------------------------------------------
// three-centre lattice sums, one wavefront per primitive pair
template <int LA, int LB>
__global__ __launch_bounds__(256) void k_three_center(DevCtx c, const int *list, int n,
                                                      int kind, int termType, int termIdx,
                                                      const double *coef, int accumulate,
                                                      double *blocks)
{
    constexpr int NA = ncart(LA), NB = ncart(LB);
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= n) return;
    const int pp = list[w];
    const PPInfo p = load_pp(c, pp);
    const double a1 = c.alpha[p.ti * c.maxA + p.ia];
    const double a2 = c.alpha[p.tj * c.maxA + p.ib];
    const PairConst q = make_pair(a1, a2, p.A, p.B);
    // screening centre and budget (integrals.F90:3211-3215, :3297-3305)
    const double ea1 = c.elemAlpha[p.ti * c.maxA + p.ia];
    const double ea2 = c.elemAlpha[p.tj * c.maxA + p.ib];
    const double acen = ea1 / (ea1 + ea2);
    const double adist = c.tau / ea2 / acen;
    double P[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) P[x] = acen * (p.A[x] - p.B[x]) + p.B[x];
    const double budget = c.tau * (1.0 - p.s2 / adist);

    unsigned nE3 = 0, nNuc = 0;
    double acc[NA][NB];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = 0.0;

    for (int site = lane; site < c.nPotSites; site += 64) {
        const int t = c.potSiteType[site];
        int n0 = 0, n1 = 0;
        bool nuc = false;
        if (kind == K3_HAM) {
            n1 = c.potNA[t];
            nuc = c.potZ[t] >= kSmallThresh;
        } else if (kind == K3_EP) {
            if (t != termType) continue;
            n0 = termIdx;
            n1 = termIdx + 1;
        } else {
            nuc = c.potZ[t] >= kSmallThresh;
        }
        if (n1 <= n0 && !nuc) continue;
        const double *pa = c.potAlpha + t * c.maxPA;
        const double nucA = c.potNuc[t];
        double a3min = (n1 > n0) ? pa[n0] : nucA;
        if (nuc && nucA < a3min) a3min = nucA;
        const double tadmax = budget * (a1 + a2 + a3min) / (a1 + a2) / a3min;
        const double tadnuc = budget * (a1 + a2 + nucA) / (a1 + a2) / nucA;
        const double *ps = c.potPos + 3 * site;
        double dv[3] = {P[0] - ps[0], P[1] - ps[1], P[2] - ps[2]};
        double lv2[3];
        nearest_lattice(c, dv, lv2);
        double cov[3] = {P[0] - ps[0] - lv2[0], P[1] - ps[1] - lv2[1], P[2] - ps[2] - lv2[2]};
        const double cos2 = cov[0] * cov[0] + cov[1] * cov[1] + cov[2] * cov[2];
        if (cos2 > tadmax) continue;
        const double mlr2 = cos2 + tadmax + 2.0 * sqrt(cos2 * tadmax);
        const double Z = c.potZ[t];
        const double *cf = coef ? coef + c.potCumul[t] : nullptr;
        for (int cell = 0; cell < c.nCells; ++cell) {
            if (c.cellSizes[cell] > mlr2) break;
            const double *L = c.cellDims + 3 * cell;
            const double r0 = cov[0] - L[0], r1 = cov[1] - L[1], r2 = cov[2] - L[2];
            const double d2 = r0 * r0 + r1 * r1 + r2 * r2;
            if (d2 > tadmax) continue;
            const double Cp[3] = {ps[0] + lv2[0] + L[0], ps[1] + lv2[1] + L[1],
                                  ps[2] + lv2[2] + L[2]};
            for (int nn = n0; nn < n1; ++nn) {
                const double a3 = pa[nn];
                const double tad = budget * (a1 + a2 + a3) / (a1 + a2) / a3;
                if (d2 > tad) break;   // ascending exponents: shrinking range
                const double wgt = (kind == K3_HAM) ? cf[nn] : 1.0;
                acc_e3c<LA, LB>(q, a3, Cp, wgt, acc);
                ++nE3;
            }
            if (nuc && d2 <= tadnuc) {
                acc_nuc<LA, LB>(q, nucA, Cp, -Z, acc);
                ++nNuc;
            }
        }
    }
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = wave_sum(acc[a][b]);
    if (c.counters) {
        const double e = wave_sum((double)nE3), u = wave_sum((double)nNuc);
        if (lane == 0) {
            atomicAdd(c.counters + 2 * (LA * 4 + LB), (unsigned long long)e);
            atomicAdd(c.counters + 2 * (LA * 4 + LB) + 1, (unsigned long long)u);
        }
    }
    double sh[nsh(LA)][nsh(LB)];
    cart_to_sh<LA, LB>(acc, sh);
    double *o = blocks + c.ppOff[pp];
#pragma unroll
    for (int i = 0; i < nsh(LA); ++i)
#pragma unroll
        for (int j = 0; j < nsh(LB); ++j) {
            const int e = i * nsh(LB) + j;
            if ((e & 63) == lane) o[e] = accumulate ? o[e] + sh[i][j] : sh[i][j];
        }
}

template <int LA, int LB>
static void three_one(const DevCtx &c, const ClassLists &cl, int kind, int tt, int ti,
                      const double *coef, int accumulate, double *blocks, hipStream_t s)
{
    const int k = LA * 4 + LB;
    const int n = cl.count[k];
    if (n == 0) return;
    const int wavesPerBlock = 4;
    hipLaunchKernelGGL((k_three_center<LA, LB>), dim3((n + wavesPerBlock - 1) / wavesPerBlock),
                       dim3(64 * wavesPerBlock), 0, s, c, cl.idx[k], n, kind, tt, ti, coef,
                       accumulate, blocks);
}


#define OLCAO_CAT2(a, b) a##b
#define OLCAO_CAT(a, b) OLCAO_CAT2(a, b)
void OLCAO_CAT(launch_3c_la, OLCAO_LA)(const DevCtx &c, const ClassLists &cl, int kind,
                                        int tt, int ti, const double *coef, int accumulate,
                                        double *blocks, hipStream_t s)
{
    three_one<OLCAO_LA, 0>(c, cl, kind, tt, ti, coef, accumulate, blocks, s);
    three_one<OLCAO_LA, 1>(c, cl, kind, tt, ti, coef, accumulate, blocks, s);
    three_one<OLCAO_LA, 2>(c, cl, kind, tt, ti, coef, accumulate, blocks, s);
    three_one<OLCAO_LA, 3>(c, cl, kind, tt, ti, coef, accumulate, blocks, s);
}

}  // namespace olcao
