// Three-centre lattice-sum kernels (electronicPE / nuclearPE / the fused
// PSCF sweep, src/uolcao/integrals.F90:2533-2856, :3211-3430), one
// translation unit per bra angular momentum (-DOLCAO_LA=0..3).
//
// Work decomposition (one wavefront per primitive pair, persistent):
//   * waves take primitive pairs from an atomic counter over a list sorted
//     by estimated cost, heaviest first;
//   * phase 1: lanes take potential sites; each lane bounds the images of
//     its site that can lie inside the screening sphere of the pair's
//     overlap centre P with a box in fractional coordinates, the candidates
//     are flattened across the wave, tested exactly (|P - C'|^2 <= range),
//     and the surviving images with their term counts are appended to an
//     LDS queue;
//   * phase 2: the queue is expanded to (image, term) items and the items
//     are spread over the 64 lanes, so every lane evaluates the same number
//     of items; each item adds its McMurchie-Davidson derivatives R_tuv
//     (t+u+v <= la+lb, 35 numbers for d-d) to the lane's W_tuv;
//   * a reduce-scatter butterfly sums W over the lanes, W is contracted once
//     with the pair's Hermite coefficients to the Cartesian block, which is
//     transformed to solid harmonics and added to the primitive-pair store.
// The image set is the one the reference sums: every image C' of every
// selected site with xi s^2 + zeta a3/(zeta+a3) |P-C'|^2 <= tau, tested term
// by term with the reference's range expression.
#include "device_common.hpp"

#ifndef OLCAO_LA
#error "compile with -DOLCAO_LA=<0..3>"
#endif

namespace olcao {

constexpr int kQCap = 256;     // queued images per wave between flushes

static __device__ __forceinline__ unsigned long long lanemask_lt(int lane)
{
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Inclusive prefix sum of an int over the wave.
static __device__ __forceinline__ int wave_scan_incl(int v, int lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

// Reduce-scatter of v[0..N) over the 64 lanes: afterwards lane l holds the
// wave sum of components [off, off + len) in v[0..len).  Fixed summation
// tree, so results are deterministic.
template <int N>
static __device__ __forceinline__ void reduce_scatter(double (&v)[N], int lane, int &off, int &len)
{
    off = 0;
    len = N;
    int S = N;
#pragma unroll
    for (int level = 0; level < 6; ++level) {
        const int m = 32 >> level;
        const bool bit = (lane & m) != 0;
        const int S2 = (S + 1) / 2;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            if (i < S2) {
                const double a = v[i];
                const double b = (i + S2 < S) ? v[(i + S2 < N) ? i + S2 : 0] : 0.0;
                const double send = bit ? a : b;
                const double keep = bit ? b : a;
                v[i] = keep + __shfl_xor(send, m, 64);
            }
        }
        if (bit) {
            off += S2;
            len = len > S2 ? len - S2 : 0;
        } else {
            len = len < S2 ? len : S2;
        }
        S = S2;
    }
}

template <int LA, int LB, bool NUC>
__global__ __launch_bounds__(64) void k_three_center(DevCtx c, const int *list, int n,
                                                     int *head, int kind, int termType,
                                                     int termIdx, const double *coef,
                                                     int accumulate, double *blocks)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = NA * NB;
    constexpr int SA = nsh(LA), SB = nsh(LB);
    constexpr int L = LA + LB, NH = nherm(L);
    __shared__ double qC[3][kQCap];
    __shared__ int qInfo[kQCap];     // type | first term << 8 | term count << 16
    __shared__ int qPre[kQCap + 1];
    __shared__ double sCov[3][64];
    __shared__ double sTad[64];
    __shared__ int sLo[3][64], sSpan[3][64], sPre[65];
    __shared__ double sRed[NC];
    __shared__ double sW[NH];
    __shared__ double sE[3][LA + 1][LB + 1][L + 1];
    __shared__ int sNext;

    const int lane = threadIdx.x;
    unsigned items = 0;
    double invn[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        invn[i] = sqrt(c.inv[3 * i] * c.inv[3 * i] + c.inv[3 * i + 1] * c.inv[3 * i + 1] +
                       c.inv[3 * i + 2] * c.inv[3 * i + 2]);

    while (true) {
        if (lane == 0) sNext = atomicAdd(head, 1);
        __syncthreads();
        const int w = sNext;
        __syncthreads();
        if (w >= n) break;
        const int pp = list[w];
        const PPInfo p = load_pp(c, pp);
        const double a1 = c.alpha[p.ti * c.maxA + p.ia];
        const double a2 = c.alpha[p.tj * c.maxA + p.ib];
        const PairConst q = make_pair(a1, a2, p.A, p.B);
        const double ea1 = c.elemAlpha[p.ti * c.maxA + p.ia];
        const double ea2 = c.elemAlpha[p.tj * c.maxA + p.ib];
        const double acen = ea1 / (ea1 + ea2);
        const double adist = c.tau / ea2 / acen;
        double P[3];
#pragma unroll
        for (int x = 0; x < 3; ++x) P[x] = acen * (p.A[x] - p.B[x]) + p.B[x];
        const double budget = c.tau * (1.0 - p.s2 / adist);

        double acc[NH];
#pragma unroll
        for (int e = 0; e < NH; ++e) acc[e] = 0.0;

        int qn = 0;
        // ---- phase 2: evaluate the queued (image, term) items ----
        auto flush = [&](int count) {
            __syncthreads();
            // exclusive prefix of term counts over the queue (4 entries/lane)
            int local[kQCap / 64];
            int sum = 0;
#pragma unroll
            for (int r = 0; r < kQCap / 64; ++r) {
                const int e = lane * (kQCap / 64) + r;
                const int cnt = e < count ? (qInfo[e] >> 16) : 0;
                local[r] = sum;
                sum += cnt;
            }
            const int incl = wave_scan_incl(sum, lane);
            const int base = incl - sum;
#pragma unroll
            for (int r = 0; r < kQCap / 64; ++r) {
                const int e = lane * (kQCap / 64) + r;
                if (e < count) qPre[e] = base + local[r];
            }
            const int total = __shfl(incl, 63, 64);
            if (lane == 0) qPre[count] = total;
            __syncthreads();
            for (int it = lane; it < total; it += 64) {
                ++items;
                int lo = 0, hi = count;   // largest e with qPre[e] <= it
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (qPre[mid] <= it) lo = mid; else hi = mid;
                }
                const int info = qInfo[lo];
                const int t = info & 255;
                const double C[3] = {qC[0][lo], qC[1][lo], qC[2][lo]};
                if (NUC) {
                    acc_nuc_herm<L>(q, c.potNuc[t], C, -c.potZ[t], acc);
                } else {
                    const int nn = ((info >> 8) & 255) + (it - qPre[lo]);
                    const double a3 = c.potAlpha[t * c.maxPA + nn];
                    const double wgt = (kind == K3_HAM) ? coef[c.potCumul[t] + nn] : 1.0;
                    acc_e3c_herm<L>(q, a3, C, wgt, acc);
                }
            }
            __syncthreads();
        };

        // ---- phase 1: candidate images per site, flattened over the wave ----
        for (int base = 0; base < c.nPotSites; base += 64) {
            const int site = base + lane;
            int cnt = 0;
            int t = 0;
            double tadmax = -1.0;
            double cov[3] = {0.0, 0.0, 0.0};
            int lo3[3] = {0, 0, 0}, span[3] = {0, 0, 0};
            if (site < c.nPotSites) {
                t = c.potSiteType[site];
                bool any = false;
                double a3min = 0.0;
                if (NUC) {
                    if (kind != K3_EP && c.potZ[t] >= kSmallThresh) {
                        any = true;
                        a3min = c.potNuc[t];
                    }
                } else if (kind == K3_HAM) {
                    if (c.potNA[t] > 0) {
                        any = true;
                        a3min = c.potAlpha[t * c.maxPA];
                    }
                } else if (kind == K3_EP && t == termType) {
                    any = true;
                    a3min = c.potAlpha[t * c.maxPA + termIdx];
                }
                if (any) {
                    tadmax = budget * (a1 + a2 + a3min) / (a1 + a2) / a3min;
                    const double r = sqrt(fmax(tadmax, 0.0)) * (1.0 + 1e-12) + 1e-12;
                    const double *ps = c.potPos + 3 * site;
#pragma unroll
                    for (int x = 0; x < 3; ++x) cov[x] = P[x] - ps[x];
                    cnt = 1;
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        const double f = c.inv[3 * i] * cov[0] + c.inv[3 * i + 1] * cov[1] +
                                         c.inv[3 * i + 2] * cov[2];
                        const double rad = r * invn[i];
                        const int l0 = (int)ceil(f - rad), h0 = (int)floor(f + rad);
                        lo3[i] = l0;
                        span[i] = h0 >= l0 ? h0 - l0 + 1 : 0;
                        cnt *= span[i];
                    }
                }
            }
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                sCov[x][lane] = cov[x];
                sLo[x][lane] = lo3[x];
                sSpan[x][lane] = span[x];
            }
            sTad[lane] = tadmax;
            const int incl = wave_scan_incl(cnt, lane);
            sPre[lane] = incl - cnt;
            const int total = __shfl(incl, 63, 64);
            if (lane == 0) sPre[64] = total;
            __syncthreads();
            for (int cb = 0; cb < total; cb += 64) {
                const int idx = cb + lane;
                bool pass = false;
                int info = 0;
                double Cp[3] = {0.0, 0.0, 0.0};
                if (idx < total) {
                    int lo = 0, hi = 64;   // owner lane: largest o with sPre[o] <= idx
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (sPre[mid] <= idx) lo = mid; else hi = mid;
                    }
                    const int o = lo;
                    int local = idx - sPre[o];
                    const int s0 = sSpan[0][o], s1 = sSpan[1][o];
                    const int nx = local % s0 + sLo[0][o];
                    local /= s0;
                    const int ny = local % s1 + sLo[1][o];
                    const int nz = local / s1 + sLo[2][o];
                    double L[3];
#pragma unroll
                    for (int x = 0; x < 3; ++x)
                        L[x] = nx * c.vec[x] + ny * c.vec[3 + x] + nz * c.vec[6 + x];
                    const double r0 = sCov[0][o] - L[0], r1 = sCov[1][o] - L[1],
                                 r2 = sCov[2][o] - L[2];
                    const double d2 = r0 * r0 + r1 * r1 + r2 * r2;
                    if (d2 <= sTad[o]) {
                        const int os = base + o;
                        const int ot = c.potSiteType[os];
                        int n0 = 0, nt = 0;
                        if (NUC) {
                            nt = 1;
                        } else {
                            int n1;
                            if (kind == K3_HAM) {
                                n0 = 0;
                                n1 = c.potNA[ot];
                            } else {
                                n0 = termIdx;
                                n1 = termIdx + 1;
                            }
                            for (int nn = n0; nn < n1; ++nn) {
                                const double a3 = c.potAlpha[ot * c.maxPA + nn];
                                const double tad = budget * (a1 + a2 + a3) / (a1 + a2) / a3;
                                if (d2 > tad) break;   // ascending exponents
                                ++nt;
                            }
                        }
                        if (nt > 0) {
                            pass = true;
                            info = ot | (n0 << 8) | (nt << 16);
                            const double *ps = c.potPos + 3 * os;
#pragma unroll
                            for (int x = 0; x < 3; ++x) Cp[x] = ps[x] + L[x];
                        }
                    }
                }
                const unsigned long long m = __ballot(pass);
                if (pass) {
                    const int slot = qn + __popcll(m & lanemask_lt(lane));
                    qC[0][slot] = Cp[0];
                    qC[1][slot] = Cp[1];
                    qC[2][slot] = Cp[2];
                    qInfo[slot] = info;
                }
                qn += __popcll(m);
                if (qn > kQCap - 64) {
                    flush(qn);
                    qn = 0;
                }
            }
            __syncthreads();
        }
        if (qn > 0) flush(qn);

        // ---- wave sum, Hermite -> Cartesian -> solid harmonics, store ----
        int off, len;
        reduce_scatter<NH>(acc, lane, off, len);
#pragma unroll
        for (int i = 0; i < NH; ++i)
            if (i < len) sW[off + i] = acc[i];
        if (lane < 3) {
            double E[LA + 1][LB + 1][L + 1];
            const double XPA = lane == 0 ? q.P[0] - q.A[0] : lane == 1 ? q.P[1] - q.A[1] : q.P[2] - q.A[2];
            const double XPB = lane == 0 ? q.P[0] - q.B[0] : lane == 1 ? q.P[1] - q.B[1] : q.P[2] - q.B[2];
            herm_e<LA, LB>(XPA, XPB, 0.5 / q.zeta, E);
#pragma unroll
            for (int i = 0; i <= LA; ++i)
#pragma unroll
                for (int j = 0; j <= LB; ++j)
#pragma unroll
                    for (int t = 0; t <= L; ++t) sE[lane][i][j][t] = E[i][j][t];
        }
        __syncthreads();
        const double kab = exp(-q.xid2);
        for (int e = lane; e < NC; e += 64) {
            const int a = e / NB, b = e - (e / NB) * NB;
            const int ax = cexp(a, 0), ay = cexp(a, 1), az = cexp(a, 2);
            const int bx = cexp(b, 0), by = cexp(b, 1), bz = cexp(b, 2);
            double sum = 0.0;
            for (int t = 0; t <= ax + bx; ++t)
                for (int u = 0; u <= ay + by; ++u) {
                    const double exy = sE[0][ax][bx][t] * sE[1][ay][by][u];
                    for (int v = 0; v <= az + bz; ++v)
                        sum += exy * sE[2][az][bz][v] * sW[herm_index(t, u, v)];
                }
            sRed[e] = kab * sum;
        }
        __syncthreads();
        double *o = blocks + c.ppOff[pp];
        for (int e = lane; e < SA * SB; e += 64) {
            const int i = e / SB, j = e - (e / SB) * SB;
            double s = 0.0;
            for (int ti = 0; ti < sh_nterm(i); ++ti)
                for (int tj = 0; tj < sh_nterm(j); ++tj)
                    s += sh_weight(i, ti) * sh_weight(j, tj) *
                         sRed[sh_cart(i, ti) * NB + sh_cart(j, tj)];
            o[e] = accumulate ? o[e] + s : s;
        }
        __syncthreads();
    }
    if (c.counters) {
        const double tot = wave_sum((double)items);
        if (lane == 0 && tot > 0.0)
            atomicAdd(c.counters + 2 * (LA * 4 + LB) + (NUC ? 1 : 0), (unsigned long long)tot);
    }
}

template <int LA, int LB>
static void three_one(const DevCtx &c, const ClassLists &cl, int *heads, int kind, int tt,
                      int ti, const double *coef, int accumulate, double *blocks, hipStream_t s)
{
    const int k = LA * 4 + LB;
    const int n = cl.count[k];
    if (n == 0) return;
    const int grid = n < 4096 ? n : 4096;
    if (kind != K3_NP)
        hipLaunchKernelGGL((k_three_center<LA, LB, false>), dim3(grid), dim3(64), 0, s, c,
                           cl.idx[k], n, heads + 2 * k, kind, tt, ti, coef, accumulate, blocks);
    if (kind != K3_EP)
        hipLaunchKernelGGL((k_three_center<LA, LB, true>), dim3(grid), dim3(64), 0, s, c,
                           cl.idx[k], n, heads + 2 * k + 1, kind, tt, ti, coef,
                           kind == K3_NP ? accumulate : 1, blocks);
}

#define OLCAO_CAT2(a, b) a##b
#define OLCAO_CAT(a, b) OLCAO_CAT2(a, b)
void OLCAO_CAT(launch_3c_la, OLCAO_LA)(const DevCtx &c, const ClassLists &cl, int *heads,
                                        int kind, int tt, int ti, const double *coef,
                                        int accumulate, double *blocks, hipStream_t s)
{
    three_one<OLCAO_LA, 0>(c, cl, heads, kind, tt, ti, coef, accumulate, blocks, s);
    three_one<OLCAO_LA, 1>(c, cl, heads, kind, tt, ti, coef, accumulate, blocks, s);
    three_one<OLCAO_LA, 2>(c, cl, heads, kind, tt, ti, coef, accumulate, blocks, s);
    three_one<OLCAO_LA, 3>(c, cl, heads, kind, tt, ti, coef, accumulate, blocks, s);
}

}  // namespace olcao
