// Blocked reduction of large Hermitian (real symmetric) matrices to real
// tridiagonal form with the matrix in global memory: the zhetrd/dsytrd
// (uplo 'L') step inside zhegv/dsygv (interfaces.F90:73-200) as zlatrd-style
// panels of kPanelB columns, each followed by the rank-2pb update of the
// trailing matrix, A -= [W V] [V W]^H in one GEMM, for the sizes whose rows do not fit the
// LDS-resident reduction of ktrd.hip.  Its output is that kernel's: V (the
// explicit reflectors, column i = v_i, v_i(i+1) = 1, zero above), tau, d, e.
//
// Panel kernel: G workgroups per matrix, launched cooperatively, walk the
// panel's columns i = j0 .. j0+pb-1 with two group barriers per column.
// Within a panel the trailing matrix A22 = A(i+1:n, i+1:n) is the panel's
// starting one; the panel's earlier reflectors enter through W (zlatrd):
//   y = A22 v                          (lower tiles of 64 x 64, each read once:
//                                       y_I += A_IJ v_J, y_J += A_IJ^H v_I)
//   a = V^H v, b = W^H v               (the panel's earlier columns)
//   p = y - V b - W a,  w = tau p + alpha_w v,
//   alpha_w = -1/2 |tau|^2 v^H p,  v^H p = v^H y - 2 Re(a^H b)  (real)
// and column i+1 takes the panel's updates before its reflector is formed:
//   A(r, i+1) -= V(r, j0:i) conj(W(i+1, :)) + W(r, :) conj(V(i+1, j0:i)).
// Step i:
//   R  every workgroup reads column i (published by the previous step, or by
//      the panel start kernel) and the partial norms, and forms the reflector
//      (zlarfg) redundantly
//   T  every workgroup takes lower tiles (stride G) and publishes each
//      tile's row and column partials and its share of v^H y; the last
//      workgroups also form the 2c dots a, b   -> barrier
//   V  the owner of each 64-row block sums its rows' partials, forms p, w,
//      stores W(:, i - j0) and V(:, i), updates its rows of column i+1 and
//      publishes them with their partial norm    -> barrier
// Publications use agent-scope relaxed atomic stores and loads (coherent
// across the device's L2 caches); the barriers give the ordering.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "numtraits.hpp"
#include "olcao_internal.h"
#include "wave.hpp"

namespace olcao {

namespace {

constexpr int kBT = 256;                   // threads per workgroup
// tiles per pass of a workgroup, whose loads are in flight together (fewer
// for complex matrices, whose tiles take twice the registers)
template <bool CPLX>
constexpr int kTiles = CPLX ? 2 : 4;
constexpr long long kBSpinLimit = 1ll << 22;

__device__ __forceinline__ void bpst(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double bpld(const double *p)
{
    return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool CPLX>
__device__ __forceinline__ void bpstT(double *p, long long i, typename Num<CPLX>::T v)
{
    if constexpr (CPLX) {
        bpst(p + 2 * i, v.x);
        bpst(p + 2 * i + 1, v.y);
    } else {
        bpst(p + i, v);
    }
}
template <bool CPLX>
__device__ __forceinline__ typename Num<CPLX>::T bpldT(const double *p, long long i)
{
    if constexpr (CPLX)
        return {bpld(p + 2 * i), bpld(p + 2 * i + 1)};
    else
        return bpld(p + i);
}

// sum over the 16 lanes of a DPP row; every lane of the row gets it
__device__ __forceinline__ double row16_sum(double a)
{
    a += dpp<kDppQuadSwap1>(a);
    a += dpp<kDppQuadSwap2>(a);
    a += dpp<kDppHalfMirror>(a);
    a += dpp<kDppMirror>(a);
    return a;
}
template <bool CPLX>
__device__ __forceinline__ typename Num<CPLX>::T row16_sumT(typename Num<CPLX>::T a)
{
    if constexpr (CPLX)
        return {row16_sum(a.x), row16_sum(a.y)};
    else
        return row16_sum(a);
}

// workgroup sum (all threads call it; every thread gets the result)
__device__ __forceinline__ double wg_sum(double v, double *red)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}
template <bool CPLX>
__device__ __forceinline__ typename Num<CPLX>::T wg_sumT(typename Num<CPLX>::T v, double *red)
{
    if constexpr (CPLX) {
        const double x = wg_sum(v.x, red);
        const double y = wg_sum(v.y, red + 4);
        return {x, y};
    } else {
        return wg_sum(v, red);
    }
}

// group barrier with a bounded spin (false: the group did not assemble)
__device__ __forceinline__ bool bgroup_barrier(unsigned *cnt, unsigned target, int *bad)
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        int b = 0;
        __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        long long sp = 0;
        while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++sp > kBSpinLimit) {
                b = 1;
                break;
            }
        }
        *bad = b;
    }
    __syncthreads();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    return *bad == 0;
}

struct TrdbArgs {
    const double *A;      // [nb] n x n, lower triangle current
    double *V, *W;        // V [nb] n x n; W [nb] n x 3 kPanelB: W, V, W of the panel
    double *d, *e, *tau;  // [nb] n (tau complex for complex matrices)
    double *pub;          // [nb] pubLen doubles
    unsigned *cnt;        // [nb] 64 barrier words
    int *info;            // [nb]
    long long pubLen;
    int n, G, j0, pb, ntiles;
};

// per-matrix publication layout (doubles, E = 1 real, 2 complex):
//   tile partials [ntiles][row, column][64] E | column [n] E | norms [G] |
//   v^H y shares [G] | dots [64] E
__host__ __device__ inline long long trdb_pub_len(int n, int G, int E)
{
    const long long Tn = (n + 63) / 64, ntiles = Tn * (Tn + 1) / 2;
    return ntiles * 128 * E + (long long)n * E + 2LL * G + 64LL * E;
}

// Panel start: column j0 (rows > j0) into the publication column, its norm
// below the first subdiagonal, d(j0), and the barrier counter cleared.  With
// j0 = n - 1 only d(n - 1) is written (after the last trailing update).
template <bool CPLX>
__global__ __launch_bounds__(kBT) void k_trdb_start(TrdbArgs p)
{
    using N = Num<CPLX>;
    using T = typename N::T;
    constexpr int E = CPLX ? 2 : 1;
    __shared__ double red[8];
    const int mat = blockIdx.x, n = p.n, j0 = p.j0, G = p.G;
    const long long nn = (long long)n * n;
    const double *A = p.A + E * nn * mat;
    double *pub = p.pub + p.pubLen * mat;
    double *cpub = pub + (long long)p.ntiles * 128 * E, *normPub = cpub + (long long)n * E;
    if (threadIdx.x == 0) p.d[(long long)n * mat + j0] = N::re(N::load(A, j0 + (long long)j0 * n));
    if (j0 >= n - 1) return;
    double part = 0.0;
    for (int r = j0 + 1 + threadIdx.x; r < n; r += kBT) {
        const T a = N::load(A, r + (long long)j0 * n);
        N::store(cpub, r, a);
        if (r >= j0 + 2) part += N::abs2(a);
    }
    const double nrm = wg_sum(part, red);
    for (int q = threadIdx.x; q < G; q += kBT) normPub[q] = q == 0 ? nrm : 0.0;
    if (threadIdx.x == 0) p.cnt[64 * mat] = 0u;
}

template <bool CPLX>
__global__ __launch_bounds__(kBT) void k_trdb_panel(TrdbArgs p)
{
    using N = Num<CPLX>;
    using T = typename N::T;
    constexpr int E = CPLX ? 2 : 1;
    extern __shared__ double lds_raw[];
    T *vraw = reinterpret_cast<T *>(lds_raw);   // column i as published, rows > i
    constexpr int kTB = kTiles<CPLX>;
    __shared__ T rowbuf[kTB][16][64];
    __shared__ T colsum[kTB][64];   // a pass's column sums, stored out by one wave
    __shared__ T sa[kPanelB], sb[kPanelB], sVs[kPanelB], sWs[kPanelB];
    __shared__ double red[8];
    __shared__ int bad;
    const int tid = threadIdx.x;
    const int n = p.n, G = p.G, j0 = p.j0;
    const int mat = blockIdx.x / G, w = blockIdx.x % G;
    if (p.info[mat] == -2) return;   // an earlier panel's barrier timed out
    const long long nn = (long long)n * n;
    const double *A = p.A + E * nn * mat;
    double *V = p.V + E * nn * mat;
    double *W = p.W + E * (long long)n * (3 * kPanelB) * mat;
    double *dd = p.d + (long long)n * mat, *ee = p.e + (long long)n * mat;
    double *tauG = p.tau + (long long)E * n * mat;
    double *pub = p.pub + p.pubLen * mat;
    double *tp = pub, *cpub = tp + (long long)p.ntiles * 128 * E;
    double *normPub = cpub + (long long)n * E, *scalPub = normPub + G, *dotPub = scalPub + G;
    unsigned *cnt = p.cnt + 64 * mat;
    unsigned epoch = 0;
    const int Tn = (n + 63) >> 6;
    const int tr = tid & 15, tc = tid >> 4;
    const int Ib0 = (j0 + 1) >> 6;
    const T one = N::mk(1.0, 0.0);
    // loads per thread of a whole column (rows > i) and of a row's tile
    // partials (at most Tn + 1, over four threads), for the largest n
    constexpr int kMaxN = CPLX ? kTrdbMaxComplex : kTrdbMaxReal;
    constexpr int kRL = (kMaxN + kBT - 1) / kBT;
    constexpr int kOL = ((kMaxN + 63) / 64 + 1 + 3) / 4;

    for (int i = j0; i < j0 + p.pb; ++i) {
        const int s = i + 1, c = i - j0;
        const int Ib = s >> 6, nAct = Tn - Ib;
        const int NT = nAct * (nAct + 1) / 2;
        const bool more = i + 1 < j0 + p.pb;

        // tiles of this step: q = w, w + G, ..., kTB per pass; tile q is
        // (I, J), Ib <= J <= I < Tn, enumerated column by column
        auto tile_ij = [&](int q, int &I, int &J) {
            int jj = 0, rem = q;
            while (rem >= nAct - jj) {
                rem -= nAct - jj;
                ++jj;
            }
            J = Ib + jj;
            I = J + rem;
        };
        T a[kTB][4][4];
        int tI[kTB], tJ[kTB];
        auto issue = [&](int q0) {
#pragma unroll
            for (int b = 0; b < kTB; ++b) {
                const int q = q0 + b * G;
                tI[b] = -1;
                tJ[b] = 0;
                if (q < NT) tile_ij(q, tI[b], tJ[b]);
                const int r0 = 64 * tI[b], c0 = 64 * tJ[b];
                const bool diag = tI[b] == tJ[b];
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int r = r0 + 4 * tr + u, cc = c0 + 4 * tc + j;
                        const bool ok = tI[b] >= 0 && r < n && cc < n && cc >= s && r >= s &&
                                        (!diag || cc <= r);
                        a[b][j][u] = ok ? N::load(A, r + (long long)cc * n) : N::zero();
                    }
            }
        };
        int q0 = w;
        issue(q0);
        // this workgroup's dot (a_k = V(:, j0+k)^H v, b_k = W(:, k)^H v over
        // rows >= s; k = G - 1 - w < 2c): its column, loaded with the tiles
        const int kDot = G - 1 - w;
        const bool hasDot = kDot < 2 * c;
        T dcol[kRL];
        {
            const double *colp = kDot < c ? V + (long long)E * (j0 + kDot) * n
                                          : W + (long long)E * (kDot - c) * n;
#pragma unroll
            for (int t = 0; t < kRL; ++t) {
                const int r = s + tid + t * kBT;
                dcol[t] = hasDot && r < n ? bpldT<CPLX>(colp, r) : N::zero();
            }
        }

        // ---- R: column i and its reflector (zlarfg), in every workgroup
        // all of a thread's loads issued before any is used
        {
            T tmp[kRL];
#pragma unroll
            for (int t = 0; t < kRL; ++t) {
                const int r = s + tid + t * kBT;
                tmp[t] = r < n ? bpldT<CPLX>(cpub, r) : N::zero();
            }
#pragma unroll
            for (int t = 0; t < kRL; ++t) {
                const int r = s + tid + t * kBT;
                if (r < n) vraw[r] = tmp[t];
            }
        }
        const double xn2 = wg_sum(tid < G ? bpld(normPub + tid) : 0.0, red);
        const T alpha = vraw[s];
        double ar, ai;
        if constexpr (CPLX) {
            ar = alpha.x;
            ai = alpha.y;
        } else {
            ar = alpha;
            ai = 0.0;
        }
        T tauv = N::zero(), scale = one;
        double beta = ar;
        if (!(xn2 == 0.0 && ai == 0.0)) {
            const double nrm = sqrt(ar * ar + ai * ai + xn2);
            beta = ar >= 0.0 ? -nrm : nrm;
            tauv = N::mk((beta - ar) / beta, -ai / beta);
            const double zr = ar - beta, zi = ai, den = zr * zr + zi * zi;
            scale = N::mk(zr / den, -zi / den);
        }
        if (w == 0 && tid == 0) {
            ee[i] = beta;
            N::store(tauG, i, tauv);
        }
        auto vat = [&](int r) -> T {
            return r == s ? one : (r > s && r < n ? N::mul(vraw[r], scale) : N::zero());
        };

        // ---- T: lower tiles of the trailing matrix
        double scal = 0.0;
        while (q0 < NT) {
#pragma unroll
            for (int b = 0; b < kTB; ++b) {
                if (tI[b] < 0) continue;
                const int I = tI[b], J = tJ[b], r0 = 64 * I, c0 = 64 * J;
                const bool diag = I == J;
                T vc[4], vr[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) vc[j] = vat(c0 + 4 * tc + j);
#pragma unroll
                for (int u = 0; u < 4; ++u) vr[u] = vat(r0 + 4 * tr + u);
                T racc[4], cacc[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) racc[u] = N::zero();
#pragma unroll
                for (int j = 0; j < 4; ++j) cacc[j] = N::zero();
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        if (diag && 4 * tc + j == 4 * tr + u) {
                            const double dv = N::re(a[b][j][u]);
                            racc[u] = N::add(racc[u], N::scale(vc[j], dv));
                            scal -= dv * N::abs2(vr[u]);
                        } else {
                            racc[u] = N::add(racc[u], N::mul(a[b][j][u], vc[j]));
                            cacc[j] = N::add(cacc[j], N::cmul(a[b][j][u], vr[u]));
                        }
                    }
#pragma unroll
                for (int j = 0; j < 4; ++j) cacc[j] = row16_sumT<CPLX>(cacc[j]);
                if (tr == 0) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) colsum[b][4 * tc + j] = cacc[j];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) rowbuf[b][tc][4 * tr + u] = racc[u];
            }
            __syncthreads();
            // one wave per tile stores its row and column sums, 64 lanes each
            if (tid < 64) {
#pragma unroll
                for (int b = 0; b < kTB; ++b) {
                    if (tI[b] < 0) continue;
                    T sum = N::zero();
#pragma unroll
                    for (int q2 = 0; q2 < 16; ++q2) sum = N::add(sum, rowbuf[b][q2][tid]);
                    const long long tix = (long long)tI[b] * (tI[b] + 1) / 2 + tJ[b];
                    bpstT<CPLX>(tp + tix * 128 * E, tid, sum);
                    bpstT<CPLX>(tp + tix * 128 * E + 64 * E, tid, colsum[b][tid]);
                    scal += 2.0 * N::re(N::cmul(vat(64 * tI[b] + tid), sum));
                }
            }
            __syncthreads();
            q0 += kTB * G;
            if (q0 < NT) issue(q0);
        }
        if (hasDot) {
            T acc = N::zero();
#pragma unroll
            for (int t = 0; t < kRL; ++t) acc = N::add(acc, N::cmul(dcol[t], vat(s + tid + t * kBT)));
            acc = wg_sumT<CPLX>(acc, red);
            if (tid == 0) bpstT<CPLX>(dotPub, kDot, acc);
        }
        // the first owned block's V and W rows and column i+1, written
        // before this step, loaded while the group assembles
        const int Iown = Ib0 + w;
        const bool ownFirst = Iown < Tn && Iown >= Ib;
        T pv[kPanelB / 4], pw[kPanelB / 4], pa = N::zero();
        {
            const int rr = tid & 63, part = tid >> 6, r = 64 * Iown + rr;
            const bool live = ownFirst && r >= s && r < n;
#pragma unroll
            for (int t = 0; t < kPanelB / 4; ++t) {
                const int k = part + 4 * t;
                const bool on = live && k < c;
                pv[t] = on ? bpldT<CPLX>(V, r + (long long)(j0 + k) * n) : N::zero();
                pw[t] = on ? bpldT<CPLX>(W, r + (long long)k * n) : N::zero();
            }
            if (live && more) pa = N::load(A, r + (long long)(i + 1) * n);
        }
        scal = wg_sum(scal, red);
        if (tid == 0) bpst(scalPub + w, scal);
        ++epoch;
        if (!bgroup_barrier(cnt, epoch * (unsigned)G, &bad)) {
            if (tid == 0) p.info[mat] = -2;
            return;
        }

        // ---- V: w, the reflector column, and column i+1 in the owned rows
        if (tid < 2 * c) {
            const T dv = bpldT<CPLX>(dotPub, tid);
            if (tid < c) sa[tid] = dv; else sb[tid - c] = dv;
        }
        if (tid < c) {
            sVs[tid] = bpldT<CPLX>(V, s + (long long)(j0 + tid) * n);
            sWs[tid] = bpldT<CPLX>(W, s + (long long)tid * n);
        }
        // y(s): row part of tile (Ib, Ib), column parts of tiles (I2, Ib)
        T ysp = N::zero();
        {
            const int ls = s - 64 * Ib;
            for (int q = tid; q <= nAct; q += kBT) {
                const long long tix = q == 0 ? (long long)Ib * (Ib + 1) / 2 + Ib
                                             : (long long)(Ib + q - 1) * (Ib + q) / 2 + Ib;
                ysp = N::add(ysp, bpldT<CPLX>(tp + tix * 128 * E + (q == 0 ? 0 : 64 * E), ls));
            }
        }
        const double vHy = wg_sum(tid < G ? bpld(scalPub + tid) : 0.0, red);
        const T ys = wg_sumT<CPLX>(ysp, red);   // (its barriers publish sa, sb, sVs, sWs)
        T ps = ys;
        double ab = 0.0;
        for (int k = 0; k < c; ++k) {
            ps = N::sub(ps, N::add(N::mul(sWs[k], sa[k]), N::mul(sVs[k], sb[k])));
            ab += N::re(N::cmul(sa[k], sb[k]));
        }
        const double aw = -0.5 * N::abs2(tauv) * (vHy - 2.0 * ab);
        const T ws = N::add(N::mul(tauv, ps), N::mk(aw, 0.0));
        T *red4 = &rowbuf[0][0][0];   // [3][4][64] reduction scratch (phase T is done)
        double npart = 0.0;
        const int rr = tid & 63, part = tid >> 6;
        for (int I = Ib0 + w; I < Tn; I += G) {
            if (I < Ib) continue;
            const int r = 64 * I + rr;
            const bool live = r >= s && r < n;
            const int nrow = I - Ib + 1, ntot = nrow + (Tn - I);
            T y = N::zero(), corr = N::zero(), cupd = N::zero();
            if (live) {
                {
                    T tmp[kOL];
#pragma unroll
                    for (int t = 0; t < kOL; ++t) {
                        const int q = part + 4 * t;
                        const long long tix = q < nrow ? (long long)I * (I + 1) / 2 + (Ib + q)
                                                       : (long long)(I + q - nrow) * (I + q - nrow + 1) / 2 + I;
                        tmp[t] = q < ntot ? bpldT<CPLX>(tp + tix * 128 * E + (q < nrow ? 0 : 64 * E), rr)
                                          : N::zero();
                    }
#pragma unroll
                    for (int t = 0; t < kOL; ++t) y = N::add(y, tmp[t]);
                }
                const bool first = I == Iown;
#pragma unroll
                for (int t = 0; t < kPanelB / 4; ++t) {
                    const int k = part + 4 * t;
                    if (k < c) {
                        const T vr = first ? pv[t] : bpldT<CPLX>(V, r + (long long)(j0 + k) * n);
                        const T wr = first ? pw[t] : bpldT<CPLX>(W, r + (long long)k * n);
                        corr = N::add(corr, N::add(N::mul(wr, sa[k]), N::mul(vr, sb[k])));
                        cupd = N::add(cupd, N::add(N::mul(vr, N::conj(sWs[k])), N::mul(wr, N::conj(sVs[k]))));
                    }
                }
            }
            red4[(0 * 4 + part) * 64 + rr] = y;
            red4[(1 * 4 + part) * 64 + rr] = corr;
            red4[(2 * 4 + part) * 64 + rr] = cupd;
            __syncthreads();
            if (part == 0 && live) {
                T ya = N::zero(), ca = N::zero(), ua = N::zero();
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    ya = N::add(ya, red4[(0 * 4 + q) * 64 + rr]);
                    ca = N::add(ca, red4[(1 * 4 + q) * 64 + rr]);
                    ua = N::add(ua, red4[(2 * 4 + q) * 64 + rr]);
                }
                const T vr = vat(r);
                const T wr = N::add(N::mul(tauv, N::sub(ya, ca)), N::scale(vr, aw));
                bpstT<CPLX>(W, r + (long long)c * n, wr);
                bpstT<CPLX>(V, r + (long long)i * n, vr);
                // [W | V | W] side by side for the trailing update's single GEMM
                N::store(W, r + (long long)(p.pb + c) * n, vr);
                N::store(W, r + (long long)(2 * p.pb + c) * n, wr);
                if (more) {
                    // column i+1, with this step's own term V(r,i) conj(w(s)) + w(r) conj(V(s,i) = 1)
                    T cn = I == Iown ? pa : N::load(A, r + (long long)(i + 1) * n);
                    cn = N::sub(cn, N::add(ua, N::add(N::mul(vr, N::conj(ws)), wr)));
                    bpstT<CPLX>(cpub, r, cn);
                    if (r == s) dd[s] = N::re(cn);
                    if (r >= s + 2) npart += N::abs2(cn);
                }
            }
            __syncthreads();
        }
        if (more) {
            npart = wg_sum(npart, red);
            if (tid == 0) bpst(normPub + w, npart);
            ++epoch;
            if (!bgroup_barrier(cnt, epoch * (unsigned)G, &bad)) {
                if (tid == 0) p.info[mat] = -2;
                return;
            }
        }
    }
}

}  // namespace

bool trdb_plan(int n, int nb, bool cplx, TrdbPlan &p)
{
    p = TrdbPlan{};
    if (n < 2 || nb < 1 || n > (cplx ? kTrdbMaxComplex : kTrdbMaxReal)) return false;
    const size_t el = cplx ? 16 : 8;
    const size_t lds = (size_t)n * el;
    int dev = 0, cus = 0, perCU = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return false;
    const void *fn = cplx ? reinterpret_cast<const void *>(&k_trdb_panel<true>)
                          : reinterpret_cast<const void *>(&k_trdb_panel<false>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, kBT, lds) != hipSuccess)
        return false;
    const int total = perCU * cus;
    if (total < 64) return false;
    const int groups = std::max(1, std::min(nb, total / 64));
    p.G = std::min(256, total / groups);
    p.nGroups = groups;
    p.lds = lds;
    p.pubLen = (size_t)trdb_pub_len(n, p.G, cplx ? 2 : 1);
    return true;
}

hipError_t launch_trdb_start(bool cplx, const TrdbPlan &pl, const double *A, double *V, double *W,
                             double *d, double *e, double *tau, double *pub, unsigned *cnt,
                             int *info, int n, int nMat, int j0, hipStream_t s)
{
    const long long Tn = (n + 63) / 64;
    TrdbArgs a{A, V, W, d, e, tau, pub, cnt, info, (long long)pl.pubLen, n, pl.G, j0, 0,
               (int)(Tn * (Tn + 1) / 2)};
    if (cplx)
        hipLaunchKernelGGL(k_trdb_start<true>, dim3(nMat), dim3(kBT), 0, s, a);
    else
        hipLaunchKernelGGL(k_trdb_start<false>, dim3(nMat), dim3(kBT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_trdb_panel(bool cplx, const TrdbPlan &pl, const double *A, double *V, double *W,
                             double *d, double *e, double *tau, double *pub, unsigned *cnt,
                             int *info, int n, int nMat, int j0, int pb, hipStream_t s)
{
    const long long Tn = (n + 63) / 64;
    TrdbArgs a{A, V, W, d, e, tau, pub, cnt, info, (long long)pl.pubLen, n, pl.G, j0, pb,
               (int)(Tn * (Tn + 1) / 2)};
    void *args[] = {(void *)&a};
    const void *fn = cplx ? reinterpret_cast<const void *>(&k_trdb_panel<true>)
                          : reinterpret_cast<const void *>(&k_trdb_panel<false>);
    return hipLaunchCooperativeKernel(fn, dim3(nMat * pl.G), dim3(kBT), args, (unsigned)pl.lds, s);
}

}  // namespace olcao

namespace olcao {

namespace {

// C = alpha sum_p P_p + beta C over an m x ncols block (column-major, leading
// dimension ldc; P: parts of m x ncols, contiguous), real or complex (E
// doubles per entry): the sum of a split-K GEMM's partial products.
__global__ void k_splitk_sum(const double *P, int parts, int m, int ncols, int E, double alpha,
                             double beta, double *C, int ldc)
{
    const long long per = (long long)m * ncols * E;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= per) return;
    const long long e = t / E, comp = t % E;
    const long long i = e % m, j = e / m;
    double acc = 0.0;
    for (int q = 0; q < parts; ++q) acc += P[per * q + t];
    double *c = C + E * (i + j * (long long)ldc) + comp;
    *c = beta == 0.0 ? alpha * acc : alpha * acc + beta * *c;
}

}  // namespace

void launch_splitk_sum(const double *P, int parts, int m, int ncols, bool cplx, double alpha,
                       double beta, double *C, int ldc, hipStream_t s)
{
    const int E = cplx ? 2 : 1;
    const long long per = (long long)m * ncols * E;
    hipLaunchKernelGGL(k_splitk_sum, dim3((unsigned)((per + 255) / 256)), dim3(256), 0, s, P, parts, m,
                       ncols, E, alpha, beta, C, ldc);
}

}  // namespace olcao
