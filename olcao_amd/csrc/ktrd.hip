// Batched reduction of Hermitian (real symmetric) matrices to real
// tridiagonal form, and the blocked back-transform of the tridiagonal
// eigenvectors: the zhetrd/dsytrd (uplo 'L') and zunmtr/dormtr steps inside
// zhegv/dsygv (interfaces.F90:73-200), for matrices too large for one
// workgroup's LDS.
//
// Reduction.  One group of G workgroups per matrix, the whole grid launched
// cooperatively; groups walk through the batch.  Workgroup b owns the rows
// r = b, b+G, ... of the full Hermitian matrix: columns 0..127 in registers,
// columns 128..n-1 in LDS, so the matrix is read from global memory once and
// never written back.  Step i builds H_i = I - tau_i v_i v_i^H with
// v_i(i+1) = 1, zeroing A(i+2:n, i) (zhetd2's order), in one pass and one
// group barrier:
//   pass     each owned row r > i takes the pending rank-2 update of step
//            i-1, A(r,:) -= v(r) w^H + w(r) v^H, in place, and accumulates
//            p(r) = tau_i A(r,:) v_i; p(r), A(r, i+1) and the wave's share of
//            p^H v_i are published
//   barrier  the group's arrival counter
//   vectors  the workgroup copies p and column i+1 into LDS, each entry
//            read from the publication buffer once per workgroup; every
//            wave then forms, redundantly, w = p - tau_i/2 (p^H v_i) v_i,
//            the updated column i+1, d(i+1) and the next reflector (zlarfg)
// v, w and the column are register vectors, lane l holding entries l, l+64, ...
//
// Back-transform.  Q = H_0 H_1 ... H_{n-2}; for blocks of kTrdBlock
// reflectors Q_b = I - V_b T_b V_b^H (zlarft, forward, columnwise).  The
// tfactor kernel forms U_b = V_b T_b; the host then applies
// X <- X - U_b (V_b^H X) block by block with batched GEMMs.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "numtraits.hpp"
#include "olcao_internal.h"
#include "wave.hpp"

namespace olcao {

namespace {

constexpr int kTrdThreads = 256;
constexpr int kTrdWaves = kTrdThreads / 64;
constexpr int kTrdK = 14;        // 64-column chunks of a row: n <= 896
constexpr int kTrdKR = 2;        // chunks held in registers (columns 0..127)
// the divide and conquer that follows the reduction keeps per-merge arrays
// of kDcMaxN entries in LDS: the reduction may not take larger matrices
static_assert(64 * kTrdK <= kDcMaxN, "tridiagonal path larger than the divide-and-conquer LDS");
constexpr int kTrdRpw = 5;       // owned rows per wave: ceil(n / G) <= 20
constexpr size_t kTrdLdsMax = 160 * 1024 - 256;
constexpr long long kSpinLimit = 1ll << 22;
constexpr int kRs = 6;           // per-row scalars: v, pending v, pending w, p, column, next column
enum { RV, RVP, RWP, PR, COLR, RCN };

template <bool CPLX>
__device__ __forceinline__ typename Num<CPLX>::T wshfl(typename Num<CPLX>::T a, int lane)
{
    if constexpr (CPLX)
        return {rlane(a.x, lane), rlane(a.y, lane)};
    else
        return rlane(a, lane);
}

template <bool CPLX, int CTRL>
__device__ __forceinline__ typename Num<CPLX>::T quad_swap(typename Num<CPLX>::T a)
{
    if constexpr (CPLX)
        return {dpp<CTRL>(a.x), dpp<CTRL>(a.y)};
    else
        return dpp<CTRL>(a);
}

template <bool CPLX>
__device__ __forceinline__ typename Num<CPLX>::T wsum(typename Num<CPLX>::T a)
{
    if constexpr (CPLX)
        return {wave_sum(a.x), wave_sum(a.y)};
    else
        return wave_sum(a);
}

// entry idx (wave-uniform) of a lane-distributed register vector, on every lane
template <bool CPLX, int KC>
__device__ __forceinline__ typename Num<CPLX>::T bcast(const typename Num<CPLX>::T (&v)[KC],
                                                       int idx)
{
    const int kk = idx >> 6;
    typename Num<CPLX>::T s = Num<CPLX>::zero();
#pragma unroll
    for (int k = 0; k < KC; ++k)
        if (k == kk) s = v[k];
    return wshfl<CPLX>(s, idx & 63);
}

// Publications (p, the column entries, the wave dots) are agent-scope relaxed
// atomic stores and loads: coherent across the device's L2 caches, so the
// barrier needs no cache write-back or invalidation.  Ordering comes from the
// barrier: every wave waits for its outstanding stores to complete
// (s_waitcnt) before the workgroup's arrival is counted, and readers issue
// their loads only after observing the full count.  (Release/acquire fences
// would write back and invalidate L2 at every column; measured 1.3x slower
// per column at n = 576.)
__device__ __forceinline__ void pst(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double pld(const double *p)
{
    return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool CPLX>
__device__ __forceinline__ void pstT(double *p, long long i, typename Num<CPLX>::T v)
{
    if constexpr (CPLX) {
        pst(p + 2 * i, v.x);
        pst(p + 2 * i + 1, v.y);
    } else {
        pst(p + i, v);
    }
}
template <bool CPLX>
__device__ __forceinline__ typename Num<CPLX>::T pldT(const double *p, long long i)
{
    if constexpr (CPLX)
        return {pld(p + 2 * i), pld(p + 2 * i + 1)};
    else
        return pld(p + i);
}

// Group barrier: the workgroup arrives once its waves' publications are
// complete and leaves when all G workgroups of the group have arrived.  The
// spin is bounded; false means the group did not assemble.
__device__ __forceinline__ bool group_barrier(unsigned *cnt, unsigned target, int *bad)
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        int b = 0;
        __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        long long sp = 0;
        while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++sp > kSpinLimit) {
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

// zlarfg/dlarfg on column i of the current matrix, held in cn (entries > i):
// alpha = cn(i+1), x = cn(i+2:n).  Returns v (v(i+1) = 1, zero above), tau
// and beta (the new subdiagonal entry e(i)).
template <bool CPLX, int KC>
__device__ __forceinline__ void reflector(int i, int n, int lane,
                                          const typename Num<CPLX>::T (&cn)[KC],
                                          typename Num<CPLX>::T (&v)[KC],
                                          typename Num<CPLX>::T &tau, double &beta,
                                          typename Num<CPLX>::T &scale)
{
    using N = Num<CPLX>;
    using T = typename N::T;
    const T alpha = bcast<CPLX>(cn, i + 1);
    double part = 0.0;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        const int c = 64 * k + lane;
        if (c > i + 1 && c < n) part += N::abs2(cn[k]);
    }
    const double xnorm2 = wave_sum(part);
    double ar, ai;
    if constexpr (CPLX) {
        ar = alpha.x;
        ai = alpha.y;
    } else {
        ar = alpha;
        ai = 0.0;
    }
    scale = N::mk(1.0, 0.0);
    if (xnorm2 == 0.0 && ai == 0.0) {
        tau = N::zero();
        beta = ar;
    } else {
        const double nrm = sqrt(ar * ar + ai * ai + xnorm2);
        beta = ar >= 0.0 ? -nrm : nrm;
        tau = N::mk((beta - ar) / beta, -ai / beta);
        // 1 / (alpha - beta)
        const double zr = ar - beta, zi = ai;
        const double den = zr * zr + zi * zi;
        scale = N::mk(zr / den, -zi / den);
    }
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        const int c = 64 * k + lane;
        v[k] = c > i + 1 && c < n ? N::mul(cn[k], scale)
               : c == i + 1       ? N::mk(1.0, 0.0)
                                  : N::zero();
    }
}

template <bool CPLX, int KC>
__global__ __launch_bounds__(kTrdThreads) void k_trd(const double *__restrict__ Ain, double *Vout,
                                                     double *dOut, double *eOut, double *tauOut,
                                                     double *pub, unsigned *cnt, int *info, int n,
                                                     int nb, int nGroups, int G)
{
    using N = Num<CPLX>;
    using T = typename N::T;
    constexpr int E = CPLX ? 2 : 1;
    extern __shared__ double lds_raw[];
    T *rows = reinterpret_cast<T *>(lds_raw);
    __shared__ int bad;
    // the wave index is uniform across a wave; readfirstlane lets the
    // compiler treat every row-dependent branch as a scalar branch
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = blockIdx.x % nGroups, b = blockIdx.x / nGroups;
    constexpr int KR = KC < kTrdKR ? KC : kTrdKR;
    const int Lr = n > 64 * KR ? n - 64 * KR : 0;
    const int rowsPer = (n + G - 1) / G;
    // wave-uniform per-row scalars of this wave's rows, after the rows:
    // [row j][RV, RVP, RWP, PR, COLR, RCN]
    T *rsc = rows + rowsPer * Lr + wv * kTrdRpw * kRs;
    // the step's p and column i+1 as the workgroup read them: [p (n) | column (n)]
    T *stP = rows + rowsPer * Lr + kTrdWaves * kTrdRpw * kRs, *stC = stP + n;
    unsigned *myCnt = cnt + 64 * g;
    // two publication buffers per group: [p (n) | column (n) | wave dots (4G)]
    const long long pubLen = (long long)(2 * n + kTrdWaves * G) * E;
    double *pubG = pub + (long long)g * 2 * pubLen;
    unsigned epoch = 0;
    const bool writer = b == 0 && wv == 0;

    for (int mat = g; mat < nb; mat += nGroups) {
        const long long per = (long long)n * n;
        const double *A = Ain + E * per * mat;
        double *V = Vout + E * per * mat;
        double *d = dOut + (long long)n * mat;
        double *e = eOut + (long long)n * mat;
        double *tauG = tauOut + (long long)E * n * mat;

        // owned rows of the full Hermitian matrix from the lower triangle;
        // row j of this wave is r = b + (wv + 4 j) G; columns 0..127 to
        // registers, the rest to LDS (one row at a time)
        T rreg[KR][kTrdRpw];
#pragma unroll
        for (int k = 0; k < KR; ++k)
#pragma unroll
            for (int j = 0; j < kTrdRpw; ++j) {
                const int r = b + (wv + kTrdWaves * j) * G, c = 64 * k + lane;
                T a = N::zero();
                if (r < n && c < n) {
                    a = c <= r ? N::load(A, r + (long long)c * n)
                               : N::conj(N::load(A, c + (long long)r * n));
                    if (c == r) a = N::mk(N::re(a), 0.0);
                }
                rreg[k][j] = a;
            }
#pragma unroll 1
        for (int j = 0; j < kTrdRpw; ++j) {
            const int s = wv + kTrdWaves * j, r = b + s * G;
            if (r >= n) break;
#pragma unroll
            for (int k = KR; k < KC; ++k) {
                const int c = 64 * k + lane;
                if (c < n) {
                    T a = c <= r ? N::load(A, r + (long long)c * n)
                                 : N::conj(N::load(A, c + (long long)r * n));
                    if (c == r) a = N::mk(N::re(a), 0.0);
                    rows[s * Lr + c - 64 * KR] = a;
                }
            }
        }
        // column 0 and reflector 0; per-row scalars of the current (rv) and
        // pending (rvp, rwp) reflector and update vectors
        T cn[KC], vc[KC], vp[KC], wp[KC];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const int c = 64 * k + lane;
            cn[k] = c < n ? N::load(A, c) : N::zero();
            vp[k] = N::zero();
            wp[k] = N::zero();
        }
        if (writer && lane == 0) d[0] = N::re(cn[0]);
        T tau, scale;
        double beta;
        reflector<CPLX, KC>(0, n, lane, cn, vc, tau, beta, scale);
        if (lane < kTrdRpw) {
            const int j = lane, r = b + (wv + kTrdWaves * j) * G;
            const T cr = r < n ? N::load(A, r) : N::zero();
            rsc[j * kRs + RV] = r > 1 ? N::mul(cr, scale) : r == 1 ? N::mk(1.0, 0.0) : N::zero();
            rsc[j * kRs + RVP] = N::zero();
            rsc[j * kRs + RWP] = N::zero();
        }

        for (int i = 0; i < n - 1; ++i) {
            double *pb = pubG + (epoch & 1) * pubLen;
            double *pP = pb, *pC = pb + (long long)E * n, *pPart = pb + 2LL * E * n;
            if (writer) {
#pragma unroll
                for (int k = 0; k < KC; ++k) {
                    const int c = 64 * k + lane;
                    if (c < n) N::store(V, c + (long long)i * n, vc[k]);
                }
                if (lane == 0) {
                    e[i] = beta;
                    N::store(tauG, i, tau);
                }
            }
            // ---- pass over the owned rows: pending update, p(r), column i+1 ----
            // Entries at columns <= i carry zero v and w, so whole live chunks
            // are processed without per-lane column tests.
            const int kLive = (i + 1) >> 6, ci = (i + 1) & 63;
            // the wave's rows side by side: per chunk, all row loads before
            // any store, so that the rows' latencies overlap
            bool live[kTrdRpw];
            T vr[kTrdRpw], wr[kTrdRpw], acc[kTrdRpw], colj[kTrdRpw];
#pragma unroll
            for (int j = 0; j < kTrdRpw; ++j) {
                const int r = b + (wv + kTrdWaves * j) * G;
                live[j] = r < n && r > i;
                vr[j] = live[j] ? rsc[j * kRs + RVP] : N::zero();
                wr[j] = live[j] ? rsc[j * kRs + RWP] : N::zero();
                acc[j] = N::zero();
                colj[j] = N::zero();
            }
#pragma unroll
            for (int k = 0; k < KC; ++k) {
                const int c = 64 * k + lane;
                if (k >= kLive && (k < KC - 1 || c < n)) {
                    T a[kTrdRpw];
#pragma unroll
                    for (int j = 0; j < kTrdRpw; ++j)
                        if (live[j])
                            a[j] = k < KR ? rreg[k < KR ? k : 0][j]
                                          : rows[(wv + kTrdWaves * j) * Lr + c - 64 * KR];
#pragma unroll
                    for (int j = 0; j < kTrdRpw; ++j)
                        if (live[j]) {
                            a[j] = N::sub(N::sub(a[j], N::cmul(wp[k], vr[j])), N::cmul(vp[k], wr[j]));
                            if (k < KR)
                                rreg[k < KR ? k : 0][j] = a[j];
                            else
                                rows[(wv + kTrdWaves * j) * Lr + c - 64 * KR] = a[j];
                            acc[j] = N::add(acc[j], N::mul(a[j], vc[k]));
                            if (k == kLive) colj[j] = wshfl<CPLX>(a[j], ci);
                        }
                }
            }
            T dotw = N::zero();
#pragma unroll
            for (int j = 0; j < kTrdRpw; ++j)
                if (live[j]) {
                    const int r = b + (wv + kTrdWaves * j) * G;
                    const T prj = N::mul(tau, wsum<CPLX>(acc[j]));
                    if (lane == 0) {
                        pstT<CPLX>(pP, r, prj);
                        pstT<CPLX>(pC, r, colj[j]);
                        rsc[j * kRs + PR] = prj;
                        rsc[j * kRs + COLR] = colj[j];
                    }
                    dotw = N::add(dotw, N::cmul(prj, rsc[j * kRs + RV]));
                }
            if (lane == 0) pstT<CPLX>(pPart, b * kTrdWaves + wv, dotw);
            ++epoch;
            if (!group_barrier(myCnt, epoch * (unsigned)G, &bad)) {
                if (threadIdx.x == 0) info[mat] = -2;
                return;
            }
            // ---- vectors: w, column i+1, next reflector ----
            // p and column i+1 (entries i+1..n-1) go through LDS, each entry
            // loaded once per workgroup; all loads are issued before any
            // store, so the step pays one memory round trip
            T part = N::zero();
            {
                constexpr int kPer = (64 * KC + kTrdThreads - 1) / kTrdThreads;
                // the wave dots (kTrdWaves G <= 4 * 64 entries) with the rest
                constexpr int kDots = (kTrdWaves * 64 + 63) / 64;
                T lp[kPer], lc[kPer], ld[kDots];
#pragma unroll
                for (int t = 0; t < kPer; ++t) {
                    const int c = i + 1 + (int)threadIdx.x + kTrdThreads * t;
                    if (c < n) {
                        lp[t] = pldT<CPLX>(pP, c);
                        lc[t] = pldT<CPLX>(pC, c);
                    }
                }
#pragma unroll
                for (int t = 0; t < kDots; ++t) {
                    const int q = lane + 64 * t;
                    ld[t] = q < kTrdWaves * G ? pldT<CPLX>(pPart, q) : N::zero();
                }
#pragma unroll
                for (int t = 0; t < kPer; ++t) {
                    const int c = i + 1 + (int)threadIdx.x + kTrdThreads * t;
                    if (c < n) {
                        stP[c] = lp[t];
                        stC[c] = lc[t];
                    }
                }
#pragma unroll
                for (int t = 0; t < kDots; ++t)
                    if (lane + 64 * t < kTrdWaves * G) part = N::add(part, ld[t]);
            }
            const T dot = wsum<CPLX>(part);
            __syncthreads();
            const T p1 = stP[i + 1];
            const T alpha = N::scale(N::mul(tau, dot), -0.5);
            // w(i+1) = p(i+1) + alpha since v_i(i+1) = 1
            const T w1 = N::add(p1, alpha);
#pragma unroll
            for (int k = 0; k < KC; ++k) {
                const int c = 64 * k + lane;
                const bool on = c < n && c > i;
                const T pk = on ? stP[c] : N::zero(), ck = on ? stC[c] : N::zero();
                const T wk = N::add(pk, N::mul(alpha, vc[k]));
                const T cv = N::sub(N::sub(ck, N::cmul(w1, vc[k])), wk);
                vp[k] = vc[k];
                wp[k] = N::sel(on, wk, N::zero());
                cn[k] = N::sel(on, cv, N::zero());
            }
            // the same for this wave's rows (lane j handles row j)
            if (lane < kTrdRpw) {
                const int j = lane, r = b + (wv + kTrdWaves * j) * G;
                const bool live = r < n && r > i;
                const T rvj = rsc[j * kRs + RV];
                const T wr = live ? N::add(rsc[j * kRs + PR], N::mul(alpha, rvj)) : N::zero();
                rsc[j * kRs + RCN] =
                    live ? N::sub(N::sub(rsc[j * kRs + COLR], N::cmul(w1, rvj)), wr) : N::zero();
                rsc[j * kRs + RVP] = rvj;
                rsc[j * kRs + RWP] = wr;
            }
            if (writer) {
                const double dNext = N::re(bcast<CPLX, KC>(cn, i + 1));
                if (lane == 0) d[i + 1] = dNext;
            }
            if (i + 1 < n - 1) {
                reflector<CPLX, KC>(i + 1, n, lane, cn, vc, tau, beta, scale);
                if (lane < kTrdRpw) {
                    const int j = lane, r = b + (wv + kTrdWaves * j) * G;
                    rsc[j * kRs + RV] = r > i + 2    ? N::mul(rsc[j * kRs + RCN], scale)
                                        : r == i + 2 ? N::mk(1.0, 0.0)
                                                     : N::zero();
                }
            }
        }
        if (threadIdx.x == 0) info[mat] = 0;
    }
}

// Partial Gram matrices of the back-transform blocks: workgroup (block,
// matrix, split) sums G(a, j) = v_a^H v_j over its share of the block's rows
// (rows j0+1..n-1 cut into S = gridDim.z contiguous ranges), so that a
// block's Gram is formed by S workgroups instead of one.  Thread t < 136
// owns the 4 x 4 tile (ta, tj), ta <= tj, of the 64 x 64 upper triangle: per
// row it reads 8 entries from LDS for 16 products.  The partial sums go to
// P[((block * nb + mat) * S + split) * 64 * 64], column-major; entries below
// the diagonal tiles are not written and not read.
constexpr int kTrdGramTiles = 16 * 17 / 2;
template <bool CPLX>
__global__ __launch_bounds__(256) void k_trd_gram(const double *__restrict__ Vall, double *Pall,
                                                  int n)
{
    using N = Num<CPLX>;
    using T = typename N::T;
    constexpr int E = CPLX ? 2 : 1;
    constexpr int NB = kTrdBlock;
    static_assert(NB == 64, "4 x 4 tiles of a 64 x 64 Gram matrix");
    __shared__ T Vs[64][NB + 1];
    const int tid = threadIdx.x;
    const int j0 = blockIdx.x * NB, mat = blockIdx.y, split = blockIdx.z;
    const int nb = gridDim.y, S = gridDim.z;
    const int mb = min(NB, n - 1 - j0);
    const long long per = (long long)n * n;
    const double *V = Vall + E * per * mat;
    const int rows = n - 1 - j0, len = (rows + S - 1) / S;
    const int rs = j0 + 1 + split * len, re = min(n, rs + len);
    // tile of this thread: tiles in row-major order of the upper triangle
    int ta = 0, tj = 0;
    {
        int q = tid;
        while (ta < 16 && q >= 16 - ta) {
            q -= 16 - ta;
            ++ta;
        }
        tj = ta + q;
    }
    const bool owner = tid < kTrdGramTiles;
    T acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = N::zero();
    for (int rb = rs; rb < re; rb += 64) {
        for (int e = tid; e < 64 * NB; e += 256) {
            const int rr = e % 64, a = e / 64, r = rb + rr;
            Vs[rr][a] = (r < re && a < mb) ? N::load(V, r + (long long)(j0 + a) * n) : N::zero();
        }
        __syncthreads();
        if (owner) {
#pragma unroll 2
            for (int rr = 0; rr < 64; ++rr) {
                T va[4], vj[4];
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    va[x] = Vs[rr][4 * ta + x];
                    vj[x] = Vs[rr][4 * tj + x];
                }
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y) acc[x][y] = N::add(acc[x][y], N::cmul(va[x], vj[y]));
            }
        }
        __syncthreads();
    }
    if (owner) {
        double *P = Pall + E * (long long)NB * NB * (((long long)blockIdx.x * nb + mat) * S + split);
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) N::store(P, (4 * ta + x) + NB * (4 * tj + y), acc[x][y]);
    }
}

// U_b = V_b T_b for one block of kTrdBlock reflectors of one matrix.
// V: explicit reflectors (column j = v_j, zero above row j+1, 1 at j+1).
// The Gram matrix G(a, j) = v_a^H v_j (a < j), summed from k_trd_gram's
// partials in a fixed order, is overwritten column by column with T (zlarft:
// T(j, j) = tau_j, T(0:j, j) = -tau_j T(0:j, 0:j) G(0:j, j)).
// SC > 0: the number of partials at compile time (S = SC); 0: S at run time
template <bool CPLX, int SC>
__global__ __launch_bounds__(256) void k_trd_tfactor(const double *__restrict__ Pall,
                                                     const double *__restrict__ tauAll,
                                                     double *Tall, int n, int Srt)
{
    const int S = SC > 0 ? SC : Srt;
    using N = Num<CPLX>;
    using T = typename N::T;
    constexpr int E = CPLX ? 2 : 1;
    constexpr int NB = kTrdBlock;
    __shared__ T Gm[NB][NB + 1];
    __shared__ T tcol[NB];
    const int tid = threadIdx.x;
    const int j0 = blockIdx.x * NB, mat = blockIdx.y;
    const int nref = n - 1;
    const int mb = min(NB, nref - j0);
    const double *tau = tauAll + (long long)E * n * mat;

    // Gram matrix of the block: Gm(a, j) = v_a^H v_j, a < j
    const double *P = Pall + E * (long long)NB * NB * (((long long)blockIdx.x * gridDim.y + mat) * S);
    for (int e = tid; e < NB * NB; e += 256) {
        const int a = e % NB, j = e / NB;
        if (a >= j) continue;
        T g = N::zero();
        if constexpr (SC > 0) {
#pragma unroll
            for (int sp = 0; sp < SC; ++sp) g = N::add(g, N::load(P + E * (long long)NB * NB * sp, e));
        } else {
            for (int sp = 0; sp < S; ++sp) g = N::add(g, N::load(P + E * (long long)NB * NB * sp, e));
        }
        Gm[a][j] = g;
    }
    __syncthreads();
    // T over G, column by column (columns < j already hold T):
    // T(a, j) = -tau_j sum_{a <= q < j} T(a, q) G(q, j); row a's sum is split
    // over the four lanes 4a .. 4a + 3 and combined within the quad
    if (tid < mb) tcol[tid] = N::load(tau, j0 + tid);
    __syncthreads();
    const int ra = tid >> 2, part = tid & 3;
    for (int j = 0; j < mb; ++j) {
        T sum = N::zero();
        if (ra < j)
            for (int q = ra + part; q < j; q += 4) sum = N::add(sum, N::mul(Gm[ra][q], Gm[q][j]));
        sum = N::add(sum, quad_swap<CPLX, kDppQuadSwap1>(sum));
        sum = N::add(sum, quad_swap<CPLX, kDppQuadSwap2>(sum));
        __syncthreads();
        const T tj = tcol[j];
        if (part == 0 && ra < j) Gm[ra][j] = N::scale(N::mul(tj, sum), -1.0);
        if (tid == 0) Gm[j][j] = tj;
        __syncthreads();
    }
    // T (upper triangle) out for the row-parallel product U = V_b T
    double *Tg = Tall + E * (long long)NB * NB * (blockIdx.x + gridDim.x * mat);
    for (int e = tid; e < NB * NB; e += 256) {
        const int a = e % NB, j = e / NB;
        N::store(Tg, e, a <= j && j < mb ? Gm[a][j] : N::zero());
    }
}

// U_b = V_b T_b: workgroup (block, matrix, 64-row chunk), thread (row, 16 columns)
template <bool CPLX>
__global__ __launch_bounds__(256) void k_trd_vt(const double *__restrict__ Vall,
                                                const double *__restrict__ Tall, double *Uall, int n)
{
    using N = Num<CPLX>;
    using T = typename N::T;
    constexpr int E = CPLX ? 2 : 1;
    constexpr int NB = kTrdBlock;
    __shared__ T Vs[64][NB + 1];
    __shared__ T Ts[NB][NB + 1];
    const int tid = threadIdx.x, blk = blockIdx.x, mat = blockIdx.y;
    const int j0 = blk * NB, mb = min(NB, n - 1 - j0);
    const int rb = j0 + 1 + 64 * blockIdx.z;
    if (rb >= n) return;
    const long long per = (long long)n * n;
    const double *V = Vall + E * per * mat;
    const double *Tg = Tall + E * (long long)NB * NB * (blk + gridDim.x * mat);
    double *U = Uall + E * per * mat;
    for (int e = tid; e < NB * NB; e += 256) Ts[e % NB][e / NB] = N::load(Tg, e);
    for (int e = tid; e < 64 * NB; e += 256) {
        const int rr = e % 64, a = e / 64, r = rb + rr;
        Vs[rr][a] = (r < n && a < mb) ? N::load(V, r + (long long)(j0 + a) * n) : N::zero();
    }
    __syncthreads();
    const int rr = tid % 64, jg = tid / 64;
    const int r = rb + rr;
    if (r >= n) return;
    for (int jj = 0; jj < NB / 4; ++jj) {
        const int j = jg * (NB / 4) + jj;
        if (j >= mb) break;
        T sum = N::zero();
        for (int a = 0; a <= j; ++a) sum = N::add(sum, N::mul(Vs[rr][a], Ts[a][j]));
        N::store(U, r + (long long)(j0 + j) * n, sum);
    }
}

// X (field of the eigenvector matrices, complex or real) = Z (real)
__global__ void k_real_to_field(const double *Z, double *X, long long count, int cplx)
{
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= count) return;
    if (cplx) {
        X[2 * t] = Z[t];
        X[2 * t + 1] = 0.0;
    } else {
        X[t] = Z[t];
    }
}

template <bool CPLX, int KC>
const void *trd_fn()
{
    return reinterpret_cast<const void *>(&k_trd<CPLX, KC>);
}

template <bool CPLX>
const void *trd_fn_kc(int kc)
{
    switch (kc) {
    case 2: return trd_fn<CPLX, 2>();
    case 3: return trd_fn<CPLX, 3>();
    case 4: return trd_fn<CPLX, 4>();
    case 5: return trd_fn<CPLX, 5>();
    case 6: return trd_fn<CPLX, 6>();
    case 7: return trd_fn<CPLX, 7>();
    case 8: return trd_fn<CPLX, 8>();
    case 9: return trd_fn<CPLX, 9>();
    case 10: return trd_fn<CPLX, 10>();
    case 11: return trd_fn<CPLX, 11>();
    case 12: return trd_fn<CPLX, 12>();
    case 13: return trd_fn<CPLX, 13>();
    default: return trd_fn<CPLX, 14>();
    }
}

// column chunks of 64: 2..kTrdK
const void *trd_kernel(bool cplx, int n)
{
    const int kc = std::max(2, (n + 63) / 64);
    return cplx ? trd_fn_kc<true>(kc) : trd_fn_kc<false>(kc);
}

}  // namespace

bool trd_plan(int n, int nb, bool cplx, TrdPlan &p)
{
    p = TrdPlan{};
    if (n < 2 || nb < 1 || n > 64 * kTrdK) return false;
    const size_t el = cplx ? 16 : 8;
    // LDS: the owned rows past the register columns, the per-row scalars,
    // the staged p and column
    const size_t Lr = n > 64 * kTrdKR ? (size_t)n - 64 * kTrdKR : 0;
    const size_t fixed = ((size_t)kTrdWaves * kTrdRpw * kRs + 2 * (size_t)n) * el;
    int G = (n + kTrdWaves * kTrdRpw - 1) / (kTrdWaves * kTrdRpw);
    for (; G <= 64; ++G) {
        const size_t rowsPer = ((size_t)n + G - 1) / G;
        if (rowsPer * Lr * el + fixed <= kTrdLdsMax) break;
    }
    if (G > 64) return false;
    const size_t lds = (((size_t)n + G - 1) / G) * Lr * el + fixed;
    int dev = 0, cus = 0, perCU = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return false;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, trd_kernel(cplx, n), kTrdThreads,
                                                     lds) != hipSuccess)
        return false;
    const int groups = std::min(nb, perCU * cus / G);
    if (groups < 1) return false;
    p.G = G;
    p.nGroups = groups;
    p.lds = lds;
    p.pubDoubles = (size_t)groups * 2 * (2 * (size_t)n + kTrdWaves * G) * (cplx ? 2 : 1);
    p.counters = (size_t)groups * 64;
    return true;
}

hipError_t launch_trd(bool cplx, const TrdPlan &p, int n, int nb, const double *A, double *V,
                      double *d, double *e, double *tau, double *pub, unsigned *cnt, int *info,
                      hipStream_t s)
{
    int G = p.G, nGroups = p.nGroups;
    void *args[] = {(void *)&A, (void *)&V, (void *)&d, (void *)&e, (void *)&tau, (void *)&pub,
                    (void *)&cnt, (void *)&info, (void *)&n, (void *)&nb, (void *)&nGroups,
                    (void *)&G};
    return hipLaunchCooperativeKernel(trd_kernel(cplx, n), dim3(nGroups * G),
                                      dim3(kTrdThreads), args, (unsigned)p.lds, s);
}

// Row splits of each block's Gram matrix: enough workgroups to fill the
// device (about kTrdGramWg), at most kTrdGramSplit, one when the blocks of
// the batch already fill it (this also bounds the partial sums' memory)
static int trd_gram_splits(int n, int nb)
{
    const long long units = (long long)((n - 1 + kTrdBlock - 1) / kTrdBlock) * nb;
    return (int)std::max(1LL, std::min<long long>(kTrdGramSplit, kTrdGramWg / std::max(units, 1LL)));
}

size_t trd_tfactor_work(int n, int nb, bool cplx)
{
    const size_t blocks = (n - 1 + kTrdBlock - 1) / kTrdBlock;
    return blocks * nb * (1 + trd_gram_splits(n, nb)) * kTrdBlock * kTrdBlock * (cplx ? 2 : 1);
}

void launch_trd_tfactor(bool cplx, int n, int nb, const double *V, const double *tau, double *Tw,
                        double *U, hipStream_t s)
{
    const int blocks = (n - 1 + kTrdBlock - 1) / kTrdBlock;
    const dim3 g2(blocks, nb, (n + 63) / 64);
    const int S = trd_gram_splits(n, nb);
    const dim3 gg(blocks, nb, S);
    // T for every block first, then the Gram partial sums
    double *P = Tw + (size_t)blocks * nb * kTrdBlock * kTrdBlock * (cplx ? 2 : 1);
    if (cplx) {
        hipLaunchKernelGGL(k_trd_gram<true>, gg, dim3(256), 0, s, V, P, n);
        if (S == kTrdGramSplit)
            hipLaunchKernelGGL((k_trd_tfactor<true, kTrdGramSplit>), dim3(blocks, nb), dim3(256), 0, s,
                               P, tau, Tw, n, S);
        else
            hipLaunchKernelGGL((k_trd_tfactor<true, 0>), dim3(blocks, nb), dim3(256), 0, s, P, tau,
                               Tw, n, S);
        hipLaunchKernelGGL(k_trd_vt<true>, g2, dim3(256), 0, s, V, Tw, U, n);
    } else {
        hipLaunchKernelGGL(k_trd_gram<false>, gg, dim3(256), 0, s, V, P, n);
        if (S == kTrdGramSplit)
            hipLaunchKernelGGL((k_trd_tfactor<false, kTrdGramSplit>), dim3(blocks, nb), dim3(256), 0,
                               s, P, tau, Tw, n, S);
        else
            hipLaunchKernelGGL((k_trd_tfactor<false, 0>), dim3(blocks, nb), dim3(256), 0, s, P, tau,
                               Tw, n, S);
        hipLaunchKernelGGL(k_trd_vt<false>, g2, dim3(256), 0, s, V, Tw, U, n);
    }
}

void launch_real_to_field(const double *Z, double *X, long long count, bool cplx, hipStream_t s)
{
    hipLaunchKernelGGL(k_real_to_field, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, Z,
                       X, count, cplx ? 1 : 0);
}

}  // namespace olcao
