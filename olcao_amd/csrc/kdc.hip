// Batched divide and conquer eigensolver for real symmetric tridiagonal
// matrices T = tridiag(e, d, e): T = Z diag(w) Z^T, w ascending.  This is the
// tridiagonal step of zheevd/dsyevd inside zhegv/dsygv's solution
// (interfaces.F90:73-200), written for many matrices at once.
//
// Every coupling e(i) is torn (Cuppen): T = diag(T1, T2) + rho u u^T with
// rho = |e|, u = e_last(T1) + sign(e) e_first(T2), T1 and T2 taking -rho on
// the touching diagonal entries, down to 1 x 1 leaves.  The merge of a node
// of size m, given its children's eigenpairs (Q = diag(Q1, Q2), D):
//   z = Q^T u / |Q^T u|, rho <- rho |Q^T u|^2, poles sorted ascending
//   deflation (tolerance 8 eps max(|d|max, |z|max)): small rho z(j), and
//     Givens rotations merging poles closer than the tolerance allows, in
//     one ascending scan (LAPACK dlaed2's rules)
//   secular equation 1 + rho sum z^2 / (d - lambda) = 0 for each kept root,
//     bracketed, in coordinates shifted to the nearer pole, with a two-pole
//     rational model step (a wave per root)
//   Gu-Eisenstat: z recomputed from the roots, eigenvectors
//     u_j(i) = zhat(i) / (d(i) - lambda_j), so they are orthogonal to
//     working precision whatever the root accuracy
//   W (m x m, rows in sorted-pole order): the new eigenvectors are Q P W,
//     P the pole sort; the deflation rotations are folded into W's rows.
// A tiled GEMM then forms Q P W and the node block is copied back.  All
// merges of one tree height run in one launch per step.
#include <hip/hip_runtime.h>

#include <cmath>

#include "olcao_internal.h"
#include "wave.hpp"

namespace olcao {

namespace {

constexpr double kEps = 1.1102230246251565e-16;   // 2^-53
constexpr int kDcMax = kDcMaxN;
constexpr int kSecularIters = 64;
constexpr int kTile = 64;
constexpr int kDcBig = 1024;     // threads of a merge of more than kDcSmallM
// threads of a whole small merge in one workgroup: its LDS state holds the
// workgroup to one per CU, so sixteen waves cost no occupancy and solve the
// secular roots (one per wave at a time) four times as wide as four would
constexpr int kDcSmall = 1024;
constexpr int kDcSmallM = 128;

// Four wave sums side by side.
__device__ __forceinline__ void wave_sum4(double &a, double &b, double &c, double &d)
{
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    d = wave_sum(d);
}

// 1/x from the hardware estimate and two Newton steps (about 1 ulp; the
// arguments here are differences of eigenvalue estimates, far from the
// denormal range)
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Block-wide reduction through LDS (all threads call it).
template <bool MAX, int NT>
__device__ double block_reduce(double v, double *red)
{
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    v = MAX ? wave_max(v) : wave_sum(v);
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double r = 0.0;
    for (int q = 0; q < NT / 64; ++q) r = MAX ? fmax(r, red[q]) : r + red[q];
    return r;
}

// Exclusive prefix count of flags over [0, m) in LDS (m <= kDcMax): per-thread
// segments, a wave scan of the segment sums, then the waves' totals.
template <int NT>
__device__ void block_scan(const int *flag, int *pre, int m, int *red)
{
    const int kSeg = (m + NT - 1) / NT;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int b0 = t * kSeg, b1 = min(m, b0 + kSeg);
    // a thread's segment holds flags other threads wrote
    __syncthreads();
    int sum = 0;
    for (int p = b0; p < b1; ++p) sum += flag[p];
    int inc = sum;   // inclusive scan within the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
    }
    __syncthreads();
    if (lane == 63) red[wv] = inc;
    __syncthreads();
    int base = 0;
    for (int q = 0; q < wv; ++q) base += red[q];
    int run = base + inc - sum;
    for (int p = b0; p < b1; ++p) {
        pre[p] = run;
        run += flag[p];
    }
    __syncthreads();
}

__global__ void k_dc_init(double *Dall, const double *Eall, double *Qall, int n)
{
    const int mat = blockIdx.y;
    double *D = Dall + (size_t)n * mat;
    const double *E = Eall + (size_t)n * mat;
    double *Q = Qall + (size_t)n * n * mat;
    const size_t nn = (size_t)n * n;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nn;
         t += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(t % n), c = (int)(t / n);
        Q[t] = r == c ? 1.0 : 0.0;
        if (r == c) {
            double v = D[r];
            if (r > 0) v -= fabs(E[r - 1]);
            if (r < n - 1) v -= fabs(E[r]);
            D[r] = v;
        }
    }
}

// A leaf block (off, m <= kDcLeaf) solved directly: its tridiagonal (the
// diagonal as k_dc_init tore it, plus the couplings inside the block, so the
// block's outer edges stay torn as the merges above expect) by cyclic Jacobi
// in one wave, the matrix in LDS, every rotation of a sweep in turn; then the
// eigenvalues ascending into D and their vectors into the block of Q.  The
// sweeps stop when the off-diagonal part is below 2^-53 of the diagonal
// (norms squared compared) or after kDcLeafSweeps.
constexpr int kDcLeafSweeps = 24;
__global__ __launch_bounds__(64) void k_dc_leaf(const int2 *leaves, int n, double *Dall,
                                                const double *Eall, double *Qall)
{
    constexpr int ML = kDcLeaf;
    const int2 lf = leaves[blockIdx.x];
    const int off = lf.x, m = lf.y, mat = blockIdx.y;
    double *D = Dall + (size_t)n * mat;
    const double *E = Eall + (size_t)n * mat;
    double *Q = Qall + (size_t)n * n * mat;
    __shared__ double A[ML][ML + 1], V[ML][ML + 1];
    __shared__ double rC[ML / 2], rS[ML / 2], rApp[ML / 2], rAqq[ML / 2];
    __shared__ int rP[ML / 2], rQ[ML / 2], rOn[ML / 2];
    static_assert(ML % 2 == 0, "round-robin pairs of an even order");
    const int lane = threadIdx.x;
    const int M = m + (m & 1), H = M / 2;
    for (int e = lane; e < ML * ML; e += 64) {
        const int r = e / ML, c = e - r * ML;
        double v = 0.0;
        if (r < m && c < m) {
            if (r == c) {
                v = D[off + r];
                if (r > 0) v += fabs(E[off + r - 1]);
                if (r < m - 1) v += fabs(E[off + r]);
            } else if (c == r + 1) {
                v = E[off + r];
            } else if (r == c + 1) {
                v = E[off + c];
            }
        }
        A[r][c] = v;
        V[r][c] = r == c ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int sweep = 0; sweep < kDcLeafSweeps; ++sweep) {
        double offn = 0.0, dn = 0.0;
        for (int e = lane; e < m * m; e += 64) {
            const int r = e / m, c = e - r * m;
            const double a = A[r][c];
            if (r == c)
                dn += a * a;
            else
                offn += a * a;
        }
        offn = wave_sum(offn);
        dn = wave_sum(dn);
        if (!(offn > kEps * kEps * dn)) break;
        // round-robin order: M - 1 rounds of M / 2 disjoint pairs (index 0
        // fixed, the others rotating; a pair with the pad index M - 1 >= m of
        // an odd m is idle), the round's rotations computed from the same A
        // and applied together: columns (with V), rows, the lower triangle
        // copied from the upper, then each pair's 2 x 2 block set exactly
        for (int r = 0; r < M - 1; ++r) {
            if (lane < H) {
                const int a = lane == 0 ? 0 : (r + lane - 1) % (M - 1) + 1;
                const int b = (r + M - 2 - lane) % (M - 1) + 1;
                const int p = min(a, b), q = max(a, b);
                const double apq = q < m ? A[p][q] : 0.0;
                double c = 1.0, s = 0.0, tt = 0.0;
                if (apq != 0.0) {
                    const double app = A[p][p], aqq = A[q][q];
                    const double th = (aqq - app) / (2.0 * apq);
                    tt = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(1.0 + th * th));
                    c = 1.0 / sqrt(1.0 + tt * tt);
                    s = tt * c;
                    rApp[lane] = app - tt * apq;
                    rAqq[lane] = aqq + tt * apq;
                }
                rP[lane] = p;
                rQ[lane] = q;
                rOn[lane] = apq != 0.0;
                rC[lane] = c;
                rS[lane] = s;
            }
            __syncthreads();
            for (int e = lane; e < m * H; e += 64) {
                const int i = e / H, k = e - i * H;
                if (!rOn[k]) continue;
                const int p = rP[k], q = rQ[k];
                const double c = rC[k], s = rS[k];
                const double ap = A[i][p], aq = A[i][q];
                A[i][p] = c * ap - s * aq;
                A[i][q] = s * ap + c * aq;
                const double vp = V[i][p], vq = V[i][q];
                V[i][p] = c * vp - s * vq;
                V[i][q] = s * vp + c * vq;
            }
            __syncthreads();
            for (int e = lane; e < H * m; e += 64) {
                const int k = e / m, j = e - k * m;
                if (!rOn[k]) continue;
                const int p = rP[k], q = rQ[k];
                const double c = rC[k], s = rS[k];
                const double ap = A[p][j], aq = A[q][j];
                A[p][j] = c * ap - s * aq;
                A[q][j] = s * ap + c * aq;
            }
            __syncthreads();
            for (int e = lane; e < m * m; e += 64) {
                const int i = e / m, j = e - i * m;
                if (i > j) A[i][j] = A[j][i];
            }
            __syncthreads();
            if (lane < H && rOn[lane]) {
                const int p = rP[lane], q = rQ[lane];
                A[p][p] = rApp[lane];
                A[q][q] = rAqq[lane];
                A[p][q] = 0.0;
                A[q][p] = 0.0;
            }
            __syncthreads();
        }
    }
    // eigenvalues ascending (ties by index), vectors with them
    if (lane < m) {
        const double lam = A[lane][lane];
        int rank = 0;
        for (int k = 0; k < m; ++k) {
            const double lk = A[k][k];
            rank += (lk < lam || (lk == lam && k < lane)) ? 1 : 0;
        }
        D[off + rank] = lam;
        for (int i = 0; i < m; ++i) Q[(size_t)(off + i) + (size_t)n * (off + rank)] = V[i][lane];
    }
}

// LDS state of one merge.
struct MergeLds {
    double sd[kDcMax], sz[kDcMax], ds[kDcMax], zs[kDcMax];
    double dk[kDcMax], zk[kDcMax], zz[kDcMax];
    int order[kDcMax], keep[kDcMax], rowKeep[kDcMax], defl[kDcMax], scratch[kDcMax];
    double dv[kDcMax];                 // deflated values, ascending
    int dp[kDcMax];                    // their sorted-pole positions
    int rotP[kDcMax], rotQ[kDcMax];
    double rotC[kDcMax], rotS[kDcMax];
    double red[16];
    int ired[16];
    int K, Nd, Nr, close, unsorted;
    double rho;
};
// The merge kernels hold MergeLds plus up to three double and two int
// arrays of kDcMax entries in LDS; all of it must fit one workgroup's
// 160 KiB, for whatever kDcMaxN is set to.
constexpr size_t kDcMergeLdsBytes = sizeof(MergeLds) + (size_t)kDcMax * (3 * 8 + 2 * 4);
static_assert(kDcMergeLdsBytes <= 160 * 1024, "divide-and-conquer merge state exceeds LDS");
// a large merge's preparation keeps 20 B per entry in LDS beside MergeView
static_assert((size_t)kDcBigMaxN * 20 + 512 <= 160 * 1024, "large-merge LDS exceeds 160 KiB");
static_assert(kTrdbMaxReal <= kDcBigMaxN && kTrdbMaxComplex <= kDcBigMaxN,
              "blocked reduction larger than the divide and conquer");

// Sort the children's poles, form z, deflate (dlaed2's rules: small rho z,
// then Givens rotations of poles closer than the tolerance, in one ascending
// scan that runs only when a parallel test finds such a pair), and leave the
// kept poles dk, z, rho z^2 and the ascending deflated values in L.
template <int NT, class ST>
__device__ void dc_prepare(int n, int off, int m1, int m, const double *D, const double *E,
                           const double *Q, int *ord, ST &L)
{
    const int tid = threadIdx.x;
    const double beta = E[off + m1 - 1];
    const double sgn = beta < 0.0 ? -1.0 : 1.0;
    // children's eigenvalues and z = (last row of Q1, sign(beta) first row of Q2)
    double nzp = 0.0;
    for (int i = tid; i < m; i += NT) {
        L.sd[i] = D[off + i];
        const double v = i < m1 ? Q[(size_t)(off + m1 - 1) + (size_t)n * (off + i)]
                                : sgn * Q[(size_t)(off + m1) + (size_t)n * (off + i)];
        L.sz[i] = v;
        nzp += v * v;
    }
    const double nz2 = block_reduce<false, NT>(nzp, L.red);
    const double nz = sqrt(nz2);
    const double rho = fabs(beta) * nz2;
    // merge order of the two sorted children (ties: first child first)
    for (int i = tid; i < m; i += NT) {
        const double v = L.sd[i];
        int lo, hi;
        if (i < m1) {   // count second-child entries < v
            lo = m1;
            hi = m;
            while (lo < hi) {
                const int md = (lo + hi) >> 1;
                if (L.sd[md] < v) lo = md + 1; else hi = md;
            }
            L.order[i + (lo - m1)] = i;
        } else {        // count first-child entries <= v
            lo = 0;
            hi = m1;
            while (lo < hi) {
                const int md = (lo + hi) >> 1;
                if (L.sd[md] <= v) lo = md + 1; else hi = md;
            }
            L.order[(i - m1) + lo] = i;
        }
    }
    __syncthreads();
    double dmx = 0.0, zmx = 0.0;
    for (int p = tid; p < m; p += NT) {
        L.ds[p] = L.sd[L.order[p]];
        L.zs[p] = nz > 0.0 ? L.sz[L.order[p]] / nz : 0.0;
        dmx = fmax(dmx, fabs(L.ds[p]));
        zmx = fmax(zmx, fabs(L.zs[p]));
        ord[off + p] = off + L.order[p];
        ord[n + off + L.order[p]] = p;   // inverse: sorted position of pole L.order[p]
    }
    dmx = block_reduce<true, NT>(dmx, L.red);
    zmx = block_reduce<true, NT>(zmx, L.red);
    const double tol = 8.0 * kEps * fmax(dmx, zmx);
    const bool allDefl = !(rho * zmx > tol);
    int *big = L.scratch, *pre = L.rowKeep;   // rowKeep is set below
    for (int p = tid; p < m; p += NT) big[p] = !allDefl && rho * fabs(L.zs[p]) > tol;
    if (tid == 0) L.close = 0;
    block_scan<NT>(big, pre, m, L.ired);
    const int nBig = m > 0 ? pre[m - 1] + big[m - 1] : 0;
    for (int p = tid; p < m; p += NT)
        if (big[p]) L.keep[pre[p]] = p;
    __syncthreads();
    // each pair of consecutive kept poles, close or not with the values as
    // they are now (L.dp, free until the deflated values are ranked below,
    // holds the verdict of the pair ending at the q-th kept pole)
    for (int q = tid + 1; q < nBig; q += NT) {
        const int pj = L.keep[q - 1], nj = L.keep[q];
        double s = L.zs[pj], c = L.zs[nj];
        const double tau = sqrt(c * c + s * s);
        c /= tau;
        s = -s / tau;
        const int cl = fabs((L.ds[nj] - L.ds[pj]) * c * s) <= tol;
        L.dp[q] = cl;
        if (cl) L.close = 1;
    }
    __syncthreads();
    if (!L.close) {
        for (int p = tid; p < m; p += NT)
            if (!big[p]) L.defl[p - pre[p]] = p;
        if (tid == 0) {
            L.K = nBig;
            L.Nd = m - nBig;
            L.Nr = 0;
        }
    } else if (tid == 0) {
        // the sequential scan: a rotation changes the values of the pair's
        // second pole, which the next pair reads; a pair after one that did
        // not rotate sees the values the parallel pass saw, and takes its
        // verdict (the same arithmetic on the same values)
        int K = 0, nd = 0, nr = 0, q = 0;
        int pj = -1;
        bool moved = false;
        for (int nj = 0; nj < m; ++nj) {
            if (!big[nj]) {
                L.defl[nd++] = nj;
                continue;
            }
            const int qn = q++;
            if (pj < 0) {
                pj = nj;
                continue;
            }
            if (!moved && !L.dp[qn]) {
                L.keep[K++] = pj;
                pj = nj;
                continue;
            }
            double s = L.zs[pj], c = L.zs[nj];
            const double tau = sqrt(c * c + s * s);
            const double t = L.ds[nj] - L.ds[pj];
            c /= tau;
            s = -s / tau;
            if (fabs(t * c * s) <= tol) {
                L.zs[nj] = tau;
                L.zs[pj] = 0.0;
                L.rotP[nr] = pj;
                L.rotQ[nr] = nj;
                L.rotC[nr] = c;
                L.rotS[nr] = s;
                ++nr;
                const double t2 = L.ds[pj] * c * c + L.ds[nj] * s * s;
                L.ds[nj] = L.ds[pj] * s * s + L.ds[nj] * c * c;
                L.ds[pj] = t2;
                L.defl[nd++] = pj;
                moved = true;
            } else {
                L.keep[K++] = pj;
                moved = false;
            }
            pj = nj;
        }
        if (pj >= 0) L.keep[K++] = pj;
        L.K = K;
        L.Nd = nd;
        L.Nr = nr;
    }
    if (tid == 0) {
        L.unsorted = 0;
        L.rho = rho;
    }
    __syncthreads();
    const int K = L.K, Nd = L.Nd;
    for (int p = tid; p < m; p += NT) L.rowKeep[p] = -1;
    __syncthreads();
    for (int ii = tid; ii < K; ii += NT) {
        L.dk[ii] = L.ds[L.keep[ii]];
        L.zk[ii] = L.zs[L.keep[ii]];
        L.zz[ii] = rho * L.zk[ii] * L.zk[ii];
        L.rowKeep[L.keep[ii]] = ii;
    }
    for (int q = tid; q + 1 < Nd; q += NT)
        if (L.ds[L.defl[q]] > L.ds[L.defl[q + 1]]) L.unsorted = 1;
    __syncthreads();
    // deflated values ascending: as scanned unless a rotation moved one
    for (int q = tid; q < Nd; q += NT) {
        const double v = L.ds[L.defl[q]];
        int rank = q;
        if (L.unsorted) {
            rank = 0;
            for (int q2 = 0; q2 < Nd; ++q2) {
                const double u = L.ds[L.defl[q2]];
                rank += (u < v) || (u == v && q2 < q);
            }
        }
        L.dv[rank] = v;
        L.dp[rank] = L.defl[q];
    }
    __syncthreads();
}

// Root j of 1 + sum zz / (dk - lambda) = 0 by one wave: bracketed, in
// coordinates shifted to the nearer pole (org), two-pole rational model
// steps; lambda = dk(org) + tau.
template <bool BIG = false>
__device__ void dc_root(int j, int K, const double *dk, const double *zz, int &orgOut,
                        double &tauOut)
{
    const int lane = threadIdx.x & 63;
    int org;
    double a, b;
    if (j < K - 1) {
        const double mid = 0.5 * (dk[j] + dk[j + 1]);
        double part = 0.0;
        for (int ii = lane; ii < K; ii += 64) part += zz[ii] / (dk[ii] - mid);
        const double g = 1.0 + wave_sum(part);
        if (g >= 0.0) {
            org = j;
            a = 0.0;
            b = mid - dk[j];
        } else {
            org = j + 1;
            a = mid - dk[j + 1];
            b = 0.0;
        }
    } else {
        double part = 0.0;
        for (int ii = lane; ii < K; ii += 64) part += zz[ii];
        org = j;
        a = 0.0;
        b = wave_sum(part);
    }
    const double dorg = dk[org];
    const double dj = dk[j] - dorg;
    const double dj1 = j < K - 1 ? dk[j + 1] - dorg : 0.0;
    double tau = 0.5 * (a + b);
    for (int it = 0; it < kSecularIters; ++it) {
        // left poles (ii <= j) give negative terms, right poles positive;
        // the lane's terms unrolled so their reciprocals overlap
        double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0;
        auto term = [&](int ii, bool on) {
            const int iu = on ? ii : 0;
            const double r = rcp_nr((dk[iu] - dorg) - tau);
            const double t = on ? zz[iu] * r : 0.0;
            const bool left = ii <= j;
            psi += left ? t : 0.0;
            dpsi += left ? t * r : 0.0;
            phi += left ? 0.0 : t;
            dphi += left ? 0.0 : t * r;
        };
        if constexpr (BIG) {
            // the same terms in the same order per lane, any K
            for (int ii = lane; ii < K; ii += 64) term(ii, true);
        } else {
#pragma unroll
            for (int q = 0; q < kDcMax / 64; ++q) {
                const int ii = lane + 64 * q;
                if (64 * q < K) term(ii, ii < K);
            }
        }
        wave_sum4(psi, dpsi, phi, dphi);
        const double g = 1.0 + psi + phi;
        if (g < 0.0) a = tau; else b = tau;
        if (b - a <= 4.0 * kEps * fmax(fabs(a), fabs(b)) || fabs(g) <= 8.0 * kEps * (1.0 + phi - psi))
            break;
        // two-pole rational model through (psi, psi') and (phi, phi')
        const double dl = dj - tau;
        const double s1 = dpsi * dl * dl, c1 = psi - s1 / dl;
        double x = 0.0;
        bool ok = false;
        if (j < K - 1) {
            const double dr = dj1 - tau;
            const double s2 = dphi * dr * dr, c2 = phi - s2 / dr;
            const double C = 1.0 + c1 + c2;
            const double B2 = -(C * (dj + dj1) + s1 + s2);
            const double C2 = C * dj * dj1 + s1 * dj1 + s2 * dj;
            if (C != 0.0) {
                const double sq = sqrt(fmax(B2 * B2 - 4.0 * C * C2, 0.0));
                const double q = -0.5 * (B2 + copysign(sq, B2));
                const double x1 = q / C;
                const double x2 = q != 0.0 ? C2 / q : INFINITY;
                const bool ok1 = a < x1 && x1 < b, ok2 = a < x2 && x2 < b;
                if (ok1 && ok2) {
                    x = fabs(x1 - tau) <= fabs(x2 - tau) ? x1 : x2;
                    ok = true;
                } else if (ok1) {
                    x = x1;
                    ok = true;
                } else if (ok2) {
                    x = x2;
                    ok = true;
                }
            } else if (B2 != 0.0) {
                x = -C2 / B2;
                ok = a < x && x < b;
            }
        } else {
            const double C = 1.0 + c1 + phi;
            if (C != 0.0) {
                x = dj + s1 / C;
                ok = a < x && x < b;
            }
        }
        if (ok && fabs(x - tau) <= 2.0 * kEps * fabs(x)) {
            tau = x;
            break;
        }
        tau = ok ? x : 0.5 * (a + b);
    }
    orgOut = org;
    tauOut = tau;
}

// Gu-Eisenstat: zhat(i)^2 = (lambda_i - d_i)/rho prod_{j!=i} (lambda_j - d_i)/(d_j - d_i),
// with d_i - lambda_j = (dk(i) - dk(org_j)) - tau_j as the secular iteration formed it.
__device__ double dc_zhat(int ii, int K, const double *dk, const double *zk, double rho,
                          const int *rorg, const double *rtau)
{
    auto delta = [&](int i, int j) { return (dk[i] - dk[rorg[j]]) - rtau[j]; };
    double p0 = -delta(ii, ii) / rho, p1 = 1.0;
    int j = 0;
    for (; j + 1 < K; j += 2) {
        const double a0 = j == ii ? 1.0 : -delta(ii, j) * rcp_nr(dk[j] - dk[ii]);
        const double a1 = j + 1 == ii ? 1.0 : -delta(ii, j + 1) * rcp_nr(dk[j + 1] - dk[ii]);
        p0 *= a0;
        p1 *= a1;
    }
    if (j < K && j != ii) p0 *= -delta(ii, j) * rcp_nr(dk[j] - dk[ii]);
    return copysign(sqrt(fmax(p0 * p1, 0.0)), zk[ii]);
}

// Final columns [c0, c1) of the merge: the eigenvalue of each (roots before
// equal deflated values), W's column in the sorted-pole basis (kept rows
// zhat(i) / (d_i - lambda_j) normalised, or the deflated unit vector), and the
// deflation rotations on W's rows.  colInfo: LDS scratch of kDcMax ints.
template <int NT>
__device__ void dc_columns(int n, int off, int m, int c0, int c1, const MergeLds &L,
                           const double *lam, const int *rorg, const double *rtau,
                           const double *zh, int *colInfo, double *D, double *W)
{
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    constexpr int NW = NT / 64;
    const int K = L.K, Nd = L.Nd, Nr = L.Nr;
    auto Wat = [&](int row, int col) -> double & { return W[(size_t)(off + row) + (size_t)n * (off + col)]; };
    for (int j = tid; j < K; j += NT) {
        const double v = lam[j];
        int lo = 0, hi = Nd;
        while (lo < hi) {
            const int md = (lo + hi) >> 1;
            if (L.dv[md] < v) lo = md + 1; else hi = md;
        }
        const int s = j + lo;
        if (s >= c0 && s < c1) {
            colInfo[s] = j;
            D[off + s] = v;
        }
    }
    for (int q = tid; q < Nd; q += NT) {
        const double v = L.dv[q];
        int lo = 0, hi = K;
        while (lo < hi) {
            const int md = (lo + hi) >> 1;
            if (lam[md] <= v) lo = md + 1; else hi = md;
        }
        const int s = q + lo;
        if (s >= c0 && s < c1) {
            colInfo[s] = -1 - L.dp[q];
            D[off + s] = v;
        }
    }
    __syncthreads();
    for (int s = c0 + wv; s < c1; s += NW) {
        const int info = colInfo[s];
        if (info >= 0) {
            const int j = info;
            const double dj = L.dk[rorg[j]], tj = rtau[j];
            double u[kDcMax / 64];
            double part = 0.0;
#pragma unroll
            for (int q = 0; q < kDcMax / 64; ++q) {
                const int ii = lane + 64 * q;
                const bool on = ii < K;
                const int iu = on ? ii : 0;
                u[q] = on ? zh[iu] * rcp_nr((L.dk[iu] - dj) - tj) : 0.0;
                part += u[q] * u[q];
            }
            const double inv = 1.0 / sqrt(wave_sum(part));
#pragma unroll
            for (int q = 0; q < kDcMax / 64; ++q) {
                const int ii = lane + 64 * q;
                if (ii < K) Wat(L.keep[ii], s) = u[q] * inv;
            }
            for (int q = lane; q < Nd; q += 64) Wat(L.defl[q], s) = 0.0;
        } else {
            const int pd = -1 - info;
            for (int p = lane; p < m; p += 64) Wat(p, s) = p == pd ? 1.0 : 0.0;
        }
    }
    __syncthreads();
    // deflation rotations on W's rows, last first; a chain (p1,p2),(p2,p3),...
    // keeps its running row in a register
    if (Nr > 0) {
        for (int s = c0 + tid; s < c1; s += NT) {
            int carryRow = -1;
            double carry = 0.0;
            for (int q = Nr - 1; q >= 0; --q) {
                const int p = L.rotP[q], r = L.rotQ[q];
                const double c = L.rotC[q], sn = L.rotS[q];
                double y;
                if (r == carryRow) {
                    y = carry;
                } else {
                    if (carryRow >= 0) Wat(carryRow, s) = carry;
                    y = Wat(r, s);
                }
                const double x = Wat(p, s);
                Wat(r, s) = sn * x + c * y;
                carry = c * x - sn * y;
                carryRow = p;
            }
            Wat(carryRow, s) = carry;
        }
    }
}

// One whole merge in one workgroup (small merges).
template <int NT>
__global__ __launch_bounds__(NT) void k_dc_merge(const int4 *nodes, int n, double *Dall,
                                                 const double *Eall, const double *Qall,
                                                 double *Wall, int *ordAll)
{
    constexpr int NW = NT / 64;
    const int4 nd = nodes[blockIdx.x];
    const int off = nd.x, m1 = nd.y, m = nd.z;
    const int mat = blockIdx.y;
    double *D = Dall + (size_t)n * mat;
    __shared__ MergeLds L;
    __shared__ double lam[kDcMax], zh[kDcMax], rtau[kDcMax];
    __shared__ int rorg[kDcMax], colInfo[kDcMax];
    dc_prepare<NT>(n, off, m1, m, D, Eall + (size_t)n * mat, Qall + (size_t)n * n * mat,
                   ordAll + (size_t)2 * n * mat, L);
    const int K = L.K, lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int j = wv; j < K; j += NW) {
        int org;
        double tau;
        dc_root(j, K, L.dk, L.zz, org, tau);
        if (lane == 0) {
            lam[j] = L.dk[org] + tau;
            rorg[j] = org;
            rtau[j] = tau;
        }
    }
    __syncthreads();
    for (int ii = threadIdx.x; ii < K; ii += NT) zh[ii] = dc_zhat(ii, K, L.dk, L.zk, L.rho, rorg, rtau);
    __syncthreads();
    dc_columns<NT>(n, off, m, 0, m, L, lam, rorg, rtau, zh, colInfo, D, Wall + (size_t)n * n * mat);
}

// Large merges in three launches, so roots and columns spread over many
// workgroups.  The node's prepared state moves through global memory:
//   st[off + i] (per matrix, n entries each): dk, zk, dv, lam, tau (double)
//   ist[...]: keep, dp, defl, org, rowKeep (int); node scalars in ns[node]
struct DcState {
    double *dk, *zk, *dv, *lam, *tau, *rotC, *rotS;
    int *keep, *dp, *defl, *org, *rowKeep, *rotP, *rotQ;
    // merges larger than kDcMaxN only: zhat, and the preparation's work arrays
    double *zh, *sd, *sz, *zz;
    int *order;
    int4 *cnt;      // per node x matrix: K, Nd, Nr
    double *rho;    // per node x matrix
};

template <int NT>
__global__ __launch_bounds__(NT) void k_dc_prep(const int4 *nodes, int nodeBase, int nNodes, int n,
                                                const double *Dall, const double *Eall,
                                                const double *Qall, int *ordAll, DcState st)
{
    const int node = nodeBase + blockIdx.x, mat = blockIdx.y;
    const int4 nd = nodes[node];
    const int off = nd.x, m1 = nd.y, m = nd.z;
    __shared__ MergeLds L;
    dc_prepare<NT>(n, off, m1, m, Dall + (size_t)n * mat, Eall + (size_t)n * mat,
                   Qall + (size_t)n * n * mat, ordAll + (size_t)2 * n * mat, L);
    const size_t b = (size_t)n * mat + off;
    for (int i = threadIdx.x; i < m; i += NT) {
        if (i < L.K) {
            st.dk[b + i] = L.dk[i];
            st.zk[b + i] = L.zk[i];
            st.keep[b + i] = L.keep[i];
        }
        if (i < L.Nd) {
            st.dv[b + i] = L.dv[i];
            st.dp[b + i] = L.dp[i];
            st.defl[b + i] = L.defl[i];
        }
        if (i < L.Nr) {
            st.rotP[b + i] = L.rotP[i];
            st.rotQ[b + i] = L.rotQ[i];
            st.rotC[b + i] = L.rotC[i];
            st.rotS[b + i] = L.rotS[i];
        }
        st.rowKeep[b + i] = L.rowKeep[i];
    }
    if (threadIdx.x == 0) {
        st.cnt[(size_t)mat * nNodes + blockIdx.x] = make_int4(L.K, L.Nd, L.Nr, 0);
        st.rho[(size_t)mat * nNodes + blockIdx.x] = L.rho;
    }
}

constexpr int kRootsPerWg = 16;
// one wave per root: a root's secular iteration is a chain of dependent
// wave reductions, so the workgroup's roots run side by side
constexpr int kRootThreads = 64 * kRootsPerWg;
constexpr int kColsPerWg = 64;

// tiles: (node index within the level, chunk)
__global__ __launch_bounds__(kRootThreads) void k_dc_roots(const int2 *tiles, const int4 *nodes,
                                                           int nodeBase, int nNodes, int n,
                                                           DcState st)
{
    const int2 tl = tiles[blockIdx.x];
    const int mat = blockIdx.y;
    const int4 nd = nodes[nodeBase + tl.x];
    const int off = nd.x;
    const int4 c = st.cnt[(size_t)mat * nNodes + tl.x];
    const int K = c.x;
    const int j0 = tl.y * kRootsPerWg;
    if (j0 >= K) return;
    const double rho = st.rho[(size_t)mat * nNodes + tl.x];
    __shared__ double dk[kDcMax], zz[kDcMax];
    const size_t b = (size_t)n * mat + off;
    for (int i = threadIdx.x; i < K; i += kRootThreads) {
        dk[i] = st.dk[b + i];
        const double z = st.zk[b + i];
        zz[i] = rho * z * z;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int j = j0 + wv; j < min(K, j0 + kRootsPerWg); j += kRootsPerWg) {
        int org;
        double tau;
        dc_root(j, K, dk, zz, org, tau);
        if (lane == 0) {
            st.lam[b + j] = dk[org] + tau;
            st.org[b + j] = org;
            st.tau[b + j] = tau;
        }
    }
}

// 512 threads: the tile's columns eight per wave, and the zhat products
// (K of K terms each) over twice as many lanes (1024 threads would spill)
constexpr int kVecThreads = 512;
__global__ __launch_bounds__(kVecThreads) void k_dc_vecs(const int2 *tiles, const int4 *nodes,
                                                         int nodeBase, int nNodes, int n,
                                                         double *Dall, double *Wall, DcState st)
{
    const int2 tl = tiles[blockIdx.x];
    const int mat = blockIdx.y;
    const int4 nd = nodes[nodeBase + tl.x];
    const int off = nd.x, m = nd.z;
    const int c0 = tl.y * kColsPerWg, c1 = min(m, c0 + kColsPerWg);
    if (c0 >= m) return;
    const int4 cn = st.cnt[(size_t)mat * nNodes + tl.x];
    __shared__ MergeLds L;
    __shared__ double lam[kDcMax], zh[kDcMax], rtau[kDcMax];
    __shared__ int rorg[kDcMax], colInfo[kDcMax];
    const size_t b = (size_t)n * mat + off;
    const int K = cn.x, Nd = cn.y, Nr = cn.z;
    if (threadIdx.x == 0) {
        L.K = K;
        L.Nd = Nd;
        L.Nr = Nr;
        L.rho = st.rho[(size_t)mat * nNodes + tl.x];
    }
    for (int i = threadIdx.x; i < m; i += kVecThreads) {
        if (i < K) {
            L.dk[i] = st.dk[b + i];
            L.zk[i] = st.zk[b + i];
            L.keep[i] = st.keep[b + i];
            lam[i] = st.lam[b + i];
            rorg[i] = st.org[b + i];
            rtau[i] = st.tau[b + i];
        }
        if (i < Nd) {
            L.dv[i] = st.dv[b + i];
            L.dp[i] = st.dp[b + i];
            L.defl[i] = st.defl[b + i];
        }
        if (i < Nr) {
            L.rotP[i] = st.rotP[b + i];
            L.rotQ[i] = st.rotQ[b + i];
            L.rotC[i] = st.rotC[b + i];
            L.rotS[i] = st.rotS[b + i];
        }
    }
    __syncthreads();
    for (int ii = threadIdx.x; ii < K; ii += kVecThreads)
        zh[ii] = dc_zhat(ii, K, L.dk, L.zk, L.rho, rorg, rtau);
    __syncthreads();
    dc_columns<kVecThreads>(n, off, m, c0, c1, L, lam, rorg, rtau, zh, colInfo,
                            Dall + (size_t)n * mat, Wall + (size_t)n * n * mat);
}

// ---- Merges larger than kDcMaxN (up to kDcBigMaxN): the same steps and the
// same arithmetic in the same order, with the merge state in the per-matrix
// global state arrays at the node's offset instead of LDS.  The sorted poles,
// their z and the deflation flags stay in (dynamic) LDS, where the sequential
// rotation scan reads them; the roots and the eigenvector columns stage
// their poles in LDS; zhat is formed once per merge by its own kernel.
struct MergeView {
    double *sd, *sz, *ds, *zs, *dk, *zk, *zz, *dv, *rotC, *rotS;
    int *order, *keep, *rowKeep, *defl, *scratch, *dp, *rotP, *rotQ;
    double red[16];
    int ired[16];
    int K, Nd, Nr, close, unsorted;
    double rho;
};

template <int NT>
__global__ __launch_bounds__(NT) void k_dc_prep_big(const int4 *nodes, int nodeBase, int nNodes,
                                                    int n, const double *Dall, const double *Eall,
                                                    const double *Qall, int *ordAll, DcState st)
{
    const int node = nodeBase + blockIdx.x, mat = blockIdx.y;
    const int4 nd = nodes[node];
    const int off = nd.x, m1 = nd.y, m = nd.z;
    extern __shared__ double big_lds[];
    __shared__ MergeView L;
    const size_t b = (size_t)n * mat + off;
    if (threadIdx.x == 0) {
        L.ds = big_lds;
        L.zs = big_lds + m;
        L.scratch = reinterpret_cast<int *>(big_lds + 2 * m);
        L.sd = st.sd + b;
        L.sz = st.sz + b;
        L.zz = st.zz + b;
        L.order = st.order + b;
        L.dk = st.dk + b;
        L.zk = st.zk + b;
        L.dv = st.dv + b;
        L.rotC = st.rotC + b;
        L.rotS = st.rotS + b;
        L.keep = st.keep + b;
        L.rowKeep = st.rowKeep + b;
        L.defl = st.defl + b;
        L.dp = st.dp + b;
        L.rotP = st.rotP + b;
        L.rotQ = st.rotQ + b;
    }
    __syncthreads();
    dc_prepare<NT>(n, off, m1, m, Dall + (size_t)n * mat, Eall + (size_t)n * mat,
                   Qall + (size_t)n * n * mat, ordAll + (size_t)2 * n * mat, L);
    if (threadIdx.x == 0) {
        st.cnt[(size_t)mat * nNodes + blockIdx.x] = make_int4(L.K, L.Nd, L.Nr, 0);
        st.rho[(size_t)mat * nNodes + blockIdx.x] = L.rho;
    }
}

__global__ __launch_bounds__(kRootThreads) void k_dc_roots_big(const int2 *tiles, const int4 *nodes,
                                                               int nodeBase, int nNodes, int n,
                                                               DcState st)
{
    const int2 tl = tiles[blockIdx.x];
    const int mat = blockIdx.y;
    const int4 nd = nodes[nodeBase + tl.x];
    const int off = nd.x;
    const int4 c = st.cnt[(size_t)mat * nNodes + tl.x];
    const int K = c.x;
    const int j0 = tl.y * kRootsPerWg;
    if (j0 >= K) return;
    const double rho = st.rho[(size_t)mat * nNodes + tl.x];
    extern __shared__ double big_lds[];
    double *dk = big_lds, *zz = big_lds + K;
    const size_t b = (size_t)n * mat + off;
    for (int i = threadIdx.x; i < K; i += kRootThreads) {
        dk[i] = st.dk[b + i];
        const double z = st.zk[b + i];
        zz[i] = rho * z * z;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int j = j0 + wv; j < min(K, j0 + kRootsPerWg); j += kRootsPerWg) {
        int org;
        double tau;
        dc_root<true>(j, K, dk, zz, org, tau);
        if (lane == 0) {
            st.lam[b + j] = dk[org] + tau;
            st.org[b + j] = org;
            st.tau[b + j] = tau;
        }
    }
}

// zhat of the kept poles, one per lane (tiles: the level's 64-column tiles)
__global__ __launch_bounds__(64) void k_dc_zhat_big(const int2 *tiles, const int4 *nodes,
                                                    int nodeBase, int nNodes, int n, DcState st)
{
    const int2 tl = tiles[blockIdx.x];
    const int mat = blockIdx.y;
    const int4 nd = nodes[nodeBase + tl.x];
    const size_t b = (size_t)n * mat + nd.x;
    const int K = st.cnt[(size_t)mat * nNodes + tl.x].x;
    const int ii = tl.y * kColsPerWg + (int)threadIdx.x;
    if (ii >= K) return;
    st.zh[b + ii] = dc_zhat(ii, K, st.dk + b, st.zk + b, st.rho[(size_t)mat * nNodes + tl.x],
                            st.org + b, st.tau + b);
}

// dc_columns for one 64-column tile of a large merge: each kept column's
// norm in a first pass over the poles, its entries in a second.
__global__ __launch_bounds__(256) void k_dc_vecs_big(const int2 *tiles, const int4 *nodes,
                                                     int nodeBase, int nNodes, int n, double *Dall,
                                                     double *Wall, DcState st)
{
    const int2 tl = tiles[blockIdx.x];
    const int mat = blockIdx.y;
    const int4 nd = nodes[nodeBase + tl.x];
    const int off = nd.x, m = nd.z;
    const int c0 = tl.y * kColsPerWg, c1 = min(m, c0 + kColsPerWg);
    if (c0 >= m) return;
    const int4 cn = st.cnt[(size_t)mat * nNodes + tl.x];
    const int K = cn.x, Nd = cn.y, Nr = cn.z;
    const size_t b = (size_t)n * mat + off;
    extern __shared__ double big_lds[];
    double *dkS = big_lds, *zhS = big_lds + K;
    __shared__ int colInfo[kColsPerWg];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int i = tid; i < K; i += 256) {
        dkS[i] = st.dk[b + i];
        zhS[i] = st.zh[b + i];
    }
    // the searches below give every column of the tile its entry; a column
    // they miss (non-finite input) stays a unit vector rather than reading
    // an index it was never given
    if (tid < kColsPerWg) colInfo[tid] = -1 - (c0 + tid);
    __syncthreads();
    const double *lam = st.lam + b, *rtau = st.tau + b, *dv = st.dv + b;
    const int *rorg = st.org + b, *dp = st.dp + b, *keep = st.keep + b, *defl = st.defl + b;
    double *D = Dall + (size_t)n * mat;
    double *W = Wall + (size_t)n * n * mat;
    auto Wat = [&](int row, int col) -> double & { return W[(size_t)(off + row) + (size_t)n * (off + col)]; };
    for (int j = tid; j < K; j += 256) {
        const double v = lam[j];
        int lo = 0, hi = Nd;
        while (lo < hi) {
            const int md = (lo + hi) >> 1;
            if (dv[md] < v) lo = md + 1; else hi = md;
        }
        const int s = j + lo;
        if (s >= c0 && s < c1) {
            colInfo[s - c0] = j;
            D[off + s] = v;
        }
    }
    for (int q = tid; q < Nd; q += 256) {
        const double v = dv[q];
        int lo = 0, hi = K;
        while (lo < hi) {
            const int md = (lo + hi) >> 1;
            if (lam[md] <= v) lo = md + 1; else hi = md;
        }
        const int s = q + lo;
        if (s >= c0 && s < c1) {
            colInfo[s - c0] = -1 - dp[q];
            D[off + s] = v;
        }
    }
    __syncthreads();
    for (int s = c0 + wv; s < c1; s += 4) {
        const int info = colInfo[s - c0];
        if (info >= 0) {
            const int j = info;
            const double dj = dkS[rorg[j]], tj = rtau[j];
            double part = 0.0;
            for (int ii = lane; ii < K; ii += 64) {
                const double u = zhS[ii] * rcp_nr((dkS[ii] - dj) - tj);
                part += u * u;
            }
            const double inv = 1.0 / sqrt(wave_sum(part));
            for (int ii = lane; ii < K; ii += 64)
                Wat(keep[ii], s) = zhS[ii] * rcp_nr((dkS[ii] - dj) - tj) * inv;
            for (int q = lane; q < Nd; q += 64) Wat(defl[q], s) = 0.0;
        } else {
            const int pd = -1 - info;
            for (int p = lane; p < m; p += 64) Wat(p, s) = p == pd ? 1.0 : 0.0;
        }
    }
    __syncthreads();
    if (Nr > 0) {
        const int *rotP = st.rotP + b, *rotQ = st.rotQ + b;
        const double *rotC = st.rotC + b, *rotS = st.rotS + b;
        for (int s = c0 + tid; s < c1; s += 256) {
            int carryRow = -1;
            double carry = 0.0;
            for (int q = Nr - 1; q >= 0; --q) {
                const int p = rotP[q], r = rotQ[q];
                const double c = rotC[q], sn = rotS[q];
                double y;
                if (r == carryRow) {
                    y = carry;
                } else {
                    if (carryRow >= 0) Wat(carryRow, s) = carry;
                    y = Wat(r, s);
                }
                const double x = Wat(p, s);
                Wat(r, s) = sn * x + c * y;
                carry = c * x - sn * y;
                carryRow = p;
            }
            Wat(carryRow, s) = carry;
        }
    }
}

// Qt(node) = Q(node) P W over node blocks; tiles: x node, y row tile, z col
// tile.  Before the merge Q(node) = diag(Q1, Q2) in natural order, so a row
// of child c takes k over that child's columns only: a tile's k range is
// the children its rows span (half the products of the full m x m x m
// product are exact zeros).  Row j of P W is W(inv[j], :), inv[j] being the
// sorted position of pole j.  The next k tile is loaded into registers
// while the current one is multiplied from LDS.
__global__ __launch_bounds__(256) void k_dc_gemm(const int4 *tiles, const int4 *nodes, int n,
                                                 const double *Qall, const double *Wall,
                                                 const int *ordAll, double *Qtall)
{
    const int4 tl = tiles[blockIdx.x];
    const int4 nd = nodes[tl.x];
    const int off = nd.x, m1 = nd.y, m = nd.z;
    const int mat = blockIdx.y;
    const double *Q = Qall + (size_t)n * n * mat;
    const double *W = Wall + (size_t)n * n * mat;
    const int *inv = ordAll + (size_t)2 * n * mat + n;
    double *Qt = Qtall + (size_t)n * n * mat;
    const int r0 = tl.y * kTile, c0 = tl.z * kTile;
    const int rLast = min(r0 + kTile, m) - 1;
    const int kb = r0 < m1 ? 0 : m1, ke = rLast < m1 ? m1 : m;
    __shared__ double As[2][16][kTile + 1];
    __shared__ double Bs[2][16][kTile + 1];
    const int tid = threadIdx.x;
    const int tr = tid % 16, tc = tid / 16;   // 4 x 4 outputs each: rows tr + 16 a, cols tc + 16 b
    double acc[4][4] = {};
    double ra[4], rb[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u;
            const int rr = e % kTile, kk = e / kTile;
            const int r = r0 + rr, k = k0 + kk;
            ra[u] = (r < m && k < ke) ? Q[(size_t)(off + r) + (size_t)n * (off + k)] : 0.0;
            const int kq = e % 16, cc = e / 16;
            const int k2 = k0 + kq, c = c0 + cc;
            rb[u] = (k2 < ke && c < m) ? W[(size_t)(off + inv[off + k2]) + (size_t)n * (off + c)] : 0.0;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u;
            As[buf][e / kTile][e % kTile] = ra[u];
            Bs[buf][e % 16][e / 16] = rb[u];
        }
    };
    int buf = 0;
    load(kb);
    store(0);
    __syncthreads();
    for (int k0 = kb; k0 < ke; k0 += 16) {
        const bool more = k0 + 16 < ke;
        if (more) load(k0 + 16);
#pragma unroll 4
        for (int kk = 0; kk < 16; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = As[buf][kk][tr + 16 * a];
#pragma unroll
            for (int bb = 0; bb < 4; ++bb) bv[bb] = Bs[buf][kk][tc + 16 * bb];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int bb = 0; bb < 4; ++bb) acc[a][bb] = fma(av[a], bv[bb], acc[a][bb]);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
            const int r = r0 + tr + 16 * a, c = c0 + tc + 16 * bb;
            if (r < m && c < m) Qt[(size_t)(off + r) + (size_t)n * (off + c)] = acc[a][bb];
        }
}

__global__ void k_dc_copy(const int4 *tiles, const int4 *nodes, int n, const double *Qtall,
                          double *Qall)
{
    const int4 tl = tiles[blockIdx.x];
    const int4 nd = nodes[tl.x];
    const int off = nd.x, m = nd.z;
    const int mat = blockIdx.y;
    const double *Qt = Qtall + (size_t)n * n * mat;
    double *Q = Qall + (size_t)n * n * mat;
    const int r0 = tl.y * kTile, c0 = tl.z * kTile;
    for (int e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
        const int r = r0 + e % kTile, c = c0 + e / kTile;
        if (r < m && c < m) {
            const size_t at = (size_t)(off + r) + (size_t)n * (off + c);
            Q[at] = Qt[at];
        }
    }
}

}  // namespace

bool dc_plan(int n, DcPlan &p)
{
    p = DcPlan{};
    // merges up to kDcMaxN keep their poles, z and rotations in LDS arrays of
    // kDcMax; larger ones (levels with maxM > kDcMaxN) keep them in global
    // memory, with the sorted poles, z and flags in up to 20 B per entry of LDS
    if (n < 1 || n > kDcBigMaxN) return false;
    p.n = n;
    if (n < 2) return true;
    // merges by height above the leaves: the blocks of at most kDcLeaf
    // entries that halving reaches first
    std::vector<std::vector<int4>> byH;
    struct Rec {
        static int go(int off, int m, std::vector<std::vector<int4>> &byH, std::vector<int2> &lv)
        {
            if (m <= kDcLeaf) {
                lv.push_back(make_int2(off, m));
                return 0;
            }
            const int m1 = m / 2;
            const int h = 1 + std::max(go(off, m1, byH, lv), go(off + m1, m - m1, byH, lv));
            if ((int)byH.size() < h) byH.resize(h);
            byH[h - 1].push_back(make_int4(off, m1, m, 0));
            return h;
        }
    };
    Rec::go(0, n, byH, p.leaves);
    for (auto &lv : byH) {
        DcLevel L;
        L.nodeBase = (int)p.nodes.size();
        L.nNodes = (int)lv.size();
        L.tileBase = (int)p.tiles.size();
        L.rootBase = (int)p.rootTiles.size();
        L.colBase = (int)p.colTiles.size();
        for (int q = 0; q < L.nNodes; ++q) {
            const int m = lv[q].z, nt = (m + kTile - 1) / kTile;
            for (int a = 0; a < nt; ++a)
                for (int b2 = 0; b2 < nt; ++b2) p.tiles.push_back(make_int4(L.nodeBase + q, a, b2, 0));
            for (int a = 0; a < (m + kRootsPerWg - 1) / kRootsPerWg; ++a) p.rootTiles.push_back(make_int2(q, a));
            for (int a = 0; a < (m + kColsPerWg - 1) / kColsPerWg; ++a) p.colTiles.push_back(make_int2(q, a));
            p.nodes.push_back(lv[q]);
            L.maxM = std::max(L.maxM, m);
        }
        L.nTiles = (int)p.tiles.size() - L.tileBase;
        L.nRoot = (int)p.rootTiles.size() - L.rootBase;
        L.nCol = (int)p.colTiles.size() - L.colBase;
        p.levels.push_back(L);
    }
    return true;
}

size_t dc_state_doubles(int n, int nb) { return (size_t)11 * n * nb; }
size_t dc_state_ints(int n, int nb) { return (size_t)8 * n * nb; }

bool launch_dc(const DcPlan &p, const DcDev &dv, int nb, double *D, const double *E, double *Q,
               double *Qt, double *W, int *ord, hipStream_t s)
{
    const int n = p.n;
    if (n < 1 || n > kDcBigMaxN) return false;
    const size_t nn = (size_t)n * nb;
    DcState st;
    st.dk = dv.stD;
    st.zk = st.dk + nn;
    st.dv = st.zk + nn;
    st.lam = st.dv + nn;
    st.tau = st.lam + nn;
    st.rotC = st.tau + nn;
    st.rotS = st.rotC + nn;
    st.keep = dv.stI;
    st.dp = st.keep + nn;
    st.defl = st.dp + nn;
    st.org = st.defl + nn;
    st.rowKeep = st.org + nn;
    st.rotP = st.rowKeep + nn;
    st.rotQ = st.rotP + nn;
    st.zh = st.rotS + nn;
    st.sd = st.zh + nn;
    st.sz = st.sd + nn;
    st.zz = st.sz + nn;
    st.order = st.rotQ + nn;
    st.cnt = dv.cnt;
    st.rho = dv.rho;
    hipLaunchKernelGGL(k_dc_init, dim3(64, nb), dim3(256), 0, s, D, E, Q, n);
    if (!p.leaves.empty())
        hipLaunchKernelGGL(k_dc_leaf, dim3((unsigned)p.leaves.size(), nb), dim3(64), 0, s,
                           dv.leaves, n, D, E, Q);
    for (const DcLevel &L : p.levels) {
        if (L.maxM > kDcMaxN) {
            const size_t mm = (size_t)L.maxM;
            hipLaunchKernelGGL(k_dc_prep_big<kDcBig>, dim3(L.nNodes, nb), dim3(kDcBig), mm * 20, s,
                               dv.nodes, L.nodeBase, L.nNodes, n, D, E, Q, ord, st);
            hipLaunchKernelGGL(k_dc_roots_big, dim3(L.nRoot, nb), dim3(kRootThreads), mm * 16, s,
                               dv.rootTiles + L.rootBase, dv.nodes, L.nodeBase, L.nNodes, n, st);
            hipLaunchKernelGGL(k_dc_zhat_big, dim3(L.nCol, nb), dim3(64), 0, s,
                               dv.colTiles + L.colBase, dv.nodes, L.nodeBase, L.nNodes, n, st);
            hipLaunchKernelGGL(k_dc_vecs_big, dim3(L.nCol, nb), dim3(256), mm * 16, s,
                               dv.colTiles + L.colBase, dv.nodes, L.nodeBase, L.nNodes, n, D, W, st);
        } else if (L.maxM <= kDcSmallM) {
            hipLaunchKernelGGL(k_dc_merge<kDcSmall>, dim3(L.nNodes, nb), dim3(kDcSmall), 0, s,
                               dv.nodes + L.nodeBase, n, D, E, Q, W, ord);
        } else {
            hipLaunchKernelGGL(k_dc_prep<kDcBig>, dim3(L.nNodes, nb), dim3(kDcBig), 0, s, dv.nodes,
                               L.nodeBase, L.nNodes, n, D, E, Q, ord, st);
            hipLaunchKernelGGL(k_dc_roots, dim3(L.nRoot, nb), dim3(kRootThreads), 0, s, dv.rootTiles + L.rootBase,
                               dv.nodes, L.nodeBase, L.nNodes, n, st);
            hipLaunchKernelGGL(k_dc_vecs, dim3(L.nCol, nb), dim3(kVecThreads), 0, s, dv.colTiles + L.colBase,
                               dv.nodes, L.nodeBase, L.nNodes, n, D, W, st);
        }
        hipLaunchKernelGGL(k_dc_gemm, dim3(L.nTiles, nb), dim3(256), 0, s, dv.tiles + L.tileBase,
                           dv.nodes, n, Q, W, ord, Qt);
        hipLaunchKernelGGL(k_dc_copy, dim3(L.nTiles, nb), dim3(256), 0, s, dv.tiles + L.tileBase,
                           dv.nodes, n, Qt, Q);
    }
    return true;
}

}  // namespace olcao
