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

namespace olcao {

namespace {

constexpr double kEps = 1.1102230246251565e-16;   // 2^-53
constexpr int kDcMax = kDcMaxN;
constexpr int kSecularIters = 64;
constexpr int kTile = 64;
constexpr int kDcBig = 1024;     // threads of a merge of more than kDcSmallM
constexpr int kDcSmall = 256;
constexpr int kDcSmallM = 128;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
// Four wave sums in seven exchanges: halves, then quarters, hold one sum each.
__device__ __forceinline__ void wave_sum4(double &a, double &b, double &c, double &d)
{
    const int lane = threadIdx.x & 63;
    const bool lo32 = lane < 32, lo16 = (lane & 16) == 0;
    double x1 = lo32 ? a : c, x2 = lo32 ? b : d;
    const double y1 = lo32 ? c : a, y2 = lo32 ? d : b;
    x1 += __shfl_xor(y1, 32, 64);
    x2 += __shfl_xor(y2, 32, 64);
    double v = lo16 ? x1 : x2;
    v += __shfl_xor(lo16 ? x2 : x1, 16, 64);
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    a = __shfl(v, 0, 64);
    b = __shfl(v, 16, 64);
    c = __shfl(v, 32, 64);
    d = __shfl(v, 48, 64);
}

// Block-wide reduction through LDS (all threads call it).
template <bool MAX, int NT>
__device__ double block_reduce(double v, double *red)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
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
    constexpr int kSeg = (kDcMax + NT - 1) / NT;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int b0 = t * kSeg, b1 = min(m, b0 + kSeg);
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

template <int NT>
__global__ __launch_bounds__(NT) void k_dc_merge(const int4 *nodes, int n, double *Dall,
                                                 const double *Eall, const double *Qall,
                                                 double *Wall, double *Sall, int *ordAll)
{
    constexpr int NW = NT / 64;
    const int4 nd = nodes[blockIdx.x];
    const int off = nd.x, m1 = nd.y, m = nd.z;
    const int mat = blockIdx.y;
    double *D = Dall + (size_t)n * mat;
    const double *E = Eall + (size_t)n * mat;
    const double *Q = Qall + (size_t)n * n * mat;
    double *W = Wall + (size_t)n * n * mat;
    double *S = Sall + (size_t)n * n * mat;
    int *ord = ordAll + (size_t)n * mat;
    auto Wat = [&](int row, int col) -> double & { return W[(size_t)(off + row) + (size_t)n * (off + col)]; };
    auto Sat = [&](int row, int col) -> double & { return S[(size_t)(off + row) + (size_t)n * (off + col)]; };

    __shared__ double sd[kDcMax], sz[kDcMax], ds[kDcMax], zs[kDcMax];
    __shared__ double dk[kDcMax], zk[kDcMax], lam[kDcMax], zh[kDcMax];
    __shared__ int order[kDcMax], keep[kDcMax], rowKeep[kDcMax], defl[kDcMax], colInfo[kDcMax];
    __shared__ int rotP[kDcMax], rotQ[kDcMax];
    __shared__ double rotC[kDcMax], rotS[kDcMax];
    __shared__ double red[NW];
    __shared__ int ired[NW];
    __shared__ int sK, sNd, sNr, sClose, sUnsorted;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const double beta = E[off + m1 - 1];
    const double sgn = beta < 0.0 ? -1.0 : 1.0;

    // children's eigenvalues and z = (last row of Q1, sign(beta) first row of Q2)
    double nzp = 0.0;
    for (int i = tid; i < m; i += NT) {
        sd[i] = D[off + i];
        const double v = i < m1 ? Q[(size_t)(off + m1 - 1) + (size_t)n * (off + i)]
                                : sgn * Q[(size_t)(off + m1) + (size_t)n * (off + i)];
        sz[i] = v;
        nzp += v * v;
    }
    const double nz2 = block_reduce<false, NT>(nzp, red);
    const double nz = sqrt(nz2);
    const double rho = fabs(beta) * nz2;
    // merge order of the two sorted children (ties: first child first)
    for (int i = tid; i < m; i += NT) {
        const double v = sd[i];
        int lo, hi;
        if (i < m1) {   // count second-child entries < v
            lo = m1;
            hi = m;
            while (lo < hi) {
                const int md = (lo + hi) >> 1;
                if (sd[md] < v) lo = md + 1; else hi = md;
            }
            order[i + (lo - m1)] = i;
        } else {        // count first-child entries <= v
            lo = 0;
            hi = m1;
            while (lo < hi) {
                const int md = (lo + hi) >> 1;
                if (sd[md] <= v) lo = md + 1; else hi = md;
            }
            order[(i - m1) + lo] = i;
        }
    }
    __syncthreads();
    double dmx = 0.0, zmx = 0.0;
    for (int p = tid; p < m; p += NT) {
        ds[p] = sd[order[p]];
        zs[p] = nz > 0.0 ? sz[order[p]] / nz : 0.0;
        dmx = fmax(dmx, fabs(ds[p]));
        zmx = fmax(zmx, fabs(zs[p]));
    }
    dmx = block_reduce<true, NT>(dmx, red);
    zmx = block_reduce<true, NT>(zmx, red);
    const double tol = 8.0 * kEps * fmax(dmx, zmx);
    const bool allDefl = !(rho * zmx > tol);

    // Deflation (dlaed2's rules) in the sorted order.  Parallel pass: small
    // rho z(p) entries, and whether any two consecutive remaining poles are
    // close enough to need a rotation; only then the sequential scan.
    int *big = colInfo, *pre = rowKeep;   // scratch until their own use below
    for (int p = tid; p < m; p += NT) big[p] = !allDefl && rho * fabs(zs[p]) > tol;
    if (tid == 0) sClose = 0;
    block_scan<NT>(big, pre, m, ired);
    const int nBig = m > 0 ? pre[m - 1] + big[m - 1] : 0;
    for (int p = tid; p < m; p += NT)
        if (big[p]) keep[pre[p]] = p;
    __syncthreads();
    for (int q = tid + 1; q < nBig; q += NT) {
        const int pj = keep[q - 1], nj = keep[q];
        double s = zs[pj], c = zs[nj];
        const double tau = sqrt(c * c + s * s);
        c /= tau;
        s = -s / tau;
        if (fabs((ds[nj] - ds[pj]) * c * s) <= tol) sClose = 1;
    }
    __syncthreads();
    if (!sClose) {
        for (int p = tid; p < m; p += NT)
            if (!big[p]) defl[p - pre[p]] = p;
        if (tid == 0) {
            sK = nBig;
            sNd = m - nBig;
            sNr = 0;
        }
    } else if (tid == 0) {
        int K = 0, nd = 0, nr = 0;
        int pj = -1;
        for (int nj = 0; nj < m; ++nj) {
            if (!big[nj]) {
                defl[nd++] = nj;
                continue;
            }
            if (pj < 0) {
                pj = nj;
                continue;
            }
            double s = zs[pj], c = zs[nj];
            const double tau = sqrt(c * c + s * s);
            const double t = ds[nj] - ds[pj];
            c /= tau;
            s = -s / tau;
            if (fabs(t * c * s) <= tol) {
                zs[nj] = tau;
                zs[pj] = 0.0;
                rotP[nr] = pj;
                rotQ[nr] = nj;
                rotC[nr] = c;
                rotS[nr] = s;
                ++nr;
                const double t2 = ds[pj] * c * c + ds[nj] * s * s;
                ds[nj] = ds[pj] * s * s + ds[nj] * c * c;
                ds[pj] = t2;
                defl[nd++] = pj;
            } else {
                keep[K++] = pj;
            }
            pj = nj;
        }
        if (pj >= 0) keep[K++] = pj;
        sK = K;
        sNd = nd;
        sNr = nr;
    }
    __syncthreads();
    const int K = sK, Nd = sNd, Nr = sNr;
    for (int p = tid; p < m; p += NT) rowKeep[p] = -1;
    if (tid == 0) sUnsorted = 0;
    __syncthreads();
    for (int ii = tid; ii < K; ii += NT) {
        dk[ii] = ds[keep[ii]];
        zk[ii] = zs[keep[ii]];
        rowKeep[keep[ii]] = ii;
    }
    for (int q = tid; q + 1 < Nd; q += NT)
        if (ds[defl[q]] > ds[defl[q + 1]]) sUnsorted = 1;
    __syncthreads();

    // secular equation, one wave per root: Delta(i, j) = dk(i) - lambda_j into S
    for (int j = wv; j < K; j += NW) {
        int org;
        double a, b;
        if (j < K - 1) {
            const double mid = 0.5 * (dk[j] + dk[j + 1]);
            double part = 0.0;
            for (int ii = lane; ii < K; ii += 64) part += rho * zk[ii] * zk[ii] / (dk[ii] - mid);
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
            for (int ii = lane; ii < K; ii += 64) part += rho * zk[ii] * zk[ii];
            org = j;
            a = 0.0;
            b = wave_sum(part);
        }
        const double dorg = dk[org];
        const double dj = dk[j] - dorg;
        const double dj1 = j < K - 1 ? dk[j + 1] - dorg : 0.0;
        double tau = 0.5 * (a + b);
        for (int it = 0; it < kSecularIters; ++it) {
            // left poles (ii <= j) give negative terms, right poles positive
            double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0;
            for (int ii = lane; ii < K; ii += 64) {
                const double r = 1.0 / ((dk[ii] - dorg) - tau);
                const double t = rho * zk[ii] * zk[ii] * r;
                if (ii <= j) {
                    psi += t;
                    dpsi += t * r;
                } else {
                    phi += t;
                    dphi += t * r;
                }
            }
            wave_sum4(psi, dpsi, phi, dphi);
            const double g = 1.0 + psi + phi;
            if (g < 0.0) a = tau; else b = tau;
            if (b - a <= 4.0 * kEps * fmax(fabs(a), fabs(b)) ||
                fabs(g) <= 8.0 * kEps * (1.0 + phi - psi))
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
        for (int ii = lane; ii < K; ii += 64) Sat(ii, j) = (dk[ii] - dorg) - tau;
        if (lane == 0) lam[j] = dorg + tau;
    }
    __syncthreads();

    // Gu-Eisenstat: zhat(i)^2 = (lambda_i - d_i)/rho prod_{j!=i} (lambda_j - d_i)/(d_j - d_i)
    for (int ii = tid; ii < K; ii += NT) {
        double prod = -Sat(ii, ii) / rho;
        for (int j = 0; j < K; ++j)
            if (j != ii) prod *= -Sat(ii, j) / (dk[j] - dk[ii]);
        zh[ii] = copysign(sqrt(fmax(prod, 0.0)), zk[ii]);
    }
    // deflated values ascending into sd (values) and sz (positions): as
    // scanned unless a rotation moved one out of order
    for (int q = tid; q < Nd; q += NT) {
        const double v = ds[defl[q]];
        int rank = q;
        if (sUnsorted) {
            rank = 0;
            for (int q2 = 0; q2 < Nd; ++q2) {
                const double u = ds[defl[q2]];
                rank += (u < v) || (u == v && q2 < q);
            }
        }
        sd[rank] = v;
        sz[rank] = (double)defl[q];
    }
    __syncthreads();
    // final column of every eigenpair: roots before equal deflated values
    for (int j = tid; j < K; j += NT) {
        const double v = lam[j];
        int lo = 0, hi = Nd;
        while (lo < hi) {
            const int md = (lo + hi) >> 1;
            if (sd[md] < v) lo = md + 1; else hi = md;
        }
        colInfo[j + lo] = j;
        D[off + j + lo] = v;
    }
    for (int q = tid; q < Nd; q += NT) {
        const double v = sd[q];
        int lo = 0, hi = K;
        while (lo < hi) {
            const int md = (lo + hi) >> 1;
            if (lam[md] <= v) lo = md + 1; else hi = md;
        }
        colInfo[q + lo] = -1 - (int)sz[q];
        D[off + q + lo] = v;
    }
    __syncthreads();
    // W, one wave per column: kept rows u(i) = zhat(i) / Delta(i, j),
    // normalised; deflated rows zero, or the unit entry of a deflated column
    for (int s = wv; s < m; s += NW) {
        const int info = colInfo[s];
        if (info >= 0) {
            const int j = info;
            double u[kDcMax / 64];
            double part = 0.0;
#pragma unroll
            for (int q = 0; q < kDcMax / 64; ++q) {
                const int ii = lane + 64 * q;
                u[q] = ii < K ? zh[ii] / Sat(ii, j) : 0.0;
                part += u[q] * u[q];
            }
            const double inv = 1.0 / sqrt(wave_sum(part));
#pragma unroll
            for (int q = 0; q < kDcMax / 64; ++q) {
                const int ii = lane + 64 * q;
                if (ii < K) Wat(keep[ii], s) = u[q] * inv;
            }
            for (int q = lane; q < Nd; q += 64) Wat(defl[q], s) = 0.0;
        } else {
            const int pd = -1 - info;
            for (int p = lane; p < m; p += 64) Wat(p, s) = p == pd ? 1.0 : 0.0;
        }
    }
    __syncthreads();
    // deflation rotations on W's rows, last first; a chain (p1,p2),(p2,p3),...
    // keeps its running row in a register
    if (Nr > 0) {
        for (int s = tid; s < m; s += NT) {
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
    for (int p = tid; p < m; p += NT) ord[off + p] = off + order[p];
}

// Qt(node) = Q(:, ord) W over node blocks; tiles: x node, y row tile, z col tile
__global__ __launch_bounds__(256) void k_dc_gemm(const int4 *tiles, const int4 *nodes, int n,
                                                 const double *Qall, const double *Wall,
                                                 const int *ordAll, double *Qtall)
{
    const int4 tl = tiles[blockIdx.x];
    const int4 nd = nodes[tl.x];
    const int off = nd.x, m = nd.z;
    const int mat = blockIdx.y;
    const double *Q = Qall + (size_t)n * n * mat;
    const double *W = Wall + (size_t)n * n * mat;
    const int *ord = ordAll + (size_t)n * mat;
    double *Qt = Qtall + (size_t)n * n * mat;
    const int r0 = tl.y * kTile, c0 = tl.z * kTile;
    __shared__ double As[16][kTile + 1];
    __shared__ double Bs[16][kTile + 1];
    const int tid = threadIdx.x;
    const int tr = tid % 16, tc = tid / 16;   // 4 x 4 outputs each: rows tr + 16 a, cols tc + 16 b
    double acc[4][4] = {};
    for (int k0 = 0; k0 < m; k0 += 16) {
        for (int e = tid; e < 16 * kTile; e += 256) {
            const int rr = e % kTile, kk = e / kTile;
            const int r = r0 + rr, k = k0 + kk;
            As[kk][rr] = (r < m && k < m) ? Q[(size_t)(off + r) + (size_t)n * ord[off + k]] : 0.0;
            const int kb = e % 16, cc = e / 16;
            const int k2 = k0 + kb, c = c0 + cc;
            Bs[kb][cc] = (k2 < m && c < m) ? W[(size_t)(off + k2) + (size_t)n * (off + c)] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = As[kk][tr + 16 * a];
#pragma unroll
            for (int bb = 0; bb < 4; ++bb) bv[bb] = Bs[kk][tc + 16 * bb];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int bb = 0; bb < 4; ++bb) acc[a][bb] = fma(av[a], bv[bb], acc[a][bb]);
        }
        __syncthreads();
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

void dc_plan(int n, DcPlan &p)
{
    p = DcPlan{};
    p.n = n;
    if (n < 2) return;
    // merges by height above the 1 x 1 leaves
    std::vector<std::vector<int4>> byH;
    struct Rec {
        static int go(int off, int m, std::vector<std::vector<int4>> &byH)
        {
            if (m == 1) return 0;
            const int m1 = m / 2;
            const int h = 1 + std::max(go(off, m1, byH), go(off + m1, m - m1, byH));
            if ((int)byH.size() < h) byH.resize(h);
            byH[h - 1].push_back(make_int4(off, m1, m, 0));
            return h;
        }
    };
    Rec::go(0, n, byH);
    for (auto &lv : byH) {
        DcLevel L;
        L.nodeBase = (int)p.nodes.size();
        L.nNodes = (int)lv.size();
        L.tileBase = (int)p.tiles.size();
        for (int q = 0; q < L.nNodes; ++q) {
            const int m = lv[q].z, nt = (m + kTile - 1) / kTile;
            for (int a = 0; a < nt; ++a)
                for (int b2 = 0; b2 < nt; ++b2) p.tiles.push_back(make_int4(L.nodeBase + q, a, b2, 0));
            p.nodes.push_back(lv[q]);
            L.maxM = std::max(L.maxM, m);
        }
        L.nTiles = (int)p.tiles.size() - L.tileBase;
        p.levels.push_back(L);
    }
}

void launch_dc(const DcPlan &p, const int4 *dNodes, const int4 *dTiles, int nb, double *D,
               const double *E, double *Q, double *Qt, double *W, double *S, int *ord,
               hipStream_t s)
{
    const int n = p.n;
    hipLaunchKernelGGL(k_dc_init, dim3(64, nb), dim3(256), 0, s, D, E, Q, n);
    for (const DcLevel &L : p.levels) {
        if (L.maxM > kDcSmallM)
            hipLaunchKernelGGL(k_dc_merge<kDcBig>, dim3(L.nNodes, nb), dim3(kDcBig), 0, s,
                               dNodes + L.nodeBase, n, D, E, Q, W, S, ord);
        else
            hipLaunchKernelGGL(k_dc_merge<kDcSmall>, dim3(L.nNodes, nb), dim3(kDcSmall), 0, s,
                               dNodes + L.nodeBase, n, D, E, Q, W, S, ord);
        hipLaunchKernelGGL(k_dc_gemm, dim3(L.nTiles, nb), dim3(256), 0, s, dTiles + L.tileBase,
                           dNodes, n, Q, W, ord, Qt);
        hipLaunchKernelGGL(k_dc_copy, dim3(L.nTiles, nb), dim3(256), 0, s, dTiles + L.tileBase,
                           dNodes, n, Qt, Q);
    }
}

}  // namespace olcao
