// Batched generalised Hermitian eigensolve for small matrices, one
// workgroup per k-point with both working matrices in LDS: what
// zhegv/dsygv(itype 1, 'V', 'U') compute (interfaces.F90:73-200):
//   S = U^H U                        Cholesky, row by row
//   C = U^-H H U^-1                  forward substitution on rows, then on columns
//   C = V diag(w) V^H                cyclic Jacobi, round-robin parallel ordering
//   sort w ascending, X = U^-1 V     back substitution
// On exit A holds the eigenvectors (columns, ascending eigenvalues), B holds
// U, w the eigenvalues; info[k] = n + j + 1 flags a B(k) that is not
// positive definite at column j (zhegv's convention), info[k] = -1 a Jacobi
// iteration that did not converge.
//
// The Jacobi rotation of the pair (p, q) with a = C(p,p), d = C(q,q),
// b = C(p,q) = |b| e^{i phi} is G = [[c, s e^{i phi}], [-s e^{-i phi}, c]],
// t = sign(z) / (|z| + sqrt(1 + z^2)), z = (d - a) / (2|b|), c = 1/sqrt(1+t^2),
// s = t c; C <- G^H C G zeroes C(p,q) and V <- V G accumulates the vectors.
// All n/2 disjoint pairs of a round are applied together: every 2x2 block
// (pair i rows, pair j columns) of C becomes G_i^H B G_j in one pass, and
// V <- V G in the same pass, so a round costs two barriers.  Matrices are
// column-major in LDS with leading dimension n + 1.
#include <hip/hip_runtime.h>

#include "numtraits.hpp"
#include "olcao_internal.h"

namespace olcao {

namespace {

constexpr int kSolveThreads = 1024;
constexpr int kMaxSweeps = 40;

// round-robin (circle method) pair i of round r among m players (m even)
__device__ __forceinline__ void rr_pair(int r, int i, int m, int &p, int &q)
{
    if (i == 0) {
        p = m - 1;
        q = r;
    } else {
        p = (r + i) % (m - 1);
        q = (r - i + (m - 1)) % (m - 1);
    }
}

}  // namespace

template <bool CPLX>
__global__ __launch_bounds__(kSolveThreads) void k_small_hegv(double *Aall, double *Ball,
                                                              double *wall, int *info, int n)
{
    using N = Num<CPLX>;
    using T = typename N::T;
    extern __shared__ double lds[];
    const int ld = n + 1;
    T *M = reinterpret_cast<T *>(lds);          // [n][ld]: U, then V
    T *C = M + (size_t)n * ld;                  // [n][ld]: H, Y, C, then U again
    double *rc = reinterpret_cast<double *>(C + (size_t)n * ld);   // [n/2+1] c
    double *rs = rc + (n / 2 + 1);                                 // s
    T *rph = reinterpret_cast<T *>(rs + (n / 2 + 1));              // e^{i phi}
    double *red = reinterpret_cast<double *>(rph + (n / 2 + 1));   // reductions / eigenvalues
    int *perm = reinterpret_cast<int *>(red + n + kSolveThreads / 64 + 1);
    __shared__ int flag;
    const int tid = threadIdx.x;
    const int k = blockIdx.x;
    const long long per = (long long)n * n;
    double *A = Aall + (CPLX ? 2 : 1) * per * k;
    double *B = Ball + (CPLX ? 2 : 1) * per * k;
    double *w = wall + (long long)n * k;

    // ---- Cholesky of S (upper triangle of B) into M ----
    for (int e = tid; e < n * n; e += kSolveThreads) {
        const int r = e % n, c = e / n;
        M[r + c * ld] = r <= c ? N::load(B, e) : N::zero();
    }
    if (tid == 0) flag = 0;
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        // U(j,j) = sqrt(S(j,j) - sum_{k<j} |U(k,j)|^2); U(j,i) for i > j
        if (tid < 64) {
            double part = 0.0;
            for (int q = tid; q < j; q += 64) part += N::abs2(M[q + j * ld]);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
            if (tid == 0) {
                double d = N::re(M[j + j * ld]) - part;
                if (!(d > 0.0)) {
                    flag = j + 1;
                    d = 1.0;
                }
                M[j + j * ld] = N::mk(sqrt(d), 0.0);
            }
        }
        __syncthreads();
        if (flag) break;
        const double inv = 1.0 / N::re(M[j + j * ld]);
        for (int i = j + 1 + tid; i < n; i += kSolveThreads) {
            T v = M[j + i * ld];
            for (int q = 0; q < j; ++q) v = N::sub(v, N::cmul(M[q + j * ld], M[q + i * ld]));
            M[j + i * ld] = N::scale(v, inv);
        }
        __syncthreads();
    }
    if (flag) {
        if (tid == 0) info[k] = n + flag;
        return;
    }
    // U to global (the caller's B) for reference and the back transform
    for (int e = tid; e < n * n; e += kSolveThreads) {
        const int r = e % n, c = e / n;
        N::store(B, e, M[r + c * ld]);
    }
    // ---- H (upper) as a full Hermitian matrix into C ----
    for (int e = tid; e < n * n; e += kSolveThreads) {
        const int r = e % n, c = e / n;
        C[r + c * ld] = r <= c ? N::load(A, e) : N::conj(N::load(A, (long long)c + (long long)r * n));
    }
    __syncthreads();
    // Y = U^-H H: row j of Y = (row j of H - sum_{q<j} conj(U(q,j)) Y(q,:)) / U(j,j)
    for (int j = 0; j < n; ++j) {
        const double inv = 1.0 / N::re(M[j + j * ld]);
        for (int c = tid; c < n; c += kSolveThreads) {
            T v = C[j + c * ld];
            for (int q = 0; q < j; ++q) v = N::sub(v, N::cmul(M[q + j * ld], C[q + c * ld]));
            C[j + c * ld] = N::scale(v, inv);
        }
        __syncthreads();
    }
    // C = Y U^-1: column j = (Y(:,j) - sum_{q<j} C(:,q) U(q,j)) / U(j,j)
    for (int j = 0; j < n; ++j) {
        const double inv = 1.0 / N::re(M[j + j * ld]);
        for (int r = tid; r < n; r += kSolveThreads) {
            T v = C[r + j * ld];
            for (int q = 0; q < j; ++q) v = N::sub(v, N::mul(C[r + q * ld], M[q + j * ld]));
            C[r + j * ld] = N::scale(v, inv);
        }
        __syncthreads();
    }
    // exact Hermitian symmetry from the upper triangle; V = I in M
    for (int e = tid; e < n * n; e += kSolveThreads) {
        const int r = e % n, c = e / n;
        if (r > c) C[r + c * ld] = N::conj(C[c + r * ld]);
        if (r == c) C[r + c * ld] = N::mk(N::re(C[r + c * ld]), 0.0);
    }
    __syncthreads();
    for (int e = tid; e < n * n; e += kSolveThreads) {
        const int r = e % n, c = e / n;
        M[r + c * ld] = N::mk(r == c ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();

    // ---- cyclic Jacobi ----
    const int m = n + (n & 1);   // players; index n is a bye when n is odd
    const int np = m / 2;
    __shared__ double normd;
    int sweep = 0;
    for (; sweep < kMaxSweeps; ++sweep) {
        if (tid == 0) {
            flag = 0;
            double mx = 0.0;
            for (int i = 0; i < n; ++i) mx = fmax(mx, fabs(N::re(C[i + i * ld])));
            normd = mx;
        }
        __syncthreads();
        for (int r = 0; r < m - 1; ++r) {
            // rotation parameters of the round's pairs
            for (int i = tid; i < np; i += kSolveThreads) {
                int p, q;
                rr_pair(r, i, m, p, q);
                double c = 1.0, s = 0.0;
                T ph = N::mk(1.0, 0.0);
                if (p < n && q < n) {
                    const T b = C[p + q * ld];
                    const double ab = sqrt(N::abs2(b));
                    const double a = N::re(C[p + p * ld]), d = N::re(C[q + q * ld]);
                    // rotate while |b| exceeds 1e-14 of the pair's diagonal scale and
                    // 1e-15 of the largest diagonal entry (rounding keeps off-diagonal
                    // entries near 1e-16 of the matrix scale)
                    const double thr = fmax(1e-14 * 0.5 * (fabs(a) + fabs(d)), 1e-15 * normd);
                    if (ab > 1e-300 && ab > thr) {
                        const double z = (d - a) / (2.0 * ab);
                        const double t = (z >= 0.0 ? 1.0 : -1.0) / (fabs(z) + sqrt(1.0 + z * z));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                        ph = N::scale(b, 1.0 / ab);
                        flag = 1;
                    }
                }
                rc[i] = c;
                rs[i] = s;
                rph[i] = ph;
            }
            __syncthreads();
            // every 2x2 block: C(pi qi, pj qj) <- G_i^H C(..) G_j; and V <- V G
            for (int e = tid; e < np * np + np * n; e += kSolveThreads) {
                if (e < np * np) {
                    const int bi = e / np, bj = e - bi * np;
                    const double ci = rc[bi], si = rs[bi], cj = rc[bj], sj = rs[bj];
                    if (si == 0.0 && sj == 0.0) continue;
                    int pi, qi, pj, qj;
                    rr_pair(r, bi, m, pi, qi);
                    rr_pair(r, bj, m, pj, qj);
                    const T phi = rph[bi], phj = rph[bj];
                    const bool ok11 = pi < n && pj < n, ok12 = pi < n && qj < n;
                    const bool ok21 = qi < n && pj < n, ok22 = qi < n && qj < n;
                    const T b11 = ok11 ? C[pi + pj * ld] : N::zero();
                    const T b12 = ok12 ? C[pi + qj * ld] : N::zero();
                    const T b21 = ok21 ? C[qi + pj * ld] : N::zero();
                    const T b22 = ok22 ? C[qi + qj * ld] : N::zero();
                    // T = B G_j
                    const T t11 = N::sub(N::scale(b11, cj), N::scale(N::cmul(phj, b12), sj));
                    const T t21 = N::sub(N::scale(b21, cj), N::scale(N::cmul(phj, b22), sj));
                    const T t12 = N::add(N::scale(N::mul(phj, b11), sj), N::scale(b12, cj));
                    const T t22 = N::add(N::scale(N::mul(phj, b21), sj), N::scale(b22, cj));
                    // N = G_i^H T
                    T n11 = N::sub(N::scale(t11, ci), N::scale(N::mul(phi, t21), si));
                    T n12 = N::sub(N::scale(t12, ci), N::scale(N::mul(phi, t22), si));
                    T n21 = N::add(N::scale(N::cmul(phi, t11), si), N::scale(t21, ci));
                    T n22 = N::add(N::scale(N::cmul(phi, t12), si), N::scale(t22, ci));
                    if (bi == bj) {   // the rotated pair itself: exact zeros, real diagonal
                        n12 = N::zero();
                        n21 = N::zero();
                        n11 = N::mk(N::re(n11), 0.0);
                        n22 = N::mk(N::re(n22), 0.0);
                    }
                    if (ok11) C[pi + pj * ld] = n11;
                    if (ok12) C[pi + qj * ld] = n12;
                    if (ok21) C[qi + pj * ld] = n21;
                    if (ok22) C[qi + qj * ld] = n22;
                } else {
                    const int e2 = e - np * np;
                    const int i = e2 / n, row = e2 - i * n;
                    const double c = rc[i], sn = rs[i];
                    if (sn == 0.0) continue;
                    int p, q;
                    rr_pair(r, i, m, p, q);
                    const T ph = rph[i];
                    const T vp = M[row + p * ld], vq = M[row + q * ld];
                    M[row + p * ld] = N::sub(N::scale(vp, c), N::scale(N::cmul(ph, vq), sn));
                    M[row + q * ld] = N::add(N::scale(N::mul(ph, vp), sn), N::scale(vq, c));
                }
            }
            __syncthreads();
        }
        if (!flag) break;
    }
    if (sweep == kMaxSweeps) {
        if (tid == 0) info[k] = -1;
        return;
    }
    // ---- eigenvalues ascending (rank sort, ties by index) ----
    for (int i = tid; i < n; i += kSolveThreads) red[i] = N::re(C[i + i * ld]);
    __syncthreads();
    for (int i = tid; i < n; i += kSolveThreads) {
        const double v = red[i];
        int rank = 0;
        for (int j = 0; j < n; ++j) {
            const double u = red[j];
            rank += (u < v) || (u == v && j < i);
        }
        perm[rank] = i;
    }
    __syncthreads();
    for (int i = tid; i < n; i += kSolveThreads) w[i] = red[perm[i]];
    // ---- X = U^-1 V[:, perm]: U back into C, back substitution in M ----
    for (int e = tid; e < n * n; e += kSolveThreads) {
        const int r = e % n, c = e / n;
        C[r + c * ld] = N::load(B, e);
    }
    __syncthreads();
    for (int j = n - 1; j >= 0; --j) {
        const double inv = 1.0 / N::re(C[j + j * ld]);
        for (int c = tid; c < n; c += kSolveThreads) {
            T v = M[j + c * ld];
            for (int q = j + 1; q < n; ++q) v = N::sub(v, N::mul(C[j + q * ld], M[q + c * ld]));
            M[j + c * ld] = N::scale(v, inv);
        }
        __syncthreads();
    }
    for (int e = tid; e < n * n; e += kSolveThreads) {
        const int r = e % n, c = e / n;
        N::store(A, e, M[r + perm[c] * ld]);
    }
    if (tid == 0) info[k] = 0;
}

size_t small_hegv_lds(int n, bool cplx)
{
    const size_t el = cplx ? 16 : 8;
    return 2 * (size_t)n * (n + 1) * el + 2 * (size_t)(n / 2 + 1) * sizeof(double) +
           (size_t)(n / 2 + 1) * el + (size_t)(n + kSolveThreads / 64 + 1) * sizeof(double) +
           (size_t)n * sizeof(int) + 64;
}

bool small_hegv_fits(int n, bool cplx)
{
    return n >= 1 && small_hegv_lds(n, cplx) <= 160 * 1024;
}

void launch_small_hegv(bool cplx, int n, int nb, double *A, double *B, double *w, int *info,
                       hipStream_t s)
{
    const size_t lds = small_hegv_lds(n, cplx);
    if (cplx)
        hipLaunchKernelGGL(k_small_hegv<true>, dim3(nb), dim3(kSolveThreads), lds, s, A, B, w,
                           info, n);
    else
        hipLaunchKernelGGL(k_small_hegv<false>, dim3(nb), dim3(kSolveThreads), lds, s, A, B, w,
                           info, n);
}

}  // namespace olcao
