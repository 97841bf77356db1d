// Batched blocked Cholesky S = U^H U (upper triangle, column-major) and the
// triangular solves of the generalised eigenproblem, the zpotrf/dpotrf and
// zhegst/zhegv back-substitution steps inside zhegv/dsygv
// (interfaces.F90:73-200).
//
// Left-looking by 64-wide block rows J (the host issues the GEMMs):
//   S(J, >=J) -= U(<J, J)^H U(<J, >=J)       GEMM
//   U(J, J) = chol(S(J, J)), Dinv_J = U(J, J)^-1     this kernel
//   U(J, >J) = Dinv_J^H S(J, >J)             GEMM
// The diagonal kernel is one workgroup per matrix eliminating on [S_JJ | I]:
// wave w keeps rows w, w + 8, ... and lane l columns l and 64 + l in
// registers.  At pivot q the wave owning row q publishes it through LDS
// (double-buffered, one barrier per pivot) and every row r > q takes
// a(r, c) -= conj(a(q, r)) a(q, c) / a(q, q), on the upper part of S (c >= r)
// and on the columns c <= q of the I part, the only ones not yet final;
// rows r <= q are skipped by whole waves.  With S_JJ = L D L^H the S part
// ends as D L^H and the identity as L^-1, so U = D^-1/2 D L^H and
// U^-1 = (D^-1/2 L^-1)^H come out of the same 64 pivot steps.  info[k]
// receives the first column (1-based) whose pivot is not positive, as potrf
// reports it.
#include <hip/hip_runtime.h>

#include "numtraits.hpp"
#include "olcao_internal.h"

namespace olcao {

namespace {

constexpr int kDiagThreads = 512;   // 8 waves: wave w rows w + 8 k, lane l columns l, 64 + l
constexpr int kDiagRows = kCholBlock / (kDiagThreads / 64);

template <bool CPLX>
__global__ __launch_bounds__(kDiagThreads) void k_chol_diag(const double *Ball, double *Uall,
                                                            double *DinvAll, int n, int j0, int jb,
                                                            int *info)
{
    using N = Num<CPLX>;
    using T = typename N::T;
    constexpr int NB = kCholBlock;
    constexpr int E = CPLX ? 2 : 1;
    constexpr int RW = kDiagRows;
    static_assert(NB == 64 && RW * (kDiagThreads / 64) == NB, "one column per lane, rows by wave");
    __shared__ T tile[NB][NB + 1];   // staging for coalesced loads and stores
    __shared__ T prow[2][2 * NB];    // the pivot row of [S_JJ | I], double-buffered
    __shared__ double sq[NB];
    const int tid = threadIdx.x, l = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int mat = blockIdx.x;
    const size_t per = (size_t)n * n;
    const double *B = Ball + E * per * mat;
    double *U = Uall + E * per * mat;
    double *Dinv = DinvAll + E * per * mat;
    for (int e = tid; e < NB * NB; e += kDiagThreads) {
        const int r = e % NB, cc = e / NB;
        tile[r][cc] = (r < jb && cc < jb && r <= cc) ? N::load(B, (size_t)(j0 + r) + (size_t)n * (j0 + cc))
                                                     : N::zero();
    }
    __syncthreads();
    T a[RW][2];
#pragma unroll
    for (int k = 0; k < RW; ++k) {
        const int r = w + 8 * k;
        a[k][0] = tile[r][l];
        a[k][1] = N::mk(r == l ? 1.0 : 0.0, 0.0);
    }
    int bad = 0;
    for (int q = 0; q < jb; ++q) {
        const int b = q & 1;
        // row q is final since pivot q - 1; its owner publishes it
        if (w == (q & 7)) {
            const int kq = q >> 3;
            T s0 = N::zero(), s1 = N::zero();
#pragma unroll
            for (int k = 0; k < RW; ++k)
                if (k == kq) {
                    s0 = a[k][0];
                    s1 = a[k][1];
                }
            prow[b][l] = s0;
            prow[b][NB + l] = s1;
        }
        __syncthreads();
        double d = N::re(prow[b][q]);
        if (!(d > 0.0)) {
            if (!bad) bad = j0 + q + 1;
            d = 1.0;
        }
        const double inv = 1.0 / d;
        const T ps = prow[b][l], pi = prow[b][NB + l];
#pragma unroll
        for (int k = 0; k < RW; ++k) {
            const int r = w + 8 * k;
            if (r > q && r < jb) {   // wave-uniform
                const T f = N::scale(prow[b][r], inv);   // a(q, r) / a(q, q)
                if (l >= r) a[k][0] = N::sub(a[k][0], N::cmul(f, ps));
                if (l <= q) a[k][1] = N::sub(a[k][1], N::cmul(f, pi));
            }
        }
    }
    // d_r = a(r, r); U = D^-1/2 (D L^H), U^-H = D^-1/2 L^-1
#pragma unroll
    for (int k = 0; k < RW; ++k) {
        const int r = w + 8 * k;
        if (l == r) {
            const double d = N::re(a[k][0]);
            sq[r] = sqrt(d > 0.0 ? d : 1.0);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < RW; ++k) {
        const int r = w + 8 * k;
        tile[r][l] = N::sel(r == l, N::mk(sq[r], 0.0),
                            N::sel(r < l, N::scale(a[k][0], 1.0 / sq[r]), N::zero()));
    }
    __syncthreads();
    for (int e = tid; e < NB * NB; e += kDiagThreads) {
        const int r = e % NB, cc = e / NB;
        if (r < jb && cc < jb && r <= cc) N::store(U, (size_t)(j0 + r) + (size_t)n * (j0 + cc), tile[r][cc]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < RW; ++k) {
        const int r = w + 8 * k;
        tile[r][l] = N::scale(a[k][1], 1.0 / sq[r]);
    }
    __syncthreads();
    // Dinv = U^-1 = (U^-H)^H: Dinv(c', r) = conj(U^-H(r, c'))
    for (int e = tid; e < NB * NB; e += kDiagThreads) {
        const int cp = e % NB, rr = e / NB;
        if (cp < jb && rr < jb) N::store(Dinv, (size_t)(j0 + cp) + (size_t)n * rr, N::conj(tile[rr][cp]));
    }
    if (tid == 0 && bad && info[mat] == 0) info[mat] = bad;
}

}  // namespace

void launch_chol_diag(bool cplx, const double *B, double *U, double *Dinv, int n, int nb, int j0,
                      int jb, int *info, hipStream_t s)
{
    if (cplx)
        hipLaunchKernelGGL(k_chol_diag<true>, dim3(nb), dim3(kDiagThreads), 0, s, B, U, Dinv, n, j0,
                           jb, info);
    else
        hipLaunchKernelGGL(k_chol_diag<false>, dim3(nb), dim3(kDiagThreads), 0, s, B, U, Dinv, n, j0,
                           jb, info);
}

}  // namespace olcao
