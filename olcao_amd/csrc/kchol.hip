// Batched blocked Cholesky S = U^H U (upper triangle, column-major) and the
// triangular solves of the generalised eigenproblem, the zpotrf/dpotrf and
// zhegst/zhegv back-substitution steps inside zhegv/dsygv
// (interfaces.F90:73-200).
//
// Left-looking by 64-wide block rows J (the host issues the GEMMs):
//   S(J, >=J) -= U(<J, J)^H U(<J, >=J)       GEMM
//   U(J, J) = chol(S(J, J)), Dinv_J = U(J, J)^-1     this kernel
//   U(J, >J) = Dinv_J^H S(J, >J)             GEMM
// The diagonal kernel is one workgroup per matrix; thread (c, g) keeps rows
// 16 g .. 16 g + 15 of column c in registers.  Elimination publishes row q
// through LDS at each pivot, a(r, c) -= conj(a(q, r)) a(q, c) / a(q, q); the
// inverse is formed by rows from the bottom, X(q, :) = R(q, :) / U(q, q),
// R(r, :) -= U(r, q) X(q, :), publishing X's row and U's column q.  info[k] receives the first column (1-based) whose
// pivot is not positive, as potrf reports it.
#include <hip/hip_runtime.h>

#include "numtraits.hpp"
#include "olcao_internal.h"

namespace olcao {

namespace {

constexpr int kDiagThreads = 256;   // 64 columns x 4 row groups of 16

template <bool CPLX>
__global__ __launch_bounds__(kDiagThreads) void k_chol_diag(const double *Ball, double *Uall,
                                                            double *DinvAll, int n, int j0, int jb,
                                                            int *info)
{
    using N = Num<CPLX>;
    using T = typename N::T;
    constexpr int NB = kCholBlock;
    constexpr int E = CPLX ? 2 : 1;
    __shared__ T tile[NB][NB + 1];   // staging for coalesced loads and stores
    __shared__ T rowq[NB], colq[NB], xrow[NB];
    __shared__ double sq[NB];
    const int tid = threadIdx.x, c = tid & 63, g = tid >> 6, mat = blockIdx.x;
    const size_t per = (size_t)n * n;
    const double *B = Ball + E * per * mat;
    double *U = Uall + E * per * mat;
    double *Dinv = DinvAll + E * per * mat;
    // thread (c, g) holds a(16 g + i, c), i < 16
    for (int e = tid; e < NB * NB; e += kDiagThreads) {
        const int r = e % NB, cc = e / NB;
        tile[r][cc] = (r < jb && cc < jb && r <= cc) ? N::load(B, (size_t)(j0 + r) + (size_t)n * (j0 + cc))
                                                     : N::zero();
    }
    __syncthreads();
    T a[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) a[i] = tile[16 * g + i][c];
    int bad = 0;
    // a(r, c) -= conj(a(q, r)) a(q, c) / a(q, q), q < r <= c; q = 16 qb + i2
    // with i2 unrolled, so that row q's register index is a constant
    for (int qb = 0; qb < 4; ++qb) {
#pragma unroll
        for (int i2 = 0; i2 < 16; ++i2) {
            const int q = 16 * qb + i2;
            if (q < jb) {
                if (g == qb) rowq[c] = a[i2];
                __syncthreads();
                double d = N::re(rowq[q]);
                if (!(d > 0.0)) {
                    if (!bad) bad = j0 + q + 1;
                    d = 1.0;
                }
                const T f = N::scale(rowq[c], 1.0 / d);
                T rq[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) rq[i] = rowq[16 * g + i];
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int r = 16 * g + i;
                    const T upd = N::sub(a[i], N::cmul(rq[i], f));
                    a[i] = N::sel(r > q && r <= c && c < jb, upd, a[i]);
                }
                __syncthreads();
            }
        }
    }
    // U(r, c) = a(r, c) / sqrt(a(r, r)), U(r, r) = sqrt(a(r, r))
#pragma unroll
    for (int i = 0; i < 16; ++i)
        if (c == 16 * g + i) {
            const double d = N::re(a[i]);
            sq[c] = sqrt(d > 0.0 ? d : 1.0);
        }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int r = 16 * g + i;
        a[i] = N::sel(r == c, N::mk(sq[r], 0.0), N::sel(r < c, N::scale(a[i], 1.0 / sq[r]), N::zero()));
        tile[r][c] = a[i];
    }
    __syncthreads();
    for (int e = tid; e < NB * NB; e += kDiagThreads) {
        const int r = e % NB, cc = e / NB;
        if (r < jb && cc < jb && r <= cc) N::store(U, (size_t)(j0 + r) + (size_t)n * (j0 + cc), tile[r][cc]);
    }
    // X = U^-1 by rows from the bottom: X(q, :) = R(q, :) / U(q, q), then
    // R(r, :) -= U(r, q) X(q, :) for r < q, starting from R = I
    T x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = N::mk(16 * g + i == c ? 1.0 : 0.0, 0.0);
    for (int qb = 3; qb >= 0; --qb) {
#pragma unroll
        for (int i2 = 15; i2 >= 0; --i2) {
            const int q = 16 * qb + i2;
            if (q < jb) {
                if (g == qb) {
                    x[i2] = N::scale(x[i2], 1.0 / sq[q]);
                    xrow[c] = x[i2];
                }
                if (c == q) {
#pragma unroll
                    for (int i = 0; i < 16; ++i) colq[16 * g + i] = a[i];
                }
                __syncthreads();
                const T xqc = xrow[c];
                T cq[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) cq[i] = colq[16 * g + i];
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const T upd = N::sub(x[i], N::mul(cq[i], xqc));
                    x[i] = N::sel(16 * g + i < q, upd, x[i]);
                }
                __syncthreads();
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) tile[16 * g + i][c] = x[i];
    __syncthreads();
    for (int e = tid; e < NB * NB; e += kDiagThreads) {
        const int r = e % NB, cc = e / NB;
        if (r < jb && cc < jb) N::store(Dinv, (size_t)(j0 + r) + (size_t)n * cc, tile[r][cc]);
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
