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
// thread (c, g) keeps rows 16 g .. 16 g + 15 of column c in registers, row q
// is published through LDS at each pivot and a(r, c) -= conj(a(q, r)) a(q, c)
// / a(q, q) for r > q.  With S_JJ = L D L^H the S part ends as D L^H and the
// identity as L^-1, so U = D^-1/2 D L^H and U^-1 = (D^-1/2 L^-1)^H come out
// of the same 64 pivot steps.  info[k] receives the first column (1-based) whose
// pivot is not positive, as potrf reports it.
#include <hip/hip_runtime.h>

#include "numtraits.hpp"
#include "olcao_internal.h"

namespace olcao {

namespace {

constexpr int kDiagThreads = 512;   // 128 columns of [S_JJ | I] x 4 row groups of 16

template <bool CPLX>
__global__ __launch_bounds__(kDiagThreads) void k_chol_diag(const double *Ball, double *Uall,
                                                            double *DinvAll, int n, int j0, int jb,
                                                            int *info)
{
    using N = Num<CPLX>;
    using T = typename N::T;
    constexpr int NB = kCholBlock;
    constexpr int E = CPLX ? 2 : 1;
    __shared__ T tile[NB][2 * NB + 1];   // staging of [S_JJ | I] for coalesced loads and stores
    __shared__ T rowq[2 * NB];
    __shared__ double sq[NB];
    const int tid = threadIdx.x, c = tid & 127, g = tid >> 7, mat = blockIdx.x;
    const size_t per = (size_t)n * n;
    const double *B = Ball + E * per * mat;
    double *U = Uall + E * per * mat;
    double *Dinv = DinvAll + E * per * mat;
    for (int e = tid; e < NB * NB; e += kDiagThreads) {
        const int r = e % NB, cc = e / NB;
        tile[r][cc] = (r < jb && cc < jb && r <= cc) ? N::load(B, (size_t)(j0 + r) + (size_t)n * (j0 + cc))
                                                     : N::zero();
        tile[r][NB + cc] = N::mk(r == cc ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    // thread (c, g) holds rows 16 g + i of column c of [S_JJ | I]
    T a[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) a[i] = tile[16 * g + i][c];
    const bool live = c < NB ? c < jb : c - NB < jb;
    int bad = 0;
    // row r -= conj(a(q, r)) / a(q, q) * row q for r > q: the S part stays
    // upper (columns c >= r), the I part becomes L^-1 of S = L D L^H; q = 16 qb + i2
    // with i2 unrolled so that row q's register index is a constant
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
                    a[i] = N::sel(r > q && (c >= NB || r <= c) && live, upd, a[i]);
                }
                __syncthreads();
            }
        }
    }
    // d_r = a(r, r); U = D^-1/2 (D L^H), U^-H = D^-1/2 L^-1
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
        T v;
        if (c < NB)
            v = N::sel(r == c, N::mk(sq[r], 0.0), N::sel(r < c, N::scale(a[i], 1.0 / sq[r]), N::zero()));
        else
            v = N::scale(a[i], 1.0 / sq[r]);
        tile[r][c] = v;
    }
    __syncthreads();
    // U (upper) and Dinv = U^-1 = (U^-H)^H: Dinv(c', r) = conj(U^-H(r, c'))
    for (int e = tid; e < NB * NB; e += kDiagThreads) {
        const int r = e % NB, cc = e / NB;
        if (r < jb && cc < jb && r <= cc) N::store(U, (size_t)(j0 + r) + (size_t)n * (j0 + cc), tile[r][cc]);
        const int cp = e % NB, rr = e / NB;
        if (cp < jb && rr < jb) N::store(Dinv, (size_t)(j0 + cp) + (size_t)n * rr, N::conj(tile[rr][NB + cp]));
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
