// Primitive-block and Boys kernels used by the parity tests of the device
// functions against the reference primitives.
#include "device_common.hpp"

namespace olcao {

// ---------------------------------------------------------------------------
// primitive blocks and Boys values for parity tests
template <int LA, int LB>
__device__ void prim_one(int op, const double *a, const double *A, const double *B,
                         const double *C, double *sh16)
{
    const PairConst q = make_pair(a[0], a[1], A, B);
    double pc[ncart(LA)][ncart(LB)];
    if (op == 0) {
        two_center_cart<LA, LB, false>(q, pc);
    } else if (op == 1) {
        two_center_cart<LA, LB, true>(q, pc);
    } else if (op == 4) {
        massvel_cart<LA, LB>(q, pc);
    } else if (op == 7 || op == 8) {
        // the k-overlap block for kq = C: real part (7), imaginary part (8)
        kov_cart<LA, LB>(q, C, op - 7, pc);
        if (op == 8)
#pragma unroll
            for (int x = 0; x < ncart(LA); ++x)
#pragma unroll
                for (int y = 0; y < ncart(LB); ++y) pc[x][y] = -pc[x][y];
    } else {
        constexpr int L = LA + LB;
        double W[nherm(L)];
#pragma unroll
        for (int e = 0; e < nherm(L); ++e) W[e] = 0.0;
        if (op == 2)
            acc_e3c_herm<L>(q, a[2], C, 1.0, W);
        else
            acc_nuc_herm<L>(q, a[2], C, 1.0, W);
        double E[3][LA + 1][LB + 1][L + 1];
#pragma unroll
        for (int x = 0; x < 3; ++x) herm_e<LA, LB>(q.P[x] - q.A[x], q.P[x] - q.B[x], 0.5 / q.zeta, E[x]);
        const double kab = exp(-q.xid2);
#pragma unroll
        for (int x = 0; x < ncart(LA); ++x)
#pragma unroll
            for (int y = 0; y < ncart(LB); ++y)
                pc[x][y] = kab * herm_contract<LA, LB>(E, W, cexp(x, 0), cexp(x, 1), cexp(x, 2),
                                                       cexp(y, 0), cexp(y, 1), cexp(y, 2));
    }
    double sh[nsh(LA)][nsh(LB)];
    cart_to_sh<LA, LB>(pc, sh);
    for (int e = 0; e < 256; ++e) sh16[e] = 0.0;
#pragma unroll
    for (int i = 0; i < nsh(LA); ++i)
#pragma unroll
        for (int j = 0; j < nsh(LB); ++j) sh16[i + 16 * j] = sh[i][j];
}

// One kernel per angular class; the host groups the inputs by class.
template <int LA, int LB>
__global__ __launch_bounds__(64) void k_prim(int op, int n, const int *idx, const double *a,
                                             const double *A, const double *B, const double *C,
                                             double *sh)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = idx[t];
    prim_one<LA, LB>(op, a + 3 * i, A + 3 * i, B + 3 * i, C + 3 * i, sh + 256LL * i);
}

__global__ void k_boys(int n, const double *T, double *F)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double f[7];
    boys<6>(T[t], f);
    for (int m = 0; m < 7; ++m) F[7 * t + m] = f[m];
}

void launch_prim_batch(int op, int n, const double *a, const double *A, const double *B,
                       const double *C, const int *const *classIdx, const int *classCount,
                       double *sh, hipStream_t s)
{
    (void)n;
#define PRIM_CLASS(X, Y)                                                                    \
    if (classCount[X * 4 + Y] > 0)                                                          \
        hipLaunchKernelGGL((k_prim<X, Y>), dim3((classCount[X * 4 + Y] + 63) / 64), dim3(64), 0, \
                           s, op, classCount[X * 4 + Y], classIdx[X * 4 + Y], a, A, B, C, sh);
    PRIM_CLASS(0, 0) PRIM_CLASS(0, 1) PRIM_CLASS(0, 2) PRIM_CLASS(0, 3)
    PRIM_CLASS(1, 0) PRIM_CLASS(1, 1) PRIM_CLASS(1, 2) PRIM_CLASS(1, 3)
    PRIM_CLASS(2, 0) PRIM_CLASS(2, 1) PRIM_CLASS(2, 2) PRIM_CLASS(2, 3)
    PRIM_CLASS(3, 0) PRIM_CLASS(3, 1) PRIM_CLASS(3, 2) PRIM_CLASS(3, 3)
#undef PRIM_CLASS
}

void launch_boys_batch(int n, const double *T, double *F, hipStream_t s)
{
    hipLaunchKernelGGL(k_boys, dim3((n + 63) / 64), dim3(64), 0, s, n, T, F);
}

}  // namespace olcao
