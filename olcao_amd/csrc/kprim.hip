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
    } else {
#pragma unroll
        for (int x = 0; x < ncart(LA); ++x)
#pragma unroll
            for (int y = 0; y < ncart(LB); ++y) pc[x][y] = 0.0;
        if (op == 2)
            acc_e3c<LA, LB>(q, a[2], C, 1.0, pc);
        else
            acc_nuc<LA, LB>(q, a[2], C, 1.0, pc);
    }
    double sh[nsh(LA)][nsh(LB)];
    cart_to_sh<LA, LB>(pc, sh);
    for (int e = 0; e < 256; ++e) sh16[e] = 0.0;
#pragma unroll
    for (int i = 0; i < nsh(LA); ++i)
#pragma unroll
        for (int j = 0; j < nsh(LB); ++j) sh16[i + 16 * j] = sh[i][j];
}

__device__ __forceinline__ int sw_l(int nib) { return nib == 1 ? 0 : nib == 2 ? 1 : nib == 4 ? 2 : 3; }

__global__ void k_prim(int op, int n, const double *a, const double *A, const double *B,
                       const double *C, const int *sw, double *sh)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int la = sw_l(sw[t] & 15), lb = sw_l((sw[t] >> 4) & 15);
    const double *aa = a + 3 * t;
    double *o = sh + 256LL * t;
#define PRIM_CASE(X, Y) \
    if (la == X && lb == Y) { prim_one<X, Y>(op, aa, A + 3 * t, B + 3 * t, C + 3 * t, o); return; }
    PRIM_CASE(0, 0) PRIM_CASE(0, 1) PRIM_CASE(0, 2) PRIM_CASE(0, 3)
    PRIM_CASE(1, 0) PRIM_CASE(1, 1) PRIM_CASE(1, 2) PRIM_CASE(1, 3)
    PRIM_CASE(2, 0) PRIM_CASE(2, 1) PRIM_CASE(2, 2) PRIM_CASE(2, 3)
    PRIM_CASE(3, 0) PRIM_CASE(3, 1) PRIM_CASE(3, 2) PRIM_CASE(3, 3)
#undef PRIM_CASE
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
                       const double *C, const int *sw, double *sh, hipStream_t s)
{
    hipLaunchKernelGGL(k_prim, dim3((n + 63) / 64), dim3(64), 0, s, op, n, a, A, B, C, sw, sh);
}

void launch_boys_batch(int n, const double *T, double *F, hipStream_t s)
{
    hipLaunchKernelGGL(k_boys, dim3((n + 63) / 64), dim3(64), 0, s, n, T, F);
}

}  // namespace olcao
