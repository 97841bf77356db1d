// Device helpers shared by the OLCAO kernels.
#pragma once

#include "olcao_internal.h"
#include "prims.hpp"

namespace olcao {

static __device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Lattice vector nearest to v: round the fractional coordinates, then try
// the candidate corrections that can bring a point closer (none for
// orthorhombic cells).
static __device__ __forceinline__ void nearest_lattice(const DevCtx &c, const double *v,
                                                double *out)
{
    double base[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double f = v[0] * c.inv[3 * i] + v[1] * c.inv[3 * i + 1] + v[2] * c.inv[3 * i + 2];
        double r = trunc(fabs(f) + 0.5);
        if (f < 0.0) r = -r;
#pragma unroll
        for (int x = 0; x < 3; ++x) base[x] += c.vec[3 * i + x] * r;
    }
    double rem[3] = {v[0] - base[0], v[1] - base[1], v[2] - base[2]};
    double best = rem[0] * rem[0] + rem[1] * rem[1] + rem[2] * rem[2];
    int bi = -1;
    for (int q = 0; q < c.nCorr; ++q) {
        double r0 = rem[0] - c.corr[3 * q], r1 = rem[1] - c.corr[3 * q + 1],
               r2 = rem[2] - c.corr[3 * q + 2];
        double s = r0 * r0 + r1 * r1 + r2 * r2;
        if (s < best * (1.0 - 1e-14)) {
            best = s;
            bi = q;
        }
    }
#pragma unroll
    for (int x = 0; x < 3; ++x) out[x] = base[x] + (bi >= 0 ? c.corr[3 * bi + x] : 0.0);
}

struct PPInfo {
    int ti, tj, ia, ib;
    double s2;
    double A[3], B[3];
};

static __device__ __forceinline__ PPInfo load_pp(const DevCtx &c, int pp)
{
    PPInfo r;
    const int pc = c.ppPc[pp];
    const int ab = c.ppAB[pp];
    const int pr = c.pcPair[pc];
    const int i = c.pairI[pr], j = c.pairJ[pr];
    r.ti = c.atomType[i];
    r.tj = c.atomType[j];
    r.ia = pp_a1(ab);
    r.ib = pp_a2(ab);
    r.s2 = c.pcS2[pc];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        r.A[x] = c.pos[3 * i + x];
        r.B[x] = c.pcB[3 * pc + x];
    }
    return r;
}

// Writes (accumulate = 0) or adds onto (accumulate = 1) the primitive-pair
// blocks of up to 64 records staged in LDS by a 64-lane workgroup: word k of
// sOut is element k % NE of record k / NE, whose block starts at sOff[k / NE].
// The records' blocks are disjoint, so DEPTH loads are issued before any
// store; a load-add-store loop over the records one by one would wait a
// full memory round trip per record.  Lanes write consecutive words.
template <int NE, int DEPTH = 16>
static __device__ __forceinline__ void store_staged_blocks(const double *sOut,
                                                           const long long *sOff, int nr,
                                                           int accumulate, double *blocks)
{
    const int lane = threadIdx.x;
    const int tot = nr * NE;
    for (int k0 = 0; k0 < tot; k0 += 64 * DEPTH) {
        double v[DEPTH], o[DEPTH];
        long long a[DEPTH];
#pragma unroll
        for (int u = 0; u < DEPTH; ++u) {
            const int k = k0 + u * 64 + lane;
            const int r = k / NE;
            const bool ok = k < tot;
            a[u] = ok ? sOff[r] + (k - r * NE) : 0;
            v[u] = ok ? sOut[k] : 0.0;
        }
        if (accumulate) {
#pragma unroll
            for (int u = 0; u < DEPTH; ++u) o[u] = k0 + u * 64 + lane < tot ? blocks[a[u]] : 0.0;
#pragma unroll
            for (int u = 0; u < DEPTH; ++u)
                if (k0 + u * 64 + lane < tot) blocks[a[u]] = o[u] + v[u];
        } else {
#pragma unroll
            for (int u = 0; u < DEPTH; ++u)
                if (k0 + u * 64 + lane < tot) blocks[a[u]] = v[u];
        }
    }
}

}  // namespace olcao
