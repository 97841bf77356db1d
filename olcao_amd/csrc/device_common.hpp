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


// Origin shift and scatter of one output element (k, s1, s2) of atom pair
// (i, j) (kPointLatticeOriginShift and saveCurrentPair, intgSaving.F90:
// 241-292, :346-459).  mode bit 0: the Bloch sum times -i (the momentum
// matrix, applyPhaseFactors runCode 1; at Gamma the kept part is -X).
// mode bit 1: the full save (saveFull = 1, valeValeSavingFull and
// coreCoreSavingFull): the block in place and its dagger at the mirrored
// place, the dagger last, so a diagonal atom block ends as its own dagger;
// otherwise the upper triangle with the strictly lower part of a diagonal
// block zeroed.  coreVale is the same in both.
__device__ __forceinline__ void bloch_store(const DevCtx &c, int i, int j, int k, int s1, int s2,
                                            int nv1, int nv2, int v1, int v2, int c1, int c2,
                                            const double *lv, double vre, double vim, int gamma,
                                            int mode, double *vv, double *cv, double *cc)
{
    const long long n = c.valeDim, nc = c.coreDim;
    const bool full = mode & 2;
    if (mode & 1) {
        const double t = vre;
        vre = gamma ? -t : vim;
        vim = -t;
    }
    if (!gamma) {
        const double *kp = c.kcart + 3 * k;
        const double dot = kp[0] * lv[0] + kp[1] * lv[1] + kp[2] * lv[2];
        if (fabs(dot) >= kSmallThresh) {
            const double cr = cos(dot), ci = sin(dot);
            const double r2 = cr * vre - ci * vim;
            vim = cr * vim + ci * vre;
            vre = r2;
        }
    }
    const int cp = gamma ? 1 : 2;
    const bool a_core = s1 >= nv1, b_core = s2 >= nv2;
    if (!a_core && !b_core) {
        const long long idx = (v1 + s1) + (long long)(v2 + s2) * n + (long long)k * n * n;
        if (full) {
            const long long idt = (v2 + s2) + (long long)(v1 + s1) * n + (long long)k * n * n;
            vv[idt * cp] = vre;
            if (!gamma) vv[idt * cp + 1] = -vim;
            if (i != j) {
                vv[idx * cp] = vre;
                if (!gamma) vv[idx * cp + 1] = vim;
            }
        } else {
            const bool zero = (i == j) && (s1 > s2);
            vv[idx * cp] = zero ? 0.0 : vre;
            if (!gamma) vv[idx * cp + 1] = zero ? 0.0 : vim;
        }
    } else if (a_core && !b_core) {
        if (i != j) {   // for i == j the dagger assignment below wins
            const long long idx = (c1 + s1 - nv1) + (long long)(v2 + s2) * nc + (long long)k * nc * n;
            cv[idx * cp] = vre;
            if (!gamma) cv[idx * cp + 1] = vim;
        }
    } else if (!a_core && b_core) {
        const long long idx = (c2 + s2 - nv2) + (long long)(v1 + s1) * nc + (long long)k * nc * n;
        cv[idx * cp] = vre;
        if (!gamma) cv[idx * cp + 1] = -vim;
    } else {
        if (!full || i != j) {
            const long long idx = (c1 + s1 - nv1) + (long long)(c2 + s2 - nv2) * nc + (long long)k * nc * nc;
            cc[idx * cp] = vre;
            if (!gamma) cc[idx * cp + 1] = vim;
        }
        if (full || i != j) {
            const long long id2 = (c2 + s2 - nv2) + (long long)(c1 + s1 - nv1) * nc + (long long)k * nc * nc;
            cc[id2 * cp] = vre;
            if (!gamma) cc[id2 * cp + 1] = -vim;
        }
    }
}

}  // namespace olcao
