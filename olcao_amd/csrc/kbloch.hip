// Bloch sum on the matrix cores: for every atom pair, the complex
// [nk x nCells(pair)] phase matrix times the real [nCells(pair) x n1 n2]
// matrix of its contracted blocks (applyPhaseFactors, intgSaving.F90:132-189,
// currentPair(:,:,k) += phaseFactor(k,cell) X_cell), followed by the
// origin-shift phase (kPointLatticeOriginShift, :241-292, skipped when
// |k.L| < 1e-8) and the scatter into the upper-triangular global matrices
// (saveCurrentPair and its core/Gamma twins, :346-700).
//
// One wavefront per (atom pair, tile of 16 k-points); the pair's outputs
// e = s2 n1 + s1 (s1 fastest, so that 16 neighbouring lanes store 16
// neighbouring rows of a column-major matrix) are taken 16 columns at a
// time; the cells are the reduction dimension.
//
//   FP64 form   v_mfma_f64_16x16x4f64, 4 cells per instruction, one
//               accumulator tile each for the real and imaginary parts.
//   mixed form  split-precision products on v_mfma_f32_16x16x32_bf16
//               (32 cells per instruction) that are exact: each phase
//               component (|P| <= 1) is cut into kSlices fixed-point
//               slices P = sum_i m_i 2^(-7-8i), each block column e, scaled
//               by its power of two 2^E >= max_cells |X_e|, likewise
//               X = 2^E sum_j n_j 2^(-7-8j), with integers |m_i|, |n_j| <=
//               128 that BF16 holds exactly.  A product m_i n_j is exact in
//               FP32 and all products of one order d = i + j share the
//               quantum 2^(E-14-8d), so an FP32 accumulator per order sums
//               them without rounding as long as it stays below 2^24; it is
//               flushed into FP64 every kFlushChunks chunks of 32 cells.
//               Orders d < kSlices = 6 are kept (21 BF16 products per chunk
//               and component); the truncation leaves a relative error of
//               about 2^-48 of the column's largest block entry per cell.
#include <cstring>
#include <vector>

#include "device_common.hpp"

namespace olcao {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int kFlushChunks = 4;   // 4 chunks x 32 cells x 6 products x 2^14 < 2^24
static_assert(kFlushChunks * 32 * kBlochSlices * (1 << 14) < (1 << 24),
              "an order's FP32 accumulator must stay exact between flushes");
static_assert(kBlochSlices <= 8, "the phase-slice table holds 8 slices");

// BF16 bits of a small integer held in a float (exact: <= 8 significant bits)
__device__ __forceinline__ unsigned short bf16_bits(float v)
{
    return (unsigned short)(__float_as_uint(v) >> 16);
}

template <bool SPLIT>
__global__ __launch_bounds__(64) void k_bloch_mfma(DevCtx c, const double *X,
                                                   const unsigned short *pslice, double *vv,
                                                   double *cv, double *cc, int gamma, int mode,
                                                   int nTiles)
{
    const int lane = threadIdx.x;
    const int pr = blockIdx.x / nTiles, tile = blockIdx.x - pr * nTiles;
    if (pr >= c.nPairs) return;
    const int i = c.pairI[pr], j = c.pairJ[pr];
    const int ti = c.atomType[i], tj = c.atomType[j];
    const int nv1 = c.nVale[ti], nv2 = c.nVale[tj];
    const int nt1 = nv1 + c.nCore[ti], nt2 = nv2 + c.nCore[tj];
    const int v1 = c.cumulVale[i], v2 = c.cumulVale[j];
    const int c1 = c.cumulCore[i], c2 = c.cumulCore[j];
    const int p0 = c.pairPc[pr], p1 = c.pairPc[pr + 1];
    const int ms = c.maxS;
    const int nk = gamma ? 1 : c.nk;
    const int k0 = tile * 16;
    const int per = nt1 * nt2;
    const double lv[3] = {c.pairLv[3 * pr], c.pairLv[3 * pr + 1], c.pairLv[3 * pr + 2]};
    const int col = lane & 15, grp = lane >> 4;
    const int krowA = k0 + col;                  // A operand: k row of this lane
    const bool kOkA = krowA < nk;
    for (int e0 = 0; e0 < per; e0 += 16) {
        const int e = e0 + col;
        const bool eOk = e < per;
        const int s2 = eOk ? e / nt1 : 0, s1 = eOk ? e - s2 * nt1 : 0;
        const long long xoff = (long long)s1 * ms + s2;
        double re[4], im[4];
        if constexpr (!SPLIT) {
            d4 aR = {0.0, 0.0, 0.0, 0.0}, aI = {0.0, 0.0, 0.0, 0.0};
            for (int q0 = p0; q0 < p1; q0 += 16) {
                double b[4], pr_[4], pi_[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {   // all loads of 16 cells first
                    const int q = q0 + 4 * u + grp;
                    const bool ok = q < p1;
                    b[u] = ok && eOk ? X[(long long)q * ms * ms + xoff] : 0.0;
                    if (gamma) {
                        pr_[u] = ok ? 1.0 : 0.0;
                        pi_[u] = 0.0;
                    } else {
                        const double *ph = c.phase + 2 * ((long long)c.pcCell[ok ? q : p0] * c.nk +
                                                          (kOkA ? krowA : 0));
                        pr_[u] = ok && kOkA ? ph[0] : 0.0;
                        pi_[u] = ok && kOkA ? ph[1] : 0.0;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    aR = __builtin_amdgcn_mfma_f64_16x16x4f64(pr_[u], b[u], aR, 0, 0, 0);
                    if (!gamma) aI = __builtin_amdgcn_mfma_f64_16x16x4f64(pi_[u], b[u], aI, 0, 0, 0);
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                re[r] = aR[r];
                im[r] = aI[r];
            }
        } else {
            constexpr int S = kBlochSlices;
            // the column's scale 2^E >= max over the pair's cells of |X_e|
            double mx = 0.0;
            for (int q = p0 + grp; q < p1; q += 4)
                if (eOk) mx = fmax(mx, fabs(X[(long long)q * ms * ms + xoff]));
            mx = fmax(mx, __shfl_xor(mx, 16, 64));
            mx = fmax(mx, __shfl_xor(mx, 32, 64));
            int E = 0;
            if (mx > 0.0) (void)frexp(mx, &E);   // mx < 2^E
            const double inv = ldexp(1.0, -E);
#pragma unroll
            for (int r = 0; r < 4; ++r) re[r] = im[r] = 0.0;
            f4 accR[S], accI[S];
#pragma unroll
            for (int d = 0; d < S; ++d) accR[d] = accI[d] = f4{0.f, 0.f, 0.f, 0.f};
            int chunk = 0;
            for (int q0 = p0; q0 < p1; q0 += 32) {
                // B operand: column e, cells q0 + 8 grp + t, sliced
                bf8 bs[S];
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const int q = q0 + 8 * grp + t;
                    double x = (q < p1 && eOk) ? X[(long long)q * ms * ms + xoff] * inv : 0.0;
#pragma unroll
                    for (int sl = 0; sl < S; ++sl) {
                        const double m = rint(x * ldexp(1.0, 7 + 8 * sl));
                        x = fma(-m, ldexp(1.0, -7 - 8 * sl), x);   // exact
                        bs[sl][t] = __builtin_bit_cast(__bf16, bf16_bits((float)m));
                    }
                }
                // A operand: phase slices of k row krowA, the same cells
                bf8 as[2][S];
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const int q = q0 + 8 * grp + t;
                    const bool ok = q < p1 && kOkA;
                    const int cell = c.pcCell[q < p1 ? q : p0];
                    const uint4 *src =
                        reinterpret_cast<const uint4 *>(pslice) + 2 * ((long long)cell * nk + (kOkA ? krowA : 0));
#pragma unroll
                    for (int comp = 0; comp < 2; ++comp) {
                        const uint4 w = ok ? src[comp] : make_uint4(0, 0, 0, 0);
                        const unsigned wv[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                        for (int sl = 0; sl < S; ++sl) {
                            const unsigned short h = (unsigned short)(wv[sl >> 1] >> (16 * (sl & 1)));
                            as[comp][sl][t] = __builtin_bit_cast(__bf16, h);
                        }
                    }
                }
#pragma unroll
                for (int d = 0; d < S; ++d)
#pragma unroll
                    for (int a = 0; a <= d; ++a) {
                        accR[d] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as[0][a], bs[d - a], accR[d], 0, 0, 0);
                        if (!gamma)
                            accI[d] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as[1][a], bs[d - a], accI[d], 0,
                                                                              0, 0);
                    }
                if (++chunk == kFlushChunks || q0 + 32 >= p1) {
                    chunk = 0;
#pragma unroll
                    for (int d = 0; d < S; ++d) {
                        const double w = ldexp(1.0, E - 14 - 8 * d);
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            re[r] = fma((double)accR[d][r], w, re[r]);
                            im[r] = fma((double)accI[d][r], w, im[r]);
                        }
                        accR[d] = accI[d] = f4{0.f, 0.f, 0.f, 0.f};
                    }
                }
            }
        }
        // lane holds rows 4 r + grp (k within the tile) of column e
        if (!eOk) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            // result rows: FP64 MFMA 4 r + grp, BF16 MFMA 4 grp + r
            const int k = k0 + (SPLIT ? 4 * grp + r : 4 * r + grp);
            if (k >= nk) continue;
            bloch_store(c, i, j, k, s1, s2, nv1, nv2, v1, v2, c1, c2, lv, re[r], im[r], gamma, mode,
                        vv, cv, cc);
        }
    }
}

}  // namespace

void launch_bloch_mfma(const DevCtx &c, const double *X, const unsigned short *pslice, double *vv,
                       double *cv, double *cc, int gamma, int mode, bool split, hipStream_t s)
{
    if (c.nPairs == 0) return;
    const int nk = gamma ? 1 : c.nk;
    const int nTiles = (nk + 15) / 16;
    if (split)
        hipLaunchKernelGGL(k_bloch_mfma<true>, dim3(c.nPairs * nTiles), dim3(64), 0, s, c, X, pslice,
                           vv, cv, cc, gamma, mode, nTiles);
    else
        hipLaunchKernelGGL(k_bloch_mfma<false>, dim3(c.nPairs * nTiles), dim3(64), 0, s, c, X, pslice,
                           vv, cv, cc, gamma, mode, nTiles);
}

void phase_slices_host(const std::vector<double> &phase, int nCells, int nk, int gamma,
                       std::vector<unsigned short> &out)
{
    // [cell][k][re, im][8] BF16 bit patterns of m_i, P = sum_i m_i 2^(-7-8i)
    const int nkk = gamma ? 1 : nk;
    out.assign((size_t)nCells * nkk * 16, 0);
    for (int q = 0; q < nCells; ++q)
        for (int k = 0; k < nkk; ++k)
            for (int comp = 0; comp < 2; ++comp) {
                double x = gamma ? (comp == 0 ? 1.0 : 0.0) : phase[2 * ((size_t)q * nk + k) + comp];
                unsigned short *o = &out[(((size_t)q * nkk + k) * 2 + comp) * 8];
                for (int sl = 0; sl < kBlochSlices; ++sl) {
                    const double m = std::rint(x * std::ldexp(1.0, 7 + 8 * sl));
                    x -= m * std::ldexp(1.0, -7 - 8 * sl);   // exact
                    const float f = (float)m;
                    unsigned u;
                    std::memcpy(&u, &f, 4);
                    o[sl] = (unsigned short)(u >> 16);
                }
            }
}

}  // namespace olcao
