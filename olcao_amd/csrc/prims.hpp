// Primitive Gaussian integrals over unnormalised real solid harmonics, as
// device (and host) functions templated on the angular-momentum class.
//
// What is computed follows the reference's primitive library
// (src/uolcao/gaussIntegrals.f90): the Cartesian block pc(a,b) over the
// components s; x y z; xx yy zz xy xz yz; xyz xxy xxz yyx yyz zzx zzy xxx yyy
// zzz (:36-38) up to the class's l, transformed to the solid harmonics
// s; x y z; xy xz yz xx-yy 2zz-xx-yy; xyz xxz-yyz xxx-3yyx 3xxy-yyy
// 2zzz-3xxz-3yyz 4zzx-xxx-yyx 4zzy-xxy-yyy (:32-34).
//
// How differs from the reference: overlap and kinetic integrals factorise
// into x, y and z; each axis runs the 1-D Obara-Saika recursion (Obara &
// Saika, J. Chem. Phys. 84, 3963 (1986))
//   E(i+1,j) = XA E(i,j) + h (i E(i-1,j) + j E(i,j-1)),  E(0,0) = 1
// and the Cartesian block is pref * Ex * Ey * Ez.  The three-centre
// integrals use the McMurchie-Davidson form about the pair centre (below),
// which lets a lattice sum over many potential images accumulate a few
// Hermite quantities per pair.  Every table index is a template constant, so
// the tables live in registers.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include <utility>

#define OHD __host__ __device__ __forceinline__

namespace olcao {

constexpr double kPi = 3.141592653589793;

OHD constexpr int ncart(int l) { return (l + 1) * (l + 2) * (l + 3) / 6; }
OHD constexpr int nsh(int l) { return (l + 1) * (l + 1); }

// Cartesian component c -> lx + 4 ly + 16 lz (reference pc order).
OHD constexpr int cart_code(int c)
{
    return c == 0 ? 0 : c == 1 ? 1 : c == 2 ? 4 : c == 3 ? 16 : c == 4 ? 2 :
           c == 5 ? 8 : c == 6 ? 32 : c == 7 ? 5 : c == 8 ? 17 : c == 9 ? 20 :
           c == 10 ? 21 : c == 11 ? 6 : c == 12 ? 18 : c == 13 ? 9 :
           c == 14 ? 24 : c == 15 ? 33 : c == 16 ? 36 : c == 17 ? 3 :
           c == 18 ? 12 : 48;
}
OHD constexpr int cart_from_code(int code)
{
    return code == 0 ? 0 : code == 1 ? 1 : code == 4 ? 2 : code == 16 ? 3 :
           code == 2 ? 4 : code == 8 ? 5 : code == 32 ? 6 : code == 5 ? 7 :
           code == 17 ? 8 : code == 20 ? 9 : code == 21 ? 10 : code == 6 ? 11 :
           code == 18 ? 12 : code == 9 ? 13 : code == 24 ? 14 : code == 33 ? 15 :
           code == 36 ? 16 : code == 3 ? 17 : code == 12 ? 18 : code == 48 ? 19 : -1;
}
OHD constexpr int cexp(int c, int axis) { return (cart_code(c) >> (2 * axis)) & 3; }
OHD constexpr int ctot(int c) { return cexp(c, 0) + cexp(c, 1) + cexp(c, 2); }
OHD constexpr int cfirst(int c)
{
    return cexp(c, 0) > 0 ? 0 : cexp(c, 1) > 0 ? 1 : cexp(c, 2) > 0 ? 2 : -1;
}
// Component with one (k = 1) or two (k = 2) fewer units along axis.
OHD constexpr int cdec(int c, int axis, int k)
{
    return cexp(c, axis) >= k ? cart_from_code(cart_code(c) - k * (1 << (2 * axis))) : 0;
}

// Solid harmonic i as up to three weighted Cartesian components.
OHD constexpr int sh_nterm(int i)
{
    return (i == 7 || i == 10 || i == 11 || i == 12) ? 2 :
           (i == 8 || i == 13 || i == 14 || i == 15) ? 3 : 1;
}
OHD constexpr int sh_cart(int i, int t)
{
    return i == 0 ? 0 : i == 1 ? 1 : i == 2 ? 2 : i == 3 ? 3 : i == 4 ? 7 :
           i == 5 ? 8 : i == 6 ? 9 :
           i == 7 ? (t == 0 ? 4 : 5) :
           i == 8 ? (t == 0 ? 6 : t == 1 ? 4 : 5) :
           i == 9 ? 10 :
           i == 10 ? (t == 0 ? 12 : 14) :
           i == 11 ? (t == 0 ? 17 : 13) :
           i == 12 ? (t == 0 ? 11 : 18) :
           i == 13 ? (t == 0 ? 19 : t == 1 ? 12 : 14) :
           i == 14 ? (t == 0 ? 15 : t == 1 ? 17 : 13) :
                     (t == 0 ? 16 : t == 1 ? 11 : 18);
}
OHD constexpr double sh_weight(int i, int t)
{
    return i == 7 ? (t == 0 ? 1.0 : -1.0) :
           i == 8 ? (t == 0 ? 2.0 : -1.0) :
           i == 10 ? (t == 0 ? 1.0 : -1.0) :
           i == 11 ? (t == 0 ? 1.0 : -3.0) :
           i == 12 ? (t == 0 ? 3.0 : -1.0) :
           i == 13 ? (t == 0 ? 2.0 : -3.0) :
           i == 14 ? (t == 0 ? 4.0 : -1.0) :
           i == 15 ? (t == 0 ? 4.0 : -1.0) : 1.0;
}

// sh(i,j) = sum_ab T(i,a) pc(a,b) T(j,b) for the class block.
template <int LA, int LB>
OHD void cart_to_sh(const double (&pc)[ncart(LA)][ncart(LB)],
                    double (&sh)[nsh(LA)][nsh(LB)])
{
    constexpr int NB = ncart(LB), SA = nsh(LA), SB = nsh(LB);
    double tmp[SA][NB];
#pragma unroll
    for (int i = 0; i < SA; ++i)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            double s = 0.0;
#pragma unroll
            for (int t = 0; t < sh_nterm(i); ++t) s += sh_weight(i, t) * pc[sh_cart(i, t)][b];
            tmp[i][b] = s;
        }
#pragma unroll
    for (int i = 0; i < SA; ++i)
#pragma unroll
        for (int j = 0; j < SB; ++j) {
            double s = 0.0;
#pragma unroll
            for (int t = 0; t < sh_nterm(j); ++t) s += tmp[i][sh_cart(j, t)] * sh_weight(j, t);
            sh[i][j] = s;
        }
}

// 1-D overlap-type Obara-Saika table, E(0,0) = 1.
template <int LA, int LB>
OHD void os1d(double XA, double XB, double h, double (&E)[LA + 1][LB + 1])
{
    E[0][0] = 1.0;
#pragma unroll
    for (int i = 1; i <= LA; ++i) {
        double v = XA * E[i - 1][0];
        if (i > 1) v += h * (i - 1) * E[i > 1 ? i - 2 : 0][0];
        E[i][0] = v;
    }
#pragma unroll
    for (int j = 1; j <= LB; ++j)
#pragma unroll
        for (int i = 0; i <= LA; ++i) {
            double v = XB * E[i][j - 1];
            double t = 0.0;
            if (i > 0) t += i * E[i > 0 ? i - 1 : 0][j - 1];
            if (j > 1) t += (j - 1) * E[i][j > 1 ? j - 2 : 0];
            E[i][j] = v + h * t;
        }
}

// 1-D kinetic table from the Obara-Saika kinetic recursion,
// T(0,0) = xi (1 - 2 xi X^2) with X = A - B along the axis.
template <int LA, int LB>
OHD void ke1d(double XA, double XB, double h, double xi, double X, double a1,
              double a2, const double (&S)[LA + 1][LB + 1],
              double (&T)[LA + 1][LB + 1])
{
    T[0][0] = xi * (1.0 - 2.0 * xi * X * X);
#pragma unroll
    for (int i = 1; i <= LA; ++i) {
        double v = XA * T[i - 1][0];
        double s = S[i][0];
        if (i > 1) {
            v += h * (i - 1) * T[i > 1 ? i - 2 : 0][0];
            s -= (i - 1) / (2.0 * a1) * S[i > 1 ? i - 2 : 0][0];
        }
        T[i][0] = v + 2.0 * xi * s;
    }
#pragma unroll
    for (int j = 1; j <= LB; ++j)
#pragma unroll
        for (int i = 0; i <= LA; ++i) {
            double v = XB * T[i][j - 1];
            double t = 0.0;
            if (i > 0) t += i * T[i > 0 ? i - 1 : 0][j - 1];
            double s = S[i][j];
            if (j > 1) {
                t += (j - 1) * T[i][j > 1 ? j - 2 : 0];
                s -= (j - 1) / (2.0 * a2) * S[i][j > 1 ? j - 2 : 0];
            }
            T[i][j] = v + h * t + 2.0 * xi * s;
        }
}

// ---------------------------------------------------------------------
// Boys function F_0..F_M with the reference's numerical character
// (gaussIntegrals.f90:16080-16221): branch point float(0.1); series branch
// with Gamma(N+1/2+k) rounded to float; closed form
//   F_m = c_m [ sqrt(pi)/(4^m T^(m+1/2)) erf(sqrt T)
//               - exp(-T) sum_p (m-p+1)!/(4^(p-1) (2(m-p+1))! T^p) ],
//   c_m = (2m)!/(2 m!)  (c_1 = 1),  F_0 = sqrt(pi)/(2 sqrt T) erf(sqrt T).
OHD constexpr double gamma_half(int k)   // Gamma(k + 1/2)
{
    double v = 1.7724538509055160273;     // sqrt(pi)
    for (int i = 0; i < k; ++i) v *= (i + 0.5);
    return v;
}
OHD constexpr double gamma_half_f32(int k) { return (double)(float)gamma_half(k); }
OHD constexpr double factd(int k)
{
    double v = 1.0;
    for (int i = 2; i <= k; ++i) v *= i;
    return v;
}
OHD constexpr double pow4(int m)
{
    double v = 1.0;
    for (int i = 0; i < m; ++i) v *= 4.0;
    return v;
}
// x^p by binary powering.
OHD double ipow_sq(double x, int p)
{
    double r = 1.0, b = x;
    bool have = false;
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        if (p & 1) { r = have ? r * b : b; have = true; }
        p >>= 1;
        if (p) b = b * b;
    }
    return r;
}

template <int M>
OHD void boys(double T, double (&F)[M + 1])
{
    const double exp_T = exp(-T);
    const double sqrt_pi = sqrt(kPi);
    const double sqT = sqrt(T);
    const double erf_T = erf(sqT);
    if (T <= (double)(float)0.1) {
        double Tp[7];
        Tp[0] = 1.0;
        Tp[1] = T;
#pragma unroll
        for (int p = 2; p <= 6; ++p) Tp[p] = ipow_sq(T, p);
#pragma unroll
        for (int N = 0; N <= M; ++N) {
            double s = (1.0 / gamma_half_f32(N + 1)) + (T / gamma_half_f32(N + 2));
#pragma unroll
            for (int p = 2; p <= 6; ++p) s += Tp[p] / gamma_half_f32(N + p + 1);
            F[N] = 0.5 * exp_T * gamma_half_f32(N) * s;
        }
    } else {
        F[0] = 0.5 * ((sqrt_pi / sqT) * erf_T);
#pragma unroll
        for (int m = 1; m <= M; ++m) {
            double lead = (sqrt_pi / (pow4(m) * (ipow_sq(T, m) * sqT))) * erf_T;
            double sum = 0.0;
#pragma unroll
            for (int p = 1; p <= m; ++p) {
                double den = pow4(p - 1) * factd(2 * (m - p + 1));
                sum += factd(m - p + 1) / (den * ipow_sq(T, p));
            }
            double v = lead - exp_T * sum;
            F[m] = (m == 1) ? v : (factd(2 * m) / (2.0 * factd(m))) * v;
        }
    }
}

// ---------------------------------------------------------------------
// Values that depend only on the primitive pair (a1 at A, a2 at B).
struct PairConst {
    double a1, a2, zeta, xi, xid2;
    double A[3], B[3], P[3], S[3];   // S = a1 A + a2 B
};

OHD PairConst make_pair(double a1, double a2, const double *A, const double *B)
{
    PairConst q;
    q.a1 = a1;
    q.a2 = a2;
    q.zeta = a1 + a2;
    q.xi = a1 * a2 / q.zeta;
    double d2 = 0.0;
    for (int x = 0; x < 3; ++x) {
        q.A[x] = A[x];
        q.B[x] = B[x];
        q.S[x] = a1 * A[x] + a2 * B[x];
        q.P[x] = q.S[x] / q.zeta;
        double d = A[x] - B[x];
        d2 += d * d;
    }
    q.xid2 = q.xi * d2;
    return q;
}

// Cartesian overlap (KIN = false) or kinetic (KIN = true) block.
template <int LA, int LB, bool KIN>
OHD void two_center_cart(const PairConst &q, double (&pc)[ncart(LA)][ncart(LB)])
{
    constexpr int NA = ncart(LA), NB = ncart(LB);
    const double t = kPi / q.zeta;
    const double pref = t * sqrt(t) * exp(-q.xid2);
    const double h = 1.0 / (2.0 * q.zeta);
    double E[3][LA + 1][LB + 1];
    double Tk[3][LA + 1][LB + 1];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        os1d<LA, LB>(q.P[x] - q.A[x], q.P[x] - q.B[x], h, E[x]);
        if (KIN)
            ke1d<LA, LB>(q.P[x] - q.A[x], q.P[x] - q.B[x], h, q.xi, q.A[x] - q.B[x],
                         q.a1, q.a2, E[x], Tk[x]);
    }
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int ax = cexp(a, 0), ay = cexp(a, 1), az = cexp(a, 2);
            const int bx = cexp(b, 0), by = cexp(b, 1), bz = cexp(b, 2);
            double v;
            if (KIN)
                v = Tk[0][ax][bx] * E[1][ay][by] * E[2][az][bz] +
                    E[0][ax][bx] * Tk[1][ay][by] * E[2][az][bz] +
                    E[0][ax][bx] * E[1][ay][by] * Tk[2][az][bz];
            else
                v = E[0][ax][bx] * E[1][ay][by] * E[2][az][bz];
            pc[a][b] = pref * v;
        }
}

// The overlap (KIN = false) or kinetic (KIN = true) block straight in the
// solid harmonics: sh(i, j) = pref sum over the Cartesian terms of harmonics
// i and j of their weights times Ex Ey Ez (or the kinetic sum).  No
// Cartesian block is held, so the sweep kernels stay small.  The 1-D tables
// first, then the harmonic rows [I0, I1) handed to emit(i, j, value).
template <int LA, int LB, bool KIN>
struct TwoCenterTables {
    double pref;
    double E[3][LA + 1][LB + 1];
    double Tk[3][LA + 1][LB + 1];
};

template <int LA, int LB, bool KIN>
OHD void two_center_tables(const PairConst &q, TwoCenterTables<LA, LB, KIN> &tb)
{
    const double t = kPi / q.zeta;
    tb.pref = t * sqrt(t) * exp(-q.xid2);
    const double h = 1.0 / (2.0 * q.zeta);
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        os1d<LA, LB>(q.P[x] - q.A[x], q.P[x] - q.B[x], h, tb.E[x]);
        if (KIN)
            ke1d<LA, LB>(q.P[x] - q.A[x], q.P[x] - q.B[x], h, q.xi, q.A[x] - q.B[x], q.a1, q.a2,
                         tb.E[x], tb.Tk[x]);
    }
}

template <int LA, int LB, bool KIN, int I0, int I1, class F>
OHD void two_center_sh_rows(const TwoCenterTables<LA, LB, KIN> &tb, F &&emit)
{
    static_for<I1 - I0>([&](auto i_) __attribute__((always_inline)) {
        constexpr int i = I0 + decltype(i_)::value;
        static_for<nsh(LB)>([&](auto j_) __attribute__((always_inline)) {
            constexpr int j = decltype(j_)::value;
            double s = 0.0;
            static_for<sh_nterm(i)>([&](auto ti_) __attribute__((always_inline)) {
                constexpr int a = sh_cart(i, decltype(ti_)::value);
                constexpr double wa = sh_weight(i, decltype(ti_)::value);
                constexpr int ax = cexp(a, 0), ay = cexp(a, 1), az = cexp(a, 2);
                static_for<sh_nterm(j)>([&](auto tj_) __attribute__((always_inline)) {
                    constexpr int b = sh_cart(j, decltype(tj_)::value);
                    constexpr double wb = sh_weight(j, decltype(tj_)::value);
                    constexpr int bx = cexp(b, 0), by = cexp(b, 1), bz = cexp(b, 2);
                    double v;
                    if constexpr (KIN)
                        v = tb.Tk[0][ax][bx] * tb.E[1][ay][by] * tb.E[2][az][bz] +
                            tb.E[0][ax][bx] * tb.Tk[1][ay][by] * tb.E[2][az][bz] +
                            tb.E[0][ax][bx] * tb.E[1][ay][by] * tb.Tk[2][az][bz];
                    else
                        v = tb.E[0][ax][bx] * tb.E[1][ay][by] * tb.E[2][az][bz];
                    s += (wa * wb) * v;
                });
            });
            emit(i, j, tb.pref * s);
        });
    });
}

// Mass-velocity block (massvel2CIntg, gaussIntegrals.f90:39070):
//   (alpha_fs^2 / 8) <a| nabla^4 |b>,  alpha_fs = fineStructure * 1e-3,
// with nabla^4 = sum_i d_i^4 + 2 sum_{i<j} d_i^2 d_j^2 applied to the more
// diffuse primitive (<nabla^4 a|b> = <a|nabla^4 b>): on the tight one its
// terms are ~(a_tight/a_diffuse)^2 larger than the result and cancel.
// Along one axis, with g_j = (x - B)^j exp(-b (x - B)^2),
//   d^2 g_j = j(j-1) g_{j-2} - 2 b (2j+1) g_j + 4 b^2 g_{j+2},
// so <i|d^2|j> and <i|d^4|j> are combinations of the 1-D overlap table
// E(i, j') for j' <= j + 4.
constexpr double kFineStructure = 7.2973525693;   // src/constants.f90 (x 1e-3)

template <int N>
OHD double d2_1d(const double (&Ei)[N], int j, double b)
{
    double v = -2.0 * b * (2 * j + 1) * Ei[j] + 4.0 * b * b * Ei[j + 2];
    if (j >= 2) v += j * (j - 1) * Ei[j >= 2 ? j - 2 : 0];
    return v;
}

template <int N>
OHD double d4_1d(const double (&Ei)[N], int j, double b)
{
    double v = -2.0 * b * (2 * j + 1) * d2_1d(Ei, j, b) + 4.0 * b * b * d2_1d(Ei, j + 2, b);
    if (j >= 2) v += j * (j - 1) * d2_1d(Ei, j >= 2 ? j - 2 : 0, b);
    return v;
}

// S, d^2 and d^4 tables [i][j] (bra power i, ket power j) of one axis, the
// derivatives taken on the ket (KET) or on the bra.
template <int LA, int LB, bool KET>
OHD void mv_axis(double XPA, double XPB, double h, double e, double (&S)[LA + 1][LB + 1],
                 double (&D2)[LA + 1][LB + 1], double (&D4)[LA + 1][LB + 1])
{
    if (KET) {
        double E[LA + 1][LB + 5];
        os1d<LA, LB + 4>(XPA, XPB, h, E);
#pragma unroll
        for (int i = 0; i <= LA; ++i)
#pragma unroll
            for (int j = 0; j <= LB; ++j) {
                S[i][j] = E[i][j];
                D2[i][j] = d2_1d(E[i], j, e);
                D4[i][j] = d4_1d(E[i], j, e);
            }
    } else {
        double E[LB + 1][LA + 5];   // roles swapped: rows ket power, columns bra power
        os1d<LB, LA + 4>(XPB, XPA, h, E);
#pragma unroll
        for (int i = 0; i <= LA; ++i)
#pragma unroll
            for (int j = 0; j <= LB; ++j) {
                S[i][j] = E[j][i];
                D2[i][j] = d2_1d(E[j], i, e);
                D4[i][j] = d4_1d(E[j], i, e);
            }
    }
}

template <int LA, int LB>
OHD void massvel_cart(const PairConst &q, double (&pc)[ncart(LA)][ncart(LB)])
{
    constexpr int NA = ncart(LA), NB = ncart(LB);
    const double t = kPi / q.zeta;
    const double fs = kFineStructure * 0.001;
    const double pref = (fs * fs / 8.0) * (t * sqrt(t) * exp(-q.xid2));
    const double h = 1.0 / (2.0 * q.zeta);
    double S[3][LA + 1][LB + 1], D2[3][LA + 1][LB + 1], D4[3][LA + 1][LB + 1];
    if (q.a1 < q.a2) {
#pragma unroll
        for (int x = 0; x < 3; ++x)
            mv_axis<LA, LB, false>(q.P[x] - q.A[x], q.P[x] - q.B[x], h, q.a1, S[x], D2[x], D4[x]);
    } else {
#pragma unroll
        for (int x = 0; x < 3; ++x)
            mv_axis<LA, LB, true>(q.P[x] - q.A[x], q.P[x] - q.B[x], h, q.a2, S[x], D2[x], D4[x]);
    }
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int bb = 0; bb < NB; ++bb) {
            const int ax = cexp(a, 0), ay = cexp(a, 1), az = cexp(a, 2);
            const int bx = cexp(bb, 0), by = cexp(bb, 1), bz = cexp(bb, 2);
            const double sx = S[0][ax][bx], sy = S[1][ay][by], sz = S[2][az][bz];
            const double dx = D2[0][ax][bx], dy = D2[1][ay][by], dz = D2[2][az][bz];
            const double v = D4[0][ax][bx] * sy * sz + sx * D4[1][ay][by] * sz +
                             sx * sy * D4[2][az][bz] +
                             2.0 * (dx * dy * sz + dx * sy * dz + sx * dy * dz);
            pc[a][bb] = pref * v;
        }
}

// One Cartesian component `comp` of the dipole integrals <a| (r - C) |b>
// (DIP, dipole3CIntg, gaussIntegrals.f90:49404) or of the gradient
// integrals <a| d/dr |b> (momentum2CIntg, :23041).  Along the component's
// axis, g_i = (x-A)^i e^{-a (x-A)^2}, g_j = (x-B)^j e^{-b (x-B)^2}:
//   dipole    <g_i| x - C |g_j> = E(i, j+1) + (B - C) E(i, j)
//   gradient  <g_i| d/dx |g_j>  = j E(i, j-1) - 2 b E(i, j+1)
//                               = 2 a E(i+1, j) - i E(i-1, j)
// the derivative taken on the more diffuse primitive (on the tight one the
// two terms cancel); the overlap table on the other two axes.
template <int LA, int LB, bool DIP>
OHD void vec_cart(const PairConst &q, int comp, const double *C,
                  double (&pc)[ncart(LA)][ncart(LB)])
{
    constexpr int NA = ncart(LA), NB = ncart(LB);
    const double t = kPi / q.zeta;
    const double pref = t * sqrt(t) * exp(-q.xid2);
    const double h = 1.0 / (2.0 * q.zeta);
    const bool bra = !DIP && q.a1 < q.a2;
    // Ek: ket-extended table [i][j <= LB+1]; Eb: bra-extended [j][i <= LA+1]
    double Ek[3][LA + 1][LB + 2], Eb[3][LB + 1][LA + 2];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        if (bra)
            os1d<LB, LA + 1>(q.P[x] - q.B[x], q.P[x] - q.A[x], h, Eb[x]);
        else
            os1d<LA, LB + 1>(q.P[x] - q.A[x], q.P[x] - q.B[x], h, Ek[x]);
    }
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            double v = pref;
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const int i = cexp(a, x), j = cexp(b, x);
                const double e0 = bra ? Eb[x][j][i] : Ek[x][i][j];
                double f;
                if (x != comp) {
                    f = e0;
                } else if (DIP) {
                    f = Ek[x][i][j + 1] + (q.B[x] - C[x]) * e0;
                } else if (bra) {
                    f = 2.0 * q.a1 * Eb[x][j][i + 1];
                    if (i > 0) f -= i * Eb[x][j][i > 0 ? i - 1 : 0];
                } else {
                    f = -2.0 * q.a2 * Ek[x][i][j + 1];
                    if (j > 0) f += j * Ek[x][i][j > 0 ? j - 1 : 0];
                }
                v *= f;
            }
            pc[a][b] = v;
        }
}

// Ket-centre derivative d/dB_X of the kinetic and overlap blocks (the
// force integrals' dkinetic2CIntg, gaussIntegrals.f90:65512, and the
// momentum term of computeForceIntg).  With g_j = (x-B)^j e^{-b (x-B)^2},
// d/dB g_j = 2 b g_{j+1} - j g_{j-1}; the ket step of the 1-D recursions
// turns 2 b E(i,j+1) - j E(i,j-1) into terms of the (i, j) tables, free of
// the cancellation between the two when b >> a:
//   dE(i,j) = 2 b XPB E(i,j) + (b/zeta) i E(i-1,j) - (a/zeta) j E(i,j-1)
//   dT(i,j) = 2 b XPB T(i,j) + (b/zeta) i T(i-1,j) - (a/zeta) j T(i,j-1)
//             + 2 xi dE(i,j)
// pc = cKE d/dB_X <a| -nabla^2/2 |b> + cS d/dB_X <a|b>.
template <int LA, int LB>
OHD void dket_cart(const PairConst &q, int X, double cKE, double cS,
                   double (&pc)[ncart(LA)][ncart(LB)])
{
    constexpr int NA = ncart(LA), NB = ncart(LB);
    const double t = kPi / q.zeta;
    const double pref = t * sqrt(t) * exp(-q.xid2);
    const double h = 1.0 / (2.0 * q.zeta);
    const double bz = q.a2 / q.zeta, az = q.a1 / q.zeta;
    double E[3][LA + 1][LB + 1], Tk[3][LA + 1][LB + 1];
    double dE[3][LA + 1][LB + 1], dT[3][LA + 1][LB + 1];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        const double xpb = q.P[x] - q.B[x];
        os1d<LA, LB>(q.P[x] - q.A[x], xpb, h, E[x]);
        ke1d<LA, LB>(q.P[x] - q.A[x], xpb, h, q.xi, q.A[x] - q.B[x], q.a1, q.a2, E[x], Tk[x]);
#pragma unroll
        for (int i = 0; i <= LA; ++i)
#pragma unroll
            for (int j = 0; j <= LB; ++j) {
                double de = 2.0 * q.a2 * xpb * E[x][i][j];
                double dt = 2.0 * q.a2 * xpb * Tk[x][i][j];
                if (i > 0) {
                    de += bz * i * E[x][i > 0 ? i - 1 : 0][j];
                    dt += bz * i * Tk[x][i > 0 ? i - 1 : 0][j];
                }
                if (j > 0) {
                    de -= az * j * E[x][i][j > 0 ? j - 1 : 0];
                    dt -= az * j * Tk[x][i][j > 0 ? j - 1 : 0];
                }
                dE[x][i][j] = de;
                dT[x][i][j] = dt + 2.0 * q.xi * de;
            }
    }
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            double s = 1.0, k = 0.0;
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const int i = cexp(a, x), j = cexp(b, x);
                s *= x == X ? dE[x][i][j] : E[x][i][j];
            }
#pragma unroll
            for (int y = 0; y < 3; ++y) {
                double p = 1.0;
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const int i = cexp(a, x), j = cexp(b, x);
                    p *= x == y ? (x == X ? dT[x][i][j] : Tk[x][i][j])
                                : (x == X ? dE[x][i][j] : E[x][i][j]);
                }
                k += p;
            }
            pc[a][b] = pref * (cKE * k + cS * s);
        }
}

// ---------------------------------------------------------------------
// Three-centre integrals in McMurchie-Davidson form about the pair centre P.
//
// The pair product is expanded in Hermite Gaussians about P,
//   g1 g2 = K_AB sum_tuv Ex^{ax bx}_t Ey^{ay by}_u Ez^{az bz}_v Lambda_tuv(r; P, zeta),
// K_AB = exp(-xi |A-B|^2), so every three-centre integral of the pair is
//   <a|O_C|b> = K_AB sum_tuv E_t E_u E_v R_tuv,   R_tuv = d^t/dPx d^u/dPy d^v/dPz f(P - C)
// with f the integral of exp(-zeta|r-P|^2) against the operator at C.  R
// depends on the image C, E only on the pair; the kernels therefore sum
// W_tuv = sum_images,terms w R_tuv (35 numbers for a d-d pair instead of
// 100 Cartesian products) and contract with E once per pair.
//   three-Gaussian overlap (operator exp(-a3|r-C|^2)):
//     f = (pi/zeta3)^(3/2) exp(-mu X^2),  mu = zeta a3/zeta3,  X = P - C;
//     R_tuv = f h_t(X_x) h_u(X_y) h_v(X_z),
//     h_0 = 1, h_1 = -2 mu x, h_{t+1} = -2 mu x h_t - 2 mu t h_{t-1}.
//   Gaussian-damped nuclear attraction (operator exp(-a3 r_C^2)/r_C):
//     f = (2 pi/zeta3) exp(-mu X^2) F_0(U),  U = kappa X^2 = zeta3 |G-C|^2,
//     kappa = zeta^2/zeta3; with G_n = exp(-mu X^2) sum_k C(n,k) mu^(n-k)
//     kappa^k F_k(U) the derivatives follow the usual recursion
//     R^(n)_000 = (2 pi/zeta3) (-2)^n G_n,
//     R^(n)_{t+1,u,v} = t R^(n+1)_{t-1,u,v} + X_x R^(n+1)_{t,u,v}
//     (same along y, z), and R_tuv = R^(0)_tuv.  The Boys values F_k(U) are
//     the reference's (boys() above), at the reference's argument U.

OHD constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
// Index of (t,u,v), t+u+v <= L, in a packed array (grouped by total order).
OHD constexpr int herm_index(int t, int u, int v)
{
    const int n = t + u + v;
    // entries of lower total order, then position within order n
    const int lower = n * (n + 1) * (n + 2) / 6;
    const int rest = u + v;          // (n - t)
    return lower + rest * (rest + 1) / 2 + v;
}

// acc += w * R_tuv of the three-Gaussian overlap for one image/term.
template <int L>
OHD void acc_e3c_herm(const PairConst &q, double a3, const double *C, double w,
                      double (&W)[nherm(L)])
{
    const double z3 = q.zeta + a3;
    const double inv = 1.0 / z3;
    const double mu = q.zeta * a3 * inv;
    double X[3], X2 = 0.0;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        X[x] = q.P[x] - C[x];
        X2 += X[x] * X[x];
    }
    const double t = kPi * inv;
    const double pref = w * t * sqrt(t) * exp(-mu * X2);
    double h[3][L + 1];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        h[x][0] = 1.0;
        if (L >= 1) h[x][L >= 1 ? 1 : 0] = -2.0 * mu * X[x];
#pragma unroll
        for (int k = 1; k < L; ++k)
            h[x][k + 1] = -2.0 * mu * (X[x] * h[x][k] + k * h[x][k - 1]);
    }
#pragma unroll
    for (int tt = 0; tt <= L; ++tt)
#pragma unroll
        for (int u = 0; u <= L - tt; ++u) {
            const double hxy = pref * h[0][tt] * h[1][u];
#pragma unroll
            for (int v = 0; v <= L - tt - u; ++v) W[herm_index(tt, u, v)] += hxy * h[2][v];
        }
}

// As acc_e3c_herm with the term's pair constants precomputed:
// mu = zeta a3 / (zeta + a3), pref0 = w (pi/(zeta + a3))^(3/2) (the same
// operations in the same order, so the sums are identical).  MIX evaluates
// exp(-mu X^2) of the distant images, where it is below e^-kMixedExpArg =
// 1e-4, from the FP64 argument with a base-2 range reduction in FP64 and the
// hardware FP32 exponential of the fractional part (relative error ~1e-7,
// so at most ~1e-11 of the item's prefactor); nearer images, which carry
// the largest terms, keep the FP64 exponential.  The polynomial part and
// the accumulation stay in FP64.
constexpr double kMixedExpArg = 9.210340371976184;   // -ln(1e-4)
OHD double exp_neg_mixed(double arg)   // exp(-arg), arg >= 0
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = -arg * 1.4426950408889634;   // log2(e)
    const double n = floor(y);
    const float f = (float)(y - n);               // in [0, 1)
    return ldexp((double)__builtin_amdgcn_exp2f(f), (int)n);
#else
    return exp(-arg);
#endif
}

// exp(-arg), arg >= 0, in FP64 from a 256-entry table T[j] = 2^(j/256):
// N = rint(-arg 256/ln2), r = -arg - N ln2/256 (two-part constant, |r| <=
// ln2/512), exp(-arg) = 2^(N div 256) T[N mod 256] (1 + r + ... + r^5/120);
// the truncation error r^6/720 is below 1e-20 and the result is within a
// few ulp.  Arguments beyond the FP64 range give 0.
OHD double exp_neg_tab(double arg, const double *T)
{
    if (arg > 700.0) return 0.0;
    const double x = -arg;
    const double Nd = rint(x * 369.3299304675746);              // 256 / ln 2
    // ln2/256 = hi + lo, hi with 22 trailing zero bits so that N hi is exact
    const double r = fma(Nd, -1.6548911585192516e-12, fma(Nd, -0.002707606172407395, x));
    const int N = (int)Nd;
    const int j = N & 255;
    const int m = (N - j) / 256;
    double p = 1.0 / 120.0;
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(T[j] * p, m);
}

template <int L, bool MIX>
OHD void acc_e3c_herm_tab(const PairConst &q, double mu, double pref0, const double *C,
                          const double *expTab, double (&W)[nherm(L)])
{
    double X[3], X2 = 0.0;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        X[x] = q.P[x] - C[x];
        X2 += X[x] * X[x];
    }
    const double arg = mu * X2;
    const double pref = pref0 * (MIX && arg > kMixedExpArg ? exp_neg_mixed(arg)
                                                             : exp_neg_tab(arg, expTab));
    double h[3][L + 1];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        h[x][0] = 1.0;
        if (L >= 1) h[x][L >= 1 ? 1 : 0] = -2.0 * mu * X[x];
#pragma unroll
        for (int k = 1; k < L; ++k)
            h[x][k + 1] = -2.0 * mu * (X[x] * h[x][k] + k * h[x][k - 1]);
    }
#pragma unroll
    for (int tt = 0; tt <= L; ++tt)
#pragma unroll
        for (int u = 0; u <= L - tt; ++u) {
            const double hxy = pref * h[0][tt] * h[1][u];
#pragma unroll
            for (int v = 0; v <= L - tt - u; ++v) W[herm_index(tt, u, v)] += hxy * h[2][v];
        }
}

// acc += w * R_tuv of the Gaussian-damped nuclear attraction.
template <int L>
OHD void acc_nuc_herm(const PairConst &q, double a3, const double *C, double w,
                      double (&W)[nherm(L)])
{
    const double z3 = q.zeta + a3;
    const double inv = 1.0 / z3;
    const double mu = q.zeta * a3 * inv;
    const double kap = q.zeta * q.zeta * inv;
    double X[3], X2 = 0.0;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        X[x] = q.P[x] - C[x];
        X2 += X[x] * X[x];
    }
    double F[L + 1];
    boys<L>(kap * X2, F);
    const double pre = w * 2.0 * (kPi * inv) * exp(-mu * X2);
    // R^(n)_000 = pre (-2)^n G_n
    double Rn[L + 1][nherm(L)];
    {
        double mup[L + 1], kp[L + 1];
        mup[0] = 1.0;
        kp[0] = 1.0;
#pragma unroll
        for (int k = 1; k <= L; ++k) {
            mup[k] = mup[k - 1] * mu;
            kp[k] = kp[k - 1] * kap;
        }
        double sgn = 1.0;
#pragma unroll
        for (int n = 0; n <= L; ++n) {
            double g = 0.0;
            double binom = 1.0;
#pragma unroll
            for (int k = 0; k <= n; ++k) {
                g += binom * mup[n - k] * kp[k] * F[k];
                binom = binom * (n - k) / (k + 1);
            }
            Rn[n][0] = pre * sgn * g;
            sgn = -2.0 * sgn;
        }
    }
#pragma unroll
    for (int n = L - 1; n >= 0; --n)
#pragma unroll
        for (int tt = 0; tt <= L - n; ++tt)
#pragma unroll
            for (int u = 0; u <= L - n - tt; ++u)
#pragma unroll
                for (int v = 0; v <= L - n - tt - u; ++v) {
                    if (tt + u + v == 0) continue;
                    double val;
                    if (tt > 0) {
                        val = X[0] * Rn[n + 1][herm_index(tt - 1, u, v)];
                        if (tt > 1) val += (tt - 1) * Rn[n + 1][herm_index(tt > 1 ? tt - 2 : 0, u, v)];
                    } else if (u > 0) {
                        val = X[1] * Rn[n + 1][herm_index(tt, u - 1, v)];
                        if (u > 1) val += (u - 1) * Rn[n + 1][herm_index(tt, u > 1 ? u - 2 : 0, v)];
                    } else {
                        val = X[2] * Rn[n + 1][herm_index(tt, u, v - 1)];
                        if (v > 1) val += (v - 1) * Rn[n + 1][herm_index(tt, u, v > 1 ? v - 2 : 0)];
                    }
                    Rn[n][herm_index(tt, u, v)] = val;
                }
#pragma unroll
    for (int e = 0; e < nherm(L); ++e) W[e] += Rn[0][e];
}

// Hermite expansion coefficients of one axis about P:
// E^{i+1,j}_t = h E^{ij}_{t-1} + XPA E^{ij}_t + (t+1) E^{ij}_{t+1}, and the
// analogue along j with XPB; E^{00}_0 = 1, h = 1/(2 zeta).
template <int LA, int LB>
OHD void herm_e(double XPA, double XPB, double h, double (&E)[LA + 1][LB + 1][LA + LB + 1])
{
    constexpr int L = LA + LB;
#pragma unroll
    for (int i = 0; i <= LA; ++i)
#pragma unroll
        for (int j = 0; j <= LB; ++j)
#pragma unroll
            for (int t = 0; t <= L; ++t) E[i][j][t] = 0.0;
    E[0][0][0] = 1.0;
#pragma unroll
    for (int i = 0; i <= LA; ++i)
#pragma unroll
        for (int j = 0; j <= LB; ++j) {
            if (i == 0 && j == 0) continue;
            const int pi = j > 0 ? i : i - 1, pj = j > 0 ? j - 1 : j;
            const double X = j > 0 ? XPB : XPA;
#pragma unroll
            for (int t = 0; t <= i + j; ++t) {
                double v = X * E[pi][pj][t];
                if (t > 0) v += h * E[pi][pj][t > 0 ? t - 1 : 0];
                if (t + 1 <= L) v += (t + 1) * E[pi][pj][t + 1 <= L ? t + 1 : L];
                E[i][j][t] = v;
            }
        }
}

// k-overlap block (KOverlap2CIntg, gaussIntegrals.f90:129408), the block of
// exp(-i kq . r) between the Gaussians, in McMurchie-Davidson form: with the
// pair expanded in Hermite Gaussians about P, each axis integrates to
//   sum_t E^{ij}_t (d/dP)^t [sqrt(pi/zeta) exp(-q^2/(4 zeta)) exp(-i q P)]
//   = sqrt(pi/zeta) exp(-q^2/(4 zeta)) exp(-i q P) sum_t E^{ij}_t (-i q)^t,
// so the block is (pi/zeta)^(3/2) exp(-xi d^2 - |q|^2/(4 zeta)) exp(-i q.P)
// times the product of the three complex axis sums.  pc receives the real
// part (part 0) or minus the imaginary part (part 1).
template <int LA, int LB>
OHD void kov_cart(const PairConst &q, const double *kq, int part,
                  double (&pc)[ncart(LA)][ncart(LB)])
{
    constexpr int NA = ncart(LA), NB = ncart(LB), L = LA + LB;
    const double h = 1.0 / (2.0 * q.zeta);
    double Gr[3][LA + 1][LB + 1], Gi[3][LA + 1][LB + 1];
    double q2 = 0.0, qP = 0.0;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        double E[LA + 1][LB + 1][L + 1];
        herm_e<LA, LB>(q.P[x] - q.A[x], q.P[x] - q.B[x], h, E);
        const double qx = kq[x];
        q2 += qx * qx;
        qP += qx * q.P[x];
        // (-i q)^t: real parts t = 0, 2, 4, 6 (1, -q^2, q^4, -q^6), imaginary
        // parts t = 1, 3, 5 (-q, q^3, -q^5)
        double pr[L + 1], pi[L + 1];
        double pw = 1.0;
#pragma unroll
        for (int t = 0; t <= L; ++t) {
            const int ph = t & 3;
            pr[t] = ph == 0 ? pw : ph == 2 ? -pw : 0.0;
            pi[t] = ph == 1 ? -pw : ph == 3 ? pw : 0.0;
            pw *= qx;
        }
#pragma unroll
        for (int i = 0; i <= LA; ++i)
#pragma unroll
            for (int j = 0; j <= LB; ++j) {
                double sr = 0.0, si = 0.0;
#pragma unroll
                for (int t = 0; t <= i + j; ++t) {
                    sr += E[i][j][t] * pr[t];
                    si += E[i][j][t] * pi[t];
                }
                Gr[x][i][j] = sr;
                Gi[x][i][j] = si;
            }
    }
    const double t = kPi / q.zeta;
    const double amp = t * sqrt(t) * exp(-q.xid2 - q2 * 0.25 / q.zeta);
    const double fr = amp * cos(qP), fi = -amp * sin(qP);
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            double vr = fr, vi = fi;
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const int i = cexp(a, x), j = cexp(b, x);
                const double ur = Gr[x][i][j], ui = Gi[x][i][j];
                const double nr = vr * ur - vi * ui, ni = vr * ui + vi * ur;
                vr = nr;
                vi = ni;
            }
            pc[a][b] = part == 0 ? vr : -vi;
        }
}

// Cartesian component (a, b) of K_AB sum_tuv E E E W.
template <int LA, int LB>
OHD double herm_contract(const double (&E)[3][LA + 1][LB + 1][LA + LB + 1], const double *W,
                         int ax, int ay, int az, int bx, int by, int bz)
{
    double s = 0.0;
    for (int t = 0; t <= ax + bx; ++t)
        for (int u = 0; u <= ay + by; ++u) {
            const double exy = E[0][ax][bx][t] * E[1][ay][by][u];
            for (int v = 0; v <= az + bz; ++v) s += exy * E[2][az][bz][v] * W[herm_index(t, u, v)];
        }
    return s;
}

// Compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1, so that
// every table index inside f is a constant and the tables stay in registers
// whatever the unroller decides.
template <class F, int... I>
OHD void static_for_impl(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
OHD void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// herm_contract with the Cartesian exponents as template constants.
template <int LA, int LB, int AX, int AY, int AZ, int BX, int BY, int BZ>
OHD double herm_contract_ct(const double (&E)[3][LA + 1][LB + 1][LA + LB + 1],
                            const double (&W)[nherm(LA + LB)])
{
    double s = 0.0;
    static_for<AX + BX + 1>([&](auto t_) __attribute__((always_inline)) {
        constexpr int t = decltype(t_)::value;
        static_for<AY + BY + 1>([&](auto u_) __attribute__((always_inline)) {
            constexpr int u = decltype(u_)::value;
            const double exy = E[0][AX][BX][t] * E[1][AY][BY][u];
            double z = 0.0;
            static_for<AZ + BZ + 1>([&](auto v_) __attribute__((always_inline)) {
                constexpr int v = decltype(v_)::value;
                z += E[2][AZ][BZ][v] * W[herm_index(t, u, v)];
            });
            s += exy * z;
        });
    });
    return s;
}

// Ket-centre derivative of the Hermite expansion along one axis:
//   d/dB [K_AB sum_t E^{ij}_t Lambda_t] = K_AB sum_t D^{ij}_t Lambda_t,
//   D^{ij}_t = 2b E^{i,j+1}_t - j E^{i,j-1}_t
//            = 2b XPB E^{ij}_t + (b/zeta) E^{ij}_{t-1} + (b/zeta) i E^{i-1,j}_t
//              - (a/zeta) j E^{i,j-1}_t,
// by the ket step E^{i,j+1}_t = XPB E_t + E_{t-1}/(2 zeta) + (t+1) E_{t+1}
// and E^{ij}_{t+1} = (i E^{i-1,j}_t + j E^{i,j-1}_t) / (2 zeta (t+1)); t runs
// to la + lb + 1, so the Hermite sums W are needed one order higher.
template <int LA, int LB>
OHD void herm_dket(const double (&E)[LA + 1][LB + 1][LA + LB + 1], double XPB, double a,
                   double b, double zeta, double (&D)[LA + 1][LB + 1][LA + LB + 2])
{
    constexpr int L = LA + LB;
    const double bz = b / zeta, az = a / zeta;
#pragma unroll
    for (int i = 0; i <= LA; ++i)
#pragma unroll
        for (int j = 0; j <= LB; ++j)
#pragma unroll
            for (int t = 0; t <= L + 1; ++t) {
                double v = t <= L ? 2.0 * b * XPB * E[i][j][t <= L ? t : 0] : 0.0;
                if (t > 0) v += bz * E[i][j][t > 0 ? t - 1 : 0];
                if (i > 0 && t <= L) v += bz * i * E[i > 0 ? i - 1 : 0][j][t <= L ? t : 0];
                if (j > 0 && t <= L) v -= az * j * E[i][j > 0 ? j - 1 : 0][t <= L ? t : 0];
                D[i][j][t] = v;
            }
}

// herm_contract_ct with axis X's coefficients replaced by its derivative
// table DX; W holds the sums up to order la + lb + 1.
template <int LA, int LB, int X, int AX, int AY, int AZ, int BX, int BY, int BZ>
OHD double herm_dcontract_ct(const double (&E)[3][LA + 1][LB + 1][LA + LB + 1],
                             const double (&DX)[LA + 1][LB + 1][LA + LB + 2],
                             const double (&W)[nherm(LA + LB + 1)])
{
    constexpr int NT = AX + BX + 1 + (X == 0), NU = AY + BY + 1 + (X == 1),
                  NV = AZ + BZ + 1 + (X == 2);
    double s = 0.0;
    static_for<NT>([&](auto t_) __attribute__((always_inline)) {
        constexpr int t = decltype(t_)::value;
        const double ex = X == 0 ? DX[AX][BX][t] : E[0][AX][BX][X == 0 ? 0 : t];
        static_for<NU>([&](auto u_) __attribute__((always_inline)) {
            constexpr int u = decltype(u_)::value;
            const double exy = ex * (X == 1 ? DX[AY][BY][u] : E[1][AY][BY][X == 1 ? 0 : u]);
            double z = 0.0;
            static_for<NV>([&](auto v_) __attribute__((always_inline)) {
                constexpr int v = decltype(v_)::value;
                z += (X == 2 ? DX[AZ][BZ][v] : E[2][AZ][BZ][X == 2 ? 0 : v]) * W[herm_index(t, u, v)];
            });
            s += exy * z;
        });
    });
    return s;
}

}  // namespace olcao
