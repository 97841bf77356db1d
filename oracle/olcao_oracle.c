/* CPU oracle for the OLCAO integral -> H(k)/S(k) path.
 *
 * TEST INFRASTRUCTURE ONLY (see olcao_oracle.h).  A plain-C restatement of
 * the reference algorithm, written from the reference's behaviour:
 *   primitives      src/uolcao/gaussIntegrals.f90 (Obara-Saika recursions,
 *                   overlap2CIntg :26, kinetic2CIntg :4422,
 *                   electron3CIntg :11624, boys :16080, nuclear3CIntg :16223)
 *   screening       src/uolcao/gaussRelations.f90:64-278
 *   lattice         src/uolcao/lattice.f90:239-443 (cell list),
 *                   :469-781 (getNumCells/makeLattice/cellBorderCheck),
 *                   :784-872 (findLatticeVector; restated as an exhaustive
 *                   nearest-lattice-point search)
 *   drivers         src/uolcao/integrals.F90:151-652 (OL), :656 (KE),
 *                   :1531-1942 + :2533-2697 (NP), :1946-2354 + :2700-2856
 *                   (EP), :2860-3605 (Hamiltonian, PSCF)
 *   contraction,    src/uolcao/intgSaving.F90:56-89 (multWithBasisFn1),
 *   Bloch sum,      :132-189 (applyPhaseFactors), :241-292
 *   scatter         (kPointLatticeOriginShift), :346-655 (saveCurrentPair)
 *   phase factors   src/uolcao/kpoints.f90:459-489
 * The primitive recursions are written generically over Cartesian
 * components (the reference unrolls them per angular-momentum class); the
 * results agree to rounding, which the golden vectors from the compiled
 * reference primitives pin.
 */
#define _GNU_SOURCE
#include "olcao_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.141592653589793
#define SMALL_THRESH 1.0e-8

/* ------------------------------------------------------------------ */
/* Cartesian components in the reference's pc(20,20) order
 * (gaussIntegrals.f90:36-38): s; x y z; xx yy zz xy xz yz;
 * xyz xxy xxz yyx yyz zzx zzy xxx yyy zzz. */
static const int CART[20][3] = {
    {0,0,0},
    {1,0,0},{0,1,0},{0,0,1},
    {2,0,0},{0,2,0},{0,0,2},{1,1,0},{1,0,1},{0,1,1},
    {1,1,1},{2,1,0},{2,0,1},{1,2,0},{0,2,1},{1,0,2},{0,1,2},
    {3,0,0},{0,3,0},{0,0,3}};
static const int NCART[4] = {1, 4, 10, 20};  /* components up to l */
static const int NSH[4] = {1, 4, 9, 16};     /* solid harmonics up to l */

static int cart_index(int lx, int ly, int lz)
{
    for (int c = 0; c < 20; c++)
        if (CART[c][0] == lx && CART[c][1] == ly && CART[c][2] == lz) return c;
    return -1;
}

/* Unnormalised real solid harmonics in terms of Cartesian components
 * (gaussIntegrals.f90:32-34): 1 s; x y z; xy xz yz xx-yy 2zz-xx-yy;
 * xyz xxz-yyz xxx-3yyx 3xxy-yyy 2zzz-3xxz-3yyz 4zzx-xxx-yyx 4zzy-xxy-yyy. */
static double SHT[16][20];
static int dec[20][3];    /* index of component minus one unit along axis */
static int dec2[20][3];   /* minus two units */
static int tables_ready = 0;

static void init_tables(void)
{
    if (tables_ready) return;
    memset(SHT, 0, sizeof(SHT));
    SHT[0][0] = 1;
    SHT[1][1] = 1; SHT[2][2] = 1; SHT[3][3] = 1;
    SHT[4][7] = 1; SHT[5][8] = 1; SHT[6][9] = 1;
    SHT[7][4] = 1; SHT[7][5] = -1;
    SHT[8][6] = 2; SHT[8][4] = -1; SHT[8][5] = -1;
    SHT[9][10] = 1;
    SHT[10][12] = 1; SHT[10][14] = -1;
    SHT[11][17] = 1; SHT[11][13] = -3;
    SHT[12][11] = 3; SHT[12][18] = -1;
    SHT[13][19] = 2; SHT[13][12] = -3; SHT[13][14] = -3;
    SHT[14][15] = 4; SHT[14][17] = -1; SHT[14][13] = -1;
    SHT[15][16] = 4; SHT[15][11] = -1; SHT[15][18] = -1;
    for (int c = 0; c < 20; c++)
        for (int i = 0; i < 3; i++) {
            int l[3] = {CART[c][0], CART[c][1], CART[c][2]};
            l[i] -= 1;
            dec[c][i] = (l[i] >= 0) ? cart_index(l[0], l[1], l[2]) : -1;
            l[i] -= 1;
            dec2[c][i] = (l[i] >= 0) ? cart_index(l[0], l[1], l[2]) : -1;
        }
    tables_ready = 1;
}

static int first_axis(int c)
{
    for (int i = 0; i < 3; i++)
        if (CART[c][i] > 0) return i;
    return -1;
}

static void decode_switch(int sw, int *la, int *lb)
{
    int lo = sw & 15, hi = (sw >> 4) & 15;
    *la = (lo == 1) ? 0 : (lo == 2) ? 1 : (lo == 4) ? 2 : 3;
    *lb = (hi == 1) ? 0 : (hi == 2) ? 1 : (hi == 4) ? 2 : 3;
}

/* sh = T pc T^T restricted to the defined block; sh column-major. */
static void cart_to_sh(int la, int lb, double pc[20][20], double *sh)
{
    double tmp[16][20];
    int na = NCART[la], nb = NCART[lb];
    for (int i = 0; i < NSH[la]; i++)
        for (int b = 0; b < nb; b++) {
            double s = 0.0;
            for (int a = 0; a < na; a++)
                if (SHT[i][a] != 0.0) s += SHT[i][a] * pc[a][b];
            tmp[i][b] = s;
        }
    for (int j = 0; j < NSH[lb]; j++)
        for (int i = 0; i < NSH[la]; i++) {
            double s = 0.0;
            for (int b = 0; b < nb; b++)
                if (SHT[j][b] != 0.0) s += tmp[i][b] * SHT[j][b];
            sh[i + 16 * j] = s;
        }
}

/* ------------------------------------------------------------------ */
/* Boys function F_0..F_6, restating gaussIntegrals.f90:16080-16221
 * including its numerical character:
 *   - the branch point is the default-real literal 10E-2, i.e. float(0.1);
 *   - the series branch uses Gamma(N+1/2+k) constants that the reference
 *     writes as default-real literals, i.e. rounded to float;
 *   - the closed-form branch keeps the reference's grouping
 *     F_m = c_m [ sqrt(pi)/(4^m T^(m+1/2)) erf(sqrt T)
 *                 - exp(-T) sum_p (m-p+1)! / (4^(p-1) (2(m-p+1))! T^p) ],
 *     c_m = (2m)!/(2 m!), F_0 = 0.5 sqrt(pi)/T^(1/2) erf(sqrt T).  */
static double boys_gamma_f32[14];
static double boys_fact[16];
static int boys_ready = 0;

static void init_boys(void)
{
    if (boys_ready) return;
    for (int k = 0; k < 14; k++)
        boys_gamma_f32[k] = (double)(float)tgamma(k + 0.5);
    boys_fact[0] = 1.0;
    for (int k = 1; k < 16; k++) boys_fact[k] = boys_fact[k - 1] * k;
    boys_ready = 1;
}

/* Integer power by binary powering (the evaluation the reference's
 * compiler uses for T**p; it reproduces the golden values more closely
 * than left-to-right multiplication). */
static double ipow_sq(double x, int p)
{
    double r = 0.0, b = x;
    int have = 0;
    while (p > 0) {
        if (p & 1) { r = have ? r * b : b; have = 1; }
        p >>= 1;
        if (p) b = b * b;
    }
    return r;
}

void orc_boys(double T, double F[7])
{
    init_boys();
    double exp_T = exp(-T);
    double sqrt_pi = sqrt(PI);
    double erf_T = erf(sqrt(T));
    if (T <= (double)(float)0.1) {
        for (int N = 0; N <= 6; N++) {
            const double *g = boys_gamma_f32 + N;
            double s = (1 / g[1]) + (T / g[2]);
            for (int p = 2; p <= 6; p++) s += ipow_sq(T, p) / g[p + 1];
            F[N] = 0.5 * exp_T * g[0] * s;
        }
    } else {
        F[0] = 0.5 * ((sqrt_pi / pow(T, 0.5)) * erf_T);
        for (int m = 1; m <= 6; m++) {
            double cm = boys_fact[2 * m] / (2.0 * boys_fact[m]);
            double pow4 = 1.0;
            for (int k = 0; k < m; k++) pow4 *= 4.0;
            double lead = (sqrt_pi / (pow4 * pow(T, m + 0.5))) * erf_T;
            double sum = 0.0, q = 1.0;
            for (int p = 1; p <= m; p++) {
                double Tp = ipow_sq(T, p);
                double num = boys_fact[m - p + 1];
                double den = q * boys_fact[2 * (m - p + 1)];
                sum += num / (den * Tp);
                q *= 4.0;
            }
            double v = lead - exp_T * sum;
            F[m] = (m == 1) ? v : cm * v;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Two-index Obara-Saika overlap-type recursion over Cartesian components:
 *   (a|b+1i) = XB_i (a|b) + h [N_i(a)(a-1i|b) + N_i(b)(a|b-1i)]
 *   (a+1i|0) = XA_i (a|0) + h N_i(a) (a-1i|0)
 * with (0|0) = pref.  Used for overlap (X=P, h=1/2zeta) and for the
 * three-Gaussian overlap (X=G, h=1/2zeta3). */
static void os_overlap(int la, int lb, double pref, const double XA[3],
                       const double XB[3], double h, double pc[20][20])
{
    int na = NCART[la], nb = NCART[lb];
    for (int b = 0; b < nb; b++)
        for (int a = 0; a < na; a++) {
            double v;
            if (b == 0 && a == 0) {
                v = pref;
            } else if (b != 0) {
                int i = first_axis(b);
                int b1 = dec[b][i];
                v = XB[i] * pc[a][b1];
                double t = 0.0;
                if (CART[a][i] > 0) t += CART[a][i] * pc[dec[a][i]][b1];
                if (CART[b][i] > 1) t += (CART[b][i] - 1) * pc[a][dec2[b][i]];
                v += h * t;
            } else {
                int i = first_axis(a);
                v = XA[i] * pc[dec[a][i]][0];
                if (CART[a][i] > 1) v += h * (CART[a][i] - 1) * pc[dec2[a][i]][0];
            }
            pc[a][b] = v;
        }
}

/* Kinetic energy, Obara-Saika (1986) kinetic recursion:
 *   T(a|b+1i) = PB_i T(a|b) + h[N_i(a)T(a-1i|b) + N_i(b)T(a|b-1i)]
 *               + 2 xi [S(a|b+1i) - N_i(b)/(2 a2) S(a|b-1i)]
 * (and the analogue along a), T(0|0) = xi (3 - 2 xi d^2) S(0|0). */
static void os_kinetic(int la, int lb, double a1, double a2, double xi,
                       double d2, const double PA[3], const double PB[3],
                       double h, double S[20][20], double T[20][20])
{
    int na = NCART[la], nb = NCART[lb];
    for (int b = 0; b < nb; b++)
        for (int a = 0; a < na; a++) {
            double v;
            if (b == 0 && a == 0) {
                v = xi * (3 - 2 * xi * d2) * S[0][0];
            } else if (b != 0) {
                int i = first_axis(b);
                int b1 = dec[b][i];
                v = PB[i] * T[a][b1];
                double t = 0.0;
                if (CART[a][i] > 0) t += CART[a][i] * T[dec[a][i]][b1];
                if (CART[b][i] > 1) t += (CART[b][i] - 1) * T[a][dec2[b][i]];
                v += h * t;
                double s = S[a][b];
                if (CART[b][i] > 1)
                    s -= (CART[b][i] - 1) / (2.0 * a2) * S[a][dec2[b][i]];
                v += 2 * xi * s;
            } else {
                int i = first_axis(a);
                v = PA[i] * T[dec[a][i]][0];
                double s = S[a][0];
                if (CART[a][i] > 1) {
                    v += h * (CART[a][i] - 1) * T[dec2[a][i]][0];
                    s -= (CART[a][i] - 1) / (2.0 * a1) * S[dec2[a][i]][0];
                }
                v += 2 * xi * s;
            }
            T[a][b] = v;
        }
}

/* Gaussian-damped nuclear attraction <g1| exp(-a3 r_C^2)/r_C |g2>,
 * Obara-Saika with auxiliary index m:
 *   th^m(a|b+1i) = GB_i th^m(a|b) - GC_i th^(m+1)(a|b)
 *       + h3 [N_i(a)(th^m - th^(m+1))(a-1i|b) + N_i(b)(th^m - th^(m+1))(a|b-1i)]
 * th^m(0|0) = F_m(U) 2 pi/zeta3 exp(-(a1a2 d12^2 + a1a3 d13^2 + a2a3 d23^2)/zeta3),
 * U = zeta3 |G - C|^2 (nuclear3CIntg :16255-16271). */
static void os_nuclear_m(int la, int lb, int L, const double pf[7], const double GA[3],
                         const double GB[3], const double GC[3], double h,
                         double th[7][20][20])
{
    int na = NCART[la], nb = NCART[lb];
    int tot[20];
    for (int c = 0; c < 20; c++) tot[c] = CART[c][0] + CART[c][1] + CART[c][2];
    for (int b = 0; b < nb; b++)
        for (int a = 0; a < na; a++) {
            int mmax = L - tot[a] - tot[b];
            for (int m = 0; m <= mmax; m++) {
                double v;
                if (b == 0 && a == 0) {
                    v = pf[m];
                } else if (b != 0) {
                    int i = first_axis(b);
                    int b1 = dec[b][i];
                    v = GB[i] * th[m][a][b1] - GC[i] * th[m + 1][a][b1];
                    double t = 0.0;
                    if (CART[a][i] > 0) {
                        int a1 = dec[a][i];
                        t += CART[a][i] * (th[m][a1][b1] - th[m + 1][a1][b1]);
                    }
                    if (CART[b][i] > 1) {
                        int b2 = dec2[b][i];
                        t += (CART[b][i] - 1) * (th[m][a][b2] - th[m + 1][a][b2]);
                    }
                    v += h * t;
                } else {
                    int i = first_axis(a);
                    int a1 = dec[a][i];
                    v = GA[i] * th[m][a1][0] - GC[i] * th[m + 1][a1][0];
                    if (CART[a][i] > 1) {
                        int a2 = dec2[a][i];
                        v += h * (CART[a][i] - 1) * (th[m][a2][0] - th[m + 1][a2][0]);
                    }
                }
                th[m][a][b] = v;
            }
        }
}

/* th^m for m <= la + lb - |a| - |b| (what the integral needs). */
static void os_nuclear(int la, int lb, const double pf[7], const double GA[3],
                       const double GB[3], const double GC[3], double h,
                       double th[7][20][20])
{
    os_nuclear_m(la, lb, la + lb, pf, GA, GB, GC, h, th);
}

/* Mass-velocity block (massvel2CIntg, gaussIntegrals.f90:39070):
 * (alpha_fs^2/8) <a|nabla^4|b>, alpha_fs = fineStructure * 1e-3
 * (src/constants.f90:47), with nabla^4 = sum_i d_i^4 + 2 sum_{i<j} d_i^2 d_j^2.
 * The operator is applied to the more diffuse primitive (<nabla^4 a|b> when
 * a1 < a2, else <a|nabla^4 b>; equal by integration by parts): on the tight
 * side its terms are ~(a_tight/a_diffuse)^2 larger than the result and
 * cancel.  One 1-D Cartesian Gaussian g_i = (x-A)^i exp(-a (x-A)^2) has
 *   d^2 g_i = i(i-1) g_{i-2} - 2a(2i+1) g_i + 4a^2 g_{i+2},
 * so the 1-D derivative integrals are combinations of the 1-D overlap table
 * extended by four along the differentiated side. */
#define MVMAX 8
static void os1d_table(int la, int lb, double XA, double XB, double h,
                       double E[MVMAX][MVMAX])
{
    E[0][0] = 1.0;
    for (int i = 1; i <= la; i++)
        E[i][0] = XA * E[i - 1][0] + (i > 1 ? h * (i - 1) * E[i - 2][0] : 0.0);
    for (int j = 1; j <= lb; j++)
        for (int i = 0; i <= la; i++) {
            double t = 0.0;
            if (i > 0) t += i * E[i - 1][j - 1];
            if (j > 1) t += (j - 1) * E[i][j - 2];
            E[i][j] = XB * E[i][j - 1] + h * t;
        }
}

/* <g_i | d^2 | g_j> with the derivative on the ket (exponent b), from E */
static double d2_ket(double E[MVMAX][MVMAX], int i, int j, double b)
{
    double v = -2.0 * b * (2 * j + 1) * E[i][j] + 4.0 * b * b * E[i][j + 2];
    if (j >= 2) v += j * (j - 1) * E[i][j - 2];
    return v;
}

static double d4_ket(double E[MVMAX][MVMAX], int i, int j, double b)
{
    double v = -2.0 * b * (2 * j + 1) * d2_ket(E, i, j, b) +
               4.0 * b * b * d2_ket(E, i, j + 2, b);
    if (j >= 2) v += j * (j - 1) * d2_ket(E, i, j - 2, b);
    return v;
}

static void os_massvel(int la, int lb, double a1, double a2, double pref,
                       const double PA[3], const double PB[3], double h,
                       double pc[20][20])
{
    static const double fs = 7.2973525693 * 0.001;
    const int bra = a1 < a2;   /* differentiate the bra (transposed tables) */
    double E[3][MVMAX][MVMAX];
    /* table T[p][q]: p the undifferentiated side's power, q the other's */
    for (int x = 0; x < 3; x++) {
        if (bra)
            os1d_table(lb, la + 4, PB[x], PA[x], h, E[x]);
        else
            os1d_table(la, lb + 4, PA[x], PB[x], h, E[x]);
    }
    const double e = bra ? a1 : a2;
    int na = NCART[la], nb = NCART[lb];
    for (int a = 0; a < na; a++)
        for (int b = 0; b < nb; b++) {
            double S[3], D2[3], D4[3];
            for (int x = 0; x < 3; x++) {
                int p = bra ? CART[b][x] : CART[a][x];
                int q = bra ? CART[a][x] : CART[b][x];
                S[x] = E[x][p][q];
                D2[x] = d2_ket(E[x], p, q, e);
                D4[x] = d4_ket(E[x], p, q, e);
            }
            double v = D4[0] * S[1] * S[2] + S[0] * D4[1] * S[2] + S[0] * S[1] * D4[2] +
                       2.0 * (D2[0] * D2[1] * S[2] + D2[0] * S[1] * D2[2] + S[0] * D2[1] * D2[2]);
            pc[a][b] = (fs * fs / 8.0) * pref * v;
        }
}

/* Dipole (op 5, dipole3CIntg, gaussIntegrals.f90:49404) and momentum
 * (op 6, momentum2CIntg, :23041) integrals, three Cartesian components each.
 * Along the component's axis, with g_i = (x-A)^i exp(-a (x-A)^2) and the
 * ket g_j = (x-B)^j exp(-b (x-B)^2):
 *   dipole    <g_i| x - C |g_j> = E(i, j+1) + (B - C) E(i, j)
 *   momentum  <g_i| d/dx  |g_j> = j E(i, j-1) - 2 b E(i, j+1)
 *                               = -(i E(i-1, j) - 2 a E(i+1, j))
 * (integration by parts), the derivative taken on the more diffuse
 * primitive: on the tight one the two terms are larger than the result and
 * cancel.  The overlap table E on the other two axes, extended by one
 * along the side that needs it. */
static void os_vec3(int op, int la, int lb, double a1, double a2, double pref,
                    const double PA[3], const double PB[3], const double BC[3], double h,
                    double pc[3][20][20])
{
    double E[3][MVMAX][MVMAX];
    const int bra = op == 6 && a1 < a2;   /* momentum: differentiate the bra */
    for (int x = 0; x < 3; x++) {
        if (bra)
            os1d_table(la + 1, lb, PA[x], PB[x], h, E[x]);
        else
            os1d_table(la, lb + 1, PA[x], PB[x], h, E[x]);
    }
    int na = NCART[la], nb = NCART[lb];
    for (int a = 0; a < na; a++)
        for (int b = 0; b < nb; b++)
            for (int c = 0; c < 3; c++) {
                double v = pref;
                for (int x = 0; x < 3; x++) {
                    int i = CART[a][x], j = CART[b][x];
                    if (x != c)
                        v *= E[x][i][j];
                    else if (op == 5)
                        v *= E[x][i][j + 1] + BC[x] * E[x][i][j];
                    else if (bra)
                        v *= 2.0 * a1 * E[x][i + 1][j] - (i > 0 ? i * E[x][i - 1][j] : 0.0);
                    else
                        v *= (j > 0 ? j * E[x][i][j - 1] : 0.0) - 2.0 * a2 * E[x][i][j + 1];
                }
                pc[c][a][b] = v;
            }
}

void orc_prim3(int op, double a1, double a2, const double *A, const double *B, const double *C,
               int sw, double *sh)
{
    init_tables();
    int la, lb;
    decode_switch(sw, &la, &lb);
    double zeta = a1 + a2, xi = a1 * a2 / zeta, d2 = 0.0;
    double PA[3], PB[3], BC[3];
    for (int x = 0; x < 3; x++) {
        double P = (a1 * A[x] + a2 * B[x]) / zeta;
        PA[x] = P - A[x];
        PB[x] = P - B[x];
        BC[x] = B[x] - C[x];
        d2 += (A[x] - B[x]) * (A[x] - B[x]);
    }
    static __thread double pc[3][20][20];
    os_vec3(op, la, lb, a1, a2, pow(PI / zeta, 1.5) * exp(-xi * d2), PA, PB, BC,
            1.0 / (2.0 * zeta), pc);
    for (int c = 0; c < 3; c++) cart_to_sh(la, lb, pc[c], sh + 256 * c);
}

/* KOverlap2CIntg (gaussIntegrals.f90:129408): the block of exp(-i dK . r)
 * between the two Gaussians, complex.  Along each axis, with y = x - P,
 * (x-A)^i (x-B)^j = sum_{k,l} C(i,k) C(j,l) PA^(i-k) PB^(j-l) y^(k+l), and the
 * moments of y against exp(-zeta y^2 - i q y) are, as the reference writes
 * them, m_n = s0 H_n(r) (i / (2 sqrt zeta))^n with r = -q / (2 sqrt zeta), H_n
 * the Hermite polynomials and s0 = sqrt(pi/zeta) exp(-xi d^2) exp(-q^2 /
 * (4 zeta)) exp(-i q P) (its preFactorko).  The Cartesian block is the
 * product of the three axes; real and imaginary parts go through the
 * solid-harmonic transform separately. */
void orc_primk(double a1, double a2, const double *A, const double *B, const double *dK, int sw,
               double *shr, double *shi)
{
    init_tables();
    int la, lb;
    decode_switch(sw, &la, &lb);
    double zeta = a1 + a2, xi = a1 * a2 / zeta;
    /* per axis: I[x][i][j] complex, i <= la, j <= lb */
    double Ir[3][4][4], Ii[3][4][4];
    for (int x = 0; x < 3; x++) {
        double P = (a1 * A[x] + a2 * B[x]) / zeta, PA = P - A[x], PB = P - B[x];
        double d = A[x] - B[x], q = dK[x];
        double amp = sqrt(PI / zeta) * exp(-xi * d * d) * exp(-q * q / (4.0 * zeta));
        double s0r = amp * cos(q * P), s0i = -amp * sin(q * P);
        double r = -q / (2.0 * sqrt(zeta)), g = 1.0 / (2.0 * sqrt(zeta));
        double H[7];
        H[0] = 1.0;
        H[1] = 2.0 * r;
        for (int m = 1; m < 6; m++) H[m + 1] = 2.0 * r * H[m] - 2.0 * m * H[m - 1];
        /* t_n = H_n (i g)^n: i^n cycles 1, i, -1, -i */
        double tr[7], ti[7];
        for (int m = 0; m <= 6; m++) {
            double mag = H[m] * pow(g, m);
            int ph = m & 3;
            tr[m] = ph == 0 ? mag : ph == 2 ? -mag : 0.0;
            ti[m] = ph == 1 ? mag : ph == 3 ? -mag : 0.0;
        }
        static const double BIN[4][4] = {{1, 0, 0, 0}, {1, 1, 0, 0}, {1, 2, 1, 0}, {1, 3, 3, 1}};
        for (int i = 0; i <= la; i++)
            for (int j = 0; j <= lb; j++) {
                double sr = 0.0, si = 0.0;
                for (int k = 0; k <= i; k++)
                    for (int l = 0; l <= j; l++) {
                        double c = BIN[i][k] * BIN[j][l] * pow(PA, i - k) * pow(PB, j - l);
                        sr += c * tr[k + l];
                        si += c * ti[k + l];
                    }
                Ir[x][i][j] = s0r * sr - s0i * si;
                Ii[x][i][j] = s0r * si + s0i * sr;
            }
    }
    double pr[20][20], pim[20][20];
    for (int a = 0; a < NCART[la]; a++)
        for (int b = 0; b < NCART[lb]; b++) {
            double vr = 1.0, vi = 0.0;
            for (int x = 0; x < 3; x++) {
                double ur = Ir[x][CART[a][x]][CART[b][x]], ui = Ii[x][CART[a][x]][CART[b][x]];
                double nr = vr * ur - vi * ui, ni = vr * ui + vi * ur;
                vr = nr;
                vi = ni;
            }
            pr[a][b] = vr;
            pim[a][b] = vi;
        }
    for (int q = 0; q < 256; q++) shr[q] = shi[q] = 0.0;
    cart_to_sh(la, lb, pr, shr);
    cart_to_sh(la, lb, pim, shi);
}

void orc_primk_batch(int n, const double *a, const double *A, const double *B, const double *dK,
                     const int *sw, double *shr, double *shi)
{
    init_tables();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
        orc_primk(a[2 * i], a[2 * i + 1], A + 3 * i, B + 3 * i, dK + 3 * i, sw[i],
                  shr + (long)256 * i, shi + (long)256 * i);
}

void orc_prim3_batch(int op, int n, const double *a, const double *A, const double *B,
                     const double *C, const int *sw, double *sh)
{
    init_tables();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
        orc_prim3(op, a[2 * i], a[2 * i + 1], A + 3 * i, B + 3 * i, C + 3 * i, sw[i],
                  sh + (long)768 * i);
}

void orc_prim(int op, double a1, double a2, double a3, const double *A,
              const double *B, const double *C, int sw, double *sh)
{
    init_tables();
    int la, lb;
    decode_switch(sw, &la, &lb);
    double zeta = a1 + a2;
    double xi = a1 * a2 / zeta;
    double P[3], PA[3], PB[3], d[3];
    double d2 = 0.0;
    for (int x = 0; x < 3; x++) {
        P[x] = (a1 * A[x] + a2 * B[x]) / zeta;
        PA[x] = P[x] - A[x];
        PB[x] = P[x] - B[x];
        d[x] = A[x] - B[x];
        d2 += d[x] * d[x];
    }
    static __thread double pc[20][20], pk[20][20], th[7][20][20];
    if (op == 4) {
        double pref = pow(PI / zeta, 1.5) * exp(-xi * d2);
        os_massvel(la, lb, a1, a2, pref, PA, PB, 1.0 / (2.0 * zeta), pc);
        cart_to_sh(la, lb, pc, sh);
        return;
    }
    if (op == 0 || op == 1) {
        double pref = pow(PI / zeta, 1.5) * exp(-xi * d2);
        double h = 1.0 / (2.0 * zeta);
        os_overlap(la, lb, pref, PA, PB, h, pc);
        if (op == 0) {
            cart_to_sh(la, lb, pc, sh);
        } else {
            os_kinetic(la, lb, a1, a2, xi, d2, PA, PB, h, pc, pk);
            cart_to_sh(la, lb, pk, sh);
        }
        return;
    }
    double z3 = zeta + a3;
    double h3 = 1.0 / (2.0 * z3);
    double G[3], GA[3], GB[3], GC[3], PCv[3];
    double pc2 = 0.0, gc2 = 0.0;
    for (int x = 0; x < 3; x++) {
        G[x] = (a1 * A[x] + a2 * B[x] + a3 * C[x]) / z3;
        GA[x] = G[x] - A[x];
        GB[x] = G[x] - B[x];
        GC[x] = G[x] - C[x];
        PCv[x] = P[x] - C[x];
        pc2 += PCv[x] * PCv[x];
        gc2 += GC[x] * GC[x];
    }
    if (op == 2) {
        double pref = pow(PI / z3, 1.5) * exp(-xi * d2 - (zeta * a3 / z3) * pc2);
        os_overlap(la, lb, pref, GA, GB, h3, pc);
        cart_to_sh(la, lb, pc, sh);
        return;
    }
    /* op == 3: nuclear */
    double F[7], pf[7];
    double U = z3 * gc2;
    orc_boys(U, F);
    double e = 0.0;
    for (int x = 0; x < 3; x++) {
        double d1 = A[x] - B[x], dd2 = A[x] - C[x], dd3 = B[x] - C[x];
        e += a1 * a2 * (d1 * d1) + a1 * a3 * (dd2 * dd2) + a2 * a3 * (dd3 * dd3);
    }
    double ex = exp(-e / z3);
    for (int m = 0; m < 7; m++) pf[m] = F[m] * 2.0 * (PI / z3) * ex;
    os_nuclear(la, lb, pf, GA, GB, GC, h3, th);
    cart_to_sh(la, lb, th[0], sh);
}

void orc_prim_batch(int op, int n, const double *a, const double *A,
                    const double *B, const double *C, const int *sw, double *sh)
{
    init_tables();
    init_boys();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double *o = sh + 256 * (long)i;
        memset(o, 0, 256 * sizeof(double));
        orc_prim(op, a[3 * i], a[3 * i + 1], a[3 * i + 2], A + 3 * i, B + 3 * i,
                 C + 3 * i, sw[i], o);
    }
}

/* Force-derivative primitives of computeForceIntg (forces.F90:60): the
 * derivative with respect to the ket centre B, C held fixed, of the kinetic
 * block (op 10, dkinetic2CIntg, gaussIntegrals.f90:65512), the
 * Gaussian-damped nuclear block (op 11, dnuclear3CIntgCB, :88076) and the
 * three-Gaussian overlap block (op 12, delectron3CIntgCB, :112289);
 * sh: [3][16][16], one column-major block per Cartesian component of B.
 * With g_b = (r-B)^b exp(-beta |r-B|^2), d/dB_i g_b = 2 beta g_{b+1i} -
 * N_i(b) g_{b-1i}.  The ket-raising step of each Obara-Saika recursion
 * turns this into terms of the (a|b) table itself, without the cancellation
 * of 2 beta (a|b+1i) against N_i(b) (a|b-1i) when beta >> alpha:
 *   overlap type (X = P, h = 1/2zeta; or X = G, h = 1/2zeta3):
 *     d_i(a|b) = 2 beta XB_i (a|b) + 2 beta h N_i(a) (a-1i|b)
 *                - (1 - 2 beta h) N_i(b) (a|b-1i)
 *   kinetic: the same recursion on T, plus 2 xi d_i S(a|b);
 *   nuclear, auxiliary index m:
 *     d_i(a|b) = 2 beta [GB_i th0(a|b) - GC_i th1(a|b)]
 *                + 2 beta h3 N_i(a) (th0 - th1)(a-1i|b)
 *                - (1 - 2 beta h3) N_i(b) th0(a|b-1i) - 2 beta h3 N_i(b) th1(a|b-1i)
 * with 1 - 2 beta h written as (the other exponents) 2h.  Returns 0, or 1
 * for a nuclear block with la + lb = 6 (it needs F_7; the reference's Boys
 * function, restated here, stops at F_6). */
int orc_primd(int op, double a1, double a2, double a3, const double *A, const double *B,
              const double *C, int sw, double *sh)
{
    init_tables();
    int la, lb;
    decode_switch(sw, &la, &lb);
    memset(sh, 0, 768 * sizeof(double));
    if (op == 11 && la + lb + 1 > 6) return 1;
    double zeta = a1 + a2;
    double xi = a1 * a2 / zeta;
    double P[3], PA[3], PB[3];
    double d2 = 0.0;
    for (int x = 0; x < 3; x++) {
        P[x] = (a1 * A[x] + a2 * B[x]) / zeta;
        PA[x] = P[x] - A[x];
        PB[x] = P[x] - B[x];
        d2 += (A[x] - B[x]) * (A[x] - B[x]);
    }
    static __thread double pc[20][20], pk[20][20], th[7][20][20], dp[20][20];
    int na = NCART[la], nb = NCART[lb];
    if (op == 10) {
        double pref = pow(PI / zeta, 1.5) * exp(-xi * d2);
        double h = 1.0 / (2.0 * zeta);
        os_overlap(la, lb, pref, PA, PB, h, pc);
        os_kinetic(la, lb, a1, a2, xi, d2, PA, PB, h, pc, pk);
        double bh = 2.0 * a2 * h, rest = 2.0 * a1 * h;
        for (int X = 0; X < 3; X++) {
            for (int b = 0; b < nb; b++)
                for (int a = 0; a < na; a++) {
                    int Na = CART[a][X], Nb = CART[b][X];
                    double dS = 2.0 * a2 * PB[X] * pc[a][b];
                    double dT = 2.0 * a2 * PB[X] * pk[a][b];
                    if (Na) {
                        dS += bh * Na * pc[dec[a][X]][b];
                        dT += bh * Na * pk[dec[a][X]][b];
                    }
                    if (Nb) {
                        dS -= rest * Nb * pc[a][dec[b][X]];
                        dT -= rest * Nb * pk[a][dec[b][X]];
                    }
                    dp[a][b] = dT + 2.0 * xi * dS;
                }
            cart_to_sh(la, lb, dp, sh + 256 * X);
        }
        return 0;
    }
    double z3 = zeta + a3;
    double h3 = 1.0 / (2.0 * z3);
    double G[3], GA[3], GB[3], GC[3];
    double pc2 = 0.0, gc2 = 0.0;
    for (int x = 0; x < 3; x++) {
        G[x] = (a1 * A[x] + a2 * B[x] + a3 * C[x]) / z3;
        GA[x] = G[x] - A[x];
        GB[x] = G[x] - B[x];
        GC[x] = G[x] - C[x];
        pc2 += (P[x] - C[x]) * (P[x] - C[x]);
        gc2 += GC[x] * GC[x];
    }
    double bh = 2.0 * a2 * h3, rest = 2.0 * (a1 + a3) * h3;
    if (op == 12) {
        double pref = pow(PI / z3, 1.5) * exp(-xi * d2 - (zeta * a3 / z3) * pc2);
        os_overlap(la, lb, pref, GA, GB, h3, pc);
        for (int X = 0; X < 3; X++) {
            for (int b = 0; b < nb; b++)
                for (int a = 0; a < na; a++) {
                    int Na = CART[a][X], Nb = CART[b][X];
                    double v = 2.0 * a2 * GB[X] * pc[a][b];
                    if (Na) v += bh * Na * pc[dec[a][X]][b];
                    if (Nb) v -= rest * Nb * pc[a][dec[b][X]];
                    dp[a][b] = v;
                }
            cart_to_sh(la, lb, dp, sh + 256 * X);
        }
        return 0;
    }
    /* op 11: nuclear, th^m up to m = la + lb + 1 */
    double F[7], pf[7];
    orc_boys(z3 * gc2, F);
    double e = 0.0;
    for (int x = 0; x < 3; x++) {
        double dd1 = A[x] - B[x], dd2 = A[x] - C[x], dd3 = B[x] - C[x];
        e += a1 * a2 * (dd1 * dd1) + a1 * a3 * (dd2 * dd2) + a2 * a3 * (dd3 * dd3);
    }
    double ex = exp(-e / z3);
    for (int m = 0; m < 7; m++) pf[m] = F[m] * 2.0 * (PI / z3) * ex;
    os_nuclear_m(la, lb, la + lb + 1, pf, GA, GB, GC, h3, th);
    for (int X = 0; X < 3; X++) {
        for (int b = 0; b < nb; b++)
            for (int a = 0; a < na; a++) {
                int Na = CART[a][X], Nb = CART[b][X];
                double v = 2.0 * a2 * (GB[X] * th[0][a][b] - GC[X] * th[1][a][b]);
                if (Na) {
                    int a1i = dec[a][X];
                    v += bh * Na * (th[0][a1i][b] - th[1][a1i][b]);
                }
                if (Nb) {
                    int b1i = dec[b][X];
                    v -= rest * Nb * th[0][a][b1i] + bh * Nb * th[1][a][b1i];
                }
                dp[a][b] = v;
            }
        cart_to_sh(la, lb, dp, sh + 256 * X);
    }
    return 0;
}

int orc_primd_batch(int op, int n, const double *a, const double *A, const double *B,
                    const double *C, const int *sw, double *sh)
{
    init_tables();
    init_boys();
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int i = 0; i < n; i++)
        bad |= orc_primd(op, a[3 * i], a[3 * i + 1], a[3 * i + 2], A + 3 * i, B + 3 * i, C + 3 * i,
                         sw[i], sh + 768 * (long)i);
    return bad;
}

/* ------------------------------------------------------------------ */
/* Lattice: lattice.f90:310-443 (negligibility radius), :469-781. */
double orc_neglig_limit(const double *vec, double tau, double minA,
                        double minPot)
{
    double nl = tau * fmax(2.0 / minA, 1.0 / minPot + 0.5 / minA);
    double maxSq = 0.0;
    for (int i = 0; i < 3; i++) {
        double s = 0.0;
        for (int x = 0; x < 3; x++) s += vec[3 * i + x] * vec[3 * i + x];
        maxSq = fmax(s, maxSq);
    }
    nl = nl + 0.75 * maxSq + sqrt(3.0 * maxSq * nl);
    nl = fmax(nl, 2.0 * tau / minA);
    return nl;
}

static double sumsq3(const double *v)
{
    return v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
}

/* cellBorderCheck: accept a cell whose lattice point lies outside the
 * sphere if one of seven box corners reached by adding the x, y, z
 * components of a1+a2+a3 falls inside (lattice.f90:698-781). */
static int border_check(const double *vec, const double *t, double nl)
{
    double sh[3] = {0, 0, 0}, c[3] = {t[0], t[1], t[2]};
    for (int i = 0; i < 3; i++)
        for (int x = 0; x < 3; x++) sh[x] += vec[3 * i + x];
    c[0] += sh[0]; if (sumsq3(c) < nl) return 1;
    c[1] += sh[1]; if (sumsq3(c) < nl) return 1;
    c[0] -= sh[0]; if (sumsq3(c) < nl) return 1;
    c[2] += sh[2]; if (sumsq3(c) < nl) return 1;
    c[0] += sh[0]; if (sumsq3(c) < nl) return 1;
    c[1] -= sh[1]; if (sumsq3(c) < nl) return 1;
    c[0] -= sh[0]; if (sumsq3(c) < nl) return 1;
    return 0;
}

typedef struct { double size; int idx; } cell_key;

static void merge_sort_keys(cell_key *k, cell_key *tmp, int n)
{
    if (n < 2) return;
    int h = n / 2;
    merge_sort_keys(k, tmp, h);
    merge_sort_keys(k + h, tmp, n - h);
    int i = 0, j = h, o = 0;
    while (i < h && j < n) tmp[o++] = (k[j].size < k[i].size) ? k[j++] : k[i++];
    while (i < h) tmp[o++] = k[i++];
    while (j < n) tmp[o++] = k[j++];
    memcpy(k, tmp, n * sizeof(cell_key));
}

int orc_make_lattice(const double *vec, double nl, int maxCells,
                     double *cellDims, double *cellSizes)
{
    int reps[3];
    for (int i = 0; i < 3; i++)
        reps[i] = 2 * (int)ceil(sqrt(nl) / sqrt(sumsq3(vec + 3 * i))) + 1;
    int cap = reps[0] * reps[1] * reps[2];
    double *dims = malloc(sizeof(double) * 3 * cap);
    cell_key *keys = malloc(sizeof(cell_key) * cap);
    cell_key *tmp = malloc(sizeof(cell_key) * cap);
    int n = 0;
    for (int i = 1; i <= reps[0]; i++)
        for (int j = 1; j <= reps[1]; j++)
            for (int k = 1; k <= reps[2]; k++) {
                int rep[3] = {i - (1 + reps[0]) / 2, j - (1 + reps[1]) / 2,
                              k - (1 + reps[2]) / 2};
                double t[3] = {0, 0, 0};
                for (int l = 0; l < 3; l++)
                    for (int m = 0; m < 3; m++) t[m] += vec[3 * l + m] * rep[l];
                double ss = sumsq3(t);
                if (ss > nl && !border_check(vec, t, nl)) continue;
                dims[3 * n] = t[0]; dims[3 * n + 1] = t[1]; dims[3 * n + 2] = t[2];
                keys[n].size = ss;
                keys[n].idx = n;
                n++;
            }
    merge_sort_keys(keys, tmp, n);
    int ret = n;
    if (n > maxCells) {
        ret = -n;
    } else {
        for (int c = 0; c < n; c++) {
            cellSizes[c] = keys[c].size;
            memcpy(cellDims + 3 * c, dims + 3 * keys[c].idx, 3 * sizeof(double));
        }
    }
    free(dims); free(keys); free(tmp);
    return ret;
}

/* Lattice vector nearest to v.  The reference rounds the fractional
 * coordinates and then consults precomputed Wigner-Seitz tables
 * (lattice.f90:784-872); here the same nearest point is found by rounding
 * and an exhaustive check of the 26 neighbouring lattice points (ties keep
 * the rounded point).  Which of two equidistant points is chosen does not
 * change any lattice sum: the summed image set is the same. */
static void inverse3(const double *vec, double inv[9])
{
    /* inv[3*i + x]: fractional coordinate i = sum_x v_x inv[3*i+x] */
    const double *a = vec, *b = vec + 3, *c = vec + 6;
    double bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
    double ca[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
    double ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    double vol = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
    for (int x = 0; x < 3; x++) {
        inv[x] = bc[x] / vol;
        inv[3 + x] = ca[x] / vol;
        inv[6 + x] = ab[x] / vol;
    }
}

static void nearest_with_inv(const double *vec, const double *inv,
                             const double *v, double *out)
{
    double base[3] = {0, 0, 0};
    for (int i = 0; i < 3; i++) {
        double f = v[0] * inv[3 * i] + v[1] * inv[3 * i + 1] + v[2] * inv[3 * i + 2];
        double r = (double)(long)(fabs(f) + 0.5);
        if (f < 0.0) r = -r;
        for (int x = 0; x < 3; x++) base[x] += vec[3 * i + x] * r;
    }
    double rem[3] = {v[0] - base[0], v[1] - base[1], v[2] - base[2]};
    double best = sumsq3(rem);
    double bestL[3] = {0, 0, 0};
    for (int i = -1; i <= 1; i++)
        for (int j = -1; j <= 1; j++)
            for (int k = -1; k <= 1; k++) {
                if (!i && !j && !k) continue;
                double L[3];
                for (int x = 0; x < 3; x++)
                    L[x] = i * vec[x] + j * vec[3 + x] + k * vec[6 + x];
                double r[3] = {rem[0] - L[0], rem[1] - L[1], rem[2] - L[2]};
                double s = sumsq3(r);
                if (s < best * (1.0 - 1e-14)) {
                    best = s;
                    memcpy(bestL, L, sizeof(bestL));
                }
            }
    for (int x = 0; x < 3; x++) out[x] = base[x] + bestL[x];
}

void orc_nearest_lattice_vector(const double *vec, const double *v, double *out)
{
    double inv[9];
    inverse3(vec, inv);
    nearest_with_inv(vec, inv, v, out);
}

/* ------------------------------------------------------------------ */
/* Integral drivers. */
static long long cnt_pc, cnt_pp, cnt_tr;
static int sample_stride = 1, sample_offset = 0;
/* Test-only factor on the three-centre image range (threeAlphaDist and its
 * nuclear and fused-sweep twins); 1 everywhere except in the test that
 * shows the brute-force lattice sums detect a range that drops images. */
static double three_range_scale = 1.0;

void orc_set_three_center_range_scale(double f) { three_range_scale = f > 0.0 ? f : 1.0; }

void orc_set_pair_sample(int stride, int offset)
{
    sample_stride = stride > 0 ? stride : 1;
    sample_offset = offset;
}

void orc_last_counts(long long *pc, long long *pp, long long *tr)
{
    *pc = cnt_pc; *pp = cnt_pp; *tr = cnt_tr;
}

typedef struct {
    const orc_sys *s;
    double inv[9];
    int valeDim, coreDim;
    int *cumulVale, *cumulCore;
    int potDim;
    int *potCumul;            /* [nPotTypes] first global term index */
    const double **elemAlpha; /* per type: alphas of first type of element */
    double *phase;            /* [nCells][nk][2] */
} drv_ctx;

static int pow2idx(int lmAlphaIndex)
{
    return lmAlphaIndex == 1 ? 1 : lmAlphaIndex == 4 ? 2 : lmAlphaIndex == 9 ? 4 : 8;
}

/* alphaDist(l,k,j,i) = tau / a_k(i) / alphaCenter, alphaCenter =
 * a_l(j)/(a_l(j)+a_k(i)) (gaussRelations.f90:119-130). */
static double alpha_center(double a1, double a2) { return a1 / (a1 + a2); }
static double alpha_dist(double tau, double a1, double a2)
{
    return tau / a2 / alpha_center(a1, a2);
}

typedef struct {
    int op, term;
    const double *coef;
    long long ntr;
} op_args;

/* One alpha pair's block for the selected operator (16x16 column-major). */
static void pair_block(const drv_ctx *c, const op_args *o, int ti, int tj,
                       int ia, int ib, double s2, const double *A,
                       const double *B, double *blk, int *contrib)
{
    const orc_sys *s = c->s;
    double a1 = s->alpha[ti * s->maxAlphas + ia];
    double a2 = s->alpha[tj * s->maxAlphas + ib];
    int sw = pow2idx(s->lmAlphaIndex[ti * s->maxAlphas + ia]) +
             16 * pow2idx(s->lmAlphaIndex[tj * s->maxAlphas + ib]);
    int n1 = s->lmAlphaIndex[ti * s->maxAlphas + ia];
    int n2 = s->lmAlphaIndex[tj * s->maxAlphas + ib];
    double tau = s->logThresh;
    /* screening uses the element-level alphas (gaussRelations.f90:89-99) */
    double ea1 = c->elemAlpha[ti][ia], ea2 = c->elemAlpha[tj][ib];
    double one[256];
    *contrib = 1;
    if (o->op == ORC_OL || o->op == ORC_KE) {
        orc_prim(o->op == ORC_OL ? 0 : 1, a1, a2, 0.0, A, B, A, sw, blk);
        return;
    }
    if (o->op == ORC_DM || o->op == ORC_MM) {
        /* one Cartesian component (term) of the dipole about coef[0..2]
         * (gaussOverlapDM, integrals3Terms.F90:64) or of the gradient
         * (gaussOverlapMM, :511) */
        double three[768];
        orc_prim3(o->op == ORC_DM ? 5 : 6, a1, a2, A, B, o->op == ORC_DM ? o->coef : A, sw, three);
        for (int q = 0; q < 256; q++) blk[q] = three[256 * o->term + q];
        return;
    }
    if (o->op == ORC_KO) {
        /* k-overlap (gaussKOverlap, integrals3Terms.F90:956) for the deltaK in
         * coef: even term the real part of the block, odd term minus its
         * imaginary part, whose Bloch sum takes the factor -i, so that the
         * two sweeps add to the complex matrix */
        double im[256];
        orc_primk(a1, a2, A, B, o->coef, sw, blk, im);
        if (o->term & 1)
            for (int q = 0; q < 256; q++) blk[q] = -im[q];
        return;
    }
    if (o->op == ORC_MV) {   /* accumulated with a minus sign, integrals.F90:1434-1447 */
        orc_prim(4, a1, a2, 0.0, A, B, A, sw, blk);
        for (int q = 0; q < 256; q++) blk[q] = -blk[q];
        return;
    }
    for (int q = 0; q < 256; q++) blk[q] = 0.0;
    double acen = alpha_center(ea1, ea2);
    double P[3];
    for (int x = 0; x < 3; x++) P[x] = acen * (A[x] - B[x]) + B[x];
    double adist = alpha_dist(tau, ea1, ea2);
    long long ntr = 0;

    /* one potential term (or nucleus) over a set of sites: the pot-cell
     * lattice sum of electronicPE/nuclearPE (integrals.F90:2608-2697,
     * :2780-2856) */
#define SITE_LOOP(SITE_COND, A3, TAD, ACCUM)                                   \
    do {                                                                       \
        for (int ps = 0; ps < s->nPotSites; ps++) {                            \
            int pt = s->potSiteType[ps];                                       \
            if (!(SITE_COND)) continue;                                        \
            const double *pp = s->potPos + 3 * ps;                             \
            double dv[3] = {P[0] - pp[0], P[1] - pp[1], P[2] - pp[2]};         \
            double lv2[3];                                                     \
            nearest_with_inv(s->vec, c->inv, dv, lv2);                         \
            double cov[3] = {P[0] - pp[0] - lv2[0], P[1] - pp[1] - lv2[1],     \
                             P[2] - pp[2] - lv2[2]};                           \
            double cos2 = sumsq3(cov);                                         \
            double tad = (TAD);                                                \
            if (cos2 > tad) continue;                                          \
            double mlr2 = cos2 + tad + 2.0 * sqrt(cos2 * tad);                 \
            for (int p = 0; p < s->nCells; p++) {                              \
                if (s->cellSizes[p] > mlr2) break;                             \
                const double *L = s->cellDims + 3 * p;                         \
                double r[3] = {cov[0] - L[0], cov[1] - L[1], cov[2] - L[2]};   \
                if (sumsq3(r) > tad) continue;                                 \
                double Cp[3] = {pp[0] + lv2[0] + L[0], pp[1] + lv2[1] + L[1],  \
                                pp[2] + lv2[2] + L[2]};                        \
                ntr++;                                                         \
                ACCUM;                                                         \
            }                                                                  \
        }                                                                      \
    } while (0)

    if (o->op == ORC_NP) {
        int any = 0;
        for (int ps = 0; ps < s->nPotSites; ps++) {
            int pt = s->potSiteType[ps];
            double Z = s->potZ[pt];
            if (Z < SMALL_THRESH) continue;
            double nuc = s->potNucAlpha[pt];
            double nucDist = tau * (ea1 + ea2 + nuc) / (ea1 + ea2) / nuc;
            double tad = (1 - s2 / adist) * nucDist * three_range_scale;
            const double *pp = s->potPos + 3 * ps;
            double dv[3] = {P[0] - pp[0], P[1] - pp[1], P[2] - pp[2]};
            double lv2[3];
            nearest_with_inv(s->vec, c->inv, dv, lv2);
            double cov[3] = {P[0] - pp[0] - lv2[0], P[1] - pp[1] - lv2[1],
                             P[2] - pp[2] - lv2[2]};
            double cos2 = sumsq3(cov);
            if (cos2 > tad) continue;
            any = 1;
            double mlr2 = cos2 + tad + 2.0 * sqrt(cos2 * tad);
            for (int p = 0; p < s->nCells; p++) {
                if (s->cellSizes[p] > mlr2) break;
                const double *L = s->cellDims + 3 * p;
                double r[3] = {cov[0] - L[0], cov[1] - L[1], cov[2] - L[2]};
                if (sumsq3(r) > tad) continue;
                double Cp[3] = {pp[0] + lv2[0] + L[0], pp[1] + lv2[1] + L[1],
                                pp[2] + lv2[2] + L[2]};
                ntr++;
                orc_prim(3, a1, a2, nuc, A, B, Cp, sw, one);
                for (int jj = 0; jj < n2; jj++)
                    for (int ii = 0; ii < n1; ii++)
                        blk[ii + 16 * jj] -= one[ii + 16 * jj] * Z;
            }
        }
        *contrib = any;
    } else if (o->op == ORC_EP) {
        /* term -> (type, alpha) */
        int t = 0;
        while (t + 1 < s->nPotTypes && c->potCumul[t + 1] <= o->term) t++;
        int n = o->term - c->potCumul[t];
        double a3 = s->potAlpha[t * s->maxPotAlphas + n];
        double potDist = tau * (ea1 + ea2 + a3) / (ea1 + ea2) / a3;
        double tadv = (1 - s2 / adist) * potDist * three_range_scale;
        int any = 0;
        SITE_LOOP(pt == t, a3, tadv, {
            any = 1;
            orc_prim(2, a1, a2, a3, A, B, Cp, sw, one);
            for (int jj = 0; jj < n2; jj++)
                for (int ii = 0; ii < n1; ii++) blk[ii + 16 * jj] += one[ii + 16 * jj];
        });
        *contrib = any;
    } else if (o->op == ORC_FORCE) {
        /* computeForceIntg (forces.F90:419-761), component term of the ket
         * derivative: minus the potential terms' three-Gaussian derivatives
         * times this spin's coefficients, plus the nuclear derivatives times
         * Z, minus the kinetic derivative, minus 2 totalEnergy (coef[potDim])
         * times the momentum block; the HAM sweep's image screening */
        int X = o->term;
        double E2 = 2.0 * o->coef[c->potDim];
        double three[768];
        for (int t = 0; t < s->nPotTypes; t++) {
            int na = s->potNAlpha[t];
            for (int n = 0; n <= na; n++) {
                double a3, cf = 0.0, Z = s->potZ[t];
                if (n < na) {
                    cf = o->coef[c->potCumul[t] + n];
                    a3 = s->potAlpha[t * s->maxPotAlphas + n];
                } else {
                    if (Z < SMALL_THRESH) continue;
                    a3 = s->potNucAlpha[t];
                }
                double tadv = tau * (1 - s2 / adist) * (a1 + a2 + a3) / (a1 + a2) / a3 *
                              three_range_scale;
                if (n < na) {
                    SITE_LOOP(pt == t, a3, tadv, {
                        orc_primd(12, a1, a2, a3, A, B, Cp, sw, three);
                        for (int jj = 0; jj < n2; jj++)
                            for (int ii = 0; ii < n1; ii++)
                                blk[ii + 16 * jj] -= three[256 * X + ii + 16 * jj] * cf;
                    });
                } else {
                    SITE_LOOP(pt == t, a3, tadv, {
                        orc_primd(11, a1, a2, a3, A, B, Cp, sw, three);
                        for (int jj = 0; jj < n2; jj++)
                            for (int ii = 0; ii < n1; ii++)
                                blk[ii + 16 * jj] += three[256 * X + ii + 16 * jj] * Z;
                    });
                }
            }
        }
        orc_primd(10, a1, a2, 0.0, A, B, A, sw, three);
        for (int jj = 0; jj < n2; jj++)
            for (int ii = 0; ii < n1; ii++) blk[ii + 16 * jj] -= three[256 * X + ii + 16 * jj];
        orc_prim3(6, a1, a2, A, B, A, sw, three);
        for (int jj = 0; jj < n2; jj++)
            for (int ii = 0; ii < n1; ii++) blk[ii + 16 * jj] -= E2 * three[256 * X + ii + 16 * jj];
    } else { /* ORC_HAM: integrals.F90:3211-3490 */
        for (int t = 0; t < s->nPotTypes; t++) {
            int na = s->potNAlpha[t];
            for (int n = 0; n <= na; n++) {
                double a3, cf = 0.0, Z = s->potZ[t];
                if (n < na) {
                    cf = o->coef[c->potCumul[t] + n];
                    a3 = s->potAlpha[t * s->maxPotAlphas + n];
                } else {
                    if (Z < SMALL_THRESH) continue;
                    a3 = s->potNucAlpha[t];
                }
                double tadv = tau * (1 - s2 / adist) * (a1 + a2 + a3) / (a1 + a2) / a3 *
                              three_range_scale;
                if (n < na) {
                    SITE_LOOP(pt == t, a3, tadv, {
                        orc_prim(2, a1, a2, a3, A, B, Cp, sw, one);
                        for (int jj = 0; jj < n2; jj++)
                            for (int ii = 0; ii < n1; ii++)
                                blk[ii + 16 * jj] += one[ii + 16 * jj] * cf;
                    });
                } else {
                    SITE_LOOP(pt == t, a3, tadv, {
                        orc_prim(3, a1, a2, a3, A, B, Cp, sw, one);
                        for (int jj = 0; jj < n2; jj++)
                            for (int ii = 0; ii < n1; ii++)
                                blk[ii + 16 * jj] -= one[ii + 16 * jj] * Z;
                    });
                }
            }
        }
        orc_prim(1, a1, a2, 0.0, A, B, A, sw, one);
        for (int jj = 0; jj < n2; jj++)
            for (int ii = 0; ii < n1; ii++) blk[ii + 16 * jj] += one[ii + 16 * jj];
        if (s->rel) {   /* integrals.F90:3463-3490 */
            orc_prim(4, a1, a2, 0.0, A, B, A, sw, one);
            for (int jj = 0; jj < n2; jj++)
                for (int ii = 0; ii < n1; ii++) blk[ii + 16 * jj] -= one[ii + 16 * jj];
        }
    }
#undef SITE_LOOP
    __atomic_fetch_add(&((op_args *)o)->ntr, ntr, __ATOMIC_RELAXED);
}

/* Store one finished atom pair into the global matrices (intgSaving.F90
 * saveCurrentPair / valeValeSaving / coreValeSaving / coreCoreSaving and
 * their Gamma twins).  cur: [nk][n1][n2] complex (or [n1][n2] real). */
static void save_pair(const drv_ctx *c, int i, int j, const double *cur,
                      int nt1, int nt2, double *vv, double *cv, double *cc, int full)
{
    const orc_sys *s = c->s;
    int ti = s->atomType[i], tj = s->atomType[j];
    int nv1 = s->nVale[ti], nv2 = s->nVale[tj];
    int nc1 = s->nCore[ti], nc2 = s->nCore[tj];
    int v1 = c->cumulVale[i], v2 = c->cumulVale[j];
    int c1 = c->cumulCore[i], c2 = c->cumulCore[j];
    int n = c->valeDim, nc = c->coreDim;
    int nk = s->gamma ? 1 : s->nk;
    int cplx = s->gamma ? 1 : 2;
    for (int k = 0; k < nk; k++) {
        const double *ck = cur + (long)k * nt1 * nt2 * cplx;
#define CUR(r, cidx, comp) ck[((r) * nt2 + (cidx)) * cplx + (comp)]
        /* valeVale block, upper triangle kept; strictly lower part of a
         * diagonal atom block zeroed (intgSaving.F90:484-509).  full
         * (valeValeSavingFull, :513-545): the block and its dagger in the
         * mirrored place, the dagger last, so that a diagonal atom block
         * ends as its own dagger */
        for (int b = 0; b < nv2; b++)
            for (int a = 0; a < nv1; a++) {
                long idx = (long)(v1 + a) + (long)(v2 + b) * n + (long)k * n * n;
                if (full) {
                    long idt = (long)(v2 + b) + (long)(v1 + a) * n + (long)k * n * n;
                    if (i != j)
                        for (int q = 0; q < cplx; q++) vv[idx * cplx + q] = CUR(a, b, q);
                    vv[idt * cplx] = CUR(a, b, 0);
                    if (cplx == 2) vv[idt * cplx + 1] = -CUR(a, b, 1);
                    continue;
                }
                int zero = (i == j) && (a > b);
                for (int q = 0; q < cplx; q++)
                    vv[idx * cplx + q] = zero ? 0.0 : CUR(a, b, q);
            }
        if (nc == 0) continue;
        /* coreVale (intgSaving.F90:555-608): pair block, then dagger */
        if (nc1)
            for (int b = 0; b < nv2; b++)
                for (int a = 0; a < nc1; a++) {
                    long idx = (long)(c1 + a) + (long)(v2 + b) * nc + (long)k * nc * n;
                    for (int q = 0; q < cplx; q++) cv[idx * cplx + q] = CUR(nv1 + a, b, q);
                }
        if (nc2)
            for (int b = 0; b < nv1; b++)
                for (int a = 0; a < nc2; a++) {
                    long idx = (long)(c2 + a) + (long)(v1 + b) * nc + (long)k * nc * n;
                    /* dagger(nv2+a, b) = conj(cur(b, nv2+a)) */
                    cv[idx * cplx] = CUR(b, nv2 + a, 0);
                    if (cplx == 2) cv[idx * cplx + 1] = -CUR(b, nv2 + a, 1);
                }
        if (nc1 == 0 || nc2 == 0) continue;
        /* coreCoreSaving; coreCoreSavingFull (intgSaving.F90) writes the
         * dagger for i == j too, over the block */
        if (!full || i != j)
            for (int b = 0; b < nc2; b++)
                for (int a = 0; a < nc1; a++) {
                    long idx = (long)(c1 + a) + (long)(c2 + b) * nc + (long)k * nc * nc;
                    for (int q = 0; q < cplx; q++) cc[idx * cplx + q] = CUR(nv1 + a, nv2 + b, q);
                }
        if (i != j || full)
            for (int b = 0; b < nc1; b++)
                for (int a = 0; a < nc2; a++) {
                    long idx = (long)(c2 + a) + (long)(c1 + b) * nc + (long)k * nc * nc;
                    cc[idx * cplx] = CUR(nv1 + b, nv2 + a, 0);
                    if (cplx == 2) cc[idx * cplx + 1] = -CUR(nv1 + b, nv2 + a, 1);
                }
#undef CUR
    }
}

int orc_integral(const orc_sys *s, int op, int term, const double *coef,
                 double *vv, double *cv, double *cc, int nthreads)
{
    init_tables();
    init_boys();
    drv_ctx c;
    memset(&c, 0, sizeof(c));
    c.s = s;
    inverse3(s->vec, c.inv);
    c.cumulVale = malloc(sizeof(int) * (s->nAtoms + 1));
    c.cumulCore = malloc(sizeof(int) * (s->nAtoms + 1));
    int v = 0, cr = 0;
    for (int a = 0; a < s->nAtoms; a++) {
        c.cumulVale[a] = v;
        c.cumulCore[a] = cr;
        v += s->nVale[s->atomType[a]];
        cr += s->nCore[s->atomType[a]];
    }
    c.valeDim = v;
    c.coreDim = cr;
    c.potCumul = malloc(sizeof(int) * (s->nPotTypes + 1));
    c.potDim = 0;
    for (int t = 0; t < s->nPotTypes; t++) {
        c.potCumul[t] = c.potDim;
        c.potDim += s->potNAlpha[t];
    }
    c.elemAlpha = malloc(sizeof(double *) * s->nTypes);
    for (int t = 0; t < s->nTypes; t++) {
        int first = t;
        for (int u = 0; u < t; u++)
            if (s->element[u] == s->element[t]) { first = u; break; }
        c.elemAlpha[t] = s->alpha + first * s->maxAlphas;
    }
    int nk = s->nk;
    int cplx = s->gamma ? 1 : 2;
    if (!s->gamma) {
        c.phase = malloc(sizeof(double) * 2 * nk * s->nCells);
        for (int cell = 0; cell < s->nCells; cell++)
            for (int k = 0; k < nk; k++) {
                const double *kp = s->kpt + 3 * k, *L = s->cellDims + 3 * cell;
                double dot = kp[0] * L[0] + kp[1] * L[1] + kp[2] * L[2];
                c.phase[2 * (cell * nk + k)] = cos(dot);
                c.phase[2 * (cell * nk + k) + 1] = sin(dot);
            }
    }
    int n = c.valeDim, nc = c.coreDim;
    int nkm = s->gamma ? 1 : nk;
    memset(vv, 0, sizeof(double) * cplx * (long)n * n * nkm);
    if (nc) {
        memset(cv, 0, sizeof(double) * cplx * (long)nc * n * nkm);
        memset(cc, 0, sizeof(double) * cplx * (long)nc * nc * nkm);
    }
    op_args o = {op, term, coef, 0};
    long long pcs = 0, pps = 0;
    int err = 0;
    int nA = s->nAtoms;
    long npairs = (long)nA * (nA + 1) / 2;
    if (nthreads <= 0) nthreads = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : omp_get_max_threads()) reduction(+ : pcs, pps)
    {
        int ms = s->maxStates, ma = s->maxAlphas;
        double *cur = malloc(sizeof(double) * cplx * nkm * ms * ms);
        double *Y = malloc(sizeof(double) * 16 * ma * ms);
        double *X = malloc(sizeof(double) * ms * ms);
        double blk[256];
#pragma omp for schedule(dynamic, 1)
        for (long pidx = 0; pidx < npairs; pidx++) {
            if ((pidx - sample_offset) % sample_stride != 0) continue;
            /* pidx -> (i, j), i <= j, row-major over i */
            int i = 0;
            long rem = pidx;
            while (rem >= nA - i) { rem -= nA - i; i++; }
            int j = i + (int)rem;
            int ti = s->atomType[i], tj = s->atomType[j];
            int nt1 = s->nVale[ti] + s->nCore[ti], nt2 = s->nVale[tj] + s->nCore[tj];
            int na1 = s->nAlpha[ti], na2 = s->nAlpha[tj];
            const double *A = s->pos + 3 * i, *Bj = s->pos + 3 * j;
            double dv[3] = {A[0] - Bj[0], A[1] - Bj[1], A[2] - Bj[2]};
            double lv[3];
            nearest_with_inv(s->vec, c.inv, dv, lv);
            double r0[3] = {A[0] - Bj[0] - lv[0], A[1] - Bj[1] - lv[1], A[2] - Bj[2] - lv[2]};
            double sepSq = sumsq3(r0);
            double lim = alpha_dist(s->logThresh, c.elemAlpha[ti][0], c.elemAlpha[tj][0]);
            if (sepSq > lim) continue;
            double maxLR = sepSq + lim + 2.0 * sqrt(sepSq * lim);
            if (maxLR > s->cellSizes[s->nCells - 1]) {
#pragma omp atomic write
                err = ORC_ERR_CELLS;
                continue;
            }
            memset(cur, 0, sizeof(double) * cplx * nkm * nt1 * nt2);
            for (int k = 0; k < s->nCells; k++) {
                if (s->cellSizes[k] > maxLR) break;
                const double *L = s->cellDims + 3 * k;
                double B[3] = {Bj[0] + lv[0] + L[0], Bj[1] + lv[1] + L[1], Bj[2] + lv[2] + L[2]};
                double d[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
                double s2 = sumsq3(d);
                if (s2 > lim) continue;
                pcs++;
                memset(Y, 0, sizeof(double) * 16 * ma * nt2);
                int maxA1 = 0;
                /* alpha walk, integrals.F90:451-570 */
                int nmin = na1 < na2 ? na1 : na2;
                for (int l = 0; l < nmin; l++) {
                    if (alpha_dist(s->logThresh, c.elemAlpha[ti][l], c.elemAlpha[tj][l]) < s2) break;
                    int mode = 0, idx1 = l, idx2 = l;
                    while (1) {
                        if (mode == 1) {
                            idx1++;
                            if (idx1 >= na1) { mode = 2; idx1 = l; continue; }
                        } else if (mode == 2) {
                            idx2++;
                            if (idx2 >= na2) break;
                        }
                        if (alpha_dist(s->logThresh, c.elemAlpha[ti][idx1], c.elemAlpha[tj][idx2]) < s2) {
                            if (mode == 1) { mode = 2; idx1 = l; continue; }
                            break;
                        }
                        pps++;
                        int contrib;
                        pair_block(&c, &o, ti, tj, idx1, idx2, s2, A, B, blk, &contrib);
                        if (contrib) {
                            int lmA = s->lmAlphaIndex[ti * ma + idx1];
                            for (int m = 0; m < nt2; m++) {
                                int lm2 = s->lmIndex[tj * ms + m] - 1;
                                double c2 = s->basis[(tj * ms + m) * ma + idx2];
                                for (int lm = 0; lm < lmA; lm++)
                                    Y[(m * ma + idx1) * 16 + lm] += blk[lm + 16 * lm2] * c2;
                            }
                            if (idx1 + 1 > maxA1) maxA1 = idx1 + 1;
                        }
                        if (mode == 0) mode = 1;
                    }
                }
                /* multWithBasisFn1 (intgSaving.F90:71-89) */
                for (int b = 0; b < nt2; b++)
                    for (int a = 0; a < nt1; a++) {
                        int lm1 = s->lmIndex[ti * ms + a] - 1;
                        double sum = 0.0;
                        for (int al = 0; al < maxA1; al++)
                            sum += s->basis[(ti * ms + a) * ma + al] * Y[(b * ma + al) * 16 + lm1];
                        X[a * nt2 + b] = sum;
                    }
                /* applyPhaseFactors / applyPhaseFactorsGamma; the momentum
                 * matrix takes a factor -i (runCode 1, intgSaving.F90:
                 * 178-187, the Gamma build keeping -X, :727-731) */
                if (s->gamma) {
                    const double sg = op == ORC_MM ? -1.0 : 1.0;
                    for (int q = 0; q < nt1 * nt2; q++) cur[q] += sg * X[q];
                } else if (op == ORC_MM || (op == ORC_KO && (term & 1))) {
                    for (int kk = 0; kk < nk; kk++) {
                        double pr = c.phase[2 * (k * nk + kk)], pi = c.phase[2 * (k * nk + kk) + 1];
                        double *ck = cur + (long)kk * nt1 * nt2 * 2;
                        for (int q = 0; q < nt1 * nt2; q++) {
                            ck[2 * q] += pi * X[q];
                            ck[2 * q + 1] -= pr * X[q];
                        }
                    }
                } else {
                    for (int kk = 0; kk < nk; kk++) {
                        double pr = c.phase[2 * (k * nk + kk)], pi = c.phase[2 * (k * nk + kk) + 1];
                        double *ck = cur + (long)kk * nt1 * nt2 * 2;
                        for (int q = 0; q < nt1 * nt2; q++) {
                            ck[2 * q] += pr * X[q];
                            ck[2 * q + 1] += pi * X[q];
                        }
                    }
                }
            }
            /* kPointLatticeOriginShift (intgSaving.F90:241-292) */
            if (!s->gamma) {
                for (int kk = 0; kk < nk; kk++) {
                    const double *kp = s->kpt + 3 * kk;
                    double dot = kp[0] * lv[0] + kp[1] * lv[1] + kp[2] * lv[2];
                    if (fabs(dot) >= SMALL_THRESH) {
                        double cr_ = cos(dot), ci_ = sin(dot);
                        double *ck = cur + (long)kk * nt1 * nt2 * 2;
                        for (int q = 0; q < nt1 * nt2; q++) {
                            double re = ck[2 * q], im = ck[2 * q + 1];
                            ck[2 * q] = cr_ * re - ci_ * im;
                            ck[2 * q + 1] = cr_ * im + ci_ * re;
                        }
                    }
                }
            }
            save_pair(&c, i, j, cur, nt1, nt2, vv, cv, cc, op == ORC_FORCE && !s->gamma);
        }
        free(cur); free(Y); free(X);
    }
    cnt_pc = pcs;
    cnt_pp = pps;
    cnt_tr = o.ntr;
    free(c.cumulVale); free(c.cumulCore); free(c.potCumul); free(c.elemAlpha);
    free(c.phase);
    return err;
}
