/*
 * mh_oracle.c — CPU FP64 restatement of MarineHydro.jl's BEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Each function
 * cites the reference file:line it restates.  Built by oracle/Makefile into
 * oracle/liboracle.so (gcc, no -ffast-math: signed zeros and atan2 branches
 * must behave exactly as in the reference).
 *
 * Parity unpinned: the two forward-speed terms of ora_hydro_coeffs_fs (U n.dgrad
 * in the radiation condition, rho U K_x sigma in the pressure) are restatements
 * of waves.jl:183-199 and solve.jl:35-45 with no golden vector or exact identity
 * to check them against; the rest of that function is checked through the
 * encountered values, the fixed-normal operator and the U = 0 reduction.
 */
#include "mh_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

static const double PI = 3.14159265358979323846;
static const double EULER_GAMMA = 0.57721566490153286060651209;

static int g_threads = 1;
void ora_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int ora_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------------- */
/* Birk (2021) constant-source panel: src/green_functions/rankine_vectorized.jl */
/* ------------------------------------------------------------------------- */

#define BIRK_ZTOL 1e-14   /* rankine_vectorized.jl:24 */
#define BIRK_DKTOL 1e-14  /* rankine_vectorized.jl:25 */

static inline double nrm3(const double v[3]) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
static inline double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* _atan_ratio (rankine_vectorized.jl:42-45): atan(num/den) without dividing;
 * `den < 0` is false for -0.0, exactly as Julia's ifelse on the primal. */
static inline double atan_ratio(double num, double den) {
    double flipped = (den < 0.0) ? -num : num;
    double ad = (den < 0.0) ? -den : den;
    return atan2(flipped, ad);
}

/* _inplane_atan (rankine_vectorized.jl:49-54): primal forced to 0 at |z|<=tol (x - x). */
static inline double inplane_atan(double z, double datan) {
    if (fabs(z) <= BIRK_ZTOL) return datan - datan;
    return datan;
}

/* birk_panel_geometry (rankine_vectorized.jl:80-137).
 * V: 4 corners x 3 (row-major).  T: 3x3 column-major [s t n].  lc: 4 x (x,y). */
void ora_birk_panel_geometry(const double V[12], double T[9], double qgc[3], double lc[8], double* area) {
    double mid[4][3];
    for (int i = 0; i < 4; i++) {
        int j = (i + 1) % 4;
        for (int c = 0; c < 3; c++) mid[i][c] = (V[3 * i + c] + V[3 * j + c]) / 2;
    }
    double mgc[3], s[3], tt[3], n[3], t[3];
    for (int c = 0; c < 3; c++) mgc[c] = (mid[1][c] + mid[3][c]) / 2;
    for (int c = 0; c < 3; c++) s[c] = mid[1][c] - mid[3][c];
    double ns = nrm3(s);
    for (int c = 0; c < 3; c++) s[c] /= ns;
    for (int c = 0; c < 3; c++) tt[c] = mid[2][c] - mid[0][c];
    cross3(s, tt, n);
    double nn = nrm3(n);
    for (int c = 0; c < 3; c++) n[c] /= nn;
    cross3(n, s, t);
    double nt = nrm3(t);
    for (int c = 0; c < 3; c++) t[c] /= nt;
    for (int c = 0; c < 3; c++) {
        T[c] = s[c];
        T[3 + c] = t[c];
        T[6 + c] = n[c];
    }
    /* compute_temporary_panel_corners (:92-97) */
    double tc[4][2];
    for (int k = 0; k < 4; k++) {
        double d[3] = {V[3 * k] - mgc[0], V[3 * k + 1] - mgc[1], V[3 * k + 2] - mgc[2]};
        tc[k][0] = dot3(d, s);
        tc[k][1] = dot3(d, t);
    }
    /* compute_area_and_centroid (:100-115) */
    double A = 0, Mx = 0, My = 0;
    for (int k = 0; k < 4; k++) {
        int k1 = (k + 1) % 4;
        double xk = tc[k][0], yk = tc[k][1], xk1 = tc[k1][0], yk1 = tc[k1][1];
        A += (yk1 - yk) * (xk + xk1) / 2;
        Mx += (xk1 - xk) * (yk * yk + yk * yk1 + yk1 * yk1) / 6;
        My += (yk1 - yk) * (xk * xk + xk * xk1 + xk1 * xk1) / 6;
    }
    double clx = My / A, cly = -Mx / A;
    /* compute_qgc (:118-122) and compute_local_corners (:125-127) */
    for (int c = 0; c < 3; c++) qgc[c] = mgc[c] + clx * s[c] + cly * t[c];
    for (int k = 0; k < 4; k++) {
        lc[2 * k] = tc[k][0] - clx;
        lc[2 * k + 1] = tc[k][1] - cly;
    }
    *area = A;
}

/* velocity_potential, eq. (46) (rankine_vectorized.jl:140-164) */
double ora_velocity_potential(double x, double y, double z, const double lc[8]) {
    double sum = 0;
    for (int k = 0; k < 4; k++) {
        int k1 = (k + 1) % 4;
        double xk = lc[2 * k], yk = lc[2 * k + 1], xk1 = lc[2 * k1], yk1 = lc[2 * k1 + 1];
        double dx = xk1 - xk, dy = yk1 - yk;
        double dk = hypot(dx, dy);
        double rk = sqrt((x - xk) * (x - xk) + (y - yk) * (y - yk) + z * z);
        double rk1 = sqrt((x - xk1) * (x - xk1) + (y - yk1) * (y - yk1) + z * z);
        double ek = (x - xk) * (x - xk) + z * z, ek1 = (x - xk1) * (x - xk1) + z * z;
        double hk = (x - xk) * (y - yk), hk1 = (x - xk1) * (y - yk1);
        double log_term = log((rk + rk1 - dk) / (rk + rk1 + dk));
        double log_part = ((x - xk) * dy - (y - yk) * dx) / dk * log_term;
        double datan = inplane_atan(z, atan_ratio(dy * ek - dx * hk, dx * z * rk) -
                                           atan_ratio(dy * ek1 - dx * hk1, dx * z * rk1));
        double contrib = (dk <= BIRK_DKTOL) ? 0.0 : -(log_part + z * datan) / (4 * PI);
        sum += contrib;
    }
    return sum;
}

/* velocity_derivatives, eqs. (47)-(49) (rankine_vectorized.jl:168-190) */
void ora_velocity_derivatives(double x, double y, double z, const double lc[8], double v[3]) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (int k = 0; k < 4; k++) {
        int k1 = (k + 1) % 4;
        double xk = lc[2 * k], yk = lc[2 * k + 1], xk1 = lc[2 * k1], yk1 = lc[2 * k1 + 1];
        double dx = xk1 - xk, dy = yk1 - yk;
        double dk = hypot(dx, dy);
        double rk = sqrt((x - xk) * (x - xk) + (y - yk) * (y - yk) + z * z);
        double rk1 = sqrt((x - xk1) * (x - xk1) + (y - yk1) * (y - yk1) + z * z);
        double ek = (x - xk) * (x - xk) + z * z, ek1 = (x - xk1) * (x - xk1) + z * z;
        double hk = (x - xk) * (y - yk), hk1 = (x - xk1) * (y - yk1);
        double log_term = log((rk + rk1 - dk) / (rk + rk1 + dk));
        double vx = (dy / dk) * log_term, vy = (-dx / dk) * log_term;
        double vz = inplane_atan(z, atan_ratio(dy * ek - dx * hk, dx * z * rk) -
                                        atan_ratio(dy * ek1 - dx * hk1, dx * z * rk1));
        if (!(dk <= BIRK_DKTOL)) {
            s0 += vx;
            s1 += vy;
            s2 += vz;
        }
    }
    double sc = -1.0 / (4 * PI);
    v[0] = sc * s0;
    v[1] = sc * s1;
    v[2] = sc * s2;
}

/* velocity_hessian, eqs. (50)-(58) (rankine_vectorized.jl:196-235); H column-major local frame */
void ora_velocity_hessian(double x, double y, double z, const double lc[8], double H[9]) {
    double acc[9] = {0};
    for (int k = 0; k < 4; k++) {
        int k1 = (k + 1) % 4;
        double xk = lc[2 * k], yk = lc[2 * k + 1], xk1 = lc[2 * k1], yk1 = lc[2 * k1 + 1];
        double dx = xk1 - xk, dy = yk1 - yk;
        double dk = hypot(dx, dy);
        if (dk <= BIRK_DKTOL) continue;
        double dxk = x - xk, dyk = y - yk, dxk1 = x - xk1, dyk1 = y - yk1;
        double rk = sqrt(dxk * dxk + dyk * dyk + z * z), rk1 = sqrt(dxk1 * dxk1 + dyk1 * dyk1 + z * z);
        double rsum = rk + rk1;
        double den = rsum * rsum - dk * dk;
        double rhok = rk * rk1 + dxk * dxk1 + dyk * dyk1 + z * z;
        double lamk = dxk * dyk1 - dxk1 * dyk;
        double xterm = dxk / rk + dxk1 / rk1, yterm = dyk / rk + dyk1 / rk1, zterm = z / rk + z / rk1;
        double rhoden = (rk * rk1) * rhok;
        double hxx = 2 * dy / den * xterm, hyx = -2 * dx / den * xterm;
        double hxy = 2 * dy / den * yterm, hyy = -2 * dx / den * yterm;
        double hxz = 2 * dy / den * zterm, hyz = -2 * dx / den * zterm;
        double hzx = z * dy * rsum / rhoden, hzy = -z * dx * rsum / rhoden, hzz = lamk * rsum / rhoden;
        double h[9] = {hxx, hyx, hzx, hxy, hyy, hzy, hxz, hyz, hzz};
        for (int q = 0; q < 9; q++) acc[q] += h[q];
    }
    double sc = -1.0 / (4 * PI);
    for (int q = 0; q < 9; q++) H[q] = sc * acc[q];
}

/* birk_phi_and_velocity, fused eq. (46)+(47)-(49) (rankine_vectorized.jl:241-267) */
void ora_birk_phi_and_velocity(double x, double y, double z, const double lc[8], double* phi, double v[3]) {
    double sx[4], sy[4], r[4], e[4], h[4];
    for (int k = 0; k < 4; k++) {
        sx[k] = x - lc[2 * k];
        sy[k] = y - lc[2 * k + 1];
        r[k] = sqrt(sx[k] * sx[k] + sy[k] * sy[k] + z * z);
        e[k] = sx[k] * sx[k] + z * z;
        h[k] = sx[k] * sy[k];
    }
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int k = 0; k < 4; k++) {
        int kn = (k + 1) % 4;
        double dx = lc[2 * kn] - lc[2 * k], dy = lc[2 * kn + 1] - lc[2 * k + 1];
        double dk = hypot(dx, dy);
        double rsum = r[k] + r[kn];
        double log_term = log((rsum - dk) / (rsum + dk));
        double lt_dk = log_term / dk;
        double datan = atan_ratio(dy * e[k] - dx * h[k], dx * z * r[k]) -
                       atan_ratio(dy * e[kn] - dx * h[kn], dx * z * r[kn]);
        double vzk = inplane_atan(z, datan);
        double phik = (sx[k] * dy - sy[k] * dx) * lt_dk + z * vzk;
        if (!(dk <= BIRK_DKTOL)) {
            s0 += phik;
            s1 += dy * lt_dk;
            s2 += -dx * lt_dk;
            s3 += vzk;
        }
    }
    double scale = -1.0 / (4 * PI);
    *phi = scale * s0;
    v[0] = scale * s1;
    v[1] = scale * s2;
    v[2] = scale * s3;
}

/* _vrankine_from_geom (rankine_vectorized.jl:273-287):
 * phi = integral of 1/|p-xi| over the panel, grad = gradient w.r.t. the source point. */
void ora_vrankine_from_geom(const double p[3], const double c[3], double radius, double area,
                            const double T[9], const double qgc[3], const double lc[8],
                            double* phi, double grad[3]) {
    double rb[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
    double r = nrm3(rb);
    if (r > 7 * radius) {
        double r3 = r * r * r;
        *phi = area / r;
        for (int q = 0; q < 3; q++) grad[q] = (rb[q] / r3) * area;
        return;
    }
    double d[3] = {p[0] - qgc[0], p[1] - qgc[1], p[2] - qgc[2]};
    double fl[3];
    for (int q = 0; q < 3; q++) fl[q] = T[3 * q] * d[0] + T[3 * q + 1] * d[1] + T[3 * q + 2] * d[2];
    double pb, vl[3];
    ora_birk_phi_and_velocity(fl[0], fl[1], fl[2], lc, &pb, vl);
    *phi = -4 * PI * pb;
    for (int q = 0; q < 3; q++) {
        double vg = T[q] * vl[0] + T[3 + q] * vl[1] + T[6 + q] * vl[2];
        grad[q] = 4 * PI * vg;
    }
}

/* ------------------------------------------------------------------------- */
/* Liang-Wu-Noblesse wave term: src/green_functions/wu.jl                     */
/* Polynomial and Struve fits are the published Liang, Wu & Noblesse (2018)  */
/* and Newman/Aarts-Janssen approximations, restated coefficient by          */
/* coefficient from wu.jl:131-435.                                           */
/* ------------------------------------------------------------------------- */

static inline double horner(const double* c, int deg, double t) {
    double acc = c[deg - 1] + c[deg] * t;
    acc *= t;
    for (int i = deg - 2; i >= 1; i--) acc = (c[i] + acc) * t;
    return c[0] + acc;
}
/* wu.jl:131-146 */
static const double WU_DC[6] = {14.19, -148.24, 847.8, -2318.58, 3168.35, -1590.27};
/* wu.jl:149-172 */
static const double WU_DB[10] = {1.11, 2.894, -76.765, 1565.35, -11336.19, 44270.15, -97014.11, 118879.26, -76209.82, 19923.28};
/* wu.jl:174-197 */
static const double WU_DA[10] = {2.948, -24.53, 249.69, -754.85, -1187.71, 16370.75, -48811.41, 68220.87, -46688.0, 12622.25};
/* wu.jl:235-258 */
static const double WU_D[10] = {0.632, -40.97, 667.16, -6072.07, 31127.39, -96293.05, 181856.75, -205690.43, 128170.2, -33744.6};
/* wu.jl:260-279 */
static const double WU_C[8] = {1.268, -9.747, 209.653, -1397.89, 5155.67, -9844.35, 9136.4, -3272.62};
/* wu.jl:281-304 */
static const double WU_B[10] = {0.938, 5.737, -67.92, 796.534, -4780.77, 17137.74, -36618.81, 44894.06, -29030.24, 7671.22};
/* wu.jl:306-329 */
static const double WU_A[10] = {1.21, -13.328, 215.896, -1763.96, 8418.94, -24314.21, 42002.57, -41592.9, 21859.0, -4838.6};

/* StruveH0 (wu.jl:368-400) */
double ora_struve_h0(double xx) {
    if (xx <= 3.0) {
        double yy = (xx / 3.0) * (xx / 3.0);
        double P0 = +1.909859164, P1 = -1.909855001, P2 = +0.687514637, P3 = -0.126164557, P4 = +0.013828813,
               P5 = -0.000876918;
        double h = P0 + (P1 + (P2 + (P3 + (P4 + P5 * yy) * yy) * yy) * yy) * yy;
        return h * (xx / 3.0);
    }
    double yy = (3.0 / xx) * (3.0 / xx);
    double a0 = 0.99999906, a1 = 4.77228920, a2 = 3.85542044, a3 = 0.32303607;
    double b1 = 4.88331068, b2 = 4.28957333, b3 = 0.52120508;
    double c1 = 2.0 * (a0 + (a1 + (a2 + a3 * yy) * yy) * yy);
    double c2 = PI * xx * (1.0 + (b1 + (b2 + b3 * yy) * yy) * yy);
    return c1 / c2 + y0(xx);
}

/* StruveH1 (wu.jl:404-435) */
double ora_struve_h1(double xx) {
    if (xx <= 3.0) {
        double yy = (xx / 3.0) * (xx / 3.0);
        double P1 = +1.909859286, P2 = -1.145914713, P3 = +0.294656958, P4 = -0.042070508, P5 = +0.003785727,
               P6 = -0.000207183;
        return (P1 + (P2 + (P3 + (P4 + (P5 + P6 * yy) * yy) * yy) * yy) * yy) * yy;
    }
    double yy = (3.0 / xx) * (3.0 / xx);
    double a0 = 1.00000004, a1 = 3.92205313, a2 = 2.64893033, a3 = 0.27450895;
    double b1 = 3.81095112, b2 = 2.26216956, b3 = 0.10885141;
    double c1 = 2.0 * (a0 + (a1 + (a2 + a3 * yy) * yy) * yy);
    double c2 = PI * (1.0 + (b1 + (b2 + b3 * yy) * yy) * yy);
    return c1 / c2 + y1(xx);
}

/* _wu_wave_term_and_derivs (wu.jl:28-37) with GF_Func_L0/Lp/Ls/Lsp/W/Wh (wu.jl:201-366) */
void ora_wu_wave_term_and_derivs(double hh, double vv, double out[6]) {
    double dd = sqrt(hh * hh + vv * vv);
    double alpha = -vv / dd, beta = hh / dd, rho = dd / (1 + dd);
    double ev = exp(vv);
    double omr3 = (1.0 - rho) * (1.0 - rho) * (1.0 - rho);
    double dd2 = dd * dd, dd3 = dd * dd * dd;
    /* GF_Func_Lp (wu.jl:339-353) */
    double A = horner(WU_A, 9, rho), B = horner(WU_B, 9, rho), C = horner(WU_C, 7, rho), D = horner(WU_D, 9, rho);
    double RR = (1.0 - beta) * A;
    RR -= beta * B;
    RR -= alpha * C / (1.0 + 6.0 * alpha * rho * (1.0 - rho));
    RR += beta * (1.0 - beta) * D;
    double Lp = rho * omr3 * RR;
    /* GF_Func_L0 (wu.jl:355-366) */
    double gama = 0.5772156649;
    double PP = log(0.5 * (dd - vv)) + gama - 2.0 * dd2;
    PP = ev * PP;
    PP = PP + dd2 - vv;
    double L0 = 2.0 * PP / (1.0 + dd3) + 2.0 * Lp;
    /* GF_Func_W (wu.jl:333-337): 2pi (H0 - i J0) e^vv */
    double H0 = ora_struve_h0(hh), J0 = j0(hh);
    double Wr = 2 * PI * H0 * ev, Wi = 2 * PI * (-J0) * ev;
    /* GF_Func_Lsp (wu.jl:207-219) */
    double dA = horner(WU_DA, 9, rho), dB = horner(WU_DB, 9, rho), dC = horner(WU_DC, 5, rho);
    double RS = beta * dA;
    RS -= (1.0 - alpha) * dB;
    RS += beta * (1.0 - beta) * rho * (1.0 - 2.0 * rho) * dC;
    double Lsp = rho * omr3 * RS;
    /* GF_Func_Ls (wu.jl:221-233) */
    double PS = (beta + hh) / (dd - vv);
    PS = PS - 2.0 * beta + 2.0 * ev * dd - hh;
    double QS = exp(-dd) * (1.0 - beta);
    QS = QS * (1.0 + dd / (1.0 + dd3));
    double Ls = 2.0 * PS / (1.0 + dd3) - 4.0 * QS + 2.0 * Lsp;
    /* GF_Func_Wh (wu.jl:201-205): 2pi (2/pi - H1 + i J1) e^vv */
    double H1 = ora_struve_h1(hh), J1 = j1(hh);
    double Whr = 2 * PI * (2 / PI - H1) * ev, Whi = 2 * PI * J1 * ev;
    double Gr = -L0 - Wr, Gi = -Wi;
    out[0] = Gr;
    out[1] = Gi;
    out[2] = -Ls - Whr;
    out[3] = -Whi;
    out[4] = Gr + 2 / dd;
    out[5] = Gi;
}

/* ------------------------------------------------------------------------- */
/* Complex exponential integral E1 (replaces SpecialFunctions.expint,        */
/* exact_Guevel_Delhommeau.jl:5).  Region rule shared with the HIP kernel:   */
/*  power series  E1 = -gamma - log z - sum a_n z^n, a_n = (-1)^n/(n n!)    */
/*  (Horner, degree from a table at ceil|z|) for |z|<=2, for                  */
/*  Re z<=0 with |z|<=4, and inside a cone around the negative real axis      */
/*  (|z|<=40, Im z <= -Re z (24/|z| - 0.1)); elsewhere the even continued     */
/*  fraction exp(z)E1(z) = 1/(z+1- 1/(z+3- 4/(z+5- ...))) by modified Lentz.  */
/* ------------------------------------------------------------------------- */

static int e1_use_series(cplx z) {
    double R = cabs(z), x = creal(z), y = fabs(cimag(z));
    if (R <= 2.0) return 1;
    if (x <= 0.0 && R <= 4.0) return 1;
    if (x < 0.0 && R <= 40.0 && y <= (-x) * (24.0 / R - 0.1)) return 1;
    return 0;
}

#define E1_QUAL static
#include "e1_table.h"

/* Horner evaluation of the truncated series, degree from the table at ceil|z| */
static cplx e1_series(cplx z) {
    double R = cabs(z);
    int ir = (int)ceil(R);
    if (ir > 40) ir = 40;
    int N = E1_DEG[ir];
    cplx p = E1_A[N - 1];
    for (int q = N - 2; q >= 0; q--) p = E1_A[q] + z * p;
    p = z * p;
    return -EULER_GAMMA - clog(z) - p;
}

static cplx expE1_cf(cplx z) {
    const double tiny = 1e-300;
    cplx f = z + 1.0;
    if (f == 0) f = tiny;
    cplx C = f, Dd = 0;
    for (int n = 1; n < 600; n++) {
        double a = -(double)n * (double)n;
        cplx b = z + (double)(2 * n + 1);
        Dd = b + a * Dd;
        if (Dd == 0) Dd = tiny;
        Dd = 1.0 / Dd;
        C = b + a / C;
        if (C == 0) C = tiny;
        cplx d = C * Dd;
        f *= d;
        if (cabs(d - 1.0) < 1e-16) break;
    }
    return 1.0 / f;
}

static cplx expE1(cplx z) {
    if (e1_use_series(z)) return cexp(z) * e1_series(z);
    return expE1_cf(z);
}

void ora_expint_e1(double zr, double zi, double out[2]) {
    cplx z = zr + I * zi;
    cplx e = e1_use_series(z) ? e1_series(z) : expE1_cf(z) * cexp(-z);
    out[0] = creal(e);
    out[1] = cimag(e);
}

void ora_expE1(double zr, double zi, double out[2]) {
    cplx w = expE1(zr + I * zi);
    out[0] = creal(w);
    out[1] = cimag(w);
}

/* ------------------------------------------------------------------------- */
/* Exact Guevel-Delhommeau wave term (exact_Guevel_Delhommeau.jl:15-59) with  */
/* the adaptive Gauss-Kronrod (7,15) bisection of HCubature's hquadrature:   */
/* pop the interval of largest error, bisect, stop when                      */
/* E <= max(rtol |I|, atol), re-sum at the end.                              */
/* ------------------------------------------------------------------------- */

/* G7/K15 nodes, outermost first; Kronrod weights; Gauss weights (QUADPACK qk15). */
static const double GK_X[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                               0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                               0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                               0.207784955007898467600689403773245, 0.0};
static const double GK_WK[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                                0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                                0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                                0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
static const double GK_WG[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                                0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

void ora_gk15_rule(double xk[8], double wk[8], double wg[4]) {
    memcpy(xk, GK_X, sizeof GK_X);
    memcpy(wk, GK_WK, sizeof GK_WK);
    memcpy(wg, GK_WG, sizeof GK_WG);
}

/* integrands of _dimless_wave_term (:16-23) and _d_dimless_wave_term_dr (:36-43) */
static cplx del_integrand(double th, double r, double z, int deriv) {
    double c = cos(th);
    cplx zeta = z + I * (r * c);
    cplx ez = cexp(zeta);
    cplx w = expE1(zeta); /* exp(zeta) * E1(zeta) */
    cplx j = w / PI + I * ez;
    if (!deriv) return creal(j) + I * creal(ez);
    cplx q = j - 1.0 / (PI * zeta);
    return (-c * cimag(q)) + I * (-c * cimag(ez));
}

static cplx gk15(double a, double b, double r, double z, int deriv, double* err) {
    double c = (a + b) * 0.5, d = (b - a) * 0.5;
    cplx f0 = del_integrand(c, r, z, deriv);
    cplx Ik = f0 * GK_WK[7], Ig = f0 * GK_WG[3];
    for (int i = 0; i < 7; i++) {
        double dx = d * GK_X[i];
        cplx fx = del_integrand(c - dx, r, z, deriv) + del_integrand(c + dx, r, z, deriv);
        Ik += fx * GK_WK[i];
        if (i & 1) Ig += fx * GK_WG[i >> 1];
    }
    Ik *= d;
    Ig *= d;
    *err = cabs(Ik - Ig);
    return Ik;
}

#define ORA_MAXBOX 4096
typedef struct {
    double a, b, E;
    cplx val;
} ora_box;

static cplx hquad(double r, double z, int deriv, double rtol, double atol, int* nbox_out) {
    ora_box* box = (ora_box*)malloc(sizeof(ora_box) * ORA_MAXBOX);
    double a = 0.0, b = PI / 2;
    double E;
    cplx I0 = gk15(a, b, r, z, deriv, &E);
    box[0].a = a;
    box[0].b = b;
    box[0].val = I0;
    box[0].E = E;
    int nb = 1;
    cplx Itot = I0;
    double Etot = E;
    while (Etot > fmax(rtol * cabs(Itot), atol) && nb < ORA_MAXBOX - 1) {
        int imax = 0;
        for (int q = 1; q < nb; q++)
            if (box[q].E > box[imax].E) imax = q;
        ora_box bx = box[imax];
        box[imax] = box[nb - 1];
        nb--;
        Itot -= bx.val;
        Etot -= bx.E;
        double w = (bx.b - bx.a) / 2;
        double m = bx.a + w;
        double E1v, E2v;
        cplx I1 = gk15(bx.a, m, r, z, deriv, &E1v);
        cplx I2 = gk15(m, bx.b, r, z, deriv, &E2v);
        Itot += I1 + I2;
        Etot += E1v + E2v;
        box[nb].a = bx.a; box[nb].b = m; box[nb].val = I1; box[nb].E = E1v; nb++;
        box[nb].a = m; box[nb].b = bx.b; box[nb].val = I2; box[nb].E = E2v; nb++;
    }
    cplx Is = 0;
    for (int q = 0; q < nb; q++) Is += box[q].val;
    free(box);
    if (nbox_out) *nbox_out = nb;
    return Is;
}

int ora_delhommeau_term(double r, double z, double rtol, double atol, double out[6], int nboxes[2]) {
    int n0 = 0, n1 = 0;
    cplx Gr = 4.0 * hquad(r, z, 1, rtol, atol, &n1); /* _d_dimless_wave_term_dr (:35-46) */
    cplx G = 4.0 * hquad(r, z, 0, rtol, atol, &n0);  /* _dimless_wave_term (:15-26) */
    cplx Gz = G + 2.0 / hypot(r, z);                  /* :56 */
    out[0] = creal(G);  out[1] = cimag(G);
    out[2] = creal(Gr); out[3] = cimag(Gr);
    out[4] = creal(Gz); out[5] = cimag(Gz);
    if (nboxes) { nboxes[0] = n0; nboxes[1] = n1; }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Assembly: src/matrix_assembly.jl                                           */
/* ------------------------------------------------------------------------- */

static void face_corners(const ora_mesh* m, int64_t j, double V[12]) {
    for (int k = 0; k < 4; k++) {
        int64_t v = m->faces[4 * j + k] - 1;
        for (int c = 0; c < 3; c++) V[3 * k + c] = m->vertices[3 * v + c];
    }
}

typedef struct {
    double T[9], qgc[3], lc[8], area;
} ora_geom;

static ora_geom* source_geoms(const ora_mesh* m) {
    ora_geom* g = (ora_geom*)malloc(sizeof(ora_geom) * (size_t)m->nf);
    for (int64_t j = 0; j < m->nf; j++) {
        double V[12];
        face_corners(m, j, V);
        ora_birk_panel_geometry(V, g[j].T, g[j].qgc, g[j].lc, &g[j].area);
    }
    return g;
}

/* one Rankine + image entry (assemble_birk_rankine inner loop, matrix_assembly.jl:301-324) */
/* fixed: the all_normals replacement of the panel normal (forward speed, matrix_assembly.jl:224-232),
 * NULL for the panel normal */
static void rankine_entry_n(const ora_mesh* m, const ora_geom* g, int64_t i, int64_t j, int direct,
                            const double* fixed, double* s_out, double* d_out);

static void rankine_entry(const ora_mesh* m, const ora_geom* g, int64_t i, int64_t j, int direct,
                          double* s_out, double* d_out) {
    rankine_entry_n(m, g, i, j, direct, NULL, s_out, d_out);
}

static void rankine_entry_n(const ora_mesh* m, const ora_geom* g, int64_t i, int64_t j, int direct,
                            const double* fixed, double* s_out, double* d_out) {
    const double* c1 = m->centers + 3 * i;
    const double* c2 = m->centers + 3 * j;
    const double* nvec = fixed ? fixed : (direct ? m->normals + 3 * j : m->normals + 3 * i);
    double phi, gs[3];
    ora_vrankine_from_geom(c1, c2, m->radii[j], m->areas[j], g->T, g->qgc, g->lc, &phi, gs);
    double s = 0.0, d = 0.0;
    s += phi;
    if (direct)
        d += nvec[0] * gs[0] + nvec[1] * gs[1] + nvec[2] * gs[2];
    else
        d += nvec[0] * (-gs[0]) + nvec[1] * (-gs[1]) + nvec[2] * (-gs[2]);
    double c1r[3] = {c1[0], c1[1], -c1[2]};
    double phir, gr[3];
    ora_vrankine_from_geom(c1r, c2, m->radii[j], m->areas[j], g->T, g->qgc, g->lc, &phir, gr);
    double grr[3];
    if (direct) {
        grr[0] = gr[0]; grr[1] = gr[1]; grr[2] = gr[2];
    } else { /* vertical_reflection(-g) (rankine_reflected.jl:113-124) */
        grr[0] = -gr[0]; grr[1] = -gr[1]; grr[2] = gr[2];
    }
    s += 1.0 * phir;
    d += 1.0 * (nvec[0] * grr[0] + nvec[1] * grr[1] + nvec[2] * grr[2]);
    double scale = -1.0 / (4 * PI);
    *s_out = scale * s;
    *d_out = scale * d;
}

/* reduced coordinates (wu.jl:7-18, abstract_greens_function.jl:117-138) */
static void reduced_hv(const double* x, const double* xi, double k, double* hh, double* vv) {
    double d2 = (x[0] - xi[0]) * (x[0] - xi[0]) + (x[1] - xi[1]) * (x[1] - xi[1]);
    double purt = 1e-18, dis;
    if (d2 == 0)
        dis = sqrt(d2 + purt * purt) - purt;
    else
        dis = sqrt(d2);
    *hh = k * dis;
    *vv = k * (x[2] + xi[2]);
}

/* one wave-term entry: Wu (_wu_sd, wu.jl:112-127) or exact Delhommeau
 * (with_reduced_coordinates_derivative, abstract_greens_function.jl:140-185 +
 *  integral_gradient, exact_Guevel_Delhommeau.jl:61-69) */
static void wave_entry_n(const ora_mesh* m, int64_t i, int64_t j, double k, int gf, int direct, const double* fixed,
                         cplx* s_out, cplx* d_out);

static void wave_entry(const ora_mesh* m, int64_t i, int64_t j, double k, int gf, int direct, cplx* s_out,
                       cplx* d_out) {
    wave_entry_n(m, i, j, k, gf, direct, NULL, s_out, d_out);
}

static void wave_entry_n(const ora_mesh* m, int64_t i, int64_t j, double k, int gf, int direct, const double* fixed,
                         cplx* s_out, cplx* d_out) {
    const double* x = m->centers + 3 * i;
    const double* xi = m->centers + 3 * j;
    const double* nvec = fixed ? fixed : (direct ? m->normals + 3 * j : m->normals + 3 * i);
    double a2 = m->areas[j];
    double hh, vv;
    reduced_hv(x, xi, k, &hh, &vv);
    double o[6];
    if (gf == ORA_GF_WU)
        ora_wu_wave_term_and_derivs(hh, vv, o);
    else
        ora_delhommeau_term(hh, vv, 1e-8, 1e-8, o, NULL);
    cplx G = o[0] + I * o[1], Gh = o[2] + I * o[3], Gv = o[4] + I * o[5];
    double dx1 = direct ? xi[0] - x[0] : x[0] - xi[0];
    double dx2 = direct ? xi[1] - x[1] : x[1] - xi[1];
    double dr1 = 0.0, dr2 = 0.0;
    if (fabs(hh) > 1e-6) {
        dr1 = k * k / hh * dx1;
        dr2 = k * k / hh * dx2;
    }
    cplx g0 = k * (dr1 * Gh), g1 = k * (dr2 * Gh), g2 = k * (k * Gv);
    cplx scale = -1.0 / (4 * PI);
    if (gf == ORA_GF_WU) {
        cplx s = (k * G) * a2;
        cplx d = (nvec[0] * g0 + nvec[1] * g1 + nvec[2] * g2) * a2;
        *s_out = scale * s;
        *d_out = scale * d;
    } else {
        cplx s = k * G * a2;
        cplx d = nvec[0] * (g0 * a2) + nvec[1] * (g1 * a2) + nvec[2] * (g2 * a2);
        *s_out = scale * s;
        *d_out = scale * d;
    }
}

int ora_assemble_rankine(const ora_mesh* m, int direct, int include_identity, double* S, double* D) {
    int64_t n = m->nf;
    ora_geom* g = source_geoms(m);
    cplx* Sc = (cplx*)S;
    cplx* Dc = (cplx*)D;
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads)
    for (int64_t j = 0; j < n; j++) {
        for (int64_t i = 0; i < n; i++) {
            double s, d;
            rankine_entry(m, &g[j], i, j, direct, &s, &d);
            Sc[i + n * j] = s;
            Dc[i + n * j] = d;
        }
    }
    if (include_identity)
        for (int64_t i = 0; i < n; i++) Dc[i + n * i] += 0.5;
    free(g);
    return 0;
}

int ora_assemble_wave(const ora_mesh* m, double k, int gf, int direct, double* S, double* D) {
    int64_t n = m->nf;
    cplx* Sc = (cplx*)S;
    cplx* Dc = (cplx*)D;
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = 0; i < n; i++) wave_entry(m, i, j, k, gf, direct, &Sc[i + n * j], &Dc[i + n * j]);
    return 0;
}

/* _assemble_cpu (matrix_assembly.jl:183-208): S = S_R + S_w, D = D_R(+1/2 I) + D_w */
int ora_assemble(const ora_mesh* m, double k, int gf, int direct, int include_identity, double* S, double* D) {
    return ora_assemble_rows(m, k, gf, direct, include_identity, 0, m->nf, S, D);
}

/* identity jump of D's diagonal: 1/2 (_add_identity!, matrix_assembly.jl:234-245); with a lid,
 * the comprehension path's rule: 1 for a panel in the free surface (matrix_assembly.jl:25-27) */
static double identity_jump(const ora_mesh* m, int64_t i, int surface_rule) {
    return (surface_rule && fabs(m->centers[3 * i + 2]) < 1e-8) ? 1.0 : 0.5;
}

static int assemble_rows_impl(const ora_mesh* m, double k, int gf, int direct, int include_identity, int surface_rule,
                              const double* fixed, int64_t i0, int64_t i1, double* S, double* D);

int ora_assemble_rows(const ora_mesh* m, double k, int gf, int direct, int include_identity, int64_t i0,
                      int64_t i1, double* S, double* D) {
    return assemble_rows_impl(m, k, gf, direct, include_identity, 0, NULL, i0, i1, S, D);
}

/* assemble_matrices(...; all_normals = fixed) (matrix_assembly.jl:183-245): the panel normal in
 * D / K replaced by the fixed vector, the identity jump scaled by (fixed . n_i) */
int ora_assemble_fixed_normal(const ora_mesh* m, double k, int gf, int direct, const double* fixed, double* S,
                              double* D) {
    return assemble_rows_impl(m, k, gf, direct, 1, 0, fixed, 0, m->nf, S, D);
}

static int assemble_rows_impl(const ora_mesh* m, double k, int gf, int direct, int include_identity, int surface_rule,
                              const double* fixed, int64_t i0, int64_t i1, double* S, double* D) {
    int64_t n = m->nf, nr = i1 - i0;
    if (i0 < 0 || i1 > n || nr < 0) return -1;
    ora_geom* g = source_geoms(m);
    cplx* Sc = (cplx*)S;
    cplx* Dc = (cplx*)D;
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
    for (int64_t j = 0; j < n; j++) {
        for (int64_t i = i0; i < i1; i++) {
            double sr, dr;
            rankine_entry_n(m, &g[j], i, j, direct, fixed, &sr, &dr);
            cplx sR = sr, dR = dr;
            if (include_identity && i == j) {
                const double* ni = m->normals + 3 * i;
                const double proj = fixed ? fixed[0] * ni[0] + fixed[1] * ni[1] + fixed[2] * ni[2] : 1.0;
                dR += identity_jump(m, i, surface_rule) * proj;
            }
            cplx sw, dw;
            wave_entry_n(m, i, j, k, gf, direct, fixed, &sw, &dw);
            Sc[(i - i0) + nr * j] = sR + sw;
            Dc[(i - i0) + nr * j] = dR + dw;
        }
    }
    free(g);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Dense complex LU (LinearSolve LUFactorization -> LAPACK zgetrf/zgetrs,     */
/* matrix_assembly.jl:411-416).  Right-looking, partial pivoting on |Re|+|Im|. */
/* ------------------------------------------------------------------------- */

static inline double cabs1(cplx a) { return fabs(creal(a)) + fabs(cimag(a)); }

int ora_zgetrf(int64_t n, double* Ad, int64_t lda, int64_t* ipiv) {
    cplx* A = (cplx*)Ad;
    int info = 0;
    for (int64_t k = 0; k < n; k++) {
        int64_t p = k;
        double best = cabs1(A[k + lda * k]);
        for (int64_t i = k + 1; i < n; i++) {
            double v = cabs1(A[i + lda * k]);
            if (v > best) { best = v; p = i; }
        }
        ipiv[k] = p;
        if (best == 0.0) {
            if (!info) info = (int)(k + 1);
            continue;
        }
        if (p != k)
            for (int64_t j = 0; j < n; j++) {
                cplx t = A[k + lda * j];
                A[k + lda * j] = A[p + lda * j];
                A[p + lda * j] = t;
            }
        cplx inv = 1.0 / A[k + lda * k];
        for (int64_t i = k + 1; i < n; i++) A[i + lda * k] *= inv;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (n - k > 256)
        for (int64_t j = k + 1; j < n; j++) {
            cplx ukj = A[k + lda * j];
            if (ukj == 0) continue;
            cplx* colj = A + lda * j;
            const cplx* colk = A + lda * k;
            for (int64_t i = k + 1; i < n; i++) colj[i] -= colk[i] * ukj;
        }
    }
    return info;
}

/* trans: 0 -> A x = b, 2 -> A^H x = b */
int ora_zgetrs(int trans, int64_t n, int64_t nrhs, const double* Ad, int64_t lda, const int64_t* ipiv, double* Bd,
               int64_t ldb) {
    const cplx* A = (const cplx*)Ad;
    cplx* B = (cplx*)Bd;
    for (int64_t c = 0; c < nrhs; c++) {
        cplx* b = B + ldb * c;
        if (trans == 0) {
            for (int64_t k = 0; k < n; k++)
                if (ipiv[k] != k) { cplx t = b[k]; b[k] = b[ipiv[k]]; b[ipiv[k]] = t; }
            for (int64_t j = 0; j < n; j++) { /* L unit lower */
                cplx bj = b[j];
                for (int64_t i = j + 1; i < n; i++) b[i] -= A[i + lda * j] * bj;
            }
            for (int64_t j = n - 1; j >= 0; j--) { /* U */
                b[j] /= A[j + lda * j];
                cplx bj = b[j];
                for (int64_t i = 0; i < j; i++) b[i] -= A[i + lda * j] * bj;
            }
        } else {
            for (int64_t j = 0; j < n; j++) { /* U^H lower */
                cplx s = b[j];
                for (int64_t i = 0; i < j; i++) s -= conj(A[i + lda * j]) * b[i];
                b[j] = s / conj(A[j + lda * j]);
            }
            for (int64_t j = n - 1; j >= 0; j--) { /* L^H unit upper */
                cplx s = b[j];
                for (int64_t i = j + 1; i < n; i++) s -= conj(A[i + lda * j]) * b[i];
                b[j] = s;
            }
            for (int64_t k = n - 1; k >= 0; k--)
                if (ipiv[k] != k) { cplx t = b[k]; b[k] = b[ipiv[k]]; b[ipiv[k]] = t; }
        }
    }
    return 0;
}

static void matmul_cols(int64_t n, int64_t nrhs, const cplx* S, const cplx* X, cplx* Y) {
    /* _mulvec (matrix_assembly.jl:380-390) */
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int64_t i = 0; i < n; i++)
        for (int64_t c = 0; c < nrhs; c++) {
            cplx s = 0;
            for (int64_t j = 0; j < n; j++) s += S[i + n * j] * X[j + n * c];
            Y[i + n * c] = s;
        }
}

/* solve(D, S, bc; direct) (matrix_assembly.jl:436-449) */
int ora_solve(int64_t n, int64_t nrhs, const double* Dd, const double* Sd, const double* bcd, int direct,
              double* phid, double* srcd) {
    const cplx* S = (const cplx*)Sd;
    const cplx* bc = (const cplx*)bcd;
    cplx* phi = (cplx*)phid;
    cplx* LU = (cplx*)malloc(sizeof(cplx) * n * n);
    int64_t* ipiv = (int64_t*)malloc(sizeof(int64_t) * n);
    memcpy(LU, Dd, sizeof(cplx) * n * n);
    int info = ora_zgetrf(n, (double*)LU, n, ipiv);
    if (direct) {
        matmul_cols(n, nrhs, S, bc, phi);
        ora_zgetrs(0, n, nrhs, (double*)LU, n, ipiv, (double*)phi, n);
    } else {
        cplx* src = (cplx*)srcd;
        memcpy(src, bc, sizeof(cplx) * n * nrhs);
        ora_zgetrs(0, n, nrhs, (double*)LU, n, ipiv, (double*)src, n);
        matmul_cols(n, nrhs, S, src, phi);
    }
    free(LU);
    free(ipiv);
    return info;
}

/* ------------------------------------------------------------------------- */
/* DOFs, boundary conditions, pressure: src/bodies.jl, src/waves.jl           */
/* ------------------------------------------------------------------------- */

/* translational_dofs / rotational_dofs (bodies.jl:74-118); which: 0..5 = Surge..Yaw */
void ora_rigid_dof(const ora_mesh* m, int which, const double rc[3], double* dof) {
    int64_t n = m->nf;
    for (int64_t i = 0; i < n; i++) {
        double v[3] = {0, 0, 0};
        if (which < 3) {
            v[which] = 1.0;
        } else {
            double ax[3] = {0, 0, 0};
            ax[which - 3] = 1.0;
            double p[3] = {m->centers[3 * i] - rc[0], m->centers[3 * i + 1] - rc[1], m->centers[3 * i + 2] - rc[2]};
            cross3(ax, p, v);
        }
        for (int c = 0; c < 3; c++) dof[i + n * c] = v[c];
    }
}

/* radiation_bc (waves.jl:183-211): bc = -i omega (n . dof) */
void ora_radiation_bc(const ora_mesh* m, const double* dof, double omega, double* bcd) {
    int64_t n = m->nf;
    cplx* bc = (cplx*)bcd;
    cplx sc = (-1.0 * I) * omega;
    for (int64_t i = 0; i < n; i++) {
        const double* nv = m->normals + 3 * i;
        double disp = nv[0] * dof[i] + nv[1] * dof[i + n] + nv[2] * dof[i + 2 * n];
        bc[i] = sc * disp;
    }
}

/* _airy_bc (waves.jl:266-282): bc = -(v . n), no conjugation */
void ora_airy_bc(const ora_mesh* m, double omega, double k, double beta, double g, double* bcd) {
    cplx* bc = (cplx*)bcd;
    for (int64_t i = 0; i < m->nf; i++) {
        const double* p = m->centers + 3 * i;
        const double* nv = m->normals + 3 * i;
        double wbar = p[0] * cos(beta) + p[1] * sin(beta);
        double cih = exp(k * p[2]);
        cplx amp = g * k / omega * cexp(I * (k * wbar));
        cplx v0 = amp * cos(beta) * cih, v1 = amp * sin(beta) * cih, v2 = (-1.0 * I) * amp * cih;
        bc[i] = -(v0 * nv[0] + v1 * nv[1] + v2 * nv[2]);
    }
}

/* _airy_pressure (waves.jl:259-264, 301-303): p = i omega rho * (-i g/omega) e^{kz} e^{i k wbar} */
void ora_airy_pressure(const ora_mesh* m, double omega, double k, double beta, double g, double rho, double* pd) {
    cplx* p = (cplx*)pd;
    for (int64_t i = 0; i < m->nf; i++) {
        const double* c = m->centers + 3 * i;
        double wbar = c[0] * cos(beta) + c[1] * sin(beta);
        cplx phi = (-1.0 * I) * g / omega * exp(k * c[2]) * cexp(I * (k * wbar));
        p[i] = I * omega * rho * phi;
    }
}

/* _integrate_dof_pressure (waves.jl:170-178): F = sum p_i * -(dof_i . n_i) * a_i */
void ora_integrate_pressure(const ora_mesh* m, const double* dof, const double* pd, double out[2]) {
    int64_t n = m->nf;
    const cplx* p = (const cplx*)pd;
    cplx s = 0;
    for (int64_t i = 0; i < n; i++) {
        const double* nv = m->normals + 3 * i;
        double disp = dof[i] * nv[0] + dof[i + n] * nv[1] + dof[i + 2 * n] * nv[2];
        s += p[i] * -disp * m->areas[i];
    }
    out[0] = creal(s);
    out[1] = cimag(s);
}

/* ------------------------------------------------------------------------- */
/* hydrodynamic_coefficients (problems_and_results.jl:464-495, solve.jl:47-78) */
/* All dofs radiate and are influenced; one DiffractionProblem per heading.  */
/* ------------------------------------------------------------------------- */

static int hydro_coeffs_impl(const ora_mesh* m, int64_t nhull, int surface_rule, int64_t ndof, const double* dofs,
                             int64_t nw, const double* omegas, int64_t nb, const double* betas, double rho, double g,
                             int gf, int direct, int mode, double* A, double* B, double* Fdd, double* Fkd);

int ora_hydro_coeffs(const ora_mesh* m, int64_t ndof, const double* dofs, int64_t nw, const double* omegas,
                     int64_t nb, const double* betas, double rho, double g, int gf, int direct, int mode, double* A,
                     double* B, double* Fdd, double* Fkd) {
    return hydro_coeffs_impl(m, m->nf, 0, ndof, dofs, nw, omegas, nb, betas, rho, g, gf, direct, mode, A, B, Fdd, Fkd);
}

/* With a lid (solve.jl:4-11, 22-28; bodies.jl:9-28): m = hull panels [0, nhull) + lid panels,
 * dofs zero on the lid; the hull's boundary conditions are padded with zeros on the lid
 * (_pad_hull_bc), the assembly uses the free-surface jump rule, and the pressure is integrated
 * over the hull (the dofs vanish on the lid). */
int ora_hydro_coeffs_lid(const ora_mesh* m, int64_t nhull, int64_t ndof, const double* dofs, int64_t nw,
                         const double* omegas, int64_t nb, const double* betas, double rho, double g, int gf,
                         int direct, int mode, double* A, double* B, double* Fdd, double* Fkd) {
    if (nhull <= 0 || nhull > m->nf) return -1;
    return hydro_coeffs_impl(m, nhull, 1, ndof, dofs, nw, omegas, nb, betas, rho, g, gf, direct, mode, A, B, Fdd, Fkd);
}

static int hydro_coeffs_impl(const ora_mesh* m, int64_t nhull, int surface_rule, int64_t ndof, const double* dofs,
                             int64_t nw, const double* omegas, int64_t nb, const double* betas, double rho, double g,
                             int gf, int direct, int mode, double* A, double* B, double* Fdd, double* Fkd) {
    int64_t n = m->nf;
    int64_t R = ndof + nb;
    cplx* Fd = (cplx*)Fdd;
    cplx* Fk = (cplx*)Fkd;
    cplx* S = (cplx*)malloc(sizeof(cplx) * n * n);
    cplx* D = (cplx*)malloc(sizeof(cplx) * n * n);
    cplx* bc = (cplx*)malloc(sizeof(cplx) * n * R);
    cplx* phi = (cplx*)malloc(sizeof(cplx) * n * R);
    cplx* src = (cplx*)malloc(sizeof(cplx) * n * R);
    cplx* p = (cplx*)malloc(sizeof(cplx) * n);
    int info = 0;
    for (int64_t iw = 0; iw < nw; iw++) {
        double w = omegas[iw];
        double k = w * w / g; /* compute_wavenumber (waves.jl:133-134) */
        for (int64_t r = 0; r < ndof; r++) ora_radiation_bc(m, dofs + 3 * n * r, w, (double*)(bc + n * r));
        for (int64_t ib = 0; ib < nb; ib++) {
            ora_airy_bc(m, w, k, betas[ib], g, (double*)(bc + n * (ndof + ib)));
            for (int64_t i = nhull; i < n; i++) bc[i + n * (ndof + ib)] = 0.0; /* _pad_hull_bc */
        }
        if (mode == 0) {
            assemble_rows_impl(m, k, gf, direct, 1, surface_rule, NULL, 0, n, (double*)S, (double*)D);
            int e = ora_solve(n, R, (double*)D, (double*)S, (double*)bc, direct, (double*)phi, (double*)src);
            if (e && !info) info = e;
        } else {
            for (int64_t r = 0; r < R; r++) {
                assemble_rows_impl(m, k, gf, direct, 1, surface_rule, NULL, 0, n, (double*)S, (double*)D);
                int e = ora_solve(n, 1, (double*)D, (double*)S, (double*)(bc + n * r), direct,
                                  (double*)(phi + n * r), (double*)(src + n * r));
                if (e && !info) info = e;
            }
        }
        for (int64_t r = 0; r < R; r++) {
            for (int64_t i = 0; i < n; i++) p[i] = I * rho * w * phi[i + n * r]; /* solve.jl:50 */
            for (int64_t d = 0; d < ndof; d++) {
                double F[2];
                ora_integrate_pressure(m, dofs + 3 * n * d, (double*)p, F);
                if (r < ndof) { /* added_mass / radiation_damping (problems_and_results.jl:207-208) */
                    A[d + ndof * (r + ndof * iw)] = F[0] / (w * w);
                    B[d + ndof * (r + ndof * iw)] = F[1] / w;
                } else {
                    Fd[d + ndof * (iw + nw * (r - ndof))] = F[0] + I * F[1];
                }
            }
        }
        for (int64_t ib = 0; ib < nb; ib++) { /* froude_krylov_force (waves.jl:294-303) */
            ora_airy_pressure(m, w, k, betas[ib], g, rho, (double*)p);
            for (int64_t d = 0; d < ndof; d++) {
                double F[2];
                ora_integrate_pressure(m, dofs + 3 * n * d, (double*)p, F);
                Fk[d + ndof * (iw + nw * ib)] = F[0] + I * F[1];
            }
        }
    }
    free(S); free(D); free(bc); free(phi); free(src); free(p);
    return info;
}


/* ------------------------------------------------------------------------- */
/* forward speed (IndirectBEM; waves.jl:136-151, 183-199, 259-282; solve.jl:35-55;        */
/* problems_and_results.jl:202-208, 236-380)                                               */
/* ------------------------------------------------------------------------- */

/* compute_encountered_values (waves.jl:136-151) */
void ora_encountered(double omega, double beta, double U, double g, double out[3]) {
    double k = omega * omega / g;
    if (U == 0) {
        out[0] = omega; out[1] = k; out[2] = beta;
        return;
    }
    double doppler = omega - k * U * cos(beta);
    double we = fabs(doppler);
    out[0] = we;
    out[1] = we * we / g;
    out[2] = doppler >= 0 ? beta : beta + PI;
}

/* Every (U, omega, beta) group: radiation problems of all dofs and the diffraction problem at
 * heading beta share the encountered (omega_e, k_e).  dgrad: ndof x 3 gradient of motion
 * (evaluate_gradient_of_motion, bodies.jl:55-66: Pitch (0,0,1), Yaw (0,-1,0), else 0).
 * Outputs: A, B (ndof, ndof, nw, nU, nb); Fd, Fk (ndof, nw, nb, nU). */
int ora_hydro_coeffs_fs(const ora_mesh* m, int64_t ndof, const double* dofs, const double* dgrad, int64_t nw,
                        const double* omegas, int64_t nb, const double* betas, int64_t nU, const double* Us,
                        double rho, double g, int gf, double* A, double* B, double* Fdd, double* Fkd) {
    int64_t n = m->nf;
    int64_t R = ndof + 1;
    cplx* Fd = (cplx*)Fdd;
    cplx* Fk = (cplx*)Fkd;
    cplx* S = (cplx*)malloc(sizeof(cplx) * n * n);
    cplx* K = (cplx*)malloc(sizeof(cplx) * n * n);
    cplx* Kx = (cplx*)malloc(sizeof(cplx) * n * n);
    cplx* bc = (cplx*)malloc(sizeof(cplx) * n * R);
    cplx* phi = (cplx*)malloc(sizeof(cplx) * n * R);
    cplx* src = (cplx*)malloc(sizeof(cplx) * n * R);
    cplx* p = (cplx*)malloc(sizeof(cplx) * n);
    const double ex[3] = {1.0, 0.0, 0.0};
    int info = 0;
    for (int64_t iu = 0; iu < nU; iu++)
        for (int64_t iw = 0; iw < nw; iw++)
            for (int64_t ib = 0; ib < nb; ib++) {
                const double U = Us[iu], w = omegas[iw], beta = betas[ib];
                const double k = w * w / g;
                double enc[3];
                ora_encountered(w, beta, U, g, enc);
                const double we = enc[0], ke = enc[1], be = enc[2];
                /* radiation_bc (waves.jl:183-199): -i omega_e (n . dof) + U (n . d dof / dx) */
                for (int64_t r = 0; r < ndof; r++) {
                    ora_radiation_bc(m, dofs + 3 * n * r, we, (double*)(bc + n * r));
                    if (U != 0)
                        for (int64_t i = 0; i < n; i++) {
                            const double* nv = m->normals + 3 * i;
                            const double* gr = dgrad + 3 * r;
                            bc[i + n * r] += U * (nv[0] * gr[0] + nv[1] * gr[1] + nv[2] * gr[2]);
                        }
                }
                /* _airy_bc (waves.jl:266-282): incident k and omega, phase along the encountered
                 * heading, velocity along the incident heading */
                for (int64_t i = 0; i < n; i++) {
                    const double* c = m->centers + 3 * i;
                    const double* nv = m->normals + 3 * i;
                    double wbar = c[0] * cos(be) + c[1] * sin(be);
                    double cih = exp(k * c[2]);
                    cplx amp = g * k / w * cexp(I * (k * wbar));
                    cplx v0 = amp * cos(beta) * cih, v1 = amp * sin(beta) * cih, v2 = (-1.0 * I) * amp * cih;
                    bc[i + n * ndof] = -(v0 * nv[0] + v1 * nv[1] + v2 * nv[2]);
                }
                assemble_rows_impl(m, ke, gf, 0, 1, 0, NULL, 0, n, (double*)S, (double*)K);
                int e = ora_solve(n, R, (double*)K, (double*)S, (double*)bc, 0, (double*)phi, (double*)src);
                if (e && !info) info = e;
                if (U != 0) /* _maybe_forward_speed_pressure (solve.jl:35-45): dphi/dx = K_x sigma */
                    assemble_rows_impl(m, ke, gf, 0, 1, 0, ex, 0, n, (double*)S, (double*)Kx);
                for (int64_t r = 0; r < R; r++) {
                    for (int64_t i = 0; i < n; i++) {
                        p[i] = I * rho * we * phi[i + n * r]; /* solve.jl:50 */
                        if (U != 0) {
                            cplx dphidx = 0;
                            for (int64_t j = 0; j < n; j++) dphidx += Kx[i + n * j] * src[j + n * r];
                            p[i] += rho * U * dphidx;
                        }
                    }
                    for (int64_t d = 0; d < ndof; d++) {
                        double F[2];
                        ora_integrate_pressure(m, dofs + 3 * n * d, (double*)p, F);
                        if (r < ndof) { /* problems_and_results.jl:202-208, 341-353 */
                            const int64_t o = d + ndof * (r + ndof * (iw + nw * (iu + nU * ib)));
                            A[o] = F[0] / (we * we);
                            B[o] = F[1] / we;
                        } else {
                            Fd[d + ndof * (iw + nw * (ib + nb * iu))] = F[0] + I * F[1];
                        }
                    }
                }
                /* froude_krylov_force (waves.jl:294-303): incident omega and k, encountered heading */
                ora_airy_pressure(m, w, k, be, g, rho, (double*)p);
                for (int64_t d = 0; d < ndof; d++) {
                    double F[2];
                    ora_integrate_pressure(m, dofs + 3 * n * d, (double*)p, F);
                    Fk[d + ndof * (iw + nw * (ib + nb * iu))] = F[0] + I * F[1];
                }
            }
    free(S); free(K); free(Kx); free(bc); free(phi); free(src); free(p);
    return info;
}
