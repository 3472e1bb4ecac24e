// Device-side math for the BEM hot path on gfx950 (FP64).
//
//  * Birk (2021) constant-source quad panel potential and velocity
//    (reference: src/green_functions/rankine_vectorized.jl:42-137, 241-287)
//  * Liang-Wu-Noblesse wave term fits (src/green_functions/wu.jl:7-37, 131-435)
//  * complex exponential integral and the exact Guevel-Delhommeau wave term
//    integrated by adaptive Gauss-Kronrod (7,15) bisection with the
//    termination rule of HCubature's hquadrature
//    (src/green_functions/exact_Guevel_Delhommeau.jl:8-59)
//
// Every formula keeps the reference's branch structure (thresholds, signed-zero
// behaviour of the atan2 rewrite, the 1e-18 perturbation of coincident
// horizontal positions) so that results agree with the CPU restatement to
// rounding.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MHX_PI 3.14159265358979323846
#define MHX_EULER 0.57721566490153286060651209

typedef double2 cplx;
typedef float2 cplxf;

__device__ __forceinline__ cplx cmk(double r, double i) { return make_double2(r, i); }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return cmk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return cmk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// acc += t x and acc -= t x, four FMAs each
__device__ __forceinline__ void cfma(cplx& acc, cplx t, cplx x) {
    acc.x = fma(t.x, x.x, acc.x);
    acc.x = fma(-t.y, x.y, acc.x);
    acc.y = fma(t.x, x.y, acc.y);
    acc.y = fma(t.y, x.x, acc.y);
}
__device__ __forceinline__ void cfms(cplx& acc, cplx t, cplx x) {
    acc.x = fma(-t.x, x.x, acc.x);
    acc.x = fma(t.y, x.y, acc.x);
    acc.y = fma(-t.x, x.y, acc.y);
    acc.y = fma(-t.y, x.x, acc.y);
}
__device__ __forceinline__ cplx cscale(double s, cplx a) { return cmk(s * a.x, s * a.y); }
__device__ __forceinline__ double cabs_(cplx a) { return hypot(a.x, a.y); }
// Smith's division
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
    if (fabs(b.x) >= fabs(b.y)) {
        double r = b.y / b.x, den = b.x + b.y * r;
        return cmk((a.x + a.y * r) / den, (a.y - a.x * r) / den);
    }
    double r = b.x / b.y, den = b.x * r + b.y;
    return cmk((a.x * r + a.y) / den, (a.y * r - a.x) / den);
}
__device__ __forceinline__ cplx cexp_(cplx z) {
    double s, c;
    sincos(z.y, &s, &c);
    double e = exp(z.x);
    return cmk(e * c, e * s);
}
__device__ __forceinline__ cplx clog_(cplx z) { return cmk(log(hypot(z.x, z.y)), atan2(z.y, z.x)); }

// ---------------------------------------------------------------------------
// Source-panel geometry, precomputed once per panel and staged in LDS.
// ---------------------------------------------------------------------------
struct PanelGeom {
    double T[9];    // column-major [s t n] local frame (bimedian construction)
    double qgc[3];  // area centroid
    double lc[8];   // local corners (x, y) relative to qgc
    double c[3];    // mesh centre (far-field source point)
    double radius;
    double area;    // mesh area (far field and wave term)
    double nrm[3];  // mesh normal
    // pad[0]: the identity jump of D's diagonal (1/2; 1 on a panel in the free surface of a body
    // with a lid), pad[1]: 1 on hull panels, 0 on lid panels (no incident-wave condition there)
    double pad[4];
};
static_assert(sizeof(PanelGeom) == 32 * sizeof(double), "PanelGeom must be 256 bytes");

__device__ __forceinline__ double atan_ratio(double num, double den) {
    double flipped = (den < 0.0) ? -num : num;
    double ad = (den < 0.0) ? -den : den;
    return atan2(flipped, ad);
}

__device__ __forceinline__ void birk_panel_geometry(const double V[12], PanelGeom& g) {
    double mid[4][3];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int j = (i + 1) & 3;
#pragma unroll
        for (int c = 0; c < 3; c++) mid[i][c] = (V[3 * i + c] + V[3 * j + c]) / 2;
    }
    double mgc[3], s[3], tt[3], n[3], t[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        mgc[c] = (mid[1][c] + mid[3][c]) / 2;
        s[c] = mid[1][c] - mid[3][c];
        tt[c] = mid[2][c] - mid[0][c];
    }
    double ns = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (int c = 0; c < 3; c++) s[c] /= ns;
    n[0] = s[1] * tt[2] - s[2] * tt[1];
    n[1] = s[2] * tt[0] - s[0] * tt[2];
    n[2] = s[0] * tt[1] - s[1] * tt[0];
    double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int c = 0; c < 3; c++) n[c] /= nn;
    t[0] = n[1] * s[2] - n[2] * s[1];
    t[1] = n[2] * s[0] - n[0] * s[2];
    t[2] = n[0] * s[1] - n[1] * s[0];
    double nt = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    for (int c = 0; c < 3; c++) t[c] /= nt;
    for (int c = 0; c < 3; c++) {
        g.T[c] = s[c];
        g.T[3 + c] = t[c];
        g.T[6 + c] = n[c];
    }
    double tc[4][2];
    for (int k = 0; k < 4; k++) {
        double d0 = V[3 * k] - mgc[0], d1 = V[3 * k + 1] - mgc[1], d2 = V[3 * k + 2] - mgc[2];
        tc[k][0] = d0 * s[0] + d1 * s[1] + d2 * s[2];
        tc[k][1] = d0 * t[0] + d1 * t[1] + d2 * t[2];
    }
    double A = 0, Mx = 0, My = 0;
    for (int k = 0; k < 4; k++) {
        int k1 = (k + 1) & 3;
        double xk = tc[k][0], yk = tc[k][1], xk1 = tc[k1][0], yk1 = tc[k1][1];
        A += (yk1 - yk) * (xk + xk1) / 2;
        Mx += (xk1 - xk) * (yk * yk + yk * yk1 + yk1 * yk1) / 6;
        My += (yk1 - yk) * (xk * xk + xk * xk1 + xk1 * xk1) / 6;
    }
    double clx = My / A, cly = -Mx / A;
    for (int c = 0; c < 3; c++) g.qgc[c] = mgc[c] + clx * s[c] + cly * t[c];
    for (int k = 0; k < 4; k++) {
        g.lc[2 * k] = tc[k][0] - clx;
        g.lc[2 * k + 1] = tc[k][1] - cly;
    }
}

// birk_phi_and_velocity (rankine_vectorized.jl:241-267), including the -1/(4 pi) scale.
__device__ __forceinline__ void birk_phi_v(double x, double y, double z, const double* lc, double& phi,
                                           double v[3]) {
    double sx[4], sy[4], r[4], e[4], h[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        sx[k] = x - lc[2 * k];
        sy[k] = y - lc[2 * k + 1];
        r[k] = sqrt(sx[k] * sx[k] + sy[k] * sy[k] + z * z);
        e[k] = sx[k] * sx[k] + z * z;
        h[k] = sx[k] * sy[k];
    }
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    const bool inplane = fabs(z) <= 1e-14;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int kn = (k + 1) & 3;
        double dx = lc[2 * kn] - lc[2 * k], dy = lc[2 * kn + 1] - lc[2 * k + 1];
        double dk = hypot(dx, dy);
        double rsum = r[k] + r[kn];
        double lt_dk = log((rsum - dk) / (rsum + dk)) / dk;
        double datan = atan_ratio(dy * e[k] - dx * h[k], dx * z * r[k]) -
                       atan_ratio(dy * e[kn] - dx * h[kn], dx * z * r[kn]);
        double vzk = inplane ? (datan - datan) : datan;
        double phik = (sx[k] * dy - sy[k] * dx) * lt_dk + z * vzk;
        if (!(dk <= 1e-14)) {
            s0 += phik;
            s1 += dy * lt_dk;
            s2 += -dx * lt_dk;
            s3 += vzk;
        }
    }
    const double scale = -1.0 / (4 * MHX_PI);
    phi = scale * s0;
    v[0] = scale * s1;
    v[1] = scale * s2;
    v[2] = scale * s3;
}

// _vrankine_from_geom (rankine_vectorized.jl:273-287): integral of 1/|p - xi| over
// the source panel and its gradient with respect to the source point.
__device__ __forceinline__ void vrankine(const double p[3], const PanelGeom& g, double& phi, double grad[3]) {
    double rb0 = p[0] - g.c[0], rb1 = p[1] - g.c[1], rb2 = p[2] - g.c[2];
    double r = sqrt(rb0 * rb0 + rb1 * rb1 + rb2 * rb2);
    if (r > 7 * g.radius) {
        double r3 = r * r * r;
        phi = g.area / r;
        grad[0] = (rb0 / r3) * g.area;
        grad[1] = (rb1 / r3) * g.area;
        grad[2] = (rb2 / r3) * g.area;
        return;
    }
    double d0 = p[0] - g.qgc[0], d1 = p[1] - g.qgc[1], d2 = p[2] - g.qgc[2];
    double fl0 = g.T[0] * d0 + g.T[1] * d1 + g.T[2] * d2;
    double fl1 = g.T[3] * d0 + g.T[4] * d1 + g.T[5] * d2;
    double fl2 = g.T[6] * d0 + g.T[7] * d1 + g.T[8] * d2;
    double pb, vl[3];
    birk_phi_v(fl0, fl1, fl2, g.lc, pb, vl);
    phi = -4 * MHX_PI * pb;
#pragma unroll
    for (int q = 0; q < 3; q++) grad[q] = 4 * MHX_PI * (g.T[q] * vl[0] + g.T[3 + q] * vl[1] + g.T[6 + q] * vl[2]);
}

// ---------------------------------------------------------------------------
// Liang-Wu-Noblesse fits (wu.jl:131-435)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double struve_h0(double xx) {
    if (xx <= 3.0) {
        double yy = (xx / 3.0) * (xx / 3.0);
        double h = 1.909859164 +
                   (-1.909855001 + (0.687514637 + (-0.126164557 + (0.013828813 + -0.000876918 * yy) * yy) * yy) * yy) * yy;
        return h * (xx / 3.0);
    }
    double yy = (3.0 / xx) * (3.0 / xx);
    double c1 = 2.0 * (0.99999906 + (4.77228920 + (3.85542044 + 0.32303607 * yy) * yy) * yy);
    double c2 = MHX_PI * xx * (1.0 + (4.88331068 + (4.28957333 + 0.52120508 * yy) * yy) * yy);
    return c1 / c2 + y0(xx);
}
__device__ __forceinline__ double struve_h1(double xx) {
    if (xx <= 3.0) {
        double yy = (xx / 3.0) * (xx / 3.0);
        return (1.909859286 +
                (-1.145914713 + (0.294656958 + (-0.042070508 + (0.003785727 + -0.000207183 * yy) * yy) * yy) * yy) * yy) * yy;
    }
    double yy = (3.0 / xx) * (3.0 / xx);
    double c1 = 2.0 * (1.00000004 + (3.92205313 + (2.64893033 + 0.27450895 * yy) * yy) * yy);
    double c2 = MHX_PI * (1.0 + (3.81095112 + (2.26216956 + 0.10885141 * yy) * yy) * yy);
    return c1 / c2 + y1(xx);
}

template <int DEG>
__device__ __forceinline__ double horner_wu(const double* c, double t) {
    double acc = (c[DEG - 1] + c[DEG] * t) * t;
#pragma unroll
    for (int i = DEG - 2; i >= 1; i--) acc = (c[i] + acc) * t;
    return c[0] + acc;
}

__constant__ static const double WU_A[10] = {1.21, -13.328, 215.896, -1763.96, 8418.94, -24314.21, 42002.57, -41592.9, 21859.0, -4838.6};
__constant__ static const double WU_B[10] = {0.938, 5.737, -67.92, 796.534, -4780.77, 17137.74, -36618.81, 44894.06, -29030.24, 7671.22};
__constant__ static const double WU_C[8] = {1.268, -9.747, 209.653, -1397.89, 5155.67, -9844.35, 9136.4, -3272.62};
__constant__ static const double WU_D[10] = {0.632, -40.97, 667.16, -6072.07, 31127.39, -96293.05, 181856.75, -205690.43, 128170.2, -33744.6};
__constant__ static const double WU_DA[10] = {2.948, -24.53, 249.69, -754.85, -1187.71, 16370.75, -48811.41, 68220.87, -46688.0, 12622.25};
__constant__ static const double WU_DB[10] = {1.11, 2.894, -76.765, 1565.35, -11336.19, 44270.15, -97014.11, 118879.26, -76209.82, 19923.28};
__constant__ static const double WU_DC[6] = {14.19, -148.24, 847.8, -2318.58, 3168.35, -1590.27};

// _wu_wave_term_and_derivs (wu.jl:28-37): G, dG/dh, dG/dv of the dimensionless wave term.  Inlined
// where the caller keeps the outputs in registers (the surrogate's pair kernel); the assembly
// kernels call the out-of-line wu_term below.
__device__ __forceinline__ void wu_term_inline(double hh, double vv, cplx& G, cplx& Gh, cplx& Gv) {
    double dd = sqrt(hh * hh + vv * vv);
    double alpha = -vv / dd, beta = hh / dd, rho = dd / (1 + dd);
    double ev = exp(vv);
    double omr3 = (1.0 - rho) * (1.0 - rho) * (1.0 - rho);
    double dd2 = dd * dd, dd3 = dd * dd * dd;
    double A = horner_wu<9>(WU_A, rho), B = horner_wu<9>(WU_B, rho), C = horner_wu<7>(WU_C, rho),
           D = horner_wu<9>(WU_D, rho);
    double RR = (1.0 - beta) * A;
    RR -= beta * B;
    RR -= alpha * C / (1.0 + 6.0 * alpha * rho * (1.0 - rho));
    RR += beta * (1.0 - beta) * D;
    double Lp = rho * omr3 * RR;
    double PP = log(0.5 * (dd - vv)) + 0.5772156649 - 2.0 * dd2;
    PP = ev * PP;
    PP = PP + dd2 - vv;
    double L0 = 2.0 * PP / (1.0 + dd3) + 2.0 * Lp;
    double H0 = struve_h0(hh), J0 = j0(hh);
    double Wr = 2 * MHX_PI * H0 * ev, Wi = 2 * MHX_PI * (-J0) * ev;
    double dA = horner_wu<9>(WU_DA, rho), dB = horner_wu<9>(WU_DB, rho), dC = horner_wu<5>(WU_DC, rho);
    double RS = beta * dA;
    RS -= (1.0 - alpha) * dB;
    RS += beta * (1.0 - beta) * rho * (1.0 - 2.0 * rho) * dC;
    double Lsp = rho * omr3 * RS;
    double PS = (beta + hh) / (dd - vv);
    PS = PS - 2.0 * beta + 2.0 * ev * dd - hh;
    double QS = exp(-dd) * (1.0 - beta);
    QS = QS * (1.0 + dd / (1.0 + dd3));
    double Ls = 2.0 * PS / (1.0 + dd3) - 4.0 * QS + 2.0 * Lsp;
    double H1 = struve_h1(hh), J1 = j1(hh);
    double Whr = 2 * MHX_PI * (2 / MHX_PI - H1) * ev, Whi = 2 * MHX_PI * J1 * ev;
    G = cmk(-L0 - Wr, -Wi);
    Gh = cmk(-Ls - Whr, -Whi);
    Gv = cmk(G.x + 2 / dd, G.y);
}

__device__ __noinline__ void wu_term(double hh, double vv, cplx& G, cplx& Gh, cplx& Gv) {
    wu_term_inline(hh, vv, G, Gh, Gv);
}

// G and dG/dh, returned in registers (an out-of-line call whose outputs do not pass through the
// caller's stack)
struct WuGh {
    cplx G, Gh;
};
__device__ __noinline__ WuGh wu_term_gh(double hh, double vv) {
    WuGh o;
    cplx Gv;
    wu_term_inline(hh, vv, o.G, o.Gh, Gv);
    return o;
}

// The same G and dG/dh with the special functions in FP32 (the learned wave term's base: its
// network residual is accurate to ~1e-5 of G, and FP32 exp, log, Bessel and Struve values move G
// by ~1e-7).  The polynomial fits A, B, C, D, dA, dB, dC stay in FP64: their alternating
// coefficients (|c| up to 2e5) cancel to O(1), which FP32 Horner would leave wrong in the third digit.
__device__ __forceinline__ float struve_h0f(float xx) {
    if (xx <= 3.0f) {
        const float yy = (xx / 3.0f) * (xx / 3.0f);
        const float h = 1.909859164f +
                        (-1.909855001f + (0.687514637f + (-0.126164557f + (0.013828813f + -0.000876918f * yy) * yy) * yy) * yy) * yy;
        return h * (xx / 3.0f);
    }
    const float yy = (3.0f / xx) * (3.0f / xx);
    const float c1 = 2.0f * (0.99999906f + (4.77228920f + (3.85542044f + 0.32303607f * yy) * yy) * yy);
    const float c2 = (float)MHX_PI * xx * (1.0f + (4.88331068f + (4.28957333f + 0.52120508f * yy) * yy) * yy);
    return c1 / c2 + y0f(xx);
}
__device__ __forceinline__ float struve_h1f(float xx) {
    if (xx <= 3.0f) {
        const float yy = (xx / 3.0f) * (xx / 3.0f);
        return (1.909859286f +
                (-1.145914713f + (0.294656958f + (-0.042070508f + (0.003785727f + -0.000207183f * yy) * yy) * yy) * yy) * yy) * yy;
    }
    const float yy = (3.0f / xx) * (3.0f / xx);
    const float c1 = 2.0f * (1.00000004f + (3.92205313f + (2.64893033f + 0.27450895f * yy) * yy) * yy);
    const float c2 = (float)MHX_PI * (1.0f + (3.81095112f + (2.26216956f + 0.10885141f * yy) * yy) * yy);
    return c1 / c2 + y1f(xx);
}
__device__ __noinline__ WuGh wu_term_gh_f32sf(double hh, double vv) {
    const double dd = sqrt(hh * hh + vv * vv), idd = 1.0 / dd;
    const double alpha = -vv * idd, beta = hh * idd, rho = dd / (1 + dd);
    const float hf = (float)hh, vf = (float)vv, ddf = (float)dd;
    const double ev = (double)expf(vf);
    const double omr3 = (1.0 - rho) * (1.0 - rho) * (1.0 - rho);
    const double dd2 = dd * dd, dd3 = dd2 * dd, iden = 1.0 / (1.0 + dd3);
    const double A = horner_wu<9>(WU_A, rho), B = horner_wu<9>(WU_B, rho), C = horner_wu<7>(WU_C, rho),
                 D = horner_wu<9>(WU_D, rho);
    double RR = (1.0 - beta) * A;
    RR -= beta * B;
    RR -= alpha * C / (1.0 + 6.0 * alpha * rho * (1.0 - rho));
    RR += beta * (1.0 - beta) * D;
    const double Lp = rho * omr3 * RR;
    double PP = (double)logf(0.5f * (ddf - vf)) + 0.5772156649 - 2.0 * dd2;
    PP = ev * PP + dd2 - vv;
    const double L0 = 2.0 * PP * iden + 2.0 * Lp;
    const double H0 = struve_h0f(hf), J0 = j0f(hf);
    const double Wr = 2 * MHX_PI * H0 * ev, Wi = 2 * MHX_PI * (-J0) * ev;
    const double dA = horner_wu<9>(WU_DA, rho), dB = horner_wu<9>(WU_DB, rho), dC = horner_wu<5>(WU_DC, rho);
    double RS = beta * dA;
    RS -= (1.0 - alpha) * dB;
    RS += beta * (1.0 - beta) * rho * (1.0 - 2.0 * rho) * dC;
    const double Lsp = rho * omr3 * RS;
    double PS = (beta + hh) / (dd - vv);
    PS = PS - 2.0 * beta + 2.0 * ev * dd - hh;
    const double QS = (double)expf(-ddf) * (1.0 - beta) * (1.0 + dd * iden);
    const double Ls = 2.0 * PS * iden - 4.0 * QS + 2.0 * Lsp;
    const double H1 = struve_h1f(hf), J1 = j1f(hf);
    const double Whr = 2 * MHX_PI * (2 / MHX_PI - H1) * ev, Whi = 2 * MHX_PI * J1 * ev;
    WuGh o;
    o.G = cmk(-L0 - Wr, -Wi);
    o.Gh = cmk(-Ls - Whr, -Whi);
    return o;
}

// reduced coordinates (wu.jl:7-18; abstract_greens_function.jl:117-138)
__device__ __forceinline__ void reduced_hv(double x0, double x1, double x2, double y0_, double y1_, double y2_,
                                           double k, double& hh, double& vv) {
    double d2 = (x0 - y0_) * (x0 - y0_) + (x1 - y1_) * (x1 - y1_);
    double dis;
    if (d2 == 0.0)
        dis = sqrt(d2 + 1e-18 * 1e-18) - 1e-18;
    else
        dis = sqrt(d2);
    hh = k * dis;
    vv = k * (x2 + y2_);
}

// ---------------------------------------------------------------------------
// exp(z) E1(z).  Region rule: power series (Horner on a_n = (-1)^n/(n n!)) for
// |z|<=2, for Re z<=0 with |z|<=4,
// and inside the cone |z|<=40, |Im z| <= -Re z (24/|z| - 0.1); elsewhere the even
// continued fraction 1/(z+1- 1/(z+3- 4/(z+5- ...))) by modified Lentz.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool e1_use_series(cplx z) {
    double R = hypot(z.x, z.y), y = fabs(z.y);
    if (R <= 2.0) return true;
    if (z.x <= 0.0 && R <= 4.0) return true;
    if (z.x < 0.0 && R <= 40.0 && y <= (-z.x) * (24.0 / R - 0.1)) return true;
    return false;
}

// per-thread tally of the quadrature work (integrand evaluations, power-series
// terms, continued-fraction steps) used for the algorithmic FLOP count
struct QuadWork {
    unsigned evals = 0, series = 0, cf = 0;
};

#define E1_QUAL __constant__ static
#include "mhx_e1_table.h"

// exp(z) E1(z); ez = exp(z) (the caller's value: the integrands need it as well)
__device__ __noinline__ cplx expE1(cplx z, cplx ez, QuadWork& wk) {
    if (e1_use_series(z)) {
        // E1 = -gamma - log z - sum_{n=1}^{N} a_n z^n, Horner, degree from the table at ceil|z|
        int ir = (int)ceil(hypot(z.x, z.y));
        ir = ir > 40 ? 40 : ir;
        const int N = E1_DEG[ir];
        // Horner from the largest degree among the wave's active lanes, with the coefficients
        // above a lane's own degree taken as zero: the lane's polynomial stays zero until its
        // leading term, so its value is bitwise that of Horner from its own degree, and the
        // coefficient index is the same for every lane (a scalar load)
        int Nw = 0;
#pragma unroll
        for (int bit = 6; bit >= 0; bit--)
            if (__ballot(N >= Nw + (1 << bit))) Nw += 1 << bit;
        Nw = __builtin_amdgcn_readfirstlane(Nw);
        double pr = 0.0, pi = 0.0;
        for (int q = Nw - 1; q >= 0; q--) {
            const double a = q < N ? E1_A[q] : 0.0;
            const double tr = a + (z.x * pr - z.y * pi);
            pi = z.x * pi + z.y * pr;
            pr = tr;
        }
        wk.series += N;
        const double sr = z.x * pr - z.y * pi, si = z.x * pi + z.y * pr;
        cplx L = clog_(z);
        cplx e1 = cmk(-MHX_EULER - L.x - sr, -L.y - si);
        return cmul(ez, e1);
    }
    const double tiny = 1e-300;
    cplx f = cmk(z.x + 1.0, z.y);
    if (f.x == 0.0 && f.y == 0.0) f = cmk(tiny, 0.0);
    cplx C = f, Dd = cmk(0.0, 0.0);
    for (int n = 1; n < 600; n++) {
        double a = -(double)n * (double)n;
        cplx b = cmk(z.x + (double)(2 * n + 1), z.y);
        Dd = cadd(b, cscale(a, Dd));
        if (Dd.x == 0.0 && Dd.y == 0.0) Dd = cmk(tiny, 0.0);
        Dd = cdiv(cmk(1.0, 0.0), Dd);
        C = cadd(b, cscale(a, cdiv(cmk(1.0, 0.0), C)));
        if (C.x == 0.0 && C.y == 0.0) C = cmk(tiny, 0.0);
        cplx d = cmul(C, Dd);
        f = cmul(f, d);
        wk.cf++;
        if (hypot(d.x - 1.0, d.y) < 1e-16) break;
    }
    return cdiv(cmk(1.0, 0.0), f);
}

// integrands of _dimless_wave_term (exact_Guevel_Delhommeau.jl:16-23) and of
// _d_dimless_wave_term_dr (:36-43)
template <bool DERIV>
__device__ __forceinline__ cplx del_integrand(double th, double r, double z, QuadWork& wk) {
    double c = cos(th);
    cplx zeta = cmk(z, r * c);
    cplx ez = cexp_(zeta);
    wk.evals++;
    cplx w = expE1(zeta, ez, wk);
    cplx j = cmk(w.x / MHX_PI - ez.y, w.y / MHX_PI + ez.x);  // w/pi + i e^zeta
    if (!DERIV) return cmk(j.x, ez.x);
    cplx inv = cdiv(cmk(1.0, 0.0), cmk(MHX_PI * zeta.x, MHX_PI * zeta.y));
    double qi = j.y - inv.y;
    return cmk(-c * qi, -c * ez.y);
}

__constant__ static const double GK_X[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                                            0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                                            0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                                            0.207784955007898467600689403773245, 0.0};
__constant__ static const double GK_WK[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                                             0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                                             0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                                             0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
__constant__ static const double GK_WG[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                                             0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

template <bool DERIV>
__device__ __forceinline__ cplx gk15(double a, double b, double r, double z, double& err, QuadWork& wk) {
    double c = (a + b) * 0.5, d = (b - a) * 0.5;
    cplx f0 = del_integrand<DERIV>(c, r, z, wk);
    cplx Ik = cscale(GK_WK[7], f0), Ig = cscale(GK_WG[3], f0);
#pragma unroll 1
    for (int i = 0; i < 7; i++) {
        double dx = d * GK_X[i];
        cplx fx = cadd(del_integrand<DERIV>(c - dx, r, z, wk), del_integrand<DERIV>(c + dx, r, z, wk));
        Ik = cadd(Ik, cscale(GK_WK[i], fx));
        if (i & 1) Ig = cadd(Ig, cscale(GK_WG[i >> 1], fx));
    }
    Ik = cscale(d, Ik);
    Ig = cscale(d, Ig);
    err = hypot(Ik.x - Ig.x, Ik.y - Ig.y);
    return Ik;
}

#define MHX_MAXBOX 48

// Both integrands at one node from a single E1 evaluation: .x/.y of G, then of dG/dr.
// Values are bitwise those of del_integrand<false> and del_integrand<true>.
__device__ __forceinline__ void del_integrand_both(double th, double r, double z, QuadWork& wk, cplx& fG, cplx& fR) {
    double c = cos(th);
    cplx zeta = cmk(z, r * c);
    cplx ez = cexp_(zeta);
    wk.evals++;
    cplx w = expE1(zeta, ez, wk);
    cplx j = cmk(w.x / MHX_PI - ez.y, w.y / MHX_PI + ez.x);
    fG = cmk(j.x, ez.x);
    cplx inv = cdiv(cmk(1.0, 0.0), cmk(MHX_PI * zeta.x, MHX_PI * zeta.y));
    fR = cmk(-c * (j.y - inv.y), -c * ez.y);
}

// del_integrand_both with cos(th) given (bitwise the same when c == cos(th))
__device__ __forceinline__ void del_integrand_both_c(double c, double r, double z, QuadWork& wk, cplx& fG, cplx& fR) {
    cplx zeta = cmk(z, r * c);
    cplx ez = cexp_(zeta);
    wk.evals++;
    cplx w = expE1(zeta, ez, wk);
    cplx j = cmk(w.x / MHX_PI - ez.y, w.y / MHX_PI + ez.x);
    fG = cmk(j.x, ez.x);
    cplx inv = cdiv(cmk(1.0, 0.0), cmk(MHX_PI * zeta.x, MHX_PI * zeta.y));
    fR = cmk(-c * (j.y - inv.y), -c * ez.y);
}

// cosines of the 15 nodes of the first interval [0, pi/2]: cosn[7] at the centre, cosn[i] and
// cosn[8 + i] at c - d x_i and c + d x_i (i < 7), computed as gk15_both computes them
__device__ __forceinline__ double first_node_cos(int q) {
    const double a = 0.0, b = MHX_PI / 2;
    const double c = (a + b) * 0.5, d = (b - a) * 0.5;
    if (q == 7) return cos(c);
    const double dx = d * GK_X[q < 7 ? q : q - 8];
    return cos(q < 7 ? c - dx : c + dx);
}

// gk15_both on [0, pi/2] from the node cosines
__device__ __forceinline__ void gk15_both_first(const double* cosn, double r, double z, cplx& IG, double& EG, cplx& IR,
                                                double& ER, QuadWork& wk) {
    const double d = (MHX_PI / 2 - 0.0) * 0.5;
    cplx g0, r0;
    del_integrand_both_c(cosn[7], r, z, wk, g0, r0);
    cplx IkG = cscale(GK_WK[7], g0), IgG = cscale(GK_WG[3], g0);
    cplx IkR = cscale(GK_WK[7], r0), IgR = cscale(GK_WG[3], r0);
#pragma unroll 1
    for (int i = 0; i < 7; i++) {
        cplx gm, rm, gp, rp;
        del_integrand_both_c(cosn[i], r, z, wk, gm, rm);
        del_integrand_both_c(cosn[8 + i], r, z, wk, gp, rp);
        cplx fG = cadd(gm, gp), fR = cadd(rm, rp);
        IkG = cadd(IkG, cscale(GK_WK[i], fG));
        IkR = cadd(IkR, cscale(GK_WK[i], fR));
        if (i & 1) {
            IgG = cadd(IgG, cscale(GK_WG[i >> 1], fG));
            IgR = cadd(IgR, cscale(GK_WG[i >> 1], fR));
        }
    }
    IkG = cscale(d, IkG); IgG = cscale(d, IgG);
    IkR = cscale(d, IkR); IgR = cscale(d, IgR);
    EG = hypot(IkG.x - IgG.x, IkG.y - IgG.y);
    ER = hypot(IkR.x - IgR.x, IkR.y - IgR.y);
    IG = IkG;
    IR = IkR;
}

// GK(7,15) of both integrands on [a, b] with shared nodes; same accumulation order as gk15.
__device__ __forceinline__ void gk15_both(double a, double b, double r, double z, cplx& IG, double& EG, cplx& IR,
                                          double& ER, QuadWork& wk) {
    double c = (a + b) * 0.5, d = (b - a) * 0.5;
    cplx g0, r0;
    del_integrand_both(c, r, z, wk, g0, r0);
    cplx IkG = cscale(GK_WK[7], g0), IgG = cscale(GK_WG[3], g0);
    cplx IkR = cscale(GK_WK[7], r0), IgR = cscale(GK_WG[3], r0);
#pragma unroll 1
    for (int i = 0; i < 7; i++) {
        double dx = d * GK_X[i];
        cplx gm, rm, gp, rp;
        del_integrand_both(c - dx, r, z, wk, gm, rm);
        del_integrand_both(c + dx, r, z, wk, gp, rp);
        cplx fG = cadd(gm, gp), fR = cadd(rm, rp);
        IkG = cadd(IkG, cscale(GK_WK[i], fG));
        IkR = cadd(IkR, cscale(GK_WK[i], fR));
        if (i & 1) {
            IgG = cadd(IgG, cscale(GK_WG[i >> 1], fG));
            IgR = cadd(IgR, cscale(GK_WG[i >> 1], fR));
        }
    }
    IkG = cscale(d, IkG); IgG = cscale(d, IgG);
    IkR = cscale(d, IkR); IgR = cscale(d, IgR);
    EG = hypot(IkG.x - IgG.x, IkG.y - IgG.y);
    ER = hypot(IkR.x - IgR.x, IkR.y - IgR.y);
    IG = IkG;
    IR = IkR;
}

// Adaptive GK(7,15) on [0, pi/2] from a given first interval estimate: pop the interval
// of largest error, bisect, stop when E <= max(rtol |I|, atol); re-sum the intervals at
// the end.  Returns false if the interval budget was exhausted.
template <bool DERIV>
__device__ bool hquad_from(double r, double z, double rtol, double atol, cplx I0, double E, cplx& out, QuadWork& wk) {
    double ba[MHX_MAXBOX], bb[MHX_MAXBOX], be[MHX_MAXBOX];
    cplx bi[MHX_MAXBOX];
    ba[0] = 0.0; bb[0] = MHX_PI / 2; bi[0] = I0; be[0] = E;
    int nb = 1;
    cplx It = I0;
    double Et = E;
    bool ok = true;
    while (Et > fmax(rtol * hypot(It.x, It.y), atol)) {
        if (nb >= MHX_MAXBOX - 1) { ok = false; break; }
        int im = 0;
        for (int q = 1; q < nb; q++)
            if (be[q] > be[im]) im = q;
        double a = ba[im], b = bb[im], e = be[im];
        cplx iv = bi[im];
        ba[im] = ba[nb - 1]; bb[im] = bb[nb - 1]; be[im] = be[nb - 1]; bi[im] = bi[nb - 1];
        nb--;
        It = csub(It, iv);
        Et -= e;
        double w = (b - a) / 2, m = a + w;
        double E1, E2;
        cplx I1 = gk15<DERIV>(a, m, r, z, E1, wk);
        cplx I2 = gk15<DERIV>(m, b, r, z, E2, wk);
        It = cadd(It, cadd(I1, I2));
        Et += E1 + E2;
        ba[nb] = a; bb[nb] = m; bi[nb] = I1; be[nb] = E1; nb++;
        ba[nb] = m; bb[nb] = b; bi[nb] = I2; be[nb] = E2; nb++;
    }
    cplx s = cmk(0.0, 0.0);
    for (int q = 0; q < nb; q++) s = cadd(s, bi[q]);
    out = s;
    return ok;
}

// The first interval [0, pi/2] of both quadratures (estimates G, Gr, errors EG, ER).  Returns true
// when both already meet the termination rule of hquad_from (E <= max(rtol |I|, atol)): G, Gr are
// then the final values, bitwise those of hquad_pair.
__device__ __forceinline__ bool hquad_first(const double* cosn, double r, double z, double rtol, double atol, cplx& G,
                                            cplx& Gr, double& EG, double& ER, QuadWork& wk) {
    gk15_both_first(cosn, r, z, G, EG, Gr, ER, wk);
    return !(EG > fmax(rtol * hypot(G.x, G.y), atol)) && !(ER > fmax(rtol * hypot(Gr.x, Gr.y), atol));
}

// ---- wave-cooperative adaptive quadrature (one pair per 32-lane group) ----
// Group sums over the 32 lanes of a group (width 32) or over its 16-lane halves (width 16).
template <int W> __device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, W);
    return v;
}
template <int W> __device__ __forceinline__ cplx group_sum(cplx v) { return cmk(group_sum<W>(v.x), group_sum<W>(v.y)); }

// Kronrod and Gauss weights of node q of the GK(7,15) ordering used here: q = 7 the centre,
// q < 7 at c - d x_q, q = 8 + i at c + d x_i
__device__ __forceinline__ void gk_weights(int q, double& wk, double& wg) {
    const int i = q == 7 ? 7 : (q < 7 ? q : q - 8);
    wk = GK_WK[i];
    wg = i == 7 ? GK_WG[3] : ((i & 1) ? GK_WG[i >> 1] : 0.0);
}

// Interval list of one group in LDS
struct QuadBoxes {
    double a[MHX_MAXBOX], b[MHX_MAXBOX], e[MHX_MAXBOX];
    cplx i[MHX_MAXBOX];
};

// One bisection of [a, b] on a group: the 15 nodes of [a, m] on lanes 0-14 and those of [m, b] on
// lanes 16-30 (nodes as gk15 places them), the GK sums of each half by shuffles within its 16
// lanes.  BOTH: both integrands from one E1 evaluation per node (del_integrand_both); otherwise
// the one of DERIV.  Results (I1, E1 of [a, m]; I2, E2 of [m, b]) are the same on every lane.
template <int MODE>  // 0: G, 1: dG/dr, 2: both
__device__ __forceinline__ void group_bisect(int gl, double a, double b, double r, double z, cplx* I1, double* E1,
                                             cplx* I2, double* E2, QuadWork& wk) {
    const int half = gl >> 4, q = gl & 15;
    double wkq = 0.0, wgq = 0.0;
    if (q < 15) gk_weights(q, wkq, wgq);
    const double xq = q == 7 ? 0.0 : GK_X[q < 7 ? q : (q < 15 ? q - 8 : 0)];
    const double w = (b - a) / 2, m = a + w;
    const double ha = half ? m : a, hb = half ? b : m;
    const double c = (ha + hb) * 0.5, d = (hb - ha) * 0.5;
    const double th = q == 7 ? c : (q < 7 ? c - d * xq : c + d * xq);
    constexpr int NI = MODE == 2 ? 2 : 1;
    cplx f[NI];
#pragma unroll
    for (int t = 0; t < NI; t++) f[t] = cmk(0.0, 0.0);
    if (q < 15) {
        if (MODE == 2) del_integrand_both(th, r, z, wk, f[0], f[1]);
        else f[0] = del_integrand<MODE == 1>(th, r, z, wk);
    }
#pragma unroll
    for (int t = 0; t < NI; t++) {
        const cplx Ik = cscale(d, group_sum<16>(cscale(wkq, f[t]))), Ig = cscale(d, group_sum<16>(cscale(wgq, f[t])));
        const double Eh = hypot(Ik.x - Ig.x, Ik.y - Ig.y);
        I1[t] = cmk(__shfl(Ik.x, 0, 32), __shfl(Ik.y, 0, 32));
        I2[t] = cmk(__shfl(Ik.x, 16, 32), __shfl(Ik.y, 16, 32));
        E1[t] = __shfl(Eh, 0, 32);
        E2[t] = __shfl(Eh, 16, 32);
    }
}

// hquad_from on one group, continued from a list of nb0 intervals that lane 0 has written to B
// (totals It, Et): the same rule (pop the interval of largest error, first index on ties, the
// last interval moved into its place; bisect; stop when E <= max(rtol |I|, atol); re-sum in list
// order), each bisection by group_bisect.  gl: lane in the group; every lane returns the same
// values.
template <bool DERIV>
__device__ bool hquad_group_loop(QuadBoxes& B, int gl, double r, double z, double rtol, double atol, int nb0, cplx It,
                                 double Et, cplx& out, QuadWork& wk) {
    int nb = nb0;
    bool ok = true;
    __asm__ volatile("" ::: "memory");
    while (Et > fmax(rtol * hypot(It.x, It.y), atol)) {
        if (nb >= MHX_MAXBOX - 1) { ok = false; break; }
        // interval of largest error (first index on ties)
        double bv = -1.0;
        int bi = 0x7fffffff;
        for (int t = gl; t < nb; t += 32)
            if (B.e[t] > bv) { bv = B.e[t]; bi = t; }
#pragma unroll
        for (int off = 16; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bv, off, 32);
            const int oi = __shfl_xor(bi, off, 32);
            if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
        }
        const int im = bi < nb ? bi : 0;  // no finite maximum (NaN errors): the first interval
        const double a = B.a[im], b = B.b[im], e = B.e[im];
        const cplx iv = B.i[im];
        const double la = B.a[nb - 1], lb = B.b[nb - 1], le = B.e[nb - 1];
        const cplx li = B.i[nb - 1];
        __asm__ volatile("" ::: "memory");
        if (gl == 0) { B.a[im] = la; B.b[im] = lb; B.e[im] = le; B.i[im] = li; }
        nb--;
        It = csub(It, iv);
        Et -= e;
        cplx I1, I2;
        double E1, E2;
        group_bisect<DERIV ? 1 : 0>(gl, a, b, r, z, &I1, &E1, &I2, &E2, wk);
        const double m = a + (b - a) / 2;
        It = cadd(It, cadd(I1, I2));
        Et += E1 + E2;
        __asm__ volatile("" ::: "memory");
        if (gl == 0) {
            B.a[nb] = a; B.b[nb] = m; B.i[nb] = I1; B.e[nb] = E1;
            B.a[nb + 1] = m; B.b[nb + 1] = b; B.i[nb + 1] = I2; B.e[nb + 1] = E2;
        }
        nb += 2;
        __asm__ volatile("" ::: "memory");
    }
    cplx s = cmk(0.0, 0.0);
    if (gl == 0)
        for (int t = 0; t < nb; t++) s = cadd(s, B.i[t]);
    out = cmk(__shfl(s.x, 0, 32), __shfl(s.y, 0, 32));
    __asm__ volatile("" ::: "memory");
    return ok;
}

// hquad_from on one group from the first interval's estimate (I0, E)
template <bool DERIV>
__device__ __forceinline__ bool hquad_from_group(QuadBoxes& B, int gl, double r, double z, double rtol, double atol,
                                                 cplx I0, double E, cplx& out, QuadWork& wk) {
    if (gl == 0) {
        B.a[0] = 0.0; B.b[0] = MHX_PI / 2; B.i[0] = I0; B.e[0] = E;
    }
    return hquad_group_loop<DERIV>(B, gl, r, z, rtol, atol, 1, I0, E, out, wk);
}

// hquad_from's state after its first step on [0, pi/2] (interval 0 popped, its halves pushed):
// the list {[0, m]: I1, E1; [m, pi/2]: I2, E2} and the totals as hquad_from forms them
template <bool DERIV>
__device__ __forceinline__ bool hquad_group_after_first(QuadBoxes& B, int gl, double r, double z, double rtol,
                                                        double atol, cplx I0, double E, cplx I1, double E1, cplx I2,
                                                        double E2, cplx& out, QuadWork& wk) {
    const double a = 0.0, b = MHX_PI / 2, m = a + (b - a) / 2;
    if (gl == 0) {
        B.a[0] = a; B.b[0] = m; B.i[0] = I1; B.e[0] = E1;
        B.a[1] = m; B.b[1] = b; B.i[1] = I2; B.e[1] = E2;
    }
    double Et = E;
    Et -= E;
    Et += E1 + E2;
    return hquad_group_loop<DERIV>(B, gl, r, z, rtol, atol, 2, cadd(csub(I0, I0), cadd(I1, I2)), Et, out, wk);
}

// The refinement of one pair on a group, from the first interval's estimates of both quadratures
// (IG, EG of G; IR, ER of dG/dr).  A quadrature that already meets the termination rule keeps
// its estimate; when both need refinement, their common first bisection evaluates both
// integrands at once.  The results are those of hquad_pair.
__device__ __forceinline__ bool refine_pair_group(QuadBoxes& B, int gl, double r, double z, double rtol, double atol,
                                                  cplx IG, double EG, cplx IR, double ER, cplx& G, cplx& Gr,
                                                  QuadWork& wk) {
    const bool needG = EG > fmax(rtol * hypot(IG.x, IG.y), atol);
    const bool needR = ER > fmax(rtol * hypot(IR.x, IR.y), atol);
    bool okG = true, okR = true;
    G = IG;
    Gr = IR;
    if (needG && needR) {
        cplx I1[2], I2[2];
        double E1[2], E2[2];
        group_bisect<2>(gl, 0.0, MHX_PI / 2, r, z, I1, E1, I2, E2, wk);
        okR = hquad_group_after_first<true>(B, gl, r, z, rtol, atol, IR, ER, I1[1], E1[1], I2[1], E2[1], Gr, wk);
        okG = hquad_group_after_first<false>(B, gl, r, z, rtol, atol, IG, EG, I1[0], E1[0], I2[0], E2[0], G, wk);
    } else if (needR) {
        okR = hquad_from_group<true>(B, gl, r, z, rtol, atol, IR, ER, Gr, wk);
    } else if (needG) {
        okG = hquad_from_group<false>(B, gl, r, z, rtol, atol, IG, EG, G, wk);
    }
    return okG && okR;
}

// The first interval [0, pi/2] of both quadratures on a group: the 15 nodes on lanes 0-14
// (cosines from cosn)
__device__ __forceinline__ void first_interval_group(int gl, const double* cosn, double r, double z, cplx& IG,
                                                     double& EG, cplx& IR, double& ER, QuadWork& wk) {
    const double d = (MHX_PI / 2 - 0.0) * 0.5;
    cplx fG = cmk(0.0, 0.0), fR = cmk(0.0, 0.0);
    double wkq = 0.0, wgq = 0.0;
    if (gl < 15) {
        gk_weights(gl, wkq, wgq);
        del_integrand_both_c(cosn[gl], r, z, wk, fG, fR);
    }
    IG = cscale(d, group_sum<32>(cscale(wkq, fG)));
    const cplx IgG = cscale(d, group_sum<32>(cscale(wgq, fG)));
    IR = cscale(d, group_sum<32>(cscale(wkq, fR)));
    const cplx IgR = cscale(d, group_sum<32>(cscale(wgq, fR)));
    EG = hypot(IG.x - IgG.x, IG.y - IgG.y);
    ER = hypot(IR.x - IgR.x, IR.y - IgR.y);
}

// G and dG/dr of the exact wave term (both without the factor 4)
__device__ __forceinline__ bool hquad_pair(double r, double z, double rtol, double atol, cplx& G, cplx& Gr,
                                           QuadWork& wk) {
    cplx IG, IR;
    double EG, ER;
    gk15_both(0.0, MHX_PI / 2, r, z, IG, EG, IR, ER, wk);
    bool ok1 = hquad_from<true>(r, z, rtol, atol, IR, ER, Gr, wk);
    bool ok0 = hquad_from<false>(r, z, rtol, atol, IG, EG, G, wk);
    return ok0 && ok1;
}

// Algorithmic FP64 operation counts of the exact wave term (add, sub, mul = 1,
// fma = 2, div, sqrt and each elementary function = 1), counted from the code
// above: per integrand evaluation outside the E1 loop, per power-series Horner
// step (one complex multiply-add), per continued-fraction step.
#define MHX_FLOPS_PER_EVAL 41.0
#define MHX_FLOPS_PER_SERIES 8.0
#define MHX_FLOPS_PER_CF 45.0

// the quantities of an entry that depend only on the (symmetric) pair: 2 / |(hh, vv)| and k^2 / hh
// (0 for hh <= 1e-6)
struct PairQ {
    double inv, kkh;
};
__device__ __forceinline__ PairQ pair_q(double k, double hh, double vv) {
    PairQ q;
    q.inv = 2.0 / hypot(hh, vv);
    q.kkh = fabs(hh) > 1e-6 ? k * k / hh : 0.0;
    return q;
}
// One entry (row panel with centre xr / normal nr, column panel g) of S and D (or K) from the
// wave term G, dG/dr and the Rankine parts sr, dr: with_reduced_coordinates_derivative
// (abstract_greens_function.jl:140-185) and the one-point integrals
// (exact_Guevel_Delhommeau.jl:61-75), scaled by -1/(4 pi); `diag` adds the identity term 1/2.
__device__ __forceinline__ void combine_entry(double k, PairQ pq, const double* xr, const double* nr,
                                              const PanelGeom& g, int direct, bool diag, cplx G, cplx Gh, double sr,
                                              double dr, cplx& S, cplx& D) {
    const double kScaleD = -1.0 / (4 * MHX_PI);
    const cplx Gv = cmk(G.x + pq.inv, G.y);
    double dx1 = direct ? g.c[0] - xr[0] : xr[0] - g.c[0];
    double dx2 = direct ? g.c[1] - xr[1] : xr[1] - g.c[1];
    const double dr1 = pq.kkh * dx1, dr2 = pq.kkh * dx2;
    const double n0 = direct ? g.nrm[0] : nr[0], n1 = direct ? g.nrm[1] : nr[1], n2 = direct ? g.nrm[2] : nr[2];
    const double a2 = g.area;
    cplx g0 = cscale(k, cscale(dr1, Gh)), g1 = cscale(k, cscale(dr2, Gh)), g2 = cscale(k, cscale(k, Gv));
    cplx s = cscale(a2, cscale(k, G));
    cplx d = cadd(cadd(cscale(n0, cscale(a2, g0)), cscale(n1, cscale(a2, g1))), cscale(n2, cscale(a2, g2)));
    if (diag) dr += g.pad[0];
    S = cmk(sr + kScaleD * s.x, kScaleD * s.y);
    D = cmk(dr + kScaleD * d.x, kScaleD * d.y);
}
__device__ __forceinline__ void combine_entry(double k, double hh, double vv, const double* xr, const double* nr,
                                              const PanelGeom& g, int direct, bool diag, cplx G, cplx Gh, double sr,
                                              double dr, cplx& S, cplx& D) {
    combine_entry(k, pair_q(k, hh, vv), xr, nr, g, direct, diag, G, Gh, sr, dr, S, D);
}
