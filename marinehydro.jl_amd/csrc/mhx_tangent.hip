// Tangent (forward-mode) derivatives of the omega-solve: the directional derivative of the
// hydrodynamic coefficients along a tangent of the frequency and of the mesh fields
// (vertices, centres, normals, areas) and dof fields.  It gives what the reference obtains
// with ForwardDiff / Enzyme through its rules (src/ad_rules.jl:215-414): the assembly is
// differentiated in dual arithmetic (value, tangent) that follows the primal kernels
// operation by operation (Birk panels and image, the Wu fit, the boundary conditions and the
// pressure integration); the exact Guevel-Delhommeau term's derivatives come from its table
// (G, dG/dr) and the relations dG/dz = G + 2/rho, d2G/dz2 = dG/dz - 2z/rho^3,
// d2G/drdz = dG/dr - 2r/rho^3 and Laplace's equation d2G/dr2 = -dG/dr / r - d2G/dz2 (the wave
// term is harmonic below the free surface); the linear solve reuses the primal LU:
// dphi = D^{-1} (dS B + S dB - dD phi) (direct), dsigma = D^{-1} (dB - dD sigma),
// dphi = dS sigma + S dsigma (indirect).
#include "mhx_tangent.h"

namespace mhx {

// ---------------------------------------------------------------------------
// dual numbers
// ---------------------------------------------------------------------------
__device__ __forceinline__ Dual dmk(double v, double d = 0.0) { return Dual{v, d}; }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return Dual{-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return Dual{q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual operator+(Dual a, double b) { return Dual{a.v + b, a.d}; }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return Dual{a + b.v, b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return Dual{a.v - b, a.d}; }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return Dual{a - b.v, -b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return Dual{a.v * b, a.d * b}; }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return Dual{a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, double b) { return Dual{a.v / b, a.d / b}; }
__device__ __forceinline__ Dual operator/(double a, Dual b) {
    const double q = a / b.v;
    return Dual{q, -q * b.d / b.v};
}
__device__ __forceinline__ Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
__device__ __forceinline__ Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
// at 0 the tangent is taken as 0 (the argument is a sum of squares whose tangent vanishes there)
__device__ __forceinline__ Dual dsqrt(Dual a) {
    const double s = sqrt(a.v);
    return Dual{s, s > 0.0 ? a.d / (2.0 * s) : 0.0};
}
__device__ __forceinline__ Dual dlog(Dual a) { return Dual{log(a.v), a.d / a.v}; }
__device__ __forceinline__ Dual dexp(Dual a) {
    const double e = exp(a.v);
    return Dual{e, e * a.d};
}
__device__ __forceinline__ Dual datan2(Dual y, Dual x) {
    const double den = x.v * x.v + y.v * y.v;
    return Dual{atan2(y.v, x.v), den > 0.0 ? (x.v * y.d - y.v * x.d) / den : 0.0};
}
__device__ __forceinline__ Dual dhypot(Dual a, Dual b) {
    const double h = hypot(a.v, b.v);
    return Dual{h, h > 0.0 ? (a.v * a.d + b.v * b.d) / h : 0.0};
}
__device__ __forceinline__ void dsincos(Dual a, Dual& s, Dual& c) {
    double sv, cv;
    sincos(a.v, &sv, &cv);
    s = Dual{sv, cv * a.d};
    c = Dual{cv, -sv * a.d};
}
// Bessel functions with their exact derivatives (J0' = -J1, J1' = J0 - J1/x, same for Y)
__device__ __forceinline__ Dual dj0(Dual x) { return Dual{j0(x.v), -j1(x.v) * x.d}; }
__device__ __forceinline__ Dual dj1(Dual x) {
    if (x.v == 0.0) return Dual{0.0, 0.5 * x.d};  // J1'(0) = 1/2
    const double a = j0(x.v), b = j1(x.v);
    return Dual{b, (a - b / x.v) * x.d};
}
__device__ __forceinline__ Dual dy0(Dual x) { return Dual{y0(x.v), -y1(x.v) * x.d}; }
__device__ __forceinline__ Dual dy1(Dual x) {
    const double a = y0(x.v), b = y1(x.v);
    return Dual{b, (a - b / x.v) * x.d};
}

// complex dual: (re, im) each a dual
struct CDual {
    Dual x, y;
};
__device__ __forceinline__ CDual cdmk(Dual x, Dual y) { return CDual{x, y}; }
__device__ __forceinline__ CDual cdadd(CDual a, CDual b) { return CDual{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ CDual cdscale(Dual s, CDual a) { return CDual{s * a.x, s * a.y}; }
__device__ __forceinline__ CDual cdmul(CDual a, CDual b) { return CDual{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cplx tangent(CDual a) { return cmk(a.x.d, a.y.d); }

// ---------------------------------------------------------------------------
// Birk panel and Rankine integrals in dual arithmetic (mirrors mhx_device.h)
// ---------------------------------------------------------------------------
__device__ __forceinline__ Dual atan_ratio_d(Dual num, Dual den) {
    Dual flipped = (den.v < 0.0) ? -num : num;
    Dual ad = (den.v < 0.0) ? -den : den;
    return datan2(flipped, ad);
}

__device__ void birk_panel_geometry_d(const Dual V[12], PanelGeomD& g) {
    Dual mid[4][3];
    for (int i = 0; i < 4; i++) {
        int j = (i + 1) & 3;
        for (int c = 0; c < 3; c++) mid[i][c] = (V[3 * i + c] + V[3 * j + c]) / 2.0;
    }
    Dual mgc[3], s[3], tt[3], n[3], t[3];
    for (int c = 0; c < 3; c++) {
        mgc[c] = (mid[1][c] + mid[3][c]) / 2.0;
        s[c] = mid[1][c] - mid[3][c];
        tt[c] = mid[2][c] - mid[0][c];
    }
    Dual ns = dsqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (int c = 0; c < 3; c++) s[c] = s[c] / ns;
    n[0] = s[1] * tt[2] - s[2] * tt[1];
    n[1] = s[2] * tt[0] - s[0] * tt[2];
    n[2] = s[0] * tt[1] - s[1] * tt[0];
    Dual nn = dsqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int c = 0; c < 3; c++) n[c] = n[c] / nn;
    t[0] = n[1] * s[2] - n[2] * s[1];
    t[1] = n[2] * s[0] - n[0] * s[2];
    t[2] = n[0] * s[1] - n[1] * s[0];
    Dual nt = dsqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    for (int c = 0; c < 3; c++) t[c] = t[c] / nt;
    for (int c = 0; c < 3; c++) {
        g.T[c] = s[c];
        g.T[3 + c] = t[c];
        g.T[6 + c] = n[c];
    }
    Dual tc[4][2];
    for (int k = 0; k < 4; k++) {
        Dual d0 = V[3 * k] - mgc[0], d1 = V[3 * k + 1] - mgc[1], d2 = V[3 * k + 2] - mgc[2];
        tc[k][0] = d0 * s[0] + d1 * s[1] + d2 * s[2];
        tc[k][1] = d0 * t[0] + d1 * t[1] + d2 * t[2];
    }
    Dual A = dmk(0), Mx = dmk(0), My = dmk(0);
    for (int k = 0; k < 4; k++) {
        int k1 = (k + 1) & 3;
        Dual xk = tc[k][0], yk = tc[k][1], xk1 = tc[k1][0], yk1 = tc[k1][1];
        A += (yk1 - yk) * (xk + xk1) / 2.0;
        Mx += (xk1 - xk) * (yk * yk + yk * yk1 + yk1 * yk1) / 6.0;
        My += (yk1 - yk) * (xk * xk + xk * xk1 + xk1 * xk1) / 6.0;
    }
    Dual clx = My / A, cly = -Mx / A;
    for (int c = 0; c < 3; c++) g.qgc[c] = mgc[c] + clx * s[c] + cly * t[c];
    for (int k = 0; k < 4; k++) {
        g.lc[2 * k] = tc[k][0] - clx;
        g.lc[2 * k + 1] = tc[k][1] - cly;
    }
}

__device__ void birk_phi_v_d(Dual x, Dual y, Dual z, const Dual* lc, Dual& phi, Dual v[3]) {
    Dual sx[4], sy[4], r[4], e[4], h[4];
    for (int k = 0; k < 4; k++) {
        sx[k] = x - lc[2 * k];
        sy[k] = y - lc[2 * k + 1];
        r[k] = dsqrt(sx[k] * sx[k] + sy[k] * sy[k] + z * z);
        e[k] = sx[k] * sx[k] + z * z;
        h[k] = sx[k] * sy[k];
    }
    Dual s0 = dmk(0), s1 = dmk(0), s2 = dmk(0), s3 = dmk(0);
    const bool inplane = fabs(z.v) <= 1e-14;
    for (int k = 0; k < 4; k++) {
        int kn = (k + 1) & 3;
        Dual dx = lc[2 * kn] - lc[2 * k], dy = lc[2 * kn + 1] - lc[2 * k + 1];
        Dual dk = dhypot(dx, dy);
        Dual rsum = r[k] + r[kn];
        Dual lt_dk = dlog((rsum - dk) / (rsum + dk)) / dk;
        Dual datan = atan_ratio_d(dy * e[k] - dx * h[k], dx * z * r[k]) -
                     atan_ratio_d(dy * e[kn] - dx * h[kn], dx * z * r[kn]);
        Dual vzk = inplane ? (datan - datan) : datan;
        Dual phik = (sx[k] * dy - sy[k] * dx) * lt_dk + z * vzk;
        if (!(dk.v <= 1e-14)) {
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

__device__ void vrankine_d(const Dual p[3], const PanelGeomD& g, Dual& phi, Dual grad[3]) {
    Dual rb0 = p[0] - g.c[0], rb1 = p[1] - g.c[1], rb2 = p[2] - g.c[2];
    Dual r = dsqrt(rb0 * rb0 + rb1 * rb1 + rb2 * rb2);
    if (r.v > 7 * g.radius) {
        Dual r3 = r * r * r;
        phi = g.area / r;
        grad[0] = (rb0 / r3) * g.area;
        grad[1] = (rb1 / r3) * g.area;
        grad[2] = (rb2 / r3) * g.area;
        return;
    }
    Dual d0 = p[0] - g.qgc[0], d1 = p[1] - g.qgc[1], d2 = p[2] - g.qgc[2];
    Dual fl0 = g.T[0] * d0 + g.T[1] * d1 + g.T[2] * d2;
    Dual fl1 = g.T[3] * d0 + g.T[4] * d1 + g.T[5] * d2;
    Dual fl2 = g.T[6] * d0 + g.T[7] * d1 + g.T[8] * d2;
    Dual pb, vl[3];
    birk_phi_v_d(fl0, fl1, fl2, g.lc, pb, vl);
    phi = -4 * MHX_PI * pb;
    for (int q = 0; q < 3; q++) grad[q] = 4 * MHX_PI * (g.T[q] * vl[0] + g.T[3 + q] * vl[1] + g.T[6 + q] * vl[2]);
}

// ---------------------------------------------------------------------------
// Wu fit in dual arithmetic (mirrors wu_term)
// ---------------------------------------------------------------------------
__device__ __forceinline__ Dual struve_h0_d(Dual xx) {
    if (xx.v <= 3.0) {
        Dual yy = (xx / 3.0) * (xx / 3.0);
        Dual h = 1.909859164 +
                 (-1.909855001 + (0.687514637 + (-0.126164557 + (0.013828813 + -0.000876918 * yy) * yy) * yy) * yy) * yy;
        return h * (xx / 3.0);
    }
    Dual yy = (3.0 / xx) * (3.0 / xx);
    Dual c1 = 2.0 * (0.99999906 + (4.77228920 + (3.85542044 + 0.32303607 * yy) * yy) * yy);
    Dual c2 = MHX_PI * xx * (1.0 + (4.88331068 + (4.28957333 + 0.52120508 * yy) * yy) * yy);
    return c1 / c2 + dy0(xx);
}
__device__ __forceinline__ Dual struve_h1_d(Dual xx) {
    if (xx.v <= 3.0) {
        Dual yy = (xx / 3.0) * (xx / 3.0);
        return (1.909859286 +
                (-1.145914713 + (0.294656958 + (-0.042070508 + (0.003785727 + -0.000207183 * yy) * yy) * yy) * yy) * yy) * yy;
    }
    Dual yy = (3.0 / xx) * (3.0 / xx);
    Dual c1 = 2.0 * (1.00000004 + (3.92205313 + (2.64893033 + 0.27450895 * yy) * yy) * yy);
    Dual c2 = MHX_PI * (1.0 + (3.81095112 + (2.26216956 + 0.10885141 * yy) * yy) * yy);
    return c1 / c2 + dy1(xx);
}

template <int DEG>
__device__ __forceinline__ Dual horner_wu_d(const double* c, Dual t) {
    Dual acc = (c[DEG - 1] + c[DEG] * t) * t;
    for (int i = DEG - 2; i >= 1; i--) acc = (c[i] + acc) * t;
    return c[0] + acc;
}

__constant__ static const double WUD_A[10] = {1.21, -13.328, 215.896, -1763.96, 8418.94, -24314.21, 42002.57, -41592.9, 21859.0, -4838.6};
__constant__ static const double WUD_B[10] = {0.938, 5.737, -67.92, 796.534, -4780.77, 17137.74, -36618.81, 44894.06, -29030.24, 7671.22};
__constant__ static const double WUD_C[8] = {1.268, -9.747, 209.653, -1397.89, 5155.67, -9844.35, 9136.4, -3272.62};
__constant__ static const double WUD_D[10] = {0.632, -40.97, 667.16, -6072.07, 31127.39, -96293.05, 181856.75, -205690.43, 128170.2, -33744.6};
__constant__ static const double WUD_DA[10] = {2.948, -24.53, 249.69, -754.85, -1187.71, 16370.75, -48811.41, 68220.87, -46688.0, 12622.25};
__constant__ static const double WUD_DB[10] = {1.11, 2.894, -76.765, 1565.35, -11336.19, 44270.15, -97014.11, 118879.26, -76209.82, 19923.28};
__constant__ static const double WUD_DC[6] = {14.19, -148.24, 847.8, -2318.58, 3168.35, -1590.27};

__device__ __noinline__ void wu_term_d(Dual hh, Dual vv, CDual& G, CDual& Gh, CDual& Gv) {
    Dual dd = dsqrt(hh * hh + vv * vv);
    Dual alpha = -vv / dd, beta = hh / dd, rho = dd / (1.0 + dd);
    Dual ev = dexp(vv);
    Dual omr3 = (1.0 - rho) * (1.0 - rho) * (1.0 - rho);
    Dual dd2 = dd * dd, dd3 = dd * dd * dd;
    Dual A = horner_wu_d<9>(WUD_A, rho), B = horner_wu_d<9>(WUD_B, rho), C = horner_wu_d<7>(WUD_C, rho),
         D = horner_wu_d<9>(WUD_D, rho);
    Dual RR = (1.0 - beta) * A;
    RR -= beta * B;
    RR -= alpha * C / (1.0 + 6.0 * alpha * rho * (1.0 - rho));
    RR += beta * (1.0 - beta) * D;
    Dual Lp = rho * omr3 * RR;
    Dual PP = dlog(0.5 * (dd - vv)) + 0.5772156649 - 2.0 * dd2;
    PP = ev * PP;
    PP = PP + dd2 - vv;
    Dual L0 = 2.0 * PP / (1.0 + dd3) + 2.0 * Lp;
    Dual H0 = struve_h0_d(hh), J0 = dj0(hh);
    Dual Wr = 2 * MHX_PI * H0 * ev, Wi = 2 * MHX_PI * (-J0) * ev;
    Dual dA = horner_wu_d<9>(WUD_DA, rho), dB = horner_wu_d<9>(WUD_DB, rho), dC = horner_wu_d<5>(WUD_DC, rho);
    Dual RS = beta * dA;
    RS -= (1.0 - alpha) * dB;
    RS += beta * (1.0 - beta) * rho * (1.0 - 2.0 * rho) * dC;
    Dual Lsp = rho * omr3 * RS;
    Dual PS = (beta + hh) / (dd - vv);
    PS = PS - 2.0 * beta + 2.0 * ev * dd - hh;
    Dual QS = dexp(-dd) * (1.0 - beta);
    QS = QS * (1.0 + dd / (1.0 + dd3));
    Dual Ls = 2.0 * PS / (1.0 + dd3) - 4.0 * QS + 2.0 * Lsp;
    Dual H1 = struve_h1_d(hh), J1 = dj1(hh);
    Dual Whr = 2 * MHX_PI * (2 / MHX_PI - H1) * ev, Whi = 2 * MHX_PI * J1 * ev;
    G = cdmk(-L0 - Wr, -Wi);
    Gh = cdmk(-Ls - Whr, -Whi);
    Gv = cdmk(G.x + 2.0 / dd, G.y);
}

__device__ __forceinline__ void reduced_hv_d(const Dual* x, const Dual* y, Dual k, Dual& hh, Dual& vv) {
    Dual d2 = (x[0] - y[0]) * (x[0] - y[0]) + (x[1] - y[1]) * (x[1] - y[1]);
    Dual dis;
    if (d2.v == 0.0)
        dis = dsqrt(d2 + 1e-18 * 1e-18) - 1e-18;
    else
        dis = dsqrt(d2);
    hh = k * dis;
    vv = k * (x[2] + y[2]);
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void tangent_geometry_kernel(int n, const double* __restrict__ verts, const double* __restrict__ dverts,
                                        const int64_t* __restrict__ faces, const double* __restrict__ centers,
                                        const double* __restrict__ dcenters, const double* __restrict__ normals,
                                        const double* __restrict__ dnormals, const double* __restrict__ areas,
                                        const double* __restrict__ dareas, const double* __restrict__ radii,
                                        PanelGeomD* __restrict__ geom) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    Dual V[12];
    for (int k = 0; k < 4; k++) {
        const int64_t v = faces[4 * j + k] - 1;
        for (int c = 0; c < 3; c++) V[3 * k + c] = dmk(verts[3 * v + c], dverts ? dverts[3 * v + c] : 0.0);
    }
    PanelGeomD g;
    birk_panel_geometry_d(V, g);
    for (int c = 0; c < 3; c++) {
        g.c[c] = dmk(centers[3 * j + c], dcenters ? dcenters[3 * j + c] : 0.0);
        g.nrm[c] = dmk(normals[3 * j + c], dnormals ? dnormals[3 * j + c] : 0.0);
    }
    g.area = dmk(areas[j], dareas ? dareas[j] : 0.0);
    g.radius = radii[j];
    geom[j] = g;
}

// dS, dD of every pair (row i = field panel, column j = source panel): Rankine parts (panel and
// image) and the wave term, as matrix_assembly.jl:274-369 / wu.jl:104-127 /
// exact_Guevel_Delhommeau.jl:61-75 in dual arithmetic.  EXACT: the wave term's value and
// first derivatives are read from the (G, dG/dr) table of the primal assembly.
template <bool EXACT>
__global__ void __launch_bounds__(256) tangent_pair_kernel(int n, const PanelGeomD* __restrict__ geom, double kv,
                                                           double kd, int direct, const cplx* __restrict__ WG,
                                                           const cplx* __restrict__ WR, cplx* __restrict__ dS,
                                                           cplx* __restrict__ dD) {
    const int j = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PanelGeomD& gj = geom[j];
    const PanelGeomD& gi = geom[i];
    const Dual k = dmk(kv, kd);
    const double kScaleD = -1.0 / (4 * MHX_PI);
    // Rankine (assemble_birk_rankine inner loop): panel and image
    const Dual c1[3] = {gi.c[0], gi.c[1], gi.c[2]};
    const Dual c1r[3] = {gi.c[0], gi.c[1], -gi.c[2]};
    const Dual* nv = direct ? gj.nrm : gi.nrm;
    Dual phi, gs[3], phir, gr[3];
    vrankine_d(c1, gj, phi, gs);
    vrankine_d(c1r, gj, phir, gr);
    Dual s = phi + phir;
    Dual d;
    if (direct) {
        d = nv[0] * gs[0] + nv[1] * gs[1] + nv[2] * gs[2];
    } else {
        d = nv[0] * (-gs[0]) + nv[1] * (-gs[1]) + nv[2] * (-gs[2]);
        gr[0] = -gr[0];
        gr[1] = -gr[1];
    }
    d += nv[0] * gr[0] + nv[1] * gr[1] + nv[2] * gr[2];
    const Dual sr = kScaleD * s, drk = kScaleD * d;
    // wave term
    Dual hh, vv;
    reduced_hv_d(gi.c, gj.c, k, hh, vv);
    CDual G, Gh, Gv;
    if (!EXACT) {
        wu_term_d(hh, vv, G, Gh, Gv);
    } else {
        const int64_t idx = i + (int64_t)n * j;
        const cplx g0 = WG[idx], gr0 = WR[idx];
        const double r = hh.v, z = vv.v, rho = hypot(r, z), r3 = rho * rho * rho;
        const cplx gz = cmk(g0.x + 2.0 / rho, g0.y);                     // dG/dz
        const cplx gzz = cmk(gz.x - 2.0 * z / r3, gz.y);                 // d2G/dz2
        const cplx grz = cmk(gr0.x - 2.0 * r / r3, gr0.y);               // d2G/drdz
        const cplx grr = r > 1e-6 ? cmk(-gr0.x / r - gzz.x, -gr0.y / r - gzz.y) : cmk(-0.5 * gzz.x, -0.5 * gzz.y);
        G = cdmk(dmk(g0.x, gr0.x * hh.d + gz.x * vv.d), dmk(g0.y, gr0.y * hh.d + gz.y * vv.d));
        Gh = cdmk(dmk(gr0.x, grr.x * hh.d + grz.x * vv.d), dmk(gr0.y, grr.y * hh.d + grz.y * vv.d));
        // Gv = G + 2/rho: its tangent is dG/dz's derivative
        Gv = cdmk(dmk(gz.x, grz.x * hh.d + gzz.x * vv.d), dmk(gz.y, grz.y * hh.d + gzz.y * vv.d));
    }
    // _wu_sd / combine (direct: Delta = xi - x, normal n_j; indirect: Delta = x - xi, normal n_i)
    Dual dx1 = direct ? gj.c[0] - gi.c[0] : gi.c[0] - gj.c[0];
    Dual dx2 = direct ? gj.c[1] - gi.c[1] : gi.c[1] - gj.c[1];
    Dual dr1 = dmk(0), dr2 = dmk(0);
    if (fabs(hh.v) > 1e-6) {
        dr1 = k * k / hh * dx1;
        dr2 = k * k / hh * dx2;
    }
    CDual g0 = cdscale(k, cdscale(dr1, Gh)), g1 = cdscale(k, cdscale(dr2, Gh)), g2 = cdscale(k, cdscale(k, Gv));
    CDual sw = cdscale(gj.area, cdscale(k, G));
    CDual dw = cdscale(gj.area, cdadd(cdadd(cdscale(nv[0], g0), cdscale(nv[1], g1)), cdscale(nv[2], g2)));
    const int64_t idx = i + (int64_t)n * j;
    dS[idx] = cmk(sr.d + kScaleD * sw.x.d, kScaleD * sw.y.d);
    dD[idx] = cmk(drk.d + kScaleD * dw.x.d, kScaleD * dw.y.d);
}

// radiation_bc (waves.jl:183-211) and _airy_bc (waves.jl:266-282) tangents
__global__ void tangent_bc_kernel(int n, const PanelGeomD* __restrict__ geom, const double* __restrict__ dofs,
                                  const double* __restrict__ ddofs, int ndof, const double* __restrict__ betas, int nb,
                                  double wv, double wd, double kv, double kd, double g, cplx* __restrict__ dB) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PanelGeomD& gi = geom[i];
    const Dual omega = dmk(wv, wd), k = dmk(kv, kd);
    for (int r = 0; r < ndof; r++) {
        const double* dof = dofs + (int64_t)3 * n * r;
        const double* ddof = ddofs ? ddofs + (int64_t)3 * n * r : nullptr;
        Dual disp = dmk(0);
        for (int c = 0; c < 3; c++) disp += gi.nrm[c] * dmk(dof[i + (int64_t)n * c], ddof ? ddof[i + (int64_t)n * c] : 0.0);
        const Dual im = -(omega * disp);  // bc = -i omega disp
        dB[i + (int64_t)n * r] = cmk(0.0, im.d);
    }
    for (int b = 0; b < nb; b++) {
        const double beta = betas[b];
        const double cb = cos(beta), sb = sin(beta);
        const Dual wbar = gi.c[0] * cb + gi.c[1] * sb;
        const Dual cih = dexp(k * gi.c[2]);
        Dual s, c;
        dsincos(k * wbar, s, c);
        const Dual am = g * k / omega;
        const CDual amp = cdmk(am * c, am * s);
        const CDual v0 = cdscale(cih, cdscale(dmk(cb), amp)), v1 = cdscale(cih, cdscale(dmk(sb), amp));
        const CDual v2 = cdscale(cih, cdmk(amp.y, -amp.x));  // -i * amp
        const CDual dot = cdadd(cdadd(cdscale(gi.nrm[0], v0), cdscale(gi.nrm[1], v1)), cdscale(gi.nrm[2], v2));
        dB[i + (int64_t)n * (ndof + b)] = cmk(-dot.x.d, -dot.y.d);
    }
}

// integrate_pressure (waves.jl:154-178) and froude_krylov_force (waves.jl:294-303) tangents:
// dF[d, r] for r < R from (phi, dphi); dFK[d, b] for the headings
__global__ void __launch_bounds__(256) tangent_forces_kernel(int n, const PanelGeomD* __restrict__ geom,
                                                             const double* __restrict__ dofs,
                                                             const double* __restrict__ ddofs, int ndof,
                                                             const cplx* __restrict__ phi, const cplx* __restrict__ dphi,
                                                             int R, double rho, double wv, double wd, double kv,
                                                             double kd, double g, const double* __restrict__ betas,
                                                             int nb, cplx* __restrict__ dF, cplx* __restrict__ dFK) {
    const int d = blockIdx.x, r = blockIdx.y;
    const double* dof = dofs + (int64_t)3 * n * d;
    const double* ddof = ddofs ? ddofs + (int64_t)3 * n * d : nullptr;
    const Dual omega = dmk(wv, wd), k = dmk(kv, kd);
    Dual accx = dmk(0), accy = dmk(0);
    for (int i = threadIdx.x; i < n; i += 256) {
        const PanelGeomD& gi = geom[i];
        Dual disp = dmk(0);
        for (int c = 0; c < 3; c++) disp += dmk(dof[i + (int64_t)n * c], ddof ? ddof[i + (int64_t)n * c] : 0.0) * gi.nrm[c];
        CDual p;
        if (r < R) {
            const cplx ph = phi[i + (int64_t)n * r], dph = dphi[i + (int64_t)n * r];
            const Dual ro = rho * omega;
            p = cdmk(-(ro * dmk(ph.y, dph.y)), ro * dmk(ph.x, dph.x));
        } else {
            const double beta = betas[r - R];
            const Dual wbar = gi.c[0] * cos(beta) + gi.c[1] * sin(beta);
            Dual s, c;
            dsincos(k * wbar, s, c);
            const Dual amp = g / omega * dexp(k * gi.c[2]);
            const CDual ph = cdmk(amp * s, -(amp * c));
            p = cdmk(-(omega * rho * ph.y), omega * rho * ph.x);
        }
        const Dual w = -(disp * gi.area);
        accx += w * p.x;
        accy += w * p.y;
    }
    __shared__ double red[2][256];
    red[0][threadIdx.x] = accx.d;
    red[1][threadIdx.x] = accy.d;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            red[0][threadIdx.x] += red[0][threadIdx.x + off];
            red[1][threadIdx.x] += red[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (r < R) dF[d + (int64_t)ndof * r] = cmk(red[0][0], red[1][0]);
        else dFK[d + (int64_t)ndof * (r - R)] = cmk(red[0][0], red[1][0]);
    }
}

// C = A + s B (n x R blocks, leading dimension n)
__global__ void axpby_kernel(int64_t cnt, const cplx* __restrict__ A, double s, const cplx* __restrict__ B,
                             cplx* __restrict__ C) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < cnt) C[t] = cmk(A[t].x + s * B[t].x, A[t].y + s * B[t].y);
}

void launch_tangent_geometry(hipStream_t st, int n, const double* verts, const double* dverts, const int64_t* faces,
                             const double* centers, const double* dcenters, const double* normals,
                             const double* dnormals, const double* areas, const double* dareas, const double* radii,
                             PanelGeomD* geom) {
    tangent_geometry_kernel<<<(n + 127) / 128, 128, 0, st>>>(n, verts, dverts, faces, centers, dcenters, normals,
                                                               dnormals, areas, dareas, radii, geom);
}

void launch_tangent_pairs(hipStream_t st, int n, const PanelGeomD* geom, double k, double dk, int direct, bool exact,
                          const cplx* WG, const cplx* WR, cplx* dS, cplx* dD) {
    dim3 grid((n + 255) / 256, n);
    if (exact)
        tangent_pair_kernel<true><<<grid, 256, 0, st>>>(n, geom, k, dk, direct, WG, WR, dS, dD);
    else
        tangent_pair_kernel<false><<<grid, 256, 0, st>>>(n, geom, k, dk, direct, WG, WR, dS, dD);
}

void launch_tangent_bc(hipStream_t st, int n, const PanelGeomD* geom, const double* dofs, const double* ddofs, int ndof,
                       const double* betas, int nb, double w, double dw, double k, double dk, double g, cplx* dB) {
    tangent_bc_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, geom, dofs, ddofs, ndof, betas, nb, w, dw, k, dk, g, dB);
}

void launch_tangent_forces(hipStream_t st, int n, const PanelGeomD* geom, const double* dofs, const double* ddofs,
                           int ndof, const cplx* phi, const cplx* dphi, int R, double rho, double w, double dw, double k,
                           double dk, double g, const double* betas, int nb, cplx* dF, cplx* dFK) {
    if (ndof <= 0) return;
    tangent_forces_kernel<<<dim3(ndof, R + nb), 256, 0, st>>>(n, geom, dofs, ddofs, ndof, phi, dphi, R, rho, w, dw, k,
                                                              dk, g, betas, nb, dF, dFK);
}

void launch_axpby(hipStream_t st, int64_t cnt, const cplx* A, double s, const cplx* B, cplx* C) {
    axpby_kernel<<<(unsigned)((cnt + 255) / 256), 256, 0, st>>>(cnt, A, s, B, C);
}

}  // namespace mhx
