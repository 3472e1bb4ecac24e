// Derivatives of the omega-solve.
//
// Forward mode: the directional derivative of the hydrodynamic coefficients along a tangent of
// the frequency and of the mesh fields (vertices, centres, normals, areas) and dof fields.
// Reverse mode: the gradient of a real functional of the coefficients with respect to all of
// those inputs at once.  Both give what the reference obtains with ForwardDiff / Enzyme through
// its rules (src/ad_rules.jl:79-414).
//
// The assembly is differentiated in dual arithmetic (a value and N partials) that follows the
// primal kernels operation by operation (Birk panels and image, the Wu fit, the boundary
// conditions and the pressure integration); the exact Guevel-Delhommeau term's derivatives come
// from its table (G, dG/dr) and the relations dG/dz = G + 2/rho, d2G/dz2 = dG/dz - 2z/rho^3,
// d2G/drdz = dG/dr - 2r/rho^3 and Laplace's equation d2G/dr2 = -dG/dr / r - d2G/dz2 (the wave
// term is harmonic below the free surface).
//
// Forward: the linear solve reuses the primal LU: dphi = D^{-1} (dS B + S dB - dD phi)
// (direct), dsigma = D^{-1} (dB - dD sigma), dphi = dS sigma + S dsigma (indirect).
// Reverse: with <a, b> = sum conj(a) b and J = Re <ybar, y>, the transposed solve on the same LU
// gives lambda = D^{-H} phibar, Dbar = -lambda phi^H, Sbar = lambda B^H, Bbar = S^H lambda
// (direct); sigmabar = S^H phibar, mu = D^{-H} sigmabar, Sbar = phibar sigma^H,
// Dbar = -mu sigma^H, Bbar = mu (indirect) -- the reference's rules for _linearsolve and
// _mulvec (ad_rules.jl:300-414).  Sbar and Dbar have rank R and are never formed: the pair
// kernel evaluates Sbar_ij = sum_r U_ir conj(V_jr) and pulls it back through the pair function
// with N-partial duals, one pass per N of the pair's 26 local inputs.
#include "mhx_tangent.h"

namespace mhx {

// ---------------------------------------------------------------------------
// dual numbers
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ DualN<N> dcst(double v) {
    DualN<N> r;
    r.v = v;
#pragma unroll
    for (int p = 0; p < N; p++) r.d[p] = 0.0;
    return r;
}
__device__ __forceinline__ Dual dmk(double v, double d = 0.0) {
    Dual r;
    r.v = v;
    r.d[0] = d;
    return r;
}
// value v with partials s * a
template <int N>
__device__ __forceinline__ DualN<N> dfun(double v, double s, const DualN<N>& a) {
    DualN<N> r;
    r.v = v;
#pragma unroll
    for (int p = 0; p < N; p++) r.d[p] = s * a.d[p];
    return r;
}
// value v with partials sa * a + sb * b
template <int N>
__device__ __forceinline__ DualN<N> dfun2(double v, double sa, const DualN<N>& a, double sb, const DualN<N>& b) {
    DualN<N> r;
    r.v = v;
#pragma unroll
    for (int p = 0; p < N; p++) r.d[p] = sa * a.d[p] + sb * b.d[p];
    return r;
}
template <int N>
__device__ __forceinline__ DualN<N> operator+(const DualN<N>& a, const DualN<N>& b) {
    DualN<N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int p = 0; p < N; p++) r.d[p] = a.d[p] + b.d[p];
    return r;
}
template <int N>
__device__ __forceinline__ DualN<N> operator-(const DualN<N>& a, const DualN<N>& b) {
    DualN<N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int p = 0; p < N; p++) r.d[p] = a.d[p] - b.d[p];
    return r;
}
template <int N>
__device__ __forceinline__ DualN<N> operator-(const DualN<N>& a) {
    return dfun(-a.v, -1.0, a);
}
template <int N>
__device__ __forceinline__ DualN<N> operator*(const DualN<N>& a, const DualN<N>& b) {
    return dfun2(a.v * b.v, b.v, a, a.v, b);
}
template <int N>
__device__ __forceinline__ DualN<N> operator/(const DualN<N>& a, const DualN<N>& b) {
    const double q = a.v / b.v;
    return dfun2(q, 1.0 / b.v, a, -q / b.v, b);
}
template <int N>
__device__ __forceinline__ DualN<N> operator+(const DualN<N>& a, double b) {
    DualN<N> r = a;
    r.v += b;
    return r;
}
template <int N>
__device__ __forceinline__ DualN<N> operator+(double a, const DualN<N>& b) {
    return b + a;
}
template <int N>
__device__ __forceinline__ DualN<N> operator-(const DualN<N>& a, double b) {
    DualN<N> r = a;
    r.v -= b;
    return r;
}
template <int N>
__device__ __forceinline__ DualN<N> operator-(double a, const DualN<N>& b) {
    return dfun(a - b.v, -1.0, b);
}
template <int N>
__device__ __forceinline__ DualN<N> operator*(const DualN<N>& a, double b) {
    return dfun(a.v * b, b, a);
}
template <int N>
__device__ __forceinline__ DualN<N> operator*(double a, const DualN<N>& b) {
    return dfun(a * b.v, a, b);
}
template <int N>
__device__ __forceinline__ DualN<N> operator/(const DualN<N>& a, double b) {
    return dfun(a.v / b, 1.0 / b, a);
}
template <int N>
__device__ __forceinline__ DualN<N> operator/(double a, const DualN<N>& b) {
    const double q = a / b.v;
    return dfun(q, -q / b.v, b);
}
template <int N>
__device__ __forceinline__ DualN<N>& operator+=(DualN<N>& a, const DualN<N>& b) {
    return a = a + b;
}
template <int N>
__device__ __forceinline__ DualN<N>& operator-=(DualN<N>& a, const DualN<N>& b) {
    return a = a - b;
}
// at 0 the partials are taken as 0 (the argument is a sum of squares whose partials vanish there)
template <int N>
__device__ __forceinline__ DualN<N> dsqrt(const DualN<N>& a) {
    const double s = sqrt(a.v);
    return dfun(s, s > 0.0 ? 0.5 / s : 0.0, a);
}
template <int N>
__device__ __forceinline__ DualN<N> dlog(const DualN<N>& a) {
    return dfun(log(a.v), 1.0 / a.v, a);
}
template <int N>
__device__ __forceinline__ DualN<N> dexp(const DualN<N>& a) {
    const double e = exp(a.v);
    return dfun(e, e, a);
}
template <int N>
__device__ __forceinline__ DualN<N> datan2(const DualN<N>& y, const DualN<N>& x) {
    const double den = x.v * x.v + y.v * y.v;
    if (!(den > 0.0)) return dcst<N>(atan2(y.v, x.v));
    return dfun2(atan2(y.v, x.v), x.v / den, y, -y.v / den, x);
}
template <int N>
__device__ __forceinline__ DualN<N> dhypot(const DualN<N>& a, const DualN<N>& b) {
    const double h = hypot(a.v, b.v);
    if (!(h > 0.0)) return dcst<N>(h);
    return dfun2(h, a.v / h, a, b.v / h, b);
}
template <int N>
__device__ __forceinline__ void dsincos(const DualN<N>& a, DualN<N>& s, DualN<N>& c) {
    double sv, cv;
    sincos(a.v, &sv, &cv);
    s = dfun(sv, cv, a);
    c = dfun(cv, -sv, a);
}
// Bessel functions with their exact derivatives (J0' = -J1, J1' = J0 - J1/x, same for Y)
template <int N>
__device__ __forceinline__ DualN<N> dj0(const DualN<N>& x) {
    return dfun(j0(x.v), -j1(x.v), x);
}
template <int N>
__device__ __forceinline__ DualN<N> dj1(const DualN<N>& x) {
    if (x.v == 0.0) return dfun(0.0, 0.5, x);  // J1'(0) = 1/2
    const double a = j0(x.v), b = j1(x.v);
    return dfun(b, a - b / x.v, x);
}
template <int N>
__device__ __forceinline__ DualN<N> dy0(const DualN<N>& x) {
    return dfun(y0(x.v), -y1(x.v), x);
}
template <int N>
__device__ __forceinline__ DualN<N> dy1(const DualN<N>& x) {
    const double a = y0(x.v), b = y1(x.v);
    return dfun(b, a - b / x.v, x);
}

// complex dual: (re, im) each a dual
template <int N>
struct CDualN {
    DualN<N> x, y;
};
template <int N>
__device__ __forceinline__ CDualN<N> cdmk(const DualN<N>& x, const DualN<N>& y) {
    return CDualN<N>{x, y};
}
template <int N>
__device__ __forceinline__ CDualN<N> cdadd(const CDualN<N>& a, const CDualN<N>& b) {
    return CDualN<N>{a.x + b.x, a.y + b.y};
}
template <int N>
__device__ __forceinline__ CDualN<N> cdscale(const DualN<N>& s, const CDualN<N>& a) {
    return CDualN<N>{s * a.x, s * a.y};
}

// ---------------------------------------------------------------------------
// Birk panel and Rankine integrals in dual arithmetic (mirrors mhx_device.h)
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ DualN<N> atan_ratio_d(const DualN<N>& num, const DualN<N>& den) {
    const bool neg = den.v < 0.0;
    return datan2(neg ? -num : num, neg ? -den : den);
}

template <int N>
__device__ void birk_panel_geometry_d(const DualN<N> V[12], PanelGeomN<N>& g) {
    DualN<N> mid[4][3];
    for (int i = 0; i < 4; i++) {
        int j = (i + 1) & 3;
        for (int c = 0; c < 3; c++) mid[i][c] = (V[3 * i + c] + V[3 * j + c]) / 2.0;
    }
    DualN<N> mgc[3], s[3], tt[3], n[3], t[3];
    for (int c = 0; c < 3; c++) {
        mgc[c] = (mid[1][c] + mid[3][c]) / 2.0;
        s[c] = mid[1][c] - mid[3][c];
        tt[c] = mid[2][c] - mid[0][c];
    }
    DualN<N> ns = dsqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (int c = 0; c < 3; c++) s[c] = s[c] / ns;
    n[0] = s[1] * tt[2] - s[2] * tt[1];
    n[1] = s[2] * tt[0] - s[0] * tt[2];
    n[2] = s[0] * tt[1] - s[1] * tt[0];
    DualN<N> nn = dsqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int c = 0; c < 3; c++) n[c] = n[c] / nn;
    t[0] = n[1] * s[2] - n[2] * s[1];
    t[1] = n[2] * s[0] - n[0] * s[2];
    t[2] = n[0] * s[1] - n[1] * s[0];
    DualN<N> nt = dsqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    for (int c = 0; c < 3; c++) t[c] = t[c] / nt;
    for (int c = 0; c < 3; c++) {
        g.T[c] = s[c];
        g.T[3 + c] = t[c];
        g.T[6 + c] = n[c];
    }
    DualN<N> tc[4][2];
    for (int k = 0; k < 4; k++) {
        DualN<N> d0 = V[3 * k] - mgc[0], d1 = V[3 * k + 1] - mgc[1], d2 = V[3 * k + 2] - mgc[2];
        tc[k][0] = d0 * s[0] + d1 * s[1] + d2 * s[2];
        tc[k][1] = d0 * t[0] + d1 * t[1] + d2 * t[2];
    }
    DualN<N> A = dcst<N>(0.0), Mx = dcst<N>(0.0), My = dcst<N>(0.0);
    for (int k = 0; k < 4; k++) {
        int k1 = (k + 1) & 3;
        const DualN<N> &xk = tc[k][0], &yk = tc[k][1], &xk1 = tc[k1][0], &yk1 = tc[k1][1];
        A += (yk1 - yk) * (xk + xk1) / 2.0;
        Mx += (xk1 - xk) * (yk * yk + yk * yk1 + yk1 * yk1) / 6.0;
        My += (yk1 - yk) * (xk * xk + xk * xk1 + xk1 * xk1) / 6.0;
    }
    DualN<N> clx = My / A, cly = -Mx / A;
    for (int c = 0; c < 3; c++) g.qgc[c] = mgc[c] + clx * s[c] + cly * t[c];
    for (int k = 0; k < 4; k++) {
        g.lc[2 * k] = tc[k][0] - clx;
        g.lc[2 * k + 1] = tc[k][1] - cly;
    }
}

template <int N>
__device__ void birk_phi_v_d(const DualN<N>& x, const DualN<N>& y, const DualN<N>& z, const DualN<N>* lc, DualN<N>& phi,
                             DualN<N> v[3]) {
    DualN<N> sx[4], sy[4], r[4], e[4], h[4];
    for (int k = 0; k < 4; k++) {
        sx[k] = x - lc[2 * k];
        sy[k] = y - lc[2 * k + 1];
        r[k] = dsqrt(sx[k] * sx[k] + sy[k] * sy[k] + z * z);
        e[k] = sx[k] * sx[k] + z * z;
        h[k] = sx[k] * sy[k];
    }
    DualN<N> s0 = dcst<N>(0.0), s1 = dcst<N>(0.0), s2 = dcst<N>(0.0), s3 = dcst<N>(0.0);
    const bool inplane = fabs(z.v) <= 1e-14;
    for (int k = 0; k < 4; k++) {
        int kn = (k + 1) & 3;
        DualN<N> dx = lc[2 * kn] - lc[2 * k], dy = lc[2 * kn + 1] - lc[2 * k + 1];
        DualN<N> dk = dhypot(dx, dy);
        if (dk.v <= 1e-14) continue;  // degenerate edge (triangle stored as a quad): no term
        DualN<N> rsum = r[k] + r[kn];
        DualN<N> lt_dk = dlog((rsum - dk) / (rsum + dk)) / dk;
        // in the panel's plane the solid-angle term is dropped with zero partials
        DualN<N> vzk = inplane ? dcst<N>(0.0)
                               : atan_ratio_d(dy * e[k] - dx * h[k], dx * z * r[k]) -
                                     atan_ratio_d(dy * e[kn] - dx * h[kn], dx * z * r[kn]);
        s0 += (sx[k] * dy - sy[k] * dx) * lt_dk + z * vzk;
        s1 += dy * lt_dk;
        s2 += -dx * lt_dk;
        s3 += vzk;
    }
    const double scale = -1.0 / (4 * MHX_PI);
    phi = scale * s0;
    v[0] = scale * s1;
    v[1] = scale * s2;
    v[2] = scale * s3;
}

template <int N>
__device__ void vrankine_d(const DualN<N> p[3], const PanelGeomN<N>& g, DualN<N>& phi, DualN<N> grad[3]) {
    DualN<N> rb0 = p[0] - g.c[0], rb1 = p[1] - g.c[1], rb2 = p[2] - g.c[2];
    DualN<N> r = dsqrt(rb0 * rb0 + rb1 * rb1 + rb2 * rb2);
    if (r.v > 7 * g.radius) {
        DualN<N> r3 = r * r * r;
        phi = g.area / r;
        grad[0] = (rb0 / r3) * g.area;
        grad[1] = (rb1 / r3) * g.area;
        grad[2] = (rb2 / r3) * g.area;
        return;
    }
    DualN<N> d0 = p[0] - g.qgc[0], d1 = p[1] - g.qgc[1], d2 = p[2] - g.qgc[2];
    DualN<N> fl0 = g.T[0] * d0 + g.T[1] * d1 + g.T[2] * d2;
    DualN<N> fl1 = g.T[3] * d0 + g.T[4] * d1 + g.T[5] * d2;
    DualN<N> fl2 = g.T[6] * d0 + g.T[7] * d1 + g.T[8] * d2;
    DualN<N> pb, vl[3];
    birk_phi_v_d(fl0, fl1, fl2, g.lc, pb, vl);
    phi = -4 * MHX_PI * pb;
    for (int q = 0; q < 3; q++) grad[q] = 4 * MHX_PI * (g.T[q] * vl[0] + g.T[3 + q] * vl[1] + g.T[6 + q] * vl[2]);
}

// ---------------------------------------------------------------------------
// Wu fit in dual arithmetic (mirrors wu_term)
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ DualN<N> struve_h0_d(const DualN<N>& xx) {
    if (xx.v <= 3.0) {
        DualN<N> yy = (xx / 3.0) * (xx / 3.0);
        DualN<N> h = 1.909859164 +
                     (-1.909855001 + (0.687514637 + (-0.126164557 + (0.013828813 + -0.000876918 * yy) * yy) * yy) * yy) *
                         yy;
        return h * (xx / 3.0);
    }
    DualN<N> yy = (3.0 / xx) * (3.0 / xx);
    DualN<N> c1 = 2.0 * (0.99999906 + (4.77228920 + (3.85542044 + 0.32303607 * yy) * yy) * yy);
    DualN<N> c2 = MHX_PI * xx * (1.0 + (4.88331068 + (4.28957333 + 0.52120508 * yy) * yy) * yy);
    return c1 / c2 + dy0(xx);
}
template <int N>
__device__ __forceinline__ DualN<N> struve_h1_d(const DualN<N>& xx) {
    if (xx.v <= 3.0) {
        DualN<N> yy = (xx / 3.0) * (xx / 3.0);
        return (1.909859286 +
                (-1.145914713 + (0.294656958 + (-0.042070508 + (0.003785727 + -0.000207183 * yy) * yy) * yy) * yy) * yy) *
               yy;
    }
    DualN<N> yy = (3.0 / xx) * (3.0 / xx);
    DualN<N> c1 = 2.0 * (1.00000004 + (3.92205313 + (2.64893033 + 0.27450895 * yy) * yy) * yy);
    DualN<N> c2 = MHX_PI * (1.0 + (3.81095112 + (2.26216956 + 0.10885141 * yy) * yy) * yy);
    return c1 / c2 + dy1(xx);
}

template <int DEG, int N>
__device__ __forceinline__ DualN<N> horner_wu_d(const double* c, const DualN<N>& t) {
    DualN<N> acc = (c[DEG - 1] + c[DEG] * t) * t;
    for (int i = DEG - 2; i >= 1; i--) acc = (c[i] + acc) * t;
    return c[0] + acc;
}

template <int N>
__device__ __noinline__ void wu_term_d(DualN<N> hh, DualN<N> vv, CDualN<N>& G, CDualN<N>& Gh, CDualN<N>& Gv) {
    DualN<N> dd = dsqrt(hh * hh + vv * vv);
    DualN<N> alpha = -vv / dd, beta = hh / dd, rho = dd / (1.0 + dd);
    DualN<N> ev = dexp(vv);
    DualN<N> omr3 = (1.0 - rho) * (1.0 - rho) * (1.0 - rho);
    DualN<N> dd2 = dd * dd, dd3 = dd * dd * dd;
    DualN<N> A = horner_wu_d<9>(WU_A, rho), B = horner_wu_d<9>(WU_B, rho), C = horner_wu_d<7>(WU_C, rho),
             D = horner_wu_d<9>(WU_D, rho);
    DualN<N> RR = (1.0 - beta) * A;
    RR -= beta * B;
    RR -= alpha * C / (1.0 + 6.0 * alpha * rho * (1.0 - rho));
    RR += beta * (1.0 - beta) * D;
    DualN<N> Lp = rho * omr3 * RR;
    DualN<N> PP = dlog(0.5 * (dd - vv)) + 0.5772156649 - 2.0 * dd2;
    PP = ev * PP;
    PP = PP + dd2 - vv;
    DualN<N> L0 = 2.0 * PP / (1.0 + dd3) + 2.0 * Lp;
    DualN<N> H0 = struve_h0_d(hh), J0 = dj0(hh);
    DualN<N> Wr = 2 * MHX_PI * H0 * ev, Wi = 2 * MHX_PI * (-J0) * ev;
    DualN<N> dA = horner_wu_d<9>(WU_DA, rho), dB = horner_wu_d<9>(WU_DB, rho), dC = horner_wu_d<5>(WU_DC, rho);
    DualN<N> RS = beta * dA;
    RS -= (1.0 - alpha) * dB;
    RS += beta * (1.0 - beta) * rho * (1.0 - 2.0 * rho) * dC;
    DualN<N> Lsp = rho * omr3 * RS;
    DualN<N> PS = (beta + hh) / (dd - vv);
    PS = PS - 2.0 * beta + 2.0 * ev * dd - hh;
    DualN<N> QS = dexp(-dd) * (1.0 - beta);
    QS = QS * (1.0 + dd / (1.0 + dd3));
    DualN<N> Ls = 2.0 * PS / (1.0 + dd3) - 4.0 * QS + 2.0 * Lsp;
    DualN<N> H1 = struve_h1_d(hh), J1 = dj1(hh);
    DualN<N> Whr = 2 * MHX_PI * (2 / MHX_PI - H1) * ev, Whi = 2 * MHX_PI * J1 * ev;
    G = cdmk(-L0 - Wr, -Wi);
    Gh = cdmk(-Ls - Whr, -Whi);
    Gv = cdmk(G.x + 2.0 / dd, G.y);
}

template <int N>
__device__ __forceinline__ void reduced_hv_d(const DualN<N>* x, const DualN<N>* y, const DualN<N>& k, DualN<N>& hh,
                                             DualN<N>& vv) {
    DualN<N> d2 = (x[0] - y[0]) * (x[0] - y[0]) + (x[1] - y[1]) * (x[1] - y[1]);
    DualN<N> dis;
    if (d2.v == 0.0)
        dis = dsqrt(d2 + 1e-18 * 1e-18) - 1e-18;
    else
        dis = dsqrt(d2);
    hh = k * dis;
    vv = k * (x[2] + y[2]);
}

// One pair (row i = field panel with centre ci and normal ni, column j = source panel gj):
// S_ij and D_ij without the identity term (out: Re S, Im S, Re D, Im D), as the Rankine
// assembly (matrix_assembly.jl:274-341), wu.jl:104-127 / exact_Guevel_Delhommeau.jl:61-75 and
// the combine of matrix_assembly.jl:343-369.  EXACT: the wave term's value and first
// derivatives come from the table entries wg = G, wr = dG/dr of the primal assembly.
template <int N, bool EXACT>
__device__ __forceinline__ void pair_eval(const DualN<N>* ci, const DualN<N>* ni, const PanelGeomN<N>& gj,
                                          const DualN<N>& k, int direct, cplx wg, cplx wr, DualN<N> out[4]) {
    const double kScaleD = -1.0 / (4 * MHX_PI);
    const DualN<N> c1r[3] = {ci[0], ci[1], -ci[2]};
    const DualN<N>* nv = direct ? gj.nrm : ni;
    DualN<N> phi, gs[3], phir, gr[3];
    vrankine_d(ci, gj, phi, gs);
    vrankine_d(c1r, gj, phir, gr);
    DualN<N> s = phi + phir;
    DualN<N> d;
    if (direct) {
        d = nv[0] * gs[0] + nv[1] * gs[1] + nv[2] * gs[2];
    } else {
        d = -(nv[0] * gs[0] + nv[1] * gs[1] + nv[2] * gs[2]);
        gr[0] = -gr[0];
        gr[1] = -gr[1];
    }
    d += nv[0] * gr[0] + nv[1] * gr[1] + nv[2] * gr[2];
    // wave term
    DualN<N> hh, vv;
    reduced_hv_d(ci, gj.c, k, hh, vv);
    CDualN<N> G, Gh, Gv;
    if (!EXACT) {
        wu_term_d(hh, vv, G, Gh, Gv);
    } else {
        const double r = hh.v, z = vv.v, rho = hypot(r, z), r3 = rho * rho * rho;
        const cplx gz = cmk(wg.x + 2.0 / rho, wg.y);        // dG/dz
        const cplx gzz = cmk(gz.x - 2.0 * z / r3, gz.y);    // d2G/dz2
        const cplx grz = cmk(wr.x - 2.0 * r / r3, wr.y);    // d2G/drdz
        const cplx grr = r > 1e-6 ? cmk(-wr.x / r - gzz.x, -wr.y / r - gzz.y) : cmk(-0.5 * gzz.x, -0.5 * gzz.y);
        G = cdmk(dfun2(wg.x, wr.x, hh, gz.x, vv), dfun2(wg.y, wr.y, hh, gz.y, vv));
        Gh = cdmk(dfun2(wr.x, grr.x, hh, grz.x, vv), dfun2(wr.y, grr.y, hh, grz.y, vv));
        // Gv = G + 2/rho = dG/dz
        Gv = cdmk(dfun2(gz.x, grz.x, hh, gzz.x, vv), dfun2(gz.y, grz.y, hh, gzz.y, vv));
    }
    // _wu_sd / combine (direct: Delta = xi - x, normal n_j; indirect: Delta = x - xi, normal n_i)
    DualN<N> dx1 = direct ? gj.c[0] - ci[0] : ci[0] - gj.c[0];
    DualN<N> dx2 = direct ? gj.c[1] - ci[1] : ci[1] - gj.c[1];
    DualN<N> dr1 = dcst<N>(0.0), dr2 = dcst<N>(0.0);
    if (fabs(hh.v) > 1e-6) {
        dr1 = k * k / hh * dx1;
        dr2 = k * k / hh * dx2;
    }
    CDualN<N> g0 = cdscale(k, cdscale(dr1, Gh)), g1 = cdscale(k, cdscale(dr2, Gh)), g2 = cdscale(k, cdscale(k, Gv));
    CDualN<N> sw = cdscale(gj.area, cdscale(k, G));
    CDualN<N> dw = cdscale(gj.area, cdadd(cdadd(cdscale(nv[0], g0), cdscale(nv[1], g1)), cdscale(nv[2], g2)));
    out[0] = kScaleD * s + kScaleD * sw.x;
    out[1] = kScaleD * sw.y;
    out[2] = kScaleD * d + kScaleD * dw.x;
    out[3] = kScaleD * dw.y;
}

// ---------------------------------------------------------------------------
// forward-mode kernels
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

// dS, dD of every pair (row i = field panel, column j = source panel)
template <bool EXACT>
__global__ void __launch_bounds__(256) tangent_pair_kernel(int n, const PanelGeomD* __restrict__ geom, double kv,
                                                           double kd, int direct, const cplx* __restrict__ WG,
                                                           const cplx* __restrict__ WR, cplx* __restrict__ dS,
                                                           cplx* __restrict__ dD) {
    const int j = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PanelGeomD& gi = geom[i];
    const int64_t idx = i + (int64_t)n * j;
    const cplx wg = EXACT ? WG[idx] : cmk(0.0, 0.0), wr = EXACT ? WR[idx] : cmk(0.0, 0.0);
    Dual out[4];
    pair_eval<1, EXACT>(gi.c, gi.nrm, geom[j], dmk(kv, kd), direct, wg, wr, out);
    dS[idx] = cmk(out[0].d[0], out[1].d[0]);
    dD[idx] = cmk(out[2].d[0], out[3].d[0]);
}

// dK_x of every pair: the indirect operator with every field-panel normal replaced by e_x (the
// forward-speed operator of solve.jl:35-45, launch_kx_combine), tangent along k only (the mesh is
// fixed); the identity term on the diagonal does not depend on k
template <bool EXACT>
__global__ void __launch_bounds__(256) tangent_kx_kernel(int n, const PanelGeomD* __restrict__ geom, double kv,
                                                         double kd, const cplx* __restrict__ WG,
                                                         const cplx* __restrict__ WR, cplx* __restrict__ dKx) {
    const int j = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PanelGeomD& gi = geom[i];
    const int64_t idx = i + (int64_t)n * j;
    const cplx wg = EXACT ? WG[idx] : cmk(0.0, 0.0), wr = EXACT ? WR[idx] : cmk(0.0, 0.0);
    const Dual ex[3] = {dmk(1.0, 0.0), dmk(0.0, 0.0), dmk(0.0, 0.0)};
    Dual out[4];
    pair_eval<1, EXACT>(gi.c, ex, geom[j], dmk(kv, kd), 0, wg, wr, out);
    dKx[idx] = cmk(out[2].d[0], out[3].d[0]);
}

// tangent of the forward-speed diffraction condition (bc_fs_kernel: the incident wave of
// wavenumber k and frequency omega, phase along the encountered heading beta_e, velocity along
// the incident heading beta; hull panels [0, nhull) only) along (omega, k)
__global__ void tangent_bc_inc_kernel(int n, int nhull, const PanelGeomD* __restrict__ geom, double wv, double wd,
                                      double kv, double kd, double beta, double beta_e, double g,
                                      cplx* __restrict__ dB) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i >= nhull) {
        dB[i] = cmk(0.0, 0.0);
        return;
    }
    const PanelGeomD& gi = geom[i];
    const Dual omega = dmk(wv, wd), k = dmk(kv, kd);
    const Dual wbar = gi.c[0] * cos(beta_e) + gi.c[1] * sin(beta_e);
    const Dual cih = dexp(k * gi.c[2]);
    Dual sn, cs;
    dsincos(k * wbar, sn, cs);
    const Dual am = g * k / omega;
    const CDualN<1> amp = cdmk(am * cs, am * sn);
    const CDualN<1> v0 = cdscale(cih, cdscale(dmk(cos(beta)), amp)), v1 = cdscale(cih, cdscale(dmk(sin(beta)), amp));
    const CDualN<1> v2 = cdscale(cih, cdmk(amp.y, -amp.x));  // -i * amp
    const CDualN<1> dot = cdadd(cdadd(cdscale(gi.nrm[0], v0), cdscale(gi.nrm[1], v1)), cdscale(gi.nrm[2], v2));
    dB[i] = cmk(-dot.x.d[0], -dot.y.d[0]);
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
        dB[i + (int64_t)n * r] = cmk(0.0, im.d[0]);
    }
    for (int b = 0; b < nb; b++) {
        const double beta = betas[b];
        const double cb = cos(beta), sb = sin(beta);
        const Dual wbar = gi.c[0] * cb + gi.c[1] * sb;
        const Dual cih = dexp(k * gi.c[2]);
        Dual s, c;
        dsincos(k * wbar, s, c);
        const Dual am = g * k / omega;
        const CDualN<1> amp = cdmk(am * c, am * s);
        const CDualN<1> v0 = cdscale(cih, cdscale(dmk(cb), amp)), v1 = cdscale(cih, cdscale(dmk(sb), amp));
        const CDualN<1> v2 = cdscale(cih, cdmk(amp.y, -amp.x));  // -i * amp
        const CDualN<1> dot = cdadd(cdadd(cdscale(gi.nrm[0], v0), cdscale(gi.nrm[1], v1)), cdscale(gi.nrm[2], v2));
        dB[i + (int64_t)n * (ndof + b)] = cmk(-dot.x.d[0], -dot.y.d[0]);
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
        CDualN<1> p;
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
            const CDualN<1> ph = cdmk(amp * s, -(amp * c));
            p = cdmk(-(omega * rho * ph.y), omega * rho * ph.x);
        }
        const Dual w = -(disp * gi.area);
        accx += w * p.x;
        accy += w * p.y;
    }
    __shared__ double red[2][256];
    red[0][threadIdx.x] = accx.d[0];
    red[1][threadIdx.x] = accy.d[0];
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

// ---------------------------------------------------------------------------
// reverse-mode kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ double re_conj_mul(cplx a, cplx b) { return a.x * b.x + a.y * b.y; }  // Re(conj(a) b)

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// integrate_pressure / froude_krylov_force pulled back (one thread per panel i):
// F[d, r] = sum_i w_id p_ir, w_id = -(dof_id . n_i) a_i, p_ir = i omega rho phi_ir (r < R) or
// p = rho g e^{k z} e^{i k (x cos b + y sin b)} (Froude-Krylov, heading b).
__global__ void __launch_bounds__(256) adjoint_forces_kernel(int n, const PanelGeom* __restrict__ geom,
                                                             const double* __restrict__ dofs, int ndof,
                                                             const cplx* __restrict__ phi, int R,
                                                             const cplx* __restrict__ Fbar,
                                                             const cplx* __restrict__ FKbar, double rho, double w,
                                                             double k, double g, const double* __restrict__ betas,
                                                             int nb, cplx* __restrict__ phibar, AdjointGrad gr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double wbar_acc = 0.0, kbar_acc = 0.0;
    if (i < n) {
        const PanelGeom& gi = geom[i];
        const double a = gi.area;
        double nbar[3] = {0, 0, 0}, cbar[3] = {0, 0, 0}, abar = 0.0;
        // phibar_ir = -i omega rho sum_d w_id Fbar[d, r]
        for (int r = 0; r < R; r++) {
            cplx acc = cmk(0.0, 0.0);
            for (int d = 0; d < ndof; d++) {
                const double* dof = dofs + (int64_t)3 * n * d;
                const double wid = -(dof[i] * gi.nrm[0] + dof[i + n] * gi.nrm[1] + dof[i + 2 * (int64_t)n] * gi.nrm[2]) * a;
                const cplx fb = Fbar[d + (int64_t)ndof * r];
                acc = cmk(acc.x + wid * fb.x, acc.y + wid * fb.y);
            }
            const double ro = w * rho;
            phibar[i + (int64_t)n * r] = cmk(ro * acc.y, -ro * acc.x);  // -i ro acc
        }
        for (int d = 0; d < ndof; d++) {
            const double* dof = dofs + (int64_t)3 * n * d;
            const double dv[3] = {dof[i], dof[i + n], dof[i + 2 * (int64_t)n]};
            const double s = dv[0] * gi.nrm[0] + dv[1] * gi.nrm[1] + dv[2] * gi.nrm[2];
            const double wid = -s * a;
            double Wbar = 0.0;
            for (int r = 0; r < R; r++) {
                const cplx ph = phi[i + (int64_t)n * r];
                const cplx p = cmk(-w * rho * ph.y, w * rho * ph.x);  // i omega rho phi
                const cplx fb = Fbar[d + (int64_t)ndof * r];
                Wbar += re_conj_mul(fb, p);
                wbar_acc += wid * re_conj_mul(fb, cmk(-rho * ph.y, rho * ph.x));  // dp/domega = i rho phi
            }
            for (int b = 0; b < nb; b++) {
                const double cb = cos(betas[b]), sb = sin(betas[b]);
                const double wb = gi.c[0] * cb + gi.c[1] * sb;
                double sn, cs;
                sincos(k * wb, &sn, &cs);
                const double amp = rho * g * exp(k * gi.c[2]);
                const cplx p = cmk(amp * cs, amp * sn);
                const cplx fb = FKbar[d + (int64_t)ndof * b];
                Wbar += re_conj_mul(fb, p);
                // Q = conj(fb) w: dJ = Re(Q dp), dp/dx = i k cb p, dp/dy = i k sb p, dp/dz = k p,
                // dp/dk = (z + i wb) p
                const cplx Q = cmk(wid * fb.x, -wid * fb.y);
                const cplx Qp = cmul(Q, p);
                cbar[0] += -k * cb * Qp.y;
                cbar[1] += -k * sb * Qp.y;
                cbar[2] += k * Qp.x;
                kbar_acc += gi.c[2] * Qp.x - wb * Qp.y;
            }
            // w = -(dof . n) a
            for (int c = 0; c < 3; c++) {
                gr.dofs[(int64_t)3 * n * d + i + (int64_t)n * c] += -Wbar * a * gi.nrm[c];
                nbar[c] += -Wbar * a * dv[c];
            }
            abar += -Wbar * s;
        }
        for (int c = 0; c < 3; c++) {
            gr.N[3 * i + c] += nbar[c];
            gr.C[3 * i + c] += cbar[c];
        }
        gr.A[i] += abar;
    }
    wbar_acc = wave_sum(wbar_acc);
    kbar_acc = wave_sum(kbar_acc);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(gr.wk, wbar_acc);
        atomicAdd(gr.wk + 1, kbar_acc);
    }
}

// radiation_bc / _airy_bc pulled back (one thread per panel i): B_ir = -i omega (n_i . dof_ir)
// and B_ib = -(n_i . v) with v = (g k / omega) e^{k z} e^{i k wb} (cos b, sin b, -i).
__global__ void __launch_bounds__(256) adjoint_bc_kernel(int n, const PanelGeom* __restrict__ geom,
                                                         const double* __restrict__ dofs, int ndof,
                                                         const double* __restrict__ betas, int nb, double w, double k,
                                                         double g, const cplx* __restrict__ Bbar, AdjointGrad gr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double wbar_acc = 0.0, kbar_acc = 0.0;
    if (i < n) {
        const PanelGeom& gi = geom[i];
        double nbar[3] = {0, 0, 0}, cbar[3] = {0, 0, 0};
        for (int r = 0; r < ndof; r++) {
            const double* dof = dofs + (int64_t)3 * n * r;
            const double dv[3] = {dof[i], dof[i + n], dof[i + 2 * (int64_t)n]};
            const double s = dv[0] * gi.nrm[0] + dv[1] * gi.nrm[1] + dv[2] * gi.nrm[2];
            const cplx bb = Bbar[i + (int64_t)n * r];
            const double sbar = -w * bb.y;  // Re(conj(bb) (-i omega))
            wbar_acc += -s * bb.y;          // Re(conj(bb) (-i s))
            for (int c = 0; c < 3; c++) {
                gr.dofs[(int64_t)3 * n * r + i + (int64_t)n * c] += sbar * gi.nrm[c];
                nbar[c] += sbar * dv[c];
            }
        }
        for (int b = 0; b < nb; b++) {
            const double cb = cos(betas[b]), sb = sin(betas[b]);
            const double wb = gi.c[0] * cb + gi.c[1] * sb;
            double sn, cs;
            sincos(k * wb, &sn, &cs);
            const double am = g * k / w * exp(k * gi.c[2]);
            const cplx E = cmk(am * cs, am * sn);
            const cplx v[3] = {cmk(cb * E.x, cb * E.y), cmk(sb * E.x, sb * E.y), cmk(E.y, -E.x)};
            const cplx B = cmk(-(gi.nrm[0] * v[0].x + gi.nrm[1] * v[1].x + gi.nrm[2] * v[2].x),
                               -(gi.nrm[0] * v[0].y + gi.nrm[1] * v[1].y + gi.nrm[2] * v[2].y));
            const cplx bb = Bbar[i + (int64_t)n * (ndof + b)];
            for (int c = 0; c < 3; c++) nbar[c] += -re_conj_mul(bb, v[c]);
            // dB/dx = i k cb B, dB/dy = i k sb B, dB/dz = k B, dB/dk = (1/k + z + i wb) B,
            // dB/domega = -B / omega
            const cplx iB = cmk(-B.y, B.x);
            cbar[0] += k * cb * re_conj_mul(bb, iB);
            cbar[1] += k * sb * re_conj_mul(bb, iB);
            cbar[2] += k * re_conj_mul(bb, B);
            kbar_acc += (1.0 / k + gi.c[2]) * re_conj_mul(bb, B) + wb * re_conj_mul(bb, iB);
            wbar_acc += -re_conj_mul(bb, B) / w;
        }
        for (int c = 0; c < 3; c++) {
            gr.N[3 * i + c] += nbar[c];
            gr.C[3 * i + c] += cbar[c];
        }
    }
    wbar_acc = wave_sum(wbar_acc);
    kbar_acc = wave_sum(kbar_acc);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(gr.wk, wbar_acc);
        atomicAdd(gr.wk + 1, kbar_acc);
    }
}

// Local inputs of one pair, in pass order: source panel j's vertices (12), centre (3),
// normal (3), area (1); field panel i's centre (3), normal (3); the wavenumber (1).
constexpr int kPairInputs = 26;
constexpr int ADJ_JT = 4;  // columns per workgroup

template <int N>
__device__ __forceinline__ DualN<N> seeded(double v, int q, int q0) {
    DualN<N> r;
    r.v = v;
#pragma unroll
    for (int p = 0; p < N; p++) r.d[p] = (q == q0 + p) ? 1.0 : 0.0;
    return r;
}

// Pull Sbar = U V^H, Dbar = U2 V2^H back through every pair.  Workgroup: 256 rows i of ADJ_JT
// columns j.  The row inputs (centre, normal of i) and k accumulate in registers; the column
// inputs are reduced over the wave and added atomically (vertices shared by several panels get
// the sum of their panels' contributions).
template <int N, bool EXACT>
__global__ void __launch_bounds__(256) adjoint_pair_kernel(int n, const double* __restrict__ verts,
                                                           const int64_t* __restrict__ faces,
                                                           const double* __restrict__ centers,
                                                           const double* __restrict__ normals,
                                                           const double* __restrict__ areas,
                                                           const double* __restrict__ radii, double kv, int direct,
                                                           const cplx* __restrict__ WG, const cplx* __restrict__ WR,
                                                           const cplx* __restrict__ U, const cplx* __restrict__ V,
                                                           const cplx* __restrict__ U2, const cplx* __restrict__ V2,
                                                           int R, AdjointGrad gr) {
    const int i_raw = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i_raw < n;
    const int i = live ? i_raw : n - 1;  // lanes past the end evaluate row n-1 with zero weight
    const int j0 = blockIdx.y * ADJ_JT;
    const int jn = min(ADJ_JT, n - j0);
    // cotangents of this row's pairs
    cplx Sb[ADJ_JT], Db[ADJ_JT];
    for (int jj = 0; jj < ADJ_JT; jj++) Sb[jj] = Db[jj] = cmk(0.0, 0.0);
    if (live) {
        for (int r = 0; r < R; r++) {
            const cplx u = U[i + (int64_t)n * r], u2 = U2[i + (int64_t)n * r];
            for (int jj = 0; jj < jn; jj++) {
                const cplx v = V[j0 + jj + (int64_t)n * r], v2 = V2[j0 + jj + (int64_t)n * r];
                // u conj(v)
                Sb[jj] = cmk(Sb[jj].x + u.x * v.x + u.y * v.y, Sb[jj].y + u.y * v.x - u.x * v.y);
                Db[jj] = cmk(Db[jj].x + u2.x * v2.x + u2.y * v2.y, Db[jj].y + u2.y * v2.x - u2.x * v2.y);
            }
        }
    }
    double gi_acc[6] = {0, 0, 0, 0, 0, 0}, gk = 0.0;
    for (int jj = 0; jj < jn; jj++) {
        const int j = j0 + jj;
        const int64_t idx = i + (int64_t)n * j;
        const cplx wg = EXACT ? WG[idx] : cmk(0.0, 0.0), wr = EXACT ? WR[idx] : cmk(0.0, 0.0);
        int64_t vid[4];
        for (int q = 0; q < 4; q++) vid[q] = faces[4 * j + q] - 1;
#pragma unroll 1
        for (int q0 = 0; q0 < kPairInputs; q0 += N) {
            DualN<N> Vj[12];
            for (int q = 0; q < 12; q++) Vj[q] = seeded<N>(verts[3 * vid[q / 3] + q % 3], q, q0);
            PanelGeomN<N> gj;
            birk_panel_geometry_d(Vj, gj);
            for (int c = 0; c < 3; c++) {
                gj.c[c] = seeded<N>(centers[3 * j + c], 12 + c, q0);
                gj.nrm[c] = seeded<N>(normals[3 * j + c], 15 + c, q0);
            }
            gj.area = seeded<N>(areas[j], 18, q0);
            gj.radius = radii[j];
            DualN<N> ci[3], ni[3];
            for (int c = 0; c < 3; c++) {
                ci[c] = seeded<N>(centers[3 * i + c], 19 + c, q0);
                ni[c] = seeded<N>(normals[3 * i + c], 22 + c, q0);
            }
            const DualN<N> k = seeded<N>(kv, 25, q0);
            DualN<N> out[4];
            pair_eval<N, EXACT>(ci, ni, gj, k, direct, wg, wr, out);
#pragma unroll
            for (int p = 0; p < N; p++) {
                const int q = q0 + p;
                if (q >= kPairInputs) break;
                double c = Sb[jj].x * out[0].d[p] + Sb[jj].y * out[1].d[p] + Db[jj].x * out[2].d[p] + Db[jj].y * out[3].d[p];
                if (!live) c = 0.0;
                if (q >= 19) {
                    if (q < 25) gi_acc[q - 19] += c;
                    else gk += c;
                } else {
                    c = wave_sum(c);
                    if ((threadIdx.x & 63) == 0) {
                        if (q < 12) atomicAdd(gr.V + 3 * vid[q / 3] + q % 3, c);
                        else if (q < 15) atomicAdd(gr.C + 3 * j + (q - 12), c);
                        else if (q < 18) atomicAdd(gr.N + 3 * j + (q - 15), c);
                        else atomicAdd(gr.A + j, c);
                    }
                }
            }
        }
    }
    if (live) {
        for (int c = 0; c < 3; c++) {
            atomicAdd(gr.C + 3 * i + c, gi_acc[c]);
            atomicAdd(gr.N + 3 * i + c, gi_acc[3 + c]);
        }
    }
    gk = wave_sum(gk);
    if ((threadIdx.x & 63) == 0) atomicAdd(gr.wk + 1, gk);
}

// Y = S^H X: one workgroup per column j of S (contiguous), R <= 16 right-hand sides
__global__ void __launch_bounds__(256) matmul_h_kernel(int n, int R, const cplx* __restrict__ S,
                                                       const cplx* __restrict__ X, cplx* __restrict__ Y) {
    const int j = blockIdx.x;
    const cplx* col = S + (int64_t)n * j;
    double acc[2 * 16];
    for (int r = 0; r < 2 * 16; r++) acc[r] = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const cplx s = col[i];
        for (int r = 0; r < R; r++) {
            const cplx x = X[i + (int64_t)n * r];
            acc[2 * r] += s.x * x.x + s.y * x.y;      // Re(conj(s) x)
            acc[2 * r + 1] += s.x * x.y - s.y * x.x;  // Im(conj(s) x)
        }
    }
    __shared__ double red[4][2 * 16];
    const int wv = threadIdx.x >> 6;
    for (int r = 0; r < 2 * R; r++) {
        const double v = wave_sum(acc[r]);
        if ((threadIdx.x & 63) == 0) red[wv][r] = v;
    }
    __syncthreads();
    if (threadIdx.x < R) {
        const int r = threadIdx.x;
        Y[j + (int64_t)n * r] = cmk(red[0][2 * r] + red[1][2 * r] + red[2][2 * r] + red[3][2 * r],
                                    red[0][2 * r + 1] + red[1][2 * r + 1] + red[2][2 * r + 1] + red[3][2 * r + 1]);
    }
}

__global__ void scale_kernel(int64_t cnt, double s, const cplx* __restrict__ A, cplx* __restrict__ C) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < cnt) C[t] = cmk(s * A[t].x, s * A[t].y);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
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

void launch_tangent_kx(hipStream_t st, int n, const PanelGeomD* geom, double k, double dk, bool exact, const cplx* WG,
                       const cplx* WR, cplx* dKx) {
    dim3 grid((n + 255) / 256, n);
    if (exact) tangent_kx_kernel<true><<<grid, 256, 0, st>>>(n, geom, k, dk, WG, WR, dKx);
    else tangent_kx_kernel<false><<<grid, 256, 0, st>>>(n, geom, k, dk, WG, WR, dKx);
}

void launch_tangent_bc_inc(hipStream_t st, int n, int nhull, const PanelGeomD* geom, double w, double dw, double k,
                           double dk, double beta, double beta_e, double g, cplx* dB) {
    tangent_bc_inc_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, nhull, geom, w, dw, k, dk, beta, beta_e, g, dB);
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

void launch_adjoint_forces(hipStream_t st, int n, const PanelGeom* geom, const double* dofs, int ndof, const cplx* phi,
                           int R, const cplx* Fbar, const cplx* FKbar, double rho, double w, double k, double g,
                           const double* betas, int nb, cplx* phibar, AdjointGrad gr) {
    adjoint_forces_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, geom, dofs, ndof, phi, R, Fbar, FKbar, rho, w, k, g,
                                                           betas, nb, phibar, gr);
}

void launch_adjoint_bc(hipStream_t st, int n, const PanelGeom* geom, const double* dofs, int ndof, const double* betas,
                       int nb, double w, double k, double g, const cplx* Bbar, AdjointGrad gr) {
    adjoint_bc_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, geom, dofs, ndof, betas, nb, w, k, g, Bbar, gr);
}

constexpr int kAdjN = 4;  // partials per pass of the pair pullback

void launch_adjoint_pairs(hipStream_t st, int n, const double* verts, const int64_t* faces, const double* centers,
                          const double* normals, const double* areas, const double* radii, double k, int direct,
                          bool exact, const cplx* WG, const cplx* WR, const cplx* U, const cplx* V, const cplx* U2,
                          const cplx* V2, int R, AdjointGrad gr) {
    dim3 grid((n + 255) / 256, (n + ADJ_JT - 1) / ADJ_JT);
    if (exact)
        adjoint_pair_kernel<kAdjN, true><<<grid, 256, 0, st>>>(n, verts, faces, centers, normals, areas, radii, k,
                                                               direct, WG, WR, U, V, U2, V2, R, gr);
    else
        adjoint_pair_kernel<kAdjN, false><<<grid, 256, 0, st>>>(n, verts, faces, centers, normals, areas, radii, k,
                                                                direct, WG, WR, U, V, U2, V2, R, gr);
}

void launch_matmul_h(hipStream_t st, int n, int R, const cplx* S, const cplx* X, cplx* Y) {
    matmul_h_kernel<<<n, 256, 0, st>>>(n, R, S, X, Y);
}

void launch_scale(hipStream_t st, int64_t cnt, double s, const cplx* A, cplx* C) {
    scale_kernel<<<(unsigned)((cnt + 255) / 256), 256, 0, st>>>(cnt, s, A, C);
}

}  // namespace mhx
