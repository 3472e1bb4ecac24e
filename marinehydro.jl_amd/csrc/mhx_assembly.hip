// Influence-matrix assembly, boundary conditions, right-hand-side products and
// pressure integration for gfx950.
//
// Reference: src/matrix_assembly.jl:274-369 (assemble_birk_rankine,
// assemble_wu_centers), :56-88 (broadcast assembly of the exact Delhommeau
// term), src/waves.jl:170-303 (BCs, Froude-Krylov, pressure integration),
// src/solve.jl:47-55.
//
// Layout: matrices are n x n column-major (Julia Matrix layout); a workgroup
// owns 256 consecutive rows (field panels, one per thread) and TJ columns
// (source panels) whose geometry is staged in LDS, so every store of a column
// is one coalesced 2 KB (real) or 4 KB (complex) segment.
#include "mhx_internal.h"

#include <algorithm>

namespace mhx {

#define TJ 8
static const double kScale = -1.0 / (4 * MHX_PI);

__global__ void panel_geometry_kernel(int n, const double* __restrict__ verts, const int64_t* __restrict__ faces,
                                      const double* __restrict__ centers, const double* __restrict__ normals,
                                      const double* __restrict__ areas, const double* __restrict__ radii,
                                      PanelGeom* __restrict__ geom) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double V[12];
    for (int k = 0; k < 4; k++) {
        int64_t v = faces[4 * j + k] - 1;
        for (int c = 0; c < 3; c++) V[3 * k + c] = verts[3 * v + c];
    }
    PanelGeom g;
    birk_panel_geometry(V, g);
    for (int c = 0; c < 3; c++) {
        g.c[c] = centers[3 * j + c];
        g.nrm[c] = normals[3 * j + c];
    }
    g.radius = radii[j];
    g.area = areas[j];
    g.pad[0] = 0.5;  // identity jump
    g.pad[1] = 1.0;  // hull panel
    g.pad[2] = g.pad[3] = 0.0;
    geom[j] = g;
}

// Lid panels (index >= nhull) take no incident-wave condition; with a lid, a panel in the free
// surface (|z| < 1e-8) has the jump 1 instead of 1/2 (solve.jl:4-11, 22-28; matrix_assembly.jl:25-27).
__global__ void mark_lid_kernel(int n, int nhull, PanelGeom* __restrict__ geom) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    geom[i].pad[1] = i < nhull ? 1.0 : 0.0;
    geom[i].pad[0] = fabs(geom[i].c[2]) < 1e-8 ? 1.0 : 0.5;
}

void launch_mark_lid(hipStream_t st, int n, int nhull, PanelGeom* geom) {
    mark_lid_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, nhull, geom);
}

void launch_panel_geometry(hipStream_t st, int n, const double* verts, const int64_t* faces, const double* centers,
                           const double* normals, const double* areas, const double* radii, PanelGeom* geom) {
    panel_geometry_kernel<<<(n + 127) / 128, 128, 0, st>>>(n, verts, faces, centers, normals, areas, radii, geom);
}

__device__ __forceinline__ void stage_columns(const PanelGeom* __restrict__ geom, int n, int j0, PanelGeom* sg) {
    const double* src = reinterpret_cast<const double*>(geom + j0);
    double* dst = reinterpret_cast<double*>(sg);
    const int cnt = min(TJ, n - j0) * 32;
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) dst[t] = src[t];
    __syncthreads();
}

// assemble_birk_rankine inner loop (matrix_assembly.jl:301-324) for the tuple
// (Rankine(), RankineReflected()), without the identity jump.
__global__ void __launch_bounds__(256) rankine_kernel(int n, const PanelGeom* __restrict__ geom, int direct,
                                                      double* __restrict__ SR, double* __restrict__ DR) {
    __shared__ PanelGeom sg[TJ];
    const int j0 = blockIdx.y * TJ;
    stage_columns(geom, n, j0, sg);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PanelGeom& gi = geom[i];
    const double c1[3] = {gi.c[0], gi.c[1], gi.c[2]};
    const double c1r[3] = {gi.c[0], gi.c[1], -gi.c[2]};
    const double ni[3] = {gi.nrm[0], gi.nrm[1], gi.nrm[2]};
    const int jn = min(TJ, n - j0);
#pragma unroll 1
    for (int jj = 0; jj < jn; jj++) {
        const PanelGeom& g = sg[jj];
        const double* nv = direct ? g.nrm : ni;
        double phi, gs[3], phir, gr[3];
        vrankine(c1, g, phi, gs);
        vrankine(c1r, g, phir, gr);
        double s = 0.0, d = 0.0;
        s += phi;
        if (direct) {
            d += nv[0] * gs[0] + nv[1] * gs[1] + nv[2] * gs[2];
        } else {
            d += nv[0] * (-gs[0]) + nv[1] * (-gs[1]) + nv[2] * (-gs[2]);
            gr[0] = -gr[0];
            gr[1] = -gr[1];
        }
        s += 1.0 * phir;
        d += 1.0 * (nv[0] * gr[0] + nv[1] * gr[1] + nv[2] * gr[2]);
        const int64_t idx = i + (int64_t)n * (j0 + jj);
        SR[idx] = kScale * s;
        DR[idx] = kScale * d;
    }
}

void launch_rankine(hipStream_t st, int n, const PanelGeom* geom, int direct, double* SR, double* DR) {
    dim3 grid((n + 255) / 256, (n + TJ - 1) / TJ);
    rankine_kernel<<<grid, 256, 0, st>>>(n, geom, direct, SR, DR);
}

// ---------------------------------------------------------------------------
// forward speed: the indirect x-derivative operator K_x (assemble_matrices with all_normals =
// [1, 0, 0], matrix_assembly.jl:183-245; used by solve.jl:35-45 for dphi/dx = K_x sigma)
// ---------------------------------------------------------------------------
// Rankine part of K_x (wavenumber independent, cached per body): the indirect gradient with
// respect to the field point dotted with e_x, panel and reflected image.
__global__ void __launch_bounds__(256) rankine_x_kernel(int n, const PanelGeom* __restrict__ geom,
                                                        double* __restrict__ KRx) {
    __shared__ PanelGeom sg[TJ];
    const int j0 = blockIdx.y * TJ;
    stage_columns(geom, n, j0, sg);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PanelGeom& gi = geom[i];
    const double c1[3] = {gi.c[0], gi.c[1], gi.c[2]};
    const double c1r[3] = {gi.c[0], gi.c[1], -gi.c[2]};
    const int jn = min(TJ, n - j0);
#pragma unroll 1
    for (int jj = 0; jj < jn; jj++) {
        const PanelGeom& g = sg[jj];
        double phi, gs[3], phir, gr[3];
        vrankine(c1, g, phi, gs);
        vrankine(c1r, g, phir, gr);
        // indirect: e_x . (-grad) + e_x . reflect(-grad of the image) = -gs_x - gr_x
        const double d = -gs[0] - gr[0];
        KRx[i + (int64_t)n * (j0 + jj)] = kScale * d;
    }
}

void launch_rankine_x(hipStream_t st, int n, const PanelGeom* geom, double* KRx) {
    dim3 grid((n + 255) / 256, (n + TJ - 1) / TJ);
    rankine_x_kernel<<<grid, 256, 0, st>>>(n, geom, KRx);
}

// K_x = KRx + wave part with the normal e_x (only the horizontal x component of the reduced
// gradient survives) + the identity jump times (e_x . n_i) on the diagonal.  EXACT: dG/dr from
// the (G, dG/dr) table of the same wavenumber, else the Wu fit.
template <bool EXACT>
__global__ void __launch_bounds__(256) kx_combine_kernel(int n, const PanelGeom* __restrict__ geom,
                                                         const double* __restrict__ KRx, double k,
                                                         const cplx* __restrict__ WR, cplx* __restrict__ Kx) {
    __shared__ PanelGeom sg[TJ];
    const int j0 = blockIdx.y * TJ;
    stage_columns(geom, n, j0, sg);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PanelGeom& gi = geom[i];
    const int jn = min(TJ, n - j0);
#pragma unroll 1
    for (int jj = 0; jj < jn; jj++) {
        const PanelGeom& g = sg[jj];
        const int j = j0 + jj;
        const int64_t idx = i + (int64_t)n * j;
        double hh, vv;
        reduced_hv(gi.c[0], gi.c[1], gi.c[2], g.c[0], g.c[1], g.c[2], k, hh, vv);
        cplx Gh;
        if (EXACT) {
            Gh = WR[idx];
        } else {
            cplx G, Gv;
            wu_term(hh, vv, G, Gh, Gv);
        }
        const double dx1 = gi.c[0] - g.c[0];  // indirect: Delta = x - xi
        double dr1 = 0.0;
        if (fabs(hh) > 1e-6) dr1 = k * k / hh * dx1;
        const cplx dw = cscale(g.area, cscale(k, cscale(dr1, Gh)));
        double dr = KRx[idx];
        if (i == j) dr += gi.pad[0] * gi.nrm[0];
        Kx[idx] = cmk(dr + kScale * dw.x, kScale * dw.y);
    }
}

void launch_kx_combine(hipStream_t st, int n, const PanelGeom* geom, const double* KRx, double k, bool exact,
                       const cplx* WR, cplx* Kx) {
    dim3 grid((n + 255) / 256, (n + TJ - 1) / TJ);
    if (exact) kx_combine_kernel<true><<<grid, 256, 0, st>>>(n, geom, KRx, k, WR, Kx);
    else kx_combine_kernel<false><<<grid, 256, 0, st>>>(n, geom, KRx, k, WR, Kx);
}

// forward-speed right-hand sides of one (U, omega, beta) group: radiation (radiation_bc,
// waves.jl:183-199) -i omega_e (n . dof) + U (n . dgrad), dgrad the constant gradient of motion of
// a rigid dof (bodies.jl:55-66); diffraction (_airy_bc, waves.jl:266-282) with the incident k and
// omega, the phase along the encountered heading and the velocity along the incident one.
__global__ void bc_fs_kernel(int n, const PanelGeom* __restrict__ geom, const double* __restrict__ dofs, int ndof,
                             const double* __restrict__ dgrad, double we, double U, double kinc, double winc,
                             double beta, double beta_e, double g, cplx* __restrict__ B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PanelGeom& gi = geom[i];
    const double* nv = gi.nrm;
    for (int r = 0; r < ndof; r++) {
        const double* dof = dofs + (int64_t)3 * n * r;
        const double disp = nv[0] * dof[i] + nv[1] * dof[i + n] + nv[2] * dof[i + 2 * n];
        const double* gr = dgrad + 3 * r;
        const double gd = nv[0] * gr[0] + nv[1] * gr[1] + nv[2] * gr[2];
        B[i + (int64_t)n * r] = cmk(U * gd * gi.pad[1], -we * disp);  // -i omega_e disp + U gd, hull panels only
    }
    const double wbar = gi.c[0] * cos(beta_e) + gi.c[1] * sin(beta_e);
    const double cih = exp(kinc * gi.c[2]);
    double s, c;
    sincos(kinc * wbar, &s, &c);
    const double am = g * kinc / winc;
    const cplx amp = cmk(am * c, am * s);
    const cplx v0 = cscale(cih, cscale(cos(beta), amp)), v1 = cscale(cih, cscale(sin(beta), amp));
    const cplx v2 = cscale(cih, cmk(amp.y, -amp.x));  // -i * amp
    const cplx dot = cadd(cadd(cscale(nv[0], v0), cscale(nv[1], v1)), cscale(nv[2], v2));
    B[i + (int64_t)n * ndof] = cmk(-dot.x * gi.pad[1], -dot.y * gi.pad[1]);
}

void launch_bc_fs(hipStream_t st, int n, const PanelGeom* geom, const double* dofs, int ndof, const double* dgrad,
                  double we, double U, double kinc, double winc, double beta, double beta_e, double g, cplx* B) {
    bc_fs_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, geom, dofs, ndof, dgrad, we, U, kinc, winc, beta, beta_e, g, B);
}

// Y += c X over cnt complex elements
__global__ void caxpy_kernel(int64_t cnt, double cr, double ci, const cplx* __restrict__ X, cplx* __restrict__ Y) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < cnt) Y[t] = cmk(Y[t].x + cr * X[t].x - ci * X[t].y, Y[t].y + cr * X[t].y + ci * X[t].x);
}

void launch_caxpy(hipStream_t st, int64_t cnt, double cr, double ci, const cplx* X, cplx* Y) {
    caxpy_kernel<<<(unsigned)((cnt + 255) / 256), 256, 0, st>>>(cnt, cr, ci, X, Y);
}

// _wu_sd (wu.jl:112-127) + the scale of _assemble_wu_centers_loop (matrix_assembly.jl:343-369)
__global__ void __launch_bounds__(256) combine_wu_kernel(int n, const PanelGeom* __restrict__ geom,
                                                         const double* __restrict__ SR, const double* __restrict__ DR,
                                                         const double* __restrict__ ks, int direct,
                                                         int include_identity, cplx* __restrict__ Sb,
                                                         cplx* __restrict__ Db, int64_t sMat) {
    const double k = ks[blockIdx.z];
    cplx* S = Sb + sMat * blockIdx.z;
    cplx* D = Db + sMat * blockIdx.z;
    __shared__ PanelGeom sg[TJ];
    const int j0 = blockIdx.y * TJ;
    stage_columns(geom, n, j0, sg);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PanelGeom& gi = geom[i];
    const double x0 = gi.c[0], x1 = gi.c[1], x2 = gi.c[2];
    const double ni0 = gi.nrm[0], ni1 = gi.nrm[1], ni2 = gi.nrm[2];
    const int jn = min(TJ, n - j0);
#pragma unroll 1
    for (int jj = 0; jj < jn; jj++) {
        const PanelGeom& g = sg[jj];
        const int j = j0 + jj;
        double hh, vv;
        reduced_hv(x0, x1, x2, g.c[0], g.c[1], g.c[2], k, hh, vv);
        cplx G, Gh, Gv;
        wu_term(hh, vv, G, Gh, Gv);
        double dx1 = direct ? g.c[0] - x0 : x0 - g.c[0];
        double dx2 = direct ? g.c[1] - x1 : x1 - g.c[1];
        double dr1 = 0.0, dr2 = 0.0;
        if (fabs(hh) > 1e-6) {
            dr1 = k * k / hh * dx1;
            dr2 = k * k / hh * dx2;
        }
        const double n0 = direct ? g.nrm[0] : ni0, n1 = direct ? g.nrm[1] : ni1, n2 = direct ? g.nrm[2] : ni2;
        cplx g0 = cscale(k, cscale(dr1, Gh)), g1 = cscale(k, cscale(dr2, Gh)), g2 = cscale(k, cscale(k, Gv));
        cplx s = cscale(g.area, cscale(k, G));
        cplx dsum = cadd(cadd(cscale(n0, g0), cscale(n1, g1)), cscale(n2, g2));
        cplx d = cscale(g.area, dsum);
        const int64_t idx = i + (int64_t)n * j;
        double sr = SR[idx], dr = DR[idx];
        if (include_identity && i == j) dr += gi.pad[0];
        S[idx] = cmk(sr + kScale * s.x, kScale * s.y);
        D[idx] = cmk(dr + kScale * d.x, kScale * d.y);
    }
}

void launch_combine_wu(hipStream_t st, int n, const PanelGeom* geom, const double* SR, const double* DR,
                       const double* ks, int batch, int direct, int include_identity, cplx* S, cplx* D,
                       int64_t sMat) {
    dim3 grid((n + 255) / 256, (n + TJ - 1) / TJ, batch);
    combine_wu_kernel<<<grid, 256, 0, st>>>(n, geom, SR, DR, ks, direct, include_identity, S, D, sMat);
}

// Exact Guevel-Delhommeau dimensionless (G, dG/dr) per unordered panel pair.
// The reduced coordinates are bitwise symmetric in (i, j), so one evaluation
// serves both matrix entries.
static int device_cus();

// Exact wave-term table, pass 1: every unordered pair gets the first GK(7,15) interval of both
// quadratures (uniform work per lane); pairs whose G and dG/dr already meet the termination rule
// are written, the others appended to a list (item << 40 | i << 20 | j) for pass 2.  Splitting
// the adaptive refinement off keeps the lanes of a wave from waiting on their slowest pair.
__device__ __forceinline__ void table_store(cplx* WG, cplx* WR, int n, int i, int j, cplx G, cplx Gr) {
    G = cscale(4.0, G);
    Gr = cscale(4.0, Gr);
    WG[i + (int64_t)n * j] = G;
    WR[i + (int64_t)n * j] = Gr;
    if (i != j) {
        WG[j + (int64_t)n * i] = G;
        WR[j + (int64_t)n * i] = Gr;
    }
}

// one atomic per wave for the work tally; tid: the thread's linear index in its block
__device__ __forceinline__ void tally_work(const QuadWork& wk, int tid, unsigned long long* status) {
    unsigned long long e = wk.evals, se = wk.series, cf = wk.cf;
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_xor(e, off);
        se += __shfl_xor(se, off);
        cf += __shfl_xor(cf, off);
    }
    if ((tid & 63) == 0 && e) {
        atomicAdd(status + 1, e);
        atomicAdd(status + 2, se);
        atomicAdd(status + 3, cf);
    }
}

__global__ void __launch_bounds__(256) delhommeau_table_kernel(int n, const PanelGeom* __restrict__ geom,
                                                               const double* __restrict__ ks, double rtol, double atol,
                                                               cplx* __restrict__ WGb, cplx* __restrict__ WRb,
                                                               int64_t sMat, unsigned long long* __restrict__ rlist,
                                                               unsigned long long* __restrict__ rcount, int64_t rcap,
                                                               unsigned long long* __restrict__ status) {
    if (blockIdx.x > blockIdx.y) return;
    const int item = blockIdx.z;
    const double k = ks[item];
    const int tid = threadIdx.y * 16 + threadIdx.x;
    __shared__ double cosn[15];
    if (tid < 15) cosn[tid] = first_node_cos(tid);
    __syncthreads();
    const int i = blockIdx.x * 16 + threadIdx.x;
    const int j = blockIdx.y * 16 + threadIdx.y;
    QuadWork wk;
    bool refine = false;
    if (i < n && j < n && i <= j) {
        const PanelGeom& gi = geom[i];
        const PanelGeom& gj = geom[j];
        double hh, vv;
        reduced_hv(gi.c[0], gi.c[1], gi.c[2], gj.c[0], gj.c[1], gj.c[2], k, hh, vv);
        cplx G, Gr;
        double EG, ER;
        cplx* WG = WGb + sMat * item;
        cplx* WR = WRb + sMat * item;
        if (hquad_first(cosn, hh, vv, rtol, atol, G, Gr, EG, ER, wk)) {
            table_store(WG, WR, n, i, j, G, Gr);
        } else {
            refine = true;
            // the first interval's estimates for pass 2, in the pair's own table slots (which pass
            // 2 overwrites): G and dG/dr at (i, j), both errors at (j, i); a diagonal pair has
            // one slot and is recomputed
            if (i != j) {
                WG[i + (int64_t)n * j] = G;
                WR[i + (int64_t)n * j] = Gr;
                WG[j + (int64_t)n * i] = cmk(EG, ER);
            }
        }
    }
    // wave-aggregated append to the refinement list
    const unsigned long long mask = __ballot(refine);
    if (mask) {
        const int lane = tid & 63;
        unsigned long long base = 0;
        const int leader = __ffsll((long long)mask) - 1;
        if (lane == leader) base = atomicAdd(rcount, (unsigned long long)__popcll(mask));
        base = __shfl(base, leader);
        if (refine) {
            const unsigned long long slot = base + __popcll(mask & ((1ull << lane) - 1ull));
            if (slot < (unsigned long long)rcap)
                rlist[slot] = ((unsigned long long)item << 40) | ((unsigned long long)i << 20) | (unsigned)j;
        }
    }
    tally_work(wk, tid, status);
}

// pass 2: the listed pairs, each the adaptive refinement on one 32-lane group from pass 1's first
// interval (two pairs per wave; grid-stride: the count is only known on the device).  The nodes of an interval are
// spread over the group's lanes, so its lanes work in step; only the two groups of a wave can
// take different numbers of bisections.
__global__ void __launch_bounds__(256) delhommeau_refine_kernel(int n, const PanelGeom* __restrict__ geom,
                                                                const double* __restrict__ ks, double rtol, double atol,
                                                                cplx* __restrict__ WGb, cplx* __restrict__ WRb,
                                                                int64_t sMat, const unsigned long long* __restrict__ rlist,
                                                                const unsigned long long* __restrict__ rcount,
                                                                int64_t rcap, unsigned long long* __restrict__ status) {
    __shared__ double cosn[15];
    __shared__ QuadBoxes boxes[8];
    if (threadIdx.x < 15) cosn[threadIdx.x] = first_node_cos(threadIdx.x);
    __syncthreads();
    const int64_t cnt = (int64_t)min(*rcount, (unsigned long long)rcap);
    const int gl = threadIdx.x & 31, grp = threadIdx.x >> 5;
    QuadBoxes& B = boxes[grp];
    QuadWork wk;
    for (int64_t t = blockIdx.x * 8LL + grp; t < cnt; t += (int64_t)gridDim.x * 8) {
        const unsigned long long e = rlist[t];
        const int item = (int)(e >> 40), i = (int)((e >> 20) & 0xFFFFF), j = (int)(e & 0xFFFFF);
        const PanelGeom& gi = geom[i];
        const PanelGeom& gj = geom[j];
        double hh, vv;
        reduced_hv(gi.c[0], gi.c[1], gi.c[2], gj.c[0], gj.c[1], gj.c[2], ks[item], hh, vv);
        cplx* WG = WGb + sMat * item;
        cplx* WR = WRb + sMat * item;
        cplx IG, IR;
        double EG, ER;
        if (i != j) {
            IG = WG[i + (int64_t)n * j];
            IR = WR[i + (int64_t)n * j];
            const cplx E = WG[j + (int64_t)n * i];
            EG = E.x;
            ER = E.y;
        } else {
            first_interval_group(gl, cosn, hh, vv, IG, EG, IR, ER, wk);
        }
        cplx G, Gr;
        const bool ok = refine_pair_group(B, gl, hh, vv, rtol, atol, IG, EG, IR, ER, G, Gr, wk);
        if (gl == 0) {
            if (!ok) atomicOr(status, 1ull);
            table_store(WG, WR, n, i, j, G, Gr);
        }
    }
    tally_work(wk, threadIdx.x, status);
}

void launch_delhommeau_table(hipStream_t st, int n, const PanelGeom* geom, const double* ks, int batch, double rtol,
                             double atol, cplx* WG, cplx* WR, int64_t sMat, unsigned long long* rlist,
                             unsigned long long* rcount, int64_t rcap, unsigned long long* status) {
    launch_zero(st, rcount, sizeof(unsigned long long));
    dim3 grid((n + 15) / 16, (n + 15) / 16, batch);
    delhommeau_table_kernel<<<grid, dim3(16, 16), 0, st>>>(n, geom, ks, rtol, atol, WG, WR, sMat, rlist, rcount, rcap,
                                                           status);
    delhommeau_refine_kernel<<<device_cus() * 8, 256, 0, st>>>(n, geom, ks, rtol, atol, WG, WR, sMat, rlist, rcount,
                                                               rcap, status);
}

// with_reduced_coordinates_derivative (abstract_greens_function.jl:140-185) and
// the one-point integrals (exact_Guevel_Delhommeau.jl:61-75) from a (G, dG/dr) table
__global__ void __launch_bounds__(256) combine_table_kernel(int n, const PanelGeom* __restrict__ geom,
                                                            const double* __restrict__ SR, const double* __restrict__ DR,
                                                            const double* __restrict__ ks, int direct,
                                                            int include_identity, const cplx* __restrict__ WGb,
                                                            const cplx* __restrict__ WRb, cplx* __restrict__ Sb,
                                                            cplx* __restrict__ Db, int64_t sMat) {
    const double k = ks[blockIdx.z];
    const cplx* WG = WGb + sMat * blockIdx.z;
    const cplx* WR = WRb + sMat * blockIdx.z;
    cplx* S = Sb + sMat * blockIdx.z;
    cplx* D = Db + sMat * blockIdx.z;
    __shared__ PanelGeom sg[TJ];
    const int j0 = blockIdx.y * TJ;
    stage_columns(geom, n, j0, sg);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PanelGeom& gi = geom[i];
    const double x0 = gi.c[0], x1 = gi.c[1], x2 = gi.c[2];
    const double ni0 = gi.nrm[0], ni1 = gi.nrm[1], ni2 = gi.nrm[2];
    const int jn = min(TJ, n - j0);
#pragma unroll 1
    for (int jj = 0; jj < jn; jj++) {
        const PanelGeom& g = sg[jj];
        const int j = j0 + jj;
        const int64_t idx = i + (int64_t)n * j;
        double hh, vv;
        reduced_hv(x0, x1, x2, g.c[0], g.c[1], g.c[2], k, hh, vv);
        const double xr[3] = {x0, x1, x2}, nr[3] = {ni0, ni1, ni2};
        cplx Sv, Dv;
        combine_entry(k, hh, vv, xr, nr, g, direct, include_identity && i == j, WG[idx], WR[idx], SR[idx], DR[idx], Sv,
                      Dv);
        S[idx] = Sv;
        D[idx] = Dv;
    }
}

void launch_combine_table(hipStream_t st, int n, const PanelGeom* geom, const double* SR, const double* DR,
                          const double* ks, int batch, int direct, int include_identity, const cplx* WG,
                          const cplx* WR, cplx* S, cplx* D, int64_t sMat) {
    dim3 grid((n + 255) / 256, (n + TJ - 1) / TJ, batch);
    combine_table_kernel<<<grid, 256, 0, st>>>(n, geom, SR, DR, ks, direct, include_identity, WG, WR, S, D, sMat);
}

// ---------------------------------------------------------------------------
// multi-body sweep: every (body, frequency) item of a sweep in one launch per stage, over a
// queue of equal-cost units (combine: 256 rows x TJ columns; table: 16 x 16 pair blocks of the
// upper triangle).  Bodies of different panel counts share the launch.
// ---------------------------------------------------------------------------
int64_t sweep_combine_units(int n) { return (int64_t)((n + 255) / 256) * ((n + TJ - 1) / TJ); }
int64_t sweep_table_units(int n) {
    const int64_t nb = (n + 15) / 16;
    return nb * (nb + 1) / 2;
}

// the item holding unit u (items sorted by their first unit)
__device__ __forceinline__ int find_item(const SweepItem* __restrict__ items, int nitems, int64_t u, bool table) {
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int64_t start = table ? items[mid].tunit0 : items[mid].unit0;
        if (start <= u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(256) queue_table_kernel(const SweepItem* __restrict__ items, int nitems,
                                                          int64_t total_units, unsigned long long* __restrict__ counter,
                                                          double rtol, double atol,
                                                          unsigned long long* __restrict__ status) {
    __shared__ unsigned long long s_unit;
    QuadWork wk;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    for (;;) {
        if (threadIdx.x == 0) s_unit = atomicAdd(counter, 1ull);
        __syncthreads();
        const unsigned long long u = s_unit;
        __syncthreads();
        if (u >= (unsigned long long)total_units) break;
        const int it = find_item(items, nitems, (int64_t)u, true);
        const SweepItem& item = items[it];
        // local unit -> block (bi, bj), bi <= bj, column-major over the upper triangle
        const int64_t local = (int64_t)u - item.tunit0;
        int64_t bj = (int64_t)((sqrt(8.0 * (double)local + 1.0) - 1.0) * 0.5);
        while (bj * (bj + 1) / 2 > local) bj--;
        while ((bj + 1) * (bj + 2) / 2 <= local) bj++;
        const int64_t bi = local - bj * (bj + 1) / 2;
        const int n = item.n;
        const int i = (int)bi * 16 + tx, j = (int)bj * 16 + ty;
        if (i < n && j < n && i <= j) {
            const PanelGeom& gi = item.geom[i];
            const PanelGeom& gj = item.geom[j];
            const double k = *item.k;
            double hh, vv;
            reduced_hv(gi.c[0], gi.c[1], gi.c[2], gj.c[0], gj.c[1], gj.c[2], k, hh, vv);
            cplx Gr, G;
            if (!hquad_pair(hh, vv, rtol, atol, G, Gr, wk)) atomicOr(status, 1ull);
            G = cscale(4.0, G);
            Gr = cscale(4.0, Gr);
            item.WG[i + (int64_t)n * j] = G;
            item.WR[i + (int64_t)n * j] = Gr;
            if (i != j) {
                item.WG[j + (int64_t)n * i] = G;
                item.WR[j + (int64_t)n * i] = Gr;
            }
        }
    }
    unsigned long long e = wk.evals, se = wk.series, cf = wk.cf;
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_xor(e, off);
        se += __shfl_xor(se, off);
        cf += __shfl_xor(cf, off);
    }
    if ((threadIdx.x & 63) == 0 && e) {
        atomicAdd(status + 1, e);
        atomicAdd(status + 2, se);
        atomicAdd(status + 3, cf);
    }
}

// RG = threads per workgroup (blockDim.x), each unit's 256 rows in 256 / RG passes; chosen per
// device and workload by the sweep's tuner
template <bool EXACT, int RG>
__global__ void __launch_bounds__(RG) queue_combine_kernel(const SweepItem* __restrict__ items, int nitems,
                                                           int64_t total_units, int direct, int include_identity) {
    __shared__ PanelGeom sg[TJ];
#pragma unroll 1
    for (int64_t u = blockIdx.x; u < total_units; u += gridDim.x) {
        const int it = find_item(items, nitems, u, false);
        const SweepItem& item = items[it];
        const int n = item.n;
        const int nrb = (n + 255) / 256;  // units are 256 rows x TJ columns whatever RG is
        const int64_t local = u - item.unit0;
        const int rb = (int)(local % nrb), cb = (int)(local / nrb);
        const int j0 = cb * TJ;
        __syncthreads();  // the previous unit's columns are no longer read
        {
            const double* src = reinterpret_cast<const double*>(item.geom + j0);
            double* dst = reinterpret_cast<double*>(sg);
            const int cnt = min(TJ, n - j0) * 32;
            for (int t = threadIdx.x; t < cnt; t += RG) dst[t] = src[t];
        }
        __syncthreads();
        const double k = *item.k;
        const int jn = min(TJ, n - j0);
#pragma unroll 1
        for (int r0 = 0; r0 < 256; r0 += RG) {
            const int i = rb * 256 + r0 + threadIdx.x;
            if (i >= n) break;
            const PanelGeom& gi = item.geom[i];
            const double xr[3] = {gi.c[0], gi.c[1], gi.c[2]}, nr[3] = {gi.nrm[0], gi.nrm[1], gi.nrm[2]};
    #pragma unroll 1
            for (int jj = 0; jj < jn; jj++) {
                const PanelGeom& g = sg[jj];
                const int j = j0 + jj;
                const int64_t idx = i + (int64_t)n * j;
                double hh, vv;
                reduced_hv(xr[0], xr[1], xr[2], g.c[0], g.c[1], g.c[2], k, hh, vv);
                cplx Sv, Dv;
                if (EXACT) {
                    combine_entry(k, hh, vv, xr, nr, g, direct, include_identity && i == j, item.WG[idx], item.WR[idx],
                                  item.SR[idx], item.DR[idx], Sv, Dv);
                } else {
                    // as combine_wu_kernel: Wu's own dG/dz
                    cplx G, Gh, Gv;
                    wu_term(hh, vv, G, Gh, Gv);
                    const double dx1 = direct ? g.c[0] - xr[0] : xr[0] - g.c[0];
                    const double dx2 = direct ? g.c[1] - xr[1] : xr[1] - g.c[1];
                    double dr1 = 0.0, dr2 = 0.0;
                    if (fabs(hh) > 1e-6) {
                        dr1 = k * k / hh * dx1;
                        dr2 = k * k / hh * dx2;
                    }
                    const double n0 = direct ? g.nrm[0] : nr[0], n1 = direct ? g.nrm[1] : nr[1], n2 = direct ? g.nrm[2] : nr[2];
                    const cplx g0 = cscale(k, cscale(dr1, Gh)), g1 = cscale(k, cscale(dr2, Gh)), g2 = cscale(k, cscale(k, Gv));
                    const cplx sw = cscale(g.area, cscale(k, G));
                    const cplx dw = cscale(g.area, cadd(cadd(cscale(n0, g0), cscale(n1, g1)), cscale(n2, g2)));
                    double dr = item.DR[idx];
                    if (include_identity && i == j) dr += g.pad[0];
                    Sv = cmk(item.SR[idx] + kScale * sw.x, kScale * sw.y);
                    Dv = cmk(dr + kScale * dw.x, kScale * dw.y);
                }
                item.S[idx] = Sv;
                item.D[idx] = Dv;
            }
        }
    }
}

static int device_cus() {
    int dev = 0, cus = 0;
    hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    return cus;
}

void launch_queue_table(hipStream_t st, const SweepItem* items, int nitems, int64_t total_units,
                        unsigned long long* counter, double rtol, double atol, unsigned long long* status) {
    if (total_units <= 0) return;
    const int64_t grid = std::min<int64_t>(total_units, (int64_t)device_cus() * 8);
    queue_table_kernel<<<(unsigned)grid, 256, 0, st>>>(items, nitems, total_units, counter, rtol, atol, status);
}

void launch_queue_combine(hipStream_t st, const SweepItem* items, int nitems, int64_t total_units, bool exact,
                          int direct, int include_identity, int rows_per_group) {
    if (total_units <= 0) return;
    const int64_t grid = std::min<int64_t>(total_units, (int64_t)device_cus() * 16);
    if (rows_per_group == 128) {
        if (exact)
            queue_combine_kernel<true, 128><<<(unsigned)grid, 128, 0, st>>>(items, nitems, total_units, direct, include_identity);
        else
            queue_combine_kernel<false, 128><<<(unsigned)grid, 128, 0, st>>>(items, nitems, total_units, direct, include_identity);
    } else {
        if (exact)
            queue_combine_kernel<true, 256><<<(unsigned)grid, 256, 0, st>>>(items, nitems, total_units, direct, include_identity);
        else
            queue_combine_kernel<false, 256><<<(unsigned)grid, 256, 0, st>>>(items, nitems, total_units, direct, include_identity);
    }
}

__global__ void zero_words_kernel(unsigned* __restrict__ p, int64_t nwords) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < nwords) p[t] = 0u;
}

void launch_zero(hipStream_t st, void* p, size_t nbytes) {
    const int64_t nw = (int64_t)(nbytes / 4);
    if (nw <= 0) return;
    zero_words_kernel<<<(unsigned)((nw + 255) / 256), 256, 0, st>>>(reinterpret_cast<unsigned*>(p), nw);
}

__global__ void sticky_info_kernel(const int* __restrict__ info, int batch, int* __restrict__ sticky) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (sticky[0]) return;
    for (int q = 0; q < batch; q++)
        if (info[q]) {
            sticky[0] = info[q];
            return;
        }
}

void launch_sticky_info(hipStream_t st, const int* info, int batch, int* sticky) {
    sticky_info_kernel<<<1, 64, 0, st>>>(info, batch, sticky);
}

// ---------------------------------------------------------------------------
// boundary conditions: radiation_bc (waves.jl:183-211), _airy_bc (waves.jl:266-282)
// ---------------------------------------------------------------------------
__global__ void bc_kernel(int n, const PanelGeom* __restrict__ geom, const double* __restrict__ dofs, int ndof,
                          const double* __restrict__ betas, int nb, const double* __restrict__ ws,
                          const double* __restrict__ ks, double g, cplx* __restrict__ Bb, int64_t sB) {
    const double omega = ws[blockIdx.z], k = ks[blockIdx.z];
    cplx* B = Bb + sB * blockIdx.z;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PanelGeom& gi = geom[i];
    const double* nv = gi.nrm;
    for (int r = 0; r < ndof; r++) {
        const double* dof = dofs + (int64_t)3 * n * r;
        double disp = nv[0] * dof[i] + nv[1] * dof[i + n] + nv[2] * dof[i + 2 * n];
        B[i + (int64_t)n * r] = cmk(-0.0 * omega * disp, -omega * disp);
    }
    for (int b = 0; b < nb; b++) {
        const double beta = betas[b];
        const double cb = cos(beta), sb = sin(beta);
        const double wbar = gi.c[0] * cb + gi.c[1] * sb;
        const double cih = exp(k * gi.c[2]);
        double s, c;
        sincos(k * wbar, &s, &c);
        const double am = g * k / omega;
        const cplx amp = cmk(am * c, am * s);
        const cplx v0 = cscale(cih, cscale(cb, amp)), v1 = cscale(cih, cscale(sb, amp));
        const cplx v2 = cscale(cih, cmk(amp.y, -amp.x));  // -i * amp
        const cplx dot = cadd(cadd(cscale(nv[0], v0), cscale(nv[1], v1)), cscale(nv[2], v2));
        B[i + (int64_t)n * (ndof + b)] = cmk(-dot.x * gi.pad[1], -dot.y * gi.pad[1]);
    }
}

void launch_bc(hipStream_t st, int n, const PanelGeom* geom, const double* dofs, int ndof, const double* betas, int nb,
               const double* ws, const double* ks, int batch, double g, cplx* B, int64_t sB) {
    bc_kernel<<<dim3((n + 255) / 256, 1, batch), 256, 0, st>>>(n, geom, dofs, ndof, betas, nb, ws, ks, g, B, sB);
}

// ---------------------------------------------------------------------------
// Y = S X for R right-hand sides (_mulvec, matrix_assembly.jl:380-390).  HBM-bound
// on S: each thread owns a row, the j range is split across NSPLIT blocks.
// ---------------------------------------------------------------------------
#define NSPLIT 16
#define XCH 64

// matrix element as FP64 complex (the matrix in FP64, or in FP32 on the surrogate path)
__device__ __forceinline__ cplx mat_d(cplx v) { return v; }
__device__ __forceinline__ cplx mat_d(cplxf v) { return cmk((double)v.x, (double)v.y); }

template <int RB, class M>
__global__ void __launch_bounds__(256) matmul_nr_kernel(int n, int R, const M* __restrict__ Sb, int64_t sS,
                                                        const cplx* __restrict__ Xb, int64_t sX,
                                                        cplx* __restrict__ partb, int64_t sP) {
    const M* S = Sb + sS * blockIdx.z;
    const cplx* X = Xb + sX * blockIdx.z;
    cplx* part = partb + sP * blockIdx.z;
    __shared__ cplx xs[XCH][RB];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int s = blockIdx.y;
    const int per = (n + NSPLIT - 1) / NSPLIT;
    const int jb = s * per, je = min(n, jb + per);
    cplx acc[RB];
#pragma unroll
    for (int c = 0; c < RB; c++) acc[c] = cmk(0.0, 0.0);
    for (int j0 = jb; j0 < je; j0 += XCH) {
        const int cnt = min(XCH, je - j0);
        __syncthreads();
        for (int t = threadIdx.x; t < XCH * RB; t += 256) {
            int jj = t % XCH, c = t / XCH;
            xs[jj][c] = (jj < cnt && c < R) ? X[(j0 + jj) + (int64_t)n * c] : cmk(0.0, 0.0);
        }
        __syncthreads();
        if (i < n)
            for (int jj = 0; jj < cnt; jj++) {
                const cplx a = mat_d(S[i + (int64_t)n * (j0 + jj)]);
#pragma unroll
                for (int c = 0; c < RB; c++) cfma(acc[c], a, xs[jj][c]);
            }
    }
    if (i < n)
#pragma unroll
        for (int c = 0; c < RB; c++)
            if (c < R) part[((int64_t)s * R + c) * n + i] = acc[c];
}

__global__ void reduce_parts_kernel(int n, int R, const cplx* __restrict__ partb, int64_t sP, cplx* __restrict__ Yb,
                                    int64_t sY) {
    const cplx* part = partb + sP * blockIdx.z;
    cplx* Y = Yb + sY * blockIdx.z;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * R) return;
    cplx s = cmk(0.0, 0.0);
    for (int q = 0; q < NSPLIT; q++) s = cadd(s, part[(int64_t)q * n * R + t]);
    Y[t] = s;
}

template <class M>
static void matmul_nr_t(hipStream_t st, int n, int R, const M* S, int64_t sS, const cplx* X, int64_t sX, cplx* Y,
                        int64_t sY, cplx* work, int64_t sW, int batch) {
    dim3 grid((n + 255) / 256, NSPLIT, batch);
    // one instantiation per column count up to 8 (the FP64 multiply-adds are as much of the cost
    // as the matrix stream, so padding columns are not free), 16 above
    switch (R) {
        case 1: matmul_nr_kernel<1, M><<<grid, 256, 0, st>>>(n, R, S, sS, X, sX, work, sW); break;
        case 2: matmul_nr_kernel<2, M><<<grid, 256, 0, st>>>(n, R, S, sS, X, sX, work, sW); break;
        case 3: matmul_nr_kernel<3, M><<<grid, 256, 0, st>>>(n, R, S, sS, X, sX, work, sW); break;
        case 4: matmul_nr_kernel<4, M><<<grid, 256, 0, st>>>(n, R, S, sS, X, sX, work, sW); break;
        case 5: matmul_nr_kernel<5, M><<<grid, 256, 0, st>>>(n, R, S, sS, X, sX, work, sW); break;
        case 6: matmul_nr_kernel<6, M><<<grid, 256, 0, st>>>(n, R, S, sS, X, sX, work, sW); break;
        case 7: matmul_nr_kernel<7, M><<<grid, 256, 0, st>>>(n, R, S, sS, X, sX, work, sW); break;
        case 8: matmul_nr_kernel<8, M><<<grid, 256, 0, st>>>(n, R, S, sS, X, sX, work, sW); break;
        default: matmul_nr_kernel<16, M><<<grid, 256, 0, st>>>(n, R, S, sS, X, sX, work, sW); break;
    }
    const int64_t tot = (int64_t)n * R;
    reduce_parts_kernel<<<dim3((tot + 255) / 256, 1, batch), 256, 0, st>>>(n, R, work, sW, Y, sY);
}

void launch_matmul_nr(hipStream_t st, int n, int R, const cplx* S, int64_t sS, const cplx* X, int64_t sX, cplx* Y,
                      int64_t sY, cplx* work, int64_t sW, int batch) {
    matmul_nr_t(st, n, R, S, sS, X, sX, Y, sY, work, sW, batch);
}

void launch_matmul_nr(hipStream_t st, int n, int R, const cplxf* S, int64_t sS, const cplx* X, int64_t sX, cplx* Y,
                      int64_t sY, cplx* work, int64_t sW, int batch) {
    matmul_nr_t(st, n, R, S, sS, X, sX, Y, sY, work, sW, batch);
}

// ---------------------------------------------------------------------------
// forces: integrate_pressure (waves.jl:154-178) with p = i rho omega phi
// (solve.jl:50) and the Froude-Krylov pressure (waves.jl:259-264, 294-303)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) forces_kernel(int n, const PanelGeom* __restrict__ geom,
                                                     const double* __restrict__ dofs, int ndof,
                                                     const cplx* __restrict__ phib, int64_t sPhi, int R, double rho,
                                                     const double* __restrict__ ws, const double* __restrict__ ks,
                                                     double g, const double* __restrict__ betas, int nb,
                                                     cplx* __restrict__ Fb, cplx* __restrict__ FKb) {
    const double omega = ws[blockIdx.z], k = ks[blockIdx.z];
    const cplx* phi = phib + sPhi * blockIdx.z;
    cplx* F = Fb + (int64_t)gridDim.x * R * blockIdx.z;
    cplx* FK = FKb + (int64_t)gridDim.x * nb * blockIdx.z;
    const int d = blockIdx.x;       // influenced dof
    const int r = blockIdx.y;       // right-hand side (r < R) or FK heading (R + b)
    const double* dof = dofs + (int64_t)3 * n * d;
    const double ro = rho * omega;
    cplx acc = cmk(0.0, 0.0);
    for (int i = threadIdx.x; i < n; i += 256) {
        const PanelGeom& gi = geom[i];
        double disp = dof[i] * gi.nrm[0] + dof[i + n] * gi.nrm[1] + dof[i + 2 * n] * gi.nrm[2];
        cplx p;
        if (r < R) {
            const cplx ph = phi[i + (int64_t)n * r];
            p = cmk(-ro * ph.y, ro * ph.x);
        } else {
            const double beta = betas[r - R];
            const double wbar = gi.c[0] * cos(beta) + gi.c[1] * sin(beta);
            double s, c;
            sincos(k * wbar, &s, &c);
            const double amp = g / omega * exp(k * gi.c[2]);
            const cplx ph = cmk(amp * s, -amp * c);  // (-i g/omega) e^{kz} e^{i k wbar}
            p = cmk(-omega * rho * ph.y, omega * rho * ph.x);
        }
        acc = cadd(acc, cscale(-disp * gi.area, p));
    }
    __shared__ double red[2][256];
    red[0][threadIdx.x] = acc.x;
    red[1][threadIdx.x] = acc.y;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            red[0][threadIdx.x] += red[0][threadIdx.x + off];
            red[1][threadIdx.x] += red[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (r < R) F[d + (int64_t)ndof * r] = cmk(red[0][0], red[1][0]);
        else FK[d + (int64_t)ndof * (r - R)] = cmk(red[0][0], red[1][0]);
    }
}

void launch_forces(hipStream_t st, int n, const PanelGeom* geom, const double* dofs, int ndof, const cplx* phi,
                   int64_t sPhi, int R, double rho, const double* ws, const double* ks, int batch, double g,
                   const double* betas, int nb, cplx* F, cplx* FK) {
    dim3 grid(ndof, R + nb, batch);
    forces_kernel<<<grid, 256, 0, st>>>(n, geom, dofs, ndof, phi, sPhi, R, rho, ws, ks, g, betas, nb, F, FK);
}

}  // namespace mhx
