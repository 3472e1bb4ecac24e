// Forward (tangent) and reverse (adjoint) derivatives of the omega-solve: shared declarations.
#pragma once
#include "mhx_internal.h"

namespace mhx {

// dual number with N partials: value and N directional derivatives
template <int N>
struct DualN {
    double v;
    double d[N];
};
using Dual = DualN<1>;

// panel geometry with partials: Birk frame, centroid and local corners from the vertices, and
// the mesh fields (centre, normal, area); the radius only selects the far-field rule
template <int N>
struct PanelGeomN {
    DualN<N> T[9], qgc[3], lc[8], c[3], nrm[3], area;
    double radius;
};
using PanelGeomD = PanelGeomN<1>;

// ---- forward mode (one direction) ----
void launch_tangent_geometry(hipStream_t st, int n, const double* verts, const double* dverts, const int64_t* faces,
                             const double* centers, const double* dcenters, const double* normals,
                             const double* dnormals, const double* areas, const double* dareas, const double* radii,
                             PanelGeomD* geom);
// dS, dD (n x n, column-major) for wavenumber k with tangent dk; exact: the wave term from the
// (G, dG/dr) table WG, WR of the primal assembly, else the Wu fit
void launch_tangent_pairs(hipStream_t st, int n, const PanelGeomD* geom, double k, double dk, int direct, bool exact,
                          const cplx* WG, const cplx* WR, cplx* dS, cplx* dD);
// dK_x (n x n): the forward-speed operator (indirect, field normals e_x) along k, mesh fixed
void launch_tangent_kx(hipStream_t st, int n, const PanelGeomD* geom, double k, double dk, bool exact, const cplx* WG,
                       const cplx* WR, cplx* dKx);
// tangent of the forward-speed diffraction column (incident omega, k; phase heading beta_e,
// velocity heading beta; zero on the lid panels [nhull, n)) along (omega, k)
void launch_tangent_bc_inc(hipStream_t st, int n, int nhull, const PanelGeomD* geom, double w, double dw, double k,
                           double dk, double beta, double beta_e, double g, cplx* dB);
// tangent of the right-hand sides (ndof radiation columns, then nb diffraction columns)
void launch_tangent_bc(hipStream_t st, int n, const PanelGeomD* geom, const double* dofs, const double* ddofs, int ndof,
                       const double* betas, int nb, double w, double dw, double k, double dk, double g, cplx* dB);
// tangent of the forces: dF (ndof x R) from phi and dphi, dFK (ndof x nb)
void launch_tangent_forces(hipStream_t st, int n, const PanelGeomD* geom, const double* dofs, const double* ddofs,
                           int ndof, const cplx* phi, const cplx* dphi, int R, double rho, double w, double dw, double k,
                           double dk, double g, const double* betas, int nb, cplx* dF, cplx* dFK);
// C = A + s B over cnt complex elements
void launch_axpby(hipStream_t st, int64_t cnt, const cplx* A, double s, const cplx* B, cplx* C);

// ---- reverse mode ----
// gradient accumulators (device, zeroed by the caller): vertices (nv x 3), centres / normals
// (n x 3), areas (n), dof fields (same layout as the dofs), and two scalars per call: d/domega
// and d/dk of the current frequency
struct AdjointGrad {
    double *V, *C, *N, *A, *dofs, *wk;
};
// pressure integration: phibar (n x R) from the force cotangents Fbar (ndof x R) and FKbar
// (ndof x nb); panel, dof, omega and k cotangents into g
void launch_adjoint_forces(hipStream_t st, int n, const PanelGeom* geom, const double* dofs, int ndof, const cplx* phi,
                           int R, const cplx* Fbar, const cplx* FKbar, double rho, double w, double k, double g,
                           const double* betas, int nb, cplx* phibar, AdjointGrad gr);
// right-hand sides: panel, dof, omega and k cotangents of Bbar (n x R)
void launch_adjoint_bc(hipStream_t st, int n, const PanelGeom* geom, const double* dofs, int ndof, const double* betas,
                       int nb, double w, double k, double g, const cplx* Bbar, AdjointGrad gr);
// assembly: Sbar = U V^H and Dbar = U2 V2^H (n x R factors, never formed); vertex, centre,
// normal, area and k cotangents of every pair
void launch_adjoint_pairs(hipStream_t st, int n, const double* verts, const int64_t* faces, const double* centers,
                          const double* normals, const double* areas, const double* radii, double k, int direct,
                          bool exact, const cplx* WG, const cplx* WR, const cplx* U, const cplx* V, const cplx* U2,
                          const cplx* V2, int R, AdjointGrad gr);
// Y = S^H X (n x n times n x R)
void launch_matmul_h(hipStream_t st, int n, int R, const cplx* S, const cplx* X, cplx* Y);
// C = s A over cnt complex elements
void launch_scale(hipStream_t st, int64_t cnt, double s, const cplx* A, cplx* C);

}  // namespace mhx
