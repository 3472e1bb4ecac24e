// Tangent (forward-mode) derivatives of the omega-solve: shared declarations.
#pragma once
#include "mhx_internal.h"

namespace mhx {

// dual number: value and directional derivative
struct Dual {
    double v, d;
};

// panel geometry with tangents: Birk frame, centroid and local corners from the vertices, and
// the mesh fields (centre, normal, area); the radius only selects the far-field rule
struct PanelGeomD {
    Dual T[9], qgc[3], lc[8], c[3], nrm[3], area;
    double radius;
};

void launch_tangent_geometry(hipStream_t st, int n, const double* verts, const double* dverts, const int64_t* faces,
                             const double* centers, const double* dcenters, const double* normals,
                             const double* dnormals, const double* areas, const double* dareas, const double* radii,
                             PanelGeomD* geom);
// dS, dD (n x n, column-major) for wavenumber k with tangent dk; exact: the wave term from the
// (G, dG/dr) table WG, WR of the primal assembly, else the Wu fit
void launch_tangent_pairs(hipStream_t st, int n, const PanelGeomD* geom, double k, double dk, int direct, bool exact,
                          const cplx* WG, const cplx* WR, cplx* dS, cplx* dD);
// tangent of the right-hand sides (ndof radiation columns, then nb diffraction columns)
void launch_tangent_bc(hipStream_t st, int n, const PanelGeomD* geom, const double* dofs, const double* ddofs, int ndof,
                       const double* betas, int nb, double w, double dw, double k, double dk, double g, cplx* dB);
// tangent of the forces: dF (ndof x R) from phi and dphi, dFK (ndof x nb)
void launch_tangent_forces(hipStream_t st, int n, const PanelGeomD* geom, const double* dofs, const double* ddofs,
                           int ndof, const cplx* phi, const cplx* dphi, int R, double rho, double w, double dw, double k,
                           double dk, double g, const double* betas, int nb, cplx* dF, cplx* dFK);
// C = A + s B over cnt complex elements
void launch_axpby(hipStream_t st, int64_t cnt, const cplx* A, double s, const cplx* B, cplx* C);

}  // namespace mhx
