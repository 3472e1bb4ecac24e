/*
 * mh_oracle — CPU FP64 restatement of MarineHydro.jl's boundary-element hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product library (libmhx.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path never does.
 *
 * Conventions (same as the reference's StaticArraysMesh, src/meshes.jl:42-51):
 *   vertices  nv x 3   row-major (x,y,z per vertex)  == Vector{SVector{3,Float64}}
 *   faces     nf x 4   row-major int64, 1-based       == Vector{SVector{4,Int}}
 *   centers / normals  nf x 3 row-major; areas / radii nf
 *   matrices  n x n complex128, column-major (Julia Matrix{ComplexF64})
 *   dofs      ndof blocks of an n x 3 column-major matrix (Julia Matrix{Float64}(n,3))
 */
#ifndef MH_ORACLE_H
#define MH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int64_t nv, nf;
    const double* vertices;
    const int64_t* faces;
    const double* centers;
    const double* normals;
    const double* areas;
    const double* radii;
} ora_mesh;

enum { ORA_GF_WU = 0, ORA_GF_DELHOMMEAU = 1 };

void ora_set_threads(int nthreads);
int ora_get_threads(void);

/* Birk panel (src/green_functions/rankine_vectorized.jl) */
void ora_birk_panel_geometry(const double V[12], double T[9], double qgc[3], double lc[8], double* area);
double ora_velocity_potential(double x, double y, double z, const double lc[8]);
void ora_velocity_derivatives(double x, double y, double z, const double lc[8], double v[3]);
void ora_velocity_hessian(double x, double y, double z, const double lc[8], double H[9]);
void ora_birk_phi_and_velocity(double x, double y, double z, const double lc[8], double* phi, double v[3]);
void ora_vrankine_from_geom(const double p[3], const double c[3], double radius, double area,
                            const double T[9], const double qgc[3], const double lc[8],
                            double* phi, double grad[3]);

/* Wu wave term (src/green_functions/wu.jl): out = [ReG, ImG, ReGh, ImGh, ReGv, ImGv] */
void ora_wu_wave_term_and_derivs(double hh, double vv, double out[6]);
double ora_struve_h0(double x);
double ora_struve_h1(double x);

/* complex exponential integral */
void ora_expint_e1(double zr, double zi, double out[2]);
void ora_expE1(double zr, double zi, double out[2]); /* exp(z)*E1(z) */

/* exact Guevel-Delhommeau (src/green_functions/exact_Guevel_Delhommeau.jl)
 * out = [ReG, ImG, ReGr, ImGr, ReGz, ImGz]; rtol/atol of the adaptive GK(7,15)
 * (1e-8/1e-8 reproduces the reference's HCubatureJL defaults).
 * nboxes (optional) returns the final interval counts of the G and Gr quadratures. */
int ora_delhommeau_term(double r, double z, double rtol, double atol, double out[6], int nboxes[2]);
void ora_gk15_rule(double xk[8], double wk[8], double wg[4]);

/* assembly (src/matrix_assembly.jl) */
int ora_assemble_rankine(const ora_mesh* m, int direct, int include_identity, double* S, double* D);
int ora_assemble_wave(const ora_mesh* m, double k, int gf, int direct, double* S, double* D);
int ora_assemble(const ora_mesh* m, double k, int gf, int direct, int include_identity, double* S, double* D);
/* assemble only rows [i0,i1) of the full n x n matrices into ldS-strided column-major buffers */
int ora_assemble_rows(const ora_mesh* m, double k, int gf, int direct, int include_identity,
                      int64_t i0, int64_t i1, double* S, double* D);

/* dense complex LU with partial pivoting (max |Re|+|Im|, as LAPACK izamax) */
int ora_zgetrf(int64_t n, double* A, int64_t lda, int64_t* ipiv);
int ora_zgetrs(int trans, int64_t n, int64_t nrhs, const double* A, int64_t lda, const int64_t* ipiv,
               double* B, int64_t ldb);
int ora_solve(int64_t n, int64_t nrhs, const double* D, const double* S, const double* bc, int direct,
              double* phi, double* sources);

/* boundary conditions, pressure (src/waves.jl) */
void ora_rigid_dof(const ora_mesh* m, int which, const double rc[3], double* dof);
void ora_radiation_bc(const ora_mesh* m, const double* dof, double omega, double* bc);
void ora_airy_bc(const ora_mesh* m, double omega, double k, double beta, double g, double* bc);
void ora_airy_pressure(const ora_mesh* m, double omega, double k, double beta, double g, double rho, double* p);
void ora_integrate_pressure(const ora_mesh* m, const double* dof, const double* p, double out[2]);

/* hydrodynamic_coefficients (src/problems_and_results.jl:464-495)
 *   A, B   (ninf, nrad, nw)  Fd, Fk (ninf, nw, nb) complex, column-major
 *   mode 0: one assembly + one LU per omega, all right-hand sides at once
 *   mode 1: reference semantics, one assembly + LU per problem (src/solve.jl:75-78) */
int ora_hydro_coeffs(const ora_mesh* m, int64_t ndof, const double* dofs, int64_t nw, const double* omegas,
                     int64_t nb, const double* betas, double rho, double g, int gf, int direct, int mode,
                     double* A, double* B, double* Fd, double* Fk);
int ora_hydro_coeffs_lid(const ora_mesh* m, int64_t nhull, int64_t ndof, const double* dofs, int64_t nw,
                         const double* omegas, int64_t nb, const double* betas, double rho, double g, int gf,
                         int direct, int mode, double* A, double* B, double* Fdd, double* Fkd);
int ora_assemble_fixed_normal(const ora_mesh* m, double k, int gf, int direct, const double* fixed, double* S,
                              double* D);
void ora_encountered(double omega, double beta, double U, double g, double out[3]);
int ora_hydro_coeffs_fs(const ora_mesh* m, int64_t ndof, const double* dofs, const double* dgrad, int64_t nw,
                        const double* omegas, int64_t nb, const double* betas, int64_t nU, const double* Us,
                        double rho, double g, int gf, double* A, double* B, double* Fdd, double* Fkd);

#ifdef __cplusplus
}
#endif
#endif
