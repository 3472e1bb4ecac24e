/*
 * mhx.h — C ABI of libmhx, the MI355X (gfx950) boundary-element hot path of
 * MarineHydro.jl: Rankine + wave Green-function panel integrals, dense complex
 * S/D assembly, dense complex LU solve, pressure integration into added mass,
 * radiation damping and excitation force.
 *
 * Plain pointers and sizes only.  All host buffers are owned by the caller;
 * device memory is owned by the context.  Every call is synchronous and
 * returns MHX_OK (0) or a negative status; mhx_last_error() gives the message.
 * A context is not thread-safe: use one per host thread (Julia task).
 *
 * Array conventions are those of the reference's Julia types, so Julia passes
 * its own arrays without a copy:
 *   vertices  nv x 3 (x,y,z per vertex)           Vector{SVector{3,Float64}}
 *   faces     nf x 4 int64, 1-based               Vector{SVector{4,Int}}
 *   centers, normals nf x 3; areas, radii nf      (src/meshes.jl:42-51)
 *   complex   interleaved (re, im) doubles        ComplexF64 / double _Complex
 *   matrices  n x n column-major                  Matrix{ComplexF64}
 *   dofs      ndof consecutive n x 3 column-major Matrix{Float64}(n, 3) each
 *             (src/bodies.jl:74-118)
 */
#ifndef MHX_H
#define MHX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes; the Julia glue maps EINVAL -> ArgumentError, EDIM -> DimensionMismatch */
#define MHX_OK 0
#define MHX_EINVAL (-1)     /* bad argument value */
#define MHX_EDIM (-2)       /* inconsistent sizes */
#define MHX_ESINGULAR (-3)  /* exactly zero pivot (LAPACK info > 0) */
#define MHX_EDEVICE (-4)    /* HIP runtime failure */
#define MHX_ENOMEM (-5)     /* device allocation failure */
#define MHX_EQUAD (-6)      /* adaptive quadrature interval budget exhausted */

/* wave Green functions: GFWu() and ExactGuevelDelhommeau() of src/formulations.jl:26-37 */
#define MHX_GF_WU 0
#define MHX_GF_DELHOMMEAU 1
#define MHX_GF_SURROGATE 2  /* learned surrogate of the exact term (Wu fit + network residual); exact
                               fallback outside the trained box and near the free surface */

/* solve schedule of mhx_hydro_coeffs */
#define MHX_MODE_PER_OMEGA 0   /* one assembly + one LU per frequency, all right-hand sides together */
#define MHX_MODE_PER_PROBLEM 1 /* reference semantics: assembly + LU per radiation/diffraction problem */
#define MHX_MODE_BATCHED 2     /* frequencies in chunks: one launch sequence (assembly, batched LU, solves,
                                  forces) per chunk; chunk size from mhx_set_batch */
#define MHX_MODE_GRAPH 3       /* MHX_MODE_BATCHED as a one-body mhx_sweep_coeffs captured into a HIP
                                  graph on the first call and replayed on the next calls with the same
                                  body, frequency count and headings (the frequencies may change) */

typedef struct mhx_ctx mhx_ctx;
typedef struct mhx_body mhx_body;

typedef struct {
    double rho;     /* fluid density (problem kwarg rho, src/problems_and_results.jl:30) */
    double g;       /* gravity (src/constants.jl:15) */
    int device;     /* HIP device ordinal */
    int timing;     /* 1: record per-phase HIP events (see mhx_phase_time) */
} mhx_options;

typedef struct {
    int64_t nv, nf;
    const double* vertices;
    const int64_t* faces;
    const double* centers;
    const double* normals;
    const double* areas;
    const double* radii;
} mhx_mesh;

/* context lifetime */
int mhx_create(mhx_ctx** ctx, const mhx_options* opt);
void mhx_destroy(mhx_ctx* ctx);
const char* mhx_last_error(const mhx_ctx* ctx);
int mhx_set_constants(mhx_ctx* ctx, double rho, double g); /* set_rho!/set_g! (src/constants.jl:21-29) */

/* assemble_matrices(gfs, mesh, k; direct, include_identity) for gfs = (Rankine(), RankineReflected(), gf)
 * (src/matrix_assembly.jl:164-208).  S_out, D_out: n x n complex column-major host buffers. */
int mhx_assemble(mhx_ctx* ctx, const mhx_mesh* mesh, double k, int gf, int direct, int include_identity,
                 double* S_out, double* D_out);

/* solve(D, S, bc; direct) (src/matrix_assembly.jl:436-449): direct phi = D \ (S bc);
 * indirect sources = D \ bc, phi = S sources.  bc, phi_out, sources_out: n x nrhs complex
 * (sources_out may be NULL; it is left untouched for the direct method).  More than 16 columns
 * are solved in chunks of 16 (each with its own factorization). */
int mhx_solve(mhx_ctx* ctx, int64_t n, int64_t nrhs, const double* D, const double* S, const double* bc, int direct,
              double* phi_out, double* sources_out);

/* zgetrf + zgetrs on the device (LinearSolve LUFactorization, src/matrix_assembly.jl:411-416):
 * X = A \ B (trans = 0) or A^H \ B (trans = 2); B overwritten by X.  Returns MHX_ESINGULAR on a zero pivot.
 * More than 16 right-hand sides are solved in column chunks of 16 (each with its own factorization). */
int mhx_lu_solve(mhx_ctx* ctx, int64_t n, int64_t nrhs, const double* A, int trans, double* B);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): mhx_lu_solve on device
 * buffers with no allocation.  A (n x n, column-major complex, device memory of the context's
 * device) is overwritten by its LU factors (LAPACK pivoting), B (n x nrhs, device) by
 * X = A \ B (trans = 0) or A^H \ B (trans = 2).  The pivots and the solve workspace live in a
 * static device array of MHX_NOALLOC_SCRATCH_BYTES per GPU (16 MiB: ipiv and the triangular-solve
 * workspace, 2 n R + 128 ceil(n / 64) 64 complex values, e.g. any order up to 4096 with 16
 * right-hand sides, 6000 with 16); a call whose workspace exceeds it returns MHX_EDIM.  Pivoting
 * is always MHX_PIVOT_ALWAYS (the original matrix is not kept).  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (several contexts or host threads) race.
 * Synchronous; MHX_ESINGULAR on an exactly zero pivot. */
#define MHX_NOALLOC_SCRATCH_BYTES (16ll << 20)
int mhx_lu_solve_noalloc(mhx_ctx* ctx, int64_t n, int64_t nrhs, double* A, int trans, double* B);

/* A batch of independent systems of one order (LinearSolve LUFactorization on each matrix, as the
 * reference's solve(problems::Vector) issues one per problem: src/matrix_assembly.jl:411-416,
 * src/solve.jl:75-78): X_q = A_q \ B_q (trans = 0) or A_q^H \ B_q (trans = 2) for q < batch.
 * A: batch consecutive n x n column-major complex matrices; B: batch consecutive n x nrhs blocks,
 * overwritten by the solutions.  The precision follows mhx_set_lu_mode (MHX_LU_AUTO: FP64; the
 * mixed modes solve trans = 0 only and fall back to FP64 per the mixed-solver rules). */
int mhx_lu_solve_batched(mhx_ctx* ctx, int64_t n, int64_t nrhs, int64_t batch, const double* A, int trans, double* B);

/* hydrodynamic_coefficients(body, (wave_frequencies, radiating_dofs=all, wave_directions), formulation)
 * (src/problems_and_results.jl:464-495) with every dof radiating and influenced:
 *   A_out, B_out  ndof x ndof x nw          (influenced, radiating, omega)
 *   Fd_out, Fk_out ndof x nw x nb complex   (influenced, omega, beta); excitation = Fd + Fk
 * One solve takes ndof + nb <= 16 right-hand sides; more headings are solved in chunks of
 * 16 - ndof (each chunk with its own assembly and factorization; ndof <= 15).  MHX_MODE_GRAPH
 * takes ndof + nb <= 16. */
int mhx_hydro_coeffs(mhx_ctx* ctx, const mhx_mesh* mesh, int64_t ndof, const double* dofs, int64_t nw,
                     const double* omegas, int64_t nb, const double* betas, int gf, int direct, int mode,
                     double* A_out, double* B_out, double* Fd_out, double* Fk_out);

/* device-resident body (FloatingBody with its rigid dofs, src/bodies.jl:9-51): geometry and the
 * wavenumber-independent Rankine part stay on the device across calls */
int mhx_body_create(mhx_ctx* ctx, const mhx_mesh* mesh, int64_t ndof, const double* dofs, mhx_body** body);

/* A body with a lid for irregular-frequency removal (the reference's FloatingBody(mesh, lid_mesh,
 * ...): solve.jl:4-11, 22-28, bodies.jl:9-28): mesh = hull panels [0, nhull) followed by the lid
 * panels, dofs over all panels with zero rows on the lid.  Lid panels take no incident-wave
 * condition, and a panel in the free surface (|z| < 1e-8) has the identity jump 1 instead of
 * 1/2 (matrix_assembly.jl:25-27); pressures integrate over the hull only (the dofs vanish on the
 * lid).  Used with mhx_body_coeffs / mhx_sweep_coeffs. */
int mhx_body_create_lid(mhx_ctx* ctx, const mhx_mesh* mesh, int64_t nhull, int64_t ndof, const double* dofs,
                        mhx_body** body);
void mhx_body_destroy(mhx_body* body);
int mhx_body_coeffs(mhx_ctx* ctx, mhx_body* body, int64_t nw, const double* omegas, int64_t nb, const double* betas,
                    int gf, int direct, int mode, double* A_out, double* B_out, double* Fd_out, double* Fk_out);

/* Forward speed (the reference's forward_speeds parameter: problems_and_results.jl:236-380,
 * waves.jl:136-151, 183-199, 266-303, solve.jl:13-19, 35-55).  For every (U, omega, beta) the
 * radiation problems of all dofs and the diffraction problem at heading beta are solved at the
 * encountered frequency omega_e = |omega - k U cos beta| (k = omega^2/g) and heading beta_e (beta,
 * or beta + pi when the Doppler shift is negative); radiation conditions
 * -i omega_e (n . dof) + U (n . dgrad), dgrad the constant gradient of motion of a rigid dof
 * (bodies.jl:55-66: ndof x 3, Pitch (0,0,1), Yaw (0,-1,0), else 0; NULL when every U is 0);
 * the pressure i rho omega_e phi + rho U dphi/dx with dphi/dx = K_x sigma (the indirect operator
 * with every panel normal replaced by e_x).  U = 0 is the plain problem (any formulation); a
 * nonzero U needs direct = 0.  gf: any; with MHX_GF_SURROGATE, K_x takes dG/dr from the same
 * surrogate table as S and D.  The dense solves follow mhx_set_lu_mode.  Outputs, column-major
 * like the reference's arrays:
 *   A_out, B_out   ndof x ndof x nw x nU x nb     Re F / omega_e^2, Im F / omega_e
 *   Fd_out, Fk_out ndof x nw x nb x nU complex    (Froude-Krylov with the incident omega, k and beta_e) */
int mhx_body_coeffs_fs(mhx_ctx* ctx, mhx_body* body, int64_t nw, const double* omegas, int64_t nb, const double* betas,
                       int64_t nU, const double* speeds, const double* dgrad, int gf, int direct, double* A_out,
                       double* B_out, double* Fd_out, double* Fk_out);

/* frequency-chunk limits of MHX_MODE_BATCHED (defaults: 64 items, 60% of the device memory free at
 * mhx_create) */
int mhx_set_batch(mhx_ctx* ctx, int64_t max_items, double max_bytes);

/* Dense solver precision (replaces LinearSolve's LUFactorization, src/matrix_assembly.jl:396-416):
 *   MHX_LU_FP64   LU in FP64 with LAPACK pivoting (the reference's zgetrf/zgetrs);
 *   MHX_LU_MIXED  LU of an FP32 copy (FP32 MFMA trailing updates), triangular solves with those
 *                 factors in FP64 arithmetic, FP64 residuals against the FP64 matrix and iterative
 *                 refinement to |b - A x| <= 1e-12 |b| per system; an item that does not reach it in
 *                 4 steps (or has a singular FP32 factor) is re-solved with the FP64 LU;
 *   MHX_LU_MIXED_BF16  as MHX_LU_MIXED with the FP32 factor's trailing updates on bf16 MFMA, each
 *                 operand split into bf16 hi + lo (products to ~2^-16), up to 8 refinement steps;
 *   MHX_LU_AUTO   (default) MHX_LU_MIXED_BF16 for MHX_GF_SURROGATE, FP64 otherwise.
 * mhx_solver_stats: counts[0] systems solved mixed, counts[1] refinement steps, counts[2] FP64
 * fallbacks; max_residual: largest accepted relative residual (since mhx_reset_timing). */
#define MHX_LU_AUTO 0
#define MHX_LU_FP64 1
#define MHX_LU_MIXED 2
#define MHX_LU_MIXED_BF16 3
int mhx_set_lu_mode(mhx_ctx* ctx, int mode);
int mhx_solver_stats(mhx_ctx* ctx, int64_t counts[3], double* max_residual);
/* Streams a batched FP32 factorization (MHX_LU_MIXED, MHX_LU_MIXED_BF16) is split over, 1 .. 4
 * (default 2): a batch of 8 or more systems of order >= 512 is factored in that many parts of at
 * least 4 systems, on as many streams, so that one part's panel phases overlap the others'
 * trailing updates.  Results do not depend on it.  No counterpart in the reference (its LU is one
 * LAPACK call per system); a tuning control of this library's batched solver. */
int mhx_set_lu_streams(mhx_ctx* ctx, int streams);
/* Autotuning of the batched FP32 factorization (on by default): on the first batched factorization
 * of a signature (order, batch size, update kind, pivoting) the library factors that batch with
 * 1-3 streams and 2-8 panels per split-bf16 update (each timed twice), keeps the fastest and uses it for every later
 * call of the signature (streams set with mhx_set_lu_streams are kept).  mhx_lu_config: the
 * streams and panels per update of the last batched FP32 factorization (0, 0 before any).  No
 * counterpart in the reference; tuning controls of this library's batched solver. */
int mhx_set_autotune(mhx_ctx* ctx, int on);
int mhx_lu_config(mhx_ctx* ctx, int* streams, int* super_panels);

/* Pivoting of the dense LU (LAPACK's partial pivoting either way, src/matrix_assembly.jl:411-416):
 *   MHX_PIVOT_VERIFIED (default)  factor with the diagonal pivots while checking LAPACK's rule on
 *                 every candidate (the diagonal is a largest |Re| + |Im| of its column after the
 *                 earlier eliminations, first index on ties); a matrix that breaks it anywhere, or
 *                 meets an exactly zero pivot, is factored again from its original values with row
 *                 interchanges.  The influence matrices of the BEM keep their diagonal pivots, so
 *                 the panel work needs no pivot search;
 *   MHX_PIVOT_ALWAYS  row interchanges with the pivot search on every column.
 * mhx_pivot_stats: counts[0] factorizations that kept the verified diagonal pivots, counts[1]
 * factorizations with the pivot search (since mhx_reset_timing). */
#define MHX_PIVOT_VERIFIED 0
#define MHX_PIVOT_ALWAYS 1
int mhx_set_pivoting(mhx_ctx* ctx, int mode);
int mhx_pivot_stats(mhx_ctx* ctx, int64_t counts[2]);

/* per-phase timing (HIP events on the context's stream, enabled by mhx_options.timing):
 * phase: "rankine", "wave_table" (wave term: exact table, or the fused surrogate pair kernel),
 * "wave" (S, D formation from the table or the Wu fit), "lu" (factorization), "lu_update" (each
 * trailing-update GEMM launch inside "lu", timed separately), "solve", "forces", "forward_speed"
 * (K_x, K_x sigma and the potential update of mhx_body_coeffs_fs);
 * returns total ms and launch count since reset */
int mhx_phase_time(mhx_ctx* ctx, const char* phase, double* total_ms, int64_t* count);
void mhx_reset_timing(mhx_ctx* ctx);
/* per-phase timing on (1) or off (0) from the next call on (the events of calls already made are
 * kept); off, no events are recorded.  No counterpart in the reference; instrumentation of this
 * library. */
int mhx_set_timing(mhx_ctx* ctx, int on);
/* Refinement steps the mixed-precision solver may take before an item is redone with the FP64 LU:
 * -1 (default) restores the solver's own limit (4 with FP32 trailing updates, 8 with split-bf16
 * ones), 0 .. 8 sets it.  The sweep (mhx_sweep_coeffs) always takes one step and redoes an item
 * that missed 1e-12 through the per-frequency FP64 path; with 0 it redoes every item.  0 is the
 * diagnostic setting that sends every system down the FP64 path.  No counterpart in the
 * reference (its solves are one LAPACK call); a control of this library's mixed-precision solver. */
int mhx_set_refine_limit(mhx_ctx* ctx, int steps);

/* work of the exact Guevel-Delhommeau wave term since the last mhx_reset_timing: counts =
 * {integrand evaluations, E1 power-series terms, E1 continued-fraction steps}; flops = the
 * algorithmic FP64 operation count derived from them (no counterpart in the reference;
 * used for roofline reporting) */
int mhx_exact_work(mhx_ctx* ctx, int64_t counts[3], double* flops);

/* Multi-body sweep (several FloatingBodies of uneven panel counts, each with its own frequency
 * list, common headings): the reference's solve(problems::Vector) / hydrodynamic_coefficients
 * over a list of bodies (solve.jl:75-78, problems_and_results.jl:464-495).  The assembly of
 * every (body, frequency) item runs as one persistent queue launch per stage (exact table,
 * then S/D; with MHX_GF_SURROGATE one fused pair launch writing S, D and the FP32 copy of D), the
 * LU is batched per body, and the frequencies are processed in as few chunks as the
 * device-memory budget allows.  Outputs per body as mhx_body_coeffs.  gf: MHX_GF_WU,
 * MHX_GF_DELHOMMEAU or MHX_GF_SURROGATE.  The dense solves follow mhx_set_lu_mode: FP64, or the
 * FP32 (split-bf16) LU with one refinement step against the FP64 matrix on the device; a
 * (body, frequency) whose residual then exceeds 1e-12 |b| is solved again in FP64 after the
 * sweep, as is one whose factorization broke LAPACK's pivoting rule (with row interchanges).
 * flags: MHX_SWEEP_GRAPH captures the whole sweep into a
 * HIP graph on the first call and replays it on later calls with the same bodies and frequency
 * counts (the frequency values may change).  A body's buffers belong to the sweep until the
 * next sweep call or mhx_body_destroy. */
typedef struct {
    mhx_body* body;
    int64_t nw;
    const double* omegas;
    double* A_out;  /* (ndof, ndof, nw) */
    double* B_out;
    double* Fd_out; /* (ndof, nw, nb) complex interleaved */
    double* Fk_out;
} mhx_sweep_body;
#define MHX_SWEEP_GRAPH 1
int mhx_sweep_coeffs(mhx_ctx* ctx, const mhx_sweep_body* bodies, int64_t nbodies, int64_t nb, const double* betas,
                     int gf, int direct, int flags);

/* Tangent (forward-mode) derivative of hydrodynamic_coefficients (problems_and_results.jl:464-495)
 * along a direction of the inputs: the mesh fields (vertices nv x 3, centers / normals nf x 3,
 * areas nf; NULL = zero), the dof fields ddofs (same layout as dofs; NULL = zero) and the
 * frequencies domegas (nw; NULL = zero).  Returns the coefficients (as mhx_hydro_coeffs, one LU per
 * frequency in FP64) and their directional derivatives dA, dB (ndof, ndof, nw), dFd, dFk
 * (ndof, nw, nb complex).  It is the derivative the reference computes with ForwardDiff/Enzyme
 * through its rules for the panel integrals and the solve (ad_rules.jl:79-414), with the mesh
 * tangent seeded as fd_mesh_rules! does (ad_rules_geometry.jl:174-230); gf: MHX_GF_WU or
 * MHX_GF_DELHOMMEAU.  The radius derivative of a body scaled about the origin is the direction
 * (vertices / r, centers / r, normals 0, 2 areas / r). */
typedef struct {
    const double* vertices;
    const double* centers;
    const double* normals;
    const double* areas;
} mhx_mesh_tangent;
int mhx_hydro_coeffs_jvp(mhx_ctx* ctx, const mhx_mesh* mesh, const mhx_mesh_tangent* dmesh, int64_t ndof,
                         const double* dofs, const double* ddofs, int64_t nw, const double* omegas,
                         const double* domegas, int64_t nb, const double* betas, int gf, int direct, double* A_out,
                         double* B_out, double* Fd_out, double* Fk_out, double* dA_out, double* dB_out,
                         double* dFd_out, double* dFk_out);

/* Frequency derivative of the coefficients of a body with a lid and/or at forward speed (the
 * reference differentiates both in omega with ForwardDiff/Enzyme: test/solve_problem_differentiation.jl:94-98
 * for a lid body, test/forward_speed_tests.jl:280-303 for a moving one).  mesh: hull panels
 * [0, nhull) then the lid panels (nhull = nf: no lid), dofs over all panels with zero rows on
 * the lid (as mhx_body_create_lid).  For every (U, omega, beta) as mhx_body_coeffs_fs (U = 0
 * included; a nonzero U needs direct = 0 and dgrad), with the directional derivatives along
 * domegas (nw; NULL = zero) of every output, through the tangent kernels: the encountered
 * frequency's tangent, dS/dD (or dK) and dK_x along the wavenumber, the tangent right-hand sides
 * and the tangent solve on the same FP64 LU.  Layouts as mhx_body_coeffs_fs:
 *   A_out, B_out, dA_out, dB_out   ndof x ndof x nw x nU x nb
 *   Fd_out, Fk_out, dFd_out, dFk_out  ndof x nw x nb x nU complex.  gf: MHX_GF_WU or MHX_GF_DELHOMMEAU. */
int mhx_hydro_coeffs_fs_jvp(mhx_ctx* ctx, const mhx_mesh* mesh, int64_t nhull, int64_t ndof, const double* dofs,
                            int64_t nw, const double* omegas, const double* domegas, int64_t nb, const double* betas,
                            int64_t nU, const double* speeds, const double* dgrad, int gf, int direct, double* A_out,
                            double* B_out, double* Fd_out, double* Fk_out, double* dA_out, double* dB_out,
                            double* dFd_out, double* dFk_out);

/* Reverse-mode gradient of the coefficients: for the real functional
 *   J = sum Abar .* A + sum Bbar .* B + Re sum conj(Fdbar) .* Fd + Re sum conj(Fkbar) .* Fk
 * (cotangents in the layouts of the outputs, Fd/Fk complex interleaved; NULL = zero), returns
 * the coefficients (FP64 LU per frequency) and dJ/d(input) for every non-NULL field of grad:
 * vertices (nv x 3), centers / normals (nf x 3), areas (nf), dofs (the dofs layout) and omegas
 * (nw).  This is what the reference's Enzyme / Zygote reverse mode returns through its rules
 * (ad_rules.jl:79-414): the transposed solve reuses the LU (z = A^H \ ybar, Abar = -z y^H) and
 * the assembly is pulled back pair by pair.  gf: MHX_GF_WU or MHX_GF_DELHOMMEAU. */
typedef struct {
    double* vertices;
    double* centers;
    double* normals;
    double* areas;
    double* dofs;
    double* omegas;
} mhx_mesh_grad;
int mhx_hydro_coeffs_vjp(mhx_ctx* ctx, const mhx_mesh* mesh, int64_t ndof, const double* dofs, int64_t nw,
                         const double* omegas, int64_t nb, const double* betas, int gf, int direct, const double* A_bar,
                         const double* B_bar, const double* Fd_bar, const double* Fk_bar, double* A_out, double* B_out,
                         double* Fd_out, double* Fk_out, const mhx_mesh_grad* grad);

/* Learned surrogate of the exact Guevel-Delhommeau wave term (no counterpart in the reference; it
 * replaces the exact term of src/green_functions/exact_Guevel_Delhommeau.jl:15-59 inside assembly).
 * The network maps (r, v) = (k R_h, k |z + zeta|) through 64 Fourier features and 4 x 128 hidden
 * layers (SiLU by default, or ReLU) to the residual (exact - GFWu) of G, dG/dr, dG/dz: real parts as 100 x residual,
 * imaginary parts as 100 exp(min(v, 5)) x residual (they decay like exp(-v)).  It is trained on
 * the device on ngrid^2 log-spaced jittered samples generated by the exact FP64 kernel.
 *   train: Adam, cosine learning rate, minibatches of `batch`; val_loss = MSE on the held-out samples
 *   save / load: versioned binary file (magic line, JSON header line, Fourier matrix, FP32 parameters)
 *   params: copies the FP32 parameters (nparams floats) and the 32 x 2 Fourier matrix to the host
 *   eval: the BF16-MFMA inference network at m points rv[2m] = (r, v); out[6m] its raw outputs
 *   loss_grad: one FP32 forward/backward of the training kernels on m samples (U m x 2, Y m x 6)
 *              with the given parameters; grad_out has nparams floats
 *   set_activation: the hidden-layer activation of the next training and of loss_grad (a loaded
 *              network of the other activation is dropped); load takes the file's activation;
 *              get_activation: the context's current one */
#define MHX_SURR_SILU 0
#define MHX_SURR_RELU 1
int mhx_surrogate_train(mhx_ctx* ctx, int64_t ngrid, int64_t steps, int64_t batch, double lr, uint64_t seed,
                        double* val_loss);
int mhx_surrogate_save(mhx_ctx* ctx, const char* path);
int mhx_surrogate_load(mhx_ctx* ctx, const char* path);
int mhx_surrogate_params(mhx_ctx* ctx, float* params_out, float* fourier_out, int64_t* nparams);
int mhx_surrogate_eval(mhx_ctx* ctx, int64_t m, const double* rv, float* out);
int mhx_surrogate_loss_grad(mhx_ctx* ctx, int64_t m, const float* U, const float* Y, const float* params,
                            const float* fourier, float* grad_out, float* loss_out);
int mhx_surrogate_set_activation(mhx_ctx* ctx, int activation);
int mhx_surrogate_get_activation(mhx_ctx* ctx, int* activation);

#ifdef __cplusplus
}
#endif
#endif
