// Internal declarations shared by the HIP translation units of libmhx.
#pragma once
#include "mhx_device.h"

#include <utility>
#include <vector>

namespace mhx {

// ---- assembly (mhx_assembly.hip)
void launch_panel_geometry(hipStream_t st, int n, const double* verts, const int64_t* faces, const double* centers,
                           const double* normals, const double* areas, const double* radii, PanelGeom* geom);
// Rankine + free-surface image parts (real, wavenumber independent): SR and D_R (direct) or K_R (indirect)
void launch_mark_lid(hipStream_t st, int n, int nhull, PanelGeom* geom);
// forward speed: Rankine part of the x-derivative operator (once per body), the full K_x for a
// wavenumber, the group's right-hand sides, Y += c X
void launch_rankine_x(hipStream_t st, int n, const PanelGeom* geom, double* KRx);
void launch_kx_combine(hipStream_t st, int n, const PanelGeom* geom, const double* KRx, double k, bool exact,
                       const cplx* WR, cplx* Kx);
void launch_bc_fs(hipStream_t st, int n, const PanelGeom* geom, const double* dofs, int ndof, const double* dgrad,
                  double we, double U, double kinc, double winc, double beta, double beta_e, double g, cplx* B);
void launch_caxpy(hipStream_t st, int64_t cnt, double cr, double ci, const cplx* X, cplx* Y);
void launch_rankine(hipStream_t st, int n, const PanelGeom* geom, int direct, double* SR, double* DR);
// Batched launches: blockIdx.z indexes a frequency; ks / ws are device arrays of
// wavenumbers / frequencies; matrices of item z start at base + z * stride.
// S = SR + S_wave, D = DR (+1/2 I) + D_wave for the Wu wave term
void launch_combine_wu(hipStream_t st, int n, const PanelGeom* geom, const double* SR, const double* DR,
                       const double* ks, int batch, int direct, int include_identity, cplx* S, cplx* D,
                       int64_t sMat);
// exact Guevel-Delhommeau pair table (G, dG/dr) for i <= j mirrored to j > i, in two passes
// (first interval for every pair; the adaptive refinement for the pairs listed in rlist, which
// holds rcap >= batch n (n + 1) / 2 entries; *rcount is cleared here).
// status[0]: nonzero if an interval budget was exhausted; status[1..3] += integrand
// evaluations, power-series terms, continued-fraction steps (status is not cleared here)
void launch_delhommeau_table(hipStream_t st, int n, const PanelGeom* geom, const double* ks, int batch, double rtol,
                             double atol, cplx* WG, cplx* WR, int64_t sMat, unsigned long long* rlist,
                             unsigned long long* rcount, int64_t rcap, unsigned long long* status);
// S, D from a pair table of (G, dG/dr) (Delhommeau or surrogate)
void launch_combine_table(hipStream_t st, int n, const PanelGeom* geom, const double* SR, const double* DR,
                          const double* ks, int batch, int direct, int include_identity, const cplx* WG,
                          const cplx* WR, cplx* S, cplx* D, int64_t sMat);

// ---- right-hand sides, products, forces (mhx_assembly.hip)
// B[:, r] for r < ndof: -i omega (n . dof_r);  B[:, ndof+b]: Airy diffraction BC for heading b
// One (body, frequency) item of a multi-body sweep: the body's geometry and Rankine cache,
// the item's output matrices and a device pointer to its wavenumber (read at run time, so a
// replayed sweep picks up new frequencies).  unit0 / tunit0: the item's first unit in the combine
// and table queues (prefix sums over the items).
struct SweepItem {
    const PanelGeom* geom;
    const double* SR;
    const double* DR;
    cplx *S, *D, *WG, *WR;
    cplxf* Df;                    // FP32 copy of D (the surrogate's mixed-precision solve), or null
    const double* k;
    int64_t unit0, tunit0, sunit0;  // first combine / exact-table / surrogate unit of the item
    int n;
    int sd32;  // the surrogate pair kernel stores S and D as FP32 (at the items' FP64 slots)
};
// combine units of an item (256 rows x 8 columns) and table units (16 x 16 blocks, upper triangle)
int64_t sweep_combine_units(int n);
int64_t sweep_table_units(int n);
// exact (G, dG/dr) tables of every item in one launch; workgroups take units from *counter
// (zeroed by the caller on the stream)
void launch_queue_table(hipStream_t st, const SweepItem* items, int nitems, int64_t total_units,
                        unsigned long long* counter, double rtol, double atol, unsigned long long* status);
// S, D of every item in one launch (exact: from the tables, else the Wu fit)
void launch_queue_combine(hipStream_t st, const SweepItem* items, int nitems, int64_t total_units, bool exact,
                          int direct, int include_identity, int rows_per_group);
// zero nbytes (a multiple of 4) with a kernel, so that the operation can sit in a captured graph
void launch_zero(hipStream_t st, void* p, size_t nbytes);
// sticky[0] = first non-zero of info[0..batch) if sticky[0] is still zero
void launch_sticky_info(hipStream_t st, const int* info, int batch, int* sticky);

void launch_bc(hipStream_t st, int n, const PanelGeom* geom, const double* dofs, int ndof, const double* betas, int nb,
               const double* ws, const double* ks, int batch, double g, cplx* B, int64_t sB);
// Y = S X (n x n times n x R); split-K partials in work (16 * n * R complex per item, stride sW);
// S in FP64, or in FP32 (the products and sums in FP64 either way)
void launch_matmul_nr(hipStream_t st, int n, int R, const cplx* S, int64_t sS, const cplx* X, int64_t sX, cplx* Y,
                      int64_t sY, cplx* work, int64_t sW, int batch);
void launch_matmul_nr(hipStream_t st, int n, int R, const cplxf* S, int64_t sS, const cplx* X, int64_t sX, cplx* Y,
                      int64_t sY, cplx* work, int64_t sW, int batch);
// F[d + ndof*r] = sum_i (i rho omega phi[i, r]) * -(dof_d,i . n_i) * a_i ; FK[d + ndof*b] for the Airy
// pressure; item z writes F + z*ndof*R and FK + z*ndof*nb
void launch_forces(hipStream_t st, int n, const PanelGeom* geom, const double* dofs, int ndof, const cplx* phi,
                   int64_t sPhi, int R, double rho, const double* ws, const double* ks, int batch, double g,
                   const double* betas, int nb, cplx* F, cplx* FK);

// ---- dense LU and solves (mhx_lu.hip)
// optional per-launch timing of the trailing-update GEMMs: event pairs recorded on the stream
struct LaunchTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
};
void zgemm(hipStream_t st, bool sub, bool accum, int m, int nn, int kk, const cplx* A, int lda, int64_t sA, const cplx* B,
           int ldb, int64_t sB, cplx* C, int ldc, int64_t sC, int batch);
// Pivoting of getrf_batched: kPivotStrict (row interchanges, LAPACK's partial pivoting) or
// kPivotVerified (diagonal pivots with LAPACK's rule checked on every candidate: a matrix that
// breaks it is flagged in work_perm[kPermWork * b] (nonzero) and must be factored again, from its
// original values, with kPivotStrict; info is then not set by the verified pass).  The verified
// pass needs inv: 2 x 64 x 64 elements per matrix (stride inv_stride) for the inverses of each
// panel's diagonal triangles.
constexpr int kPivotStrict = 0;
constexpr int kPivotVerified = 1;
constexpr int kMaxOrder = 16384;  // largest order of the factorization (leaf register tiles)
constexpr int kSplitSuperMax = 8;  // most panels per split-bf16 trailing update (split_work is sized for it)
constexpr int kSplitSuperDefault = 4;  // panels per split-bf16 update when the LU tuner has not chosen
// inv64 (verified pass, optional): where the FP64 copies of the diagonal blocks' triangular
// inverses go, in the layout of getrs_batched's inverse tables (getrs_inverses(work, n, R) of the
// getrs work of each matrix, stride inv64_stride), so that the solves need not form them again.
int getrf_batched(hipStream_t st, int n, cplx* A, int lda, int64_t strideA, int batch, int* ipiv, int64_t ipiv_stride,
                  int* work_perm, int* info, LaunchTimer* update_timer = nullptr, int pivot = kPivotStrict,
                  cplx* inv = nullptr, int64_t inv_stride = 0, cplx* inv64 = nullptr, int64_t inv64_stride = 0);
// complex elements of getrs work per matrix (sW must be at least this).  getrs_batched keeps the
// diagonal-block inverses and the row permutation in work / perm.  prepared: kGetrsFresh forms
// both; kGetrsPrepared reuses both (a later call on the same factors); kGetrsInverses reuses the
// inverses written by the verified factorization (inv64) and forms the permutation.
constexpr int kGetrsFresh = 0;
constexpr int kGetrsPrepared = 1;
constexpr int kGetrsInverses = 2;
int64_t getrs_work(int n, int R);
cplx* getrs_inverses(cplx* work, int n, int R);  // the inverse tables inside one matrix's getrs work
int getrs_batched(hipStream_t st, int trans, int n, int R, const cplx* LU, int lda, int64_t sA, const int* ipiv,
                  int64_t ipiv_stride, int* perm, int64_t perm_stride, cplx* B, int64_t sB, cplx* work, int64_t sW,
                  int batch, int prepared = kGetrsFresh);
// FP32 factorization (mixed-precision solver); the solves read FP32 factors and compute in FP64
// split: workspace of split_work(n) bf16 per matrix for the split-bf16 trailing updates (nullptr:
// FP32 MFMA updates)
int getrf_batched(hipStream_t st, int n, cplxf* A, int lda, int64_t strideA, int batch, int* ipiv, int64_t ipiv_stride,
                  int* work_perm, int* info, LaunchTimer* update_timer = nullptr, unsigned short* split = nullptr,
                  int pivot = kPivotStrict, cplxf* inv = nullptr, int64_t inv_stride = 0, cplx* inv64 = nullptr,
                  int64_t inv64_stride = 0, int super_panels = kSplitSuperDefault);
// super_panels: 64-column panels factored left-looking before each split-bf16 trailing update
// (1 .. kSplitSuperMax; the update's k is 64 times that), split path only
int64_t split_work(int n);
int getrs_batched(hipStream_t st, int trans, int n, int R, const cplxf* LU, int lda, int64_t sA, const int* ipiv,
                  int64_t ipiv_stride, int* perm, int64_t perm_stride, cplx* B, int64_t sB, cplx* work, int64_t sW,
                  int batch, int prepared = kGetrsFresh);
void convert_to_f32(hipStream_t st, int64_t count, const cplx* src, cplxf* dst);
// per item z < batch: dst[z sD + e] = (FP64) src[z sS + e], e < count
void widen_to_f64(hipStream_t st, int64_t count, int batch, const cplxf* src, int64_t sS, cplx* dst, int64_t sD);
// T <- Bv - T (n R elements per item) and per-item squared norms |T|^2, |Bv|^2 added into norms[2z], norms[2z+1]
void launch_residual(hipStream_t st, int64_t cnt, const cplx* Bv, int64_t sB, cplx* T, int64_t sT, double* norms,
                     int batch);
void launch_add(hipStream_t st, int64_t cnt, const cplx* D, int64_t sD, cplx* X, int64_t sX, int batch);
constexpr int kPermWork = 1 + 4 * 64;  // ints per matrix for getrf_batched's work_perm
// device-to-device copy by a kernel (bytes a multiple of 16; graph-safe)
void launch_copy(hipStream_t st, const void* src, void* dst, size_t bytes);
// out[q] = the verified-pivot flag of item q of a batch factored with kPivotVerified
void launch_gather_pivot_flags(hipStream_t st, const int* wperm, int batch, int* out);

}  // namespace mhx
