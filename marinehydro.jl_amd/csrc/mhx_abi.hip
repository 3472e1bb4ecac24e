// C ABI of libmhx (declared in include/mhx.h): contexts, device-resident bodies,
// and the host orchestration of one frequency solve
// (reference: src/solve.jl:47-78, src/problems_and_results.jl:464-495).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mhx.h"
#include "mhx_surrogate.h"
#include "mhx_tangent.h"



struct SweepState;  // multi-body sweep: item queue, device outputs, captured graph

constexpr int kMaxLuStreams = 4;

struct PhaseAcc {
    double ms = 0.0;
    int64_t count = 0;
};

struct mhx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // side streams (default priority) for the other parts of a batched factorization, with the
    // events that fork them from and join them back to `stream`; lu_streams: streams a batched
    // FP32 factorization is split over (1 .. kMaxLuStreams)
    hipStream_t side[kMaxLuStreams - 1] = {};
    hipEvent_t fork = nullptr, join[kMaxLuStreams - 1] = {};
    int lu_streams = 2;
    // panels per split-bf16 trailing update of a batched FP32 factorization (1 .. 8)
    int lu_super = mhx::kSplitSuperDefault;
    // autotuning of the batched FP32 factorization's streams and super-panel width (per order,
    // batch, update kind and pivoting, on the first factorization of that signature); streams set
    // by mhx_set_lu_streams are kept
    bool autotune = true, lu_streams_fixed = false;
    int last_streams = 0, last_super = 0;  // the configuration of the last batched FP32 factorization
    double rho = 1000.0, g = 9.81;
    int timing = 0;
    int refine_limit = -1;  // refinement steps before an FP64 redo (-1: the solver's default)
    std::string err;
    std::map<std::string, PhaseAcc> phases;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    unsigned long long quad_work[3] = {0, 0, 0};  // exact wave term: evaluations, series terms, CF steps
    double batch_bytes = 64.0 * (1ull << 30);      // device memory budget of one frequency chunk (set at creation)
    int64_t max_batch = 64;
    // learned wave-term surrogate
    bool surr_ready = false;
    int surr_act = MHX_SURR_SILU;  // hidden-layer activation of the network (next training / loaded)
    float *surr_params = nullptr, *surr_B = nullptr, *surr_bias = nullptr;
    unsigned short* surr_W = nullptr;
    double surr_val_loss = -1.0;
    int64_t surr_steps = 0;
    // dense solver: precision and pivoting policies, mixed-precision statistics
    int lu_mode = MHX_LU_AUTO;
    int pivoting = MHX_PIVOT_VERIFIED;
    int64_t lu_diagonal = 0, lu_interchanged = 0;  // factorizations that kept / did not keep the diagonal pivots
    int64_t refine_steps = 0, mixed_solves = 0, fp64_fallbacks = 0;
    double max_residual = 0.0;
    // multi-body sweeps: the last sweep's state (kept for graph replay) and the tuner's choices
    SweepState* sweep = nullptr;
    std::map<std::string, int> tuned;
    // this device's copy of the static scratch array of mhx_lu_solve_noalloc (null if its address
    // could not be taken)
    unsigned char* scratch = nullptr;
};

struct mhx_body {
    mhx_ctx* ctx = nullptr;
    int n = 0, ndof = 0;
    PanelGeom* geom = nullptr;
    double* dofs = nullptr;
    double* SR = nullptr;
    double* DR[2] = {nullptr, nullptr};  // [direct] : K_R (indirect) or D_R (direct)
    bool have_rankine[2] = {false, false};
    double* KRx = nullptr;  // Rankine part of the forward-speed x-derivative operator (built on first use)
    // per-frequency work, cap_batch items
    int cap_batch = 0, cap_R = 0, cap_nb = 0;
    bool cap_table = false;
    cplx *S = nullptr, *D = nullptr, *WG = nullptr, *WR = nullptr;
    cplx *B = nullptr, *Y = nullptr, *work = nullptr, *F = nullptr, *FK = nullptr;
    cplx* D0 = nullptr;  // copy of D kept for the pivoted refactorization of a flagged item (FP64 path)
    double *betas = nullptr, *d_ws = nullptr, *d_ks = nullptr;
    int *ipiv = nullptr, *perm = nullptr, *wperm = nullptr, *info = nullptr;
    unsigned long long* status = nullptr;
    unsigned long long *fb_list = nullptr, *fb_count = nullptr;  // surrogate fallback pairs
    int64_t fb_cap = 0;
    // mixed-precision solver: FP32 factors, right-hand-side copy, residuals, norms
    cplxf* Df = nullptr;
    unsigned short* Ps = nullptr;  // split-bf16 panels of the FP32 factorization (MHX_LU_MIXED_BF16)
    cplx *Rr = nullptr, *Tr = nullptr;
    double* norms = nullptr;
    bool df_ready = false;  // Df already holds the FP32 copy of D (written by the fused surrogate assembly)
    // S and D of the assembled batch are stored as FP32 complex, item q's matrix at the start of its
    // FP64 slot ((cplxf*)(S + q nn)): the surrogate path with the mixed-precision solver
    bool sd32 = false;
    // generation of the per-frequency buffers above: a new value whenever any of them is freed or
    // reallocated, so that a captured sweep graph holding their addresses is not replayed
    uint64_t gen = 0;
    int64_t nn() const { return (int64_t)n * n; }
    int64_t nR() const { return (int64_t)n * cap_R; }
    int64_t sW() const { return 16 * nR() + mhx::getrs_work(n, cap_R); }  // split-K partials + getrs work
};

namespace {

std::atomic<uint64_t> g_buffer_gen{0};
void bump_gen(mhx_body* b) { b->gen = ++g_buffer_gen; }

int fail(mhx_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HCHK(call)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return fail(ctx, MHX_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

template <typename T>
int dalloc(mhx_ctx* ctx, T** p, size_t count) {
    if (*p) return MHX_OK;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * (count ? count : 1));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(ctx, MHX_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return MHX_OK;
}

template <typename T>
void dfree(T*& p) {
    if (p) hipFree(p);
    p = nullptr;
}

struct Phase {
    mhx_ctx* c;
    std::string name;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    Phase(mhx_ctx* c_, const char* nm) : c(c_), name(nm) {
        if (!c->timing) return;
        hipEventCreate(&e0);
        hipEventCreate(&e1);
        hipEventRecord(e0, c->stream);
    }
    ~Phase() {
        if (!c->timing) return;
        hipEventRecord(e1, c->stream);
        c->pending.push_back({name, {e0, e1}});
    }
};

// per-launch events of the trailing-update GEMMs -> phase "lu_update" (one count per launch)
void take_launches(mhx_ctx* c, mhx::LaunchTimer& t) {
    for (auto& e : t.ev) c->pending.emplace_back("lu_update", e);
    t.ev.clear();
}

void flush_timing(mhx_ctx* c) {
    for (auto& p : c->pending) {
        float ms = 0.f;
        hipEventSynchronize(p.second.second);
        hipEventElapsedTime(&ms, p.second.first, p.second.second);
        auto& a = c->phases[p.first];
        a.ms += ms;
        a.count += 1;
        hipEventDestroy(p.second.first);
        hipEventDestroy(p.second.second);
    }
    c->pending.clear();
}

int check_mesh(mhx_ctx* ctx, const mhx_mesh* m) {
    if (!m || !m->vertices || !m->faces || !m->centers || !m->normals || !m->areas || !m->radii)
        return fail(ctx, MHX_EINVAL, "mesh: null array");
    if (m->nf <= 0 || m->nv <= 0) return fail(ctx, MHX_EDIM, "mesh: empty");
    if (m->nf > (1 << 30)) return fail(ctx, MHX_EDIM, "mesh: too many faces");
    for (int64_t i = 0; i < 4 * m->nf; i++)
        if (m->faces[i] < 1 || m->faces[i] > m->nv) return fail(ctx, MHX_EDIM, "mesh: face index out of 1..nv");
    return MHX_OK;
}

int ensure_rankine(mhx_ctx* ctx, mhx_body* b, int direct) {
    if (b->have_rankine[direct]) return MHX_OK;
    const size_t nn = (size_t)b->n * b->n;
    int rc;
    if ((rc = dalloc(ctx, &b->SR, nn))) return rc;
    if ((rc = dalloc(ctx, &b->DR[direct], nn))) return rc;
    {
        Phase ph(ctx, "rankine");
        mhx::launch_rankine(ctx->stream, b->n, b->geom, direct, b->SR, b->DR[direct]);
    }
    HCHK(hipGetLastError());
    b->have_rankine[direct] = true;
    return MHX_OK;
}

void free_work(mhx_body* b) {
    dfree(b->S); dfree(b->D); dfree(b->WG); dfree(b->WR); dfree(b->B); dfree(b->Y); dfree(b->work);
    dfree(b->F); dfree(b->FK); dfree(b->betas); dfree(b->d_ws); dfree(b->d_ks); dfree(b->ipiv); dfree(b->perm);
    dfree(b->wperm); dfree(b->info); dfree(b->status); dfree(b->fb_list); dfree(b->fb_count);
    dfree(b->Df); dfree(b->Ps); dfree(b->Rr); dfree(b->Tr); dfree(b->norms); dfree(b->D0);
    bump_gen(b);
    b->fb_cap = 0;
    b->cap_batch = b->cap_R = b->cap_nb = 0;
    b->cap_table = false;
}

// work buffers for `batch` frequency items with up to R right-hand sides and nb headings
// (with_S = false: no S matrices, for a temporary body that only solves systems)
int ensure_work(mhx_ctx* ctx, mhx_body* b, int R, int nb, bool table, int batch = 1, bool with_S = true) {
    if (batch <= b->cap_batch && R <= b->cap_R && nb <= b->cap_nb && (!table || b->cap_table)) return MHX_OK;
    const int nbt = batch > b->cap_batch ? batch : b->cap_batch;
    const int cr = R > 16 ? R : 16, cnb = nb > 8 ? nb : 8;
    const bool tab = table || b->cap_table;
    free_work(b);
    b->cap_batch = nbt;
    b->cap_R = cr;
    b->cap_nb = cnb;
    b->cap_table = tab;
    const size_t n = b->n, nn = n * n, B = nbt;
    int rc;
    if ((with_S && (rc = dalloc(ctx, &b->S, nn * B))) || (rc = dalloc(ctx, &b->D, nn * B))) return rc;
    if (tab && ((rc = dalloc(ctx, &b->WG, nn * B)) || (rc = dalloc(ctx, &b->WR, nn * B)))) return rc;
    if ((rc = dalloc(ctx, &b->B, n * cr * B)) || (rc = dalloc(ctx, &b->Y, n * cr * B)) ||
        (rc = dalloc(ctx, &b->work, (size_t)b->sW() * B)) || (rc = dalloc(ctx, &b->F, (size_t)(b->ndof ? b->ndof : 1) * cr * B)) ||
        (rc = dalloc(ctx, &b->FK, (size_t)(b->ndof ? b->ndof : 1) * cnb * B)) || (rc = dalloc(ctx, &b->betas, (size_t)cnb)) ||
        (rc = dalloc(ctx, &b->d_ws, B)) || (rc = dalloc(ctx, &b->d_ks, B)) || (rc = dalloc(ctx, &b->ipiv, n * B)) ||
        (rc = dalloc(ctx, &b->perm, n * B)) || (rc = dalloc(ctx, &b->wperm, (size_t)mhx::kPermWork * B)) ||
        (rc = dalloc(ctx, &b->info, B)) || (rc = dalloc(ctx, &b->status, 4))) {
        free_work(b);
        return rc;
    }
    return MHX_OK;
}

bool use_mixed(const mhx_ctx* ctx, int gf) {
    return ctx->lu_mode == MHX_LU_MIXED || ctx->lu_mode == MHX_LU_MIXED_BF16 ||
           (ctx->lu_mode == MHX_LU_AUTO && gf == MHX_GF_SURROGATE);
}
// split-bf16 trailing updates: asked for, or the default precision of the surrogate path
bool use_split(const mhx_ctx* ctx, int gf) {
    return ctx->lu_mode == MHX_LU_MIXED_BF16 || (ctx->lu_mode == MHX_LU_AUTO && gf == MHX_GF_SURROGATE);
}

// device bytes one batch item of a solve holds (the chunk size of the batched schedules)
double item_bytes(const mhx_ctx* ctx, int n, int gf, bool table, int rhs_cols) {
    double per = (double)n * n * sizeof(cplx) * (table ? 4 : 2) + (double)rhs_cols * n * 16 * sizeof(cplx);
    if (gf != MHX_GF_WU) per += 8.0 * (double)mhx::surr_fallback_capacity(n, 1);  // pair list
    if (use_mixed(ctx, gf)) per += (double)n * n * sizeof(cplxf) + 2.0 * n * 16 * sizeof(cplx);  // FP32 factors
    else if (ctx->pivoting == MHX_PIVOT_VERIFIED) per += (double)n * n * sizeof(cplx);  // copy of D
    if (use_split(ctx, gf)) per += 2.0 * mhx::split_work(n);  // split-bf16 panels
    return per;
}

// the body's pair list (one entry per unordered pair of each of `batch` items): the surrogate's
// exact-fallback pairs, or the exact term's pairs left for adaptive refinement
int ensure_pair_list(mhx_ctx* ctx, mhx_body* b, int batch) {
    const int64_t cap = mhx::surr_fallback_capacity(b->n, batch);
    if (b->fb_cap >= cap) return MHX_OK;
    int rc;
    dfree(b->fb_list);
    bump_gen(b);
    b->fb_list = nullptr;
    b->fb_cap = 0;
    if (!b->fb_count && (rc = dalloc(ctx, &b->fb_count, 1))) return rc;
    if ((rc = dalloc(ctx, &b->fb_list, (size_t)cap))) return rc;
    b->fb_cap = cap;
    return MHX_OK;
}

// wave assembly of `batch` items (wavenumbers already in b->d_ks) into b->S, b->D.  keep_table:
// the surrogate writes its (G, dG/dr) table into b->WG, b->WR (allocated by the caller) and S, D
// are combined from it, as for the exact term, so that later kernels can read dG/dr.
int assemble_into(mhx_ctx* ctx, mhx_body* b, int batch, int gf, int direct, int include_identity,
                  bool keep_table = false, bool sd32_ok = false) {
    b->df_ready = false;
    b->sd32 = false;
    if (gf == MHX_GF_WU) {
        Phase ph(ctx, "wave");
        mhx::launch_combine_wu(ctx->stream, b->n, b->geom, b->SR, b->DR[direct], b->d_ks, batch, direct,
                               include_identity, b->S, b->D, b->nn());
    } else if (gf == MHX_GF_SURROGATE) {
        if (!ctx->surr_ready) return fail(ctx, MHX_EINVAL, "surrogate: no trained network (train or load one)");
        int rc;
        if ((rc = ensure_pair_list(ctx, b, batch))) return rc;
        // fused: the pair kernel writes S, D (and the FP32 copy of D for the mixed-precision solver)
        mhx::SurrOut out;
        out.SR = b->SR;
        out.DR = b->DR[direct];
        out.direct = direct;
        out.include_identity = include_identity;
        out.sMat = b->nn();
        if (keep_table) {
            out.WG = b->WG;
            out.WR = b->WR;
        } else {
            out.S = b->S;
            out.D = b->D;
        }
        if (!keep_table && use_mixed(ctx, gf)) {
            int rc = dalloc(ctx, &b->Df, (size_t)b->nn() * b->cap_batch);
            if (rc) return rc;
            out.Df = b->Df;
            // the caller solves with the mixed-precision solver: S and D in FP32 (their entries carry
            // the network's BF16-level error; the residual products then read half the bytes)
            out.sd32 = sd32_ok ? 1 : 0;
            b->sd32 = sd32_ok;
        }
        mhx::launch_zero(ctx->stream, b->status, 4 * sizeof(unsigned long long));
        {
            Phase ph(ctx, "wave_table");
            mhx::launch_surr_table(ctx->stream, b->n, b->geom, b->d_ks, batch, ctx->surr_W, ctx->surr_bias,
                                   ctx->surr_B, out, b->fb_list, b->fb_count, b->fb_cap, b->status, ctx->surr_act);
        }
        b->df_ready = out.Df != nullptr;
        if (keep_table) {
            Phase ph(ctx, "wave");
            mhx::launch_combine_table(ctx->stream, b->n, b->geom, b->SR, b->DR[direct], b->d_ks, batch, direct,
                                      include_identity, b->WG, b->WR, b->S, b->D, b->nn());
        }
    } else {
        int rc;
        if ((rc = ensure_pair_list(ctx, b, batch))) return rc;
        mhx::launch_zero(ctx->stream, b->status, 4 * sizeof(unsigned long long));
        {
            Phase ph(ctx, "wave_table");
            mhx::launch_delhommeau_table(ctx->stream, b->n, b->geom, b->d_ks, batch, 1e-8, 1e-8, b->WG, b->WR, b->nn(),
                                         b->fb_list, b->fb_count, b->fb_cap, b->status);
        }
        Phase ph(ctx, "wave");
        mhx::launch_combine_table(ctx->stream, b->n, b->geom, b->SR, b->DR[direct], b->d_ks, batch, direct,
                                  include_identity, b->WG, b->WR, b->S, b->D, b->nn());
    }
    HCHK(hipGetLastError());
    return MHX_OK;
}

int check_status(mhx_ctx* ctx, mhx_body* b, int gf) {
    if (gf == MHX_GF_WU) return MHX_OK;
    unsigned long long st[4] = {0, 0, 0, 0};
    HCHK(hipMemcpyAsync(st, b->status, sizeof st, hipMemcpyDeviceToHost, ctx->stream));
    HCHK(hipStreamSynchronize(ctx->stream));
    for (int q = 0; q < 3; q++) ctx->quad_work[q] += st[q + 1];
    if (st[0]) return fail(ctx, MHX_EQUAD, "exact Delhommeau quadrature exceeded its interval budget");
    return MHX_OK;
}

// factor the `batch` matrices b->D in place and solve for the R columns of b->B; phi into b->Y
int check_info(mhx_ctx* ctx, mhx_body* b, int batch, bool* singular) {
    std::vector<int> info(batch, 0);
    HCHK(hipMemcpyAsync(info.data(), b->info, sizeof(int) * batch, hipMemcpyDeviceToHost, ctx->stream));
    HCHK(hipStreamSynchronize(ctx->stream));
    *singular = false;
    for (int q = 0; q < batch; q++)
        if (info[q]) {
            *singular = true;
            ctx->err = "LU: exactly zero pivot at column " + std::to_string(info[q]);
        }
    return MHX_OK;
}

// Per-item flags of a verified factorization (nonzero: LAPACK's rule failed somewhere, the item
// must be factored again with row interchanges); the flag of item q is wperm[kPermWork q].
int read_pivot_flags(mhx_ctx* ctx, const int* wperm, int batch, std::vector<int>& flags) {
    flags.assign(batch, 0);
    HCHK(hipMemcpy2DAsync(flags.data(), sizeof(int), wperm, sizeof(int) * mhx::kPermWork, sizeof(int), batch,
                          hipMemcpyDeviceToHost, ctx->stream));
    HCHK(hipStreamSynchronize(ctx->stream));
    return MHX_OK;
}

// FP64 LU of D (in place) with the context's pivoting.  Verified diagonal pivots: D is copied to
// D0 first, and every item whose factorization breaks LAPACK's rule is copied back and factored
// again with row interchanges (so the pivots are LAPACK's either way).
// Y = M X for a batch's S or D (M = b->S or b->D), in the precision the assembly stored them in
void matmul_sd(hipStream_t st, const mhx_body* b, int R, const cplx* M, const cplx* X, cplx* Y, int batch) {
    if (b->sd32)
        mhx::launch_matmul_nr(st, b->n, R, reinterpret_cast<const cplxf*>(M), 2 * b->nn(), X, b->nR(), Y, b->nR(),
                              b->work, b->sW(), batch);
    else
        mhx::launch_matmul_nr(st, b->n, R, M, b->nn(), X, b->nR(), Y, b->nR(), b->work, b->sW(), batch);
}

// the FP32 working copy Df of items [q0, q0 + batch) of D: converted from the FP64 D, or copied from
// its FP32 storage
int df_from_d(mhx_ctx* ctx, mhx_body* b, int64_t q0, int batch) {
    const int64_t nn = b->nn();
    if (!b->sd32) {
        mhx::convert_to_f32(ctx->stream, nn * batch, b->D + q0 * nn, b->Df + q0 * nn);
        return MHX_OK;
    }
    HCHK(hipMemcpy2DAsync(b->Df + q0 * nn, sizeof(cplxf) * nn, reinterpret_cast<const cplxf*>(b->D + q0 * nn),
                          sizeof(cplx) * nn, sizeof(cplxf) * nn, batch, hipMemcpyDeviceToDevice, ctx->stream));
    return MHX_OK;
}

// S and D of a batch stored in FP32 widened back to FP64 in place (through D0), for the FP64 path
int widen_sd(mhx_ctx* ctx, mhx_body* b, int batch) {
    if (!b->sd32) return MHX_OK;
    int rc;
    if ((rc = dalloc(ctx, &b->D0, (size_t)b->nn() * b->cap_batch))) return rc;
    for (cplx* m : {b->S, b->D}) {
        if (!m) continue;
        mhx::widen_to_f64(ctx->stream, b->nn(), batch, reinterpret_cast<const cplxf*>(m), 2 * b->nn(), b->D0, b->nn());
        HCHK(hipMemcpyAsync(m, b->D0, sizeof(cplx) * b->nn() * batch, hipMemcpyDeviceToDevice, ctx->stream));
    }
    b->sd32 = false;
    return MHX_OK;
}

int getrf_fp64(mhx_ctx* ctx, mhx_body* b, int batch, mhx::LaunchTimer* lt) {
    const int n = b->n;
    hipStream_t st = ctx->stream;
    if (ctx->pivoting != MHX_PIVOT_VERIFIED) {
        if (mhx::getrf_batched(st, n, b->D, n, b->nn(), batch, b->ipiv, n, b->wperm, b->info, lt))
            return fail(ctx, MHX_EDIM, "LU: order above 16384 is not supported");
        ctx->lu_interchanged += batch;
        return MHX_OK;
    }
    int rc;
    if ((rc = dalloc(ctx, &b->D0, (size_t)b->nn() * b->cap_batch))) return rc;
    HCHK(hipMemcpyAsync(b->D0, b->D, sizeof(cplx) * b->nn() * batch, hipMemcpyDeviceToDevice, st));
    if (mhx::getrf_batched(st, n, b->D, n, b->nn(), batch, b->ipiv, n, b->wperm, b->info, lt, mhx::kPivotVerified,
                           b->work, b->sW()))
        return fail(ctx, MHX_EDIM, "LU: order above 16384 is not supported");
    std::vector<int> fl;
    if ((rc = read_pivot_flags(ctx, b->wperm, batch, fl))) return rc;
    for (int q = 0; q < batch; q++) {
        if (!fl[q]) {
            ctx->lu_diagonal++;
            continue;
        }
        ctx->lu_interchanged++;
        HCHK(hipMemcpyAsync(b->D + q * b->nn(), b->D0 + q * b->nn(), sizeof(cplx) * b->nn(), hipMemcpyDeviceToDevice, st));
        mhx::getrf_batched(st, n, b->D + q * b->nn(), n, b->nn(), 1, b->ipiv + (size_t)q * n, n,
                           b->wperm + (size_t)q * mhx::kPermWork, b->info + q, lt);
    }
    return MHX_OK;
}

// FP64 LU of D (in place) and the solves (rhs in b->B; solution potentials in b->Y).
// form_phi = false (indirect only): the system's solution stays in b->B and S is not used;
// trans = 2 then solves with D^H.
int solve_fp64(mhx_ctx* ctx, mhx_body* b, int R, int direct, int batch, bool form_phi = true, int trans = 0) {
    const int n = b->n;
    if (b->sd32) {
        const int rc = widen_sd(ctx, b, batch);
        if (rc) return rc;
    }
    mhx::launch_zero(ctx->stream, b->info, sizeof(int) * batch);
    {
        Phase ph(ctx, "lu");
        mhx::LaunchTimer lt;
        int rc = getrf_fp64(ctx, b, batch, ctx->timing ? &lt : nullptr);
        take_launches(ctx, lt);
        if (rc) return rc;
    }
    HCHK(hipGetLastError());
    {
        Phase ph(ctx, "solve");
        cplx* W2 = b->work + 16 * b->nR();
        if (direct) {
            mhx::launch_matmul_nr(ctx->stream, n, R, b->S, b->nn(), b->B, b->nR(), b->Y, b->nR(), b->work, b->sW(), batch);
            mhx::getrs_batched(ctx->stream, 0, n, R, b->D, n, b->nn(), b->ipiv, n, b->perm, n, b->Y, b->nR(), W2,
                               b->sW(), batch);
        } else {
            mhx::getrs_batched(ctx->stream, trans, n, R, b->D, n, b->nn(), b->ipiv, n, b->perm, n, b->B, b->nR(), W2,
                               b->sW(), batch);
            if (form_phi)
                mhx::launch_matmul_nr(ctx->stream, n, R, b->S, b->nn(), b->B, b->nR(), b->Y, b->nR(), b->work, b->sW(),
                                      batch);
        }
    }
    HCHK(hipGetLastError());
    bool singular = false;
    int rc = check_info(ctx, b, batch, &singular);
    if (rc) return rc;
    return singular ? MHX_ESINGULAR : MHX_OK;
}

// Mixed precision: FP32 LU of a copy of D, solves with the FP32 factors in FP64 arithmetic,
// FP64 residuals against the FP64 matrix and iterative refinement until the relative
// residual |r| / |rhs| <= kRefineTol on every item.  Items that do not converge in
// kRefineMax steps (or whose FP32 factor is singular) are redone with the FP64 LU.
constexpr double kRefineTol = 1e-12;
constexpr int kRefineMax = 4;       // FP32 trailing updates
constexpr int kRefineMaxSplit = 8;  // split-bf16 trailing updates (factors to ~2^-16)

// Parts a batched FP32 factorization and its solves are split into, one stream each: a batch of 8
// or more items of order >= 512 in up to ctx->lu_streams parts of at least 4 items.
int lu_parts(const mhx_ctx* ctx, int n, int batch) {
    return (batch >= 8 && n >= 512) ? std::max(1, std::min(ctx->lu_streams, batch / 4)) : 1;
}

// fn(stream, first, count) for the items [first, first + count) of each of `parts` parts of a batch:
// part 0 on the library's stream, part p on side[p - 1] (default priority), forked from and joined
// back to the library's stream by events.  The parts share nothing but the device.
template <class F>
int on_parts(mhx_ctx* ctx, int parts, int batch, F fn) {
    if (parts > 1) HCHK(hipEventRecord(ctx->fork, ctx->stream));
    int rc = 0;
    for (int p = parts - 1; p >= 0 && !rc; p--) {
        const int first = (int)((int64_t)batch * p / parts), cnt = (int)((int64_t)batch * (p + 1) / parts) - first;
        hipStream_t st = p == 0 ? ctx->stream : ctx->side[p - 1];
        if (p > 0) HCHK(hipStreamWaitEvent(st, ctx->fork, 0));
        rc = fn(st, first, cnt);
        if (p > 0) HCHK(hipEventRecord(ctx->join[p - 1], st));
    }
    for (int p = 1; p < parts; p++) HCHK(hipStreamWaitEvent(ctx->stream, ctx->join[p - 1], 0));
    return rc;
}

// FP32 LU of the batch's b->Df in place, in lu_parts parts on their streams: the panel phases of
// one part (launches of one workgroup per item, latency-bound) run beside the trailing updates of
// the others, on compute units the panels leave idle.
int getrf_f32_two_streams(mhx_ctx* ctx, mhx_body* b, int batch, mhx::LaunchTimer* lt, unsigned short* split,
                          int pivot, cplx* inv64 = nullptr) {
    const int n = b->n;
    return on_parts(ctx, lu_parts(ctx, n, batch), batch, [&](hipStream_t st, int first, int cnt) {
        return mhx::getrf_batched(st, n, b->Df + (size_t)first * b->nn(), n, b->nn(), cnt, b->ipiv + (size_t)first * n,
                                  n, b->wperm + (size_t)first * mhx::kPermWork, b->info + first, lt,
                                  split ? split + (size_t)first * mhx::split_work(n) : nullptr, pivot,
                                  reinterpret_cast<cplxf*>(b->work + (size_t)first * b->sW()), 2 * b->sW(),
                                  inv64 ? inv64 + (size_t)first * b->sW() : nullptr, b->sW(), ctx->lu_super);
    });
}

// The LU tuner: on the first batched FP32 factorization of a signature (order, batch, update kind,
// pivoting) it factors the batch's own matrices (FP32 copies made again from the intact FP64 D
// before every trial) with 1-3 streams and 2-8 panels per split update, twice each, timed with
// events on the library's stream, and keeps the fastest for every later call of that signature.  Results do not
// depend on the choice beyond the FP32 rounding of the factors (the refinement takes the solutions
// to the same tolerance).  Returns whether Df still holds the FP32 copy of D.
bool tune_lu(mhx_ctx* ctx, mhx_body* b, int batch, unsigned short* split, int pivot, cplx* inv64) {
    const int n = b->n;
    if (batch < 8 || n < 512) return true;
    const std::string key = "lu:" + std::to_string(n) + ":" + std::to_string(batch) + ":" + (split ? "s" : "f") + ":" +
                            std::to_string(pivot) + (ctx->lu_streams_fixed ? ":fixed" + std::to_string(ctx->lu_streams) : "");
    auto f = ctx->tuned.find(key);
    if (f != ctx->tuned.end()) {
        if (!ctx->lu_streams_fixed) ctx->lu_streams = f->second / 16;
        ctx->lu_super = f->second % 16;
        return true;
    }
    if (!ctx->autotune) return true;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return true;
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return true;
    }
    const int keep_streams = ctx->lu_streams;
    // two rounds over the candidates, each configuration keeping its shorter time, so that a
    // trial slowed by clock ramp-up or a first launch does not decide
    std::map<int, float> trial_ms;
    for (int round = 0; round < 2; round++)
        for (int streams = 1; streams <= 3; streams++) {
            if (ctx->lu_streams_fixed && streams != keep_streams) continue;
            if (streams > 1 && batch / streams < 4) continue;
            for (int sp = split ? 2 : mhx::kSplitSuperMax; sp <= mhx::kSplitSuperMax; sp++) {
                ctx->lu_streams = streams;
                ctx->lu_super = sp;
                if (df_from_d(ctx, b, 0, batch)) continue;
                hipEventRecord(e0, ctx->stream);
                getrf_f32_two_streams(ctx, b, batch, nullptr, split, pivot, inv64);
                hipEventRecord(e1, ctx->stream);
                if (hipEventSynchronize(e1) != hipSuccess) continue;
                float ms = 0.f;
                hipEventElapsedTime(&ms, e0, e1);
                const int cfg = 16 * streams + sp;
                auto t = trial_ms.find(cfg);
                if (t == trial_ms.end() || ms < t->second) trial_ms[cfg] = ms;
            }
        }
    int best = -1;
    float best_ms = 0.f;
    for (const auto& t : trial_ms)
        if (best < 0 || t.second < best_ms) {
            best = t.first;
            best_ms = t.second;
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (best < 0) {
        ctx->lu_streams = keep_streams;
        ctx->lu_super = mhx::kSplitSuperDefault;
        return false;
    }
    ctx->tuned[key] = best;
    ctx->lu_streams = ctx->lu_streams_fixed ? keep_streams : best / 16;
    ctx->lu_super = best % 16;
    return false;
}

int solve_mixed(mhx_ctx* ctx, mhx_body* b, int R, int direct, int batch, int gf, bool form_phi = true) {
    const int n = b->n;
    hipStream_t st = ctx->stream;
    int rc;
    const size_t nRb = (size_t)b->nR() * batch;
    if ((rc = dalloc(ctx, &b->Df, (size_t)b->nn() * b->cap_batch)) || (rc = dalloc(ctx, &b->Rr, (size_t)b->nR() * b->cap_batch)) ||
        (rc = dalloc(ctx, &b->Tr, (size_t)b->nR() * b->cap_batch)) || (rc = dalloc(ctx, &b->norms, 2 * (size_t)b->cap_batch)))
        return rc;
    const bool split = use_split(ctx, gf);
    if (split && (rc = dalloc(ctx, &b->Ps, (size_t)mhx::split_work(n) * b->cap_batch))) return rc;
    const int refine_max = ctx->refine_limit >= 0 ? ctx->refine_limit : split ? kRefineMaxSplit : kRefineMax;
    bool inverses_ready = true;  // the verified factorization wrote the solves' diagonal-block inverses
    mhx::launch_zero(st, b->info, sizeof(int) * batch);
    {
        const bool verified = ctx->pivoting == MHX_PIVOT_VERIFIED;
        cplx* inv64 = verified ? mhx::getrs_inverses(b->work + 16 * b->nR(), n, R) : nullptr;
        if (!tune_lu(ctx, b, batch, split ? b->Ps : nullptr, verified ? mhx::kPivotVerified : mhx::kPivotStrict,
                     inv64))
            b->df_ready = false;  // the tuner's trials factored Df
        ctx->last_streams = lu_parts(ctx, n, batch);
        ctx->last_super = split ? ctx->lu_super : 1;
        Phase ph(ctx, "lu");
        if (!b->df_ready && (rc = df_from_d(ctx, b, 0, batch))) return rc;
        b->df_ready = false;
        mhx::LaunchTimer lt;
        // with verified pivots the factorization also leaves the FP64 diagonal-block inverses in
        // the solves' tables (work after the split-K partials, for R right-hand sides)
        rc = getrf_f32_two_streams(ctx, b, batch, ctx->timing ? &lt : nullptr, split ? b->Ps : nullptr,
                                   verified ? mhx::kPivotVerified : mhx::kPivotStrict, inv64);
        if (rc) {
            take_launches(ctx, lt);
            return fail(ctx, MHX_EDIM, "LU: order above 16384 is not supported");
        }
        // verified diagonal pivots: items that broke LAPACK's rule are converted again from the
        // FP64 D (intact) and factored with row interchanges
        if (verified) {
            std::vector<int> fl;
            if ((rc = read_pivot_flags(ctx, b->wperm, batch, fl))) return rc;
            for (int q = 0; q < batch; q++) {
                if (!fl[q]) {
                    ctx->lu_diagonal++;
                    continue;
                }
                ctx->lu_interchanged++;
                inverses_ready = false;  // the refactored item's blocks differ from the verified pass
                if ((rc = df_from_d(ctx, b, q, 1))) return rc;
                mhx::getrf_batched(st, n, b->Df + q * b->nn(), n, b->nn(), 1, b->ipiv + (size_t)q * n, n,
                                   b->wperm + (size_t)q * mhx::kPermWork, b->info + q, ctx->timing ? &lt : nullptr,
                                   split ? b->Ps + (size_t)q * mhx::split_work(n) : nullptr);
            }
        } else {
            ctx->lu_interchanged += batch;
            inverses_ready = false;
        }
        take_launches(ctx, lt);
    }
    HCHK(hipGetLastError());
    bool singular = false;
    if ((rc = check_info(ctx, b, batch, &singular))) return rc;
    bool converged = false;
    if (!singular) {
        Phase ph(ctx, "solve");
        cplx* W2 = b->work + 16 * b->nR();
        cplx* X = direct ? b->Y : b->B;  // the system's unknowns, solved in place
        if (direct) matmul_sd(st, b, R, b->S, b->B, b->Y, batch);
        HCHK(hipMemcpyAsync(b->Rr, X, sizeof(cplx) * nRb, hipMemcpyDeviceToDevice, st));
        mhx::getrs_batched(st, 0, n, R, b->Df, n, b->nn(), b->ipiv, n, b->perm, n, X, b->nR(), W2, b->sW(), batch,
                           inverses_ready ? mhx::kGetrsInverses : mhx::kGetrsFresh);
        std::vector<double> nr(2 * batch);
        for (int it = 0; it <= refine_max; it++) {
            matmul_sd(st, b, R, b->D, X, b->Tr, batch);
            mhx::launch_zero(st, b->norms, sizeof(double) * 2 * batch);
            mhx::launch_residual(st, (int64_t)n * R, b->Rr, b->nR(), b->Tr, b->nR(), b->norms, batch);
            HCHK(hipMemcpyAsync(nr.data(), b->norms, sizeof(double) * 2 * batch, hipMemcpyDeviceToHost, st));
            HCHK(hipStreamSynchronize(st));
            double worst = 0.0;
            for (int q = 0; q < batch; q++) {
                const double rel = nr[2 * q + 1] > 0 ? std::sqrt(nr[2 * q] / nr[2 * q + 1]) : std::sqrt(nr[2 * q]);
                worst = std::max(worst, std::isfinite(rel) ? rel : 1e300);
            }
            if (worst <= kRefineTol) {
                converged = true;
                ctx->max_residual = std::max(ctx->max_residual, worst);
                break;
            }
            if (it == refine_max) break;
            mhx::getrs_batched(st, 0, n, R, b->Df, n, b->nn(), b->ipiv, n, b->perm, n, b->Tr, b->nR(), W2, b->sW(), batch,
                               mhx::kGetrsPrepared);
            mhx::launch_add(st, (int64_t)n * R, b->Tr, b->nR(), X, b->nR(), batch);
            ctx->refine_steps += batch;
        }
        if (converged && !direct && form_phi) matmul_sd(st, b, R, b->S, b->B, b->Y, batch);
        HCHK(hipGetLastError());
    }
    if (converged) {
        ctx->mixed_solves += batch;
        return MHX_OK;
    }
    // FP64 fallback (D is intact): restore the right-hand side and factor in FP64
    ctx->fp64_fallbacks += batch;
    if (!singular && !direct) HCHK(hipMemcpyAsync(b->B, b->Rr, sizeof(cplx) * nRb, hipMemcpyDeviceToDevice, st));
    return solve_fp64(ctx, b, R, direct, batch, form_phi);
}

int factor_and_solve(mhx_ctx* ctx, mhx_body* b, int R, int direct, int batch = 1, int gf = MHX_GF_WU) {
    int rc = use_mixed(ctx, gf) ? solve_mixed(ctx, b, R, direct, batch, gf) : solve_fp64(ctx, b, R, direct, batch);
    if (rc == MHX_ESINGULAR) return fail(ctx, MHX_ESINGULAR, ctx->err);
    return rc;
}

int set_freqs(mhx_ctx* ctx, mhx_body* b, const double* ws, int batch) {
    std::vector<double> h(2 * batch);
    for (int q = 0; q < batch; q++) {
        h[q] = ws[q];
        h[batch + q] = ws[q] * ws[q] / ctx->g;  // compute_wavenumber (src/waves.jl:133-134)
    }
    HCHK(hipMemcpyAsync(b->d_ws, h.data(), sizeof(double) * batch, hipMemcpyHostToDevice, ctx->stream));
    HCHK(hipMemcpyAsync(b->d_ks, h.data() + batch, sizeof(double) * batch, hipMemcpyHostToDevice, ctx->stream));
    HCHK(hipStreamSynchronize(ctx->stream));
    return MHX_OK;
}

int body_upload(mhx_ctx* ctx, const mhx_mesh* m, int64_t ndof, const double* dofs, mhx_body** out) {
    int rc = check_mesh(ctx, m);
    if (rc) return rc;
    if (ndof < 0 || (ndof > 0 && !dofs)) return fail(ctx, MHX_EINVAL, "dofs: null with ndof > 0");
    mhx_body* b = new mhx_body();
    b->ctx = ctx;
    b->n = (int)m->nf;
    b->ndof = (int)ndof;
    const size_t n = m->nf, nv = m->nv;
    double *dv = nullptr, *dc = nullptr, *dn = nullptr, *da = nullptr, *dr = nullptr;
    int64_t* df = nullptr;
    auto cleanup = [&]() { dfree(dv); dfree(dc); dfree(dn); dfree(da); dfree(dr); dfree(df); };
    if ((rc = dalloc(ctx, &dv, 3 * nv)) || (rc = dalloc(ctx, &df, 4 * n)) || (rc = dalloc(ctx, &dc, 3 * n)) ||
        (rc = dalloc(ctx, &dn, 3 * n)) || (rc = dalloc(ctx, &da, n)) || (rc = dalloc(ctx, &dr, n)) ||
        (rc = dalloc(ctx, &b->geom, n)) || (rc = dalloc(ctx, &b->dofs, 3 * n * (ndof ? ndof : 1)))) {
        cleanup();
        mhx_body_destroy(b);
        return rc;
    }
    hipStream_t st = ctx->stream;
    hipMemcpyAsync(dv, m->vertices, sizeof(double) * 3 * nv, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(df, m->faces, sizeof(int64_t) * 4 * n, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dc, m->centers, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dn, m->normals, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(da, m->areas, sizeof(double) * n, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dr, m->radii, sizeof(double) * n, hipMemcpyHostToDevice, st);
    if (ndof) hipMemcpyAsync(b->dofs, dofs, sizeof(double) * 3 * n * ndof, hipMemcpyHostToDevice, st);
    mhx::launch_panel_geometry(st, (int)n, dv, df, dc, dn, da, dr, b->geom);
    hipError_t e = hipStreamSynchronize(st);
    cleanup();
    if (e != hipSuccess || (e = hipGetLastError()) != hipSuccess) {
        mhx_body_destroy(b);
        return fail(ctx, MHX_EDEVICE, std::string("body upload: ") + hipGetErrorString(e));
    }
    *out = b;
    return MHX_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// scratch of mhx_lu_solve_noalloc: a file-scope device array, one copy per GPU (each GPU holds
// its own copy at its own address until the process ends).  Host code never passes the array's
// name to a kernel or copy: it uses the address hipGetSymbolAddress returns on each device.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[MHX_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch;  // device ordinal -> address of `scratch` on it

// once per device, with that device current (called by mhx_create)
unsigned char* scratch_address(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch.find(device);
    if (it != g_scratch.end()) return it->second;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
    }
    g_scratch[device] = (unsigned char*)p;
    return (unsigned char*)p;
}

}  // namespace

// ---------------------------------------------------------------------------
// multi-body sweep
// ---------------------------------------------------------------------------
struct SweepState {
    std::vector<mhx_body*> bodies;
    std::vector<int64_t> nws;        // frequencies per body
    std::vector<int> per_chunk;      // items per chunk per body
    int K = 1;                       // chunks
    int64_t nb = 0;
    int gf = 0, direct = 0;
    double rho = 0.0, g = 0.0;       // constants the captured kernels carry by value
    std::vector<uint64_t> gens;      // buffer generation of each body at capture
    std::vector<double*> w_all, k_all;  // device, per body: the sweep's frequencies and wavenumbers
    std::vector<cplx*> F_all, FK_all;   // device, per body: forces of every frequency
    std::vector<mhx::SweepItem> h_items;
    mhx::SweepItem* items = nullptr;    // device, all chunks
    std::vector<int> chunk_first, chunk_count;
    std::vector<int64_t> chunk_units, chunk_tunits;
    unsigned long long* counters = nullptr;  // table-queue counters, one per chunk
    unsigned long long* status = nullptr;    // quadrature status and work tally of the sweep
    int* sticky = nullptr;                   // first zero pivot per body
    bool verified = false;                   // LU with verified diagonal pivots
    int* pflags = nullptr;                   // their per-item flags: body ib, frequency w at flag_off[ib] + w
    std::vector<int64_t> flag_off;
    int rows = 256;                          // tuner's choice of rows per combine workgroup
    bool mixed = false, split = false;       // FP32 LU (split-bf16 updates) with one refinement step
    double* rnorms = nullptr;                // mixed: |r|^2, |b|^2 after the refinement, per (body, frequency)
    std::vector<int64_t> chunk_sunits;       // surrogate units per chunk
    unsigned long long *fb_list = nullptr, *fb_count = nullptr;  // surrogate: exact-fallback pairs
    int64_t fb_cap = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void release() {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        for (auto& p : w_all) dfree(p);
        for (auto& p : k_all) dfree(p);
        for (auto& p : F_all) dfree(p);
        for (auto& p : FK_all) dfree(p);
        dfree(items);
        dfree(counters);
        dfree(status);
        dfree(sticky);
        dfree(pflags);
        dfree(rnorms);
        dfree(fb_list);
        dfree(fb_count);
    }
};

namespace {

void drop_sweep(mhx_ctx* c) {
    if (!c->sweep) return;
    c->sweep->release();
    delete c->sweep;
    c->sweep = nullptr;
}

bool sweep_uses(mhx_ctx* c, const mhx_body* b) {
    if (!c->sweep) return false;
    for (auto* x : c->sweep->bodies)
        if (x == b) return true;
    return false;
}

// Everything of one sweep on the stream, with no host synchronisation and no memset nodes, so the
// sequence can be captured into a graph: frequencies from the per-body arrays, right-hand sides,
// one table launch and one combine launch per chunk for all bodies, then per body the batched
// LU, the solves and the pressure integration into the per-body force arrays.
void sweep_enqueue(mhx_ctx* ctx, SweepState& s) {
    hipStream_t st = ctx->stream;
    const bool exact = s.gf == MHX_GF_DELHOMMEAU, surr = s.gf == MHX_GF_SURROGATE;
    const double g = ctx->g;
    mhx::launch_zero(st, s.status, 4 * sizeof(unsigned long long));
    mhx::launch_zero(st, s.sticky, s.bodies.size() * sizeof(int));
    for (int c = 0; c < s.K; c++) {
        for (size_t ib = 0; ib < s.bodies.size(); ib++) {
            mhx_body* b = s.bodies[ib];
            const int64_t w0 = (int64_t)c * s.per_chunk[ib];
            const int bt = (int)std::max<int64_t>(0, std::min<int64_t>(s.per_chunk[ib], s.nws[ib] - w0));
            if (!bt) continue;
            hipMemcpyAsync(b->d_ws, s.w_all[ib] + w0, sizeof(double) * bt, hipMemcpyDeviceToDevice, st);
            hipMemcpyAsync(b->d_ks, s.k_all[ib] + w0, sizeof(double) * bt, hipMemcpyDeviceToDevice, st);
            mhx::launch_bc(st, b->n, b->geom, b->dofs, b->ndof, b->betas, (int)s.nb, b->d_ws, b->d_ks, bt, g, b->B,
                           b->nR());
        }
        const mhx::SweepItem* it = s.items + s.chunk_first[c];
        if (exact) {
            mhx::launch_zero(st, s.counters + c, sizeof(unsigned long long));
            mhx::launch_queue_table(st, it, s.chunk_count[c], s.chunk_tunits[c], s.counters + c, 1e-8, 1e-8, s.status);
        }
        if (surr)  // the fused pair kernel writes S, D and the FP32 copy of D
            mhx::launch_queue_surr(st, it, s.chunk_count[c], s.chunk_sunits[c], s.direct, 1, ctx->surr_W,
                                   ctx->surr_bias, ctx->surr_B, s.fb_list, s.fb_count, s.fb_cap, s.status,
                                   ctx->surr_act);
        else
            mhx::launch_queue_combine(st, it, s.chunk_count[c], s.chunk_units[c], exact, s.direct, 1, s.rows);
        for (size_t ib = 0; ib < s.bodies.size(); ib++) {
            mhx_body* b = s.bodies[ib];
            const int64_t w0 = (int64_t)c * s.per_chunk[ib];
            const int bt = (int)std::max<int64_t>(0, std::min<int64_t>(s.per_chunk[ib], s.nws[ib] - w0));
            if (!bt) continue;
            const int n = b->n, R = b->ndof + (int)s.nb;
            mhx::launch_zero(st, b->info, sizeof(int) * bt);
            if (s.mixed) {
                // FP32 LU (verified diagonal pivots, their flags), the solves with the FP32 factors,
                // one refinement step against the FP64 D and the residual norms for the host's check
                if (!surr) mhx::convert_to_f32(st, b->nn() * bt, b->D, b->Df);
                getrf_f32_two_streams(ctx, b, bt, nullptr, s.split ? b->Ps : nullptr,
                                      s.verified ? mhx::kPivotVerified : mhx::kPivotStrict,
                                      s.verified ? mhx::getrs_inverses(b->work + 16 * b->nR(), n, R) : nullptr);
                if (s.verified) mhx::launch_gather_pivot_flags(st, b->wperm, bt, s.pflags + s.flag_off[ib] + w0);
                mhx::launch_sticky_info(st, b->info, bt, s.sticky + ib);
                cplx* W2 = b->work + 16 * b->nR();
                cplx* X = s.direct ? b->Y : b->B;
                const int64_t nRb = b->nR() * bt;
                if (s.direct) matmul_sd(st, b, R, b->S, b->B, b->Y, bt);
                mhx::launch_copy(st, X, b->Rr, sizeof(cplx) * nRb);
                // (a flagged item's results are replaced by the pivoted redo after the sweep; its
                // verified factors and their inverses are consistent with each other meanwhile)
                mhx::getrs_batched(st, 0, n, R, b->Df, n, b->nn(), b->ipiv, n, b->perm, n, X, b->nR(), W2, b->sW(), bt,
                                   s.verified ? mhx::kGetrsInverses : mhx::kGetrsFresh);
                matmul_sd(st, b, R, b->D, X, b->Tr, bt);
                mhx::launch_zero(st, b->norms, sizeof(double) * 2 * bt);
                mhx::launch_residual(st, (int64_t)n * R, b->Rr, b->nR(), b->Tr, b->nR(), b->norms, bt);
                mhx::getrs_batched(st, 0, n, R, b->Df, n, b->nn(), b->ipiv, n, b->perm, n, b->Tr, b->nR(), W2, b->sW(),
                                   bt, true);
                mhx::launch_add(st, (int64_t)n * R, b->Tr, b->nR(), X, b->nR(), bt);
                double* rn = s.rnorms + 2 * (s.flag_off[ib] + w0);
                matmul_sd(st, b, R, b->D, X, b->Tr, bt);
                mhx::launch_zero(st, rn, sizeof(double) * 2 * bt);
                mhx::launch_residual(st, (int64_t)n * R, b->Rr, b->nR(), b->Tr, b->nR(), rn, bt);
                if (!s.direct) matmul_sd(st, b, R, b->S, b->B, b->Y, bt);
                if (b->ndof > 0)
                    mhx::launch_forces(st, n, b->geom, b->dofs, b->ndof, b->Y, b->nR(), R, ctx->rho, b->d_ws, b->d_ks, bt,
                                       g, b->betas, (int)s.nb, s.F_all[ib] + w0 * b->ndof * R,
                                       s.FK_all[ib] + w0 * b->ndof * (s.nb ? s.nb : 1));
                continue;
            }
            if (s.verified) {
                // the copy of D for a pivoted redo of flagged items (after the sweep), then the
                // factorization with verified diagonal pivots and its flags
                mhx::launch_copy(st, b->D, b->D0, sizeof(cplx) * b->nn() * bt);
                mhx::getrf_batched(st, n, b->D, n, b->nn(), bt, b->ipiv, n, b->wperm, b->info, nullptr,
                                   mhx::kPivotVerified, b->work, b->sW());
                mhx::launch_gather_pivot_flags(st, b->wperm, bt, s.pflags + s.flag_off[ib] + w0);
            } else {
                mhx::getrf_batched(st, n, b->D, n, b->nn(), bt, b->ipiv, n, b->wperm, b->info, nullptr);
            }
            mhx::launch_sticky_info(st, b->info, bt, s.sticky + ib);
            cplx* W2 = b->work + 16 * b->nR();
            if (s.direct) {
                mhx::launch_matmul_nr(st, n, R, b->S, b->nn(), b->B, b->nR(), b->Y, b->nR(), b->work, b->sW(), bt);
                mhx::getrs_batched(st, 0, n, R, b->D, n, b->nn(), b->ipiv, n, b->perm, n, b->Y, b->nR(), W2, b->sW(), bt);
            } else {
                mhx::getrs_batched(st, 0, n, R, b->D, n, b->nn(), b->ipiv, n, b->perm, n, b->B, b->nR(), W2, b->sW(), bt);
                mhx::launch_matmul_nr(st, n, R, b->S, b->nn(), b->B, b->nR(), b->Y, b->nR(), b->work, b->sW(), bt);
            }
            if (b->ndof > 0)
                mhx::launch_forces(st, n, b->geom, b->dofs, b->ndof, b->Y, b->nR(), R, ctx->rho, b->d_ws, b->d_ks, bt, g,
                                   b->betas, (int)s.nb, s.F_all[ib] + w0 * b->ndof * R,
                                   s.FK_all[ib] + w0 * b->ndof * (s.nb ? s.nb : 1));
        }
    }
}

// rows per combine workgroup: time both shapes once per workload signature on the first chunk
// (the outputs are rewritten by the sweep itself) and keep the faster
int sweep_tune(mhx_ctx* ctx, SweepState& s, const std::string& key) {
    auto f = ctx->tuned.find(key);
    if (f != ctx->tuned.end()) return f->second;
    hipStream_t st = ctx->stream;
    const bool exact = s.gf == MHX_GF_DELHOMMEAU;
    hipEvent_t e[3];
    for (auto& x : e) hipEventCreate(&x);
    const int cand[2] = {128, 256};
    float ms[2] = {0.f, 0.f};
    for (int v = 0; v < 2; v++) {  // warm, then time
        mhx::launch_queue_combine(st, s.items, s.chunk_count[0], s.chunk_units[0], exact, s.direct, 1, cand[v]);
        hipEventRecord(e[0], st);
        mhx::launch_queue_combine(st, s.items, s.chunk_count[0], s.chunk_units[0], exact, s.direct, 1, cand[v]);
        hipEventRecord(e[1], st);
        hipEventSynchronize(e[1]);
        hipEventElapsedTime(&ms[v], e[0], e[1]);
    }
    for (auto& x : e) hipEventDestroy(x);
    const int best = ms[0] < ms[1] ? cand[0] : cand[1];
    ctx->tuned[key] = best;
    return best;
}

}  // namespace

extern "C" {

int mhx_create(mhx_ctx** out, const mhx_options* opt) {
    if (!out) return MHX_EINVAL;
    *out = nullptr;
    mhx_ctx* c = new mhx_ctx();
    if (opt) {
        c->rho = opt->rho;
        c->g = opt->g;
        c->device = opt->device;
        c->timing = opt->timing;
    }
    if (!(c->g > 0) || !(c->rho > 0)) {
        delete c;
        return MHX_EINVAL;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || c->device < 0 || c->device >= ndev) {
        delete c;
        return MHX_EDEVICE;
    }
    bool ok = hipSetDevice(c->device) == hipSuccess &&
              hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < kMaxLuStreams - 1; i++)
        ok = hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&c->join[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        for (int i = 0; i < kMaxLuStreams - 1; i++) {
            if (c->side[i]) hipStreamDestroy(c->side[i]);
            if (c->join[i]) hipEventDestroy(c->join[i]);
        }
        if (c->stream) hipStreamDestroy(c->stream);
        if (c->fork) hipEventDestroy(c->fork);
        delete c;
        return MHX_EDEVICE;
    }
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0) c->batch_bytes = 0.6 * (double)free_b;
    c->scratch = scratch_address(c->device);
    *out = c;
    return MHX_OK;
}

void mhx_destroy(mhx_ctx* c) {
    if (!c) return;
    flush_timing(c);
    if (c->stream) hipStreamSynchronize(c->stream);
    drop_sweep(c);
    dfree(c->surr_params); dfree(c->surr_B); dfree(c->surr_bias); dfree(c->surr_W);
    for (int i = 0; i < kMaxLuStreams - 1; i++)
        if (c->side[i]) hipStreamSynchronize(c->side[i]);
    if (c->stream) hipStreamDestroy(c->stream);
    for (int i = 0; i < kMaxLuStreams - 1; i++) {
        if (c->side[i]) hipStreamDestroy(c->side[i]);
        if (c->join[i]) hipEventDestroy(c->join[i]);
    }
    if (c->fork) hipEventDestroy(c->fork);
    delete c;
}

const char* mhx_last_error(const mhx_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mhx_set_constants(mhx_ctx* ctx, double rho, double g) {
    if (!ctx) return MHX_EINVAL;
    if (!(rho > 0) || !(g > 0)) return fail(ctx, MHX_EINVAL, "rho and g must be positive");
    ctx->rho = rho;
    ctx->g = g;
    return MHX_OK;
}

int mhx_body_create(mhx_ctx* ctx, const mhx_mesh* mesh, int64_t ndof, const double* dofs, mhx_body** body) {
    if (!ctx || !body) return MHX_EINVAL;
    hipSetDevice(ctx->device);
    return body_upload(ctx, mesh, ndof, dofs, body);
}

int mhx_body_create_lid(mhx_ctx* ctx, const mhx_mesh* mesh, int64_t nhull, int64_t ndof, const double* dofs,
                        mhx_body** body) {
    if (!ctx || !body) return MHX_EINVAL;
    *body = nullptr;
    if (!mesh || nhull <= 0 || nhull > mesh->nf) return fail(ctx, MHX_EDIM, "body: nhull must be in 1..nf");
    hipSetDevice(ctx->device);
    int rc = body_upload(ctx, mesh, ndof, dofs, body);
    if (rc) return rc;
    mhx::launch_mark_lid(ctx->stream, (*body)->n, (int)nhull, (*body)->geom);
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        mhx_body_destroy(*body);
        *body = nullptr;
        return fail(ctx, MHX_EDEVICE, std::string("body lid: ") + hipGetErrorString(e));
    }
    return MHX_OK;
}

void mhx_body_destroy(mhx_body* b) {
    if (!b) return;
    if (b->ctx && sweep_uses(b->ctx, b)) {
        hipStreamSynchronize(b->ctx->stream);
        drop_sweep(b->ctx);
    }
    dfree(b->geom); dfree(b->dofs); dfree(b->SR); dfree(b->DR[0]); dfree(b->DR[1]); dfree(b->KRx);
    free_work(b);
    delete b;
}

int mhx_assemble(mhx_ctx* ctx, const mhx_mesh* mesh, double k, int gf, int direct, int include_identity, double* S_out,
                 double* D_out) {
    if (!ctx) return MHX_EINVAL;
    if (!S_out || !D_out) return fail(ctx, MHX_EINVAL, "assemble: null output");
    if (gf != MHX_GF_WU && gf != MHX_GF_DELHOMMEAU && gf != MHX_GF_SURROGATE)
        return fail(ctx, MHX_EINVAL, "unknown Green function");
    if (!(k > 0) || !std::isfinite(k)) return fail(ctx, MHX_EINVAL, "wavenumber must be positive and finite");
    hipSetDevice(ctx->device);
    direct = direct ? 1 : 0;
    mhx_body* b = nullptr;
    int rc = body_upload(ctx, mesh, 0, nullptr, &b);
    if (rc) return rc;
    const double w = std::sqrt(k * ctx->g);
    if (!(rc = ensure_rankine(ctx, b, direct)) && !(rc = ensure_work(ctx, b, 1, 0, gf == MHX_GF_DELHOMMEAU)) &&
        !(rc = set_freqs(ctx, b, &w, 1))) {
        // the wavenumber itself, not k recomputed from omega
        if (hipMemcpyAsync(b->d_ks, &k, sizeof(double), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
            rc = fail(ctx, MHX_EDEVICE, "assemble: copy k");
        if (!rc && !(rc = assemble_into(ctx, b, 1, gf, direct, include_identity ? 1 : 0)) &&
            !(rc = check_status(ctx, b, gf))) {
            const size_t bytes = sizeof(cplx) * (size_t)b->n * b->n;
            hipError_t e1 = hipMemcpyAsync(S_out, b->S, bytes, hipMemcpyDeviceToHost, ctx->stream);
            hipError_t e2 = hipMemcpyAsync(D_out, b->D, bytes, hipMemcpyDeviceToHost, ctx->stream);
            hipError_t e3 = hipStreamSynchronize(ctx->stream);
            if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) rc = fail(ctx, MHX_EDEVICE, "assemble: copy back");
        }
    }
    flush_timing(ctx);
    mhx_body_destroy(b);
    return rc;
}

static int lu_solve_host(mhx_ctx* ctx, int64_t n, int64_t nrhs, int64_t batch, const double* A, int trans,
                         double* Bh, bool fp64_only);

int mhx_lu_solve(mhx_ctx* ctx, int64_t n, int64_t nrhs, const double* A, int trans, double* Bh) {
    if (!ctx) return MHX_EINVAL;
    if (n <= 0 || nrhs <= 0) return fail(ctx, MHX_EDIM, "lu_solve: empty system");
    if (nrhs > 16 && A && Bh) {
        // the right-hand sides in column chunks of 16 (contiguous in the column-major B), each
        // with its own factorization of A
        for (int64_t c0 = 0; c0 < nrhs; c0 += 16) {
            const int rc = mhx_lu_solve(ctx, n, std::min<int64_t>(16, nrhs - c0), A, trans, Bh + 2 * n * c0);
            if (rc) return rc;
        }
        return MHX_OK;
    }
    if (n > mhx::kMaxOrder) return fail(ctx, MHX_EDIM, "lu_solve: order above 16384 is not supported");
    if (!A || !Bh) return fail(ctx, MHX_EINVAL, "lu_solve: null buffer");
    if (trans != 0 && trans != 2) return fail(ctx, MHX_EINVAL, "lu_solve: trans must be 0 or 2");
    return lu_solve_host(ctx, n, nrhs, 1, A, trans, Bh, true);
}

// Layout of the scratch for one system of order n with R right-hand sides: ipiv, perm (n ints
// each), the interchange workspace, info, then (256-byte aligned) the solve workspace.
static int64_t noalloc_ints(int64_t n) { return 2 * n + mhx::kPermWork + 1; }
static int64_t noalloc_bytes(int64_t n, int64_t R) {
    const int64_t head = (noalloc_ints(n) * (int64_t)sizeof(int) + 255) / 256 * 256;
    return head + mhx::getrs_work((int)n, (int)R) * (int64_t)sizeof(cplx);
}

int mhx_lu_solve_noalloc(mhx_ctx* ctx, int64_t n, int64_t nrhs, double* A, int trans, double* B) {
    if (!ctx) return MHX_EINVAL;
    if (n <= 0 || nrhs <= 0) return fail(ctx, MHX_EDIM, "lu_solve_noalloc: empty system");
    if (nrhs > 16) return fail(ctx, MHX_EDIM, "lu_solve_noalloc: at most 16 right-hand sides per call");
    if (n > mhx::kMaxOrder) return fail(ctx, MHX_EDIM, "lu_solve_noalloc: order above 16384 is not supported");
    if (!A || !B) return fail(ctx, MHX_EINVAL, "lu_solve_noalloc: null buffer");
    if (trans != 0 && trans != 2) return fail(ctx, MHX_EINVAL, "lu_solve_noalloc: trans must be 0 or 2");
    const int64_t need = noalloc_bytes(n, nrhs);
    if (need > (int64_t)MHX_NOALLOC_SCRATCH_BYTES)
        return fail(ctx, MHX_EDIM, "lu_solve_noalloc: the workspace of order " + std::to_string(n) + " with " +
                                       std::to_string(nrhs) + " right-hand sides (" + std::to_string(need) +
                                       " bytes) exceeds the " + std::to_string((long long)MHX_NOALLOC_SCRATCH_BYTES) +
                                       "-byte scratch");
    if (!ctx->scratch) return fail(ctx, MHX_EDEVICE, "lu_solve_noalloc: scratch address unavailable on this device");
    hipSetDevice(ctx->device);
    for (const void* p : {(const void*)A, (const void*)B}) {
        hipPointerAttribute_t at;
        if (hipPointerGetAttributes(&at, p) != hipSuccess ||
            (at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged) || at.device != ctx->device) {
            (void)hipGetLastError();
            return fail(ctx, MHX_EINVAL, "lu_solve_noalloc: A and B must be device memory of the context's device");
        }
    }
    hipStream_t st = ctx->stream;
    int* ipiv = reinterpret_cast<int*>(ctx->scratch);
    int* perm = ipiv + n;
    int* wperm = perm + n;
    int* info = wperm + mhx::kPermWork;
    cplx* W = reinterpret_cast<cplx*>(ctx->scratch + (noalloc_ints(n) * (int64_t)sizeof(int) + 255) / 256 * 256);
    cplx* dA = reinterpret_cast<cplx*>(A);
    cplx* dB = reinterpret_cast<cplx*>(B);
    mhx::launch_zero(st, info, sizeof(int));  // a fill kernel, not a memset node
    if (mhx::getrf_batched(st, (int)n, dA, (int)n, n * n, 1, ipiv, n, wperm, info, nullptr))
        return fail(ctx, MHX_EDIM, "lu_solve_noalloc: order not supported by the factorization");
    mhx::getrs_batched(st, trans, (int)n, (int)nrhs, dA, (int)n, 0, ipiv, n, perm, n, dB, 0, W, 0, 1);
    int hinfo = 0;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(ctx, MHX_EDEVICE, std::string("lu_solve_noalloc: ") + hipGetErrorString(e));
    if (hinfo) return fail(ctx, MHX_ESINGULAR, "LU: exactly zero pivot at column " + std::to_string(hinfo));
    return MHX_OK;
}

int mhx_lu_solve_batched(mhx_ctx* ctx, int64_t n, int64_t nrhs, int64_t batch, const double* A, int trans, double* Bh) {
    if (!ctx) return MHX_EINVAL;
    if (n <= 0 || nrhs <= 0 || batch <= 0) return fail(ctx, MHX_EDIM, "lu_solve_batched: empty system or batch");
    if (nrhs > 16) return fail(ctx, MHX_EDIM, "lu_solve_batched: at most 16 right-hand sides per system");
    if (n > mhx::kMaxOrder) return fail(ctx, MHX_EDIM, "lu_solve_batched: order above 16384 is not supported");
    if (batch > (1 << 16)) return fail(ctx, MHX_EDIM, "lu_solve_batched: batch too large");
    if (!A || !Bh) return fail(ctx, MHX_EINVAL, "lu_solve_batched: null buffer");
    if (trans != 0 && trans != 2) return fail(ctx, MHX_EINVAL, "lu_solve_batched: trans must be 0 or 2");
    return lu_solve_host(ctx, n, nrhs, batch, A, trans, Bh, false);
}

// systems from host buffers through a temporary body: FP64 (fp64_only, or the context's LU mode
// says so, or trans = 2) or the mixed-precision solver
static int lu_solve_host(mhx_ctx* ctx, int64_t n, int64_t nrhs, int64_t batch, const double* A, int trans,
                         double* Bh, bool fp64_only) {
    hipSetDevice(ctx->device);
    mhx_body b;
    b.ctx = ctx;
    b.n = (int)n;
    b.ndof = 0;
    int rc = ensure_work(ctx, &b, (int)nrhs, 0, false, (int)batch, false);
    if (!rc) {
        hipStream_t st = ctx->stream;
        const size_t nn = (size_t)n * n, nR = (size_t)n * nrhs;
        for (int64_t q = 0; q < batch && !rc; q++) {
            if (hipMemcpyAsync(b.D + q * b.nn(), A + 2 * nn * q, sizeof(cplx) * nn, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(b.B + q * b.nR(), Bh + 2 * nR * q, sizeof(cplx) * nR, hipMemcpyHostToDevice, st) != hipSuccess)
                rc = fail(ctx, MHX_EDEVICE, "lu_solve_batched: upload");
        }
        // the mixed-precision solvers take A x = B only; A^H x = B is solved in FP64
        const bool mixed =
            !fp64_only && trans == 0 && (ctx->lu_mode == MHX_LU_MIXED || ctx->lu_mode == MHX_LU_MIXED_BF16);
        if (!rc)
            rc = mixed ? solve_mixed(ctx, &b, (int)nrhs, 0, (int)batch,
                                     ctx->lu_mode == MHX_LU_MIXED_BF16 ? MHX_GF_SURROGATE : MHX_GF_WU, false)
                       : solve_fp64(ctx, &b, (int)nrhs, 0, (int)batch, false, trans);
        if (rc == MHX_ESINGULAR) rc = fail(ctx, MHX_ESINGULAR, ctx->err);
        for (int64_t q = 0; q < batch && !rc; q++)
            if (hipMemcpyAsync(Bh + 2 * nR * q, b.B + q * b.nR(), sizeof(cplx) * nR, hipMemcpyDeviceToHost, st) != hipSuccess)
                rc = fail(ctx, MHX_EDEVICE, "lu_solve_batched: copy back");
        if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail(ctx, MHX_EDEVICE, "lu_solve_batched: sync");
    }
    flush_timing(ctx);
    free_work(&b);
    return rc;
}

int mhx_solve(mhx_ctx* ctx, int64_t n, int64_t nrhs, const double* D, const double* S, const double* bc, int direct,
              double* phi_out, double* sources_out) {
    if (!ctx) return MHX_EINVAL;
    if (n <= 0 || nrhs <= 0) return fail(ctx, MHX_EDIM, "solve: empty system");
    if (nrhs > 16 && D && S && bc && phi_out) {
        // the boundary conditions in column chunks of 16 (contiguous in the column-major bc,
        // phi and sources), each chunk a full solve
        for (int64_t c0 = 0; c0 < nrhs; c0 += 16) {
            const int rc = mhx_solve(ctx, n, std::min<int64_t>(16, nrhs - c0), D, S, bc + 2 * n * c0, direct,
                                     phi_out + 2 * n * c0, sources_out ? sources_out + 2 * n * c0 : nullptr);
            if (rc) return rc;
        }
        return MHX_OK;
    }
    if (n > mhx::kMaxOrder) return fail(ctx, MHX_EDIM, "solve: order above 16384 is not supported");
    if (!D || !S || !bc || !phi_out) return fail(ctx, MHX_EINVAL, "solve: null buffer");
    hipSetDevice(ctx->device);
    mhx_body b;
    b.ctx = ctx;
    b.n = (int)n;
    b.ndof = 0;
    int rc = ensure_work(ctx, &b, (int)nrhs, 0, false);
    if (!rc) {
        hipStream_t st = ctx->stream;
        hipMemcpyAsync(b.D, D, sizeof(cplx) * n * n, hipMemcpyHostToDevice, st);
        hipMemcpyAsync(b.S, S, sizeof(cplx) * n * n, hipMemcpyHostToDevice, st);
        hipMemcpyAsync(b.B, bc, sizeof(cplx) * n * nrhs, hipMemcpyHostToDevice, st);
        rc = factor_and_solve(ctx, &b, (int)nrhs, direct ? 1 : 0);
        if (!rc) {
            hipMemcpyAsync(phi_out, b.Y, sizeof(cplx) * n * nrhs, hipMemcpyDeviceToHost, st);
            if (!direct && sources_out)
                hipMemcpyAsync(sources_out, b.B, sizeof(cplx) * n * nrhs, hipMemcpyDeviceToHost, st);
            if (hipStreamSynchronize(st) != hipSuccess) rc = fail(ctx, MHX_EDEVICE, "solve: copy back");
        }
    }
    flush_timing(ctx);
    free_work(&b);
    return rc;
}

// packing of added_mass / radiation_damping (src/problems_and_results.jl:207-208, 325-380)
static void pack(const std::vector<cplx>& F, const std::vector<cplx>& FK, int ndof, int R, int nb, int64_t iw,
                 int64_t nw, double w, double* A_out, double* B_out, double* Fd_out, double* Fk_out) {
    for (int r = 0; r < R; r++)
        for (int d = 0; d < ndof; d++) {
            const cplx f = F[d + (size_t)ndof * r];
            if (r < ndof) {
                A_out[d + (size_t)ndof * (r + (size_t)ndof * iw)] = f.x / (w * w);
                B_out[d + (size_t)ndof * (r + (size_t)ndof * iw)] = f.y / w;
            } else {
                const size_t o = d + (size_t)ndof * (iw + (size_t)nw * (r - ndof));
                Fd_out[2 * o] = f.x;
                Fd_out[2 * o + 1] = f.y;
            }
        }
    for (int bb = 0; bb < nb; bb++)
        for (int d = 0; d < ndof; d++) {
            const size_t o = d + (size_t)ndof * (iw + (size_t)nw * bb);
            Fk_out[2 * o] = FK[d + (size_t)ndof * bb].x;
            Fk_out[2 * o + 1] = FK[d + (size_t)ndof * bb].y;
        }
}

int mhx_body_coeffs(mhx_ctx* ctx, mhx_body* b, int64_t nw, const double* omegas, int64_t nb, const double* betas, int gf,
                    int direct, int mode, double* A_out, double* B_out, double* Fd_out, double* Fk_out) {
    if (!ctx || !b) return MHX_EINVAL;
    if (nw < 0 || nb < 0 || (nw && !omegas) || (nb && !betas)) return fail(ctx, MHX_EINVAL, "coeffs: bad frequency/heading arrays");
    if (gf != MHX_GF_WU && gf != MHX_GF_DELHOMMEAU && gf != MHX_GF_SURROGATE)
        return fail(ctx, MHX_EINVAL, "unknown Green function");
    if (mode != MHX_MODE_PER_OMEGA && mode != MHX_MODE_PER_PROBLEM && mode != MHX_MODE_BATCHED &&
        mode != MHX_MODE_GRAPH)
        return fail(ctx, MHX_EINVAL, "unknown mode");
    if (mode == MHX_MODE_GRAPH) {
        // the batched schedule as a one-body sweep captured into a HIP graph (replayed while the
        // body, frequency count, headings, constants and precision stay the same)
        mhx_sweep_body sb{b, nw, omegas, A_out, B_out, Fd_out, Fk_out};
        return mhx_sweep_coeffs(ctx, &sb, 1, nb, betas, gf, direct, MHX_SWEEP_GRAPH);
    }
    const int ndof = b->ndof, n = b->n;
    if (ndof + nb > 16 && ndof < 16) {
        // more right-hand sides than one solve takes (16): the headings in chunks of 16 - ndof,
        // each chunk a full pass (its own assembly and factorization; the radiation problems are
        // solved again and only the first chunk's A and B are kept).  The excitation outputs are
        // (dof, omega, heading) with the heading slowest, so a chunk's are one contiguous slab.
        const int64_t per = 16 - ndof;
        std::vector<double> Ad, Bd;
        for (int64_t b0 = 0; b0 < nb; b0 += per) {
            const int64_t nbg = std::min<int64_t>(per, nb - b0);
            const size_t off = 2 * (size_t)ndof * (size_t)nw * (size_t)b0;
            double *Ao = A_out, *Bo = B_out;
            if (b0 > 0 && ndof > 0) {
                Ad.resize((size_t)ndof * ndof * nw);
                Bd.resize((size_t)ndof * ndof * nw);
                Ao = Ad.data();
                Bo = Bd.data();
            }
            const int rc = mhx_body_coeffs(ctx, b, nw, omegas, nbg, betas + b0, gf, direct, mode, Ao, Bo,
                                           Fd_out ? Fd_out + off : nullptr, Fk_out ? Fk_out + off : nullptr);
            if (rc) return rc;
        }
        return MHX_OK;
    }
    const int R = ndof + (int)nb;
    if (R <= 0) return fail(ctx, MHX_EDIM, "coeffs: no radiating dof and no heading");
    if (R > 16) return fail(ctx, MHX_EDIM, "coeffs: at most 15 dofs");
    if (n > mhx::kMaxOrder) return fail(ctx, MHX_EDIM, "coeffs: more than 16384 panels is not supported");
    if ((ndof && (!A_out || !B_out)) || (nb && ndof && (!Fd_out || !Fk_out)))
        return fail(ctx, MHX_EINVAL, "coeffs: null output");
    for (int64_t w = 0; w < nw; w++)
        if (!(omegas[w] > 0) || !std::isfinite(omegas[w])) return fail(ctx, MHX_EINVAL, "omega must be positive and finite");
    hipSetDevice(ctx->device);
    direct = direct ? 1 : 0;
    const bool table = gf == MHX_GF_DELHOMMEAU;  // (G, dG/dr) table of the exact term
    int chunk = 1;
    if (mode == MHX_MODE_BATCHED && nw > 1) {
        const double per = item_bytes(ctx, n, gf, table, 20);
        chunk = (int)std::min<double>((double)std::min<int64_t>(nw, ctx->max_batch), std::max(1.0, ctx->batch_bytes / per));
    }
    int rc;
    if ((rc = ensure_work(ctx, b, R, (int)nb, table, chunk))) return rc;
    if ((rc = ensure_rankine(ctx, b, direct))) return rc;
    hipStream_t st = ctx->stream;
    if (nb) HCHK(hipMemcpyAsync(b->betas, betas, sizeof(double) * nb, hipMemcpyHostToDevice, st));
    std::vector<cplx> F((size_t)ndof * R * chunk), FK((size_t)ndof * (nb ? nb : 1) * chunk);
    for (int64_t w0 = 0; w0 < nw; w0 += chunk) {
        const int bt = (int)std::min<int64_t>(chunk, nw - w0);
        if ((rc = set_freqs(ctx, b, omegas + w0, bt))) break;
        mhx::launch_bc(st, n, b->geom, b->dofs, ndof, b->betas, (int)nb, b->d_ws, b->d_ks, bt, ctx->g, b->B, b->nR());
        if (mode != MHX_MODE_PER_PROBLEM) {
            if ((rc = assemble_into(ctx, b, bt, gf, direct, 1, false, true)) || (rc = check_status(ctx, b, gf))) break;
            if ((rc = factor_and_solve(ctx, b, R, direct, bt, gf))) break;
        } else {
            // reference semantics: one assembly (Rankine part included, as the reference recomputes
            // it) and one LU per radiation / diffraction problem
            std::vector<cplx> hB((size_t)n * R), hphi((size_t)n * R);
            HCHK(hipMemcpyAsync(hB.data(), b->B, sizeof(cplx) * n * R, hipMemcpyDeviceToHost, st));
            HCHK(hipStreamSynchronize(st));
            for (int r = 0; r < R && !rc; r++) {
                b->have_rankine[direct] = false;
                if ((rc = ensure_rankine(ctx, b, direct))) break;
                if ((rc = assemble_into(ctx, b, 1, gf, direct, 1, false, true)) || (rc = check_status(ctx, b, gf))) break;
                HCHK(hipMemcpyAsync(b->B, hB.data() + (size_t)n * r, sizeof(cplx) * n, hipMemcpyHostToDevice, st));
                if ((rc = factor_and_solve(ctx, b, 1, direct, 1, gf))) break;
                HCHK(hipMemcpyAsync(hphi.data() + (size_t)n * r, b->Y, sizeof(cplx) * n, hipMemcpyDeviceToHost, st));
                HCHK(hipStreamSynchronize(st));
            }
            if (rc) break;
            HCHK(hipMemcpyAsync(b->Y, hphi.data(), sizeof(cplx) * n * R, hipMemcpyHostToDevice, st));
        }
        if (ndof > 0) {
            Phase ph(ctx, "forces");
            mhx::launch_forces(st, n, b->geom, b->dofs, ndof, b->Y, b->nR(), R, ctx->rho, b->d_ws, b->d_ks, bt, ctx->g,
                               b->betas, (int)nb, b->F, b->FK);
        }
        HCHK(hipGetLastError());
        HCHK(hipMemcpyAsync(F.data(), b->F, sizeof(cplx) * ndof * R * bt, hipMemcpyDeviceToHost, st));
        if (nb) HCHK(hipMemcpyAsync(FK.data(), b->FK, sizeof(cplx) * ndof * nb * bt, hipMemcpyDeviceToHost, st));
        HCHK(hipStreamSynchronize(st));
        for (int q = 0; q < bt; q++) {
            std::vector<cplx> Fq(F.begin() + (size_t)q * ndof * R, F.begin() + (size_t)(q + 1) * ndof * R);
            std::vector<cplx> FKq(FK.begin() + (size_t)q * ndof * nb, FK.begin() + (size_t)(q + 1) * ndof * nb);
            pack(Fq, FKq, ndof, R, (int)nb, w0 + q, nw, omegas[w0 + q], A_out, B_out, Fd_out, Fk_out);
        }
    }
    flush_timing(ctx);
    return rc;
}

// Forward speed.  U = 0: the plain problems, through mhx_body_coeffs (one batched solve per
// frequency chunk shared by every heading).  U != 0: per (U, beta), the frequencies of a chunk
// batched, each item at its own encountered (omega_e, k_e): indirect assembly at k_e (the
// surrogate keeps its (G, dG/dr) table, as the exact term does), LU in the precision
// mhx_set_lu_mode picks for the Green function, then K_x at k_e written over the dead LU factors, dphi/dx = K_x sigma, and
// phi <- phi - i (U / omega_e) dphi/dx so that the force kernel's i rho omega_e phi carries the
// rho U dphi/dx term (solve.jl:35-55).
int mhx_body_coeffs_fs(mhx_ctx* ctx, mhx_body* b, int64_t nw, const double* omegas, int64_t nb, const double* betas,
                       int64_t nU, const double* Us, const double* dgrad, int gf, int direct, double* A_out,
                       double* B_out, double* Fd_out, double* Fk_out) {
    if (!ctx || !b) return MHX_EINVAL;
    if (nw < 0 || nb < 1 || nU < 1 || (nw && !omegas) || !betas || !Us)
        return fail(ctx, MHX_EINVAL, "forward speed: bad frequency, heading or speed arrays (at least one heading and one speed)");
    if (gf != MHX_GF_WU && gf != MHX_GF_DELHOMMEAU && gf != MHX_GF_SURROGATE)
        return fail(ctx, MHX_EINVAL, "unknown Green function");
    bool moving = false;
    for (int64_t u = 0; u < nU; u++) {
        if (!std::isfinite(Us[u])) return fail(ctx, MHX_EINVAL, "forward speed must be finite");
        if (Us[u] != 0) moving = true;
    }
    for (int64_t w = 0; w < nw; w++)
        if (!(omegas[w] > 0) || !std::isfinite(omegas[w])) return fail(ctx, MHX_EINVAL, "omega must be positive and finite");
    for (int64_t q = 0; q < nb; q++)
        if (!std::isfinite(betas[q])) return fail(ctx, MHX_EINVAL, "wave direction must be finite");
    const int ndof = b->ndof, n = b->n;
    if (ndof == 0 || nw == 0) return MHX_OK;  // every output is empty
    if (!A_out || !B_out || !Fd_out || !Fk_out) return fail(ctx, MHX_EINVAL, "coeffs: null output");
    if (moving) {
        if (direct)
            return fail(ctx, MHX_EINVAL,
                        "nonzero forward speed is not supported with the direct method; use IndirectBEM (direct = false)");
        if (!dgrad) return fail(ctx, MHX_EINVAL, "nonzero forward speed: null gradient of motion");
    }
    if (ndof + 1 > 16) return fail(ctx, MHX_EDIM, "forward speed: at most 15 dofs");
    if (n > mhx::kMaxOrder) return fail(ctx, MHX_EDIM, "coeffs: more than 16384 panels is not supported");
    hipSetDevice(ctx->device);
    const size_t nd = ndof, nd2 = nd * nd;
    auto aidx = [&](int d, int r, int64_t iw, int64_t iu, int64_t ib) {
        return d + nd * (r + nd * (iw + (size_t)nw * (iu + (size_t)nU * ib)));
    };
    auto fidx = [&](int d, int64_t iw, int64_t ib, int64_t iu) {
        return d + nd * (iw + (size_t)nw * (ib + (size_t)nb * iu));
    };
    int rc = MHX_OK;
    // ---- U = 0: the plain problems, every heading sharing the solve ----
    for (int64_t iu = 0; iu < nU && !rc; iu++) {
        if (Us[iu] != 0) continue;
        const int64_t cb = 16 - ndof;
        std::vector<double> tA(nd2 * nw), tB(nd2 * nw), tFd(2 * nd * nw * cb), tFk(2 * nd * nw * cb);
        for (int64_t b0 = 0; b0 < nb && !rc; b0 += cb) {
            const int64_t bc = std::min<int64_t>(cb, nb - b0);
            rc = mhx_body_coeffs(ctx, b, nw, omegas, bc, betas + b0, gf, direct, MHX_MODE_BATCHED, tA.data(), tB.data(),
                                 tFd.data(), tFk.data());
            if (rc) break;
            for (int64_t q = 0; q < bc; q++)
                for (int64_t iw = 0; iw < nw; iw++)
                    for (int d = 0; d < ndof; d++) {
                        const size_t src = d + nd * (iw + (size_t)nw * q), dst = fidx(d, iw, b0 + q, iu);
                        Fd_out[2 * dst] = tFd[2 * src];
                        Fd_out[2 * dst + 1] = tFd[2 * src + 1];
                        Fk_out[2 * dst] = tFk[2 * src];
                        Fk_out[2 * dst + 1] = tFk[2 * src + 1];
                    }
        }
        if (rc) break;
        for (int64_t ib = 0; ib < nb; ib++)
            for (int64_t iw = 0; iw < nw; iw++)
                for (int r = 0; r < ndof; r++)
                    for (int d = 0; d < ndof; d++) {
                        const size_t src = d + nd * (r + nd * iw);
                        A_out[aidx(d, r, iw, iu, ib)] = tA[src];
                        B_out[aidx(d, r, iw, iu, ib)] = tB[src];
                    }
    }
    if (rc || !moving) return rc;
    // ---- U != 0 ----
    const int R = ndof + 1;
    const bool table = gf != MHX_GF_WU;  // K_x reads dG/dr of the exact term (or the surrogate's)
    const double g = ctx->g;
    int chunk = 1;
    {
        const double per = item_bytes(ctx, n, gf, table, 21);
        chunk = (int)std::min<double>((double)std::min<int64_t>(nw, ctx->max_batch), std::max(1.0, ctx->batch_bytes / per));
    }
    if ((rc = ensure_work(ctx, b, R, 1, table, chunk)) || (rc = ensure_rankine(ctx, b, 0))) return rc;
    hipStream_t st = ctx->stream;
    if (!b->KRx) {
        if ((rc = dalloc(ctx, &b->KRx, (size_t)b->nn()))) return rc;
        Phase ph(ctx, "rankine");
        mhx::launch_rankine_x(st, n, b->geom, b->KRx);
    }
    double *d_grad = nullptr, *d_inc = nullptr;
    cplx* psi = nullptr;
    auto cleanup = [&]() { dfree(d_grad); dfree(d_inc); dfree(psi); };
    if ((rc = dalloc(ctx, &d_grad, 3 * nd)) || (rc = dalloc(ctx, &d_inc, 3 * (size_t)chunk)) ||
        (rc = dalloc(ctx, &psi, (size_t)b->nR() * chunk))) {
        cleanup();
        return rc;
    }
    // every exit of the loops below goes through cleanup (HCHK returns from the lambda)
    auto run = [&]() -> int {
        HCHK(hipMemcpyAsync(d_grad, dgrad, sizeof(double) * 3 * nd, hipMemcpyHostToDevice, st));
        std::vector<double> we(chunk), ke(chunk), inc(3 * (size_t)chunk);
        std::vector<cplx> F((size_t)ndof * R * chunk), FK((size_t)ndof * chunk);
        int e;
        for (int64_t iu = 0; iu < nU; iu++) {
            const double U = Us[iu];
            if (U == 0) continue;
            for (int64_t ib = 0; ib < nb; ib++) {
                const double beta = betas[ib];
                for (int64_t w0 = 0; w0 < nw; w0 += chunk) {
                    const int bt = (int)std::min<int64_t>(chunk, nw - w0);
                    // compute_encountered_values (waves.jl:136-151)
                    for (int q = 0; q < bt; q++) {
                        const double w = omegas[w0 + q], k = w * w / g;
                        const double dop = w - k * U * cos(beta);
                        we[q] = fabs(dop);
                        ke[q] = we[q] * we[q] / g;
                        inc[q] = w;
                        inc[chunk + q] = k;
                        inc[2 * chunk + q] = dop >= 0 ? beta : beta + MHX_PI;
                        if (!(we[q] > 0)) return fail(ctx, MHX_EINVAL, "forward speed: zero encountered frequency");
                    }
                    if ((e = set_freqs(ctx, b, we.data(), bt))) return e;
                    HCHK(hipMemcpyAsync(d_inc, inc.data(), sizeof(double) * 3 * chunk, hipMemcpyHostToDevice, st));
                    for (int q = 0; q < bt; q++)
                        mhx::launch_bc_fs(st, n, b->geom, b->dofs, ndof, d_grad, we[q], U, inc[chunk + q], inc[q], beta,
                                          inc[2 * chunk + q], g, b->B + q * b->nR());
                    if ((e = assemble_into(ctx, b, bt, gf, 0, 1, true)) || (e = check_status(ctx, b, gf))) return e;
                    // sigma stays in b->B, phi = S sigma in b->Y
                    if ((e = factor_and_solve(ctx, b, R, 0, bt, gf))) return e;
                    {
                        Phase ph(ctx, "forward_speed");
                        for (int q = 0; q < bt; q++)
                            mhx::launch_kx_combine(st, n, b->geom, b->KRx, ke[q], table, table ? b->WR + q * b->nn() : nullptr,
                                                   b->D + q * b->nn());
                        mhx::launch_matmul_nr(st, n, R, b->D, b->nn(), b->B, b->nR(), psi, b->nR(), b->work, b->sW(), bt);
                        for (int q = 0; q < bt; q++)
                            mhx::launch_caxpy(st, (int64_t)n * R, 0.0, -U / we[q], psi + q * b->nR(), b->Y + q * b->nR());
                    }
                    {
                        Phase ph(ctx, "forces");
                        mhx::launch_forces(st, n, b->geom, b->dofs, ndof, b->Y, b->nR(), R, ctx->rho, b->d_ws, b->d_ks, bt, g,
                                           nullptr, 0, b->F, b->FK);
                        // froude_krylov_force (waves.jl:294-303): incident omega and k, encountered heading
                        for (int q = 0; q < bt; q++)
                            mhx::launch_forces(st, n, b->geom, b->dofs, ndof, b->Y, b->nR(), 0, ctx->rho, d_inc + q,
                                               d_inc + chunk + q, 1, g, d_inc + 2 * chunk + q, 1, b->F, b->FK + q * ndof);
                    }
                    HCHK(hipGetLastError());
                    HCHK(hipMemcpyAsync(F.data(), b->F, sizeof(cplx) * ndof * R * bt, hipMemcpyDeviceToHost, st));
                    HCHK(hipMemcpyAsync(FK.data(), b->FK, sizeof(cplx) * ndof * bt, hipMemcpyDeviceToHost, st));
                    HCHK(hipStreamSynchronize(st));
                    // problems_and_results.jl:202-208, 341-380
                    for (int q = 0; q < bt; q++) {
                        const int64_t iw = w0 + q;
                        for (int r = 0; r < R; r++)
                            for (int d = 0; d < ndof; d++) {
                                const cplx f = F[d + (size_t)ndof * (r + (size_t)R * q)];
                                if (r < ndof) {
                                    A_out[aidx(d, r, iw, iu, ib)] = f.x / (we[q] * we[q]);
                                    B_out[aidx(d, r, iw, iu, ib)] = f.y / we[q];
                                } else {
                                    const size_t o = fidx(d, iw, ib, iu);
                                    Fd_out[2 * o] = f.x;
                                    Fd_out[2 * o + 1] = f.y;
                                }
                            }
                        for (int d = 0; d < ndof; d++) {
                            const size_t o = fidx(d, iw, ib, iu);
                            Fk_out[2 * o] = FK[d + (size_t)ndof * q].x;
                            Fk_out[2 * o + 1] = FK[d + (size_t)ndof * q].y;
                        }
                    }
                }
            }
        }
        return MHX_OK;
    };
    rc = run();
    cleanup();
    flush_timing(ctx);
    return rc;
}

// Directional derivative of the coefficients (forward mode): per frequency, the primal omega-solve
// (FP64 LU) and its tangent from the tangent assembly, the tangent right-hand sides and the
// same LU factors.
int mhx_hydro_coeffs_jvp(mhx_ctx* ctx, const mhx_mesh* mesh, const mhx_mesh_tangent* dmesh, int64_t ndof,
                         const double* dofs, const double* ddofs, int64_t nw, const double* omegas,
                         const double* domegas, int64_t nb, const double* betas, int gf, int direct, double* A_out,
                         double* B_out, double* Fd_out, double* Fk_out, double* dA_out, double* dB_out,
                         double* dFd_out, double* dFk_out) {
    if (!ctx) return MHX_EINVAL;
    if (gf != MHX_GF_WU && gf != MHX_GF_DELHOMMEAU)
        return fail(ctx, MHX_EINVAL, "coeffs_jvp: the tangent is available for GFWu and ExactGuevelDelhommeau");
    if (nw < 0 || nb < 0 || (nw && !omegas) || (nb && !betas)) return fail(ctx, MHX_EINVAL, "coeffs_jvp: bad frequency/heading arrays");
    if ((ndof && (!A_out || !B_out || !dA_out || !dB_out)) || (nb && ndof && (!Fd_out || !Fk_out || !dFd_out || !dFk_out)))
        return fail(ctx, MHX_EINVAL, "coeffs_jvp: null output");
    for (int64_t w = 0; w < nw; w++)
        if (!(omegas[w] > 0) || !std::isfinite(omegas[w])) return fail(ctx, MHX_EINVAL, "omega must be positive and finite");
    const int64_t R64 = ndof + nb;
    if (R64 <= 0) return fail(ctx, MHX_EDIM, "coeffs_jvp: no radiating dof and no heading");
    if (R64 > 16) return fail(ctx, MHX_EDIM, "coeffs_jvp: at most 16 right-hand sides (dofs + headings)");
    if (mesh && mesh->nf > mhx::kMaxOrder) return fail(ctx, MHX_EDIM, "coeffs_jvp: more than 16384 panels is not supported");
    hipSetDevice(ctx->device);
    mhx_body* b = nullptr;
    int rc = body_upload(ctx, mesh, ndof, dofs, &b);
    if (rc) return rc;
    const int n = b->n, R = (int)R64;
    direct = direct ? 1 : 0;
    hipStream_t st = ctx->stream;
    double *dv = nullptr, *ddv = nullptr, *dc = nullptr, *ddc = nullptr, *dn = nullptr, *ddn = nullptr, *da = nullptr,
           *dda = nullptr, *dr = nullptr, *d_ddofs = nullptr, *d_betas = nullptr;
    int64_t* df = nullptr;
    mhx::PanelGeomD* gd = nullptr;
    cplx *dS = nullptr, *dD = nullptr, *dBv = nullptr, *t1 = nullptr, *t2 = nullptr, *t3 = nullptr, *dphi = nullptr,
         *dF = nullptr, *dFK = nullptr;
    auto cleanup = [&]() {
        dfree(dv); dfree(ddv); dfree(dc); dfree(ddc); dfree(dn); dfree(ddn); dfree(da); dfree(dda); dfree(dr);
        dfree(d_ddofs); dfree(d_betas); dfree(df); dfree(gd); dfree(dS); dfree(dD); dfree(dBv); dfree(t1); dfree(t2);
        dfree(t3); dfree(dphi); dfree(dF); dfree(dFK);
        mhx_body_destroy(b);
    };
    const size_t nn = (size_t)n * n, nR = (size_t)n * R, nv = mesh->nv;
    const bool tab = gf == MHX_GF_DELHOMMEAU;
    if ((rc = ensure_work(ctx, b, R, (int)nb, tab, 1)) || (rc = ensure_rankine(ctx, b, direct)) ||
        (rc = dalloc(ctx, &dv, 3 * nv)) || (rc = dalloc(ctx, &ddv, 3 * nv)) || (rc = dalloc(ctx, &df, 4 * (size_t)n)) ||
        (rc = dalloc(ctx, &dc, 3 * (size_t)n)) || (rc = dalloc(ctx, &ddc, 3 * (size_t)n)) || (rc = dalloc(ctx, &dn, 3 * (size_t)n)) ||
        (rc = dalloc(ctx, &ddn, 3 * (size_t)n)) || (rc = dalloc(ctx, &da, n)) || (rc = dalloc(ctx, &dda, n)) ||
        (rc = dalloc(ctx, &dr, n)) || (rc = dalloc(ctx, &d_ddofs, 3 * (size_t)n * (ndof ? ndof : 1))) ||
        (rc = dalloc(ctx, &d_betas, (size_t)(nb ? nb : 1))) || (rc = dalloc(ctx, &gd, n)) || (rc = dalloc(ctx, &dS, nn)) ||
        (rc = dalloc(ctx, &dD, nn)) || (rc = dalloc(ctx, &dBv, nR)) || (rc = dalloc(ctx, &t1, nR)) ||
        (rc = dalloc(ctx, &t2, nR)) || (rc = dalloc(ctx, &t3, nR)) || (rc = dalloc(ctx, &dphi, nR)) ||
        (rc = dalloc(ctx, &dF, (size_t)(ndof ? ndof : 1) * R)) || (rc = dalloc(ctx, &dFK, (size_t)(ndof ? ndof : 1) * (nb ? nb : 1)))) {
        cleanup();
        return rc;
    }
    // mesh and tangent fields on the device (a null tangent array is zero)
    auto up = [&](double* dst, const double* src, size_t cnt) {
        if (src) hipMemcpyAsync(dst, src, sizeof(double) * cnt, hipMemcpyHostToDevice, st);
        else mhx::launch_zero(st, dst, sizeof(double) * cnt);
    };
    up(dv, mesh->vertices, 3 * nv);
    up(dc, mesh->centers, 3 * (size_t)n);
    up(dn, mesh->normals, 3 * (size_t)n);
    up(da, mesh->areas, n);
    up(dr, mesh->radii, n);
    up(ddv, dmesh ? dmesh->vertices : nullptr, 3 * nv);
    up(ddc, dmesh ? dmesh->centers : nullptr, 3 * (size_t)n);
    up(ddn, dmesh ? dmesh->normals : nullptr, 3 * (size_t)n);
    up(dda, dmesh ? dmesh->areas : nullptr, n);
    if (ndof) up(d_ddofs, ddofs, 3 * (size_t)n * ndof);
    if (nb) {
        hipMemcpyAsync(d_betas, betas, sizeof(double) * nb, hipMemcpyHostToDevice, st);
        hipMemcpyAsync(b->betas, betas, sizeof(double) * nb, hipMemcpyHostToDevice, st);
    }
    hipMemcpyAsync(df, mesh->faces, sizeof(int64_t) * 4 * n, hipMemcpyHostToDevice, st);
    // device errors inside the sweep end it with rc set, so that the buffers are released below
    auto ok = [&](hipError_t e, const char* what) {
        if (e == hipSuccess) return true;
        rc = fail(ctx, MHX_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
        return false;
    };
    mhx::launch_tangent_geometry(st, n, dv, ddv, df, dc, ddc, dn, ddn, da, dda, dr, gd);
    if (!ok(hipGetLastError(), "tangent geometry")) {
        cleanup();
        return rc;
    }
    std::vector<cplx> F((size_t)ndof * R), FK((size_t)ndof * (nb ? nb : 1)), hdF((size_t)ndof * R),
        hdFK((size_t)ndof * (nb ? nb : 1));
    const double g = ctx->g;
    for (int64_t iw = 0; iw < nw && !rc; iw++) {
        const double w = omegas[iw], dw = domegas ? domegas[iw] : 0.0;
        const double k = w * w / g, dk = 2.0 * w * dw / g;  // compute_wavenumber (src/waves.jl:133-134)
        if ((rc = set_freqs(ctx, b, omegas + iw, 1))) break;
        mhx::launch_bc(st, n, b->geom, b->dofs, ndof, b->betas, (int)nb, b->d_ws, b->d_ks, 1, g, b->B, b->nR());
        mhx::launch_tangent_bc(st, n, gd, b->dofs, ndof ? d_ddofs : nullptr, ndof, d_betas, (int)nb, w, dw, k, dk, g, dBv);
        if ((rc = assemble_into(ctx, b, 1, gf, direct, 1)) || (rc = check_status(ctx, b, gf))) break;
        mhx::launch_tangent_pairs(st, n, gd, k, dk, direct, tab, b->WG, b->WR, dS, dD);
        if ((rc = solve_fp64(ctx, b, R, direct, 1))) {
            if (rc == MHX_ESINGULAR) fail(ctx, rc, ctx->err);
            break;
        }
        cplx* W2 = b->work + 16 * b->nR();
        const int64_t cnt = (int64_t)n * R;
        {
            Phase ph(ctx, "tangent");
            if (direct) {
                // dphi = D^{-1} (dS B + S dB - dD phi)
                mhx::launch_matmul_nr(st, n, R, dS, 0, b->B, 0, t1, 0, b->work, 0, 1);
                mhx::launch_matmul_nr(st, n, R, b->S, 0, dBv, 0, t2, 0, b->work, 0, 1);
                mhx::launch_matmul_nr(st, n, R, dD, 0, b->Y, 0, t3, 0, b->work, 0, 1);
                mhx::launch_axpby(st, cnt, t1, 1.0, t2, dphi);
                mhx::launch_axpby(st, cnt, dphi, -1.0, t3, dphi);
                mhx::getrs_batched(st, 0, n, R, b->D, n, 0, b->ipiv, n, b->perm, n, dphi, 0, W2, 0, 1,
                                   true);
            } else {
                // sigma in B; dsigma = D^{-1} (dB - dD sigma); dphi = dS sigma + S dsigma
                mhx::launch_matmul_nr(st, n, R, dD, 0, b->B, 0, t3, 0, b->work, 0, 1);
                mhx::launch_axpby(st, cnt, dBv, -1.0, t3, t2);
                mhx::getrs_batched(st, 0, n, R, b->D, n, 0, b->ipiv, n, b->perm, n, t2, 0, W2, 0, 1,
                                   true);
                mhx::launch_matmul_nr(st, n, R, dS, 0, b->B, 0, t1, 0, b->work, 0, 1);
                mhx::launch_matmul_nr(st, n, R, b->S, 0, t2, 0, t3, 0, b->work, 0, 1);
                mhx::launch_axpby(st, cnt, t1, 1.0, t3, dphi);
            }
        }
        if (ndof > 0) {
            mhx::launch_forces(st, n, b->geom, b->dofs, ndof, b->Y, b->nR(), R, ctx->rho, b->d_ws, b->d_ks, 1, g, b->betas,
                               (int)nb, b->F, b->FK);
            mhx::launch_tangent_forces(st, n, gd, b->dofs, d_ddofs, ndof, b->Y, dphi, R, ctx->rho, w, dw, k, dk, g, d_betas,
                                       (int)nb, dF, dFK);
        }
        if (!ok(hipGetLastError(), "tangent launch") ||
            !ok(hipMemcpyAsync(F.data(), b->F, sizeof(cplx) * ndof * R, hipMemcpyDeviceToHost, st), "copy F") ||
            !ok(hipMemcpyAsync(hdF.data(), dF, sizeof(cplx) * ndof * R, hipMemcpyDeviceToHost, st), "copy dF") ||
            (nb && !ok(hipMemcpyAsync(FK.data(), b->FK, sizeof(cplx) * ndof * nb, hipMemcpyDeviceToHost, st), "copy FK")) ||
            (nb && !ok(hipMemcpyAsync(hdFK.data(), dFK, sizeof(cplx) * ndof * nb, hipMemcpyDeviceToHost, st), "copy dFK")) ||
            !ok(hipStreamSynchronize(st), "sync"))
            break;
        pack(F, FK, ndof, R, (int)nb, iw, nw, w, A_out, B_out, Fd_out, Fk_out);
        // tangents: A = Re F / w^2, B = Im F / w
        for (int r = 0; r < R; r++)
            for (int d = 0; d < ndof; d++) {
                const cplx f = F[d + (size_t)ndof * r], tf = hdF[d + (size_t)ndof * r];
                if (r < ndof) {
                    dA_out[d + (size_t)ndof * (r + (size_t)ndof * iw)] = tf.x / (w * w) - 2.0 * f.x * dw / (w * w * w);
                    dB_out[d + (size_t)ndof * (r + (size_t)ndof * iw)] = tf.y / w - f.y * dw / (w * w);
                } else {
                    const size_t o = d + (size_t)ndof * (iw + (size_t)nw * (r - ndof));
                    dFd_out[2 * o] = tf.x;
                    dFd_out[2 * o + 1] = tf.y;
                }
            }
        for (int bb = 0; bb < (int)nb; bb++)
            for (int d = 0; d < ndof; d++) {
                const size_t o = d + (size_t)ndof * (iw + (size_t)nw * bb);
                dFk_out[2 * o] = hdFK[d + (size_t)ndof * bb].x;
                dFk_out[2 * o + 1] = hdFK[d + (size_t)ndof * bb].y;
            }
    }
    flush_timing(ctx);
    cleanup();
    return rc;
}

// Frequency tangent of the coefficients of a body with an optional lid (panels [nhull, nf)) at
// any forward speeds (U = 0 included), one (U, beta, omega) group at a time: the encountered
// frequency and wavenumber with their tangents, the tangent pairs, right-hand sides, solve and
// pressure (for a moving body also dK_x sigma + K_x dsigma), then A = Re F / omega_e^2,
// B = Im F / omega_e differentiated through omega_e.
int mhx_hydro_coeffs_fs_jvp(mhx_ctx* ctx, const mhx_mesh* mesh, int64_t nhull, int64_t ndof, const double* dofs,
                            int64_t nw, const double* omegas, const double* domegas, int64_t nb, const double* betas,
                            int64_t nU, const double* Us, const double* dgrad, int gf, int direct, double* A_out,
                            double* B_out, double* Fd_out, double* Fk_out, double* dA_out, double* dB_out,
                            double* dFd_out, double* dFk_out) {
    if (!ctx) return MHX_EINVAL;
    if (!mesh) return fail(ctx, MHX_EINVAL, "fs_jvp: null mesh");
    if (gf != MHX_GF_WU && gf != MHX_GF_DELHOMMEAU)
        return fail(ctx, MHX_EINVAL, "fs_jvp: the tangent is available for GFWu and ExactGuevelDelhommeau");
    if (nhull <= 0 || nhull > mesh->nf) return fail(ctx, MHX_EDIM, "fs_jvp: nhull must be in 1..nf");
    if (nw < 0 || nb < 1 || nU < 1 || (nw && !omegas) || !betas || !Us)
        return fail(ctx, MHX_EINVAL, "fs_jvp: bad frequency, heading or speed arrays (at least one heading and one speed)");
    if (ndof < 1 || ndof + 1 > 16) return fail(ctx, MHX_EDIM, "fs_jvp: 1..15 dofs");
    bool moving = false;
    for (int64_t u = 0; u < nU; u++) {
        if (!std::isfinite(Us[u])) return fail(ctx, MHX_EINVAL, "forward speed must be finite");
        moving = moving || Us[u] != 0;
    }
    if (moving && direct)
        return fail(ctx, MHX_EINVAL, "nonzero forward speed is not supported with the direct method; use IndirectBEM (direct = false)");
    if (moving && !dgrad) return fail(ctx, MHX_EINVAL, "nonzero forward speed: null gradient of motion");
    for (int64_t w = 0; w < nw; w++)
        if (!(omegas[w] > 0) || !std::isfinite(omegas[w])) return fail(ctx, MHX_EINVAL, "omega must be positive and finite");
    if (nw && (!A_out || !B_out || !Fd_out || !Fk_out || !dA_out || !dB_out || !dFd_out || !dFk_out))
        return fail(ctx, MHX_EINVAL, "fs_jvp: null output");
    if (mesh->nf > mhx::kMaxOrder) return fail(ctx, MHX_EDIM, "fs_jvp: more than 16384 panels is not supported");
    hipSetDevice(ctx->device);
    mhx_body* b = nullptr;
    int rc = body_upload(ctx, mesh, ndof, dofs, &b);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    if (nhull < mesh->nf) mhx::launch_mark_lid(st, b->n, (int)nhull, b->geom);
    const int n = b->n, R = (int)ndof + 1;
    direct = direct ? 1 : 0;
    const bool tab = gf == MHX_GF_DELHOMMEAU;
    const size_t nn = (size_t)n * n, nR = (size_t)n * R, nv = mesh->nv;
    double *dv = nullptr, *dc = nullptr, *dn = nullptr, *da = nullptr, *dr = nullptr, *d_grad = nullptr,
           *d_inc = nullptr;
    int64_t* df = nullptr;
    mhx::PanelGeomD* gd = nullptr;
    cplx *dS = nullptr, *dD = nullptr, *dBv = nullptr, *t1 = nullptr, *t2 = nullptr, *t3 = nullptr, *dphi = nullptr,
         *dsig = nullptr, *Kx = nullptr, *psi = nullptr, *dpsi = nullptr, *dF = nullptr, *dFK = nullptr;
    auto cleanup = [&]() {
        dfree(dv); dfree(dc); dfree(dn); dfree(da); dfree(dr); dfree(d_grad); dfree(d_inc); dfree(df); dfree(gd);
        dfree(dS); dfree(dD); dfree(dBv); dfree(t1); dfree(t2); dfree(t3); dfree(dphi); dfree(dsig); dfree(Kx);
        dfree(psi); dfree(dpsi); dfree(dF); dfree(dFK);
        mhx_body_destroy(b);
    };
    if ((rc = ensure_work(ctx, b, R, 1, tab, 1)) || (rc = ensure_rankine(ctx, b, direct)) ||
        (rc = dalloc(ctx, &dv, 3 * nv)) || (rc = dalloc(ctx, &df, 4 * (size_t)n)) || (rc = dalloc(ctx, &dc, 3 * (size_t)n)) ||
        (rc = dalloc(ctx, &dn, 3 * (size_t)n)) || (rc = dalloc(ctx, &da, n)) || (rc = dalloc(ctx, &dr, n)) ||
        (rc = dalloc(ctx, &d_grad, 3 * (size_t)ndof)) || (rc = dalloc(ctx, &d_inc, 3)) || (rc = dalloc(ctx, &gd, n)) ||
        (rc = dalloc(ctx, &dS, nn)) || (rc = dalloc(ctx, &dD, nn)) || (rc = dalloc(ctx, &dBv, nR)) ||
        (rc = dalloc(ctx, &t1, nR)) || (rc = dalloc(ctx, &t2, nR)) || (rc = dalloc(ctx, &t3, nR)) ||
        (rc = dalloc(ctx, &dphi, nR)) || (rc = dalloc(ctx, &dsig, nR)) || (rc = dalloc(ctx, &psi, nR)) ||
        (rc = dalloc(ctx, &dpsi, nR)) || (rc = dalloc(ctx, &dF, (size_t)ndof * R)) || (rc = dalloc(ctx, &dFK, (size_t)ndof)) ||
        (moving && ((rc = dalloc(ctx, &Kx, nn)) || (!b->KRx && (rc = dalloc(ctx, &b->KRx, nn)))))) {
        cleanup();
        return rc;
    }
    auto up = [&](double* dst, const double* src, size_t cnt) {
        hipMemcpyAsync(dst, src, sizeof(double) * cnt, hipMemcpyHostToDevice, st);
    };
    up(dv, mesh->vertices, 3 * nv);
    up(dc, mesh->centers, 3 * (size_t)n);
    up(dn, mesh->normals, 3 * (size_t)n);
    up(da, mesh->areas, n);
    up(dr, mesh->radii, n);
    hipMemcpyAsync(df, mesh->faces, sizeof(int64_t) * 4 * n, hipMemcpyHostToDevice, st);
    if (dgrad) up(d_grad, dgrad, 3 * (size_t)ndof);
    else mhx::launch_zero(st, d_grad, sizeof(double) * 3 * ndof);
    // the mesh is fixed: geometry with zero tangents
    mhx::launch_tangent_geometry(st, n, dv, nullptr, df, dc, nullptr, dn, nullptr, da, nullptr, dr, gd);
    if (moving) mhx::launch_rankine_x(st, n, b->geom, b->KRx);
    auto ok = [&](hipError_t e, const char* what) {
        if (e == hipSuccess) return true;
        rc = fail(ctx, MHX_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
        return false;
    };
    if (!ok(hipGetLastError(), "fs_jvp set-up")) {
        cleanup();
        return rc;
    }
    const double g = ctx->g;
    const size_t nd = ndof;
    std::vector<cplx> F(nd * R), FK(nd), hdF(nd * R), hdFK(nd);
    auto aidx = [&](size_t d, size_t r, int64_t iw, int64_t iu, int64_t ib) {
        return d + nd * (r + nd * (iw + (size_t)nw * (iu + (size_t)nU * ib)));
    };
    auto fidx = [&](size_t d, int64_t iw, int64_t ib, int64_t iu) { return d + nd * (iw + (size_t)nw * (ib + (size_t)nb * iu)); };
    const int64_t cnt = (int64_t)n * R;
    cplx* W2 = b->work + 16 * b->nR();
    for (int64_t iu = 0; iu < nU && !rc; iu++) {
        const double U = Us[iu];
        for (int64_t ib = 0; ib < nb && !rc; ib++) {
            const double beta = betas[ib];
            for (int64_t iw = 0; iw < nw && !rc; iw++) {
                // encountered frequency (waves.jl:136-151) and its tangent
                const double w = omegas[iw], dw = domegas ? domegas[iw] : 0.0;
                const double k = w * w / g, dk = 2.0 * w * dw / g;
                const double dop = w - k * U * cos(beta), sgn = dop >= 0 ? 1.0 : -1.0;
                const double we = fabs(dop), dwe = sgn * (dw - dk * U * cos(beta));
                const double ke = we * we / g, dke = 2.0 * we * dwe / g;
                const double beta_e = dop >= 0 ? beta : beta + MHX_PI;
                if (!(we > 0)) {
                    rc = fail(ctx, MHX_EINVAL, "forward speed: zero encountered frequency");
                    break;
                }
                if ((rc = set_freqs(ctx, b, &we, 1))) break;
                mhx::launch_bc_fs(st, n, b->geom, b->dofs, (int)ndof, d_grad, we, U, k, w, beta, beta_e, g, b->B);
                // radiation columns: d(-i omega_e n.dof) (the U n.dgrad term is constant); the
                // diffraction column along the incident (omega, k)
                mhx::launch_tangent_bc(st, n, gd, b->dofs, nullptr, (int)ndof, nullptr, 0, we, dwe, ke, dke, g, dBv);
                mhx::launch_tangent_bc_inc(st, n, (int)nhull, gd, w, dw, k, dk, beta, beta_e, g, dBv + (size_t)n * ndof);
                if ((rc = assemble_into(ctx, b, 1, gf, direct, 1)) || (rc = check_status(ctx, b, gf))) break;
                mhx::launch_tangent_pairs(st, n, gd, ke, dke, direct, tab, b->WG, b->WR, dS, dD);
                if ((rc = solve_fp64(ctx, b, R, direct, 1))) {
                    if (rc == MHX_ESINGULAR) fail(ctx, rc, ctx->err);
                    break;
                }
                if (direct) {
                    // dphi = D^{-1} (dS B + S dB - dD phi)
                    mhx::launch_matmul_nr(st, n, R, dS, 0, b->B, 0, t1, 0, b->work, 0, 1);
                    mhx::launch_matmul_nr(st, n, R, b->S, 0, dBv, 0, t2, 0, b->work, 0, 1);
                    mhx::launch_matmul_nr(st, n, R, dD, 0, b->Y, 0, t3, 0, b->work, 0, 1);
                    mhx::launch_axpby(st, cnt, t1, 1.0, t2, dphi);
                    mhx::launch_axpby(st, cnt, dphi, -1.0, t3, dphi);
                    mhx::getrs_batched(st, 0, n, R, b->D, n, 0, b->ipiv, n, b->perm, n, dphi, 0, W2, 0, 1, true);
                } else {
                    // sigma in B: dsigma = K^{-1} (dB - dK sigma); dphi = dS sigma + S dsigma
                    mhx::launch_matmul_nr(st, n, R, dD, 0, b->B, 0, t3, 0, b->work, 0, 1);
                    mhx::launch_axpby(st, cnt, dBv, -1.0, t3, dsig);
                    mhx::getrs_batched(st, 0, n, R, b->D, n, 0, b->ipiv, n, b->perm, n, dsig, 0, W2, 0, 1, true);
                    mhx::launch_matmul_nr(st, n, R, dS, 0, b->B, 0, t1, 0, b->work, 0, 1);
                    mhx::launch_matmul_nr(st, n, R, b->S, 0, dsig, 0, t3, 0, b->work, 0, 1);
                    mhx::launch_axpby(st, cnt, t1, 1.0, t3, dphi);
                }
                if (U != 0) {
                    // phi_eff = phi - i (U / omega_e) psi, psi = K_x sigma (solve.jl:35-45);
                    // dpsi = dK_x sigma + K_x dsigma
                    mhx::launch_kx_combine(st, n, b->geom, b->KRx, ke, tab, tab ? b->WR : nullptr, Kx);
                    mhx::launch_tangent_kx(st, n, gd, ke, dke, tab, b->WG, b->WR, dS);
                    mhx::launch_matmul_nr(st, n, R, Kx, 0, b->B, 0, psi, 0, b->work, 0, 1);
                    mhx::launch_matmul_nr(st, n, R, dS, 0, b->B, 0, t1, 0, b->work, 0, 1);
                    mhx::launch_matmul_nr(st, n, R, Kx, 0, dsig, 0, t2, 0, b->work, 0, 1);
                    mhx::launch_axpby(st, cnt, t1, 1.0, t2, dpsi);
                    // dphi_eff = dphi - i U (dpsi / omega_e - psi domega_e / omega_e^2)
                    mhx::launch_caxpy(st, cnt, 0.0, -U / we, dpsi, dphi);
                    mhx::launch_caxpy(st, cnt, 0.0, U * dwe / (we * we), psi, dphi);
                    mhx::launch_caxpy(st, cnt, 0.0, -U / we, psi, b->Y);
                }
                // forces at omega_e; Froude-Krylov with the incident omega, k along beta_e
                // (d_inc = {beta_e, omega, k})
                const double hinc[3] = {beta_e, w, k};
                hipMemcpyAsync(d_inc, hinc, sizeof hinc, hipMemcpyHostToDevice, st);
                mhx::launch_forces(st, n, b->geom, b->dofs, (int)ndof, b->Y, b->nR(), R, ctx->rho, b->d_ws, b->d_ks, 1, g,
                                   nullptr, 0, b->F, b->FK);
                mhx::launch_tangent_forces(st, n, gd, b->dofs, nullptr, (int)ndof, b->Y, dphi, R, ctx->rho, we, dwe, ke,
                                           dke, g, nullptr, 0, dF, dFK);
                mhx::launch_forces(st, n, b->geom, b->dofs, (int)ndof, b->Y, b->nR(), 0, ctx->rho, d_inc + 1, d_inc + 2, 1, g,
                                   d_inc, 1, b->F, b->FK);
                mhx::launch_tangent_forces(st, n, gd, b->dofs, nullptr, (int)ndof, b->Y, dphi, 0, ctx->rho, w, dw, k, dk,
                                           g, d_inc, 1, dF, dFK);
                const bool copied =
                    ok(hipGetLastError(), "fs_jvp launch") &&
                    ok(hipMemcpyAsync(F.data(), b->F, sizeof(cplx) * nd * R, hipMemcpyDeviceToHost, st), "copy F") &&
                    ok(hipMemcpyAsync(hdF.data(), dF, sizeof(cplx) * nd * R, hipMemcpyDeviceToHost, st), "copy dF") &&
                    ok(hipMemcpyAsync(FK.data(), b->FK, sizeof(cplx) * nd, hipMemcpyDeviceToHost, st), "copy FK") &&
                    ok(hipMemcpyAsync(hdFK.data(), dFK, sizeof(cplx) * nd, hipMemcpyDeviceToHost, st), "copy dFK") &&
                    ok(hipStreamSynchronize(st), "sync");
                if (!copied) break;
                for (size_t r = 0; r < (size_t)R; r++)
                    for (size_t d = 0; d < nd; d++) {
                        const cplx f = F[d + nd * r], tf = hdF[d + nd * r];
                        if (r < nd) {
                            A_out[aidx(d, r, iw, iu, ib)] = f.x / (we * we);
                            B_out[aidx(d, r, iw, iu, ib)] = f.y / we;
                            dA_out[aidx(d, r, iw, iu, ib)] = tf.x / (we * we) - 2.0 * f.x * dwe / (we * we * we);
                            dB_out[aidx(d, r, iw, iu, ib)] = tf.y / we - f.y * dwe / (we * we);
                        } else {
                            const size_t o = fidx(d, iw, ib, iu);
                            Fd_out[2 * o] = f.x;
                            Fd_out[2 * o + 1] = f.y;
                            dFd_out[2 * o] = tf.x;
                            dFd_out[2 * o + 1] = tf.y;
                            Fk_out[2 * o] = FK[d].x;
                            Fk_out[2 * o + 1] = FK[d].y;
                            dFk_out[2 * o] = hdFK[d].x;
                            dFk_out[2 * o + 1] = hdFK[d].y;
                        }
                    }
            }
        }
    }
    flush_timing(ctx);
    cleanup();
    return rc;
}

// Gradient of J = <Abar, A> + <Bbar, B> + Re<Fdbar, Fd> + Re<Fkbar, Fk> (reverse mode): per
// frequency the primal omega-solve (FP64 LU), the pressure pullback, the transposed solve on the
// same LU factors, and the right-hand-side and pair pullbacks accumulated into the gradients.
int mhx_hydro_coeffs_vjp(mhx_ctx* ctx, const mhx_mesh* mesh, int64_t ndof, const double* dofs, int64_t nw,
                         const double* omegas, int64_t nb, const double* betas, int gf, int direct, const double* A_bar,
                         const double* B_bar, const double* Fd_bar, const double* Fk_bar, double* A_out, double* B_out,
                         double* Fd_out, double* Fk_out, const mhx_mesh_grad* grad) {
    if (!ctx) return MHX_EINVAL;
    if (gf != MHX_GF_WU && gf != MHX_GF_DELHOMMEAU)
        return fail(ctx, MHX_EINVAL, "coeffs_vjp: the adjoint is available for GFWu and ExactGuevelDelhommeau");
    if (!grad) return fail(ctx, MHX_EINVAL, "coeffs_vjp: null gradient struct");
    if (nw < 0 || nb < 0 || (nw && !omegas) || (nb && !betas)) return fail(ctx, MHX_EINVAL, "coeffs_vjp: bad frequency/heading arrays");
    if ((ndof && (!A_out || !B_out)) || (nb && ndof && (!Fd_out || !Fk_out)))
        return fail(ctx, MHX_EINVAL, "coeffs_vjp: null output");
    for (int64_t w = 0; w < nw; w++)
        if (!(omegas[w] > 0) || !std::isfinite(omegas[w])) return fail(ctx, MHX_EINVAL, "omega must be positive and finite");
    const int64_t R64 = ndof + nb;
    if (R64 <= 0) return fail(ctx, MHX_EDIM, "coeffs_vjp: no radiating dof and no heading");
    if (R64 > 16) return fail(ctx, MHX_EDIM, "coeffs_vjp: at most 16 right-hand sides (dofs + headings)");
    if (mesh && mesh->nf > mhx::kMaxOrder) return fail(ctx, MHX_EDIM, "coeffs_vjp: more than 16384 panels is not supported");
    hipSetDevice(ctx->device);
    mhx_body* b = nullptr;
    int rc = body_upload(ctx, mesh, ndof, dofs, &b);
    if (rc) return rc;
    const int n = b->n, R = (int)R64;
    direct = direct ? 1 : 0;
    hipStream_t st = ctx->stream;
    const size_t nv = mesh->nv, nR = (size_t)n * R;
    const bool tab = gf == MHX_GF_DELHOMMEAU;
    double *dv = nullptr, *dc = nullptr, *dn = nullptr, *da = nullptr, *dr = nullptr;
    int64_t* df = nullptr;
    double *gV = nullptr, *gC = nullptr, *gN = nullptr, *gA = nullptr, *gD = nullptr, *wk = nullptr;
    cplx *Fb = nullptr, *FKb = nullptr, *phib = nullptr, *t1 = nullptr, *t2 = nullptr;
    auto cleanup = [&]() {
        dfree(dv); dfree(dc); dfree(dn); dfree(da); dfree(dr); dfree(df); dfree(gV); dfree(gC); dfree(gN); dfree(gA);
        dfree(gD); dfree(wk); dfree(Fb); dfree(FKb); dfree(phib); dfree(t1); dfree(t2);
        mhx_body_destroy(b);
    };
    const size_t nd1 = ndof ? ndof : 1, nb1 = nb ? nb : 1;
    if ((rc = ensure_work(ctx, b, R, (int)nb, tab, 1)) || (rc = ensure_rankine(ctx, b, direct)) ||
        (rc = dalloc(ctx, &dv, 3 * nv)) || (rc = dalloc(ctx, &df, 4 * (size_t)n)) || (rc = dalloc(ctx, &dc, 3 * (size_t)n)) ||
        (rc = dalloc(ctx, &dn, 3 * (size_t)n)) || (rc = dalloc(ctx, &da, n)) || (rc = dalloc(ctx, &dr, n)) ||
        (rc = dalloc(ctx, &gV, 3 * nv)) || (rc = dalloc(ctx, &gC, 3 * (size_t)n)) || (rc = dalloc(ctx, &gN, 3 * (size_t)n)) ||
        (rc = dalloc(ctx, &gA, n)) || (rc = dalloc(ctx, &gD, 3 * (size_t)n * nd1)) || (rc = dalloc(ctx, &wk, 2)) ||
        (rc = dalloc(ctx, &Fb, nd1 * R)) || (rc = dalloc(ctx, &FKb, nd1 * nb1)) || (rc = dalloc(ctx, &phib, nR)) ||
        (rc = dalloc(ctx, &t1, nR)) || (rc = dalloc(ctx, &t2, nR))) {
        cleanup();
        return rc;
    }
    auto ok = [&](hipError_t e, const char* what) {
        if (e == hipSuccess) return true;
        rc = fail(ctx, MHX_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
        return false;
    };
    hipMemcpyAsync(dv, mesh->vertices, sizeof(double) * 3 * nv, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(df, mesh->faces, sizeof(int64_t) * 4 * n, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dc, mesh->centers, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dn, mesh->normals, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(da, mesh->areas, sizeof(double) * n, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dr, mesh->radii, sizeof(double) * n, hipMemcpyHostToDevice, st);
    mhx::launch_zero(st, gV, sizeof(double) * 3 * nv);
    mhx::launch_zero(st, gC, sizeof(double) * 3 * n);
    mhx::launch_zero(st, gN, sizeof(double) * 3 * n);
    mhx::launch_zero(st, gA, sizeof(double) * n);
    mhx::launch_zero(st, gD, sizeof(double) * 3 * n * nd1);
    if (nb) hipMemcpyAsync(b->betas, betas, sizeof(double) * nb, hipMemcpyHostToDevice, st);
    const mhx::AdjointGrad gr{gV, gC, gN, gA, gD, wk};
    std::vector<cplx> F((size_t)ndof * R), FK((size_t)ndof * nb1), hFb(nd1 * R), hFKb(nd1 * nb1);
    const double g = ctx->g;
    for (int64_t iw = 0; iw < nw && !rc; iw++) {
        const double w = omegas[iw], k = w * w / g;  // compute_wavenumber (src/waves.jl:133-134)
        if ((rc = set_freqs(ctx, b, omegas + iw, 1))) break;
        mhx::launch_bc(st, n, b->geom, b->dofs, ndof, b->betas, (int)nb, b->d_ws, b->d_ks, 1, g, b->B, b->nR());
        if ((rc = assemble_into(ctx, b, 1, gf, direct, 1)) || (rc = check_status(ctx, b, gf))) break;
        if ((rc = solve_fp64(ctx, b, R, direct, 1))) {
            if (rc == MHX_ESINGULAR) fail(ctx, rc, ctx->err);
            break;
        }
        if (ndof > 0)
            mhx::launch_forces(st, n, b->geom, b->dofs, ndof, b->Y, b->nR(), R, ctx->rho, b->d_ws, b->d_ks, 1, g, b->betas,
                               (int)nb, b->F, b->FK);
        if (!ok(hipMemcpyAsync(F.data(), b->F, sizeof(cplx) * ndof * R, hipMemcpyDeviceToHost, st), "copy F") ||
            (nb && !ok(hipMemcpyAsync(FK.data(), b->FK, sizeof(cplx) * ndof * nb, hipMemcpyDeviceToHost, st), "copy FK")) ||
            !ok(hipStreamSynchronize(st), "sync"))
            break;
        pack(F, FK, ndof, R, (int)nb, iw, nw, w, A_out, B_out, Fd_out, Fk_out);
        // force cotangents from the packing (problems_and_results.jl:207-208): A = Re F / w^2,
        // B = Im F / w, Fd = F (diffraction columns), Fk = FK
        double wbar = 0.0;
        for (int r = 0; r < R; r++)
            for (int d = 0; d < ndof; d++) {
                cplx fb = make_double2(0.0, 0.0);
                const cplx f = F[d + (size_t)ndof * r];
                if (r < ndof) {
                    const size_t o = d + (size_t)ndof * (r + (size_t)ndof * iw);
                    const double ab = A_bar ? A_bar[o] : 0.0, bb = B_bar ? B_bar[o] : 0.0;
                    fb = make_double2(ab / (w * w), bb / w);
                    wbar += -2.0 * ab * f.x / (w * w * w) - bb * f.y / (w * w);
                } else if (Fd_bar) {
                    const size_t o = d + (size_t)ndof * (iw + (size_t)nw * (r - ndof));
                    fb = make_double2(Fd_bar[2 * o], Fd_bar[2 * o + 1]);
                }
                hFb[d + (size_t)ndof * r] = fb;
            }
        for (int bb = 0; bb < (int)nb; bb++)
            for (int d = 0; d < ndof; d++) {
                const size_t o = d + (size_t)ndof * (iw + (size_t)nw * bb);
                hFKb[d + (size_t)ndof * bb] = Fk_bar ? make_double2(Fk_bar[2 * o], Fk_bar[2 * o + 1]) : make_double2(0.0, 0.0);
            }
        double hwk[2] = {0.0, 0.0};
        cplx* W2 = b->work + 16 * b->nR();
        const int64_t cnt = (int64_t)n * R;
        hipMemcpyAsync(Fb, hFb.data(), sizeof(cplx) * nd1 * R, hipMemcpyHostToDevice, st);
        hipMemcpyAsync(FKb, hFKb.data(), sizeof(cplx) * nd1 * nb1, hipMemcpyHostToDevice, st);
        mhx::launch_zero(st, wk, 2 * sizeof(double));
        {
            Phase ph(ctx, "adjoint");
            mhx::launch_adjoint_forces(st, n, b->geom, b->dofs, ndof, b->Y, R, Fb, FKb, ctx->rho, w, k, g, b->betas,
                                       (int)nb, phib, gr);
            const cplx *U, *V, *V2;
            cplx* Bbar;
            if (direct) {
                // lambda = D^{-H} phibar; Bbar = S^H lambda; Sbar = lambda B^H; Dbar = -lambda phi^H
                mhx::getrs_batched(st, 2, n, R, b->D, n, 0, b->ipiv, n, b->perm, n, phib, 0, W2, 0, 1,
                                   true);
                mhx::launch_matmul_h(st, n, R, b->S, phib, t1);
                mhx::launch_scale(st, cnt, -1.0, phib, t2);
                U = phib;
                V = b->B;
                V2 = b->Y;
                Bbar = t1;
            } else {
                // sigmabar = S^H phibar; mu = D^{-H} sigmabar; Sbar = phibar sigma^H; Dbar = -mu sigma^H; Bbar = mu
                mhx::launch_matmul_h(st, n, R, b->S, phib, t1);
                mhx::getrs_batched(st, 2, n, R, b->D, n, 0, b->ipiv, n, b->perm, n, t1, 0, W2, 0, 1,
                                   true);
                mhx::launch_scale(st, cnt, -1.0, t1, t2);
                U = phib;
                V = b->B;
                V2 = b->B;
                Bbar = t1;
            }
            mhx::launch_adjoint_bc(st, n, b->geom, b->dofs, ndof, b->betas, (int)nb, w, k, g, Bbar, gr);
            mhx::launch_adjoint_pairs(st, n, dv, df, dc, dn, da, dr, k, direct, tab, b->WG, b->WR, U, V, t2, V2, R, gr);
        }
        if (!ok(hipGetLastError(), "adjoint launch") ||
            !ok(hipMemcpyAsync(hwk, wk, 2 * sizeof(double), hipMemcpyDeviceToHost, st), "copy wk") ||
            !ok(hipStreamSynchronize(st), "sync"))
            break;
        if (grad->omegas) grad->omegas[iw] = wbar + hwk[0] + hwk[1] * 2.0 * w / g;
    }
    if (!rc) {
        const bool good = (!grad->vertices || ok(hipMemcpyAsync(grad->vertices, gV, sizeof(double) * 3 * nv, hipMemcpyDeviceToHost, st), "copy")) &&
                          (!grad->centers || ok(hipMemcpyAsync(grad->centers, gC, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st), "copy")) &&
                          (!grad->normals || ok(hipMemcpyAsync(grad->normals, gN, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st), "copy")) &&
                          (!grad->areas || ok(hipMemcpyAsync(grad->areas, gA, sizeof(double) * n, hipMemcpyDeviceToHost, st), "copy")) &&
                          (!grad->dofs || !ndof || ok(hipMemcpyAsync(grad->dofs, gD, sizeof(double) * 3 * n * ndof, hipMemcpyDeviceToHost, st), "copy"));
        if (good) ok(hipStreamSynchronize(st), "sync");
    }
    flush_timing(ctx);
    cleanup();
    return rc;
}

// Several bodies of uneven panel counts, each with its own frequency list, in one sweep: the
// assembly of every (body, frequency) item runs as one queue launch per stage, the LU is batched
// per body, and with MHX_SWEEP_GRAPH the whole sequence is captured once and replayed by later
// calls with the same bodies and frequency counts.
int mhx_sweep_coeffs(mhx_ctx* ctx, const mhx_sweep_body* in, int64_t nbodies, int64_t nb, const double* betas, int gf,
                     int direct, int flags) {
    if (!ctx) return MHX_EINVAL;
    if (!in || nbodies <= 0) return fail(ctx, MHX_EINVAL, "sweep: no bodies");
    if (gf != MHX_GF_WU && gf != MHX_GF_DELHOMMEAU && gf != MHX_GF_SURROGATE)
        return fail(ctx, MHX_EINVAL, "sweep: the multi-body sweep runs GFWu, ExactGuevelDelhommeau or the surrogate");
    if (gf == MHX_GF_SURROGATE && !ctx->surr_ready)
        return fail(ctx, MHX_EINVAL, "sweep: the surrogate needs a trained network (mhx_surrogate_train/load)");
    if (nb < 0 || (nb && !betas)) return fail(ctx, MHX_EINVAL, "sweep: bad heading array");
    if (flags & ~MHX_SWEEP_GRAPH) return fail(ctx, MHX_EINVAL, "sweep: unknown flag");
    direct = direct ? 1 : 0;
    const bool table = gf == MHX_GF_DELHOMMEAU, surr = gf == MHX_GF_SURROGATE, graph = (flags & MHX_SWEEP_GRAPH) != 0;
    const bool verified = ctx->pivoting == MHX_PIVOT_VERIFIED;
    const bool mixed = use_mixed(ctx, gf), split = use_split(ctx, gf);
    for (int64_t ib = 0; ib < nbodies; ib++) {
        const mhx_sweep_body& e = in[ib];
        if (!e.body || e.body->ctx != ctx) return fail(ctx, MHX_EINVAL, "sweep: body is null or from another context");
        for (int64_t jb = 0; jb < ib; jb++)
            if (in[jb].body == e.body) return fail(ctx, MHX_EINVAL, "sweep: a body is listed twice");
        const int64_t R = e.body->ndof + nb;
        if (R <= 0 || R > 16) return fail(ctx, MHX_EDIM, "sweep: ndof + nbeta must be in 1..16 for every body");
        if (e.body->n > mhx::kMaxOrder) return fail(ctx, MHX_EDIM, "sweep: more than 16384 panels is not supported");
        if (e.nw < 0 || (e.nw && !e.omegas)) return fail(ctx, MHX_EINVAL, "sweep: bad frequency array");
        for (int64_t w = 0; w < e.nw; w++)
            if (!(e.omegas[w] > 0) || !std::isfinite(e.omegas[w]))
                return fail(ctx, MHX_EINVAL, "omega must be positive and finite");
        if (e.body->ndof && e.nw && (!e.A_out || !e.B_out || (nb && (!e.Fd_out || !e.Fk_out))))
            return fail(ctx, MHX_EINVAL, "sweep: null output");
    }
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    // chunk plan: K chunks, each body contributing ceil(nw / K) items per chunk
    double total = 0.0;
    int64_t K = 1;
    for (int64_t ib = 0; ib < nbodies; ib++) {
        const double n = in[ib].body->n;
        double per = n * n * sizeof(cplx) * ((table ? 4 : 2) + (verified && !mixed ? 1 : 0)) + 20.0 * n * 16 * sizeof(cplx);
        if (mixed) per += n * n * sizeof(cplxf) + 2.0 * n * 16 * sizeof(cplx);   // FP32 factors, Rr, Tr
        if (surr) per += 8.0 * mhx::surr_fallback_capacity((int)n, 1);            // fallback pair list
        total += per * in[ib].nw;
        K = std::max<int64_t>(K, (in[ib].nw + ctx->max_batch - 1) / ctx->max_batch);
    }
    K = std::max<int64_t>(K, (int64_t)std::ceil(total / ctx->batch_bytes));
    std::vector<int> per(nbodies);
    int rc;
    for (int64_t ib = 0; ib < nbodies; ib++) {
        mhx_body* b = in[ib].body;
        per[ib] = (int)((in[ib].nw + K - 1) / K);
        if ((rc = ensure_work(ctx, b, b->ndof + (int)nb, (int)nb, table, std::max(1, per[ib]))) ||
            (rc = ensure_rankine(ctx, b, direct)) ||
            (verified && !mixed && (rc = dalloc(ctx, &b->D0, (size_t)b->nn() * b->cap_batch))))
            return rc;
        if (mixed && ((rc = dalloc(ctx, &b->Df, (size_t)b->nn() * b->cap_batch)) ||
                      (rc = dalloc(ctx, &b->Rr, (size_t)b->nR() * b->cap_batch)) ||
                      (rc = dalloc(ctx, &b->Tr, (size_t)b->nR() * b->cap_batch)) ||
                      (rc = dalloc(ctx, &b->norms, 2 * (size_t)b->cap_batch)) ||
                      (split && (rc = dalloc(ctx, &b->Ps, (size_t)mhx::split_work(b->n) * b->cap_batch)))))
            return rc;
        if (nb && (rc = hipMemcpyAsync(b->betas, betas, sizeof(double) * nb, hipMemcpyHostToDevice, st)) != hipSuccess)
            return fail(ctx, MHX_EDEVICE, "sweep: heading upload");
    }
    // the item queue (pointers into each body's buffers)
    std::vector<mhx::SweepItem> items;
    std::vector<int> cfirst, ccount;
    std::vector<int64_t> cunits, ctunits, csunits;
    int64_t fb_cap = 1;
    for (int64_t c = 0; c < K; c++) {
        cfirst.push_back((int)items.size());
        int64_t u = 0, tu = 0, su = 0, pairs = 0;
        for (int64_t ib = 0; ib < nbodies; ib++) {
            mhx_body* b = in[ib].body;
            const int64_t w0 = c * per[ib];
            const int64_t bt = std::max<int64_t>(0, std::min<int64_t>(per[ib], in[ib].nw - w0));
            for (int64_t q = 0; q < bt; q++) {
                mhx::SweepItem it{};
                it.geom = b->geom;
                it.SR = b->SR;
                it.DR = b->DR[direct];
                it.S = b->S + q * b->nn();
                it.D = b->D + q * b->nn();
                it.WG = table ? b->WG + q * b->nn() : nullptr;
                it.WR = table ? b->WR + q * b->nn() : nullptr;
                it.Df = surr && mixed ? b->Df + q * b->nn() : nullptr;
                it.sd32 = surr && mixed ? 1 : 0;  // the mixed solver reads S and D in FP32
                it.k = b->d_ks + q;
                it.n = b->n;
                it.unit0 = u;
                it.tunit0 = tu;
                it.sunit0 = su;
                u += mhx::sweep_combine_units(b->n);
                tu += mhx::sweep_table_units(b->n);
                su += mhx::sweep_surr_units(b->n);
                pairs += mhx::surr_fallback_capacity(b->n, 1);
                items.push_back(it);
            }
        }
        ccount.push_back((int)items.size() - cfirst.back());
        cunits.push_back(u);
        ctunits.push_back(tu);
        csunits.push_back(su);
        fb_cap = std::max(fb_cap, pairs);
    }
    std::vector<mhx_body*> bodies(nbodies);
    std::vector<int64_t> nws(nbodies);
    std::vector<uint64_t> gens(nbodies);
    for (int64_t ib = 0; ib < nbodies; ib++) {
        in[ib].body->sd32 = surr && mixed;  // the queue's pair kernel stores S and D in FP32 (items' sd32)
        bodies[ib] = in[ib].body;
        nws[ib] = in[ib].nw;
        gens[ib] = in[ib].body->gen;
    }
    // graph mode keeps its state (and graph) in the context for replay; an eager sweep uses its own
    // state and leaves a captured sweep in place
    SweepState local;
    SweepState* s = graph ? ctx->sweep : &local;
    const bool reuse = graph && s && s->exec && s->bodies == bodies && s->nws == nws && s->gens == gens &&
                       s->rho == ctx->rho && s->g == ctx->g && s->verified == verified && s->mixed == mixed &&
                       s->split == split && s->nb == nb && s->gf == gf &&
                       s->direct == direct && s->K == K && s->h_items.size() == items.size() &&
                       std::memcmp(s->h_items.data(), items.data(), items.size() * sizeof(mhx::SweepItem)) == 0;
    auto abandon = [&]() {
        if (graph) drop_sweep(ctx);
        else local.release();
    };
    if (!reuse) {
        if (graph) {
            drop_sweep(ctx);
            s = ctx->sweep = new SweepState();
        }
        s->bodies = bodies;
        s->nws = nws;
        s->gens = gens;
        s->rho = ctx->rho;
        s->g = ctx->g;
        s->verified = verified;
        s->mixed = mixed;
        s->split = split;
        s->flag_off.assign(nbodies, 0);
        int64_t nflags = 0;
        for (int64_t ib = 0; ib < nbodies; ib++) {
            s->flag_off[ib] = nflags;
            nflags += in[ib].nw;
        }
        s->per_chunk = per;
        s->K = (int)K;
        s->nb = nb;
        s->gf = gf;
        s->direct = direct;
        s->h_items = items;
        s->chunk_first = cfirst;
        s->chunk_count = ccount;
        s->chunk_units = cunits;
        s->chunk_tunits = ctunits;
        s->chunk_sunits = csunits;
        s->fb_cap = surr ? fb_cap : 0;
        s->w_all.assign(nbodies, nullptr);
        s->k_all.assign(nbodies, nullptr);
        s->F_all.assign(nbodies, nullptr);
        s->FK_all.assign(nbodies, nullptr);
        rc = MHX_OK;
        for (int64_t ib = 0; ib < nbodies && !rc; ib++) {
            const size_t nw1 = in[ib].nw ? in[ib].nw : 1, nd1 = bodies[ib]->ndof ? bodies[ib]->ndof : 1;
            if ((rc = dalloc(ctx, &s->w_all[ib], nw1)) || (rc = dalloc(ctx, &s->k_all[ib], nw1)) ||
                (rc = dalloc(ctx, &s->F_all[ib], nd1 * (bodies[ib]->ndof + nb) * nw1)) ||
                (rc = dalloc(ctx, &s->FK_all[ib], nd1 * (nb ? nb : 1) * nw1)))
                break;
        }
        if (rc || (rc = dalloc(ctx, &s->items, items.size())) || (rc = dalloc(ctx, &s->counters, (size_t)K)) ||
            (rc = dalloc(ctx, &s->status, 4)) || (rc = dalloc(ctx, &s->sticky, (size_t)nbodies)) ||
            (verified && (rc = dalloc(ctx, &s->pflags, (size_t)std::max<int64_t>(1, nflags)))) ||
            (mixed && (rc = dalloc(ctx, &s->rnorms, 2 * (size_t)std::max<int64_t>(1, nflags)))) ||
            (surr && ((rc = dalloc(ctx, &s->fb_list, (size_t)fb_cap)) || (rc = dalloc(ctx, &s->fb_count, 1))))) {
            abandon();
            return rc;
        }
        hipMemcpyAsync(s->items, items.data(), sizeof(mhx::SweepItem) * items.size(), hipMemcpyHostToDevice, st);
    }
    // this call's frequencies (outside any graph: a replay reads them from these arrays)
    for (int64_t ib = 0; ib < nbodies; ib++) {
        if (!in[ib].nw) continue;
        std::vector<double> k(in[ib].nw);
        for (int64_t w = 0; w < in[ib].nw; w++) k[w] = in[ib].omegas[w] * in[ib].omegas[w] / ctx->g;  // waves.jl:133-134
        hipMemcpyAsync(s->w_all[ib], in[ib].omegas, sizeof(double) * in[ib].nw, hipMemcpyHostToDevice, st);
        hipMemcpyAsync(s->k_all[ib], k.data(), sizeof(double) * in[ib].nw, hipMemcpyHostToDevice, st);
        if (hipStreamSynchronize(st) != hipSuccess) return fail(ctx, MHX_EDEVICE, "sweep: frequency upload");
    }
    if (!reuse && !items.empty() && !surr) {
        std::string key = std::string(table ? "exact" : "wu") + (direct ? "/d" : "/i");
        for (int64_t ib = 0; ib < nbodies; ib++) key += "/" + std::to_string(bodies[ib]->n) + "x" + std::to_string(per[ib]);
        // wavenumbers of chunk 0 in place for the tuner's trial launches
        for (int64_t ib = 0; ib < nbodies; ib++)
            if (per[ib] && in[ib].nw)
                hipMemcpyAsync(bodies[ib]->d_ks, s->k_all[ib], sizeof(double) * std::min<int64_t>(per[ib], in[ib].nw),
                               hipMemcpyDeviceToDevice, st);
        s->rows = sweep_tune(ctx, *s, key);
    }
    hipError_t e = hipSuccess;
    if (graph) {
        if (!reuse) {
            e = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
            if (e == hipSuccess) {
                sweep_enqueue(ctx, *s);
                e = hipStreamEndCapture(st, &s->graph);
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                abandon();
                return fail(ctx, MHX_EDEVICE, std::string("sweep: graph capture: ") + hipGetErrorString(e));
            }
        }
        Phase ph(ctx, "sweep");
        e = hipGraphLaunch(s->exec, st);
    } else {
        Phase ph(ctx, "sweep");
        sweep_enqueue(ctx, *s);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        abandon();
        return fail(ctx, MHX_EDEVICE, std::string("sweep: launch: ") + hipGetErrorString(e));
    }
    // statuses, then the coefficients
    unsigned long long stq[4] = {0, 0, 0, 0};
    std::vector<int> sticky(nbodies, 0);
    hipMemcpyAsync(stq, s->status, sizeof stq, hipMemcpyDeviceToHost, st);
    hipMemcpyAsync(sticky.data(), s->sticky, sizeof(int) * nbodies, hipMemcpyDeviceToHost, st);
    if ((e = hipStreamSynchronize(st)) != hipSuccess) {
        abandon();
        return fail(ctx, MHX_EDEVICE, std::string("sweep: ") + hipGetErrorString(e));
    }
    flush_timing(ctx);
    for (int q = 0; q < 3; q++) ctx->quad_work[q] += stq[q + 1];
    rc = MHX_OK;
    if (stq[0]) rc = fail(ctx, MHX_EQUAD, "exact Delhommeau quadrature exceeded its interval budget");
    for (int64_t ib = 0; ib < nbodies && !rc; ib++) {
        if (sticky[ib]) rc = fail(ctx, MHX_ESINGULAR, "LU: exactly zero pivot at column " + std::to_string(sticky[ib]));
    }
    for (int64_t ib = 0; ib < nbodies && !rc; ib++) {
        mhx_body* b = bodies[ib];
        const int ndof = b->ndof, R = ndof + (int)nb, nw = (int)in[ib].nw;
        if (!ndof || !nw) continue;
        std::vector<cplx> F((size_t)ndof * R * nw), FK((size_t)ndof * (nb ? nb : 1) * nw);
        hipMemcpyAsync(F.data(), s->F_all[ib], sizeof(cplx) * F.size(), hipMemcpyDeviceToHost, st);
        if (nb) hipMemcpyAsync(FK.data(), s->FK_all[ib], sizeof(cplx) * ndof * nb * nw, hipMemcpyDeviceToHost, st);
        if (hipStreamSynchronize(st) != hipSuccess) {
            rc = fail(ctx, MHX_EDEVICE, "sweep: copy back");
            break;
        }
        for (int q = 0; q < nw; q++) {
            std::vector<cplx> Fq(F.begin() + (size_t)q * ndof * R, F.begin() + (size_t)(q + 1) * ndof * R);
            std::vector<cplx> FKq(FK.begin() + (size_t)q * ndof * nb, FK.begin() + (size_t)(q + 1) * ndof * nb);
            pack(Fq, FKq, ndof, R, (int)nb, q, nw, in[ib].omegas[q], in[ib].A_out, in[ib].B_out, in[ib].Fd_out,
                 in[ib].Fk_out);
        }
    }
    // every (body, frequency) whose factorization broke LAPACK's rule (verified diagonal pivots),
    // or whose mixed-precision solution missed the residual tolerance after its refinement step,
    // is solved again on its own (row interchanges; FP64 LU for a missed residual) and its
    // coefficients replaced
    if (!rc && (s->verified || s->mixed)) {
        int64_t nflags = 0;
        for (int64_t ib = 0; ib < nbodies; ib++) nflags += in[ib].nw;
        std::vector<int> fl((size_t)std::max<int64_t>(1, nflags), 0);
        std::vector<double> rn(2 * (size_t)std::max<int64_t>(1, nflags), 0.0);
        if (nflags && ((s->verified && hipMemcpyAsync(fl.data(), s->pflags, sizeof(int) * nflags, hipMemcpyDeviceToHost,
                                                      st) != hipSuccess) ||
                       (s->mixed && hipMemcpyAsync(rn.data(), s->rnorms, sizeof(double) * 2 * nflags,
                                                   hipMemcpyDeviceToHost, st) != hipSuccess) ||
                       hipStreamSynchronize(st) != hipSuccess))
            rc = fail(ctx, MHX_EDEVICE, "sweep: pivot flags and residuals");
        const int saved_piv = ctx->pivoting, saved_lu = ctx->lu_mode;
        for (int64_t ib = 0; ib < nbodies && !rc; ib++) {
            mhx_body* b = bodies[ib];
            const int ndof = b->ndof, nw = (int)in[ib].nw;
            for (int w = 0; w < nw && !rc; w++) {
                const int64_t f = s->flag_off[ib] + w;
                const bool pivot = s->verified && fl[f];
                bool missed = false;
                if (s->mixed) {
                    const double rel = rn[2 * f + 1] > 0 ? std::sqrt(rn[2 * f] / rn[2 * f + 1]) : std::sqrt(rn[2 * f]);
                    missed = !(rel <= kRefineTol) || ctx->refine_limit == 0;
                    if (!missed) {
                        ctx->mixed_solves++;
                        ctx->refine_steps++;
                        ctx->max_residual = std::max(ctx->max_residual, rel);
                    }
                }
                if (s->verified && !pivot) ctx->lu_diagonal++;
                if (!pivot && !missed) continue;
                if (missed) ctx->fp64_fallbacks++;
                if (!ndof) continue;
                const size_t nd = ndof, nb1 = nb ? nb : 1;
                std::vector<double> A1(nd * nd), B1(nd * nd), Fd1(2 * nd * nb1), Fk1(2 * nd * nb1);
                if (pivot) ctx->pivoting = MHX_PIVOT_ALWAYS;
                if (missed) ctx->lu_mode = MHX_LU_FP64;
                rc = mhx_body_coeffs(ctx, b, 1, in[ib].omegas + w, nb, betas, gf, direct, MHX_MODE_PER_OMEGA, A1.data(),
                                     B1.data(), Fd1.data(), Fk1.data());
                ctx->pivoting = saved_piv;
                ctx->lu_mode = saved_lu;
                if (rc) break;
                for (size_t e = 0; e < nd * nd; e++) {
                    in[ib].A_out[e + nd * nd * w] = A1[e];
                    in[ib].B_out[e + nd * nd * w] = B1[e];
                }
                for (int64_t bb = 0; bb < nb; bb++)
                    for (size_t d = 0; d < nd; d++) {
                        const size_t o = d + nd * (w + (size_t)nw * bb), o1 = d + nd * bb;
                        in[ib].Fd_out[2 * o] = Fd1[2 * o1];
                        in[ib].Fd_out[2 * o + 1] = Fd1[2 * o1 + 1];
                        in[ib].Fk_out[2 * o] = Fk1[2 * o1];
                        in[ib].Fk_out[2 * o + 1] = Fk1[2 * o1 + 1];
                    }
            }
        }
    }
    if (!graph || rc) abandon();
    return rc;
}

int mhx_hydro_coeffs(mhx_ctx* ctx, const mhx_mesh* mesh, int64_t ndof, const double* dofs, int64_t nw,
                     const double* omegas, int64_t nb, const double* betas, int gf, int direct, int mode, double* A_out,
                     double* B_out, double* Fd_out, double* Fk_out) {
    if (!ctx) return MHX_EINVAL;
    mhx_body* b = nullptr;
    int rc = mhx_body_create(ctx, mesh, ndof, dofs, &b);
    if (rc) return rc;
    rc = mhx_body_coeffs(ctx, b, nw, omegas, nb, betas, gf, direct, mode, A_out, B_out, Fd_out, Fk_out);
    mhx_body_destroy(b);
    return rc;
}

int mhx_phase_time(mhx_ctx* ctx, const char* phase, double* total_ms, int64_t* count) {
    if (!ctx || !phase) return MHX_EINVAL;
    flush_timing(ctx);
    auto it = ctx->phases.find(phase);
    if (total_ms) *total_ms = it == ctx->phases.end() ? 0.0 : it->second.ms;
    if (count) *count = it == ctx->phases.end() ? 0 : it->second.count;
    return MHX_OK;
}

int mhx_set_timing(mhx_ctx* ctx, int on) {
    if (!ctx) return MHX_EINVAL;
    ctx->timing = on ? 1 : 0;
    return MHX_OK;
}

int mhx_set_refine_limit(mhx_ctx* ctx, int steps) {
    if (!ctx) return MHX_EINVAL;
    if (steps < -1 || steps > kRefineMaxSplit) return fail(ctx, MHX_EINVAL, "refinement limit outside -1 .. 8");
    ctx->refine_limit = steps;
    return MHX_OK;
}

void mhx_reset_timing(mhx_ctx* ctx) {
    if (!ctx) return;
    flush_timing(ctx);
    ctx->phases.clear();
    ctx->quad_work[0] = ctx->quad_work[1] = ctx->quad_work[2] = 0;
    ctx->refine_steps = ctx->mixed_solves = ctx->fp64_fallbacks = 0;
    ctx->lu_diagonal = ctx->lu_interchanged = 0;
    ctx->max_residual = 0.0;
}

int mhx_set_lu_mode(mhx_ctx* ctx, int mode) {
    if (!ctx) return MHX_EINVAL;
    if (mode != MHX_LU_AUTO && mode != MHX_LU_FP64 && mode != MHX_LU_MIXED && mode != MHX_LU_MIXED_BF16)
        return fail(ctx, MHX_EINVAL, "lu mode must be MHX_LU_AUTO, MHX_LU_FP64, MHX_LU_MIXED or MHX_LU_MIXED_BF16");
    ctx->lu_mode = mode;
    return MHX_OK;
}

int mhx_set_lu_streams(mhx_ctx* ctx, int streams) {
    if (!ctx) return MHX_EINVAL;
    if (streams < 1 || streams > kMaxLuStreams) return fail(ctx, MHX_EINVAL, "lu streams must be 1 .. 4");
    ctx->lu_streams = streams;
    ctx->lu_streams_fixed = true;
    return MHX_OK;
}

int mhx_set_autotune(mhx_ctx* ctx, int on) {
    if (!ctx) return MHX_EINVAL;
    ctx->autotune = on != 0;
    return MHX_OK;
}

int mhx_lu_config(mhx_ctx* ctx, int* streams, int* super_panels) {
    if (!ctx || !streams || !super_panels) return MHX_EINVAL;
    *streams = ctx->last_streams;
    *super_panels = ctx->last_super;
    return MHX_OK;
}

int mhx_set_pivoting(mhx_ctx* ctx, int mode) {
    if (!ctx) return MHX_EINVAL;
    if (mode != MHX_PIVOT_VERIFIED && mode != MHX_PIVOT_ALWAYS)
        return fail(ctx, MHX_EINVAL, "pivoting must be MHX_PIVOT_VERIFIED or MHX_PIVOT_ALWAYS");
    ctx->pivoting = mode;
    return MHX_OK;
}

int mhx_pivot_stats(mhx_ctx* ctx, int64_t counts[2]) {
    if (!ctx || !counts) return MHX_EINVAL;
    counts[0] = ctx->lu_diagonal;
    counts[1] = ctx->lu_interchanged;
    return MHX_OK;
}

int mhx_solver_stats(mhx_ctx* ctx, int64_t counts[3], double* max_residual) {
    if (!ctx) return MHX_EINVAL;
    if (counts) {
        counts[0] = ctx->mixed_solves;
        counts[1] = ctx->refine_steps;
        counts[2] = ctx->fp64_fallbacks;
    }
    if (max_residual) *max_residual = ctx->max_residual;
    return MHX_OK;
}

int mhx_set_batch(mhx_ctx* ctx, int64_t max_items, double max_bytes) {
    if (!ctx) return MHX_EINVAL;
    if (max_items < 1 || !(max_bytes > 0)) return fail(ctx, MHX_EINVAL, "batch limits must be positive");
    ctx->max_batch = max_items;
    ctx->batch_bytes = max_bytes;
    return MHX_OK;
}

int mhx_exact_work(mhx_ctx* ctx, int64_t counts[3], double* flops) {
    if (!ctx) return MHX_EINVAL;
    if (counts)
        for (int q = 0; q < 3; q++) counts[q] = (int64_t)ctx->quad_work[q];
    if (flops)
        *flops = MHX_FLOPS_PER_EVAL * ctx->quad_work[0] + MHX_FLOPS_PER_SERIES * ctx->quad_work[1] +
                 MHX_FLOPS_PER_CF * ctx->quad_work[2];
    return MHX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// learned surrogate of the wave term
// ---------------------------------------------------------------------------
namespace {

struct HostRng {
    uint64_t s;
    explicit HostRng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

int surr_alloc(mhx_ctx* ctx) {
    int rc;
    const int64_t P = mhx::surr_param_count();
    if ((rc = dalloc(ctx, &ctx->surr_params, (size_t)P)) || (rc = dalloc(ctx, &ctx->surr_B, 2 * SURR_NF)) ||
        (rc = dalloc(ctx, &ctx->surr_W, (size_t)SURR_PACKED_W)) || (rc = dalloc(ctx, &ctx->surr_bias, SURR_BIAS)))
        return rc;
    return MHX_OK;
}

// deterministic initial parameters (uniform +-1/sqrt(fan_in)) and Fourier frequencies (N(0, 4^2))
void surr_init_host(uint64_t seed, std::vector<float>& params, std::vector<float>& B) {
    HostRng rng(seed);
    const int64_t P = mhx::surr_param_count();
    params.assign(P, 0.f);
    const int in_dim[5] = {64, 128, 128, 128, 128}, out_dim[5] = {128, 128, 128, 128, SURR_NOUT};
    int64_t o = 0;
    for (int l = 0; l < 5; l++) {
        const double bound = 1.0 / std::sqrt((double)in_dim[l]);
        for (int64_t q = 0; q < (int64_t)out_dim[l] * in_dim[l] + out_dim[l]; q++)
            params[o++] = (float)((2.0 * rng.uniform() - 1.0) * bound);
    }
    B.assign(2 * SURR_NF, 0.f);
    for (auto& b : B) b = (float)(4.0 * rng.normal());
}

struct TrainBufs {
    mhx::SurrTrainWork w;
    float *grad = nullptr, *m1 = nullptr, *m2 = nullptr, *loss = nullptr;
    int* idx = nullptr;
    void release() {
        for (int q = 0; q < 6; q++) dfree(w.act[q]);
        for (int q = 0; q < 5; q++) { dfree(w.z[q]); dfree(w.dz[q]); }
        dfree(w.part); dfree(grad); dfree(m1); dfree(m2); dfree(loss); dfree(idx);
    }
};

int train_bufs(mhx_ctx* ctx, TrainBufs& t, int M) {
    const int64_t P = mhx::surr_param_count();
    int rc;
    t.w.M = M;
    const int widths[6] = {64, 128, 128, 128, 128, 8};
    for (int q = 0; q < 6; q++)
        if ((rc = dalloc(ctx, &t.w.act[q], (size_t)M * widths[q]))) return rc;
    for (int q = 0; q < 5; q++)
        if ((rc = dalloc(ctx, &t.w.z[q], (size_t)M * 128)) || (rc = dalloc(ctx, &t.w.dz[q], (size_t)M * 128))) return rc;
    if ((rc = dalloc(ctx, &t.w.part, (size_t)32 * 128 * 128)) || (rc = dalloc(ctx, &t.grad, (size_t)P)) ||
        (rc = dalloc(ctx, &t.m1, (size_t)P)) || (rc = dalloc(ctx, &t.m2, (size_t)P)) || (rc = dalloc(ctx, &t.loss, 1)) ||
        (rc = dalloc(ctx, &t.idx, (size_t)M)))
        return rc;
    t.w.Bf = ctx->surr_B;
    return MHX_OK;
}

__global__ void strided_index_kernel(int M, int64_t start, int64_t stride, int* idx) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < M) idx[t] = (int)(start + stride * t);
}

}  // namespace

extern "C" {

int mhx_surrogate_train(mhx_ctx* ctx, int64_t ngrid, int64_t steps, int64_t batch, double lr, uint64_t seed,
                        double* val_loss) {
    if (!ctx) return MHX_EINVAL;
    if (ngrid < 16 || ngrid > 4096 || steps < 1 || batch < 64 || batch > 65536 || batch % 64 || !(lr > 0))
        return fail(ctx, MHX_EINVAL, "surrogate_train: bad grid, steps, batch (multiple of 64) or learning rate");
    hipSetDevice(ctx->device);
    int rc;
    if ((rc = surr_alloc(ctx))) return rc;
    hipStream_t st = ctx->stream;
    std::vector<float> hp, hB;
    surr_init_host(seed, hp, hB);
    const int64_t P = mhx::surr_param_count();
    HCHK(hipMemcpyAsync(ctx->surr_params, hp.data(), sizeof(float) * P, hipMemcpyHostToDevice, st));
    HCHK(hipMemcpyAsync(ctx->surr_B, hB.data(), sizeof(float) * hB.size(), hipMemcpyHostToDevice, st));
    const int64_t ns = ngrid * ngrid;
    float *U = nullptr, *Y = nullptr;
    unsigned long long* status = nullptr;
    TrainBufs t;
    auto cleanup = [&]() { dfree(U); dfree(Y); dfree(status); t.release(); };
    if ((rc = dalloc(ctx, &U, 2 * (size_t)ns)) || (rc = dalloc(ctx, &Y, 6 * (size_t)ns)) || (rc = dalloc(ctx, &status, 4)) ||
        (rc = train_bufs(ctx, t, (int)batch))) {
        cleanup();
        return rc;
    }
    HCHK(hipMemsetAsync(status, 0, 4 * sizeof(unsigned long long), st));
    {
        Phase ph(ctx, "surrogate_samples");
        mhx::launch_surr_samples(st, (int)ngrid, seed ^ 0x5DEECE66Dull, U, Y, status);
    }
    HCHK(hipMemsetAsync(t.m1, 0, sizeof(float) * P, st));
    HCHK(hipMemsetAsync(t.m2, 0, sizeof(float) * P, st));
    {
        Phase ph(ctx, "surrogate_train");
        for (int64_t step = 1; step <= steps; step++) {
            mhx::launch_batch_index(st, (int)batch, ns, seed * 1315423911ull + (uint64_t)step, t.idx);
            mhx::surr_forward_backward(st, t.w, (int)batch, t.idx, U, Y, ctx->surr_params, t.grad, nullptr, ctx->surr_act);
            const double frac = (double)(step - 1) / (double)steps;
            const float lrs = (float)(lr * 0.5 * (1.0 + std::cos(3.141592653589793 * frac)));
            mhx::surr_adam(st, P, ctx->surr_params, t.grad, t.m1, t.m2, lrs, (int)step);
        }
    }
    HCHK(hipGetLastError());
    // validation loss over the held-out stride
    const int64_t nval = ns / SURR_VAL_STRIDE;
    double vsum = 0.0;
    int64_t vcnt = 0;
    for (int64_t v0 = 0; v0 < nval; v0 += batch) {
        const int m = (int)std::min<int64_t>(batch, nval - v0);
        if (m < 64) break;
        strided_index_kernel<<<(m + 255) / 256, 256, 0, st>>>(m, v0 * SURR_VAL_STRIDE, SURR_VAL_STRIDE, t.idx);
        mhx::surr_forward_backward(st, t.w, m, t.idx, U, Y, ctx->surr_params, t.grad, t.loss, ctx->surr_act);
        float l = 0.f;
        HCHK(hipMemcpyAsync(&l, t.loss, sizeof(float), hipMemcpyDeviceToHost, st));
        HCHK(hipStreamSynchronize(st));
        vsum += (double)l * m;
        vcnt += m;
    }
    unsigned long long hst[4] = {0, 0, 0, 0};
    HCHK(hipMemcpyAsync(hst, status, sizeof hst, hipMemcpyDeviceToHost, st));
    mhx::launch_surr_pack(st, ctx->surr_params, ctx->surr_W, ctx->surr_bias);
    HCHK(hipStreamSynchronize(st));
    cleanup();
    flush_timing(ctx);
    if (hst[0]) return fail(ctx, MHX_EQUAD, "surrogate samples: exact quadrature exceeded its interval budget");
    ctx->surr_val_loss = vcnt ? vsum / vcnt : -1.0;
    ctx->surr_steps = steps;
    ctx->surr_ready = true;
    if (val_loss) *val_loss = ctx->surr_val_loss;
    return MHX_OK;
}

int mhx_surrogate_save(mhx_ctx* ctx, const char* path) {
    if (!ctx || !path) return MHX_EINVAL;
    if (!ctx->surr_ready) return fail(ctx, MHX_EINVAL, "surrogate_save: no trained network");
    const int64_t P = mhx::surr_param_count();
    std::vector<float> hp(P), hB(2 * SURR_NF);
    HCHK(hipMemcpy(hp.data(), ctx->surr_params, sizeof(float) * P, hipMemcpyDeviceToHost));
    HCHK(hipMemcpy(hB.data(), ctx->surr_B, sizeof(float) * hB.size(), hipMemcpyDeviceToHost));
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(ctx, MHX_EINVAL, std::string("surrogate_save: cannot open ") + path);
    char hdr[512];
    std::snprintf(hdr, sizeof hdr,
                  "{\"format\": \"mhx-surrogate\", \"version\": 2, \"activation\": \"%s\", \"params\": %lld, "
                  "\"fourier\": %d, \"outputs\": %d, "
                  "\"scale\": %.1f, \"r_max\": %.1f, \"v_min\": %g, \"v_max\": %.1f, \"steps\": %lld, "
                  "\"val_mse\": %.6e}\n",
                  ctx->surr_act == MHX_SURR_RELU ? "relu" : "silu", (long long)P, SURR_NF, SURR_NOUT, SURR_SCALE,
                  SURR_RMAX, SURR_VMIN, SURR_VMAX,
                  (long long)ctx->surr_steps, ctx->surr_val_loss);
    std::fputs("MHXSURR1\n", f);
    std::fputs(hdr, f);
    std::fwrite(hB.data(), sizeof(float), hB.size(), f);
    std::fwrite(hp.data(), sizeof(float), hp.size(), f);
    std::fclose(f);
    return MHX_OK;
}

int mhx_surrogate_load(mhx_ctx* ctx, const char* path) {
    if (!ctx || !path) return MHX_EINVAL;
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(ctx, MHX_EINVAL, std::string("surrogate_load: cannot open ") + path);
    char line[1024];
    const int64_t P = mhx::surr_param_count();
    std::vector<float> hp(P), hB(2 * SURR_NF);
    bool ok = std::fgets(line, sizeof line, f) && std::strcmp(line, "MHXSURR1\n") == 0 && std::fgets(line, sizeof line, f);
    long long pc = -1, steps = 0;
    double vl = -1.0;
    int act = MHX_SURR_SILU;
    bool act_ok = true;
    if (ok) {
        const char* q = std::strstr(line, "\"params\": ");
        ok = q && std::sscanf(q + 10, "%lld", &pc) == 1 && pc == P;
        // the hidden layers' activation the weights were trained with (version-1 files, without
        // the field, hold SiLU networks)
        const char* qa = std::strstr(line, "\"activation\": \"");
        if (qa) {
            qa += 15;
            if (std::strncmp(qa, "silu\"", 5) == 0) act = MHX_SURR_SILU;
            else if (std::strncmp(qa, "relu\"", 5) == 0) act = MHX_SURR_RELU;
            else act_ok = false;
        }
        const char* q2 = std::strstr(line, "\"steps\": ");
        if (q2) std::sscanf(q2 + 9, "%lld", &steps);
        const char* q3 = std::strstr(line, "\"val_mse\": ");
        if (q3) std::sscanf(q3 + 11, "%lf", &vl);
    }
    ok = ok && std::fread(hB.data(), sizeof(float), hB.size(), f) == hB.size() &&
         std::fread(hp.data(), sizeof(float), hp.size(), f) == hp.size();
    std::fclose(f);
    if (!ok) return fail(ctx, MHX_EINVAL, std::string("surrogate_load: not an mhx surrogate file: ") + path);
    if (!act_ok)
        return fail(ctx, MHX_EINVAL, std::string("surrogate_load: unknown activation in ") + path);
    hipSetDevice(ctx->device);
    int rc;
    if ((rc = surr_alloc(ctx))) return rc;
    HCHK(hipMemcpyAsync(ctx->surr_params, hp.data(), sizeof(float) * P, hipMemcpyHostToDevice, ctx->stream));
    HCHK(hipMemcpyAsync(ctx->surr_B, hB.data(), sizeof(float) * hB.size(), hipMemcpyHostToDevice, ctx->stream));
    mhx::launch_surr_pack(ctx->stream, ctx->surr_params, ctx->surr_W, ctx->surr_bias);
    HCHK(hipStreamSynchronize(ctx->stream));
    ctx->surr_ready = true;
    ctx->surr_act = act;
    ctx->surr_steps = steps;
    ctx->surr_val_loss = vl;
    return MHX_OK;
}

int mhx_surrogate_set_activation(mhx_ctx* ctx, int activation) {
    if (!ctx) return MHX_EINVAL;
    if (activation != MHX_SURR_SILU && activation != MHX_SURR_RELU)
        return fail(ctx, MHX_EINVAL, "surrogate activation must be MHX_SURR_SILU or MHX_SURR_RELU");
    if (activation != ctx->surr_act) ctx->surr_ready = false;  // a network of the other activation is dropped
    ctx->surr_act = activation;
    return MHX_OK;
}

int mhx_surrogate_get_activation(mhx_ctx* ctx, int* activation) {
    if (!ctx || !activation) return MHX_EINVAL;
    *activation = ctx->surr_act;
    return MHX_OK;
}

int mhx_surrogate_params(mhx_ctx* ctx, float* params_out, float* fourier_out, int64_t* nparams) {
    if (!ctx) return MHX_EINVAL;
    const int64_t P = mhx::surr_param_count();
    if (nparams) *nparams = P;
    if (!params_out && !fourier_out) return MHX_OK;
    if (!ctx->surr_ready) return fail(ctx, MHX_EINVAL, "surrogate_params: no trained network");
    if (params_out) HCHK(hipMemcpy(params_out, ctx->surr_params, sizeof(float) * P, hipMemcpyDeviceToHost));
    if (fourier_out) HCHK(hipMemcpy(fourier_out, ctx->surr_B, sizeof(float) * 2 * SURR_NF, hipMemcpyDeviceToHost));
    return MHX_OK;
}

int mhx_surrogate_eval(mhx_ctx* ctx, int64_t m, const double* rv, float* out) {
    if (!ctx) return MHX_EINVAL;
    if (!ctx->surr_ready) return fail(ctx, MHX_EINVAL, "surrogate_eval: no trained network");
    if (m <= 0 || !rv || !out) return fail(ctx, MHX_EINVAL, "surrogate_eval: bad arguments");
    hipSetDevice(ctx->device);
    double* drv = nullptr;
    float* dout = nullptr;
    int rc;
    if ((rc = dalloc(ctx, &drv, 2 * (size_t)m)) || (rc = dalloc(ctx, &dout, 6 * (size_t)m))) {
        dfree(drv); dfree(dout);
        return rc;
    }
    hipMemcpyAsync(drv, rv, sizeof(double) * 2 * m, hipMemcpyHostToDevice, ctx->stream);
    mhx::launch_surr_eval(ctx->stream, (int)m, drv, ctx->surr_W, ctx->surr_bias, ctx->surr_B, dout, ctx->surr_act);
    hipMemcpyAsync(out, dout, sizeof(float) * 6 * m, hipMemcpyDeviceToHost, ctx->stream);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    dfree(drv); dfree(dout);
    if (e != hipSuccess) return fail(ctx, MHX_EDEVICE, "surrogate_eval: device error");
    return MHX_OK;
}

int mhx_surrogate_loss_grad(mhx_ctx* ctx, int64_t m, const float* U, const float* Yt, const float* params,
                            const float* fourier, float* grad_out, float* loss_out) {
    if (!ctx) return MHX_EINVAL;
    if (m < 1 || m > 65536 || !U || !Yt || !params || !fourier || !grad_out)
        return fail(ctx, MHX_EINVAL, "surrogate_loss_grad: bad arguments");
    hipSetDevice(ctx->device);
    const int64_t P = mhx::surr_param_count();
    float *dU = nullptr, *dY = nullptr, *dp = nullptr, *dB = nullptr;
    TrainBufs t;
    int rc;
    auto cleanup = [&]() { dfree(dU); dfree(dY); dfree(dp); dfree(dB); t.release(); };
    if ((rc = dalloc(ctx, &dU, 2 * (size_t)m)) || (rc = dalloc(ctx, &dY, 6 * (size_t)m)) || (rc = dalloc(ctx, &dp, (size_t)P)) ||
        (rc = dalloc(ctx, &dB, 2 * SURR_NF)) || (rc = train_bufs(ctx, t, (int)m))) {
        cleanup();
        return rc;
    }
    t.w.Bf = dB;
    hipStream_t st = ctx->stream;
    hipMemcpyAsync(dU, U, sizeof(float) * 2 * m, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dY, Yt, sizeof(float) * 6 * m, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dp, params, sizeof(float) * P, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dB, fourier, sizeof(float) * 2 * SURR_NF, hipMemcpyHostToDevice, st);
    mhx::surr_forward_backward(st, t.w, (int)m, nullptr, dU, dY, dp, t.grad, t.loss, ctx->surr_act);
    float l = 0.f;
    hipMemcpyAsync(&l, t.loss, sizeof(float), hipMemcpyDeviceToHost, st);
    hipMemcpyAsync(grad_out, t.grad, sizeof(float) * P, hipMemcpyDeviceToHost, st);
    hipError_t e = hipStreamSynchronize(st);
    cleanup();
    if (e != hipSuccess) return fail(ctx, MHX_EDEVICE, "surrogate_loss_grad: device error");
    if (loss_out) *loss_out = l;
    return MHX_OK;
}

}  // extern "C"
