// Dense complex LU with partial pivoting and triangular solves for gfx950.
//
// Replaces LinearSolve's LUFactorization -> LAPACK zgetrf/zgetrs
// (reference: src/matrix_assembly.jl:396-416, 436-449).  Pivot choice follows
// LAPACK izamax (largest |Re|+|Im|, first index on ties), so the factors agree
// with the CPU LU to rounding.
//
// Blocking (right-looking):
//   outer panels of NB=64 columns; each outer panel is factored in leaves of
//   W columns (4 for FP64, 8 for FP32) held in registers of one 1024-thread
//   workgroup (the pivot search is a block reduction), each leaf followed by a
//   multi-workgroup rank-W update of the rest of the outer panel;
//   then the 64 row interchanges are applied to the other columns, the
//   64-row U block is solved (unit-lower TRSM), and the trailing matrix is
//   updated by a complex GEMM on FP64 MFMA (v_mfma_f64_16x16x4f64).
// Every kernel takes a batch index (blockIdx.z) so many matrices of the same
// order are factored by one launch sequence (frequency sweeps).
//
// The factorization is generic in the element type: complex FP64 (cplx) or
// complex FP32 (cplxf, trailing updates on v_mfma_f32_16x16x4f32).  The FP32
// factors serve the mixed-precision solver (FP32 LU, FP64 residuals and
// iterative refinement); the triangular solves always run in FP64 arithmetic.
#include "mhx_internal.h"

#include <type_traits>

namespace mhx {

template <class C> struct Cx;
template <> struct Cx<cplx> { typedef double R; };
template <> struct Cx<cplxf> { typedef float R; };

template <class C>
__device__ __forceinline__ C tmk(typename Cx<C>::R r, typename Cx<C>::R i) {
    C c;
    c.x = r;
    c.y = i;
    return c;
}
template <class C> __device__ __forceinline__ C tsub(C a, C b) { return tmk<C>(a.x - b.x, a.y - b.y); }
template <class C> __device__ __forceinline__ C tadd(C a, C b) { return tmk<C>(a.x + b.x, a.y + b.y); }
template <class C> __device__ __forceinline__ C tmul(C a, C b) {
    return tmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// 1 / u (Smith's algorithm in the element's precision)
template <class C> __device__ __forceinline__ C trecip(C u) {
    typedef typename Cx<C>::R R;
    if (fabs(u.x) >= fabs(u.y)) {
        const R r = u.y / u.x, den = u.x + u.y * r;
        return tmk<C>(R(1) / den, -r / den);
    }
    const R r = u.x / u.y, den = u.x * r + u.y;
    return tmk<C>(r / den, R(-1) / den);
}
template <class C> __device__ __forceinline__ cplx to_d(C v) { return cmk((double)v.x, (double)v.y); }

#define NB 64
#define LEAF_THREADS 1024

template <class C>
struct MatBatch {
    C* A;           // base pointer of matrix 0
    int64_t stride; // elements between consecutive matrices
    int lda;
};

// ---------------------------------------------------------------------------
// leaf: factor columns [c0, c0+W) over rows [c0, n); swap the pivot rows in
// the other columns of the outer panel [K, KE); solve the W-row U block of the
// outer panel columns right of the leaf.
// ---------------------------------------------------------------------------
template <int W, int ROWS, class C>
__global__ void __launch_bounds__(LEAF_THREADS) getf2_leaf_kernel(MatBatch<C> mb, int n, int c0, int K, int KE,
                                                                  int* __restrict__ ipiv, int64_t ipiv_stride,
                                                                  int* __restrict__ info) {
    typedef typename Cx<C>::R R;
    const int b = blockIdx.z;
    C* A = mb.A + mb.stride * b;
    int* piv = ipiv + ipiv_stride * b;
    const int lda = mb.lda;
    const int tid = threadIdx.x;
    // per column parity (the LDS of column c is reused at c + 2, after the two barriers of c + 1)
    __shared__ R red_v[2][LEAF_THREADS / 64];
    __shared__ int red_r[2][LEAF_THREADS / 64];
    __shared__ C xrow[2][W];  // row col before the interchange
    __shared__ C prow[2][W];  // the pivot row
    __shared__ int s_pivs[W];
    __shared__ C L11[W][W];

    const int wcols = min(W, n - c0);
    C x[ROWS][W];
#pragma unroll
    for (int q = 0; q < ROWS; q++) {
        int r = c0 + tid + LEAF_THREADS * q;
#pragma unroll
        for (int c = 0; c < W; c++) x[q][c] = (r < n && c < wcols) ? A[r + (int64_t)lda * (c0 + c)] : tmk<C>(0, 0);
    }
#pragma unroll
    for (int c = 0; c < W; c++) {
        if (c >= wcols) break;
        const int col = c0 + c, par = c & 1;
        // --- pivot search: max |Re|+|Im|, smallest row on ties
        R bv = -1;
        int br = 0x7fffffff;
#pragma unroll
        for (int q = 0; q < ROWS; q++) {
            int r = c0 + tid + LEAF_THREADS * q;
            R v = fabs(x[q][c].x) + fabs(x[q][c].y);
            if (r >= col && r < n && (v > bv || (v == bv && r < br))) { bv = v; br = r; }
        }
        for (int off = 32; off > 0; off >>= 1) {
            R ov = __shfl_xor(bv, off);
            int orow = __shfl_xor(br, off);
            if (ov > bv || (ov == bv && orow < br)) { bv = ov; br = orow; }
        }
        if ((tid & 63) == 0) { red_v[par][tid >> 6] = bv; red_r[par][tid >> 6] = br; }
        __syncthreads();
        // every thread combines the per-wave candidates in wave order (the same rule as above)
        R v = red_v[par][0];
        int p = red_r[par][0];
#pragma unroll
        for (int w = 1; w < LEAF_THREADS / 64; w++)
            if (red_v[par][w] > v || (red_v[par][w] == v && red_r[par][w] < p)) { v = red_v[par][w]; p = red_r[par][w]; }
        if (tid == 0) {
            piv[col] = p;
            s_pivs[c] = p;
        }
        // --- interchange rows col <-> p inside the leaf, and the pivot row for the update
        const int tc = (col - c0) % LEAF_THREADS, qc = (col - c0) / LEAF_THREADS;
        const int tp = (p - c0) % LEAF_THREADS, qp = (p - c0) / LEAF_THREADS;
#pragma unroll
        for (int q = 0; q < ROWS; q++) {
            if (tid == tc && q == qc)
                for (int cc = 0; cc < W; cc++) xrow[par][cc] = x[q][cc];
            if (tid == tp && q == qp)
                for (int cc = 0; cc < W; cc++) prow[par][cc] = x[q][cc];
        }
        __syncthreads();
        if (p != col) {
#pragma unroll
            for (int q = 0; q < ROWS; q++) {
                if (tid == tc && q == qc)
                    for (int cc = 0; cc < W; cc++) x[q][cc] = prow[par][cc];
                if (tid == tp && q == qp)
                    for (int cc = 0; cc < W; cc++) x[q][cc] = xrow[par][cc];
            }
        }
        const C u = prow[par][c];
        if (u.x != 0 || u.y != 0) {
            const C inv = trecip(u);
#pragma unroll
            for (int q = 0; q < ROWS; q++) {
                int r = c0 + tid + LEAF_THREADS * q;
                if (r > col && r < n) {
                    x[q][c] = tmul(x[q][c], inv);
#pragma unroll
                    for (int cc = 0; cc < W; cc++)
                        if (cc > c) x[q][cc] = tsub(x[q][cc], tmul(x[q][c], prow[par][cc]));
                }
            }
        } else if (tid == 0) {
            atomicCAS(info + b, 0, col + 1);
        }
    }
    // store back
#pragma unroll
    for (int q = 0; q < ROWS; q++) {
        int r = c0 + tid + LEAF_THREADS * q;
        if (r < n)
#pragma unroll
            for (int c = 0; c < W; c++)
                if (c < wcols) A[r + (int64_t)lda * (c0 + c)] = x[q][c];
    }
    // L11 (unit lower 4x4 of this leaf) to LDS
    if (tid < wcols)
        for (int c = 0; c < W; c++) L11[tid][c] = x[0][c];
    __syncthreads();
    // other outer-panel columns: apply the leaf's interchanges, then solve U rows
    const int nother = (KE - K) - wcols;
    // The at most 2 W rows a column touches (c0 + c and the pivot rows) are loaded together, the
    // interchanges and the substitution done on the register copies (an entry stands for its row;
    // a row listed twice keeps both copies equal), and all of them stored back.
    // (the rows are the same for every thread: scalar registers)
    int idx[2 * W];
#pragma unroll
    for (int c = 0; c < W; c++) {
        idx[c] = c0 + min(c, wcols - 1);
        idx[W + c] = __builtin_amdgcn_readfirstlane(c < wcols ? s_pivs[c] : c0);
    }
    for (int t = tid; t < nother; t += LEAF_THREADS) {
        // L11 is read from LDS inside the loop: hoisted out of it, the W x W block would occupy
        // 2 W^2 registers for the whole loop (it runs once per thread: nother < LEAF_THREADS)
        asm volatile("" ::: "memory");
        int j = K + t;
        if (j >= c0) j += wcols;
        C* colj = A + (int64_t)lda * j;
        C v[2 * W];
#pragma unroll
        for (int k = 0; k < 2 * W; k++) v[k] = colj[idx[k]];
        auto get = [&](int r) {
            C w = v[0];
#pragma unroll
            for (int k = 2 * W - 1; k >= 0; k--)
                if (idx[k] == r) w = v[k];
            return w;
        };
        auto set = [&](int r, C w) {
#pragma unroll
            for (int k = 0; k < 2 * W; k++)
                if (idx[k] == r) v[k] = w;
        };
#pragma unroll
        for (int c = 0; c < W; c++) {
            if (c >= wcols) break;
            const int r0 = c0 + c, r1 = s_pivs[c];
            if (r1 != r0) {
                const C t0 = get(r0), t1 = get(r1);
                set(r0, t1);
                set(r1, t0);
            }
        }
        if (j > c0) {  // U12 rows c0..c0+wcols-1: unit-lower forward substitution
            C u[W];
#pragma unroll
            for (int c = 0; c < W; c++) {
                if (c >= wcols) break;
                C s = get(c0 + c);
#pragma unroll
                for (int s2 = 0; s2 < c; s2++) s = tsub(s, tmul(L11[c][s2], u[s2]));
                u[c] = s;
                set(c0 + c, s);
            }
        }
#pragma unroll
        for (int k = 0; k < 2 * W; k++) colj[idx[k]] = v[k];
    }
}

// rank-W update of the outer panel: A[r, j] -= sum_s L[r, c0+s] U[c0+s, j],
// rows r in [c0+W, n), columns j in [c0+W, KE)
template <int W, class C>
__global__ void __launch_bounds__(256) panel_update_kernel(MatBatch<C> mb, int n, int c0, int KE) {
    const int b = blockIdx.z;
    C* A = mb.A + mb.stride * b;
    const int lda = mb.lda;
    __shared__ C U[W][NB];
    const int j0 = c0 + W, nj = KE - j0;
    for (int t = threadIdx.x; t < W * nj; t += blockDim.x) {
        int s = t / nj, jj = t % nj;
        U[s][jj] = A[(c0 + s) + (int64_t)lda * (j0 + jj)];
    }
    __syncthreads();
    const int r = c0 + W + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    C l[W];
#pragma unroll
    for (int s = 0; s < W; s++) l[s] = A[r + (int64_t)lda * (c0 + s)];
    // 16 columns per round trip: their loads are issued before any of their stores (element by
    // element, each load would wait for the previous store through the same pointer)
    constexpr int JC = 16;
    for (int jb = 0; jb < nj; jb += JC) {
        C a[JC];
#pragma unroll
        for (int q = 0; q < JC; q++)
            if (jb + q < nj) a[q] = A[r + (int64_t)lda * (j0 + jb + q)];
#pragma unroll
        for (int q = 0; q < JC; q++)
            if (jb + q < nj) {
#pragma unroll
                for (int s = 0; s < W; s++) a[q] = tsub(a[q], tmul(l[s], U[s][jb + q]));
                A[r + (int64_t)lda * (j0 + jb + q)] = a[q];
            }
    }
}

// permutation of the <=2*NB rows touched by an outer panel's interchanges:
// perm_dst[e] receives the old contents of row perm_src[e]
__global__ void build_perm_kernel(const int* __restrict__ ipiv, int64_t ipiv_stride, int K, int KE,
                                  int* __restrict__ perm, int64_t perm_stride) {
    const int b = blockIdx.z;
    const int* piv = ipiv + ipiv_stride * b;
    int* out = perm + perm_stride * b;
    if (threadIdx.x != 0) return;
    int rows[2 * NB], src[2 * NB];
    int m = 0;
    for (int k = K; k < KE; k++) {
        int a = k, c = piv[k];
        if (a == c) continue;
        int ia = -1, ic = -1;
        for (int e = 0; e < m; e++) {
            if (rows[e] == a) ia = e;
            if (rows[e] == c) ic = e;
        }
        if (ia < 0) { rows[m] = a; src[m] = a; ia = m++; }
        if (ic < 0) { rows[m] = c; src[m] = c; ic = m++; }
        int t = src[ia]; src[ia] = src[ic]; src[ic] = t;
    }
    out[0] = m;
    for (int e = 0; e < m; e++) { out[1 + e] = rows[e]; out[1 + 2 * NB + e] = src[e]; }
}

// apply the permutation to columns outside [K, KE); one column per block-row of threads
template <class C>
__global__ void __launch_bounds__(256) apply_perm_kernel(MatBatch<C> mb, int n, int K, int KE,
                                                         const int* __restrict__ perm, int64_t perm_stride) {
    const int b = blockIdx.z;
    C* A = mb.A + mb.stride * b;
    const int* pm = perm + perm_stride * b;
    const int m = pm[0];
    const int e = threadIdx.x % (2 * NB);
    const int cl = threadIdx.x / (2 * NB);
    int j = (blockIdx.x * (256 / (2 * NB))) + cl;
    const int nout = n - (KE - K);
    C v = tmk<C>(0, 0);
    bool act = e < m && j < nout;
    if (j >= K) j += KE - K;
    if (act) v = A[pm[1 + 2 * NB + e] + (int64_t)mb.lda * j];
    __syncthreads();
    if (act) A[pm[1 + e] + (int64_t)mb.lda * j] = v;
}

// U12 = L11^{-1} A12 for the outer panel rows [K, KE), columns [KE, n): forward substitution with
// the unit-lower block L11 staged once in LDS; TRSM_COLS columns per workgroup, thread (row i,
// column quarter q) owns TRSM_COLS / 4 columns of row i.
#define TRSM_COLS 16
template <class C>
__global__ void __launch_bounds__(256) trsm_panel_kernel(MatBatch<C> mb, int n, int K, int KE) {
    constexpr int CPQ = TRSM_COLS / 4;
    const int b = blockIdx.z;
    C* A = mb.A + mb.stride * b;
    const int lda = mb.lda;
    const int nb = KE - K;
    __shared__ C L[NB][NB + 1];         // L[s][i] = L11[i][s]
    __shared__ C X[NB][TRSM_COLS + 1];
    const int tid = threadIdx.x, i = tid & (NB - 1), q = tid / NB;
    const int j0 = KE + blockIdx.x * TRSM_COLS;
    for (int e = tid; e < NB * NB; e += 256) {
        const int r = e & (NB - 1), s = e / NB;
        L[s][r] = (r < nb && s < nb) ? A[(K + r) + (int64_t)lda * (K + s)] : tmk<C>(0, 0);
    }
    for (int e = tid; e < NB * TRSM_COLS; e += 256) {
        const int r = e & (NB - 1), c = e / NB;
        X[r][c] = (r < nb && j0 + c < n) ? A[(K + r) + (int64_t)lda * (j0 + c)] : tmk<C>(0, 0);
    }
    __syncthreads();
    C x[CPQ];
#pragma unroll
    for (int cc = 0; cc < CPQ; cc++) x[cc] = X[i][q * CPQ + cc];
    for (int r = 0; r < nb - 1; r++) {
        // row r is final; rows below subtract L[i][r] x_r
        if (i > r) {
            const C l = L[r][i];
#pragma unroll
            for (int cc = 0; cc < CPQ; cc++) x[cc] = tsub(x[cc], tmul(l, X[r][q * CPQ + cc]));
        }
        if (i == r + 1) {
#pragma unroll
            for (int cc = 0; cc < CPQ; cc++) X[i][q * CPQ + cc] = x[cc];
        }
        __syncthreads();
    }
#pragma unroll
    for (int cc = 0; cc < CPQ; cc++) {
        const int j = j0 + q * CPQ + cc;
        if (i < nb && j < n) A[(K + i) + (int64_t)lda * j] = x[cc];
    }
}

// step r of the forward substitution with a compile-time r (the column stays in registers)
template <int r>
__device__ __forceinline__ void trsm_cols_step(cplxf (&x)[NB], const cplxf (*Lt)[NB]) {
    if constexpr (r < NB - 1) {
        // the step's L11 reads in groups of 8, each behind a compiler memory barrier, so that at
        // most 8 of them are live next to the 64 elements of the column
#pragma unroll
        for (int i0 = r + 1; i0 < NB; i0 += 8) {
            __asm__ volatile("" ::: "memory");
#pragma unroll
            for (int i = i0; i < i0 + 8 && i < NB; i++) x[i] = tsub(x[i], tmul(Lt[r][i], x[r]));
        }
        trsm_cols_step<r + 1>(x, Lt);
    }
}

// FP32 factor: one thread per column of U12, the 64-row forward substitution held in registers
// (x_i -= L11[i][r] x_r for r = 0.. in order, as above), L11 read from LDS with uniform addresses
// (broadcast), so no barrier inside the solve.  Launched only for full panels (KE = K + NB < n).
__global__ void __launch_bounds__(256) trsm_cols_kernel(MatBatch<cplxf> mb, int n, int K) {
    const int b = blockIdx.z;
    cplxf* A = mb.A + mb.stride * b;
    const int lda = mb.lda, KE = K + NB;
    __shared__ cplxf Lt[NB][NB];  // Lt[r][i] = L11[i][r]
    for (int e = threadIdx.x; e < NB * NB; e += 256) {
        const int i = e & (NB - 1), r = e / NB;
        Lt[r][i] = A[(K + i) + (int64_t)lda * (K + r)];
    }
    __syncthreads();
    const int j = KE + blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    cplxf* col = A + K + (int64_t)lda * j;
    cplxf x[NB];
#pragma unroll
    for (int i = 0; i < NB; i++) x[i] = col[i];
    trsm_cols_step<0>(x, Lt);
#pragma unroll
    for (int i = 0; i < NB; i++) col[i] = x[i];
}

// ---------------------------------------------------------------------------
// Verified diagonal pivoting.  For the influence matrices of the BEM (1/2 I plus small
// off-diagonal couplings) LAPACK's partial pivoting keeps the diagonal pivot in every column:
// after the earlier eliminations the diagonal is a largest |Re| + |Im| of the column's candidates
// (and the first on ties).  These kernels factor with the diagonal pivots and check exactly that
// rule on every candidate of every column; a matrix that breaks it anywhere, or meets an exactly
// zero pivot, is flagged and factored again by the caller with row interchanges (the leaf kernels
// above).  The pivot sequence is therefore LAPACK's in both cases, and a panel of 64 columns
// costs one workgroup per matrix for its diagonal block plus one multi-workgroup launch for its
// L21 rows and U12 columns.
// Flag of matrix b: flags[flag_stride * b] (0: the rule held so far; set to 1 on a failure;
// cleared by the diagonal kernel of the first panel).
// ---------------------------------------------------------------------------
template <class C> __device__ __forceinline__ typename Cx<C>::R tabs1(C a) { return fabs(a.x) + fabs(a.y); }

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

template <class R> struct Mfma16;
template <> struct Mfma16<double> {
    typedef double4_t V;
    static __device__ __forceinline__ V run(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int v, int lk) { return 4 * v + lk; }
};
template <> struct Mfma16<float> {
    typedef float4_t V;
    static __device__ __forceinline__ V run(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int v, int lk) { return 4 * lk + v; }
};

// value of lane `src` (uniform) of a wave
__device__ __forceinline__ double rlane(double v, int src) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), src),
                            __builtin_amdgcn_readlane(__double2loint(v), src));
}
__device__ __forceinline__ float rlane(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
template <class C> __device__ __forceinline__ C rlane(C v, int src) { return tmk<C>(rlane(v.x, src), rlane(v.y, src)); }

// 1 / u as conj(u) / |u|^2 (one real division) where |u|^2 is far from overflow and underflow,
// else Smith's algorithm
template <class C> __device__ __forceinline__ C crecip(C u) {
    typedef typename Cx<C>::R R;
    const R d = u.x * u.x + u.y * u.y;
    const R lo = sizeof(R) == 8 ? R(1e-280) : R(1e-35), hi = sizeof(R) == 8 ? R(1e280) : R(1e35);
    if (d > lo && d < hi) {
        const R sc = R(1) / d;
        return tmk<C>(u.x * sc, -u.y * sc);
    }
    return trecip(u);
}

// Diagonal block A[K:K+nb, K:K+nb] (nb <= 64; padded with the identity to 64), one workgroup of
// 512 threads per matrix.  LU with the diagonal pivots in four slabs of 16 columns: wave s holds
// columns 16 s .. 16 s + 15 (lane = row) and factors them alone, wave-synchronously (row k
// broadcast with readlane: no barrier per column), checking LAPACK's rule on the candidates of
// each column; its multipliers go through LDS (one barrier per slab) to waves s + 1 .. 3, which
// apply the slab's 16 rank-1 updates to their columns in column order.  Then waves 0-3 form
// L11^{-1} and waves 4-7 U11^{-1} (16 right-hand-side columns of the identity per wave,
// wave-synchronous substitution), stored in inv (L11^{-1} at 0, U11^{-1} at 64 x 64, column-major)
// for the L21 / U12 products of panel_inv_gemm_kernel.  Rows below the block are checked there.
template <class C>
__global__ void __launch_bounds__(512) diag_lu_kernel(MatBatch<C> mb, int n, int K, int* __restrict__ ipiv,
                                                      int64_t ipiv_stride, int* __restrict__ flags, int64_t flag_stride,
                                                      C* __restrict__ inv, int64_t inv_stride, cplx* __restrict__ inv64,
                                                      int64_t inv64_stride, int64_t inv64_u) {
    typedef typename Cx<C>::R R;
    const int b = blockIdx.z;
    C* A = mb.A + mb.stride * b;
    const int lda = mb.lda;
    const int nb = min(NB, n - K);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int* flag = flags + flag_stride * b;
    __shared__ C LU[NB][NB + 1];  // the factored block
    // the slab multipliers (LU phase) and then the triangle being inverted share this buffer
    __shared__ __align__(16) unsigned char xbuf[sizeof(C) * NB * (NB + 1)];
    __shared__ __align__(16) unsigned char tbuf[sizeof(C) * 48 * 17];
    C(*Ls)[NB][17] = reinterpret_cast<C(*)[NB][17]>(xbuf);  // Ls[2][NB][17]
    if (K == 0 && tid == 0) *flag = 0;
    C x[16];
    bool bad = false;
    if (w < 4) {
#pragma unroll
        for (int c = 0; c < 16; c++) {
            const int col = 16 * w + c;
            x[c] = (lane < nb && col < nb) ? A[(K + lane) + (int64_t)lda * (K + col)]
                                           : tmk<C>(lane == col ? R(1) : R(0), R(0));
        }
    }
    for (int s = 0; s < 4; s++) {
        if (w == s) {
#pragma unroll
            for (int kk = 0; kk < 16; kk++) {
                const int k = 16 * s + kk;
                const C u = rlane(x[kk], k);
                const bool zero = u.x == 0 && u.y == 0;
                const C iv = zero ? tmk<C>(0, 0) : crecip(u);
                const R au = tabs1(u);
                bad = bad || zero;
                if (lane > k) {
                    bad = bad || tabs1(x[kk]) > au;
                    x[kk] = tmul(x[kk], iv);
                }
#pragma unroll
                for (int cc = kk + 1; cc < 16; cc++) {
                    const C ukc = rlane(x[cc], k);
                    if (lane > k) x[cc] = tsub(x[cc], tmul(x[kk], ukc));
                }
            }
#pragma unroll
            for (int c = 0; c < 16; c++) Ls[s & 1][lane][c] = x[c];
        }
        __syncthreads();
        if (w > s && w < 4) {
            C l[16];
#pragma unroll
            for (int kk = 0; kk < 16; kk++) l[kk] = Ls[s & 1][lane][kk];
#pragma unroll
            for (int kk = 0; kk < 16; kk++) {
                const int k = 16 * s + kk;
#pragma unroll
                for (int c = 0; c < 16; c++) {
                    const C ukc = rlane(x[c], k);
                    if (lane > k) x[c] = tsub(x[c], tmul(l[kk], ukc));
                }
            }
        }
    }
    if (w < 4) {
#pragma unroll
        for (int c = 0; c < 16; c++) {
            const int col = 16 * w + c;
            LU[lane][col] = x[c];
            if (lane < nb && col < nb) A[(K + lane) + (int64_t)lda * (K + col)] = x[c];
        }
    }
    if (tid < nb) ipiv[ipiv_stride * b + K + tid] = K + tid;
    if (bad) *flag = 1;
    __syncthreads();
    // inverses, one triangle at a time in X (the slab buffer's space): the four 16 x 16 diagonal
    // blocks by substitution (wave w: block w % 4, right-hand-side columns 8 (w / 4) .. + 7; lane =
    // row r (lane % 16) and column pair g (lane / 16)), then the off-diagonal blocks block row by
    // block row: X_ij = -X_ii sum_k T_ik X_kj with T = L (k = j .. i - 1, rows downwards) or U
    // (k = i + 1 .. j, rows upwards).
    C(*X)[NB + 1] = reinterpret_cast<C(*)[NB + 1]>(xbuf);
    C(*T)[17] = reinterpret_cast<C(*)[17]>(tbuf);
    const int blk = w & 3, r16 = lane & 15, g = lane >> 4, cb = 8 * (w >> 2) + 2 * g;
    for (int tri = 0; tri < 2; tri++) {
        C y[2];
#pragma unroll
        for (int c = 0; c < 2; c++) y[c] = tmk<C>(r16 == cb + c ? R(1) : R(0), R(0));
        if (tri == 0) {
            for (int k = 0; k < 15; k++) {
                const C lrk = LU[16 * blk + r16][16 * blk + k];
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    const C yk = tmk<C>(__shfl(y[c].x, k + 16 * g), __shfl(y[c].y, k + 16 * g));
                    if (r16 > k) y[c] = tsub(y[c], tmul(lrk, yk));
                }
            }
        } else {
            for (int k = 15; k >= 0; k--) {
                const C iv = crecip(LU[16 * blk + k][16 * blk + k]);
                const C urk = LU[16 * blk + r16][16 * blk + k];
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    if (r16 == k) y[c] = tmul(y[c], iv);
                    const C yk = tmk<C>(__shfl(y[c].x, k + 16 * g), __shfl(y[c].y, k + 16 * g));
                    if (r16 < k) y[c] = tsub(y[c], tmul(urk, yk));
                }
            }
        }
        // the whole triangle buffer: zero, then the diagonal blocks
        for (int e = tid; e < NB * NB; e += 512) X[e & 63][e >> 6] = tmk<C>(0, 0);
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 2; c++) X[16 * blk + r16][16 * blk + cb + c] = y[c];
        __syncthreads();
        for (int step = 0; step < 3; step++) {
            const int i = tri == 0 ? step + 1 : 2 - step;  // block row
            const int nbj = tri == 0 ? i : 3 - i;           // blocks j in this row
            // T_j = sum_k LU_ik X_kj
            for (int o = tid; o < nbj * 256; o += 512) {
                const int jj = o >> 8, r = (o >> 4) & 15, c = o & 15;
                const int j = tri == 0 ? jj : i + 1 + jj;
                const int k0 = tri == 0 ? j : i + 1, k1 = tri == 0 ? i - 1 : j;
                C acc = tmk<C>(0, 0);
                for (int k = k0; k <= k1; k++)
#pragma unroll 4
                    for (int t = 0; t < 16; t++)
                        acc = tadd(acc, tmul(LU[16 * i + r][16 * k + t], X[16 * k + t][16 * j + c]));
                T[16 * jj + r][c] = acc;
            }
            __syncthreads();
            // X_ij = -X_ii T_j
            for (int o = tid; o < nbj * 256; o += 512) {
                const int jj = o >> 8, r = (o >> 4) & 15, c = o & 15;
                const int j = tri == 0 ? jj : i + 1 + jj;
                C acc = tmk<C>(0, 0);
#pragma unroll 4
                for (int t = 0; t < 16; t++) acc = tadd(acc, tmul(X[16 * i + r][16 * i + t], T[16 * jj + t][c]));
                X[16 * i + r][16 * j + c] = tmk<C>(-acc.x, -acc.y);
            }
            __syncthreads();
        }
        C* out = inv + inv_stride * b + (tri == 0 ? 0 : NB * NB);
        for (int e = tid; e < NB * NB; e += 512) out[e] = X[e & 63][e >> 6];
        // the same inverse in FP64, as block K / 64 of the triangular solves' inverse tables
        if (inv64) {
            cplx* o64 = inv64 + inv64_stride * b + (int64_t)(K / NB) * NB * NB + (tri == 0 ? 0 : inv64_u);
            for (int e = tid; e < NB * NB; e += 512) o64[e] = to_d(X[e & 63][e >> 6]);
        }
        __syncthreads();
    }
}

// L21 = A21 U11^{-1} (part 0: rows [K + nb, n) in 64-row tiles) and U12 = L11^{-1} A12 (part 1:
// columns [K + nb, n) in 64-column tiles), both in place, from the inverses of diag_lu_kernel:
// 64 x 64 x 64 complex products on MFMA (as zgemm_mfma_kernel: FP64 four real products per
// complex product, FP32 three).  Part 0 also checks LAPACK's rule on the rows below the block:
// the updated candidate of column k is l_ik u_kk, whose |Re| + |Im| must not exceed |u_kk|'s.
template <class C>
__global__ void __launch_bounds__(256) panel_inv_gemm_kernel(MatBatch<C> mb, int n, int K, const C* inv,
                                                             int64_t inv_stride, int* flags, int64_t flag_stride) {
    typedef typename Cx<C>::R R;
    typedef Mfma16<R> M;
    constexpr bool THREE = sizeof(R) == 4;
    const int b = blockIdx.z, part = blockIdx.y;
    C* Am = mb.A + mb.stride * b;
    const int lda = mb.lda;
    const int nb = min(NB, n - K), KE = K + nb;
    const C* iv = inv + inv_stride * b;
    // operands: X (64 rows x 64 k) and Y (64 k x 64 cols); the product written over X (part 0) or Y
    const int r0 = KE + 64 * blockIdx.x;  // part 0: first row; part 1: first column
    const int rows = part == 0 ? min(64, n - r0) : nb;
    const int cols = part == 0 ? nb : min(64, n - r0);
    const C* X = part == 0 ? Am + r0 + (int64_t)lda * K : iv;       // A21 tile or L^{-1}
    const int ldx = part == 0 ? lda : NB;
    const C* Y = part == 0 ? iv + NB * NB : Am + K + (int64_t)lda * r0;  // U^{-1} or A12 tile
    const int ldy = part == 0 ? NB : lda;
    C* Out = part == 0 ? Am + r0 + (int64_t)lda * K : Am + K + (int64_t)lda * r0;
    constexpr int KC = 16;
    __shared__ R Ar[KC][80], Ai[KC][80], Br[KC][74], Bi[KC][74];
    __shared__ R au[NB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    if (part == 0 && tid < NB) au[tid] = tid < nb ? tabs1(Am[(K + tid) + (int64_t)lda * (K + tid)]) : R(0);
    typename M::V t1[2][2], t2[2][2], t3[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++)
#pragma unroll
            for (int v = 0; v < 4; v++) t1[bi][bj][v] = t2[bi][bj][v] = t3[bi][bj][v] = 0;
    for (int k0 = 0; k0 < NB; k0 += KC) {
#pragma unroll
        for (int e = 0; e < KC / 4; e++) {
            const int idx = tid + 256 * e;
            const int r = idx & 63, ka = idx >> 6;
            const C xa = (r < rows && k0 + ka < nb) ? X[r + (int64_t)ldx * (k0 + ka)] : tmk<C>(0, 0);
            Ar[ka][r] = xa.x;
            Ai[ka][r] = xa.y;
            const int kb = idx % KC, c = idx / KC;
            const C yb = (c < cols && k0 + kb < nb) ? Y[(k0 + kb) + (int64_t)ldy * c] : tmk<C>(0, 0);
            Br[kb][c] = yb.x;
            Bi[kb][c] = yb.y;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < KC; ks += 4) {
            R ar[2], ai[2], br[2], bi_[2];
#pragma unroll
            for (int q = 0; q < 2; q++) {
                ar[q] = Ar[ks + lk][wm + q * 16 + li];
                ai[q] = Ai[ks + lk][wm + q * 16 + li];
                br[q] = Br[ks + lk][wn + q * 16 + li];
                bi_[q] = Bi[ks + lk][wn + q * 16 + li];
            }
#pragma unroll
            for (int bi = 0; bi < 2; bi++)
#pragma unroll
                for (int bj = 0; bj < 2; bj++) {
                    if (THREE) {
                        t1[bi][bj] = M::run(ar[bi], br[bj], t1[bi][bj]);
                        t2[bi][bj] = M::run(ai[bi], bi_[bj], t2[bi][bj]);
                        t3[bi][bj] = M::run(ar[bi] + ai[bi], br[bj] + bi_[bj], t3[bi][bj]);
                    } else {
                        t1[bi][bj] = M::run(ar[bi], br[bj], t1[bi][bj]);
                        t1[bi][bj] = M::run(-ai[bi], bi_[bj], t1[bi][bj]);
                        t3[bi][bj] = M::run(ar[bi], bi_[bj], t3[bi][bj]);
                        t3[bi][bj] = M::run(ai[bi], br[bj], t3[bi][bj]);
                    }
                }
        }
        __syncthreads();
    }
    bool bad = false;
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = wm + bi * 16 + M::row(v, lk), c = wn + bj * 16 + li;
                if (r < rows && c < cols) {
                    R re, im;
                    if (THREE) {
                        re = t1[bi][bj][v] - t2[bi][bj][v];
                        im = t3[bi][bj][v] - t1[bi][bj][v] - t2[bi][bj][v];
                    } else {
                        re = t1[bi][bj][v];
                        im = t3[bi][bj][v];
                    }
                    const C val = tmk<C>(re, im);
                    if (part == 0) {
                        const C ukk = Am[(K + c) + (int64_t)lda * (K + c)];
                        bad = bad || tabs1(tmul(val, ukk)) > au[c];
                    }
                    Out[r + (int64_t)(part == 0 ? lda : lda) * c] = val;
                }
            }
    if (bad) flags[flag_stride * b] = 1;
}

// ---------------------------------------------------------------------------
// C[m x nn] (+)= (-)A[m x kk] * B[kk x nn], complex, column-major, on MFMA
// (v_mfma_f64_16x16x4f64 for cplx, v_mfma_f32_16x16x4f32 for cplxf).
// 256 threads (4 waves in 2x2), 64x64 tile, K chunks of 16 staged in LDS as
// separate real / imaginary planes.  Lane layout of both instructions:
//   A: lane l holds A[l%16][l/16], B: B[l/16][l%16];
//   C: 4 values per lane, f64: C[4v + l/16][l%16], f32: C[4(l/16) + v][l%16].
// ---------------------------------------------------------------------------
// Complex product per k-step: FP64 uses the four real products (Re = ar br - ai bi, Im = ar bi +
// ai br), rounding as the reference's zgemm; FP32 (the mixed-precision factor) uses three,
// T1 = ar br, T2 = ai bi, T3 = (ar + ai)(br + bi): Re = T1 - T2, Im = T3 - T1 - T2.
template <bool SUB, bool ACCUM, class C>
__global__ void __launch_bounds__(256) zgemm_mfma_kernel(int m, int nn, int kk, const C* __restrict__ Ab, int lda,
                                                         int64_t sA, const C* __restrict__ Bb, int ldb, int64_t sB,
                                                         C* __restrict__ Cb, int ldc, int64_t sC) {
    typedef typename Cx<C>::R R;
    typedef Mfma16<R> M;
    constexpr bool THREE = sizeof(R) == 4;
    const int bz = blockIdx.z;
    const C* A = Ab + sA * bz;
    const C* B = Bb + sB * bz;
    C* Cm = Cb + sC * bz;
    // Row strides padded against LDS bank conflicts: the B tile is stored by lanes that run down
    // k (16 rows) for 4 columns, which with a 64-word row stride put 16 lanes on one bank; 74
    // spreads them, and 80 puts the four k-rows of an A read on distinct bank groups.
    constexpr int KC = 16;  // K chunk (32 raises the registers and halves the occupancy: slower)
    __shared__ R Ar[KC][80], Ai[KC][80], Br[KC][74], Bi[KC][74];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int row0 = blockIdx.x * 64, col0 = blockIdx.y * 64;
    typename M::V t1[2][2], t2[2][2], t3[2][2];
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++)
#pragma unroll
            for (int v = 0; v < 4; v++) t1[bi][bj][v] = t2[bi][bj][v] = t3[bi][bj][v] = 0;
    // A tile (64 rows x 16 k, rows contiguous) and B tile (16 k x 64 cols, k contiguous): each
    // thread holds 4 + 4 elements.  The next chunk's loads are issued before the current chunk's
    // MFMAs, so their latency hides behind the compute.
    constexpr int EPT = KC / 4;  // elements of each tile per thread
    C ra[EPT], rb[EPT];
    auto load_chunk = [&](int k0) {
#pragma unroll
        for (int e = 0; e < EPT; e++) {
            const int idx = tid + 256 * e;
            const int r = idx & 63, ka = idx >> 6;
            ra[e] = (row0 + r < m && k0 + ka < kk) ? A[(row0 + r) + (int64_t)lda * (k0 + ka)] : tmk<C>(0, 0);
            const int kb = idx % KC, c = idx / KC;
            rb[e] = (col0 + c < nn && k0 + kb < kk) ? B[(k0 + kb) + (int64_t)ldb * (col0 + c)] : tmk<C>(0, 0);
        }
    };
    load_chunk(0);
    for (int k0 = 0; k0 < kk; k0 += KC) {
#pragma unroll
        for (int e = 0; e < EPT; e++) {
            const int idx = tid + 256 * e;
            const int r = idx & 63, ka = idx >> 6;
            Ar[ka][r] = SUB ? -ra[e].x : ra[e].x;
            Ai[ka][r] = SUB ? -ra[e].y : ra[e].y;
            const int kb = idx % KC, c = idx / KC;
            Br[kb][c] = rb[e].x;
            Bi[kb][c] = rb[e].y;
        }
        __syncthreads();
        if (k0 + KC < kk) load_chunk(k0 + KC);
#pragma unroll
        for (int ks = 0; ks < KC; ks += 4) {
            R ar[2], ai[2], br[2], bi_[2];
#pragma unroll
            for (int q = 0; q < 2; q++) {
                ar[q] = Ar[ks + lk][wm + q * 16 + li];
                ai[q] = Ai[ks + lk][wm + q * 16 + li];
                br[q] = Br[ks + lk][wn + q * 16 + li];
                bi_[q] = Bi[ks + lk][wn + q * 16 + li];
            }
#pragma unroll
            for (int bi = 0; bi < 2; bi++)
#pragma unroll
                for (int bj = 0; bj < 2; bj++) {
                    if (THREE) {
                        t1[bi][bj] = M::run(ar[bi], br[bj], t1[bi][bj]);
                        t2[bi][bj] = M::run(ai[bi], bi_[bj], t2[bi][bj]);
                        t3[bi][bj] = M::run(ar[bi] + ai[bi], br[bj] + bi_[bj], t3[bi][bj]);
                    } else {  // t1: real part, t3: imaginary part
                        t1[bi][bj] = M::run(ar[bi], br[bj], t1[bi][bj]);
                        t1[bi][bj] = M::run(-ai[bi], bi_[bj], t1[bi][bj]);
                        t3[bi][bj] = M::run(ar[bi], bi_[bj], t3[bi][bj]);
                        t3[bi][bj] = M::run(ai[bi], br[bj], t3[bi][bj]);
                    }
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int r = row0 + wm + bi * 16 + M::row(v, lk), c = col0 + wn + bj * 16 + li;
                if (r < m && c < nn) {
                    R re, im;
                    if (THREE) {
                        re = t1[bi][bj][v] - t2[bi][bj][v];
                        im = t3[bi][bj][v] - t1[bi][bj][v] - t2[bi][bj][v];
                    } else {
                        re = t1[bi][bj][v];
                        im = t3[bi][bj][v];
                    }
                    if (ACCUM) {
                        const C old = Cm[r + (int64_t)ldc * c];
                        re += old.x;
                        im += old.y;
                    }
                    Cm[r + (int64_t)ldc * c] = tmk<C>(re, im);
                }
            }
}

template <class C>
static void gemm_t(hipStream_t st, bool sub, bool accum, int m, int nn, int kk, const C* A, int lda, int64_t sA,
                   const C* B, int ldb, int64_t sB, C* Cm, int ldc, int64_t sC, int batch) {
    if (m <= 0 || nn <= 0) return;
    dim3 grid((m + 63) / 64, (nn + 63) / 64, batch);
    if (sub && accum)
        zgemm_mfma_kernel<true, true, C><<<grid, 256, 0, st>>>(m, nn, kk, A, lda, sA, B, ldb, sB, Cm, ldc, sC);
    else if (!sub && accum)
        zgemm_mfma_kernel<false, true, C><<<grid, 256, 0, st>>>(m, nn, kk, A, lda, sA, B, ldb, sB, Cm, ldc, sC);
    else
        zgemm_mfma_kernel<false, false, C><<<grid, 256, 0, st>>>(m, nn, kk, A, lda, sA, B, ldb, sB, Cm, ldc, sC);
}

void zgemm(hipStream_t st, bool sub, bool accum, int m, int nn, int kk, const cplx* A, int lda, int64_t sA,
           const cplx* B, int ldb, int64_t sB, cplx* C, int ldc, int64_t sC, int batch) {
    gemm_t(st, sub, accum, m, nn, kk, A, lda, sA, B, ldb, sB, C, ldc, sC, batch);
}

// ---------------------------------------------------------------------------
// Split-bf16 trailing update of the FP32 factor (mixed-precision solver).  Each panel element
// x is stored as two bf16 numbers, hi = bf16(x) and lo = bf16(x - hi) (round to nearest even),
// and a product x y is taken as hi*hi + hi*lo + lo*hi: bf16 products are exact in FP32, so the
// error per product is about 2^-16 relative (the dropped lo*lo and the rounding of lo), against
// 2^-24 for FP32; the solver's FP64 refinement removes the difference.  The products run on
// v_mfma_f32_32x32x16_bf16 with FP32 accumulation: Re = Br Ar - Bi Ai and Im = Bi Ar + Br Ai
// (six MFMAs each).  The MFMA's M dimension is the column of C and N its row, so that a
// register of the 32x32 accumulator holds 32 consecutive rows of one column (coalesced
// read-modify-write of the column-major C).
//
// Split arrays of one matrix (bf16, padded to SPLIT_TILE rows/cols, ks = 64 S k slots for the
// S <= SPLIT_SUPER panels of one update), chunk-major: k in chunks of 16, and within a chunk the
// 16 slots of a row contiguous (32 bytes), rows after each other.  Array q, row r, slot k is bf16
// number ((q * (ks / 16) + k / 16) * mpad + r) * 16 + k % 16.  A wave's load of one chunk of 32
// rows is then 1 KB contiguous (with rows k-contiguous over all of ks, the same load touched 32
// cache lines and used a quarter of each: 22 % slower updates).
//   q = 0 .. 3  L21 rows:    re hi, re lo, im hi, im lo
//   q = 4 .. 7  U12 columns: re hi, re lo, im hi, im lo
// ---------------------------------------------------------------------------
constexpr int SPLIT_TILE = 128;
#ifndef SPLIT_GROUP
#define SPLIT_GROUP 8
#endif
constexpr int SPLIT_SUPER = kSplitSuperMax;  // most panels per trailing update on the split path (k <= 512)
constexpr int SPLIT_KC = 16;    // k chunk of the layout and of the update's LDS stages
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

static inline int split_pad(int n) { return (n + SPLIT_TILE - 1) / SPLIT_TILE * SPLIT_TILE; }
int64_t split_work(int n) { return (int64_t)8 * split_pad(n) * SPLIT_SUPER * NB; }

__device__ __forceinline__ unsigned short bf16_rne(float f) {
    unsigned int u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

// thread (kc, r), r fastest: k slots 16 kc .. 16 kc + 15 of row r of L21 (rows [KE, n), columns
// [K, KE); consecutive threads read consecutive rows of a column) and of column r of U12 (rows
// [K, KE), columns [KE, n)); rows past n - KE and k past KE - K are zero
__global__ void split_panels_kernel(MatBatch<cplxf> mb, int n, int K, int KE, int mpad, int ks,
                                    unsigned short* __restrict__ Sb, int64_t sS) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nkc = ks / SPLIT_KC;
    const int kc = t / mpad, r = t - kc * mpad;
    if (kc >= nkc) return;
    const cplxf* A = mb.A + mb.stride * blockIdx.z;
    unsigned short* S = Sb + sS * blockIdx.z;
    const int m = n - KE, kk = KE - K, lda = mb.lda;
#pragma unroll
    for (int p = 0; p < 2; p++) {
        unsigned int w[4][SPLIT_KC / 2];  // [array][slot pair]
#pragma unroll
        for (int e = 0; e < SPLIT_KC; e++) {
            const int k = SPLIT_KC * kc + e;
            cplxf a = make_float2(0.f, 0.f);
            if (r < m && k < kk) a = p == 0 ? A[(KE + r) + (int64_t)lda * (K + k)] : A[(K + k) + (int64_t)lda * (KE + r)];
            const float v[2] = {a.x, a.y};
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const unsigned short hi = bf16_rne(v[c]);
                const unsigned short lo = bf16_rne(v[c] - bf16_val(hi));
                if (e & 1) {
                    w[2 * c][e / 2] |= (unsigned)hi << 16;
                    w[2 * c + 1][e / 2] |= (unsigned)lo << 16;
                } else {
                    w[2 * c][e / 2] = hi;
                    w[2 * c + 1][e / 2] = lo;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint4* dst = reinterpret_cast<uint4*>(S + (((int64_t)(4 * p + q) * nkc + kc) * mpad + r) * SPLIT_KC);
            dst[0] = make_uint4(w[q][0], w[q][1], w[q][2], w[q][3]);
            dst[1] = make_uint4(w[q][4], w[q][5], w[q][6], w[q][7]);
        }
    }
}

__device__ __forceinline__ bf16x8_t bf16x8_neg(bf16x8_t v) {
    uint4 u = __builtin_bit_cast(uint4, v);
    u.x ^= 0x80008000u;
    u.y ^= 0x80008000u;
    u.z ^= 0x80008000u;
    u.w ^= 0x80008000u;
    return __builtin_bit_cast(bf16x8_t, u);
}

// C[0:m, 0:m] -= L21 U12 from the split arrays.  Workgroup tile 128 x 128 of C, four waves of
// 64 x 64 (2 x 2 accumulator tiles of 32 x 32, real and imaginary); k in chunks of 16.  The
// operands of a chunk go from HBM straight into LDS (global_load_lds_dwordx4: a wave's 64 lanes
// write 1 KB contiguous, a chunk of 128 rows of one array being 4 KB contiguous in the
// chunk-major layout), into two LDS stages: chunk c + 1 is in flight while chunk c's MFMAs run,
// one barrier per chunk.  A chunk's operands are read into registers before the next chunk's
// loads are issued.  Rows are 32 B (two 16-byte halves): the eight lanes of a 16-byte read phase
// read eight consecutive rows and hit distinct banks.  Two workgroups per CU (64 KB of LDS each;
// C read in the epilogue): one workgroup's C read-modify-write overlaps the other's MFMAs.
typedef __attribute__((address_space(3))) void* lds_void_ptr;

__global__ void __launch_bounds__(256, 2) zgemm_split_kernel(int m, int mpad, int ks, const unsigned short* __restrict__ Sb,
                                                             int64_t sS, cplxf* __restrict__ Cb, int ldc, int64_t sC) {
    // [stage][L21 rows, U12 columns][array q][row * 2 + half]; L21 rows are the MFMA B operand (C
    // rows), U12 columns the A operand (C columns)
    __shared__ bf16x8_t Ls[2 * 2 * 4 * 2 * SPLIT_TILE];
    // XCD-aware tile order: workgroups are dealt to the 8 XCDs round-robin by linear id, so the
    // id is remapped to give each XCD a contiguous run of tiles; within a matrix the run walks
    // groups of SPLIT_GROUP row tiles, row tile fastest, so that the 64 tiles an XCD holds at once
    // form a SPLIT_GROUP x (64 / SPLIT_GROUP) block and share its row and column panels in the L2
    int L = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    {
        const int T = gridDim.x * gridDim.y * gridDim.z, q = T / 8;
        if (L < 8 * q) L = (L % 8) * q + L / 8;
    }
    const int per = gridDim.x * gridDim.y, tz = L / per, t = L - tz * per;
    const int grp = t / (SPLIT_GROUP * gridDim.y), first = grp * SPLIT_GROUP;
    const int gs = min((int)gridDim.x - first, SPLIT_GROUP), r = t - grp * SPLIT_GROUP * gridDim.y;
    const int tx = first + r % gs, ty = r / gs;
    const unsigned short* S = Sb + sS * tz;
    cplxf* Cm = Cb + sC * tz;
    const int tid = threadIdx.x, lane = tid & 63, h = lane >> 5, l32 = lane & 31;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int row0 = tx * SPLIT_TILE, col0 = ty * SPLIT_TILE;
    const int wr = (wave & 1) * 64, wc = (wave >> 1) * 64;
    f32x16_t acc_re[2][2], acc_im[2][2];  // [column tile][row tile]
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc_re[a][b][v] = acc_im[a][b][v] = 0.f;
    const int nkc = ks / SPLIT_KC;
    const bf16x8_t* Sa = reinterpret_cast<const bf16x8_t*>(S);                              // arrays 0..3
    const bf16x8_t* Sbv = reinterpret_cast<const bf16x8_t*>(S + (int64_t)4 * mpad * ks);  // arrays 4..7
    constexpr int kArr = 2 * SPLIT_TILE;  // 16-byte units of one array's chunk
    // chunk kc into stage st: thread tid loads unit tid of each array's chunk, wave w's lanes
    // landing at units 64 w .. 64 w + 63 of the stage's array
    auto issue = [&](int kc, int st) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const bf16x8_t* ga = Sa + (((int64_t)q * nkc + kc) * mpad + row0) * 2 + tid;
            const bf16x8_t* gb = Sbv + (((int64_t)q * nkc + kc) * mpad + col0) * 2 + tid;
            __builtin_amdgcn_global_load_lds((const void*)ga, (lds_void_ptr)(Ls + ((st * 2 + 0) * 4 + q) * kArr + 64 * wave),
                                             16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void*)gb, (lds_void_ptr)(Ls + ((st * 2 + 1) * 4 + q) * kArr + 64 * wave),
                                             16, 0, 0);
        }
    };
    issue(0, 0);
    for (int kc = 0; kc < nkc; kc++) {
        const int st = kc & 1;
        // this wave's loads of chunk kc have landed; after the barrier every wave's have, and every
        // wave has read stage st ^ 1 (chunk kc - 1) into registers
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        const bf16x8_t* La = Ls + (st * 2 + 0) * 4 * kArr;
        const bf16x8_t* Lb = Ls + (st * 2 + 1) * 4 * kArr;
        bf16x8_t ar[4][2], br[4][2];  // [array: re hi, re lo, im hi, im lo][32-row tile]
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int t = 0; t < 2; t++) {
                ar[q][t] = La[q * kArr + (wr + 32 * t + l32) * 2 + h];
                br[q][t] = Lb[q * kArr + (wc + 32 * t + l32) * 2 + h];
            }
        if (kc + 1 < nkc) issue(kc + 1, st ^ 1);
#pragma unroll
        for (int a = 0; a < 2; a++) {
            const bf16x8_t brh = br[0][a], brl = br[1][a], bih = br[2][a], bil = br[3][a];
            const bf16x8_t nbih = bf16x8_neg(bih), nbil = bf16x8_neg(bil);
#pragma unroll
            for (int b = 0; b < 2; b++) {
                f32x16_t re = acc_re[a][b], im = acc_im[a][b];
                re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(brh, ar[0][b], re, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(brh, ar[1][b], re, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(brl, ar[0][b], re, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nbih, ar[2][b], re, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nbih, ar[3][b], re, 0, 0, 0);
                re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nbil, ar[2][b], re, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bih, ar[0][b], im, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bih, ar[1][b], im, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bil, ar[0][b], im, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(brh, ar[2][b], im, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(brh, ar[3][b], im, 0, 0, 0);
                im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(brl, ar[2][b], im, 0, 0, 0);
                acc_re[a][b] = re;
                acc_im[a][b] = im;
            }
        }
    }
    // epilogue: accumulator register v of lane (l32 + 32 h) is C row l32 (N index), column
    // 8 (v / 4) + 4 h + v % 4 (M index) of its 32 x 32 tile; each access covers 32 consecutive
    // rows of two columns.  A wave tile inside C (the common case) reads all 32 values of a
    // column tile at once; edge tiles check every element.
    if (row0 + wr + 64 <= m && col0 + wc + 64 <= m) {
#pragma unroll
        for (int a = 0; a < 2; a++) {
            cplxf old[2][16];
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int r = row0 + wr + 32 * b + l32, c = col0 + wc + 32 * a + 8 * (v >> 2) + 4 * h + (v & 3);
                    old[b][v] = Cm[r + (int64_t)ldc * c];
                }
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int r = row0 + wr + 32 * b + l32, c = col0 + wc + 32 * a + 8 * (v >> 2) + 4 * h + (v & 3);
                    Cm[r + (int64_t)ldc * c] =
                        make_float2(old[b][v].x - acc_re[a][b][v], old[b][v].y - acc_im[a][b][v]);
                }
        }
        return;
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const int r = row0 + wr + 32 * b + l32;
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int c = col0 + wc + 32 * a + 8 * (v >> 2) + 4 * h + (v & 3);
                if (r < m && c < m) {
                    const cplxf o = Cm[r + (int64_t)ldc * c];
                    Cm[r + (int64_t)ldc * c] = make_float2(o.x - acc_re[a][b][v], o.y - acc_im[a][b][v]);
                }
            }
        }
}

// ---------------------------------------------------------------------------
// Split-bf16 products of FP32 complex operands read straight from the matrix (the split path's
// panel-side products: the left-looking updates inside a super-panel and the L21 / U12 products
// with the diagonal-block inverses).  Each operand element is split into bf16 hi + lo while it is
// staged into LDS, and the products run as in zgemm_split_kernel (hi*hi + hi*lo + lo*hi on
// v_mfma_f32_32x32x16_bf16, FP32 accumulation): 1,024 flops per cycle and SIMD against 64 for the
// FP32 MFMA of zgemm_mfma_kernel<cplxf>, whose three-product form left these small-k products
// compute-bound.  Their precision is the trailing updates' (products to ~2^-16).
// Tile 64 x 64 per workgroup of 256 threads: wave w owns the 32 x 32 block of rows 32 (w & 1) and
// columns 32 (w >> 1); k in chunks of 16, the next chunk's FP32 values loaded into registers
// during the current chunk's MFMAs.
// ---------------------------------------------------------------------------
struct FsplitLds {
    // [array re hi, re lo, im hi, im lo][row or column][2 units of 8 k + pad] as 8-byte halves
    __align__(16) uint2 a[4][64][6];
    __align__(16) uint2 b[4][64][6];
};

__device__ __forceinline__ uint2 split_pack4(float x0, float x1, float x2, float x3, bool lo) {
    const float v[4] = {x0, x1, x2, x3};
    unsigned short h[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const unsigned short hi = bf16_rne(v[e]);
        h[e] = lo ? bf16_rne(v[e] - bf16_val(hi)) : hi;
    }
    return make_uint2(h[0] | ((unsigned)h[1] << 16), h[2] | ((unsigned)h[3] << 16));
}

// acc = A[0:rows, 0:kk] B[0:kk, 0:cols] for the 64 x 64 tile whose first row of A / column of B
// the pointers address (rows, cols <= 64); acc register v of lane (l32, h) of wave w is element
// (32 (w & 1) + l32, 32 (w >> 1) + 8 (v / 4) + 4 h + v % 4)
__device__ __forceinline__ void fsplit_tile64(const cplxf* __restrict__ A, int lda, const cplxf* __restrict__ B,
                                              int ldb, int rows, int cols, int kk, FsplitLds& L, f32x16_t& acc_re,
                                              f32x16_t& acc_im) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, h = lane >> 5, l32 = lane & 31;
    const int wr = 32 * (wave & 1), wc = 32 * (wave >> 1);
#pragma unroll
    for (int v = 0; v < 16; v++) acc_re[v] = acc_im[v] = 0.f;
    // staging: A element (row ar, k 4 akq + e), B element (k 4 bkq + e, column bc), e = 0..3
    const int ar = tid & 63, akq = tid >> 6, bc = tid >> 2, bkq = tid & 3;
    cplxf ra[4], rb[4];
    auto load_chunk = [&](int k0) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int ka = k0 + 4 * akq + e, kb = k0 + 4 * bkq + e;
            ra[e] = (ar < rows && ka < kk) ? A[ar + (int64_t)lda * ka] : make_float2(0.f, 0.f);
            rb[e] = (bc < cols && kb < kk) ? B[kb + (int64_t)ldb * bc] : make_float2(0.f, 0.f);
        }
    };
    auto store_chunk = [&]() {
        const int ua = 2 * (akq >> 1) + (akq & 1), ub = 2 * (bkq >> 1) + (bkq & 1);  // 8-byte slot in the row
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int lo = 0; lo < 2; lo++) {
                L.a[2 * c + lo][ar][ua] = c == 0 ? split_pack4(ra[0].x, ra[1].x, ra[2].x, ra[3].x, lo)
                                                 : split_pack4(ra[0].y, ra[1].y, ra[2].y, ra[3].y, lo);
                L.b[2 * c + lo][bc][ub] = c == 0 ? split_pack4(rb[0].x, rb[1].x, rb[2].x, rb[3].x, lo)
                                                 : split_pack4(rb[0].y, rb[1].y, rb[2].y, rb[3].y, lo);
            }
    };
    load_chunk(0);
    for (int k0 = 0; k0 < kk; k0 += 16) {
        store_chunk();
        __syncthreads();
        if (k0 + 16 < kk) load_chunk(k0 + 16);
        const bf16x8_t arh = *reinterpret_cast<const bf16x8_t*>(&L.a[0][wr + l32][2 * h]);
        const bf16x8_t arl = *reinterpret_cast<const bf16x8_t*>(&L.a[1][wr + l32][2 * h]);
        const bf16x8_t aih = *reinterpret_cast<const bf16x8_t*>(&L.a[2][wr + l32][2 * h]);
        const bf16x8_t ail = *reinterpret_cast<const bf16x8_t*>(&L.a[3][wr + l32][2 * h]);
        const bf16x8_t brh = *reinterpret_cast<const bf16x8_t*>(&L.b[0][wc + l32][2 * h]);
        const bf16x8_t brl = *reinterpret_cast<const bf16x8_t*>(&L.b[1][wc + l32][2 * h]);
        const bf16x8_t bih = *reinterpret_cast<const bf16x8_t*>(&L.b[2][wc + l32][2 * h]);
        const bf16x8_t bil = *reinterpret_cast<const bf16x8_t*>(&L.b[3][wc + l32][2 * h]);
        const bf16x8_t nbih = bf16x8_neg(bih), nbil = bf16x8_neg(bil);
        acc_re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(brh, arh, acc_re, 0, 0, 0);
        acc_im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bih, arh, acc_im, 0, 0, 0);
        acc_re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(brh, arl, acc_re, 0, 0, 0);
        acc_im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bih, arl, acc_im, 0, 0, 0);
        acc_re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(brl, arh, acc_re, 0, 0, 0);
        acc_im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bil, arh, acc_im, 0, 0, 0);
        acc_re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nbih, aih, acc_re, 0, 0, 0);
        acc_im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(brh, aih, acc_im, 0, 0, 0);
        acc_re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nbih, ail, acc_re, 0, 0, 0);
        acc_im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(brh, ail, acc_im, 0, 0, 0);
        acc_re = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nbil, aih, acc_re, 0, 0, 0);
        acc_im = __builtin_amdgcn_mfma_f32_32x32x16_bf16(brl, aih, acc_im, 0, 0, 0);
        __syncthreads();
    }
}

// C[m x nn] (+)= (-)A[m x kk] B[kk x nn], FP32 complex column-major, split-bf16 products
template <bool SUB, bool ACCUM>
__global__ void __launch_bounds__(256) zgemm_fsplit_kernel(int m, int nn, int kk, const cplxf* __restrict__ Ab, int lda,
                                                           int64_t sA, const cplxf* __restrict__ Bb, int ldb, int64_t sB,
                                                           cplxf* __restrict__ Cb, int ldc, int64_t sC) {
    __shared__ FsplitLds L;
    const int bz = blockIdx.z;
    const int row0 = blockIdx.x * 64, col0 = blockIdx.y * 64;
    f32x16_t acc_re, acc_im;
    fsplit_tile64(Ab + sA * bz + row0, lda, Bb + sB * bz + (int64_t)ldb * col0, ldb, min(64, m - row0),
                  min(64, nn - col0), kk, L, acc_re, acc_im);
    cplxf* Cm = Cb + sC * bz;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5, l32 = lane & 31;
    const int r = row0 + 32 * (wave & 1) + l32;
    if (r >= m) return;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int c = col0 + 32 * (wave >> 1) + 8 * (v >> 2) + 4 * h + (v & 3);
        if (c < nn) {
            float re = SUB ? -acc_re[v] : acc_re[v], im = SUB ? -acc_im[v] : acc_im[v];
            if (ACCUM) {
                const cplxf old = Cm[r + (int64_t)ldc * c];
                re += old.x;
                im += old.y;
            }
            Cm[r + (int64_t)ldc * c] = make_float2(re, im);
        }
    }
}

static void gemm_fsplit(hipStream_t st, bool sub, bool accum, int m, int nn, int kk, const cplxf* A, int lda,
                        int64_t sA, const cplxf* B, int ldb, int64_t sB, cplxf* Cm, int ldc, int64_t sC, int batch) {
    if (m <= 0 || nn <= 0) return;
    dim3 grid((m + 63) / 64, (nn + 63) / 64, batch);
    if (sub && accum)
        zgemm_fsplit_kernel<true, true><<<grid, 256, 0, st>>>(m, nn, kk, A, lda, sA, B, ldb, sB, Cm, ldc, sC);
    else if (!sub && accum)
        zgemm_fsplit_kernel<false, true><<<grid, 256, 0, st>>>(m, nn, kk, A, lda, sA, B, ldb, sB, Cm, ldc, sC);
    else
        zgemm_fsplit_kernel<false, false><<<grid, 256, 0, st>>>(m, nn, kk, A, lda, sA, B, ldb, sB, Cm, ldc, sC);
}

// Where panel_inv_fsplit_kernel also writes its L21 rows and U12 columns as the split arrays of
// the coming trailing update (S null: not written): the update of A[KS:, KS:] by the panels of the
// super-panel starting at column K0, nkc k chunks of 16, rows / columns padded to mpad.  Only rows
// (L21) and columns (U12) >= KS go there; L21 and U12 of a panel are final once this kernel has
// run, so the update needs no separate split pass over them.
struct SplitDst {
    unsigned short* S;
    int64_t sS;
    int KS, K0, nkc, mpad;
};

// panel_inv_gemm_kernel on split-bf16 products: L21 = A21 U11^{-1} (part 0, with the check of
// LAPACK's rule on the rows below the block) and U12 = L11^{-1} A12 (part 1), in place
__global__ void __launch_bounds__(256) panel_inv_fsplit_kernel(MatBatch<cplxf> mb, int n, int K, const cplxf* inv,
                                                               int64_t inv_stride, int* flags, int64_t flag_stride,
                                                               SplitDst sd) {
    __shared__ FsplitLds L;
    __shared__ float au[NB];
    const int b = blockIdx.z, part = blockIdx.y;
    cplxf* Am = mb.A + mb.stride * b;
    const int lda = mb.lda;
    const int nb = min(NB, n - K), KE = K + nb;
    const cplxf* iv = inv + inv_stride * b;
    const int r0 = KE + 64 * blockIdx.x;  // part 0: first row; part 1: first column
    const int rows = part == 0 ? min(64, n - r0) : nb;
    const int cols = part == 0 ? nb : min(64, n - r0);
    const cplxf* X = part == 0 ? Am + r0 + (int64_t)lda * K : iv;
    const int ldx = part == 0 ? lda : NB;
    const cplxf* Y = part == 0 ? iv + NB * NB : Am + K + (int64_t)lda * r0;
    const int ldy = part == 0 ? NB : lda;
    cplxf* Out = part == 0 ? Am + r0 + (int64_t)lda * K : Am + K + (int64_t)lda * r0;
    if (part == 0 && threadIdx.x < NB)
        au[threadIdx.x] = threadIdx.x < nb ? tabs1(Am[(K + threadIdx.x) + (int64_t)lda * (K + threadIdx.x)]) : 0.f;
    f32x16_t acc_re, acc_im;
    fsplit_tile64(X, ldx, Y, ldy, rows, cols, nb, L, acc_re, acc_im);  // its barriers order the au writes
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5, l32 = lane & 31;
    const int r = 32 * (wave & 1) + l32;
    bool bad = false;
    if (r < rows) {
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int c = 32 * (wave >> 1) + 8 * (v >> 2) + 4 * h + (v & 3);
            if (c < cols) {
                const cplxf val = make_float2(acc_re[v], acc_im[v]);
                if (part == 0) {
                    const cplxf ukk = Am[(K + c) + (int64_t)lda * (K + c)];
                    bad = bad || tabs1(tmul(val, ukk)) > au[c];
                }
                Out[r + (int64_t)lda * c] = val;
            }
        }
    }
    if (bad) flags[flag_stride * b] = 1;
    if (!sd.S || r >= rows) return;
    unsigned short* Sb = sd.S + sd.sS * b;
    if (part == 0) {
        // row r0 + r of L21: k = K - K0 + c, four consecutive k per accumulator group (8-byte stores)
        const int rr = r0 + r - sd.KS;
        if (rr < 0) return;
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const int c0 = 32 * (wave >> 1) + 8 * g + 4 * h;
            if (c0 >= cols) continue;
            const int k = K - sd.K0 + c0;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint2 w = q < 2 ? split_pack4(acc_re[4 * g], acc_re[4 * g + 1], acc_re[4 * g + 2], acc_re[4 * g + 3], q & 1)
                                      : split_pack4(acc_im[4 * g], acc_im[4 * g + 1], acc_im[4 * g + 2], acc_im[4 * g + 3], q & 1);
                *reinterpret_cast<uint2*>(Sb + (((int64_t)q * sd.nkc + k / 16) * sd.mpad + rr) * 16 + k % 16) = w;
            }
        }
    } else {
        // row r of U12 (k = K - K0 + r), columns r0 + c
        const int k = K - sd.K0 + r;
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int c = 32 * (wave >> 1) + 8 * (v >> 2) + 4 * h + (v & 3);
            const int cc = r0 + c - sd.KS;
            if (c >= cols || cc < 0) continue;
            const float re = acc_re[v], im = acc_im[v];
            const unsigned short rh = bf16_rne(re), ih = bf16_rne(im);
            const unsigned short part4[4] = {rh, bf16_rne(re - bf16_val(rh)), ih, bf16_rne(im - bf16_val(ih))};
#pragma unroll
            for (int q = 0; q < 4; q++)
                Sb[(((int64_t)(4 + q) * sd.nkc + k / 16) * sd.mpad + cc) * 16 + k % 16] = part4[q];
        }
    }
}

// The two left-looking products of panel j of a super-panel in one launch (they write disjoint
// blocks and read only the earlier panels): product 0, its columns A[Kj:n, Kj:Kq] -= A[Kj:n, K:Kj]
// A[K:Kj, Kj:Kq] (row tile blockIdx.x); product 1, its U rows A[Kj:Kq, Kq:n] -= A[Kj:Kq, K:Kj]
// A[K:Kj, Kq:n] (column tile blockIdx.x).  blockIdx.y selects the product.
__global__ void __launch_bounds__(256) narrow_pair_fsplit_kernel(MatBatch<cplxf> mb, int n, int K, int Kj) {
    __shared__ FsplitLds L;
    const int b = blockIdx.z, lda = mb.lda, Kq = Kj + NB, kk = Kj - K;
    cplxf* A = mb.A + mb.stride * b;
    const int prod = blockIdx.y;
    const int m = prod == 0 ? n - Kj : NB, nn = prod == 0 ? NB : n - Kq;
    const int row0 = Kj + (prod == 0 ? 64 * blockIdx.x : 0), col0 = prod == 0 ? Kj : Kq + 64 * blockIdx.x;
    const int rows = min(64, Kj + m - row0), cols = min(64, (prod == 0 ? Kq : n) - col0);
    if (rows <= 0 || cols <= 0) return;
    f32x16_t acc_re, acc_im;
    fsplit_tile64(A + row0 + (int64_t)lda * K, lda, A + K + (int64_t)lda * col0, lda, rows, cols, kk, L, acc_re, acc_im);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5, l32 = lane & 31;
    const int r = 32 * (wave & 1) + l32;
    if (r >= rows) return;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int c = 32 * (wave >> 1) + 8 * (v >> 2) + 4 * h + (v & 3);
        if (c < cols) {
            cplxf* p = A + (row0 + r) + (int64_t)lda * (col0 + c);
            const cplxf o = *p;
            *p = make_float2(o.x - acc_re[v], o.y - acc_im[v]);
        }
    }
}

// ---------------------------------------------------------------------------
// getrf driver.  ipiv: 0-based global row indices, int32, per matrix stride.
// work_perm: int workspace of (1 + 4*NB) per matrix.  info: int per matrix (0 / first zero pivot, 1-based).
// ---------------------------------------------------------------------------
// leaf of W columns over rows of ROWS x 1024 (the column count shrinks as the rows grow, to keep
// the register tile at 16 (FP64) or 32 (FP32) elements per thread)
template <int W, int ROWS, class C>
static void leaf_t(hipStream_t st, MatBatch<C> mb, int n, int c0, int K, int KE, int* ipiv, int64_t ipiv_stride,
                   int* info, int batch) {
    getf2_leaf_kernel<W, ROWS, C><<<dim3(1, 1, batch), LEAF_THREADS, 0, st>>>(mb, n, c0, K, KE, ipiv, ipiv_stride, info);
    if (c0 + W < KE && c0 + W < n)
        panel_update_kernel<W, C><<<dim3((n - c0 - W + 255) / 256, 1, batch), 256, 0, st>>>(mb, n, c0, KE);
}

// One leaf at c0 of the panel [K, KE).  The leaf width is fixed for the whole panel by the rows
// below K (every width divides the panel's 64 columns, so no leaf crosses its end); the rows per
// thread of each leaf follow the rows below c0.
template <class C>
static int leaf_dispatch(hipStream_t st, MatBatch<C> mb, int n, int c0, int K, int KE, int* ipiv, int64_t ipiv_stride,
                         int* info, int batch) {
    constexpr bool kF32 = std::is_same<C, cplxf>::value;
    constexpr int W = kF32 ? 8 : 4;  // rows per thread <= 4
    const int rp_panel = (n - K + LEAF_THREADS - 1) / LEAF_THREADS;
    const int rp = (n - c0 + LEAF_THREADS - 1) / LEAF_THREADS;
    if (rp_panel > 8) { leaf_t<W / 4, 16, C>(st, mb, n, c0, K, KE, ipiv, ipiv_stride, info, batch); return W / 4; }
    if (rp_panel > 4) { leaf_t<W / 2, 8, C>(st, mb, n, c0, K, KE, ipiv, ipiv_stride, info, batch); return W / 2; }
    if (rp <= 1) { leaf_t<W, 1, C>(st, mb, n, c0, K, KE, ipiv, ipiv_stride, info, batch); return W; }
    if (rp <= 2) { leaf_t<W, 2, C>(st, mb, n, c0, K, KE, ipiv, ipiv_stride, info, batch); return W; }
    leaf_t<W, 4, C>(st, mb, n, c0, K, KE, ipiv, ipiv_stride, info, batch);
    return W;
}

template <class C>
static int getrf_t(hipStream_t st, int n, C* A, int lda, int64_t strideA, int batch, int* ipiv, int64_t ipiv_stride,
                   int* work_perm, int* info, LaunchTimer* timer, unsigned short* split, int pivot, C* inv,
                   int64_t inv_stride, cplx* inv64, int64_t inv64_stride, int super_panels) {
    if (n <= 0) return 0;
    if (n > kMaxOrder) return -1;
    if (pivot == kPivotVerified && !inv) return -2;
    MatBatch<C> mb{A, strideA, lda};
    const int64_t pstride = kPermWork;
    constexpr bool kF32 = std::is_same<C, cplxf>::value;
    const bool use_split = kF32 && split != nullptr;
    const bool verified = pivot == kPivotVerified;
    // pivoted: leaves over [K, KE), each followed by the rank-W update of the panel's rest
    auto factor_panel = [&](int K, int KE) {
        for (int c0 = K; c0 < KE;) c0 += leaf_dispatch<C>(st, mb, n, c0, K, KE, ipiv, ipiv_stride, info, batch);
    };
    // pivoted: the panel's row interchanges applied to every column outside it
    auto permute_others = [&](int K, int KE) {
        if (n - (KE - K) > 0) {
            build_perm_kernel<<<dim3(1, 1, batch), 64, 0, st>>>(ipiv, ipiv_stride, K, KE, work_perm, pstride);
            const int nout = n - (KE - K);
            apply_perm_kernel<C><<<dim3((nout + 1) / 2, 1, batch), 256, 0, st>>>(mb, n, K, KE, work_perm, pstride);
        }
    };
    // pivoted: U12 = L11^{-1} A12 for rows [K, KE = K + NB) and columns [KE, n)
    auto trsm = [&](int K, int KE) {
        if constexpr (kF32)
            trsm_cols_kernel<<<dim3((n - KE + 255) / 256, 1, batch), 256, 0, st>>>(mb, n, K);
        else
            trsm_panel_kernel<C><<<dim3((n - KE + TRSM_COLS - 1) / TRSM_COLS, 1, batch), 256, 0, st>>>(mb, n, K, KE);
    };
    // verified diagonal pivots: the panel's diagonal block (and its triangular inverses), then its
    // L21 rows and U12 columns as products with those inverses
    auto factor_verified = [&](int K, SplitDst sd) {
        diag_lu_kernel<C><<<dim3(1, 1, batch), 512, 0, st>>>(mb, n, K, ipiv, ipiv_stride, work_perm, pstride, inv,
                                                             inv_stride, inv64, inv64_stride,
                                                             (int64_t)((n + NB - 1) / NB) * NB * NB);
        const int m = n - min(n, K + NB);
        if (m <= 0) return;
        if constexpr (kF32)
            if (use_split) {
                panel_inv_fsplit_kernel<<<dim3((m + 63) / 64, 2, batch), 256, 0, st>>>(mb, n, K, inv, inv_stride,
                                                                                     work_perm, pstride, sd);
                return;
            }
        panel_inv_gemm_kernel<C><<<dim3((m + 63) / 64, 2, batch), 256, 0, st>>>(mb, n, K, inv, inv_stride, work_perm,
                                                                                 pstride);
    };
    // the left-looking products inside a super-panel: A[r0:r0+m, c0:c0+nn] -= A[r0:, K:K+kk] A[K:K+kk, c0:]
    // (split-bf16 products on the split path, like its trailing updates)
    auto narrow = [&](int r0, int c0, int m, int nn, int K, int kk) {
        const C* Ap = A + r0 + (int64_t)lda * K;
        const C* Bp = A + K + (int64_t)lda * c0;
        C* Cp = A + r0 + (int64_t)lda * c0;
        if constexpr (kF32)
            if (use_split) {
                gemm_fsplit(st, true, true, m, nn, kk, Ap, lda, strideA, Bp, lda, strideA, Cp, lda, strideA, batch);
                return;
            }
        gemm_t<C>(st, true, true, m, nn, kk, Ap, lda, strideA, Bp, lda, strideA, Cp, lda, strideA, batch);
    };
    // A[KE:, KE:] -= A[KE:, K:KE] A[K:KE, KE:] (the timed launch: the GEMM alone)
    // the split arrays of the update of A[KS:, KS:] by the panels from K0 on (written by the
    // verified path's panel products; none otherwise)
    auto split_dst = [&](int K0, int KS) {
        SplitDst sd{nullptr, 0, KS, K0, 0, 0};
        if (use_split && verified && KS < n) {
            sd.S = split;
            sd.sS = split_work(n);
            sd.nkc = (KS - K0) / SPLIT_KC;
            sd.mpad = split_pad(n - KS);
        }
        return sd;
    };
    auto update = [&](int K, int KE) {
        const int m = n - KE, mpad = split_pad(m), ks = (KE - K + NB - 1) / NB * NB;
        const int64_t sS = split_work(n);
        if constexpr (kF32)
            if (use_split && !verified)
                split_panels_kernel<<<dim3((mpad * (ks / SPLIT_KC) + 255) / 256, 1, batch), 256, 0, st>>>(mb, n, K, KE, mpad,
                                                                                                  ks, split, sS);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timer) {
            hipEventCreate(&e0);
            hipEventCreate(&e1);
            hipEventRecord(e0, st);
        }
        if constexpr (kF32)
            if (use_split)
                zgemm_split_kernel<<<dim3(mpad / SPLIT_TILE, mpad / SPLIT_TILE, batch), 256, 0, st>>>(
                    m, mpad, ks, split, sS, A + KE + (int64_t)lda * KE, lda, strideA);
        if (!use_split)
            gemm_t<C>(st, true, true, m, m, KE - K, A + KE + (int64_t)lda * K, lda, strideA,
                      A + K + (int64_t)lda * KE, lda, strideA, A + KE + (int64_t)lda * KE, lda, strideA, batch);
        if (timer) {
            hipEventRecord(e1, st);
            timer->ev.emplace_back(e0, e1);
        }
    };
    for (int K = 0; K < n;) {
        // panels of this step: S full panels with a trailing matrix left after them (split path)
        int S = 1;
        const int smax = super_panels < 1 ? 1 : (super_panels > SPLIT_SUPER ? SPLIT_SUPER : super_panels);
        if (use_split)
            while (S < smax && K + (S + 1) * NB < n) S++;
        if (S == 1) {  // one panel
            const int KE = min(n, K + NB);
            if (verified) {
                factor_verified(K, split_dst(K, KE));
            } else {
                factor_panel(K, KE);
                permute_others(K, KE);
                if (KE < n) trsm(K, KE);
            }
            if (KE < n) update(K, KE);
            K = KE;
            continue;
        }
        // Multi-level step (split updates): panels P_j = [K + 64 j, K + 64 (j + 1)), j < S.  The
        // rest of the matrix receives their S rank-64 updates in one k = 64 S pass, read and
        // written once.  Panel j, left-looking over the panels before it:
        //   its columns (rows >= Kj) -= L21(P_<j) U12(P_<j)[:, P_j];  factor it; its interchanges
        //   (also in the earlier panels' columns);  its U rows right of it -= L21(P_<j)[P_j rows]
        //   U12(P_<j)[:, right of P_j];  U12(P_j) = L_jj^{-1} (those rows).
        // Then A[KS:, KS:] -= A[KS:, K:KS] A[K:KS, KS:].  The same operations as S one-panel
        // steps; only the trailing matrix's FP32 rounding differs (one accumulation instead of S).
        // With verified diagonal pivots there are no interchanges, so both left-looking products
        // of panel j come before its factorization.
        const int KS = K + S * NB;
        for (int j = 0; j < S; j++) {
            const int Kj = K + j * NB, Kq = Kj + NB;
            if (verified) {
                if (j > 0) {
                    if constexpr (kF32) {
                        if (use_split) {
                            const int tiles = max((n - Kj + 63) / 64, (n - Kq + 63) / 64);
                            narrow_pair_fsplit_kernel<<<dim3(tiles, 2, batch), 256, 0, st>>>(mb, n, K, Kj);
                        } else {
                            narrow(Kj, Kj, n - Kj, NB, K, Kj - K);
                            narrow(Kj, Kq, NB, n - Kq, K, Kj - K);
                        }
                    } else {
                        narrow(Kj, Kj, n - Kj, NB, K, Kj - K);
                        narrow(Kj, Kq, NB, n - Kq, K, Kj - K);
                    }
                }
                factor_verified(Kj, split_dst(K, KS));
                continue;
            }
            if (j > 0) narrow(Kj, Kj, n - Kj, NB, K, Kj - K);
            factor_panel(Kj, Kq);
            permute_others(Kj, Kq);
            if (j > 0) narrow(Kj, Kq, NB, n - Kq, K, Kj - K);
            trsm(Kj, Kq);
        }
        update(K, KS);
        K = KS;
    }
    return 0;
}

int getrf_batched(hipStream_t st, int n, cplx* A, int lda, int64_t strideA, int batch, int* ipiv, int64_t ipiv_stride,
                  int* work_perm, int* info, LaunchTimer* update_timer, int pivot, cplx* inv, int64_t inv_stride,
                  cplx* inv64, int64_t inv64_stride) {
    return getrf_t(st, n, A, lda, strideA, batch, ipiv, ipiv_stride, work_perm, info, update_timer, nullptr, pivot, inv,
                   inv_stride, inv64, inv64_stride, 1);
}

int getrf_batched(hipStream_t st, int n, cplxf* A, int lda, int64_t strideA, int batch, int* ipiv, int64_t ipiv_stride,
                  int* work_perm, int* info, LaunchTimer* update_timer, unsigned short* split, int pivot, cplxf* inv,
                  int64_t inv_stride, cplx* inv64, int64_t inv64_stride, int super_panels) {
    return getrf_t(st, n, A, lda, strideA, batch, ipiv, ipiv_stride, work_perm, info, update_timer, split, pivot, inv,
                   inv_stride, inv64, inv64_stride, super_panels);
}

// ---------------------------------------------------------------------------
// getrs: X = A^{-1} B (trans=0) or A^{-H} B (trans=2) for R <= MAXR right-hand sides.
// ---------------------------------------------------------------------------
#define TRS_B 64
#define MAXR 16

// idx[i] = original row of the element at position i after the interchanges
__global__ void ipiv_to_perm_kernel(const int* __restrict__ ipiv, int64_t ipiv_stride, int n, int* __restrict__ perm,
                                    int64_t perm_stride) {
    const int b = blockIdx.z;
    const int* piv = ipiv + ipiv_stride * b;
    int* idx = perm + perm_stride * b;
    // the interchanges are applied in order by one thread; the pivots reach it through LDS in
    // chunks (read one by one from global memory, a batch of order-4096 matrices took 265 us)
    constexpr int CH = 2048;
    __shared__ int ps[CH];
    for (int i = threadIdx.x; i < n; i += blockDim.x) idx[i] = i;
    for (int k0 = 0; k0 < n; k0 += CH) {
        const int cnt = min(CH, n - k0);
        __syncthreads();  // the previous chunk's interchanges (and the identity) are done
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) ps[i] = piv[k0 + i];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int k = 0; k < cnt; k++) {
                const int p = ps[k], kk = k0 + k;
                if (p != kk) { int t = idx[kk]; idx[kk] = idx[p]; idx[p] = t; }
            }
    }
}

// gather (inverse=0): Y[i] = X[idx[i]];  scatter (inverse=1): Y[idx[i]] = X[i]
__global__ void permute_rows_kernel(int n, int R, const int* __restrict__ perm, int64_t perm_stride,
                                    const cplx* __restrict__ X, int64_t sX, cplx* __restrict__ Y, int64_t sY, int inverse) {
    const int b = blockIdx.z;
    const int* idx = perm + perm_stride * b;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < R; c++) {
        if (!inverse)
            Y[sY * b + i + (int64_t)n * c] = X[sX * b + idx[i] + (int64_t)n * c];
        else
            Y[sY * b + idx[i] + (int64_t)n * c] = X[sX * b + i + (int64_t)n * c];
    }
}

// Inverses of the TRS_B x TRS_B diagonal blocks of L (unit lower) and U (upper),
// in place in LDS (LAPACK trti2 order), one workgroup per (block, triangle, matrix).
// Output: Inv[(blk * TRS_B + i) + TRS_B * nblk_rows...] stored as nblk consecutive
// TRS_B x TRS_B column-major blocks per triangle and matrix.
template <class C>
__global__ void __launch_bounds__(TRS_B) invert_diag_kernel(const C* __restrict__ Ab, int lda, int64_t sA, int n,
                                                            cplx* __restrict__ Linv, cplx* __restrict__ Uinv,
                                                            int64_t sInv) {
    const C* A = Ab + sA * blockIdx.z;
    const int blk = blockIdx.x, upper = blockIdx.y;
    const int b0 = blk * TRS_B, nbk = min(TRS_B, n - b0);
    cplx* out = (upper ? Uinv : Linv) + sInv * blockIdx.z + (int64_t)blk * TRS_B * TRS_B;
    __shared__ cplx T[TRS_B][TRS_B];  // T[i][j]
    const int t = threadIdx.x;
    for (int j = 0; j < TRS_B; j++) {
        cplx v = cmk(0.0, 0.0);
        if (t < nbk && j < nbk) v = to_d(A[(b0 + t) + (int64_t)lda * (b0 + j)]);
        T[t][j] = v;
    }
    __syncthreads();
    if (!upper) {
        // unit lower: columns from the right; Inv[i][j] = -sum_{s=j+1..i} Inv[i][s] L[s][j]
        for (int j = nbk - 2; j >= 0; j--) {
            cplx acc = cmk(0.0, 0.0);
            if (t > j && t < nbk) {
                acc = T[t][j];  // s = i term with Inv[i][i] = 1
                for (int s2 = j + 1; s2 < t; s2++) acc = cadd(acc, cmul(T[t][s2], T[s2][j]));
            }
            __syncthreads();
            if (t > j && t < nbk) T[t][j] = cmk(-acc.x, -acc.y);
            __syncthreads();
        }
        if (t < nbk) T[t][t] = cmk(1.0, 0.0);
    } else {
        // upper: columns from the left; Inv[j][j] = 1/U[j][j], Inv[i][j] = -Inv[j][j] sum_{s=i..j-1} Inv[i][s] U[s][j]
        for (int j = 0; j < nbk; j++) {
            const cplx djj = cdiv(cmk(1.0, 0.0), T[j][j]);
            cplx acc = cmk(0.0, 0.0);
            if (t < j)
                for (int s2 = t; s2 < j; s2++) acc = cadd(acc, cmul(T[t][s2], T[s2][j]));
            __syncthreads();
            if (t < j) T[t][j] = cmk(-(djj.x * acc.x - djj.y * acc.y), -(djj.x * acc.y + djj.y * acc.x));
            if (t == j) T[j][j] = djj;
            __syncthreads();
        }
    }
    __syncthreads();
    // the tile holds the packed LU: keep only the inverted triangle
    for (int j = 0; j < TRS_B; j++) {
        const bool keep = t < nbk && j < nbk && (upper ? j >= t : j <= t);
        out[t + TRS_B * j] = keep ? T[t][j] : cmk(0.0, 0.0);
    }
}

// ---------------------------------------------------------------------------
// Blocked triangular solve, one launch per TRS_B-row block step (trs_step_kernel below):
//   forward (lower=1): L (unit lower), or U^H when conj;  backward (lower=0): U, or L^H when conj.
// Block blk's x is op(Inv_blk) times its y after the updates by every block solved before it.
// ---------------------------------------------------------------------------

// element (a, q) of the dependency tile (a: index contiguous in memory), e = 0..15 per thread;
// FULL: the tile lies inside the matrix (no bounds checks)
template <class C, bool CONJ, bool FULL>
__device__ __forceinline__ void trs_load_tile(const C* __restrict__ A, int lda, int b0, int nbk, int c0, int ncb,
                                              int tid, C (&reg)[TRS_B * TRS_B / 256]) {
#pragma unroll
    for (int e = 0; e < TRS_B * TRS_B / 256; e++) {
        const int idx = tid + 256 * e;
        const int a = idx & (TRS_B - 1), q = idx / TRS_B;
        if (!CONJ) {
            reg[e] = (FULL || (a < nbk && q < ncb)) ? A[(b0 + a) + (int64_t)lda * (c0 + q)] : tmk<C>(0, 0);
        } else {
            reg[e] = (FULL || (q < nbk && a < ncb)) ? A[(c0 + a) + (int64_t)lda * (b0 + q)] : tmk<C>(0, 0);
        }
    }
}

// One block step of the blocked triangular solve X = op(T)^{-1} Y, all items of the batch
// (blockIdx.z), in sweep order (lower: block rows upwards from 0, upper: downwards from the last).
//   first = 1: x_blk = op(Inv_blk) y_blk (one workgroup per item; the first block of the sweep);
//   first = 0: workgroup j updates block row rb = blk +/- (j + 1):  y_rb -= op(T)[rb, blk] x_blk,
//              and the workgroup of the next block of the sweep (j = 0), whose y is then final,
//              forms x_rb = op(Inv_rb) y_rb.
// The launches' order carries the dependencies (x_blk and y_rb were written by earlier launches),
// so there are no waits inside a launch.  Each element of y is reduced in the order of the block
// steps, tile by tile, in FP64.  CONJ: op(T) = T^H (else T); FULL: every block is TRS_B rows (the
// order is a multiple of TRS_B), so no element needs a bounds check.
template <int RB, class C, bool CONJ, bool FULL>
__global__ void __launch_bounds__(256) trs_step_kernel(const C* __restrict__ Ab, int lda, int64_t sA, int n, int R,
                                                       int lower, int blk, int first,
                                                       const cplx* __restrict__ Invb, int64_t sInv,
                                                       cplx* __restrict__ Yb, int64_t sY, cplx* __restrict__ Xb,
                                                       int64_t sX) {
    constexpr int CPT = RB / 4;  // right-hand-side columns per thread
    constexpr int TE = TRS_B * TRS_B / 256;
    // the factor tile (tileT[s][i] = op(T)[r0 + i][b0 + s]) and, once the update is done, the
    // diagonal-block inverse (invs[s][i] = op(Inv)[i][s]) share one LDS buffer
    constexpr size_t kTile = sizeof(C) * TRS_B * (TRS_B + 1), kInv = sizeof(cplx) * TRS_B * TRS_B;
    __shared__ __align__(16) unsigned char smem[kTile > kInv ? kTile : kInv];
    __shared__ cplx xs[TRS_B][RB];
    C(*tileT)[TRS_B + 1] = reinterpret_cast<C(*)[TRS_B + 1]>(smem);
    cplx(*invs)[TRS_B] = reinterpret_cast<cplx(*)[TRS_B]>(smem);
    const int tid = threadIdx.x, i = tid & (TRS_B - 1), rq = tid / TRS_B;
    const int item = blockIdx.z, j = blockIdx.x;
    const int b0 = blk * TRS_B, nbk = FULL ? TRS_B : min(TRS_B, n - b0);
    const C* A = Ab + sA * item;
    cplx* Y = Yb + sY * item;
    cplx* X = Xb + sX * item;
    const int rb = first ? blk : (lower ? blk + 1 + j : blk - 1 - j);
    const int r0 = rb * TRS_B, nrb = FULL ? TRS_B : min(TRS_B, n - r0);
    const bool diag = first || j == 0;  // this workgroup forms x of block rb
    // the inverse of block rb, loaded at once (its latency hides behind the update)
    cplx ireg[TE];
    if (diag) {
        const cplx* Iv = Invb + sInv * item + (int64_t)rb * TRS_B * TRS_B;
#pragma unroll
        for (int e = 0; e < TE; e++) ireg[e] = Iv[tid + 256 * e];
    }
    // x_rb = op(Inv_rb) xs (xs holds the final y of block rb); stored to X
    auto solve_diag = [&]() {
#pragma unroll
        for (int e = 0; e < TE; e++) {
            const int idx = tid + 256 * e;
            const int a = idx & (TRS_B - 1), q = idx / TRS_B;  // element (a, q) of Inv_rb
            const cplx v = ireg[e];
            if (!CONJ) invs[q][a] = v;
            else invs[a][q] = cmk(v.x, -v.y);
        }
        __syncthreads();
        if (FULL || i < nrb) {
            cplx x[CPT];
#pragma unroll
            for (int cc = 0; cc < CPT; cc++) x[cc] = cmk(0.0, 0.0);
#pragma unroll 4
            for (int s = 0; s < (FULL ? TRS_B : nrb); s++) {
                const cplx v = invs[s][i];
#pragma unroll
                for (int cc = 0; cc < CPT; cc++) cfma(x[cc], v, xs[s][rq * CPT + cc]);
            }
#pragma unroll
            for (int cc = 0; cc < CPT; cc++) {
                const int col = rq * CPT + cc;
                if (col < R) X[(r0 + i) + (int64_t)n * col] = x[cc];
            }
        }
    };
    if (first) {
#pragma unroll
        for (int cc = 0; cc < CPT; cc++) {
            const int col = rq * CPT + cc;
            xs[i][col] = ((FULL || i < nrb) && col < R) ? Y[(r0 + i) + (int64_t)n * col] : cmk(0.0, 0.0);
        }
        solve_diag();
        return;
    }
    C reg[TE];
    trs_load_tile<C, CONJ, FULL>(A, lda, r0, nrb, b0, nbk, tid, reg);
#pragma unroll
    for (int cc = 0; cc < CPT; cc++) {
        const int col = rq * CPT + cc;
        xs[i][col] = ((FULL || i < nbk) && col < R) ? X[(b0 + i) + (int64_t)n * col] : cmk(0.0, 0.0);
    }
    cplx acc[CPT];
#pragma unroll
    for (int cc = 0; cc < CPT; cc++) {
        const int col = rq * CPT + cc;
        acc[cc] = ((FULL || i < nrb) && col < R) ? Y[(r0 + i) + (int64_t)n * col] : cmk(0.0, 0.0);
    }
#pragma unroll
    for (int e = 0; e < TE; e++) {
        const int idx = tid + 256 * e;
        const int a = idx & (TRS_B - 1), q = idx / TRS_B;
        C w = reg[e];
        if (!CONJ) {
            tileT[q][a] = w;
        } else {
            w.y = -w.y;
            tileT[a][q] = w;
        }
    }
    __syncthreads();
#pragma unroll 8
    for (int s = 0; s < TRS_B; s++) {
        const cplx t = to_d(tileT[s][i]);
#pragma unroll
        for (int cc = 0; cc < CPT; cc++) cfms(acc[cc], t, xs[s][rq * CPT + cc]);
    }
    if (diag) {
        __syncthreads();  // every thread is done with tileT and with xs of block blk
#pragma unroll
        for (int cc = 0; cc < CPT; cc++) xs[i][rq * CPT + cc] = acc[cc];
        solve_diag();
    } else if (FULL || i < nrb) {
#pragma unroll
        for (int cc = 0; cc < CPT; cc++) {
            const int col = rq * CPT + cc;
            if (col < R) Y[(r0 + i) + (int64_t)n * col] = acc[cc];
        }
    }
}

template <int RB, class C, bool CONJ, bool FULL>
static void trs_pass_k(hipStream_t st, const C* A, int lda, int64_t sA, int n, int R, int lower,
                       const cplx* Inv, int64_t sInv, cplx* Y, int64_t sY, cplx* X, int64_t sX, int batch) {
    const int nblk = (n + TRS_B - 1) / TRS_B;
    const int blk0 = lower ? 0 : nblk - 1;
    trs_step_kernel<RB, C, CONJ, FULL><<<dim3(1, 1, batch), 256, 0, st>>>(A, lda, sA, n, R, lower, blk0, 1, Inv, sInv,
                                                                          Y, sY, X, sX);
    for (int step = 0; step + 1 < nblk; step++) {
        const int blk = lower ? step : nblk - 1 - step;
        trs_step_kernel<RB, C, CONJ, FULL><<<dim3(nblk - 1 - step, 1, batch), 256, 0, st>>>(
            A, lda, sA, n, R, lower, blk, 0, Inv, sInv, Y, sY, X, sX);
    }
}

template <int RB, class C>
static void trs_pass_t(hipStream_t st, const C* A, int lda, int64_t sA, int n, int R, int lower, int conj,
                       const cplx* Inv, int64_t sInv, cplx* Y, int64_t sY, cplx* X, int64_t sX, int batch) {
    const bool full = n % TRS_B == 0;
    if (conj) {
        if (full) trs_pass_k<RB, C, true, true>(st, A, lda, sA, n, R, lower, Inv, sInv, Y, sY, X, sX, batch);
        else trs_pass_k<RB, C, true, false>(st, A, lda, sA, n, R, lower, Inv, sInv, Y, sY, X, sX, batch);
    } else {
        if (full) trs_pass_k<RB, C, false, true>(st, A, lda, sA, n, R, lower, Inv, sInv, Y, sY, X, sX, batch);
        else trs_pass_k<RB, C, false, false>(st, A, lda, sA, n, R, lower, Inv, sInv, Y, sY, X, sX, batch);
    }
}

// X = op(T)^{-1} Y; Y is overwritten (block rows are reduced in place)
template <class C>
static void trs_pass(hipStream_t st, const C* A, int lda, int64_t sA, int n, int R, int lower, int conj,
                     const cplx* Inv, int64_t sInv, cplx* Y, int64_t sY, cplx* X, int64_t sX, int batch) {
    if (R <= 4) trs_pass_t<4, C>(st, A, lda, sA, n, R, lower, conj, Inv, sInv, Y, sY, X, sX, batch);
    else if (R <= 8) trs_pass_t<8, C>(st, A, lda, sA, n, R, lower, conj, Inv, sInv, Y, sY, X, sX, batch);
    else trs_pass_t<16, C>(st, A, lda, sA, n, R, lower, conj, Inv, sInv, Y, sY, X, sX, batch);
}

// B (n x R, leading dimension n) is overwritten with the solution.  work per matrix (stride sW):
// a reserved header of ceil(n/64) + 1 ints (rounded up to whole complex elements; first, so that
// the rest does not move with R); W1, W2 (n R complex each); the
// diagonal-block inverses of L and U (2 ceil(n/64) 64^2 complex).
static int64_t flag_words(int n) { return ((n + TRS_B - 1) / TRS_B + 1 + 3) / 4; }

int64_t getrs_work(int n, int R) {
    const int64_t nblk = (n + TRS_B - 1) / TRS_B;
    return flag_words(n) + 2 * (int64_t)n * R + 2 * nblk * TRS_B * TRS_B;
}

cplx* getrs_inverses(cplx* work, int n, int R) { return work + flag_words(n) + 2 * (int64_t)n * R; }

template <class C>
static int getrs_t(hipStream_t st, int trans, int n, int R, const C* LU, int lda, int64_t sA, const int* ipiv,
                   int64_t ipiv_stride, int* perm, int64_t perm_stride, cplx* B, int64_t sB, cplx* work, int64_t sW,
                   int batch, int prepared) {
    if (R > MAXR) return -1;
    if (n <= 0 || R <= 0) return 0;
    const int nblk = (n + TRS_B - 1) / TRS_B;
    cplx* W1 = work + flag_words(n);
    cplx* W2 = W1 + (int64_t)n * R;
    cplx* Linv = W2 + (int64_t)n * R;
    cplx* Uinv = Linv + (int64_t)nblk * TRS_B * TRS_B;
    if (prepared == kGetrsFresh)
        invert_diag_kernel<C><<<dim3(nblk, 2, batch), TRS_B, 0, st>>>(LU, lda, sA, n, Linv, Uinv, sW);
    if (prepared != kGetrsPrepared)
        ipiv_to_perm_kernel<<<dim3(1, 1, batch), 256, 0, st>>>(ipiv, ipiv_stride, n, perm, perm_stride);
    dim3 gp((n + 255) / 256, 1, batch);
    if (trans == 0) {
        permute_rows_kernel<<<gp, 256, 0, st>>>(n, R, perm, perm_stride, B, sB, W1, sW, 0);
        trs_pass(st, LU, lda, sA, n, R, 1, 0, Linv, sW, W1, sW, W2, sW, batch);  // L
        trs_pass(st, LU, lda, sA, n, R, 0, 0, Uinv, sW, W2, sW, B, sB, batch);   // U
    } else {
        for (int b = 0; b < batch; b++)
            hipMemcpyAsync(W1 + sW * b, B + sB * b, sizeof(cplx) * (size_t)n * R, hipMemcpyDeviceToDevice, st);
        trs_pass(st, LU, lda, sA, n, R, 1, 1, Uinv, sW, W1, sW, W2, sW, batch);  // U^H
        trs_pass(st, LU, lda, sA, n, R, 0, 1, Linv, sW, W2, sW, W1, sW, batch);  // L^H
        permute_rows_kernel<<<gp, 256, 0, st>>>(n, R, perm, perm_stride, W1, sW, B, sB, 1);
    }
    return 0;
}

int getrs_batched(hipStream_t st, int trans, int n, int R, const cplx* LU, int lda, int64_t sA, const int* ipiv,
                  int64_t ipiv_stride, int* perm, int64_t perm_stride, cplx* B, int64_t sB, cplx* work, int64_t sW,
                  int batch, int prepared) {
    return getrs_t(st, trans, n, R, LU, lda, sA, ipiv, ipiv_stride, perm, perm_stride, B, sB, work, sW, batch, prepared);
}


int getrs_batched(hipStream_t st, int trans, int n, int R, const cplxf* LU, int lda, int64_t sA, const int* ipiv,
                  int64_t ipiv_stride, int* perm, int64_t perm_stride, cplx* B, int64_t sB, cplx* work, int64_t sW,
                  int batch, int prepared) {
    return getrs_t(st, trans, n, R, LU, lda, sA, ipiv, ipiv_stride, perm, perm_stride, B, sB, work, sW, batch, prepared);
}



// ---------------------------------------------------------------------------
// mixed-precision helpers: FP64 -> FP32 copy of a batch of matrices, residual
// R = B - T with per-item squared norms, X += D
// ---------------------------------------------------------------------------
__global__ void to_f32_kernel(int64_t count, const cplx* __restrict__ src, cplxf* __restrict__ dst) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count; t += (int64_t)gridDim.x * blockDim.x)
        dst[t] = make_float2((float)src[t].x, (float)src[t].y);
}

void convert_to_f32(hipStream_t st, int64_t count, const cplx* src, cplxf* dst) {
    const int64_t blocks = std::min<int64_t>((count + 255) / 256, 8192);
    to_f32_kernel<<<(int)blocks, 256, 0, st>>>(count, src, dst);
}

// dst[z * sD + e] = src[z * sS + e] widened, e < count, item z = blockIdx.y
__global__ void widen_kernel(int64_t count, const cplxf* __restrict__ src, int64_t sS, cplx* __restrict__ dst,
                             int64_t sD) {
    const cplxf* a = src + sS * blockIdx.y;
    cplx* o = dst + sD * blockIdx.y;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        o[i] = to_d(a[i]);
}

void widen_to_f64(hipStream_t st, int64_t count, int batch, const cplxf* src, int64_t sS, cplx* dst, int64_t sD) {
    if (count <= 0 || batch <= 0) return;
    const int64_t blocks = std::min<int64_t>((count + 255) / 256, 2048);
    widen_kernel<<<dim3((unsigned)blocks, (unsigned)batch), 256, 0, st>>>(count, src, sS, dst, sD);
}

// per item z: T <- Bv - T over n*R elements; norms[2z] += |T|^2, norms[2z+1] += |Bv|^2
__global__ void residual_kernel(int64_t cnt, const cplx* __restrict__ Bvb, int64_t sB, cplx* __restrict__ Tb,
                                int64_t sT, double* __restrict__ norms) {
    const int z = blockIdx.z;
    const cplx* Bv = Bvb + sB * z;
    cplx* T = Tb + sT * z;
    double rr = 0.0, bb = 0.0;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < cnt; t += (int64_t)gridDim.x * blockDim.x) {
        const cplx b = Bv[t], y = T[t];
        const cplx r = cmk(b.x - y.x, b.y - y.y);
        T[t] = r;
        rr += r.x * r.x + r.y * r.y;
        bb += b.x * b.x + b.y * b.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
        rr += __shfl_xor(rr, off);
        bb += __shfl_xor(bb, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(norms + 2 * z, rr);
        atomicAdd(norms + 2 * z + 1, bb);
    }
}

void launch_residual(hipStream_t st, int64_t cnt, const cplx* Bv, int64_t sB, cplx* T, int64_t sT, double* norms,
                     int batch) {
    const int64_t blocks = std::min<int64_t>((cnt + 255) / 256, 64);
    residual_kernel<<<dim3((int)blocks, 1, batch), 256, 0, st>>>(cnt, Bv, sB, T, sT, norms);
}

__global__ void axpy_kernel(int64_t cnt, const cplx* __restrict__ Db, int64_t sD, cplx* __restrict__ Xb, int64_t sX) {
    const int z = blockIdx.z;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < cnt; t += (int64_t)gridDim.x * blockDim.x) {
        const cplx d = Db[sD * z + t];
        cplx& x = Xb[sX * z + t];
        x = cmk(x.x + d.x, x.y + d.y);
    }
}

void launch_add(hipStream_t st, int64_t cnt, const cplx* D, int64_t sD, cplx* X, int64_t sX, int batch) {
    const int64_t blocks = std::min<int64_t>((cnt + 255) / 256, 256);
    axpy_kernel<<<dim3((int)blocks, 1, batch), 256, 0, st>>>(cnt, D, sD, X, sX);
}

}  // namespace mhx

namespace mhx {

// device-to-device copy with a kernel (16-byte units), so that it can sit in a captured graph
__global__ void copy16_kernel(int64_t count, const uint4* __restrict__ src, uint4* __restrict__ dst) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count; t += (int64_t)gridDim.x * blockDim.x)
        dst[t] = src[t];
}

void launch_copy(hipStream_t st, const void* src, void* dst, size_t bytes) {
    const int64_t count = (int64_t)(bytes / 16);
    if (count <= 0) return;
    const int64_t blocks = std::min<int64_t>((count + 255) / 256, 8192);
    copy16_kernel<<<(int)blocks, 256, 0, st>>>(count, reinterpret_cast<const uint4*>(src), reinterpret_cast<uint4*>(dst));
}

// out[q] = the verified-pivot flag of item q (wperm[kPermWork q])
__global__ void gather_flags_kernel(const int* __restrict__ wperm, int batch, int* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < batch) out[t] = wperm[(int64_t)kPermWork * t];
}

void launch_gather_pivot_flags(hipStream_t st, const int* wperm, int batch, int* out) {
    if (batch > 0) gather_flags_kernel<<<(batch + 255) / 256, 256, 0, st>>>(wperm, batch, out);
}

}  // namespace mhx
