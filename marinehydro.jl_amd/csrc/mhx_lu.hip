// Dense complex LU with partial pivoting and triangular solves for gfx950.
//
// Replaces LinearSolve's LUFactorization -> LAPACK zgetrf/zgetrs
// (reference: src/matrix_assembly.jl:396-416, 436-449).  Pivot choice follows
// LAPACK izamax (largest |Re|+|Im|, first index on ties), so the factors agree
// with the CPU LU to rounding.
//
// Blocking (right-looking):
//   outer panels of NB=64 columns; each outer panel is factored in leaves of
//   LW=4 columns held in registers of one 1024-thread workgroup (the pivot
//   search is a block reduction), each leaf followed by a multi-workgroup
//   rank-4 update of the rest of the outer panel;
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

#define LW 4
#define NB 64
#define LEAF_THREADS 1024

template <class C>
struct MatBatch {
    C* A;           // base pointer of matrix 0
    int64_t stride; // elements between consecutive matrices
    int lda;
};

// ---------------------------------------------------------------------------
// leaf: factor columns [c0, c0+LW) over rows [c0, n); swap the pivot rows in
// the other columns of the outer panel [K, KE); solve the 4-row U block of the
// outer panel columns right of the leaf.
// ---------------------------------------------------------------------------
template <int ROWS, class C>
__global__ void __launch_bounds__(LEAF_THREADS) getf2_leaf_kernel(MatBatch<C> mb, int n, int c0, int K, int KE,
                                                                  int* __restrict__ ipiv, int64_t ipiv_stride,
                                                                  int* __restrict__ info) {
    typedef typename Cx<C>::R R;
    const int b = blockIdx.z;
    C* A = mb.A + mb.stride * b;
    int* piv = ipiv + ipiv_stride * b;
    const int lda = mb.lda;
    const int tid = threadIdx.x;
    __shared__ R red_v[LEAF_THREADS / 64];
    __shared__ int red_r[LEAF_THREADS / 64];
    __shared__ C xrow[2][LW];
    __shared__ C prow[LW];
    __shared__ int s_piv;
    __shared__ int s_pivs[LW];
    __shared__ C L11[LW][LW];

    const int wcols = min(LW, n - c0);
    C x[ROWS][LW];
#pragma unroll
    for (int q = 0; q < ROWS; q++) {
        int r = c0 + tid + LEAF_THREADS * q;
#pragma unroll
        for (int c = 0; c < LW; c++) x[q][c] = (r < n && c < wcols) ? A[r + (int64_t)lda * (c0 + c)] : tmk<C>(0, 0);
    }
#pragma unroll
    for (int c = 0; c < LW; c++) {
        if (c >= wcols) break;
        const int col = c0 + c;
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
        if ((tid & 63) == 0) { red_v[tid >> 6] = bv; red_r[tid >> 6] = br; }
        __syncthreads();
        if (tid == 0) {
            R v = red_v[0];
            int r = red_r[0];
            for (int w = 1; w < LEAF_THREADS / 64; w++)
                if (red_v[w] > v || (red_v[w] == v && red_r[w] < r)) { v = red_v[w]; r = red_r[w]; }
            s_piv = r;
            piv[col] = r;
            s_pivs[c] = r;
        }
        __syncthreads();
        const int p = s_piv;
        // --- swap rows col <-> p inside the leaf
        const int tc = (col - c0) % LEAF_THREADS, qc = (col - c0) / LEAF_THREADS;
        const int tp = (p - c0) % LEAF_THREADS, qp = (p - c0) / LEAF_THREADS;
        if (p != col) {
#pragma unroll
            for (int q = 0; q < ROWS; q++) {
                if (tid == tc && q == qc)
                    for (int cc = 0; cc < LW; cc++) xrow[0][cc] = x[q][cc];
                if (tid == tp && q == qp)
                    for (int cc = 0; cc < LW; cc++) xrow[1][cc] = x[q][cc];
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < ROWS; q++) {
                if (tid == tc && q == qc)
                    for (int cc = 0; cc < LW; cc++) x[q][cc] = xrow[1][cc];
                if (tid == tp && q == qp)
                    for (int cc = 0; cc < LW; cc++) x[q][cc] = xrow[0][cc];
            }
        }
        // --- broadcast the pivot row
#pragma unroll
        for (int q = 0; q < ROWS; q++)
            if (tid == tc && q == qc)
                for (int cc = 0; cc < LW; cc++) prow[cc] = x[q][cc];
        __syncthreads();
        const C u = prow[c];
        if (u.x != 0 || u.y != 0) {
            const C inv = trecip(u);
#pragma unroll
            for (int q = 0; q < ROWS; q++) {
                int r = c0 + tid + LEAF_THREADS * q;
                if (r > col && r < n) {
                    x[q][c] = tmul(x[q][c], inv);
#pragma unroll
                    for (int cc = 0; cc < LW; cc++)
                        if (cc > c) x[q][cc] = tsub(x[q][cc], tmul(x[q][c], prow[cc]));
                }
            }
        } else if (tid == 0) {
            atomicCAS(info + b, 0, col + 1);
        }
        __syncthreads();
    }
    // store back
#pragma unroll
    for (int q = 0; q < ROWS; q++) {
        int r = c0 + tid + LEAF_THREADS * q;
        if (r < n)
#pragma unroll
            for (int c = 0; c < LW; c++)
                if (c < wcols) A[r + (int64_t)lda * (c0 + c)] = x[q][c];
    }
    // L11 (unit lower 4x4 of this leaf) to LDS
    if (tid < wcols)
        for (int c = 0; c < LW; c++) L11[tid][c] = x[0][c];
    __syncthreads();
    // other outer-panel columns: apply the leaf's interchanges, then solve U rows
    const int nother = (KE - K) - wcols;
    for (int t = tid; t < nother; t += LEAF_THREADS) {
        int j = K + t;
        if (j >= c0) j += wcols;
        C* colj = A + (int64_t)lda * j;
        for (int c = 0; c < wcols; c++) {
            int r0 = c0 + c, r1 = s_pivs[c];
            if (r1 != r0) {
                C t0 = colj[r0];
                colj[r0] = colj[r1];
                colj[r1] = t0;
            }
        }
        if (j > c0) {  // U12 rows c0..c0+wcols-1: unit-lower forward substitution
            C u[LW];
            for (int c = 0; c < wcols; c++) {
                C s = colj[c0 + c];
                for (int s2 = 0; s2 < c; s2++) s = tsub(s, tmul(L11[c][s2], u[s2]));
                u[c] = s;
                colj[c0 + c] = s;
            }
        }
    }
}

// rank-LW update of the outer panel: A[r, j] -= sum_s L[r, c0+s] U[c0+s, j],
// rows r in [c0+LW, n), columns j in [c0+LW, KE)
template <class C>
__global__ void __launch_bounds__(256) panel_update_kernel(MatBatch<C> mb, int n, int c0, int KE) {
    const int b = blockIdx.z;
    C* A = mb.A + mb.stride * b;
    const int lda = mb.lda;
    __shared__ C U[LW][NB];
    const int j0 = c0 + LW, nj = KE - j0;
    for (int t = threadIdx.x; t < LW * nj; t += blockDim.x) {
        int s = t / nj, jj = t % nj;
        U[s][jj] = A[(c0 + s) + (int64_t)lda * (j0 + jj)];
    }
    __syncthreads();
    const int r = c0 + LW + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    C l[LW];
#pragma unroll
    for (int s = 0; s < LW; s++) l[s] = A[r + (int64_t)lda * (c0 + s)];
    for (int jj = 0; jj < nj; jj++) {
        C* p = A + r + (int64_t)lda * (j0 + jj);
        C a = *p;
#pragma unroll
        for (int s = 0; s < LW; s++) a = tsub(a, tmul(l[s], U[s][jj]));
        *p = a;
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

// ---------------------------------------------------------------------------
// C[m x nn] (+)= (-)A[m x kk] * B[kk x nn], complex, column-major, on MFMA
// (v_mfma_f64_16x16x4f64 for cplx, v_mfma_f32_16x16x4f32 for cplxf).
// 256 threads (4 waves in 2x2), 64x64 tile, K chunks of 16 staged in LDS as
// separate real / imaginary planes.  Lane layout of both instructions:
//   A: lane l holds A[l%16][l/16], B: B[l/16][l%16];
//   C: 4 values per lane, f64: C[4v + l/16][l%16], f32: C[4(l/16) + v][l%16].
// ---------------------------------------------------------------------------
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
// getrf driver.  ipiv: 0-based global row indices, int32, per matrix stride.
// work_perm: int workspace of (1 + 4*NB) per matrix.  info: int per matrix (0 / first zero pivot, 1-based).
// ---------------------------------------------------------------------------
template <class C>
static int getrf_t(hipStream_t st, int n, C* A, int lda, int64_t strideA, int batch, int* ipiv, int64_t ipiv_stride,
                   int* work_perm, int* info, LaunchTimer* timer) {
    if (n <= 0) return 0;
    const int rows_per = (n + LEAF_THREADS - 1) / LEAF_THREADS;
    if (rows_per > 4) return -1;  // orders above 4096 need a larger register tile
    MatBatch<C> mb{A, strideA, lda};
    const int64_t pstride = 1 + 4 * NB;
    for (int K = 0; K < n; K += NB) {
        const int KE = min(n, K + NB);
        for (int c0 = K; c0 < KE; c0 += LW) {
            const int m_rows = n - c0;
            const int rp = (m_rows + LEAF_THREADS - 1) / LEAF_THREADS;
            dim3 g1(1, 1, batch);
            if (rp <= 1)
                getf2_leaf_kernel<1, C><<<g1, LEAF_THREADS, 0, st>>>(mb, n, c0, K, KE, ipiv, ipiv_stride, info);
            else if (rp <= 2)
                getf2_leaf_kernel<2, C><<<g1, LEAF_THREADS, 0, st>>>(mb, n, c0, K, KE, ipiv, ipiv_stride, info);
            else
                getf2_leaf_kernel<4, C><<<g1, LEAF_THREADS, 0, st>>>(mb, n, c0, K, KE, ipiv, ipiv_stride, info);
            if (c0 + LW < KE && c0 + LW < n) {
                dim3 g2((n - c0 - LW + 255) / 256, 1, batch);
                panel_update_kernel<C><<<g2, 256, 0, st>>>(mb, n, c0, KE);
            }
        }
        if (n - (KE - K) > 0) {
            build_perm_kernel<<<dim3(1, 1, batch), 64, 0, st>>>(ipiv, ipiv_stride, K, KE, work_perm, pstride);
            const int nout = n - (KE - K);
            apply_perm_kernel<C><<<dim3((nout + 1) / 2, 1, batch), 256, 0, st>>>(mb, n, K, KE, work_perm, pstride);
        }
        if (KE < n) {
            trsm_panel_kernel<C><<<dim3((n - KE + TRSM_COLS - 1) / TRSM_COLS, 1, batch), 256, 0, st>>>(mb, n, K, KE);
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (timer) {
                hipEventCreate(&e0);
                hipEventCreate(&e1);
                hipEventRecord(e0, st);
            }
            gemm_t<C>(st, true, true, n - KE, n - KE, KE - K, A + KE + (int64_t)lda * K, lda, strideA,
                      A + K + (int64_t)lda * KE, lda, strideA, A + KE + (int64_t)lda * KE, lda, strideA, batch);
            if (timer) {
                hipEventRecord(e1, st);
                timer->ev.emplace_back(e0, e1);
            }
        }
    }
    return 0;
}

int getrf_batched(hipStream_t st, int n, cplx* A, int lda, int64_t strideA, int batch, int* ipiv, int64_t ipiv_stride,
                  int* work_perm, int* info, LaunchTimer* update_timer) {
    return getrf_t(st, n, A, lda, strideA, batch, ipiv, ipiv_stride, work_perm, info, update_timer);
}

int getrf_batched(hipStream_t st, int n, cplxf* A, int lda, int64_t strideA, int batch, int* ipiv, int64_t ipiv_stride,
                  int* work_perm, int* info, LaunchTimer* update_timer) {
    return getrf_t(st, n, A, lda, strideA, batch, ipiv, ipiv_stride, work_perm, info, update_timer);
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
    for (int i = threadIdx.x; i < n; i += blockDim.x) idx[i] = i;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < n; k++) {
            int p = piv[k];
            if (p != k) { int t = idx[k]; idx[k] = idx[p]; idx[p] = t; }
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
// Sync-free blocked triangular solve: one persistent launch per pass.
//   forward (lower=1): L (unit lower), or U^H when conj;  backward (lower=0): U, or L^H when conj.
// Task (step, item) solves the TRS_B-row block blk = step (forward) or nblk-1-step (backward):
//   acc = y_blk - sum over the blocks solved before it of op(T)[blk, c] x_c,  x_blk = op(Inv_blk) acc.
// Tasks are dealt to workgroups in increasing step order and every workgroup runs its tasks in
// that order; a task waits on the per-(item, block) flag of each dependency, which its producer
// sets to this pass's epoch after its x stores have completed.  Flags and x are accessed only
// with device-coherent (agent-scope) atomic loads and stores, so no cache-wide fence is needed:
// the consumer issues its x loads after it has seen the flag, the producer raises the flag after
// s_waitcnt vmcnt(0) on its x stores.  The smallest unfinished task therefore never waits on
// anything unfinished, so the pass completes whenever the grid fits on the device at once (it is
// sized to half of that).  A wait that exceeds TRS_SPIN_LIMIT polls (tens of milliseconds, far
// beyond a pass) sets *err and gives up, and every later wait of the launch then gives up at
// once (the result is invalid and the caller reports an error): no launch can hang.
// ---------------------------------------------------------------------------
#define TRS_SPIN_LIMIT (1 << 17)
// s_waitcnt vmcnt(0) (gfx9-family encoding: expcnt and lgkmcnt left at their maxima)
constexpr int kWaitVmcnt0 = 0x0F70;

__device__ __forceinline__ cplx ld_agent(const cplx* p) {
    const double* q = reinterpret_cast<const double*>(p);
    return cmk(__hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
               __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ void st_agent(cplx* p, cplx v) {
    double* q = reinterpret_cast<double*>(p);
    __hip_atomic_store(q, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// element (a, q) of the dependency tile (a: index contiguous in memory), e = 0..15 per thread
template <class C>
__device__ __forceinline__ void trs_load_tile(const C* __restrict__ A, int lda, int conj, int b0, int nbk, int c0,
                                              int ncb, int tid, C (&reg)[TRS_B * TRS_B / 256]) {
#pragma unroll
    for (int e = 0; e < TRS_B * TRS_B / 256; e++) {
        const int idx = tid + 256 * e;
        const int a = idx & (TRS_B - 1), q = idx / TRS_B;
        C w = tmk<C>(0, 0);
        if (!conj) {
            if (a < nbk && q < ncb) w = A[(b0 + a) + (int64_t)lda * (c0 + q)];
        } else {
            if (q < nbk && a < ncb) w = A[(c0 + a) + (int64_t)lda * (b0 + q)];
        }
        reg[e] = w;
    }
}

template <int RB, class C>
__global__ void __launch_bounds__(256) trs_sweep_kernel(const C* __restrict__ Ab, int lda, int64_t sA, int n, int R,
                                                        int lower, int conj, const cplx* __restrict__ Invb,
                                                        int64_t sInv, const cplx* __restrict__ Yb, int64_t sY,
                                                        cplx* __restrict__ Xb, int64_t sX, int* __restrict__ flags,
                                                        int64_t sF, int epoch, int batch, int* __restrict__ err) {
    constexpr int CPT = RB / 4;  // right-hand-side columns per thread
    constexpr int TE = TRS_B * TRS_B / 256;
    constexpr int XE = RB * TRS_B / 256;
    __shared__ C tileT[TRS_B][TRS_B + 1];  // tileT[s][i] = op(T)[b0 + i][c0 + s]
    __shared__ cplx xs[TRS_B][RB];
    const int tid = threadIdx.x, i = tid & (TRS_B - 1), rq = tid / TRS_B;
    const int nblk = (n + TRS_B - 1) / TRS_B;
    const int ntask = nblk * batch;
    for (int task = blockIdx.x; task < ntask; task += gridDim.x) {
        const int step = task / batch, item = task - step * batch;
        const int blk = lower ? step : nblk - 1 - step;
        const int b0 = blk * TRS_B, nbk = min(TRS_B, n - b0);
        const C* A = Ab + sA * item;
        const cplx* Y = Yb + sY * item;
        cplx* X = Xb + sX * item;
        int* fl = flags + sF * item;
        cplx acc[CPT];
#pragma unroll
        for (int cc = 0; cc < CPT; cc++) {
            const int col = rq * CPT + cc;
            acc[cc] = (i < nbk && col < R) ? Y[(b0 + i) + (int64_t)n * col] : cmk(0.0, 0.0);
        }
        // the factor tiles do not depend on the flags: the next one is loaded while this one is used
        C reg[TE];
        if (step > 0) {
            const int cb = lower ? 0 : nblk - 1;
            trs_load_tile(A, lda, conj, b0, nbk, cb * TRS_B, min(TRS_B, n - cb * TRS_B), tid, reg);
        }
        for (int d = 0; d < step; d++) {
            const int cb = lower ? d : nblk - 1 - d;  // dependency blocks in the order they are produced
            const int c0 = cb * TRS_B, ncb = min(TRS_B, n - c0);
            if (tid == 0) {
                int spins = 0;
                while (__hip_atomic_load(fl + cb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
                    // give up after TRS_SPIN_LIMIT polls, or at once when another wait already gave up
                    if ((++spins & 63) == 1 &&
                        (spins > TRS_SPIN_LIMIT || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                        atomicOr(err, 1);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            __syncthreads();  // flag seen; every thread is done with the previous tile
#pragma unroll
            for (int e = 0; e < TE; e++) {
                const int idx = tid + 256 * e;
                const int a = idx & (TRS_B - 1), q = idx / TRS_B;
                C w = reg[e];
                if (!conj) {
                    tileT[q][a] = w;
                } else {
                    w.y = -w.y;
                    tileT[a][q] = w;
                }
            }
            cplx xv[XE];
#pragma unroll
            for (int e = 0; e < XE; e++) {
                const int idx = tid + 256 * e;
                const int s = idx % TRS_B, r = idx / TRS_B;
                xv[e] = (s < ncb && r < R) ? ld_agent(X + (c0 + s) + (int64_t)n * r) : cmk(0.0, 0.0);
            }
#pragma unroll
            for (int e = 0; e < XE; e++) {
                const int idx = tid + 256 * e;
                xs[idx % TRS_B][idx / TRS_B] = xv[e];
            }
            if (d + 1 < step) {
                const int nb2 = lower ? d + 1 : nblk - 2 - d;
                trs_load_tile(A, lda, conj, b0, nbk, nb2 * TRS_B, min(TRS_B, n - nb2 * TRS_B), tid, reg);
            }
            __syncthreads();
#pragma unroll 8
            for (int s = 0; s < TRS_B; s++) {
                const cplx t = to_d(tileT[s][i]);
#pragma unroll
                for (int cc = 0; cc < CPT; cc++) acc[cc] = csub(acc[cc], cmul(t, xs[s][rq * CPT + cc]));
            }
        }
        __syncthreads();  // all threads done with xs of the last dependency
        // x_blk = op(Inv_blk) acc
#pragma unroll
        for (int cc = 0; cc < CPT; cc++) xs[i][rq * CPT + cc] = acc[cc];
        __syncthreads();
        const cplx* Inv = Invb + sInv * item + (int64_t)blk * TRS_B * TRS_B;
        cplx x[CPT];
#pragma unroll
        for (int cc = 0; cc < CPT; cc++) x[cc] = cmk(0.0, 0.0);
        if (i < nbk) {
#pragma unroll 4
            for (int s = 0; s < nbk; s++) {
                cplx v;
                if (!conj) v = Inv[i + TRS_B * s];
                else {
                    const cplx w = Inv[s + TRS_B * i];
                    v = cmk(w.x, -w.y);
                }
#pragma unroll
                for (int cc = 0; cc < CPT; cc++) x[cc] = cadd(x[cc], cmul(v, xs[s][rq * CPT + cc]));
            }
#pragma unroll
            for (int cc = 0; cc < CPT; cc++) {
                const int col = rq * CPT + cc;
                if (col < R) st_agent(X + (b0 + i) + (int64_t)n * col, x[cc]);
            }
        }
        // publish: the x stores above are device-coherent (agent-scope atomics); wait until every
        // thread's stores have completed, then raise the flag
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        __syncthreads();  // also: every thread is done with tileT and xs before the next task
        if (tid == 0) __hip_atomic_store(fl + blk, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// persistent grid: at most half of the workgroups the device can hold at once
template <int RB, class C>
static int trs_grid(int ntask) {
    int dev = 0, cus = 0, occ = 0;
    hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, trs_sweep_kernel<RB, C>, 256, 0) != hipSuccess || occ <= 0)
        occ = 1;
    return std::max(1, std::min(ntask, std::max(1, cus * occ / 2)));
}

template <class C>
static void trs_pass(hipStream_t st, const C* A, int lda, int64_t sA, int n, int R, int lower, int conj,
                     const cplx* Inv, int64_t sInv, const cplx* Y, int64_t sY, cplx* X, int64_t sX, int* flags,
                     int64_t sF, int epoch, int batch, int* err) {
    const int ntask = ((n + TRS_B - 1) / TRS_B) * batch;
    if (R <= 4)
        trs_sweep_kernel<4, C><<<trs_grid<4, C>(ntask), 256, 0, st>>>(A, lda, sA, n, R, lower, conj, Inv, sInv, Y, sY,
                                                                      X, sX, flags, sF, epoch, batch, err);
    else if (R <= 8)
        trs_sweep_kernel<8, C><<<trs_grid<8, C>(ntask), 256, 0, st>>>(A, lda, sA, n, R, lower, conj, Inv, sInv, Y, sY,
                                                                      X, sX, flags, sF, epoch, batch, err);
    else
        trs_sweep_kernel<16, C><<<trs_grid<16, C>(ntask), 256, 0, st>>>(A, lda, sA, n, R, lower, conj, Inv, sInv, Y,
                                                                        sY, X, sX, flags, sF, epoch, batch, err);
}

// B (n x R, leading dimension n) is overwritten with the solution.  work per matrix (stride sW):
// the wait flags of the triangular sweeps plus one error word (ceil(n/64) + 1 ints, rounded up to
// whole complex elements; first, so that their place does not depend on R), W1, W2 (n R complex
// each) and the diagonal-block inverses of L and U (2 ceil(n/64) 64^2 complex).  The flags must
// be zero when the work memory is first used; *epoch (host counter, one per work buffer) is
// advanced by every sweep.
static int64_t flag_words(int n) { return ((n + TRS_B - 1) / TRS_B + 1 + 3) / 4; }
int64_t getrs_flag_words(int n) { return flag_words(n); }

int64_t getrs_work(int n, int R) {
    const int64_t nblk = (n + TRS_B - 1) / TRS_B;
    return flag_words(n) + 2 * (int64_t)n * R + 2 * nblk * TRS_B * TRS_B;
}

template <class C>
static int getrs_t(hipStream_t st, int trans, int n, int R, const C* LU, int lda, int64_t sA, const int* ipiv,
                   int64_t ipiv_stride, int* perm, int64_t perm_stride, cplx* B, int64_t sB, cplx* work, int64_t sW,
                   int batch, bool prepared, int* epoch) {
    if (R > MAXR) return -1;
    if (n <= 0 || R <= 0) return 0;
    const int nblk = (n + TRS_B - 1) / TRS_B;
    int* flags = reinterpret_cast<int*>(work);
    int* err = flags + nblk;  // error word of item 0
    cplx* W1 = work + flag_words(n);
    cplx* W2 = W1 + (int64_t)n * R;
    cplx* Linv = W2 + (int64_t)n * R;
    cplx* Uinv = Linv + (int64_t)nblk * TRS_B * TRS_B;
    const int64_t sF = sW * (int64_t)(sizeof(cplx) / sizeof(int));
    if (!prepared) {
        invert_diag_kernel<C><<<dim3(nblk, 2, batch), TRS_B, 0, st>>>(LU, lda, sA, n, Linv, Uinv, sW);
        ipiv_to_perm_kernel<<<dim3(1, 1, batch), 256, 0, st>>>(ipiv, ipiv_stride, n, perm, perm_stride);
    }
    dim3 gp((n + 255) / 256, 1, batch);
    if (trans == 0) {
        permute_rows_kernel<<<gp, 256, 0, st>>>(n, R, perm, perm_stride, B, sB, W1, sW, 0);
        trs_pass(st, LU, lda, sA, n, R, 1, 0, Linv, sW, W1, sW, W2, sW, flags, sF, ++*epoch, batch, err);  // L
        trs_pass(st, LU, lda, sA, n, R, 0, 0, Uinv, sW, W2, sW, B, sB, flags, sF, ++*epoch, batch, err);   // U
    } else {
        for (int b = 0; b < batch; b++)
            hipMemcpyAsync(W1 + sW * b, B + sB * b, sizeof(cplx) * (size_t)n * R, hipMemcpyDeviceToDevice, st);
        trs_pass(st, LU, lda, sA, n, R, 1, 1, Uinv, sW, W1, sW, W2, sW, flags, sF, ++*epoch, batch, err);  // U^H
        trs_pass(st, LU, lda, sA, n, R, 0, 1, Linv, sW, W2, sW, W1, sW, flags, sF, ++*epoch, batch, err);  // L^H
        permute_rows_kernel<<<gp, 256, 0, st>>>(n, R, perm, perm_stride, W1, sW, B, sB, 1);
    }
    return 0;
}

int getrs_batched(hipStream_t st, int trans, int n, int R, const cplx* LU, int lda, int64_t sA, const int* ipiv,
                  int64_t ipiv_stride, int* perm, int64_t perm_stride, cplx* B, int64_t sB, cplx* work, int64_t sW,
                  int batch, int* epoch, bool prepared) {
    return getrs_t(st, trans, n, R, LU, lda, sA, ipiv, ipiv_stride, perm, perm_stride, B, sB, work, sW, batch, prepared,
                   epoch);
}


int getrs_batched(hipStream_t st, int trans, int n, int R, const cplxf* LU, int lda, int64_t sA, const int* ipiv,
                  int64_t ipiv_stride, int* perm, int64_t perm_stride, cplx* B, int64_t sB, cplx* work, int64_t sW,
                  int batch, int* epoch, bool prepared) {
    return getrs_t(st, trans, n, R, LU, lda, sA, ipiv, ipiv_stride, perm, perm_stride, B, sB, work, sW, batch, prepared,
                   epoch);
}

int* getrs_error_word(int n, cplx* work) { return reinterpret_cast<int*>(work) + (n + TRS_B - 1) / TRS_B; }


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
