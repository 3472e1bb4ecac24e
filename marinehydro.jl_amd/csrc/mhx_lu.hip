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
#include "mhx_internal.h"

namespace mhx {

#define LW 4
#define NB 64
#define LEAF_THREADS 1024

struct MatBatch {
    cplx* A;        // base pointer of matrix 0
    int64_t stride; // elements between consecutive matrices
    int lda;
};

// ---------------------------------------------------------------------------
// leaf: factor columns [c0, c0+LW) over rows [c0, n); swap the pivot rows in
// the other columns of the outer panel [K, KE); solve the 4-row U block of the
// outer panel columns right of the leaf.
// ---------------------------------------------------------------------------
template <int ROWS>
__global__ void __launch_bounds__(LEAF_THREADS) getf2_leaf_kernel(MatBatch mb, int n, int c0, int K, int KE,
                                                                  int* __restrict__ ipiv, int64_t ipiv_stride,
                                                                  int* __restrict__ info) {
    const int b = blockIdx.z;
    cplx* A = mb.A + mb.stride * b;
    int* piv = ipiv + ipiv_stride * b;
    const int lda = mb.lda;
    const int tid = threadIdx.x;
    __shared__ double red_v[LEAF_THREADS / 64];
    __shared__ int red_r[LEAF_THREADS / 64];
    __shared__ cplx xrow[2][LW];
    __shared__ cplx prow[LW];
    __shared__ int s_piv;
    __shared__ int s_pivs[LW];
    __shared__ cplx L11[LW][LW];

    const int wcols = min(LW, n - c0);
    cplx x[ROWS][LW];
#pragma unroll
    for (int q = 0; q < ROWS; q++) {
        int r = c0 + tid + LEAF_THREADS * q;
#pragma unroll
        for (int c = 0; c < LW; c++) x[q][c] = (r < n && c < wcols) ? A[r + (int64_t)lda * (c0 + c)] : cmk(0.0, 0.0);
    }
#pragma unroll
    for (int c = 0; c < LW; c++) {
        if (c >= wcols) break;
        const int col = c0 + c;
        // --- pivot search: max |Re|+|Im|, smallest row on ties
        double bv = -1.0;
        int br = 0x7fffffff;
#pragma unroll
        for (int q = 0; q < ROWS; q++) {
            int r = c0 + tid + LEAF_THREADS * q;
            double v = fabs(x[q][c].x) + fabs(x[q][c].y);
            if (r >= col && r < n && (v > bv || (v == bv && r < br))) { bv = v; br = r; }
        }
        for (int off = 32; off > 0; off >>= 1) {
            double ov = __shfl_xor(bv, off);
            int orow = __shfl_xor(br, off);
            if (ov > bv || (ov == bv && orow < br)) { bv = ov; br = orow; }
        }
        if ((tid & 63) == 0) { red_v[tid >> 6] = bv; red_r[tid >> 6] = br; }
        __syncthreads();
        if (tid == 0) {
            double v = red_v[0];
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
        const cplx u = prow[c];
        if (u.x != 0.0 || u.y != 0.0) {
            const cplx inv = cdiv(cmk(1.0, 0.0), u);
#pragma unroll
            for (int q = 0; q < ROWS; q++) {
                int r = c0 + tid + LEAF_THREADS * q;
                if (r > col && r < n) {
                    x[q][c] = cmul(x[q][c], inv);
#pragma unroll
                    for (int cc = 0; cc < LW; cc++)
                        if (cc > c) x[q][cc] = csub(x[q][cc], cmul(x[q][c], prow[cc]));
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
        cplx* colj = A + (int64_t)lda * j;
        for (int c = 0; c < wcols; c++) {
            int r0 = c0 + c, r1 = s_pivs[c];
            if (r1 != r0) {
                cplx t0 = colj[r0];
                colj[r0] = colj[r1];
                colj[r1] = t0;
            }
        }
        if (j > c0) {  // U12 rows c0..c0+wcols-1: unit-lower forward substitution
            cplx u[LW];
            for (int c = 0; c < wcols; c++) {
                cplx s = colj[c0 + c];
                for (int s2 = 0; s2 < c; s2++) s = csub(s, cmul(L11[c][s2], u[s2]));
                u[c] = s;
                colj[c0 + c] = s;
            }
        }
    }
}

// rank-LW update of the outer panel: A[r, j] -= sum_s L[r, c0+s] U[c0+s, j],
// rows r in [c0+LW, n), columns j in [c0+LW, KE)
__global__ void __launch_bounds__(256) panel_update_kernel(MatBatch mb, int n, int c0, int KE) {
    const int b = blockIdx.z;
    cplx* A = mb.A + mb.stride * b;
    const int lda = mb.lda;
    __shared__ cplx U[LW][NB];
    const int j0 = c0 + LW, nj = KE - j0;
    for (int t = threadIdx.x; t < LW * nj; t += blockDim.x) {
        int s = t / nj, jj = t % nj;
        U[s][jj] = A[(c0 + s) + (int64_t)lda * (j0 + jj)];
    }
    __syncthreads();
    const int r = c0 + LW + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    cplx l[LW];
#pragma unroll
    for (int s = 0; s < LW; s++) l[s] = A[r + (int64_t)lda * (c0 + s)];
    for (int jj = 0; jj < nj; jj++) {
        cplx* p = A + r + (int64_t)lda * (j0 + jj);
        cplx a = *p;
#pragma unroll
        for (int s = 0; s < LW; s++) a = csub(a, cmul(l[s], U[s][jj]));
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
__global__ void __launch_bounds__(256) apply_perm_kernel(MatBatch mb, int n, int K, int KE, const int* __restrict__ perm,
                                                         int64_t perm_stride) {
    const int b = blockIdx.z;
    cplx* A = mb.A + mb.stride * b;
    const int* pm = perm + perm_stride * b;
    const int m = pm[0];
    const int e = threadIdx.x % (2 * NB);
    const int cl = threadIdx.x / (2 * NB);
    int j = (blockIdx.x * (256 / (2 * NB))) + cl;
    const int nout = n - (KE - K);
    cplx v = cmk(0.0, 0.0);
    bool act = e < m && j < nout;
    if (j >= K) j += KE - K;
    if (act) v = A[pm[1 + 2 * NB + e] + (int64_t)mb.lda * j];
    __syncthreads();
    if (act) A[pm[1 + e] + (int64_t)mb.lda * j] = v;
}

// U12 = L11^{-1} A12 for the outer panel rows [K, KE), columns [KE, n); 4 columns per block
__global__ void __launch_bounds__(256) trsm_panel_kernel(MatBatch mb, int n, int K, int KE) {
    const int b = blockIdx.z;
    cplx* A = mb.A + mb.stride * b;
    const int lda = mb.lda;
    const int nb = KE - K;
    __shared__ cplx X[NB][4];
    const int i = threadIdx.x % NB, c = threadIdx.x / NB;
    const int j = KE + blockIdx.x * 4 + c;
    const bool act = i < nb && j < n;
    if (act) X[i][c] = A[(K + i) + (int64_t)lda * j];
    __syncthreads();
    for (int r = 0; r < nb - 1; r++) {
        if (act && i > r) {
            cplx l = A[(K + i) + (int64_t)lda * (K + r)];
            X[i][c] = csub(X[i][c], cmul(l, X[r][c]));
        }
        __syncthreads();
    }
    if (act) A[(K + i) + (int64_t)lda * j] = X[i][c];
}

// ---------------------------------------------------------------------------
// C[m x nn] (-)= A[m x kk] * B[kk x nn], complex, column-major, on FP64 MFMA.
// 256 threads (4 waves in 2x2), 64x64 tile, K chunks of 16 staged in LDS as
// separate real / imaginary planes.  v_mfma_f64_16x16x4f64 lane layout:
//   A: lane l holds A[l%16][l/16], B: B[l/16][l%16],
//   C: 4 values per lane, C[4*(l/16)+v][l%16].
// ---------------------------------------------------------------------------
typedef double double4_t __attribute__((ext_vector_type(4)));

template <bool SUB, bool ACCUM>
__global__ void __launch_bounds__(256) zgemm_mfma_kernel(int m, int nn, int kk, const cplx* __restrict__ Ab, int lda,
                                                         int64_t sA, const cplx* __restrict__ Bb, int ldb, int64_t sB,
                                                         cplx* __restrict__ Cb, int ldc, int64_t sC) {
    const int bz = blockIdx.z;
    const cplx* A = Ab + sA * bz;
    const cplx* B = Bb + sB * bz;
    cplx* C = Cb + sC * bz;
    __shared__ double Ar[16][64], Ai[16][64], Br[16][64], Bi[16][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int row0 = blockIdx.x * 64, col0 = blockIdx.y * 64;
    double4_t accr[2][2], acci[2][2];
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                double cr = 0.0, ci = 0.0;
                if (ACCUM) {
                    int r = row0 + wm + bi * 16 + 4 * v + lk, c = col0 + wn + bj * 16 + li;
                    if (r < m && c < nn) {
                        cplx t = C[r + (int64_t)ldc * c];
                        cr = t.x; ci = t.y;
                    }
                }
                accr[bi][bj][v] = cr;
                acci[bi][bj][v] = ci;
            }
        }
    for (int k0 = 0; k0 < kk; k0 += 16) {
        // A tile: 64 rows x 16 k; thread loads 4 elements (rows contiguous)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            int idx = tid + 256 * e;
            int r = idx & 63, k = idx >> 6;
            cplx t = cmk(0.0, 0.0);
            if (row0 + r < m && k0 + k < kk) t = A[(row0 + r) + (int64_t)lda * (k0 + k)];
            Ar[k][r] = SUB ? -t.x : t.x;
            Ai[k][r] = SUB ? -t.y : t.y;
        }
        // B tile: 16 k x 64 cols; k contiguous in memory
#pragma unroll
        for (int e = 0; e < 4; e++) {
            int idx = tid + 256 * e;
            int k = idx & 15, c = idx >> 4;
            cplx t = cmk(0.0, 0.0);
            if (col0 + c < nn && k0 + k < kk) t = B[(k0 + k) + (int64_t)ldb * (col0 + c)];
            Br[k][c] = t.x;
            Bi[k][c] = t.y;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 16; ks += 4) {
            double ar[2], ai[2], br[2], bi_[2];
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
                    accr[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[bi], br[bj], accr[bi][bj], 0, 0, 0);
                    accr[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[bi], bi_[bj], accr[bi][bj], 0, 0, 0);
                    acci[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[bi], bi_[bj], acci[bi][bj], 0, 0, 0);
                    acci[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[bi], br[bj], acci[bi][bj], 0, 0, 0);
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
                int r = row0 + wm + bi * 16 + 4 * v + lk, c = col0 + wn + bj * 16 + li;
                if (r < m && c < nn) C[r + (int64_t)ldc * c] = cmk(accr[bi][bj][v], acci[bi][bj][v]);
            }
}

void zgemm(hipStream_t st, bool sub, bool accum, int m, int nn, int kk, const cplx* A, int lda, int64_t sA,
           const cplx* B, int ldb, int64_t sB, cplx* C, int ldc, int64_t sC, int batch) {
    if (m <= 0 || nn <= 0) return;
    dim3 grid((m + 63) / 64, (nn + 63) / 64, batch);
    if (sub && accum)
        zgemm_mfma_kernel<true, true><<<grid, 256, 0, st>>>(m, nn, kk, A, lda, sA, B, ldb, sB, C, ldc, sC);
    else if (!sub && accum)
        zgemm_mfma_kernel<false, true><<<grid, 256, 0, st>>>(m, nn, kk, A, lda, sA, B, ldb, sB, C, ldc, sC);
    else
        zgemm_mfma_kernel<false, false><<<grid, 256, 0, st>>>(m, nn, kk, A, lda, sA, B, ldb, sB, C, ldc, sC);
}

// ---------------------------------------------------------------------------
// getrf driver.  ipiv: 0-based global row indices, int32, per matrix stride.
// work_perm: int workspace of (1 + 4*NB) per matrix.  info: int per matrix (0 / first zero pivot, 1-based).
// ---------------------------------------------------------------------------
int getrf_batched(hipStream_t st, int n, cplx* A, int lda, int64_t strideA, int batch, int* ipiv, int64_t ipiv_stride,
                  int* work_perm, int* info) {
    if (n <= 0) return 0;
    const int rows_per = (n + LEAF_THREADS - 1) / LEAF_THREADS;
    if (rows_per > 4) return -1;  // orders above 4096 need a larger register tile
    MatBatch mb{A, strideA, lda};
    const int64_t pstride = 1 + 4 * NB;
    for (int K = 0; K < n; K += NB) {
        const int KE = min(n, K + NB);
        for (int c0 = K; c0 < KE; c0 += LW) {
            const int m_rows = n - c0;
            const int rp = (m_rows + LEAF_THREADS - 1) / LEAF_THREADS;
            dim3 g1(1, 1, batch);
            if (rp <= 1)
                getf2_leaf_kernel<1><<<g1, LEAF_THREADS, 0, st>>>(mb, n, c0, K, KE, ipiv, ipiv_stride, info);
            else if (rp <= 2)
                getf2_leaf_kernel<2><<<g1, LEAF_THREADS, 0, st>>>(mb, n, c0, K, KE, ipiv, ipiv_stride, info);
            else
                getf2_leaf_kernel<4><<<g1, LEAF_THREADS, 0, st>>>(mb, n, c0, K, KE, ipiv, ipiv_stride, info);
            if (c0 + LW < KE && c0 + LW < n) {
                dim3 g2((n - c0 - LW + 255) / 256, 1, batch);
                panel_update_kernel<<<g2, 256, 0, st>>>(mb, n, c0, KE);
            }
        }
        if (n - (KE - K) > 0) {
            build_perm_kernel<<<dim3(1, 1, batch), 64, 0, st>>>(ipiv, ipiv_stride, K, KE, work_perm, pstride);
            const int nout = n - (KE - K);
            apply_perm_kernel<<<dim3((nout + 1) / 2, 1, batch), 256, 0, st>>>(mb, n, K, KE, work_perm, pstride);
        }
        if (KE < n) {
            trsm_panel_kernel<<<dim3((n - KE + 3) / 4, 1, batch), 256, 0, st>>>(mb, n, K, KE);
            zgemm(st, true, true, n - KE, n - KE, KE - K, A + KE + (int64_t)lda * K, lda, strideA,
                  A + K + (int64_t)lda * KE, lda, strideA, A + KE + (int64_t)lda * KE, lda, strideA, batch);
        }
    }
    return 0;
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
__global__ void __launch_bounds__(TRS_B) invert_diag_kernel(const cplx* __restrict__ Ab, int lda, int64_t sA, int n,
                                                            cplx* __restrict__ Linv, cplx* __restrict__ Uinv,
                                                            int64_t sInv) {
    const cplx* A = Ab + sA * blockIdx.z;
    const int blk = blockIdx.x, upper = blockIdx.y;
    const int b0 = blk * TRS_B, nbk = min(TRS_B, n - b0);
    cplx* out = (upper ? Uinv : Linv) + sInv * blockIdx.z + (int64_t)blk * TRS_B * TRS_B;
    __shared__ cplx T[TRS_B][TRS_B];  // T[i][j]
    const int t = threadIdx.x;
    for (int j = 0; j < TRS_B; j++) {
        cplx v = cmk(0.0, 0.0);
        if (t < nbk && j < nbk) v = A[(b0 + t) + (int64_t)lda * (b0 + j)];
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

__device__ __forceinline__ cplx tri_elem(const cplx* A, int lda, int i, int j, int conj) {
    // element (i, j) of op(A): A or A^H
    if (!conj) return A[i + (int64_t)lda * j];
    cplx v = A[j + (int64_t)lda * i];
    return cmk(v.x, -v.y);
}

// One block-step of a triangular solve with TRS_B-row diagonal blocks.
//  lower=1: forward (L, or U^H when conj); rows [b0,b1) solved, rows below updated
//  lower=0: backward (U, or L^H when conj); rows [b0,b1) solved, rows above updated
// x_b = op(Inv_b) y_b with the precomputed diagonal-block inverse (Inv = L_bb^{-1} or
// U_bb^{-1}; op = conjugate transpose when conj).  Every block forms x_b from Bw (not
// written in this launch); block 0 stores it into X, the others update their Bw rows.
template <int RB>
__global__ void __launch_bounds__(256) trs_step_kernel(const cplx* __restrict__ Ab, int lda, int64_t sA, int n, int R,
                                                       int b0, int b1, int lower, int conj,
                                                       const cplx* __restrict__ Invb, int64_t sInv,
                                                       cplx* __restrict__ Bwb, int64_t sB, cplx* __restrict__ Xb,
                                                       int64_t sX) {
    const int bz = blockIdx.z;
    const cplx* A = Ab + sA * bz;
    const cplx* Inv = Invb + sInv * bz + (int64_t)(b0 / TRS_B) * TRS_B * TRS_B;
    cplx* Bw = Bwb + sB * bz;
    cplx* X = Xb + sX * bz;
    const int nbk = b1 - b0;
    __shared__ cplx ys[TRS_B][RB];
    __shared__ cplx xs[TRS_B][RB];
    const int tid = threadIdx.x;
    for (int t = tid; t < TRS_B * RB; t += 256) {
        int i = t % TRS_B, c = t / TRS_B;
        ys[i][c] = (i < nbk && c < R) ? Bw[(b0 + i) + (int64_t)n * c] : cmk(0.0, 0.0);
    }
    __syncthreads();
    for (int t = tid; t < TRS_B * RB; t += 256) {
        int i = t % TRS_B, c = t / TRS_B;
        cplx acc = cmk(0.0, 0.0);
        for (int s2 = 0; s2 < nbk; s2++) {
            cplx v;
            if (!conj) v = Inv[i + TRS_B * s2];
            else {
                cplx w = Inv[s2 + TRS_B * i];
                v = cmk(w.x, -w.y);
            }
            acc = cadd(acc, cmul(v, ys[s2][c]));
        }
        xs[i][c] = acc;
    }
    __syncthreads();
    if (blockIdx.x == 0) {
        for (int t = tid; t < nbk * R; t += 256) {
            int i = t % nbk, c = t / nbk;
            X[(b0 + i) + (int64_t)n * c] = xs[i][c];
        }
        return;
    }
    const bool fwd = lower;
    const int row = fwd ? b1 + (blockIdx.x - 1) * 256 + tid : (blockIdx.x - 1) * 256 + tid;
    if (fwd ? (row >= n) : (row >= b0)) return;
    cplx acc[RB];
#pragma unroll
    for (int c = 0; c < RB; c++) acc[c] = (c < R) ? Bw[row + (int64_t)n * c] : cmk(0.0, 0.0);
    for (int s2 = 0; s2 < nbk; s2++) {
        cplx a = tri_elem(A, lda, row, b0 + s2, conj);
#pragma unroll
        for (int c = 0; c < RB; c++) acc[c] = csub(acc[c], cmul(a, xs[s2][c]));
    }
#pragma unroll
    for (int c = 0; c < RB; c++)
        if (c < R) Bw[row + (int64_t)n * c] = acc[c];
}

static void trs_pass(hipStream_t st, const cplx* A, int lda, int64_t sA, int n, int R, int lower, int conj,
                     const cplx* Inv, int64_t sInv, cplx* Bw, int64_t sB, cplx* X, int64_t sX, int batch) {
    const int nblk = (n + TRS_B - 1) / TRS_B;
    for (int s = 0; s < nblk; s++) {
        const int blk = lower ? s : nblk - 1 - s;
        const int b0 = blk * TRS_B, b1 = min(n, b0 + TRS_B);
        const int rest = lower ? n - b1 : b0;
        dim3 g(1 + (rest + 255) / 256, 1, batch);
        if (R <= 4)
            trs_step_kernel<4><<<g, 256, 0, st>>>(A, lda, sA, n, R, b0, b1, lower, conj, Inv, sInv, Bw, sB, X, sX);
        else if (R <= 8)
            trs_step_kernel<8><<<g, 256, 0, st>>>(A, lda, sA, n, R, b0, b1, lower, conj, Inv, sInv, Bw, sB, X, sX);
        else
            trs_step_kernel<16><<<g, 256, 0, st>>>(A, lda, sA, n, R, b0, b1, lower, conj, Inv, sInv, Bw, sB, X, sX);
    }
}

// B (n x R, leading dimension n) is overwritten with the solution.  work: 2 n R complex
// plus 2 * ceil(n/64) * 64 * 64 complex (diagonal-block inverses) per matrix, stride sW.
int64_t getrs_work(int n, int R) {
    const int64_t nblk = (n + TRS_B - 1) / TRS_B;
    return 2 * (int64_t)n * R + 2 * nblk * TRS_B * TRS_B;
}

int getrs_batched(hipStream_t st, int trans, int n, int R, const cplx* LU, int lda, int64_t sA, const int* ipiv,
                  int64_t ipiv_stride, int* perm, int64_t perm_stride, cplx* B, int64_t sB, cplx* work, int64_t sW,
                  int batch) {
    if (R > MAXR) return -1;
    if (n <= 0 || R <= 0) return 0;
    const int nblk = (n + TRS_B - 1) / TRS_B;
    cplx* W1 = work;
    cplx* W2 = work + (int64_t)n * R;
    cplx* Linv = work + 2 * (int64_t)n * R;
    cplx* Uinv = Linv + (int64_t)nblk * TRS_B * TRS_B;
    invert_diag_kernel<<<dim3(nblk, 2, batch), TRS_B, 0, st>>>(LU, lda, sA, n, Linv, Uinv, sW);
    ipiv_to_perm_kernel<<<dim3(1, 1, batch), 256, 0, st>>>(ipiv, ipiv_stride, n, perm, perm_stride);
    dim3 gp((n + 255) / 256, 1, batch);
    if (trans == 0) {
        permute_rows_kernel<<<gp, 256, 0, st>>>(n, R, perm, perm_stride, B, sB, W1, sW, 0);
        trs_pass(st, LU, lda, sA, n, R, 1, 0, Linv, sW, W1, sW, W2, sW, batch);  // L: W1 -> W2
        trs_pass(st, LU, lda, sA, n, R, 0, 0, Uinv, sW, W2, sW, B, sB, batch);   // U: W2 -> B
    } else {
        for (int b = 0; b < batch; b++)
            hipMemcpyAsync(W1 + sW * b, B + sB * b, sizeof(cplx) * (size_t)n * R, hipMemcpyDeviceToDevice, st);
        trs_pass(st, LU, lda, sA, n, R, 1, 1, Uinv, sW, W1, sW, W2, sW, batch);  // U^H (lower)
        trs_pass(st, LU, lda, sA, n, R, 0, 1, Linv, sW, W2, sW, W1, sW, batch);  // L^H (unit upper)
        permute_rows_kernel<<<gp, 256, 0, st>>>(n, R, perm, perm_stride, W1, sW, B, sB, 1);
    }
    return 0;
}

}  // namespace mhx
