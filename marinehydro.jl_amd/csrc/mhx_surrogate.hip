// Learned surrogate of the exact Guevel-Delhommeau wave term.
//
// The network maps the reduced coordinates (r, v) = (k R_h, k |z + zeta|) to six
// outputs: 100 x the residual (exact - Liang-Wu-Noblesse fit) of Re/Im G, Re/Im dG/dr
// and Re/Im dG/dz.  Architecture: 2 inputs -> 64 Fourier features
// [sin(2 pi B u), cos(2 pi B u)] of u = (log-scaled r, log-scaled v) -> 4 x 128 SiLU
// -> 6 linear outputs.  The fit it corrects is the reference's GFWu
// (src/green_functions/wu.jl); the exact term is exact_Guevel_Delhommeau.jl:15-59.
//
//  * samples: ngrid^2 points on a log-spaced grid with seeded jitter, targets from the
//    exact FP64 kernel (hquad_pair) and the Wu fit, on the device; imaginary residuals
//    are scaled by exp(min(v, 5)) (the imaginary parts decay like exp(-v));
//  * training: FP32 kernels (tiled GEMMs with fused bias/SiLU epilogues, split-K
//    weight gradients, MSE loss, Adam);
//  * inference: BF16 weights and activations on v_mfma_f32_32x32x8bf16_1k with FP32
//    accumulation, fused into the pair-table assembly; pairs outside the trained box
//    or in the free-surface region (v < 1e-3) go to the exact kernel.
#include "mhx_surrogate.h"

#include <algorithm>

namespace mhx {

// ---------------------------------------------------------------------------
// input map
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ void surr_inputs(double r, double v, float& u1, float& u2) {
    const double lr0 = -6.907755278982137;   // log(1e-3)
    const double lr1 = 4.6051801859294744;    // log(100 + 1e-3)
    const double lv0 = -7.600902459542082;   // log(5e-4)
    const double lv1 = 3.912023005428146;    // log(50)
    double vv = v < 5e-4 ? 5e-4 : v;
    u1 = (float)((log(r + 1e-3) - lr0) / (lr1 - lr0));
    u2 = (float)((log(vv) - lv0) / (lv1 - lv0));
}

__device__ __forceinline__ bool surr_in_box(double r, double v) {
    return r <= SURR_RMAX && v <= SURR_VMAX && v >= SURR_VMIN;
}

// ---------------------------------------------------------------------------
// samples: exact and Wu at (r, v) -> inputs u (2 floats) and scaled residual targets (6)
// ---------------------------------------------------------------------------
__global__ void surr_samples_kernel(int ngrid, uint64_t seed, float* __restrict__ U, float* __restrict__ Y,
                                    unsigned long long* __restrict__ status) {
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)ngrid * ngrid;
    if (idx >= total) return;
    const int ir = (int)(idx % ngrid), iv = (int)(idx / ngrid);
    // seeded jitter (splitmix64)
    uint64_t x = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(idx + 1);
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    const double j1 = (double)(x >> 11) * (1.0 / 9007199254740992.0);
    x = x * 0x2545F4914F6CDD1Dull + 1;
    const double j2 = (double)(x >> 11) * (1.0 / 9007199254740992.0);
    double r = ir == 0 ? 0.0 : exp(log(1e-4) + (log(SURR_RMAX) - log(1e-4)) * ((ir - 1 + j1) / (ngrid - 1)));
    double v = exp(log(SURR_VMIN * 0.5) + (log(SURR_VMAX) - log(SURR_VMIN * 0.5)) * ((iv + j2) / ngrid));
    QuadWork wk;
    cplx G, Gr;
    if (!hquad_pair(r, -v, 1e-8, 1e-8, G, Gr, wk)) atomicOr(status, 1ull);
    G = cscale(4.0, G);
    Gr = cscale(4.0, Gr);
    cplx W, Wh, Wv;
    wu_term(r > 0 ? r : 1e-300, -v, W, Wh, Wv);
    const cplx Gz = cmk(G.x + 2.0 / hypot(r, v), G.y);
    float u1, u2;
    surr_inputs(r, v, u1, u2);
    U[2 * idx] = u1;
    U[2 * idx + 1] = u2;
    const double S = SURR_SCALE;
    // the imaginary parts decay like exp(-v) (2 pi e^z J0(r), 2 pi e^z J1(r)); their residuals are
    // learned relative to that envelope so that the network's own error decays with it
    const double Si = S * exp(fmin(v, SURR_IM_VC));
    Y[6 * idx + 0] = (float)(S * (G.x - W.x));
    Y[6 * idx + 1] = (float)(Si * (G.y - W.y));
    Y[6 * idx + 2] = (float)(S * (Gr.x - Wh.x));
    Y[6 * idx + 3] = (float)(Si * (Gr.y - Wh.y));
    Y[6 * idx + 4] = (float)(S * (Gz.x - Wv.x));
    Y[6 * idx + 5] = (float)(Si * (Gz.y - Wv.y));
}

void launch_surr_samples(hipStream_t st, int ngrid, uint64_t seed, float* U, float* Y, unsigned long long* status) {
    const int64_t total = (int64_t)ngrid * ngrid;
    surr_samples_kernel<<<(total + 255) / 256, 256, 0, st>>>(ngrid, seed, U, Y, status);
}

// ---------------------------------------------------------------------------
// training kernels (FP32)
// ---------------------------------------------------------------------------
// features F[m][0..63] = [sin(2 pi B u), cos(2 pi B u)] for batch rows idx[m]
__global__ void surr_features_kernel(int M, const int* __restrict__ idx, const float* __restrict__ U,
                                     const float* __restrict__ Bf, float* __restrict__ F) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M * SURR_NF) return;
    const int m = t / SURR_NF, f = t % SURR_NF;
    const int64_t s = idx ? idx[m] : m;
    const float u1 = U[2 * s], u2 = U[2 * s + 1];
    const float p = 6.283185307179586f * (Bf[2 * f] * u1 + Bf[2 * f + 1] * u2);
    F[(int64_t)m * 64 + f] = sinf(p);
    F[(int64_t)m * 64 + SURR_NF + f] = cosf(p);
}

void launch_surr_features(hipStream_t st, int M, const int* idx, const float* U, const float* Bf, float* F) {
    surr_features_kernel<<<(M * SURR_NF + 255) / 256, 256, 0, st>>>(M, idx, U, Bf, F);
}

__device__ __forceinline__ float silu_f(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu_f(float z) {
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// C[M x N] = sum_k A(m, k) B(k, n).  A(m,k) = TA ? A[k*lda+m] : A[m*lda+k];
// B(k,n) = TB ? B[n*ldb+k] : B[k*ldb+n].  Epilogues:
//  EPI 0: C = acc (split-K partial: C + blockIdx.z * M*N, k range split over gridDim.z)
//  EPI 1: Z = acc + bias[n], H = silu(Z)   (C = Z, C2 = H)
//  EPI 2: Z = acc + bias[n]                 (C = Z)
//  EPI 3: C = acc * silu'(Zp[m*N+n])        (backward through the activation)
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) sgemm_kernel(int M, int N, int K, const float* __restrict__ A, int lda,
                                                    const float* __restrict__ B, int ldb, float* __restrict__ C,
                                                    float* __restrict__ C2, const float* __restrict__ bias,
                                                    const float* __restrict__ Zp) {
    __shared__ float As[16][64 + 4];
    __shared__ float Bs[16][64 + 4];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    int kb = 0, ke = K;
    if (EPI == 0) {
        const int per = ((K + gridDim.z - 1) / gridDim.z + 15) / 16 * 16;
        kb = blockIdx.z * per;
        ke = min(K, kb + per);
    }
    float acc[4][4] = {};
    for (int k0 = kb; k0 < ke; k0 += 16) {
        for (int e = threadIdx.x; e < 16 * 64; e += 256) {
            const int kk = e & 15, mm = e >> 4;
            const int gm = m0 + mm, gk = k0 + kk;
            float a = 0.f;
            if (gm < M && gk < ke) a = TA ? A[(int64_t)gk * lda + gm] : A[(int64_t)gm * lda + gk];
            As[kk][mm] = a;
            const int gn = n0 + mm;
            float bv = 0.f;
            if (gn < N && gk < ke) bv = TB ? B[(int64_t)gn * ldb + gk] : B[(int64_t)gk * ldb + gn];
            Bs[kk][mm] = bv;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
            float a[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; j++) bv[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] += a[i] * bv[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int m = m0 + ty * 4 + i, nn = n0 + tx * 4 + j;
            if (m >= M || nn >= N) continue;
            const int64_t o = (int64_t)m * N + nn;
            if (EPI == 0) {
                C[(int64_t)blockIdx.z * M * N + o] = acc[i][j];
            } else if (EPI == 1) {
                const float z = acc[i][j] + bias[nn];
                C[o] = z;
                C2[o] = silu_f(z);
            } else if (EPI == 2) {
                C[o] = acc[i][j] + bias[nn];
            } else {
                C[o] = acc[i][j] * dsilu_f(Zp[o]);
            }
        }
}

// sum of nsplit partials (each n elements) into out; col-sums for bias gradients
__global__ void sum_parts_kernel(int n, int nsplit, const float* __restrict__ parts, float* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float s = 0.f;
    for (int q = 0; q < nsplit; q++) s += parts[(int64_t)q * n + t];
    out[t] = s;
}

__global__ void colsum_kernel(int M, int N, const float* __restrict__ X, float* __restrict__ out) {
    const int nn = blockIdx.x;
    float s = 0.f;
    for (int m = threadIdx.x; m < M; m += blockDim.x) s += X[(int64_t)m * N + nn];
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[nn] = red[0];
}

// dZ = 2 w (Z - Y) / (M * 6), loss partial sums
__global__ void mse_grad_kernel(int M, const int* __restrict__ idx, const float* __restrict__ Z,
                                const float* __restrict__ Yall, float* __restrict__ dZ, float* __restrict__ loss) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (t < M * SURR_NOUT) {
        const int m = t / SURR_NOUT, o = t % SURR_NOUT;
        const int64_t s = idx ? idx[m] : m;
        const float d = Z[t] - Yall[6 * s + o];
        dZ[t] = 2.f * d / (float)(M * SURR_NOUT);
        l = d * d / (float)(M * SURR_NOUT);
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, l);
}

__global__ void adam_kernel(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                            float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    const float a = b1 * m1[t] + (1.f - b1) * gt;
    const float b = b2 * m2[t] + (1.f - b2) * gt * gt;
    m1[t] = a;
    m2[t] = b;
    p[t] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// random minibatch indices (excluding the validation stride)
__global__ void batch_index_kernel(int M, int64_t nsamples, uint64_t seed, int* __restrict__ idx) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M) return;
    uint64_t x = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(t + 1);
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    int64_t s = (int64_t)(x % (uint64_t)nsamples);
    if (s % SURR_VAL_STRIDE == 0) s = (s + 1) % nsamples;
    idx[t] = (int)s;
}

// ---- training driver pieces (host) ----
struct SgemmArgs {
    int M, N, K;
    const float* A; int lda;
    const float* B; int ldb;
    float* C; float* C2; const float* bias; const float* Zp;
};

template <bool TA, bool TB, int EPI>
static void sgemm(hipStream_t st, const SgemmArgs& a, int splits = 1) {
    dim3 grid((a.N + 63) / 64, (a.M + 63) / 64, splits);
    sgemm_kernel<TA, TB, EPI><<<grid, 256, 0, st>>>(a.M, a.N, a.K, a.A, a.lda, a.B, a.ldb, a.C, a.C2, a.bias, a.Zp);
}

// parameter offsets inside the flat FP32 parameter vector
static const int64_t OFF_W[5] = {0, 128 * 64 + 128, 128 * 64 + 128 + 128 * 128 + 128,
                                 128 * 64 + 128 + 2 * (128 * 128 + 128), 128 * 64 + 128 + 3 * (128 * 128 + 128)};
static const int IN_DIM[5] = {64, 128, 128, 128, 128};
static const int OUT_DIM[5] = {128, 128, 128, 128, SURR_NOUT};

int64_t surr_param_count() { return OFF_W[4] + (int64_t)SURR_NOUT * 128 + SURR_NOUT; }

// one forward + backward pass on M rows (idx or identity), gradients into grad (flat, zeroed here)
void surr_forward_backward(hipStream_t st, SurrTrainWork& w, int M, const int* idx, const float* U, const float* Y,
                           const float* params, float* grad, float* loss) {
    launch_surr_features(st, M, idx, U, w.Bf, w.act[0]);
    const float* hin = w.act[0];
    for (int l = 0; l < 5; l++) {
        const float* W = params + OFF_W[l];
        const float* b = W + (int64_t)OUT_DIM[l] * IN_DIM[l];
        SgemmArgs a{M, OUT_DIM[l], IN_DIM[l], hin, IN_DIM[l], W, IN_DIM[l], w.z[l], w.act[l + 1], b, nullptr};
        if (l < 4) sgemm<false, true, 1>(st, a);
        else sgemm<false, true, 2>(st, a);
        hin = w.act[l + 1];
    }
    if (loss) hipMemsetAsync(loss, 0, sizeof(float), st);
    mse_grad_kernel<<<(M * SURR_NOUT + 255) / 256, 256, 0, st>>>(M, idx, w.z[4], Y, w.dz[4], loss);
    for (int l = 4; l >= 0; l--) {
        float* gW = grad + OFF_W[l];
        float* gb = gW + (int64_t)OUT_DIM[l] * IN_DIM[l];
        const float* hprev = w.act[l];  // features for l = 0
        // dW[o][i] = sum_m dz[m][o] h[m][i]
        const int splits = 32;
        SgemmArgs a{OUT_DIM[l], IN_DIM[l], M, w.dz[l], OUT_DIM[l], hprev, IN_DIM[l], w.part, nullptr, nullptr, nullptr};
        sgemm<true, false, 0>(st, a, splits);
        const int nW = OUT_DIM[l] * IN_DIM[l];
        sum_parts_kernel<<<(nW + 255) / 256, 256, 0, st>>>(nW, splits, w.part, gW);
        colsum_kernel<<<OUT_DIM[l], 256, 0, st>>>(M, OUT_DIM[l], w.dz[l], gb);
        if (l > 0) {
            // dz[l-1] = (dz[l] W_l) * silu'(z[l-1])
            const float* W = params + OFF_W[l];
            SgemmArgs d{M, IN_DIM[l], OUT_DIM[l], w.dz[l], OUT_DIM[l], W, IN_DIM[l], w.dz[l - 1], nullptr, nullptr,
                        w.z[l - 1]};
            sgemm<false, false, 3>(st, d);
        }
    }
}

void surr_adam(hipStream_t st, int64_t n, float* p, const float* g, float* m1, float* m2, float lr, int step) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    adam_kernel<<<(n + 255) / 256, 256, 0, st>>>((int)n, p, g, m1, m2, lr, b1, b2, eps, c1, c2);
}

void launch_batch_index(hipStream_t st, int M, int64_t nsamples, uint64_t seed, int* idx) {
    batch_index_kernel<<<(M + 255) / 256, 256, 0, st>>>(M, nsamples, seed, idx);
}

// ---------------------------------------------------------------------------
// inference: BF16 MFMA, fused into the pair table
// ---------------------------------------------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__host__ __device__ __forceinline__ unsigned short f2bf(float f) {
    union {
        float f;
        unsigned int u;
    } c;
    c.f = f;
    unsigned int u = c.u;
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ int IN_DIM_D(int l) { return l == 0 ? 64 : 128; }
__device__ __forceinline__ int OUT_DIM_D(int l) { return l == 4 ? SURR_NOUT : 128; }
__device__ __forceinline__ int64_t OFF_W_D(int l) { return l == 0 ? 0 : 128 * 64 + 128 + (int64_t)(l - 1) * (128 * 128 + 128); }

// FP32 parameters -> BF16 weight blocks [l][n][k] (n padded to 32 for the output layer)
__global__ void surr_pack_kernel(const float* __restrict__ params, unsigned short* __restrict__ Wb,
                                 float* __restrict__ bias) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    // layer offsets in Wb: 128*64, 3 * 128*128, 32*128
    if (t < SURR_PACKED_W) {
        int l, o = t;
        int64_t base;
        if (o < 128 * 64) { l = 0; base = 0; }
        else if (o < 128 * 64 + 3 * 128 * 128) { l = 1 + (o - 128 * 64) / (128 * 128); base = 128 * 64 + (int64_t)(l - 1) * 128 * 128; }
        else { l = 4; base = 128 * 64 + 3 * 128 * 128; }
        const int in = IN_DIM_D(l);
        const int r = (int)((o - base) / in), k = (int)((o - base) % in);
        float v = 0.f;
        if (r < OUT_DIM_D(l)) v = params[OFF_W_D(l) + (int64_t)r * in + k];
        Wb[t] = f2bf(v);
    }
    if (t < 4 * 128 + 32) {
        const int l = t / 128 < 4 ? t / 128 : 4;
        const int r = t - 128 * l;
        float v = 0.f;
        if (r < OUT_DIM_D(l)) v = params[OFF_W_D(l) + (int64_t)OUT_DIM_D(l) * IN_DIM_D(l) + r];
        bias[t] = v;
    }
}

void launch_surr_pack(hipStream_t st, const float* params, unsigned short* Wb, float* bias) {
    surr_pack_kernel<<<(SURR_PACKED_W + 255) / 256, 256, 0, st>>>(params, Wb, bias);
}

// one wave: 32 pairs x 128 neurons on MFMA 32x32x8 (four 32-column tiles)
template <int KIN, int NT>
__device__ __forceinline__ void surr_layer(const unsigned short* __restrict__ act, int act_ld,
                                           const unsigned short* __restrict__ W, f32x16 acc[NT], int lane) {
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int q = 0; q < 16; q++) acc[t][q] = 0.f;
    const int row = lane & 31, kq = (lane >> 5) * 4;
#pragma unroll 2
    for (int k0 = 0; k0 < KIN; k0 += 8) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(act + row * act_ld + k0 + kq);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const bf16x4 b = *reinterpret_cast<const bf16x4*>(W + (t * 32 + row) * KIN + k0 + kq);
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc[t], 0, 0, 0);
        }
    }
}

#define SURR_TILE_I 16
#define SURR_TILE_J 8
#define SURR_ACT_LD 136  // 128 + 8 bf16 padding

// One workgroup (4 waves) handles a 16 x 8 block of (i, j) pairs per step (each wave
// 32 pairs: 4 rows i x 8 columns j), for unordered pairs i <= j.  Pairs in the box go
// through the network; others are appended to the exact fallback list.
__global__ void __launch_bounds__(256) surr_table_kernel(int n, const PanelGeom* __restrict__ geom,
                                                         const double* __restrict__ ks,
                                                         const unsigned short* __restrict__ Wb,
                                                         const float* __restrict__ bias, const float* __restrict__ Bf,
                                                         cplx* __restrict__ WGb, cplx* __restrict__ WRb, int64_t sMat,
                                                         int* __restrict__ fb_list, int* __restrict__ fb_count,
                                                         int fb_cap) {
    const int bi = blockIdx.x, bj = blockIdx.y;
    const int i0 = bi * SURR_TILE_I, j0 = bj * SURR_TILE_J;
    if (i0 > j0 + SURR_TILE_J - 1) return;  // whole tile below the diagonal
    const double k = ks[blockIdx.z];
    cplx* WG = WGb + sMat * blockIdx.z;
    cplx* WR = WRb + sMat * blockIdx.z;
    __shared__ unsigned short act[128 * SURR_ACT_LD];
    __shared__ float outv[128][SURR_NOUT];
    __shared__ float sB[2 * SURR_NF];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < 2 * SURR_NF) sB[tid] = Bf[tid];
    // pair of this thread (threads 0..127): local p = tid; i = i0 + p / 8, j = j0 + p % 8
    const int p = tid & 127;
    const int i = i0 + p / SURR_TILE_J, j = j0 + p % SURR_TILE_J;
    double hh = 0.0, vv = 0.0;
    bool valid = false, net = false;
    if (i < n && j < n && i <= j) {
        const PanelGeom& gi = geom[i];
        const PanelGeom& gj = geom[j];
        reduced_hv(gi.c[0], gi.c[1], gi.c[2], gj.c[0], gj.c[1], gj.c[2], k, hh, vv);
        valid = true;
        net = surr_in_box(hh, -vv);
    }
    __syncthreads();
    // features: each of the 256 threads writes 32 of the 64 features of its pair (bf16)
    {
        float u1, u2;
        surr_inputs(hh, -vv, u1, u2);
        const int half = tid >> 7;  // features [32 half, 32 half + 32)
        for (int f = half * 32; f < half * 32 + 32; f++) {
            const int ff = f % SURR_NF;
            const float ph = 6.283185307179586f * (sB[2 * ff] * u1 + sB[2 * ff + 1] * u2);
            const float val = f < SURR_NF ? sinf(ph) : cosf(ph);
            act[p * SURR_ACT_LD + f] = f2bf(val);
        }
    }
    __syncthreads();
    // four waves x 32 pairs: pairs [32 wave, 32 wave + 32)
    unsigned short* myact = act + (32 * wave) * SURR_ACT_LD;
    const unsigned short* Wl = Wb;
    f32x16 acc[4];
    // hidden layers
#pragma unroll 1
    for (int l = 0; l < 4; l++) {
        if (l == 0) surr_layer<64, 4>(myact, SURR_ACT_LD, Wl, acc, lane);
        else surr_layer<128, 4>(myact, SURR_ACT_LD, Wl, acc, lane);
        Wl += (l == 0 ? 128 * 64 : 128 * 128);
        // SiLU, back to bf16 in place (same wave: its reads above are complete)
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int row = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                const int col = t * 32 + (lane & 31);
                float z = acc[t][q] + bias[l * 128 + col];
                z = z / (1.0f + __expf(-z));
                myact[row * SURR_ACT_LD + col] = f2bf(z);
            }
    }
    // output layer (6 of 32 columns used)
    {
        f32x16 o[1];
        surr_layer<128, 1>(myact, SURR_ACT_LD, Wl, o, lane);
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int row = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
            const int col = lane & 31;
            if (col < SURR_NOUT) outv[32 * wave + row][col] = o[0][q] + bias[4 * 128 + col];
        }
    }
    __syncthreads();
    if (tid >= 128 || !valid) return;
    if (!net) {
        const int slot = atomicAdd(fb_count, 1);
        if (slot < fb_cap) {
            fb_list[2 * slot] = i;
            fb_list[2 * slot + 1] = j;
        }
        return;
    }
    cplx W, Wh, Wv;
    wu_term(hh > 0 ? hh : 1e-300, vv, W, Wh, Wv);
    const double is = 1.0 / SURR_SCALE;
    const double ii = is * exp(-fmin(-vv, SURR_IM_VC));
    const cplx G = cmk(W.x + is * outv[p][0], W.y + ii * outv[p][1]);
    const cplx Gr = cmk(Wh.x + is * outv[p][2], Wh.y + ii * outv[p][3]);
    WG[i + (int64_t)n * j] = G;
    WR[i + (int64_t)n * j] = Gr;
    if (i != j) {
        WG[j + (int64_t)n * i] = G;
        WR[j + (int64_t)n * i] = Gr;
    }
}

// exact term for the fallback pairs of one item
__global__ void surr_fallback_kernel(int n, const PanelGeom* __restrict__ geom, const double* __restrict__ ks,
                                     int item, const int* __restrict__ fb_list, const int* __restrict__ fb_count,
                                     int fb_cap, cplx* __restrict__ WGb, cplx* __restrict__ WRb, int64_t sMat,
                                     unsigned long long* __restrict__ status) {
    const int cnt = min(*fb_count, fb_cap);
    const double k = ks[item];
    cplx* WG = WGb + sMat * item;
    cplx* WR = WRb + sMat * item;
    // grid-stride over the listed pairs (the list length is only known on the device)
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < cnt; t += gridDim.x * blockDim.x) {
        const int i = fb_list[2 * t], j = fb_list[2 * t + 1];
        const PanelGeom& gi = geom[i];
        const PanelGeom& gj = geom[j];
        double hh, vv;
        reduced_hv(gi.c[0], gi.c[1], gi.c[2], gj.c[0], gj.c[1], gj.c[2], k, hh, vv);
        QuadWork wk;
        cplx G, Gr;
        if (!hquad_pair(hh, vv, 1e-8, 1e-8, G, Gr, wk)) atomicOr(status, 1ull);
        G = cscale(4.0, G);
        Gr = cscale(4.0, Gr);
        WG[i + (int64_t)n * j] = G;
        WR[i + (int64_t)n * j] = Gr;
        if (i != j) {
            WG[j + (int64_t)n * i] = G;
            WR[j + (int64_t)n * i] = Gr;
        }
    }
}

void launch_surr_table(hipStream_t st, int n, const PanelGeom* geom, const double* ks, int batch,
                       const unsigned short* Wb, const float* bias, const float* Bf, cplx* WG, cplx* WR, int64_t sMat,
                       int* fb_list, int* fb_count, int fb_cap, unsigned long long* status) {
    const int ti = (n + SURR_TILE_I - 1) / SURR_TILE_I, tj = (n + SURR_TILE_J - 1) / SURR_TILE_J;
    for (int b = 0; b < batch; b++) {
        hipMemsetAsync(fb_count, 0, sizeof(int), st);
        surr_table_kernel<<<dim3(ti, tj, 1), 256, 0, st>>>(n, geom, ks + b, Wb, bias, Bf, WG + sMat * b, WR + sMat * b,
                                                           sMat, fb_list, fb_count, fb_cap);
        // the fallback list of item b holds at most n(n+1)/2 <= fb_cap pairs
        const int64_t maxp = (int64_t)n * (n + 1) / 2;
        const int64_t blocks = std::min<int64_t>((maxp + 255) / 256, 2048);
        surr_fallback_kernel<<<(int)blocks, 256, 0, st>>>(n, geom, ks, b, fb_list, fb_count, fb_cap, WG, WR, sMat,
                                                           status);
    }
}

// evaluate the network (MFMA path) at arbitrary (r, v): out[6 m + o], the raw outputs (real
// parts: SURR_SCALE x residual; imaginary parts: SURR_SCALE exp(min(v, SURR_IM_VC)) x residual)
__global__ void __launch_bounds__(256) surr_eval_kernel(int M, const double* __restrict__ rv,
                                                        const unsigned short* __restrict__ Wb,
                                                        const float* __restrict__ bias, const float* __restrict__ Bf,
                                                        float* __restrict__ out) {
    __shared__ unsigned short act[128 * SURR_ACT_LD];
    __shared__ float sB[2 * SURR_NF];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < 2 * SURR_NF) sB[tid] = Bf[tid];
    const int p = tid & 127;
    const int m = blockIdx.x * 128 + p;
    double r = 0.0, v = 0.005;
    if (m < M) { r = rv[2 * m]; v = rv[2 * m + 1]; }
    __syncthreads();
    float u1, u2;
    surr_inputs(r, v, u1, u2);
    const int half = tid >> 7;
    for (int f = half * 32; f < half * 32 + 32; f++) {
        const int ff = f % SURR_NF;
        const float ph = 6.283185307179586f * (sB[2 * ff] * u1 + sB[2 * ff + 1] * u2);
        act[p * SURR_ACT_LD + f] = f2bf(f < SURR_NF ? sinf(ph) : cosf(ph));
    }
    __syncthreads();
    unsigned short* myact = act + (32 * wave) * SURR_ACT_LD;
    const unsigned short* Wl = Wb;
    f32x16 acc[4];
#pragma unroll 1
    for (int l = 0; l < 4; l++) {
        if (l == 0) surr_layer<64, 4>(myact, SURR_ACT_LD, Wl, acc, lane);
        else surr_layer<128, 4>(myact, SURR_ACT_LD, Wl, acc, lane);
        Wl += (l == 0 ? 128 * 64 : 128 * 128);
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int row = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                const int col = t * 32 + (lane & 31);
                float z = acc[t][q] + bias[l * 128 + col];
                z = z / (1.0f + __expf(-z));
                myact[row * SURR_ACT_LD + col] = f2bf(z);
            }
    }
    f32x16 o[1];
    surr_layer<128, 1>(myact, SURR_ACT_LD, Wl, o, lane);
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int row = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
        const int col = lane & 31;
        const int mm = blockIdx.x * 128 + 32 * wave + row;
        if (col < SURR_NOUT && mm < M) out[(int64_t)6 * mm + col] = o[0][q] + bias[4 * 128 + col];
    }
}

void launch_surr_eval(hipStream_t st, int M, const double* rv, const unsigned short* Wb, const float* bias,
                      const float* Bf, float* out) {
    surr_eval_kernel<<<(M + 127) / 128, 256, 0, st>>>(M, rv, Wb, bias, Bf, out);
}

}  // namespace mhx
