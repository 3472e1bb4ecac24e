// Learned surrogate of the exact Guevel-Delhommeau wave term.
//
// The network maps the reduced coordinates (r, v) = (k R_h, k |z + zeta|) to six
// outputs: 100 x the residual (exact - Liang-Wu-Noblesse fit) of Re/Im G, Re/Im dG/dr
// and Re/Im dG/dz.  Architecture: 2 inputs -> 64 Fourier features
// [sin(2 pi B u), cos(2 pi B u)] of u = (log-scaled r, log-scaled v) -> 4 x 128 SiLU (or ReLU)
// -> 6 linear outputs.  The fit it corrects is the reference's GFWu
// (src/green_functions/wu.jl); the exact term is exact_Guevel_Delhommeau.jl:15-59.
//
//  * samples: ngrid^2 points on a log-spaced grid with seeded jitter, targets from the
//    exact FP64 kernel (hquad_pair) and the Wu fit, on the device; imaginary residuals
//    are scaled by exp(min(v, 5)) (the imaginary parts decay like exp(-v));
//  * training: FP32 kernels (tiled GEMMs with fused bias/activation epilogues, split-K
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

// the same mapping at inference, in FP32 with base-2 logarithms (the ratios of logarithms do not
// depend on the base): u differs from the FP64 mapping by a few 1e-8, far below the BF16
// rounding of the Fourier features it feeds
__device__ __forceinline__ void surr_inputs_f32(double r, double v, float& u1, float& u2) {
    const float l2r0 = -9.965784284662087f;            // log2(1e-3)
    const float inv_r = 1.0f / (6.6438706166529995f + 9.965784284662087f);   // 1 / (log2(100 + 1e-3) - log2(1e-3))
    const float l2v0 = -10.965784284662087f;           // log2(5e-4)
    const float inv_v = 1.0f / (5.643856189774724f + 10.965784284662087f);  // 1 / (log2(50) - log2(5e-4))
    const float vv = fmaxf((float)v, 5e-4f);
    u1 = (__builtin_amdgcn_logf((float)r + 1e-3f) - l2r0) * inv_r;
    u2 = (__builtin_amdgcn_logf(vv) - l2v0) * inv_v;
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

// hidden-layer activation of the network (SURR_ACT_SILU: z / (1 + e^-z); SURR_ACT_RELU: max(z, 0))
// and its derivative (ReLU: 0 at z = 0)
template <int ACT> __device__ __forceinline__ float act_f(float z) {
    if (ACT == SURR_ACT_RELU) return fmaxf(z, 0.0f);
    return z / (1.0f + expf(-z));
}
template <int ACT> __device__ __forceinline__ float dact_f(float z) {
    if (ACT == SURR_ACT_RELU) return z > 0.0f ? 1.0f : 0.0f;
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// C[M x N] = sum_k A(m, k) B(k, n).  A(m,k) = TA ? A[k*lda+m] : A[m*lda+k];
// B(k,n) = TB ? B[n*ldb+k] : B[k*ldb+n].  Epilogues:
//  EPI 0: C = acc (split-K partial: C + blockIdx.z * M*N, k range split over gridDim.z)
//  EPI 1: Z = acc + bias[n], H = act(Z)    (C = Z, C2 = H)
//  EPI 2: Z = acc + bias[n]                 (C = Z)
//  EPI 3: C = acc * act'(Zp[m*N+n])         (backward through the activation)
template <bool TA, bool TB, int EPI, int ACT>
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
                C2[o] = act_f<ACT>(z);
            } else if (EPI == 2) {
                C[o] = acc[i][j] + bias[nn];
            } else {
                C[o] = acc[i][j] * dact_f<ACT>(Zp[o]);
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
static void sgemm(hipStream_t st, const SgemmArgs& a, int splits = 1, int act = SURR_ACT_SILU) {
    dim3 grid((a.N + 63) / 64, (a.M + 63) / 64, splits);
    if (act == SURR_ACT_RELU)
        sgemm_kernel<TA, TB, EPI, SURR_ACT_RELU><<<grid, 256, 0, st>>>(a.M, a.N, a.K, a.A, a.lda, a.B, a.ldb, a.C, a.C2,
                                                                       a.bias, a.Zp);
    else
        sgemm_kernel<TA, TB, EPI, SURR_ACT_SILU><<<grid, 256, 0, st>>>(a.M, a.N, a.K, a.A, a.lda, a.B, a.ldb, a.C, a.C2,
                                                                       a.bias, a.Zp);
}

// parameter offsets inside the flat FP32 parameter vector
static const int64_t OFF_W[5] = {0, 128 * 64 + 128, 128 * 64 + 128 + 128 * 128 + 128,
                                 128 * 64 + 128 + 2 * (128 * 128 + 128), 128 * 64 + 128 + 3 * (128 * 128 + 128)};
static const int IN_DIM[5] = {64, 128, 128, 128, 128};
static const int OUT_DIM[5] = {128, 128, 128, 128, SURR_NOUT};

int64_t surr_param_count() { return OFF_W[4] + (int64_t)SURR_NOUT * 128 + SURR_NOUT; }

// one forward + backward pass on M rows (idx or identity), gradients into grad (flat, zeroed here)
void surr_forward_backward(hipStream_t st, SurrTrainWork& w, int M, const int* idx, const float* U, const float* Y,
                           const float* params, float* grad, float* loss, int act) {
    launch_surr_features(st, M, idx, U, w.Bf, w.act[0]);
    const float* hin = w.act[0];
    for (int l = 0; l < 5; l++) {
        const float* W = params + OFF_W[l];
        const float* b = W + (int64_t)OUT_DIM[l] * IN_DIM[l];
        SgemmArgs a{M, OUT_DIM[l], IN_DIM[l], hin, IN_DIM[l], W, IN_DIM[l], w.z[l], w.act[l + 1], b, nullptr};
        if (l < 4) sgemm<false, true, 1>(st, a, 1, act);
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
            // dz[l-1] = (dz[l] W_l) * act'(z[l-1])
            const float* W = params + OFF_W[l];
            SgemmArgs d{M, IN_DIM[l], OUT_DIM[l], w.dz[l], OUT_DIM[l], W, IN_DIM[l], w.dz[l - 1], nullptr, nullptr,
                        w.z[l - 1]};
            sgemm<false, false, 3>(st, d, 1, act);
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
// inference: BF16 MFMA (v_mfma_f32_32x32x16_bf16), weights resident in LDS
// ---------------------------------------------------------------------------
// Each wave evaluates the network on 64 points (two column tiles of 32) in the
// transposed form  H_next^T = W H^T: the weights are the A operand (32 neurons x 16
// inputs per lane group, read from LDS), the activations the B operand (16 inputs x
// 32 points).  With the 32x32 accumulator layout (lane l, register 4g + i holds row
// 8g + 4(l/32) + i, column l%32) the eight accumulator values a lane holds for rows
// 16b + {0..3, 8..11} + 4(l/32) are exactly the eight B-operand inputs that lane needs
// for input block 2t + b of the next layer, so activations stay in registers from
// layer to layer: the packed weights carry the matching input permutation
//     input slot (block kb, lane half h, element e) <-> neuron 16 kb + 8 (e/4) + 4 h + e%4.
// LDS layout of the packed weights (bf16x8 units, one per lane):
//   layer 0  [tile 0..3][kb 0..3][lane 0..63]   inputs: Fourier features
//   layer 1-3 [tile 0..3][kb 0..7][lane]
//   output   [kb 0..7][lane]                    rows 0..5 used, 6..31 zero
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

#define SW_UNITS (16 * 64 + 3 * 32 * 64 + 8 * 64)  // 7680 bf16x8 = 122,880 B
__host__ __device__ constexpr int sw_layer_off(int l) { return l == 0 ? 0 : (l <= 3 ? 16 * 64 + (l - 1) * 32 * 64 : 16 * 64 + 96 * 64); }

struct SurrLds {
    bf16x8 w[SW_UNITS];
    float bias[4 * 128 + 8];
    float fourier[2 * SURR_NF];
};

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

// input index (in the FP32 parameter layout) feeding slot (kb, h, e) of layer l
__device__ __forceinline__ int surr_slot_input(int l, int kb, int h, int e) {
    if (l == 0) {  // feature: sin (e < 4) or cos of frequency 16 h + 4 kb + e % 4
        const int f = 16 * h + 4 * kb + (e & 3);
        return (e < 4 ? 0 : SURR_NF) + f;
    }
    return 16 * kb + 8 * (e >> 2) + 4 * h + (e & 3);
}

// FP32 parameters -> packed BF16 units (layout above) and FP32 biases [l][128] + output[8]
// Inference works on y = -log2(e) z instead of the pre-activation z: the hidden layers'
// biases, and the first layer's weights, carry the factor -log2(e); the activation gives
// h' = -log2(e) act(z) from y (SiLU: y / (1 + 2^y), one v_exp_f32 with no scaling multiply;
// ReLU: min(y, 0)), and the next layer's weights would undo the factor (-1 / log2(e)) while
// also applying their own (-log2(e)): the middle layers' weights are unchanged and only the
// output layer's carry -1 / log2(e).
constexpr float kLog2e = 1.4426950408889634f;
__global__ void surr_pack_kernel(const float* __restrict__ params, unsigned short* __restrict__ Wb,
                                 float* __restrict__ bias) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < SW_UNITS * 8) {
        const int unit = t >> 3, e = t & 7;
        const int lane = unit & 63, blk = unit >> 6;
        int l, tile, kb;
        if (blk < 16) { l = 0; tile = blk / 4; kb = blk % 4; }
        else if (blk < 16 + 96) { l = 1 + (blk - 16) / 32; tile = ((blk - 16) % 32) / 8; kb = (blk - 16) % 8; }
        else { l = 4; tile = 0; kb = blk - 112; }
        const int row = 32 * tile + (lane & 31), h = lane >> 5;
        const int in = surr_slot_input(l, kb, h, e);
        float v = 0.f;
        if (row < OUT_DIM_D(l)) v = params[OFF_W_D(l) + (int64_t)row * IN_DIM_D(l) + in];
        Wb[t] = f2bf(v * (l == 0 ? -kLog2e : l == 4 ? -1.0f / kLog2e : 1.0f));
    }
    if (t < 4 * 128 + 8) {
        const int l = t < 512 ? t / 128 : 4;
        const int r = t - 128 * l;
        float v = 0.f;
        if (r < OUT_DIM_D(l)) v = params[OFF_W_D(l) + (int64_t)OUT_DIM_D(l) * IN_DIM_D(l) + r];
        bias[t] = l < 4 ? -kLog2e * v : v;
    }
}

void launch_surr_pack(hipStream_t st, const float* params, unsigned short* Wb, float* bias) {
    surr_pack_kernel<<<(SW_UNITS * 8 + 255) / 256, 256, 0, st>>>(params, Wb, bias);
}

// cooperative copy of the packed network into LDS (whole block)
__device__ __forceinline__ void surr_load_lds(SurrLds& L, const unsigned short* __restrict__ Wb,
                                              const float* __restrict__ bias, const float* __restrict__ Bf) {
    const bf16x8* src = reinterpret_cast<const bf16x8*>(Wb);
    for (int q = threadIdx.x; q < SW_UNITS; q += blockDim.x) L.w[q] = src[q];
    for (int q = threadIdx.x; q < 4 * 128 + 8; q += blockDim.x) L.bias[q] = bias[q];
    for (int q = threadIdx.x; q < 2 * SURR_NF; q += blockDim.x) L.fourier[q] = Bf[q];
    __syncthreads();
}

// Part of the products of neuron tile t of one layer on the two column tiles: input blocks
// kb0 .. kb1 - 1 of `act` (B operands) into acc0, acc1, the bias put in first when kb0 = 0.
template <int KB>
__device__ __forceinline__ void surr_mfma_part(const bf16x8* W, const float* bias, int t, int kb0, int kb1, int lane,
                                               int h, const bf16x8 act[2][8], f32x16& acc0, f32x16& acc1) {
    if (kb0 == 0) {
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const float4 b = *reinterpret_cast<const float4*>(bias + 32 * t + 8 * g + 4 * h);
            acc0[4 * g + 0] = acc1[4 * g + 0] = b.x;
            acc0[4 * g + 1] = acc1[4 * g + 1] = b.y;
            acc0[4 * g + 2] = acc1[4 * g + 2] = b.z;
            acc0[4 * g + 3] = acc1[4 * g + 3] = b.w;
        }
    }
#pragma unroll
    for (int kb = kb0; kb < kb1; kb++) {
        const bf16x8 a = W[(t * KB + kb) * 64 + lane];
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, act[0][kb], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, act[1][kb], acc1, 0, 0, 0);
    }
}

// -log2(e) act(z) from y = -log2(e) z (SiLU: 2^y overflows to inf for large y, h' -> 0)
template <int ACT> __device__ __forceinline__ float act_log2(float y) {
    if (ACT == SURR_ACT_RELU) return fminf(y, 0.0f);
    return y * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(y));
}

// activation of neuron tile t's accumulators into blocks 2t, 2t + 1 of the next layer's B operands
template <int ACT>
__device__ __forceinline__ void surr_act_tile(int t, const f32x16& acc0, const f32x16& acc1, bf16x8 nxt[2][8]) {
#pragma unroll
    for (int b = 0; b < 2; b++) {
        bf16x8 v0, v1;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int q = 4 * (2 * b + (e >> 2)) + (e & 3);
            v0[e] = (__bf16)act_log2<ACT>(acc0[q]);
            v1[e] = (__bf16)act_log2<ACT>(acc1[q]);
        }
        nxt[0][2 * t + b] = v0;
        nxt[1][2 * t + b] = v1;
    }
}

// Hidden layer `layer` (KB input blocks in X, outputs to Y), software-pipelined over its four
// neuron tiles: the MFMAs of tile t + 1 run between the activation instructions of tile t, and
// the last tile's activations run beside the next layer's first tile on its input blocks 0 .. 5
// (which the activations of tiles 0 .. 2 have finished; blocks 6, 7 follow in the next call).
// On entry a0, a1 hold tile 0 of this layer with input blocks 0 .. kb_done - 1 already added
// (kb_done = 0: none, the bias not yet in); on return they hold the next layer's tile 0 (NKB
// input blocks per tile, blocks 0 .. 5 added) and Y this layer's outputs.
template <int KB, int NKB, int ACT>
__device__ __forceinline__ void surr_layer(const SurrLds& L, int layer, int kb_done, int lane, const bf16x8 X[2][8],
                                           bf16x8 Y[2][8], f32x16& a0, f32x16& a1, f32x16& b0, f32x16& b1) {
    const int h = lane >> 5;
    const bf16x8* W = L.w + sw_layer_off(layer);
    const float* bias = L.bias + 128 * layer;
    const bf16x8* Wn = L.w + sw_layer_off(layer + 1);
    const float* bn = L.bias + 128 * (layer + 1);
    // scheduling regions; the barriers keep a later region's LDS loads from being hoisted, which
    // would raise the register pressure
    __builtin_amdgcn_sched_barrier(0);
    surr_mfma_part<KB>(W, bias, 0, kb_done, KB, lane, h, X, a0, a1);
    __builtin_amdgcn_sched_barrier(0);
    surr_mfma_part<KB>(W, bias, 1, 0, KB, lane, h, X, b0, b1);
    surr_act_tile<ACT>(0, a0, a1, Y);
    __builtin_amdgcn_sched_barrier(0);
    surr_mfma_part<KB>(W, bias, 2, 0, KB, lane, h, X, a0, a1);
    surr_act_tile<ACT>(1, b0, b1, Y);
    __builtin_amdgcn_sched_barrier(0);
    surr_mfma_part<KB>(W, bias, 3, 0, KB, lane, h, X, b0, b1);
    surr_act_tile<ACT>(2, a0, a1, Y);
    __builtin_amdgcn_sched_barrier(0);
    surr_mfma_part<NKB>(Wn, bn, 0, 0, 6, lane, h, Y, a0, a1);
    surr_act_tile<ACT>(3, b0, b1, Y);
}

// the network on the lane's two points (u of column tile 0 and 1); out[pt] holds the
// output rows: lanes 0-31 outputs 0..3 (registers 0..3), lanes 32-63 outputs 4, 5 (registers 0, 1)
template <int ACT>
__device__ __forceinline__ void surr_mlp64(const SurrLds& L, int lane, const float u1[2], const float u2[2],
                                           f32x16 out[2]) {
    const int h = lane >> 5;
    bf16x8 P[2][8], Q[2][8];
#pragma unroll
    for (int pt = 0; pt < 2; pt++)
#pragma unroll
        for (int kb = 0; kb < 4; kb++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int f = 16 * h + 4 * kb + e;
                // v_sin/v_cos take revolutions: sin(2 pi x)
                const float x = L.fourier[2 * f] * u1[pt] + L.fourier[2 * f + 1] * u2[pt];
                P[pt][kb][e] = (__bf16)__builtin_amdgcn_sinf(x);
                P[pt][kb][4 + e] = (__bf16)__builtin_amdgcn_cosf(x);
            }
    f32x16 a0, a1, b0, b1;
    surr_layer<4, 8, ACT>(L, 0, 0, lane, P, Q, a0, a1, b0, b1);  // features P -> Q
    surr_layer<8, 8, ACT>(L, 1, 6, lane, Q, P, a0, a1, b0, b1);  // Q -> P
    surr_layer<8, 8, ACT>(L, 2, 6, lane, P, Q, a0, a1, b0, b1);  // P -> Q
    // layer 3 (Q -> P), its last tile's activations beside the output layer's blocks 0 .. 5; the
    // output layer's bias: rows 0..3 on lanes 0-31, rows 4, 5 on lanes 32-63 (the rest padding)
    {
        const bf16x8* W = L.w + sw_layer_off(3);
        const float* bias = L.bias + 128 * 3;
        const bf16x8* Wo = L.w + sw_layer_off(4);
        __builtin_amdgcn_sched_barrier(0);
        surr_mfma_part<8>(W, bias, 0, 6, 8, lane, h, Q, a0, a1);
        __builtin_amdgcn_sched_barrier(0);
        surr_mfma_part<8>(W, bias, 1, 0, 8, lane, h, Q, b0, b1);
        surr_act_tile<ACT>(0, a0, a1, P);
        __builtin_amdgcn_sched_barrier(0);
        surr_mfma_part<8>(W, bias, 2, 0, 8, lane, h, Q, a0, a1);
        surr_act_tile<ACT>(1, b0, b1, P);
        __builtin_amdgcn_sched_barrier(0);
        surr_mfma_part<8>(W, bias, 3, 0, 8, lane, h, Q, b0, b1);
        surr_act_tile<ACT>(2, a0, a1, P);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int pt = 0; pt < 2; pt++)
#pragma unroll
            for (int q = 0; q < 16; q++) out[pt][q] = 0.f;
        if (h == 0) {
#pragma unroll
            for (int q = 0; q < 4; q++) out[0][q] = out[1][q] = L.bias[512 + q];
        } else {
            out[0][0] = out[1][0] = L.bias[512 + 4];
            out[0][1] = out[1][1] = L.bias[512 + 5];
        }
#pragma unroll
        for (int kb = 0; kb < 6; kb++) {
            const bf16x8 a = Wo[kb * 64 + lane];
            out[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, P[0][kb], out[0], 0, 0, 0);
            out[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, P[1][kb], out[1], 0, 0, 0);
        }
        surr_act_tile<ACT>(3, b0, b1, P);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int kb = 6; kb < 8; kb++) {
            const bf16x8 a = Wo[kb * 64 + lane];
            out[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, P[0][kb], out[0], 0, 0, 0);
            out[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, P[1][kb], out[1], 0, 0, 0);
        }
    }
}

// triangular tile index -> (bi, bj), bi <= bj, t = bj (bj + 1) / 2 + bi
__device__ __forceinline__ void tri_tile(int64_t t, int& bi, int& bj) {
    int b = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((int64_t)(b + 1) * (b + 2) / 2 <= t) b++;
    while ((int64_t)b * (b + 1) / 2 > t) b--;
    bj = b;
    bi = (int)(t - (int64_t)b * (b + 1) / 2);
}

// unit index within an item -> (bi, bj): the units come in fours, the 2 x 2 blocks of one 16 x 16
// super-tile (BI <= BJ, in triangular order; bit 0 of t selects bi, bit 1 bj), so that the four units
// whose 64-byte column segments share 128-byte lines of the column-major outputs (the FP32 copy of D,
// the Rankine parts read) run in the same workgroup, on one XCD's L2.  False for the units below the
// diagonal (bi > bj) or past the matrix: the wave skips them.
__device__ __forceinline__ bool quad_tile(int64_t t, int nb, int& bi, int& bj) {
    int BI, BJ;
    tri_tile(t >> 2, BI, BJ);
    bi = 2 * BI + (int)(t & 1);
    bj = 2 * BJ + (int)((t >> 1) & 1);
    return bi <= bj && bj < nb;
}

// units per item of order n: 4 per 16 x 16 super-tile of the upper triangle (a multiple of 4)
__host__ __device__ __forceinline__ int64_t surr_item_units(int n) {
    const int64_t nsb = ((n + 7) / 8 + 1) / 2;
    return 2 * nsb * (nsb + 1);
}

#define SURR_WAVES 8

// destination of one item's pair table, pointers offset to the item: the (G, dG/dr) table (S
// null), or the finished S and D with their Rankine parts SR, DR (and the FP32 copy Df of D when
// set)
struct SurrDst {
    cplx *WG, *WR;
    const double *SR, *DR;
    cplx *S, *D;
    cplxf* Df;
    int direct, include_identity;
    int sd32;  // S and D as FP32 complex at the item's slots
};

__device__ __forceinline__ SurrDst surr_dst(const SurrOut& o, int item) {
    SurrDst d;
    d.WG = o.WG ? o.WG + o.sMat * item : nullptr;
    d.WR = o.WR ? o.WR + o.sMat * item : nullptr;
    d.SR = o.SR;
    d.DR = o.DR;
    d.S = o.S ? o.S + o.sMat * item : nullptr;
    d.D = o.D ? o.D + o.sMat * item : nullptr;
    d.Df = o.Df ? o.Df + o.sMat * item : nullptr;
    d.direct = o.direct;
    d.include_identity = o.include_identity;
    d.sd32 = o.sd32;
    return d;
}

__device__ __forceinline__ SurrDst surr_dst(const SweepItem& it, int direct, int include_identity) {
    SurrDst d;
    d.WG = d.WR = nullptr;
    d.SR = it.SR;
    d.DR = it.DR;
    d.S = it.S;
    d.D = it.D;
    d.Df = it.Df;
    d.direct = direct;
    d.include_identity = include_identity;
    d.sd32 = it.sd32;
    return d;
}

typedef double nt_double2 __attribute__((ext_vector_type(2)));
typedef float nt_float2 __attribute__((ext_vector_type(2)));

// non-temporal (streaming) stores of the outputs: written once and read by the next kernel, they
// would otherwise evict the panel geometry and the Rankine lines that the units re-read from L2
__device__ __forceinline__ void store_stream(cplx* p, cplx v) {
    __builtin_nontemporal_store(nt_double2{v.x, v.y}, reinterpret_cast<nt_double2*>(p));
}
__device__ __forceinline__ void store_stream(cplxf* p, double re, double im) {
    __builtin_nontemporal_store(nt_float2{(float)re, (float)im}, reinterpret_cast<nt_float2*>(p));
}

// store one entry (row r, column c) of the item's output from the pair's wave term: the (G, dG/dr)
// table, or (fused) the S and D entries with their Rankine parts, and the FP32 copy of D for the
// mixed-precision solver
__device__ __forceinline__ void surr_store_entry(const SurrDst& o, int n, int r, int c, double k, PairQ pq,
                                                 const PanelGeom* __restrict__ geom, cplx G, cplx Gr) {
    const int64_t rc = r + (int64_t)n * c;
    if (!o.S) {
        store_stream(o.WG + rc, G);
        store_stream(o.WR + rc, Gr);
        return;
    }
    const PanelGeom& gr = geom[r];
    cplx Sv, Dv;
    combine_entry(k, pq, gr.c, gr.nrm, geom[c], o.direct, o.include_identity && r == c, G, Gr, o.SR[rc],
                  o.DR[rc], Sv, Dv);
    if (o.sd32) {
        store_stream(reinterpret_cast<cplxf*>(o.S) + rc, Sv.x, Sv.y);
        store_stream(reinterpret_cast<cplxf*>(o.D) + rc, Dv.x, Dv.y);
    } else {
        store_stream(o.S + rc, Sv);
        store_stream(o.D + rc, Dv);
    }
    if (o.Df) store_stream(o.Df + rc, Dv.x, Dv.y);
}

// both entries (i, j) and (j, i) of a pair from one thread (the exact fallback)
__device__ __forceinline__ void surr_store_pair(const SurrDst& o, int n, int i, int j, double k, double hh, double vv,
                                                const PanelGeom* __restrict__ geom, cplx G, cplx Gr) {
    const PairQ pq = pair_q(k, hh, vv);
    surr_store_entry(o, n, i, j, k, pq, geom, G, Gr);
    if (i != j) surr_store_entry(o, n, j, i, k, pq, geom, G, Gr);
}

// One work unit of a wave: the 8 x 8 block of panel pairs (i in 8 bi.., j in 8 bj.., bi <= bj;
// pairs with i > j are skipped) of one item (wavenumber k).  Pairs inside the trained box get Wu +
// network; the others are appended to the fallback list as (tag | i << 20 | j), tag = item << 40.
template <int ACT>
__device__ __forceinline__ void surr_unit(const SurrLds& L, int lane, int n, int bi, int bj, double k,
                                          const PanelGeom* __restrict__ geom, const SurrDst& dst,
                                          unsigned long long tag, unsigned long long* __restrict__ fb_list,
                                          unsigned long long* __restrict__ fb_count, int64_t fb_cap) {
    const int h = lane >> 5, c = lane & 31;
    // the lane's own pair 32 h + c; column tile pt of the network holds pairs 32 pt + c, so a
    // lane takes the other tile's inputs from its partner lane (lane ^ 32)
    double hh_me = 0.0, vv_me = 0.0;
    bool valid_me;
    float u1[2], u2[2];
    {
        const int p = 32 * h + c;
        const int i = 8 * bi + (p & 7), j = 8 * bj + (p >> 3);
        valid_me = i < n && j < n && i <= j;
        if (valid_me) {
            const PanelGeom& gi = geom[i];
            const PanelGeom& gj = geom[j];
            reduced_hv(gi.c[0], gi.c[1], gi.c[2], gj.c[0], gj.c[1], gj.c[2], k, hh_me, vv_me);
        }
        float a1, a2;
        surr_inputs_f32(hh_me, -vv_me, a1, a2);
        const float b1 = __shfl_xor(a1, 32), b2 = __shfl_xor(a2, 32);
        u1[0] = h ? b1 : a1;
        u1[1] = h ? a1 : b1;
        u2[0] = h ? b2 : a2;
        u2[1] = h ? a2 : b2;
    }
    f32x16 out[2];
    surr_mlp64<ACT>(L, lane, u1, u2, out);
    // outputs 0..3 of this lane's pair (32 h + c): own column tile 0, or lane c's tile 1
    float o[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float other = __shfl_xor(out[1][q], 32);
        o[q] = h == 0 ? out[0][q] : other;
    }
    const int p = 32 * h + c;
    const int i = 8 * bi + (p & 7), j = 8 * bj + (p >> 3);
    const bool net = valid_me && surr_in_box(hh_me, -vv_me);
    const bool fb = valid_me && !net;
    // wave-aggregated append to the fallback list
    const unsigned long long mask = __ballot(fb);
    if (mask) {
        unsigned long long base = 0;
        const int leader = __ffsll((long long)mask) - 1;
        if (lane == leader) base = atomicAdd(fb_count, (unsigned long long)__popcll(mask));
        base = __shfl(base, leader);
        if (fb) {
            // fallback lanes below this one: mbcnt counts the mask's bits under the lane
            const unsigned below = __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
            const unsigned long long slot = base + below;
            if (slot < (unsigned long long)fb_cap) fb_list[slot] = tag | ((unsigned long long)i << 20) | (unsigned)j;
        }
    }
    if (!__ballot(net)) return;
    cplx G = cmk(0.0, 0.0), Gr = cmk(0.0, 0.0);
    PairQ pq = {0.0, 0.0};
    if (net) {
        const WuGh w = wu_term_gh_f32sf(hh_me > 0 ? hh_me : 1e-300, vv_me);
        const double is = 1.0 / SURR_SCALE;
        // the imaginary outputs' scale in FP32 (the network's outputs carry BF16 rounding; an FP64
        // exp here costs constants that the compiler keeps in registers across the unit loop)
        const double ii = is * (double)expf(-fminf((float)-vv_me, (float)SURR_IM_VC));
        G = cmk(w.G.x + is * o[0], w.G.y + ii * o[1]);
        Gr = cmk(w.Gh.x + is * o[2], w.Gh.y + ii * o[3]);
        // entry (i, j): lanes 8m .. 8m + 7 write rows 8 bi .. 8 bi + 7 of column 8 bj + m
        pq = pair_q(k, hh_me, vv_me);
        surr_store_entry(dst, n, i, j, k, pq, geom, G, Gr);
    }
    // entry (j, i) from the transposed lane t, whose pair is (8 bi + lane / 8, 8 bj + lane % 8): lanes
    // 8m .. 8m + 7 again write eight consecutive rows of one column (hh and vv are symmetric in the pair)
    const int t = ((lane & 7) << 3) | (lane >> 3);
    const PairQ pq_t = {__shfl(pq.inv, t), __shfl(pq.kkh, t)};
    const cplx G_t = cmk(__shfl(G.x, t), __shfl(G.y, t)), Gr_t = cmk(__shfl(Gr.x, t), __shfl(Gr.y, t));
    const bool net_t = __shfl((int)net, t) != 0;
    const int mi = 8 * bi + (lane >> 3), mj = 8 * bj + (lane & 7);
    if (net_t && mi != mj) surr_store_entry(dst, n, mj, mi, k, pq_t, geom, G_t, Gr_t);
}

// Persistent pair-table kernel over all items of the batch: each wave takes units in a strided
// order.
template <int ACT>
__global__ void __launch_bounds__(64 * SURR_WAVES) surr_table_kernel(
    int n, int batch, const PanelGeom* __restrict__ geom, const double* __restrict__ ks,
    const unsigned short* __restrict__ Wb, const float* __restrict__ bias, const float* __restrict__ Bf, SurrOut out,
    unsigned long long* __restrict__ fb_list, unsigned long long* __restrict__ fb_count, int64_t fb_cap) {
    __shared__ SurrLds L;
    surr_load_lds(L, Wb, bias, Bf);
    const int lane = threadIdx.x & 63;
    const int nb = (n + 7) / 8;
    const int64_t per_item = surr_item_units(n);
    const int64_t units = per_item * batch;
    const int64_t stride = (int64_t)gridDim.x * SURR_WAVES;
    // the wave index through readfirstlane: the unit, its item and tile live in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    for (int64_t unit = (int64_t)blockIdx.x * SURR_WAVES + wave; unit < units; unit += stride) {
        const int item = (int)(unit / per_item);
        int bi, bj;
        if (!quad_tile(unit - (int64_t)item * per_item, nb, bi, bj)) continue;
        surr_unit<ACT>(L, lane, n, bi, bj, ks[item], geom, surr_dst(out, item), (unsigned long long)item << 40, fb_list,
                  fb_count, fb_cap);
    }
}

// the same over a queue of items of different bodies (SweepItem::sunit0: an item's first unit)
template <int ACT>
__global__ void __launch_bounds__(64 * SURR_WAVES) surr_queue_kernel(
    const SweepItem* __restrict__ items, int nitems, int64_t units, int direct, int include_identity,
    const unsigned short* __restrict__ Wb, const float* __restrict__ bias, const float* __restrict__ Bf,
    unsigned long long* __restrict__ fb_list, unsigned long long* __restrict__ fb_count, int64_t fb_cap) {
    __shared__ SurrLds L;
    surr_load_lds(L, Wb, bias, Bf);
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * SURR_WAVES;
    // the wave index through readfirstlane: the unit, its item and tile live in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    for (int64_t unit = (int64_t)blockIdx.x * SURR_WAVES + wave; unit < units; unit += stride) {
        int lo = 0, hi = nitems - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (items[mid].sunit0 <= unit) lo = mid;
            else hi = mid - 1;
        }
        const SweepItem& it = items[lo];
        int bi, bj;
        if (!quad_tile(unit - it.sunit0, (it.n + 7) / 8, bi, bj)) continue;
        surr_unit<ACT>(L, lane, it.n, bi, bj, *it.k, it.geom, surr_dst(it, direct, include_identity),
                  (unsigned long long)lo << 40, fb_list, fb_count, fb_cap);
    }
}

// exact term for the listed fallback pairs (grid-stride: the count is only known on the device)
__device__ __forceinline__ void surr_exact_pair(const SurrDst& dst, int n, const PanelGeom* __restrict__ geom,
                                                double k, int i, int j, unsigned long long* __restrict__ status) {
    const PanelGeom& gi = geom[i];
    const PanelGeom& gj = geom[j];
    double hh, vv;
    reduced_hv(gi.c[0], gi.c[1], gi.c[2], gj.c[0], gj.c[1], gj.c[2], k, hh, vv);
    QuadWork wk;
    cplx G, Gr;
    if (!hquad_pair(hh, vv, 1e-8, 1e-8, G, Gr, wk)) atomicOr(status, 1ull);
    G = cscale(4.0, G);
    Gr = cscale(4.0, Gr);
    surr_store_pair(dst, n, i, j, k, hh, vv, geom, G, Gr);
}

__global__ void surr_fallback_kernel(int n, const PanelGeom* __restrict__ geom, const double* __restrict__ ks,
                                     const unsigned long long* __restrict__ fb_list,
                                     const unsigned long long* __restrict__ fb_count, int64_t fb_cap, SurrOut out,
                                     unsigned long long* __restrict__ status) {
    const int64_t cnt = (int64_t)min(*fb_count, (unsigned long long)fb_cap);
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < cnt; t += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long e = fb_list[t];
        const int item = (int)(e >> 40), i = (int)((e >> 20) & 0xFFFFF), j = (int)(e & 0xFFFFF);
        surr_exact_pair(surr_dst(out, item), n, geom, ks[item], i, j, status);
    }
}

__global__ void surr_queue_fallback_kernel(const SweepItem* __restrict__ items, int direct, int include_identity,
                                           const unsigned long long* __restrict__ fb_list,
                                           const unsigned long long* __restrict__ fb_count, int64_t fb_cap,
                                           unsigned long long* __restrict__ status) {
    const int64_t cnt = (int64_t)min(*fb_count, (unsigned long long)fb_cap);
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < cnt; t += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long e = fb_list[t];
        const SweepItem& it = items[(int)(e >> 40)];
        const int i = (int)((e >> 20) & 0xFFFFF), j = (int)(e & 0xFFFFF);
        surr_exact_pair(surr_dst(it, direct, include_identity), it.n, it.geom, *it.k, i, j, status);
    }
}

static int device_cus() {
    int dev = 0, cus = 0;
    hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    return cus;
}

int64_t surr_fallback_capacity(int n, int batch) { return (int64_t)n * (n + 1) / 2 * batch; }

void launch_surr_table(hipStream_t st, int n, const PanelGeom* geom, const double* ks, int batch,
                       const unsigned short* Wb, const float* bias, const float* Bf, const SurrOut& out,
                       unsigned long long* fb_list, unsigned long long* fb_count, int64_t fb_cap,
                       unsigned long long* status, int act) {
    launch_zero(st, fb_count, sizeof(unsigned long long));
    const int64_t units = surr_item_units(n) * batch;
    const int64_t blocks = std::min<int64_t>(device_cus(), (units + SURR_WAVES - 1) / SURR_WAVES);
    if (act == SURR_ACT_RELU)
        surr_table_kernel<SURR_ACT_RELU><<<(int)blocks, 64 * SURR_WAVES, 0, st>>>(n, batch, geom, ks, Wb, bias, Bf, out,
                                                                                 fb_list, fb_count, fb_cap);
    else
        surr_table_kernel<SURR_ACT_SILU><<<(int)blocks, 64 * SURR_WAVES, 0, st>>>(n, batch, geom, ks, Wb, bias, Bf, out,
                                                                                 fb_list, fb_count, fb_cap);
    surr_fallback_kernel<<<2048, 256, 0, st>>>(n, geom, ks, fb_list, fb_count, fb_cap, out, status);
}

int64_t sweep_surr_units(int n) { return surr_item_units(n); }

void launch_queue_surr(hipStream_t st, const SweepItem* items, int nitems, int64_t total_units, int direct,
                       int include_identity, const unsigned short* Wb, const float* bias, const float* Bf,
                       unsigned long long* fb_list, unsigned long long* fb_count, int64_t fb_cap,
                       unsigned long long* status, int act) {
    launch_zero(st, fb_count, sizeof(unsigned long long));
    if (total_units <= 0) return;
    const int64_t blocks = std::min<int64_t>(device_cus(), (total_units + SURR_WAVES - 1) / SURR_WAVES);
    if (act == SURR_ACT_RELU)
        surr_queue_kernel<SURR_ACT_RELU><<<(int)blocks, 64 * SURR_WAVES, 0, st>>>(
            items, nitems, total_units, direct, include_identity, Wb, bias, Bf, fb_list, fb_count, fb_cap);
    else
        surr_queue_kernel<SURR_ACT_SILU><<<(int)blocks, 64 * SURR_WAVES, 0, st>>>(
            items, nitems, total_units, direct, include_identity, Wb, bias, Bf, fb_list, fb_count, fb_cap);
    surr_queue_fallback_kernel<<<2048, 256, 0, st>>>(items, direct, include_identity, fb_list, fb_count, fb_cap, status);
}

// the network (MFMA path) at arbitrary (r, v): out[6 m + o], the raw outputs (real parts:
// SURR_SCALE x residual; imaginary parts: SURR_SCALE exp(min(v, SURR_IM_VC)) x residual)
template <int ACT>
__global__ void __launch_bounds__(64 * SURR_WAVES) surr_eval_kernel(int M, const double* __restrict__ rv,
                                                                    const unsigned short* __restrict__ Wb,
                                                                    const float* __restrict__ bias,
                                                                    const float* __restrict__ Bf,
                                                                    float* __restrict__ out) {
    __shared__ SurrLds L;
    surr_load_lds(L, Wb, bias, Bf);
    const int lane = threadIdx.x & 63, h = lane >> 5, c = lane & 31;
    const int units = (M + 63) / 64;
    for (int unit = blockIdx.x * SURR_WAVES + (threadIdx.x >> 6); unit < units; unit += gridDim.x * SURR_WAVES) {
        float u1[2], u2[2];
#pragma unroll
        for (int pt = 0; pt < 2; pt++) {
            const int m = 64 * unit + 32 * pt + c;
            double r = 0.0, v = 0.005;
            if (m < M) { r = rv[2 * m]; v = rv[2 * m + 1]; }
            surr_inputs_f32(r, v, u1[pt], u2[pt]);
        }
        f32x16 o[2];
        surr_mlp64<ACT>(L, lane, u1, u2, o);
#pragma unroll
        for (int pt = 0; pt < 2; pt++) {
            const int m = 64 * unit + 32 * pt + c;
            if (m >= M) continue;
            if (h == 0) {
#pragma unroll
                for (int q = 0; q < 4; q++) out[(int64_t)6 * m + q] = o[pt][q];
            } else {
                out[(int64_t)6 * m + 4] = o[pt][0];
                out[(int64_t)6 * m + 5] = o[pt][1];
            }
        }
    }
}

void launch_surr_eval(hipStream_t st, int M, const double* rv, const unsigned short* Wb, const float* bias,
                      const float* Bf, float* out, int act) {
    const int units = (M + 63) / 64;
    const int blocks = std::max(1, std::min(device_cus(), (units + SURR_WAVES - 1) / SURR_WAVES));
    if (act == SURR_ACT_RELU)
        surr_eval_kernel<SURR_ACT_RELU><<<blocks, 64 * SURR_WAVES, 0, st>>>(M, rv, Wb, bias, Bf, out);
    else
        surr_eval_kernel<SURR_ACT_SILU><<<blocks, 64 * SURR_WAVES, 0, st>>>(M, rv, Wb, bias, Bf, out);
}

}  // namespace mhx
