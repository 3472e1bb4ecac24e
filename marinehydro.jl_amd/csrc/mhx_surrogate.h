// Learned surrogate of the exact wave term: shared declarations.
#pragma once
#include "mhx_internal.h"

#define SURR_NF 32          // Fourier frequencies (64 features)
#define SURR_ACT_SILU 0     // hidden-layer activations (MHX_SURR_SILU / MHX_SURR_RELU)
#define SURR_ACT_RELU 1
#define SURR_NOUT 6         // Re/Im of dG, dG/dr, dG/dz residuals
#define SURR_SCALE 100.0    // target = SURR_SCALE * (exact - Wu)
#define SURR_RMAX 100.0     // trained box: r = k R_h in [0, 100]
#define SURR_VMAX 50.0      //              v = k |z + zeta| in [SURR_VMIN, 50]
#define SURR_VMIN 1e-3      // below: free-surface region, exact fallback
#define SURR_IM_VC 5.0      // imaginary outputs are residuals relative to exp(-min(v, SURR_IM_VC))
#define SURR_VAL_STRIDE 64  // every 64th sample is held out for validation
#define SURR_PACKED_W (128 * 64 + 3 * 128 * 128 + 32 * 128)  // bf16 weights, output layer padded to 32 rows
#define SURR_BIAS (4 * 128 + 8)

namespace mhx {

struct SurrTrainWork {
    int M = 0;
    float* act[6] = {};  // features (M x 64), hidden outputs (M x 128), output (M x 6)
    float* z[5] = {};    // pre-activations
    float* dz[5] = {};
    float* part = nullptr;  // split-K partials (32 x 128 x 128)
    const float* Bf = nullptr;
};

int64_t surr_param_count();
void launch_surr_samples(hipStream_t st, int ngrid, uint64_t seed, float* U, float* Y, unsigned long long* status);
void launch_surr_features(hipStream_t st, int M, const int* idx, const float* U, const float* Bf, float* F);
void surr_forward_backward(hipStream_t st, SurrTrainWork& w, int M, const int* idx, const float* U, const float* Y,
                           const float* params, float* grad, float* loss, int act);
void surr_adam(hipStream_t st, int64_t n, float* p, const float* g, float* m1, float* m2, float lr, int step);
void launch_batch_index(hipStream_t st, int M, int64_t nsamples, uint64_t seed, int* idx);
void launch_surr_pack(hipStream_t st, const float* params, unsigned short* Wb, float* bias);
int64_t surr_fallback_capacity(int n, int batch);
// Destination of the surrogate pair table: with S null, the symmetric (G, dG/dr) table WG, WR
// (combined by launch_combine_table); otherwise the finished S and D (and, when Df is set, the
// FP32 copy of D) from the Rankine parts SR, DR.  Strides sMat per item.
struct SurrOut {
    cplx *WG = nullptr, *WR = nullptr;
    const double *SR = nullptr, *DR = nullptr;
    cplx *S = nullptr, *D = nullptr;
    cplxf* Df = nullptr;
    int direct = 1, include_identity = 1;
    int64_t sMat = 0;
    // S and D stored as FP32 complex: entry rc of item q at ((cplxf*)(S + sMat q))[rc] (the first half
    // of the item's FP64 slot)
    int sd32 = 0;
};
void launch_surr_table(hipStream_t st, int n, const PanelGeom* geom, const double* ks, int batch,
                       const unsigned short* Wb, const float* bias, const float* Bf, const SurrOut& out,
                       unsigned long long* fb_list, unsigned long long* fb_count, int64_t fb_cap,
                       unsigned long long* status, int act);
void launch_surr_eval(hipStream_t st, int M, const double* rv, const unsigned short* Wb, const float* bias,
                      const float* Bf, float* out, int act);
// the fused pair kernel over a queue of items of several bodies (S, D and, when SweepItem::Df is
// set, the FP32 copy of D), units of 8 x 8 pair blocks counted from SweepItem::sunit0; fallback
// pairs (outside the trained box) get the exact term; fb_count is zeroed here
int64_t sweep_surr_units(int n);
void launch_queue_surr(hipStream_t st, const SweepItem* items, int nitems, int64_t total_units, int direct,
                       int include_identity, const unsigned short* Wb, const float* bias, const float* Bf,
                       unsigned long long* fb_list, unsigned long long* fb_count, int64_t fb_cap,
                       unsigned long long* status, int act);

}  // namespace mhx
