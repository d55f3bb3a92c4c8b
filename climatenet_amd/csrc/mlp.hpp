// Contact MLP shapes, parameter layout and the device view used by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pdfc.h"

namespace pdfc {

constexpr int kMlpFeatures = 12;  // per-contact inputs
constexpr int kMlpIn = 16;        // padded input width (one MFMA k-block)
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 3;        // impulse guess in the contact frame
constexpr int kMlpNorm = 2 * kMlpFeatures + kMlpOut;  // x mean, x std, y scale

// FP32 master parameters, flat: W1 (16 x 64) b1 | W2 (64 x 64) b2 | W3 (64 x 64) b3 | W4 (64 x 16) b4 (16).
// Rows 12..15 of W1 and columns 3..15 of W4 are padding and stay zero.
__host__ __device__ constexpr int mlp_k(int l) { return l == 0 ? kMlpIn : kMlpHidden; }
__host__ __device__ constexpr int mlp_n(int l) { return l == 3 ? 16 : kMlpHidden; }
__host__ __device__ constexpr int mlp_off_w(int l) {
    return l == 0 ? 0 : (l == 1 ? 16 * 64 + 64 : (l == 2 ? 16 * 64 + 64 + 64 * 64 + 64 : 16 * 64 + 64 + 2 * (64 * 64 + 64)));
}
__host__ __device__ constexpr int mlp_off_b(int l) { return mlp_off_w(l) + mlp_k(l) * mlp_n(l); }
constexpr int kMlpParams = mlp_off_b(3) + 16;  // 10448
// BF16 operand copies per layer: W (K x N) then W^T (N x K).
__host__ __device__ constexpr int mlp_bf_size(int l) { return mlp_k(l) * mlp_n(l); }
__host__ __device__ constexpr int mlp_bf_off(int l) {
    return l == 0 ? 0 : mlp_bf_off(l - 1) + 2 * mlp_bf_size(l - 1);
}
constexpr int kBfTotal = mlp_bf_off(3) + 2 * mlp_bf_size(3);

struct MlpDev {
    const float* b[4];
    unsigned short* w_bf[4];
    unsigned short* wt_bf[4];
    const float* x_mean;
    const float* x_std;
    const float* y_scale;
};

static_assert(kMlpParams == PDFC_MLP_PARAMS, "parameter count of the C ABI");
static_assert(kMlpNorm == PDFC_MLP_NORM, "normalisation size of the C ABI");
static_assert(kMlpFeatures == PDFC_MLP_FEATURES && kMlpOut == PDFC_MLP_OUTPUTS, "shapes of the C ABI");

// y[rows x 3] = y_scale * MLP((x - mean) / std), one wave per 16 rows, on stream st.
void launch_mlp_add_previous(const float* feat, int rows, float* y, hipStream_t st);
void launch_mlp_forward(const MlpDev& P, const float* x, int rows, float* y, hipStream_t st);
// Message of the last failure, read by pdfc_last_error() (defined with the simulation).
void set_last_error(const char* msg);

}  // namespace pdfc

const pdfc::MlpDev* pdfc_mlp_device(const pdfc_mlp* M);
