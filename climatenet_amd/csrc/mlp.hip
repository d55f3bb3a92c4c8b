// Contact MLP for the learned warm start: 12 -> 64 -> 64 -> 64 -> 3, SiLU,
// BF16 operands on MFMA with FP32 accumulation: gfx950's v_mfma_f32_16x16x32_bf16 for the
// 64-deep products (the hidden and output layers, the gradients through 64-wide layers), and
// v_mfma_f32_16x16x16_bf16 where the depth is 16 (the 12 input features padded to 16, the
// weight gradients summed over a 16-row tile, the gradient through the 16-wide padded output).
//
// One wave handles a tile of 16 samples through all layers; activations live
// in LDS, weight operands are read from global memory (10 K parameters, cache
// resident). Training fuses forward, MSE gradient and backward per tile, writes
// per-wave partial weight gradients, sums them in a fixed order and applies
// Adam to FP32 master weights, from which the BF16 operand copies are refreshed.
//
// Operand placement: A (16 x K tile) lane l holds row l%16 and B (K x 16) lane l
// column l%16, both for k = 4*(l/16)..+3 (16-deep) or k = 8*(l/16)..+7 (32-deep); A and B
// take the same k of a lane, so the product does not depend on how the instruction orders
// its k. D lane l holds column l%16, rows 4*(l/16)..+3 (both forms).
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "mlp.hpp"

namespace pdfc {
namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 bit pattern, round to nearest even (gfx950's v_cvt_pk_bf16_f32)
__device__ __forceinline__ unsigned short to_bf16(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }

// sigmoid through the hardware exponential and reciprocal (v_exp_f32, v_rcp_f32: 1 ulp each)
__device__ __forceinline__ float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float dsilu(float z) {
    float s = sigmoid(z);
    return s * (1.f + z * (1.f - s));
}

__device__ __forceinline__ f4 mfma32(bf8 a, bf8 b, f4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma16(s4 a, s4 b, f4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
}

// D(16 x 16 tile n0) += A(16 x K, bf16 row-major in LDS, leading dim lda) * B, where the
// B operand for column n comes from Bt (N x K row-major bf16, leading dim K). K is 16 or 64:
// 32-deep steps while they fit, then a 16-deep one.
__device__ __forceinline__ f4 mfma_tile(const unsigned short* A, int lda, const unsigned short* Bt, int K, int n0,
                                        f4 acc) {
    const int l = threadIdx.x & 63;
    const int r = l & 15;
    int k0 = 0;
    for (; k0 + 32 <= K; k0 += 32) {
        // 8 consecutive bf16 (16-byte aligned: lda, K, k0 and kq are multiples of 8)
        const int kq = (l >> 4) * 8;
        const bf8 a = *reinterpret_cast<const bf8*>(A + r * lda + k0 + kq);
        const bf8 b = *reinterpret_cast<const bf8*>(Bt + (size_t)(n0 + r) * K + k0 + kq);
        acc = mfma32(a, b, acc);
    }
    if (k0 < K) {
        const int kq = (l >> 4) * 4;
        const s4 a = *reinterpret_cast<const s4*>(A + r * lda + k0 + kq);
        const s4 b = *reinterpret_cast<const s4*>(Bt + (size_t)(n0 + r) * K + k0 + kq);
        acc = mfma16(a, b, acc);
    }
    return acc;
}

struct alignas(16) Tile {
    unsigned short X[16 * 16];      // normalised input, bf16
    unsigned short H[3][16 * 64];   // silu(Z_l), bf16
    float Z[3][16 * 64];            // pre-activations
    float Y[16 * 16];               // outputs (fp32)
    unsigned short G[16 * 64];      // bf16 gradient operand of the current layer
    float dH[16 * 64];              // gradient w.r.t. the activation of the current layer
};

// Forward of one tile; X must be loaded. Writes Z, H, Y.
__device__ void tile_forward(Tile& T, const MlpDev& P) {
    const int l = threadIdx.x & 63;
    const int col = l & 15, rq = (l >> 4) * 4;
    const unsigned short* in = T.X;
    int K = kMlpIn;
    for (int layer = 0; layer < 3; ++layer) {
        for (int nt = 0; nt < 4; ++nt) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma_tile(in, K, P.wt_bf[layer], K, nt * 16, acc);
            const float bias = P.b[layer][nt * 16 + col];
            for (int j = 0; j < 4; ++j) {
                float z = acc[j] + bias;
                int idx = (rq + j) * 64 + nt * 16 + col;
                T.Z[layer][idx] = z;
                T.H[layer][idx] = to_bf16(silu(z));
            }
        }
        __syncthreads();
        in = T.H[layer];
        K = kMlpHidden;
    }
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma_tile(T.H[2], kMlpHidden, P.wt_bf[3], kMlpHidden, 0, acc);
    for (int j = 0; j < 4; ++j) T.Y[(rq + j) * 16 + col] = acc[j] + P.b[3][col];
    __syncthreads();
}

__device__ void load_inputs(Tile& T, const MlpDev& P, const float* __restrict__ x, int row0, int nrows) {
    const int l = threadIdx.x & 63;
    for (int e = l; e < 16 * 16; e += 64) {
        int r = e >> 4, c = e & 15;
        float v = 0.f;
        if (r < nrows && c < kMlpFeatures) v = (x[(size_t)(row0 + r) * kMlpFeatures + c] - P.x_mean[c]) / P.x_std[c];
        T.X[e] = to_bf16(v);
    }
    __syncthreads();
}

// Inference: y = out_scale * MLP((x - mean) / std). One wave per workgroup walks 16-row
// tiles with a grid stride; it loads every weight operand fragment (40 x 4 bf16, the B
// operands of all layers) and its biases into registers once and reuses them for each
// tile. Activations pass through LDS between layers (D layout -> A layout). The
// arithmetic (k-block order, bias, SiLU, rounding) is that of tile_forward.
constexpr int kMlpForwardGrid = 2048;

__global__ void __launch_bounds__(64) k_mlp_forward(MlpDev P, const float* __restrict__ x, int nrows,
                                                    float* __restrict__ y) {
    __shared__ alignas(16) unsigned short X[16 * 16];
    __shared__ alignas(16) unsigned short H[2][16 * 64];
    const int l = threadIdx.x;
    const int r = l & 15, kq = (l >> 4) * 4, kq8 = (l >> 4) * 8, col = l & 15, rq = (l >> 4) * 4;
    s4 w0[4];
    bf8 w1[4][2], w2[4][2], w3[2];
    float b0[4], b1[4], b2[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        w0[nt] = *reinterpret_cast<const s4*>(P.wt_bf[0] + (size_t)(nt * 16 + r) * kMlpIn + kq);
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            w1[nt][ks] = *reinterpret_cast<const bf8*>(P.wt_bf[1] + (size_t)(nt * 16 + r) * kMlpHidden + ks * 32 + kq8);
            w2[nt][ks] = *reinterpret_cast<const bf8*>(P.wt_bf[2] + (size_t)(nt * 16 + r) * kMlpHidden + ks * 32 + kq8);
        }
        b0[nt] = P.b[0][nt * 16 + col];
        b1[nt] = P.b[1][nt * 16 + col];
        b2[nt] = P.b[2][nt * 16 + col];
    }
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
        w3[ks] = *reinterpret_cast<const bf8*>(P.wt_bf[3] + (size_t)r * kMlpHidden + ks * 32 + kq8);
    const float b3 = P.b[3][col];
    const float ys = col < kMlpOut ? P.y_scale[col] : 0.f;
    const int ntiles = (nrows + 15) / 16;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int row0 = tile * 16;
        const int nr = min(16, nrows - row0);
        for (int e = l; e < 16 * 16; e += 64) {
            const int rr = e >> 4, c = e & 15;
            float v = 0.f;
            if (rr < nr && c < kMlpFeatures) v = (x[(size_t)(row0 + rr) * kMlpFeatures + c] - P.x_mean[c]) / P.x_std[c];
            X[e] = to_bf16(v);
        }
        __syncthreads();
        {  // layer 1
            const s4 a = *reinterpret_cast<const s4*>(X + r * 16 + kq);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                f4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mfma16(a, w0[nt], acc);
#pragma unroll
                for (int j = 0; j < 4; ++j) H[0][(rq + j) * 64 + nt * 16 + col] = to_bf16(silu(acc[j] + b0[nt]));
            }
        }
        __syncthreads();
#pragma unroll
        for (int layer = 1; layer < 3; ++layer) {  // layers 2 and 3
            const unsigned short* in = H[layer - 1];
            unsigned short* out = H[layer & 1 ? 1 : 0];
            bf8 a[2];
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) a[ks] = *reinterpret_cast<const bf8*>(in + r * 64 + ks * 32 + kq8);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) acc = mfma32(a[ks], layer == 1 ? w1[nt][ks] : w2[nt][ks], acc);
                const float bb = layer == 1 ? b1[nt] : b2[nt];
#pragma unroll
                for (int j = 0; j < 4; ++j) out[(rq + j) * 64 + nt * 16 + col] = to_bf16(silu(acc[j] + bb));
            }
            __syncthreads();
        }
        {  // output layer from H[0] (layer 3's activations)
            bf8 a[2];
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) a[ks] = *reinterpret_cast<const bf8*>(H[0] + r * 64 + ks * 32 + kq8);
            f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) acc = mfma32(a[ks], w3[ks], acc);
            if (col < kMlpOut) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (rq + j < nr) y[(size_t)(row0 + rq + j) * kMlpOut + col] = (acc[j] + b3) * ys;
            }
        }
        __syncthreads();  // X and H are rewritten by the next tile
    }
}

// Training tile: forward, dL/dY of L = mean over rows and outputs of (y - t)^2 (in
// normalised units), backward; partial gradients of this tile into part[tile][param].
__global__ void __launch_bounds__(64) k_mlp_grad_tile(MlpDev P, const float* __restrict__ x,
                                                      const float* __restrict__ t, int nrows, float inv_count,
                                                      float* __restrict__ part, float* __restrict__ loss_part) {
    __shared__ Tile T;
    const int row0 = blockIdx.x * 16;
    const int nr = min(16, nrows - row0);
    const int l = threadIdx.x;
    const int col = l & 15, rq = (l >> 4) * 4;
    load_inputs(T, P, x, row0, nr);
    tile_forward(T, P);
    float* g = part + (size_t)blockIdx.x * kMlpParams;
    // output gradient (fp32 in dH[0..], bf16 operand in G)
    float loss = 0.f;
    for (int e = l; e < 16 * 16; e += 64) {
        int r = e >> 4, c = e & 15;
        float d = 0.f;
        if (r < nr && c < kMlpOut) {
            float diff = T.Y[e] - t[(size_t)(row0 + r) * kMlpOut + c] / P.y_scale[c];
            loss += diff * diff;
            d = 2.f * diff * inv_count;
        }
        T.dH[e] = d;
        T.G[e] = to_bf16(d);
    }
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o, 64);
    if (l == 0) loss_part[blockIdx.x] = loss;
    __syncthreads();
    // bias gradient of layer 4 (fp32 sums over rows)
    if (l < 16) {
        float s = 0.f;
        for (int r = 0; r < 16; ++r) s += T.dH[r * 16 + l];
        g[mlp_off_b(3) + l] = l < kMlpOut ? s : 0.f;
    }
    // layers 4..1: weight gradient = H_prev^T G, then gradient of H_prev = G W^T
    for (int layer = 3; layer >= 0; --layer) {
        const int Kin = layer == 0 ? kMlpIn : kMlpHidden;
        const int Nout = layer == 3 ? 16 : kMlpHidden;
        const unsigned short* Hprev = layer == 0 ? T.X : T.H[layer - 1];
        const int ldh = layer == 0 ? 16 : 64;
        float* gW = g + mlp_off_w(layer);
        // dW (Kin x Nout) = Hprev^T (Kin x 16 rows) * G (16 rows x Nout): M = Kin, K = 16 rows
        for (int mt = 0; mt < Kin / 16; ++mt)
            for (int nt = 0; nt < Nout / 16; ++nt) {
                s4 a, b;
                for (int j = 0; j < 4; ++j) {
                    a[j] = (short)Hprev[(rq + j) * ldh + mt * 16 + col];   // A[i=feature][k=row]
                    b[j] = (short)T.G[(rq + j) * Nout + nt * 16 + col];    // B[k=row][n=feature]
                }
                f4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mfma16(a, b, acc);
                for (int j = 0; j < 4; ++j) {
                    int i = mt * 16 + rq + j, n = nt * 16 + col;
                    float v = acc[j];
                    if (layer == 0 && i >= kMlpFeatures) v = 0.f;   // padded inputs stay zero
                    if (layer == 3 && n >= kMlpOut) v = 0.f;        // padded outputs stay zero
                    gW[i * Nout + n] = v;
                }
            }
        if (layer == 0) break;
        __syncthreads();
        // dH_prev (16 x 64) = G (16 x Nout) * W^T: B operand [k=out][n=in] = W[in][out] (row-major, ld Nout)
        for (int nt = 0; nt < 4; ++nt) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma_tile(T.G, Nout, P.w_bf[layer], Nout, nt * 16, acc);
            for (int j = 0; j < 4; ++j) T.dH[(rq + j) * 64 + nt * 16 + col] = acc[j];
        }
        __syncthreads();
        // through the SiLU of layer-1: dZ = dH * silu'(Z); bias gradient; bf16 operand
        float* Zp = T.Z[layer - 1];
        for (int e = l; e < 16 * 64; e += 64) {
            int r = e >> 6;
            float dz = r < nr ? T.dH[e] * dsilu(Zp[e]) : 0.f;
            T.dH[e] = dz;
            T.G[e] = to_bf16(dz);
        }
        __syncthreads();
        {
            float s = 0.f;
            for (int r = 0; r < 16; ++r) s += T.dH[r * 64 + l];
            g[mlp_off_b(layer - 1) + l] = s;
        }
        __syncthreads();
    }
}

// grad[p] = sum over tiles in order; loss = (sum of tile losses) / loss_div.
// With loss_acc: loss_acc[0] += sum * loss_norm (mean over an epoch's batches), loss_acc[1] += 1.
__global__ void k_mlp_reduce(const float* __restrict__ part, int tiles, float* __restrict__ grad,
                             const float* __restrict__ loss_part, float* __restrict__ loss, float loss_norm,
                             float* __restrict__ loss_acc, float loss_div) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kMlpParams) {
        float s = 0.f;
        for (int k = 0; k < tiles; ++k) s += part[(size_t)k * kMlpParams + p];
        grad[p] = s;
    }
    if (p == 0) {
        float s = 0.f;
        for (int k = 0; k < tiles; ++k) s += loss_part[k];
        *loss = s / loss_div;
        if (loss_acc) {
            loss_acc[0] += s * loss_norm;
            loss_acc[1] += 1.f;
        }
    }
}

// Adam on the FP32 master weights.
__global__ void k_mlp_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1,
                           float c2) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    float g = grad[p];
    float mm = b1 * m[p] + (1.f - b1) * g;
    float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    w[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// BF16 operand copies: W (K x N row-major) and W^T (N x K row-major) per layer.
__global__ void k_mlp_refresh(const float* __restrict__ w, MlpDev P) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    for (int layer = 0; layer < 4; ++layer) {
        const int K = layer == 0 ? kMlpIn : kMlpHidden;
        const int N = layer == 3 ? 16 : kMlpHidden;
        if (e < K * N) {
            int k = e / N, n = e % N;
            unsigned short h = to_bf16(w[mlp_off_w(layer) + e]);
            P.w_bf[layer][e] = h;
            P.wt_bf[layer][n * K + k] = h;
        }
    }
}

}  // namespace

// Learned warm start: y (the network's change of the reaction per unit mass) += the previous
// step's reaction per unit mass, features 6..8 of each contact. A contact without a previous
// reaction gets y = 0: it starts plain, as with the previous-step warm start.
__global__ void k_mlp_add_previous(const float* __restrict__ feat, int rows, float* __restrict__ y) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const float* p = feat + (size_t)r * kMlpFeatures + 6;
    const bool has_prev = p[0] != 0.f || p[1] != 0.f || p[2] != 0.f;
    for (int k = 0; k < 3; ++k) y[3 * (size_t)r + k] = has_prev ? y[3 * (size_t)r + k] + p[k] : 0.f;
}

void launch_mlp_add_previous(const float* feat, int rows, float* y, hipStream_t st) {
    if (rows > 0) hipLaunchKernelGGL(k_mlp_add_previous, dim3((rows + 255) / 256), dim3(256), 0, st, feat, rows, y);
}

void launch_mlp_forward(const MlpDev& P, const float* x, int rows, float* y, hipStream_t st) {
    if (rows > 0)
        hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)std::min((rows + 15) / 16, kMlpForwardGrid)), dim3(64), 0, st,
                           P, x, rows, y);
}

}  // namespace pdfc

// ------------------------------------------------------------------- host
struct pdfc_mlp {
    pdfc::MlpDev dev{};
    float* w = nullptr;      // master weights (device)
    float* m = nullptr;
    float* v = nullptr;
    float* grad = nullptr;
    float* loss = nullptr;     // [0] last batch loss sum, [1..2] epoch accumulator
    float* norm = nullptr;   // x_mean[12], x_std[12], y_scale[3]
    unsigned short* bf = nullptr;
    float* part = nullptr;
    float* loss_part = nullptr;
    int part_tiles = 0;
    float* xbuf = nullptr;
    float* tbuf = nullptr;
    size_t xcap = 0;
    int64_t adam_t = 0;
    hipStream_t stream = nullptr;
};

namespace {
void chk(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
void bind(pdfc_mlp* M) {
    pdfc::MlpDev& D = M->dev;
    for (int l = 0; l < 4; ++l) {
        D.b[l] = M->w + pdfc::mlp_off_b(l);
        D.w_bf[l] = M->bf + pdfc::mlp_bf_off(l);
        D.wt_bf[l] = M->bf + pdfc::mlp_bf_off(l) + pdfc::mlp_bf_size(l);
    }
    D.x_mean = M->norm;
    D.x_std = M->norm + pdfc::kMlpFeatures;
    D.y_scale = M->norm + 2 * pdfc::kMlpFeatures;
}
void refresh(pdfc_mlp* M) {
    hipLaunchKernelGGL(pdfc::k_mlp_refresh, dim3((64 * 64 + 255) / 256), dim3(256), 0, M->stream, M->w, M->dev);
}
void ensure_io(pdfc_mlp* M, size_t rows) {
    if (rows <= M->xcap) return;
    if (M->xbuf) (void)hipFree(M->xbuf);
    if (M->tbuf) (void)hipFree(M->tbuf);
    M->xcap = rows;
    chk(hipMalloc(&M->xbuf, rows * pdfc::kMlpFeatures * sizeof(float)), "hipMalloc");
    chk(hipMalloc(&M->tbuf, rows * pdfc::kMlpOut * sizeof(float)), "hipMalloc");
}
void ensure_parts(pdfc_mlp* M, int tiles) {
    if (tiles <= M->part_tiles) return;
    if (M->part) (void)hipFree(M->part);
    if (M->loss_part) (void)hipFree(M->loss_part);
    M->part_tiles = tiles;
    chk(hipMalloc(&M->part, (size_t)tiles * pdfc::kMlpParams * sizeof(float)), "hipMalloc");
    chk(hipMalloc(&M->loss_part, (size_t)tiles * sizeof(float)), "hipMalloc");
}
// gradient of the batch rows [0, rows) already in xbuf/tbuf
void batch_gradient(pdfc_mlp* M, const float* x, const float* t, int rows) {
    const int tiles = (rows + 15) / 16;
    ensure_parts(M, tiles);
    float inv = 1.f / (float)(rows * pdfc::kMlpOut);
    hipLaunchKernelGGL(pdfc::k_mlp_grad_tile, dim3(tiles), dim3(64), 0, M->stream, M->dev, x, t, rows, inv, M->part,
                       M->loss_part);
    hipLaunchKernelGGL(pdfc::k_mlp_reduce, dim3((pdfc::kMlpParams + 255) / 256), dim3(256), 0, M->stream, M->part,
                       tiles, M->grad, M->loss_part, M->loss, inv, M->loss + 1, 1.f);
}

// ---- the allocation-free gradient (pdfc_mlp_gradient_noalloc) ----
// Its per-tile partial gradients and tile losses live in this fixed array, one copy per GPU,
// held until the process ends: kNoAllocTiles x kMlpParams partial gradients, then the tile
// losses. In host code `scratch` only names the symbol; its device address on the current GPU
// comes from hipGetSymbolAddress (noalloc_scratch below) and is what the kernels receive.
constexpr int kNoAllocTiles = PDFC_MLP_NOALLOC_MAX_ROWS / 16;
constexpr size_t kScratchFloats = (size_t)kNoAllocTiles * pdfc::kMlpParams + kNoAllocTiles;
static_assert(PDFC_MLP_NOALLOC_MAX_ROWS % 16 == 0, "whole tiles");
static __device__ __align__(256) float scratch[kScratchFloats];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current GPU, looked up once per device.
float* noalloc_scratch() {
    int dev = 0;
    chk(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        chk(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<float*>(ptr);
    }
    return g_scratch_ptr[dev];
}
template <typename F>
int guard(F&& f) {
    try {
        f();
        return PDFC_OK;
    } catch (const std::invalid_argument& e) {
        pdfc::set_last_error(e.what());
        return PDFC_ERR_INVALID;
    } catch (const std::exception& e) {
        pdfc::set_last_error(e.what());
        return PDFC_ERR_DEVICE;
    }
}
}  // namespace

const pdfc::MlpDev* pdfc_mlp_device(const pdfc_mlp* M) { return M ? &M->dev : nullptr; }

extern "C" {

int pdfc_mlp_create(pdfc_mlp** out, uint64_t seed) {
    return guard([&] {
        if (!out) throw std::invalid_argument("null argument");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw std::runtime_error("no HIP device available");
        auto* M = new pdfc_mlp();
        chk(hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking), "stream");
        chk(hipMalloc(&M->w, pdfc::kMlpParams * sizeof(float)), "hipMalloc");
        chk(hipMalloc(&M->m, pdfc::kMlpParams * sizeof(float)), "hipMalloc");
        chk(hipMalloc(&M->v, pdfc::kMlpParams * sizeof(float)), "hipMalloc");
        chk(hipMalloc(&M->grad, pdfc::kMlpParams * sizeof(float)), "hipMalloc");
        chk(hipMalloc(&M->loss, 3 * sizeof(float)), "hipMalloc");
        chk(hipMalloc(&M->norm, pdfc::kMlpNorm * sizeof(float)), "hipMalloc");
        chk(hipMalloc(&M->bf, pdfc::kBfTotal * sizeof(unsigned short)), "hipMalloc");
        bind(M);
        // Glorot-uniform weights, zero biases, identity normalisation
        std::mt19937_64 rng(seed);
        std::vector<float> w(pdfc::kMlpParams, 0.f), norm(pdfc::kMlpNorm, 1.f);
        for (int i = 0; i < pdfc::kMlpFeatures; ++i) norm[i] = 0.f;
        const int fan_in[4] = {pdfc::kMlpFeatures, pdfc::kMlpHidden, pdfc::kMlpHidden, pdfc::kMlpHidden};
        const int fan_out[4] = {pdfc::kMlpHidden, pdfc::kMlpHidden, pdfc::kMlpHidden, pdfc::kMlpOut};
        for (int l = 0; l < 4; ++l) {
            const int K = l == 0 ? pdfc::kMlpIn : pdfc::kMlpHidden, N = l == 3 ? 16 : pdfc::kMlpHidden;
            double a = std::sqrt(6.0 / (fan_in[l] + fan_out[l]));
            for (int k = 0; k < K; ++k)
                for (int n = 0; n < N; ++n) {
                    double u = (double)(rng() >> 11) * (1.0 / 9007199254740992.0);
                    float val = (float)((2. * u - 1.) * a);
                    if ((l == 0 && k >= pdfc::kMlpFeatures) || (l == 3 && n >= pdfc::kMlpOut)) val = 0.f;
                    w[pdfc::mlp_off_w(l) + k * N + n] = val;
                }
        }
        chk(hipMemcpy(M->w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice), "copy");
        chk(hipMemcpy(M->norm, norm.data(), norm.size() * sizeof(float), hipMemcpyHostToDevice), "copy");
        chk(hipMemset(M->m, 0, pdfc::kMlpParams * sizeof(float)), "memset");
        chk(hipMemset(M->v, 0, pdfc::kMlpParams * sizeof(float)), "memset");
        refresh(M);
        chk(hipStreamSynchronize(M->stream), "sync");
        (void)noalloc_scratch();  // set-up of the allocation-free gradient on this device
        *out = M;
    });
}

void pdfc_mlp_destroy(pdfc_mlp* M) {
    if (!M) return;
    if (M->stream) (void)hipStreamSynchronize(M->stream);
    for (void* p : {(void*)M->w, (void*)M->m, (void*)M->v, (void*)M->grad, (void*)M->loss, (void*)M->norm,
                    (void*)M->bf, (void*)M->part, (void*)M->loss_part, (void*)M->xbuf, (void*)M->tbuf})
        if (p) (void)hipFree(p);
    if (M->stream) (void)hipStreamDestroy(M->stream);
    delete M;
}

int pdfc_mlp_set_params(pdfc_mlp* M, const float* params, const float* norm) {
    return guard([&] {
        if (!M) throw std::invalid_argument("null mlp");
        if (params) chk(hipMemcpy(M->w, params, pdfc::kMlpParams * sizeof(float), hipMemcpyHostToDevice), "copy");
        if (norm) chk(hipMemcpy(M->norm, norm, pdfc::kMlpNorm * sizeof(float), hipMemcpyHostToDevice), "copy");
        chk(hipMemset(M->m, 0, pdfc::kMlpParams * sizeof(float)), "memset");
        chk(hipMemset(M->v, 0, pdfc::kMlpParams * sizeof(float)), "memset");
        M->adam_t = 0;
        refresh(M);
        chk(hipStreamSynchronize(M->stream), "sync");
    });
}

int pdfc_mlp_get_params(pdfc_mlp* M, float* params, float* norm) {
    return guard([&] {
        if (!M) throw std::invalid_argument("null mlp");
        chk(hipStreamSynchronize(M->stream), "sync");
        if (params) chk(hipMemcpy(params, M->w, pdfc::kMlpParams * sizeof(float), hipMemcpyDeviceToHost), "copy");
        if (norm) chk(hipMemcpy(norm, M->norm, pdfc::kMlpNorm * sizeof(float), hipMemcpyDeviceToHost), "copy");
    });
}

int pdfc_mlp_forward(pdfc_mlp* M, const float* x, size_t rows, float* y) {
    return guard([&] {
        if (!M || (!x && rows) || (!y && rows)) throw std::invalid_argument("null argument");
        if (rows == 0) return;
        ensure_io(M, rows);
        chk(hipMemcpyAsync(M->xbuf, x, rows * pdfc::kMlpFeatures * sizeof(float), hipMemcpyHostToDevice, M->stream),
            "copy");
        pdfc::launch_mlp_forward(M->dev, M->xbuf, (int)rows, M->tbuf, M->stream);
        chk(hipMemcpyAsync(y, M->tbuf, rows * pdfc::kMlpOut * sizeof(float), hipMemcpyDeviceToHost, M->stream), "copy");
        chk(hipStreamSynchronize(M->stream), "sync");
    });
}

int pdfc_mlp_gradient(pdfc_mlp* M, const float* x, const float* t, size_t rows, float* grad, float* loss) {
    return guard([&] {
        if (!M || !x || !t || rows == 0) throw std::invalid_argument("bad argument");
        ensure_io(M, rows);
        chk(hipMemcpyAsync(M->xbuf, x, rows * pdfc::kMlpFeatures * sizeof(float), hipMemcpyHostToDevice, M->stream),
            "copy");
        chk(hipMemcpyAsync(M->tbuf, t, rows * pdfc::kMlpOut * sizeof(float), hipMemcpyHostToDevice, M->stream), "copy");
        batch_gradient(M, M->xbuf, M->tbuf, (int)rows);
        if (grad)
            chk(hipMemcpyAsync(grad, M->grad, pdfc::kMlpParams * sizeof(float), hipMemcpyDeviceToHost, M->stream), "copy");
        float l = 0.f;
        chk(hipMemcpyAsync(&l, M->loss, sizeof(float), hipMemcpyDeviceToHost, M->stream), "copy");
        chk(hipStreamSynchronize(M->stream), "sync");
        if (loss) *loss = l / (float)(rows * pdfc::kMlpOut);
    });
}

int pdfc_mlp_gradient_noalloc(pdfc_mlp* M, const float* x, const float* t, size_t rows, float* grad, float* loss) {
    return guard([&] {
        if (!M || !x || !t || !grad || !loss || rows == 0) throw std::invalid_argument("bad argument");
        if (rows > (size_t)PDFC_MLP_NOALLOC_MAX_ROWS)
            throw std::invalid_argument("pdfc_mlp_gradient_noalloc: " + std::to_string(rows) +
                                        " rows exceed the maximum of " + std::to_string(PDFC_MLP_NOALLOC_MAX_ROWS));
        float* part = noalloc_scratch();
        float* loss_part = part + (size_t)kNoAllocTiles * pdfc::kMlpParams;
        const int n = (int)rows, tiles = (n + 15) / 16;
        const float inv = 1.f / (float)(n * pdfc::kMlpOut);
        // every tile writes all its partials and its loss, so nothing is cleared
        hipLaunchKernelGGL(pdfc::k_mlp_grad_tile, dim3(tiles), dim3(64), 0, M->stream, M->dev, x, t, n, inv, part,
                           loss_part);
        hipLaunchKernelGGL(pdfc::k_mlp_reduce, dim3((pdfc::kMlpParams + 255) / 256), dim3(256), 0, M->stream, part,
                           tiles, grad, loss_part, loss, inv, (float*)nullptr, (float)(n * pdfc::kMlpOut));
        chk(hipGetLastError(), "launch");
        chk(hipStreamSynchronize(M->stream), "sync");
    });
}

int pdfc_mlp_train(pdfc_mlp* M, const float* x, const float* t, size_t rows, int epochs, int batch, float lr,
                   uint64_t seed, float* loss_per_epoch) {
    return guard([&] {
        if (!M || !x || !t || rows < 16 || epochs < 1 || batch < 16) throw std::invalid_argument("bad argument");
        batch = (batch / 16) * 16;
        ensure_io(M, rows);
        // shuffled copy per epoch on the host, uploaded once per epoch
        std::vector<float> xs(rows * pdfc::kMlpFeatures), ts(rows * pdfc::kMlpOut);
        std::vector<size_t> idx(rows);
        for (size_t i = 0; i < rows; ++i) idx[i] = i;
        std::mt19937_64 rng(seed);
        for (int e = 0; e < epochs; ++e) {
            for (size_t i = rows - 1; i > 0; --i) std::swap(idx[i], idx[rng() % (i + 1)]);
            for (size_t i = 0; i < rows; ++i) {
                std::memcpy(&xs[i * pdfc::kMlpFeatures], &x[idx[i] * pdfc::kMlpFeatures], pdfc::kMlpFeatures * sizeof(float));
                std::memcpy(&ts[i * pdfc::kMlpOut], &t[idx[i] * pdfc::kMlpOut], pdfc::kMlpOut * sizeof(float));
            }
            chk(hipMemcpyAsync(M->xbuf, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice, M->stream), "copy");
            chk(hipMemcpyAsync(M->tbuf, ts.data(), ts.size() * sizeof(float), hipMemcpyHostToDevice, M->stream), "copy");
            chk(hipMemsetAsync(M->loss + 1, 0, 2 * sizeof(float), M->stream), "memset");
            for (size_t r0 = 0; r0 + batch <= rows; r0 += batch) {
                batch_gradient(M, M->xbuf + r0 * pdfc::kMlpFeatures, M->tbuf + r0 * pdfc::kMlpOut, batch);
                ++M->adam_t;
                const float b1 = 0.9f, b2 = 0.999f;
                float c1 = 1.f - std::pow(b1, (float)M->adam_t), c2 = 1.f - std::pow(b2, (float)M->adam_t);
                hipLaunchKernelGGL(pdfc::k_mlp_adam, dim3((pdfc::kMlpParams + 255) / 256), dim3(256), 0, M->stream,
                                   M->w, M->m, M->v, M->grad, lr, b1, b2, 1e-8f, c1, c2);
                refresh(M);
            }
            if (loss_per_epoch) {
                float acc[2] = {0.f, 0.f};
                chk(hipMemcpyAsync(acc, M->loss + 1, 2 * sizeof(float), hipMemcpyDeviceToHost, M->stream), "copy");
                chk(hipStreamSynchronize(M->stream), "sync");
                loss_per_epoch[e] = acc[1] > 0.f ? acc[0] / acc[1] : 0.f;
            }
        }
        chk(hipStreamSynchronize(M->stream), "sync");
    });
}

}  // extern "C"
