// mc_mlp.h - the warm-start network of the per-cell balance: a 72 -> 128 ->
// 128 -> 128 -> 31 ReLU perceptron, BF16 weights and activations, FP32
// accumulation on the matrix cores (v_mfma_f32_16x16x32_bf16), trained on
// the device by its own forward, backward and Adam kernels.
//
// Features of a cell (72): log10 of J summed over 64 contiguous groups of
// frequency bins, log10 nH, log10 of the abundances of H, He, C, N, O, Ne, S.
// Targets (31): log10 Te and log10 of the ionic fractions of those seven
// elements, stages 1..min(Z+1, 5) (2 + 3 + 5*5 = 30).  Both are standardised
// with per-column mean and standard deviation held with the network.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcm {

constexpr int kBands = 64;
constexpr int kIn = 72, kInP = 96;       // inputs, padded to a multiple of 32
constexpr int kHid = 128;
constexpr int kOut = 31, kOutP = 32;
constexpr int kLayers = 4;
constexpr int kSlotsZ[7] = { 1, 2, 6, 7, 8, 10, 16 };
constexpr int kStageMax = 5;

__host__ __device__ constexpr int layer_in(int l) { return l == 0 ? kInP : kHid; }
__host__ __device__ constexpr int layer_out(int l) { return l == kLayers - 1 ? kOutP : kHid; }
// flat FP32 parameter vector: W_l [out][in] then b_l [out], layer after layer
__host__ __device__ constexpr int w_off(int l)
{
    return l == 0 ? 0 : w_off(l - 1) + layer_out(l - 1) * layer_in(l - 1) + layer_out(l - 1);
}
__host__ __device__ constexpr int b_off(int l) { return w_off(l) + layer_out(l) * layer_in(l); }
constexpr int kParams = w_off(kLayers);  // 49,568
constexpr int kWeightsBf16 = kParams;     // bf16 copies of W (the b slots unused)

struct Net {
    const uint16_t *W;    // bf16 copy of the parameters (W_l at w_off(l), [out][in])
    const float *P;       // fp32 master parameters (biases read from here)
};

// column scaling of features and targets: x_std = (x - mean) / std
struct Scale {
    const float *xmean, *xstd, *ymean, *ystd;
};

}  // namespace mcm

// launchers (mc_mlp.hip); every pointer is device memory.
// forward: X0 bf16 [n][kInP] -> Y fp32 [n][ldy] (standardised outputs);
// H (nullable) bf16 [3][n][kHid] keeps the hidden activations for backward.
hipError_t mcm_forward(const uint16_t *X0, int n, mcm::Net net, float *Y, int ldy, uint16_t *H,
                       int blocks, hipStream_t s);
// backward of L = sum M (Y - T)^2 over a batch: grad (fp32 [kParams]) gets
// the unnormalised gradient 2 sum M (Y - T) dY/dP; sums[0] += sum M (Y-T)^2,
// sums[1] += sum M (zero them first).  dA, dB: bf16 scratch [n][kHid].
hipError_t mcm_backward(const uint16_t *X0, const uint16_t *H, const float *Y, const float *T,
                        const float *M, int n, mcm::Net net, uint16_t *dA, uint16_t *dB,
                        float *grad, double *sums, hipStream_t s);
// Adam step t (1-based) on grad / sums[1]; refreshes the bf16 copy
hipError_t mcm_adam(float *P, const float *grad, float *m, float *v, const double *sums, int t,
                    float lr, uint16_t *W, hipStream_t s);
hipError_t mcm_publish(const float *P, uint16_t *W, hipStream_t s);
// batch rows idx[0..n) of a raw data set: standardised bf16 features, standardised targets, mask
hipError_t mcm_gather(const float *Xs, const float *Ts, const float *Ms, const int *idx, int n,
                      mcm::Scale sc, uint16_t *X0, float *T, float *M, hipStream_t s);
hipError_t mcm_standardise(const float *Xraw, int n, mcm::Scale sc, uint16_t *X0, hipStream_t s);
hipError_t mcm_unscale(float *Y, int n, int ldy, mcm::Scale sc, hipStream_t s);
