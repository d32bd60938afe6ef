// mc_mlp.hip - kernels of the warm-start network (see mc_mlp.h): fused
// forward, backward GEMMs, loss, Adam, and the data movers around them.
//
// Matrix products use v_mfma_f32_16x16x32_bf16 (one 64-lane wavefront
// computes a 16x16 FP32 tile from 16x32 and 32x16 BF16 operands).  Operand
// layout of that instruction: lane l supplies A[l%16][(l/16)*8 + 0..7] and
// B[(l/16)*8 + 0..7][l%16] and receives D[(l/16)*4 + 0..3][l%16].
#include "mc_mlp.h"

namespace mcm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even (operands are finite)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 ld8(const uint16_t *p)
{
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(p));
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------- forward
// LDS: the four weight matrices (row stride in + 8 to spread banks), the
// biases, and per wavefront two 16-row activation buffers (stride 136).
constexpr int kActS = kHid + 8;
constexpr int kTile = 64;                       // rows per workgroup tile (4 wavefronts x 16)
__host__ __device__ constexpr int lw_stride(int l) { return layer_in(l) + 8; }
__host__ __device__ constexpr int lw_off(int l)
{
    return l == 0 ? 0 : lw_off(l - 1) + layer_out(l - 1) * lw_stride(l - 1);
}
constexpr int kLdsW = lw_off(kLayers);                                   // bf16 elements
constexpr int kLdsBias = kHid * 3 + kOutP;                               // floats
constexpr int kLdsAct = 4 * 2 * 16 * kActS;                              // bf16 elements
constexpr size_t kFwdLds = (size_t)kLdsW * 2 + (size_t)kLdsBias * 4 + (size_t)kLdsAct * 2;

__global__ void __launch_bounds__(256) forward_kernel(const uint16_t *__restrict__ X0, int n, Net net,
                                                      float *__restrict__ Y, int ldy,
                                                      uint16_t *__restrict__ H)
{
    extern __shared__ __align__(16) unsigned char lds[];
    uint16_t *Ws = reinterpret_cast<uint16_t *>(lds);
    float *bias = reinterpret_cast<float *>(lds + (size_t)kLdsW * 2);
    uint16_t *act = reinterpret_cast<uint16_t *>(lds + (size_t)kLdsW * 2 + (size_t)kLdsBias * 4);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    // weights and biases into LDS, 16 bytes at a time
    for (int l = 0; l < kLayers; ++l) {
        const int K = layer_in(l), N = layer_out(l), vpr = K / 8;
        const uint16_t *src = net.W + w_off(l);
        for (int e = tid; e < N * vpr; e += 256) {
            const int r = e / vpr, v = e % vpr;
            *reinterpret_cast<uint4 *>(Ws + lw_off(l) + r * lw_stride(l) + v * 8) =
                *reinterpret_cast<const uint4 *>(src + (size_t)r * K + v * 8);
        }
    }
    for (int e = tid; e < kLdsBias; e += 256) {
        const int l = e < 3 * kHid ? e / kHid : 3, j = e < 3 * kHid ? e % kHid : e - 3 * kHid;
        bias[e] = net.P[b_off(l) + j];
    }
    uint16_t *a0 = act + (size_t)(w * 2) * 16 * kActS, *a1 = a0 + 16 * kActS;
    const int ntiles = (n + kTile - 1) / kTile;
    const int lr = lane & 15, lk = (lane >> 4) * 8, lrow = (lane >> 4) * 4;
    __syncthreads();
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int r0 = tile * kTile + w * 16;
        for (int e = lane; e < 16 * (kInP / 8); e += 64) {
            const int r = e / (kInP / 8), v = e % (kInP / 8);
            uint4 val = make_uint4(0, 0, 0, 0);
            if (r0 + r < n) val = *reinterpret_cast<const uint4 *>(X0 + (size_t)(r0 + r) * kInP + v * 8);
            *reinterpret_cast<uint4 *>(a0 + r * kActS + v * 8) = val;
        }
        __syncthreads();
        for (int l = 0; l < kLayers; ++l) {
            const uint16_t *in = (l & 1) ? a1 : a0;
            uint16_t *out = (l & 1) ? a0 : a1;
            const int K = layer_in(l), N = layer_out(l);
            const uint16_t *Wl = Ws + lw_off(l);
            const float *bl = bias + l * kHid;
            for (int t = 0; t < N / 16; ++t) {
                f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
                for (int s = 0; s < K / 32; ++s) {
                    const bf16x8 a = ld8(in + lr * kActS + 32 * s + lk);
                    const bf16x8 b = ld8(Wl + (16 * t + lr) * lw_stride(l) + 32 * s + lk);
                    acc = mfma(a, b, acc);
                }
                const int col = 16 * t + lr;
                for (int i = 0; i < 4; ++i) {
                    const int row = lrow + i;
                    float v = acc[i] + bl[col];
                    if (l < kLayers - 1) {
                        v = v > 0.f ? v : 0.f;
                        const uint16_t h = f2bf(v);
                        out[row * kActS + col] = h;
                        if (H && r0 + row < n) H[((size_t)l * n + r0 + row) * kHid + col] = h;
                    } else if (col < kOut && r0 + row < n) {
                        Y[(size_t)(r0 + row) * ldy + col] = v;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------- GEMM
// C[M][N] = sum_k A(m,k) B(k,n), A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn], BF16 operands staged in LDS (k contiguous),
// 64x64 tile per workgroup, K step 32.  EPI_F32_ADD splits K over
// blockIdx.z (kc rows each) and adds into a zeroed C32.
enum { EPI_F32_ADD = 0, EPI_RELU_MASK = 1 };

template <int EPI>
__global__ void __launch_bounds__(256) gemm_kernel(const uint16_t *__restrict__ A, long sam, long sak,
                                                   const uint16_t *__restrict__ B, long sbk, long sbn,
                                                   int M, int N, int K, int kc, float *__restrict__ C32,
                                                   uint16_t *__restrict__ C16, int ldc,
                                                   const uint16_t *__restrict__ Hm, int ldh)
{
    constexpr int S = 40;
    __shared__ __align__(16) uint16_t As[64 * S], Bs[64 * S];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int lr = lane & 15, lk = (lane >> 4) * 8, lrow = (lane >> 4) * 4;
    f32x4 acc[4];
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{ 0.f, 0.f, 0.f, 0.f };
    const int rr = tid >> 2, kk = (tid & 3) * 8;
    const int kb = blockIdx.z * kc, ke = kb + kc < K ? kb + kc : K;
    for (int k0 = kb; k0 < ke; k0 += 32) {
        for (int i = 0; i < 8; ++i) {
            const int k = k0 + kk + i, m = m0 + rr, nn = n0 + rr;
            As[rr * S + kk + i] = (m < M && k < ke) ? A[(long)m * sam + (long)k * sak] : (uint16_t)0;
            Bs[rr * S + kk + i] = (nn < N && k < ke) ? B[(long)k * sbk + (long)nn * sbn] : (uint16_t)0;
        }
        __syncthreads();
        const bf16x8 a = ld8(As + (16 * w + lr) * S + lk);
        for (int t = 0; t < 4; ++t) acc[t] = mfma(a, ld8(Bs + (16 * t + lr) * S + lk), acc[t]);
        __syncthreads();
    }
    for (int t = 0; t < 4; ++t)
        for (int i = 0; i < 4; ++i) {
            const int row = m0 + 16 * w + lrow + i, col = n0 + 16 * t + lr;
            if (row >= M || col >= N) continue;
            if (EPI == EPI_F32_ADD) {
                atomicAdd(&C32[(long)row * ldc + col], acc[t][i]);
            } else {
                const float v = bf2f(Hm[(long)row * ldh + col]) > 0.f ? acc[t][i] : 0.f;
                C16[(long)row * ldc + col] = f2bf(v);
            }
        }
}

// dY = 2 M (Y - T) as bf16 [n][kOutP] (column kOutP-1 zero); sums of M (Y-T)^2 and M
__global__ void __launch_bounds__(256) loss_kernel(const float *__restrict__ Y, const float *__restrict__ T,
                                                   const float *__restrict__ Mk, int n,
                                                   uint16_t *__restrict__ dY, double *__restrict__ sums)
{
    __shared__ double red[2][256];
    double l2 = 0.0, mm = 0.0;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < (long)n * kOutP; e += (long)gridDim.x * 256) {
        const long r = e / kOutP;
        const int j = (int)(e % kOutP);
        float g = 0.f;
        if (j < kOut) {
            const float m = Mk[r * kOut + j], d = Y[r * kOutP + j] - T[r * kOut + j];
            g = 2.f * m * d;
            l2 += (double)m * (double)d * (double)d;
            mm += (double)m;
        }
        dY[e] = f2bf(g);
    }
    red[0][threadIdx.x] = l2;
    red[1][threadIdx.x] = mm;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            red[0][threadIdx.x] += red[0][threadIdx.x + o];
            red[1][threadIdx.x] += red[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&sums[0], red[0][0]);
        atomicAdd(&sums[1], red[1][0]);
    }
}

// out[j] = sum_r dY[r][j], one workgroup per column
__global__ void __launch_bounds__(256) colsum_kernel(const uint16_t *__restrict__ dY, int n, int ld,
                                                     float *__restrict__ out)
{
    __shared__ float red[256];
    const int j = blockIdx.x;
    float s = 0.f;
    for (int r = threadIdx.x; r < n; r += 256) s += bf2f(dY[(long)r * ld + j]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
}

__global__ void adam_kernel(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ m,
                            float *__restrict__ v, const double *__restrict__ sums, float c1, float c2,
                            float lr, uint16_t *__restrict__ W)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    const float scale = sums[1] > 0.0 ? (float)(1.0 / sums[1]) : 0.f;
    const float g = G[i] * scale;
    const float mi = 0.9f * m[i] + 0.1f * g;
    const float vi = 0.999f * v[i] + 0.001f * g * g;
    m[i] = mi;
    v[i] = vi;
    const float p = P[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
    P[i] = p;
    W[i] = f2bf(p);
}

__global__ void publish_kernel(const float *__restrict__ P, uint16_t *__restrict__ W)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kParams) W[i] = f2bf(P[i]);
}

__device__ __forceinline__ float stdz(float x, float mean, float sd) { return (x - mean) / sd; }

__global__ void gather_kernel(const float *__restrict__ Xs, const float *__restrict__ Ts,
                              const float *__restrict__ Ms, const int *__restrict__ idx, int n, Scale sc,
                              uint16_t *__restrict__ X0, float *__restrict__ T, float *__restrict__ M)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * kInP) return;
    const long r = e / kInP;
    const int j = (int)(e % kInP);
    const long src = idx[r];
    X0[e] = j < kIn ? f2bf(stdz(Xs[src * kIn + j], sc.xmean[j], sc.xstd[j])) : (uint16_t)0;
    if (j < kOut) {
        T[r * kOut + j] = stdz(Ts[src * kOut + j], sc.ymean[j], sc.ystd[j]);
        M[r * kOut + j] = Ms[src * kOut + j];
    }
}

__global__ void standardise_kernel(const float *__restrict__ X, int n, Scale sc, uint16_t *__restrict__ X0)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * kInP) return;
    const long r = e / kInP;
    const int j = (int)(e % kInP);
    X0[e] = j < kIn ? f2bf(stdz(X[r * kIn + j], sc.xmean[j], sc.xstd[j])) : (uint16_t)0;
}

__global__ void unscale_kernel(float *__restrict__ Y, int n, int ldy, Scale sc)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * kOut) return;
    const long r = e / kOut;
    const int j = (int)(e % kOut);
    Y[r * ldy + j] = Y[r * ldy + j] * sc.ystd[j] + sc.ymean[j];
}

}  // namespace mcm

using namespace mcm;

static unsigned grid1(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t mcm_forward(const uint16_t *X0, int n, Net net, float *Y, int ldy, uint16_t *H, int blocks,
                       hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)forward_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFwdLds);
        if (e != hipSuccess) return e;
        attr = true;
    }
    const int ntiles = (n + kTile - 1) / kTile;
    const int g = ntiles < blocks ? ntiles : blocks;
    hipLaunchKernelGGL(forward_kernel, dim3(g), dim3(256), kFwdLds, s, X0, n, net, Y, ldy, H);
    return hipGetLastError();
}

template <int EPI>
static hipError_t gemm(const uint16_t *A, long sam, long sak, const uint16_t *B, long sbk, long sbn, int M,
                       int N, int K, float *C32, uint16_t *C16, int ldc, const uint16_t *Hm, int ldh,
                       hipStream_t s)
{
    // split K in chunks of 128 rows when accumulating (the weight gradients)
    const int kc = EPI == EPI_F32_ADD ? 128 : K;
    hipLaunchKernelGGL(gemm_kernel<EPI>, dim3(grid1(N, 64), grid1(M, 64), grid1(K, kc)), dim3(256), 0, s, A,
                       sam, sak, B, sbk, sbn, M, N, K, kc, C32, C16, ldc, Hm, ldh);
    return hipGetLastError();
}

hipError_t mcm_backward(const uint16_t *X0, const uint16_t *H, const float *Y, const float *T, const float *M,
                        int n, Net net, uint16_t *dA, uint16_t *dB, float *grad, double *sums, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(loss_kernel, dim3(grid1((long)n * kOutP, 256) < 1024 ? grid1((long)n * kOutP, 256) : 1024),
                       dim3(256), 0, s, Y, T, M, n, dA, sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if ((e = hipMemsetAsync(grad, 0, kParams * sizeof(float), s)) != hipSuccess) return e;
    // gradient rows of layer l flow in cur (ld = its width); prev activation
    // of layer l: H[l-1] (or X0)
    uint16_t *cur = dA, *nxt = dB;
    for (int l = kLayers - 1; l >= 0; --l) {
        const int K = layer_in(l), N = layer_out(l);
        const uint16_t *prev = l == 0 ? X0 : H + (size_t)(l - 1) * n * kHid;
        // dW_l [N][K] = cur^T prev
        e = gemm<EPI_F32_ADD>(cur, 1, N, prev, K, 1, N, K, n, grad + w_off(l), nullptr, K, nullptr, 0, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(colsum_kernel, dim3(N), dim3(256), 0, s, cur, n, N, grad + b_off(l));
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (l == 0) break;
        // d(prev) [n][K] = (cur W_l) masked by prev > 0
        e = gemm<EPI_RELU_MASK>(cur, N, 1, net.W + w_off(l), K, 1, n, K, N, nullptr, nxt, K, prev, K, s);
        if (e != hipSuccess) return e;
        uint16_t *t = cur;
        cur = nxt;
        nxt = t;
    }
    return hipSuccess;
}

hipError_t mcm_adam(float *P, const float *grad, float *m, float *v, const double *sums, int t, float lr,
                    uint16_t *W, hipStream_t s)
{
    const float c1 = 1.f - powf(0.9f, (float)t), c2 = 1.f - powf(0.999f, (float)t);
    hipLaunchKernelGGL(adam_kernel, dim3(grid1(kParams, 256)), dim3(256), 0, s, P, grad, m, v, sums, c1, c2, lr, W);
    return hipGetLastError();
}

hipError_t mcm_publish(const float *P, uint16_t *W, hipStream_t s)
{
    hipLaunchKernelGGL(publish_kernel, dim3(grid1(kParams, 256)), dim3(256), 0, s, P, W);
    return hipGetLastError();
}

hipError_t mcm_gather(const float *Xs, const float *Ts, const float *Ms, const int *idx, int n, Scale sc,
                      uint16_t *X0, float *T, float *M, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_kernel, dim3(grid1((long)n * kInP, 256)), dim3(256), 0, s, Xs, Ts, Ms, idx, n, sc,
                       X0, T, M);
    return hipGetLastError();
}

hipError_t mcm_standardise(const float *Xraw, int n, Scale sc, uint16_t *X0, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(standardise_kernel, dim3(grid1((long)n * kInP, 256)), dim3(256), 0, s, Xraw, n, sc, X0);
    return hipGetLastError();
}

hipError_t mcm_unscale(float *Y, int n, int ldy, Scale sc, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(unscale_kernel, dim3(grid1((long)n * kOut, 256)), dim3(256), 0, s, Y, n, ldy, sc);
    return hipGetLastError();
}
