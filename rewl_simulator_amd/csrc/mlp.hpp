// Warm-start MLP: forward, backward and Adam as gfx950 kernels.
//
// Network and conventions are those of include/rewl_mlp.h.  One workgroup of
// four wavefronts takes tiles of 32 points; in every hidden layer wavefront w
// owns output columns [32w, 32w + 32) and computes them for all 32 points with
// v_mfma_f32_32x32x16_bf16 (8 MFMAs for K = 128).  Operand layout of that
// instruction (64 lanes, l = lane):
//   A (32 x 16): lane l holds A[l % 32][8 (l / 32) + e], e = 0..7
//   B (16 x 32): lane l holds B[8 (l / 32) + e][l % 32]
//   D (32 x 32): lane l, register v holds D[(v % 4) + 4 (l / 32) + 8 (v / 4)][l % 32]
// so each lane owns one column j of every hidden layer and keeps z2 and z3 for
// its 16 points in registers through the backward pass (z1 is recomputed from
// the four inputs).  The weight operands (rows of W2, W3 and of their
// transposes for the backward products) are read from L1/L2 as the MFMAs need
// them, which leaves the registers to the dW accumulators; activations and
// back-propagated errors go
// through LDS in row ([point][feature]) and transposed ([feature][point])
// layouts, both padded so that the 16-byte operand reads spread over the banks.
// dW2 and dW3 accumulate in MFMA accumulators across a workgroup's tiles (K =
// points); every workgroup writes one partial gradient and a second kernel
// sums the partials in a fixed order (deterministic) and applies Adam.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "philox.hpp"

namespace rewl {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4v __attribute__((ext_vector_type(4)));

constexpr int NIN = 4, H = 128, TILE = 32;
// FP32 master parameter vector
constexpr int O_W1 = 0, O_B1 = O_W1 + H * NIN, O_W2 = O_B1 + H, O_B2 = O_W2 + H * H, O_W3 = O_B2 + H,
              O_B3 = O_W3 + H * H, O_W4 = O_B3 + H, O_B4 = O_W4 + H, NP = O_B4 + 1;
constexpr int NPART = NP + 1;   // partial gradient + partial loss
// bf16 weight copies (row-major, plus transposes of W2 and W3)
constexpr int Q_W1 = 0, Q_W2 = Q_W1 + H * NIN, Q_W2T = Q_W2 + H * H, Q_W3 = Q_W2T + H * H, Q_W3T = Q_W3 + H * H,
              Q_W4 = Q_W3T + H * H, NQ = Q_W4 + H;
constexpr int RS = H + 8;      // row tile stride, bf16 elements (272 bytes)
constexpr int TS = TILE + 8;   // transposed tile stride (80 bytes)
constexpr uint32_t DOMAIN_MLP = 3;

// FP32 -> bf16, round to nearest even (NaN kept quiet).
__host__ __device__ __forceinline__ uint16_t bf16_bits(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf16_float(uint16_t b) {
    return __builtin_bit_cast(float, (uint32_t)b << 16);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.707106781186547524f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.707106781186547524f));
    const float pdf = expf(-0.5f * z * z) * 0.398942280401432678f;
    return cdf + z * pdf;
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4v*>(p));
}

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

struct Args {
    const float* theta;     // FP32 master parameters (biases are read from here)
    const uint16_t* wq;     // bf16 weight copies
    const float* x;         // [n][4]
    const float* t;         // [n] targets (training)
    const float* w;         // [n] weights, or nullptr for 1
    int64_t n;              // data points
    int64_t npts;           // points of this launch
    int32_t sample;         // 1: launch point k reads data point philox(k) mod n
    uint32_t k0, k1;        // Philox key of the minibatch stream
    uint64_t step;
    float inv_wsum;         // loss normaliser
    float* y;               // outputs per launch point, or nullptr
    float* part;            // [gridDim.x][NPART] (training)
};

__device__ __forceinline__ int64_t source_point(const Args& a, int64_t k) {
    if (!a.sample) return k;
    const u32x4 r = philox4x32_10(u32x4{(uint32_t)k, (uint32_t)(k >> 32), DOMAIN_MLP, (uint32_t)a.step}, a.k0, a.k1);
    return (int64_t)(((uint64_t)r.x * (uint64_t)a.n) >> 32);
}

// row of the D layout held in register v by lane half hf
__device__ __forceinline__ int drow(int v, int hf) { return (v & 3) + 4 * hf + 8 * (v >> 2); }

template <bool TRAIN>
__global__ void __launch_bounds__(256) k_mlp(Args a) {
    __shared__ float xs[TILE][NIN];
    __shared__ float ts[TILE], ws[TILE], dys[TILE];
    __shared__ float yred[4][TILE];
    __shared__ __align__(16) uint16_t a1r[TILE * RS], a2r[TILE * RS], dzr[TILE * RS];
    __shared__ __align__(16) uint16_t a1t[H * TS], a2t[H * TS], dzt[H * TS];

    const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, hf = l >> 5, lc = l & 31;
    const int j = 32 * wv + lc;   // this lane's column in every hidden layer

    float w1j[NIN];
#pragma unroll
    for (int c = 0; c < NIN; ++c) w1j[c] = bf16_float(a.wq[Q_W1 + j * NIN + c]);
    const float b1j = a.theta[O_B1 + j], b2j = a.theta[O_B2 + j], b3j = a.theta[O_B3 + j];
    const float w4j = bf16_float(a.wq[Q_W4 + j]), b4 = a.theta[O_B4];
    // weight operands of this lane's column (row j of W, or of W^T for the
    // backward products), re-read per tile from L1/L2 to keep registers free
    // for the dW accumulators
    const uint16_t* qW2 = a.wq + Q_W2 + j * H + 8 * hf;
    const uint16_t* qW3 = a.wq + Q_W3 + j * H + 8 * hf;
    const uint16_t* qW2T = a.wq + Q_W2T + j * H + 8 * hf;
    const uint16_t* qW3T = a.wq + Q_W3T + j * H + 8 * hf;
    f32x16 G2[4], G3[4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int v = 0; v < 16; ++v) G2[c][v] = G3[c][v] = 0.f;
    float gw4 = 0.f, gb3 = 0.f, gb2 = 0.f, gb1 = 0.f, gb4 = 0.f, lossacc = 0.f;
    float gw1[NIN] = {0.f, 0.f, 0.f, 0.f};

    const int64_t ntiles = (a.npts + TILE - 1) / TILE;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t base = tile * TILE;
        __syncthreads();   // the previous tile's readers are done with every buffer
        if (tid < TILE) {
            const int64_t k = base + tid;
            float xv[NIN] = {0.f, 0.f, 0.f, 0.f}, tv = 0.f, wt = 0.f;
            if (k < a.npts) {
                const int64_t src = source_point(a, k);
#pragma unroll
                for (int c = 0; c < NIN; ++c) xv[c] = a.x[src * NIN + c];
                if (TRAIN) {
                    tv = a.t[src];
                    wt = a.w ? a.w[src] : 1.f;
                }
            }
#pragma unroll
            for (int c = 0; c < NIN; ++c) xs[tid][c] = xv[c];
            ts[tid] = tv;
            ws[tid] = wt;
        }
        __syncthreads();

        // ---- layer 1 (VALU, D layout) ------------------------------------------
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int b = drow(v, hf);
            float z = b1j;
#pragma unroll
            for (int c = 0; c < NIN; ++c) z = z + w1j[c] * xs[b][c];
            const uint16_t q = bf16_bits(gelu(z));
            a1r[b * RS + j] = q;
            a1t[j * TS + b] = q;
        }
        __syncthreads();

        // ---- layer 2 ----------------------------------------------------------
        f32x16 acc;
        bf16x8 wb[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) wb[s] = load8(qW2 + 16 * s);
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[v] = 0.f;
#pragma unroll
        for (int s = 0; s < 8; ++s) acc = mfma(load8(a1r + lc * RS + 16 * s + 8 * hf), wb[s], acc);
        float z2[16];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int b = drow(v, hf);
            z2[v] = acc[v] + b2j;
            const uint16_t q = bf16_bits(gelu(z2[v]));
            a2r[b * RS + j] = q;
            a2t[j * TS + b] = q;
        }
        __syncthreads();

        // ---- layer 3 and the output unit -------------------------------------
#pragma unroll
        for (int s = 0; s < 8; ++s) wb[s] = load8(qW3 + 16 * s);
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[v] = 0.f;
#pragma unroll
        for (int s = 0; s < 8; ++s) acc = mfma(load8(a2r + lc * RS + 16 * s + 8 * hf), wb[s], acc);
        float z3[16];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            z3[v] = acc[v] + b3j;
            float p = w4j * gelu(z3[v]);
#pragma unroll
            for (int o = 16; o > 0; o >>= 1) p += __shfl_xor(p, o);   // over the 32 columns of this half
            if (lc == 0) yred[wv][drow(v, hf)] = p;
        }
        __syncthreads();
        if (tid < TILE) {
            const float yv = b4 + ((yred[0][tid] + yred[1][tid]) + (yred[2][tid] + yred[3][tid]));
            const int64_t k = base + tid;
            if (a.y && k < a.npts) a.y[k] = yv;
            if (TRAIN) {
                const float d = yv - ts[tid];
                dys[tid] = 2.f * ws[tid] * d * a.inv_wsum;
                lossacc += ws[tid] * d * d;
            }
        }
        if (!TRAIN) continue;
        __syncthreads();

        // ---- backward: output unit and layer 3 -------------------------------
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int b = drow(v, hf);
            const float dy = dys[b];
            gw4 += dy * gelu(z3[v]);
            const float dz = dy * w4j * gelu_grad(z3[v]);
            gb3 += dz;
            const uint16_t q = bf16_bits(dz);
            dzr[b * RS + j] = q;
            dzt[j * TS + b] = q;
        }
        if (tid == 0) {
            float sdy = 0.f;
            for (int b = 0; b < TILE; ++b) sdy += dys[b];
            gb4 += sdy;
        }
        __syncthreads();
        // dW3[i][jj] += sum_b dz3[b][i] a2[b][jj]: this wavefront owns rows i in [32 wv, 32 wv + 32)
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int s = 0; s < 2; ++s)
                G3[c] = mfma(load8(dzt + (32 * wv + lc) * TS + 16 * s + 8 * hf),
                             load8(a2t + (32 * c + lc) * TS + 16 * s + 8 * hf), G3[c]);
        // dL/da2 = dz3 W3
#pragma unroll
        for (int s = 0; s < 8; ++s) wb[s] = load8(qW3T + 16 * s);
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[v] = 0.f;
#pragma unroll
        for (int s = 0; s < 8; ++s) acc = mfma(load8(dzr + lc * RS + 16 * s + 8 * hf), wb[s], acc);
        float dz2[16];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            dz2[v] = acc[v] * gelu_grad(z2[v]);
            gb2 += dz2[v];
        }
        __syncthreads();   // dz3 fully consumed
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int b = drow(v, hf);
            const uint16_t q = bf16_bits(dz2[v]);
            dzr[b * RS + j] = q;
            dzt[j * TS + b] = q;
        }
        __syncthreads();

        // ---- backward: layer 2 and layer 1 -----------------------------------
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int s = 0; s < 2; ++s)
                G2[c] = mfma(load8(dzt + (32 * wv + lc) * TS + 16 * s + 8 * hf),
                             load8(a1t + (32 * c + lc) * TS + 16 * s + 8 * hf), G2[c]);
#pragma unroll
        for (int s = 0; s < 8; ++s) wb[s] = load8(qW2T + 16 * s);
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[v] = 0.f;
#pragma unroll
        for (int s = 0; s < 8; ++s) acc = mfma(load8(dzr + lc * RS + 16 * s + 8 * hf), wb[s], acc);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int b = drow(v, hf);
            float z1 = b1j;   // recomputed: the same operations as the forward pass
#pragma unroll
            for (int c = 0; c < NIN; ++c) z1 = z1 + w1j[c] * xs[b][c];
            const float dz1 = acc[v] * gelu_grad(z1);
            gb1 += dz1;
#pragma unroll
            for (int c = 0; c < NIN; ++c) gw1[c] += dz1 * xs[b][c];
        }
    }
    if (!TRAIN) return;

    // ---- this workgroup's partial gradient ------------------------------------
    float* P = a.part + (size_t)blockIdx.x * NPART;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int i = 32 * wv + drow(v, hf);
            P[O_W2 + i * H + 32 * c + lc] = G2[c][v];
            P[O_W3 + i * H + 32 * c + lc] = G3[c][v];
        }
    // lanes l and l + 32 hold the same column
    gw4 += __shfl_xor(gw4, 32);
    gb3 += __shfl_xor(gb3, 32);
    gb2 += __shfl_xor(gb2, 32);
    gb1 += __shfl_xor(gb1, 32);
#pragma unroll
    for (int c = 0; c < NIN; ++c) gw1[c] += __shfl_xor(gw1[c], 32);
    if (hf == 0) {
        P[O_W4 + j] = gw4;
        P[O_B3 + j] = gb3;
        P[O_B2 + j] = gb2;
        P[O_B1 + j] = gb1;
#pragma unroll
        for (int c = 0; c < NIN; ++c) P[O_W1 + j * NIN + c] = gw1[c];
    }
    if (wv == 0) {
        // loss: lanes 0..31 of wavefront 0 hold the per-point sums
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) lossacc += __shfl_xor(lossacc, o);
        if (l == 0) {
            P[O_B4] = gb4;
            P[NP] = lossacc;
        }
    }
}

// grad[p] = sum over partials g = 0..G-1 (in that order); p = NP is the loss.
__global__ void k_mlp_reduce(const float* part, int G, float* grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPART) return;
    float s = 0.f;
    for (int g = 0; g < G; ++g) s += part[(size_t)g * NPART + p];
    grad[p] = s;
}

__device__ __forceinline__ void store_quantised(uint16_t* wq, int p, float th) {
    const uint16_t q = bf16_bits(th);
    if (p < O_B1) {
        wq[Q_W1 + (p - O_W1)] = q;
    } else if (p >= O_W2 && p < O_B2) {
        const int i = (p - O_W2) / H, jj = (p - O_W2) % H;
        wq[Q_W2 + i * H + jj] = q;
        wq[Q_W2T + jj * H + i] = q;
    } else if (p >= O_W3 && p < O_B3) {
        const int i = (p - O_W3) / H, jj = (p - O_W3) % H;
        wq[Q_W3 + i * H + jj] = q;
        wq[Q_W3T + jj * H + i] = q;
    } else if (p >= O_W4 && p < O_B4) {
        wq[Q_W4 + (p - O_W4)] = q;
    }
}

__global__ void k_mlp_quantise(const float* theta, uint16_t* wq) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < NP) store_quantised(wq, p, theta[p]);
}

// Adam (Kingma & Ba) with bias correction; refreshes the bf16 copies.
__global__ void k_mlp_adam(float* theta, float* m, float* v, const float* grad, float lr, float beta1, float beta2,
                           float eps, float bc1, float bc2, uint16_t* wq) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NP) return;
    const float g = grad[p];
    const float mp = beta1 * m[p] + (1.f - beta1) * g;
    const float vp = beta2 * v[p] + (1.f - beta2) * g * g;
    m[p] = mp;
    v[p] = vp;
    const float th = theta[p] - lr * (mp / bc1) / (sqrtf(vp / bc2) + eps);
    theta[p] = th;
    store_quantised(wq, p, th);
}

}  // namespace mlp
}  // namespace rewl
