// nsfv1_mlp.hip -- learned Newton initial guess: a per-node MLP
// 24 -> 128 -> 128 -> 4 (ReLU), BF16 weights and activations with FP32
// accumulation on the MFMA units (v_mfma_f32_16x16x16_bf16), plus the
// backward pass on the MFMA units with FP32 operands
// (v_mfma_f32_16x16x4_f32), deterministic gradient reduction and Adam used to train it
// on trajectories of the plain solver, and the feature / guess kernels that
// connect it to the discretisation.  No counterpart in the reference.
//
// Parameter blob (FP32 master copy, MLP_NPARAM floats):
//   W1 [128][32] (inputs padded 24 -> 32 with zero columns), b1 [128],
//   W2 [128][128], b2 [128], W3 [16][128] (outputs padded 4 -> 16 with zero
//   rows), b3 [16].
// The BF16 copy (MLP_NPARAM uint16) uses the same offsets; only its weight
// matrices are read.
//
// MFMA mapping (one wavefront = 16 nodes).  The layers are evaluated in
// transposed form, H^T = W X^T, so that the accumulator layout of one layer
// (lane l holds rows 4*(l/16)+j, column l%16) is exactly the B-operand
// layout of the next (k = 4*(l/16)+j, column l%16): activations never leave
// registers.  A operands (weights) come from LDS.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nsfv1_mlp.hpp"

namespace nsfv1 {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __host__ inline uint16_t f2bf(float f)
{
    // round to nearest even (finite inputs)
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

__device__ inline float relu_bf(float z) { return bf2f(f2bf(z > 0.f ? z : 0.f)); }

// ---------------------------------------------------------------------
// forward: y[n][0..3] = W3 relu(W2 relu(W1 x + b1) + b2) + b3
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256)
mlp_forward(const float* __restrict__ P, const uint16_t* __restrict__ WB, const float* __restrict__ X, int64_t n,
            float* __restrict__ Y, float* __restrict__ Z1, float* __restrict__ Z2)
{
    __shared__ __align__(16) uint16_t w1[MLP_H * MLP_KIN];
    __shared__ __align__(16) uint16_t w2[MLP_H * MLP_H];
    __shared__ __align__(16) uint16_t w3[MLP_KOUT * MLP_H];
    __shared__ float bias[2 * MLP_H + MLP_KOUT];
    for (int i = threadIdx.x; i < MLP_H * MLP_KIN; i += blockDim.x) w1[i] = WB[MLP_OFF_W1 + i];
    for (int i = threadIdx.x; i < MLP_H * MLP_H; i += blockDim.x) w2[i] = WB[MLP_OFF_W2 + i];
    for (int i = threadIdx.x; i < MLP_KOUT * MLP_H; i += blockDim.x) w3[i] = WB[MLP_OFF_W3 + i];
    for (int i = threadIdx.x; i < MLP_H; i += blockDim.x) {
        bias[i] = P[MLP_OFF_B1 + i];
        bias[MLP_H + i] = P[MLP_OFF_B2 + i];
    }
    if (threadIdx.x < MLP_KOUT) bias[2 * MLP_H + threadIdx.x] = P[MLP_OFF_B3 + threadIdx.x];
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, kq = (lane >> 4) * 4;
    // the weights are staged once per workgroup (the compiler keeps each
    // lane's A operands in registers across the loop); the workgroup then
    // walks 64-node blocks (grid-stride), so the staging is amortised
    const int64_t nblk = (n + 63) / 64;
    for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int64_t node = (blk * 4 + wave) * 16 + col;
    const bool valid = node < n;

    // layer 1: B = X^T tiles (k = input feature, column = node), two k-blocks
    bf16x4 xb[MLP_KIN / 16];
#pragma unroll
    for (int ks = 0; ks < MLP_KIN / 16; ++ks) {
        float xv[4] = {0.f, 0.f, 0.f, 0.f};
        if (valid) {
            const float4 t = *reinterpret_cast<const float4*>(X + node * MLP_KIN + ks * 16 + kq);
            xv[0] = t.x; xv[1] = t.y; xv[2] = t.z; xv[3] = t.w;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) xb[ks][j] = (short)f2bf(xv[j]);
    }
    bf16x4 h1[MLP_H / 16];
#pragma unroll
    for (int ot = 0; ot < MLP_H / 16; ++ot) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < MLP_KIN / 16; ++ks) {
            const bf16x4 a = *reinterpret_cast<const bf16x4*>(&w1[(ot * 16 + col) * MLP_KIN + ks * 16 + kq]);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, xb[ks], acc, 0, 0, 0);
        }
        float z[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            z[j] = acc[j] + bias[ot * 16 + kq + j];
            h1[ot][j] = (short)f2bf(z[j] > 0.f ? z[j] : 0.f);
        }
        if (Z1 && valid) *reinterpret_cast<float4*>(Z1 + node * MLP_H + ot * 16 + kq) = make_float4(z[0], z[1], z[2], z[3]);
    }
    // layer 2
    bf16x4 h2[MLP_H / 16];
#pragma unroll
    for (int ot = 0; ot < MLP_H / 16; ++ot) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < MLP_H / 16; ++ks) {
            const bf16x4 a = *reinterpret_cast<const bf16x4*>(&w2[(ot * 16 + col) * MLP_H + ks * 16 + kq]);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, h1[ks], acc, 0, 0, 0);
        }
        float z[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            z[j] = acc[j] + bias[MLP_H + ot * 16 + kq + j];
            h2[ot][j] = (short)f2bf(z[j] > 0.f ? z[j] : 0.f);
        }
        if (Z2 && valid) *reinterpret_cast<float4*>(Z2 + node * MLP_H + ot * 16 + kq) = make_float4(z[0], z[1], z[2], z[3]);
    }
    // layer 3 (rows 0..3 of the padded 16-row output)
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < MLP_H / 16; ++ks) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(&w3[col * MLP_H + ks * 16 + kq]);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, h2[ks], acc, 0, 0, 0);
    }
    if (valid && kq == 0)
        *reinterpret_cast<float4*>(Y + node * MLP_OUT) =
            make_float4(acc[0] + bias[2 * MLP_H + 0], acc[1] + bias[2 * MLP_H + 1], acc[2] + bias[2 * MLP_H + 2],
                        acc[3] + bias[2 * MLP_H + 3]);
    }
}

// ---------------------------------------------------------------------
// backward through the activations on the MFMA units (FP32 operands,
// v_mfma_f32_16x16x4_f32), one wavefront per 16 nodes, transposed form:
//   dZ2^T = (W3^T dY^T) * [z2 > 0],  dZ1^T = (W2^T dZ2^T) * [z1 > 0]
// with the BF16-rounded weights the forward used (straight-through for the
// activation rounding).  16x16x4 layouts: A[i][k] in lane (i + 16 k),
// B[k][j] in lane (j + 16 k), D[4 g + r][j] in lane (j + 16 g), element r.
// The dZ2 accumulators feed the W2^T product as B operands without leaving
// registers: MFMA (ot, r) sums over the hidden rows k = 16 ot + 4 g + r
// held by lane group g, and its A operand is gathered for the same k.
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256)
mlp_backward_act(const uint16_t* __restrict__ WB, const float* __restrict__ Z1, const float* __restrict__ Z2,
                 const float* __restrict__ DY, int64_t n, float* __restrict__ DZ1, float* __restrict__ DZ2)
{
    __shared__ __align__(16) uint16_t w2[MLP_H * MLP_H];
    __shared__ __align__(16) uint16_t w3[MLP_OUT * MLP_H];
    for (int i = threadIdx.x; i < MLP_H * MLP_H; i += blockDim.x) w2[i] = WB[MLP_OFF_W2 + i];
    for (int i = threadIdx.x; i < MLP_OUT * MLP_H; i += blockDim.x) w3[i] = WB[MLP_OFF_W3 + i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, g = lane >> 4;
    const int64_t node = ((int64_t)blockIdx.x * 4 + wave) * 16 + col;
    const bool valid = node < n;
    // B of the first product: dY^T (k = output 0..3 = g, column = node)
    const float dyv = valid ? DY[node * MLP_OUT + g] : 0.f;
    f32x4 dz2[MLP_H / 16];
#pragma unroll
    for (int ot = 0; ot < MLP_H / 16; ++ot) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(bf2f(w3[g * MLP_H + ot * 16 + col]), dyv, acc, 0, 0, 0);
        float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        if (valid) z = *reinterpret_cast<const float4*>(Z2 + node * MLP_H + ot * 16 + 4 * g);
        dz2[ot][0] = z.x > 0.f ? acc[0] : 0.f;
        dz2[ot][1] = z.y > 0.f ? acc[1] : 0.f;
        dz2[ot][2] = z.z > 0.f ? acc[2] : 0.f;
        dz2[ot][3] = z.w > 0.f ? acc[3] : 0.f;
        if (valid)
            *reinterpret_cast<float4*>(DZ2 + node * MLP_H + ot * 16 + 4 * g) =
                make_float4(dz2[ot][0], dz2[ot][1], dz2[ot][2], dz2[ot][3]);
    }
#pragma unroll
    for (int o1 = 0; o1 < MLP_H / 16; ++o1) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ot = 0; ot < MLP_H / 16; ++ot) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float a = bf2f(w2[(ot * 16 + 4 * g + r) * MLP_H + o1 * 16 + col]);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, dz2[ot][r], acc, 0, 0, 0);
            }
        }
        float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        if (valid) z = *reinterpret_cast<const float4*>(Z1 + node * MLP_H + o1 * 16 + 4 * g);
        if (valid)
            *reinterpret_cast<float4*>(DZ1 + node * MLP_H + o1 * 16 + 4 * g) =
                make_float4(z.x > 0.f ? acc[0] : 0.f, z.y > 0.f ? acc[1] : 0.f, z.z > 0.f ? acc[2] : 0.f,
                            z.w > 0.f ? acc[3] : 0.f);
    }
}

// ---------------------------------------------------------------------
// weight gradients on the MFMA units: for a chunk of nodes per workgroup,
//   dW1 = dZ1^T X,  dW2 = dZ2^T H1,  dW3 = dY^T H2
// as 16x16 output tiles accumulated over the nodes (k = node, 4 per MFMA):
// W1 8x2 tiles, W2 8x8, W3 1x8 (rows 4..15 zero): 88 tiles, 22 per
// wavefront.  Nodes are staged in LDS 16 at a time; the biases' sums are
// taken by 144 threads on the VALU.  Partials go to part[block][MLP_NPARAM]
// and are reduced in block order afterwards (deterministic).
// ---------------------------------------------------------------------
constexpr int WG_SUB = 16;                      // nodes staged in LDS at a time
constexpr int WG_TILES = 16 + 64 + 8;           // output tiles of W1, W2, W3
constexpr int WG_TPW = WG_TILES / 4;            // tiles per wavefront
constexpr int WG_LD = MLP_H + 4;                // padded LDS row (floats)

__global__ void __launch_bounds__(256)
mlp_wgrad_partial(const float* __restrict__ X, const float* __restrict__ Z1, const float* __restrict__ Z2,
                  const float* __restrict__ DY, const float* __restrict__ DZ1, const float* __restrict__ DZ2,
                  int64_t n, int64_t chunk, float* __restrict__ part)
{
    __shared__ float sx[WG_SUB][MLP_KIN + 4];
    __shared__ float sh1[WG_SUB][WG_LD];
    __shared__ float sh2[WG_SUB][WG_LD];
    __shared__ float sdz1[WG_SUB][WG_LD];
    __shared__ float sdz2[WG_SUB][WG_LD];
    __shared__ float sdy[WG_SUB][MLP_KOUT + 4];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int col = lane & 15, g = lane >> 4;
    f32x4 acc[WG_TPW];
#pragma unroll
    for (int q = 0; q < WG_TPW; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;   // b1 (threads 0..127) and b2 (threads 128..255) gradient sums
    float b3sum = 0.f;  // b3 (threads 0..15)
    const int64_t a = (int64_t)blockIdx.x * chunk;
    const int64_t b = a + chunk < n ? a + chunk : n;
    for (int64_t s0 = a; s0 < b; s0 += WG_SUB) {
        const int cnt = (int)((b - s0) < WG_SUB ? (b - s0) : WG_SUB);
        __syncthreads();
        for (int i = threadIdx.x; i < WG_SUB * MLP_H; i += blockDim.x) {
            const int t = i / MLP_H, f = i % MLP_H;
            const bool ok = t < cnt;
            const int64_t nd = s0 + t;
            sh1[t][f] = ok ? relu_bf(Z1[nd * MLP_H + f]) : 0.f;
            sh2[t][f] = ok ? relu_bf(Z2[nd * MLP_H + f]) : 0.f;
            sdz1[t][f] = ok ? DZ1[nd * MLP_H + f] : 0.f;
            sdz2[t][f] = ok ? DZ2[nd * MLP_H + f] : 0.f;
        }
        for (int i = threadIdx.x; i < WG_SUB * MLP_KIN; i += blockDim.x) {
            const int t = i / MLP_KIN, f = i % MLP_KIN;
            sx[t][f] = t < cnt ? bf2f(f2bf(X[(s0 + t) * MLP_KIN + f])) : 0.f;
        }
        for (int i = threadIdx.x; i < WG_SUB * MLP_KOUT; i += blockDim.x) {
            const int t = i / MLP_KOUT, f = i % MLP_KOUT;
            sdy[t][f] = (t < cnt && f < MLP_OUT) ? DY[(s0 + t) * MLP_OUT + f] : 0.f;
        }
        __syncthreads();
        if (threadIdx.x < MLP_H) {
            for (int t = 0; t < WG_SUB; ++t) bsum += sdz1[t][threadIdx.x];
        } else if (threadIdx.x < 2 * MLP_H) {
            for (int t = 0; t < WG_SUB; ++t) bsum += sdz2[t][threadIdx.x - MLP_H];
        }
        if (threadIdx.x < MLP_KOUT) {
            for (int t = 0; t < WG_SUB; ++t) b3sum += sdy[t][threadIdx.x];
        }
#pragma unroll
        for (int q = 0; q < WG_TPW; ++q) {
            const int tile = wave * WG_TPW + q;
#pragma unroll
            for (int ks = 0; ks < WG_SUB / 4; ++ks) {
                const int nd = ks * 4 + g;
                float av, bv;
                if (tile < 16) {  // W1: rows o (128), columns = inputs (32)
                    av = sdz1[nd][(tile >> 1) * 16 + col];
                    bv = sx[nd][(tile & 1) * 16 + col];
                } else if (tile < 80) {  // W2
                    av = sdz2[nd][((tile - 16) >> 3) * 16 + col];
                    bv = sh1[nd][((tile - 16) & 7) * 16 + col];
                } else {  // W3 (rows = outputs, padded to 16)
                    av = sdy[nd][col];
                    bv = sh2[nd][(tile - 80) * 16 + col];
                }
                acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[q], 0, 0, 0);
            }
        }
    }
    float* out = part + (int64_t)blockIdx.x * MLP_NPARAM;
#pragma unroll
    for (int q = 0; q < WG_TPW; ++q) {
        const int tile = wave * WG_TPW + q;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int64_t p;
            if (tile < 16)
                p = MLP_OFF_W1 + (int64_t)((tile >> 1) * 16 + 4 * g + r) * MLP_KIN + (tile & 1) * 16 + col;
            else if (tile < 80)
                p = MLP_OFF_W2 + (int64_t)(((tile - 16) >> 3) * 16 + 4 * g + r) * MLP_H + ((tile - 16) & 7) * 16 + col;
            else
                p = MLP_OFF_W3 + (int64_t)(4 * g + r) * MLP_H + (tile - 80) * 16 + col;
            out[p] = acc[q][r];
        }
    }
    if (threadIdx.x < MLP_H) out[MLP_OFF_B1 + threadIdx.x] = bsum;
    else if (threadIdx.x < 2 * MLP_H) out[MLP_OFF_B2 + threadIdx.x - MLP_H] = bsum;
    if (threadIdx.x < MLP_KOUT) out[MLP_OFF_B3 + threadIdx.x] = b3sum;
}

__global__ void mlp_reduce(const float* __restrict__ part, int nblk, float* __restrict__ grad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    float s = 0.f;
    for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * MLP_NPARAM + p];
    grad[p] = s;
}

// MSE loss gradient on the 4 outputs: dy = 2 (y - t) / (4 n); loss partials
__global__ void mlp_mse(const float* __restrict__ Y, const float* __restrict__ T, int64_t n, float* __restrict__ DY,
                        float* __restrict__ lpart)
{
    __shared__ float sh[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < n * MLP_OUT) {
        const float e = Y[i] - T[i];
        DY[i] = 2.f * e / (float)(n * MLP_OUT);
        l = e * e;
    }
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = sh[0];
}

// Adam on the FP32 master copy; refreshes the BF16 copy.  Padding entries
// (zero columns of W1, zero rows of W3 and b3) stay zero.
__global__ void mlp_adam(float* __restrict__ P, float* __restrict__ M, float* __restrict__ V,
                         const float* __restrict__ G, float lr, float b1, float b2, float eps, float c1, float c2,
                         uint16_t* __restrict__ WB)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    bool pad = false;
    if (p < MLP_OFF_B1) pad = (p % MLP_KIN) >= MLP_IN;
    else if (p >= MLP_OFF_W3 && p < MLP_OFF_B3) pad = (p - MLP_OFF_W3) / MLP_H >= MLP_OUT;
    else if (p >= MLP_OFF_B3) pad = (p - MLP_OFF_B3) >= MLP_OUT;
    float w = P[p];
    if (!pad) {
        const float g = G[p];
        const float m = b1 * M[p] + (1.f - b1) * g;
        const float v = b2 * V[p] + (1.f - b2) * g * g;
        M[p] = m;
        V[p] = v;
        w -= lr * (m * c1) / (sqrtf(v * c2) + eps);
    } else {
        w = 0.f;
    }
    P[p] = w;
    WB[p] = f2bf(w);
}

__global__ void mlp_pack(const float* __restrict__ P, uint16_t* __restrict__ WB)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < MLP_NPARAM) WB[p] = f2bf(P[p]);
}

// ---------------------------------------------------------------------
// features, guess and target (node-major states [nv][4], FP64)
//
// With the three previous states u1 = u^{n-1}, u2 = u^{n-2}, u3 = u^{n-3},
// the backward differences are evaluated as nested differences,
//   d1 = u - u1,  e1 = u1 - u2,  f1 = u2 - u3,
//   d2 = d1 - e1 (= u - 2 u1 + u2),  d3 = d2 - (e1 - f1) (= u - 3 u1 + 3 u2 - u3),
// (exactly 0 for equal states, no products to contract), the quadratic
// extrapolation is q = u + d1 + d2 (= 3 u - 3 u1 + u2),
// and the step's scale sigma = max(rms of d3 over the velocity components
// of all nodes, 1e-30).  Differences enter the features divided by sigma,
// so one network sees slow and fast phases of a trajectory on one scale.
// ---------------------------------------------------------------------
__device__ inline float clampf(float x) { return fminf(fmaxf(x, -PRED_CLAMP), PRED_CLAMP); }

// partial sums of d3^2 over the velocity components, one per workgroup
// (grid-stride over the nodes, fixed grid: the sum order is fixed)
__global__ void __launch_bounds__(256)
predict_scale_partial(const double* __restrict__ u, const double* __restrict__ u1, const double* __restrict__ u2,
                      const double* __restrict__ u3, int64_t nv, double* __restrict__ part)
{
    __shared__ double sh[256];
    double acc = 0.0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int64_t i = v * 4 + c;
            const double a = u1[i], b = u2[i], e1 = a - b;
            const double d3 = ((u[i] - a) - e1) - (e1 - (b - u3[i]));
            acc += d3 * d3;
        }
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(256)
predict_scale_final(const double* __restrict__ part, int nparts, int64_t nv, double* __restrict__ sigma)
{
    __shared__ double sh[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) acc += part[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double r = nv > 0 ? sqrt(sh[0] / (3.0 * (double)nv)) : 0.0;
        *sigma = r > 1e-30 ? r : 1e-30;
    }
}

// x[v] = clamp(( u, d1 / sigma, d2 / sigma, d3 / sigma,
//                1-ring mean of u - u, (1-ring mean of d1 - d1) / sigma ) * scale),
// 24 values padded to 32.  The 1-ring is the node's column blocks in the
// CSR pattern (rows of component 0), the node itself excluded.
__global__ void predict_features(const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                                 const double* __restrict__ u, const double* __restrict__ u1,
                                 const double* __restrict__ u2, const double* __restrict__ u3, int64_t nv,
                                 const float* __restrict__ scale, const double* __restrict__ sigma,
                                 float* __restrict__ X)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const double is = 1.0 / *sigma;
    double un[4], d1[4], d2[4], d3[4], m0[4] = {0, 0, 0, 0}, m1[4] = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int64_t i = v * 4 + c;
        un[c] = u[i];
        const double a = u1[i], b = u2[i], e1 = a - b;
        d1[c] = un[c] - a;
        d2[c] = d1[c] - e1;
        d3[c] = d2[c] - (e1 - (b - u3[i]));
    }
    const int64_t r0 = rowptr[v * 4], r1 = rowptr[v * 4 + 1];
    int cnt = 0;
    for (int64_t k = r0; k < r1; k += 4) {
        const int64_t w = colidx[k] / 4;
        if (w == v) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double x = u[w * 4 + c];
            m0[c] += x;
            m1[c] += x - u1[w * 4 + c];
        }
        ++cnt;
    }
    const double ic = cnt > 0 ? 1.0 / cnt : 0.0;
    float out[MLP_KIN];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        out[c] = clampf((float)un[c] * scale[c]);
        out[4 + c] = clampf((float)(d1[c] * is) * scale[4 + c]);
        out[8 + c] = clampf((float)(d2[c] * is) * scale[8 + c]);
        out[12 + c] = clampf((float)(d3[c] * is) * scale[12 + c]);
        out[16 + c] = clampf((float)(cnt > 0 ? m0[c] * ic - un[c] : 0.0) * scale[16 + c]);
        out[20 + c] = clampf((float)(cnt > 0 ? (m1[c] * ic - d1[c]) * is : 0.0) * scale[20 + c]);
    }
#pragma unroll
    for (int c = MLP_IN; c < MLP_KIN; ++c) out[c] = 0.f;
#pragma unroll
    for (int q = 0; q < MLP_KIN; q += 4)
        *reinterpret_cast<float4*>(X + v * MLP_KIN + q) = make_float4(out[q], out[q + 1], out[q + 2], out[q + 3]);
}

// guess = u + d1 + d2 + y * out_scale * sigma (per component)
__global__ void predict_guess(const double* __restrict__ u, const double* __restrict__ u1,
                              const double* __restrict__ u2, const float* __restrict__ Y,
                              const float* __restrict__ oscale, const double* __restrict__ sigma, int64_t nv,
                              double* __restrict__ g)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv * 4) return;
    const int c = (int)(t & 3);
    const double d1 = u[t] - u1[t];
    const double d2 = d1 - (u1[t] - u2[t]);
    g[t] = u[t] + d1 + d2 + (double)Y[t] * (double)oscale[c] * *sigma;
}

// training target: t = clamp(((u^{n+1} - u) - d1 - d2) / (sigma * out_scale)),
// 0 for a component whose out_scale is 0 (not corrected)
__global__ void predict_target(const double* __restrict__ un1, const double* __restrict__ u,
                               const double* __restrict__ u1, const double* __restrict__ u2,
                               const float* __restrict__ oscale, const double* __restrict__ sigma, int64_t nv,
                               float* __restrict__ T)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nv * 4) return;
    const int c = (int)(t & 3);
    const double os = (double)oscale[c];
    const double d1 = u[t] - u1[t];
    const double d2 = d1 - (u1[t] - u2[t]);
    const double r = ((un1[t] - u[t]) - d1) - d2;
    T[t] = os > 0.0 ? clampf((float)(r / (*sigma * os))) : 0.f;
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
static inline unsigned nblocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

hipError_t launch_mlp_forward(const float* P, const uint16_t* WB, const float* X, int64_t n, float* Y, float* Z1,
                              float* Z2, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    // one workgroup per CU (256 CUs): each holds the weights in registers
    // (one wave per SIMD) and walks its share of the 64-node blocks
    const unsigned g = nblocks(n, 64) < 256u ? nblocks(n, 64) : 256u;
    mlp_forward<<<g, 256, 0, s>>>(P, WB, X, n, Y, Z1, Z2);
    return hipGetLastError();
}

int64_t mlp_work_floats(int64_t n) { return 2 * n * MLP_H + (int64_t)MLP_WG_BLOCKS * MLP_NPARAM + n * MLP_OUT + 1024; }

hipError_t launch_mlp_backward(const uint16_t* WB, const float* X, const float* Z1, const float* Z2, const float* DY,
                               int64_t n, float* grad, float* work, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    float* dz1 = work;
    float* dz2 = dz1 + n * MLP_H;
    float* part = dz2 + n * MLP_H;
    mlp_backward_act<<<nblocks(n, 64), 256, 0, s>>>(WB, Z1, Z2, DY, n, dz1, dz2);
    const int64_t chunk = (n + MLP_WG_BLOCKS - 1) / MLP_WG_BLOCKS;
    mlp_wgrad_partial<<<MLP_WG_BLOCKS, 256, 0, s>>>(X, Z1, Z2, DY, dz1, dz2, n, chunk, part);
    mlp_reduce<<<nblocks(MLP_NPARAM, 256), 256, 0, s>>>(part, MLP_WG_BLOCKS, grad);
    return hipGetLastError();
}

hipError_t launch_mlp_mse(const float* Y, const float* T, int64_t n, float* DY, float* lpart, int* nparts,
                          hipStream_t s)
{
    const unsigned nb = nblocks(n * MLP_OUT, 256);
    *nparts = (int)nb;
    mlp_mse<<<nb, 256, 0, s>>>(Y, T, n, DY, lpart);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float* P, float* M, float* V, const float* G, int step, float lr, float b1, float b2,
                           float eps, uint16_t* WB, hipStream_t s)
{
    const float c1 = 1.f / (1.f - powf(b1, (float)step)), c2 = 1.f / (1.f - powf(b2, (float)step));
    mlp_adam<<<nblocks(MLP_NPARAM, 256), 256, 0, s>>>(P, M, V, G, lr, b1, b2, eps, c1, c2, WB);
    return hipGetLastError();
}

hipError_t launch_mlp_pack(const float* P, uint16_t* WB, hipStream_t s)
{
    mlp_pack<<<nblocks(MLP_NPARAM, 256), 256, 0, s>>>(P, WB);
    return hipGetLastError();
}

hipError_t launch_predict_scale(const double* u, const double* u1, const double* u2, const double* u3, int64_t nv,
                                double* part, double* sigma, hipStream_t s)
{
    predict_scale_partial<<<PRED_SCALE_BLOCKS, 256, 0, s>>>(u, u1, u2, u3, nv, part);
    predict_scale_final<<<1, 256, 0, s>>>(part, PRED_SCALE_BLOCKS, nv, sigma);
    return hipGetLastError();
}

hipError_t launch_predict_features(const int64_t* rowptr, const int32_t* colidx, const double* u, const double* u1,
                                   const double* u2, const double* u3, int64_t nv, const float* scale,
                                   const double* sigma, float* X, hipStream_t s)
{
    if (nv <= 0) return hipSuccess;
    predict_features<<<nblocks(nv, 256), 256, 0, s>>>(rowptr, colidx, u, u1, u2, u3, nv, scale, sigma, X);
    return hipGetLastError();
}

hipError_t launch_predict_guess(const double* u, const double* u1, const double* u2, const float* Y,
                                const float* oscale, const double* sigma, int64_t nv, double* g, hipStream_t s)
{
    if (nv <= 0) return hipSuccess;
    predict_guess<<<nblocks(nv * 4, 256), 256, 0, s>>>(u, u1, u2, Y, oscale, sigma, nv, g);
    return hipGetLastError();
}

hipError_t launch_predict_target(const double* un1, const double* u, const double* u1, const double* u2,
                                 const float* oscale, const double* sigma, int64_t nv, float* T, hipStream_t s)
{
    if (nv <= 0) return hipSuccess;
    predict_target<<<nblocks(nv * 4, 256), 256, 0, s>>>(un1, u, u1, u2, oscale, sigma, nv, T);
    return hipGetLastError();
}

}  // namespace nsfv1
