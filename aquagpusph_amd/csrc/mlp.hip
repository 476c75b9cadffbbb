// The eddy-viscosity network of the LES closure on gfx950: a 7 -> 64 -> 64
// -> 1 perceptron with GELU hidden units, run and trained with BF16 matrix
// instructions (v_mfma_f32_16x16x16_bf16) and FP32 accumulation.
//
// Parameters (FP32 master copy, one array of NPARAM = 5313 floats):
//   W1[64][16]  inputs padded from 7 to 16 (columns 7..15 stay zero)
//   b1[64], W2[64][64] (W2[out][in]), b2[64], w3[64], b3
// The matrix operands (inputs, W1, W2, hidden activations and the back-
// propagated pre-activation gradients) are rounded to BF16 where they enter
// a matrix instruction; everything else (biases, GELU, the output layer,
// the loss, the Adam update) is FP32.
//
// One wavefront handles 16 samples at a time.  With lane l = 16 g + r, the
// 16x16x16 instruction takes A[r][4g + j], B[4g + j][r] and returns
// D[4g + j][r], j = 0..3; so the accumulator of a layer (samples x units)
// is already the A operand of the transposed products of the backward pass
// and the B operand of the next weight gradient; only two transposes go
// through LDS (hidden activations into layer 2, dZ2 into dH1).
//
// Training is deterministic: every wave sums the gradients of its own
// fixed tiles in registers and writes one partial per wave; a second
// kernel adds the partials in wave order; Adam then updates the master copy.
#include "aq_device.hpp"
#include "aq_internal.hpp"

namespace aq {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NIN = 7;     // network inputs
constexpr int KIN = 16;    // padded input width
constexpr int NH = 64;     // hidden units per layer
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NH * KIN;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NH;
constexpr int NPARAM = OFF_B3 + 1;
constexpr int WAVES = 4;  // per workgroup

static_assert(NPARAM == 5313, "parameter layout");

// FP32 -> BF16, round to nearest even (finite inputs)
__device__ __forceinline__ short bf16(float x)
{
    uint32_t b = __float_as_uint(x);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (short)(b >> 16);
}

// GELU, tanh form, and its derivative
__device__ __forceinline__ float gelu(float x)
{
    const float c = 0.7978845608028654f;  // sqrt(2/pi)
    const float u = c * (x + 0.044715f * x * x * x);
    return 0.5f * x * (1.f + tanhf(u));
}

__device__ __forceinline__ float gelu_d(float x)
{
    const float c = 0.7978845608028654f;
    const float u = c * (x + 0.044715f * x * x * x);
    const float t = tanhf(u);
    const float du = c * (1.f + 3.f * 0.044715f * x * x);
    return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * du;
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// sum over the 16 lanes of a row group (lanes 16 g .. 16 g + 15)
__device__ __forceinline__ float sum16(float v)
{
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    return v;
}

// sum over the 4 row groups (lanes r, r + 16, r + 32, r + 48)
__device__ __forceinline__ float sum4g(float v)
{
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

struct Fwd {
    float z1[4][4];  // Z1[4g + j][16 t + r] (pre-activation, bias added)
    float h1[4][4];  // GELU(Z1)
    float z2[4][4];
    float y[4];      // outputs of samples 4 g + j (every lane of the group)
};

// Forward pass of samples s0 .. s0 + 15 (X rows beyond M are zero).  lds:
// 16 x 64 floats of this wave.
__device__ __forceinline__ void forward_tile(const float* __restrict__ X,
                                             uint32_t M, uint32_t s0,
                                             const float* __restrict__ P,
                                             float* lds, Fwd& f)
{
    const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    // layer 1: A = X[s0 + r][4 g + j]
    bf16x4 a;
    {
        const uint32_t s = s0 + r;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = 4 * g + j;
            const float x = (s < M && k < NIN) ? X[(size_t)s * 8 + k] : 0.f;
            a[j] = bf16(x);
        }
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        bf16x4 b;
#pragma unroll
        for (int j = 0; j < 4; j++)
            b[j] = bf16(P[OFF_W1 + (16 * t + r) * KIN + 4 * g + j]);
        f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
        acc = mfma(a, b, acc);
        const float bias = P[OFF_B1 + 16 * t + r];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            f.z1[t][j] = acc[j] + bias;
            f.h1[t][j] = gelu(f.z1[t][j]);
            lds[(4 * g + j) * NH + 16 * t + r] = f.h1[t][j];
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // layer 2: A chunk c = H1[r][16 c + 4 g + j]
    f32x4 acc2[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc2[t] = f32x4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
    for (int c = 0; c < 4; c++) {
        bf16x4 ac;
#pragma unroll
        for (int j = 0; j < 4; j++)
            ac[j] = bf16(lds[r * NH + 16 * c + 4 * g + j]);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            bf16x4 b;
#pragma unroll
            for (int j = 0; j < 4; j++)
                b[j] = bf16(
                    P[OFF_W2 + (16 * t + r) * NH + 16 * c + 4 * g + j]);
            acc2[t] = mfma(ac, b, acc2[t]);
        }
    }
    __builtin_amdgcn_wave_barrier();
    // output layer in FP32
    float part[4] = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const float bias = P[OFF_B2 + 16 * t + r];
        const float w = P[OFF_W3 + 16 * t + r];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            f.z2[t][j] = acc2[t][j] + bias;
            part[j] += gelu(f.z2[t][j]) * w;
        }
    }
    const float b3 = P[OFF_B3];
#pragma unroll
    for (int j = 0; j < 4; j++) f.y[j] = sum16(part[j]) + b3;
}

__global__ void __launch_bounds__(64 * WAVES)
    mlp_forward_k(const float* __restrict__ X, uint32_t M,
                  const float* __restrict__ P, float* __restrict__ y)
{
    __shared__ float lds_all[WAVES][16 * NH];
    const int wv = threadIdx.x >> 6;
    const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    const uint32_t ntiles = (M + 15) / 16;
    const uint32_t gw = blockIdx.x * WAVES + wv, nw = gridDim.x * WAVES;
    for (uint32_t tile = gw; tile < ntiles; tile += nw) {
        Fwd f;
        forward_tile(X, M, tile * 16, P, lds_all[wv], f);
        if (r == 0) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t s = tile * 16 + 4 * g + j;
                if (s < M) y[s] = f.y[j];
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Loss (mean squared error over M samples) and its gradient: every wave
// writes the gradient of its tiles to part[wave][0 .. NPARAM) and its share
// of the loss to part[wave][NPARAM].
__global__ void __launch_bounds__(64 * WAVES)
    mlp_grad_k(const float* __restrict__ X, const float* __restrict__ target,
               uint32_t M, const float* __restrict__ P,
               float* __restrict__ part)
{
    __shared__ float lds_all[WAVES][16 * NH];
    const int wv = threadIdx.x >> 6;
    const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    float* lds = lds_all[wv];
    const uint32_t ntiles = (M + 15) / 16;
    const uint32_t gw = blockIdx.x * WAVES + wv, nw = gridDim.x * WAVES;
    const float inv_m = 2.f / (float)M;
    f32x4 dw2[4][4], dw1[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
        dw1[a] = f32x4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int b = 0; b < 4; b++) dw2[a][b] = f32x4{ 0.f, 0.f, 0.f, 0.f };
    }
    float db1[4] = { 0.f, 0.f, 0.f, 0.f }, db2[4] = { 0.f, 0.f, 0.f, 0.f };
    float dw3[4] = { 0.f, 0.f, 0.f, 0.f };
    float db3 = 0.f, loss = 0.f;
    for (uint32_t tile = gw; tile < ntiles; tile += nw) {
        const uint32_t s0 = tile * 16;
        Fwd f;
        forward_tile(X, M, s0, P, lds, f);
        // dL/dy for the samples 4 g + j
        float dy[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t s = s0 + 4 * g + j;
            const float e = (s < M) ? f.y[j] - target[s] : 0.f;
            dy[j] = e * inv_m;
            if (r == 0) loss += e * e / (float)M;
        }
        // output layer, dZ2 = dy w3 GELU'(Z2)
        float dz2[4][4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const float w = P[OFF_W3 + 16 * t + r];
            float pw = 0.f, pb = 0.f;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                pw += dy[j] * gelu(f.z2[t][j]);
                dz2[t][j] = dy[j] * w * gelu_d(f.z2[t][j]);
                pb += dz2[t][j];
                lds[(4 * g + j) * NH + 16 * t + r] = dz2[t][j];
            }
            dw3[t] += pw;
            db2[t] += pb;
        }
        if (r == 0) db3 += dy[0] + dy[1] + dy[2] + dy[3];
        // dW2[o][i] += sum_s dZ2[s][o] H1[s][i]: A = dZ2 tile (o),
        // B = H1 tile (i), both in their accumulator layout
#pragma unroll
        for (int to = 0; to < 4; to++) {
            bf16x4 a;
#pragma unroll
            for (int j = 0; j < 4; j++) a[j] = bf16(dz2[to][j]);
#pragma unroll
            for (int ti = 0; ti < 4; ti++) {
                bf16x4 b;
#pragma unroll
                for (int j = 0; j < 4; j++) b[j] = bf16(f.h1[ti][j]);
                dw2[to][ti] = mfma(a, b, dw2[to][ti]);
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        // dH1 = dZ2 W2: A chunk c = dZ2[r][16 c + 4 g + j] (from LDS),
        // B = W2[16 c + 4 g + j][16 ti + r]
        f32x4 dh1[4];
#pragma unroll
        for (int t = 0; t < 4; t++) dh1[t] = f32x4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int c = 0; c < 4; c++) {
            bf16x4 a;
#pragma unroll
            for (int j = 0; j < 4; j++)
                a[j] = bf16(lds[r * NH + 16 * c + 4 * g + j]);
#pragma unroll
            for (int ti = 0; ti < 4; ti++) {
                bf16x4 b;
#pragma unroll
                for (int j = 0; j < 4; j++)
                    b[j] = bf16(
                        P[OFF_W2 + (16 * c + 4 * g + j) * NH + 16 * ti + r]);
                dh1[ti] = mfma(a, b, dh1[ti]);
            }
        }
        __builtin_amdgcn_wave_barrier();
        // dZ1 = dH1 GELU'(Z1); dW1[h][f] += sum_s dZ1[s][h] X[s][f]
        bf16x4 xb;  // B = X[s0 + 4 g + j][r]
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t s = s0 + 4 * g + j;
            xb[j] = bf16((s < M && r < NIN) ? X[(size_t)s * 8 + r] : 0.f);
        }
#pragma unroll
        for (int t = 0; t < 4; t++) {
            bf16x4 a;
            float pb = 0.f;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float dz1 = dh1[t][j] * gelu_d(f.z1[t][j]);
                pb += dz1;
                a[j] = bf16(dz1);
            }
            db1[t] += pb;
            dw1[t] = mfma(a, xb, dw1[t]);
        }
    }
    // the wave's partial gradient
    float* out = part + (size_t)gw * (NPARAM + 1);
#pragma unroll
    for (int t = 0; t < 4; t++) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            out[OFF_W1 + (16 * t + 4 * g + j) * KIN + r] = dw1[t][j];
#pragma unroll
            for (int ti = 0; ti < 4; ti++)
                out[OFF_W2 + (16 * t + 4 * g + j) * NH + 16 * ti + r] =
                    dw2[t][ti][j];
        }
        const float b1s = sum4g(db1[t]), b2s = sum4g(db2[t]);
        const float w3s = sum4g(dw3[t]);
        if (g == 0) {
            out[OFF_B1 + 16 * t + r] = b1s;
            out[OFF_B2 + 16 * t + r] = b2s;
            out[OFF_W3 + 16 * t + r] = w3s;
        }
    }
    // db3 and the loss live in the r == 0 lanes
    float b3s = __shfl_xor(db3, 16) + db3;
    b3s += __shfl_xor(b3s, 32);
    float ls = __shfl_xor(loss, 16) + loss;
    ls += __shfl_xor(ls, 32);
    if (l == 0) {
        out[OFF_B3] = b3s;
        out[NPARAM] = ls;
    }
}

// grad[k] = sum over waves of part[w][k] (k < NPARAM + 1: the last is the
// loss), in wave order
__global__ void mlp_reduce_k(const float* __restrict__ part, uint32_t nw,
                             float* __restrict__ grad)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > (uint32_t)NPARAM) return;
    float s = 0.f;
    for (uint32_t w = 0; w < nw; w++) s += part[(size_t)w * (NPARAM + 1) + k];
    grad[k] = s;
}

// Adam (Kingma & Ba), step t >= 1
__global__ void mlp_adam_k(float* __restrict__ P, const float* __restrict__ G,
                           float* __restrict__ m, float* __restrict__ v,
                           float lr, float beta1, float beta2, float eps,
                           float c1, float c2)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (uint32_t)NPARAM) return;
    const float gk = G[k];
    const float mk = beta1 * m[k] + (1.f - beta1) * gk;
    const float vk = beta2 * v[k] + (1.f - beta2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    P[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

}  // namespace

int mlp_nparam() { return NPARAM; }

// waves used by the gradient kernel for M samples
unsigned mlp_grad_waves(uint32_t M)
{
    const uint32_t ntiles = std::max(1u, (M + 15) / 16);
    const uint32_t nw = std::min<uint32_t>(1024u, ntiles);
    return (nw + WAVES - 1) / WAVES * WAVES;
}

size_t mlp_workspace(uint32_t M)
{
    return (size_t)mlp_grad_waves(M) * (NPARAM + 1) * sizeof(float);
}

void launch_mlp_forward(const float* X, uint32_t M, const float* P, float* y,
                        hipStream_t s)
{
    if (!M) return;
    const uint32_t ntiles = (M + 15) / 16;
    const unsigned nb =
        std::max(1u, std::min<unsigned>(2048, (ntiles + WAVES - 1) / WAVES));
    hipLaunchKernelGGL(mlp_forward_k, dim3(nb), dim3(64 * WAVES), 0, s, X, M,
                       P, y);
}

// grad_out[NPARAM + 1]: the gradient and, last, the loss
void launch_mlp_grad(const float* X, const float* target, uint32_t M,
                     const float* P, float* grad_out, void* ws, hipStream_t s)
{
    AQ_REQUIRE(M > 0, "mlp: no samples");
    const unsigned nw = mlp_grad_waves(M);
    hipLaunchKernelGGL(mlp_grad_k, dim3(nw / WAVES), dim3(64 * WAVES), 0, s,
                       X, target, M, P, (float*)ws);
    hipLaunchKernelGGL(mlp_reduce_k, dim3((NPARAM + 1 + 255) / 256),
                       dim3(256), 0, s, (const float*)ws, nw, grad_out);
}

void launch_mlp_adam(float* P, const float* G, float* m, float* v, int t,
                     float lr, float beta1, float beta2, float eps,
                     hipStream_t s)
{
    AQ_REQUIRE(t >= 1, "mlp: Adam steps count from 1");
    const float c1 = 1.f - powf(beta1, (float)t);
    const float c2 = 1.f - powf(beta2, (float)t);
    hipLaunchKernelGGL(mlp_adam_k, dim3((NPARAM + 255) / 256), dim3(256), 0,
                       s, P, G, m, v, lr, beta1, beta2, eps, c1, c2);
}

}  // namespace aq

using namespace aq;

extern "C" {

AQ_API int aq_mlp_nparam(void) { return mlp_nparam(); }

AQ_API size_t aq_mlp_workspace(uint32_t m) { return mlp_workspace(m); }

AQ_API int aq_mlp_forward(const float* x, uint32_t m, const float* params,
                          float* y, void* stream)
{
    return guard([&] {
        AQ_REQUIRE(x && params && y, "aq_mlp_forward: null pointer");
        launch_mlp_forward(x, m, params, y, S(stream));
    });
}

AQ_API int aq_mlp_grad(const float* x, const float* target, uint32_t m,
                       const float* params, float* grad, void* workspace,
                       void* stream)
{
    return guard([&] {
        AQ_REQUIRE(x && target && params && grad && workspace,
                   "aq_mlp_grad: null pointer");
        launch_mlp_grad(x, target, m, params, grad, workspace, S(stream));
    });
}

AQ_API int aq_mlp_adam(float* params, const float* grad, float* adam_m,
                       float* adam_v, int step, float lr, void* stream)
{
    return guard([&] {
        AQ_REQUIRE(params && grad && adam_m && adam_v,
                   "aq_mlp_adam: null pointer");
        launch_mlp_adam(params, grad, adam_m, adam_v, step, lr, 0.9f, 0.999f,
                        1e-8f, S(stream));
    });
}

}  // extern "C"
