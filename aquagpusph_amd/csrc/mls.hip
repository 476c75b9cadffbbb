// MLS kernel-gradient correction matrices and the full / simple delta-SPH
// Laplacian corrections on gfx950 (resources/Scripts/basic/MLS.cl,
// basic/deltaSPH.cl:57-190 and 261-313; cfd/deltaSPH.cl restricts them to
// the fluid).  A "matrix" is float4 (2x2, row major) in 2D and float16 (4x4,
// row major, the 3x3 block used) in 3D, with the operations of
// types/{2D,3D}.h: outer products, MATRIX_MUL, inv() by the adjugate
// (identity when |1/det| > 1e16) and the pseudo-inverse
// MATRIX_INV(M) = inv(M^T M) M^T.  One thread per particle; the pair
// kernels walk the link-list in the reference order.
#include "aq_device.hpp"
#include "aq_internal.hpp"

namespace aq {

namespace {

constexpr int BS = 256;

struct Cells {
    const uint32_t* icell;
    const uint32_t* ihoc;
    const uint32_t* n_cells;
    const uint32_t* status;
    uint32_t N;
};

__device__ __forceinline__ bool skip(const Cells& c)
{
    return c.status && (*c.status & 1u);
}

// EXCLUDED_PARTICLE: cfd (imove != 1) or basic (imove <= 0)
__device__ __forceinline__ bool excluded(int im, int basic)
{
    return basic ? (im <= 0) : (im != 1);
}

template <int D> struct MatT {
    static constexpr int S = (D == 2) ? 4 : 16;
};

template <int D>
__device__ __forceinline__ void mat_mul(const float* a, const float* b,
                                        float* c)
{
    if constexpr (D == 2) {
        c[0] = a[0] * b[0] + a[1] * b[2];
        c[1] = a[0] * b[1] + a[1] * b[3];
        c[2] = a[2] * b[0] + a[3] * b[2];
        c[3] = a[2] * b[1] + a[3] * b[3];
    } else {
#pragma unroll
        for (int k = 0; k < 16; k++) c[k] = 0.f;
#pragma unroll
        for (int r_ = 0; r_ < 3; r_++)
#pragma unroll
            for (int c_ = 0; c_ < 3; c_++) {
                float s = a[4 * r_ + 0] * b[c_];
                s = s + a[4 * r_ + 1] * b[4 + c_];
                s = s + a[4 * r_ + 2] * b[8 + c_];
                c[4 * r_ + c_] = s;
            }
    }
}

template <int D>
__device__ __forceinline__ void mat_transpose(const float* a, float* t)
{
    constexpr int n = (D == 2) ? 2 : 4;
#pragma unroll
    for (int r_ = 0; r_ < n; r_++)
#pragma unroll
        for (int c_ = 0; c_ < n; c_++) t[n * r_ + c_] = a[n * c_ + r_];
}

template <int D>
__device__ __forceinline__ void mat_inv(const float* m, float* o)
{
    if constexpr (D == 2) {
        const float d = 1.f / (m[0] * m[3] - m[1] * m[2]);
        if (fabsf(d) > 1.e16f) {
            o[0] = 1.f, o[1] = 0.f, o[2] = 0.f, o[3] = 1.f;
            return;
        }
        o[0] = m[3] * d;
        o[1] = -m[1] * d;
        o[2] = -m[2] * d;
        o[3] = m[0] * d;
    } else {
        const float det = m[0] * (m[5] * m[10] - m[6] * m[9]) +
                          m[1] * (m[6] * m[8] - m[4] * m[10]) +
                          m[2] * (m[4] * m[9] - m[5] * m[8]);
        const float d = 1.f / det;
#pragma unroll
        for (int k = 0; k < 16; k++) o[k] = (k % 5 == 0) ? 1.f : 0.f;
        if (fabsf(d) > 1.e16f) return;
        o[0] = (m[5] * m[10] - m[6] * m[9]) * d;
        o[1] = (m[2] * m[9] - m[1] * m[10]) * d;
        o[2] = (m[1] * m[6] - m[2] * m[5]) * d;
        o[4] = (m[6] * m[8] - m[4] * m[10]) * d;
        o[5] = (m[0] * m[10] - m[2] * m[8]) * d;
        o[6] = (m[2] * m[4] - m[0] * m[6]) * d;
        o[8] = (m[4] * m[9] - m[5] * m[8]) * d;
        o[9] = (m[1] * m[8] - m[0] * m[9]) * d;
        o[10] = (m[0] * m[5] - m[1] * m[4]) * d;
    }
}

// basic/MLS.cl:entry
template <int D>
__global__ void __launch_bounds__(BS)
    mls_entry_k(const int* __restrict__ imove, const float* __restrict__ r,
                const float* __restrict__ rho, const float* __restrict__ m,
                float* __restrict__ mls, const uint32_t* mls_imove, Cells c,
                float H, float CONF, float smax)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    if (skip(c)) return;
    const int mi = (int)*mls_imove;
    if (imove[i] != mi) return;
    constexpr int S = MatT<D>::S;
    constexpr int n = (D == 2) ? 2 : 4;
    const f3 r_i = ldv<D>(r, i);
    float acc[S];
#pragma unroll
    for (int k = 0; k < S; k++) acc[k] = 0.f;
    walk<D>(c.icell[i], c.N, c.n_cells[0], c.n_cells[1], c.icell, c.ihoc,
            [&](uint32_t j) {
                if (i == j) return;
                if (imove[j] != mi) return;
                const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
                const float s = dot<D>(r_ij, r_ij);
                if (s > smax) return;
                const float q = sqrtf(s) / H;
                const float f = kernelF<D>(q) * CONF * m[j] / rho[j];
                const float a[3] = { r_ij.x, r_ij.y, r_ij.z };
                const float v2[3] = { f * r_ij.x, f * r_ij.y, f * r_ij.z };
#pragma unroll
                for (int p_ = 0; p_ < D; p_++)
#pragma unroll
                    for (int q_ = 0; q_ < D; q_++)
                        acc[n * p_ + q_] += a[p_] * v2[q_];
            });
#pragma unroll
    for (int k = 0; k < S; k++) mls[(size_t)i * S + k] = acc[k];
}

// basic/MLS.cl:mls_inv
template <int D>
__global__ void mls_inv_k(const int* __restrict__ imove,
                          float* __restrict__ mls, uint32_t N,
                          const uint32_t* mls_imove)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != (int)*mls_imove) return;
    constexpr int S = MatT<D>::S;
    float m[S], t[S], mtm[S], iv[S], o[S];
#pragma unroll
    for (int k = 0; k < S; k++) m[k] = mls[(size_t)i * S + k];
    mat_transpose<D>(m, t);
    mat_mul<D>(t, m, mtm);
    mat_inv<D>(mtm, iv);
    mat_mul<D>(iv, t, o);
#pragma unroll
    for (int k = 0; k < S; k++) mls[(size_t)i * S + k] = o[k];
}

// deltaSPH.cl:simple (the whole vec)
template <int D>
__global__ void lapp_simple_k(const uint32_t* __restrict__ iset,
                              const int* __restrict__ imove,
                              float* __restrict__ lap_p_corr,
                              const float* __restrict__ refd, uint32_t N,
                              const float* g, int basic)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (excluded(imove[i], basic)) return;
    constexpr int VS = (D == 2) ? 2 : 4;
    const float rd = refd[iset[i]];
#pragma unroll
    for (int k = 0; k < VS; k++) lap_p_corr[(size_t)i * VS + k] = rd * g[k];
}

// deltaSPH.cl:full
template <int D>
__global__ void __launch_bounds__(BS)
    lapp_full_k(const int* __restrict__ imove, const float* __restrict__ r,
                const float* __restrict__ rho, const float* __restrict__ m,
                const float* __restrict__ p, float* __restrict__ lap_p_corr,
                Cells c, float H, float CONF, float smax, int basic)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    if (skip(c)) return;
    if (excluded(imove[i], basic)) return;
    const f3 r_i = ldv<D>(r, i);
    const float p_i = p[i];
    f3 g = ldv<D>(lap_p_corr, i);
    walk<D>(c.icell[i], c.N, c.n_cells[0], c.n_cells[1], c.icell, c.ihoc,
            [&](uint32_t j) {
                if ((i == j) || excluded(imove[j], basic)) return;
                const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
                const float s = dot<D>(r_ij, r_ij);
                if (s > smax) return;
                const float q = sqrtf(s) / H;
                const float f = kernelF<D>(q) * CONF * m[j] / rho[j];
                axpy<D>(g, (p[j] - p_i) * f, r_ij);
            });
    stxyz<D>(lap_p_corr, i, g);
}

// deltaSPH.cl:full_mls (MATRIX_DOT; w = 0 in 3D)
template <int D>
__global__ void lapp_full_mls_k(const int* __restrict__ imove,
                                const float* __restrict__ mls,
                                float* __restrict__ lap_p_corr, uint32_t N,
                                int basic)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (excluded(imove[i], basic)) return;
    constexpr int S = MatT<D>::S;
    constexpr int n = (D == 2) ? 2 : 4;
    const float* M = mls + (size_t)i * S;
    const f3 v = ldv<D>(lap_p_corr, i);
    float o[3];
#pragma unroll
    for (int a = 0; a < D; a++) {
        float s = M[n * a + 0] * v.x;
        s = s + M[n * a + 1] * v.y;
        if (D == 3) s = s + M[n * a + 2] * v.z;
        o[a] = s;
    }
    if constexpr (D == 2) {
        reinterpret_cast<float2*>(lap_p_corr)[i] = make_float2(o[0], o[1]);
    } else {
        reinterpret_cast<float4*>(lap_p_corr)[i] =
            make_float4(o[0], o[1], o[2], 0.f);
    }
}

// deltaSPH.cl:lapp_corr
template <int D>
__global__ void __launch_bounds__(BS)
    lapp_corr_k(const int* __restrict__ imove, const float* __restrict__ r,
                const float* __restrict__ rho, const float* __restrict__ m,
                const float* __restrict__ lap_p_corr,
                float* __restrict__ lap_p, Cells c, float H, float CONF,
                float smax, int basic)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    if (skip(c)) return;
    if (excluded(imove[i], basic)) return;
    const f3 r_i = ldv<D>(r, i);
    const f3 ci = ldv<D>(lap_p_corr, i);
    float lp = lap_p[i];
    walk<D>(c.icell[i], c.N, c.n_cells[0], c.n_cells[1], c.icell, c.ihoc,
            [&](uint32_t j) {
                if ((i == j) || excluded(imove[j], basic)) return;
                const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
                const float s = dot<D>(r_ij, r_ij);
                if (s > smax) return;
                const float q = sqrtf(s) / H;
                const f3 cj = ldv<D>(lap_p_corr, j);
                f3 gij;
                gij.x = cj.x + ci.x;
                gij.y = cj.y + ci.y;
                gij.z = (D == 3) ? cj.z + ci.z : 0.f;
                const float f = kernelF<D>(q) * CONF * m[j] / rho[j];
                lp -= 0.5f * dot<D>(gij, r_ij) * f;
            });
    lap_p[i] = lp;
}

}  // namespace

#define AQ_G(N) dim3(blocks_for((N), BS)), dim3(BS), 0, s
#define AQ_D(dims, KER, ...)                                                   \
    do {                                                                       \
        if ((dims) == 2)                                                       \
            hipLaunchKernelGGL(KER<2>, __VA_ARGS__);                           \
        else                                                                   \
            hipLaunchKernelGGL(KER<3>, __VA_ARGS__);                           \
    } while (0)

void launch_mls_entry(int dims, const int32_t* imove, const float* r,
                      const float* rho, const float* m, float* mls,
                      const uint32_t* mls_imove, uint32_t N,
                      const uint32_t* icell, const uint32_t* ihoc,
                      const uint32_t* n_cells, const uint32_t* status,
                      aq_defs d, hipStream_t s)
{
    Cells c{ icell, ihoc, n_cells, status, N };
    AQ_D(dims, mls_entry_k, AQ_G(N), imove, r, rho, m, mls, mls_imove, c, d.H,
         d.CONF, support_threshold(d.H));
}

void launch_mls_inv(int dims, const int32_t* imove, float* mls, uint32_t N,
                    const uint32_t* mls_imove, hipStream_t s)
{
    AQ_D(dims, mls_inv_k, AQ_G(N), imove, mls, N, mls_imove);
}

void launch_lapp_simple(int dims, const uint32_t* iset, const int32_t* imove,
                        float* lap_p_corr, const float* refd, uint32_t N,
                        const float* g, int basic, hipStream_t s)
{
    AQ_D(dims, lapp_simple_k, AQ_G(N), iset, imove, lap_p_corr, refd, N, g,
         basic);
}

void launch_lapp_full(int dims, const int32_t* imove, const float* r,
                      const float* rho, const float* m, const float* p,
                      float* lap_p_corr, uint32_t N, const uint32_t* icell,
                      const uint32_t* ihoc, const uint32_t* n_cells,
                      const uint32_t* status, aq_defs d, int basic,
                      hipStream_t s)
{
    Cells c{ icell, ihoc, n_cells, status, N };
    AQ_D(dims, lapp_full_k, AQ_G(N), imove, r, rho, m, p, lap_p_corr, c, d.H,
         d.CONF, support_threshold(d.H), basic);
}

void launch_lapp_full_mls(int dims, const int32_t* imove, const float* mls,
                          float* lap_p_corr, uint32_t N, int basic,
                          hipStream_t s)
{
    AQ_D(dims, lapp_full_mls_k, AQ_G(N), imove, mls, lap_p_corr, N, basic);
}

void launch_lapp_corr(int dims, const int32_t* imove, const float* r,
                      const float* rho, const float* m,
                      const float* lap_p_corr, float* lap_p, uint32_t N,
                      const uint32_t* icell, const uint32_t* ihoc,
                      const uint32_t* n_cells, const uint32_t* status,
                      aq_defs d, int basic, hipStream_t s)
{
    Cells c{ icell, ihoc, n_cells, status, N };
    AQ_D(dims, lapp_corr_k, AQ_G(N), imove, r, rho, m, lap_p_corr, lap_p, c,
         d.H, d.CONF, support_threshold(d.H), basic);
}

#undef AQ_D
#undef AQ_G

}  // namespace aq
