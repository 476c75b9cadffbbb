// SPH large-eddy closure on gfx950 (no counterpart in the reference): an
// eddy viscosity nu_t per fluid particle, either Smagorinsky's
// (C_s H)^2 |S| or a learned coefficient c of H^2 |S| given by the network
// of mlp.hip from invariants of the local velocity gradient, added to the
// momentum rate with the variable-viscosity form of the Morris Laplacian.
//   grad_u   (matrix*): G_i = sum_j m_j / rho_j (u_j - u_i) (x) F_ij r_ij,
//            the SPH gradient (grad_u[a][b] = du_a / dx_b; exact for a
//            linear field in a full neighbourhood)
//   les_in   (8 floats per particle): the network inputs -- with S, W the
//            symmetric and antisymmetric parts of G and n = S:S + W:W,
//            S:S/n, W:W/n, tr(S^3)/n^1.5, tr(W^2 S)/n^1.5, tr(W^2 S^2)/n^2,
//            tr(W^2 S W S^2)/n^3, and h/dr
//   nu_t     max(c, c_min) H^2 |S| or (C_s H)^2 |S|, |S| = sqrt(2 S:S)
//            (c_min = les_cmin, 0 by default; a negative floor lets the
//            network return energy to the resolved scales, backscatter)
//   dudt_i  += sum_j F_ij m_j (rho_i nu_i + rho_j nu_j) / (rho_i rho_j)
//              (u_j - u_i)
// Pairs are fluid-fluid within the support, walked in the reference order.
#include "aq_device.hpp"
#include "aq_internal.hpp"

namespace aq {

namespace {

constexpr int BS = 256;

template <int D> struct MS_ {
    static constexpr int n = (D == 2) ? 2 : 4;   // row stride
    static constexpr int s = (D == 2) ? 4 : 16;  // stored floats
};

__device__ __forceinline__ float comp(const f3& v, int k)
{
    return k == 0 ? v.x : (k == 1 ? v.y : v.z);
}

template <int D>
__global__ void __launch_bounds__(BS)
    les_gradu_k(const int* __restrict__ imove, const float* __restrict__ r,
                const float* __restrict__ u, const float* __restrict__ rho,
                const float* __restrict__ m, float* __restrict__ grad_u,
                uint32_t N, const uint32_t* __restrict__ icell,
                const uint32_t* __restrict__ ihoc,
                const uint32_t* __restrict__ nc, float H, float CONF,
                float smax)
{
    constexpr int n = MS_<D>::n, S = MS_<D>::s;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float G[S];
#pragma unroll
    for (int k = 0; k < S; k++) G[k] = 0.f;
    if (imove[i] == 1) {
        const f3 r_i = ldv<D>(r, i), u_i = ldv<D>(u, i);
        walk<D>(icell[i], N, nc[0], nc[1], icell, ihoc, [&](uint32_t j) {
            if (j == i || imove[j] != 1) return;
            const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
            const float s = dot<D>(r_ij, r_ij);
            if (s > smax) return;
            const float q = sqrtf(s) / H;
            const float w = kernelF<D>(q) * CONF * m[j] / rho[j];
            const f3 du = sub<D>(ldv<D>(u, j), u_i);
            for (int a = 0; a < D; a++)
                for (int b = 0; b < D; b++)
                    G[a * n + b] += comp(du, a) * (w * comp(r_ij, b));
        });
    }
#pragma unroll
    for (int k = 0; k < S; k++) grad_u[(size_t)i * S + k] = G[k];
}

// y = A B for D x D blocks of row stride 3
template <int D>
__device__ __forceinline__ void mm(const float* A, const float* B, float* Y)
{
    for (int a = 0; a < D; a++)
        for (int b = 0; b < D; b++) {
            float s = 0.f;
            for (int c = 0; c < D; c++) s += A[a * 3 + c] * B[c * 3 + b];
            Y[a * 3 + b] = s;
        }
}

template <int D> __device__ __forceinline__ float tr(const float* A)
{
    float s = 0.f;
    for (int a = 0; a < D; a++) s += A[a * 3 + a];
    return s;
}

template <int D>
__global__ void __launch_bounds__(BS)
    les_features_k(const int* __restrict__ imove,
                   const float* __restrict__ grad_u, float* __restrict__ x,
                   uint32_t N, float hdr)
{
    constexpr int n = MS_<D>::n, SS = MS_<D>::s;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float S[9], W[9];
    float s2 = 0.f, w2 = 0.f;
    for (int a = 0; a < D; a++)
        for (int b = 0; b < D; b++) {
            const float gab = grad_u[(size_t)i * SS + a * n + b];
            const float gba = grad_u[(size_t)i * SS + b * n + a];
            S[a * 3 + b] = 0.5f * (gab + gba);
            W[a * 3 + b] = 0.5f * (gab - gba);
        }
    for (int a = 0; a < D; a++)
        for (int b = 0; b < D; b++) {
            s2 += S[a * 3 + b] * S[a * 3 + b];
            w2 += W[a * 3 + b] * W[a * 3 + b];
        }
    const float nn = s2 + w2 + 1e-30f;
    const float n15 = nn * sqrtf(nn);
    float S2[9], W2[9], T[9], U[9];
    mm<D>(S, S, S2);
    mm<D>(W, W, W2);
    mm<D>(S2, S, T);
    const float trS3 = tr<D>(T);
    mm<D>(W2, S, T);
    const float trW2S = tr<D>(T);
    mm<D>(W2, S2, U);
    const float trW2S2 = tr<D>(U);
    mm<D>(T, W, U);   // W^2 S W
    mm<D>(U, S2, T);  // W^2 S W S^2
    const float trW2SWS2 = tr<D>(T);
    float* o = x + (size_t)i * 8;
    o[0] = s2 / nn;
    o[1] = w2 / nn;
    o[2] = trS3 / n15;
    o[3] = trW2S / n15;
    o[4] = trW2S2 / (nn * nn);
    o[5] = trW2SWS2 / (nn * nn * nn);
    o[6] = hdr;
    o[7] = 0.f;
    (void)imove;
}

// mode 0: nu_t = max(c, cmin) H^2 |S|; mode 1: (C_s H)^2 |S|
template <int D>
__global__ void __launch_bounds__(BS)
    les_nut_k(const int* __restrict__ imove, const float* __restrict__ grad_u,
              const float* __restrict__ c, float* __restrict__ nu_t,
              uint32_t N, float H, int mode, float Cs, float cmin)
{
    constexpr int n = MS_<D>::n, SS = MS_<D>::s;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != 1) {
        nu_t[i] = 0.f;
        return;
    }
    float s2 = 0.f;
    for (int a = 0; a < D; a++)
        for (int b = 0; b < D; b++) {
            const float sab = 0.5f * (grad_u[(size_t)i * SS + a * n + b] +
                                      grad_u[(size_t)i * SS + b * n + a]);
            s2 += sab * sab;
        }
    const float smag = sqrtf(2.f * s2);
    const float k = mode == 0 ? fmaxf(c[i], cmin) : Cs * Cs;
    nu_t[i] = k * H * H * smag;
}

template <int D>
__global__ void __launch_bounds__(BS)
    les_interactions_k(const int* __restrict__ imove,
                       const float* __restrict__ r,
                       const float* __restrict__ u,
                       const float* __restrict__ rho,
                       const float* __restrict__ m,
                       const float* __restrict__ nu_t,
                       float* __restrict__ dudt, uint32_t N,
                       const uint32_t* __restrict__ icell,
                       const uint32_t* __restrict__ ihoc,
                       const uint32_t* __restrict__ nc, float H, float CONF,
                       float smax)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != 1) return;
    const f3 r_i = ldv<D>(r, i), u_i = ldv<D>(u, i);
    const float rho_i = rho[i], mu_i = rho_i * nu_t[i];
    f3 acc = { 0.f, 0.f, 0.f };
    walk<D>(icell[i], N, nc[0], nc[1], icell, ihoc, [&](uint32_t j) {
        if (j == i || imove[j] != 1) return;
        const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
        const float s = dot<D>(r_ij, r_ij);
        if (s > smax) return;
        const float q = sqrtf(s) / H;
        const float rho_j = rho[j];
        const float f = kernelF<D>(q) * CONF * m[j];
        const float sc = f * (mu_i + rho_j * nu_t[j]) / (rho_i * rho_j);
        axpy<D>(acc, sc, sub<D>(ldv<D>(u, j), u_i));
    });
    const f3 d = ldv<D>(dudt, i);
    stxyz<D>(dudt, i, f3{ d.x + acc.x, d.y + acc.y, d.z + acc.z });
}

}  // namespace

#define AQ_LG(N) dim3(blocks_for((N), BS)), dim3(BS), 0, s
#define AQ_LD(dims, KER, ...)                                                  \
    do {                                                                       \
        if ((dims) == 2)                                                       \
            hipLaunchKernelGGL(KER<2>, __VA_ARGS__);                           \
        else                                                                   \
            hipLaunchKernelGGL(KER<3>, __VA_ARGS__);                           \
    } while (0)

void launch_les_gradu(int dims, const int32_t* imove, const float* r,
                      const float* u, const float* rho, const float* m,
                      float* grad_u, uint32_t N, const uint32_t* icell,
                      const uint32_t* ihoc, const uint32_t* n_cells,
                      aq_defs d, float smax, hipStream_t s)
{
    if (!N) return;
    AQ_LD(dims, les_gradu_k, AQ_LG(N), imove, r, u, rho, m, grad_u, N, icell,
          ihoc, n_cells, d.H, d.CONF, smax);
}
void launch_les_features(int dims, const int32_t* imove, const float* grad_u,
                         float* x, uint32_t N, float hdr, hipStream_t s)
{
    if (!N) return;
    AQ_LD(dims, les_features_k, AQ_LG(N), imove, grad_u, x, N, hdr);
}
void launch_les_nut(int dims, const int32_t* imove, const float* grad_u,
                    const float* c, float* nu_t, uint32_t N, float H,
                    int mode, float Cs, float cmin, hipStream_t s)
{
    if (!N) return;
    AQ_LD(dims, les_nut_k, AQ_LG(N), imove, grad_u, c, nu_t, N, H, mode, Cs,
          cmin);
}
void launch_les_interactions(int dims, const int32_t* imove, const float* r,
                             const float* u, const float* rho, const float* m,
                             const float* nu_t, float* dudt, uint32_t N,
                             const uint32_t* icell, const uint32_t* ihoc,
                             const uint32_t* n_cells, aq_defs d, float smax,
                             hipStream_t s)
{
    if (!N) return;
    AQ_LD(dims, les_interactions_k, AQ_LG(N), imove, r, u, rho, m, nu_t, dudt,
          N, icell, ihoc, n_cells, d.H, d.CONF, smax);
}

#undef AQ_LD
#undef AQ_LG

}  // namespace aq
