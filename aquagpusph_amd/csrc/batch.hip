// Batched ensembles on gfx950: many independent cases in one particle
// array, one link-list, one walk and one launch per tool.  Each case is a
// particle set (iset); the cases sit apart in space (further than the
// support), so the shared cell grid keeps them independent, and all pair
// kernels serve every case in the same launch.  What differs between cases
// lives in per-set arrays: the sound speed cs_set, the gravity g_set, a
// horizontal harmonic excitation of the case's frame (exc_amp, exc_omega:
// the frame moves as exc_amp sin(exc_omega t), which adds
// exc_amp exc_omega^2 sin(exc_omega t) to the gravity along x), the time
// t_set and the time step dt_set.  These kernels are the
// per-set forms of the scalar-parameter ones (basic/EOS.cl,
// cfd/Rates.cl, cfd/TimeStep.cl, basic/time_scheme/improved_euler.cl,
// cfd/deltaSPH.cl, cfd/Boundary/BI/{GradP,InterpolationShepard}.cl); the
// scalar expressions the reference evaluates on the host in double (the
// fixed time step, t + dt, the excitation) are evaluated here in double
// and narrowed to float the same way.
#include "aq_device.hpp"
#include "aq_internal.hpp"

namespace aq {

namespace {

constexpr int BS = 256;
constexpr uint32_t MAX_SETS = 4096;

// x gravity of set s at its time t: g_set.x + a w^2 sin(w t) (a = exc_amp,
// w = exc_omega), in double, narrowed
template <int D>
__device__ __forceinline__ float gx_eff(const float* g_set, const float* exc_a,
                                        const float* exc_w, const float* t_set,
                                        uint32_t s)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const double a = exc_a[s], w = exc_w[s], t = t_set[s];
    return (float)((double)g_set[(size_t)s * VS] + a * w * w * sin(w * t));
}

__global__ void __launch_bounds__(BS)
    bt_eos_k(const uint32_t* __restrict__ iset, const int* __restrict__ imove,
             const float* __restrict__ rho, float* __restrict__ p,
             const float* __restrict__ refd,
             const float* __restrict__ cs_set, uint32_t N,
             const float* __restrict__ p0p)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int im = imove[i];
    if ((im <= 0) && (im != -1)) return;
    const uint32_t s = iset[i];
    const float cs = cs_set[s];
    p[i] = (*p0p) + cs * cs * (rho[i] - refd[s]);
}

template <int D>
__global__ void __launch_bounds__(BS)
    bt_rates_k(const uint32_t* __restrict__ iset,
               const int* __restrict__ imove,
               const float* __restrict__ grad_p,
               const float* __restrict__ lap_u,
               const float* __restrict__ div_u, float* __restrict__ dudt,
               float* __restrict__ drhodt, const float* __restrict__ visc_dyn,
               const float* __restrict__ g_set,
               const float* __restrict__ exc_a,
               const float* __restrict__ exc_w,
               const float* __restrict__ t_set, uint32_t N)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != 1) return;
    const uint32_t s = iset[i];
    const float mu = visc_dyn[s];
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        const float g = k == 0 ? gx_eff<D>(g_set, exc_a, exc_w, t_set, s)
                               : g_set[(size_t)s * VS + k];
        dudt[a] = -grad_p[a] + mu * lap_u[a] + g;
    }
    drhodt[i] = -div_u[i];
}

// cfd/TimeStep.cl with the fixed step courant h / cs_set[s] of the case
template <int D>
__global__ void __launch_bounds__(BS)
    bt_timestep_k(const uint32_t* __restrict__ iset,
                  const int* __restrict__ imove, const float* __restrict__ u,
                  float* __restrict__ dt_var, uint32_t N,
                  const float* __restrict__ cs_set,
                  const float* __restrict__ dt_minp,
                  const float* __restrict__ courantp,
                  const float* __restrict__ dt_Map,
                  const float* __restrict__ hp)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float dt =
        (float)((double)(*courantp) * (double)(*hp) / (double)cs_set[iset[i]]);
    if (imove[i] <= 0) {
        dt_var[i] = dt;
        return;
    }
    const float dr_max = (*dt_Map) * (*hp);
    float s = u[(size_t)i * VS] * u[(size_t)i * VS];
#pragma unroll
    for (int k = 1; k < VS; k++)
        s = s + u[(size_t)i * VS + k] * u[(size_t)i * VS + k];
    const float dt_u = (*courantp) * dr_max / sqrtf(s);
    dt_var[i] = cl_max(cl_min(dt, dt_u), *dt_minp);
}

// dt_set[s] = min of dt_var over set s (dt_set cleared to +inf before):
// a block-local minimum per set in LDS, then one global atomic per set
// and block.  Positive floats order like their bit patterns.
__global__ void __launch_bounds__(BS)
    bt_dt_min_k(const uint32_t* __restrict__ iset,
                const float* __restrict__ dt_var, uint32_t N,
                uint32_t n_sets, uint32_t* __restrict__ dt_set_bits)
{
    __shared__ uint32_t lmin[MAX_SETS];
    for (uint32_t s = threadIdx.x; s < n_sets; s += BS) lmin[s] = 0x7F800000u;
    __syncthreads();
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < N;
         i += gridDim.x * BS) {
        const float v = dt_var[i];
        const uint32_t b = (v > 0.f) ? __float_as_uint(v) : 0u;
        atomicMin(&lmin[iset[i]], b);
    }
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < n_sets; s += BS)
        if (lmin[s] != 0x7F800000u) atomicMin(&dt_set_bits[s], lmin[s]);
}

template <int D>
__global__ void __launch_bounds__(BS)
    bt_predictor_k(const uint32_t* __restrict__ iset,
                   const int* __restrict__ imove, const float* __restrict__ r,
                   const float* __restrict__ u, const float* __restrict__ dudt,
                   const float* __restrict__ rho,
                   const float* __restrict__ drhodt, float* __restrict__ r_in,
                   float* __restrict__ u_in, float* __restrict__ dudt_in,
                   float* __restrict__ rho_in, float* __restrict__ drhodt_in,
                   uint32_t N, const float* __restrict__ dt_set)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float DT = dt_set[iset[i]];
    if (imove[i] <= 0) DT = 0.f;
    const float hdt2 = 0.5f * DT * DT;
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        const float a_ = dudt[a];
        const float u_ = u[a];
        dudt_in[a] = a_;
        u_in[a] = u_ + DT * a_;
        r_in[a] = r[a] + DT * u_ + hdt2 * a_;
    }
    drhodt_in[i] = drhodt[i];
    rho_in[i] = rho[i] + DT * drhodt[i];
}

template <int D>
__global__ void __launch_bounds__(BS)
    bt_corrector_k(const uint32_t* __restrict__ iset,
                   const int* __restrict__ imove, float* __restrict__ r,
                   float* __restrict__ u, const float* __restrict__ dudt,
                   float* __restrict__ rho, const float* __restrict__ drhodt,
                   const float* __restrict__ dudt_in,
                   const float* __restrict__ drhodt_in, uint32_t N,
                   const float* __restrict__ dt_set)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] <= 0) return;
    const float DT = 0.5f * dt_set[iset[i]];
    const float DT2 = DT * DT;
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        const float da = dudt[a] - dudt_in[a];
        u[a] += DT * da;
        r[a] += DT2 * da;
    }
    rho[i] += DT * (drhodt[i] - drhodt_in[i]);
}

// cfd/deltaSPH.cl:29 + basic/deltaSPH.cl:330-350 with the case's dt
__global__ void __launch_bounds__(BS)
    bt_delta_k(const uint32_t* __restrict__ iset,
               const int* __restrict__ imove, const float* __restrict__ rho,
               const float* __restrict__ lap_p, float* __restrict__ drhodt,
               const float* __restrict__ refd,
               const float* __restrict__ delta, uint32_t N,
               const float* __restrict__ dt_set)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != 1) return;
    const uint32_t s = iset[i];
    const float delta_f = delta[s] * dt_set[s] * rho[i] / refd[s];
    drhodt[i] += delta_f * lap_p[i];
}

template <int D>
__global__ void __launch_bounds__(BS)
    bt_bi_gradp_k(const uint32_t* __restrict__ iset,
                  const int* __restrict__ imove,
                  const float* __restrict__ shepard,
                  const float* __restrict__ lap_u,
                  const float* __restrict__ dudt, float* __restrict__ grad_p,
                  const float* __restrict__ visc_dyn,
                  const float* __restrict__ refd,
                  const float* __restrict__ g_set,
                  const float* __restrict__ exc_a,
                  const float* __restrict__ exc_w,
                  const float* __restrict__ t_set, uint32_t N)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != -3) return;
    const uint32_t s = iset[i];
    float sh = shepard[i];
    if (sh < 1.0E-6f) sh = 1.f;
    const float f = visc_dyn[s] / refd[s];
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        const float g = k == 0 ? gx_eff<D>(g_set, exc_a, exc_w, t_set, s)
                               : g_set[(size_t)s * VS + k];
        grad_p[a] = f * lap_u[a] / sh + g - dudt[a];
    }
}

__global__ void __launch_bounds__(BS)
    bt_bi_interp_shepard_k(const uint32_t* __restrict__ iset,
                           const int* __restrict__ imove,
                           const float* __restrict__ shepard,
                           float* __restrict__ rho, float* __restrict__ p,
                           const float* __restrict__ refd,
                           const float* __restrict__ cs_set, uint32_t N,
                           const float* __restrict__ p0p)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != -3) return;
    float sh = shepard[i];
    if (sh < 1.0E-6f) sh = 1.f;
    const float pi = p[i] / sh;
    p[i] = pi;
    const uint32_t s = iset[i];
    const float cs = cs_set[s];
    rho[i] = refd[s] + (pi - (*p0p)) / (cs * cs);
}

// t_set[s] = t_set[s] + dt_set[s] (in double, narrowed, as "t = t + dt")
__global__ void bt_time_k(float* __restrict__ t_set,
                          const float* __restrict__ dt_set, uint32_t n_sets)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_sets) return;
    t_set[s] = (float)((double)t_set[s] + (double)dt_set[s]);
}

}  // namespace

#define AQ_BG(N) dim3(blocks_for((N), BS)), dim3(BS), 0, s
#define AQ_BD(dims, KER, ...)                                                  \
    do {                                                                       \
        if ((dims) == 2)                                                       \
            hipLaunchKernelGGL(KER<2>, __VA_ARGS__);                           \
        else                                                                   \
            hipLaunchKernelGGL(KER<3>, __VA_ARGS__);                           \
    } while (0)

void launch_bt_eos(const uint32_t* iset, const int32_t* imove,
                   const float* rho, float* p, const float* refd,
                   const float* cs_set, uint32_t N, const float* p0,
                   hipStream_t s)
{
    if (!N) return;
    hipLaunchKernelGGL(bt_eos_k, AQ_BG(N), iset, imove, rho, p, refd, cs_set,
                       N, p0);
}
void launch_bt_rates(int dims, const uint32_t* iset, const int32_t* imove,
                     const float* grad_p, const float* lap_u,
                     const float* div_u, float* dudt, float* drhodt,
                     const float* visc, const float* g_set,
                     const float* exc_a, const float* exc_w,
                     const float* t_set, uint32_t N, hipStream_t s)
{
    if (!N) return;
    AQ_BD(dims, bt_rates_k, AQ_BG(N), iset, imove, grad_p, lap_u, div_u, dudt,
          drhodt, visc, g_set, exc_a, exc_w, t_set, N);
}
void launch_bt_timestep(int dims, const uint32_t* iset, const int32_t* imove,
                        const float* u, float* dt_var, uint32_t N,
                        const float* cs_set, const float* dt_min,
                        const float* courant, const float* dt_Ma,
                        const float* h, hipStream_t s)
{
    if (!N) return;
    AQ_BD(dims, bt_timestep_k, AQ_BG(N), iset, imove, u, dt_var, N, cs_set,
          dt_min, courant, dt_Ma, h);
}
void launch_bt_dt_min(const uint32_t* iset, const float* dt_var, uint32_t N,
                      uint32_t n_sets, float* dt_set, hipStream_t s)
{
    AQ_REQUIRE(n_sets <= MAX_SETS, "batched time step: more than 4096 sets");
    if (!N || !n_sets) return;
    const unsigned nb =
        std::max(1u, std::min<unsigned>(1024, blocks_for(N, BS * 8)));
    hipLaunchKernelGGL(bt_dt_min_k, dim3(nb), dim3(BS), 0, s, iset, dt_var, N,
                       n_sets, (uint32_t*)dt_set);
}
void launch_bt_predictor(int dims, const uint32_t* iset, const int32_t* imove,
                         const float* r, const float* u, const float* dudt,
                         const float* rho, const float* drhodt, float* r_in,
                         float* u_in, float* dudt_in, float* rho_in,
                         float* drhodt_in, uint32_t N, const float* dt_set,
                         hipStream_t s)
{
    if (!N) return;
    AQ_BD(dims, bt_predictor_k, AQ_BG(N), iset, imove, r, u, dudt, rho,
          drhodt, r_in, u_in, dudt_in, rho_in, drhodt_in, N, dt_set);
}
void launch_bt_corrector(int dims, const uint32_t* iset, const int32_t* imove,
                         float* r, float* u, const float* dudt, float* rho,
                         const float* drhodt, const float* dudt_in,
                         const float* drhodt_in, uint32_t N,
                         const float* dt_set, hipStream_t s)
{
    if (!N) return;
    AQ_BD(dims, bt_corrector_k, AQ_BG(N), iset, imove, r, u, dudt, rho,
          drhodt, dudt_in, drhodt_in, N, dt_set);
}
void launch_bt_delta(const uint32_t* iset, const int32_t* imove,
                     const float* rho, const float* lap_p, float* drhodt,
                     const float* refd, const float* delta, uint32_t N,
                     const float* dt_set, hipStream_t s)
{
    if (!N) return;
    hipLaunchKernelGGL(bt_delta_k, AQ_BG(N), iset, imove, rho, lap_p, drhodt,
                       refd, delta, N, dt_set);
}
void launch_bt_bi_gradp(int dims, const uint32_t* iset, const int32_t* imove,
                        const float* shepard, const float* lap_u,
                        const float* dudt, float* grad_p, const float* visc,
                        const float* refd, const float* g_set,
                        const float* exc_a, const float* exc_w,
                        const float* t_set, uint32_t N, hipStream_t s)
{
    if (!N) return;
    AQ_BD(dims, bt_bi_gradp_k, AQ_BG(N), iset, imove, shepard, lap_u, dudt,
          grad_p, visc, refd, g_set, exc_a, exc_w, t_set, N);
}
void launch_bt_bi_interp_shepard(const uint32_t* iset, const int32_t* imove,
                                 const float* shepard, float* rho, float* p,
                                 const float* refd, const float* cs_set,
                                 uint32_t N, const float* p0, hipStream_t s)
{
    if (!N) return;
    hipLaunchKernelGGL(bt_bi_interp_shepard_k, AQ_BG(N), iset, imove, shepard,
                       rho, p, refd, cs_set, N, p0);
}
void launch_bt_time(float* t_set, const float* dt_set, uint32_t n_sets,
                    hipStream_t s)
{
    if (!n_sets) return;
    hipLaunchKernelGGL(bt_time_k, dim3(blocks_for(n_sets, 64)), dim3(64), 0,
                       s, t_set, dt_set, n_sets);
}

#undef AQ_BD
#undef AQ_BG

}  // namespace aq
