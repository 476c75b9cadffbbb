// Mixed-precision time integration on gfx950: the improved-Euler position
// update carried in double, everything else in float.
//
// A particle's position is r + dr64: r is the float position every other
// tool reads and writes (walks, sort keys, savers, portals), dr64 the
// double residual that float rounding dropped.  The predictor and the
// corrector integrate r + dr64 in double and split the result again into
// its float rounding and the exact double remainder, so tools that move r
// by a whole offset (teleports) keep the residual meaningful, and a case
// that starts (or resumes) with dr64 = 0 starts from the float positions.
// The velocity, density and their rates stay float, as in the reference
// (basic/time_scheme/improved_euler.cl:75-103, 125-147).
#include "aq_device.hpp"
#include "aq_internal.hpp"

namespace aq {

namespace {

constexpr int BS = 256;

// improved_euler.cl:75-103 with the position in double
template <int D>
__global__ void __launch_bounds__(BS)
    predictor64_k(const int* __restrict__ imove, const float* __restrict__ r,
                  const double* __restrict__ dr64,
                  const float* __restrict__ u, const float* __restrict__ dudt,
                  const float* __restrict__ rho,
                  const float* __restrict__ drhodt, float* __restrict__ r_in,
                  double* __restrict__ dr64_in, float* __restrict__ u_in,
                  float* __restrict__ dudt_in, float* __restrict__ rho_in,
                  float* __restrict__ drhodt_in, uint32_t N,
                  const float* __restrict__ dtp, const uint32_t* status)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float DT = *dtp;
    if ((imove[i] <= 0) || (status && *status)) DT = 0.f;
    const float hdt2 = 0.5f * DT * DT;
    const double dDT = DT, dh = hdt2;
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        const float a_ = dudt[a];
        const float u_ = u[a];
        dudt_in[a] = a_;
        u_in[a] = u_ + DT * a_;
        const double x = ((double)r[a] + dr64[a]) + dDT * (double)u_ +
                         dh * (double)a_;
        const float xf = (float)x;
        r_in[a] = xf;
        dr64_in[a] = x - (double)xf;
    }
    drhodt_in[i] = drhodt[i];
    rho_in[i] = rho[i] + DT * drhodt[i];
}

// improved_euler.cl:125-147 with the position in double
template <int D>
__global__ void __launch_bounds__(BS)
    corrector64_k(const int* __restrict__ imove, float* __restrict__ r,
                  double* __restrict__ dr64, float* __restrict__ u,
                  const float* __restrict__ dudt, float* __restrict__ rho,
                  const float* __restrict__ drhodt,
                  const float* __restrict__ dudt_in,
                  const float* __restrict__ drhodt_in, uint32_t N,
                  const float* __restrict__ dtp, const uint32_t* status)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (status && *status) return;  // a failure was flagged: state kept
    if (imove[i] <= 0) return;
    const float DT = 0.5f * (*dtp);
    const double DT2 = (double)DT * (double)DT;
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        const float da = dudt[a] - dudt_in[a];
        u[a] += DT * da;
        const double x = ((double)r[a] + dr64[a]) + DT2 * (double)da;
        const float xf = (float)x;
        r[a] = xf;
        dr64[a] = x - (double)xf;
    }
    rho[i] += DT * (drhodt[i] - drhodt_in[i]);
}

// the fluid particles' masses (0 elsewhere), for a double sum
__global__ void __launch_bounds__(BS)
    fluid_mass_k(const int* __restrict__ imove, const float* __restrict__ m,
                 float* __restrict__ m_fluid, uint32_t N)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    m_fluid[i] = (imove[i] == 1) ? m[i] : 0.f;
}

}  // namespace

void launch_predictor64(int dims, const int32_t* imove, const float* r,
                        const double* dr64, const float* u, const float* dudt,
                        const float* rho, const float* drhodt, float* r_in,
                        double* dr64_in, float* u_in, float* dudt_in,
                        float* rho_in, float* drhodt_in, uint32_t N,
                        const float* dt, const uint32_t* status, hipStream_t s)
{
    if (!N) return;
    const dim3 g(blocks_for(N, BS)), b(BS);
    if (dims == 2)
        hipLaunchKernelGGL(predictor64_k<2>, g, b, 0, s, imove, r, dr64, u,
                           dudt, rho, drhodt, r_in, dr64_in, u_in, dudt_in,
                           rho_in, drhodt_in, N, dt, status);
    else
        hipLaunchKernelGGL(predictor64_k<3>, g, b, 0, s, imove, r, dr64, u,
                           dudt, rho, drhodt, r_in, dr64_in, u_in, dudt_in,
                           rho_in, drhodt_in, N, dt, status);
}

void launch_corrector64(int dims, const int32_t* imove, float* r,
                        double* dr64, float* u, const float* dudt,
                        float* rho, const float* drhodt, const float* dudt_in,
                        const float* drhodt_in, uint32_t N, const float* dt,
                        const uint32_t* status, hipStream_t s)
{
    if (!N) return;
    const dim3 g(blocks_for(N, BS)), b(BS);
    if (dims == 2)
        hipLaunchKernelGGL(corrector64_k<2>, g, b, 0, s, imove, r, dr64, u,
                           dudt, rho, drhodt, dudt_in, drhodt_in, N, dt,
                           status);
    else
        hipLaunchKernelGGL(corrector64_k<3>, g, b, 0, s, imove, r, dr64, u,
                           dudt, rho, drhodt, dudt_in, drhodt_in, N, dt,
                           status);
}

void launch_fluid_mass(const int32_t* imove, const float* m, float* m_fluid,
                       uint32_t N, hipStream_t s)
{
    if (!N) return;
    hipLaunchKernelGGL(fluid_mass_k, dim3(blocks_for(N, BS)), dim3(BS), 0, s,
                       imove, m, m_fluid, N);
}

}  // namespace aq
