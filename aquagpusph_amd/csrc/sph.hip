// SPH particle kernels of the cfd preset stack on gfx950.
//
// Streaming kernels (one thread per particle) and neighbour-walk kernels
// (one thread per particle walking the 3^D link-list cells in the reference
// order).  Each kernel cites the reference script it replaces and keeps its
// expression trees, early exits and accumulation order, so FP32 results
// match the reference kernel operation by operation.
//
// Candidate rejection: the reference tests q = length(r_ij)/H >= 2 for every
// candidate.  With correctly rounded sqrt and division q is a monotone
// function of s = dot(r_ij, r_ij), so the same test is "s > s_max" with
// s_max precomputed on the host (support_threshold); sqrt and division run
// only for accepted pairs.
#include <cfloat>
#include <cstring>
#include <cmath>
#include <mutex>
#include <unordered_map>

#include "aq_device.hpp"
#include <type_traits>

#include "aq_internal.hpp"

namespace aq {

constexpr int BS = 256;

#define AQ_STATUS_SKIP                                                         \
    if (ll.status && (*ll.status & 1u)) return

float support_threshold(float H)
{
    static std::mutex mtx;
    static std::unordered_map<uint32_t, float> cache;
    uint32_t key;
    std::memcpy(&key, &H, 4);
    std::lock_guard<std::mutex> lk(mtx);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    // q(s) = fl(fl(sqrt(s)) / H) is non-decreasing in s; binary search the
    // largest float s >= 0 with q(s) < 2 over the ordered bit patterns.
    volatile float vH = H;
    auto accept = [&](uint32_t bits) {
        float s;
        std::memcpy(&s, &bits, 4);
        volatile float root = std::sqrt(s);
        volatile float q = root / vH;
        return q < 2.f;
    };
    uint32_t lo = 0, hi = 0x7F800000u;  // accept(lo) true, accept(hi) false
    while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (accept(mid))
            lo = mid;
        else
            hi = mid;
    }
    float smax;
    std::memcpy(&smax, &lo, 4);
    cache[key] = smax;
    return smax;
}

// ===========================================================================
// Streaming kernels
// ===========================================================================

// basic/time_scheme/improved_euler.cl:75-103
template <int D>
__global__ void predictor_k(const int* __restrict__ imove,
                            const float* __restrict__ r,
                            const float* __restrict__ u,
                            const float* __restrict__ dudt,
                            const float* __restrict__ rho,
                            const float* __restrict__ drhodt,
                            float* __restrict__ r_in, float* __restrict__ u_in,
                            float* __restrict__ dudt_in,
                            float* __restrict__ rho_in,
                            float* __restrict__ drhodt_in, uint32_t N,
                            const float* dtp, const uint32_t* status)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float DT = *dtp;
    // a failure flagged earlier: predict no motion (the state is kept)
    if ((imove[i] <= 0) || (status && *status)) DT = 0.f;
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

// ---- semi-implicit midpoint scheme (basic/time_scheme/midpoint.cl);
// vectors on every stored component, as the reference's float4 arithmetic
// midpoint.cl:54-75 (predictor): the state at the start of the step
template <int D>
__global__ void mp_predictor_k(const float* __restrict__ r,
                               const float* __restrict__ u,
                               const float* __restrict__ dudt,
                               const float* __restrict__ rho,
                               const float* __restrict__ drhodt,
                               float* __restrict__ r_in,
                               float* __restrict__ u_in,
                               float* __restrict__ dudt_in,
                               float* __restrict__ rho_in,
                               float* __restrict__ drhodt_in, uint32_t N)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        dudt_in[a] = dudt[a];
        u_in[a] = u[a];
        r_in[a] = r[a];
    }
    drhodt_in[i] = drhodt[i];
    rho_in[i] = rho[i];
}

// midpoint.cl:95-112 (midpoint): half a step with the current rates
template <int D>
__global__ void mp_advance_k(const int* __restrict__ imove,
                             const float* __restrict__ u_in,
                             float* __restrict__ u,
                             const float* __restrict__ dudt,
                             const float* __restrict__ rho_in,
                             float* __restrict__ rho,
                             const float* __restrict__ drhodt, uint32_t N,
                             const float* __restrict__ dtp)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] <= 0) return;
    const float hdt = 0.5f * (*dtp);
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        u[a] = u_in[a] + hdt * dudt[a];
    }
    rho[i] = rho_in[i] + hdt * drhodt[i];
}

// midpoint.cl:129-142 (midpoint_r)
template <int D>
__global__ void mp_advance_r_k(const int* __restrict__ imove,
                               const float* __restrict__ r_in,
                               float* __restrict__ r,
                               const float* __restrict__ u, uint32_t N,
                               const float* __restrict__ dtp)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] <= 0) return;
    const float hdt = 0.5f * (*dtp);
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        r[a] = r_in[a] + hdt * u[a];
    }
}

// midpoint.cl:164-179 (relax)
template <int D>
__global__ void mp_relax_k(const int* __restrict__ imove,
                           const float* __restrict__ dudt_in,
                           float* __restrict__ dudt,
                           const float* __restrict__ drhodt_in,
                           float* __restrict__ drhodt, uint32_t N,
                           const float* __restrict__ fp)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] <= 0) return;
    const float f = *fp;
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        dudt[a] = f * dudt_in[a] + (1.f - f) * dudt[a];
    }
    drhodt[i] = f * drhodt_in[i] + (1.f - f) * drhodt[i];
}

// midpoint.cl:203-226 (residuals)
template <int D>
__global__ void mp_residuals_k(const int* __restrict__ imove,
                               const float* __restrict__ m,
                               const float* __restrict__ u,
                               const float* __restrict__ dudt_in,
                               const float* __restrict__ dudt,
                               const float* __restrict__ rho,
                               const float* __restrict__ p,
                               const float* __restrict__ drhodt_in,
                               const float* __restrict__ drhodt,
                               float* __restrict__ res, uint32_t N)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] <= 0) {
        res[i] = 0.f;
        return;
    }
    const float rho2 = rho[i] * rho[i];
    const f3 dd = sub<D>(ldv<D>(dudt, i), ldv<D>(dudt_in, i));
    res[i] = m[i] * (fabsf(dot<D>(ldv<D>(u, i), dd)) +
                     fabsf(p[i] / rho2 * (drhodt[i] - drhodt_in[i])));
}

// midpoint.cl:252-270 (corrector): the whole step with the last rates
template <int D>
__global__ void mp_corrector_k(const int* __restrict__ imove,
                               const float* __restrict__ r_in,
                               float* __restrict__ r,
                               const float* __restrict__ u_in,
                               float* __restrict__ u,
                               const float* __restrict__ dudt,
                               const float* __restrict__ rho_in,
                               float* __restrict__ rho,
                               const float* __restrict__ drhodt, uint32_t N,
                               const float* __restrict__ dtp)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] <= 0) return;
    const float dt = *dtp;
    const float hdt2 = 0.5f * dt * dt;
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        r[a] = r_in[a] + dt * u_in[a] + hdt2 * dudt[a];
        u[a] = u_in[a] + dt * dudt[a];
    }
    rho[i] = rho_in[i] + dt * drhodt[i];
}

// basic/time_scheme/improved_euler.cl:125-147
template <int D>
__global__ void corrector_k(const int* __restrict__ imove,
                            float* __restrict__ r, float* __restrict__ u,
                            const float* __restrict__ dudt,
                            float* __restrict__ rho,
                            const float* __restrict__ drhodt,
                            const float* __restrict__ dudt_in,
                            const float* __restrict__ drhodt_in, uint32_t N,
                            const float* dtp, const uint32_t* status)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    // a failure flagged earlier (ihoc overflow, failed assert): the state
    // stays where it was
    if (status && *status) return;
    if (imove[i] > 0) {
        const float DT = 0.5f * (*dtp);
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
}

// basic/time_scheme/euler.cl (predictor): back up the state
template <int D>
__global__ void euler_predictor_k(const float* __restrict__ r,
                                  const float* __restrict__ u,
                                  const float* __restrict__ dudt,
                                  const float* __restrict__ rho,
                                  const float* __restrict__ drhodt,
                                  float* __restrict__ r_in,
                                  float* __restrict__ u_in,
                                  float* __restrict__ dudt_in,
                                  float* __restrict__ rho_in,
                                  float* __restrict__ drhodt_in, uint32_t N)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        dudt_in[a] = dudt[a];
        u_in[a] = u[a];
        r_in[a] = r[a];
    }
    drhodt_in[i] = drhodt[i];
    rho_in[i] = rho[i];
}

// basic/time_scheme/euler.cl (corrector)
template <int D>
__global__ void euler_corrector_k(const int* __restrict__ imove,
                                  float* __restrict__ r, float* __restrict__ u,
                                  const float* __restrict__ dudt,
                                  float* __restrict__ rho,
                                  const float* __restrict__ drhodt, uint32_t N,
                                  const float* dtp)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] > 0) {
        const float dt = *dtp;
        const float hdt2 = 0.5f * dt * dt;
#pragma unroll
        for (int k = 0; k < VS; k++) {
            const size_t a = (size_t)i * VS + k;
            const float u0 = u[a];
            r[a] += dt * u0 + hdt2 * dudt[a];
            u[a] = u0 + dt * dudt[a];
        }
        rho[i] += dt * drhodt[i];
    }
}

// basic/Domain.cl:48-90
template <int D>
__global__ void domain_k(int* __restrict__ imove, float* __restrict__ r_in,
                         float* __restrict__ u_in, float* __restrict__ dudt_in,
                         float* __restrict__ m, uint32_t N,
                         const float* dmin, const float* dmax)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] <= -255) return;
    bool out = false;
#pragma unroll
    for (int k = 0; k < D; k++) {
        const float c = r_in[(size_t)i * VS + k];
        if (isnan(c) || isinf(c) || c < dmin[k] || c > dmax[k]) out = true;
    }
    if (out) {
        imove[i] = -256;
        m[i] = 0.f;
#pragma unroll
        for (int k = 0; k < VS; k++) {
            u_in[(size_t)i * VS + k] = 0.f;
            dudt_in[(size_t)i * VS + k] = 0.f;
            r_in[(size_t)i * VS + k] = dmax[k];
        }
    }
}

// The head of the step in one pass: the improved-Euler predictor, then
// basic/Domain.cl on the positions it predicts.  Each part is the
// expression of its own kernel (predictor_k, domain_k) in the same order;
// a particle the domain check sends out gets the domain kernel's values
// in place of the predicted ones, as it would overwrite them, so the
// results are those of the two launches bit for bit.
template <int D>
__global__ void step_head_k(int* __restrict__ imove,
                            const float* __restrict__ r,
                            const float* __restrict__ u,
                            const float* __restrict__ dudt,
                            const float* __restrict__ rho,
                            const float* __restrict__ drhodt,
                            float* __restrict__ r_in, float* __restrict__ u_in,
                            float* __restrict__ dudt_in,
                            float* __restrict__ rho_in,
                            float* __restrict__ drhodt_in,
                            float* __restrict__ m, uint32_t N,
                            const float* dtp, const uint32_t* status,
                            const float* dmin, const float* dmax)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int im = imove[i];
    float DT = *dtp;
    // a failure flagged earlier: predict no motion (the state is kept)
    if ((im <= 0) || (status && *status)) DT = 0.f;
    const float hdt2 = 0.5f * DT * DT;
    float ai[VS], ui[VS], ri[VS];
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        const float a_ = dudt[a];
        const float u_ = u[a];
        ai[k] = a_;
        ui[k] = u_ + DT * a_;
        ri[k] = r[a] + DT * u_ + hdt2 * a_;
    }
    const float dr_ = drhodt[i];
    drhodt_in[i] = dr_;
    rho_in[i] = rho[i] + DT * dr_;
    bool out = false;
    if (im > -255) {
#pragma unroll
        for (int k = 0; k < D; k++) {
            const float c = ri[k];
            if (isnan(c) || isinf(c) || c < dmin[k] || c > dmax[k]) out = true;
        }
    }
    if (out) {
        imove[i] = -256;
        m[i] = 0.f;
#pragma unroll
        for (int k = 0; k < VS; k++) {
            ui[k] = 0.f;
            ai[k] = 0.f;
            ri[k] = dmax[k];
        }
    }
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        dudt_in[a] = ai[k];
        u_in[a] = ui[k];
        r_in[a] = ri[k];
    }
}

// The reorder and what follows it in one pass: basic/Sort.cl stage 1 and
// stage 2 (gathers by id_unsorted: id, iset, imove, r, normal, tangent;
// rho, m, u, and dudt, drhodt into their _in copies), then basic/EOS.cl and
// basic/Binormal.cl on the gathered values.  Each part is the expression
// of its own kernel (permute_k, eos_k, binormal_k) on the values the
// gathers left in registers, so the results are those of the separate
// launches bit for bit.
template <int D>
__global__ void sort_chain_k(
    const uint32_t* __restrict__ perm, uint32_t N,
    const uint32_t* __restrict__ id_in, uint32_t* __restrict__ id,
    const uint32_t* __restrict__ iset_in, uint32_t* __restrict__ iset,
    const int* __restrict__ imove_in, int* __restrict__ imove,
    const float* __restrict__ r_in, float* __restrict__ r,
    const float* __restrict__ normal_in, float* __restrict__ normal,
    const float* __restrict__ tangent_in, float* __restrict__ tangent,
    const float* __restrict__ rho_in, float* __restrict__ rho,
    const float* __restrict__ m_in, float* __restrict__ m,
    const float* __restrict__ u_in, float* __restrict__ u,
    const float* __restrict__ dudt, float* __restrict__ dudt_in,
    const float* __restrict__ drhodt, float* __restrict__ drhodt_in,
    float* __restrict__ p, const float* __restrict__ refd, const float* csp,
    const float* p0p, float* __restrict__ binormal)
{
    using V = typename std::conditional<D == 2, float2, float4>::type;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint32_t j = perm[i];
    const uint32_t id_j = id_in[j], st = iset_in[j];
    const int im = imove_in[j];
    const V r_j = reinterpret_cast<const V*>(r_in)[j];
    const V n_j = reinterpret_cast<const V*>(normal_in)[j];
    const V t_j = reinterpret_cast<const V*>(tangent_in)[j];
    const float rho_j = rho_in[j], m_j = m_in[j], dr_j = drhodt[j];
    const V u_j = reinterpret_cast<const V*>(u_in)[j];
    const V a_j = reinterpret_cast<const V*>(dudt)[j];
    id[i] = id_j;
    iset[i] = st;
    imove[i] = im;
    reinterpret_cast<V*>(r)[i] = r_j;
    reinterpret_cast<V*>(normal)[i] = n_j;
    rho[i] = rho_j;
    m[i] = m_j;
    reinterpret_cast<V*>(u)[i] = u_j;
    reinterpret_cast<V*>(dudt_in)[i] = a_j;
    drhodt_in[i] = dr_j;
    // basic/EOS.cl
    if (!((im <= 0) && (im != -1))) {
        const float cs = *csp;
        p[i] = (*p0p) + cs * cs * (rho_j - refd[st]);
    }
    // basic/Binormal.cl
    if constexpr (D == 2) {
        reinterpret_cast<float2*>(binormal)[i] = make_float2(0.f, 0.f);
        reinterpret_cast<float2*>(tangent)[i] = make_float2(n_j.y, -n_j.x);
    } else {
        reinterpret_cast<float4*>(tangent)[i] = t_j;
        reinterpret_cast<float4*>(binormal)[i] = make_float4(
            n_j.y * t_j.z - n_j.z * t_j.y, n_j.z * t_j.x - n_j.x * t_j.z,
            n_j.x * t_j.y - n_j.y * t_j.x, 0.f);
    }
}

// basic/EOS.cl:57-73
__global__ void eos_k(const uint32_t* __restrict__ iset,
                      const int* __restrict__ imove,
                      const float* __restrict__ rho, float* __restrict__ p,
                      const float* __restrict__ refd, uint32_t N,
                      const float* csp, const float* p0p)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int im = imove[i];
    if ((im <= 0) && (im != -1)) return;
    const float cs = *csp;
    p[i] = (*p0p) + cs * cs * (rho[i] - refd[iset[i]]);
}

// Tait equation of state (no counterpart in the reference, whose only
// liquid EOS is the linear one of basic/EOS.cl): with B = rho0 cs^2 / gamma,
//   p = p0 + B ((rho / rho0)^gamma - 1),
// for the particles basic/EOS.cl:35 computes (imove > 0 or imove == -1).
// An integral gamma in [1, 16] is raised by binary powering in a fixed
// order (the same products on the host), any other through powf.
__device__ __forceinline__ float tait_pow(float x, float g)
{
    const float gi = rintf(g);
    if (gi == g && gi >= 1.f && gi <= 16.f) {
        unsigned n = (unsigned)gi;
        float r = 1.f, b = x;
        bool first = true;
        while (n) {
            if (n & 1u) {
                r = first ? b : r * b;
                first = false;
            }
            n >>= 1;
            if (n) b = b * b;
        }
        return r;
    }
    return powf(x, g);
}

__global__ void tait_eos_k(const uint32_t* __restrict__ iset,
                           const int* __restrict__ imove,
                           const float* __restrict__ rho, float* __restrict__ p,
                           const float* __restrict__ refd, uint32_t N,
                           const float* csp, const float* p0p,
                           const float* gammap)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int im = imove[i];
    if ((im <= 0) && (im != -1)) return;
    const float cs = *csp, g = *gammap;
    const float rho0 = refd[iset[i]];
    const float B = rho0 * cs * cs / g;
    p[i] = (*p0p) + B * (tait_pow(rho[i] / rho0, g) - 1.f);
}

// cfd/Boundary/BI/InterpolationShepard.cl:48-76 with the inverse of the
// Tait EOS for the element density:
//   rho = rho0 ((p - p0) / B + 1)^(1 / gamma)
__global__ void bi_interp_shepard_tait_k(const uint32_t* __restrict__ iset,
                                         const int* __restrict__ imove,
                                         const float* __restrict__ shepard,
                                         float* __restrict__ rho,
                                         float* __restrict__ p,
                                         const float* __restrict__ refd,
                                         uint32_t N, const float* csp,
                                         const float* p0p, const float* gammap)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != -3) return;
    float s = shepard[i];
    if (s < 1.0E-6f) s = 1.f;
    const float pi = p[i] / s;
    p[i] = pi;
    const float cs = *csp, g = *gammap;
    const float rho0 = refd[iset[i]];
    const float B = rho0 * cs * cs / g;
    rho[i] = rho0 * powf((pi - (*p0p)) / B + 1.f, 1.f / g);
}

// basic/Binormal.cl:37-53
template <int D>
__global__ void binormal_k(const float* __restrict__ normal,
                           float* __restrict__ tangent,
                           float* __restrict__ binormal, uint32_t N)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if constexpr (D == 2) {
        const float2 n = reinterpret_cast<const float2*>(normal)[i];
        reinterpret_cast<float2*>(binormal)[i] = make_float2(0.f, 0.f);
        reinterpret_cast<float2*>(tangent)[i] = make_float2(n.y, -n.x);
    } else {
        const float4 a = reinterpret_cast<const float4*>(normal)[i];
        const float4 b = reinterpret_cast<const float4*>(tangent)[i];
        reinterpret_cast<float4*>(binormal)[i] =
            make_float4(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
                        a.x * b.y - a.y * b.x, 0.f);
    }
}

// cfd/Rates.cl:55-77
template <int D>
__global__ void rates_k(const uint32_t* __restrict__ iset,
                        const int* __restrict__ imove,
                        const float* __restrict__ grad_p,
                        const float* __restrict__ lap_u,
                        const float* __restrict__ div_u,
                        float* __restrict__ dudt, float* __restrict__ drhodt,
                        const float* __restrict__ visc_dyn, uint32_t N,
                        const float* g)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != 1) return;
    const float mu = visc_dyn[iset[i]];
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        dudt[a] = -grad_p[a] + mu * lap_u[a] + g[k];
    }
    drhodt[i] = -div_u[i];
}

// cfd/Energy/EnergyKin.cl:40-56
template <int D>
__global__ void energy_kin_k(float* __restrict__ ek,
                             const int* __restrict__ imove,
                             const float* __restrict__ u,
                             const float* __restrict__ m, uint32_t N)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != 1) {
        ek[i] = 0.f;
        return;
    }
    const f3 u_i = ldv<D>(u, i);
    ek[i] = 0.5f * m[i] * dot<D>(u_i, u_i);
}

// cfd/Energy/Energy.cl:59-86 (power)
template <int D>
__global__ void energy_power_k(float* __restrict__ dek,
                               float* __restrict__ dep,
                               float* __restrict__ dec,
                               const int* __restrict__ imove,
                               const float* __restrict__ u,
                               const float* __restrict__ rho,
                               const float* __restrict__ m,
                               const float* __restrict__ p,
                               const float* __restrict__ dudt,
                               const float* __restrict__ drhodt, uint32_t N,
                               const float* __restrict__ g)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != 1) {
        dek[i] = dep[i] = dec[i] = 0.f;
        return;
    }
    const f3 gv = { g[0], g[1], (D == 3) ? g[2] : 0.f };
    const f3 u_i = ldv<D>(u, i);
    dep[i] = -m[i] * dot<D>(gv, u_i);
    dek[i] = m[i] * dot<D>(u_i, ldv<D>(dudt, i));
    dec[i] = m[i] * p[i] / (rho[i] * rho[i]) * drhodt[i];
}

// cfd/Energy/Energy.cl:113-142 (energy)
template <int D>
__global__ void energy_parts_k(float* __restrict__ ek, float* __restrict__ ep,
                               float* __restrict__ ec,
                               const uint32_t* __restrict__ iset,
                               const int* __restrict__ imove,
                               const float* __restrict__ r,
                               const float* __restrict__ u,
                               const float* __restrict__ rho,
                               const float* __restrict__ m,
                               const float* __restrict__ refd, uint32_t N,
                               const float* __restrict__ g,
                               const float* __restrict__ csp)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != 1) {
        ek[i] = ep[i] = ec[i] = 0.f;
        return;
    }
    const float cs = *csp;
    const f3 gv = { g[0], g[1], (D == 3) ? g[2] : 0.f };
    const f3 u_i = ldv<D>(u, i);
    ek[i] = 0.5f * m[i] * dot<D>(u_i, u_i);
    ep[i] = -m[i] * dot<D>(gv, ldv<D>(r, i));
    const float rho0 = refd[iset[i]];
    ec[i] = m[i] * cs * cs * (rho0 / rho[i] + logf(rho[i] / rho0) - 1.f);
}

// basic/deltaSPH.cl:330-350, cfd/deltaSPH.cl:29 (fluid only)
__global__ void delta_rate_k(const uint32_t* __restrict__ iset,
                             const int* __restrict__ imove,
                             const float* __restrict__ rho,
                             const float* __restrict__ lap_p,
                             float* __restrict__ drhodt,
                             const float* __restrict__ refd,
                             const float* __restrict__ delta, uint32_t N,
                             const float* dtp)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != 1) return;
    const uint32_t s = iset[i];
    const float delta_f = delta[s] * (*dtp) * rho[i] / refd[s];
    drhodt[i] += delta_f * lap_p[i];
}

// cfd/TimeStep.cl:56-77 (length of the whole vec, w included)
template <int D>
__global__ void timestep_k(const int* __restrict__ imove,
                           const float* __restrict__ u,
                           float* __restrict__ dt_var, uint32_t N,
                           const float* dtp, const float* dt_minp,
                           const float* courantp, const float* dt_Map,
                           const float* hp)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float dt = *dtp;
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

// cfd/SensorsRenormalization.cl:42-67
template <int D>
__global__ void sensors_renorm_k(const int* __restrict__ imove,
                                 const float* __restrict__ shepard,
                                 float* __restrict__ u, float* __restrict__ rho,
                                 float* __restrict__ p, uint32_t N)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != 0) return;
    float s = shepard[i];
    if (s < 1.0E-6f) s = 1.f;
#pragma unroll
    for (int k = 0; k < VS; k++) u[(size_t)i * VS + k] /= s;
    rho[i] /= s;
    p[i] /= s;
}

// cfd/Boundary/BI/Shepard.cl:173-195 (apply)
template <int D>
__global__ void bi_shepard_apply_k(const int* __restrict__ imove,
                                   const float* __restrict__ shepard,
                                   float* __restrict__ grad_p,
                                   float* __restrict__ lap_u,
                                   float* __restrict__ div_u, uint32_t N)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != 1) return;
    const float s = shepard[i];
#pragma unroll
    for (int k = 0; k < VS; k++) grad_p[(size_t)i * VS + k] /= s;
#pragma unroll
    for (int k = 0; k < VS; k++) lap_u[(size_t)i * VS + k] /= s;
    div_u[i] /= s;
}

// The per-particle tail of the step in one pass: BI Shepard apply
// (APPLY), cfd Rates, the delta-SPH rate (DELTA) and the improved-Euler
// corrector, which run one after the other and each touch their own
// particle only.  Each part is the expression of its own kernel above
// (bi_shepard_apply_k, rates_k, delta_rate_k, corrector_k), in the same
// order, on the values the part before produced (kept in registers instead
// of read back), so the results are those of the four launches bit for bit.
template <int D, bool APPLY, bool DELTA>
__global__ void step_tail_k(const uint32_t* __restrict__ iset,
                            const int* __restrict__ imove,
                            const float* __restrict__ shepard,
                            float* __restrict__ grad_p,
                            float* __restrict__ lap_u,
                            float* __restrict__ div_u,
                            float* __restrict__ dudt,
                            float* __restrict__ drhodt,
                            const float* __restrict__ visc_dyn,
                            const float* g, const float* __restrict__ lap_p,
                            const float* __restrict__ refd,
                            const float* __restrict__ delta,
                            float* __restrict__ r, float* __restrict__ u,
                            float* __restrict__ rho,
                            const float* __restrict__ dudt_in,
                            const float* __restrict__ drhodt_in, uint32_t N,
                            const float* dtp, const uint32_t* status)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    // every operand is loaded up front (one round trip; the rows of the
    // particles a part skips are read and left unused)
    const int im = imove[i];
    const bool stopped = status && *status;
    const uint32_t st = iset[i];
    float gp[VS], lu[VS], du_st[VS], du_in[VS], uu[VS], rr[VS];
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        gp[k] = grad_p[a];
        lu[k] = lap_u[a];
        du_st[k] = dudt[a];
        du_in[k] = dudt_in[a];
        uu[k] = u[a];
        rr[k] = r[a];
    }
    float dv = div_u[i];
    const float sh = APPLY ? shepard[i] : 1.f;
    const float lp = DELTA ? lap_p[i] : 0.f;
    const float rho_i = rho[i];
    const float dr_st = drhodt[i];
    const float dr_in = drhodt_in[i];
    const float dt = *dtp;
    const bool fluid = im == 1;
    float du[VS];
    float dr = dr_st;
#pragma unroll
    for (int k = 0; k < VS; k++) du[k] = du_st[k];
    if (fluid) {
        if (APPLY) {
#pragma unroll
            for (int k = 0; k < VS; k++) {
                gp[k] /= sh;
                grad_p[(size_t)i * VS + k] = gp[k];
            }
#pragma unroll
            for (int k = 0; k < VS; k++) {
                lu[k] /= sh;
                lap_u[(size_t)i * VS + k] = lu[k];
            }
            dv /= sh;
            div_u[i] = dv;
        }
        const float mu = visc_dyn[st];
#pragma unroll
        for (int k = 0; k < VS; k++) {
            du[k] = -gp[k] + mu * lu[k] + g[k];
            dudt[(size_t)i * VS + k] = du[k];
        }
        dr = -dv;
        if (DELTA) {
            const float delta_f = delta[st] * dt * rho_i / refd[st];
            dr += delta_f * lp;
        }
        drhodt[i] = dr;
    }
    // the corrector: a failure flagged earlier leaves the state as it was
    if (stopped) return;
    if (im > 0) {
        const float DT = 0.5f * dt;
        const float DT2 = DT * DT;
#pragma unroll
        for (int k = 0; k < VS; k++) {
            const size_t a = (size_t)i * VS + k;
            const float da = du[k] - du_in[k];
            u[a] = uu[k] + DT * da;
            r[a] = rr[k] + DT2 * da;
        }
        rho[i] = rho_i + DT * (dr - dr_in);
    }
}

// cfd/Boundary/BI/GradP.cl:55-83 (freeslip)
template <int D>
__global__ void bi_gradp_k(const uint32_t* __restrict__ iset,
                           const int* __restrict__ imove,
                           const float* __restrict__ shepard,
                           const float* __restrict__ lap_u,
                           const float* __restrict__ dudt,
                           float* __restrict__ grad_p,
                           const float* __restrict__ visc_dyn,
                           const float* __restrict__ refd, uint32_t N,
                           const float* g)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != -3) return;
    float s = shepard[i];
    if (s < 1.0E-6f) s = 1.f;
    const float f = visc_dyn[iset[i]] / refd[iset[i]];
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        grad_p[a] = f * lap_u[a] / s + g[k] - dudt[a];
    }
}

// cfd/Boundary/BI/InterpolationShepard.cl:48-76
__global__ void bi_interp_shepard_k(const uint32_t* __restrict__ iset,
                                    const int* __restrict__ imove,
                                    const float* __restrict__ shepard,
                                    float* __restrict__ rho,
                                    float* __restrict__ p,
                                    const float* __restrict__ refd,
                                    uint32_t N, const float* csp,
                                    const float* p0p)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != -3) return;
    float s = shepard[i];
    if (s < 1.0E-6f) s = 1.f;
    const float pi = p[i] / s;
    p[i] = pi;
    const float cs = *csp;
    rho[i] = refd[iset[i]] + (pi - (*p0p)) / (cs * cs);
}

// ===========================================================================
// Neighbour-walk kernels
// ===========================================================================
struct LLArgs {
    const uint32_t* icell;
    const uint32_t* ihoc;
    const uint32_t* n_cells;
    const uint32_t* status;
    uint32_t N;
    // optional cell list: cell ends (AQ_CELL_END bits) and, when `flagged`,
    // the kinds each cell holds
    const uint32_t* cell_end = nullptr;
    uint32_t flagged = 0;
};

// walk() over the cell list when one is given: the same candidates in the
// same order, the cell end read once per cell instead of icell once per
// candidate, cells holding none of the `want` kinds (AQ_CELL_FLUID /
// AQ_CELL_ELEM) skipped, which only drops candidates the caller rejects, and
// the positions and imove of WB candidates loaded together before their
// bodies run (one memory latency per batch instead of per candidate).
// body(j, r_j, imove_j).
template <int D, class Body>
__device__ __forceinline__ void walk_ll(const LLArgs& ll, uint32_t c_i,
                                        uint32_t want,
                                        const float* __restrict__ r,
                                        const int* __restrict__ imove,
                                        Body body)
{
    if (!ll.cell_end) {
        walk<D>(c_i, ll.N, ll.n_cells[0], ll.n_cells[1], ll.icell, ll.ihoc,
                [&](uint32_t j) { body(j, ldv<D>(r, j), imove[j]); });
        return;
    }
    constexpr uint32_t WB = 4;
    const uint32_t nx = ll.n_cells[0], ny = ll.n_cells[1];
    if (!ll.flagged) want = 0;
    constexpr int NC = (D == 2) ? 9 : 27;
#pragma unroll 1
    for (int k = 0; k < NC; k++) {
        int ci, cj, ck;
        if constexpr (D == 2) {
            ci = k / 3 - 1;
            cj = k % 3 - 1;
            ck = 0;
        } else {
            ci = k / 9 - 1;
            cj = (k / 3) % 3 - 1;
            ck = k % 3 - 1;
        }
        const uint32_t c_j = c_i + (uint32_t)ci + (uint32_t)cj * nx +
                             (uint32_t)ck * nx * ny;
        const uint32_t j0 = ll.ihoc[c_j];
        if (j0 >= ll.N) continue;
        const uint32_t ce = ll.cell_end[c_j];
        if (want && !(ce & want)) continue;
        const uint32_t j1 = ce & AQ_CELL_END;
        uint32_t j = j0;
        for (; j + WB <= j1; j += WB) {
            f3 rb[WB];
            int ib[WB];
#pragma unroll
            for (uint32_t t = 0; t < WB; t++) {
                rb[t] = ldv<D>(r, j + t);
                ib[t] = imove[j + t];
            }
#pragma unroll
            for (uint32_t t = 0; t < WB; t++) body(j + t, rb[t], ib[t]);
        }
        for (; j < j1; j++) body(j, ldv<D>(r, j), imove[j]);
    }
}

// The particles a per-particle kernel runs on: every index below N, or the
// compacted list[0 .. *count) (in any order: each particle's result is
// independent of the others).
struct Sel {
    const uint32_t* list;
    const uint32_t* count;
    __device__ uint32_t n(uint32_t N) const { return list ? *count : N; }
    __device__ uint32_t at(uint32_t k) const { return list ? list[k] : k; }
};

// Indices i < N with lo <= imove[i] <= hi (and near[i] != 0 when near is
// given), appended to list in per-wave runs; *count is zero on entry and
// ends as their number.
__global__ void __launch_bounds__(BS)
    select_k(const int* __restrict__ imove, const uint8_t* __restrict__ near,
             int lo, int hi, uint32_t N, uint32_t* __restrict__ list,
             uint32_t* __restrict__ count)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    int im = 0;
    if (i < N) im = imove[i];
    const bool take = (i < N) && (im >= lo) && (im <= hi) &&
                      (!near || near[i]);
    const uint64_t mk = __ballot(take);
    if (mk == 0) return;
    uint32_t base = 0;
    if ((threadIdx.x & 63) == 0)
        base = atomicAdd(count, (uint32_t)__popcll(mk));
    base = __builtin_amdgcn_readfirstlane(base);
    const uint32_t below = __builtin_amdgcn_mbcnt_hi(
        (uint32_t)(mk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk, 0u));
    if (take) list[base + below] = i;
}

__global__ void set_zero_k(uint32_t* __restrict__ x) { *x = 0u; }

// basic/Shepard.cl:76-125 (cfd/Shepard.cl when cfd != 0)
template <int D>
__global__ void __launch_bounds__(BS)
    shepard_k(const int* __restrict__ imove, const float* __restrict__ r,
              const float* __restrict__ rho, const float* __restrict__ m,
              float* __restrict__ shepard, LLArgs ll, float H, float CONW,
              float smax, int cfd, Sel sel)
{
    AQ_STATUS_SKIP;
    auto excl = [&](int v) { return (cfd & 1) ? (v != 1) : (v >= 3); };
    // cfd: fluid candidates only
    const uint32_t want = (cfd & 1) ? AQ_CELL_FLUID : 0u;
    const uint32_t n = sel.n(ll.N);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += gridDim.x * blockDim.x) {
        const uint32_t i = sel.at(k);
        const int im = imove[i];
        if ((cfd & 2) && im == 1) continue;  // fluid done by the fused walk
        if ((im < -3) || ((im > 0) && excl(im))) continue;
        const f3 r_i = ldv<D>(r, i);
        float acc = shepard[i];
        walk_ll<D>(ll, ll.icell[i], want, r, imove,
                   [&](uint32_t j, const f3& r_j, int im_j) {
            const f3 r_ij = sub<D>(r_j, r_i);
            const float s = dot<D>(r_ij, r_ij);
            if (s > smax) return;
            if (excl(im_j)) return;
            const float q = sqrtf(s) / H;
            acc += kernelW<D>(q) * CONW * m[j] / rho[j];
        });
        shepard[i] = acc;
    }
}

// cfd/Interactions.cl:60-145
template <int D>
__global__ void __launch_bounds__(BS)
    interactions_k(const int* __restrict__ imove, const float* __restrict__ r,
                   const float* __restrict__ u, const float* __restrict__ rho,
                   const float* __restrict__ m, const float* __restrict__ p,
                   float* __restrict__ grad_p, float* __restrict__ lap_u,
                   float* __restrict__ div_u, LLArgs ll, float H, float CONF,
                   float smax, int lap)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ll.N) return;
    AQ_STATUS_SKIP;
    if (imove[i] != 1) return;
    constexpr float CLEARY = (D == 2) ? 8.f : 10.f;
    const f3 r_i = ldv<D>(r, i);
    const f3 u_i = ldv<D>(u, i);
    const float p_i = p[i];
    const float rho_i = rho[i];
    f3 gp = ldv<D>(grad_p, i), lu = ldv<D>(lap_u, i);
    float dv = div_u[i];
    walk<D>(ll.icell[i], ll.N, ll.n_cells[0], ll.n_cells[1], ll.icell,
            ll.ihoc, [&](uint32_t j) {
                const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
                const float s = dot<D>(r_ij, r_ij);
                if (s > smax) return;
                if (i == j) return;
                if (imove[j] != 1) return;
                const float q = sqrtf(s) / H;
                const float rho_j = rho[j];
                const float p_j = p[j];
                const f3 u_ij = sub<D>(ldv<D>(u, j), u_i);
                const float udr = dot<D>(u_ij, r_ij);
                const float f_ij = kernelF<D>(q) * CONF * m[j];
                axpy<D>(gp, (p_i + p_j) / (rho_i * rho_j) * f_ij, r_ij);
                if (lap == AQ_LAP_MONAGHAN) {
                    const float r2 = (q * q + 0.01f) * H * H;
                    axpy<D>(lu,
                            f_ij * CLEARY * udr / (r2 * rho_i * rho_j), r_ij);
                } else {
                    axpy<D>(lu, f_ij * 2.f / (rho_i * rho_j), u_ij);
                }
                dv += udr * f_ij * rho_i / rho_j;
            });
    stxyz<D>(grad_p, i, gp);
    stxyz<D>(lap_u, i, lu);
    div_u[i] = dv;
}

// basic/deltaSPH.cl:191-242 (lapp), fluid only
template <int D>
__global__ void __launch_bounds__(BS)
    lapp_k(const int* __restrict__ imove, const float* __restrict__ r,
           const float* __restrict__ rho, const float* __restrict__ m,
           const float* __restrict__ p, float* __restrict__ lap_p, LLArgs ll,
           float H, float CONF, float smax)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ll.N) return;
    AQ_STATUS_SKIP;
    if (imove[i] != 1) return;
    const f3 r_i = ldv<D>(r, i);
    const float p_i = p[i];
    float acc = lap_p[i];
    walk<D>(ll.icell[i], ll.N, ll.n_cells[0], ll.n_cells[1], ll.icell,
            ll.ihoc, [&](uint32_t j) {
                const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
                const float s = dot<D>(r_ij, r_ij);
                if (s > smax) return;
                if ((i == j) || (imove[j] != 1)) return;
                const float q = sqrtf(s) / H;
                const float f_ij = kernelF<D>(q) * CONF * m[j] / rho[j];
                acc += (p[j] - p_i) * f_ij;
            });
    lap_p[i] = acc;
}

// cfd/Sensors.cl:57-130
template <int D>
__global__ void __launch_bounds__(BS)
    sensors_k(const int* __restrict__ imove, const float* __restrict__ r,
              const float* __restrict__ m, float* __restrict__ u,
              float* __restrict__ rho, float* __restrict__ p, LLArgs ll,
              float H, float CONW, float smax, const float* per)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ll.N) return;
    AQ_STATUS_SKIP;
    if (imove[i] != 0) return;
    const f3 r_i = ldv<D>(r, i);
    f3 uu = { 0.f, 0.f, 0.f };
    float rr = 0.f, pp = 0.f;
    auto body = [&](uint32_t j, const f3& r_j) {
        const f3 r_ij = sub<D>(r_j, r_i);
        const float s = dot<D>(r_ij, r_ij);
        if (s > smax) return;
        if (i == j) return;
        if (imove[j] != 1) return;
        const float q = sqrtf(s) / H;
        const float rho_j = rho[j];
        const float w_ij = kernelW<D>(q) * CONW * m[j] / rho_j;
        const f3 u_j = ldv<D>(u, j);
        axpy<D>(uu, w_ij, u_j);
        rr += rho_j * w_ij;
        pp += p[j] * w_ij;
    };
    if (per)  // a periodic box: wrapped cells, images at r + shift
        walk_per<D>(ll.icell[i], ll.N, ll.n_cells, per, ll.icell, ll.ihoc,
                    [&](uint32_t j, const f3& sh) {
                        body(j, per_shift<D>(ldv<D>(r, j), sh, per));
                    });
    else
        walk<D>(ll.icell[i], ll.N, ll.n_cells[0], ll.n_cells[1], ll.icell,
                ll.ihoc, [&](uint32_t j) { body(j, ldv<D>(r, j)); });
    stxyz<D>(u, i, uu);
    rho[i] = rr;
    p[i] = pp;
}

// cfd/Boundary/BI/Shepard.cl:61-150 (compute)
template <int D>
__global__ void __launch_bounds__(BS)
    bi_shepard_compute_k(const int* __restrict__ imove,
                         const float* __restrict__ r,
                         const float* __restrict__ normal,
                         const float* __restrict__ tangent,
                         const float* __restrict__ binormal,
                         const float* __restrict__ m,
                         float* __restrict__ shepard, LLArgs ll, float H,
                         float CONW, float DIMS, float smax, int skip_fluid,
                         Sel sel)
{
    AQ_STATUS_SKIP;
    const float eps = 1E-8f * H;
    const uint32_t n = sel.n(ll.N);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += gridDim.x * blockDim.x) {
        const uint32_t i = sel.at(k);
        const int im = imove[i];
        if ((im < -3) || (im > 1)) continue;
        if (skip_fluid && im == 1) continue;  // fluid done by the fused walk
        const f3 r_i = ldv<D>(r, i);
        float acc = 1.f;
        bool self_added = false;
        walk_ll<D>(ll, ll.icell[i], AQ_CELL_ELEM, r, imove,
                   [&](uint32_t j, const f3& r_j, int im_j) {
            if (im_j != -3) return;
            if (i == j) {
                if (!self_added) {
                    self_added = true;
                    acc -= 0.5f;
                }
                return;
            }
            const f3 r_ij = sub<D>(r_j, r_i);
            const float s = dot<D>(r_ij, r_ij);
            if (s > smax) return;
            const float q = sqrtf(s) / H;
            const f3 n_j = ldv<D>(normal, j);
            const f3 t_j = ldv<D>(tangent, j);
            const f3 b_j = ldv<D>(binormal, j);
            const float r_n = dot<D>(r_ij, n_j);
            const float area_j = m[j];
            if ((r_n > -eps) && (r_n < eps)) {
                if (!self_added) {
                    const float r_t = fabsf(dot<D>(r_ij, t_j));
                    const float r_b = fabsf(dot<D>(r_ij, b_j));
                    const float dr =
                        0.55f * pow_cl(area_j, 1.f / (DIMS - 1.f));
                    if ((r_t <= dr) && (r_b <= dr)) {
                        self_added = true;
                        acc -= 0.5f;
                    }
                }
                return;
            }
            const float r_t = fabsf(dot<D>(r_ij, t_j));
            const float r_b = fabsf(dot<D>(r_ij, b_j));
            acc += r_n * CONW * kernelS_P<D>(q) * area_j +
                   kernelS_D<D>(fabsf(r_n), r_t, r_b, area_j);
        });
        shepard[i] = acc;
    }
}

// cfd/Boundary/BI/LapU.cl:58-123 (freeslip)
template <int D>
__global__ void __launch_bounds__(BS)
    bi_lapu_k(const int* __restrict__ imove, const float* __restrict__ r,
              const float* __restrict__ u, const float* __restrict__ rho,
              const float* __restrict__ m, float* __restrict__ lap_u,
              LLArgs ll, float H, float CONF, float smax, int lap, Sel sel)
{
    AQ_STATUS_SKIP;
    const uint32_t n = sel.n(ll.N);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += gridDim.x * blockDim.x) {
        const uint32_t i = sel.at(k);
        if (imove[i] != -3) continue;
        constexpr float CLEARY = (D == 2) ? 8.f : 10.f;
        const f3 r_i = ldv<D>(r, i);
        const f3 u_i = ldv<D>(u, i);
        f3 lu = ldv<D>(lap_u, i);
        walk_ll<D>(ll, ll.icell[i], AQ_CELL_FLUID, r, imove,
                   [&](uint32_t j, const f3& r_j, int im_j) {
            const f3 r_ij = sub<D>(r_j, r_i);
            const float s = dot<D>(r_ij, r_ij);
            if (s > smax) return;
            if (i == j) return;
            if (im_j != 1) return;
            const float q = sqrtf(s) / H;
            const float rho_j = rho[j];
            const float f_ij = kernelF<D>(q) * CONF * m[j];
            const f3 u_ij = sub<D>(ldv<D>(u, j), u_i);
            if (lap == AQ_LAP_MONAGHAN) {
                const float udr = dot<D>(u_ij, r_ij);
                const float r2 = (q * q + 0.01f) * H * H;
                axpy<D>(lu, f_ij * CLEARY * udr / (r2 * rho_j), r_ij);
            } else {
                axpy<D>(lu, f_ij * 2.f / rho_j, u_ij);
            }
        });
        stxyz<D>(lap_u, i, lu);
    }
}

// cfd/Boundary/BI/Interpolation.cl:54-107
template <int D>
__global__ void __launch_bounds__(BS)
    bi_interp_k(const int* __restrict__ imove, const float* __restrict__ r,
                const float* __restrict__ m, const float* __restrict__ rho,
                const float* __restrict__ grad_p, float* __restrict__ p,
                LLArgs ll, float H, float CONW, float smax, Sel sel)
{
    AQ_STATUS_SKIP;
    const uint32_t n = sel.n(ll.N);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += gridDim.x * blockDim.x) {
        const uint32_t i = sel.at(k);
        if (imove[i] != -3) continue;
        const f3 r_i = ldv<D>(r, i);
        const float rho_i = rho[i];
        f3 gp_i = ldv<D>(grad_p, i);
        gp_i.x = rho_i * gp_i.x;
        gp_i.y = rho_i * gp_i.y;
        gp_i.z = rho_i * gp_i.z;
        float acc = 0.f;
        walk_ll<D>(ll, ll.icell[i], AQ_CELL_FLUID, r, imove,
                   [&](uint32_t j, const f3& r_j, int im_j) {
            const f3 r_ij = sub<D>(r_j, r_i);
            const float s = dot<D>(r_ij, r_ij);
            if (s > smax) return;
            if (im_j != 1) return;
            const float q = sqrtf(s) / H;
            const float w_ij = kernelW<D>(q) * CONW * m[j] / rho[j];
            acc += (p[j] - dot<D>(gp_i, r_ij)) * w_ij;
        });
        p[i] = acc;
    }
}

// cfd/Boundary/BI/Interactions.cl:51-133
template <int D>
__global__ void __launch_bounds__(BS)
    bi_interactions_k(const int* __restrict__ imove,
                      const float* __restrict__ r,
                      const float* __restrict__ normal,
                      const float* __restrict__ u,
                      const float* __restrict__ rho,
                      const float* __restrict__ m,
                      const float* __restrict__ p, float* __restrict__ grad_p,
                      float* __restrict__ div_u, LLArgs ll, float H,
                      float CONW, float smax, const uint8_t* __restrict__ near,
                      Sel sel)
{
    AQ_STATUS_SKIP;
    const uint32_t n = sel.n(ll.N);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += gridDim.x * blockDim.x) {
        const uint32_t i = sel.at(k);
        if (imove[i] != 1) continue;
        // no element inside the support: nothing is added (bit-identical skip)
        if (near && !near[i]) continue;
        const f3 r_i = ldv<D>(r, i);
        const f3 u_i = ldv<D>(u, i);
        const float p_i = p[i];
        const float rho_i = rho[i];
        f3 gp = ldv<D>(grad_p, i);
        float dv = div_u[i];
        walk_ll<D>(ll, ll.icell[i], AQ_CELL_ELEM, r, imove,
                   [&](uint32_t j, const f3& r_j, int im_j) {
            if (im_j != -3) return;
            const f3 r_ij = sub<D>(r_j, r_i);
            const float s = dot<D>(r_ij, r_ij);
            if (s > smax) return;
            const float q = sqrtf(s) / H;
            const f3 n_j = ldv<D>(normal, j);
            const float area_j = m[j];
            const float p_j = p[j];
            const f3 du = sub<D>(ldv<D>(u, j), u_i);
            const float w_ij = kernelW<D>(q) * CONW * area_j;
            axpy<D>(gp, (p_i + p_j) / rho_i * w_ij, n_j);
            dv += rho_i * dot<D>(du, n_j) * w_ij;
        });
        stxyz<D>(grad_p, i, gp);
        div_u[i] = dv;
    }
}

// ---------------------------------------------------------------------------
// Fused fluid walk.  For a fluid particle i one traversal of the neighbour
// cells evaluates, each in its own accumulator and in the reference order,
//   * the Shepard factor: cfd/Shepard.cl (SHEP == 1, fluid j incl. self) or
//     the boundary-integral compute of cfd/Boundary/BI/Shepard.cl:61-150
//     (SHEP == 2, element j),
//   * cfd/Interactions.cl:60-145 (fluid j != i),
//   * basic/deltaSPH.cl:191-242 lapp (LAPP, fluid j != i),
// and flags particles with an element inside the support (the only ones
// cfd/Boundary/BI/Interactions.cl adds anything to).  Every sum sees the
// same terms in the same order as in the separate kernels, so the results
// are bit-identical; the walk and the candidate tests are done once.
// ---------------------------------------------------------------------------
template <int D, int SHEP, bool LAPP>
__global__ void __launch_bounds__(BS)
    fluid_pass_k(const int* __restrict__ imove, const float* __restrict__ r,
                 const float* __restrict__ u, const float* __restrict__ rho,
                 const float* __restrict__ m, const float* __restrict__ p,
                 const float* __restrict__ normal,
                 const float* __restrict__ tangent,
                 const float* __restrict__ binormal,
                 float* __restrict__ shepard, float* __restrict__ grad_p,
                 float* __restrict__ lap_u, float* __restrict__ div_u,
                 float* __restrict__ lap_p, uint8_t* __restrict__ near,
                 LLArgs ll, float H, float CONW, float CONF, float DIMS,
                 float smax, int lap)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ll.N) return;
    AQ_STATUS_SKIP;
    if (imove[i] != 1) return;
    constexpr float CLEARY = (D == 2) ? 8.f : 10.f;
    const f3 r_i = ldv<D>(r, i);
    const f3 u_i = ldv<D>(u, i);
    const float p_i = p[i];
    const float rho_i = rho[i];
    f3 gp = ldv<D>(grad_p, i), lu = ldv<D>(lap_u, i);
    float dv = div_u[i];
    float lp = LAPP ? lap_p[i] : 0.f;
    float sh = (SHEP == 2) ? 1.f : (SHEP == 1 ? shepard[i] : 0.f);
    bool self_added = false;
    bool nearw = false;
    const float eps = 1E-8f * H;
    walk<D>(ll.icell[i], ll.N, ll.n_cells[0], ll.n_cells[1], ll.icell,
            ll.ihoc, [&](uint32_t j) {
                const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
                const float s = dot<D>(r_ij, r_ij);
                if (s > smax) return;
                const int im_j = imove[j];
                if (im_j == 1) {
                    const float q = sqrtf(s) / H;
                    const float m_j = m[j];
                    const float rho_j = rho[j];
                    if (SHEP == 1) sh += kernelW<D>(q) * CONW * m_j / rho_j;
                    if (i == j) return;
                    const float p_j = p[j];
                    const f3 u_ij = sub<D>(ldv<D>(u, j), u_i);
                    const float udr = dot<D>(u_ij, r_ij);
                    const float fk = kernelF<D>(q) * CONF;
                    const float f_ij = fk * m_j;
                    axpy<D>(gp, (p_i + p_j) / (rho_i * rho_j) * f_ij, r_ij);
                    if (lap == AQ_LAP_MONAGHAN) {
                        const float r2 = (q * q + 0.01f) * H * H;
                        axpy<D>(lu,
                                f_ij * CLEARY * udr / (r2 * rho_i * rho_j),
                                r_ij);
                    } else {
                        axpy<D>(lu, f_ij * 2.f / (rho_i * rho_j), u_ij);
                    }
                    dv += udr * f_ij * rho_i / rho_j;
                    if (LAPP) lp += (p_j - p_i) * (f_ij / rho_j);
                } else if (im_j == -3) {
                    nearw = true;
                    if (SHEP == 2) {
                        const float q = sqrtf(s) / H;
                        const f3 n_j = ldv<D>(normal, j);
                        const f3 t_j = ldv<D>(tangent, j);
                        const f3 b_j = ldv<D>(binormal, j);
                        const float r_n = dot<D>(r_ij, n_j);
                        const float area_j = m[j];
                        if ((r_n > -eps) && (r_n < eps)) {
                            if (!self_added) {
                                const float r_t = fabsf(dot<D>(r_ij, t_j));
                                const float r_b = fabsf(dot<D>(r_ij, b_j));
                                const float dr =
                                    0.55f * pow_cl(area_j, 1.f / (DIMS - 1.f));
                                if ((r_t <= dr) && (r_b <= dr)) {
                                    self_added = true;
                                    sh -= 0.5f;
                                }
                            }
                            return;
                        }
                        const float r_t = fabsf(dot<D>(r_ij, t_j));
                        const float r_b = fabsf(dot<D>(r_ij, b_j));
                        sh += r_n * CONW * kernelS_P<D>(q) * area_j +
                              kernelS_D<D>(fabsf(r_n), r_t, r_b, area_j);
                    }
                }
            });
    stxyz<D>(grad_p, i, gp);
    stxyz<D>(lap_u, i, lu);
    div_u[i] = dv;
    if (LAPP) lap_p[i] = lp;
    if (SHEP) shepard[i] = sh;
    if (near) near[i] = nearw ? 1 : 0;
}

// basic/neighs.cl:52-91
template <int D>
__global__ void __launch_bounds__(BS)
    neighs_k(const int* __restrict__ imove, uint32_t* __restrict__ n_neighs,
             const uint32_t* limitp, LLArgs ll)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ll.N) return;
    AQ_STATUS_SKIP;
    if (imove[i] <= -255) {
        n_neighs[i] = 0;
        return;
    }
    const uint32_t limit = *limitp;
    uint32_t n = 0;
    bool stop = false;
    walk<D>(ll.icell[i], ll.N, ll.n_cells[0], ll.n_cells[1], ll.icell,
            ll.ihoc, [&](uint32_t) {
                if (stop) return;
                n += 1;
                if (n >= limit) stop = true;
            });
    n_neighs[i] = n;
}

// ===========================================================================
// Launchers (also used by the runtime tools)
// ===========================================================================
#define AQ_GRID(N) dim3(blocks_for((N), BS)), dim3(BS), 0, s
#define AQ_DISPATCH(dims, KER, ...)                                            \
    do {                                                                       \
        if ((dims) == 2)                                                       \
            hipLaunchKernelGGL(KER<2>, __VA_ARGS__);                           \
        else                                                                   \
            hipLaunchKernelGGL(KER<3>, __VA_ARGS__);                           \
    } while (0)

static LLArgs mkll(const uint32_t* icell, const uint32_t* ihoc,
                   const uint32_t* n_cells, const uint32_t* status, uint32_t N)
{
    LLArgs a;
    a.icell = icell;
    a.ihoc = ihoc;
    a.n_cells = n_cells;
    a.status = status;
    a.N = N;
    return a;
}

// Link-list arguments of a per-particle walk kernel, with the cell list of
// `pw` when there is one.
static LLArgs mkll_pw(const uint32_t* icell, const uint32_t* ihoc,
                      const uint32_t* n_cells, const uint32_t* status,
                      uint32_t N, const PartWalk* pw)
{
    LLArgs a = mkll(icell, ihoc, n_cells, status, N);
    if (pw && pw->cell_end) {
        a.cell_end = pw->cell_end;
        a.flagged = pw->flagged;
    }
    return a;
}

// The particles a per-particle kernel runs on: all N, or (with the scratch
// of `pw`) the compacted indices with lo <= imove <= hi (and near != 0).
static Sel select_parts(const PartWalk* pw, const int32_t* imove,
                        const uint8_t* near, int lo, int hi, uint32_t N,
                        hipStream_t s)
{
    if (!pw || !pw->list || !pw->count) return Sel{ nullptr, nullptr };
    hipLaunchKernelGGL(set_zero_k, dim3(1), dim3(1), 0, s, pw->count);
    hipLaunchKernelGGL(select_k, AQ_GRID(N), imove, near, lo, hi, N,
                       pw->list, pw->count);
    return Sel{ pw->list, pw->count };
}

// One thread per particle, or a resident grid striding over a list.
static dim3 sel_grid(const void* ker, const Sel& sel, uint32_t N)
{
    return sel.list ? dim3(resident_blocks(ker, BS)) : dim3(blocks_for(N, BS));
}

void launch_predictor(int dims, const int32_t* imove, const float* r,
                      const float* u, const float* dudt, const float* rho,
                      const float* drhodt, float* r_in, float* u_in,
                      float* dudt_in, float* rho_in, float* drhodt_in,
                      uint32_t N, const float* dt, const uint32_t* status,
                      hipStream_t s)
{
    AQ_DISPATCH(dims, predictor_k, AQ_GRID(N), imove, r, u, dudt, rho, drhodt,
                r_in, u_in, dudt_in, rho_in, drhodt_in, N, dt, status);
}
void launch_corrector(int dims, const int32_t* imove, float* r, float* u,
                      const float* dudt, float* rho, const float* drhodt,
                      const float* dudt_in, const float* drhodt_in,
                      uint32_t N, const float* dt, const uint32_t* status,
                      hipStream_t s)
{
    AQ_DISPATCH(dims, corrector_k, AQ_GRID(N), imove, r, u, dudt, rho, drhodt,
                dudt_in, drhodt_in, N, dt, status);
}
void launch_euler_predictor(int dims, const float* r, const float* u,
                            const float* dudt, const float* rho,
                            const float* drhodt, float* r_in, float* u_in,
                            float* dudt_in, float* rho_in, float* drhodt_in,
                            uint32_t N, hipStream_t s)
{
    AQ_DISPATCH(dims, euler_predictor_k, AQ_GRID(N), r, u, dudt, rho, drhodt,
                r_in, u_in, dudt_in, rho_in, drhodt_in, N);
}
void launch_euler_corrector(int dims, const int32_t* imove, float* r,
                            float* u, const float* dudt, float* rho,
                            const float* drhodt, uint32_t N, const float* dt,
                            hipStream_t s)
{
    AQ_DISPATCH(dims, euler_corrector_k, AQ_GRID(N), imove, r, u, dudt, rho,
                drhodt, N, dt);
}
void launch_step_head(int dims, int32_t* imove, const float* r,
                      const float* u, const float* dudt, const float* rho,
                      const float* drhodt, float* r_in, float* u_in,
                      float* dudt_in, float* rho_in, float* drhodt_in,
                      float* m, uint32_t N, const float* dt,
                      const uint32_t* status, const float* dmin,
                      const float* dmax, hipStream_t s)
{
    AQ_DISPATCH(dims, step_head_k, AQ_GRID(N), imove, r, u, dudt, rho, drhodt,
                r_in, u_in, dudt_in, rho_in, drhodt_in, m, N, dt, status,
                dmin, dmax);
}
void launch_sort_chain(int dims, const uint32_t* perm, uint32_t N,
                       const uint32_t* id_in, uint32_t* id,
                       const uint32_t* iset_in, uint32_t* iset,
                       const int* imove_in, int* imove, const float* r_in,
                       float* r, const float* normal_in, float* normal,
                       const float* tangent_in, float* tangent,
                       const float* rho_in, float* rho, const float* m_in,
                       float* m, const float* u_in, float* u,
                       const float* dudt, float* dudt_in,
                       const float* drhodt, float* drhodt_in, float* p,
                       const float* refd, const float* cs, const float* p0,
                       float* binormal, hipStream_t s)
{
    AQ_DISPATCH(dims, sort_chain_k, AQ_GRID(N), perm, N, id_in, id, iset_in,
                iset, imove_in, imove, r_in, r, normal_in, normal, tangent_in,
                tangent, rho_in, rho, m_in, m, u_in, u, dudt, dudt_in, drhodt,
                drhodt_in, p, refd, cs, p0, binormal);
}
void launch_domain(int dims, int32_t* imove, float* r_in, float* u_in,
                   float* dudt_in, float* m, uint32_t N, const float* dmin,
                   const float* dmax, hipStream_t s)
{
    AQ_DISPATCH(dims, domain_k, AQ_GRID(N), imove, r_in, u_in, dudt_in, m, N,
                dmin, dmax);
}
void launch_eos(const uint32_t* iset, const int32_t* imove, const float* rho,
                float* p, const float* refd, uint32_t N, const float* cs,
                const float* p0, hipStream_t s)
{
    hipLaunchKernelGGL(eos_k, AQ_GRID(N), iset, imove, rho, p, refd, N, cs,
                       p0);
}
void launch_tait_eos(const uint32_t* iset, const int32_t* imove,
                     const float* rho, float* p, const float* refd, uint32_t N,
                     const float* cs, const float* p0, const float* gamma,
                     hipStream_t s)
{
    hipLaunchKernelGGL(tait_eos_k, AQ_GRID(N), iset, imove, rho, p, refd, N,
                       cs, p0, gamma);
}
void launch_bi_interp_shepard_tait(const uint32_t* iset, const int32_t* imove,
                                   const float* shepard, float* rho, float* p,
                                   const float* refd, uint32_t N,
                                   const float* cs, const float* p0,
                                   const float* gamma, hipStream_t s)
{
    hipLaunchKernelGGL(bi_interp_shepard_tait_k, AQ_GRID(N), iset, imove,
                       shepard, rho, p, refd, N, cs, p0, gamma);
}
void launch_binormal(int dims, const float* normal, float* tangent,
                     float* binormal, uint32_t N, hipStream_t s)
{
    AQ_DISPATCH(dims, binormal_k, AQ_GRID(N), normal, tangent, binormal, N);
}
void launch_rates(int dims, const uint32_t* iset, const int32_t* imove,
                  const float* grad_p, const float* lap_u, const float* div_u,
                  float* dudt, float* drhodt, const float* visc_dyn,
                  uint32_t N, const float* g, hipStream_t s)
{
    AQ_DISPATCH(dims, rates_k, AQ_GRID(N), iset, imove, grad_p, lap_u, div_u,
                dudt, drhodt, visc_dyn, N, g);
}
void launch_mp_predictor(int dims, const float* r, const float* u,
                         const float* dudt, const float* rho,
                         const float* drhodt, float* r_in, float* u_in,
                         float* dudt_in, float* rho_in, float* drhodt_in,
                         uint32_t N, hipStream_t s)
{
    AQ_DISPATCH(dims, mp_predictor_k, AQ_GRID(N), r, u, dudt, rho, drhodt,
                r_in, u_in, dudt_in, rho_in, drhodt_in, N);
}
void launch_mp_advance(int dims, const int32_t* imove, const float* u_in,
                       float* u, const float* dudt, const float* rho_in,
                       float* rho, const float* drhodt, uint32_t N,
                       const float* dt, hipStream_t s)
{
    AQ_DISPATCH(dims, mp_advance_k, AQ_GRID(N), imove, u_in, u, dudt, rho_in,
                rho, drhodt, N, dt);
}
void launch_mp_advance_r(int dims, const int32_t* imove, const float* r_in,
                         float* r, const float* u, uint32_t N,
                         const float* dt, hipStream_t s)
{
    AQ_DISPATCH(dims, mp_advance_r_k, AQ_GRID(N), imove, r_in, r, u, N, dt);
}
void launch_mp_relax(int dims, const int32_t* imove, const float* dudt_in,
                     float* dudt, const float* drhodt_in, float* drhodt,
                     uint32_t N, const float* f, hipStream_t s)
{
    AQ_DISPATCH(dims, mp_relax_k, AQ_GRID(N), imove, dudt_in, dudt, drhodt_in,
                drhodt, N, f);
}
void launch_mp_residuals(int dims, const int32_t* imove, const float* m,
                         const float* u, const float* dudt_in,
                         const float* dudt, const float* rho, const float* p,
                         const float* drhodt_in, const float* drhodt,
                         float* res, uint32_t N, hipStream_t s)
{
    AQ_DISPATCH(dims, mp_residuals_k, AQ_GRID(N), imove, m, u, dudt_in, dudt,
                rho, p, drhodt_in, drhodt, res, N);
}
void launch_mp_corrector(int dims, const int32_t* imove, const float* r_in,
                         float* r, const float* u_in, float* u,
                         const float* dudt, const float* rho_in, float* rho,
                         const float* drhodt, uint32_t N, const float* dt,
                         hipStream_t s)
{
    AQ_DISPATCH(dims, mp_corrector_k, AQ_GRID(N), imove, r_in, r, u_in, u,
                dudt, rho_in, rho, drhodt, N, dt);
}
void launch_energy_kin(int dims, float* ek, const int32_t* imove,
                       const float* u, const float* m, uint32_t N,
                       hipStream_t s)
{
    AQ_DISPATCH(dims, energy_kin_k, AQ_GRID(N), ek, imove, u, m, N);
}
void launch_energy_power(int dims, float* dek, float* dep, float* dec,
                         const int32_t* imove, const float* u,
                         const float* rho, const float* m, const float* p,
                         const float* dudt, const float* drhodt, uint32_t N,
                         const float* g, hipStream_t s)
{
    AQ_DISPATCH(dims, energy_power_k, AQ_GRID(N), dek, dep, dec, imove, u, rho,
                m, p, dudt, drhodt, N, g);
}
void launch_energy_parts(int dims, float* ek, float* ep, float* ec,
                         const uint32_t* iset, const int32_t* imove,
                         const float* r, const float* u, const float* rho,
                         const float* m, const float* refd, uint32_t N,
                         const float* g, const float* cs, hipStream_t s)
{
    AQ_DISPATCH(dims, energy_parts_k, AQ_GRID(N), ek, ep, ec, iset, imove, r,
                u, rho, m, refd, N, g, cs);
}
void launch_delta_rate(const uint32_t* iset, const int32_t* imove,
                       const float* rho, const float* lap_p, float* drhodt,
                       const float* refd, const float* delta, uint32_t N,
                       const float* dt, hipStream_t s)
{
    hipLaunchKernelGGL(delta_rate_k, AQ_GRID(N), iset, imove, rho, lap_p,
                       drhodt, refd, delta, N, dt);
}
void launch_timestep(int dims, const int32_t* imove, const float* u,
                     float* dt_var, uint32_t N, const float* dt,
                     const float* dt_min, const float* courant,
                     const float* dt_Ma, const float* h, hipStream_t s)
{
    AQ_DISPATCH(dims, timestep_k, AQ_GRID(N), imove, u, dt_var, N, dt, dt_min,
                courant, dt_Ma, h);
}
void launch_sensors_renorm(int dims, const int32_t* imove,
                           const float* shepard, float* u, float* rho,
                           float* p, uint32_t N, hipStream_t s)
{
    AQ_DISPATCH(dims, sensors_renorm_k, AQ_GRID(N), imove, shepard, u, rho, p,
                N);
}
void launch_step_tail(int dims, bool apply, bool delta_sph,
                      const uint32_t* iset, const int32_t* imove,
                      const float* shepard, float* grad_p, float* lap_u,
                      float* div_u, float* dudt, float* drhodt,
                      const float* visc_dyn, const float* g,
                      const float* lap_p, const float* refd,
                      const float* delta, float* r, float* u, float* rho,
                      const float* dudt_in, const float* drhodt_in,
                      uint32_t N, const float* dt, const uint32_t* status,
                      hipStream_t s)
{
#define AQ_TAIL(DD, AP, DL)                                                    \
    hipLaunchKernelGGL((step_tail_k<DD, AP, DL>), AQ_GRID(N), iset, imove,    \
                       shepard, grad_p, lap_u, div_u, dudt, drhodt, visc_dyn,  \
                       g, lap_p, refd, delta, r, u, rho, dudt_in, drhodt_in,   \
                       N, dt, status)
    if (dims == 2) {
        if (apply && delta_sph) AQ_TAIL(2, true, true);
        else if (apply) AQ_TAIL(2, true, false);
        else if (delta_sph) AQ_TAIL(2, false, true);
        else AQ_TAIL(2, false, false);
    } else {
        if (apply && delta_sph) AQ_TAIL(3, true, true);
        else if (apply) AQ_TAIL(3, true, false);
        else if (delta_sph) AQ_TAIL(3, false, true);
        else AQ_TAIL(3, false, false);
    }
#undef AQ_TAIL
}
void launch_bi_shepard_apply(int dims, const int32_t* imove,
                             const float* shepard, float* grad_p,
                             float* lap_u, float* div_u, uint32_t N,
                             hipStream_t s)
{
    AQ_DISPATCH(dims, bi_shepard_apply_k, AQ_GRID(N), imove, shepard, grad_p,
                lap_u, div_u, N);
}
void launch_bi_gradp(int dims, const uint32_t* iset, const int32_t* imove,
                     const float* shepard, const float* lap_u,
                     const float* dudt, float* grad_p, const float* visc_dyn,
                     const float* refd, uint32_t N, const float* g,
                     hipStream_t s)
{
    AQ_DISPATCH(dims, bi_gradp_k, AQ_GRID(N), iset, imove, shepard, lap_u,
                dudt, grad_p, visc_dyn, refd, N, g);
}
void launch_bi_interp_shepard(const uint32_t* iset, const int32_t* imove,
                              const float* shepard, float* rho, float* p,
                              const float* refd, uint32_t N, const float* cs,
                              const float* p0, hipStream_t s)
{
    hipLaunchKernelGGL(bi_interp_shepard_k, AQ_GRID(N), iset, imove, shepard,
                       rho, p, refd, N, cs, p0);
}

void launch_shepard(int dims, const int32_t* imove, const float* r,
                    const float* rho, const float* m, float* shepard,
                    uint32_t N, const uint32_t* icell, const uint32_t* ihoc,
                    const uint32_t* n_cells, const uint32_t* status,
                    aq_defs d, int cfd, hipStream_t s, const PartWalk* pw)
{
    // with the fluid done elsewhere (cfd & 2) it runs on -3 <= imove <= 0
    const LLArgs ll = mkll_pw(icell, ihoc, n_cells, status, N, pw);
    const Sel sel = select_parts((cfd & 2) ? pw : nullptr, imove, nullptr,
                                 -3, 0, N, s);
    auto* ker = (dims == 2) ? shepard_k<2> : shepard_k<3>;
    hipLaunchKernelGGL(ker, sel_grid((const void*)ker, sel, N), dim3(BS), 0,
                       s, imove, r, rho, m, shepard, ll, d.H, d.CONW,
                       support_threshold(d.H), cfd, sel);
}
void launch_interactions(int dims, const int32_t* imove, const float* r,
                         const float* u, const float* rho, const float* m,
                         const float* p, float* grad_p, float* lap_u,
                         float* div_u, uint32_t N, const uint32_t* icell,
                         const uint32_t* ihoc, const uint32_t* n_cells,
                         const uint32_t* status, aq_defs d, hipStream_t s)
{
    AQ_DISPATCH(dims, interactions_k, AQ_GRID(N), imove, r, u, rho, m, p,
                grad_p, lap_u, div_u, mkll(icell, ihoc, n_cells, status, N),
                d.H, d.CONF, support_threshold(d.H), d.lap_formulation);
}
void launch_lapp(int dims, const int32_t* imove, const float* r,
                 const float* rho, const float* m, const float* p,
                 float* lap_p, uint32_t N, const uint32_t* icell,
                 const uint32_t* ihoc, const uint32_t* n_cells,
                 const uint32_t* status, aq_defs d, hipStream_t s)
{
    AQ_DISPATCH(dims, lapp_k, AQ_GRID(N), imove, r, rho, m, p, lap_p,
                mkll(icell, ihoc, n_cells, status, N), d.H, d.CONF,
                support_threshold(d.H));
}
void launch_sensors(int dims, const int32_t* imove, const float* r,
                    const float* m, float* u, float* rho, float* p,
                    uint32_t N, const uint32_t* icell, const uint32_t* ihoc,
                    const uint32_t* n_cells, const uint32_t* status,
                    aq_defs d, hipStream_t s, const float* per)
{
    AQ_DISPATCH(dims, sensors_k, AQ_GRID(N), imove, r, m, u, rho, p,
                mkll(icell, ihoc, n_cells, status, N), d.H, d.CONW,
                support_threshold(d.H), per);
}
void launch_bi_shepard_compute(int dims, const int32_t* imove, const float* r,
                               const float* normal, const float* tangent,
                               const float* binormal, const float* m,
                               float* shepard, uint32_t N,
                               const uint32_t* icell, const uint32_t* ihoc,
                               const uint32_t* n_cells,
                               const uint32_t* status, aq_defs d,
                               hipStream_t s, int skip_fluid,
                               const PartWalk* pw)
{
    const LLArgs ll = mkll_pw(icell, ihoc, n_cells, status, N, pw);
    const Sel sel = select_parts(pw, imove, nullptr, -3, skip_fluid ? 0 : 1,
                                 N, s);
    auto* ker =
        (dims == 2) ? bi_shepard_compute_k<2> : bi_shepard_compute_k<3>;
    hipLaunchKernelGGL(ker, sel_grid((const void*)ker, sel, N), dim3(BS), 0,
                       s, imove, r, normal, tangent, binormal, m, shepard, ll,
                       d.H, d.CONW, d.DIMS, support_threshold(d.H),
                       skip_fluid, sel);
}
void launch_bi_lapu(int dims, const int32_t* imove, const float* r,
                    const float* u, const float* rho, const float* m,
                    float* lap_u, uint32_t N, const uint32_t* icell,
                    const uint32_t* ihoc, const uint32_t* n_cells,
                    const uint32_t* status, aq_defs d, hipStream_t s,
                    const PartWalk* pw)
{
    const LLArgs ll = mkll_pw(icell, ihoc, n_cells, status, N, pw);
    const Sel sel = select_parts(pw, imove, nullptr, -3, -3, N, s);
    auto* ker = (dims == 2) ? bi_lapu_k<2> : bi_lapu_k<3>;
    hipLaunchKernelGGL(ker, sel_grid((const void*)ker, sel, N), dim3(BS), 0,
                       s, imove, r, u, rho, m, lap_u, ll, d.H, d.CONF,
                       support_threshold(d.H), d.lap_formulation, sel);
}
void launch_bi_interp(int dims, const int32_t* imove, const float* r,
                      const float* m, const float* rho, const float* grad_p,
                      float* p, uint32_t N, const uint32_t* icell,
                      const uint32_t* ihoc, const uint32_t* n_cells,
                      const uint32_t* status, aq_defs d, hipStream_t s,
                      const PartWalk* pw)
{
    const LLArgs ll = mkll_pw(icell, ihoc, n_cells, status, N, pw);
    const Sel sel = select_parts(pw, imove, nullptr, -3, -3, N, s);
    auto* ker = (dims == 2) ? bi_interp_k<2> : bi_interp_k<3>;
    hipLaunchKernelGGL(ker, sel_grid((const void*)ker, sel, N), dim3(BS), 0,
                       s, imove, r, m, rho, grad_p, p, ll, d.H, d.CONW,
                       support_threshold(d.H), sel);
}
void launch_bi_interactions(int dims, const int32_t* imove, const float* r,
                            const float* normal, const float* u,
                            const float* rho, const float* m, const float* p,
                            float* grad_p, float* div_u, uint32_t N,
                            const uint32_t* icell, const uint32_t* ihoc,
                            const uint32_t* n_cells, const uint32_t* status,
                            aq_defs d, hipStream_t s, const uint8_t* near,
                            const PartWalk* pw)
{
    const LLArgs ll = mkll_pw(icell, ihoc, n_cells, status, N, pw);
    const Sel sel = select_parts(pw, imove, near, 1, 1, N, s);
    auto* ker = (dims == 2) ? bi_interactions_k<2> : bi_interactions_k<3>;
    hipLaunchKernelGGL(ker, sel_grid((const void*)ker, sel, N), dim3(BS), 0,
                       s, imove, r, normal, u, rho, m, p, grad_p, div_u, ll,
                       d.H, d.CONW, support_threshold(d.H), near, sel);
}

void launch_fluid_pass(int dims, int shep, bool lapp, const int32_t* imove,
                       const float* r, const float* u, const float* rho,
                       const float* m, const float* p, const float* normal,
                       const float* tangent, const float* binormal,
                       float* shepard, float* grad_p, float* lap_u,
                       float* div_u, float* lap_p, uint8_t* near, uint32_t N,
                       const uint32_t* icell, const uint32_t* ihoc,
                       const uint32_t* n_cells, const uint32_t* status,
                       aq_defs d, hipStream_t s)
{
    const LLArgs ll = mkll(icell, ihoc, n_cells, status, N);
    const float smax = support_threshold(d.H);
#define AQ_FP(DD, SH, LP)                                                      \
    hipLaunchKernelGGL((fluid_pass_k<DD, SH, LP>), AQ_GRID(N), imove, r, u,    \
                       rho, m, p, normal, tangent, binormal, shepard, grad_p,  \
                       lap_u, div_u, lap_p, near, ll, d.H, d.CONW, d.CONF,     \
                       d.DIMS, smax, d.lap_formulation)
#define AQ_FP_D(SH, LP)                                                        \
    do {                                                                       \
        if (dims == 2) AQ_FP(2, SH, LP);                                       \
        else AQ_FP(3, SH, LP);                                                 \
    } while (0)
    if (shep == 0 && !lapp) AQ_FP_D(0, false);
    else if (shep == 0 && lapp) AQ_FP_D(0, true);
    else if (shep == 1 && !lapp) AQ_FP_D(1, false);
    else if (shep == 1 && lapp) AQ_FP_D(1, true);
    else if (shep == 2 && !lapp) AQ_FP_D(2, false);
    else AQ_FP_D(2, true);
#undef AQ_FP_D
#undef AQ_FP
}
void launch_neighs(int dims, const int32_t* imove, uint32_t* n_neighs,
                   const uint32_t* limit, uint32_t N, const uint32_t* icell,
                   const uint32_t* ihoc, const uint32_t* n_cells,
                   const uint32_t* status, hipStream_t s)
{
    AQ_DISPATCH(dims, neighs_k, AQ_GRID(N), imove, n_neighs, limit,
                mkll(icell, ihoc, n_cells, status, N));
}

}  // namespace aq

// ===========================================================================
// C ABI
// ===========================================================================
using namespace aq;

#define AQ_PTRS(...) AQ_REQUIRE(all_set({ __VA_ARGS__ }), "null pointer argument")
static bool all_set(std::initializer_list<const void*> l)
{
    for (auto p : l)
        if (!p) return false;
    return true;
}
static void check_defs(const aq_defs& d)
{
    AQ_REQUIRE(d.H > 0.f && std::isfinite(d.H), "defs.H must be > 0");
    AQ_REQUIRE(d.lap_formulation == AQ_LAP_MONAGHAN ||
                   d.lap_formulation == AQ_LAP_MORRIS,
               "defs.lap_formulation must be AQ_LAP_MONAGHAN or AQ_LAP_MORRIS");
}

extern "C" {

AQ_API int aq_predictor(int dims, const int32_t* imove, const float* r,
                        const float* u, const float* dudt, const float* rho,
                        const float* drhodt, float* r_in, float* u_in,
                        float* dudt_in, float* rho_in, float* drhodt_in,
                        uint32_t N, const float* dt, void* stream)
{
    return guard([&] {
        check_dims(dims);
        if (!N) return;
        AQ_PTRS(imove, r, u, dudt, rho, drhodt, r_in, u_in, dudt_in, rho_in,
                drhodt_in, dt);
        launch_predictor(dims, imove, r, u, dudt, rho, drhodt, r_in, u_in,
                         dudt_in, rho_in, drhodt_in, N, dt, nullptr,
                         S(stream));
    });
}

AQ_API int aq_corrector(int dims, const int32_t* imove, float* r, float* u,
                        const float* dudt, float* rho, const float* drhodt,
                        const float* dudt_in, const float* drhodt_in,
                        uint32_t N, const float* dt, void* stream)
{
    return guard([&] {
        check_dims(dims);
        if (!N) return;
        AQ_PTRS(imove, r, u, dudt, rho, drhodt, dudt_in, drhodt_in, dt);
        launch_corrector(dims, imove, r, u, dudt, rho, drhodt, dudt_in,
                         drhodt_in, N, dt, nullptr, S(stream));
    });
}

AQ_API int aq_domain(int dims, int32_t* imove, float* r_in, float* u_in,
                     float* dudt_in, float* m, uint32_t N,
                     const float* domain_min, const float* domain_max,
                     void* stream)
{
    return guard([&] {
        check_dims(dims);
        if (!N) return;
        AQ_PTRS(imove, r_in, u_in, dudt_in, m, domain_min, domain_max);
        launch_domain(dims, imove, r_in, u_in, dudt_in, m, N, domain_min,
                      domain_max, S(stream));
    });
}

AQ_API int aq_eos(const uint32_t* iset, const int32_t* imove,
                  const float* rho, float* p, const float* refd, uint32_t N,
                  const float* cs, const float* p0, void* stream)
{
    return guard([&] {
        if (!N) return;
        AQ_PTRS(iset, imove, rho, p, refd, cs, p0);
        launch_eos(iset, imove, rho, p, refd, N, cs, p0, S(stream));
    });
}

AQ_API int aq_binormal(int dims, const float* normal, float* tangent,
                       float* binormal, uint32_t N, void* stream)
{
    return guard([&] {
        check_dims(dims);
        if (!N) return;
        AQ_PTRS(normal, tangent, binormal);
        launch_binormal(dims, normal, tangent, binormal, N, S(stream));
    });
}

AQ_API int aq_rates(int dims, const uint32_t* iset, const int32_t* imove,
                    const float* rho, const float* grad_p, const float* lap_u,
                    const float* div_u, float* dudt, float* drhodt,
                    const float* visc_dyn, uint32_t N, const float* g,
                    void* stream)
{
    (void)rho;
    return guard([&] {
        check_dims(dims);
        if (!N) return;
        AQ_PTRS(iset, imove, grad_p, lap_u, div_u, dudt, drhodt, visc_dyn, g);
        launch_rates(dims, iset, imove, grad_p, lap_u, div_u, dudt, drhodt,
                     visc_dyn, N, g, S(stream));
    });
}

AQ_API int aq_energy_kin(int dims, float* energy_kin, const int32_t* imove,
                         const float* u, const float* m, uint32_t N,
                         void* stream)
{
    return guard([&] {
        check_dims(dims);
        if (!N) return;
        AQ_PTRS(energy_kin, imove, u, m);
        launch_energy_kin(dims, energy_kin, imove, u, m, N, S(stream));
    });
}

AQ_API int aq_energy_power(int dims, float* energy_dekdt, float* energy_depdt,
                           float* energy_decdt, const int32_t* imove,
                           const float* u, const float* rho, const float* m,
                           const float* p, const float* dudt,
                           const float* drhodt, uint32_t N, const float* g,
                           void* stream)
{
    return guard([&] {
        check_dims(dims);
        if (!N) return;
        AQ_PTRS(energy_dekdt, energy_depdt, energy_decdt, imove, u, rho, m, p,
                dudt, drhodt, g);
        launch_energy_power(dims, energy_dekdt, energy_depdt, energy_decdt,
                            imove, u, rho, m, p, dudt, drhodt, N, g,
                            S(stream));
    });
}

AQ_API int aq_energy(int dims, float* energy_ek, float* energy_ep,
                     float* energy_ec, const uint32_t* iset,
                     const int32_t* imove, const float* r, const float* u,
                     const float* rho, const float* m, const float* refd,
                     uint32_t N, const float* g, const float* cs,
                     void* stream)
{
    return guard([&] {
        check_dims(dims);
        if (!N) return;
        AQ_PTRS(energy_ek, energy_ep, energy_ec, iset, imove, r, u, rho, m,
                refd, g, cs);
        launch_energy_parts(dims, energy_ek, energy_ep, energy_ec, iset,
                            imove, r, u, rho, m, refd, N, g, cs, S(stream));
    });
}

AQ_API int aq_delta_rate(const uint32_t* iset, const int32_t* imove,
                         const float* rho, const float* lap_p, float* drhodt,
                         const float* refd, const float* delta, uint32_t N,
                         const float* dt, void* stream)
{
    return guard([&] {
        if (!N) return;
        AQ_PTRS(iset, imove, rho, lap_p, drhodt, refd, delta, dt);
        launch_delta_rate(iset, imove, rho, lap_p, drhodt, refd, delta, N, dt,
                          S(stream));
    });
}

AQ_API int aq_timestep(int dims, const int32_t* imove, const float* u,
                       float* dt_var, uint32_t N, const float* dt,
                       const float* dt_min, const float* courant,
                       const float* dt_Ma, const float* h, void* stream)
{
    return guard([&] {
        check_dims(dims);
        if (!N) return;
        AQ_PTRS(imove, u, dt_var, dt, dt_min, courant, dt_Ma, h);
        launch_timestep(dims, imove, u, dt_var, N, dt, dt_min, courant, dt_Ma,
                        h, S(stream));
    });
}

AQ_API int aq_sensors_renorm(int dims, const int32_t* imove,
                             const float* shepard, float* u, float* rho,
                             float* p, uint32_t N, void* stream)
{
    return guard([&] {
        check_dims(dims);
        if (!N) return;
        AQ_PTRS(imove, shepard, u, rho, p);
        launch_sensors_renorm(dims, imove, shepard, u, rho, p, N, S(stream));
    });
}

AQ_API int aq_bi_shepard_apply(int dims, const int32_t* imove,
                               const float* shepard, float* grad_p,
                               float* lap_u, float* div_u, uint32_t N,
                               void* stream)
{
    return guard([&] {
        check_dims(dims);
        if (!N) return;
        AQ_PTRS(imove, shepard, grad_p, lap_u, div_u);
        launch_bi_shepard_apply(dims, imove, shepard, grad_p, lap_u, div_u, N,
                                S(stream));
    });
}

AQ_API int aq_bi_gradp(int dims, const uint32_t* iset, const int32_t* imove,
                       const float* shepard, const float* rho,
                       const float* lap_u, const float* dudt, float* grad_p,
                       const float* visc_dyn, const float* refd, uint32_t N,
                       const float* g, void* stream)
{
    (void)rho;
    return guard([&] {
        check_dims(dims);
        if (!N) return;
        AQ_PTRS(iset, imove, shepard, lap_u, dudt, grad_p, visc_dyn, refd, g);
        launch_bi_gradp(dims, iset, imove, shepard, lap_u, dudt, grad_p,
                        visc_dyn, refd, N, g, S(stream));
    });
}

AQ_API int aq_bi_interp_shepard(const uint32_t* iset, const int32_t* imove,
                                const float* shepard, float* rho, float* p,
                                const float* refd, uint32_t N,
                                const float* cs, const float* p0,
                                void* stream)
{
    return guard([&] {
        if (!N) return;
        AQ_PTRS(iset, imove, shepard, rho, p, refd, cs, p0);
        launch_bi_interp_shepard(iset, imove, shepard, rho, p, refd, N, cs,
                                 p0, S(stream));
    });
}

AQ_API int aq_shepard(int dims, const int32_t* imove, const float* r,
                      const float* rho, const float* m, float* shepard,
                      uint32_t N, const uint32_t* icell, const uint32_t* ihoc,
                      const uint32_t* n_cells, aq_defs defs, int cfd_variant,
                      void* stream)
{
    return guard([&] {
        check_dims(dims);
        check_defs(defs);
        if (!N) return;
        AQ_PTRS(imove, r, rho, m, shepard, icell, ihoc, n_cells);
        launch_shepard(dims, imove, r, rho, m, shepard, N, icell, ihoc,
                       n_cells, nullptr, defs, cfd_variant ? 1 : 0,
                       S(stream), nullptr);
    });
}

AQ_API int aq_interactions(int dims, const int32_t* imove, const float* r,
                           const float* u, const float* rho, const float* m,
                           const float* p, float* grad_p, float* lap_u,
                           float* div_u, uint32_t N, const uint32_t* icell,
                           const uint32_t* ihoc, const uint32_t* n_cells,
                           aq_defs defs, void* stream)
{
    return guard([&] {
        check_dims(dims);
        check_defs(defs);
        if (!N) return;
        AQ_PTRS(imove, r, u, rho, m, p, grad_p, lap_u, div_u, icell, ihoc,
                n_cells);
        launch_interactions(dims, imove, r, u, rho, m, p, grad_p, lap_u, div_u,
                            N, icell, ihoc, n_cells, nullptr, defs,
                            S(stream));
    });
}

AQ_API int aq_lapp(int dims, const int32_t* imove, const float* r,
                   const float* rho, const float* m, const float* p,
                   float* lap_p, uint32_t N, const uint32_t* icell,
                   const uint32_t* ihoc, const uint32_t* n_cells,
                   aq_defs defs, void* stream)
{
    return guard([&] {
        check_dims(dims);
        check_defs(defs);
        if (!N) return;
        AQ_PTRS(imove, r, rho, m, p, lap_p, icell, ihoc, n_cells);
        launch_lapp(dims, imove, r, rho, m, p, lap_p, N, icell, ihoc, n_cells,
                    nullptr, defs, S(stream));
    });
}

AQ_API int aq_sensors(int dims, const int32_t* imove, const float* r,
                      const float* m, float* u, float* rho, float* p,
                      uint32_t N, const uint32_t* icell, const uint32_t* ihoc,
                      const uint32_t* n_cells, aq_defs defs, void* stream)
{
    return guard([&] {
        check_dims(dims);
        check_defs(defs);
        if (!N) return;
        AQ_PTRS(imove, r, m, u, rho, p, icell, ihoc, n_cells);
        launch_sensors(dims, imove, r, m, u, rho, p, N, icell, ihoc, n_cells,
                       nullptr, defs, S(stream), nullptr);
    });
}

AQ_API int aq_bi_shepard_compute(int dims, const int32_t* imove,
                                 const float* r, const float* normal,
                                 const float* tangent, const float* binormal,
                                 const float* m, float* shepard, uint32_t N,
                                 const uint32_t* icell, const uint32_t* ihoc,
                                 const uint32_t* n_cells, aq_defs defs,
                                 void* stream)
{
    return guard([&] {
        check_dims(dims);
        check_defs(defs);
        if (!N) return;
        AQ_PTRS(imove, r, normal, tangent, binormal, m, shepard, icell, ihoc,
                n_cells);
        launch_bi_shepard_compute(dims, imove, r, normal, tangent, binormal, m,
                                  shepard, N, icell, ihoc, n_cells, nullptr,
                                  defs, S(stream), 0, nullptr);
    });
}

AQ_API int aq_bi_lapu(int dims, const int32_t* imove, const float* r,
                      const float* u, const float* rho, const float* m,
                      float* lap_u, uint32_t N, const uint32_t* icell,
                      const uint32_t* ihoc, const uint32_t* n_cells,
                      aq_defs defs, void* stream)
{
    return guard([&] {
        check_dims(dims);
        check_defs(defs);
        if (!N) return;
        AQ_PTRS(imove, r, u, rho, m, lap_u, icell, ihoc, n_cells);
        launch_bi_lapu(dims, imove, r, u, rho, m, lap_u, N, icell, ihoc,
                       n_cells, nullptr, defs, S(stream), nullptr);
    });
}

AQ_API int aq_bi_interp(int dims, const int32_t* imove, const float* r,
                        const float* m, const float* rho,
                        const float* grad_p, float* p, uint32_t N,
                        const uint32_t* icell, const uint32_t* ihoc,
                        const uint32_t* n_cells, aq_defs defs, void* stream)
{
    return guard([&] {
        check_dims(dims);
        check_defs(defs);
        if (!N) return;
        AQ_PTRS(imove, r, m, rho, grad_p, p, icell, ihoc, n_cells);
        launch_bi_interp(dims, imove, r, m, rho, grad_p, p, N, icell, ihoc,
                         n_cells, nullptr, defs, S(stream), nullptr);
    });
}

AQ_API int aq_bi_interactions(int dims, const int32_t* imove, const float* r,
                              const float* normal, const float* u,
                              const float* rho, const float* m,
                              const float* p, float* grad_p, float* div_u,
                              uint32_t N, const uint32_t* icell,
                              const uint32_t* ihoc, const uint32_t* n_cells,
                              aq_defs defs, void* stream)
{
    return guard([&] {
        check_dims(dims);
        check_defs(defs);
        if (!N) return;
        AQ_PTRS(imove, r, normal, u, rho, m, p, grad_p, div_u, icell, ihoc,
                n_cells);
        launch_bi_interactions(dims, imove, r, normal, u, rho, m, p, grad_p,
                               div_u, N, icell, ihoc, n_cells, nullptr, defs,
                               S(stream), nullptr, nullptr);
    });
}

AQ_API int aq_neighs(int dims, const int32_t* imove, uint32_t* n_neighs,
                     const uint32_t* neighs_limit, uint32_t N,
                     const uint32_t* icell, const uint32_t* ihoc,
                     const uint32_t* n_cells, void* stream)
{
    return guard([&] {
        check_dims(dims);
        if (!N) return;
        AQ_PTRS(imove, n_neighs, neighs_limit, icell, ihoc, n_cells);
        launch_neighs(dims, imove, n_neighs, neighs_limit, N, icell, ihoc,
                      n_cells, nullptr, S(stream));
    });
}

}  // extern "C"
