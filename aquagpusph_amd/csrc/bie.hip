// Conservative boundary integrals (resources/Scripts/cfd/Boundary/BIe/*.cl)
// on gfx950: wall elements imove == -3 (area in m, unit outward normal).
// One thread per particle; the pair kernels walk the link-list cells in the
// reference order.  Candidates beyond the support are rejected with
// s > s_max, the exact form of the reference's q >= SUPPORT test.
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

template <int D> __device__ __forceinline__ float4 ld4(const float* a, uint32_t i)
{
    if constexpr (D == 2) {
        const float2 t = reinterpret_cast<const float2*>(a)[i];
        return make_float4(t.x, t.y, 0.f, 0.f);
    } else {
        return reinterpret_cast<const float4*>(a)[i];
    }
}

// BIe/Interactions.cl:entry
template <int D>
__global__ void __launch_bounds__(BS)
    bie_interactions_k(const int* __restrict__ imove,
                       const float* __restrict__ r,
                       const float* __restrict__ normal,
                       const float* __restrict__ u,
                       const float* __restrict__ m,
                       float* __restrict__ grad_w_bi,
                       float* __restrict__ div_u_bi, Cells c, float H,
                       float CONW, float smax)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    if (skip(c)) return;
    if (imove[i] != 1) return;
    const f3 r_i = ldv<D>(r, i);
    f3 gw = ldv<D>(grad_w_bi, i);
    float dv = div_u_bi[i];
    walk<D>(c.icell[i], c.N, c.n_cells[0], c.n_cells[1], c.icell, c.ihoc,
            [&](uint32_t j) {
                if (imove[j] != -3) return;
                const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
                const float s = dot<D>(r_ij, r_ij);
                if (s > smax) return;
                const float q = sqrtf(s) / H;
                const f3 n_j = ldv<D>(normal, j);
                const f3 u_j = ldv<D>(u, j);
                const float w = kernelW<D>(q) * CONW * m[j];
                f3 g;
                g.x = n_j.x * w;
                g.y = n_j.y * w;
                g.z = (D == 3) ? n_j.z * w : 0.f;
                gw.x += g.x;
                gw.y += g.y;
                if (D == 3) gw.z += g.z;
                dv -= dot<D>(u_j, g);
            });
    stxyz<D>(grad_w_bi, i, gw);
    div_u_bi[i] = dv;
}

// BIe/Interactions.cl:p_boundary
template <int D>
__global__ void __launch_bounds__(BS)
    bie_p_boundary_k(const int* __restrict__ imove,
                     const float* __restrict__ r,
                     const float* __restrict__ m,
                     const float* __restrict__ rho, float* __restrict__ p,
                     Cells c, float H, float CONW, float smax)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    if (skip(c)) return;
    if (imove[i] != -3) return;
    const f3 r_i = ldv<D>(r, i);
    float acc = 0.f;
    walk<D>(c.icell[i], c.N, c.n_cells[0], c.n_cells[1], c.icell, c.ihoc,
            [&](uint32_t j) {
                if (imove[j] != 1) return;
                const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
                const float s = dot<D>(r_ij, r_ij);
                if (s > smax) return;
                const float q = sqrtf(s) / H;
                acc += 2.f * p[j] * kernelW<D>(q) * CONW * m[j] / rho[j];
            });
    p[i] = acc;
}

// BIe/Rates.cl:entry (whole-vec operations, w included in 3D)
template <int D>
__global__ void bie_rates_k(const int* __restrict__ imove,
                            const float* __restrict__ rho,
                            const float* __restrict__ p,
                            const float* __restrict__ u,
                            const float* __restrict__ grad_w_bi,
                            const float* __restrict__ div_u_bi,
                            float* __restrict__ grad_p,
                            float* __restrict__ div_u, uint32_t N)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != 1) return;
    const float s = 2.f * p[i] / rho[i];
    const float4 gw = ld4<D>(grad_w_bi, i);
    const float4 ui = ld4<D>(u, i);
    float d = ui.x * gw.x;
    d = d + ui.y * gw.y;
    if (D == 3) {
        d = d + ui.z * gw.z;
        d = d + ui.w * gw.w;
        float4* g = reinterpret_cast<float4*>(grad_p) + i;
        float4 v = *g;
        v.x += s * gw.x;
        v.y += s * gw.y;
        v.z += s * gw.z;
        v.w += s * gw.w;
        *g = v;
    } else {
        float2* g = reinterpret_cast<float2*>(grad_p) + i;
        float2 v = *g;
        v.x += s * gw.x;
        v.y += s * gw.y;
        *g = v;
    }
    div_u[i] -= 2.f * rho[i] * (d + div_u_bi[i]);
}

// BIe/ElasticBounce.cl:entry
template <int D>
__global__ void __launch_bounds__(BS)
    bie_bounce_k(const int* __restrict__ imove, const float* __restrict__ r_in,
                 const float* __restrict__ normal, const float* __restrict__ m,
                 const float* __restrict__ u_in, float* __restrict__ dudt,
                 Cells c, const float* dtp, float DR_FACTOR,
                 float MIN_BOUND_DIST)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    if (skip(c)) return;
    if (imove[i] != 1) return;
    const float dt = *dtp;
    if (!(dt != 0.f)) return;
    const f3 r_i = ldv<D>(r_in, i);
    const f3 ui = ldv<D>(u_in, i);
    f3 a = ldv<D>(dudt, i), uu;
    uu.x = ui.x + 0.5f * dt * a.x;
    uu.y = ui.y + 0.5f * dt * a.y;
    uu.z = (D == 3) ? ui.z + 0.5f * dt * a.z : 0.f;
    walk<D>(c.icell[i], c.N, c.n_cells[0], c.n_cells[1], c.icell, c.ihoc,
            [&](uint32_t j) {
                if (imove[j] != -3) return;
                const f3 r_ij = sub<D>(ldv<D>(r_in, j), r_i);
                const f3 n_j = ldv<D>(normal, j);
                const float rn = dot<D>(r_ij, n_j);
                if (rn < 0.f) return;
                const float dr = (D == 3) ? sqrtf(m[j]) : m[j];
                const float Rr = DR_FACTOR * dr;
                f3 rt;
                rt.x = r_ij.x - rn * n_j.x;
                rt.y = r_ij.y - rn * n_j.y;
                rt.z = (D == 3) ? r_ij.z - rn * n_j.z : 0.f;
                if (dot<D>(rt, rt) >= Rr * Rr) return;
                const float drn = dt * dot<D>(uu, n_j);
                if (drn < 0.f) return;
                if (rn - drn <= MIN_BOUND_DIST * dr) {
                    f3 v;
                    v.x = ui.x + dt * a.x;
                    v.y = ui.y + dt * a.y;
                    v.z = (D == 3) ? ui.z + dt * a.z : 0.f;
                    const float vn = dot<D>(v, n_j);
                    a.x = (v.x - 2.f * vn * n_j.x - ui.x) / dt;
                    a.y = (v.y - 2.f * vn * n_j.y - ui.y) / dt;
                    if (D == 3) a.z = (v.z - 2.f * vn * n_j.z - ui.z) / dt;
                    uu.x = ui.x + 0.5f * dt * a.x;
                    uu.y = ui.y + 0.5f * dt * a.y;
                    if (D == 3) uu.z = ui.z + 0.5f * dt * a.z;
                }
            });
    stxyz<D>(dudt, i, a);
}

// cfd/Boundary/ElasticBounce.cl:entry: the plain elastic bounce of the
// fluid on ghost-particle (-2) and boundary-integral (-3) wall elements.  A
// fluid particle within R = DR_FACTOR dr of an element (tangentially), on
// the fluid side of it, that would come closer than MIN_BOUND_DIST dr in
// the next dt loses (1 + ELASTIC_FACTOR) times its normal relative velocity
// and acceleration.  u and dudt are updated in place; the next elements of
// the walk see the corrected values (the reference's order).
template <int D>
__global__ void __launch_bounds__(BS)
    elastic_bounce_k(const int* __restrict__ imove, const float* __restrict__ r,
                     const float* __restrict__ normal, float* __restrict__ u,
                     float* __restrict__ dudt, Cells c, const float* drp,
                     const float* dtp, float DR_FACTOR, float MIN_BOUND_DIST,
                     float ELASTIC_FACTOR)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    if (skip(c)) return;
    if (imove[i] != 1) return;
    const float dt = *dtp, dr = *drp;
    const float Rr = DR_FACTOR * dr;
    const float e1 = 1.f + ELASTIC_FACTOR;
    const f3 r_i = ldv<D>(r, i);
    f3 ui = ldv<D>(u, i);
    f3 ai = ldv<D>(dudt, i);
    bool hit = false;
    walk<D>(c.icell[i], c.N, c.n_cells[0], c.n_cells[1], c.icell, c.ihoc,
            [&](uint32_t j) {
                const int mj = imove[j];
                if ((mj != -2) && (mj != -3)) return;
                const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
                const f3 n_j = ldv<D>(normal, j);
                const float r0 = dot<D>(r_ij, n_j);
                if (r0 < 0.f) return;  // behind the element
                f3 rt;
                rt.x = r_ij.x - r0 * n_j.x;
                rt.y = r_ij.y - r0 * n_j.y;
                rt.z = (D == 3) ? r_ij.z - r0 * n_j.z : 0.f;
                if (dot<D>(rt, rt) >= Rr * Rr) return;  // passing too far
                const float u_n = dot<D>(sub<D>(ui, ldv<D>(u, j)), n_j);
                const float a_n = dot<D>(sub<D>(ai, ldv<D>(dudt, j)), n_j);
                const float dist = dt * u_n + 0.5f * dt * dt * a_n;
                if (dist < 0.f) return;  // already moving away
                if (r0 - dist <= MIN_BOUND_DIST * dr) {
                    const float sa = e1 * a_n, su = e1 * u_n;
                    ai.x = ai.x - sa * n_j.x;
                    ai.y = ai.y - sa * n_j.y;
                    ui.x = ui.x - su * n_j.x;
                    ui.y = ui.y - su * n_j.y;
                    if (D == 3) {
                        ai.z = ai.z - sa * n_j.z;
                        ui.z = ui.z - su * n_j.z;
                    }
                    hit = true;
                }
            });
    if (hit) {
        stxyz<D>(dudt, i, ai);
        stxyz<D>(u, i, ui);
    }
}

// BIe/PST.cl:entry (r of the fluid particle is moved in place)
template <int D>
__global__ void __launch_bounds__(BS)
    bie_pst_k(const int* __restrict__ imove, float* __restrict__ r,
              const float* __restrict__ normal, const float* __restrict__ m,
              const float* __restrict__ rho, Cells c, float DIMS,
              float DR_FACTOR)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    if (skip(c)) return;
    if (imove[i] != 1) return;
    const float Ri = 0.5f * powf(m[i] / rho[i], 1.f / DIMS);
    f3 ri = ldv<D>(r, i);
    walk<D>(c.icell[i], c.N, c.n_cells[0], c.n_cells[1], c.icell, c.ihoc,
            [&](uint32_t j) {
                if (imove[j] != -3) return;
                const f3 r_ij = sub<D>(ldv<D>(r, j), ri);
                const f3 n_j = ldv<D>(normal, j);
                const float rn = dot<D>(r_ij, n_j);
                if (fabsf(rn) > Ri) return;
                const float dr = (D == 3) ? sqrtf(m[j]) : m[j];
                const float Rj = DR_FACTOR * dr;
                f3 rt;
                rt.x = r_ij.x - rn * n_j.x;
                rt.y = r_ij.y - rn * n_j.y;
                rt.z = (D == 3) ? r_ij.z - rn * n_j.z : 0.f;
                if (dot<D>(rt, rt) >= Rj * Rj) return;
                ri.x += (rn - Ri) * n_j.x;
                ri.y += (rn - Ri) * n_j.y;
                if (D == 3) ri.z += (rn - Ri) * n_j.z;
            });
    stxyz<D>(r, i, ri);
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

void launch_bie_interactions(int dims, const int32_t* imove, const float* r,
                             const float* normal, const float* u,
                             const float* m, float* grad_w_bi, float* div_u_bi,
                             uint32_t N, const uint32_t* icell,
                             const uint32_t* ihoc, const uint32_t* n_cells,
                             const uint32_t* status, aq_defs d, hipStream_t s)
{
    Cells c{ icell, ihoc, n_cells, status, N };
    AQ_D(dims, bie_interactions_k, AQ_G(N), imove, r, normal, u, m, grad_w_bi,
         div_u_bi, c, d.H, d.CONW, support_threshold(d.H));
}

void launch_bie_p_boundary(int dims, const int32_t* imove, const float* r,
                           const float* m, const float* rho, float* p,
                           uint32_t N, const uint32_t* icell,
                           const uint32_t* ihoc, const uint32_t* n_cells,
                           const uint32_t* status, aq_defs d, hipStream_t s)
{
    Cells c{ icell, ihoc, n_cells, status, N };
    AQ_D(dims, bie_p_boundary_k, AQ_G(N), imove, r, m, rho, p, c, d.H, d.CONW,
         support_threshold(d.H));
}

void launch_bie_rates(int dims, const int32_t* imove, const float* rho,
                      const float* p, const float* u, const float* grad_w_bi,
                      const float* div_u_bi, float* grad_p, float* div_u,
                      uint32_t N, hipStream_t s)
{
    AQ_D(dims, bie_rates_k, AQ_G(N), imove, rho, p, u, grad_w_bi, div_u_bi,
         grad_p, div_u, N);
}

void launch_bie_bounce(int dims, const int32_t* imove, const float* r_in,
                       const float* normal, const float* m, const float* u_in,
                       float* dudt, uint32_t N, const float* dt,
                       const uint32_t* icell, const uint32_t* ihoc,
                       const uint32_t* n_cells, const uint32_t* status,
                       float dr_factor, float min_bound_dist, hipStream_t s)
{
    Cells c{ icell, ihoc, n_cells, status, N };
    AQ_D(dims, bie_bounce_k, AQ_G(N), imove, r_in, normal, m, u_in, dudt, c, dt,
         dr_factor, min_bound_dist);
}

void launch_elastic_bounce(int dims, const int32_t* imove, const float* r,
                           const float* normal, float* u, float* dudt,
                           uint32_t N, const float* dr, const float* dt,
                           const uint32_t* icell, const uint32_t* ihoc,
                           const uint32_t* n_cells, const uint32_t* status,
                           float dr_factor, float min_bound_dist,
                           float elastic_factor, hipStream_t s)
{
    Cells c{ icell, ihoc, n_cells, status, N };
    AQ_D(dims, elastic_bounce_k, AQ_G(N), imove, r, normal, u, dudt, c, dr, dt,
         dr_factor, min_bound_dist, elastic_factor);
}

void launch_bie_pst(int dims, const int32_t* imove, float* r,
                    const float* normal, const float* m, const float* rho,
                    uint32_t N, const uint32_t* icell, const uint32_t* ihoc,
                    const uint32_t* n_cells, const uint32_t* status, aq_defs d,
                    float dr_factor, hipStream_t s)
{
    Cells c{ icell, ihoc, n_cells, status, N };
    AQ_D(dims, bie_pst_k, AQ_G(N), imove, r, normal, m, rho, c, d.DIMS,
         dr_factor);
}

#undef AQ_D
#undef AQ_G

}  // namespace aq
