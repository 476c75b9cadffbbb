// Plane portals on gfx950 (resources/Presets/cfd/portal.xml,
// resources/Scripts/cfd/Boundary/Portal/*.cl): periodic and inflow-outflow
// boundaries made of an out plane and an in plane.  Particles within
// SUPPORT h of the out plane are moved by in_r - out_r and re-hashed into the
// cell of their new place ("mirrored"); the mirrored fluid particles then
// add the terms of the unmirrored particles around the in plane, walking
// the cells of their new place with the link-list walk (a cell run ends at
// a particle whose icell differs, mirrored ones included, as in the
// reference; cells outside the grid are skipped); finally they are moved
// back.  Particles past the out plane
// are teleported to the in plane.  One thread per particle: only the
// particles next to a plane do any work.
#include "aq_device.hpp"
#include "aq_internal.hpp"

namespace aq {

namespace {

constexpr int BS = 256;

struct PCells {
    const uint32_t* icell;
    const uint32_t* ihoc;
    const uint32_t* n_cells;
    const uint32_t* status;
    uint32_t N;
};

template <int D> struct VS_ {
    static constexpr int n = (D == 2) ? 2 : 4;
};

__device__ __forceinline__ bool skip(const PCells& c)
{
    return c.status && (*c.status & 1u);
}

// Mirror.cl:35-52: the icell formula of LinkList.cl.in on the new place;
// the mirrored points of an out plane at the particle bounding box lie up to
// 2 h below r_min, where the saturating conversion gives cell column 3
template <int D>
__device__ __forceinline__ uint32_t portal_cell(const float* rv,
                                                const float* r_min,
                                                const uint32_t* nc, float H)
{
    const float idist = 1.f / (2.f * H);
    const uint32_t cx = cvt_u32((rv[0] - r_min[0]) * idist) + 3u;
    const uint32_t cy = cvt_u32((rv[1] - r_min[1]) * idist) + 3u;
    if constexpr (D == 3) {
        const uint32_t cz = cvt_u32((rv[2] - r_min[2]) * idist) + 3u;
        return cx - 1u + (cy - 1u) * nc[0] + (cz - 1u) * nc[0] * nc[1];
    }
    return cx - 1u + (cy - 1u) * nc[0];
}

// Mirror.cl:71-95 (mirror); vectors on every stored component, as float4
template <int D>
__global__ void __launch_bounds__(BS)
    portal_mirror_k(float* __restrict__ r, int* __restrict__ imirrored,
                    uint32_t* __restrict__ icell, uint32_t N,
                    const float* __restrict__ in_r,
                    const float* __restrict__ out_r,
                    const float* __restrict__ pn,
                    const float* __restrict__ r_min,
                    const uint32_t* __restrict__ nc, float H)
{
    constexpr int VS = VS_<D>::n;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float rij[VS];
#pragma unroll
    for (int k = 0; k < VS; k++) rij[k] = r[(size_t)i * VS + k] - out_r[k];
    float dist = rij[0] * pn[0];
#pragma unroll
    for (int k = 1; k < VS; k++) dist = dist + rij[k] * pn[k];
    if (fabsf(dist) > 2.f * H) {
        imirrored[i] = 0;
        return;
    }
    imirrored[i] = 1;
    float rn[VS];
#pragma unroll
    for (int k = 0; k < VS; k++) {
        rn[k] = in_r[k] + rij[k];
        r[(size_t)i * VS + k] = rn[k];
    }
    icell[i] = portal_cell<D>(rn, r_min, nc, H);
}

// Mirror.cl:108-124 (unmirror)
template <int D>
__global__ void __launch_bounds__(BS)
    portal_unmirror_k(float* __restrict__ r,
                      const int* __restrict__ imirrored, uint32_t N,
                      const float* __restrict__ in_r,
                      const float* __restrict__ out_r)
{
    constexpr int VS = VS_<D>::n;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (!imirrored[i]) return;
#pragma unroll
    for (int k = 0; k < VS; k++) {
        const size_t a = (size_t)i * VS + k;
        r[a] = out_r[k] + (r[a] - in_r[k]);
    }
}

// Mirror.cl:136-153 (teleport)
template <int D>
__global__ void __launch_bounds__(BS)
    portal_teleport_k(float* __restrict__ r, uint32_t N,
                      const float* __restrict__ in_r,
                      const float* __restrict__ out_r,
                      const float* __restrict__ pn)
{
    constexpr int VS = VS_<D>::n;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float rij[VS];
#pragma unroll
    for (int k = 0; k < VS; k++) rij[k] = r[(size_t)i * VS + k] - out_r[k];
    float dist = rij[0] * pn[0];
#pragma unroll
    for (int k = 1; k < VS; k++) dist = dist + rij[k] * pn[k];
    if (dist < 0.f) return;
#pragma unroll
    for (int k = 0; k < VS; k++) r[(size_t)i * VS + k] = in_r[k] + rij[k];
}

#define AQ_PORTAL_WALK(...)                                                    \
    walk_in<D>(c.icell[i], c.N, c.n_cells, c.icell, c.ihoc,                    \
               [&](uint32_t j) __VA_ARGS__)

// Shepard.cl:58-100
template <int D>
__global__ void __launch_bounds__(BS)
    portal_shepard_k(const int* __restrict__ imove,
                     const int* __restrict__ imirrored,
                     const float* __restrict__ r,
                     const float* __restrict__ rho,
                     const float* __restrict__ m, float* __restrict__ shepard,
                     PCells c, float H, float CONW, float smax)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N || skip(c)) return;
    const int im = imove[i];
    if ((im < -3) || (im > 1) || !imirrored[i]) return;
    const f3 r_i = ldv<D>(r, i);
    float acc = shepard[i];
    AQ_PORTAL_WALK({
        if ((imove[j] != 1) || imirrored[j]) return;
        const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
        const float s = dot<D>(r_ij, r_ij);
        if (s > smax) return;
        const float q = sqrtf(s) / H;
        acc += kernelW<D>(q) * CONW * m[j] / rho[j];
    });
    shepard[i] = acc;
}

// Interactions.cl:67-140 (fluid and ghost-particle candidates)
template <int D>
__global__ void __launch_bounds__(BS)
    portal_interactions_k(const int* __restrict__ imove,
                          const int* __restrict__ imirrored,
                          const float* __restrict__ r,
                          const float* __restrict__ u,
                          const float* __restrict__ rho,
                          const float* __restrict__ m,
                          const float* __restrict__ p,
                          float* __restrict__ grad_p,
                          float* __restrict__ lap_u,
                          float* __restrict__ div_u, PCells c, float H,
                          float CONF, float smax, int lap)
{
    constexpr float CLEARY = (D == 2) ? 8.f : 10.f;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N || skip(c)) return;
    if (!imirrored[i] || (imove[i] != 1)) return;
    const f3 r_i = ldv<D>(r, i), u_i = ldv<D>(u, i);
    const float p_i = p[i], rho_i = rho[i];
    f3 gp = ldv<D>(grad_p, i), lu = ldv<D>(lap_u, i);
    float dv = div_u[i];
    AQ_PORTAL_WALK({
        const int im_j = imove[j];
        if (imirrored[j] || ((im_j != 1) && (im_j != -1))) return;
        const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
        const float s = dot<D>(r_ij, r_ij);
        if (s > smax) return;
        const float q = sqrtf(s) / H;
        const float rho_j = rho[j], p_j = p[j];
        const f3 du = sub<D>(ldv<D>(u, j), u_i);
        const float udr = dot<D>(du, r_ij);
        const float f_ij = kernelF<D>(q) * CONF * m[j];
        axpy<D>(gp, (p_i + p_j) / (rho_i * rho_j) * f_ij, r_ij);
        if (lap == AQ_LAP_MONAGHAN) {
            const float r2 = (q * q + 0.01f) * H * H;
            axpy<D>(lu, f_ij * CLEARY * udr / (r2 * rho_i * rho_j), r_ij);
        } else {
            axpy<D>(lu, f_ij * 2.f / (rho_i * rho_j), du);
        }
        dv += udr * f_ij * rho_i / rho_j;
    });
    stxyz<D>(grad_p, i, gp);
    stxyz<D>(lap_u, i, lu);
    div_u[i] = dv;
}

// BoundaryIntegrals.cl:55-128 (the element normal is the `normal` argument)
template <int D>
__global__ void __launch_bounds__(BS)
    portal_bi_k(const int* __restrict__ imove,
                const int* __restrict__ imirrored,
                const float* __restrict__ r,
                const float* __restrict__ normal,
                const float* __restrict__ u, const float* __restrict__ rho,
                const float* __restrict__ m, const float* __restrict__ p,
                float* __restrict__ grad_p, float* __restrict__ div_u,
                PCells c, float H, float CONW, float smax)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N || skip(c)) return;
    if (!imirrored[i] || (imove[i] != 1)) return;
    const f3 r_i = ldv<D>(r, i), u_i = ldv<D>(u, i);
    const float p_i = p[i], rho_i = rho[i];
    f3 gp = ldv<D>(grad_p, i);
    float dv = div_u[i];
    AQ_PORTAL_WALK({
        if (imirrored[j] || (imove[j] != -3)) return;
        const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
        const float s = dot<D>(r_ij, r_ij);
        if (s > smax) return;
        const float q = sqrtf(s) / H;
        const f3 n_j = ldv<D>(normal, j);
        const f3 du = sub<D>(ldv<D>(u, j), u_i);
        const float w_ij = kernelW<D>(q) * CONW * m[j];
        axpy<D>(gp, (p_i + p[j]) / rho_i * w_ij, n_j);
        dv += rho_i * dot<D>(du, n_j) * w_ij;
    });
    stxyz<D>(grad_p, i, gp);
    div_u[i] = dv;
}

// deltaSPH.cl:121-191 (lapp) when lap_p_corr is null, 50-119 (full) when
// lap_p is null
template <int D>
__global__ void __launch_bounds__(BS)
    portal_lapp_k(const int* __restrict__ imove,
                  const int* __restrict__ imirrored,
                  const float* __restrict__ r, const float* __restrict__ rho,
                  const float* __restrict__ m, const float* __restrict__ p,
                  float* __restrict__ lap_p, float* __restrict__ lap_p_corr,
                  PCells c, float H, float CONF, float smax)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N || skip(c)) return;
    if (!imirrored[i] || (imove[i] != 1)) return;
    const f3 r_i = ldv<D>(r, i);
    const float p_i = p[i];
    float lp = lap_p ? lap_p[i] : 0.f;
    f3 g = lap_p_corr ? ldv<D>(lap_p_corr, i) : f3{ 0.f, 0.f, 0.f };
    AQ_PORTAL_WALK({
        if (imirrored[j] || (imove[j] != 1)) return;
        const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
        const float s = dot<D>(r_ij, r_ij);
        if (s > smax) return;
        const float q = sqrtf(s) / H;
        const float f_ij = kernelF<D>(q) * CONF * m[j] / rho[j];
        if (lap_p) lp += (p[j] - p_i) * f_ij;
        if (lap_p_corr) axpy<D>(g, (p[j] - p_i) * f_ij, r_ij);
    });
    if (lap_p) lap_p[i] = lp;
    if (lap_p_corr) stxyz<D>(lap_p_corr, i, g);
}

// deltaSPH.cl:193-246 (lapp_corr)
template <int D>
__global__ void __launch_bounds__(BS)
    portal_lapp_corr_k(const int* __restrict__ imove,
                       const int* __restrict__ imirrored,
                       const float* __restrict__ r,
                       const float* __restrict__ rho,
                       const float* __restrict__ m,
                       const float* __restrict__ lap_p_corr,
                       float* __restrict__ lap_p, PCells c, float H,
                       float CONF, float smax)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N || skip(c)) return;
    if (!imirrored[i] || (imove[i] != 1)) return;
    const f3 r_i = ldv<D>(r, i);
    const f3 g_i = ldv<D>(lap_p_corr, i);
    float lp = lap_p[i];
    AQ_PORTAL_WALK({
        if (imirrored[j] || (imove[j] != 1)) return;
        const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
        const float s = dot<D>(r_ij, r_ij);
        if (s > smax) return;
        const float q = sqrtf(s) / H;
        const f3 g_j = ldv<D>(lap_p_corr, j);
        const f3 g = { g_j.x + g_i.x, g_j.y + g_i.y, g_j.z + g_i.z };
        const float f_ij = kernelF<D>(q) * CONF * m[j] / rho[j];
        lp -= 0.5f * dot<D>(g, r_ij) * f_ij;
    });
    lap_p[i] = lp;
}

// MLS.cl:60-110 (entry): mls += outer(r_ij, f_ij r_ij)
template <int D>
__global__ void __launch_bounds__(BS)
    portal_mls_k(const int* __restrict__ imove,
                 const int* __restrict__ imirrored,
                 const float* __restrict__ r, const float* __restrict__ rho,
                 const float* __restrict__ m, float* __restrict__ mls,
                 const uint32_t* __restrict__ mls_imove, PCells c, float H,
                 float CONF, float smax)
{
    constexpr int n = (D == 2) ? 2 : 4, S = (D == 2) ? 4 : 16;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N || skip(c)) return;
    if ((imove[i] != 1) || !imirrored[i]) return;
    const int mi = (int)*mls_imove;
    const f3 r_i = ldv<D>(r, i);
    float M[S];
#pragma unroll
    for (int k = 0; k < S; k++) M[k] = mls[(size_t)i * S + k];
    AQ_PORTAL_WALK({
        if ((imove[j] != mi) || imirrored[j]) return;
        const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
        const float s = dot<D>(r_ij, r_ij);
        if (s > smax) return;
        const float q = sqrtf(s) / H;
        const float f_ij = kernelF<D>(q) * CONF * m[j] / rho[j];
        const float a[3] = { r_ij.x, r_ij.y, r_ij.z };
        for (int x = 0; x < D; x++)
            for (int y = 0; y < D; y++) M[x * n + y] += a[x] * (f_ij * a[y]);
    });
#pragma unroll
    for (int k = 0; k < S; k++) mls[(size_t)i * S + k] = M[k];
}

#undef AQ_PORTAL_WALK

}  // namespace

#define AQ_PG(N) dim3(((N) + BS - 1) / BS), dim3(BS), 0, s
#define AQ_PD(dims, KER, ...)                                                  \
    do {                                                                       \
        if ((dims) == 2)                                                       \
            hipLaunchKernelGGL(KER<2>, __VA_ARGS__);                           \
        else                                                                   \
            hipLaunchKernelGGL(KER<3>, __VA_ARGS__);                           \
    } while (0)

static PCells pcells(const uint32_t* icell, const uint32_t* ihoc,
                     const uint32_t* n_cells, const uint32_t* status,
                     uint32_t N)
{
    return PCells{ icell, ihoc, n_cells, status, N };
}

void launch_portal_mirror(int dims, float* r, int32_t* imirrored,
                          uint32_t* icell, uint32_t N, const float* in_r,
                          const float* out_r, const float* pn,
                          const float* r_min, const uint32_t* n_cells,
                          aq_defs d, hipStream_t s)
{
    if (!N) return;
    AQ_PD(dims, portal_mirror_k, AQ_PG(N), r, imirrored, icell, N, in_r, out_r,
          pn, r_min, n_cells, d.H);
}
void launch_portal_unmirror(int dims, float* r, const int32_t* imirrored,
                            uint32_t N, const float* in_r, const float* out_r,
                            hipStream_t s)
{
    if (!N) return;
    AQ_PD(dims, portal_unmirror_k, AQ_PG(N), r, imirrored, N, in_r, out_r);
}
void launch_portal_teleport(int dims, float* r, uint32_t N, const float* in_r,
                            const float* out_r, const float* pn,
                            hipStream_t s)
{
    if (!N) return;
    AQ_PD(dims, portal_teleport_k, AQ_PG(N), r, N, in_r, out_r, pn);
}
void launch_portal_shepard(int dims, const int32_t* imove,
                           const int32_t* imirrored, const float* r,
                           const float* rho, const float* m, float* shepard,
                           uint32_t N, const uint32_t* icell,
                           const uint32_t* ihoc, const uint32_t* n_cells,
                           const uint32_t* status, aq_defs d, float smax,
                           hipStream_t s)
{
    if (!N) return;
    AQ_PD(dims, portal_shepard_k, AQ_PG(N), imove, imirrored, r, rho, m,
          shepard, pcells(icell, ihoc, n_cells, status, N), d.H, d.CONW,
          smax);
}
void launch_portal_interactions(int dims, const int32_t* imove,
                                const int32_t* imirrored, const float* r,
                                const float* u, const float* rho,
                                const float* m, const float* p, float* grad_p,
                                float* lap_u, float* div_u, uint32_t N,
                                const uint32_t* icell, const uint32_t* ihoc,
                                const uint32_t* n_cells,
                                const uint32_t* status, aq_defs d, float smax,
                                hipStream_t s)
{
    if (!N) return;
    AQ_PD(dims, portal_interactions_k, AQ_PG(N), imove, imirrored, r, u, rho, m,
          p, grad_p, lap_u, div_u, pcells(icell, ihoc, n_cells, status, N),
          d.H, d.CONF, smax, d.lap_formulation);
}
void launch_portal_bi(int dims, const int32_t* imove,
                      const int32_t* imirrored, const float* r,
                      const float* normal, const float* u, const float* rho,
                      const float* m, const float* p, float* grad_p,
                      float* div_u, uint32_t N, const uint32_t* icell,
                      const uint32_t* ihoc, const uint32_t* n_cells,
                      const uint32_t* status, aq_defs d, float smax,
                      hipStream_t s)
{
    if (!N) return;
    AQ_PD(dims, portal_bi_k, AQ_PG(N), imove, imirrored, r, normal, u, rho, m,
          p, grad_p, div_u, pcells(icell, ihoc, n_cells, status, N), d.H,
          d.CONW, smax);
}
void launch_portal_lapp(int dims, const int32_t* imove,
                        const int32_t* imirrored, const float* r,
                        const float* rho, const float* m, const float* p,
                        float* lap_p, float* lap_p_corr, uint32_t N,
                        const uint32_t* icell, const uint32_t* ihoc,
                        const uint32_t* n_cells, const uint32_t* status,
                        aq_defs d, float smax, hipStream_t s)
{
    if (!N) return;
    AQ_PD(dims, portal_lapp_k, AQ_PG(N), imove, imirrored, r, rho, m, p, lap_p,
          lap_p_corr, pcells(icell, ihoc, n_cells, status, N), d.H, d.CONF,
          smax);
}
void launch_portal_lapp_corr(int dims, const int32_t* imove,
                             const int32_t* imirrored, const float* r,
                             const float* rho, const float* m,
                             const float* lap_p_corr, float* lap_p,
                             uint32_t N, const uint32_t* icell,
                             const uint32_t* ihoc, const uint32_t* n_cells,
                             const uint32_t* status, aq_defs d, float smax,
                             hipStream_t s)
{
    if (!N) return;
    AQ_PD(dims, portal_lapp_corr_k, AQ_PG(N), imove, imirrored, r, rho, m,
          lap_p_corr, lap_p, pcells(icell, ihoc, n_cells, status, N), d.H,
          d.CONF, smax);
}
void launch_portal_mls(int dims, const int32_t* imove,
                       const int32_t* imirrored, const float* r,
                       const float* rho, const float* m, float* mls,
                       const uint32_t* mls_imove, uint32_t N,
                       const uint32_t* icell, const uint32_t* ihoc,
                       const uint32_t* n_cells, const uint32_t* status,
                       aq_defs d, float smax, hipStream_t s)
{
    if (!N) return;
    AQ_PD(dims, portal_mls_k, AQ_PG(N), imove, imirrored, r, rho, m, mls,
          mls_imove, pcells(icell, ihoc, n_cells, status, N), d.H, d.CONF,
          smax);
}

#undef AQ_PD
#undef AQ_PG

}  // namespace aq
