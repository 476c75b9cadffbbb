// Ghost-particle walls (resources/Scripts/cfd/Boundary/GP/*.cl) on gfx950.
//
// Ghosts are imove == -1, each associated with a mirroring boundary element
// (imove == -2, unit outward normal, area in m) by associations[i] (>= N:
// none).  One thread per particle; the pair kernels walk the 3^D link-list
// cells in the reference order, so every sum associates as in the reference
// kernels.  Candidates are rejected with s > s_max, the exact equivalent of
// the reference's q >= SUPPORT test (support_threshold in sph.hip).
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

// GP/Sort.cl:35-53: associations into the new sorted order
__global__ void gp_sort_k(const uint32_t* __restrict__ assoc_in,
                          uint32_t* __restrict__ assoc,
                          const uint32_t* __restrict__ id_sorted, uint32_t N)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint32_t a = assoc_in[i];
    assoc[id_sorted[i]] = (a >= N) ? N : id_sorted[a];
}

// GP/Mirror.cl:47-90
template <int D>
__global__ void gp_mirror_k(const int* __restrict__ imove,
                            const uint32_t* __restrict__ assoc,
                            const float* __restrict__ normal,
                            float* __restrict__ r,
                            uint32_t* __restrict__ gp_icell, uint32_t N,
                            const float* r_min, const uint32_t* n_cells,
                            float H, float SUPPORT)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != -1) return;
    const uint32_t iref = assoc[i];
    if (iref >= N) return;
    const f3 r_ref = ldv<D>(r, iref);
    const f3 n_ref = ldv<D>(normal, iref);
    f3 r_i = ldv<D>(r, i);
    const float d = dot<D>(sub<D>(r_ref, r_i), n_ref);
    r_i.x += 2.f * (d * n_ref.x);
    r_i.y += 2.f * (d * n_ref.y);
    if (D == 3) r_i.z += 2.f * (d * n_ref.z);
    stxyz<D>(r, i, r_i);
    const float idist = 1.f / (SUPPORT * H);
    const uint32_t nx = n_cells[0], ny = n_cells[1];
    const uint32_t cx = cvt_u32((r_i.x - r_min[0]) * idist) + 3u;
    const uint32_t cy = cvt_u32((r_i.y - r_min[1]) * idist) + 3u;
    if constexpr (D == 3) {
        const uint32_t cz = cvt_u32((r_i.z - r_min[2]) * idist) + 3u;
        gp_icell[i] = cx - 1u + (cy - 1u) * nx + (cz - 1u) * nx * ny;
    } else {
        gp_icell[i] = cx - 1u + (cy - 1u) * nx;
    }
}

// GP/Interpolation.cl:58-140: interpolation at the mirrored position,
// walking around gp_icell[i].  Neighbour cells beyond the grid
// (c_j >= n_cells.w) are skipped: the reference would read ihoc out of
// range there.
template <int D>
__global__ void __launch_bounds__(BS)
    gp_interp_k(const int* __restrict__ imove, const float* __restrict__ r,
                const float* __restrict__ m, const float* __restrict__ rho,
                const float* __restrict__ p, const float* __restrict__ u,
                float* __restrict__ gp_rho, float* __restrict__ gp_p,
                float* __restrict__ gp_u, float* __restrict__ shepard,
                const uint32_t* __restrict__ gp_icell, Cells c, float H,
                float CONW, float smax)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    if (skip(c)) return;
    if (imove[i] != -1) return;
    const f3 r_i = ldv<D>(r, i);
    float a_rho = gp_rho[i], a_p = gp_p[i], sh = 0.f;
    f3 a_u = ldv<D>(gp_u, i);
    const uint32_t c_i = gp_icell[i];
    const uint32_t nx = c.n_cells[0], ny = c.n_cells[1], nw = c.n_cells[3];
#pragma unroll 1
    for (int ci = -1; ci <= 1; ci++)
#pragma unroll 1
        for (int cj = -1; cj <= 1; cj++)
#pragma unroll 1
            for (int ck = (D == 3 ? -1 : 0); ck <= (D == 3 ? 1 : 0); ck++) {
                const uint32_t c_j = c_i + (uint32_t)ci + (uint32_t)cj * nx +
                                     (uint32_t)ck * nx * ny;
                if (c_j >= nw) continue;
                uint32_t j = c.ihoc[c_j];
                while ((j < c.N) && (c.icell[j] == c_j)) {
                    if (imove[j] == 1) {
                        const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
                        const float s = dot<D>(r_ij, r_ij);
                        if (!(s > smax)) {
                            const float q = sqrtf(s) / H;
                            const float rho_j = rho[j];
                            const float w_ij = kernelW<D>(q) * CONW * m[j] / rho_j;
                            a_rho += w_ij * rho_j;
                            a_p += w_ij * p[j];
                            const f3 u_j = ldv<D>(u, j);
                            a_u.x += w_ij * u_j.x;
                            a_u.y += w_ij * u_j.y;
                            if (D == 3) a_u.z += w_ij * u_j.z;
                            sh += w_ij;
                        }
                    }
                    j++;
                }
            }
    gp_rho[i] = a_rho;
    gp_p[i] = a_p;
    stxyz<D>(gp_u, i, a_u);
    shepard[i] = sh;
}

// GP/Renormalization.cl:44-69 (the whole vec, w included, is divided)
template <int D>
__global__ void gp_renorm_k(const int* __restrict__ imove,
                            const float* __restrict__ shepard,
                            float* __restrict__ gp_rho,
                            float* __restrict__ gp_p,
                            float* __restrict__ gp_u, uint32_t N)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != -1) return;
    float s = shepard[i];
    if (s < 1.0E-6f) s = 1.f;
    gp_p[i] /= s;
    gp_rho[i] /= s;
    constexpr int VS = (D == 2) ? 2 : 4;
    for (int k = 0; k < VS; k++) gp_u[(size_t)i * VS + k] /= s;
}

// GP/PreLapU.cl:51-90 (model 1: Takeda extension, otherwise frozen)
template <int D>
__global__ void gp_prelapu_k(const int* __restrict__ imove,
                             const uint32_t* __restrict__ assoc,
                             const float* __restrict__ normal,
                             const float* __restrict__ gp_rho,
                             const float* __restrict__ gp_u,
                             float* __restrict__ rho, float* __restrict__ u,
                             uint32_t N, int model)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != -1) return;
    const uint32_t iref = assoc[i];
    if (iref >= N) return;
    rho[i] = gp_rho[i];
    if (model != 1) return;
    const f3 n = ldv<D>(normal, iref);
    const f3 ur = ldv<D>(u, iref);
    const float dr = dot<D>(ur, n);
    const f3 gu = ldv<D>(gp_u, i);
    const float dg = dot<D>(gu, n);
    float o[3];
    const float nn[3] = { n.x, n.y, n.z };
    const float uu[3] = { ur.x, ur.y, ur.z };
    const float gg[3] = { gu.x, gu.y, gu.z };
    for (int k = 0; k < D; k++) {
        const float u_ref = uu[k] - dr * nn[k];
        const float u_n = dg * nn[k];
        const float u_t = 2.f * u_ref - (gg[k] - u_n);
        o[k] = u_n + u_t;
    }
    stxyz<D>(u, i, f3{ o[0], o[1], D == 3 ? o[2] : 0.f });
}

// GP/FreeSlip.cl:46-66
template <int D>
__global__ void gp_freeslip_k(const uint32_t* __restrict__ iset,
                              const int* __restrict__ imove,
                              const uint32_t* __restrict__ assoc,
                              const float* __restrict__ gp_u,
                              float* __restrict__ u, uint32_t N,
                              const uint32_t* freeslip_iset)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if ((imove[i] != -1) || (iset[i] != *freeslip_iset)) return;
    if (assoc[i] >= N) return;
    stxyz<D>(u, i, ldv<D>(gp_u, i));
}

// GP/LapU.cl:58-127
template <int D>
__global__ void __launch_bounds__(BS)
    gp_lapu_k(const int* __restrict__ imove, const float* __restrict__ r,
              const float* __restrict__ u, const float* __restrict__ rho,
              const float* __restrict__ m, float* __restrict__ lap_u, Cells c,
              float H, float CONF, float smax, int lap)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    if (skip(c)) return;
    const int im_i = imove[i];
    if ((im_i != 1) && (im_i != -2)) return;
    constexpr float CLEARY = (D == 2) ? 8.f : 10.f;
    const f3 r_i = ldv<D>(r, i);
    const f3 u_i = ldv<D>(u, i);
    const float rho_i = rho[i];
    f3 lu = ldv<D>(lap_u, i);
    walk<D>(c.icell[i], c.N, c.n_cells[0], c.n_cells[1], c.icell, c.ihoc,
            [&](uint32_t j) {
                const int im_j = imove[j];
                if ((im_j != 1) && (im_j != -1)) return;
                if (im_i == im_j) return;
                const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
                const float s = dot<D>(r_ij, r_ij);
                if (s > smax) return;
                const float q = sqrtf(s) / H;
                const float rho_j = rho[j];
                const f3 u_ij = sub<D>(ldv<D>(u, j), u_i);
                const float udr = dot<D>(u_ij, r_ij);
                const float f_ij = kernelF<D>(q) * CONF * m[j];
                if (lap == AQ_LAP_MONAGHAN) {
                    const float r2 = (q * q + 0.01f) * H * H;
                    axpy<D>(lu, f_ij * CLEARY * udr / (r2 * rho_i * rho_j),
                            r_ij);
                } else {
                    axpy<D>(lu, f_ij * 2.f / (rho_i * rho_j), u_ij);
                }
            });
    stxyz<D>(lap_u, i, lu);
}

// GP/PreInteractions.cl:66-104
template <int D>
__global__ void gp_preinter_k(const uint32_t* __restrict__ iset,
                              const int* __restrict__ imove,
                              const uint32_t* __restrict__ assoc,
                              const float* __restrict__ r,
                              const float* __restrict__ normal,
                              const float* __restrict__ dudt,
                              const float* __restrict__ lap_u,
                              const float* __restrict__ gp_p,
                              const float* __restrict__ gp_u,
                              const float* __restrict__ rho,
                              float* __restrict__ p, float* __restrict__ u,
                              const float* __restrict__ visc_dyn, uint32_t N,
                              const float* g)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (imove[i] != -1) return;
    const uint32_t iref = assoc[i];
    if (iref >= N) return;
    const f3 n = ldv<D>(normal, iref);
    const f3 gu = ldv<D>(gp_u, i);
    const float dg = dot<D>(gu, n);
    const float dr = dot<D>(ldv<D>(u, iref), n);
    const float nn[3] = { n.x, n.y, n.z };
    const float gg[3] = { gu.x, gu.y, gu.z };
    float o[3];
    for (int k = 0; k < D; k++) {
        const float u_n = dg * nn[k];
        const float u_t = gg[k] - u_n;
        o[k] = 2.f * dr * nn[k] - u_n + u_t;
    }
    stxyz<D>(u, i, f3{ o[0], o[1], D == 3 ? o[2] : 0.f });
    const float mu = visc_dyn[iset[iref]];
    const f3 dd = ldv<D>(dudt, iref);
    const f3 lu = ldv<D>(lap_u, iref);
    f3 a;
    a.x = g[0] - dd.x + mu * lu.x;
    a.y = g[1] - dd.y + mu * lu.y;
    a.z = (D == 3) ? g[2] - dd.z + mu * lu.z : 0.f;
    const float dpdn = rho[iref] * dot<D>(a, n);
    const float dn = dot<D>(sub<D>(ldv<D>(r, i), ldv<D>(r, iref)), n);
    p[i] = gp_p[i] + 2.f * dpdn * dn;
}

// GP/Interactions.cl:62-132
template <int D>
__global__ void __launch_bounds__(BS)
    gp_interactions_k(const int* __restrict__ imove,
                      const float* __restrict__ r,
                      const float* __restrict__ u,
                      const float* __restrict__ rho,
                      const float* __restrict__ m,
                      const float* __restrict__ p, float* __restrict__ grad_p,
                      float* __restrict__ div_u, Cells c, float H, float CONF,
                      float smax)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    if (skip(c)) return;
    if (imove[i] != 1) return;
    const f3 r_i = ldv<D>(r, i);
    const f3 u_i = ldv<D>(u, i);
    const float p_i = p[i];
    const float rho_i = rho[i];
    f3 gp = ldv<D>(grad_p, i);
    float dv = div_u[i];
    walk<D>(c.icell[i], c.N, c.n_cells[0], c.n_cells[1], c.icell, c.ihoc,
            [&](uint32_t j) {
                if (imove[j] != -1) return;
                const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
                const float s = dot<D>(r_ij, r_ij);
                if (s > smax) return;
                const float q = sqrtf(s) / H;
                const float rho_j = rho[j];
                const float p_j = p[j];
                const float udr = dot<D>(sub<D>(ldv<D>(u, j), u_i), r_ij);
                const float f_ij = kernelF<D>(q) * CONF * m[j];
                axpy<D>(gp, (p_i + p_j) / (rho_i * rho_j) * f_ij, r_ij);
                dv += udr * f_ij * rho_i / rho_j;
            });
    stxyz<D>(grad_p, i, gp);
    div_u[i] = dv;
}

}  // namespace

// ---- launchers (used by the kernel-tool registry and the C ABI) ------------
#define AQ_G(N) dim3(blocks_for((N), BS)), dim3(BS), 0, s
#define AQ_D(dims, KER, ...)                                                   \
    do {                                                                       \
        if ((dims) == 2)                                                       \
            hipLaunchKernelGGL(KER<2>, __VA_ARGS__);                           \
        else                                                                   \
            hipLaunchKernelGGL(KER<3>, __VA_ARGS__);                           \
    } while (0)

void launch_gp_sort(const uint32_t* assoc_in, uint32_t* assoc,
                    const uint32_t* id_sorted, uint32_t N, hipStream_t s)
{
    hipLaunchKernelGGL(gp_sort_k, AQ_G(N), assoc_in, assoc, id_sorted, N);
}

void launch_gp_mirror(int dims, const int32_t* imove, const uint32_t* assoc,
                      const float* normal, float* r, uint32_t* gp_icell,
                      uint32_t N, const float* r_min, const uint32_t* n_cells,
                      float H, float SUPPORT, hipStream_t s)
{
    AQ_D(dims, gp_mirror_k, AQ_G(N), imove, assoc, normal, r, gp_icell, N,
         r_min, n_cells, H, SUPPORT);
}

void launch_gp_interp(int dims, const int32_t* imove, const float* r,
                      const float* m, const float* rho, const float* p,
                      const float* u, float* gp_rho, float* gp_p, float* gp_u,
                      float* shepard, uint32_t N, const uint32_t* gp_icell,
                      const uint32_t* icell, const uint32_t* ihoc,
                      const uint32_t* n_cells, const uint32_t* status,
                      aq_defs d, hipStream_t s)
{
    Cells c{ icell, ihoc, n_cells, status, N };
    AQ_D(dims, gp_interp_k, AQ_G(N), imove, r, m, rho, p, u, gp_rho, gp_p,
         gp_u, shepard, gp_icell, c, d.H, d.CONW, support_threshold(d.H));
}

void launch_gp_renorm(int dims, const int32_t* imove, const float* shepard,
                      float* gp_rho, float* gp_p, float* gp_u, uint32_t N,
                      hipStream_t s)
{
    AQ_D(dims, gp_renorm_k, AQ_G(N), imove, shepard, gp_rho, gp_p, gp_u, N);
}

void launch_gp_prelapu(int dims, const int32_t* imove, const uint32_t* assoc,
                       const float* normal, const float* gp_rho,
                       const float* gp_u, float* rho, float* u, uint32_t N,
                       int model, hipStream_t s)
{
    AQ_D(dims, gp_prelapu_k, AQ_G(N), imove, assoc, normal, gp_rho, gp_u, rho,
         u, N, model);
}

void launch_gp_freeslip(int dims, const uint32_t* iset, const int32_t* imove,
                        const uint32_t* assoc, const float* gp_u, float* u,
                        uint32_t N, const uint32_t* freeslip_iset,
                        hipStream_t s)
{
    AQ_D(dims, gp_freeslip_k, AQ_G(N), iset, imove, assoc, gp_u, u, N,
         freeslip_iset);
}

void launch_gp_lapu(int dims, const int32_t* imove, const float* r,
                    const float* u, const float* rho, const float* m,
                    float* lap_u, uint32_t N, const uint32_t* icell,
                    const uint32_t* ihoc, const uint32_t* n_cells,
                    const uint32_t* status, aq_defs d, hipStream_t s)
{
    Cells c{ icell, ihoc, n_cells, status, N };
    AQ_D(dims, gp_lapu_k, AQ_G(N), imove, r, u, rho, m, lap_u, c, d.H, d.CONF,
         support_threshold(d.H), d.lap_formulation);
}

void launch_gp_preinteractions(int dims, const uint32_t* iset,
                               const int32_t* imove, const uint32_t* assoc,
                               const float* r, const float* normal,
                               const float* dudt, const float* lap_u,
                               const float* gp_p, const float* gp_u,
                               const float* rho, float* p, float* u,
                               const float* visc_dyn, uint32_t N,
                               const float* g, hipStream_t s)
{
    AQ_D(dims, gp_preinter_k, AQ_G(N), iset, imove, assoc, r, normal, dudt,
         lap_u, gp_p, gp_u, rho, p, u, visc_dyn, N, g);
}

void launch_gp_interactions(int dims, const int32_t* imove, const float* r,
                            const float* u, const float* rho, const float* m,
                            const float* p, float* grad_p, float* div_u,
                            uint32_t N, const uint32_t* icell,
                            const uint32_t* ihoc, const uint32_t* n_cells,
                            const uint32_t* status, aq_defs d, hipStream_t s)
{
    Cells c{ icell, ihoc, n_cells, status, N };
    AQ_D(dims, gp_interactions_k, AQ_G(N), imove, r, u, rho, m, p, grad_p,
         div_u, c, d.H, d.CONF, support_threshold(d.H));
}

#undef AQ_D
#undef AQ_G

}  // namespace aq
