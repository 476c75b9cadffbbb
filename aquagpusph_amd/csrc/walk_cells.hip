// Cell-parallel neighbour walk with per-lane candidate bitsets.
//
// One wavefront owns one non-empty cell (64 of its particles at a time, one
// per lane); the candidates of the 27 (3D) / 9 (2D) neighbour cells are
// enumerated in the reference order (x offset outermost, then y, then z;
// sorted index inside a cell, types/3D.h:197-219), which is the same list
// for every particle of the cell.
//
//   phase 1  the lanes load 64 candidates at a time (coalesced), the
//            candidate loop reads them back from the registers of the lane
//            that loaded them (readlane, no memory latency in the loop) and
//            every lane records in LDS one bit per candidate it accepts
//            (q < 2, fluid or boundary element);
//   phase 2  every lane walks its own accepted bits in increasing candidate
//            order and evaluates the pair terms.
//
// Each lane therefore adds exactly the terms the reference kernel adds, in
// the same order, and the results are bit-identical; but the pair
// arithmetic runs with the lanes that have work instead of under a mask of
// the lanes that accepted the current candidate.  Candidates are processed
// in segments of SEG (bits per lane in LDS: SEG/8 bytes).
#include "aq_device.hpp"
#include "aq_internal.hpp"

namespace aq {

constexpr int WC_WAVES = 4;           // wavefronts per workgroup
constexpr int WC_SEGW = 16;           // 32-bit words per lane per segment
constexpr int WC_SEG = WC_SEGW * 32;  // candidates per segment

template <int D> struct NCell {
    static constexpr int n = (D == 2) ? 9 : 27;
};

__device__ __forceinline__ uint32_t uni(uint32_t x)
{
    return __builtin_amdgcn_readfirstlane(x);
}

struct WalkArgs {
    const uint32_t* icell;
    const uint32_t* ihoc;
    const uint32_t* n_cells;
    const uint32_t* cell_end;
    const uint32_t* heads;
    const uint32_t* n_heads;
    const uint32_t* status;
    uint32_t N;
};

// Shared state of one wavefront: neighbour-cell table and the bitsets.
struct WaveLds {
    uint32_t start[27];
    uint32_t cum[28];
    uint32_t cand[WC_SEG];  // particle index of each segment candidate
    uint32_t bits[WC_SEGW][64];
};

template <int D>
__device__ __forceinline__ uint32_t neighbour_table(const WalkArgs& w,
                                                    uint32_t c_i, WaveLds& L,
                                                    int lane)
{
    // lane k < ncell fills entry k; the prefix sum runs on lane 0
    constexpr int NC = NCell<D>::n;
    const uint32_t nx = w.n_cells[0], ny = w.n_cells[1];
    if (lane < NC) {
        int ci, cj, ck;
        if (D == 2) {
            ci = lane / 3 - 1;
            cj = lane % 3 - 1;
            ck = 0;
        } else {
            ci = lane / 9 - 1;
            cj = (lane / 3) % 3 - 1;
            ck = lane % 3 - 1;
        }
        const uint32_t c = c_i + (uint32_t)ci + (uint32_t)cj * nx +
                           (uint32_t)ck * nx * ny;
        const uint32_t s = w.ihoc[c];
        L.start[lane] = s;
        L.cum[lane + 1] = (s < w.N) ? w.cell_end[c] - s : 0u;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_s_waitcnt(0);
    if (lane == 0) {
        uint32_t acc = 0;
        L.cum[0] = 0;
        for (int k = 0; k < NC; k++) {
            acc += L.cum[k + 1];
            L.cum[k + 1] = acc;
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_s_waitcnt(0);
    return uni(L.cum[NC]);
}

// ---------------------------------------------------------------------------
// Fused fluid walk (same terms as fluid_pass_k in sph.hip).
// ---------------------------------------------------------------------------
template <int D, int SHEP, bool LAPP>
__global__ void __launch_bounds__(64 * WC_WAVES)
    fluid_pass_cells_k(const int* __restrict__ imove,
                       const float* __restrict__ r,
                       const float* __restrict__ u,
                       const float* __restrict__ rho,
                       const float* __restrict__ m,
                       const float* __restrict__ p,
                       const float* __restrict__ normal,
                       const float* __restrict__ tangent,
                       const float* __restrict__ binormal,
                       float* __restrict__ shepard, float* __restrict__ grad_p,
                       float* __restrict__ lap_u, float* __restrict__ div_u,
                       float* __restrict__ lap_p, uint8_t* __restrict__ near,
                       WalkArgs w, float H, float CONW, float CONF, float DIMS,
                       float smax, int lap)
{
    constexpr int NC = NCell<D>::n;
    constexpr float CLEARY = (D == 2) ? 8.f : 10.f;
    __shared__ WaveLds lds[WC_WAVES];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    WaveLds& L = lds[wv];
    if (w.status && (*w.status & 1u)) return;
    const uint32_t n_heads = uni(*w.n_heads);
    const uint32_t gwave = blockIdx.x * WC_WAVES + wv;
    const uint32_t nwaves = gridDim.x * WC_WAVES;
    const float eps = 1E-8f * H;

    for (uint32_t hk = gwave; hk < n_heads; hk += nwaves) {
        const uint32_t cstart = uni(w.heads[hk]);
        const uint32_t c_i = uni(w.icell[cstart]);
        const uint32_t cend = uni(w.cell_end[c_i]);
        const uint32_t total = neighbour_table<D>(w, c_i, L, lane);
        for (uint32_t i0 = cstart; i0 < cend; i0 += 64) {
            const uint32_t i = i0 + lane;
            const bool act = (i < cend) && (imove[i] == 1);
            f3 r_i = { 0.f, 0.f, 0.f }, u_i = { 0.f, 0.f, 0.f };
            float p_i = 0.f, rho_i = 1.f;
            f3 gp = { 0.f, 0.f, 0.f }, lu = { 0.f, 0.f, 0.f };
            float dv = 0.f, lp = 0.f, sh = 0.f;
            if (act) {
                r_i = ldv<D>(r, i);
                u_i = ldv<D>(u, i);
                p_i = p[i];
                rho_i = rho[i];
                gp = ldv<D>(grad_p, i);
                lu = ldv<D>(lap_u, i);
                dv = div_u[i];
                if (LAPP) lp = lap_p[i];
                sh = (SHEP == 2) ? 1.f : (SHEP == 1 ? shepard[i] : 0.f);
            }
            bool self_added = false;
            bool nearw = false;
            int kl = 0;  // per-lane cursor: neighbour cell of candidate t
            for (uint32_t seg = 0; seg < total; seg += WC_SEG) {
                const uint32_t nseg = min((uint32_t)WC_SEG, total - seg);
                // ---- phase 1: accept bits --------------------------------
                // chunks of 64 candidates: lane k loads candidate cb + k
                // (coalesced), the uniform loop reads them back by readlane
                for (uint32_t cb = 0; cb < nseg; cb += 64) {
                    const uint32_t t = seg + cb + lane;
                    float cx = 0.f, cy = 0.f, cz = 0.f;
                    int cim = 0;
                    if (cb + lane < nseg) {
                        while (t >= L.cum[kl + 1]) kl++;
                        const uint32_t j = L.start[kl] + (t - L.cum[kl]);
                        const f3 rj = ldv<D>(r, j);
                        cx = rj.x;
                        cy = rj.y;
                        cz = rj.z;
                        cim = imove[j];
                        L.cand[cb + lane] = j;
                    }
                    const uint32_t nk = min(64u, nseg - cb);
                    uint32_t lo = 0, hi = 0;
                    for (uint32_t k = 0; k < nk; k++) {
                        f3 r_j;
                        r_j.x = __int_as_float(__builtin_amdgcn_readlane(
                            __float_as_int(cx), k));
                        r_j.y = __int_as_float(__builtin_amdgcn_readlane(
                            __float_as_int(cy), k));
                        r_j.z = (D == 3) ? __int_as_float(
                                               __builtin_amdgcn_readlane(
                                                   __float_as_int(cz), k))
                                         : 0.f;
                        const int im_j = __builtin_amdgcn_readlane(cim, k);
                        const f3 r_ij = sub<D>(r_j, r_i);
                        const float s = dot<D>(r_ij, r_ij);
                        const bool in = act && !(s > smax);
                        if (in && im_j == -3) nearw = true;
                        const bool acc =
                            in && (im_j == 1 || (SHEP == 2 && im_j == -3));
                        const uint32_t bit = (acc ? 1u : 0u) << (k & 31);
                        if (k < 32)
                            lo |= bit;
                        else
                            hi |= bit;
                    }
                    L.bits[cb >> 5][lane] = lo;
                    if (nk > 32) L.bits[(cb >> 5) + 1][lane] = hi;
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_s_waitcnt(0);
                // ---- phase 2: each lane walks its accepted candidates ------
                const int nw = (int)((nseg + 31) >> 5);
                int wp = 0;
                uint32_t bw = (nw > 0) ? L.bits[0][lane] : 0u;
                while (true) {
                    while (bw == 0 && wp + 1 < nw) {
                        wp++;
                        bw = L.bits[wp][lane];
                    }
                    if (bw == 0) break;
                    const int b = __builtin_ctz(bw);
                    bw &= bw - 1;
                    const uint32_t j = L.cand[wp * 32 + b];
                    const f3 r_ij = sub<D>(ldv<D>(r, j), r_i);
                    const float s = dot<D>(r_ij, r_ij);
                    const int im_j = imove[j];
                    if (im_j == 1) {
                        const float q = sqrtf(s) / H;
                        const float m_j = m[j];
                        const float rho_j = rho[j];
                        if (SHEP == 1)
                            sh += kernelW<D>(q) * CONW * m_j / rho_j;
                        if (i == j) continue;
                        const float p_j = p[j];
                        const f3 u_ij = sub<D>(ldv<D>(u, j), u_i);
                        const float udr = dot<D>(u_ij, r_ij);
                        const float fk = kernelF<D>(q) * CONF;
                        const float f_ij = fk * m_j;
                        axpy<D>(gp, (p_i + p_j) / (rho_i * rho_j) * f_ij,
                                r_ij);
                        if (lap == AQ_LAP_MONAGHAN) {
                            const float r2 = (q * q + 0.01f) * H * H;
                            axpy<D>(lu,
                                    f_ij * CLEARY * udr /
                                        (r2 * rho_i * rho_j),
                                    r_ij);
                        } else {
                            axpy<D>(lu, f_ij * 2.f / (rho_i * rho_j), u_ij);
                        }
                        dv += udr * f_ij * rho_i / rho_j;
                        if (LAPP) lp += (p_j - p_i) * (f_ij / rho_j);
                    } else if (SHEP == 2) {
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
                                    0.55f * powf(area_j, 1.f / (DIMS - 1.f));
                                if ((r_t <= dr) && (r_b <= dr)) {
                                    self_added = true;
                                    sh -= 0.5f;
                                }
                            }
                            continue;
                        }
                        const float r_t = fabsf(dot<D>(r_ij, t_j));
                        const float r_b = fabsf(dot<D>(r_ij, b_j));
                        sh += r_n * CONW * kernelS_P<D>(q) * area_j +
                              kernelS_D<D>(fabsf(r_n), r_t, r_b, area_j);
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
            if (act) {
                stxyz<D>(grad_p, i, gp);
                stxyz<D>(lap_u, i, lu);
                div_u[i] = dv;
                if (LAPP) lap_p[i] = lp;
                if (SHEP) shepard[i] = sh;
                if (near) near[i] = nearw ? 1 : 0;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

void launch_fluid_pass_cells(int dims, int shep, bool lapp,
                             const int32_t* imove, const float* r,
                             const float* u, const float* rho, const float* m,
                             const float* p, const float* normal,
                             const float* tangent, const float* binormal,
                             float* shepard, float* grad_p, float* lap_u,
                             float* div_u, float* lap_p, uint8_t* near,
                             uint32_t N, const uint32_t* icell,
                             const uint32_t* ihoc, const uint32_t* n_cells,
                             const uint32_t* cell_end, const uint32_t* heads,
                             const uint32_t* n_heads, const uint32_t* status,
                             aq_defs d, float smax, unsigned nblocks,
                             hipStream_t s)
{
    WalkArgs w{ icell, ihoc, n_cells, cell_end, heads, n_heads, status, N };
#define AQ_FC(DD, SH, LP)                                                      \
    hipLaunchKernelGGL((fluid_pass_cells_k<DD, SH, LP>), dim3(nblocks),        \
                       dim3(64 * WC_WAVES), 0, s, imove, r, u, rho, m, p,      \
                       normal, tangent, binormal, shepard, grad_p, lap_u,      \
                       div_u, lap_p, near, w, d.H, d.CONW, d.CONF, d.DIMS,     \
                       smax, d.lap_formulation)
#define AQ_FC_D(SH, LP)                                                        \
    do {                                                                       \
        if (dims == 2) AQ_FC(2, SH, LP);                                       \
        else AQ_FC(3, SH, LP);                                                 \
    } while (0)
    if (shep == 0 && !lapp) AQ_FC_D(0, false);
    else if (shep == 0 && lapp) AQ_FC_D(0, true);
    else if (shep == 1 && !lapp) AQ_FC_D(1, false);
    else if (shep == 1 && lapp) AQ_FC_D(1, true);
    else if (shep == 2 && !lapp) AQ_FC_D(2, false);
    else AQ_FC_D(2, true);
#undef AQ_FC_D
#undef AQ_FC
}

}  // namespace aq
