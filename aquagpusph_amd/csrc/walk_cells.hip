// Cell-parallel neighbour walk with per-lane candidate bitsets.
//
// One wavefront owns one non-empty cell (64 of its particles at a time, one
// per lane); the candidates of the 27 (3D) / 9 (2D) neighbour cells are
// enumerated in the reference order (x offset outermost, then y, then z;
// sorted index inside a cell, types/3D.h:197-219), which is the same list
// for every particle of the cell.  Candidates are processed in segments of
// WC_SEG:
//
//   phase 1  the lanes load 64 candidates at a time (coalesced) into LDS;
//            the candidate loop reads them back by broadcast and every lane
//            records one bit per candidate it accepts (q <= 2, fluid
//            particle or wall element);
//   phase 2  every lane walks its own accepted fluid candidates in
//            increasing order and evaluates the pair terms, then (only in
//            wavefronts that met one) its accepted wall elements.
//
// The fluid sums (grad p, lap u, div u, lap p, and the cfd Shepard sum) and
// the boundary-integral Shepard sum take their terms from disjoint candidate
// sets, so splitting the walk in two passes keeps each sum's terms and
// their order: every lane adds exactly what the reference kernels add, in
// the same order, and the results are bit-identical.  The pair arithmetic
// runs with the lanes that have work instead of under a mask of the lanes
// that accepted the current candidate.
#include "aq_device.hpp"
#include "aq_internal.hpp"

namespace aq {

constexpr int WC_WAVES = 4;  // wavefronts per workgroup
// a segment holds SEGW 32-bit words of accept bits per lane, i.e.
// SEGW * 32 candidates

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
constexpr uint32_t WC_ELEM = 0x80000000u;  // candidate flag: wall element

template <int SEGW> struct WaveLds {
    uint32_t start[27];
    uint32_t cum[28];
    float4 stage[64];            // candidate chunk: x, y, z, imove bits
    uint32_t cand[SEGW * 32];    // particle index of each segment candidate
    uint32_t bits[SEGW][64];
};

template <int D, class Lds>
__device__ __forceinline__ uint32_t neighbour_table(const WalkArgs& w,
                                                    uint32_t c_i, Lds& L,
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
template <int D, int SHEP, bool LAPP, int SEGW>
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
    constexpr uint32_t WC_SEG = SEGW * 32;
    __shared__ WaveLds<SEGW> lds[WC_WAVES];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    WaveLds<SEGW>& L = lds[wv];
    if (w.status && (*w.status & 1u)) return;
    const uint32_t n_heads = uni(*w.n_heads);
    const uint32_t gwave = blockIdx.x * WC_WAVES + wv;
    const uint32_t nwaves = gridDim.x * WC_WAVES;
    const float eps = 1E-8f * H;
    // wall elements enter the accept bits for the boundary-integral Shepard
    // sum and for the near-wall flags
    const bool want_el = (SHEP == 2) || (near != nullptr);

    for (uint32_t hk = gwave; hk < n_heads; hk += nwaves) {
        const uint32_t cstart = uni(w.heads[hk]);
        const uint32_t c_i = uni(w.icell[cstart]);
        const uint32_t cend = uni(w.cell_end[c_i]);
        bool table = false;
        uint32_t total = 0;
        for (uint32_t i0 = cstart; i0 < cend; i0 += 64) {
            const uint32_t i = i0 + lane;
            const bool act = (i < cend) && (imove[i] == 1);
            if (__ballot(act) == 0) continue;  // no fluid particle here
            if (!table) {
                total = neighbour_table<D>(w, c_i, L, lane);
                table = true;
            }
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
                // Bits are shifted in from the right: candidate k of a chunk
                // ends up at bit 31 - (k & 31) of word k >> 5 (first
                // candidate = most significant bit).
                bool elem = false;  // this lane accepted a wall element
                for (uint32_t cb = 0; cb < nseg; cb += 64) {
                    const uint32_t t = seg + cb + lane;
                    const bool valid = cb + lane < nseg;
                    float4 st = make_float4(0.f, 0.f, 0.f, 0.f);
                    bool take = false, is_el = false;
                    if (valid) {
                        while (t >= L.cum[kl + 1]) kl++;
                        const uint32_t j = L.start[kl] + (t - L.cum[kl]);
                        const f3 rj = ldv_o<D>(r, j);
                        const int im_j = ldi_o(imove, j);
                        is_el = (im_j == -3);
                        take = (im_j == 1) || (want_el && is_el);
                        st = make_float4(rj.x, rj.y, rj.z, 0.f);
                        L.cand[cb + lane] = j | (is_el ? WC_ELEM : 0u);
                    }
                    L.stage[lane] = st;
                    // candidates a lane may accept / wall elements among them
                    const uint64_t tm = __ballot(take);
                    const uint64_t em = __ballot(take && is_el);
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_s_waitcnt(0);
                    uint32_t lo = 0, hi = 0;
#pragma unroll 8
                    for (int k = 0; k < 32; k++) {
                        const float4 c = L.stage[k];
                        const f3 r_ij = sub<D>(f3{ c.x, c.y, c.z }, r_i);
                        const float s = dot<D>(r_ij, r_ij);
                        lo = (lo << 1) | (!(s > smax) ? 1u : 0u);
                    }
                    if (nseg - cb > 32) {
#pragma unroll 8
                        for (int k = 32; k < 64; k++) {
                            const float4 c = L.stage[k];
                            const f3 r_ij = sub<D>(f3{ c.x, c.y, c.z }, r_i);
                            const float s = dot<D>(r_ij, r_ij);
                            hi = (hi << 1) | (!(s > smax) ? 1u : 0u);
                        }
                    }
                    lo &= __builtin_bitreverse32((uint32_t)tm);
                    hi &= __builtin_bitreverse32((uint32_t)(tm >> 32));
                    if (!act) lo = hi = 0;
                    elem = elem ||
                           ((lo & __builtin_bitreverse32((uint32_t)em)) |
                            (hi & __builtin_bitreverse32(
                                      (uint32_t)(em >> 32)))) != 0;
                    L.bits[cb >> 5][lane] = lo;
                    if (nseg - cb > 32) L.bits[(cb >> 5) + 1][lane] = hi;
                    __builtin_amdgcn_wave_barrier();
                }
                nearw = nearw || elem;
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_s_waitcnt(0);
                const int nw = (int)((nseg + 31) >> 5);
                // ---- phase 2a: fluid candidates --------------------------
                {
                    int wp = 0;
                    uint32_t bw = L.bits[0][lane];
                    while (true) {
                        while (bw == 0 && wp + 1 < nw) {
                            wp++;
                            bw = L.bits[wp][lane];
                        }
                        if (bw == 0) break;
                        const int b = __builtin_clz(bw);
                        bw ^= 0x80000000u >> b;
                        const uint32_t cj = L.cand[wp * 32 + b];
                        if (cj & WC_ELEM) continue;
                        const uint32_t j = cj;
                        const f3 r_ij = sub<D>(ldv_o<D>(r, j), r_i);
                        const float s = dot<D>(r_ij, r_ij);
                        const float q = sqrtf(s) / H;
                        const float m_j = ldf_o(m, j);
                        const float rho_j = ldf_o(rho, j);
                        if (SHEP == 1)
                            sh += kernelW<D>(q) * CONW * m_j / rho_j;
                        if (i == j) continue;
                        const float p_j = ldf_o(p, j);
                        const f3 u_ij = sub<D>(ldv_o<D>(u, j), u_i);
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
                    }
                }
                // ---- phase 2b: wall elements (boundary-integral Shepard) --
                if (SHEP == 2 && __ballot(elem) != 0) {
                    int wp = 0;
                    uint32_t bw = elem ? L.bits[0][lane] : 0u;
                    while (elem) {
                        while (bw == 0 && wp + 1 < nw) {
                            wp++;
                            bw = L.bits[wp][lane];
                        }
                        if (bw == 0) break;
                        const int b = __builtin_clz(bw);
                        bw ^= 0x80000000u >> b;
                        const uint32_t cj = L.cand[wp * 32 + b];
                        if (!(cj & WC_ELEM)) continue;
                        const uint32_t j = cj & ~WC_ELEM;
                        const f3 r_ij = sub<D>(ldv_o<D>(r, j), r_i);
                        const float s = dot<D>(r_ij, r_ij);
                        const float q = sqrtf(s) / H;
                        const f3 n_j = ldv_o<D>(normal, j);
                        const f3 t_j = ldv_o<D>(tangent, j);
                        const f3 b_j = ldv_o<D>(binormal, j);
                        const float r_n = dot<D>(r_ij, n_j);
                        const float area_j = ldf_o(m, j);
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
                             aq_defs d, float smax, unsigned nblocks, int segw,
                             hipStream_t s)
{
    // candidate loads use 32-bit byte offsets (16 B per particle at most)
    AQ_REQUIRE(N < (1u << 28), "cell-parallel walk: more than 2^28 particles");
    WalkArgs w{ icell, ihoc, n_cells, cell_end, heads, n_heads, status, N };
#define AQ_FC(DD, SH, LP, SW)                                                  \
    hipLaunchKernelGGL((fluid_pass_cells_k<DD, SH, LP, SW>), dim3(nblocks),    \
                       dim3(64 * WC_WAVES), 0, s, imove, r, u, rho, m, p,      \
                       normal, tangent, binormal, shepard, grad_p, lap_u,      \
                       div_u, lap_p, near, w, d.H, d.CONW, d.CONF, d.DIMS,     \
                       smax, d.lap_formulation)
#define AQ_FC_D(SH, LP)                                                        \
    do {                                                                       \
        if (dims == 2 && segw == 8) AQ_FC(2, SH, LP, 8);                       \
        else if (dims == 2) AQ_FC(2, SH, LP, 16);                              \
        else if (segw == 8) AQ_FC(3, SH, LP, 8);                               \
        else AQ_FC(3, SH, LP, 16);                                             \
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
