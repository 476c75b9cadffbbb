// Cell-parallel neighbour walk with per-lane candidate bitsets.
//
// One wavefront owns one non-empty cell (64 of its particles at a time, one
// per lane); the candidates of the 27 (3D) / 9 (2D) neighbour cells are
// enumerated in the reference order (x offset outermost, then y, then z;
// sorted index inside a cell, types/3D.h:197-219), which is the same list
// for every particle of the cell.  The list is cut in segments of SEG
// candidates, tested into a ring of two LDS slots:
//
//   test     the lanes load 64 candidates at a time (coalesced) into LDS;
//            the candidate loop reads them back by broadcast and every lane
//            shifts in one accept bit per candidate (q <= 2 and the wanted
//            particle type);
//   consume  every lane walks its own accepted candidates in increasing
//            order and evaluates the pair terms.  A lane that finishes a
//            segment goes on with the next one while slower lanes finish
//            theirs; the oldest slot is refilled once no lane needs it.
//
// The fluid sums (grad p, lap u, div u, lap p, and the cfd Shepard sum) and
// the boundary-integral Shepard sum take their terms from disjoint candidate
// sets, so they are walked in two passes (the second only in wavefronts
// with a wall element in range), each keeping its terms and their order:
// every lane adds exactly what the reference kernels add, in the same
// order, and the results are bit-identical.  The pair arithmetic runs with
// the lanes that have work instead of under a mask of the lanes that
// accepted the current candidate.
#include <map>
#include <mutex>

#include "aq_device.hpp"
#include "aq_internal.hpp"

namespace aq {

constexpr int WC_WAVES = 4;  // wavefronts per workgroup

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
    uint32_t* next;  // work counter: next cell to hand out (zero at launch)
    const uint32_t* status;
    uint32_t N;
};

// Shared state of one wavefront: neighbour-cell table, the candidate chunk
// being tested and the two segment slots (candidate indices, accept bits;
// SEGW 32-bit words of bits per lane and slot).
template <int SEGW> struct WaveLds {
    uint32_t start[27];
    uint32_t cum[28];
    float4 stage[64];
    uint32_t cand[2][SEGW * 32];
    uint32_t bits[2][SEGW][64];
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

// Walk the accepted candidates of one particle per lane.  `kind` selects
// the candidates (0: fluid particles, imove == 1; 1: wall elements,
// imove == -3); with kind 0 and `flag_el`, wall elements in range are also
// tested and reported through `el_near`.  pair(j) is called for the
// accepted candidates of each lane in increasing candidate order.
template <int D, int SEGW, class Pair>
__device__ __forceinline__ void ring_walk(WaveLds<SEGW>& L, int lane,
                                          uint32_t total, const f3& r_i,
                                          bool act, float smax,
                                          const float* __restrict__ r,
                                          const int* __restrict__ imove,
                                          int kind, bool flag_el,
                                          bool& el_near, Pair pair)
{
    constexpr uint32_t SEG = SEGW * 32;
    const uint32_t nsegs = (total + SEG - 1) / SEG;
    const int nwords = (int)((total + 31) >> 5);
    int kl = 0;        // per-lane cursor: neighbour cell of a staged candidate
    int gw = -1;       // word being consumed (over the whole list)
    uint32_t bw = 0;   // its remaining bits
    uint32_t tested = 0;
    bool eln = false;
    while (true) {
        const int tw = min((int)(tested * SEGW), nwords);
        while (bw == 0 && gw + 1 < tw) {
            gw++;
            bw = L.bits[(gw / SEGW) & 1][gw % SEGW][lane];
        }
        // refill the oldest slot when no lane still reads from it
        if (tested < nsegs &&
            (tested < 2 ||
             __ballot(bw != 0 && gw < (int)((tested - 1) * SEGW)) == 0)) {
            const uint32_t seg = tested * SEG;
            const uint32_t nseg = min(SEG, total - seg);
            const int slot = tested & 1;
            for (uint32_t cb = 0; cb < nseg; cb += 64) {
                const uint32_t t = seg + cb + lane;
                float4 st = make_float4(0.f, 0.f, 0.f, 0.f);
                bool take = false, is_el = false;
                if (cb + lane < nseg) {
                    while (t >= L.cum[kl + 1]) kl++;
                    const uint32_t j = L.start[kl] + (t - L.cum[kl]);
                    const f3 rj = ldv_o<D>(r, j);
                    const int im_j = ldi_o(imove, j);
                    is_el = (im_j == -3);
                    take = kind ? is_el : (im_j == 1 || (flag_el && is_el));
                    st = make_float4(rj.x, rj.y, rj.z, 0.f);
                    L.cand[slot][cb + lane] = j;
                }
                const uint64_t tm = __ballot(take);
                const uint64_t em = __ballot(take && is_el);
                uint32_t lo = 0, hi = 0;
                if (tm != 0) {
                    L.stage[lane] = st;
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_s_waitcnt(0);
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
                            const f3 r_ij =
                                sub<D>(f3{ c.x, c.y, c.z }, r_i);
                            const float s = dot<D>(r_ij, r_ij);
                            hi = (hi << 1) | (!(s > smax) ? 1u : 0u);
                        }
                    }
                    lo &= __builtin_bitreverse32((uint32_t)tm);
                    hi &= __builtin_bitreverse32((uint32_t)(tm >> 32));
                    if (!act) lo = hi = 0;
                    if (!kind && em != 0) {
                        // wall elements in range: flagged, then dropped
                        const uint32_t el =
                            __builtin_bitreverse32((uint32_t)em);
                        const uint32_t eh =
                            __builtin_bitreverse32((uint32_t)(em >> 32));
                        eln = eln || ((lo & el) | (hi & eh)) != 0;
                        lo &= ~el;
                        hi &= ~eh;
                    }
                }
                L.bits[slot][cb >> 5][lane] = lo;
                if (nseg - cb > 32) L.bits[slot][(cb >> 5) + 1][lane] = hi;
                __builtin_amdgcn_wave_barrier();
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_s_waitcnt(0);
            tested++;
            continue;
        }
        if (__ballot(bw != 0) == 0) break;
        if (bw != 0) {
            const int b = __builtin_clz(bw);
            bw ^= 0x80000000u >> b;
            pair(L.cand[(gw / SEGW) & 1][(gw % SEGW) * 32 + b]);
        }
    }
    el_near = eln;
    __builtin_amdgcn_wave_barrier();
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
    constexpr float CLEARY = (D == 2) ? 8.f : 10.f;
    __shared__ WaveLds<SEGW> lds[WC_WAVES];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    WaveLds<SEGW>& L = lds[wv];
    if (w.status && (*w.status & 1u)) return;
    const uint32_t n_heads = uni(*w.n_heads);
    const float eps = 1E-8f * H;
    // wall elements in range are needed for the boundary-integral Shepard
    // sum and for the near-wall flags
    const bool flag_el = (SHEP == 2) || (near != nullptr);

    while (true) {
        // cells are handed out one at a time (a vector atomic from lane 0)
        uint32_t hk = 0;
        if (lane == 0) hk = atomicAdd(w.next, 1u);
        hk = uni(hk);
        if (hk >= n_heads) break;
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
            // ---- fluid neighbours ------------------------------------------
            bool nearw = false;
            ring_walk<D, SEGW>(
                L, lane, total, r_i, act, smax, r, imove, 0, flag_el, nearw,
                [&](uint32_t j) {
                    const f3 r_ij = sub<D>(ldv_o<D>(r, j), r_i);
                    const float s = dot<D>(r_ij, r_ij);
                    const float q = sqrtf(s) / H;
                    const float m_j = ldf_o(m, j);
                    const float rho_j = ldf_o(rho, j);
                    if (SHEP == 1) sh += kernelW<D>(q) * CONW * m_j / rho_j;
                    if (i == j) return;
                    const float p_j = ldf_o(p, j);
                    const f3 u_ij = sub<D>(ldv_o<D>(u, j), u_i);
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
                });
            // ---- wall elements (boundary-integral Shepard) ------------------
            if (SHEP == 2 && __ballot(nearw) != 0) {
                bool self_added = false;
                bool unused = false;
                ring_walk<D, SEGW>(
                    L, lane, total, r_i, act && nearw, smax, r, imove, 1,
                    false, unused, [&](uint32_t j) {
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
                            return;
                        }
                        const float r_t = fabsf(dot<D>(r_ij, t_j));
                        const float r_b = fabsf(dot<D>(r_ij, b_j));
                        sh += r_n * CONW * kernelS_P<D>(q) * area_j +
                              kernelS_D<D>(fabsf(r_n), r_t, r_b, area_j);
                    });
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

__global__ void zero_u32_k(uint32_t* p)
{
    if (threadIdx.x == 0) *p = 0u;
}

// Workgroups of `ker` that fit on the device at once (all CUs), cached per
// kernel: the walk hands cells out dynamically, so a grid of exactly that
// many workgroups keeps every slot busy until the cells run out.
static unsigned resident_blocks(const void* ker, int block)
{
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, unsigned> cache;
    int dev = 0;
    AQ_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    auto key = std::make_pair(dev, ker);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per_cu = 0, ncu = 0;
    AQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ker, block,
                                                        0));
    AQ_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                                 dev));
    const unsigned nb = (unsigned)std::max(1, per_cu) * (unsigned)std::max(1, ncu);
    cache[key] = nb;
    return nb;
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
    // the work counter lives next to the non-empty cell count
    uint32_t* next = const_cast<uint32_t*>(n_heads) + 1;
    hipLaunchKernelGGL(zero_u32_k, dim3(1), dim3(64), 0, s, next);
    WalkArgs w{ icell, ihoc,  n_cells, cell_end, heads,
                n_heads, next, status,  N };
#define AQ_FC(DD, SH, LP, SW)                                                  \
    do {                                                                       \
        auto* ker = fluid_pass_cells_k<DD, SH, LP, SW>;                        \
        const unsigned nb = nblocks ? nblocks : resident_blocks((const void*)ker, \
                                                    64 * WC_WAVES);            \
        hipLaunchKernelGGL(ker, dim3(nb), dim3(64 * WC_WAVES), 0, s, imove, r, \
                           u, rho, m, p, normal, tangent, binormal, shepard,   \
                           grad_p, lap_u, div_u, lap_p, near, w, d.H, d.CONW,  \
                           d.CONF, d.DIMS, smax, d.lap_formulation);           \
    } while (0)
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
