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
//            shifts in one accept bit per candidate (q < 2 and the wanted
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
#include <type_traits>

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
// a wave-uniform float, moved to a scalar register
__device__ __forceinline__ float uni_f(float x)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

// Cells are handed out in NRANGE contiguous ranges of the non-empty cell
// list, one per XCD: workgroups are dispatched to the XCDs round-robin, so
// workgroup b runs on XCD b % NRANGE and takes the cells of range b % NRANGE
// first; the neighbourhoods in flight on one XCD then lie in one slab of
// the domain and share that XCD's L2.  A workgroup whose range is done
// helps with the next ones.  Each range has its own counter, RSTRIDE words
// apart (one cache line each).
constexpr uint32_t NRANGE = 8;
constexpr uint32_t RSTRIDE = 32;

struct WalkArgs {
    const uint32_t* icell;
    const uint32_t* ihoc;
    const uint32_t* n_cells;
    const uint32_t* cell_end;
    const uint32_t* heads;
    const uint32_t* n_heads;
    uint32_t* next;  // range counters (zero at launch), RSTRIDE words apart
    const uint32_t* status;
    uint32_t N;
    uint32_t flagged;  // cell_end carries AQ_CELL_FLUID / AQ_CELL_ELEM flags
    uint32_t ranges;   // 1: one global counter; NRANGE: per-XCD ranges
    // periodic_L of a periodic box (the walks instantiated with PER wrap
    // their neighbour cells), nullptr otherwise
    const float* per = nullptr;
    // near-cell list the fused walk fills: (first, cell, cell_end)
    // triplets of the cells with a fluid particle near a wall element,
    // counted in n_heads[AQ_CL_NNCELLS]; nullptr: not built
    uint32_t* nlist = nullptr;
    // the non-empty cells in blocked order, as (first, cell, cell_end)
    // triplets (n_heads[0] of them), for the fused walk's hand-out;
    // nullptr: the row-major heads list
    const uint32_t* blist = nullptr;
};

// Next cell (index into heads) for the calling wave, n_heads when all are
// taken.  `q` is the wave's current range (starts at its XCD's).
__device__ __forceinline__ uint32_t next_cell(const WalkArgs& w,
                                              uint32_t n_heads, int lane,
                                              uint32_t& q)
{
    uint32_t hk = n_heads;
    if (lane == 0) {
        const uint32_t nr = w.ranges;
        const uint32_t home = q;
        while (q < home + nr) {
            const uint32_t k = q % nr;
            const uint32_t lo = (uint32_t)(((uint64_t)n_heads * k) / nr);
            const uint32_t hi = (uint32_t)(((uint64_t)n_heads * (k + 1)) / nr);
            const uint32_t t = atomicAdd(w.next + k * RSTRIDE, 1u);
            if (t < hi - lo) {
                hk = lo + t;
                break;
            }
            q++;
        }
    }
    return uni(hk);
}

// Shared state of one wavefront: neighbour-cell table, the candidate chunk
// being tested and the two segment slots (candidate indices, accept bits;
// SEGW 32-bit words of bits per lane and slot).
template <int SEGW, bool PER = false> struct WaveLds {
    uint32_t start[27];
    uint32_t cum[28];
    float4 stage[64];
    uint32_t cand[2][SEGW * 32];
    uint32_t bits[2][SEGW][64];
    // periodic box: the shift of each neighbour cell's particles
    float4 shift[PER ? 27 : 1];
};

// Candidate words of the periodic walks: the particle index in the low
// bits and the neighbour-table entry (0..26) in the top 5.
constexpr uint32_t PER_KL_SHIFT = 27;
constexpr uint32_t PER_J_MASK = (1u << PER_KL_SHIFT) - 1u;

// Candidate table of cell c_i: start and prefix count of each neighbour
// cell in walk order.  Cells known to hold none of the `want` kinds
// (AQ_CELL_FLUID / AQ_CELL_ELEM) count as empty: they have no candidate the
// caller accepts, so the sums are unchanged.
template <int D, bool PER = false, class Lds>
__device__ __forceinline__ uint32_t neighbour_table(const WalkArgs& w,
                                                    uint32_t c_i, Lds& L,
                                                    int lane, uint32_t want)
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
        uint32_t c;
        if constexpr (PER) {
            f3 sh;
            c = per_neighbour<D>(c_i, ci, cj, ck, w.n_cells, w.per, sh);
            L.shift[lane] = make_float4(sh.x, sh.y, sh.z, 0.f);
        } else {
            c = c_i + (uint32_t)ci + (uint32_t)cj * nx +
                (uint32_t)ck * nx * ny;
        }
        const uint32_t s = w.ihoc[c];
        uint32_t n = 0;
        if (s < w.N) {
            const uint32_t ce = w.cell_end[c];
            if (!w.flagged || (ce & want)) n = (ce & AQ_CELL_END) - s;
        }
        L.start[lane] = s;
        L.cum[lane + 1] = n;
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

// Bounding box of the active lanes' positions (uniform; +-inf if none)
template <int D>
__device__ __forceinline__ void wave_box(const f3& r_i, bool act, f3& bmin,
                                         f3& bmax)
{
    const float inf = __builtin_huge_valf();
    float lo[3] = { act ? r_i.x : inf, act ? r_i.y : inf,
                    act ? r_i.z : inf };
    float hi[3] = { act ? r_i.x : -inf, act ? r_i.y : -inf,
                    act ? r_i.z : -inf };
#pragma unroll
    for (int k = 0; k < D; k++) {
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], o));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], o));
        }
    }
    // uniform: held in scalar registers
    bmin = f3{ uni_f(lo[0]), uni_f(lo[1]), D == 3 ? uni_f(lo[2]) : 0.f };
    bmax = f3{ uni_f(hi[0]), uni_f(hi[1]), D == 3 ? uni_f(hi[2]) : 0.f };
}

// Walk the accepted candidates of one particle per lane.  `kind` selects
// the candidates (0: fluid particles, imove == 1; 1: wall elements,
// imove == -3); with kind 0 and `flag_el`, wall elements in range are also
// tested and reported through `el_near`.  pair(j) is called for the
// accepted candidates of each lane in increasing candidate order.
//
// (A branch-free variant, calling pair() on every lane at every step with a
// dummy index on the idle lanes, measured 25% slower on the 3D dam break:
// the idle lanes' loads and the longer dependent chains cost more than the
// exec-mask changes they save.)
//
// Box prefilter: a candidate farther than the support from the bounding box
// of the active lanes' positions is accepted by no lane, so it is not
// staged; the survivors of each 64-candidate chunk are compacted in their
// order (each lane's accepted candidates keep their order, so the sums are
// unchanged).  The box distance is computed like the lanes' own test, one
// rounded subtraction, square and sum per axis in the same order, and
// rounding is monotone, so it never exceeds a lane's s: the prefilter
// drops only candidates every lane would reject.
// pair(j, kl): kl is the neighbour-table entry the candidate came from
// (its shift in a periodic walk; 0 otherwise).
//
// NB > 1 (batched consume, not periodic): `pair` has load(j), returning
// the candidate's fields, and use(j, fields); each step takes up to NB
// accepted candidates of the lane's current bit word, issues all their
// loads, then uses them in increasing candidate order.  A lane's pair
// terms form a chain of dependent loads (next accepted candidate, its
// fields, the terms), so a walk with few lanes busy (wall cells) is bound
// by the memory latency of that chain; batching takes one latency per NB
// candidates instead of one per candidate, with the same terms in the same
// order.
//
// LOOSE (the fused fluid walk, not periodic): the per-lane test runs on
// coordinates relative to the box centre c, s' = |r_i - c|^2 + |r_j - c|^2
// - 2 (r_i - c).(r_j - c), with |r_j - c|^2 computed once per staged
// candidate: 5 operations per lane and candidate instead of 8.  s' is
// within 64 ulp of R^2 (R bounds |r - c| for every owner and survivor:
// the half diagonal of the box plus the support radius) of the reference's
// s, so the test s' <= smax + 2^-14 R^2 accepts every candidate the exact
// test accepts, and a few within a relative 2^-14 R^2 / smax beyond; the
// consume repeats the exact test (s <= smax) before adding a pair, so the
// sums are unchanged.  The staged record's w carries |r_j - c|^2 with the
// sign bit set for wall elements.
template <int D, int SEGW, int NB = 1, bool LOOSE = false, bool PER,
          class Pair>
__device__ __forceinline__ void ring_walk(WaveLds<SEGW, PER>& L, int lane,
                                          uint32_t total, const f3& r_i,
                                          bool act, const f3& bmin,
                                          const f3& bmax, float smax,
                                          const float* __restrict__ r,
                                          const int* __restrict__ imove,
                                          int kind, bool flag_el,
                                          bool& el_near, Pair pair,
                                          const float* __restrict__ per)
{
    constexpr uint32_t SEG = SEGW * 32;
    const uint32_t nsegs = (total + SEG - 1) / SEG;
    const uint32_t nwords = (total + 31) >> 5;
    // the two slots form a ring of 2 SEGW words: word w of the candidate
    // list sits at ring position w % (2 SEGW)
    constexpr uint32_t RING = 2 * SEGW;
    const uint32_t* bits = &L.bits[0][0][0];
    const uint32_t* cand = &L.cand[0][0];
    static_assert(!(LOOSE && PER), "relative test: not for periodic walks");
    // LOOSE: the reference point, the owner's relative position and the
    // widened threshold
    f3 cen = { 0.f, 0.f, 0.f }, rio = { 0.f, 0.f, 0.f };
    float ai = 0.f, smaxl = smax;
    if constexpr (LOOSE) {
        if (bmin.x <= bmax.x) {
            cen = f3{ uni_f(0.5f * (bmin.x + bmax.x)),
                      uni_f(0.5f * (bmin.y + bmax.y)),
                      D == 3 ? uni_f(0.5f * (bmin.z + bmax.z)) : 0.f };
            const f3 hd = sub<D>(bmax, cen);
            const float rb = sqrtf(dot<D>(hd, hd)) * 1.0001f + sqrtf(smax);
            smaxl = uni_f(smax + rb * rb * 0x1p-14f);
        }
        rio = sub<D>(r_i, cen);
        ai = dot<D>(rio, rio);
    }
    int kl = 0;          // per-lane cursor: neighbour cell of a staged candidate
    uint32_t nw = 0;     // next word to load (the current one is nw - 1)
    uint32_t bw = 0;     // remaining bits of the current word
    uint32_t tested = 0;
    bool eln = false;
    auto advance = [&]() {
        const uint32_t tw = min(tested * SEGW, nwords);
        while (bw == 0 && nw < tw) {
            bw = bits[(nw % RING) * 64 + lane];
            nw++;
        }
    };
    // consume while no slot can be refilled (a lane still reads from the
    // oldest) and some lane has a candidate left
    auto go = [&]() {
        advance();
        if (tested < nsegs &&
            (tested < 2 ||
             __ballot(bw != 0 && nw <= (tested - 1) * SEGW) == 0))
            return false;
        return __ballot(bw != 0) != 0;
    };
    while (true) {
        // the test sits at the bottom of the consume loop, so the pair sums
        // leave it from where they are updated (no copies at the back edge)
        for (bool more = go(); more; more = go()) {
            if constexpr (NB == 1) {
                if (bw != 0) {
                    const uint32_t b = __builtin_clz(bw);
                    bw ^= 0x80000000u >> b;
                    const uint32_t cw = cand[((nw - 1) % RING) * 32 + b];
                    if constexpr (PER)
                        pair(cw & PER_J_MASK, cw >> PER_KL_SHIFT);
                    else
                        pair(cw, 0u);
                }
            } else {
                static_assert(!PER, "batched consume: not for periodic walks");
                uint32_t jb[NB];
                int nb = 0;
#pragma unroll
                for (int k = 0; k < NB; k++) {
                    jb[k] = 0;
                    if (bw != 0) {
                        const uint32_t b = __builtin_clz(bw);
                        bw ^= 0x80000000u >> b;
                        jb[k] = cand[((nw - 1) % RING) * 32 + b];
                        nb = k + 1;
                    }
                }
                decltype(pair.load(0u)) cd[NB];
#pragma unroll
                for (int k = 0; k < NB; k++)
                    if (k < nb) cd[k] = pair.load(jb[k]);
#pragma unroll
                for (int k = 0; k < NB; k++)
                    if (k < nb) pair.use(jb[k], cd[k]);
            }
        }
        if (tested >= nsegs) break;  // and no lane has a candidate left
        // refill the oldest slot: chunk(cb, ...) tests the candidates
        // cb .. cb + 63 of the segment (lane: candidate cb + lane, when
        // `have`, at j, rj, of kind im_j)
        const uint32_t seg = tested * SEG;
        const uint32_t nseg = min(SEG, total - seg);
        const int slot = tested & 1;
        auto chunk = [&](uint32_t cb, bool have, uint32_t j, f3 rj,
                         int im_j) {
            float4 st = make_float4(0.f, 0.f, 0.f, 0.f);
            bool take = false, is_el = false;
            if (have) {
                is_el = (im_j == -3);
                take = kind ? is_el : (im_j == 1 || (flag_el && is_el));
                if (take) {
                    const float dx =
                        fmaxf(fmaxf(bmin.x - rj.x, rj.x - bmax.x), 0.f);
                    const float dy =
                        fmaxf(fmaxf(bmin.y - rj.y, rj.y - bmax.y), 0.f);
                    float d2 = dx * dx;
                    d2 = d2 + dy * dy;
                    if constexpr (D == 3) {
                        const float dz =
                            fmaxf(fmaxf(bmin.z - rj.z, rj.z - bmax.z), 0.f);
                        d2 = d2 + dz * dz;
                    }
                    take = !(d2 > smax);
                }
                if constexpr (LOOSE) {
                    const f3 rjo = sub<D>(rj, cen);
                    float bj = rjo.x * rjo.x;
                    bj = __builtin_fmaf(rjo.y, rjo.y, bj);
                    if constexpr (D == 3) bj = __builtin_fmaf(rjo.z, rjo.z, bj);
                    st = make_float4(rjo.x, rjo.y, rjo.z, is_el ? -bj : bj);
                } else {
                    st = make_float4(rj.x, rj.y, rj.z, is_el ? 1.f : 0.f);
                }
            }
            // the per-lane test of a staged candidate
            auto accept = [&](const float4& c) {
                if constexpr (LOOSE) {
                    float t = rio.x * c.x;
                    t = __builtin_fmaf(rio.y, c.y, t);
                    if constexpr (D == 3) t = __builtin_fmaf(rio.z, c.z, t);
                    const float s = __builtin_fmaf(-2.f, t, ai + fabsf(c.w));
                    return !(s > smaxl);
                } else {
                    const f3 r_ij = sub<D>(f3{ c.x, c.y, c.z }, r_i);
                    const float s = dot<D>(r_ij, r_ij);
                    return !(s > smax);
                }
            };
            const uint64_t tm = __ballot(take);
            const uint32_t cnt = (uint32_t)__popcll(tm);
            uint32_t lo = 0, hi = 0;
            if (cnt != 0) {
                // compact the survivors into stage[0 .. cnt), in order
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(tm >> 32),
                    __builtin_amdgcn_mbcnt_lo((uint32_t)tm, 0u));
                if (take) {
                    L.stage[rank] = st;
                    L.cand[slot][cb + rank] =
                        PER ? (j | ((uint32_t)kl << PER_KL_SHIFT)) : j;
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_s_waitcnt(0);
                const uint32_t c1 = cnt < 32u ? cnt : 32u;
#pragma unroll 8
                for (uint32_t k = 0; k < c1; k++)
                    lo = (lo << 1) | (accept(L.stage[k]) ? 1u : 0u);
                if (c1 < 32u) lo <<= 32u - c1;
                if (cnt > 32u) {
#pragma unroll 8
                    for (uint32_t k = 32; k < cnt; k++)
                        hi = (hi << 1) | (accept(L.stage[k]) ? 1u : 0u);
                    if (cnt < 64u) hi <<= 64u - cnt;
                }
                if (!act) lo = hi = 0;
                if (!kind && __ballot(take && is_el) != 0) {
                    // wall elements in range: flagged, then dropped
                    uint32_t el = 0, eh = 0;
                    for (uint32_t k = 0; k < cnt; k++) {
                        const float w = L.stage[k].w;
                        if (LOOSE ? !(__float_as_uint(w) >> 31) : w == 0.f)
                            continue;
                        if (k < 32u) el |= 0x80000000u >> k;
                        else eh |= 0x80000000u >> (k - 32u);
                    }
                    eln = eln || ((lo & el) | (hi & eh)) != 0;
                    lo &= ~el;
                    hi &= ~eh;
                }
            }
            L.bits[slot][cb >> 5][lane] = lo;
            if (nseg - cb > 32) L.bits[slot][(cb >> 5) + 1][lane] = hi;
            __builtin_amdgcn_wave_barrier();
        };
        if constexpr (NB > 1) {
            // the whole segment's candidate loads issued together
            constexpr uint32_t NCH = SEG / 64;
            uint32_t pj[NCH];
            f3 prj[NCH];
            int pim[NCH];
#pragma unroll
            for (uint32_t ch = 0; ch < NCH; ch++) {
                pj[ch] = 0;
                const uint32_t t = seg + ch * 64 + lane;
                if (ch * 64 + lane < nseg) {
                    while (t >= L.cum[kl + 1]) kl++;
                    pj[ch] = L.start[kl] + (t - L.cum[kl]);
                }
            }
#pragma unroll
            for (uint32_t ch = 0; ch < NCH; ch++) {
                prj[ch] = f3{ 0.f, 0.f, 0.f };
                pim[ch] = 0;
                if (ch * 64 + lane < nseg) {
                    prj[ch] = ldv_o<D>(r, pj[ch]);
                    pim[ch] = ldi_o(imove, pj[ch]);
                }
            }
#pragma unroll
            for (uint32_t ch = 0; ch < NCH; ch++) {
                if (ch * 64 >= nseg) break;
                chunk(ch * 64, ch * 64 + lane < nseg, pj[ch], prj[ch],
                      pim[ch]);
            }
        } else {
            for (uint32_t cb = 0; cb < nseg; cb += 64) {
                const uint32_t t = seg + cb + lane;
                uint32_t j = 0;
                f3 rj = { 0.f, 0.f, 0.f };
                int im_j = 0;
                const bool have = cb + lane < nseg;
                if (have) {
                    while (t >= L.cum[kl + 1]) kl++;
                    j = L.start[kl] + (t - L.cum[kl]);
                    rj = ldv_o<D>(r, j);
                    if constexpr (PER) {
                        const float4 sh = L.shift[kl];
                        rj = per_shift<D>(rj, f3{ sh.x, sh.y, sh.z }, per);
                    }
                    im_j = ldi_o(imove, j);
                }
                chunk(cb, have, j, rj, im_j);
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_s_waitcnt(0);
        tested++;
    }
    el_near = eln;
    __builtin_amdgcn_wave_barrier();
}

// The fused walk's form of ring_walk (one candidate per consume step).
template <int D, int SEGW, bool LOOSE = false, bool PER, class Pair>
__device__ __forceinline__ void ring_walk_l(WaveLds<SEGW, PER>& L, int lane,
                                            uint32_t total, const f3& r_i,
                                            bool act, const f3& bmin,
                                            const f3& bmax, float smax,
                                            const float* __restrict__ r,
                                            const int* __restrict__ imove,
                                            int kind, bool flag_el,
                                            bool& el_near, Pair pair,
                                            const float* __restrict__ per)
{
    ring_walk<D, SEGW, 1, LOOSE>(L, lane, total, r_i, act, bmin, bmax, smax,
                                 r, imove, kind, flag_el, el_near, pair, per);
}
// ---------------------------------------------------------------------------
// Fused fluid walk (same terms as fluid_pass_k in sph.hip).
//   MATH 0: IEEE-754 terms and sums, bit-identical to the oracle;
//   MATH 1: OpenCL relaxed-math terms (reciprocals, hardware sqrt, fma);
//   MATH 2: the IEEE terms with compensated (Neumaier) sums for the
//           momentum (grad_p, lap_u) and density (div_u, lap_p) rates.
// ---------------------------------------------------------------------------
// s + x with the rounding error of every addition collected in c (Kahan-
// Babuska / Neumaier: the larger operand's low bits are the ones lost)
__device__ __forceinline__ void nsum(float& s, float& c, float x)
{
    const float t = s + x;
    const bool big = fabsf(s) >= fabsf(x);
    const float lo = big ? ((s - t) + x) : ((x - t) + s);
    c += lo;
    s = t;
}

template <int D>
__device__ __forceinline__ void naxpy(f3& acc, f3& c, float s, const f3& v)
{
    nsum(acc.x, c.x, s * v.x);
    nsum(acc.y, c.y, s * v.y);
    if (D == 3) nsum(acc.z, c.z, s * v.z);
}

// ---------------------------------------------------------------------------
// Range record of the fluid fields, for the division and square root path
// without range scaling (div_nr, sqrt_nr): 7 words from RREC words after
// the non-empty cell count on, RSTRIDE words apart (one cache line each, so
// that the workgroups' atomics on different words proceed in parallel), bit
// patterns of non-negative floats compared as
// integers (a negative rho or m, an infinity or a NaN lands above every
// finite positive pattern): min |p| over p != 0, max |p|, min and max rho,
// min and max m, max |u| component over the fluid particles.
// ---------------------------------------------------------------------------
constexpr uint32_t RREC = 288;
// the cell list's count buffer holds 512 words (runtime/tools.cpp)
static_assert(RREC + 6 * RSTRIDE < 512, "range record beyond the count buffer");
constexpr uint32_t NREC = 7;
__constant__ const uint32_t rec_init[NREC] = { 0x7F800000u, 0u, 0xFFFFFFFFu,
                                               0u, 0xFFFFFFFFu, 0u, 0u };

__device__ __forceinline__ uint32_t wave_min(uint32_t x)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
        x = min(x, (uint32_t)__shfl_xor((int)x, o));
    return x;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t x)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
        x = max(x, (uint32_t)__shfl_xor((int)x, o));
    return x;
}

template <int D>
__global__ void __launch_bounds__(256)
    fluid_ranges_k(const int* __restrict__ imove, const float* __restrict__ p,
                   const float* __restrict__ rho, const float* __restrict__ m,
                   const float* __restrict__ u, uint32_t N,
                   uint32_t* __restrict__ rec)
{
    uint32_t v[NREC];
#pragma unroll
    for (uint32_t k = 0; k < NREC; k++) v[k] = rec_init[k];
    // four particles per thread and pass, every load issued before the
    // fluid tests
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < N;
         i0 += 4 * stride) {
        int im[4];
        uint32_t pa[4], rb[4], mb[4];
        f3 uu[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const uint32_t i = i0 + t * stride;
            im[t] = 0;
            if (i < N) {
                im[t] = imove[i];
                pa[t] = __float_as_uint(p[i]) & 0x7FFFFFFFu;
                rb[t] = __float_as_uint(rho[i]);
                mb[t] = __float_as_uint(m[i]);
                uu[t] = ldv<D>(u, i);
            }
        }
#pragma unroll
        for (int t = 0; t < 4; t++) {
            if (im[t] != 1) continue;
            const uint32_t pa_ = pa[t], rb_ = rb[t], mb_ = mb[t];
            const f3 ui = uu[t];
            if (pa_) v[0] = min(v[0], pa_);
            v[1] = max(v[1], pa_);
            v[2] = min(v[2], rb_);
            v[3] = max(v[3], rb_);
            v[4] = min(v[4], mb_);
            v[5] = max(v[5], mb_);
            uint32_t ua = max(__float_as_uint(ui.x) & 0x7FFFFFFFu,
                              __float_as_uint(ui.y) & 0x7FFFFFFFu);
            if constexpr (D == 3) ua = max(ua, __float_as_uint(ui.z) & 0x7FFFFFFFu);
            v[6] = max(v[6], ua);
        }
    }
    __shared__ uint32_t part[4][NREC];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (uint32_t k = 0; k < NREC; k++) {
        const bool is_min = (k == 0 || k == 2 || k == 4);
        const uint32_t x = is_min ? wave_min(v[k]) : wave_max(v[k]);
        if (lane == 0) part[wv][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < NREC) {
        const uint32_t k = threadIdx.x;
        const bool is_min = (k == 0 || k == 2 || k == 4);
        uint32_t x = part[0][k];
        for (int w2 = 1; w2 < 4; w2++)
            x = is_min ? min(x, part[w2][k]) : max(x, part[w2][k]);
        if (is_min) atomicMin(rec + k * RSTRIDE, x);
        else atomicMax(rec + k * RSTRIDE, x);
    }
}

// The threshold T (a power of two, 0 when the path is not usable) such
// that every division of the fused walk's pair terms meets div_nr's
// conditions (b in [2^-100, 2^100]; a == 0 or |a| >= 2^-100 and a / b in
// [2^-120, 2^90]) for all fluid pairs with s == 0 or s >= 2^-96 and
// udr == 0 or |udr| >= T (both tested per pair), from the range record.
// Each operand's bounds follow its expression with a factor 2 of margin
// for the roundings:
//   sqrt(s) / H          sqrt(s) in {0} U [2^-48, sqrt(smax)]
//   f_ij = kF(q) CONF m_j, kF = wcon a^3, a = 2 - q: a == 0 or
//                        |a| in [2^-23, 2] (a nonzero difference of 2 and
//                        a float in [0, 4) is a multiple of 2^-23)
//   (p_i + p_j) / (rho_i rho_j)   a nonzero sum of two floats of magnitude
//                                 >= Pmin is >= Pmin 2^-24
//   f_ij CLEARY udr / (r2 rho_i rho_j), r2 = (q^2 + 0.01) H^2   (Monaghan)
//   2 f_ij / (rho_i rho_j)                                     (Morris)
//   udr f_ij rho_i / rho_j,  f_ij / rho_j,  |udr| <= 3 (2 Umax) sqrt(smax)
// A zero numerator (f_ij == 0 at q == 2, udr == 0, p_i + p_j == 0) needs
// no bound.
template <int D, int LAPF, bool LAPP>
__device__ float fast_threshold(const uint32_t* __restrict__ rec, float H,
                                float CONF, float smax)
{
    uint32_t b[NREC];
#pragma unroll
    for (uint32_t k = 0; k < NREC; k++) b[k] = uni(rec[k * RSTRIDE]);
    if (b[1] >= 0x7F800000u || b[3] >= 0x7F800000u ||
        b[5] >= 0x7F800000u || b[6] >= 0x7F800000u)
        return 0.f;  // an infinity, a NaN, a negative rho or m
    if (b[2] > b[3] || b[4] > b[5]) return 0.f;  // no fluid particle
    const double Pmin = __uint_as_float(b[0]), Pmax = __uint_as_float(b[1]);
    const double Rmin = __uint_as_float(b[2]), Rmax = __uint_as_float(b[3]);
    const double Mmin = __uint_as_float(b[4]), Mmax = __uint_as_float(b[5]);
    const double Umax = __uint_as_float(b[6]);
    const double h = H, conf = CONF;
    if (!(h > 0.0) || !(conf > 0.0) || !(smax > 0.f) || !(Mmin > 0.0) ||
        !(Rmin > 0.0))
        return 0.f;
    auto win = [](double nlo, double nhi, double dlo, double dhi) {
        return dlo >= 0x1p-100 && dhi <= 0x1p100 && nlo >= 0x1p-100 &&
               nhi <= 0x1p100 && nlo / dhi >= 0x1p-120 &&
               nhi / dlo <= 0x1p90;
    };
    constexpr double CLEARY = (D == 2) ? 8.0 : 10.0;
    const double wcon = ((D == 2) ? 1.09375 : 0.8203125) * 0.318309886;
    const double rs = 2.0 * sqrt((double)smax);
    const double flo = wcon * 0x1p-69 * conf * Mmin * 0.5;
    const double fhi = wcon * 8.0 * conf * Mmax * 2.0;
    const double ud = 12.0 * Umax * rs;
    const double rrlo = Rmin * Rmin * 0.5, rrhi = Rmax * Rmax * 2.0;
    // q <= sqrt(smax) / H, about 2: r2 <= 4.2 H^2
    const double slo = 0.01 * h * h * 0.5 * rrlo, shi = 4.2 * h * h * 2.0 * rrhi;
    // smallest T for the two terms with udr in the numerator
    double need = fmax(0x1p-100 / (flo * Rmin * 0.5),
                       0x1p-120 * Rmax * 2.0 / (flo * Rmin * 0.5));
    if constexpr (LAPF == AQ_LAP_MONAGHAN)
        need = fmax(need, fmax(0x1p-100 / (flo * CLEARY * 0.5),
                               0x1p-120 * shi / (flo * CLEARY * 0.5)));
    if (!(need <= 0x1p-20)) return 0.f;  // not worth the per-pair fall-backs
    int e = 0;
    frexp(need, &e);
    const double T = ldexp(1.0, e);  // a power of two >= need
    bool ok = win(0x1p-49, rs, h, h);
    if (b[0] != 0x7F800000u)  // some p != 0
        ok = ok && win(Pmin * 0x1p-25, 4.0 * Pmax, rrlo, rrhi);
    if constexpr (LAPF == AQ_LAP_MONAGHAN)
        ok = ok && win(flo * CLEARY * T * 0.5, fhi * CLEARY * ud * 2.0, slo,
                       shi);
    else
        ok = ok && win(flo, 4.0 * fhi, rrlo, rrhi);
    ok = ok && win(T * flo * Rmin * 0.5, ud * fhi * Rmax * 2.0, Rmin * 0.5,
                   Rmax * 2.0);
    if (LAPP) ok = ok && win(flo, fhi, Rmin * 0.5, Rmax * 2.0);
    return ok ? (float)T : 0.f;
}

template <int D, int SHEP, bool LAPP, int SEGW, int LAPF, int MATH,
          bool PER = false>
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
                       float smax, float iH, const uint32_t* __restrict__ rec)
{
    constexpr float CLEARY = (D == 2) ? 8.f : 10.f;
    __shared__ WaveLds<SEGW, PER> lds[WC_WAVES];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    WaveLds<SEGW, PER>& L = lds[wv];
    // a candidate's position, at its nearest image in a periodic box
    auto rjof = [&](uint32_t j, uint32_t kl) {
        f3 v = ldv_o<D>(r, j);
        if constexpr (PER) {
            const float4 sh = L.shift[kl];
            v = per_shift<D>(v, f3{ sh.x, sh.y, sh.z }, w.per);
        }
        return v;
    };
    if (w.status && (*w.status & 1u)) return;
    const uint32_t n_heads = uni(*w.n_heads);
    const float eps = 1E-8f * H;
    // wall elements in range are needed for the boundary-integral Shepard
    // sum and for the near-wall flags
    const bool flag_el = (SHEP == 2) || (near != nullptr);
    // divisions and square roots without range scaling (div_nr, sqrt_nr)
    // when the range record allows them; IEEE sequences otherwise
    float fdT = 0.f;
    if constexpr (MATH != 1)
        if (rec) fdT = fast_threshold<D, LAPF, LAPP>(rec, H, CONF, smax);
    const bool fastdiv = fdT > 0.f;
    const Rcp RH = rcp_nr(H);

    uint32_t range = blockIdx.x % w.ranges;
    while (true) {
        // cells are handed out one at a time (a vector atomic from lane 0)
        const uint32_t hk = next_cell(w, n_heads, lane, range);
        if (hk >= n_heads) break;
        uint32_t cstart, c_i, ce;
        if (w.blist) {
            cstart = uni(w.blist[3 * hk]);
            c_i = uni(w.blist[3 * hk + 1]);
            ce = uni(w.blist[3 * hk + 2]);
        } else {
            cstart = uni(w.heads[hk]);
            c_i = uni(w.icell[cstart]);
            ce = uni(w.cell_end[c_i]);
        }
        const uint32_t cend = ce & AQ_CELL_END;
        if (w.flagged && !(ce & AQ_CELL_FLUID)) continue;  // no fluid here
        bool cell_near = false;  // some fluid particle of the cell is near
        for (uint32_t i0 = cstart; i0 < cend; i0 += 64) {
            const uint32_t i = i0 + lane;
            const bool act = (i < cend) && (imove[i] == 1);
            if (__ballot(act) == 0) continue;  // no fluid particle here
            const uint32_t total = neighbour_table<D, PER>(
                w, c_i, L, lane,
                AQ_CELL_FLUID | (flag_el ? AQ_CELL_ELEM : 0u));
            f3 r_i = { 0.f, 0.f, 0.f }, u_i = { 0.f, 0.f, 0.f };
            float p_i = 0.f, rho_i = 1.f;
            f3 gp = { 0.f, 0.f, 0.f }, lu = { 0.f, 0.f, 0.f };
            float dv = 0.f, lp = 0.f, sh = 0.f;
            // compensations (MATH 2)
            f3 cgp = { 0.f, 0.f, 0.f }, clu = { 0.f, 0.f, 0.f };
            float cdv = 0.f, clp = 0.f;
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
            constexpr bool FAST = MATH == 1;
            const float irho_i = FAST ? __builtin_amdgcn_rcpf(rho_i) : 0.f;
            // ---- fluid neighbours ------------------------------------------
            bool nearw = false;
            f3 bmin, bmax;
            wave_box<D>(r_i, act, bmin, bmax);
            // pair terms; FD selects div_nr for the divisions
            auto terms = [&](auto fd, const f3& r_ij, float q, float m_j,
                             float rho_j, float p_j, const f3& u_ij,
                             float udr) {
                constexpr bool FD = decltype(fd)::value;
                if (SHEP == 1) sh += kernelW<D>(q) * CONW * m_j / rho_j;
                // The reference skips j == i here.  For i itself r_ij,
                // u_ij and p_j - p_i are exact zeros, so every term below
                // is +-0 and leaves the accumulators unchanged (they are
                // never -0: they start at +0 and an exact cancellation
                // rounds to +0); with finite fields the sums are
                // bit-identical without the test.
                const float fk = kernelF<D>(q) * CONF;
                const float f_ij = fk * m_j;
                float sg, sl, td, tl = 0.f;
                if constexpr (FD) {
                    const Rcp Rr = rcp_nr(rho_i * rho_j);
                    const Rcp Rj = rcp_nr(rho_j);
                    sg = div_nr(p_i + p_j, Rr) * f_ij;
                    if constexpr (LAPF == AQ_LAP_MONAGHAN) {
                        const float r2 = (q * q + 0.01f) * H * H;
                        sl = div_nr(f_ij * CLEARY * udr,
                                    rcp_nr(r2 * rho_i * rho_j));
                    } else {
                        sl = div_nr(f_ij * 2.f, Rr);
                    }
                    td = div_nr(udr * f_ij * rho_i, Rj);
                    if (LAPP) tl = (p_j - p_i) * div_nr(f_ij, Rj);
                } else {
                    sg = (p_i + p_j) / (rho_i * rho_j) * f_ij;
                    if constexpr (LAPF == AQ_LAP_MONAGHAN) {
                        const float r2 = (q * q + 0.01f) * H * H;
                        sl = f_ij * CLEARY * udr / (r2 * rho_i * rho_j);
                    } else {
                        sl = f_ij * 2.f / (rho_i * rho_j);
                    }
                    td = udr * f_ij * rho_i / rho_j;
                    if (LAPP) tl = (p_j - p_i) * (f_ij / rho_j);
                }
                if constexpr (MATH == 2) {
                    naxpy<D>(gp, cgp, sg, r_ij);
                    naxpy<D>(lu, clu, sl,
                             LAPF == AQ_LAP_MONAGHAN ? r_ij : u_ij);
                    nsum(dv, cdv, td);
                    if (LAPP) nsum(lp, clp, tl);
                } else {
                    axpy<D>(gp, sg, r_ij);
                    axpy<D>(lu, sl, LAPF == AQ_LAP_MONAGHAN ? r_ij : u_ij);
                    dv += td;
                    if (LAPP) lp += tl;
                }
            };
            // div_nr / sqrt_nr where the per-pair bounds hold (s == 0 or
            // s >= 2^-96, udr == 0 or |udr| >= T), the IEEE sequences for
            // the other pairs
            // The relative-coordinate test (LOOSE) accepts a thin shell
            // beyond the support.  Such a candidate is given zero mass here
            // (the reference's test, s <= smax, as a select rather than a
            // branch, so every load of the pair is issued at once): its
            // terms are then +-0 (finite fields), and adding +-0 leaves the
            // sums unchanged, since they are never -0 (they start at +0, or
            // 1, and an exact cancellation rounds to +0).
            constexpr bool LOOSE = (MATH != 1) && !PER;
            auto pair_fd = [&](uint32_t j, uint32_t kl) {
                const f3 r_ij = sub<D>(rjof(j, kl), r_i);
                const float s = dot<D>(r_ij, r_ij);
                const float qf = div_nr(sqrt_nr(s), RH);
                float m_j = ldf_o(m, j);
                if (LOOSE) m_j = (s > smax) ? 0.f : m_j;
                const float rho_j = ldf_o(rho, j);
                const float p_j = ldf_o(p, j);
                const f3 u_ij = sub<D>(ldv_o<D>(u, j), u_i);
                const float udr = dot<D>(u_ij, r_ij);
                const bool bad = ((s < 0x1p-96f) & (s != 0.f)) |
                                 ((udr != 0.f) & (fabsf(udr) < fdT));
                if (!bad)
                    terms(std::true_type{}, r_ij, qf, m_j, rho_j, p_j, u_ij,
                          udr);
                else
                    terms(std::false_type{}, r_ij, sqrtf(s) / H, m_j, rho_j,
                          p_j, u_ij, udr);
            };
            auto pair_ieee = [&](uint32_t j, uint32_t kl) {
                    if constexpr (FAST) {
                        // OpenCL C default accuracy: reciprocal-based
                        // division, hardware square root, contracted
                        // multiply-adds
                        const f3 r_ij = sub<D>(rjof(j, kl), r_i);
                        const float s = fdot<D>(r_ij, r_ij);
                        const float q = __builtin_amdgcn_sqrtf(s) * iH;
                        const float m_j = ldf_o(m, j);
                        const float irho_j =
                            __builtin_amdgcn_rcpf(ldf_o(rho, j));
                        if (SHEP == 1)
                            sh = fmaf(kernelW<D>(q) * CONW * m_j, irho_j, sh);
                        const float p_j = ldf_o(p, j);
                        const f3 u_ij = sub<D>(ldv_o<D>(u, j), u_i);
                        const float udr = fdot<D>(u_ij, r_ij);
                        const float f_ij = kernelF<D>(q) * CONF * m_j;
                        const float fr = f_ij * irho_i * irho_j;
                        faxpy<D>(gp, (p_i + p_j) * fr, r_ij);
                        if constexpr (LAPF == AQ_LAP_MONAGHAN) {
                            const float r2 = fmaf(q, q, 0.01f) * (H * H);
                            faxpy<D>(lu,
                                     CLEARY * udr * fr *
                                         __builtin_amdgcn_rcpf(r2),
                                     r_ij);
                        } else {
                            faxpy<D>(lu, 2.f * fr, u_ij);
                        }
                        dv = fmaf(udr * f_ij * rho_i, irho_j, dv);
                        if (LAPP) lp = fmaf(p_j - p_i, f_ij * irho_j, lp);
                        return;
                    }
                    const f3 r_ij = sub<D>(rjof(j, kl), r_i);
                    const float s = dot<D>(r_ij, r_ij);
                    const float q = sqrtf(s) / H;
                    float m_j = ldf_o(m, j);
                    if (LOOSE) m_j = (s > smax) ? 0.f : m_j;
                    const float rho_j = ldf_o(rho, j);
                    const float p_j = ldf_o(p, j);
                    const f3 u_ij = sub<D>(ldv_o<D>(u, j), u_i);
                    const float udr = dot<D>(u_ij, r_ij);
                    terms(std::false_type{}, r_ij, q, m_j, rho_j, p_j, u_ij,
                          udr);
            };
            if constexpr (MATH != 1) {
                if (fastdiv) {
                    ring_walk_l<D, SEGW, LOOSE>(
                        L, lane, total, r_i, act, bmin, bmax, smax, r, imove,
                        0, flag_el, nearw, pair_fd, w.per);
                } else {
                    ring_walk_l<D, SEGW, LOOSE>(
                        L, lane, total, r_i, act, bmin, bmax, smax, r, imove,
                        0, flag_el, nearw, pair_ieee, w.per);
                }
            } else {
                ring_walk_l<D, SEGW, false>(L, lane, total, r_i, act, bmin,
                                           bmax, smax, r, imove, 0, flag_el,
                                           nearw, pair_ieee, w.per);
            }
            // ---- wall elements (boundary-integral Shepard) ------------------
            if (SHEP == 2 && __ballot(nearw) != 0) {
                bool self_added = false;
                bool unused = false;
                const uint32_t total_el =
                    neighbour_table<D, PER>(w, c_i, L, lane, AQ_CELL_ELEM);
                ring_walk_l<D, SEGW, LOOSE>(
                    L, lane, total_el, r_i, act && nearw, bmin, bmax, smax,
                    r, imove, 1, false, unused, [&](uint32_t j, uint32_t kl) {
                        const f3 r_ij = sub<D>(rjof(j, kl), r_i);
                        const float s = dot<D>(r_ij, r_ij);
                        if (LOOSE && s > smax) return;
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
                    },
                    w.per);
            }
            if (act) {
                if constexpr (MATH == 2) {
                    gp = { gp.x + cgp.x, gp.y + cgp.y, gp.z + cgp.z };
                    lu = { lu.x + clu.x, lu.y + clu.y, lu.z + clu.z };
                    dv += cdv;
                    lp += clp;
                }
                stxyz<D>(grad_p, i, gp);
                stxyz<D>(lap_u, i, lu);
                div_u[i] = dv;
                if (LAPP) lap_p[i] = lp;
                if (SHEP) shepard[i] = sh;
                if (near) near[i] = nearw ? 1 : 0;
            }
            if (w.nlist) cell_near = cell_near || __ballot(act && nearw) != 0;
            __builtin_amdgcn_wave_barrier();
        }
        if (w.nlist && cell_near && lane == 0) {
            uint32_t* cnt = const_cast<uint32_t*>(w.n_heads) + AQ_CL_NNCELLS;
            const uint32_t k = atomicAdd(cnt, 1u);
            w.nlist[3 * k] = cstart;
            w.nlist[3 * k + 1] = c_i;
            w.nlist[3 * k + 2] = ce;
        }
    }
}

// ---------------------------------------------------------------------------
// Generic cell-parallel pair kernel: one policy per reference kernel.
//   Pol::KIND           candidates walked (0 fluid, 1 wall elements)
//   pol.owner(i, im_i)  the particles that run the kernel (the reference
//                       kernel's early exits)
//   Pol::State          per-particle accumulators
//   pol.init(i, r_i, st), pol.pair<NRQ>(i, j, r_i, st), pol.finish(i, st)
//                       (NRQ: q by div_nr, see qof)
// pair() is called for the accepted candidates of each particle in the
// reference walk order, so every sum is bit-identical to the per-particle
// kernel of sph.hip.
// ---------------------------------------------------------------------------
// q = sqrt(s) / H.  NRQ: the quotient by div_nr, which is the IEEE quotient
// for H in [2^-40, 2^40] and every accepted pair: sqrtf(s) is 0 or at
// least 2^-75, and at most about 2 H.
template <bool NRQ> __device__ __forceinline__ float qof(float s, float H)
{
    if constexpr (NRQ) return div_nr(sqrtf(s), rcp_nr(H));
    else return sqrtf(s) / H;
}

// Candidate fields loaded ahead of their use (NB > 1): the policy's
// Cand load(j) and use<NRQ>(i, j, cand, r_i, st), which together are its
// pair<NRQ>(i, j, r_i, st).
template <class Pol, bool NRQ> struct BatchPair {
    const Pol& pol;
    uint32_t i;
    const f3& r_i;
    typename Pol::State& st;
    __device__ __forceinline__ typename Pol::Cand load(uint32_t j) const
    {
        return pol.load(j);
    }
    __device__ __forceinline__ void use(uint32_t j,
                                        const typename Pol::Cand& c) const
    {
        pol.template use<NRQ>(i, j, c, r_i, st);
    }
};

// Pol::STATIC_CELLS when the policy declares it (else false): static
// hand-out of the listed cells in cell_kernel_k
template <class P, class = void> struct StaticCells : std::false_type {};
template <class P>
struct StaticCells<P, std::void_t<decltype(P::STATIC_CELLS)>>
    : std::integral_constant<bool, P::STATIC_CELLS> {};

// PER: a periodic box; the policy's pair_p<NRQ>(i, j, shift, r_i, st) gets
// the shift of j's neighbour cell.  NB > 1: batched consume (ring_walk).
// olist: walk the cells it lists ((first particle, cell, cell_end)
// triplets, counted in n_heads[count_word]: the owner-cell lists, the
// near-cell list) instead of every non-empty cell.
template <int D, int SEGW, class Pol, bool PER = false, int NB = 1>
__global__ void __launch_bounds__(64 * WC_WAVES)
    cell_kernel_k(Pol pol, const int* __restrict__ imove,
                  const float* __restrict__ r, WalkArgs w, float smax,
                  int nrq, const uint32_t* __restrict__ olist,
                  uint32_t count_word)
{
    __shared__ WaveLds<SEGW, PER> lds[WC_WAVES];
    const int lane = threadIdx.x & 63;
    WaveLds<SEGW, PER>& L = lds[threadIdx.x >> 6];
    if (w.status && (*w.status & 1u)) return;
    const uint32_t n_heads =
        olist ? uni(w.n_heads[count_word]) : uni(*w.n_heads);
    uint32_t range = blockIdx.x % w.ranges;
    // static hand-out of a policy's listed cells (wave v takes v, v + W,
    // ...; no atomic per cell)
    const bool stat = StaticCells<Pol>::value && olist != nullptr;
    const uint32_t nwv = gridDim.x * WC_WAVES;
    uint32_t hs = blockIdx.x * WC_WAVES + (threadIdx.x >> 6);
    while (true) {
        uint32_t hk;
        if (stat) {
            hk = hs;
            hs += nwv;
        } else {
            hk = next_cell(w, n_heads, lane, range);
        }
        if (hk >= n_heads) break;
        uint32_t cstart, c_i, ce;
        if (olist) {
            // (first particle, cell, cell_end) triplets: one round trip
            cstart = uni(olist[3 * hk]);
            c_i = uni(olist[3 * hk + 1]);
            ce = uni(olist[3 * hk + 2]);
        } else {
            cstart = uni(w.heads[hk]);
            c_i = uni(w.icell[cstart]);
            ce = uni(w.cell_end[c_i]);
        }
        const uint32_t cend = ce & AQ_CELL_END;
        if (w.flagged && Pol::OWNERS && !(ce & Pol::OWNERS)) continue;
        bool table = false;
        uint32_t total = 0;
        for (uint32_t i0 = cstart; i0 < cend; i0 += 64) {
            const uint32_t i = i0 + lane;
            const bool act = (i < cend) && pol.owner(i, imove[i]);
            if (__ballot(act) == 0) continue;
            if (!table) {
                total = neighbour_table<D, PER>(
                    w, c_i, L, lane,
                    Pol::KIND ? AQ_CELL_ELEM : AQ_CELL_FLUID);
                table = true;
            }
            typename Pol::State st{};
            f3 r_i = { 0.f, 0.f, 0.f };
            if (act) {
                r_i = ldv<D>(r, i);
                pol.init(i, r_i, st);
            }
            bool unused = false;
            f3 bmin, bmax;
            wave_box<D>(r_i, act, bmin, bmax);
            auto call = [&](auto nrqc, uint32_t j, uint32_t kl) {
                constexpr bool NRQ = decltype(nrqc)::value;
                if constexpr (PER) {
                    const float4 sh = L.shift[kl];
                    pol.template pair_p<NRQ>(i, j, f3{ sh.x, sh.y, sh.z },
                                             r_i, st);
                } else {
                    pol.template pair<NRQ>(i, j, r_i, st);
                }
            };
            if constexpr (NB > 1) {
                if (nrq)
                    ring_walk<D, SEGW, NB>(
                        L, lane, total, r_i, act, bmin, bmax, smax, r, imove,
                        Pol::KIND, false, unused,
                        BatchPair<Pol, true>{ pol, i, r_i, st }, w.per);
                else
                    ring_walk<D, SEGW, NB>(
                        L, lane, total, r_i, act, bmin, bmax, smax, r, imove,
                        Pol::KIND, false, unused,
                        BatchPair<Pol, false>{ pol, i, r_i, st }, w.per);
            } else if (nrq)
                ring_walk<D, SEGW>(
                    L, lane, total, r_i, act, bmin, bmax, smax, r, imove,
                    Pol::KIND, false, unused,
                    [&](uint32_t j, uint32_t kl) {
                        call(std::true_type{}, j, kl);
                    },
                    w.per);
            else
                ring_walk<D, SEGW>(
                    L, lane, total, r_i, act, bmin, bmax, smax, r, imove,
                    Pol::KIND, false, unused,
                    [&](uint32_t j, uint32_t kl) {
                        call(std::false_type{}, j, kl);
                    },
                    w.per);
            if (act) pol.finish(i, st);
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// ---------------------------------------------------------------------------
// Cooperative cell kernel: Q lanes per owner.
//
// A wall cell holds few owners (the 16 elements of a 3D wall plane in a cell
// of side 2h at h = 2 dr), so a wavefront with one owner per lane leaves
// most lanes idle, and the few wet wall cells (about a fifth) each form a
// long chain of candidate tests and pair terms.  Here a pass takes up to
// OW = 64 / Q owners of the cell and gives each Q lanes (lane q OW + o
// serves owner o):
//   test     every chunk of 64 candidates is staged as in ring_walk; lane
//            (q, o) tests the q-th 64/Q of the chunk's survivors against
//            owner o, and the Q pieces are assembled into the owner's
//            accept bits, which all Q lanes hold;
//   consume  in each round the Q lanes of an owner take its next Q accepted
//            candidates, lane q evaluates the q-th one's term (Pol::term, a
//            pure function of the owner and the candidate), the terms go
//            through LDS and lane 0 of the owner adds them to the
//            accumulator in candidate order (Pol::add).
// Every owner's accumulator thus takes the terms of the one-lane walk in the
// same order: the sums are bit-identical.
// ---------------------------------------------------------------------------
// B: candidates per lane and round (Q B per owner), whose loads are issued
// together; tsl holds B x 64 terms.
template <int D, int SEGW, int Q, int B, class Term, class LoadFn,
          class TermFn, class AddFn>
__device__ __forceinline__ void coop_walk(WaveLds<SEGW, false>& L,
                                          Term* tsl, int lane,
                                          uint32_t total, const f3& r_i,
                                          bool act, const f3& bmin,
                                          const f3& bmax, float smax,
                                          const float* __restrict__ r,
                                          const int* __restrict__ imove,
                                          int kind, LoadFn load, TermFn term,
                                          AddFn add)
{
    static_assert(Q == 1 || Q == 2 || Q == 4, "lanes per owner");
    static_assert(B >= 1 && B <= 4, "candidates per lane and round");
    constexpr int OW = 64 / Q;   // owners per pass
    constexpr int TPL = 64 / Q;  // survivors tested per lane and chunk
    constexpr uint32_t SEG = SEGW * 32;
    constexpr uint32_t RING = 2 * SEGW;
    const int q = lane / OW, o = lane % OW;
    const uint32_t nsegs = (total + SEG - 1) / SEG;
    const uint32_t nwords = (total + 31) >> 5;
    const uint32_t* bits = &L.bits[0][0][0];
    const uint32_t* cand = &L.cand[0][0];
    int kl = 0;
    uint32_t nw = 0, bw = 0, tested = 0;
    auto advance = [&]() {
        const uint32_t tw = min(tested * SEGW, nwords);
        while (bw == 0 && nw < tw) {
            bw = bits[(nw % RING) * 64 + lane];
            nw++;
        }
    };
    auto go = [&]() {
        advance();
        if (tested < nsegs &&
            (tested < 2 ||
             __ballot(bw != 0 && nw <= (tested - 1) * SEGW) == 0))
            return false;
        return __ballot(bw != 0) != 0;
    };
    while (true) {
        for (bool more = go(); more; more = go()) {
            // one round: the next Q B accepted candidates of every owner
            // (the Q lanes of an owner hold the same bits and step alike);
            // lane q takes the candidates k = q + Q b
            uint32_t myj[B];
            bool mine[B];
#pragma unroll
            for (int b = 0; b < B; b++) {
                myj[b] = 0;
                mine[b] = false;
            }
            int n = 0;
#pragma unroll
            for (int k = 0; k < Q * B; k++) {
                if (bw == 0) advance();
                if (bw != 0) {
                    const uint32_t b = __builtin_clz(bw);
                    bw ^= 0x80000000u >> b;
                    const uint32_t jj = cand[((nw - 1) % RING) * 32 + b];
                    if (k % Q == q) {
                        myj[k / Q] = jj;
                        mine[k / Q] = true;
                    }
                    n = k + 1;
                }
            }
            decltype(load(0u)) cd[B];
#pragma unroll
            for (int b = 0; b < B; b++)
                if (mine[b]) cd[b] = load(myj[b]);
            if constexpr (Q == 1 && B == 1) {
                if (mine[0]) add(term(myj[0], cd[0]));
            } else {
#pragma unroll
                for (int b = 0; b < B; b++) {
                    Term t{};
                    if (mine[b]) t = term(myj[b], cd[b]);
                    tsl[b * 64 + lane] = t;
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_s_waitcnt(0);
                if (q == 0) {
#pragma unroll
                    for (int k = 0; k < Q * B; k++)
                        if (k < n) add(tsl[(k / Q) * 64 + (k % Q) * OW + o]);
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        if (tested >= nsegs) break;  // and no lane has a candidate left
        // refill the oldest slot: the segment's candidate loads are issued
        // together, then its chunks of 64 are tested
        const uint32_t seg = tested * SEG;
        const uint32_t nseg = min(SEG, total - seg);
        const int slot = tested & 1;
        constexpr uint32_t NCH = SEG / 64;
        uint32_t pj[NCH];
        f3 prj[NCH];
        int pim[NCH];
#pragma unroll
        for (uint32_t ch = 0; ch < NCH; ch++) {
            pj[ch] = 0;
            const uint32_t t = seg + ch * 64 + lane;
            if (ch * 64 + lane < nseg) {
                while (t >= L.cum[kl + 1]) kl++;
                pj[ch] = L.start[kl] + (t - L.cum[kl]);
            }
        }
#pragma unroll
        for (uint32_t ch = 0; ch < NCH; ch++) {
            prj[ch] = f3{ 0.f, 0.f, 0.f };
            pim[ch] = 0;
            if (ch * 64 + lane < nseg) {
                prj[ch] = ldv_o<D>(r, pj[ch]);
                pim[ch] = ldi_o(imove, pj[ch]);
            }
        }
#pragma unroll
        for (uint32_t ch = 0; ch < NCH; ch++) {
            const uint32_t cb = ch * 64;
            if (cb >= nseg) break;
            float4 st = make_float4(0.f, 0.f, 0.f, 0.f);
            bool take = false;
            const uint32_t j = pj[ch];
            if (cb + lane < nseg) {
                const f3 rj = prj[ch];
                const int im_j = pim[ch];
                take = kind ? (im_j == -3) : (im_j == 1);
                if (take) {
                    const float dx =
                        fmaxf(fmaxf(bmin.x - rj.x, rj.x - bmax.x), 0.f);
                    const float dy =
                        fmaxf(fmaxf(bmin.y - rj.y, rj.y - bmax.y), 0.f);
                    float d2 = dx * dx;
                    d2 = d2 + dy * dy;
                    if constexpr (D == 3) {
                        const float dz =
                            fmaxf(fmaxf(bmin.z - rj.z, rj.z - bmax.z), 0.f);
                        d2 = d2 + dz * dz;
                    }
                    take = !(d2 > smax);
                }
                st = make_float4(rj.x, rj.y, rj.z, 0.f);
            }
            const uint64_t tm = __ballot(take);
            const uint32_t cnt = (uint32_t)__popcll(tm);
            uint32_t lo = 0, hi = 0;
            if (cnt != 0) {
                // compact the survivors into stage[0 .. cnt), in order
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(tm >> 32),
                    __builtin_amdgcn_mbcnt_lo((uint32_t)tm, 0u));
                if (take) {
                    L.stage[rank] = st;
                    L.cand[slot][cb + rank] = j;
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_s_waitcnt(0);
                // lane (q, o): survivors q TPL .. q TPL + TPL - 1
                const uint32_t k0 = (uint32_t)(q * TPL);
                const uint32_t kn =
                    cnt > k0 ? min(cnt - k0, (uint32_t)TPL) : 0u;
                uint64_t piece = 0;
                for (uint32_t k = 0; k < kn; k++) {
                    const float4 c = L.stage[k0 + k];
                    const f3 r_ij = sub<D>(f3{ c.x, c.y, c.z }, r_i);
                    const float s = dot<D>(r_ij, r_ij);
                    piece = (piece << 1) | (!(s > smax) ? 1u : 0u);
                }
                if (kn == 0u || !act) piece = 0;
                else if (kn < (uint32_t)TPL) piece <<= (uint32_t)TPL - kn;
                // the owner's accept bits of the whole chunk, MSB first
                uint64_t mask = 0;
#pragma unroll
                for (int qq = 0; qq < Q; qq++) {
                    const uint64_t pc =
                        (Q == 1) ? piece
                                 : (uint64_t)(uint32_t)__shfl(
                                       (int)(uint32_t)piece, qq * OW + o);
                    mask |= pc << (64 - TPL * (qq + 1));
                }
                lo = (uint32_t)(mask >> 32);
                hi = (uint32_t)mask;
            }
            L.bits[slot][cb >> 5][lane] = lo;
            if (nseg - cb > 32) L.bits[slot][(cb >> 5) + 1][lane] = hi;
            __builtin_amdgcn_wave_barrier();
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_s_waitcnt(0);
        tested++;
    }
    __builtin_amdgcn_wave_barrier();
}

// Owners of each listed cell in passes of up to 64 / Q; the policy splits
// its pair into load(j) (the candidate's fields), term<NRQ>(i, j, cand,
// r_i, st) (pure) and add(st, term), applied in candidate order.  olist as
// in cell_kernel_k, its count in n_heads[count_word].
template <int D, int SEGW, int Q, class Pol>
__global__ void __launch_bounds__(64 * WC_WAVES)
    cell_coop_k(Pol pol, const int* __restrict__ imove,
                const float* __restrict__ r, WalkArgs w, float smax, int nrq,
                const uint32_t* __restrict__ olist, uint32_t count_word)
{
    constexpr int OW = 64 / Q;
    using Term = typename Pol::Term;
    constexpr int B = Pol::COOP_B;
    __shared__ WaveLds<SEGW, false> lds[WC_WAVES];
    __shared__ uint32_t owns[WC_WAVES][OW];
    __shared__ Term tsl[WC_WAVES][B * 64];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int q = lane / OW, o = lane % OW;
    WaveLds<SEGW, false>& L = lds[wv];
    if (w.status && (*w.status & 1u)) return;
    const uint32_t n_heads =
        olist ? uni(w.n_heads[count_word]) : uni(*w.n_heads);
    uint32_t range = blockIdx.x % w.ranges;
    // static hand-out (Pol::STATIC_CELLS): wave v takes the cells
    // v, v + W, v + 2W, ... (no atomic per cell), for lists whose cells cost
    // alike
    const uint32_t nwv = gridDim.x * WC_WAVES;
    uint32_t hs = blockIdx.x * WC_WAVES + (uint32_t)wv;
    while (true) {
        uint32_t hk;
        if (Pol::STATIC_CELLS) {
            hk = hs;
            hs += nwv;
        } else {
            hk = next_cell(w, n_heads, lane, range);
        }
        if (hk >= n_heads) break;
        uint32_t cstart, c_i, ce;
        if (olist) {
            cstart = uni(olist[3 * hk]);
            c_i = uni(olist[3 * hk + 1]);
            ce = uni(olist[3 * hk + 2]);
        } else {
            cstart = uni(w.heads[hk]);
            c_i = uni(w.icell[cstart]);
            ce = uni(w.cell_end[c_i]);
        }
        const uint32_t cend = ce & AQ_CELL_END;
        if (w.flagged && Pol::OWNERS && !(ce & Pol::OWNERS)) continue;
        bool table = false;
        uint32_t total = 0;
        uint32_t pos = cstart;  // scan cursor (uniform)
        while (pos < cend) {
            // collect up to OW owners
            uint32_t nf = 0;
            while (nf < (uint32_t)OW && pos < cend) {
                const uint32_t ii = pos + (uint32_t)lane;
                const bool own = ii < cend && pol.owner(ii, imove[ii]);
                const uint64_t bm = __ballot(own);
                const uint32_t c = (uint32_t)__popcll(bm);
                const uint32_t room = (uint32_t)OW - nf;
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(bm >> 32),
                    __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u));
                if (own && rank < room) owns[wv][nf + rank] = ii;
                if (c <= room) {
                    nf += c;
                    pos += 64;
                } else {
                    // resume at the first owner not taken
                    uint64_t m = bm;
                    for (uint32_t k = 0; k < room; k++) m &= m - 1ull;
                    pos += (uint32_t)__builtin_ctzll(m);
                    nf = OW;
                }
            }
            if (nf == 0) break;
            if (!table) {
                total = neighbour_table<D, false>(
                    w, c_i, L, lane,
                    Pol::KIND ? AQ_CELL_ELEM : AQ_CELL_FLUID);
                table = true;
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_s_waitcnt(0);
            const bool act = (uint32_t)o < nf;
            const uint32_t i = act ? owns[wv][o] : 0u;
            typename Pol::State st{};
            f3 r_i = { 0.f, 0.f, 0.f };
            if (act) {
                r_i = ldv<D>(r, i);
                pol.init(i, r_i, st);
            }
            f3 bmin, bmax;
            wave_box<D>(r_i, act, bmin, bmax);
            auto add = [&](const Term& t) { pol.add(st, t); };
            auto load = [&](uint32_t j) { return pol.load(j); };
            using Cand = typename Pol::Cand;
            if (nrq)
                coop_walk<D, SEGW, Q, B>(
                    L, &tsl[wv][0], lane, total, r_i, act, bmin, bmax, smax,
                    r, imove, Pol::KIND, load,
                    [&](uint32_t j, const Cand& c) {
                        return pol.template term<true>(i, j, c, r_i, st);
                    },
                    add);
            else
                coop_walk<D, SEGW, Q, B>(
                    L, &tsl[wv][0], lane, total, r_i, act, bmin, bmax, smax,
                    r, imove, Pol::KIND, load,
                    [&](uint32_t j, const Cand& c) {
                        return pol.template term<false>(i, j, c, r_i, st);
                    },
                    add);
            if (act && q == 0) pol.finish(i, st);
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// The element-owner policies below split each pair into load(j) (the
// candidate's fields), term<NRQ>(i, j, cand, r_i, st) (a pure function of
// the owner's state and the candidate) and add(st, term); pair and use
// apply them in that order, so every walk form adds the same terms.
#define AQ_POL_PAIR_FROM_TERM                                                  \
    template <bool NRQ>                                                        \
    __device__ void use(uint32_t i, uint32_t j, const Cand& c,                 \
                        const f3& r_i, State& st) const                        \
    {                                                                          \
        add(st, term<NRQ>(i, j, c, r_i, st));                                  \
    }                                                                          \
    template <bool NRQ>                                                        \
    __device__ void pair(uint32_t i, uint32_t j, const f3& r_i, State& st)     \
        const                                                                  \
    {                                                                          \
        use<NRQ>(i, j, load(j), r_i, st);                                      \
    }

// cfd/Boundary/BI/Shepard.cl:61-150 (compute) for the non-fluid particles
// (skip_fluid) or all of them
template <int D> struct PolBiShepard {
    static constexpr int BATCH = 4;  // candidates loaded together
    static constexpr bool STATIC_CELLS = true;  // coop walk hand-out
    static constexpr int COOP_B = 1;  // per lane and round (coop walk)
    static constexpr int KIND = 1;
    static constexpr uint32_t OWNERS = 0;  // elements, sensors, ...
    const int* imove;
    const float *normal, *tangent, *binormal, *m;
    float* shepard;
    float H, CONW, DIMS;
    int skip_fluid;
    struct State {
        float acc;
        bool self_added;
    };
    struct Cand {
        f3 r, n, t, b;
        float area;
    };
    // kind 0: add v; 1: the particle's own element (i itself, or a
    // coplanar element within half a spacing), which takes 1/2 once;
    // 2: nothing
    struct Term {
        float v;
        uint32_t kind;
    };
    __device__ bool owner(uint32_t, int im) const
    {
        return !((im < -3) || (im > 1)) && !(skip_fluid && im == 1);
    }
    __device__ void init(uint32_t, const f3&, State& st) const
    {
        st.acc = 1.f;
        st.self_added = false;
    }
    __device__ Cand load(uint32_t j) const
    {
        return Cand{ ldv_o<D>(r_, j), ldv_o<D>(normal, j),
                     ldv_o<D>(tangent, j), ldv_o<D>(binormal, j),
                     ldf_o(m, j) };
    }
    template <bool NRQ>
    __device__ Term term(uint32_t i, uint32_t j, const Cand& c,
                         const f3& r_i, const State&) const
    {
        if (i == j) return Term{ 0.f, 1u };
        const float eps = 1E-8f * H;
        const f3 r_ij = sub<D>(c.r, r_i);
        const float r_n = dot<D>(r_ij, c.n);
        const float r_t = fabsf(dot<D>(r_ij, c.t));
        const float r_b = fabsf(dot<D>(r_ij, c.b));
        if ((r_n > -eps) && (r_n < eps)) {
            // a coplanar element (most candidates on a flat wall): no q
            const float dr = 0.55f * pow_cl(c.area, 1.f / (DIMS - 1.f));
            return Term{ 0.f, ((r_t <= dr) && (r_b <= dr)) ? 1u : 2u };
        }
        const float q = qof<NRQ>(dot<D>(r_ij, r_ij), H);
        return Term{ r_n * CONW * kernelS_P<D>(q) * c.area +
                         kernelS_D<D>(fabsf(r_n), r_t, r_b, c.area),
                     0u };
    }
    __device__ void add(State& st, const Term& t) const
    {
        if (t.kind == 0u) {
            st.acc += t.v;
        } else if (t.kind == 1u && !st.self_added) {
            st.self_added = true;
            st.acc -= 0.5f;
        }
    }
    AQ_POL_PAIR_FROM_TERM
    __device__ void finish(uint32_t i, const State& st) const
    {
        shepard[i] = st.acc;
    }
    const float* r_;
};

// cfd/Boundary/BI/LapU.cl:58-123 (freeslip)
template <int D> struct PolBiLapU {
    static constexpr int BATCH = 4;  // candidates loaded together
    static constexpr bool STATIC_CELLS = true;  // coop walk hand-out
    static constexpr int COOP_B = 2;  // per lane and round (coop walk)
    static constexpr int KIND = 0;
    static constexpr uint32_t OWNERS = AQ_CELL_ELEM;
    const float *r_, *u, *rho, *m;
    float* lap_u;
    float H, CONF;
    int lap;
    struct State {
        f3 u_i, lu;
    };
    struct Cand {
        f3 r, u;
        float rho, m;
    };
    struct Term {
        f3 v;  // the rounded products of the axpy
        uint32_t use;
    };
    __device__ bool owner(uint32_t, int im) const { return im == -3; }
    __device__ void init(uint32_t i, const f3&, State& st) const
    {
        st.u_i = ldv<D>(u, i);
        st.lu = ldv<D>(lap_u, i);
    }
    __device__ Cand load(uint32_t j) const
    {
        return Cand{ ldv_o<D>(r_, j), ldv_o<D>(u, j), ldf_o(rho, j),
                     ldf_o(m, j) };
    }
    template <bool NRQ>
    __device__ Term term(uint32_t i, uint32_t j, const Cand& c,
                         const f3& r_i, const State& st) const
    {
        constexpr float CLEARY = (D == 2) ? 8.f : 10.f;
        if (i == j) return Term{ f3{ 0.f, 0.f, 0.f }, 0u };
        const f3 r_ij = sub<D>(c.r, r_i);
        const float s = dot<D>(r_ij, r_ij);
        const float q = qof<NRQ>(s, H);
        const float rho_j = c.rho;
        const float f_ij = kernelF<D>(q) * CONF * c.m;
        const f3 u_ij = sub<D>(c.u, st.u_i);
        float sc;
        f3 v;
        if (lap == AQ_LAP_MONAGHAN) {
            const float udr = dot<D>(u_ij, r_ij);
            const float r2 = (q * q + 0.01f) * H * H;
            sc = f_ij * CLEARY * udr / (r2 * rho_j);
            v = r_ij;
        } else {
            sc = f_ij * 2.f / rho_j;
            v = u_ij;
        }
        return Term{ f3{ sc * v.x, sc * v.y, sc * v.z }, 1u };
    }
    __device__ void add(State& st, const Term& t) const
    {
        if (!t.use) return;
        st.lu.x += t.v.x;
        st.lu.y += t.v.y;
        if constexpr (D == 3) st.lu.z += t.v.z;
    }
    AQ_POL_PAIR_FROM_TERM
    __device__ void finish(uint32_t i, const State& st) const
    {
        stxyz<D>(lap_u, i, st.lu);
    }
};

// cfd/Boundary/BI/Interpolation.cl:54-107
template <int D> struct PolBiInterp {
    static constexpr int BATCH = 4;  // candidates loaded together
    static constexpr bool STATIC_CELLS = true;  // coop walk hand-out
    static constexpr int COOP_B = 2;  // per lane and round (coop walk)
    static constexpr int KIND = 0;
    static constexpr uint32_t OWNERS = AQ_CELL_ELEM;
    const float *r_, *m, *rho, *grad_p;
    float* p;
    float H, CONW;
    struct State {
        f3 gp_i;
        float acc;
    };
    struct Cand {
        f3 r;
        float m, rho, p;
    };
    using Term = float;
    __device__ bool owner(uint32_t, int im) const { return im == -3; }
    __device__ void init(uint32_t i, const f3&, State& st) const
    {
        const float rho_i = rho[i];
        st.gp_i = ldv<D>(grad_p, i);
        st.gp_i.x = rho_i * st.gp_i.x;
        st.gp_i.y = rho_i * st.gp_i.y;
        st.gp_i.z = rho_i * st.gp_i.z;
        st.acc = 0.f;
    }
    __device__ Cand load(uint32_t j) const
    {
        return Cand{ ldv_o<D>(r_, j), ldf_o(m, j), ldf_o(rho, j),
                     ldf_o(p, j) };
    }
    template <bool NRQ>
    __device__ Term term(uint32_t, uint32_t, const Cand& c, const f3& r_i,
                         const State& st) const
    {
        const f3 r_ij = sub<D>(c.r, r_i);
        const float s = dot<D>(r_ij, r_ij);
        const float q = qof<NRQ>(s, H);
        const float w_ij = kernelW<D>(q) * CONW * c.m / c.rho;
        return (c.p - dot<D>(st.gp_i, r_ij)) * w_ij;
    }
    __device__ void add(State& st, const Term& t) const { st.acc += t; }
    AQ_POL_PAIR_FROM_TERM
    __device__ void finish(uint32_t i, const State& st) const
    {
        p[i] = st.acc;
    }
};
#undef AQ_POL_PAIR_FROM_TERM

// cfd/Boundary/BI/Interactions.cl:51-133, for the fluid particles with an
// element inside the support (near flags of the fused walk) or all of them
template <int D> struct PolBiInteractions {
    static constexpr int BATCH = 4;  // candidates loaded together
    // the near cells cost about alike: static hand-out over the near list
    static constexpr bool STATIC_CELLS = true;
    static constexpr int KIND = 1;
    static constexpr uint32_t OWNERS = AQ_CELL_FLUID;
    const float *r_, *normal, *u, *rho, *m, *p;
    float *grad_p, *div_u;
    const uint8_t* near;
    float H, CONW;
    struct State {
        f3 u_i, gp;
        float p_i, rho_i, dv;
    };
    __device__ bool owner(uint32_t i, int im) const
    {
        return im == 1 && (!near || near[i]);
    }
    __device__ void init(uint32_t i, const f3&, State& st) const
    {
        st.u_i = ldv<D>(u, i);
        st.p_i = p[i];
        st.rho_i = rho[i];
        st.gp = ldv<D>(grad_p, i);
        st.dv = div_u[i];
    }
    struct Cand {
        f3 r, n, u;
        float area, p;
    };
    __device__ Cand load(uint32_t j) const
    {
        return Cand{ ldv_o<D>(r_, j), ldv_o<D>(normal, j), ldv_o<D>(u, j),
                     ldf_o(m, j), ldf_o(p, j) };
    }
    template <bool NRQ>
    __device__ void pair(uint32_t i, uint32_t j, const f3& r_i,
                         State& st) const
    {
        use<NRQ>(i, j, load(j), r_i, st);
    }
    template <bool NRQ>
    __device__ void use(uint32_t, uint32_t, const Cand& c, const f3& r_i,
                        State& st) const
    {
        const f3 r_ij = sub<D>(c.r, r_i);
        const float s = dot<D>(r_ij, r_ij);
        const float q = qof<NRQ>(s, H);
        const f3 n_j = c.n;
        const float area_j = c.area;
        const float p_j = c.p;
        const f3 du = sub<D>(c.u, st.u_i);
        const float w_ij = kernelW<D>(q) * CONW * area_j;
        axpy<D>(st.gp, (st.p_i + p_j) / st.rho_i * w_ij, n_j);
        st.dv += st.rho_i * dot<D>(du, n_j) * w_ij;
    }
    __device__ void finish(uint32_t i, const State& st) const
    {
        stxyz<D>(grad_p, i, st.gp);
        div_u[i] = st.dv;
    }
};

// ---- LES closure pair kernels (cfd/LES/GradU.cl, cfd/LES/Interactions.cl
// of Presets/cfd/LES.xml; no counterpart in the reference) -------------------
// The SPH velocity gradient of the fluid particles,
//   G_i = sum_j m_j / rho_j (u_j - u_i) (x) F_ij r_ij,
// grad_u[a n + b] = du_a / dx_b (row stride n = 2 in 2D, 4 in 3D), summed
// in the walk order; pair_p: the same at a periodic image.  The non-fluid
// particles' rows are cleared by the launcher.
template <int D> struct PolLesGradU {
    static constexpr int KIND = 0;
    static constexpr uint32_t OWNERS = AQ_CELL_FLUID;
    static constexpr int NR = (D == 2) ? 2 : 4;
    const float *r_, *u, *rho, *m;
    float* grad_u;
    float H, CONF;
    const float* per;
    struct State {
        f3 u_i;
        float G[D * D];
    };
    __device__ bool owner(uint32_t, int im) const { return im == 1; }
    __device__ void init(uint32_t i, const f3&, State& st) const
    {
        st.u_i = ldv<D>(u, i);
#pragma unroll
        for (int k = 0; k < D * D; k++) st.G[k] = 0.f;
    }
    __device__ void add(uint32_t j, const f3& r_j, const f3& r_i, float q,
                        State& st) const
    {
        const f3 r_ij = sub<D>(r_j, r_i);
        const float w = kernelF<D>(q) * CONF * ldf_o(m, j) / ldf_o(rho, j);
        const f3 du = sub<D>(ldv_o<D>(u, j), st.u_i);
        const float dv[3] = { du.x, du.y, du.z };
        const float rv[3] = { r_ij.x, r_ij.y, r_ij.z };
#pragma unroll
        for (int a = 0; a < D; a++)
#pragma unroll
            for (int b = 0; b < D; b++)
                st.G[a * D + b] += dv[a] * (w * rv[b]);
    }
    template <bool NRQ>
    __device__ void pair(uint32_t i, uint32_t j, const f3& r_i,
                         State& st) const
    {
        if (i == j) return;
        const f3 r_j = ldv_o<D>(r_, j);
        const f3 r_ij = sub<D>(r_j, r_i);
        add(j, r_j, r_i, qof<NRQ>(dot<D>(r_ij, r_ij), H), st);
    }
    template <bool NRQ>
    __device__ void pair_p(uint32_t i, uint32_t j, const f3& sh,
                           const f3& r_i, State& st) const
    {
        if (i == j) return;
        const f3 r_j = per_shift<D>(ldv_o<D>(r_, j), sh, per);
        const f3 r_ij = sub<D>(r_j, r_i);
        add(j, r_j, r_i, qof<NRQ>(dot<D>(r_ij, r_ij), H), st);
    }
    __device__ void finish(uint32_t i, const State& st) const
    {
        constexpr int S = (D == 2) ? 4 : 16;
        float o[S];
#pragma unroll
        for (int k = 0; k < S; k++) o[k] = 0.f;
#pragma unroll
        for (int a = 0; a < D; a++)
#pragma unroll
            for (int b = 0; b < D; b++) o[a * NR + b] = st.G[a * D + b];
#pragma unroll
        for (int k = 0; k < S; k++) grad_u[(size_t)i * S + k] = o[k];
    }
};

// The eddy-viscosity term in the variable-viscosity Morris form,
//   dudt_i += sum_j F_ij m_j (rho_i nu_i + rho_j nu_j) / (rho_i rho_j)
//             (u_j - u_i),
// summed in the walk order and added to dudt at the end.
template <int D> struct PolLesInteractions {
    static constexpr int KIND = 0;
    static constexpr uint32_t OWNERS = AQ_CELL_FLUID;
    const float *r_, *u, *rho, *m, *nu_t;
    float* dudt;
    float H, CONF;
    const float* per;
    struct State {
        f3 u_i, acc;
        float rho_i, mu_i;
    };
    __device__ bool owner(uint32_t, int im) const { return im == 1; }
    __device__ void init(uint32_t i, const f3&, State& st) const
    {
        st.u_i = ldv<D>(u, i);
        st.rho_i = rho[i];
        st.mu_i = st.rho_i * nu_t[i];
        st.acc = f3{ 0.f, 0.f, 0.f };
    }
    __device__ void add(uint32_t j, float q, State& st) const
    {
        const float rho_j = ldf_o(rho, j);
        const float f = kernelF<D>(q) * CONF * ldf_o(m, j);
        const float sc =
            f * (st.mu_i + rho_j * ldf_o(nu_t, j)) / (st.rho_i * rho_j);
        axpy<D>(st.acc, sc, sub<D>(ldv_o<D>(u, j), st.u_i));
    }
    template <bool NRQ>
    __device__ void pair(uint32_t i, uint32_t j, const f3& r_i,
                         State& st) const
    {
        if (i == j) return;
        const f3 r_ij = sub<D>(ldv_o<D>(r_, j), r_i);
        add(j, qof<NRQ>(dot<D>(r_ij, r_ij), H), st);
    }
    template <bool NRQ>
    __device__ void pair_p(uint32_t i, uint32_t j, const f3& sh,
                           const f3& r_i, State& st) const
    {
        if (i == j) return;
        const f3 r_ij = sub<D>(per_shift<D>(ldv_o<D>(r_, j), sh, per), r_i);
        add(j, qof<NRQ>(dot<D>(r_ij, r_ij), H), st);
    }
    __device__ void finish(uint32_t i, const State& st) const
    {
        const f3 d = ldv<D>(dudt, i);
        stxyz<D>(dudt, i, f3{ d.x + st.acc.x, d.y + st.acc.y,
                              d.z + st.acc.z });
    }
};

// grad_u of the particles that are not fluid: zero (the per-particle
// kernel writes zeros for them)
template <int D>
__global__ void les_gradu_clear_k(const int* __restrict__ imove,
                                  float* __restrict__ grad_u, uint32_t N)
{
    constexpr int S = (D == 2) ? 4 : 16;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || imove[i] == 1) return;
#pragma unroll
    for (int k = 0; k < S; k++) grad_u[(size_t)i * S + k] = 0.f;
}

// Clears the NRANGE range counters.
__global__ void zero_u32_k(uint32_t* p, uint32_t* rec,
                           uint32_t* extra = nullptr)
{
    if (threadIdx.x < NRANGE) p[threadIdx.x * RSTRIDE] = 0u;
    if (rec && threadIdx.x < NREC)
        rec[threadIdx.x * RSTRIDE] = rec_init[threadIdx.x];
    if (extra && threadIdx.x == 0) *extra = 0u;
}

// Cell hand-out ranges: NRANGE (one per XCD) unless AQ_WALK_RANGES=1.
static uint32_t walk_ranges()
{
    static const uint32_t nr = [] {
        const char* e = getenv("AQ_WALK_RANGES");
        return (e && atoi(e) == 1) ? 1u : NRANGE;
    }();
    return nr;
}

// Workgroups of `ker` that fit on the device at once (all CUs), cached per
// kernel: the walk hands cells out dynamically, so a grid of exactly that
// many workgroups keeps every slot busy until the cells run out.
unsigned resident_blocks(const void* ker, int block)
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

// Compute units of the current device (cached).
static unsigned cu_count()
{
    static std::mutex mu;
    static std::map<int, unsigned> cache;
    int dev = 0;
    AQ_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(dev);
    if (it != cache.end()) return it->second;
    int ncu = 0;
    AQ_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                                 dev));
    return cache[dev] = (unsigned)std::max(1, ncu);
}

// Workgroups per compute unit the walks of the side stream take
// (AQ_SIDE_WGS, default 1); the fused walk beside them leaves them that
// room, so both stay resident together.
unsigned side_wgs_per_cu()
{
    static const unsigned n = [] {
        const char* e = getenv("AQ_SIDE_WGS");
        const int v = e ? atoi(e) : 1;
        return (unsigned)std::min(4, std::max(0, v));
    }();
    return n;
}

// A resident grid less `reserve` workgroups per compute unit (one per
// compute unit at least).
static unsigned reserved_grid(unsigned resident, unsigned reserve)
{
    const unsigned ncu = cu_count();
    const unsigned cut = reserve * ncu;
    return resident > cut + ncu ? resident - cut : std::min(resident, ncu);
}

// Grid of a walk: the resident grid, or the side stream's share.
static unsigned walk_grid(const void* ker, const CellArgs& c)
{
    const unsigned res = resident_blocks(ker, 64 * WC_WAVES);
    if (c.next_off != AQ_CL_SIDE) return res;
    return std::min(res, std::max(1u, side_wgs_per_cu()) * cu_count());
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
                             int flagged, int math, unsigned ranges,
                             int fast_div, hipStream_t s, const float* per,
                             unsigned reserve_per_cu, uint32_t* nlist,
                             const uint32_t* blist)
{
    // a periodic box has no walls: the boundary-integral Shepard pass is
    // not built for it
    AQ_REQUIRE(!per || shep != 2,
               "periodic box: boundary-integral walls are not supported");
    // candidate loads use 32-bit byte offsets (16 B per particle at most)
    AQ_REQUIRE(N < (1u << 28), "cell-parallel walk: more than 2^28 particles");
    // the hand-out counters live after the non-empty cell count
    uint32_t* next = const_cast<uint32_t*>(n_heads) + RSTRIDE;
    // the range record (relaxed arithmetic has no use for it)
    uint32_t* rec =
        (fast_div && math != 1) ? const_cast<uint32_t*>(n_heads) + RREC
                                : nullptr;
    // the near-cell list is filled from the walk's near flags
    if (!near || shep != 2) nlist = nullptr;
    hipLaunchKernelGGL(zero_u32_k, dim3(1), dim3(64), 0, s, next, rec,
                       nlist ? const_cast<uint32_t*>(n_heads) + AQ_CL_NNCELLS
                             : nullptr);
    if (rec && N) {
        const unsigned nb = std::min(512u, (N + 1023u) / 1024u);
        if (dims == 2)
            hipLaunchKernelGGL(fluid_ranges_k<2>, dim3(nb), dim3(256), 0, s,
                               imove, p, rho, m, u, N, rec);
        else
            hipLaunchKernelGGL(fluid_ranges_k<3>, dim3(nb), dim3(256), 0, s,
                               imove, p, rho, m, u, N, rec);
    }
    WalkArgs w{ icell,  ihoc,   n_cells, cell_end, heads,
                n_heads, next, status, N,       (uint32_t)flagged,
                ranges ? std::min<unsigned>(ranges, NRANGE) : walk_ranges(),
                per, nlist, flagged ? blist : nullptr };
#define AQ_FC3(DD, SH, LP, SW, LF, PR)                                         \
    do {                                                                       \
        auto* ker = math == 1                                                  \
                        ? fluid_pass_cells_k<DD, SH, LP, SW, LF, 1, PR>        \
                    : math == 2 ? fluid_pass_cells_k<DD, SH, LP, 8, LF, 2, PR> \
                                : fluid_pass_cells_k<DD, SH, LP, SW, LF, 0, PR>; \
        const unsigned nb =                                                    \
            nblocks ? nblocks                                                  \
                    : reserved_grid(resident_blocks((const void*)ker,          \
                                                    64 * WC_WAVES),            \
                                    reserve_per_cu);                           \
        hipLaunchKernelGGL(ker, dim3(nb), dim3(64 * WC_WAVES), 0, s, imove, r, \
                           u, rho, m, p, normal, tangent, binormal, shepard,   \
                           grad_p, lap_u, div_u, lap_p, near, w, d.H, d.CONW,  \
                           d.CONF, d.DIMS, smax, 1.f / d.H, rec);              \
    } while (0)
#define AQ_FC2(DD, SH, LP, SW, LF)                                             \
    do {                                                                       \
        if (per && SH != 2) AQ_FC3(DD, SH, LP, SW, LF, (SH != 2));             \
        else AQ_FC3(DD, SH, LP, SW, LF, false);                                \
    } while (0)
#define AQ_FC(DD, SH, LP, SW)                                                  \
    do {                                                                       \
        if (d.lap_formulation == AQ_LAP_MONAGHAN)                              \
            AQ_FC2(DD, SH, LP, SW, AQ_LAP_MONAGHAN);                           \
        else                                                                   \
            AQ_FC2(DD, SH, LP, SW, AQ_LAP_MORRIS);                             \
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
#undef AQ_FC2
#undef AQ_FC3
}


// ---- launchers of the generic cell kernels ---------------------------------
template <int D, class Pol, bool PER = false, int NB = 1>
static void launch_cells(const Pol& pol, const int32_t* imove, const float* r,
                         const CellArgs& c, float smax, hipStream_t s,
                         const float* per = nullptr,
                         const uint32_t* olist = nullptr,
                         uint32_t count_word = AQ_CL_NOCELLS)
{
    AQ_REQUIRE(c.N < (1u << 28), "cell-parallel walk: more than 2^28 particles");
    uint32_t* next = const_cast<uint32_t*>(c.n_heads) +
                     (c.next_off ? c.next_off : RSTRIDE);
    hipLaunchKernelGGL(zero_u32_k, dim3(1), dim3(64), 0, s, next,
                       (uint32_t*)nullptr);
    WalkArgs w{ c.icell,   c.ihoc,   c.n_cells, c.cell_end, c.heads,
                c.n_heads, next,     c.status,  c.N,        c.flagged,
                walk_ranges(), per };
    if (!olist && c.blist) {
        // every non-empty cell, in blocked order (n_heads[0] of them)
        olist = c.blist;
        count_word = 0;
    }
    auto* ker = cell_kernel_k<D, 8, Pol, PER, NB>;
    const unsigned nb = walk_grid((const void*)ker, c);
    const int nrq = (pol.H >= 0x1p-40f && pol.H <= 0x1p40f) ? 1 : 0;
    hipLaunchKernelGGL(ker, dim3(nb), dim3(64 * WC_WAVES), 0, s, pol, imove,
                       r, w, smax, nrq, olist, count_word);
}

// The boundary-integral element kernels: batched consume; the kernels whose
// owners are wall elements or other non-fluid particles (owner_cells) walk
// the owner-cell list when the link-list built it.
// Kernels whose owners are wall elements or other non-fluid particles
// (owner_cells) run the cooperative walk over the owner-cell list when the
// link-list built it (4 lanes per owner); the others, and all of them
// without the list, the one-owner-per-lane walk with batched loads.
template <int D, class Pol>
static void launch_bi_cells(const Pol& pol, const int32_t* imove,
                            const float* r, const CellArgs& c, float smax,
                            hipStream_t s)
{
    launch_cells<D, Pol, false, Pol::BATCH>(pol, imove, r, c, smax, s);
}

// list: 0 every non-empty cell (one owner per lane), 1 the owner cells,
// 2 the wet owner cells (cooperative walk)
template <int D, class Pol>
static void launch_bi_coop(const Pol& pol, const int32_t* imove,
                           const float* r, const CellArgs& c, float smax,
                           hipStream_t s, int list)
{
    const uint32_t* ol = !c.flagged ? nullptr
                         : list == 1 ? c.ocells
                         : list == 2 ? c.wcells
                                     : nullptr;
    if (!ol) {
        launch_bi_cells<D>(pol, imove, r, c, smax, s);
        return;
    }
    AQ_REQUIRE(c.N < (1u << 28), "cell-parallel walk: more than 2^28 particles");
    uint32_t* next = const_cast<uint32_t*>(c.n_heads) +
                     (c.next_off ? c.next_off : RSTRIDE);
    hipLaunchKernelGGL(zero_u32_k, dim3(1), dim3(64), 0, s, next,
                       (uint32_t*)nullptr);
    WalkArgs w{ c.icell,   c.ihoc,   c.n_cells, c.cell_end, c.heads,
                c.n_heads, next,     c.status,  c.N,        c.flagged,
                walk_ranges(), nullptr };
    auto* ker = cell_coop_k<D, 8, 4, Pol>;
    const unsigned nb = walk_grid((const void*)ker, c);
    const int nrq = (pol.H >= 0x1p-40f && pol.H <= 0x1p40f) ? 1 : 0;
    hipLaunchKernelGGL(ker, dim3(nb), dim3(64 * WC_WAVES), 0, s, pol, imove,
                       r, w, smax, nrq, ol,
                       list == 1 ? AQ_CL_NOCELLS : AQ_CL_NWCELLS);
}

// p = 0 for the wall elements (imove == -3): the value BI Interpolation
// gives the elements without fluid in range, which the walk over the wet
// owner cells does not visit
__global__ void elem_zero_k(const int* __restrict__ imove,
                            float* __restrict__ p, uint32_t N)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N && imove[i] == -3) p[i] = 0.f;
}

void launch_bi_shepard_compute_cells(int dims, const int32_t* imove,
                                     const float* r, const float* normal,
                                     const float* tangent,
                                     const float* binormal, const float* m,
                                     float* shepard, const CellArgs& c,
                                     aq_defs d, float smax, int skip_fluid,
                                     hipStream_t s)
{
    // with the fluid among the owners (skip_fluid == 0) every cell owns
    if (dims == 2) {
        PolBiShepard<2> p{ imove, normal, tangent, binormal, m, shepard,
                           d.H,   d.CONW, d.DIMS,  skip_fluid, r };
        launch_bi_coop<2>(p, imove, r, c, smax, s, skip_fluid ? 1 : 0);
    } else {
        PolBiShepard<3> p{ imove, normal, tangent, binormal, m, shepard,
                           d.H,   d.CONW, d.DIMS,  skip_fluid, r };
        launch_bi_coop<3>(p, imove, r, c, smax, s, skip_fluid ? 1 : 0);
    }
}

void launch_bi_lapu_cells(int dims, const int32_t* imove, const float* r,
                          const float* u, const float* rho, const float* m,
                          float* lap_u, const CellArgs& c, aq_defs d,
                          float smax, hipStream_t s)
{
    if (dims == 2) {
        PolBiLapU<2> p{ r, u, rho, m, lap_u, d.H, d.CONF, d.lap_formulation };
        launch_bi_coop<2>(p, imove, r, c, smax, s, 2);
    } else {
        PolBiLapU<3> p{ r, u, rho, m, lap_u, d.H, d.CONF, d.lap_formulation };
        launch_bi_coop<3>(p, imove, r, c, smax, s, 2);
    }
}

void launch_bi_interp_cells(int dims, const int32_t* imove, const float* r,
                            const float* m, const float* rho,
                            const float* grad_p, float* p, const CellArgs& c,
                            aq_defs d, float smax, hipStream_t s)
{
    if (c.flagged && c.wcells)
        hipLaunchKernelGGL(elem_zero_k, dim3(blocks_for(c.N, 256)), dim3(256),
                           0, s, imove, p, c.N);
    if (dims == 2) {
        PolBiInterp<2> q{ r, m, rho, grad_p, p, d.H, d.CONW };
        launch_bi_coop<2>(q, imove, r, c, smax, s, 2);
    } else {
        PolBiInterp<3> q{ r, m, rho, grad_p, p, d.H, d.CONW };
        launch_bi_coop<3>(q, imove, r, c, smax, s, 2);
    }
}

void launch_bi_interactions_cells(int dims, const int32_t* imove,
                                  const float* r, const float* normal,
                                  const float* u, const float* rho,
                                  const float* m, const float* p,
                                  float* grad_p, float* div_u,
                                  const uint8_t* near, const CellArgs& c,
                                  aq_defs d, float smax, hipStream_t s)
{
    // with the near flags, only the cells the fused walk listed as near
    const uint32_t* ol = (near && c.flagged) ? c.nlist : nullptr;
    if (dims == 2) {
        PolBiInteractions<2> q{ r, normal, u, rho, m, p, grad_p, div_u,
                                near, d.H, d.CONW };
        launch_cells<2, PolBiInteractions<2>, false,
                     PolBiInteractions<2>::BATCH>(q, imove, r, c, smax, s,
                                                  nullptr, ol, AQ_CL_NNCELLS);
    } else {
        PolBiInteractions<3> q{ r, normal, u, rho, m, p, grad_p, div_u,
                                near, d.H, d.CONW };
        launch_cells<3, PolBiInteractions<3>, false,
                     PolBiInteractions<3>::BATCH>(q, imove, r, c, smax, s,
                                                  nullptr, ol, AQ_CL_NNCELLS);
    }
}

void launch_les_gradu_cells(int dims, const int32_t* imove, const float* r,
                            const float* u, const float* rho, const float* m,
                            float* grad_u, const CellArgs& c, aq_defs d,
                            float smax, const float* per, hipStream_t s)
{
    const dim3 g(blocks_for(c.N, 256)), b(256);
    if (dims == 2) {
        hipLaunchKernelGGL(les_gradu_clear_k<2>, g, b, 0, s, imove, grad_u,
                           c.N);
        PolLesGradU<2> p{ r, u, rho, m, grad_u, d.H, d.CONF, per };
        if (per) launch_cells<2, PolLesGradU<2>, true>(p, imove, r, c, smax, s, per);
        else launch_cells<2>(p, imove, r, c, smax, s);
    } else {
        hipLaunchKernelGGL(les_gradu_clear_k<3>, g, b, 0, s, imove, grad_u,
                           c.N);
        PolLesGradU<3> p{ r, u, rho, m, grad_u, d.H, d.CONF, per };
        if (per) launch_cells<3, PolLesGradU<3>, true>(p, imove, r, c, smax, s, per);
        else launch_cells<3>(p, imove, r, c, smax, s);
    }
}

void launch_les_interactions_cells(int dims, const int32_t* imove,
                                   const float* r, const float* u,
                                   const float* rho, const float* m,
                                   const float* nu_t, float* dudt,
                                   const CellArgs& c, aq_defs d, float smax,
                                   const float* per, hipStream_t s)
{
    if (dims == 2) {
        PolLesInteractions<2> p{ r, u, rho, m, nu_t, dudt, d.H, d.CONF, per };
        if (per)
            launch_cells<2, PolLesInteractions<2>, true>(p, imove, r, c, smax,
                                                         s, per);
        else launch_cells<2>(p, imove, r, c, smax, s);
    } else {
        PolLesInteractions<3> p{ r, u, rho, m, nu_t, dudt, d.H, d.CONF, per };
        if (per)
            launch_cells<3, PolLesInteractions<3>, true>(p, imove, r, c, smax,
                                                         s, per);
        else launch_cells<3>(p, imove, r, c, smax, s);
    }
}

}  // namespace aq
