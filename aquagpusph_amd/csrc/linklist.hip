// Link-list: bounds, cell hash and head-of-chain table.
//
// Replaces the reference LinkList tool (aquagpusph/CalcServer/LinkList.cpp
// :326-494, kernels LinkList.cl.in:32-113).  The reference reads r_min and
// r_max back to the host every step to size the grid and reallocate ihoc;
// here the grid size n_cells is computed on the device with the same float
// arithmetic and kept in device memory, and ihoc lives in a buffer of fixed
// capacity, so the whole link-list can be replayed from a HIP graph.  An
// overflowing grid sets bit 0 of a sticky status word and the later kernels
// skip their work; the caller grows the capacity and re-runs.
#include <cmath>
#include <mutex>
#include <vector>

#include "aq_device.hpp"
#include "aq_internal.hpp"

namespace aq {

void radix_sort(uint32_t* keys, uint32_t* perm, uint32_t* inv, uint32_t n,
                unsigned key_bits, void* workspace, hipStream_t s);
size_t radix_sort_workspace(uint32_t n);
unsigned radix_key_bits(uint64_t max_val);

constexpr int MM_BS = 256;
constexpr int MM_MAXB = 1024;

template <int D>
__global__ void __launch_bounds__(MM_BS)
    minmax_partial(const float* __restrict__ r, uint32_t N,
                   float* __restrict__ part)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    float mn[VS], mx[VS];
#pragma unroll
    for (int k = 0; k < VS; k++) {
        mn[k] = (k == 3) ? 0.f : INFINITY;
        mx[k] = (k == 3) ? -0.f : -INFINITY;
    }
    for (uint32_t i = blockIdx.x * MM_BS + threadIdx.x; i < N;
         i += gridDim.x * MM_BS) {
#pragma unroll
        for (int k = 0; k < VS; k++) {
            const float v = r[(size_t)i * VS + k];
            mn[k] = cl_min(mn[k], v);
            mx[k] = cl_max(mx[k], v);
        }
    }
    __shared__ float smn[VS][MM_BS], smx[VS][MM_BS];
#pragma unroll
    for (int k = 0; k < VS; k++) {
        smn[k][threadIdx.x] = mn[k];
        smx[k][threadIdx.x] = mx[k];
    }
    __syncthreads();
    for (int o = MM_BS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
#pragma unroll
            for (int k = 0; k < VS; k++) {
                smn[k][threadIdx.x] =
                    cl_min(smn[k][threadIdx.x], smn[k][threadIdx.x + o]);
                smx[k][threadIdx.x] =
                    cl_max(smx[k][threadIdx.x], smx[k][threadIdx.x + o]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < VS; k++) {
            part[(size_t)blockIdx.x * 2 * VS + k] = smn[k][0];
            part[(size_t)blockIdx.x * 2 * VS + VS + k] = smx[k][0];
        }
    }
}

// final min/max + LinkList::nCells (LinkList.cpp:185-232) + capacity check
template <int D>
__global__ void __launch_bounds__(MM_BS)
    grid_final(const float* __restrict__ part, uint32_t nb, float* r_min,
               float* r_max, const float* support, const float* h,
               uint32_t* n_cells, uint32_t capacity, uint32_t* status,
               const float* per_min, const float* per_L)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    __shared__ float smn[VS][MM_BS], smx[VS][MM_BS];
    float mn[VS], mx[VS];
#pragma unroll
    for (int k = 0; k < VS; k++) {
        mn[k] = (k == 3) ? 0.f : INFINITY;
        mx[k] = (k == 3) ? -0.f : -INFINITY;
    }
    for (uint32_t b = threadIdx.x; b < nb; b += MM_BS) {
#pragma unroll
        for (int k = 0; k < VS; k++) {
            mn[k] = cl_min(mn[k], part[(size_t)b * 2 * VS + k]);
            mx[k] = cl_max(mx[k], part[(size_t)b * 2 * VS + VS + k]);
        }
    }
#pragma unroll
    for (int k = 0; k < VS; k++) {
        smn[k][threadIdx.x] = mn[k];
        smx[k][threadIdx.x] = mx[k];
    }
    __syncthreads();
    for (int o = MM_BS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
#pragma unroll
            for (int k = 0; k < VS; k++) {
                smn[k][threadIdx.x] =
                    cl_min(smn[k][threadIdx.x], smn[k][threadIdx.x + o]);
                smx[k][threadIdx.x] =
                    cl_max(smx[k][threadIdx.x], smx[k][threadIdx.x + o]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float lo[VS], hi[VS];
#pragma unroll
        for (int k = 0; k < VS; k++) {
            lo[k] = smn[k][0];
            hi[k] = smx[k][0];
            r_min[k] = lo[k];
            r_max[k] = hi[k];
        }
        const float cell_length = (*support) * (*h);
        uint64_t n[3] = { 1, 1, 1 };
        bool small = false;
#pragma unroll
        for (int k = 0; k < D; k++) {
            if (per_L && per_L[k] > 0.f) {
                // periodic axis: n cells of L / n >= the support tile the
                // box exactly (at least 3, so the wrapped neighbour cells
                // are distinct)
                n[k] = (uint64_t)cvt_u32(per_L[k] / cell_length);
                small = small || n[k] < 3;
                r_min[k] = per_min[k];
                r_max[k] = per_min[k] + per_L[k];
                continue;
            }
            const float v = (hi[k] - lo[k]) / cell_length;
            n[k] = (uint64_t)cvt_u32(v) + 6;
        }
        const uint64_t w = n[0] * n[1] * n[2];
        if (small) {
            *status |= 1u | 4u;  // a periodic box under 3 support lengths
        } else if (w > (uint64_t)capacity) {
            // keep the previous, consistent grid; flag the overflow
            *status |= 1u;
        } else {
            n_cells[0] = (uint32_t)n[0];
            n_cells[1] = (uint32_t)n[1];
            n_cells[2] = (uint32_t)n[2];
            n_cells[3] = (uint32_t)w;
        }
    }
}

// iCell, LinkList.cl.in:54-85
// Along a periodic axis (per_L[k] > 0) the cell coordinate is
// min(trunc((r - periodic_min) n / L), n - 1), with no margin cells.
template <int D>
__global__ void icell_k(uint32_t* __restrict__ icell,
                        const float* __restrict__ r, uint32_t N,
                        const float* r_min, const float* support,
                        const float* h, const uint32_t* n_cells,
                        const uint32_t* status, const float* per_min,
                        const float* per_L)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (status && (*status & 1u)) return;
    const float idist = 1.f / ((*support) * (*h));
    const uint32_t nx = n_cells[0], ny = n_cells[1];
    if (per_L) {
        constexpr int VS = (D == 2) ? 2 : 4;
        uint32_t q[3] = { 0u, 0u, 0u };
#pragma unroll
        for (int k = 0; k < D; k++) {
            const float x = r[(size_t)i * VS + k];
            if (per_L[k] > 0.f) {
                const uint32_t n = n_cells[k];
                const float inv = (float)n / per_L[k];
                q[k] = min(cvt_u32((x - per_min[k]) * inv), n - 1u);
            } else {
                q[k] = cvt_u32((x - r_min[k]) * idist) + 2u;
            }
        }
        icell[i] = q[0] + q[1] * nx + (D == 3 ? q[2] * nx * ny : 0u);
        return;
    }
    if constexpr (D == 2) {
        const float2 p = reinterpret_cast<const float2*>(r)[i];
        const uint32_t cx =
            cvt_u32((p.x - r_min[0]) * idist) + 3u;
        const uint32_t cy =
            cvt_u32((p.y - r_min[1]) * idist) + 3u;
        icell[i] = cx - 1u + (cy - 1u) * nx;
    } else {
        const float4 p = reinterpret_cast<const float4*>(r)[i];
        const uint32_t cx =
            cvt_u32((p.x - r_min[0]) * idist) + 3u;
        const uint32_t cy =
            cvt_u32((p.y - r_min[1]) * idist) + 3u;
        const uint32_t cz =
            cvt_u32((p.z - r_min[2]) * idist) + 3u;
        icell[i] = cx - 1u + (cy - 1u) * nx + (cz - 1u) * nx * ny;
    }
}

// iHoc, LinkList.cl.in:32-42
__global__ void ihoc_fill(uint32_t* __restrict__ ihoc, uint32_t N,
                          const uint32_t* n_cells, uint32_t capacity,
                          const uint32_t* status)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= capacity || c >= n_cells[3]) return;
    if (status && (*status & 1u)) return;
    ihoc[c] = N;
}

// linkList, LinkList.cl.in:92-113
__global__ void ihoc_heads(const uint32_t* __restrict__ icell,
                           uint32_t* __restrict__ ihoc, uint32_t N,
                           const uint32_t* status)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N - 1) return;
    if (status && (*status & 1u)) return;
    const uint32_t c = icell[i];
    const uint32_t c2 = icell[i + 1];
    if (i == 0) ihoc[c] = 0;
    if (c2 != c) ihoc[c2] = i + 1;
}

// ---- cell ranges for the cell-parallel walks --------------------------------
// head[i] = 1 when i starts a run of equal icell; cell_end[c] = one past the
// last particle of cell c (only written for non-empty cells); kword (the
// per-cell kind words cell_list_k accumulates) cleared.
__global__ void cell_heads_k(const uint32_t* __restrict__ icell, uint32_t N,
                             uint32_t* __restrict__ head,
                             uint32_t* __restrict__ cell_end,
                             uint32_t* __restrict__ kword,
                             const uint32_t* status)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (status && (*status & 1u)) {
        head[i] = 0;
        return;
    }
    if (kword) kword[i] = 0;
    const uint32_t c = icell[i];
    head[i] = (i == 0 || icell[i - 1] != c) ? 1u : 0u;
    if (i == N - 1 || icell[i + 1] != c) cell_end[c] = i + 1;
}

// heads[k] = first particle of the k-th non-empty cell.  With kword: the
// kinds of particle of each non-empty cell (AQ_CELL_FLUID / ELEM / OTHER
// bits) or-ed into kword[k] and its number of non-fluid particles added to
// the low bits (below 2^28), one pair of atomics per run of a cell's
// particles within a wavefront; imove is read through perm (id_unsorted):
// the cell list is built right after the sort, before the fields are
// reordered.
__global__ void cell_list_k(const uint32_t* __restrict__ head,
                            const uint32_t* __restrict__ pos, uint32_t N,
                            uint32_t* __restrict__ heads,
                            uint32_t* __restrict__ n_heads,
                            const uint32_t* __restrict__ perm,
                            const int* __restrict__ imove,
                            uint32_t* __restrict__ kword,
                            const uint32_t* status)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < N;
    uint32_t hd = 0, ps = 0;
    if (valid) {
        hd = head[i];
        ps = pos[i];
        if (hd) heads[ps] = i;
        if (i == N - 1) {
            *n_heads = ps + hd;
            // the owner-cell lists are refilled next
            n_heads[AQ_CL_NOCELLS] = 0;
            n_heads[AQ_CL_NWCELLS] = 0;
        }
    }
    if (!kword || (status && (*status & 1u))) return;  // uniform
    const int lane = threadIdx.x & 63;
    const int im = valid ? imove[perm[i]] : 1;
    const uint64_t vm = __ballot(valid);
    // a run starts at a cell's first particle or at the wavefront's first
    const uint64_t lead = __ballot(valid && (hd || lane == 0));
    const uint64_t fl = __ballot(valid && im == 1);
    const uint64_t el = __ballot(valid && im == -3);
    const uint64_t ot = __ballot(valid && im != 1 && im != -3);
    if (!valid || !((lead >> lane) & 1ull)) return;
    const uint64_t after = lane == 63 ? 0ull : lead >> (lane + 1);
    const int ln = after ? lane + 1 + __builtin_ctzll(after) : 64;
    const uint64_t run =
        (ln == 64 ? ~0ull : (1ull << ln) - 1ull) & ~((1ull << lane) - 1ull) & vm;
    uint32_t f = 0;
    if (fl & run) f |= AQ_CELL_FLUID;
    if (el & run) f |= AQ_CELL_ELEM;
    if (ot & run) f |= AQ_CELL_OTHER;
    const uint32_t k = ps + hd - 1u;
    atomicOr(kword + k, f);
    const uint32_t nonf = (uint32_t)__popcll((el | ot) & run);
    if (nonf) atomicAdd(kword + k, nonf);
}

// Reserves nent consecutive entries for each thread of the block in a list
// counted by *counter, with one atomic per block (the lists' counters are
// single words every block adds to); returns the thread's first entry.
// Every thread of the block calls it; blockDim.x is a multiple of 64, at
// most 1024.
__device__ uint32_t block_reserve(uint32_t nent, uint32_t* counter)
{
    __shared__ uint32_t wtot[16];
    __shared__ uint32_t bbase;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t x = nent;  // inclusive scan over the wavefront
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) wtot[wv] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int v = 0; v < (int)(blockDim.x >> 6); v++) {
            const uint32_t w = wtot[v];
            wtot[v] = t;
            t += w;
        }
        bbase = t ? atomicAdd(counter, t) : 0u;
    }
    __syncthreads();
    return bbase + wtot[wv] + x - nent;
}

// Kind flags of each non-empty cell into the top bits of cell_end
// (AQ_CELL_FLUID: imove == 1, AQ_CELL_ELEM: imove == -3, AQ_CELL_OTHER: any
// other imove; cell ends stay below 2^28), from the kind words of
// cell_list_k.  The cells with elements or other particles are appended to
// the owner-cell list as (first particle, cell, end particle with the cell's
// flags) triplets, one atomic per block (the list is ordered within a
// block's cells, the blocks in no set order): the cells of at most
// AQ_OWNER_CHUNK non-fluid particles as one entry, the others as one entry
// per run of particles ending at every AQ_OWNER_CHUNK-th non-fluid one (the
// last run ends at the cell's end).
__global__ void cell_kinds_k(const uint32_t* __restrict__ heads,
                             uint32_t* __restrict__ n_heads,
                             const uint32_t* __restrict__ icell,
                             const uint32_t* __restrict__ perm,
                             const int* __restrict__ imove,
                             uint32_t* __restrict__ cell_end,
                             const uint32_t* __restrict__ kword,
                             const uint32_t* status,
                             uint32_t* __restrict__ ocells)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nh = *n_heads;
    if (status && (*status & 1u)) return;           // uniform
    if (blockIdx.x * blockDim.x >= nh) return;      // uniform per block
    constexpr uint32_t CH = AQ_OWNER_CHUNK;
    uint32_t f = 0, start = 0, c = 0, end = 0, nonf = 0;
    if (k < nh) {
        start = heads[k];
        c = icell[start];
        end = cell_end[c];
        const uint32_t kw = kword[k];
        f = kw & (AQ_CELL_FLUID | AQ_CELL_ELEM | AQ_CELL_OTHER);
        nonf = kw & AQ_CELL_END;
        cell_end[c] = end | f;
    }
    if (!ocells) return;
    const bool own = (f & (AQ_CELL_ELEM | AQ_CELL_OTHER)) != 0u;
    const uint32_t nent = !own ? 0u : (nonf > CH ? (nonf + CH - 1u) / CH : 1u);
    const uint32_t e0 = block_reserve(nent, n_heads + AQ_CL_NOCELLS);
    if (own && nonf <= CH) {
        ocells[3 * e0] = start;
        ocells[3 * e0 + 1] = c;
        ocells[3 * e0 + 2] = end | f;
    }
    // the cells split into runs: each lane scans its cell, 16 particles a
    // round trip; run m ends at the (m + 1) CH-th non-fluid particle
    if (own && nonf > CH) {
        const uint32_t elast = e0 + nent - 1u;
        uint32_t e = e0, cnt = 0;
        ocells[3 * e] = start;
        ocells[3 * e + 1] = c;
        for (uint32_t p0 = start; p0 < end && e < elast; p0 += 16) {
            int im[16];
#pragma unroll
            for (int u = 0; u < 16; u++)
                im[u] = p0 + u < end ? imove[perm[p0 + u]] : 1;
#pragma unroll
            for (int u = 0; u < 16; u++) {
                if (im[u] == 1 || e >= elast || ++cnt < CH) continue;
                ocells[3 * e + 2] = (p0 + u + 1u) | f;
                e++;
                ocells[3 * e] = p0 + u + 1u;
                ocells[3 * e + 1] = c;
                cnt = 0;
            }
        }
        ocells[3 * elast + 2] = end | f;
    }
}

// Wet owner cells: the owner-list entries (ocells) whose cell has fluid in
// its 3^d neighbourhood, appended to wcells as the same triplets (count at
// n_heads[AQ_CL_NWCELLS]).  The element kernels whose candidates are fluid
// (BI LapU, BI Interpolation) have no term to add in the other cells.
template <int D>
__global__ void owner_wet_k(const uint32_t* __restrict__ ocells,
                            uint32_t* __restrict__ n_heads,
                            const uint32_t* __restrict__ ihoc,
                            const uint32_t* __restrict__ cell_end,
                            const uint32_t* __restrict__ n_cells, uint32_t N,
                            uint32_t* __restrict__ wcells,
                            const uint32_t* status)
{
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (status && (*status & 1u)) return;  // uniform
    const uint32_t n = n_heads[AQ_CL_NOCELLS];
    if (blockIdx.x * blockDim.x >= n) return;  // uniform per block
    bool wet = false;
    uint32_t t0 = 0, t1 = 0, t2 = 0;
    if (m < n) {
        t0 = ocells[3 * m];
        t1 = ocells[3 * m + 1];
        t2 = ocells[3 * m + 2];
        const uint32_t nx = n_cells[0], ny = n_cells[1];
        constexpr int NC = (D == 2) ? 9 : 27;
        for (int k = 0; k < NC && !wet; k++) {
            int ci, cj, ck;
            if (D == 2) {
                ci = k / 3 - 1;
                cj = k % 3 - 1;
                ck = 0;
            } else {
                ci = k / 9 - 1;
                cj = (k / 3) % 3 - 1;
                ck = k % 3 - 1;
            }
            const uint32_t c = t1 + (uint32_t)ci + (uint32_t)cj * nx +
                               (uint32_t)ck * nx * ny;
            if (ihoc[c] < N && (cell_end[c] & AQ_CELL_FLUID)) wet = true;
        }
    }
    const uint32_t e = block_reserve(wet ? 1u : 0u, n_heads + AQ_CL_NWCELLS);
    if (wet) {
        const uint32_t o = 3u * e;
        wcells[o] = t0;
        wcells[o + 1] = t1;
        wcells[o + 2] = t2;
    }
}

// Blocked order of the non-empty cells for the fused walk: the cells are
// bucketed by blocks of BLK^d grid cells (8^3 in 3D, 16^2 in 2D) into
// bcells, as (first particle, cell, cell_end) triplets, so the cells handed
// out together lie in one compact region and the neighbourhoods in flight
// on one XCD share its L2 (the row-major cell list puts a few long rows of
// cells in flight instead).  Within a block the order is arbitrary; the
// walk's sums do not depend on the order in which cells are visited.
template <int D> struct Blk {
    static constexpr uint32_t B = (D == 2) ? 16u : 8u;
};
template <int D>
__device__ __forceinline__ uint32_t cell_block(uint32_t c,
                                               const uint32_t* n_cells,
                                               uint32_t nbmax)
{
    constexpr uint32_t B = Blk<D>::B;
    const uint32_t nx = n_cells[0], ny = n_cells[1];
    const uint32_t cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
    const uint32_t nbx = (nx + B - 1) / B, nby = (ny + B - 1) / B;
    const uint32_t b = (cz / B * nby + cy / B) * nbx + cx / B;
    return b < nbmax ? b : nbmax - 1u;
}

// Runs of consecutive cells of one block within a wavefront (the cells come
// in row-major order, so a row crosses a block every B cells): the lanes
// below nv hold cells; lane l is in the run led by lane *lead_lane, of
// *len cells (set on the leading lane only).
__device__ __forceinline__ bool block_run(uint32_t b, int lane, uint32_t nv,
                                          int* lead_lane, uint32_t* len)
{
    const uint32_t prev = __shfl_up(b, 1);
    const bool valid = (uint32_t)lane < nv;
    const bool lead = valid && (lane == 0 || prev != b);
    const uint64_t L = __ballot(lead);
    const uint64_t upto = lane == 63 ? ~0ull : (2ull << lane) - 1ull;
    *lead_lane = 63 - __builtin_clzll((L & upto) | 1ull);
    const uint64_t after = lane == 63 ? 0ull : L >> (lane + 1);
    const uint32_t ln = after ? (uint32_t)(lane + 1 + __builtin_ctzll(after))
                              : 64u;
    *len = (ln < nv ? ln : nv) - (uint32_t)lane;
    return lead;
}

template <int D>
__global__ void blk_count_k(const uint32_t* __restrict__ heads,
                            const uint32_t* __restrict__ n_heads,
                            const uint32_t* __restrict__ icell,
                            const uint32_t* __restrict__ n_cells,
                            uint32_t* __restrict__ bcount, uint32_t nbmax,
                            const uint32_t* status)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (status && (*status & 1u)) return;            // uniform
    const uint32_t nh = *n_heads;
    const uint32_t k0 = k - (uint32_t)(threadIdx.x & 63);
    if (k0 >= nh) return;                            // uniform per wavefront
    const int lane = threadIdx.x & 63;
    const uint32_t nv = nh - k0 < 64u ? nh - k0 : 64u;
    const uint32_t b =
        k < nh ? cell_block<D>(icell[heads[k]], n_cells, nbmax) : 0u;
    int ll;
    uint32_t len;
    if (block_run(b, lane, nv, &ll, &len)) atomicAdd(bcount + b, len);
}

// bcount counts down to zero again, ready for the next step
template <int D>
__global__ void blk_scatter_k(const uint32_t* __restrict__ heads,
                              const uint32_t* __restrict__ n_heads,
                              const uint32_t* __restrict__ icell,
                              const uint32_t* __restrict__ n_cells,
                              const uint32_t* __restrict__ cell_end,
                              uint32_t* __restrict__ bcount,
                              const uint32_t* __restrict__ bbase,
                              uint32_t nbmax, uint32_t* __restrict__ bcells,
                              const uint32_t* status)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (status && (*status & 1u)) return;            // uniform
    const uint32_t nh = *n_heads;
    const uint32_t k0 = k - (uint32_t)(threadIdx.x & 63);
    if (k0 >= nh) return;                            // uniform per wavefront
    const int lane = threadIdx.x & 63;
    const uint32_t nv = nh - k0 < 64u ? nh - k0 : 64u;
    uint32_t start = 0, c = 0, b = 0;
    if (k < nh) {
        start = heads[k];
        c = icell[start];
        b = cell_block<D>(c, n_cells, nbmax);
    }
    int ll;
    uint32_t len, base = 0;
    // a run takes the top len places still free in its block
    if (block_run(b, lane, nv, &ll, &len))
        base = bbase[b] + atomicSub(bcount + b, len) - len;
    base = __shfl(base, ll);
    if (k >= nh) return;
    const uint32_t pos = base + (uint32_t)(lane - ll);
    bcells[3 * pos] = start;
    bcells[3 * pos + 1] = c;
    bcells[3 * pos + 2] = cell_end[c];
}

// Blocks of the blocked cell order a grid of `cap` cells can hold (a bound
// on the product of ceil(n_i / B) over the axes).
size_t cell_blocks_bound(int dims, size_t cap)
{
    const double c = (double)cap;
    if (dims == 2) return (size_t)(c / 256.0 + 2.0 * std::sqrt(c) / 16.0 + 2.0);
    return (size_t)(c / 512.0 + std::pow(c, 2.0 / 3.0) * 3.0 / 64.0 +
                    std::cbrt(c) * 3.0 / 8.0 + 2.0);
}

void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t m,
                    uint32_t* sums, hipStream_t s);
size_t scan_workspace(uint32_t m);

size_t cell_list_workspace(uint32_t N)
{
    return (((size_t)N * 4 + 255) & ~(size_t)255) * 3 + scan_workspace(N) +
           256;
}

// Non-empty cells of a sorted icell: heads[k] = first particle of the k-th
// non-empty cell, *n_heads their number, cell_end[c] its end and, when
// imove and perm are given, the kinds of particle it holds and (with
// ocells, n_heads of AQ_CL_WORDS words) the owner-cell list.
void cell_list(const uint32_t* icell, uint32_t N, uint32_t* cell_end,
               uint32_t* heads, uint32_t* n_heads, const uint32_t* status,
               void* ws, const uint32_t* perm, const int* imove,
               hipStream_t s, uint32_t* ocells, uint32_t* wcells, int dims,
               const uint32_t* ihoc, const uint32_t* n_cells, size_t ocap,
               uint32_t* bcells, uint32_t* bcount, uint32_t* bbase,
               uint32_t* bsums, uint32_t nbmax)
{
    const size_t a = ((size_t)N * 4 + 255) & ~(size_t)255;
    uint32_t* head = (uint32_t*)ws;
    uint32_t* pos = (uint32_t*)((char*)ws + a);
    uint32_t* kword = (perm && imove) ? (uint32_t*)((char*)ws + 2 * a)
                                      : nullptr;
    uint32_t* sums = (uint32_t*)((char*)ws + 3 * a);
    const dim3 g(blocks_for(N, 256)), b(256);
    hipLaunchKernelGGL(cell_heads_k, g, b, 0, s, icell, N, head, cell_end,
                       kword, status);
    exclusive_scan(head, pos, N, sums, s);
    hipLaunchKernelGGL(cell_list_k, g, b, 0, s, head, pos, N, heads, n_heads,
                       perm, imove, kword, status);
    // 1024-thread blocks: the list appends take one atomic per block
    const dim3 b4(1024);
    if (kword)
        hipLaunchKernelGGL(cell_kinds_k, dim3(blocks_for(N, 1024)), b4, 0, s,
                           heads, n_heads, icell, perm, imove, cell_end,
                           kword, status, ocells);
    if (perm && imove && ocells && wcells && ocap) {
        const dim3 gw(blocks_for(aq_owner_list_cap(ocap, N), 1024));
        if (dims == 2)
            hipLaunchKernelGGL(owner_wet_k<2>, gw, b4, 0, s, ocells, n_heads,
                               ihoc, cell_end, n_cells, N, wcells, status);
        else
            hipLaunchKernelGGL(owner_wet_k<3>, gw, b4, 0, s, ocells, n_heads,
                               ihoc, cell_end, n_cells, N, wcells, status);
    }
    if (bcells && ocap && nbmax) {
        // the blocked order (after the kind flags, which it copies)
        const dim3 gb(blocks_for(ocap, 256));
        if (dims == 2)
            hipLaunchKernelGGL(blk_count_k<2>, gb, b, 0, s, heads, n_heads,
                               icell, n_cells, bcount, nbmax, status);
        else
            hipLaunchKernelGGL(blk_count_k<3>, gb, b, 0, s, heads, n_heads,
                               icell, n_cells, bcount, nbmax, status);
        exclusive_scan(bcount, bbase, nbmax, bsums, s);
        if (dims == 2)
            hipLaunchKernelGGL(blk_scatter_k<2>, gb, b, 0, s, heads, n_heads,
                               icell, n_cells, cell_end, bcount, bbase, nbmax,
                               bcells, status);
        else
            hipLaunchKernelGGL(blk_scatter_k<3>, gb, b, 0, s, heads, n_heads,
                               icell, n_cells, cell_end, bcount, bbase, nbmax,
                               bcells, status);
    }
}

// ---- host drivers ---------------------------------------------------------
size_t linklist_workspace(uint32_t N)
{
    const size_t mm = (size_t)MM_MAXB * 2 * 4 * sizeof(float);
    return ((mm + 255) & ~(size_t)255) + radix_sort_workspace(N);
}

void linklist_grid(int dims, const float* r, uint32_t N, float* r_min,
                   float* r_max, const float* support, const float* h,
                   uint32_t* n_cells, uint32_t cap, uint32_t* status,
                   void* ws, hipStream_t s, const float* per_min = nullptr,
                   const float* per_L = nullptr)
{
    const unsigned nb =
        std::min<unsigned>(MM_MAXB, blocks_for(N, MM_BS * 8));
    float* part = (float*)ws;
    if (dims == 2) {
        hipLaunchKernelGGL(minmax_partial<2>, dim3(nb), dim3(MM_BS), 0, s, r,
                           N, part);
        hipLaunchKernelGGL(grid_final<2>, dim3(1), dim3(MM_BS), 0, s, part,
                           nb, r_min, r_max, support, h, n_cells, cap,
                           status, per_min, per_L);
    } else {
        hipLaunchKernelGGL(minmax_partial<3>, dim3(nb), dim3(MM_BS), 0, s, r,
                           N, part);
        hipLaunchKernelGGL(grid_final<3>, dim3(1), dim3(MM_BS), 0, s, part,
                           nb, r_min, r_max, support, h, n_cells, cap,
                           status, per_min, per_L);
    }
}

void linklist_icell(int dims, uint32_t* icell, const float* r, uint32_t N,
                    const float* r_min, const float* support, const float* h,
                    const uint32_t* n_cells, const uint32_t* status,
                    hipStream_t s, const float* per_min = nullptr,
                    const float* per_L = nullptr)
{
    const dim3 g(blocks_for(N, 256)), b(256);
    if (dims == 2)
        hipLaunchKernelGGL(icell_k<2>, g, b, 0, s, icell, r, N, r_min,
                           support, h, n_cells, status, per_min, per_L);
    else
        hipLaunchKernelGGL(icell_k<3>, g, b, 0, s, icell, r, N, r_min,
                           support, h, n_cells, status, per_min, per_L);
}

void linklist_ihoc(const uint32_t* icell, uint32_t* ihoc, uint32_t N,
                   const uint32_t* n_cells, uint32_t cap,
                   const uint32_t* status, hipStream_t s)
{
    hipLaunchKernelGGL(ihoc_fill, dim3(blocks_for(cap, 256)), dim3(256), 0, s,
                       ihoc, N, n_cells, cap, status);
    if (N > 1)
        hipLaunchKernelGGL(ihoc_heads, dim3(blocks_for(N - 1, 256)),
                           dim3(256), 0, s, icell, ihoc, N, status);
}

void linklist_all(int dims, const float* r, uint32_t N, float* r_min,
                  float* r_max, const float* support, const float* h,
                  uint32_t* n_cells, uint32_t* icell, uint32_t* id_unsorted,
                  uint32_t* id_sorted, uint32_t* ihoc, uint32_t cap,
                  uint32_t* status, void* ws, hipStream_t s,
                  const float* per_min, const float* per_L)
{
    linklist_grid(dims, r, N, r_min, r_max, support, h, n_cells, cap, status,
                  ws, s, per_min, per_L);
    linklist_icell(dims, icell, r, N, r_min, support, h, n_cells, status, s,
                   per_min, per_L);
    const size_t mm = (size_t)MM_MAXB * 2 * 4 * sizeof(float);
    void* sws = (char*)ws + ((mm + 255) & ~(size_t)255);
    // icell < n_cells.w <= cap, so sorting on the bits of cap gives the
    // reference result (it sorts on the bits of n_cells.w)
    radix_sort(icell, id_unsorted, id_sorted, N, radix_key_bits(cap), sws, s);
    linklist_ihoc(icell, ihoc, N, n_cells, cap, status, s);
}

// ---- the link-list without allocation -------------------------------------
// Scratch of aq_linklist_noalloc: the workspace aq_linklist takes as an
// argument (min/max partials, the radix sort's key and value buffers, digit
// counts and scan sums) for up to AQ_LINKLIST_NOALLOC_MAX_N particles, in a
// file-scope device array.  Each GPU holds its own copy, at its own
// address, until the process ends; host code uses that address, looked up
// once per device, never the array's name.  Every call on one GPU shares
// the array and keeps the last call's contents: calls that overlap on one
// GPU race.  Nothing in it has to start at zero: each kernel of the
// link-list writes what a later one reads.
constexpr size_t NOALLOC_BYTES = (size_t)9 << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];

static void* scratch_ptr()
{
    static std::mutex mu;
    static std::vector<void*> ptrs;
    int dev = 0;
    AQ_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((size_t)dev >= ptrs.size()) ptrs.resize(dev + 1, nullptr);
    if (!ptrs[dev]) {
        void* p = nullptr;
        AQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        AQ_REQUIRE(p, "aq_linklist_noalloc: no address for the scratch "
                      "array on this device");
        ptrs[dev] = p;
    }
    return ptrs[dev];
}

}  // namespace aq

using namespace aq;

extern "C" {

AQ_API size_t aq_linklist_workspace(uint32_t N) { return linklist_workspace(N); }

AQ_API int aq_linklist_noalloc(int dims, const float* r, uint32_t N,
                               float* r_min, float* r_max,
                               const float* support, const float* h,
                               uint32_t* n_cells, uint32_t* icell,
                               uint32_t* id_unsorted, uint32_t* id_sorted,
                               uint32_t* ihoc, uint32_t ihoc_capacity,
                               uint32_t* status, void* stream)
{
    return guard([&] {
        check_dims(dims);
        AQ_REQUIRE(N > 0, "aq_linklist_noalloc: N must be > 0");
        if (N > AQ_LINKLIST_NOALLOC_MAX_N)
            throw Error(AQ_EINVAL,
                        "aq_linklist_noalloc: N = " + std::to_string(N) +
                            " exceeds the no-allocation limit of " +
                            std::to_string(AQ_LINKLIST_NOALLOC_MAX_N) +
                            " particles (use aq_linklist with a workspace)");
        AQ_REQUIRE(r && r_min && r_max && support && h && n_cells && icell &&
                       id_unsorted && id_sorted && ihoc && status,
                   "aq_linklist_noalloc: null pointer");
        AQ_REQUIRE(linklist_workspace(N) <= NOALLOC_BYTES,
                   "aq_linklist_noalloc: scratch smaller than the workspace");
        linklist_all(dims, r, N, r_min, r_max, support, h, n_cells, icell,
                     id_unsorted, id_sorted, ihoc, ihoc_capacity, status,
                     scratch_ptr(), S(stream), nullptr, nullptr);
    });
}

AQ_API int aq_linklist_grid(int dims, const float* r, uint32_t N,
                            float* r_min, float* r_max, const float* support,
                            const float* h, uint32_t* n_cells,
                            uint32_t ihoc_capacity, uint32_t* status,
                            void* workspace, void* stream)
{
    return guard([&] {
        check_dims(dims);
        AQ_REQUIRE(N > 0, "aq_linklist_grid: N must be > 0");
        AQ_REQUIRE(r && r_min && r_max && support && h && n_cells && status &&
                       workspace,
                   "aq_linklist_grid: null pointer");
        linklist_grid(dims, r, N, r_min, r_max, support, h, n_cells,
                      ihoc_capacity, status, workspace, S(stream));
    });
}

AQ_API int aq_linklist_icell(int dims, uint32_t* icell, const float* r,
                             uint32_t N, const float* r_min,
                             const float* support, const float* h,
                             const uint32_t* n_cells, const uint32_t* status,
                             void* stream)
{
    return guard([&] {
        check_dims(dims);
        AQ_REQUIRE(icell && r && r_min && support && h && n_cells,
                   "aq_linklist_icell: null pointer");
        if (N) linklist_icell(dims, icell, r, N, r_min, support, h, n_cells,
                              status, S(stream));
    });
}

AQ_API int aq_linklist_ihoc(const uint32_t* icell, uint32_t* ihoc,
                            uint32_t N, const uint32_t* n_cells,
                            uint32_t ihoc_capacity, const uint32_t* status,
                            void* stream)
{
    return guard([&] {
        AQ_REQUIRE(icell && ihoc && n_cells, "aq_linklist_ihoc: null pointer");
        AQ_REQUIRE(N > 0, "aq_linklist_ihoc: N must be > 0");
        linklist_ihoc(icell, ihoc, N, n_cells, ihoc_capacity, status,
                      S(stream));
    });
}

AQ_API int aq_linklist(int dims, const float* r, uint32_t N, float* r_min,
                       float* r_max, const float* support, const float* h,
                       uint32_t* n_cells, uint32_t* icell,
                       uint32_t* id_unsorted, uint32_t* id_sorted,
                       uint32_t* ihoc, uint32_t ihoc_capacity,
                       uint32_t* status, void* workspace, void* stream)
{
    return guard([&] {
        check_dims(dims);
        AQ_REQUIRE(N > 0, "aq_linklist: N must be > 0");
        AQ_REQUIRE(r && r_min && r_max && support && h && n_cells && icell &&
                       id_unsorted && id_sorted && ihoc && status &&
                       workspace,
                   "aq_linklist: null pointer");
        linklist_all(dims, r, N, r_min, r_max, support, h, n_cells, icell,
                     id_unsorted, id_sorted, ihoc, ihoc_capacity, status,
                     workspace, S(stream), nullptr, nullptr);
    });
}

}  // extern "C"
