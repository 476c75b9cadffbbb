// Stable LSD radix sort (8-bit digits) and the permutation helpers.
//
// Replaces the reference RadixSort tool (aquagpusph/CalcServer/RadixSort.cpp
// :129-303, kernels RadixSort.cl.in:35-323), which sorts 4-bit digits with
// 1024 work-items that each scan a contiguous chunk serially.  Here each
// pass is reduce-then-scan over tiles of 4096 keys:
//   1. rs_hist    : per-tile digit histogram (wave-aggregated LDS counters),
//                   stored digit-major so one exclusive scan yields, for each
//                   (digit, tile), the first output slot;
//   2. scan       : three-kernel exclusive scan of the digit-major table;
//   3. rs_scatter : tiles re-read their keys in index order; a key's rank
//                   among equal digits of its 256-key row comes from eight
//                   wavefront ballots (match-any), rows are processed in
//                   order, so equal digits keep their input order (stable).
// No kernel waits on another workgroup, so every wave always finishes.
#include <algorithm>

#include "aq_internal.hpp"

namespace aq {

constexpr int RS_BS = 256;  // threads per tile (4 wave64)
constexpr int RS_IPT = 16;  // keys per thread
constexpr int RS_TILE = RS_BS * RS_IPT;
constexpr int SC_BS = 256;
constexpr int SC_IPT = 4;
constexpr int SC_TILE = SC_BS * SC_IPT;

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// lanes of the wavefront whose 8-bit digit equals this lane's
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t active)
{
    uint64_t m = active;
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

__global__ void __launch_bounds__(RS_BS)
    rs_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift,
            uint32_t* __restrict__ counts, uint32_t ntiles)
{
    __shared__ uint32_t h[256];
    const int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * RS_TILE;
    const uint64_t lt = (1ull << lane_id()) - 1ull;
#pragma unroll 4
    for (int it = 0; it < RS_IPT; it++) {
        const size_t idx = base + (size_t)it * RS_BS + tid;
        const bool valid = idx < n;
        const uint32_t d = valid ? (keys[idx] >> shift) & 255u : 0u;
        const uint64_t act = __ballot(valid);
        const uint64_t peers = match_digit(d, act);
        if (valid && (peers & lt) == 0)
            atomicAdd(&h[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    counts[(size_t)tid * ntiles + blockIdx.x] = h[tid];
}

__global__ void __launch_bounds__(RS_BS)
    rs_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
               uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
               uint32_t n, int shift, const uint32_t* __restrict__ offs,
               uint32_t ntiles)
{
    __shared__ uint32_t base[256];
    __shared__ uint32_t wc[4][256];
    const int tid = threadIdx.x;
    const int w = tid >> 6;
    const uint64_t lt = (1ull << lane_id()) - 1ull;
    base[tid] = offs[(size_t)tid * ntiles + blockIdx.x];
    const size_t tile = (size_t)blockIdx.x * RS_TILE;
    for (int it = 0; it < RS_IPT; it++) {
#pragma unroll
        for (int k = 0; k < 4; k++) wc[k][tid] = 0;
        __syncthreads();
        const size_t idx = tile + (size_t)it * RS_BS + tid;
        const bool valid = idx < n;
        const uint32_t key = valid ? kin[idx] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        const uint64_t act = __ballot(valid);
        const uint64_t peers = match_digit(d, act);
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && (peers & lt) == 0) wc[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = base[d] + rank;
            for (int k = 0; k < w; k++) pos += wc[k][d];
            kout[pos] = key;
            vout[pos] = vin ? vin[idx] : (uint32_t)idx;
        }
        __syncthreads();
        base[tid] += wc[0][tid] + wc[1][tid] + wc[2][tid] + wc[3][tid];
    }
}

// ---- exclusive scan of uint32 ----------------------------------------------
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* ws,
                                                    uint32_t* total)
{
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    const int nw = blockDim.x >> 6;
    for (int k = 0; k < nw; k++) {
        if (k < w) off += ws[k];
        tot += ws[k];
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

__global__ void __launch_bounds__(SC_BS)
    scan_reduce(const uint32_t* __restrict__ in, uint32_t m,
                uint32_t* __restrict__ sums)
{
    __shared__ uint32_t ws[SC_BS / 64];
    const size_t base = (size_t)blockIdx.x * SC_TILE + threadIdx.x * SC_IPT;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SC_IPT; k++)
        if (base + k < m) s += in[base + k];
    uint32_t tot;
    block_excl_scan(s, ws, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// single workgroup: exclusive scan of the per-tile sums, any length
__global__ void __launch_bounds__(1024) scan_top(uint32_t* sums, uint32_t nb)
{
    __shared__ uint32_t ws[16];
    uint32_t carry = 0;
    for (uint32_t b = 0; b < nb; b += 1024) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t v = (i < nb) ? sums[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(v, ws, &tot);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
    }
}

__global__ void __launch_bounds__(SC_BS)
    scan_down(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
              uint32_t m, const uint32_t* __restrict__ sums)
{
    __shared__ uint32_t ws[SC_BS / 64];
    const size_t base = (size_t)blockIdx.x * SC_TILE + threadIdx.x * SC_IPT;
    uint32_t v[SC_IPT];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SC_IPT; k++) {
        v[k] = (base + k < m) ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t x = block_excl_scan(s, ws, &tot) + sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SC_IPT; k++) {
        if (base + k < m) out[base + k] = x;
        x += v[k];
    }
}

// single workgroup: exclusive scan of a short array in chunks of 1024
// (in == out allowed: a chunk is read before any of it is written)
__global__ void __launch_bounds__(1024)
    scan_one(const uint32_t* in, uint32_t* out, uint32_t m)
{
    __shared__ uint32_t ws[16];
    uint32_t carry = 0;
    for (uint32_t b = 0; b < m; b += 1024) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t v = (i < m) ? in[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(v, ws, &tot);
        if (i < m) out[i] = carry + ex;
        carry += tot;
    }
}

size_t scan_workspace(uint32_t m) { return (size_t)blocks_for(m, SC_TILE) * 4; }

void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t m,
                    uint32_t* sums, hipStream_t s)
{
    if (m <= 16384) {
        // one launch for short arrays (the blocked cell order's counts)
        if (m) hipLaunchKernelGGL(scan_one, dim3(1), dim3(1024), 0, s, in, out, m);
        return;
    }
    const unsigned nb = blocks_for(m, SC_TILE);
    hipLaunchKernelGGL(scan_reduce, dim3(nb), dim3(SC_BS), 0, s, in, m, sums);
    hipLaunchKernelGGL(scan_top, dim3(1), dim3(1024), 0, s, sums, nb);
    hipLaunchKernelGGL(scan_down, dim3(nb), dim3(SC_BS), 0, s, in, out, m,
                       sums);
}

// ---- permutations -----------------------------------------------------------
__global__ void inverse_perm(const uint32_t* __restrict__ perm,
                             uint32_t* __restrict__ inv, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inv[perm[i]] = i;
}

template <typename T>
__global__ void permute_k(const T* __restrict__ src, T* __restrict__ dst,
                          const uint32_t* __restrict__ perm, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

// ---- driver -------------------------------------------------------------
static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t radix_sort_workspace(uint32_t n)
{
    const size_t nt = blocks_for(n, RS_TILE);
    const size_t m = nt * 256;
    return align256((size_t)n * 4) * 2 + align256(m * 4) +
           align256(scan_workspace((uint32_t)m));
}

void radix_sort(uint32_t* keys, uint32_t* perm, uint32_t* inv, uint32_t n,
                unsigned key_bits, void* workspace, hipStream_t s)
{
    if (n == 0) return;
    const uint32_t nt = blocks_for(n, RS_TILE);
    const uint32_t m = nt * 256;
    char* w = (char*)workspace;
    uint32_t* kb = (uint32_t*)w;
    w += align256((size_t)n * 4);
    uint32_t* vb = (uint32_t*)w;
    w += align256((size_t)n * 4);
    uint32_t* counts = (uint32_t*)w;
    w += align256((size_t)m * 4);
    uint32_t* sums = (uint32_t*)w;
    const int passes = key_bits == 0 ? 1 : (int)((key_bits + 7) / 8);
    // The keys alternate between kb (even passes: the first cannot write
    // keys in place) and keys; the values between vb and perm, chosen so
    // that the last pass writes perm (the first pass reads no values).
    uint32_t *kin = keys, *vin = nullptr, *kout = kb, *vout = nullptr;
    for (int p = 0; p < passes; p++) {
        const int shift = 8 * p;
        kout = (p % 2 == 0) ? kb : keys;
        vout = ((passes - 1 - p) % 2 == 0) ? perm : vb;
        hipLaunchKernelGGL(rs_hist, dim3(nt), dim3(RS_BS), 0, s, kin, n,
                           shift, counts, nt);
        exclusive_scan(counts, counts, m, sums, s);
        hipLaunchKernelGGL(rs_scatter, dim3(nt), dim3(RS_BS), 0, s, kin, vin,
                           kout, vout, n, shift, counts, nt);
        kin = kout;
        vin = vout;
    }
    // the sorted keys are in kb after an odd number of passes
    if (kin != keys)
        AQ_HIP(hipMemcpyAsync(keys, kin, (size_t)n * 4,
                              hipMemcpyDeviceToDevice, s));
    if (inv)
        hipLaunchKernelGGL(inverse_perm, dim3(blocks_for(n, 256)), dim3(256),
                           0, s, perm, inv, n);
}

struct alignas(16) Words8 {
    uint4 a, b;
};

// Several gathers by one permutation in one pass (the perm entry read
// once): dst_f[i] = src_f[perm[i]] for the fields f < n, of 4, 8 or 16
// bytes per element.  Every field's loads are issued before its stores.
constexpr int PERM_BATCH = 8;
struct PermBatch {
    const void* src[PERM_BATCH];
    void* dst[PERM_BATCH];
    int esize[PERM_BATCH];
    int n;
};

__global__ void permute_batch_k(PermBatch b, const uint32_t* __restrict__ perm,
                                uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = perm[i];
    uint4 v[PERM_BATCH];
#pragma unroll
    for (int f = 0; f < PERM_BATCH; f++) {
        if (f >= b.n) break;
        const int es = b.esize[f];
        if (es == 16) {
            v[f] = ((const uint4*)b.src[f])[p];
        } else if (es == 8) {
            const uint2 t = ((const uint2*)b.src[f])[p];
            v[f] = make_uint4(t.x, t.y, 0u, 0u);
        } else {
            v[f] = make_uint4(((const uint32_t*)b.src[f])[p], 0u, 0u, 0u);
        }
    }
#pragma unroll
    for (int f = 0; f < PERM_BATCH; f++) {
        if (f >= b.n) break;
        const int es = b.esize[f];
        if (es == 16)
            ((uint4*)b.dst[f])[i] = v[f];
        else if (es == 8)
            ((uint2*)b.dst[f])[i] = make_uint2(v[f].x, v[f].y);
        else
            ((uint32_t*)b.dst[f])[i] = v[f].x;
    }
}

// Gathers (src[f] -> dst[f], esize[f] bytes per element) by one
// permutation.  Fields of 4, 8 or 16 bytes, none of whose destinations is
// another field's source, go through permute_batch_k, up to PERM_BATCH at
// a time; the others one by one, in order after them.
void permute(const void* src, void* dst, size_t esize, const uint32_t* perm,
             uint32_t n, hipStream_t s);
void permute_many(const void* const* src, void* const* dst,
                  const size_t* esize, int nf, const uint32_t* perm,
                  uint32_t n, hipStream_t s)
{
    if (n == 0 || nf <= 0) return;
    bool ok = true;
    for (int f = 0; f < nf && ok; f++) {
        ok = (esize[f] == 4 || esize[f] == 8 || esize[f] == 16) &&
             src[f] != dst[f];
        for (int g = 0; g < nf && ok; g++) ok = dst[f] != src[g];
    }
    if (!ok) {
        for (int f = 0; f < nf; f++) permute(src[f], dst[f], esize[f], perm, n, s);
        return;
    }
    for (int f0 = 0; f0 < nf; f0 += PERM_BATCH) {
        PermBatch b{};
        b.n = std::min(nf - f0, PERM_BATCH);
        for (int f = 0; f < b.n; f++) {
            b.src[f] = src[f0 + f];
            b.dst[f] = dst[f0 + f];
            b.esize[f] = (int)esize[f0 + f];
        }
        hipLaunchKernelGGL(permute_batch_k, dim3(blocks_for(n, 256)),
                           dim3(256), 0, s, b, perm, n);
    }
}

void permute(const void* src, void* dst, size_t esize, const uint32_t* perm,
             uint32_t n, hipStream_t s)
{
    const dim3 g(blocks_for(n, 256)), b(256);
    switch (esize) {
    case 1:
        hipLaunchKernelGGL(permute_k<uint8_t>, g, b, 0, s, (const uint8_t*)src,
                           (uint8_t*)dst, perm, n);
        break;
    case 2:
        hipLaunchKernelGGL(permute_k<uint16_t>, g, b, 0, s,
                           (const uint16_t*)src, (uint16_t*)dst, perm, n);
        break;
    case 4:
        hipLaunchKernelGGL(permute_k<uint32_t>, g, b, 0, s,
                           (const uint32_t*)src, (uint32_t*)dst, perm, n);
        break;
    case 8:
        hipLaunchKernelGGL(permute_k<uint2>, g, b, 0, s, (const uint2*)src,
                           (uint2*)dst, perm, n);
        break;
    case 16:
        hipLaunchKernelGGL(permute_k<uint4>, g, b, 0, s, (const uint4*)src,
                           (uint4*)dst, perm, n);
        break;
    case 32:
        hipLaunchKernelGGL(permute_k<Words8>, g, b, 0, s, (const Words8*)src,
                           (Words8*)dst, perm, n);
        break;
    default:
        throw Error(AQ_EINVAL, "aq_permute: element size must be 1,2,4,8,16");
    }
}

unsigned radix_key_bits(uint64_t max_val)
{
    uint64_t p = 1;
    while (p < max_val) p <<= 1;
    unsigned bits = 1;
    for (; (p & 1) == 0; p >>= 1, bits++)
        ;
    return (bits + 3u) / 4u * 4u;
}

}  // namespace aq

using namespace aq;

extern "C" {

AQ_API size_t aq_radix_sort_workspace(uint32_t n)
{
    return radix_sort_workspace(n);
}

AQ_API unsigned aq_radix_key_bits(uint64_t max_val)
{
    return radix_key_bits(max_val);
}

AQ_API int aq_radix_sort(uint32_t* keys, uint32_t* perm, uint32_t* inv_perm,
                         uint32_t n, unsigned key_bits, void* workspace,
                         void* stream)
{
    return guard([&] {
        AQ_REQUIRE(keys && perm && (workspace || n == 0),
                   "aq_radix_sort: null pointer");
        AQ_REQUIRE(key_bits <= 32, "aq_radix_sort: key_bits > 32");
        radix_sort(keys, perm, inv_perm, n, key_bits, workspace, S(stream));
    });
}

AQ_API int aq_permute(const void* src, void* dst, size_t esize,
                      const uint32_t* perm, uint32_t N, void* stream)
{
    return guard([&] {
        AQ_REQUIRE(src && dst && perm, "aq_permute: null pointer");
        AQ_REQUIRE(src != dst, "aq_permute: src and dst must differ");
        if (N) permute(src, dst, esize, perm, N, S(stream));
    });
}

}  // extern "C"
