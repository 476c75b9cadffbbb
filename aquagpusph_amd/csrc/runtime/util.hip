// Small device helpers of the runtime: array fill (Set tool), the batched
// fills and copies of consecutive set / copy tools, and the failure stamp
// of a step.
#include <algorithm>

#include "sim.hpp"

namespace aq {
namespace rt {

// dst[i*stored + k] = pattern[k] for i < n (Set.cl.in:29-47)
__global__ void fill_words(uint32_t* __restrict__ dst,
                           const uint32_t* __restrict__ pattern, int stored,
                           size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < stored; k++) dst[i * stored + k] = pattern[k];
}

void fill_launch(void* dst, const uint32_t* pattern_dev, int stored, size_t n,
                 hipStream_t s)
{
    if (!n) return;
    hipLaunchKernelGGL(fill_words, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                       (uint32_t*)dst, pattern_dev, stored, n);
}

// Several fills in one launch: blockIdx.y picks the array.  A pattern of 4
// words goes out as one 16-byte store per element when the array is
// 16-byte aligned (every array the runtime allocates is).
__global__ void fill_batch_k(FillBatch b)
{
    const int f = blockIdx.y;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.n[f]) return;
    const int st = b.stored[f];
    const uint32_t* pat = b.pat[f];
    uint32_t* dst = (uint32_t*)b.dst[f];
    if (st == 4 && (((uintptr_t)dst) & 15u) == 0) {
        ((uint4*)dst)[i] = make_uint4(pat[0], pat[1], pat[2], pat[3]);
        return;
    }
    for (int k = 0; k < st; k++) dst[i * st + k] = pat[k];
}

void fill_batch(const FillBatch& b, hipStream_t s)
{
    size_t nmax = 0;
    for (int f = 0; f < b.cnt; f++) nmax = std::max(nmax, b.n[f]);
    if (!nmax) return;
    hipLaunchKernelGGL(fill_batch_k, dim3(blocks_for(nmax, 256), b.cnt),
                       dim3(256), 0, s, b);
}

// Several copies of whole words in one launch: blockIdx.y picks the pair;
// 16 bytes per thread where both ends are 16-byte aligned.
__global__ void copy_batch_k(CopyBatch b)
{
    const int f = blockIdx.y;
    const size_t w = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    const size_t nw = b.words[f];
    if (w >= nw) return;
    const uint32_t* src = (const uint32_t*)b.src[f];
    uint32_t* dst = (uint32_t*)b.dst[f];
    if (w + 4 <= nw && ((((uintptr_t)src) | ((uintptr_t)dst)) & 15u) == 0) {
        ((uint4*)dst)[w / 4] = ((const uint4*)src)[w / 4];
        return;
    }
    for (size_t k = w; k < nw && k < w + 4; k++) dst[k] = src[k];
}

void copy_batch(const CopyBatch& b, hipStream_t s)
{
    size_t nmax = 0;
    for (int f = 0; f < b.cnt; f++) nmax = std::max(nmax, b.words[f]);
    if (!nmax) return;
    hipLaunchKernelGGL(copy_batch_k,
                       dim3(blocks_for((nmax + 3) / 4, 256), b.cnt),
                       dim3(256), 0, s, b);
}

// The first step that ends with a failure flagged in status[0] records the
// iteration counter in status[1] (0xFFFFFFFF until then).
__global__ void status_stamp_k(uint32_t* status, const uint32_t* iter)
{
    if (status[0] != 0u && status[1] == 0xFFFFFFFFu)
        status[1] = iter ? *iter : 0u;
}

void status_stamp(uint32_t* status, const uint32_t* iter, hipStream_t s)
{
    hipLaunchKernelGGL(status_stamp_k, dim3(1), dim3(1), 0, s, status, iter);
}

}  // namespace rt
}  // namespace aq
