// Small device helpers of the runtime: array fill (Set tool).
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

}  // namespace rt
}  // namespace aq
