// Small device helpers of the runtime: array fill (Set tool) and the
// failure stamp of a step.
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
