// UnSort (aquagpusph/CalcServer/UnSort.cl.in:31-42): scatter an array in
// sorted particle order back to the original order, out[id[i]] = in[i],
// for elements of any whole number of 32-bit words.
#include "aq_device.hpp"
#include "aq_internal.hpp"

namespace aq {

template <int W>
__global__ void unsort_k(const uint32_t* __restrict__ id,
                         const uint32_t* __restrict__ in,
                         uint32_t* __restrict__ out, uint32_t N)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const size_t d = (size_t)id[i] * W, s = (size_t)i * W;
#pragma unroll
    for (int w = 0; w < W; w++) out[d + w] = in[s + w];
}

void launch_unsort(const uint32_t* id, const void* in, void* out, uint32_t N,
                   uint32_t words, hipStream_t s)
{
    const dim3 g((N + 255) / 256), b(256);
    const uint32_t* a = (const uint32_t*)in;
    uint32_t* o = (uint32_t*)out;
    switch (words) {
    case 1: hipLaunchKernelGGL(unsort_k<1>, g, b, 0, s, id, a, o, N); break;
    case 2: hipLaunchKernelGGL(unsort_k<2>, g, b, 0, s, id, a, o, N); break;
    case 4: hipLaunchKernelGGL(unsort_k<4>, g, b, 0, s, id, a, o, N); break;
    case 16: hipLaunchKernelGGL(unsort_k<16>, g, b, 0, s, id, a, o, N); break;
    default:
        throw Error(AQ_EUNSUP, "unsort: elements of " + std::to_string(words) +
                                   " words are not supported");
    }
}

}  // namespace aq

using namespace aq;

AQ_API int aq_unsort(const uint32_t* id, const void* in, void* out,
                     uint32_t N, uint32_t elem_bytes, void* stream)
{
    return guard([&] {
        if (!N) return;
        AQ_REQUIRE(id && in && out, "aq_unsort: null pointer");
        AQ_REQUIRE(elem_bytes % 4 == 0, "aq_unsort: elem_bytes must be a "
                                        "multiple of 4");
        launch_unsort(id, in, out, N, elem_bytes / 4, (hipStream_t)stream);
    });
}
