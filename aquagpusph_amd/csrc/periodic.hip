// Periodic boxes (no counterpart in the reference, which makes a domain
// periodic with plane portals, cfd/portal.xml, mirroring the particles near
// each face every step): the axes whose length periodic_L[k] is positive
// wrap.  Positions are brought back into [periodic_min, periodic_min + L)
// before the link-list, whose grid then tiles the box exactly, and the
// neighbour walks wrap their cell offsets and shift the wrapped particles
// by the box length (per_neighbour in aq_device.hpp).
#include "aq_device.hpp"
#include "aq_internal.hpp"

namespace aq {

namespace {

constexpr int BS = 256;

// One wrap per axis: a particle moves less than L in a time step.
template <int D>
__global__ void __launch_bounds__(BS)
    periodic_wrap_k(float* __restrict__ r, uint32_t N,
                    const float* __restrict__ per_min,
                    const float* __restrict__ per_L)
{
    constexpr int VS = (D == 2) ? 2 : 4;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
#pragma unroll
    for (int k = 0; k < D; k++) {
        const float L = per_L[k];
        if (!(L > 0.f)) continue;
        const float lo = per_min[k];
        float x = r[(size_t)i * VS + k];
        if (x < lo) x = x + L;
        else if (x >= lo + L) x = x - L;
        r[(size_t)i * VS + k] = x;
    }
}

}  // namespace

void launch_periodic_wrap(int dims, float* r, uint32_t N,
                          const float* per_min, const float* per_L,
                          hipStream_t s)
{
    if (!N) return;
    const dim3 g(blocks_for(N, BS)), b(BS);
    if (dims == 2)
        hipLaunchKernelGGL(periodic_wrap_k<2>, g, b, 0, s, r, N, per_min,
                           per_L);
    else
        hipLaunchKernelGGL(periodic_wrap_k<3>, g, b, 0, s, r, N, per_min,
                           per_L);
}

}  // namespace aq
