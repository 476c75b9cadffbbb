// Hardware check of the division and square root sequences without range
// scaling (div_nr, sqrt_nr in aq_device.hpp) against the compiler's IEEE
// a / b and sqrtf: operands drawn over the whole window where the fused
// walk uses them (b in [2^-100, 2^100]; a == +-0 or a / b in
// [2^-120, 2^90] with |a| >= 2^-99; s == 0 or s in [2^-96, 2^128)), random
// signs and mantissas, counted where the bit patterns differ.
#include "aq_device.hpp"
#include "aq_internal.hpp"

namespace aq {

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256)
    nr_check_k(uint64_t n, uint64_t seed, unsigned long long* bad)
{
    unsigned long long nd = 0, ns = 0;
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t x = mix64(seed + 0x9E3779B97F4A7C15ull * (k + 1));
        const uint64_t y = mix64(x);
        // denominator: exponent in [-100, 100], positive
        const int eb = (int)(x % 201u) - 100;
        const float b = __uint_as_float(((uint32_t)(eb + 127) << 23) |
                                        (uint32_t)((x >> 8) & 0x7FFFFFu));
        // numerator: +-0 one time in 32, else exponent with
        // ea - eb in [-119, 89] and ea in [-99, 127]
        float a;
        if (((y >> 59) & 31u) == 0u) {
            a = (y & 1u) ? -0.f : 0.f;
        } else {
            const int lo = max(eb - 119, -99), hi = min(eb + 89, 127);
            const int ea = lo + (int)((y >> 32) % (uint64_t)(hi - lo + 1));
            a = __uint_as_float(((uint32_t)(y & 1u) << 31) |
                                ((uint32_t)(ea + 127) << 23) |
                                (uint32_t)((y >> 1) & 0x7FFFFFu));
        }
        if (__float_as_uint(div_nr(a, rcp_nr(b))) != __float_as_uint(a / b))
            nd++;
        // square root: 0 one time in 32, else exponent in [-96, 127]
        float s = 0.f;
        if (((x >> 59) & 31u) != 0u) {
            const int es = (int)((x >> 32) % 224u) - 96;
            s = __uint_as_float(((uint32_t)(es + 127) << 23) |
                                (uint32_t)(y >> 40 & 0x7FFFFFu));
        }
        if (__float_as_uint(sqrt_nr(s)) != __float_as_uint(sqrtf(s))) ns++;
    }
    if (nd) atomicAdd(bad, nd);
    if (ns) atomicAdd(bad + 1, ns);
}

__global__ void __launch_bounds__(256)
    clmath_k(int op, const float* __restrict__ x, float* __restrict__ y,
             uint32_t n)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    y[k] = op == 0 ? atan_cl(x[k]) : acospi_cl(x[k]);
}

}  // namespace

}  // namespace aq

extern "C" {

AQ_API int aq_check_nr_arith(uint64_t n, uint64_t seed, uint64_t* mismatches)
{
    using namespace aq;
    return guard([&] {
        AQ_REQUIRE(mismatches, "aq_check_nr_arith: null pointer");
        unsigned long long* d = nullptr;
        AQ_HIP(hipMalloc(&d, 2 * sizeof(unsigned long long)));
        unsigned long long h[2] = { 0, 0 };
        try {
            AQ_HIP(hipMemset(d, 0, 2 * sizeof(unsigned long long)));
            hipLaunchKernelGGL(nr_check_k, dim3(2048), dim3(256), 0, nullptr,
                               n, seed, d);
            AQ_HIP(hipGetLastError());
            AQ_HIP(hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost));
        } catch (...) {
            (void)hipFree(d);
            throw;
        }
        AQ_HIP(hipFree(d));
        mismatches[0] = h[0];
        mismatches[1] = h[1];
    });
}

AQ_API int aq_clmath_eval(int op, const float* x, float* y, uint32_t n,
                          void* stream)
{
    using namespace aq;
    return guard([&] {
        AQ_REQUIRE(op == 0 || op == 1, "aq_clmath_eval: op is 0 or 1");
        AQ_REQUIRE(n == 0 || (x && y), "aq_clmath_eval: null pointer");
        if (n == 0) return;
        hipLaunchKernelGGL(clmath_k, dim3((n + 255) / 256), dim3(256), 0,
                           (hipStream_t)stream, op, x, y, n);
        AQ_HIP(hipGetLastError());
    });
}

}  // extern "C"
