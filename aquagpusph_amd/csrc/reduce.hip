// Generic reductions (min / max / sum) over scalar or vec arrays.
//
// Replaces the reference Reduction tool (aquagpusph/CalcServer/Reduction.cpp
// :142-437, kernel Reduction.cl.in:35-66), which chains tree-reduction
// launches down to one element and then reads the result back to the host.
// Here a fixed number of workgroups each reduce a grid-strided range into
// one partial, a single workgroup reduces the partials and writes the result
// to device memory; nothing is read back, so the step stays on the device.
// min/max are exact in any order; FP32 sums differ from the reference's tree
// order within rounding.
#include "aq_device.hpp"
#include "aq_internal.hpp"

namespace aq {

constexpr int RD_BS = 256;
constexpr int RD_MAXB = 512;

template <int OP, typename T> __device__ __forceinline__ T rop(T a, T b)
{
    if constexpr (OP == AQ_RED_MIN) return (b < a) ? b : a;
    if constexpr (OP == AQ_RED_MAX) return (a < b) ? b : a;
    return a + b;
}

template <int OP, typename T, int NC>
__device__ void block_reduce_store(T (&v)[NC], T* out)
{
    __shared__ T sm[NC][RD_BS];
#pragma unroll
    for (int k = 0; k < NC; k++) sm[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int o = RD_BS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
#pragma unroll
            for (int k = 0; k < NC; k++)
                sm[k][threadIdx.x] =
                    rop<OP, T>(sm[k][threadIdx.x], sm[k][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NC; k++) out[k] = sm[k][0];
    }
}

// TI: element type read; T: type the reduction runs and writes in (a float
// array summed in double when T = double)
template <int OP, typename TI, typename T, int NC>
__global__ void __launch_bounds__(RD_BS)
    reduce_partial(const TI* __restrict__ in, uint32_t n, T* __restrict__ part,
                   const T* __restrict__ ident)
{
    T v[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) v[k] = ident[k];
    for (uint32_t i = blockIdx.x * RD_BS + threadIdx.x; i < n;
         i += gridDim.x * RD_BS) {
#pragma unroll
        for (int k = 0; k < NC; k++)
            v[k] = rop<OP, T>(v[k], (T)in[(size_t)i * NC + k]);
    }
    block_reduce_store<OP, T, NC>(v, part + (size_t)blockIdx.x * NC);
}

template <int OP, typename T, int NC>
__global__ void __launch_bounds__(RD_BS)
    reduce_final(const T* __restrict__ part, uint32_t nb, T* __restrict__ out,
                 const T* __restrict__ ident)
{
    T v[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) v[k] = ident[k];
    for (uint32_t b = threadIdx.x; b < nb; b += RD_BS) {
#pragma unroll
        for (int k = 0; k < NC; k++)
            v[k] = rop<OP, T>(v[k], part[(size_t)b * NC + k]);
    }
    block_reduce_store<OP, T, NC>(v, out);
}

size_t reduce_workspace(uint32_t n, int ncomp)
{
    (void)n;
    // partials + identity
    return (size_t)(RD_MAXB + 1) * ncomp * 8 + 256;
}

template <int OP, typename TI, typename T, int NC>
static void run_reduce(const void* in, uint32_t n, void* out,
                       const void* ident_dev, void* ws, hipStream_t s)
{
    const unsigned nb = std::max(
        1u, std::min<unsigned>(RD_MAXB, blocks_for(n, RD_BS * 8)));
    T* part = (T*)ws;
    hipLaunchKernelGGL((reduce_partial<OP, TI, T, NC>), dim3(nb), dim3(RD_BS),
                       0, s, (const TI*)in, n, part, (const T*)ident_dev);
    hipLaunchKernelGGL((reduce_final<OP, T, NC>), dim3(1), dim3(RD_BS), 0, s,
                       part, nb, (T*)out, (const T*)ident_dev);
}

template <int OP, typename TI, typename T = TI>
static void run_nc(int nc, const void* in, uint32_t n, void* out,
                   const void* ident_dev, void* ws, hipStream_t s)
{
    switch (nc) {
    case 1:
        run_reduce<OP, TI, T, 1>(in, n, out, ident_dev, ws, s);
        break;
    case 2:
        run_reduce<OP, TI, T, 2>(in, n, out, ident_dev, ws, s);
        break;
    case 4:
        run_reduce<OP, TI, T, 4>(in, n, out, ident_dev, ws, s);
        break;
    default:
        throw Error(AQ_EINVAL, "reduction: ncomp must be 1, 2 or 4");
    }
}

template <int OP>
static void run_type(int type, int nc, const void* in, uint32_t n, void* out,
                     const void* ident_dev, void* ws, hipStream_t s)
{
    switch (type) {
    case AQ_T_FLOAT:
        run_nc<OP, float>(nc, in, n, out, ident_dev, ws, s);
        break;
    case AQ_T_INT:
        run_nc<OP, int32_t>(nc, in, n, out, ident_dev, ws, s);
        break;
    case AQ_T_UINT:
        run_nc<OP, uint32_t>(nc, in, n, out, ident_dev, ws, s);
        break;
    case AQ_T_DOUBLE:
        run_nc<OP, double>(nc, in, n, out, ident_dev, ws, s);
        break;
    case AQ_T_FLOAT_F64:
        run_nc<OP, float, double>(nc, in, n, out, ident_dev, ws, s);
        break;
    default:
        throw Error(AQ_EINVAL, "reduction: unknown element type");
    }
}

// ident_dev: device copy of the identity (ncomp elements)
void reduce(int op, int type, int nc, const void* in, uint32_t n, void* out,
            const void* ident_dev, void* ws, hipStream_t s)
{
    switch (op) {
    case AQ_RED_MIN:
        run_type<AQ_RED_MIN>(type, nc, in, n, out, ident_dev, ws, s);
        break;
    case AQ_RED_MAX:
        run_type<AQ_RED_MAX>(type, nc, in, n, out, ident_dev, ws, s);
        break;
    case AQ_RED_SUM:
        run_type<AQ_RED_SUM>(type, nc, in, n, out, ident_dev, ws, s);
        break;
    default:
        throw Error(AQ_EINVAL, "reduction: unknown operation");
    }
}

}  // namespace aq

using namespace aq;

extern "C" {

AQ_API size_t aq_reduce_workspace(uint32_t n, int ncomp)
{
    return reduce_workspace(n, ncomp);
}

AQ_API int aq_reduce(int op, int type, int ncomp, const void* in, uint32_t n,
                     void* out, const void* identity, void* workspace,
                     void* stream)
{
    return guard([&] {
        AQ_REQUIRE(in && out && identity && workspace,
                   "aq_reduce: null pointer");
        AQ_REQUIRE(ncomp == 1 || ncomp == 2 || ncomp == 4,
                   "aq_reduce: ncomp must be 1, 2 or 4");
        // identity goes at the end of the workspace, after the partials
        char* ident_dev = (char*)workspace + (size_t)RD_MAXB * ncomp * 8;
        const size_t ib =
            (type == AQ_T_DOUBLE || type == AQ_T_FLOAT_F64) ? 8 : 4;
        AQ_HIP(hipMemcpyAsync(ident_dev, identity, (size_t)ncomp * ib,
                              hipMemcpyHostToDevice, S(stream)));
        reduce(op, type, ncomp, in, n, out, ident_dev, workspace, S(stream));
    });
}

}  // extern "C"
