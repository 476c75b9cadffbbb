// Device-side building blocks: vec storage, the Wendland C2 kernel and the
// link-list neighbour walk.  All FP32 expressions keep the association of the
// reference OpenCL sources (the library is compiled with -ffp-contract=off).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace aq {

// ---- vec storage: float2 in 2D, float4 in 3D (types/2D.h:23, 3D.h:23) ----
template <int D> struct VecT;
template <> struct VecT<2> {
    typedef float2 T;
};
template <> struct VecT<3> {
    typedef float4 T;
};

struct f3 {
    float x, y, z;
};

// float -> uint with saturation (NaN and negatives -> 0, >= 2^32 ->
// UINT_MAX); the OpenCL (uint) cast is implementation-defined outside
// [0, 2^32), which the cell hash of a point below r_min reaches
__device__ __forceinline__ uint32_t cvt_u32(float x)
{
    if (!(x > 0.f)) return 0u;
    if (x >= 4294967296.f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

template <int D> __device__ __forceinline__ f3 ldv(const float* a, uint32_t i)
{
    f3 v;
    if constexpr (D == 2) {
        const float2 t = reinterpret_cast<const float2*>(a)[i];
        v.x = t.x;
        v.y = t.y;
        v.z = 0.f;
    } else {
        const float4 t = reinterpret_cast<const float4*>(a)[i];
        v.x = t.x;
        v.y = t.y;
        v.z = t.z;
    }
    return v;
}

// Loads at a 32-bit byte offset from a wave-uniform base pointer: the
// address stays "SGPR base + VGPR offset" (no 64-bit address arithmetic per
// load).  Callers guarantee that the offsets fit in 32 bits.
__device__ __forceinline__ float ldf_o(const float* a, uint32_t j)
{
    return *reinterpret_cast<const float*>(
        reinterpret_cast<const char*>(a) + (j << 2));
}
__device__ __forceinline__ int ldi_o(const int* a, uint32_t j)
{
    return *reinterpret_cast<const int*>(
        reinterpret_cast<const char*>(a) + (j << 2));
}
template <int D>
__device__ __forceinline__ f3 ldv_o(const float* a, uint32_t j)
{
    f3 v;
    const char* b = reinterpret_cast<const char*>(a);
    if constexpr (D == 2) {
        const float2 t = *reinterpret_cast<const float2*>(b + (j << 3));
        v.x = t.x;
        v.y = t.y;
        v.z = 0.f;
    } else {
        const float4 t = *reinterpret_cast<const float4*>(b + (j << 4));
        v.x = t.x;
        v.y = t.y;
        v.z = t.z;
    }
    return v;
}

// store the XYZ part, leaving w untouched (the reference writes .XYZ)
template <int D>
__device__ __forceinline__ void stxyz(float* a, uint32_t i, const f3& v)
{
    if constexpr (D == 2) {
        reinterpret_cast<float2*>(a)[i] = make_float2(v.x, v.y);
    } else {
        float* p = a + 4 * (size_t)i;
        p[0] = v.x;
        p[1] = v.y;
        p[2] = v.z;
    }
}

template <int D> __device__ __forceinline__ f3 sub(const f3& a, const f3& b)
{
    f3 v;
    v.x = a.x - b.x;
    v.y = a.y - b.y;
    v.z = (D == 3) ? a.z - b.z : 0.f;
    return v;
}

// dot over the first D components, left to right
// s * v and a + b per component (the z component only in 3D)
template <int D> __device__ __forceinline__ f3 scl(float s, const f3& v)
{
    return f3{ s * v.x, s * v.y, (D == 3) ? s * v.z : 0.f };
}
template <int D> __device__ __forceinline__ f3 add3(const f3& a, const f3& b)
{
    return f3{ a.x + b.x, a.y + b.y, (D == 3) ? a.z + b.z : a.z };
}

// Contracted forms (one rounding per multiply-add), for the kernels built
// with the OpenCL C default floating-point rules.
template <int D> __device__ __forceinline__ float fdot(const f3& a, const f3& b)
{
    float s = a.x * b.x;
    s = fmaf(a.y, b.y, s);
    if constexpr (D == 3) s = fmaf(a.z, b.z, s);
    return s;
}
template <int D>
__device__ __forceinline__ void faxpy(f3& acc, float s, const f3& v)
{
    acc.x = fmaf(s, v.x, acc.x);
    acc.y = fmaf(s, v.y, acc.y);
    if constexpr (D == 3) acc.z = fmaf(s, v.z, acc.z);
}

// ---- correctly rounded division and square root without range scaling ----
// hipcc expands an IEEE float a / b (b > 0) to
//   y  = fma(fma(-b, rcp(b), 1), rcp(b), rcp(b))       reciprocal, one Newton step
//   q0 = a y;  q1 = fma(a - b q0, y, q0);  q2 = fma(a - b q1, y, q1)
// wrapped in v_div_scale (operands rescaled by 2^+-64 near the exponent
// limits), v_div_fmas (the last fma, rescaled) and v_div_fixup (zeros,
// infinities, NaN, overflow, and the quotient's sign).  For b in
// [2^-100, 2^100] and a == 0 or a / b in [2^-120, 2^90] with
// |a| >= 2^-100, v_div_scale leaves both operands as they are, v_div_fmas
// is a plain fma and v_div_fixup returns the quotient unchanged, so div_nr
// below, the same instructions without the three range steps, is a / b bit
// for bit.  The remainders are negated (b q - a, exact under the fma like
// a - b q), which gives a zero numerator's sign to the quotient as
// v_div_fixup does.  Callers check the operand ranges before taking this
// path (the fused fluid walk bounds them once per launch from the fields'
// ranges, plus two per-pair tests).
struct Rcp {
    float b, y;
};
__device__ __forceinline__ Rcp rcp_nr(float b)
{
    const float y0 = __builtin_amdgcn_rcpf(b);
    return { b, fmaf(fmaf(-b, y0, 1.f), y0, y0) };
}
__device__ __forceinline__ float div_nr(float a, const Rcp& d)
{
    const float q0 = a * d.y;
    const float q1 = fmaf(-fmaf(d.b, q0, -a), d.y, q0);
    return fmaf(-fmaf(d.b, q1, -a), d.y, q1);
}
// sqrtf(s) for s == 0 or s >= 2^-96: hipcc's sequence (hardware estimate,
// then one ulp down or up by the signs of the fma residuals) without the
// 2^32 prescaling of s < 2^-96 and the zero / infinity pass-through; zero
// goes through the residual tests unchanged.
__device__ __forceinline__ float sqrt_nr(float s)
{
    const float y = __builtin_amdgcn_sqrtf(s);
    const float yd = __uint_as_float(__float_as_uint(y) - 1u);
    const float yu = __uint_as_float(__float_as_uint(y) + 1u);
    const float rd = fmaf(-yd, y, s), ru = fmaf(-yu, y, s);
    const float z = (rd <= 0.f) ? yd : y;
    return (ru > 0.f) ? yu : z;
}

template <int D> __device__ __forceinline__ float dot(const f3& a, const f3& b)
{
    float s = a.x * b.x;
    s = s + a.y * b.y;
    if constexpr (D == 3) s = s + a.z * b.z;
    return s;
}

template <int D>
__device__ __forceinline__ void axpy(f3& acc, float s, const f3& v)
{
    acc.x += s * v.x;
    acc.y += s * v.y;
    if constexpr (D == 3) acc.z += s * v.z;
}

// ---- Wendland C2 (KernelFunctions/Wendland{2D,3D}.hcl) --------------------
#define AQ_IPI 0.318309886f

template <int D> __device__ __forceinline__ float kernelW(float q)
{
    const float wcon = (D == 2) ? 0.109375f * AQ_IPI : 0.08203125f * AQ_IPI;
    const float a = 2.f - q;
    return wcon * (1.f + 2.f * q) * a * a * a * a;
}

template <int D> __device__ __forceinline__ float kernelF(float q)
{
    const float wcon = (D == 2) ? 1.09375f * AQ_IPI : 0.8203125f * AQ_IPI;
    const float a = 2.f - q;
    return wcon * a * a * a;
}

template <int D> __device__ __forceinline__ float kernelS_P(float q)
{
    const float q2 = q * q;
    const float q3 = q2 * q;
    if constexpr (D == 2) {
        const float wcon = 0.109375f * AQ_IPI;
        return wcon * (0.285714f * q3 * q2 - 2.5f * q2 * q2 + 8.f * q3 -
                       10.f * q2 + 8.f);
    } else {
        const float wcon = 0.08203125f * AQ_IPI;
        return wcon * (0.25f * q3 * q2 - 2.142857f * q2 * q2 +
                       6.666667f * q3 - 8.f * q2 + 5.333333f);
    }
}

__device__ __forceinline__ float cl_sign(float x)
{
    if (x > 0.f) return 1.f;
    if (x < 0.f) return -1.f;
    if (x != x) return 0.f;
    return x;
}
__device__ __forceinline__ float cl_min(float x, float y)
{
    return (y < x) ? y : x;
}
__device__ __forceinline__ float cl_max(float x, float y)
{
    return (x < y) ? y : x;
}

// ---- atan, acospi and pow of the boundary-integral kernel terms ----------
// The reference evaluates these with its OpenCL runtime's atan / acospi /
// pow (Wendland2D.hcl / Wendland3D.hcl:137-185), whose only contract is the
// OpenCL C ULP bounds (atan 5, acospi 5, pow 16).  Here they are one fixed
// float sequence, the Cephes method (S. L. Moshier, Cephes Mathematical
// Library: atanf.c, asinf.c, acosf.c): range reduction by tan(pi/8) and
// tan(3pi/8) for atan, by 1/2 for asin, and the Cephes minimax polynomials
// evaluated with fused multiply-adds.  The CPU oracle runs the same
// sequence, so the two agree bit for bit; both are checked against the
// bounds over the float range.
__device__ __forceinline__ float atan_cl(float x)
{
    const float ax = fabsf(x);
    float a = ax, base = 0.f;
    if (ax > 2.414213562373095f) {  // tan(3 pi / 8)
        base = 1.5707963267948966f;
        a = -1.f / ax;
    } else if (ax > 0.4142135623730950f) {  // tan(pi / 8)
        base = 0.7853981633974483f;
        a = (ax - 1.f) / (ax + 1.f);
    }
    const float z = a * a;
    float p = __builtin_fmaf(8.05374449538e-2f, z, -1.38776856032e-1f);
    p = __builtin_fmaf(p, z, 1.99777106478e-1f);
    p = __builtin_fmaf(p, z, -3.33329491539e-1f);
    const float r = base + __builtin_fmaf(p * z, a, a);
    return copysignf(r, x);
}

// asin on [0, 1/2] (Cephes asinf polynomial)
__device__ __forceinline__ float asin_half_cl(float a)
{
    const float z = a * a;
    float p = __builtin_fmaf(4.2163199048e-2f, z, 2.4181311049e-2f);
    p = __builtin_fmaf(p, z, 4.5470025998e-2f);
    p = __builtin_fmaf(p, z, 7.4953002686e-2f);
    p = __builtin_fmaf(p, z, 1.6666752422e-1f);
    return __builtin_fmaf(p * z, a, a);
}

// acos(x)/pi for x in [0, 1] (the only range the kernel terms reach):
// acos x = 2 asin(sqrt((1 - x)/2)) above 1/2, pi/2 - asin x below
__device__ __forceinline__ float acospi_cl(float x)
{
    float ac;
    if (x > 0.5f) {
        const float s = sqrtf(0.5f * (1.f - x));
        ac = 2.f * asin_half_cl(s);
    } else {
        ac = 1.5707963267948966f - asin_half_cl(x);
    }
    return ac * 0.318309886183791f;
}

// pow(area, 1 / (DIMS - 1)): the exponent is 1 in 2D and 1/2 in 3D, where
// the correctly rounded identity and square root are the exact results
__device__ __forceinline__ float pow_cl(float x, float y)
{
    if (y == 1.f) return x;
    if (y == 0.5f) return sqrtf(x);
    return powf(x, y);
}

__device__ __forceinline__ float omega3(float a, float b)
{
    const float a2 = a * a;
    const float b2 = b * b;
    return acospi_cl(cl_min(
        sqrtf((1.f + a2 + b2) / ((1.f + a2) * (1.f + b2))), 1.f));
}

template <int D>
__device__ __forceinline__ float kernelS_D(float d, float t, float b, float s)
{
    if constexpr (D == 2) {
        const float wcon = 0.5f * AQ_IPI;
        const float dr = 0.5f * s;
        return -wcon * (atan_cl((t + dr) / d) - atan_cl((t - dr) / d));
    } else {
        const float wcon = 0.25f;
        const float dr = 0.5f * sqrtf(s);
        const float t1 = (t - dr) / d, t2 = (t + dr) / d;
        const float b1 = (b - dr) / d, b2 = (b + dr) / d;
        const float st = cl_sign(t1), sb = cl_sign(b1);
        return -wcon * (omega3(t2, b2) - st * omega3(t1, b2) -
                        sb * omega3(t2, b1) + st * sb * omega3(t1, b1));
    }
}

// ---- neighbour walk (types/3D.h:197-219, types/2D.h:174-193) --------------
// Calls body(j) for every particle j of the 3^D cells around c_i, in the
// reference order: x offset outermost, then y, then z; in each cell the
// sorted indices ascending while icell[j] == cell.
template <int D, class Body>
__device__ __forceinline__ void walk(uint32_t c_i, uint32_t N, uint32_t nx,
                                     uint32_t ny, const uint32_t* __restrict__ icell,
                                     const uint32_t* __restrict__ ihoc, Body body)
{
#pragma unroll 1
    for (int ci = -1; ci <= 1; ci++) {
#pragma unroll 1
        for (int cj = -1; cj <= 1; cj++) {
            if constexpr (D == 2) {
                const uint32_t c_j = c_i + (uint32_t)ci + (uint32_t)cj * nx;
                uint32_t j = ihoc[c_j];
                while ((j < N) && (icell[j] == c_j)) {
                    body(j);
                    j++;
                }
            } else {
#pragma unroll 1
                for (int ck = -1; ck <= 1; ck++) {
                    const uint32_t c_j = c_i + (uint32_t)ci +
                                         (uint32_t)cj * nx +
                                         (uint32_t)ck * nx * ny;
                    uint32_t j = ihoc[c_j];
                    while ((j < N) && (icell[j] == c_j)) {
                        body(j);
                        j++;
                    }
                }
            }
        }
    }
}

// ---- periodic boxes ----------------------------------------------------------
// A box periodic along the axes whose length per[k] is positive (the vec
// variable periodic_L; per == nullptr: no periodic axis).  Along such an
// axis the link-list grid has n_cells[k] cells tiling [periodic_min,
// periodic_min + L) exactly (no margin cells), and the neighbour cell
// c + o wraps modulo n_cells[k]; the particles of a wrapped cell are seen
// at r + shift, shift = -L (wrapped below 0) or +L (wrapped above n - 1),
// so every pair within the support is found once, at its nearest image
// (each periodic axis has at least 3 cells of at least the support).
// Neighbour cell (offset o) of cell c and its shift.  The walk order is the
// reference's (x offset outermost, then y, then z), over the wrapped cells.
template <int D>
__device__ __forceinline__ uint32_t per_neighbour(uint32_t c, int oi, int oj,
                                                  int ok,
                                                  const uint32_t* __restrict__ nc,
                                                  const float* __restrict__ per,
                                                  f3& sh)
{
    const uint32_t nx = nc[0], ny = nc[1];
    int q[3] = { (int)(c % nx), (int)((c / nx) % ny),
                 D == 3 ? (int)(c / (nx * ny)) : 0 };
    const int o[3] = { oi, oj, ok };
    float s[3] = { 0.f, 0.f, 0.f };
#pragma unroll
    for (int k = 0; k < D; k++) {
        const int n = (int)nc[k];
        int v = q[k] + o[k];
        const float L = per[k];
        if (L > 0.f) {
            if (v < 0) {
                v += n;
                s[k] = -L;
            } else if (v >= n) {
                v -= n;
                s[k] = L;
            }
        }
        q[k] = v;
    }
    sh = f3{ s[0], s[1], s[2] };
    return (uint32_t)q[0] + (uint32_t)q[1] * nx +
           (D == 3 ? (uint32_t)q[2] * nx * ny : 0u);
}

// r + shift, componentwise, only along the periodic axes (exact: the
// other components are left as they are, not added a zero)
template <int D>
__device__ __forceinline__ f3 per_shift(const f3& r, const f3& sh,
                                        const float* __restrict__ per)
{
    f3 o = r;
    if (per[0] > 0.f) o.x = r.x + sh.x;
    if (per[1] > 0.f) o.y = r.y + sh.y;
    if (D == 3 && per[2] > 0.f) o.z = r.z + sh.z;
    return o;
}

// walk<D> over a periodic grid: body(j, shift) for every particle j of the
// 3^D wrapped cells around c_i, in the reference order.
template <int D, class Body>
__device__ __forceinline__ void walk_per(uint32_t c_i, uint32_t N,
                                         const uint32_t* __restrict__ nc,
                                         const float* __restrict__ per,
                                         const uint32_t* __restrict__ icell,
                                         const uint32_t* __restrict__ ihoc,
                                         Body body)
{
#pragma unroll 1
    for (int ci = -1; ci <= 1; ci++)
#pragma unroll 1
        for (int cj = -1; cj <= 1; cj++)
#pragma unroll 1
            for (int ck = (D == 3 ? -1 : 0); ck <= (D == 3 ? 1 : 0); ck++) {
                f3 sh;
                const uint32_t c_j = per_neighbour<D>(c_i, ci, cj, ck, nc,
                                                      per, sh);
                uint32_t j = ihoc[c_j];
                while ((j < N) && (icell[j] == c_j)) {
                    body(j, sh);
                    j++;
                }
            }
}

// The walk around a cell given explicitly (a moved point, e.g. a mirrored
// particle): neighbour cells outside the grid (c_j >= n_cells.w, a point
// beyond the particle bounding box) are skipped, so every ihoc read stays
// in range whatever the point.
template <int D, class Body>
__device__ __forceinline__ void walk_in(uint32_t c_i, uint32_t N,
                                        const uint32_t* __restrict__ nc,
                                        const uint32_t* __restrict__ icell,
                                        const uint32_t* __restrict__ ihoc,
                                        Body body)
{
    const uint32_t nx = nc[0], ny = nc[1], nw = nc[3];
#pragma unroll 1
    for (int ci = -1; ci <= 1; ci++)
#pragma unroll 1
        for (int cj = -1; cj <= 1; cj++)
#pragma unroll 1
            for (int ck = (D == 3 ? -1 : 0); ck <= (D == 3 ? 1 : 0); ck++) {
                const uint32_t c_j = c_i + (uint32_t)ci + (uint32_t)cj * nx +
                                     (uint32_t)ck * nx * ny;
                if (c_j >= nw) continue;
                uint32_t j = ihoc[c_j];
                while ((j < N) && (icell[j] == c_j)) {
                    body(j);
                    j++;
                }
            }
}

}  // namespace aq
