// CPU oracle for the AQUAgpusph per-timestep SPH pipeline.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, kernel by kernel, the
// OpenCL scripts and native tools of the reference (AQUAgpusph v5.0.4) so
// that the HIP implementation can be checked against it.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
// product library never does.
//
// Every function is templated on the spatial dimension D (2 or 3) and on the
// real type R (float for parity runs, double for the "FP64 oracle" used to
// state the error of every FP32 result).  Arithmetic is written in the same
// association order as the reference kernels and is compiled without FP
// contraction, so that the FP32 instantiation reproduces the reference
// expression trees operation by operation.
//
// Layout follows the reference variables: one array per field; `vec` fields
// hold VS(D) reals per particle (2 in 2D, 4 in 3D with w carried along);
// indices are 32-bit (`usize` = uint, the reference default addr_bits=32).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

namespace orc {

typedef uint32_t usize;

template <int D> struct Dim {
    static constexpr int VS = (D == 2) ? 2 : 4;  // stored components per vec
};

// ---------------------------------------------------------------------------
// Small vector helpers over the first D ("XYZ") components.
// dot() is evaluated left to right: ((x*x' + y*y') + z*z').
// ---------------------------------------------------------------------------
template <int D, typename R> struct V {
    R c[3];
};

template <int D, typename R>
inline V<D, R> ld(const R* a, usize i)
{
    V<D, R> v;
    for (int k = 0; k < D; k++) v.c[k] = a[(size_t)i * Dim<D>::VS + k];
    if (D == 2) v.c[2] = R(0);
    return v;
}

template <int D, typename R>
inline V<D, R> sub(const V<D, R>& a, const V<D, R>& b)
{
    V<D, R> v;
    for (int k = 0; k < D; k++) v.c[k] = a.c[k] - b.c[k];
    v.c[2] = (D == 2) ? R(0) : v.c[2];
    return v;
}

template <int D, typename R>
inline R dot(const V<D, R>& a, const V<D, R>& b)
{
    R s = a.c[0] * b.c[0];
    s = s + a.c[1] * b.c[1];
    if (D == 3) s = s + a.c[2] * b.c[2];
    return s;
}

template <int D, typename R> inline R length(const V<D, R>& a)
{
    return std::sqrt(dot<D, R>(a, a));
}

// acc[i].XYZ += s * v, component by component
template <int D, typename R>
inline void addsv(R* acc, usize i, R s, const V<D, R>& v)
{
    for (int k = 0; k < D; k++) acc[(size_t)i * Dim<D>::VS + k] += s * v.c[k];
}

// ---------------------------------------------------------------------------
// Wendland C2 kernel (reference: resources/Scripts/KernelFunctions/
// Wendland2D.hcl:44-66 and Wendland3D.hcl:44-66, boundary-integral parts at
// Wendland2D.hcl:100-154 and Wendland3D.hcl:100-187).  Constants are the
// single-precision literals of the reference; products are kept in their
// left-to-right order.
// ---------------------------------------------------------------------------
static const float kIPI = 0.318309886f;  // Wendland*.hcl iM_PI

template <int D, typename R> inline R kernelW(R q)
{
    const R wcon = (D == 2) ? R(0.109375f * kIPI) : R(0.08203125f * kIPI);
    const R a = R(2) - q;
    return wcon * (R(1) + R(2) * q) * a * a * a * a;
}

template <int D, typename R> inline R kernelF(R q)
{
    const R wcon = (D == 2) ? R(1.09375f * kIPI) : R(0.8203125f * kIPI);
    const R a = R(2) - q;
    return wcon * a * a * a;
}

template <int D, typename R> inline R kernelS_P(R q)
{
    const R q2 = q * q;
    const R q3 = q2 * q;
    if (D == 2) {
        const R wcon = R(0.109375f * kIPI);
        return wcon * (R(0.285714f) * q3 * q2 - R(2.5f) * q2 * q2 +
                       R(8.f) * q3 - R(10.f) * q2 + R(8.f));
    }
    const R wcon = R(0.08203125f * kIPI);
    return wcon * (R(0.25f) * q3 * q2 - R(2.142857f) * q2 * q2 +
                   R(6.666667f) * q3 - R(8.f) * q2 + R(5.333333f));
}

// OpenCL sign(): 1, -1, or the (signed) zero / 0 for NaN
template <typename R> inline R cl_sign(R x)
{
    if (x > R(0)) return R(1);
    if (x < R(0)) return R(-1);
    if (x != x) return R(0);
    return x;
}

template <typename R> inline R cl_min(R x, R y) { return (y < x) ? y : x; }
template <typename R> inline R cl_max(R x, R y) { return (x < y) ? y : x; }

// atan, acospi and pow of the boundary-integral terms
// (KernelFunctions/Wendland2D.hcl, Wendland3D.hcl:137-185).  The reference
// takes them from its OpenCL runtime, bound only by the OpenCL C ULP limits
// (atan 5, acospi 5, pow 16).  The float instantiation runs one fixed
// sequence, the Cephes method (S. L. Moshier, Cephes Mathematical Library:
// atanf.c, asinf.c, acosf.c): range reduction by tan(pi/8) and tan(3pi/8)
// for atan, by 1/2 for asin, and the Cephes minimax polynomials evaluated
// with fused multiply-adds (std::fma rounds once, like the device fma).
// The double instantiation uses libm.
inline float atan_cl(float x)
{
    const float ax = std::fabs(x);
    float a = ax, base = 0.f;
    if (ax > 2.414213562373095f) {
        base = 1.5707963267948966f;
        a = -1.f / ax;
    } else if (ax > 0.4142135623730950f) {
        base = 0.7853981633974483f;
        a = (ax - 1.f) / (ax + 1.f);
    }
    const float z = a * a;
    float p = std::fma(8.05374449538e-2f, z, -1.38776856032e-1f);
    p = std::fma(p, z, 1.99777106478e-1f);
    p = std::fma(p, z, -3.33329491539e-1f);
    const float r = base + std::fma(p * z, a, a);
    return std::copysign(r, x);
}

inline float asin_half_cl(float a)
{
    const float z = a * a;
    float p = std::fma(4.2163199048e-2f, z, 2.4181311049e-2f);
    p = std::fma(p, z, 4.5470025998e-2f);
    p = std::fma(p, z, 7.4953002686e-2f);
    p = std::fma(p, z, 1.6666752422e-1f);
    return std::fma(p * z, a, a);
}

// acos(x)/pi on [0, 1]
inline float acospi_cl(float x)
{
    float ac;
    if (x > 0.5f) {
        const float s = std::sqrt(0.5f * (1.f - x));
        ac = 2.f * asin_half_cl(s);
    } else {
        ac = 1.5707963267948966f - asin_half_cl(x);
    }
    return ac * 0.318309886183791f;
}

template <typename R> inline R cl_atan(R x) { return std::atan(x); }
template <> inline float cl_atan<float>(float x) { return atan_cl(x); }

template <typename R> inline R cl_acospi(R x)
{
    return std::acos(x) / R(3.14159265358979323846);
}
template <> inline float cl_acospi<float>(float x) { return acospi_cl(x); }

// pow(area, 1 / (DIMS - 1)): exponent 1 (2D) or 1/2 (3D), whose exact
// results are the identity and the correctly rounded square root
template <typename R> inline R cl_pow(R x, R y)
{
    if (y == R(1)) return x;
    if (y == R(0.5)) return std::sqrt(x);
    return std::pow(x, y);
}

template <typename R> inline R Omega3(R a, R b)
{
    const R a2 = a * a;
    const R b2 = b * b;
    return cl_acospi<R>(cl_min<R>(
        std::sqrt((R(1) + a2 + b2) / ((R(1) + a2) * (R(1) + b2))), R(1)));
}

template <int D, typename R> inline R kernelS_D(R d, R t, R b, R s)
{
    if (D == 2) {
        const R wcon = R(0.5f * kIPI);
        const R dr = R(0.5f) * s;
        return -wcon * (cl_atan<R>((t + dr) / d) - cl_atan<R>((t - dr) / d));
    }
    const R wcon = R(0.25f);
    const R dr = R(0.5f) * std::sqrt(s);
    const R t1 = (t - dr) / d, t2 = (t + dr) / d;
    const R b1 = (b - dr) / d, b2 = (b + dr) / d;
    const R st = cl_sign(t1), sb = cl_sign(b1);
    return -wcon * (Omega3(t2, b2) - st * Omega3(t1, b2) - sb * Omega3(t2, b1) +
                    st * sb * Omega3(t1, b1));
}

// ---------------------------------------------------------------------------
// Neighbour walk (reference: resources/Scripts/types/3D.h:197-219 and
// types/2D.h:174-193): x offset outermost, then y, then z; within each cell
// the particles in ascending sorted index while icell[j] == cell.  Cell ids
// are computed in 32-bit unsigned arithmetic like the reference.
// ---------------------------------------------------------------------------
template <int D, class F>
inline void walk(usize c_i, usize N, const usize* nc, const usize* icell,
                 const usize* ihoc, F f)
{
    for (int ci = -1; ci <= 1; ci++) {
        for (int cj = -1; cj <= 1; cj++) {
            if (D == 2) {
                const usize c_j = c_i + (usize)ci + (usize)cj * nc[0];
                usize j = ihoc[c_j];
                while ((j < N) && (icell[j] == c_j)) {
                    f(j);
                    j++;
                }
            } else {
                for (int ck = -1; ck <= 1; ck++) {
                    const usize c_j = c_i + (usize)ci + (usize)cj * nc[0] +
                                      (usize)ck * nc[0] * nc[1];
                    usize j = ihoc[c_j];
                    while ((j < N) && (icell[j] == c_j)) {
                        f(j);
                        j++;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Periodic box (Presets/cfd/periodic.xml; no counterpart in the reference,
// whose periodic domains are plane portals): the lengths along the periodic
// axes, 0 along the others, set by orc_set_periodic for the walks below.
// Along a periodic axis the grid tiles the box with n cells (no margins),
// the neighbour cell c + o wraps modulo n and its particles are seen at
// r + shift, shift = -L (wrapped below 0) or +L (above n - 1), added only
// along the periodic axes.
// ---------------------------------------------------------------------------
inline double g_perL[3] = { 0.0, 0.0, 0.0 };
inline bool per_on() { return g_perL[0] > 0 || g_perL[1] > 0 || g_perL[2] > 0; }

// walk<D> handing each candidate's position (at its periodic image) to
// f(j, r_j); without a periodic box the plain walk and r[j]
template <int D, typename R, class F>
inline void walk_r(usize c_i, usize N, const usize* nc, const usize* icell,
                   const usize* ihoc, const R* r, F f)
{
    const int VS = Dim<D>::VS;
    auto pos = [&](usize j) {
        V<D, R> v;
        for (int k = 0; k < 3; k++)
            v.c[k] = (k < D) ? r[(size_t)j * VS + k] : R(0);
        return v;
    };
    if (!per_on()) {
        walk<D>(c_i, N, nc, icell, ihoc, [&](usize j) { f(j, pos(j)); });
        return;
    }
    const int q[3] = { (int)(c_i % nc[0]), (int)((c_i / nc[0]) % nc[1]),
                       D == 3 ? (int)(c_i / (nc[0] * nc[1])) : 0 };
    for (int ci = -1; ci <= 1; ci++)
        for (int cj = -1; cj <= 1; cj++)
            for (int ck = (D == 3 ? -1 : 0); ck <= (D == 3 ? 1 : 0); ck++) {
                const int o[3] = { ci, cj, ck };
                int v[3];
                R sh[3] = { R(0), R(0), R(0) };
                for (int k = 0; k < 3; k++) {
                    v[k] = q[k] + o[k];
                    if (k < D && g_perL[k] > 0) {
                        const int n = (int)nc[k];
                        const R L = R(g_perL[k]);
                        if (v[k] < 0) {
                            v[k] += n;
                            sh[k] = -L;
                        } else if (v[k] >= n) {
                            v[k] -= n;
                            sh[k] = L;
                        }
                    }
                }
                const usize c_j = (usize)v[0] + (usize)v[1] * nc[0] +
                                  (D == 3 ? (usize)v[2] * nc[0] * nc[1] : 0u);
                usize j = ihoc[c_j];
                while ((j < N) && (icell[j] == c_j)) {
                    V<D, R> rj = pos(j);
                    for (int k = 0; k < D; k++)
                        if (g_perL[k] > 0) rj.c[k] = rj.c[k] + sh[k];
                    f(j, rj);
                    j++;
                }
            }
}

// float -> unsigned conversion with the AMDGPU v_cvt_u32_f32 saturation
// (NaN and negatives -> 0, too large -> UINT_MAX); the OpenCL cast is
// implementation-defined outside [0, 2^32).
template <typename R> inline usize cvt_u32(R x)
{
    if (!(x > R(0))) return 0u;
    if (x >= R(4294967296.0)) return 0xFFFFFFFFu;
    return (usize)x;
}

// ---------------------------------------------------------------------------
// Link-list (reference: aquagpusph/CalcServer/LinkList.cpp:185-232 for
// n_cells, LinkList.cl.in:32-113 for the kernels, RadixSort.cpp:129-303 and
// RadixSort.cl.in for the stable LSD sort).
// ---------------------------------------------------------------------------

// component-wise min/max of a vec array with the reference identities
// (Reduction.hcl.in VEC_INFINITY / -VEC_INFINITY)
template <int D, typename R>
void minmax_vec(const R* r, usize N, R* rmin, R* rmax)
{
    const int VS = Dim<D>::VS;
    for (int k = 0; k < VS; k++) {
        const bool w = (k == 3);
        rmin[k] = w ? R(0) : std::numeric_limits<R>::infinity();
        rmax[k] = w ? -R(0) : -std::numeric_limits<R>::infinity();
    }
    for (usize i = 0; i < N; i++)
        for (int k = 0; k < VS; k++) {
            rmin[k] = cl_min(rmin[k], r[(size_t)i * VS + k]);
            rmax[k] = cl_max(rmax[k], r[(size_t)i * VS + k]);
        }
}

// LinkList::nCells: n = (ulong)((max-min)/cell_length) + 6, w = x*y*z.
// cell_length = support*h computed in float at setup.
template <int D, typename R>
void n_cells(const R* rmin, const R* rmax, R cell_length, usize* nc)
{
    uint64_t n[3] = { 1, 1, 1 };
    for (int k = 0; k < D; k++) {
        const R v = (rmax[k] - rmin[k]) / cell_length;
        n[k] = (uint64_t)cvt_u32(v) + 6;
    }
    nc[0] = (usize)n[0];
    nc[1] = (usize)n[1];
    nc[2] = (usize)n[2];
    nc[3] = (usize)(n[0] * n[1] * n[2]);
}

// iCell kernel, LinkList.cl.in:54-85
template <int D, typename R>
void icell(usize* ic, const R* r, usize N, const R* rmin, R support, R h,
           const usize* nc)
{
    const int VS = Dim<D>::VS;
    const R idist = R(1) / (support * h);
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        const usize cx = cvt_u32((r[(size_t)i * VS + 0] - rmin[0]) * idist) + 3u;
        const usize cy = cvt_u32((r[(size_t)i * VS + 1] - rmin[1]) * idist) + 3u;
        if (D == 3) {
            const usize cz =
                cvt_u32((r[(size_t)i * VS + 2] - rmin[2]) * idist) + 3u;
            ic[i] = cx - 1u + (cy - 1u) * nc[0] + (cz - 1u) * nc[0] * nc[1];
        } else {
            ic[i] = cx - 1u + (cy - 1u) * nc[0];
        }
    }
}

// Number of key bits the reference radix sort processes
// (RadixSort.cpp:152-176): bits of nextPowerOf2(max_val), rounded up to the
// 4-bit digit.  max_val = n_cells.w for the link-list, UINT_MAX>>1 otherwise.
inline unsigned radix_key_bits(uint64_t max_val)
{
    uint64_t p = 1;
    while (p < max_val) p <<= 1;  // AuxiliarMethods nextPowerOf2
    unsigned bits = 1;
    for (; (p & 1) == 0; p >>= 1, bits++)
        ;
    return (bits + 3u) / 4u * 4u;
}

// Stable LSD radix sort result (RadixSort.cl.in:35-323): keys sorted on their
// low `key_bits` bits, perm[k] = original index of the k-th sorted key
// (id_unsorted) and inv[perm[k]] = k (id_sorted).  A stable LSD sort over
// contiguous per-thread chunks with a radix-major exclusive scan equals a
// stable sort by the masked key.
inline void stable_sort(usize* keys, usize* perm, usize* inv, usize n,
                        unsigned key_bits)
{
    const uint64_t mask =
        (key_bits >= 32) ? 0xFFFFFFFFull : ((1ull << key_bits) - 1ull);
    std::vector<usize> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    std::vector<usize> k(keys, keys + n);
    std::stable_sort(idx.begin(), idx.end(), [&](usize a, usize b) {
        return (k[a] & mask) < (k[b] & mask);
    });
    for (usize s = 0; s < n; s++) {
        keys[s] = k[idx[s]];
        perm[s] = idx[s];
        inv[idx[s]] = s;
    }
}

// iHoc + linkList kernels, LinkList.cl.in:32-42 and 92-113
inline void ihoc(const usize* ic, usize* hoc, usize N, const usize* nc)
{
    for (usize c = 0; c < nc[3]; c++) hoc[c] = N;
    if (N < 1) return;
    for (usize i = 0; i + 1 < N; i++) {
        const usize c = ic[i], c2 = ic[i + 1];
        if (i == 0) hoc[c] = 0;
        if (c2 != c) hoc[c2] = i + 1;
    }
}

// ---------------------------------------------------------------------------
// Reorder, resources/Scripts/basic/Sort.cl:57-124 (stage1 and stage2 after
// the "Backup *" copies of Presets/src/basic.xml:141-146): x[id_sorted[i]] =
// x_in[i].  Generic element-size scatter.
// ---------------------------------------------------------------------------
inline void scatter(const void* src, void* dst, size_t esize,
                    const usize* id_sorted, usize N)
{
    const char* s = (const char*)src;
    char* d = (char*)dst;
    for (usize i = 0; i < N; i++)
        std::memcpy(d + (size_t)id_sorted[i] * esize, s + (size_t)i * esize,
                    esize);
}

// ---------------------------------------------------------------------------
// Streaming kernels
// ---------------------------------------------------------------------------

// improved Euler predictor, basic/time_scheme/improved_euler.cl:75-103
template <int D, typename R>
void predictor(const int* imove, const R* r, const R* u, const R* dudt,
               const R* rho, const R* drhodt, R* r_in, R* u_in, R* dudt_in,
               R* rho_in, R* drhodt_in, usize N, R dt)
{
    const int VS = Dim<D>::VS;
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const size_t i = (size_t)ii;
        R DT = dt;
        if (imove[i] <= 0) DT = R(0);
        for (int k = 0; k < VS; k++) {
            const size_t a = i * VS + k;
            dudt_in[a] = dudt[a];
            u_in[a] = u[a] + DT * dudt[a];
            r_in[a] = r[a] + DT * u[a] + R(0.5f) * DT * DT * dudt[a];
        }
        drhodt_in[i] = drhodt[i];
        rho_in[i] = rho[i] + DT * drhodt[i];
    }
}

// improved Euler corrector, basic/time_scheme/improved_euler.cl:125-147
template <int D, typename R>
void corrector(const int* imove, R* r, R* u, const R* dudt, R* rho,
               const R* drhodt, const R* dudt_in, const R* drhodt_in, usize N,
               R dt)
{
    const int VS = Dim<D>::VS;
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const size_t i = (size_t)ii;
        if (imove[i] > 0) {
            const R DT = R(0.5f) * dt;
            for (int k = 0; k < VS; k++) {
                const size_t a = i * VS + k;
                u[a] += DT * (dudt[a] - dudt_in[a]);
            }
            for (int k = 0; k < VS; k++) {
                const size_t a = i * VS + k;
                r[a] += DT * DT * (dudt[a] - dudt_in[a]);
            }
            rho[i] += DT * (drhodt[i] - drhodt_in[i]);
        }
    }
}

// The precision mode's improved Euler (Presets/precision.xml): the same
// update with the position carried in double as r + dr64 (r float, dr64 the
// double remainder); velocity and density as above.
template <int D>
void predictor64(const int* imove, const float* r, const double* dr64,
                 const float* u, const float* dudt, const float* rho,
                 const float* drhodt, float* r_in, double* dr64_in,
                 float* u_in, float* dudt_in, float* rho_in,
                 float* drhodt_in, usize N, float dt)
{
    const int VS = Dim<D>::VS;
    for (size_t i = 0; i < N; i++) {
        float DT = dt;
        if (imove[i] <= 0) DT = 0.f;
        const float hdt2 = 0.5f * DT * DT;
        const double dDT = DT, dh = hdt2;
        for (int k = 0; k < VS; k++) {
            const size_t a = i * VS + k;
            dudt_in[a] = dudt[a];
            u_in[a] = u[a] + DT * dudt[a];
            const double x = ((double)r[a] + dr64[a]) + dDT * (double)u[a] +
                             dh * (double)dudt[a];
            r_in[a] = (float)x;
            dr64_in[a] = x - (double)r_in[a];
        }
        drhodt_in[i] = drhodt[i];
        rho_in[i] = rho[i] + DT * drhodt[i];
    }
}

template <int D>
void corrector64(const int* imove, float* r, double* dr64, float* u,
                 const float* dudt, float* rho, const float* drhodt,
                 const float* dudt_in, const float* drhodt_in, usize N,
                 float dt)
{
    const int VS = Dim<D>::VS;
    for (size_t i = 0; i < N; i++) {
        if (imove[i] <= 0) continue;
        const float DT = 0.5f * dt;
        const double DT2 = (double)DT * (double)DT;
        for (int k = 0; k < VS; k++) {
            const size_t a = i * VS + k;
            const float da = dudt[a] - dudt_in[a];
            u[a] += DT * da;
            const double x = ((double)r[a] + dr64[a]) + DT2 * (double)da;
            r[a] = (float)x;
            dr64[a] = x - (double)r[a];
        }
        rho[i] += DT * (drhodt[i] - drhodt_in[i]);
    }
}

// basic/Domain.cl:48-90
template <int D, typename R>
void domain(int* imove, R* r_in, R* u_in, R* dudt_in, R* m, usize N,
            const R* dmin, const R* dmax)
{
    const int VS = Dim<D>::VS;
    for (usize i = 0; i < N; i++) {
        if (imove[i] <= -255) continue;
        const R* c = r_in + (size_t)i * VS;
        bool out = false;
        for (int k = 0; k < D; k++) {
            if (std::isnan(c[k]) || std::isinf(c[k]) || c[k] < dmin[k] ||
                c[k] > dmax[k])
                out = true;
        }
        if (out) {
            imove[i] = -256;
            m[i] = R(0);
            for (int k = 0; k < VS; k++) {
                u_in[(size_t)i * VS + k] = R(0);
                dudt_in[(size_t)i * VS + k] = R(0);
                r_in[(size_t)i * VS + k] = dmax[k];
            }
        }
    }
}

// basic/EOS.cl:57-73 (excluded: imove <= 0 && imove != -1)
template <typename R>
void eos(const unsigned* iset, const int* imove, const R* rho, R* p,
         const R* refd, usize N, R cs, R p0)
{
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const size_t i = (size_t)ii;
        if ((imove[i] <= 0) && (imove[i] != -1)) continue;
        p[i] = p0 + cs * cs * (rho[i] - refd[iset[i]]);
    }
}

// Tait EOS (no counterpart in the reference; the linear EOS above is its
// only liquid law): p = p0 + B ((rho / rho0)^gamma - 1), B = rho0 cs^2 /
// gamma, for the particles of basic/EOS.cl:35.  An integral gamma in
// [1, 16] is raised by binary powering in the device's order of products.
template <typename R> R tait_pow(R x, R g)
{
    const R gi = std::nearbyint(g);
    if (gi == g && gi >= R(1) && gi <= R(16)) {
        unsigned n = (unsigned)gi;
        R r = R(1), b = x;
        bool first = true;
        while (n) {
            if (n & 1u) {
                r = first ? b : r * b;
                first = false;
            }
            n >>= 1;
            if (n) b = b * b;
        }
        return r;
    }
    return std::pow(x, g);
}

template <typename R>
void tait_eos(const unsigned* iset, const int* imove, const R* rho, R* p,
              const R* refd, usize N, R cs, R p0, R gamma)
{
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const size_t i = (size_t)ii;
        if ((imove[i] <= 0) && (imove[i] != -1)) continue;
        const R rho0 = refd[iset[i]];
        const R B = rho0 * cs * cs / gamma;
        p[i] = p0 + B * (tait_pow(rho[i] / rho0, gamma) - R(1));
    }
}

// cfd/Boundary/BI/InterpolationShepard.cl:48-76 with the Tait inverse for
// the element density, rho = rho0 ((p - p0) / B + 1)^(1 / gamma)
template <typename R>
void bi_interp_shepard_tait(const unsigned* iset, const int* imove,
                            const R* shepard, R* rho, R* p, const R* refd,
                            usize N, R cs, R p0, R gamma)
{
    for (usize i = 0; i < N; i++) {
        if (imove[i] != -3) continue;
        R s = shepard[i];
        if (s < R(1.0E-6f)) s = R(1);
        p[i] /= s;
        const R rho0 = refd[iset[i]];
        const R B = rho0 * cs * cs / gamma;
        rho[i] = rho0 * std::pow((p[i] - p0) / B + R(1), R(1) / gamma);
    }
}

// basic/Binormal.cl:37-53
template <int D, typename R>
void binormal(const R* normal, R* tangent, R* binormal, usize N)
{
    for (usize i = 0; i < N; i++) {
        if (D == 2) {
            binormal[2 * i] = R(0);
            binormal[2 * i + 1] = R(0);
            tangent[2 * i] = normal[2 * i + 1];
            tangent[2 * i + 1] = -normal[2 * i];
        } else {
            const R* a = normal + 4 * (size_t)i;
            const R* b = tangent + 4 * (size_t)i;
            R* c = binormal + 4 * (size_t)i;
            const R x = a[1] * b[2] - a[2] * b[1];
            const R y = a[2] * b[0] - a[0] * b[2];
            const R z = a[0] * b[1] - a[1] * b[0];
            c[0] = x;
            c[1] = y;
            c[2] = z;
            c[3] = R(0);
        }
    }
}

// cfd/Rates.cl:55-77
template <int D, typename R>
void rates(const unsigned* iset, const int* imove, const R* rho,
           const R* grad_p, const R* lap_u, const R* div_u, R* dudt,
           R* drhodt, const R* visc_dyn, usize N, const R* g)
{
    (void)rho;
    const int VS = Dim<D>::VS;
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const size_t i = (size_t)ii;
        if (imove[i] != 1) continue;
        const R mu = visc_dyn[iset[i]];
        for (int k = 0; k < VS; k++) {
            const size_t a = i * VS + k;
            dudt[a] = -grad_p[a] + mu * lap_u[a] + g[k];
        }
        drhodt[i] = -div_u[i];
    }
}

// basic/deltaSPH.cl:330-350 with cfd/deltaSPH.cl:29 (fluid only)
template <typename R>
void delta_rate(const unsigned* iset, const int* imove, const R* rho,
                const R* lap_p, R* drhodt, const R* refd, const R* delta,
                usize N, R dt)
{
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const size_t i = (size_t)ii;
        if (imove[i] != 1) continue;
        const unsigned s = iset[i];
        const R delta_f = delta[s] * dt * rho[i] / refd[s];
        drhodt[i] += delta_f * lap_p[i];
    }
}

// cfd/TimeStep.cl:56-77; length() of the full vec (w included)
template <int D, typename R>
void timestep(const int* imove, const R* u, R* dt_var, usize N, R dt,
              R dt_min, R courant, R dt_Ma, R h)
{
    const int VS = Dim<D>::VS;
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const size_t i = (size_t)ii;
        if (imove[i] <= 0) {
            dt_var[i] = dt;
            continue;
        }
        const R dr_max = dt_Ma * h;
        R s = u[i * VS] * u[i * VS];
        for (int k = 1; k < VS; k++) s = s + u[i * VS + k] * u[i * VS + k];
        const R dt_u = courant * dr_max / std::sqrt(s);
        dt_var[i] = cl_max(cl_min(dt, dt_u), dt_min);
    }
}

// cfd/Boundary/BI/Shepard.cl:173-195 (apply)
template <int D, typename R>
void bi_shepard_apply(const int* imove, const R* shepard, R* grad_p,
                      R* lap_u, R* div_u, usize N)
{
    const int VS = Dim<D>::VS;
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const size_t i = (size_t)ii;
        if (imove[i] != 1) continue;
        const R s = shepard[i];
        for (int k = 0; k < VS; k++) grad_p[i * VS + k] /= s;
        for (int k = 0; k < VS; k++) lap_u[i * VS + k] /= s;
        div_u[i] /= s;
    }
}

// cfd/Boundary/BI/GradP.cl:55-83 (freeslip)
template <int D, typename R>
void bi_gradp(const unsigned* iset, const int* imove, const R* shepard,
              const R* rho, const R* lap_u, const R* dudt, R* grad_p,
              const R* visc_dyn, const R* refd, usize N, const R* g)
{
    (void)rho;
    const int VS = Dim<D>::VS;
    for (usize i = 0; i < N; i++) {
        if (imove[i] != -3) continue;
        R s = shepard[i];
        if (s < R(1.0E-6f)) s = R(1);
        const R f = visc_dyn[iset[i]] / refd[iset[i]];
        for (int k = 0; k < VS; k++) {
            const size_t a = (size_t)i * VS + k;
            grad_p[a] = f * lap_u[a] / s + g[k] - dudt[a];
        }
    }
}

// cfd/Boundary/BI/InterpolationShepard.cl:48-76
template <typename R>
void bi_interp_shepard(const unsigned* iset, const int* imove,
                       const R* shepard, R* rho, R* p, const R* refd, usize N,
                       R cs, R p0)
{
    for (usize i = 0; i < N; i++) {
        if (imove[i] != -3) continue;
        R s = shepard[i];
        if (s < R(1.0E-6f)) s = R(1);
        p[i] /= s;
        rho[i] = refd[iset[i]] + (p[i] - p0) / (cs * cs);
    }
}

// cfd/SensorsRenormalization.cl:42-67
template <int D, typename R>
void sensors_renorm(const int* imove, const R* shepard, R* u, R* rho, R* p,
                    usize N)
{
    const int VS = Dim<D>::VS;
    for (usize i = 0; i < N; i++) {
        if (imove[i] != 0) continue;
        R s = shepard[i];
        if (s < R(1.0E-6f)) s = R(1);
        for (int k = 0; k < VS; k++) u[(size_t)i * VS + k] /= s;
        rho[i] /= s;
        p[i] /= s;
    }
}

// The grid and the cell hash of a periodic box (periodic_min pmin,
// lengths L; the axes with L == 0 as in n_cells / icell above): n =
// trunc(L / cell_length) cells along a periodic axis (at least 3), the cell
// coordinate min(trunc((r - pmin) n / L), n - 1).
template <int D, typename R>
bool n_cells_per(const R* rmin, const R* rmax, const R* pmin, const R* L,
                 R cell_length, usize* nc, R* rmin_out)
{
    uint64_t n[3] = { 1, 1, 1 };
    bool ok = true;
    for (int k = 0; k < D; k++) {
        if (L[k] > R(0)) {
            n[k] = (uint64_t)cvt_u32(L[k] / cell_length);
            ok = ok && n[k] >= 3;
            rmin_out[k] = pmin[k];
            continue;
        }
        const R v = (rmax[k] - rmin[k]) / cell_length;
        n[k] = (uint64_t)cvt_u32(v) + 6;
        rmin_out[k] = rmin[k];
    }
    nc[0] = (usize)n[0];
    nc[1] = (usize)n[1];
    nc[2] = (usize)n[2];
    nc[3] = (usize)(n[0] * n[1] * n[2]);
    return ok;
}

template <int D, typename R>
void icell_per(usize* ic, const R* r, usize N, const R* rmin, R support, R h,
               const usize* nc, const R* pmin, const R* L)
{
    const int VS = Dim<D>::VS;
    const R idist = R(1) / (support * h);
    for (usize i = 0; i < N; i++) {
        usize q[3] = { 0, 0, 0 };
        for (int k = 0; k < D; k++) {
            const R x = r[(size_t)i * VS + k];
            if (L[k] > R(0)) {
                const R inv = R((R)nc[k]) / L[k];
                q[k] = std::min<usize>(cvt_u32((x - pmin[k]) * inv),
                                       nc[k] - 1u);
            } else {
                q[k] = cvt_u32((x - rmin[k]) * idist) + 2u;
            }
        }
        ic[i] = q[0] + q[1] * nc[0] + (D == 3 ? q[2] * nc[0] * nc[1] : 0u);
    }
}

// cfd periodic wrap: r back into [pmin, pmin + L) along the periodic axes
// (one wrap: a particle moves less than L in a step)
template <int D, typename R>
void periodic_wrap(R* r, usize N, const R* pmin, const R* L)
{
    const int VS = Dim<D>::VS;
    for (usize i = 0; i < N; i++)
        for (int k = 0; k < D; k++) {
            if (!(L[k] > R(0))) continue;
            R x = r[(size_t)i * VS + k];
            if (x < pmin[k]) x = x + L[k];
            else if (x >= pmin[k] + L[k]) x = x - L[k];
            r[(size_t)i * VS + k] = x;
        }
}

// ---------------------------------------------------------------------------
// Pair kernels.  H, CONW, CONF are the preset definitions after the
// reference's "%#G" formatting (CalcServer.cpp:241-260).
// ---------------------------------------------------------------------------
struct LL {
    const usize* icell;
    const usize* ihoc;
    const usize* nc;
};

// basic/Shepard.cl:76-125; `cfd_excl` selects the cfd variant
// (cfd/Shepard.cl:29: EXCLUDED_PARTICLE = imove != 1) over the basic default
// (imove >= 3).
template <int D, typename R>
void shepard(const int* imove, const R* r, const R* rho, const R* m,
             R* shep, usize N, LL ll, R H, R CONW, int cfd_excl)
{
    auto excl = [&](usize k) {
        return cfd_excl ? (imove[k] != 1) : (imove[k] >= 3);
    };
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if ((imove[i] < -3) || ((imove[i] > 0) && excl(i))) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        R acc = shep[i];
        walk_r<D, R>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, r,
                        [&](usize j, const V<D, R>& r_j) {
            if (excl(j)) return;
            const V<D, R> r_ij = sub<D, R>(r_j, r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            acc += kernelW<D, R>(q) * CONW * m[j] / rho[j];
        });
        shep[i] = acc;
    }
}

// s + x with every addition's rounding error collected in c (Kahan-Babuska /
// Neumaier compensated summation: the larger operand's low bits are the
// ones lost).  The precision mode's walk sums with it; the result is s + c.
template <typename R> inline void nsum(R& s, R& c, R x)
{
    const R t = s + x;
    const R lo = (std::fabs(s) >= std::fabs(x)) ? ((s - t) + x) : ((x - t) + s);
    c += lo;
    s = t;
}

// cfd/Interactions.cl:60-145 (Monaghan-Cleary Laplacian when lap_form==1,
// Morris when lap_form==2); kahan: the same terms summed with nsum
template <int D, typename R>
void interactions(const int* imove, const R* r, const R* u, const R* rho,
                  const R* m, const R* p, R* grad_p, R* lap_u, R* div_u,
                  usize N, LL ll, R H, R CONF, int lap_form, bool kahan = false)
{
    const int VS = Dim<D>::VS;
    const R CLEARY = (D == 2) ? R(8.f) : R(10.f);
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if (imove[i] != 1) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        const V<D, R> u_i = ld<D, R>(u, i);
        const R p_i = p[i];
        const R rho_i = rho[i];
        V<D, R> gp = ld<D, R>(grad_p, i), lu = ld<D, R>(lap_u, i);
        R du = div_u[i];
        R cg[3] = { 0, 0, 0 }, cl[3] = { 0, 0, 0 }, cd = 0;
        walk_r<D, R>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, r,
                        [&](usize j, const V<D, R>& r_j) {
            if (i == j) return;
            if (imove[j] != 1) return;
            const V<D, R> r_ij = sub<D, R>(r_j, r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const R rho_j = rho[j];
            const R p_j = p[j];
            const V<D, R> u_ij = sub<D, R>(ld<D, R>(u, j), u_i);
            const R udr = dot<D, R>(u_ij, r_ij);
            const R f_ij = kernelF<D, R>(q) * CONF * m[j];
            const R sg = (p_i + p_j) / (rho_i * rho_j) * f_ij;
            R sl;
            const V<D, R>* lv;
            if (lap_form == 1) {
                const R r2 = (q * q + R(0.01f)) * H * H;
                sl = f_ij * CLEARY * udr / (r2 * rho_i * rho_j);
                lv = &r_ij;
            } else {
                sl = f_ij * R(2.f) / (rho_i * rho_j);
                lv = &u_ij;
            }
            const R td = udr * f_ij * rho_i / rho_j;
            if (kahan) {
                for (int k = 0; k < D; k++) {
                    nsum(gp.c[k], cg[k], sg * r_ij.c[k]);
                    nsum(lu.c[k], cl[k], sl * lv->c[k]);
                }
                nsum(du, cd, td);
            } else {
                for (int k = 0; k < D; k++) {
                    gp.c[k] += sg * r_ij.c[k];
                    lu.c[k] += sl * lv->c[k];
                }
                du += td;
            }
        });
        for (int k = 0; k < D; k++) {
            grad_p[(size_t)i * VS + k] = kahan ? gp.c[k] + cg[k] : gp.c[k];
            lap_u[(size_t)i * VS + k] = kahan ? lu.c[k] + cl[k] : lu.c[k];
        }
        div_u[i] = kahan ? du + cd : du;
    }
}

// basic/deltaSPH.cl:191-242 (lapp) with cfd/deltaSPH.cl:29 (fluid only);
// kahan: summed with nsum
template <int D, typename R>
void lapp(const int* imove, const R* r, const R* rho, const R* m, const R* p,
          R* lap_p, usize N, LL ll, R H, R CONF, bool kahan = false)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if (imove[i] != 1) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        const R p_i = p[i];
        R acc = lap_p[i], c = 0;
        walk_r<D, R>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, r,
                        [&](usize j, const V<D, R>& r_j) {
            if ((i == j) || (imove[j] != 1)) return;
            const V<D, R> r_ij = sub<D, R>(r_j, r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const R f_ij = kernelF<D, R>(q) * CONF * m[j] / rho[j];
            if (kahan) nsum(acc, c, (p[j] - p_i) * f_ij);
            else acc += (p[j] - p_i) * f_ij;
        });
        lap_p[i] = kahan ? acc + c : acc;
    }
}

// cfd/Sensors.cl:57-130
template <int D, typename R>
void sensors(const int* imove, const R* r, const R* m, R* u, R* rho, R* p,
             usize N, LL ll, R H, R CONW)
{
    const int VS = Dim<D>::VS;
    for (usize i = 0; i < N; i++) {
        if (imove[i] != 0) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        R uu[3] = { 0, 0, 0 }, rr = 0, pp = 0;
        walk_r<D, R>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, r,
                     [&](usize j, const V<D, R>& r_j) {
            if (i == j) return;
            if (imove[j] != 1) return;
            const V<D, R> r_ij = sub<D, R>(r_j, r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const R rho_j = rho[j];
            const R w_ij = kernelW<D, R>(q) * CONW * m[j] / rho_j;
            for (int k = 0; k < D; k++) uu[k] += u[(size_t)j * VS + k] * w_ij;
            rr += rho_j * w_ij;
            pp += p[j] * w_ij;
        });
        for (int k = 0; k < D; k++) u[(size_t)i * VS + k] = uu[k];
        rho[i] = rr;
        p[i] = pp;
    }
}

// cfd/Boundary/BI/Shepard.cl:61-150 (compute)
template <int D, typename R>
void bi_shepard_compute(const int* imove, const R* r, const R* normal,
                        const R* tangent, const R* binormal, const R* m,
                        R* shep, usize N, LL ll, R H, R CONW, R DIMS)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if ((imove[i] < -3) || (imove[i] > 1)) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        R acc = R(1);
        bool self_added = false;
        walk<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if (imove[j] != -3) return;
            if (i == j) {
                if (!self_added) {
                    self_added = true;
                    acc -= R(0.5f);
                }
                return;
            }
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const V<D, R> n_j = ld<D, R>(normal, j);
            const V<D, R> t_j = ld<D, R>(tangent, j);
            const V<D, R> b_j = ld<D, R>(binormal, j);
            const R r_n = dot<D, R>(r_ij, n_j);
            const R area_j = m[j];
            if ((r_n > R(-1E-8f) * H) && (r_n < R(1E-8f) * H)) {
                if (!self_added) {
                    const R r_t = std::fabs(dot<D, R>(r_ij, t_j));
                    const R r_b = std::fabs(dot<D, R>(r_ij, b_j));
                    const R dr = R(0.55f) *
                                 cl_pow<R>(area_j, R(1.f) / (DIMS - R(1.f)));
                    if ((r_t <= dr) && (r_b <= dr)) {
                        self_added = true;
                        acc -= R(0.5f);
                    }
                }
                return;
            }
            const R r_t = std::fabs(dot<D, R>(r_ij, t_j));
            const R r_b = std::fabs(dot<D, R>(r_ij, b_j));
            acc += r_n * CONW * kernelS_P<D, R>(q) * area_j +
                   kernelS_D<D, R>(std::fabs(r_n), r_t, r_b, area_j);
        });
        shep[i] = acc;
    }
}

// cfd/Boundary/BI/LapU.cl:58-123 (freeslip)
template <int D, typename R>
void bi_lapu(const int* imove, const R* r, const R* u, const R* rho,
             const R* m, R* lap_u, usize N, LL ll, R H, R CONF, int lap_form)
{
    const int VS = Dim<D>::VS;
    const R CLEARY = (D == 2) ? R(8.f) : R(10.f);
#pragma omp parallel for schedule(dynamic, 64)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if (imove[i] != -3) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        const V<D, R> u_i = ld<D, R>(u, i);
        V<D, R> lu = ld<D, R>(lap_u, i);
        walk<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if (i == j) return;
            if (imove[j] != 1) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const R rho_j = rho[j];
            const R f_ij = kernelF<D, R>(q) * CONF * m[j];
            const V<D, R> u_ij = sub<D, R>(ld<D, R>(u, j), u_i);
            if (lap_form == 1) {
                const R udr = dot<D, R>(u_ij, r_ij);
                const R r2 = (q * q + R(0.01f)) * H * H;
                const R sl = f_ij * CLEARY * udr / (r2 * rho_j);
                for (int k = 0; k < D; k++) lu.c[k] += sl * r_ij.c[k];
            } else {
                const R sl = f_ij * R(2.f) / rho_j;
                for (int k = 0; k < D; k++) lu.c[k] += sl * u_ij.c[k];
            }
        });
        for (int k = 0; k < D; k++) lap_u[(size_t)i * VS + k] = lu.c[k];
    }
}

// cfd/Boundary/BI/Interpolation.cl:54-107
template <int D, typename R>
void bi_interp(const int* imove, const R* r, const R* m, const R* rho,
               const R* grad_p, R* p, usize N, LL ll, R H, R CONW)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if (imove[i] != -3) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        V<D, R> gp_i = ld<D, R>(grad_p, i);
        for (int k = 0; k < D; k++) gp_i.c[k] = rho[i] * gp_i.c[k];
        R acc = R(0);
        walk<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if (imove[j] != 1) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const R w_ij = kernelW<D, R>(q) * CONW * m[j] / rho[j];
            acc += (p[j] - dot<D, R>(gp_i, r_ij)) * w_ij;
        });
        p[i] = acc;
    }
}

// cfd/Boundary/BI/Interactions.cl:51-133
template <int D, typename R>
void bi_interactions(const int* imove, const R* r, const R* normal,
                     const R* u, const R* rho, const R* m, const R* p,
                     R* grad_p, R* div_u, usize N, LL ll, R H, R CONW)
{
    const int VS = Dim<D>::VS;
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if (imove[i] != 1) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        const V<D, R> u_i = ld<D, R>(u, i);
        const R p_i = p[i];
        const R rho_i = rho[i];
        V<D, R> gp = ld<D, R>(grad_p, i);
        R dv = div_u[i];
        walk<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if (imove[j] != -3) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const V<D, R> n_j = ld<D, R>(normal, j);
            const R area_j = m[j];
            const R p_j = p[j];
            const V<D, R> du = sub<D, R>(ld<D, R>(u, j), u_i);
            const R w_ij = kernelW<D, R>(q) * CONW * area_j;
            const R sg = (p_i + p_j) / rho_i * w_ij;
            for (int k = 0; k < D; k++) gp.c[k] += sg * n_j.c[k];
            dv += rho_i * dot<D, R>(du, n_j) * w_ij;
        });
        for (int k = 0; k < D; k++) grad_p[(size_t)i * VS + k] = gp.c[k];
        div_u[i] = dv;
    }
}

// ---------------------------------------------------------------------------
// Ghost-particle walls (resources/Scripts/cfd/Boundary/GP/*.cl): ghosts are
// imove == -1, their mirroring elements imove == -2 (unit outward normal,
// area in m), associations[i] the element of ghost i (>= N: none).
// ---------------------------------------------------------------------------

// The walk of a particle whose cell is given explicitly (GP/Interpolation.cl
// walks around the mirrored position, gp_icell).  Neighbour cells outside
// the grid (c_j >= n_cells.w, e.g. a mirrored point beyond the particle
// bounding box) are skipped; the reference reads ihoc out of range there.
template <int D, class F>
inline void walk_cell(usize c_i, usize N, const usize* nc, const usize* icell,
                      const usize* ihoc, F f)
{
    for (int ci = -1; ci <= 1; ci++)
        for (int cj = -1; cj <= 1; cj++)
            for (int ck = (D == 3 ? -1 : 0); ck <= (D == 3 ? 1 : 0); ck++) {
                const usize c_j = c_i + (usize)ci + (usize)cj * nc[0] +
                                  (usize)ck * nc[0] * nc[1];
                if (c_j >= nc[3]) continue;
                usize j = ihoc[c_j];
                while ((j < N) && (icell[j] == c_j)) {
                    f(j);
                    j++;
                }
            }
}

// GP/Sort.cl:35-53: associations into the new sorted order
inline void gp_sort(const usize* assoc_in, usize* assoc, const usize* id_sorted,
                    usize N)
{
    for (usize i = 0; i < N; i++) {
        const usize i_out = id_sorted[i];
        assoc[i_out] = (assoc_in[i] >= N) ? N : id_sorted[assoc_in[i]];
    }
}

// GP/Mirror.cl:47-90: reflect each ghost across its element and hash the
// mirrored point (iCell formula with the SUPPORT and H definitions)
template <int D, typename R>
void gp_mirror(const int* imove, const usize* assoc, const R* normal, R* r,
               usize* gp_icell, usize N, const R* rmin, const usize* nc, R H,
               R SUPPORT)
{
    const int VS = Dim<D>::VS;
    const R idist = R(1) / (SUPPORT * H);
    for (usize i = 0; i < N; i++) {
        if (imove[i] != -1) continue;
        const usize iref = assoc[i];
        if (iref >= N) continue;
        const V<D, R> r_ref = ld<D, R>(r, iref);
        const V<D, R> n_ref = ld<D, R>(normal, iref);
        const R d = dot<D, R>(sub<D, R>(r_ref, ld<D, R>(r, i)), n_ref);
        for (int k = 0; k < D; k++)
            r[(size_t)i * VS + k] += R(2.f) * (d * n_ref.c[k]);
        const usize cx = cvt_u32((r[(size_t)i * VS + 0] - rmin[0]) * idist) + 3u;
        const usize cy = cvt_u32((r[(size_t)i * VS + 1] - rmin[1]) * idist) + 3u;
        if (D == 3) {
            const usize cz =
                cvt_u32((r[(size_t)i * VS + 2] - rmin[2]) * idist) + 3u;
            gp_icell[i] = cx - 1u + (cy - 1u) * nc[0] + (cz - 1u) * nc[0] * nc[1];
        } else {
            gp_icell[i] = cx - 1u + (cy - 1u) * nc[0];
        }
    }
}

// GP/Interpolation.cl:58-140: SPH interpolation of rho, p, u (and the
// Shepard sum) at the mirrored ghost position
template <int D, typename R>
void gp_interp(const int* imove, const R* r, const R* m, const R* rho,
               const R* p, const R* u, R* gp_rho, R* gp_p, R* gp_u,
               R* shepard, usize N, const usize* gp_icell, LL ll, R H, R CONW)
{
    const int VS = Dim<D>::VS;
#pragma omp parallel for schedule(dynamic, 64)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if (imove[i] != -1) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        R a_rho = gp_rho[i], a_p = gp_p[i], sh = R(0);
        V<D, R> a_u = ld<D, R>(gp_u, i);
        walk_cell<D>(gp_icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if (imove[j] != 1) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const R rho_j = rho[j];
            const R w_ij = kernelW<D, R>(q) * CONW * m[j] / rho_j;
            a_rho += w_ij * rho_j;
            a_p += w_ij * p[j];
            const V<D, R> u_j = ld<D, R>(u, j);
            for (int k = 0; k < D; k++) a_u.c[k] += w_ij * u_j.c[k];
            sh += w_ij;
        });
        gp_rho[i] = a_rho;
        gp_p[i] = a_p;
        for (int k = 0; k < D; k++) gp_u[(size_t)i * VS + k] = a_u.c[k];
        shepard[i] = sh;
    }
}

// GP/Renormalization.cl:44-69 (the whole vec, w included, is divided)
template <int D, typename R>
void gp_renorm(const int* imove, const R* shepard, R* gp_rho, R* gp_p,
               R* gp_u, usize N)
{
    const int VS = Dim<D>::VS;
    for (usize i = 0; i < N; i++) {
        if (imove[i] != -1) continue;
        R s = shepard[i];
        if (s < R(1.0E-6f)) s = R(1);
        gp_p[i] /= s;
        gp_rho[i] /= s;
        for (int k = 0; k < VS; k++) gp_u[(size_t)i * VS + k] /= s;
    }
}

// GP/PreLapU.cl:51-90: interpolated density; the velocity for the wall
// Laplacian by the Takeda extension (model 1) or frozen (model 2)
template <int D, typename R>
void gp_prelapu(const int* imove, const usize* assoc, const R* normal,
                const R* gp_rho, const R* gp_u, R* rho, R* u, usize N,
                int model)
{
    const int VS = Dim<D>::VS;
    for (usize i = 0; i < N; i++) {
        if (imove[i] != -1) continue;
        const usize iref = assoc[i];
        if (iref >= N) continue;
        rho[i] = gp_rho[i];
        if (model != 1) continue;
        const V<D, R> n = ld<D, R>(normal, iref);
        const V<D, R> ur = ld<D, R>(u, iref);
        const R dr = dot<D, R>(ur, n);
        const V<D, R> gu = ld<D, R>(gp_u, i);
        const R dg = dot<D, R>(gu, n);
        for (int k = 0; k < D; k++) {
            const R u_ref = ur.c[k] - dr * n.c[k];
            const R u_n = dg * n.c[k];
            const R u_t = R(2.f) * u_ref - (gu.c[k] - u_n);
            u[(size_t)i * VS + k] = u_n + u_t;
        }
    }
}

// GP/FreeSlip.cl:46-66
template <int D, typename R>
void gp_freeslip(const unsigned* iset, const int* imove, const usize* assoc,
                 const R* gp_u, R* u, usize N, unsigned freeslip_iset)
{
    const int VS = Dim<D>::VS;
    for (usize i = 0; i < N; i++) {
        if ((imove[i] != -1) || (iset[i] != freeslip_iset)) continue;
        if (assoc[i] >= N) continue;
        for (int k = 0; k < D; k++)
            u[(size_t)i * VS + k] = gp_u[(size_t)i * VS + k];
    }
}

// GP/LapU.cl:58-127: fluid (and elements) against ghosts (and fluid), the
// fluid-fluid and ghost-ghost pairs excluded
template <int D, typename R>
void gp_lapu(const int* imove, const R* r, const R* u, const R* rho,
             const R* m, R* lap_u, usize N, LL ll, R H, R CONF, int lap_form)
{
    const int VS = Dim<D>::VS;
    const R CLEARY = (D == 2) ? R(8.f) : R(10.f);
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        const int im_i = imove[i];
        if ((im_i != 1) && (im_i != -2)) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        const V<D, R> u_i = ld<D, R>(u, i);
        const R rho_i = rho[i];
        V<D, R> lu = ld<D, R>(lap_u, i);
        walk<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            const int im_j = imove[j];
            if ((im_j != 1) && (im_j != -1)) return;
            if (im_i == im_j) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const R rho_j = rho[j];
            const V<D, R> u_ij = sub<D, R>(ld<D, R>(u, j), u_i);
            const R udr = dot<D, R>(u_ij, r_ij);
            const R f_ij = kernelF<D, R>(q) * CONF * m[j];
            if (lap_form == 1) {
                const R r2 = (q * q + R(0.01f)) * H * H;
                const R sl = f_ij * CLEARY * udr / (r2 * rho_i * rho_j);
                for (int k = 0; k < D; k++) lu.c[k] += sl * r_ij.c[k];
            } else {
                const R sl = f_ij * R(2.f) / (rho_i * rho_j);
                for (int k = 0; k < D; k++) lu.c[k] += sl * u_ij.c[k];
            }
        });
        for (int k = 0; k < D; k++) lap_u[(size_t)i * VS + k] = lu.c[k];
    }
}

// GP/PreInteractions.cl:66-104: ghost velocity with the mirrored normal
// component, ghost pressure extended with the element's normal gradient
// dp/dn = rho (g - dudt + mu lap_u) . n
template <int D, typename R>
void gp_preinteractions(const unsigned* iset, const int* imove,
                        const usize* assoc, const R* r, const R* normal,
                        const R* dudt, const R* lap_u, const R* gp_p,
                        const R* gp_u, const R* rho, R* p, R* u,
                        const R* visc_dyn, usize N, const R* g)
{
    const int VS = Dim<D>::VS;
    for (usize i = 0; i < N; i++) {
        if (imove[i] != -1) continue;
        const usize iref = assoc[i];
        if (iref >= N) continue;
        const V<D, R> n = ld<D, R>(normal, iref);
        const V<D, R> gu = ld<D, R>(gp_u, i);
        const R dg = dot<D, R>(gu, n);
        const R dr = dot<D, R>(ld<D, R>(u, iref), n);
        for (int k = 0; k < D; k++) {
            const R u_n = dg * n.c[k];
            const R u_t = gu.c[k] - u_n;
            u[(size_t)i * VS + k] = R(2.f) * dr * n.c[k] - u_n + u_t;
        }
        const R mu = visc_dyn[iset[iref]];
        V<D, R> a;
        for (int k = 0; k < D; k++)
            a.c[k] = g[k] - dudt[(size_t)iref * VS + k] +
                     mu * lap_u[(size_t)iref * VS + k];
        if (D == 2) a.c[2] = R(0);
        const R dpdn = rho[iref] * dot<D, R>(a, n);
        const R dn = dot<D, R>(sub<D, R>(ld<D, R>(r, i), ld<D, R>(r, iref)), n);
        p[i] = gp_p[i] + R(2.f) * dpdn * dn;
    }
}

// GP/Interactions.cl:62-132: fluid against ghosts
template <int D, typename R>
void gp_interactions(const int* imove, const R* r, const R* u, const R* rho,
                     const R* m, const R* p, R* grad_p, R* div_u, usize N,
                     LL ll, R H, R CONF)
{
    const int VS = Dim<D>::VS;
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if (imove[i] != 1) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        const V<D, R> u_i = ld<D, R>(u, i);
        const R p_i = p[i];
        const R rho_i = rho[i];
        V<D, R> gp = ld<D, R>(grad_p, i);
        R dv = div_u[i];
        walk<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if (imove[j] != -1) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const R rho_j = rho[j];
            const R p_j = p[j];
            const R udr = dot<D, R>(sub<D, R>(ld<D, R>(u, j), u_i), r_ij);
            const R f_ij = kernelF<D, R>(q) * CONF * m[j];
            const R sg = (p_i + p_j) / (rho_i * rho_j) * f_ij;
            for (int k = 0; k < D; k++) gp.c[k] += sg * r_ij.c[k];
            dv += udr * f_ij * rho_i / rho_j;
        });
        for (int k = 0; k < D; k++) grad_p[(size_t)i * VS + k] = gp.c[k];
        div_u[i] = dv;
    }
}

// ---------------------------------------------------------------------------
// Conservative boundary integrals (resources/Scripts/cfd/Boundary/BIe/*.cl):
// wall elements imove == -3 (area in m, outward normal).
// ---------------------------------------------------------------------------

// BIe/Interactions.cl:entry: grad_w_bi = sum n_j W CONW area_j,
// div_u_bi -= u_j . (that term), over the elements in the support
template <int D, typename R>
void bie_interactions(const int* imove, const R* r, const R* normal,
                      const R* u, const R* m, R* grad_w_bi, R* div_u_bi,
                      usize N, LL ll, R H, R CONW)
{
    const int VS = Dim<D>::VS;
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if (imove[i] != 1) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        V<D, R> gw = ld<D, R>(grad_w_bi, i);
        R dv = div_u_bi[i];
        walk<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if (imove[j] != -3) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const V<D, R> n_j = ld<D, R>(normal, j);
            const V<D, R> u_j = ld<D, R>(u, j);
            const R w = kernelW<D, R>(q) * CONW * m[j];
            V<D, R> g;
            for (int k = 0; k < 3; k++) g.c[k] = (k < D) ? n_j.c[k] * w : R(0);
            for (int k = 0; k < D; k++) gw.c[k] += g.c[k];
            dv -= dot<D, R>(u_j, g);
        });
        for (int k = 0; k < D; k++) grad_w_bi[(size_t)i * VS + k] = gw.c[k];
        div_u_bi[i] = dv;
    }
}

// BIe/Interactions.cl:p_boundary: element pressure 2 sum p_j W CONW m_j/rho_j
template <int D, typename R>
void bie_p_boundary(const int* imove, const R* r, const R* m, const R* rho,
                    R* p, usize N, LL ll, R H, R CONW)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if (imove[i] != -3) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        R acc = R(0);
        walk<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if (imove[j] != 1) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            acc += R(2.f) * p[j] * kernelW<D, R>(q) * CONW * m[j] / rho[j];
        });
        p[i] = acc;
    }
}

// BIe/Rates.cl:entry (whole-vec operations: the w component takes part in
// the 3D sum and dot product)
template <int D, typename R>
void bie_rates(const int* imove, const R* rho, const R* p, const R* u,
               const R* grad_w_bi, const R* div_u_bi, R* grad_p, R* div_u,
               usize N)
{
    const int VS = Dim<D>::VS;
    for (usize i = 0; i < N; i++) {
        if (imove[i] != 1) continue;
        const R s = R(2.f) * p[i] / rho[i];
        R d = u[(size_t)i * VS] * grad_w_bi[(size_t)i * VS];
        for (int k = 1; k < VS; k++)
            d = d + u[(size_t)i * VS + k] * grad_w_bi[(size_t)i * VS + k];
        for (int k = 0; k < VS; k++)
            grad_p[(size_t)i * VS + k] += s * grad_w_bi[(size_t)i * VS + k];
        div_u[i] -= R(2.f) * rho[i] * (d + div_u_bi[i]);
    }
}

// cfd/Boundary/ElasticBounce.cl:77-148: the plain elastic bounce on ghost
// (-2) and boundary-integral (-3) elements.  Particle i (fluid) within
// R = DR_FACTOR dr of element j tangentially and on its fluid side, whose
// normal approach dist = dt u_n + dt^2/2 a_n (relative to the element)
// would bring it within MIN_BOUND_DIST dr, loses (1 + ELASTIC_FACTOR) times
// its normal relative velocity and acceleration; u and dudt are updated in
// place and the following elements see the corrected values.
template <int D, typename R>
void elastic_bounce(const int* imove, const R* r, const R* normal, R* u,
                    R* dudt, usize N, R dr, R dt, LL ll, R DR_FACTOR,
                    R MIN_BOUND_DIST, R ELASTIC_FACTOR)
{
    const int VS = Dim<D>::VS;
    const R Rr = DR_FACTOR * dr;
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if (imove[i] != 1) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        V<D, R> ui = ld<D, R>(u, i);
        V<D, R> ai = ld<D, R>(dudt, i);
        bool hit = false;
        walk<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if ((imove[j] != -2) && (imove[j] != -3)) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const V<D, R> n_j = ld<D, R>(normal, j);
            const R r0 = dot<D, R>(r_ij, n_j);
            if (r0 < R(0)) return;
            V<D, R> rt;
            for (int k = 0; k < 3; k++)
                rt.c[k] = (k < D) ? r_ij.c[k] - r0 * n_j.c[k] : R(0);
            if (dot<D, R>(rt, rt) >= Rr * Rr) return;
            const R u_n = dot<D, R>(sub<D, R>(ui, ld<D, R>(u, j)), n_j);
            const R a_n = dot<D, R>(sub<D, R>(ai, ld<D, R>(dudt, j)), n_j);
            const R dist = dt * u_n + R(0.5f) * dt * dt * a_n;
            if (dist < R(0)) return;
            if (r0 - dist <= MIN_BOUND_DIST * dr) {
                const R e1 = R(1.f) + ELASTIC_FACTOR;
                const R sa = e1 * a_n, su = e1 * u_n;
                for (int k = 0; k < D; k++) {
                    ai.c[k] = ai.c[k] - sa * n_j.c[k];
                    ui.c[k] = ui.c[k] - su * n_j.c[k];
                }
                hit = true;
            }
        });
        if (hit)
            for (int k = 0; k < D; k++) {
                dudt[(size_t)i * VS + k] = ai.c[k];
                u[(size_t)i * VS + k] = ui.c[k];
            }
    }
}

// BIe/ElasticBounce.cl:entry: reflect the velocity of fluid particles about
// to cross an element (within DR_FACTOR dr of its centre, tangentially)
template <int D, typename R>
void bie_elastic_bounce(const int* imove, const R* r_in, const R* normal,
                        const R* m, const R* u_in, R* dudt, usize N, R dt,
                        LL ll, R DR_FACTOR, R MIN_BOUND_DIST)
{
    const int VS = Dim<D>::VS;
    if (!(dt != R(0))) return;
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if (imove[i] != 1) continue;
        const V<D, R> r_i = ld<D, R>(r_in, i);
        const V<D, R> ui = ld<D, R>(u_in, i);
        V<D, R> a = ld<D, R>(dudt, i), uu;
        for (int k = 0; k < 3; k++)
            uu.c[k] = (k < D) ? ui.c[k] + R(0.5f) * dt * a.c[k] : R(0);
        walk<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if (imove[j] != -3) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r_in, j), r_i);
            const V<D, R> n_j = ld<D, R>(normal, j);
            const R rn = dot<D, R>(r_ij, n_j);
            if (rn < R(0)) return;
            const R dr = (D == 3) ? std::sqrt(m[j]) : m[j];
            const R Rr = DR_FACTOR * dr;
            V<D, R> rt;
            for (int k = 0; k < 3; k++)
                rt.c[k] = (k < D) ? r_ij.c[k] - rn * n_j.c[k] : R(0);
            if (dot<D, R>(rt, rt) >= Rr * Rr) return;
            const R drn = dt * dot<D, R>(uu, n_j);
            if (drn < R(0)) return;
            if (rn - drn <= MIN_BOUND_DIST * dr) {
                V<D, R> v, vr;
                for (int k = 0; k < 3; k++)
                    v.c[k] = (k < D) ? ui.c[k] + dt * a.c[k] : R(0);
                const R vn = dot<D, R>(v, n_j);
                for (int k = 0; k < D; k++) {
                    vr.c[k] = v.c[k] - R(2.f) * vn * n_j.c[k];
                    a.c[k] = (vr.c[k] - ui.c[k]) / dt;
                    uu.c[k] = ui.c[k] + R(0.5f) * dt * a.c[k];
                }
            }
        });
        for (int k = 0; k < D; k++) dudt[(size_t)i * VS + k] = a.c[k];
    }
}

// BIe/PST.cl:entry: push fluid particles that overlap an element back to
// the distance Ri = 0.5 (m/rho)^(1/DIMS); r is updated in place and the
// next elements see the moved particle
template <int D, typename R>
void bie_pst(const int* imove, R* r, const R* normal, const R* m,
             const R* rho, usize N, LL ll, R DIMS, R DR_FACTOR)
{
    const int VS = Dim<D>::VS;
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if (imove[i] != 1) continue;
        const R Ri = R(0.5f) * std::pow(m[i] / rho[i], R(1.f) / DIMS);
        walk<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if (imove[j] != -3) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), ld<D, R>(r, i));
            const V<D, R> n_j = ld<D, R>(normal, j);
            const R rn = dot<D, R>(r_ij, n_j);
            if (std::fabs(rn) > Ri) return;
            const R dr = (D == 3) ? std::sqrt(m[j]) : m[j];
            const R Rj = DR_FACTOR * dr;
            V<D, R> rt;
            for (int k = 0; k < 3; k++)
                rt.c[k] = (k < D) ? r_ij.c[k] - rn * n_j.c[k] : R(0);
            if (dot<D, R>(rt, rt) >= Rj * Rj) return;
            for (int k = 0; k < D; k++)
                r[(size_t)i * VS + k] += (rn - Ri) * n_j.c[k];
        });
    }
}

// ---------------------------------------------------------------------------
// MLS kernel-gradient correction and the full / simple delta-SPH Laplacian
// corrections (resources/Scripts/basic/MLS.cl, basic/deltaSPH.cl:57-190 and
// 261-313, cfd/deltaSPH.cl: EXCLUDED_PARTICLE = imove != 1).  A "matrix"
// is float4 (2x2, row major) in 2D and float16 (4x4, row major, the 3x3
// block used) in 3D, with the helpers of types/{2D,3D}.h.
// ---------------------------------------------------------------------------
template <int D> struct Mat {
    static constexpr int S = (D == 2) ? 4 : 16;
};

// MATRIX_MUL (3x3 block in 3D, the rest zero); dot products left to right
template <int D, typename R>
inline void mat_mul(const R* a, const R* b, R* c)
{
    if (D == 2) {
        c[0] = a[0] * b[0] + a[1] * b[2];
        c[1] = a[0] * b[1] + a[1] * b[3];
        c[2] = a[2] * b[0] + a[3] * b[2];
        c[3] = a[2] * b[1] + a[3] * b[3];
        return;
    }
    for (int k = 0; k < 16; k++) c[k] = R(0);
    for (int r_ = 0; r_ < 3; r_++)
        for (int c_ = 0; c_ < 3; c_++) {
            R s = a[4 * r_ + 0] * b[c_];
            s = s + a[4 * r_ + 1] * b[4 + c_];
            s = s + a[4 * r_ + 2] * b[8 + c_];
            c[4 * r_ + c_] = s;
        }
}

template <int D, typename R> inline void mat_transpose(const R* a, R* t)
{
    const int n = (D == 2) ? 2 : 4;
    for (int r_ = 0; r_ < n; r_++)
        for (int c_ = 0; c_ < n; c_++) t[n * r_ + c_] = a[n * c_ + r_];
}

// inv(): adjugate times 1/det, the identity when |1/det| > 1e16
template <int D, typename R> inline void mat_inv(const R* m, R* o)
{
    if (D == 2) {
        const R d = R(1) / (m[0] * m[3] - m[1] * m[2]);
        if (std::fabs(d) > R(1.e16f)) {
            o[0] = R(1), o[1] = R(0), o[2] = R(0), o[3] = R(1);
            return;
        }
        o[0] = m[3] * d;
        o[1] = -m[1] * d;
        o[2] = -m[2] * d;
        o[3] = m[0] * d;
        return;
    }
    const R det = m[0] * (m[5] * m[10] - m[6] * m[9]) +
                  m[1] * (m[6] * m[8] - m[4] * m[10]) +
                  m[2] * (m[4] * m[9] - m[5] * m[8]);
    const R d = R(1) / det;
    for (int k = 0; k < 16; k++) o[k] = (k % 5 == 0) ? R(1) : R(0);
    if (std::fabs(d) > R(1.e16f)) return;
    o[0] = (m[5] * m[10] - m[6] * m[9]) * d;
    o[1] = (m[2] * m[9] - m[1] * m[10]) * d;
    o[2] = (m[1] * m[6] - m[2] * m[5]) * d;
    o[4] = (m[6] * m[8] - m[4] * m[10]) * d;
    o[5] = (m[0] * m[10] - m[2] * m[8]) * d;
    o[6] = (m[2] * m[4] - m[0] * m[6]) * d;
    o[8] = (m[4] * m[9] - m[5] * m[8]) * d;
    o[9] = (m[1] * m[8] - m[0] * m[9]) * d;
    o[10] = (m[0] * m[5] - m[1] * m[4]) * d;
}

// MATRIX_INV: the pseudo-inverse inv(M^T M) M^T
template <int D, typename R> inline void mat_pinv(const R* m, R* o)
{
    R t[16], mtm[16], iv[16];
    mat_transpose<D, R>(m, t);
    mat_mul<D, R>(t, m, mtm);
    mat_inv<D, R>(mtm, iv);
    mat_mul<D, R>(iv, t, o);
}

// basic/MLS.cl:entry: sum over the mls_imove neighbours of
// outer(r_ij, F CONF m_j / rho_j r_ij)
template <int D, typename R>
void mls_entry(const int* imove, const R* r, const R* rho, const R* m,
               R* mls, usize N, unsigned mls_imove, LL ll, R H, R CONF)
{
    const int S = Mat<D>::S;
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if (imove[i] != (int)mls_imove) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        R acc[16];
        for (int k = 0; k < S; k++) acc[k] = R(0);
        walk<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if (i == j) return;
            if (imove[j] != (int)mls_imove) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const R f = kernelF<D, R>(q) * CONF * m[j] / rho[j];
            R v2[3];
            for (int k = 0; k < D; k++) v2[k] = f * r_ij.c[k];
            const int n = (D == 2) ? 2 : 4;
            for (int a = 0; a < D; a++)
                for (int b = 0; b < D; b++)
                    acc[n * a + b] += r_ij.c[a] * v2[b];
        });
        for (int k = 0; k < S; k++) mls[(size_t)i * S + k] = acc[k];
    }
}

// basic/MLS.cl:mls_inv
template <int D, typename R>
void mls_inv(const int* imove, R* mls, usize N, unsigned mls_imove)
{
    const int S = Mat<D>::S;
    for (usize i = 0; i < N; i++) {
        if (imove[i] != (int)mls_imove) continue;
        R o[16];
        mat_pinv<D, R>(mls + (size_t)i * S, o);
        for (int k = 0; k < S; k++) mls[(size_t)i * S + k] = o[k];
    }
}

// deltaSPH.cl:simple: lap_p_corr = refd g (the whole vec)
template <int D, typename R>
void lapp_simple(const unsigned* iset, const int* imove, R* lap_p_corr,
                 const R* refd, usize N, const R* g)
{
    const int VS = Dim<D>::VS;
    for (usize i = 0; i < N; i++) {
        if (imove[i] != 1) continue;
        for (int k = 0; k < VS; k++)
            lap_p_corr[(size_t)i * VS + k] = refd[iset[i]] * g[k];
    }
}

// deltaSPH.cl:full: lap_p_corr += sum (p_j - p_i) F CONF m_j/rho_j r_ij
template <int D, typename R>
void lapp_full(const int* imove, const R* r, const R* rho, const R* m,
               const R* p, R* lap_p_corr, usize N, LL ll, R H, R CONF)
{
    const int VS = Dim<D>::VS;
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if (imove[i] != 1) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        const R p_i = p[i];
        V<D, R> g = ld<D, R>(lap_p_corr, i);
        walk<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if ((i == j) || (imove[j] != 1)) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const R f = kernelF<D, R>(q) * CONF * m[j] / rho[j];
            const R s = (p[j] - p_i) * f;
            for (int k = 0; k < D; k++) g.c[k] += s * r_ij.c[k];
        });
        for (int k = 0; k < D; k++) lap_p_corr[(size_t)i * VS + k] = g.c[k];
    }
}

// deltaSPH.cl:full_mls: lap_p_corr = MATRIX_DOT(mls, lap_p_corr) (w = 0)
template <int D, typename R>
void lapp_full_mls(const int* imove, const R* mls, R* lap_p_corr, usize N)
{
    const int VS = Dim<D>::VS, S = Mat<D>::S;
    for (usize i = 0; i < N; i++) {
        if (imove[i] != 1) continue;
        const R* M = mls + (size_t)i * S;
        R* v = lap_p_corr + (size_t)i * VS;
        R o[4] = { R(0), R(0), R(0), R(0) };
        const int n = (D == 2) ? 2 : 4;
        for (int a = 0; a < D; a++) {
            R s = M[n * a + 0] * v[0];
            s = s + M[n * a + 1] * v[1];
            if (D == 3) s = s + M[n * a + 2] * v[2];
            o[a] = s;
        }
        for (int k = 0; k < VS; k++) v[k] = o[k];
    }
}

// deltaSPH.cl:lapp_corr: lap_p -= 0.5 (corr_j + corr_i) . r_ij F CONF m_j/rho_j
template <int D, typename R>
void lapp_corr(const int* imove, const R* r, const R* rho, const R* m,
               const R* lap_p_corr, R* lap_p, usize N, LL ll, R H, R CONF)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if (imove[i] != 1) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        const V<D, R> c_i = ld<D, R>(lap_p_corr, i);
        R lp = lap_p[i];
        walk<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if ((i == j) || (imove[j] != 1)) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            V<D, R> gij;
            const V<D, R> c_j = ld<D, R>(lap_p_corr, j);
            for (int k = 0; k < 3; k++)
                gij.c[k] = (k < D) ? c_j.c[k] + c_i.c[k] : R(0);
            const R f = kernelF<D, R>(q) * CONF * m[j] / rho[j];
            lp -= R(0.5f) * dot<D, R>(gij, r_ij) * f;
        });
        lap_p[i] = lp;
    }
}

// basic/neighs.cl:52-91: every candidate of the 3^d cells (self included,
// no distance or type filter); the walk stops as soon as the count reaches
// neighs_limit; buffer particles (imove <= -255) get 0.
template <int D>
void neighs(const int* imove, usize* n_neighs, usize limit, usize N, LL ll)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const usize i = (usize)ii;
        if (imove[i] <= -255) {
            n_neighs[i] = 0;
            continue;
        }
        usize n = 0;
        bool stop = false;
        walk<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize) {
            if (stop) return;
            n += 1;
            if (n >= limit) stop = true;
        });
        n_neighs[i] = n;
    }
}

// ---------------------------------------------------------------------------
// Plane portals (cfd/portal.xml, cfd/Boundary/Portal/*.cl).  Particles within
// SUPPORT h of the out plane are moved by in_r - out_r (and re-hashed into
// the cell of their new place); those "mirrored" particles then add the
// terms of the unmirrored particles around the in plane, walking the cells
// of their new place; a cell run ends at a mirrored particle whose icell
// changed (as in the reference) and cells outside the grid are skipped.
// ---------------------------------------------------------------------------
// Portal/Mirror.cl:35-52 (cell)
template <int D, typename R>
usize portal_cell(const R* r, const R* r_min, const usize* nc, R H)
{
    const R idist = R(1) / (R(2) * H);
    const usize cx = cvt_u32((r[0] - r_min[0]) * idist) + 3u;
    const usize cy = cvt_u32((r[1] - r_min[1]) * idist) + 3u;
    if (D == 3) {
        const usize cz = cvt_u32((r[2] - r_min[2]) * idist) + 3u;
        return cx - 1u + (cy - 1u) * nc[0] + (cz - 1u) * nc[0] * nc[1];
    }
    return cx - 1u + (cy - 1u) * nc[0];
}

// Portal/Mirror.cl:71-95 (mirror), on every stored component
template <int D, typename R>
void portal_mirror(R* r, int* imirrored, usize* icell, usize N,
                   const R* in_r, const R* out_r, const R* n, const R* r_min,
                   const usize* nc, R H)
{
    const int VS = Dim<D>::VS;
    for (usize i = 0; i < N; i++) {
        R rij[4] = { 0, 0, 0, 0 };
        R dist = R(0);
        for (int k = 0; k < VS; k++) rij[k] = r[(size_t)i * VS + k] - out_r[k];
        dist = rij[0] * n[0];
        for (int k = 1; k < VS; k++) dist = dist + rij[k] * n[k];
        if (std::fabs(dist) > R(2) * H) {
            imirrored[i] = 0;
            continue;
        }
        imirrored[i] = 1;
        for (int k = 0; k < VS; k++) r[(size_t)i * VS + k] = in_r[k] + rij[k];
        icell[i] = portal_cell<D, R>(r + (size_t)i * VS, r_min, nc, H);
    }
}

// Portal/Mirror.cl:108-124 (unmirror)
template <int D, typename R>
void portal_unmirror(R* r, const int* imirrored, usize N, const R* in_r,
                     const R* out_r)
{
    const int VS = Dim<D>::VS;
    for (usize i = 0; i < N; i++) {
        if (!imirrored[i]) continue;
        for (int k = 0; k < VS; k++) {
            const size_t a = (size_t)i * VS + k;
            r[a] = out_r[k] + (r[a] - in_r[k]);
        }
    }
}

// Portal/Mirror.cl:136-153 (teleport): past the out plane -> in plane
template <int D, typename R>
void portal_teleport(R* r, usize N, const R* in_r, const R* out_r,
                     const R* n)
{
    const int VS = Dim<D>::VS;
    for (usize i = 0; i < N; i++) {
        R rij[4] = { 0, 0, 0, 0 };
        for (int k = 0; k < VS; k++) rij[k] = r[(size_t)i * VS + k] - out_r[k];
        R dist = rij[0] * n[0];
        for (int k = 1; k < VS; k++) dist = dist + rij[k] * n[k];
        if (dist < R(0)) continue;
        for (int k = 0; k < VS; k++) r[(size_t)i * VS + k] = in_r[k] + rij[k];
    }
}

// Portal/Shepard.cl:58-100
template <int D, typename R>
void portal_shepard(const int* imove, const int* imirrored, const R* r,
                    const R* rho, const R* m, R* shep, usize N, LL ll, R H,
                    R CONW)
{
    for (usize i = 0; i < N; i++) {
        if ((imove[i] < -3) || (imove[i] > 1) || !imirrored[i]) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        R acc = shep[i];
        walk_cell<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if ((imove[j] != 1) || imirrored[j]) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            acc += kernelW<D, R>(q) * CONW * m[j] / rho[j];
        });
        shep[i] = acc;
    }
}

// Portal/Interactions.cl:67-140 (fluid and ghost candidates)
template <int D, typename R>
void portal_interactions(const int* imove, const int* imirrored, const R* r,
                         const R* u, const R* rho, const R* m, const R* p,
                         R* grad_p, R* lap_u, R* div_u, usize N, LL ll, R H,
                         R CONF, int lap)
{
    const R CLEARY = (D == 2) ? R(8) : R(10);
    for (usize i = 0; i < N; i++) {
        if (!imirrored[i] || (imove[i] != 1)) continue;
        const V<D, R> r_i = ld<D, R>(r, i), u_i = ld<D, R>(u, i);
        const R p_i = p[i], rho_i = rho[i];
        walk_cell<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if (imirrored[j] || ((imove[j] != 1) && (imove[j] != -1)))
                return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const R rho_j = rho[j], p_j = p[j];
            const V<D, R> du = sub<D, R>(ld<D, R>(u, j), u_i);
            const R udr = dot<D, R>(du, r_ij);
            const R f_ij = kernelF<D, R>(q) * CONF * m[j];
            addsv<D, R>(grad_p, i, (p_i + p_j) / (rho_i * rho_j) * f_ij, r_ij);
            if (lap == 1) {
                const R r2 = (q * q + R(0.01f)) * H * H;
                addsv<D, R>(lap_u, i,
                            f_ij * CLEARY * udr / (r2 * rho_i * rho_j), r_ij);
            } else {
                addsv<D, R>(lap_u, i, f_ij * R(2) / (rho_i * rho_j), du);
            }
            div_u[i] += udr * f_ij * rho_i / rho_j;
        });
    }
}

// Portal/BoundaryIntegrals.cl:55-128 (the normal is the `normal` argument)
template <int D, typename R>
void portal_bi(const int* imove, const int* imirrored, const R* r,
               const R* normal, const R* u, const R* rho, const R* m,
               const R* p, R* grad_p, R* div_u, usize N, LL ll, R H, R CONW)
{
    for (usize i = 0; i < N; i++) {
        if (!imirrored[i] || (imove[i] != 1)) continue;
        const V<D, R> r_i = ld<D, R>(r, i), u_i = ld<D, R>(u, i);
        const R p_i = p[i], rho_i = rho[i];
        walk_cell<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if (imirrored[j] || (imove[j] != -3)) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const V<D, R> n_j = ld<D, R>(normal, j);
            const V<D, R> du = sub<D, R>(ld<D, R>(u, j), u_i);
            const R w_ij = kernelW<D, R>(q) * CONW * m[j];
            addsv<D, R>(grad_p, i, (p_i + p[j]) / rho_i * w_ij, n_j);
            div_u[i] += rho_i * dot<D, R>(du, n_j) * w_ij;
        });
    }
}

// Portal/deltaSPH.cl:121-191 (lapp) and 50-119 (full)
template <int D, typename R>
void portal_lapp(const int* imove, const int* imirrored, const R* r,
                 const R* rho, const R* m, const R* p, R* lap_p, R* lap_p_corr,
                 usize N, LL ll, R H, R CONF)
{
    for (usize i = 0; i < N; i++) {
        if (!imirrored[i] || (imove[i] != 1)) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        const R p_i = p[i];
        walk_cell<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if (imirrored[j] || (imove[j] != 1)) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const R f_ij = kernelF<D, R>(q) * CONF * m[j] / rho[j];
            if (lap_p) lap_p[i] += (p[j] - p_i) * f_ij;
            if (lap_p_corr) addsv<D, R>(lap_p_corr, i, (p[j] - p_i) * f_ij, r_ij);
        });
    }
}

// Portal/deltaSPH.cl:193-246 (lapp_corr)
template <int D, typename R>
void portal_lapp_corr(const int* imove, const int* imirrored, const R* r,
                      const R* rho, const R* m, const R* lap_p_corr, R* lap_p,
                      usize N, LL ll, R H, R CONF)
{
    for (usize i = 0; i < N; i++) {
        if (!imirrored[i] || (imove[i] != 1)) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        const V<D, R> g_i = ld<D, R>(lap_p_corr, i);
        walk_cell<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if (imirrored[j] || (imove[j] != 1)) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            V<D, R> g;
            const V<D, R> g_j = ld<D, R>(lap_p_corr, j);
            for (int k = 0; k < 3; k++) g.c[k] = g_j.c[k] + g_i.c[k];
            const R f_ij = kernelF<D, R>(q) * CONF * m[j] / rho[j];
            lap_p[i] -= R(0.5f) * dot<D, R>(g, r_ij) * f_ij;
        });
    }
}

// Portal/MLS.cl:60-110 (entry): mls += outer(r_ij, f_ij r_ij)
template <int D, typename R>
void portal_mls(const int* imove, const int* imirrored, const R* r,
                const R* rho, const R* m, R* mls, usize N, unsigned mls_imove,
                LL ll, R H, R CONF)
{
    const int n = (D == 2) ? 2 : 4, S = (D == 2) ? 4 : 16;
    for (usize i = 0; i < N; i++) {
        if ((imove[i] != 1) || !imirrored[i]) continue;
        const V<D, R> r_i = ld<D, R>(r, i);
        R* M = mls + (size_t)i * S;
        walk_cell<D>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, [&](usize j) {
            if ((imove[j] != (int)mls_imove) || imirrored[j]) return;
            const V<D, R> r_ij = sub<D, R>(ld<D, R>(r, j), r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const R f_ij = kernelF<D, R>(q) * CONF * m[j] / rho[j];
            for (int a = 0; a < D; a++)
                for (int b = 0; b < D; b++)
                    M[a * n + b] += r_ij.c[a] * (f_ij * r_ij.c[b]);
        });
    }
}

// ---------------------------------------------------------------------------
// LES closure (csrc/les.hip; no counterpart in the reference): the SPH
// velocity gradient G_i = sum_j m_j/rho_j (u_j - u_i) (x) F_ij r_ij (matrix
// layout of MLS: row stride 2 in 2D, 4 in 3D) and the eddy-viscosity term
// dudt_i += sum_j F_ij m_j (rho_i nu_i + rho_j nu_j)/(rho_i rho_j) u_ij.
// ---------------------------------------------------------------------------
template <int D, typename R>
void les_gradu(const int* imove, const R* r, const R* u, const R* rho,
               const R* m, R* grad_u, usize N, LL ll, R H, R CONF)
{
    const int n = (D == 2) ? 2 : 4, S = (D == 2) ? 4 : 16;
    for (usize i = 0; i < N; i++) {
        R G[16] = { 0 };
        if (imove[i] == 1) {
            const V<D, R> r_i = ld<D, R>(r, i), u_i = ld<D, R>(u, i);
            walk_r<D, R>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, r,
                        [&](usize j, const V<D, R>& r_j) {
                if (j == i || imove[j] != 1) return;
                const V<D, R> r_ij = sub<D, R>(r_j, r_i);
                const R q = length<D, R>(r_ij) / H;
                if (q >= R(2)) return;
                const R w = kernelF<D, R>(q) * CONF * m[j] / rho[j];
                const V<D, R> du = sub<D, R>(ld<D, R>(u, j), u_i);
                for (int a = 0; a < D; a++)
                    for (int b = 0; b < D; b++)
                        G[a * n + b] += du.c[a] * (w * r_ij.c[b]);
            });
        }
        for (int k = 0; k < S; k++) grad_u[(size_t)i * S + k] = G[k];
    }
}

template <int D, typename R>
void les_interactions(const int* imove, const R* r, const R* u, const R* rho,
                      const R* m, const R* nu_t, R* dudt, usize N, LL ll,
                      R H, R CONF)
{
    const int VS = Dim<D>::VS;
    for (usize i = 0; i < N; i++) {
        if (imove[i] != 1) continue;
        const V<D, R> r_i = ld<D, R>(r, i), u_i = ld<D, R>(u, i);
        const R rho_i = rho[i], mu_i = rho_i * nu_t[i];
        R acc[3] = { 0, 0, 0 };
        walk_r<D, R>(ll.icell[i], N, ll.nc, ll.icell, ll.ihoc, r,
                        [&](usize j, const V<D, R>& r_j) {
            if (j == i || imove[j] != 1) return;
            const V<D, R> r_ij = sub<D, R>(r_j, r_i);
            const R q = length<D, R>(r_ij) / H;
            if (q >= R(2)) return;
            const R rho_j = rho[j];
            const R f = kernelF<D, R>(q) * CONF * m[j];
            const R sc = f * (mu_i + rho_j * nu_t[j]) / (rho_i * rho_j);
            const V<D, R> du = sub<D, R>(ld<D, R>(u, j), u_i);
            for (int k = 0; k < D; k++) acc[k] += sc * du.c[k];
        });
        for (int k = 0; k < D; k++) dudt[(size_t)i * VS + k] += acc[k];
    }
}

// ---------------------------------------------------------------------------
// Semi-implicit midpoint scheme (basic/time_scheme/midpoint.cl).  Vectors
// are updated on every stored component, as the float4 arithmetic of the
// reference.
// ---------------------------------------------------------------------------
// midpoint.cl:54-75 (predictor): the state at the start of the step
template <int D, typename R>
void midpoint_predictor(const R* r, const R* u, const R* dudt, const R* rho,
                        const R* drhodt, R* r_in, R* u_in, R* dudt_in,
                        R* rho_in, R* drhodt_in, usize N)
{
    const int VS = Dim<D>::VS;
    for (size_t a = 0; a < (size_t)N * VS; a++) {
        dudt_in[a] = dudt[a];
        u_in[a] = u[a];
        r_in[a] = r[a];
    }
    for (size_t i = 0; i < N; i++) {
        drhodt_in[i] = drhodt[i];
        rho_in[i] = rho[i];
    }
}

// midpoint.cl:95-112 (midpoint): half a step with the current rates
template <int D, typename R>
void midpoint_advance(const int* imove, const R* u_in, R* u, const R* dudt,
                      const R* rho_in, R* rho, const R* drhodt, usize N,
                      R dt)
{
    const int VS = Dim<D>::VS;
    for (size_t i = 0; i < N; i++) {
        if (imove[i] <= 0) continue;
        const R hdt = R(0.5f) * dt;
        for (int k = 0; k < VS; k++) {
            const size_t a = i * VS + k;
            u[a] = u_in[a] + hdt * dudt[a];
        }
        rho[i] = rho_in[i] + hdt * drhodt[i];
    }
}

// midpoint.cl:129-142 (midpoint_r)
template <int D, typename R>
void midpoint_r(const int* imove, const R* r_in, R* r, const R* u, usize N,
                R dt)
{
    const int VS = Dim<D>::VS;
    for (size_t i = 0; i < N; i++) {
        if (imove[i] <= 0) continue;
        const R hdt = R(0.5f) * dt;
        for (int k = 0; k < VS; k++)
            r[i * VS + k] = r_in[i * VS + k] + hdt * u[i * VS + k];
    }
}

// midpoint.cl:164-179 (relax)
template <int D, typename R>
void midpoint_relax(const int* imove, const R* dudt_in, R* dudt,
                    const R* drhodt_in, R* drhodt, usize N, R f)
{
    const int VS = Dim<D>::VS;
    for (size_t i = 0; i < N; i++) {
        if (imove[i] <= 0) continue;
        for (int k = 0; k < VS; k++) {
            const size_t a = i * VS + k;
            dudt[a] = f * dudt_in[a] + (R(1) - f) * dudt[a];
        }
        drhodt[i] = f * drhodt_in[i] + (R(1) - f) * drhodt[i];
    }
}

// midpoint.cl:203-226 (residuals): m (|u.(dudt - dudt_in)| +
// |p/rho^2 (drhodt - drhodt_in)|)
template <int D, typename R>
void midpoint_residuals(const int* imove, const R* m, const R* u,
                        const R* dudt_in, const R* dudt, const R* rho,
                        const R* p, const R* drhodt_in, const R* drhodt,
                        R* res, usize N)
{
    for (size_t i = 0; i < N; i++) {
        if (imove[i] <= 0) {
            res[i] = R(0);
            continue;
        }
        const R rho2 = rho[i] * rho[i];
        const V<D, R> dd = sub<D, R>(ld<D, R>(dudt, i), ld<D, R>(dudt_in, i));
        res[i] = m[i] * (std::fabs(dot<D, R>(ld<D, R>(u, i), dd)) +
                         std::fabs(p[i] / rho2 * (drhodt[i] - drhodt_in[i])));
    }
}

// midpoint.cl:252-270 (corrector): the whole step with the last rates
template <int D, typename R>
void midpoint_corrector(const int* imove, const R* r_in, R* r, const R* u_in,
                        R* u, const R* dudt, const R* rho_in, R* rho,
                        const R* drhodt, usize N, R dt)
{
    const int VS = Dim<D>::VS;
    for (size_t i = 0; i < N; i++) {
        if (imove[i] <= 0) continue;
        const R hdt2 = R(0.5f) * dt * dt;
        for (int k = 0; k < VS; k++) {
            const size_t a = i * VS + k;
            r[a] = r_in[a] + dt * u_in[a] + hdt2 * dudt[a];
            u[a] = u_in[a] + dt * dudt[a];
        }
        rho[i] = rho_in[i] + dt * drhodt[i];
    }
}

// ---------------------------------------------------------------------------
// Energy of the fluid (cfd/energy_kin.xml, cfd/energy.xml).  Non-fluid
// particles get zeros.
// ---------------------------------------------------------------------------
// cfd/Energy/EnergyKin.cl:40-56: E_kin = m u.u / 2
template <int D, typename R>
void energy_kin(R* ek, const int* imove, const R* u, const R* m, usize N)
{
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const size_t i = (size_t)ii;
        if (imove[i] != 1) {
            ek[i] = R(0);
            continue;
        }
        const V<D, R> u_i = ld<D, R>(u, i);
        ek[i] = R(0.5f) * m[i] * dot<D, R>(u_i, u_i);
    }
}

// cfd/Energy/Energy.cl:59-86 (power): dEk/dt = m u.du/dt,
// dEp/dt = -m g.u, dEc/dt = m p / rho^2 drho/dt
template <int D, typename R>
void energy_power(R* dek, R* dep, R* dec, const int* imove, const R* u,
                  const R* rho, const R* m, const R* p, const R* dudt,
                  const R* drhodt, usize N, const R* g)
{
    V<D, R> gv;
    for (int k = 0; k < 3; k++) gv.c[k] = k < D ? g[k] : R(0);
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const size_t i = (size_t)ii;
        if (imove[i] != 1) {
            dek[i] = dep[i] = dec[i] = R(0);
            continue;
        }
        const V<D, R> u_i = ld<D, R>(u, i);
        dep[i] = -m[i] * dot<D, R>(gv, u_i);
        dek[i] = m[i] * dot<D, R>(u_i, ld<D, R>(dudt, i));
        dec[i] = m[i] * p[i] / (rho[i] * rho[i]) * drhodt[i];
    }
}

// cfd/Energy/Energy.cl:113-142 (energy): E_k = m u.u / 2, E_p = -m g.r,
// E_c = m cs^2 (rho0 / rho + log(rho / rho0) - 1)
template <int D, typename R>
void energy_parts(R* ek, R* ep, R* ec, const unsigned* iset,
                  const int* imove, const R* r, const R* u, const R* rho,
                  const R* m, const R* refd, usize N, const R* g, R cs)
{
    V<D, R> gv;
    for (int k = 0; k < 3; k++) gv.c[k] = k < D ? g[k] : R(0);
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)N; ii++) {
        const size_t i = (size_t)ii;
        if (imove[i] != 1) {
            ek[i] = ep[i] = ec[i] = R(0);
            continue;
        }
        const V<D, R> u_i = ld<D, R>(u, i);
        ek[i] = R(0.5f) * m[i] * dot<D, R>(u_i, u_i);
        ep[i] = -m[i] * dot<D, R>(gv, ld<D, R>(r, i));
        const R rho0 = refd[iset[i]];
        ec[i] = m[i] * cs * cs *
                (rho0 / rho[i] + std::log(rho[i] / rho0) - R(1));
    }
}

}  // namespace orc
