"""Pins the float atan / acospi sequence of the boundary-integral kernel
terms (the Cephes method; KernelFunctions/Wendland2D.hcl and
Wendland3D.hcl:137-185 take these functions from the OpenCL runtime, whose
contract is the OpenCL C ULP bounds: atan 5 ulp, acospi 5 ulp).  The GPU
runs the same sequence (checked bit for bit in test_nr_arith_gpu.py).

Over every positive finite float the sequence stays within 2.81 ulp for
atan and, over [0, 1] (the only range the kernel terms reach), 2.45 ulp for
acospi, against libm in double; the test scans a strided sample of the
same ranges so that it runs in a fraction of a second."""
import ctypes

import numpy as np

from oracle import pipeline as op


def _ulp(opc, lo, hi, stride):
    L = op.lib()
    f = L.orc_clmath_ulp
    f.argtypes = [ctypes.c_int, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                  ctypes.c_void_p]
    out = np.zeros(2)
    f(opc, lo, hi, stride, out.ctypes.data)
    return out


def _eval(opc, x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros_like(x)
    op.lib().orc_clmath_f32(ctypes.c_int(opc), ctypes.c_void_p(x.ctypes.data),
                            ctypes.c_void_p(y.ctypes.data),
                            ctypes.c_uint(x.size))
    return y


def _bits(v):
    return int(np.float32(v).view(np.uint32))


def test_atan_within_opencl_bound():
    err, at = _ulp(0, 0, 0x7F7FFFFF, 251)
    assert err <= 2.81, (err, at)  # OpenCL C: 5 ulp
    # negative arguments by symmetry, exactly
    x = np.float32(np.random.default_rng(0).uniform(-50, 50, 4096))
    np.testing.assert_array_equal(_eval(0, -x), -_eval(0, x))


def test_acospi_within_opencl_bound():
    err, at = _ulp(1, 0, _bits(1.0), 53)
    assert err <= 2.45, (err, at)  # OpenCL C: 5 ulp


def test_special_values():
    y = _eval(0, [0.0, -0.0, np.inf, -np.inf, 1.0, 3.0e38, 1e-30])
    assert y[0] == 0 and not np.signbit(y[0])
    assert y[1] == 0 and np.signbit(y[1])
    assert y[2] == np.float32(np.pi / 2) and y[3] == -np.float32(np.pi / 2)
    assert abs(float(y[4]) - np.pi / 4) <= 6e-8
    assert y[5] == np.float32(np.pi / 2)
    assert y[6] == np.float32(1e-30)  # atan x = x below the float spacing
    a = _eval(1, [1.0, 0.0, 0.5, np.float32(np.sqrt(0.5))])
    assert a[0] == 0.0 and a[1] == 0.5
    assert abs(float(a[2]) - 1 / 3) <= 3e-8
    assert abs(float(a[3]) - 0.25) <= 3e-8


def test_wall_term_float_against_double():
    """kernelS_D (Wendland2D.hcl / Wendland3D.hcl:137-185) composed in float
    from this sequence, against the same float composition from correctly
    rounded atan / acospi and against the double form.  The float formula
    itself loses up to ~1e-5 to cancellation in 3D (differences of acospi
    values); the sequence adds at most 1e-7 to that, i.e. a couple of ulp
    of the term's O(1) scale."""
    L = op.lib()
    L.orc_kernel_sd.restype = ctypes.c_double
    L.orc_kernel_sd.argtypes = [ctypes.c_int] + [ctypes.c_double] * 4
    rng = np.random.default_rng(2)
    H = 0.02

    def cr(opc, x):
        x = np.asarray(x, np.float32).astype(np.float64)
        y = np.arctan(x) if opc == 0 else np.arccos(x) / np.pi
        return y.astype(np.float32)

    for dims in (2, 3):
        for _ in range(300):
            d = rng.uniform(1e-6, 2.0) * H
            t, b = rng.uniform(0, 2.0, 2) * H
            s = (rng.uniform(0.3, 1.0) * H / 2) ** (dims - 1)
            ref = L.orc_kernel_sd(dims, d, t, b, s)
            got = _kernel_sd_f32(dims, d, t, b, s, _eval)
            rnd = _kernel_sd_f32(dims, d, t, b, s, cr)
            assert abs(got - rnd) <= 1e-7, (dims, d, t, b, s, got, rnd)
            assert abs(got - ref) <= abs(rnd - ref) + 1e-7


def _kernel_sd_f32(dims, d, t, b, s, fn):
    """The float form of kernelS_D, composed from the float sequence in the
    reference's operation order (Wendland2D.hcl / Wendland3D.hcl)."""
    f = np.float32
    d, t, b, s = f(d), f(t), f(b), f(s)
    if dims == 2:
        wcon = f(0.5) * f(0.318309886)
        dr = f(0.5) * s
        a = fn(0, [(t + dr) / d, (t - dr) / d])
        return float(-wcon * (a[0] - a[1]))
    dr = f(0.5) * np.sqrt(s)
    t1, t2 = (t - dr) / d, (t + dr) / d
    b1, b2 = (b - dr) / d, (b + dr) / d

    def om(x, y):
        x2, y2 = x * x, y * y
        v = np.sqrt((f(1) + x2 + y2) / ((f(1) + x2) * (f(1) + y2)))
        return fn(1, [min(v, f(1))])[0]
    st, sb = np.sign(t1), np.sign(b1)
    return float(-f(0.25) * (om(t2, b2) - st * om(t1, b2) - sb * om(t2, b1) +
                            st * sb * om(t1, b1)))
