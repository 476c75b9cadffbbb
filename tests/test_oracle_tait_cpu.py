"""Pins the oracle's Tait equation of state (basic/tait.xml; no counterpart
in the reference) by its closed form p = p0 + B ((rho/rho0)^gamma - 1),
B = rho0 cs^2 / gamma: the reference state gives p0 exactly, the slope at
rho0 is cs^2, the float binary powering of an integral gamma stays within a
few ulp of the exact power, gamma = 1 is the linear EOS, the inverse of the
wall interpolation returns the density the law started from, and the
particles basic/EOS.cl:35 skips keep their pressure.  Planted slips (B
without the 1/gamma, the exponent gamma - 1, the reference density of the
wrong set) are shown to fail on the CPU."""
import ctypes

import numpy as np
import pytest

from oracle import pipeline as op
from oracle.pipeline import _p

CS, P0 = 20.0, 1.5e2
REFD = np.array([998.0, 1025.0])


def _tait(rho, iset, imove, gamma, dtype=np.float64, p_init=-7.0):
    L = op.lib()
    sfx = "f32_3d" if dtype == np.float32 else "f64_3d"
    R = ctypes.c_float if dtype == np.float32 else ctypes.c_double
    rho = np.ascontiguousarray(rho, dtype)
    p = np.full(rho.shape, p_init, dtype)
    refd = REFD.astype(dtype)
    iset = np.ascontiguousarray(iset, np.uint32)
    imove = np.ascontiguousarray(imove, np.int32)
    getattr(L, "orc_tait_eos_" + sfx)(
        _p(iset), _p(imove), _p(rho), _p(p), _p(refd),
        ctypes.c_uint(rho.size), R(CS), R(P0), R(gamma))
    return p


def _closed(rho, iset, gamma, B_fac=1.0, expo=None):
    rho0 = REFD[iset]
    B = rho0 * CS ** 2 / gamma * B_fac
    return P0 + B * ((rho / rho0) ** (gamma if expo is None else expo) - 1.0)


@pytest.mark.parametrize("gamma", [7.0, 1.0, 3.0, 7.5])
def test_tait_closed_form(gamma):
    rng = np.random.default_rng(1)
    n = 4000
    iset = rng.integers(0, 2, n)
    rho = REFD[iset] * (1.0 + rng.uniform(-0.03, 0.03, n))
    imove = np.ones(n, np.int32)
    p = _tait(rho, iset, imove, gamma)
    np.testing.assert_allclose(p, _closed(rho, iset, gamma), rtol=1e-12,
                               atol=1e-9 * CS ** 2 * REFD.max())
    # reference state and slope
    p_ref = _tait(REFD[iset], iset, imove, gamma)
    np.testing.assert_array_equal(p_ref, np.full(n, P0))
    eps = 1e-6
    dp = (_tait(REFD[iset] * (1 + eps), iset, imove, gamma) -
          _tait(REFD[iset] * (1 - eps), iset, imove, gamma)) / \
        (2 * eps * REFD[iset])
    np.testing.assert_allclose(dp, CS ** 2, rtol=1e-6)
    # the slips fail
    for slip in (dict(B_fac=gamma), dict(expo=gamma - 1.0)):
        if gamma == 1.0 and "B_fac" in slip:
            continue
        assert not np.allclose(p, _closed(rho, iset, gamma, **slip),
                               rtol=1e-9)
    wrong_set = _closed(rho, 1 - iset, gamma)
    assert not np.allclose(p, wrong_set, rtol=1e-9)


def test_tait_float_power_and_linear_limit():
    """float32: the binary powering of (rho/rho0)^7 is within 8 ulp of the
    exact power (|p - p0| relative), and gamma = 1 is the linear EOS
    p0 + cs^2 (rho - rho0) within rounding of the two float sequences."""
    rng = np.random.default_rng(2)
    n = 20000
    iset = np.zeros(n, np.int64)
    rho = (REFD[0] * (1.0 + rng.uniform(-0.05, 0.05, n))).astype(np.float32)
    imove = np.ones(n, np.int32)
    p32 = _tait(rho, iset, imove, 7.0, np.float32).astype(np.float64)
    exact = _closed(rho.astype(np.float64), iset, 7.0)
    # B x^7 - B: the error is that of B x^7 (a few ulp) over |p - p0|
    x7 = (rho.astype(np.float64) / REFD[0]) ** 7
    B = REFD[0] * CS ** 2 / 7.0
    tol = 8 * 2.0 ** -24 * (B * x7 + B + P0)
    assert np.all(np.abs(p32 - exact) <= tol)
    p1 = _tait(rho, iset, imove, 1.0, np.float32).astype(np.float64)
    lin = P0 + CS ** 2 * (rho.astype(np.float64) - REFD[0])
    np.testing.assert_allclose(p1, lin, rtol=0,
                               atol=4 * 2.0 ** -24 * CS ** 2 * REFD[0])


def test_tait_skips_what_the_linear_eos_skips():
    """basic/EOS.cl:35: imove <= 0 except -1 keeps its pressure."""
    imove = np.array([1, 0, -1, -2, -3, 2], np.int32)
    iset = np.zeros(6, np.int64)
    rho = np.full(6, REFD[0] * 1.01)
    p = _tait(rho, iset, imove, 7.0, p_init=-7.0)
    done = (imove > 0) | (imove == -1)
    np.testing.assert_array_equal(p[~done], -7.0)
    assert np.all(p[done] > P0)


def test_tait_wall_inverse_returns_the_density():
    """cfd BI interpolation Shepard with the Tait inverse: an element whose
    interpolated pressure is gamma_s * EOS(rho) gets back rho (and the
    pressure divided by its Shepard factor)."""
    L = op.lib()
    rng = np.random.default_rng(3)
    n = 500
    iset = rng.integers(0, 2, n).astype(np.uint32)
    rho = REFD[iset] * (1.0 + rng.uniform(-0.02, 0.02, n))
    imove = np.full(n, -3, np.int32)
    imove[:50] = 1
    p_true = _closed(rho, iset, 7.0)
    shep = rng.uniform(0.4, 1.0, n)
    p = p_true * shep
    r_out = np.full(n, 5.0)
    getattr(L, "orc_bi_interp_shepard_tait_f64_3d")(
        _p(iset), _p(imove), _p(shep), _p(r_out), _p(p), _p(REFD),
        ctypes.c_uint(n), ctypes.c_double(CS), ctypes.c_double(P0),
        ctypes.c_double(7.0))
    el = imove == -3
    np.testing.assert_allclose(p[el], p_true[el], rtol=1e-13)
    np.testing.assert_allclose(r_out[el], rho[el], rtol=1e-13)
    np.testing.assert_array_equal(r_out[~el], 5.0)
