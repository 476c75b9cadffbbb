"""Pins the midpoint restatement of the oracle (basic/time_scheme/
midpoint.cl and the midpoint.xml inner loop) with closed forms: constant
acceleration is integrated exactly; the half step, the relaxation and the
residual follow their formulas; and the inner iteration is a contraction,
so on a dam break at the working dt its residual falls geometrically and the
step converges to the fixed point whatever the iteration count beyond it."""
import ctypes

import numpy as np
import pytest

from oracle import pipeline as op
from oracle.pipeline import _p

from aquagpusph_amd import cases


def _k(name, dims):
    return getattr(op.lib(), "orc_%s_f64_%dd" % (name, dims))


@pytest.mark.parametrize("dims", [2, 3])
def test_midpoint_kernels_closed_forms(dims):
    rng = np.random.default_rng(1)
    n = 50
    vs = 2 if dims == 2 else 4
    imove = np.ones(n, np.int32)
    imove[::5] = -3
    mv = imove > 0
    r_in, u_in, a = (rng.normal(0, 1, (n, vs)) for _ in range(3))
    rho_in = 1000 + rng.normal(0, 1, n)
    drho = rng.normal(0, 1, n)
    dt = 1e-3
    # half step
    u = np.zeros((n, vs))
    rho = np.zeros(n)
    _k("midpoint_advance", dims)(_p(imove), _p(u_in), _p(u), _p(a),
                                 _p(rho_in), _p(rho), _p(drho),
                                 ctypes.c_uint(n), ctypes.c_double(dt))
    np.testing.assert_allclose(u[mv], u_in[mv] + 0.5 * dt * a[mv], rtol=1e-15)
    assert np.all(u[~mv] == 0) and np.all(rho[~mv] == 0)
    np.testing.assert_allclose(rho[mv], rho_in[mv] + 0.5 * dt * drho[mv])
    # constant acceleration: the corrector is exact
    r = np.zeros((n, vs))
    u = np.zeros((n, vs))
    _k("midpoint_corrector", dims)(_p(imove), _p(r_in), _p(r), _p(u_in),
                                   _p(u), _p(a), _p(rho_in), _p(rho),
                                   _p(drho), ctypes.c_uint(n),
                                   ctypes.c_double(dt))
    np.testing.assert_allclose(r[mv], r_in[mv] + dt * u_in[mv] +
                               0.5 * dt * dt * a[mv], rtol=1e-14)
    np.testing.assert_allclose(u[mv], u_in[mv] + dt * a[mv], rtol=1e-14)
    # relaxation: convex combination of the old and new rates
    a_new = rng.normal(0, 1, (n, vs))
    d_new = rng.normal(0, 1, n)
    for f in (0.0, 0.3, 1.0):
        aa, dd = a_new.copy(), d_new.copy()
        _k("midpoint_relax", dims)(_p(imove), _p(a), _p(aa), _p(drho),
                                   _p(dd), ctypes.c_uint(n),
                                   ctypes.c_double(f))
        np.testing.assert_allclose(aa[mv], f * a[mv] + (1 - f) * a_new[mv],
                                   rtol=1e-14, atol=1e-15)
        np.testing.assert_array_equal(aa[~mv], a_new[~mv])
    # residual
    m = rng.uniform(1, 2, n)
    p = rng.normal(0, 100, n)
    res = np.full(n, -1.0)
    _k("midpoint_residuals", dims)(_p(imove), _p(m), _p(u_in), _p(a),
                                   _p(a_new), _p(rho_in), _p(p), _p(drho),
                                   _p(d_new), _p(res), ctypes.c_uint(n))
    want = m * (np.abs(((a_new - a)[:, :dims] * u_in[:, :dims]).sum(1)) +
                np.abs(p / rho_in ** 2 * (d_new - drho)))
    np.testing.assert_allclose(res[mv], want[mv], rtol=1e-12)
    assert np.all(res[~mv] == 0)


def test_inner_iteration_is_a_contraction():
    """FP64 dam break without delta-SPH (whose lap p is reset outside the
    inner loop when deltaSPH.xml follows midpoint.xml, so that its sum grows
    with every iteration, as in the reference): the inner residual falls by
    a roughly constant factor per iteration, and more iterations change the
    step's result only at the level of the last residual."""
    c = cases.dam_break(2, 0.3 / 12)

    def run(iters):
        o = op.OracleSim(c, dtype=np.float64, delta=False,
                         midpoint=dict(iter_max=iters, relax=0.0,
                                       residual_max=0.0))
        for _ in range(6):
            o.step()
        return o

    o = run(8)
    res = np.array(o.residuals[-8:])
    assert np.all(res[1:] < res[:-1])
    assert res[-1] < 1e-8 * res[0]
    a, b = run(8), run(12)
    scale = np.abs(b.f["u"]).max()
    assert np.abs(a.f["u"] - b.f["u"]).max() < 1e-6 * scale
