"""Pins the MLS and delta-SPH correction restatement of the oracle
(resources/Scripts/basic/MLS.cl, basic/deltaSPH.cl simple/full/full_mls/
lapp_corr) with the identities they are built on: the MLS matrix inverts
the kernel moment matrix, the MLS-renormalised gradient of a linear field
is exact on every particle (free surface included), and the corrected
Laplacian of a linear field vanishes."""
import ctypes

import numpy as np
import pytest

from oracle import pipeline as op
from oracle.pipeline import _p

import orc


def _block(dims, n=(20, 14, 12), dr=0.01, seed=0):
    """A jittered block of fluid (free surface on every side)."""
    rng = np.random.default_rng(seed)
    axes = [(np.arange(k) + 0.5) * dr for k in n[:dims]]
    x = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, dims)
    x += rng.uniform(-0.15, 0.15, x.shape) * dr
    vs = 2 if dims == 2 else 4
    r = np.zeros((x.shape[0], vs))
    r[:, :dims] = x
    h = 2 * dr
    ll = orc.link(dims, r, h)
    N = x.shape[0]
    return dict(ll=ll, N=N, vs=vs, h=h, dr=dr, dims=dims,
                imove=np.ones(N, np.int32), rho=np.full(N, 1000.0),
                m=np.full(N, 1000.0 * dr ** dims))


def _defs(h, dims):
    return (op.definition(h), op.definition(1.0 / h ** dims),
            op.definition(1.0 / h ** (dims + 2)))


def _mls(s):
    ll, N = s["ll"], s["N"]
    H, _, CONF = _defs(s["h"], s["dims"])
    S = 4 if s["dims"] == 2 else 16
    M = np.zeros((N, S))
    orc.call(ll, "mls_entry", _p(s["imove"]), _p(ll["r"]), _p(s["rho"]),
             _p(s["m"]), _p(M), ctypes.c_uint(N), ctypes.c_uint(1),
             *orc.ll_args(ll), ctypes.c_double(H), ctypes.c_double(CONF))
    L = M.copy()
    orc.call(ll, "mls_inv", _p(s["imove"]), _p(L), ctypes.c_uint(N),
             ctypes.c_uint(1))
    return M, L


def _sq(A, dims):
    n = 2 if dims == 2 else 4
    return A.reshape(-1, n, n)[:, :dims, :dims]


@pytest.mark.parametrize("dims", [2, 3])
def test_mls_matrix_is_the_inverse_moment(dims):
    s = _block(dims)
    M, L = _mls(s)
    Mq, Lq = _sq(M, dims), _sq(L, dims)
    eye = np.eye(dims)
    np.testing.assert_allclose(Lq @ Mq, np.broadcast_to(eye, Mq.shape),
                               atol=1e-9)
    # unused rows/columns of the 3D float16 stay zero
    if dims == 3:
        assert np.all(L.reshape(-1, 4, 4)[:, 3, :] == 0)
        assert np.all(L.reshape(-1, 4, 4)[:, :, 3] == 0)
    # interior particles: the moment matrix is the identity (the kernel
    # gradient normalisation) up to the discretisation
    x = s["ll"]["r"][:, :dims]
    c = x.mean(0)
    inner = np.all(np.abs(x - c) < (np.array([20, 14, 12])[:dims] / 2 - 5) *
                   s["dr"], axis=1)
    assert inner.sum() > 4
    np.testing.assert_allclose(Mq[inner], np.broadcast_to(eye, Mq[inner].shape),
                               atol=0.08)


@pytest.mark.parametrize("dims", [2, 3])
def test_full_correction_gradient_and_laplacian_of_a_linear_field(dims):
    """lap_p_corr (full + full_mls) is grad(p) on every particle and the
    corrected Laplacian (lapp + lapp_corr) of a linear p vanishes, where the
    plain one does not at the free surface; the simple correction is
    rho0 g."""
    s = _block(dims)
    ll, N, vs = s["ll"], s["N"], s["vs"]
    H, _, CONF = _defs(s["h"], dims)
    M, L = _mls(s)
    a = np.array([3000.0, -1500.0, 700.0])[:dims]
    x = ll["r"][:, :dims]
    p = 2.0e4 + x @ a
    corr = np.zeros((N, vs))
    orc.call(ll, "lapp_full", _p(s["imove"]), _p(ll["r"]), _p(s["rho"]),
             _p(s["m"]), _p(p), _p(corr), ctypes.c_uint(N), *orc.ll_args(ll),
             ctypes.c_double(H), ctypes.c_double(CONF))
    orc.call(ll, "lapp_full_mls", _p(s["imove"]), _p(L), _p(corr),
             ctypes.c_uint(N))
    np.testing.assert_allclose(corr[:, :dims], np.broadcast_to(a, (N, dims)),
                               rtol=1e-9, atol=1e-9 * np.abs(a).max())
    if dims == 3:
        assert np.all(corr[:, 3] == 0)
    lp = np.zeros(N)
    orc.call(ll, "lapp", _p(s["imove"]), _p(ll["r"]), _p(s["rho"]),
             _p(s["m"]), _p(p), _p(lp), ctypes.c_uint(N), *orc.ll_args(ll),
             ctypes.c_double(H), ctypes.c_double(CONF))
    plain = np.abs(lp).max()
    assert plain > 1e3 * np.abs(a).max() * s["dr"] / s["h"] ** 2 * 1e-3
    orc.call(ll, "lapp_corr", _p(s["imove"]), _p(ll["r"]), _p(s["rho"]),
             _p(s["m"]), _p(corr), _p(lp), ctypes.c_uint(N),
             *orc.ll_args(ll), ctypes.c_double(H), ctypes.c_double(CONF))
    assert np.abs(lp).max() < 1e-8 * plain
    # simple: rho0 g
    g = np.zeros(vs)
    g[dims - 1] = -9.81
    sc = np.zeros((N, vs))
    iset = np.zeros(N, np.uint32)
    refd = np.array([998.0])
    orc.call(ll, "lapp_simple", _p(iset), _p(s["imove"]), _p(sc), _p(refd),
             ctypes.c_uint(N), _p(g))
    np.testing.assert_array_equal(sc, np.broadcast_to(998.0 * g, (N, vs)))
