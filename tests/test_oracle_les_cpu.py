"""Pins the LES closure restatement (oracle/les.py, orc_les_*):

- the SPH velocity gradient reproduces a linear field's gradient inside a
  lattice (to the lattice quadrature error of the kernel moment) and is
  zero for a uniform field;
- the network inputs are invariants: unchanged by a rotation of the frame,
  (1, 0) for a pure strain and (0, 1) for a pure rotation;
- Smagorinsky's |S| is the shear rate of a simple shear;
- the eddy-viscosity term conserves momentum (antisymmetric pairs), never
  adds kinetic energy for nu_t >= 0, and for a uniform nu_t equals the
  viscous term of the (pinned) Morris interactions with mu = rho nu_t."""
import ctypes

import numpy as np
import pytest

from oracle import les as olq
from oracle import pipeline as op
from oracle.pipeline import _p

from aquagpusph_amd import cases


def _lattice_sim(dims, n, seed=0):
    rng = np.random.default_rng(seed)
    case = cases.random_box(dims, n ** dims, seed=seed)
    side = np.arange(n) * (1.0 / n)
    g = np.meshgrid(*([side] * dims), indexing="ij")
    vs = 2 if dims == 2 else 4
    r = np.zeros((n ** dims, vs))
    r[:, :dims] = np.stack([a.ravel() for a in g], axis=1)
    case["r"] = r
    case["dr"] = 1.0 / n
    N = n ** dims
    case["params"] = dict(h=2.0 / n, cs=10.0, courant=0.1,
                          g=[0.0] * dims, refd=[1000.0], visc_dyn=[1e-3],
                          delta=[0.0], p0=0.0)
    case["m"] = np.full(N, 1000.0 * (1.0 / n) ** dims)
    case["rho"] = np.full(N, 1000.0)
    o = op.OracleSim(case, dtype=np.float64, domain=False, bi=False,
                     delta=False, les={"Cs": 0.15})
    f = o.f
    for k in ("r", "u", "rho"):
        f[k + "_in"][...] = f[k]
    o.link_list("r_in")
    o.reorder()
    return o, rng


def _gradu(o):
    f = o.f
    o._k("les_gradu")(_p(f["imove"]), _p(f["r"]), _p(f["u"]), _p(f["rho"]),
                      _p(f["m"]), _p(f["grad_u"]), ctypes.c_uint(o.N),
                      *o._ll(), o._r(o.H), o._r(o.CONF))
    return olq.matrices(f["grad_u"], o.dims)


@pytest.mark.parametrize("dims", [2, 3])
def test_gradient_of_a_linear_field(dims):
    o, rng = _lattice_sim(dims, 24 if dims == 2 else 12)
    A = rng.normal(0, 1, (dims, dims))
    x = o.f["r"][:, :dims]
    o.f["u"][...] = 0
    o.f["u"][:, :dims] = x @ A.T
    G = _gradu(o)
    inner = np.all((x > 0.25) & (x < 0.7), axis=1)
    err = np.abs(G[inner] - A).max() / np.abs(A).max()
    assert err < 5e-3, err
    o.f["u"][:, :dims] = 1.5
    assert np.abs(_gradu(o)).max() == 0.0


def test_features_are_invariants():
    rng = np.random.default_rng(3)
    for dims in (2, 3):
        n = 2 if dims == 2 else 4
        G = rng.normal(0, 1, (50, dims, dims))
        Q, _ = np.linalg.qr(rng.normal(0, 1, (dims, dims)))
        GR = Q @ G @ Q.T

        def pack(M):
            a = np.zeros((len(M), n, n))
            a[:, :dims, :dims] = M
            return a.reshape(len(M), -1)
        x1 = olq.features(pack(G), dims, 2.0)
        x2 = olq.features(pack(GR), dims, 2.0)
        np.testing.assert_allclose(x1, x2, rtol=1e-9, atol=1e-12)
        strain = pack(np.array([np.diag([1.0] + [-1.0 / (dims - 1)] *
                                        (dims - 1))]))
        rot = np.zeros((1, dims, dims))
        rot[0, 0, 1], rot[0, 1, 0] = 1.0, -1.0
        assert olq.features(strain, dims, 2.0)[0, :2] == pytest.approx([1, 0])
        assert olq.features(pack(rot), dims, 2.0)[0, :2] == pytest.approx(
            [0, 1])
    # simple shear du/dy = gamma: |S| = gamma
    shear = np.zeros((1, 4))
    shear[0, 1] = 3.0
    assert olq.strain_rate(shear, 2)[0] == pytest.approx(3.0)
    nu = olq.eddy_viscosity(np.array([1]), shear, 2, 0.01, Cs=0.15)
    assert nu[0] == pytest.approx((0.15 * 0.01) ** 2 * 3.0)


@pytest.mark.parametrize("dims", [2, 3])
def test_eddy_term_conserves_and_dissipates(dims):
    o, rng = _lattice_sim(dims, 16 if dims == 2 else 8, seed=2)
    f, N = o.f, o.N
    f["r"][:, :dims] += rng.uniform(-0.1, 0.1, (N, dims)) / 16
    f["u"][:, :dims] = rng.normal(0, 1, (N, dims))
    f["rho"][...] = 1000.0 * (1 + rng.uniform(-0.01, 0.01, N))
    f["nu_t"][...] = rng.uniform(0, 1e-3, N)
    f["dudt"][...] = 0
    o._k("les_interactions")(_p(f["imove"]), _p(f["r"]), _p(f["u"]),
                             _p(f["rho"]), _p(f["m"]), _p(f["nu_t"]),
                             _p(f["dudt"]), ctypes.c_uint(N), *o._ll(),
                             o._r(o.H), o._r(o.CONF))
    a = f["dudt"][:, :dims]
    mom = (f["m"][:, None] * a).sum(axis=0)
    scale = (f["m"][:, None] * np.abs(a)).sum()
    assert np.abs(mom).max() <= 1e-12 * scale
    assert np.sum(f["m"][:, None] * f["u"][:, :dims] * a) < 0
    # uniform nu_t and rho: the Morris viscous term with mu = rho nu_t
    f["rho"][...] = 1000.0
    f["nu_t"][...] = 2e-4
    f["dudt"][...] = 0
    o._k("les_interactions")(_p(f["imove"]), _p(f["r"]), _p(f["u"]),
                             _p(f["rho"]), _p(f["m"]), _p(f["nu_t"]),
                             _p(f["dudt"]), ctypes.c_uint(N), *o._ll(),
                             o._r(o.H), o._r(o.CONF))
    lu = np.zeros_like(f["u"])
    gp = np.zeros_like(f["u"])
    dv = np.zeros(N)
    o._k("interactions")(_p(f["imove"]), _p(f["r"]), _p(f["u"]),
                         _p(f["rho"]), _p(f["m"]), _p(f["p"]), _p(gp),
                         _p(lu), _p(dv), ctypes.c_uint(N), *o._ll(),
                         o._r(o.H), o._r(o.CONF), ctypes.c_int(2))
    np.testing.assert_allclose(f["dudt"][:, :dims], 0.2 * lu[:, :dims],
                               rtol=1e-10, atol=1e-12 * np.abs(lu).max())
