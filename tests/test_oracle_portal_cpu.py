"""Pins the plane-portal restatement of the oracle (cfd/portal.xml,
cfd/Boundary/Portal/*.cl).

A portal pair makes the particles within SUPPORT h of the out plane see the
particles next to the in plane as if both planes were one.  So, with the
pair x = +L/2 -> -L/2 and its reverse, the terms the portal kernels add
must be exactly what the ordinary kernels (pinned elsewhere) add when the
box is explicitly surrounded by its periodic images r +- L e_x.  That is
checked kernel by kernel, in float64, on a jittered lattice with a wall of
boundary elements: the portal terms are compared with the ordinary kernels
on (box + images) minus the ordinary kernels on the box alone.  A wrong
shift, sign, plane test, candidate filter or cell hash fails it.

On top: the teleport is a wrap into the box (closed form), the mirror and
unmirror round-trip, and on the eight-portal Taylor-Green lattice (the
reference's examples/2D/taylor_green) the Shepard sum is the same for every
particle and the pressure gradient of p = cos kx + cos ky is -C (sin kx,
sin ky) for one C (translation invariance of the periodic lattice, corners
included through the diagonal portals)."""
import ctypes

import numpy as np
import pytest

from oracle import pipeline as op
from oracle.pipeline import _p

from aquagpusph_amd import cases

C = ctypes.c_uint


def _box(dims, seed=5, n=(16, 8, 4), dr=1.0, hfac=1.3):
    """Jittered fluid lattice in [-L/2, L/2) x [0, ..) with a row of boundary
    elements (imove -3, normal -y / -z) under it; random u, rho, p, and a
    random lap_p_corr field."""
    rng = np.random.default_rng(seed)
    vs = 2 if dims == 2 else 4
    nx = n[0]
    L = nx * dr
    axes = [-0.5 * L + (np.arange(nx) + 0.5) * dr]
    axes += [(np.arange(n[k]) + 0.5) * dr for k in range(1, dims)]
    g = np.meshgrid(*axes, indexing="ij")
    fl = np.stack([a.ravel() for a in g], axis=1)
    fl += rng.uniform(-0.2, 0.2, fl.shape) * dr
    fl[:, 0] = np.clip(fl[:, 0], -0.5 * L + 1e-3, 0.5 * L - 1e-3)
    we = [axes[0]] + [axes[k] for k in range(1, dims - 1)] + [[0.0]]
    g = np.meshgrid(*we, indexing="ij")
    wl = np.stack([a.ravel() for a in g], axis=1)
    nf, nw = fl.shape[0], wl.shape[0]
    N = nf + nw
    r = np.zeros((N, vs))
    r[:nf, :dims] = fl
    r[nf:, :dims] = wl
    normal = np.zeros((N, vs))
    normal[nf:, dims - 1] = -1.0
    u = np.zeros((N, vs))
    u[:, :dims] = rng.normal(0, 1, (N, dims))
    gcorr = np.zeros((N, vs))
    gcorr[:, :dims] = rng.normal(0, 1, (N, dims))
    rho = 1000.0 * (1 + rng.uniform(-0.01, 0.01, N))
    m = np.full(N, 1000.0 * dr ** dims)
    m[nf:] = dr ** (dims - 1)
    imove = np.ones(N, np.int32)
    imove[nf:] = -3
    p = rng.normal(0, 100.0, N)
    h = hfac * dr
    params = dict(h=h, cs=10.0, courant=0.1, g=[0.0] * dims, refd=[1000.0],
                  visc_dyn=[1e-3], delta=[0.1], p0=0.0,
                  domain_min=[-L] * dims, domain_max=[L] * dims)
    return dict(dims=dims, N=N, r=r, normal=normal, tangent=np.zeros((N, vs)),
                u=u, dudt=np.zeros((N, vs)), rho=rho, drhodt=np.zeros(N), m=m,
                imove=imove, iset=np.zeros(N, np.uint32), params=params,
                p=p, gcorr=gcorr, L=L)


def _images(case):
    """The box with its two images along x appended (originals first)."""
    L, N = case["L"], case["N"]
    out = dict(case)
    for k in ("r", "normal", "tangent", "u", "dudt", "gcorr"):
        out[k] = np.concatenate([case[k]] * 3)
    for k in ("rho", "drhodt", "m", "imove", "iset", "p"):
        out[k] = np.concatenate([case[k]] * 3)
    out["r"][N:2 * N, 0] += L
    out["r"][2 * N:, 0] -= L
    out["N"] = 3 * N
    return out


def _sorted_sim(case, **kw):
    """An OracleSim over `case` with its link-list built and its arrays in
    cell order; p and lap_p_corr carried along, accumulators zeroed."""
    o = op.OracleSim(case, dtype=np.float64, **kw)
    f = o.f
    for k in ("r", "u", "rho"):
        f[k + "_in"][...] = f[k]
    o.link_list("r_in")
    o.reorder()
    ids = f["id"].astype(np.int64)
    f["p"][...] = case["p"][ids]
    f["lap_p_corr"] = np.ascontiguousarray(case["gcorr"][ids], np.float64)
    _zero(o)
    return o, ids


def _zero(o):
    for k in ("shepard", "grad_p", "lap_u", "div_u", "lap_p"):
        o.f[k][...] = 0
    if "mls" in o.f:
        o.f["mls"][...] = 0


def _standard(o, stage):
    """The ordinary kernels the portal kernels of `stage` mirror."""
    f, N = o.f, C(o.N)
    k, R, ll = o._k, o._r, o._ll()
    H, CONW, CONF = R(o.H), R(o.CONW), R(o.CONF)
    if stage == "mls":
        k("mls_entry")(_p(f["imove"]), _p(f["r"]), _p(f["rho"]), _p(f["m"]),
                       _p(f["mls"]), N, C(1), *ll, H, CONF)
    elif stage == "forces":
        k("shepard")(_p(f["imove"]), _p(f["r"]), _p(f["rho"]), _p(f["m"]),
                     _p(f["shepard"]), N, *ll, H, CONW, ctypes.c_int(1))
        k("interactions")(_p(f["imove"]), _p(f["r"]), _p(f["u"]),
                          _p(f["rho"]), _p(f["m"]), _p(f["p"]),
                          _p(f["grad_p"]), _p(f["lap_u"]), _p(f["div_u"]), N,
                          *ll, H, CONF, ctypes.c_int(o.opts["lap_form"]))
        k("bi_interactions")(_p(f["imove"]), _p(f["r"]), _p(f["normal"]),
                             _p(f["u"]), _p(f["rho"]), _p(f["m"]),
                             _p(f["p"]), _p(f["grad_p"]), _p(f["div_u"]), N,
                             *ll, H, CONW)
        k("lapp")(_p(f["imove"]), _p(f["r"]), _p(f["rho"]), _p(f["m"]),
                  _p(f["p"]), _p(f["lap_p"]), N, *ll, H, CONF)
        f["lap_p_corr"][...] = 0
        k("lapp_full")(_p(f["imove"]), _p(f["r"]), _p(f["rho"]), _p(f["m"]),
                       _p(f["p"]), _p(f["lap_p_corr"]), N, *ll, H, CONF)
    else:
        k("lapp_corr")(_p(f["imove"]), _p(f["r"]), _p(f["rho"]), _p(f["m"]),
                       _p(f["lap_p_corr"]), _p(f["lap_p"]), N, *ll, H, CONF)


OUT = {"mls": ("mls",),
       "forces": ("shepard", "grad_p", "lap_u", "div_u", "lap_p",
                  "lap_p_corr"),
       "lapp_corr": ("lap_p",)}


def _by_id(o, ids, key, n):
    """Field `key` of the first n original particles, in original order."""
    a = o.f[key]
    out = np.zeros((n,) + a.shape[1:], a.dtype)
    sel = ids < n
    out[ids[sel]] = a[sel]
    return out


@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("stage,lap_form", [("forces", 1), ("forces", 2),
                                            ("mls", 1), ("lapp_corr", 1)])
def test_portal_terms_equal_periodic_images(dims, stage, lap_form):
    case = _box(dims)
    a = 0.5 * case["L"]
    z = [0.0] * (dims - 1)
    portals = [("right_", [a] + z, [-a] + z, [1.0] + z),
               ("left_", [-a] + z, [a] + z, [-1.0] + z)]
    kw = dict(bi=True, delta=True, delta_corr="full", lap_form=lap_form)
    n = case["N"]
    # portal terms alone: zero accumulators, then the portal stage
    o, ids = _sorted_sim(case, portals=portals, **kw)
    if stage == "lapp_corr":
        # the stage reads lap_p_corr: keep the random field
        pass
    elif stage == "forces":
        o.f["lap_p_corr"][...] = 0
    r0 = o.f["r"].copy()
    ic0 = o.f["icell"].copy()
    o._portal_stage(stage, o._ll())
    # mirror + unmirror leave r (to rounding) and icell (exactly) in place
    np.testing.assert_allclose(o.f["r"], r0, rtol=0, atol=1e-12 * a)
    np.testing.assert_array_equal(o.f["icell"], ic0)
    got = {k: _by_id(o, ids, k, n) for k in OUT[stage]}
    # ordinary kernels with and without the images
    e, eids = _sorted_sim(_images(case), **kw)
    _standard(e, stage)
    b, bids = _sorted_sim(case, **kw)
    _standard(b, stage)
    for k in OUT[stage]:
        want = _by_id(e, eids, k, n) - _by_id(b, bids, k, n)
        scale = np.abs(_by_id(e, eids, k, n)).max()
        assert np.abs(want).max() > 1e-3 * scale, (k, "no portal terms")
        err = np.abs(got[k] - want).max() / scale
        assert err <= 1e-12, (k, err)


def test_teleport_wraps_into_the_box():
    """The eight Taylor-Green portals applied in order send a point just
    outside [-L/2, L/2)^2 to its periodic image inside (closed form)."""
    L = 2.0 * np.pi
    a = 0.5 * L
    rng = np.random.default_rng(1)
    pts = rng.uniform(-a - 0.05, a + 0.05, (4000, 2))
    pts = np.concatenate([pts, [[a + 0.01, a + 0.02], [-a - 0.01, a + 0.03],
                                [a + 0.02, -a - 0.01], [-a - 0.03, -a - 0.02]]])
    r = np.ascontiguousarray(pts, np.float64)
    tel = getattr(op.lib(), "orc_portal_teleport_f64_2d")
    for _, out_r, in_r, n in cases.taylor_green_portals(L):
        # (arrays held by name: _p keeps only their address)
        pi, po, pn = (np.asarray(v, np.float64) for v in (in_r, out_r, n))
        tel(_p(r), C(len(r)), _p(pi), _p(po), _p(pn))
    want = np.mod(pts + a, L) - a
    # a point exactly on an out plane goes to the in plane
    np.testing.assert_allclose(r, want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("dims", [2, 3])
def test_mirror_selects_the_support_strip(dims):
    """imirrored = (|(r - out_r).n| <= 2H), mirrored points are moved by
    in_r - out_r and re-hashed with the link-list cell formula."""
    case = _box(dims, seed=9)
    a = 0.5 * case["L"]
    z = [0.0] * (dims - 1)
    o, _ = _sorted_sim(case)
    f = o.f
    vs = o.vs
    out_r = np.zeros(vs)
    in_r = np.zeros(vs)
    n = np.zeros(vs)
    out_r[0], in_r[0], n[0] = a, -a, 1.0
    r0 = f["r"].copy()
    im = np.zeros(o.N, np.int32)
    icell = f["icell"].copy()
    o._k("portal_mirror")(_p(f["r"]), _p(im), _p(icell), C(o.N), _p(in_r),
                          _p(out_r), _p(n), _p(o.r_min), _p(o.n_cells),
                          o._r(o.H))
    dist = (r0[:, 0] - a)
    sel = np.abs(dist) <= 2.0 * o.H
    np.testing.assert_array_equal(im.astype(bool), sel)
    np.testing.assert_allclose(f["r"][sel, 0], r0[sel, 0] - 2 * a, atol=1e-12)
    np.testing.assert_array_equal(f["r"][~sel], r0[~sel])
    # cell hash of the moved points (the link-list formula, r_min-clamped)
    want = np.zeros(o.N, np.uint32)
    o._k("icell")(_p(want), _p(f["r"]), C(o.N), _p(o.r_min),
                  o._r(o.support), o._r(o.h), _p(o.n_cells))
    np.testing.assert_array_equal(icell[sel], want[sel])
    np.testing.assert_array_equal(icell[~sel], f["icell"][~sel])
    assert sel.any() and (~sel).any()
    del z


def test_taylor_green_lattice_is_translation_invariant():
    """Eight portals close the Taylor-Green box: on its uniform lattice
    (uniform rho) the Shepard sum is the same for every particle and the
    pressure gradient of p = cos kx + cos ky is -C (sin kx, sin ky) with
    one C, corners included."""
    # a box side that float32 holds exactly: the portal planes are float
    # variables, so their shift is then exactly the lattice period
    case = cases.taylor_green(n=24, hfac=2.0,
                              L=float(np.float32(2.0 * np.pi)))
    N = case["N"]
    case["rho"] = np.full(N, 1.0)
    L = case["params"]["L"]
    k = 2.0 * np.pi / L
    x, y = case["r"][:, 0], case["r"][:, 1]
    case["p"] = np.cos(k * x) + np.cos(k * y)
    case["gcorr"] = np.zeros((N, 2))
    case["L"] = L
    portals = cases.taylor_green_portals(L)
    o, ids = _sorted_sim(case, portals=portals, delta=False, lap_form=2)
    _standard_shepard_interactions(o)
    o._portal_stage("forces", o._ll())
    sh = _by_id(o, ids, "shepard", N)
    assert np.ptp(sh) <= 1e-12 * sh.mean(), np.ptp(sh) / sh.mean()
    gp = _by_id(o, ids, "grad_p", N)
    s = np.stack([np.sin(k * x), np.sin(k * y)], axis=1)
    Cfit = -np.sum(gp * s) / np.sum(s * s)
    assert Cfit > 0
    np.testing.assert_allclose(gp, -Cfit * s, rtol=0,
                               atol=1e-10 * np.abs(gp).max())
    # without the diagonal portals the corner particles lose terms
    o2, ids2 = _sorted_sim(case, portals=portals[:4], delta=False,
                           lap_form=2)
    _standard_shepard_interactions(o2)
    o2._portal_stage("forces", o2._ll())
    sh2 = _by_id(o2, ids2, "shepard", N)
    assert np.ptp(sh2) > 1e-3 * sh2.mean()


def _standard_shepard_interactions(o):
    f, N = o.f, C(o.N)
    k, R, ll = o._k, o._r, o._ll()
    k("shepard")(_p(f["imove"]), _p(f["r"]), _p(f["rho"]), _p(f["m"]),
                 _p(f["shepard"]), N, *ll, R(o.H), R(o.CONW), ctypes.c_int(1))
    k("interactions")(_p(f["imove"]), _p(f["r"]), _p(f["u"]), _p(f["rho"]),
                      _p(f["m"]), _p(f["p"]), _p(f["grad_p"]), _p(f["lap_u"]),
                      _p(f["div_u"]), N, *ll, R(o.H), R(o.CONF),
                      ctypes.c_int(2))
