"""Pins the oracle's periodic box (Presets/cfd/periodic.xml; no counterpart
in the reference, whose periodic domains are plane portals) against the
same fluid surrounded by its explicit periodic images r + (a, b, c) L,
a, b, c in {-1, 0, 1}: the ordinary walk over the box plus its 3^D - 1
images adds, for every particle of the box, exactly the pairs the
periodic walk finds through its wrapped cells and shifted positions.  In
float64 the two agree to 1e-12 (their summation orders differ) for the
cfd Shepard sum, the cfd interactions (both Laplacian forms), lap p, the
LES velocity gradient and the LES viscous term, and the sensors.  Images
placed without their shift, or the same box without periodicity, fail the
comparison; the wrapped grid, the cell hash and the wrap of the
positions are checked by their closed forms, and on the uniform lattice
the Shepard sum is the same for every particle and sum_i m_i grad_p_i
vanishes (antisymmetry over whole-box pairs)."""
import ctypes

import numpy as np
import pytest

from aquagpusph_amd import cases
from oracle import pipeline as op
from oracle.pipeline import _p

import orc

U_ = ctypes.c_uint
D_ = ctypes.c_double


def _set_per(L):
    L = list(L) + [0.0] * (3 - len(L))
    op.lib().orc_set_periodic(D_(L[0]), D_(L[1]), D_(L[2]))


def _link_per(dims, r, h, pmin, L):
    """The periodic link-list of positions r (float64): grid, hash, sort."""
    lib = op.lib()
    sfx = "f64_%dd" % dims
    vs = 2 if dims == 2 else 4
    r = np.ascontiguousarray(r, np.float64).reshape(-1, vs)
    N = r.shape[0]
    rmin = np.zeros(vs)
    rmax = np.zeros(vs)
    getattr(lib, "orc_minmax_vec_" + sfx)(_p(r), U_(N), _p(rmin), _p(rmax))
    pm = np.zeros(vs)
    pl = np.zeros(vs)
    pm[:dims] = pmin
    pl[:dims] = L
    nc = np.zeros(4, np.uint32)
    ro = rmin.copy()
    cl = float(np.float32(2.0) * np.float32(h))
    ok = getattr(lib, "orc_n_cells_per_" + sfx)(_p(rmin), _p(rmax), _p(pm),
                                                _p(pl), D_(cl), _p(nc),
                                                _p(ro))
    assert ok
    ic = np.zeros(N, np.uint32)
    getattr(lib, "orc_icell_per_" + sfx)(_p(ic), _p(r), U_(N), _p(ro),
                                         D_(2.0), D_(float(np.float32(h))),
                                         _p(nc), _p(pm), _p(pl))
    perm = np.zeros(N, np.uint32)
    inv = np.zeros(N, np.uint32)
    lib.orc_stable_sort(_p(ic), _p(perm), _p(inv), U_(N),
                        U_(lib.orc_radix_key_bits(ctypes.c_ulonglong(int(nc[3])))))
    hoc = np.zeros(int(nc[3]), np.uint32)
    lib.orc_ihoc(_p(ic), _p(hoc), U_(N), _p(nc))
    return dict(r=r[perm].copy(), perm=perm, inv=inv, icell=ic, ihoc=hoc,
                n_cells=nc, sfx=sfx, R=np.float64)


def _fields(dims, n, seed, sensors=0):
    """A jittered periodic lattice in the unit box with random fields."""
    rng = np.random.default_rng(seed)
    c = cases.periodic_box(dims, n, seed=seed)
    vs = 2 if dims == 2 else 4
    N = c["N"]
    r = c["r"].copy()
    r[:, :dims] += rng.uniform(-0.3, 0.3, (N, dims)) / n
    r[:, :dims] %= 1.0
    imove = np.ones(N, np.int32)
    if sensors:
        s = np.zeros((sensors, vs))
        s[:, :dims] = rng.uniform(0.0, 1.0, (sensors, dims))
        r = np.concatenate([r, s])
        imove = np.concatenate([imove, np.zeros(sensors, np.int32)])
    M = r.shape[0]
    u = np.zeros((M, vs))
    u[:, :dims] = rng.normal(size=(M, dims))
    rho = 1.0 + 0.01 * rng.normal(size=M)
    p = 1.0 + 0.1 * rng.normal(size=M)
    m = np.full(M, 1.0 / N) * (1.0 + 0.05 * rng.normal(size=M))
    nu = 1e-3 * (1.0 + rng.uniform(size=M))
    return dict(r=r, u=u, rho=rho, p=p, m=m, nu=nu, imove=imove,
                h=c["params"]["h"], N=M)


def _images(dims, F, skip_shift=False):
    """The box followed by its 3^dims - 1 images (r + offset, the fields
    copied), those within reach of the box; skip_shift places the images
    at the box positions (the slip of a walk that forgets the shift)."""
    offs = [o for o in np.ndindex(*([3] * dims)) if any(k != 1 for k in o)]
    parts = {k: [F[k]] for k in ("r", "u", "rho", "p", "m", "nu", "imove")}
    # only the images within the support (2h) of the box can interact with
    # it; the others are left out
    pad = 2.2 * F["h"]
    for o in offs:
        rr = F["r"].copy()
        sh = np.array(o, float) - 1.0
        near = np.all((F["r"][:, :dims] + sh > -pad) &
                      (F["r"][:, :dims] + sh < 1.0 + pad), axis=1)
        if not skip_shift:
            rr[:, :dims] += sh
        parts["r"].append(rr[near])
        for k in ("u", "rho", "p", "m", "nu", "imove"):
            parts[k].append(F[k][near])
    return {k: np.concatenate(v) for k, v in parts.items()}


def _run(dims, F, ll, lap_form):
    """cfd Shepard, interactions, lap p, LES gradient and viscous term and
    the sensors on the sorted set of `ll`; results in the input order."""
    perm = ll["perm"]
    N = ll["r"].shape[0]
    vs = 2 if dims == 2 else 4
    S = 4 if dims == 2 else 16
    g = {k: np.ascontiguousarray(F[k][perm]) for k in
         ("u", "rho", "p", "m", "nu", "imove")}
    r = ll["r"]
    H, CONW, CONF = orc.defs(F["h"], dims)
    args = (_p(ll["icell"]), _p(ll["ihoc"]), _p(ll["n_cells"]))
    out = {"shepard": np.zeros(N), "grad_p": np.zeros((N, vs)),
           "lap_u": np.zeros((N, vs)), "div_u": np.zeros(N),
           "lap_p": np.zeros(N), "grad_u": np.zeros((N, S)),
           "dudt": np.zeros((N, vs))}
    call = lambda name, *a: orc.call(ll, name, *a)
    call("shepard", _p(g["imove"]), _p(r), _p(g["rho"]), _p(g["m"]),
         _p(out["shepard"]), U_(N), *args, D_(H), D_(CONW), ctypes.c_int(1))
    call("interactions", _p(g["imove"]), _p(r), _p(g["u"]), _p(g["rho"]),
         _p(g["m"]), _p(g["p"]), _p(out["grad_p"]), _p(out["lap_u"]),
         _p(out["div_u"]), U_(N), *args, D_(H), D_(CONF),
         ctypes.c_int(lap_form))
    call("lapp", _p(g["imove"]), _p(r), _p(g["rho"]), _p(g["m"]),
         _p(g["p"]), _p(out["lap_p"]), U_(N), *args, D_(H), D_(CONF))
    call("les_gradu", _p(g["imove"]), _p(r), _p(g["u"]), _p(g["rho"]),
         _p(g["m"]), _p(out["grad_u"]), U_(N), *args, D_(H), D_(CONF))
    call("les_interactions", _p(g["imove"]), _p(r), _p(g["u"]),
         _p(g["rho"]), _p(g["m"]), _p(g["nu"]), _p(out["dudt"]), U_(N),
         *args, D_(H), D_(CONF))
    su, srho, sp = g["u"].copy(), g["rho"].copy(), g["p"].copy()
    call("sensors", _p(g["imove"]), _p(r), _p(g["m"]), _p(su), _p(srho),
         _p(sp), U_(N), *args, D_(H), D_(CONW))
    out["sens_u"], out["sens_rho"], out["sens_p"] = su, srho, sp
    inv = ll["inv"]
    return {k: v[inv] for k, v in out.items()}


def _periodic(dims, F, lap_form):
    ll = _link_per(dims, F["r"], F["h"], [0.0] * dims, [1.0] * dims)
    _set_per([1.0] * dims)
    try:
        return _run(dims, F, ll, lap_form)
    finally:
        _set_per([0.0, 0.0, 0.0])


def _plain(dims, F, lap_form, skip_shift=False):
    E = _images(dims, F, skip_shift)
    E["h"] = F["h"]
    ll = orc.link(dims, E["r"], F["h"])
    res = _run(dims, E, ll, lap_form)
    return {k: v[:F["N"]] for k, v in res.items()}


def _close(a, b, tol=1e-12):
    s = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= tol * s


@pytest.mark.parametrize("dims,n", [(2, 16), (3, 13)])
@pytest.mark.parametrize("lap_form", [1, 2])
def test_periodic_walk_equals_explicit_images(dims, n, lap_form):
    F = _fields(dims, n, seed=dims + lap_form, sensors=5)
    per = _periodic(dims, F, lap_form)
    img = _plain(dims, F, lap_form)
    for k in per:
        assert _close(per[k], img[k]), (k, np.abs(per[k] - img[k]).max())
    # the slips fail: neighbours taken from the wrapped cells without the
    # shift of their positions, and the box without periodicity
    wrong = _plain(dims, F, lap_form, skip_shift=True)
    assert not _close(wrong["grad_p"], img["grad_p"], 1e-6)
    bare = {}
    ll = orc.link(dims, F["r"], F["h"])
    bare = _run(dims, F, ll, lap_form)
    assert not _close(bare["shepard"], img["shepard"], 1e-6)


@pytest.mark.parametrize("dims,n", [(2, 16), (3, 13)])
def test_periodic_grid_hash_and_wrap(dims, n):
    """n = trunc(L / (2h)) cells per axis (no margins), the cell coordinate
    trunc((r - pmin) n / L) clamped below n, and the wrap brings any
    position moved by less than L back into [pmin, pmin + L)."""
    F = _fields(dims, n, seed=7)
    h = F["h"]
    ll = _link_per(dims, F["r"], h, [0.0] * dims, [1.0] * dims)
    nc = ll["n_cells"]
    ncell = int(1.0 / (2.0 * h))
    assert list(nc[:dims]) == [ncell] * dims
    x = ll["r"][:, :dims]
    q = np.minimum(np.floor(x * ncell).astype(np.int64), ncell - 1)
    want = q[:, 0] + q[:, 1] * ncell + \
        (q[:, 2] * ncell * ncell if dims == 3 else 0)
    np.testing.assert_array_equal(np.sort(ll["icell"]), np.sort(want))
    # the wrap
    vs = 2 if dims == 2 else 4
    rng = np.random.default_rng(1)
    r = np.zeros((500, vs))
    r[:, :dims] = rng.uniform(-0.9, 1.9, (500, dims))
    r0 = r.copy()
    pm = np.zeros(vs)
    pl = np.zeros(vs)
    pl[:dims] = 1.0
    getattr(op.lib(), "orc_periodic_wrap_f64_%dd" % dims)(
        _p(r), U_(500), _p(pm), _p(pl))
    assert np.all((r[:, :dims] >= 0.0) & (r[:, :dims] < 1.0))
    d = r[:, :dims] - r0[:, :dims]
    assert np.all(np.isin(np.round(d), [-1.0, 0.0, 1.0]))
    np.testing.assert_allclose(d, np.round(d), atol=1e-15)
    # a box under 3 cells is refused
    ro = np.zeros(vs)
    bad = getattr(op.lib(), "orc_n_cells_per_f64_%dd" % dims)(
        _p(pm), _p(pl), _p(pm), _p(pl), D_(0.4), _p(np.zeros(4, np.uint32)),
        _p(ro))
    assert bad == 0


@pytest.mark.parametrize("dims,n", [(2, 16), (3, 13)])
def test_periodic_lattice_is_uniform(dims, n):
    """On the unperturbed periodic lattice every particle sees the same
    neighbourhood: the Shepard sum is one value, and a uniform pressure
    gives no pressure gradient; with random pressures the momentum of the
    pressure term, sum_i m_i grad_p_i, vanishes (every pair acts on both
    particles with opposite signs)."""
    c = cases.periodic_box(dims, n)
    N = c["N"]
    vs = 2 if dims == 2 else 4
    rng = np.random.default_rng(3)
    F = dict(r=c["r"], u=np.zeros((N, vs)), rho=np.ones(N),
             p=np.ones(N), m=c["m"], nu=np.zeros(N),
             imove=np.ones(N, np.int32), h=c["params"]["h"], N=N)
    res = _periodic(dims, F, 1)
    sh = res["shepard"]
    assert np.ptp(sh) <= 1e-13 * sh.max()
    assert np.abs(res["grad_p"]).max() <= 1e-12
    F["p"] = rng.normal(size=N)
    res = _periodic(dims, F, 1)
    mom = (F["m"][:, None] * res["grad_p"][:, :dims]).sum(0)
    scale = np.abs(F["m"][:, None] * res["grad_p"][:, :dims]).sum()
    assert np.abs(mom).max() <= 1e-12 * scale
