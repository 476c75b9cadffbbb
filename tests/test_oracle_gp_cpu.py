"""Pins the ghost-particle (GP) wall restatement of the oracle
(resources/Scripts/cfd/Boundary/GP/*.cl) with closed forms: exact
reflection, Shepard-renormalised interpolation of uniform fields, the
hydrostatic pressure extension, the velocity extensions of PreLapU /
PreInteractions, and the hydrostatic balance of the wall particles that
only the ghost terms provide."""
import ctypes

import numpy as np
import pytest

from oracle import pipeline as op
from oracle.pipeline import _p

import orc

G = 9.81


def _wall_setup(dims, nx=24, ny=12, dr=0.01, nz=16):
    """Fluid lattice above a bottom wall y = 0 (normal -y): elements at the
    wall, four ghost layers below, each associated with its element."""
    h = 2 * dr
    xs = (np.arange(nx) + 0.5) * dr
    ys = (np.arange(ny) + 0.5) * dr
    zs = (np.arange(nz) + 0.5) * dr
    if dims == 2:
        fl = np.stack(np.meshgrid(xs, ys, indexing="ij"), -1).reshape(-1, 2)
        el = np.stack(np.meshgrid(xs, [0.0], indexing="ij"), -1).reshape(-1, 2)
    else:  # y is the wall-normal axis in 3D too
        fl = np.stack(np.meshgrid(xs, ys, zs, indexing="ij"),
                      -1).reshape(-1, 3)
        el = np.stack(np.meshgrid(xs, [0.0], zs, indexing="ij"),
                      -1).reshape(-1, 3)
    nf, ne = fl.shape[0], el.shape[0]
    layers = [(k + 0.5) * dr for k in range(4)]
    gh = np.concatenate([el - d * np.eye(dims)[1] for d in layers])
    N = nf + ne + gh.shape[0]
    vs = 2 if dims == 2 else 4
    r = np.zeros((N, vs))
    r[:nf, :dims] = fl
    r[nf:nf + ne, :dims] = el
    r[nf + ne:, :dims] = gh
    imove = np.full(N, -1, np.int32)
    imove[:nf] = 1
    imove[nf:nf + ne] = -2
    assoc = np.full(N, N, np.uint32)
    assoc[nf + ne:] = np.tile(np.arange(nf, nf + ne), 4)
    normal = np.zeros((N, vs))
    normal[nf:, 1] = -1.0
    return dict(dims=dims, r=r, imove=imove, assoc=assoc, normal=normal,
                N=N, nf=nf, ne=ne, h=h, dr=dr, H=ny * dr, vs=vs)


def _linked(s):
    """Sort the set like the link-list and remap the associations."""
    ll = orc.link(s["dims"], s["r"], s["h"])
    perm, inv = ll["perm"], ll["inv"]
    N = s["N"]
    out = dict(ll=ll)
    for k in ("imove", "normal"):
        out[k] = np.ascontiguousarray(s[k][perm])
    a = s["assoc"][perm]
    out["assoc"] = np.where(a >= N, N, inv[np.minimum(a, N - 1)]).astype(
        np.uint32)
    out["r"] = ll["r"]
    return out


def _defs(h, dims):
    return (op.definition(h), op.definition(1.0 / h ** dims),
            op.definition(1.0 / h ** (dims + 2)))


@pytest.mark.parametrize("dims", [2, 3])
def test_gp_mirror_is_an_exact_reflection(dims):
    s = _wall_setup(dims)
    L = _linked(s)
    ll, N = L["ll"], s["N"]
    r = L["r"].copy()
    gic = np.zeros(N, np.uint32)
    H = op.definition(s["h"])
    orc.call(ll, "gp_mirror", _p(L["imove"]), _p(L["assoc"]),
             _p(L["normal"]), _p(r), _p(gic), ctypes.c_uint(N),
             _p(ll["rmin"]), _p(ll["n_cells"]), ctypes.c_double(H),
             ctypes.c_double(2.0))
    g = L["imove"] == -1
    want = L["r"][g].copy()
    want[:, 1] = -want[:, 1]  # the wall is y = 0
    np.testing.assert_allclose(r[g], want, atol=1e-15)
    np.testing.assert_array_equal(r[~g], L["r"][~g])
    # gp_icell is the link-list cell of the mirrored point
    ic = np.zeros(N, np.uint32)
    orc.call(ll, "icell", _p(ic), _p(r), ctypes.c_uint(N), _p(ll["rmin"]),
             ctypes.c_double(2.0), ctypes.c_double(float(np.float32(s["h"]))),
             _p(ll["n_cells"]))
    np.testing.assert_array_equal(gic[g], ic[g])


def _interp(s, L, rho, p, u, m):
    ll, N = L["ll"], s["N"]
    H, CONW, _ = _defs(s["h"], s["dims"])
    r = L["r"].copy()
    gic = np.zeros(N, np.uint32)
    orc.call(ll, "gp_mirror", _p(L["imove"]), _p(L["assoc"]),
             _p(L["normal"]), _p(r), _p(gic), ctypes.c_uint(N),
             _p(ll["rmin"]), _p(ll["n_cells"]), ctypes.c_double(H),
             ctypes.c_double(2.0))
    g_rho, g_p, shep = np.zeros(N), np.zeros(N), np.zeros(N)
    g_u = np.zeros((N, s["vs"]))
    orc.call(ll, "gp_interp", _p(L["imove"]), _p(r), _p(m), _p(rho), _p(p),
             _p(u), _p(g_rho), _p(g_p), _p(g_u), _p(shep), ctypes.c_uint(N),
             _p(gic), *orc.ll_args(ll), ctypes.c_double(H),
             ctypes.c_double(CONW))
    raw = shep.copy()
    orc.call(ll, "gp_renorm", _p(L["imove"]), _p(shep), _p(g_rho), _p(g_p),
             _p(g_u), ctypes.c_uint(N))
    return g_rho, g_p, g_u, raw


@pytest.mark.parametrize("dims", [2, 3])
def test_gp_interpolation_of_uniform_fields(dims):
    """Shepard-renormalised interpolation at the mirrored points returns a
    uniform field exactly; the raw Shepard sum is the kernel integral over
    the fluid side (between 0.5 and 1)."""
    s = _wall_setup(dims)
    L = _linked(s)
    N, vs = s["N"], s["vs"]
    rho = np.full(N, 1000.0)
    p = np.full(N, 123.0)
    u = np.zeros((N, vs))
    u[:, 0], u[:, 1] = 0.3, -0.2
    m = np.full(N, 1000.0 * s["dr"] ** dims)
    g_rho, g_p, g_u, raw = _interp(s, L, rho, p, u, m)
    g = L["imove"] == -1
    np.testing.assert_allclose(g_rho[g], 1000.0, rtol=1e-12)
    np.testing.assert_allclose(g_p[g], 123.0, rtol=1e-12)
    np.testing.assert_allclose(g_u[g, 0], 0.3, rtol=1e-12)
    np.testing.assert_allclose(g_u[g, 1], -0.2, rtol=1e-12)
    # away from the block's lateral edges the support is the fluid half
    # space (and more for the deeper layers)
    r = L["r"]
    inner = g.copy()
    for k in ([0] if dims == 2 else [0, 2]):
        c = r[L["imove"] == 1, k]
        inner &= (r[:, k] > c.min() + 1.9 * s["h"]) & \
            (r[:, k] < c.max() - 1.9 * s["h"])
    assert inner.sum() > 10
    assert np.all(raw[inner] > 0.45) and np.all(raw[inner] < 1.05)


@pytest.mark.parametrize("dims", [2, 3])
def test_gp_pressure_and_velocity_extension(dims):
    """PreInteractions: with the hydrostatic pressure interpolated at the
    mirrored point, the ghost gets the hydrostatic pressure of its own depth
    (dp/dn = rho g.n); the normal velocity is mirrored against the wall
    velocity, the tangential one kept.  PreLapU: Takeda extension
    u = u_n + 2 u_wall,t - u_t, or the frozen velocity."""
    s = _wall_setup(dims)
    L = _linked(s)
    ll, N, vs = L["ll"], s["N"], s["vs"]
    rho0 = 1000.0
    g = L["imove"] == -1
    d = -L["r"][:, 1]  # ghost depth below the wall
    gp_p = rho0 * G * (s["H"] - d)  # hydrostatic p at the mirrored point
    gp_u = np.zeros((N, vs))
    gp_u[:, 0], gp_u[:, 1] = 0.7, 0.25
    rho = np.full(N, rho0)
    u = np.zeros((N, vs))
    uw = 0.1  # tangential wall velocity
    u[L["imove"] == -2, 0] = uw
    dudt = np.zeros((N, vs))
    lap_u = np.zeros((N, vs))
    p = np.zeros(N)
    gv = np.zeros(vs)
    gv[1] = -G
    iset = np.zeros(N, np.uint32)
    visc = np.array([1e-3])
    u_pi = u.copy()
    orc.call(ll, "gp_preinteractions", _p(iset), _p(L["imove"]),
             _p(L["assoc"]), _p(L["r"]), _p(L["normal"]), _p(dudt),
             _p(lap_u), _p(gp_p), _p(gp_u), _p(rho), _p(p), _p(u_pi),
             _p(visc), ctypes.c_uint(N), _p(gv))
    np.testing.assert_allclose(p[g], rho0 * G * (s["H"] + d[g]), rtol=1e-12)
    np.testing.assert_allclose(u_pi[g, 0], 0.7, rtol=1e-12)
    np.testing.assert_allclose(u_pi[g, 1], -0.25, rtol=1e-12)
    # Takeda extension for the Laplacian (model 1) and frozen (model 2)
    for model, want in ((1, (2 * uw - 0.7, 0.25)), (2, (0.0, 0.0))):
        uu = u.copy()
        rr = rho.copy()
        g_rho = np.full(N, 990.0)
        orc.call(ll, "gp_prelapu", _p(L["imove"]), _p(L["assoc"]),
                 _p(L["normal"]), _p(g_rho), _p(gp_u), _p(rr), _p(uu),
                 ctypes.c_uint(N), ctypes.c_int(model))
        np.testing.assert_allclose(rr[g], 990.0)
        np.testing.assert_allclose(uu[g, 0], want[0], atol=1e-15)
        np.testing.assert_allclose(uu[g, 1], want[1], atol=1e-15)


@pytest.mark.parametrize("dims", [2, 3])
def test_gp_hydrostatic_balance_at_the_wall(dims):
    """Fluid at rest over the wall with the ghost fields of the GP chain:
    grad_p (fluid terms + ghost terms) approximates grad(p)/rho = -g for
    the particles next to the wall, to the accuracy of the method: the
    Shepard-renormalised interpolation at mirrored points next to the wall
    averages over the fluid side only and under-estimates the pressure
    there (13 % low on the first row at h = 2 dr, 7 % on the second, < 2 %
    beyond).  Without the ghost terms the first rows are off by tens of
    percent, and a wrong sign or factor in the extension is off by more."""
    s = _wall_setup(dims)
    L = _linked(s)
    ll, N, vs = L["ll"], s["N"], s["vs"]
    H, CONW, CONF = _defs(s["h"], dims)
    rho0 = 1000.0
    y = L["r"][:, 1]
    p = rho0 * G * (s["H"] - y)
    rho = np.full(N, rho0)
    m = np.full(N, rho0 * s["dr"] ** dims)
    u = np.zeros((N, vs))
    g_rho, g_p, g_u, _ = _interp(s, L, rho, p, u, m)
    gv = np.zeros(vs)
    gv[1] = -G
    iset = np.zeros(N, np.uint32)
    z = np.zeros((N, vs))
    pp = p.copy()
    orc.call(ll, "gp_preinteractions", _p(iset), _p(L["imove"]),
             _p(L["assoc"]), _p(L["r"]), _p(L["normal"]), _p(z), _p(z),
             _p(g_p), _p(g_u), _p(rho), _p(pp), _p(u.copy()),
             _p(np.array([0.0])), ctypes.c_uint(N), _p(gv))
    gp = np.zeros((N, vs))
    lu = np.zeros((N, vs))
    dv = np.zeros(N)
    orc.call(ll, "interactions", _p(L["imove"]), _p(L["r"]), _p(u), _p(rho),
             _p(m), _p(pp), _p(gp), _p(lu), _p(dv), ctypes.c_uint(N),
             *orc.ll_args(ll), ctypes.c_double(H), ctypes.c_double(CONF),
             ctypes.c_int(1))
    fluid_only = gp.copy()
    orc.call(ll, "gp_interactions", _p(L["imove"]), _p(L["r"]), _p(u),
             _p(rho), _p(m), _p(pp), _p(gp), _p(dv), ctypes.c_uint(N),
             *orc.ll_args(ll), ctypes.c_double(H), ctypes.c_double(CONF))
    x = L["r"][:, 0]
    sel = (L["imove"] == 1) & (y < 2 * s["h"]) & \
        (np.abs(x - x[L["imove"] == 1].mean()) < 4 * s["dr"])
    if dims == 3:
        zc = L["r"][:, 2]
        sel &= np.abs(zc - zc[L["imove"] == 1].mean()) < 3 * s["dr"]
    assert sel.sum() >= 4
    want = -rho0 * G  # d p / dy
    err = np.abs(gp[sel, 1] * rho0 - want) / abs(want)
    assert err.max() < 0.15, err
    y1 = y[sel]
    assert err[y1 > 0.6 * s["h"]].max() < 0.08
    assert err[y1 > 1.1 * s["h"]].max() < 0.02
    assert np.all(np.abs(gp[sel, 0]) * rho0 < 0.03 * rho0 * G)
    assert np.abs(fluid_only[sel, 1] * rho0 - want).max() > 0.2 * rho0 * G


@pytest.mark.parametrize("dims", [2, 3])
def test_gp_lapu_wall_drag(dims):
    """A uniform stream over the wall with frozen (zero) ghost velocities:
    the ghost terms of lap_u decelerate the first rows along the stream and
    vanish beyond the support."""
    s = _wall_setup(dims)
    L = _linked(s)
    ll, N, vs = L["ll"], s["N"], s["vs"]
    H, _, CONF = _defs(s["h"], dims)
    rho = np.full(N, 1000.0)
    m = np.full(N, 1000.0 * s["dr"] ** dims)
    u = np.zeros((N, vs))
    u[L["imove"] == 1, 0] = 0.5
    lu = np.zeros((N, vs))
    orc.call(ll, "gp_lapu", _p(L["imove"]), _p(L["r"]), _p(u), _p(rho),
             _p(m), _p(lu), ctypes.c_uint(N), *orc.ll_args(ll),
             ctypes.c_double(H), ctypes.c_double(CONF), ctypes.c_int(2))
    y = L["r"][:, 1]
    fl = L["imove"] == 1
    near = fl & (y < 0.6 * s["h"])
    far = fl & (y > 2.1 * s["h"])
    assert np.all(lu[near, 0] < 0)
    assert np.all(lu[far] == 0)
    # the elements see the stream through the fluid terms (Morris form)
    el = L["imove"] == -2
    assert np.all(lu[el, 0] > 0)
