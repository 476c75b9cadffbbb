"""Pins the oracle's plain elastic bounce (cfd/Boundary/ElasticBounce.cl:77-148)
by its closed form: a fluid particle about to come within MIN_BOUND_DIST dr
of a wall element loses (1 + e) times its normal velocity and acceleration
relative to the element (e = 0 stops it, e = 1 mirrors it), its tangential
motion is kept, and particles moving away, passing beside the element
(beyond DR_FACTOR dr), behind it, or too far to reach it this step are left
alone.  Each element of the walk sees the velocity the previous ones left,
so a particle over several elements of one wall is reflected once; the
slips this catches (the original velocity at every element, a dropped
relative velocity, a flipped normal) are planted in an oracle composition
on the CPU and must fail."""
import ctypes

import numpy as np
import pytest

from oracle.pipeline import _p

import orc

DR = 0.01
H = 2 * DR
DT = 1e-3


def _at(dims, x, z):
    p = [0.0] * dims
    p[0], p[dims - 1] = x, z
    return p


def _vel(dims, ux, uz):
    return _at(dims, ux, uz)


def _elements(dims, corner=False):
    """Floor elements z = 0 spaced dr (outward normal -e_z, the last axis),
    5 (2D) or 5 x 5 (3D) around the origin; with `corner`, a left wall
    x = -2.5 dr of 5 (5 x 5) elements too (outward normal -e_x)."""
    v = dims - 1
    xe = (np.arange(5) - 2.0) * DR
    pos, nrm = [], []
    if dims == 2:
        for x in xe:
            pos.append(_at(dims, x, 0.0))
            nrm.append(_vel(dims, 0.0, -1.0))
        if corner:
            for z in xe + 2.5 * DR:
                pos.append(_at(dims, -2.5 * DR, z))
                nrm.append(_vel(dims, -1.0, 0.0))
    else:
        for x in xe:
            for y in xe:
                pos.append([x, y, 0.0])
                nrm.append([0.0, 0.0, -1.0])
        if corner:
            for y in xe:
                for z in xe + 2.5 * DR:
                    pos.append([-2.5 * DR, y, z])
                    nrm.append([-1.0, 0.0, 0.0])
    assert all(len(q) == dims for q in pos) and v == dims - 1
    return np.array(pos), np.array(nrm)


def _floor(dims, parts, u, a, u_wall=None, a_wall=None, corner=False):
    """The fluid particles `parts` (velocities u, accelerations a) over the
    wall elements of _elements; returns the link-list and the permuted
    arrays."""
    vs = 2 if dims == 2 else 4
    pos, nrm = _elements(dims, corner)
    npart = len(parts)
    ne = pos.shape[0]
    r = np.zeros((npart + ne, vs))
    for k, q in enumerate(parts):
        r[k, :dims] = q
    r[npart:, :dims] = pos
    N = r.shape[0]
    ll = orc.link(dims, r, H)
    perm = ll["perm"]
    imove = np.where(np.arange(N) < npart, 1, -3).astype(np.int32)
    normal = np.zeros((N, vs))
    normal[npart:, :dims] = nrm
    uu = np.zeros((N, vs))
    aa = np.zeros((N, vs))
    uu[:npart, :dims] = u
    aa[:npart, :dims] = a
    if u_wall is not None:
        uu[npart:, :dims] = u_wall
    if a_wall is not None:
        aa[npart:, :dims] = a_wall
    return dict(ll=ll, imove=imove[perm].copy(), normal=normal[perm].copy(),
                u=uu[perm].copy(), a=aa[perm].copy(), N=N,
                k=[int(ll["inv"][q]) for q in range(npart)])


def _bounce(s, e=0.0, dr_factor=1.5, mbd=0.3):
    u, a = s["u"].copy(), s["a"].copy()
    orc.call(s["ll"], "elastic_bounce", _p(s["imove"]), _p(s["ll"]["r"]),
             _p(s["normal"]), _p(u), _p(a), ctypes.c_uint(s["N"]),
             ctypes.c_double(DR), ctypes.c_double(DT), *orc.ll_args(s["ll"]),
             ctypes.c_double(dr_factor), ctypes.c_double(mbd),
             ctypes.c_double(e))
    return u, a


@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("e", [0.0, 1.0, 0.5])
def test_bounce_removes_or_mirrors_the_normal_motion(dims, e):
    v = dims - 1
    parts = [_at(dims, 0.001, 0.1 * DR),   # over the wall, approaching
             _at(dims, 0.002, 0.2 * DR),   # moving away
             _at(dims, 0.0, 0.5 * DR),     # approaching, out of reach
             _at(dims, 0.0, -0.2 * DR)]    # behind the wall
    u = [_vel(dims, 0.3, -1.0), _vel(dims, 0.3, 1.0),
         _vel(dims, 0.3, -1.0), _vel(dims, 0.3, -1.0)]
    a = [_vel(dims, 0.1, -9.81)] * 4
    s = _floor(dims, parts, u, a)
    un, an = _bounce(s, e)
    k = s["k"]
    # particle 0: the normal component scaled by -e, tangential kept
    np.testing.assert_allclose(un[k[0], 0], 0.3, rtol=0, atol=0)
    np.testing.assert_allclose(un[k[0], v], e * 1.0, rtol=1e-15)
    np.testing.assert_allclose(an[k[0], 0], 0.1, rtol=0, atol=0)
    np.testing.assert_allclose(an[k[0], v], e * 9.81, rtol=1e-15)
    # moving away / out of reach this step / behind: untouched
    for q in (1, 2, 3):
        np.testing.assert_array_equal(un[k[q]], s["u"][k[q]])
        np.testing.assert_array_equal(an[k[q]], s["a"][k[q]])
    # the reach is dist = dt u_n + dt^2/2 a_n against r0 - 0.3 dr: a faster
    # particle 0.5 dr away reaches and bounces
    s2 = _floor(dims, [parts[2]], [_vel(dims, 0.0, -0.2 * DR / DT)],
                [_vel(dims, 0.0, 0.0)])
    u2, _ = _bounce(s2, e)
    np.testing.assert_allclose(u2[s2["k"][0], v], e * 0.2 * DR / DT,
                               rtol=1e-14)


@pytest.mark.parametrize("dims", [2, 3])
def test_bounce_patch_is_dr_factor_wide(dims):
    """Tangentially, the element acts within R = DR_FACTOR dr of its centre:
    a particle beyond the last element by more than R is not bounced, one
    within R is."""
    out = _at(dims, 2 * DR + 1.6 * DR, 0.1 * DR)
    inside = _at(dims, 2 * DR + 1.4 * DR, 0.1 * DR)
    u = [_vel(dims, 0.0, -1.0)] * 2
    a = [_vel(dims, 0.0, 0.0)] * 2
    s = _floor(dims, [out, inside], u, a)
    un, _ = _bounce(s)
    v = dims - 1
    assert un[s["k"][0], v] == -1.0
    assert un[s["k"][1], v] == 0.0
    # a smaller factor leaves the second one alone too
    un, _ = _bounce(s, dr_factor=1.0)
    assert un[s["k"][1], v] == -1.0


@pytest.mark.parametrize("dims", [2, 3])
def test_bounce_is_relative_to_the_wall(dims):
    """An element moving with the particle (same normal velocity and
    acceleration) does not bounce it; one moving towards it does, and
    the particle is left with the wall's normal velocity when e = 0."""
    v = dims - 1
    p = [_at(dims, 0.0, 0.1 * DR)]
    u = [_vel(dims, 0.2, -1.0)]
    a = [_vel(dims, 0.0, -2.0)]
    s = _floor(dims, p, u, a, u_wall=_vel(dims, 0.0, -1.0),
               a_wall=_vel(dims, 0.0, -2.0))
    un, an = _bounce(s)
    np.testing.assert_array_equal(un[s["k"][0]], s["u"][s["k"][0]])
    s = _floor(dims, p, u, a, u_wall=_vel(dims, 0.0, 0.5),
               a_wall=_vel(dims, 0.0, 0.0))
    un, an = _bounce(s)
    np.testing.assert_allclose(un[s["k"][0], v], 0.5, rtol=1e-15)
    np.testing.assert_allclose(an[s["k"][0], v], 0.0, atol=1e-15)
    assert un[s["k"][0], 0] == 0.2


def _python_bounce(dims, s, e, stale=False, relative=True, sign=1.0):
    """The closed form composed element by element in Python, elements in
    the oracle's walk order (the 3^d cells, sorted index ascending in each),
    with a slip switch: `stale` tests every element with the particle's
    original velocity (while still subtracting from the updated one),
    `relative=False` drops the wall velocity, `sign` flips the normal."""
    ll = s["ll"]
    r = ll["r"][:, :dims]
    N = s["N"]
    nc = ll["n_cells"]
    u = s["u"].copy()
    a = s["a"].copy()
    for i in range(N):
        if s["imove"][i] != 1:
            continue
        ui, ai = u[i, :dims].copy(), a[i, :dims].copy()
        u0, a0 = ui.copy(), ai.copy()
        c = int(ll["icell"][i])
        cx, cy = c % int(nc[0]), (c // int(nc[0])) % int(nc[1])
        cz = c // (int(nc[0]) * int(nc[1]))
        order = []
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in ((-1, 0, 1) if dims == 3 else (0,)):
                    cc = (cx + dx) + (cy + dy) * int(nc[0]) + \
                        (cz + dz) * int(nc[0]) * int(nc[1])
                    order += [j for j in range(N) if ll["icell"][j] == cc]
        for j in order:
            if s["imove"][j] not in (-2, -3):
                continue
            n = sign * s["normal"][j, :dims]
            rij = r[j] - r[i]
            r0 = rij @ n
            if r0 < 0:
                continue
            rt = rij - r0 * n
            if rt @ rt >= (1.5 * DR) ** 2:
                continue
            uc, ac = (u0, a0) if stale else (ui, ai)
            uw = u[j, :dims] if relative else 0.0
            aw = a[j, :dims] if relative else 0.0
            un = (uc - uw) @ n
            an = (ac - aw) @ n
            dist = DT * un + 0.5 * DT * DT * an
            if dist < 0:
                continue
            if r0 - dist <= 0.3 * DR:
                ai = ai - (1 + e) * an * n
                ui = ui - (1 + e) * un * n
        u[i, :dims], a[i, :dims] = ui, ai
    return u, a


@pytest.mark.parametrize("dims", [2, 3])
def test_bounce_sees_the_corrected_velocity(dims):
    """With e = 1 a particle over several elements of the floor is mirrored
    once: after the first element its normal velocity points away, so the
    others skip it; in a corner the floor and the wall each mirror their
    own component.  Testing every element with the original velocity
    mirrors the floor component once per element (back and forth), a
    dropped wall velocity or a flipped normal gives another answer: each
    slip fails the comparison with the oracle."""
    # one particle over the floor, one in the corner with the left wall
    p = [_at(dims, 0.5 * DR, 0.1 * DR), _at(dims, -2.4 * DR, 0.05 * DR)]
    u = [_vel(dims, 0.3, -1.0), _vel(dims, -0.7, -0.5)]
    a = [_vel(dims, 0.0, -9.81), _vel(dims, 0.0, 1.0)]
    s = _floor(dims, p, u, a, u_wall=_vel(dims, 0.0, 0.25),
               a_wall=_vel(dims, 0.0, 0.0), corner=True)
    uo, ao = _bounce(s, e=1.0)
    ur, ar = _python_bounce(dims, s, 1.0)
    np.testing.assert_allclose(uo, ur, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(ao, ar, rtol=1e-13, atol=1e-15)
    v = dims - 1
    k = s["k"][0]
    # mirrored about the moving wall: u_n' = 2 u_w - u_n
    np.testing.assert_allclose(uo[k, v], 2 * 0.25 + 1.0, rtol=1e-14)
    for slip in (dict(stale=True), dict(relative=False), dict(sign=-1.0)):
        us, as_ = _python_bounce(dims, s, 1.0, **slip)
        assert not (np.allclose(us, uo, rtol=1e-13, atol=1e-15) and
                    np.allclose(as_, ao, rtol=1e-13, atol=1e-15)), slip
