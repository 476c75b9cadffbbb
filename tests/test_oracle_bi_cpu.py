"""Pins the oracle's boundary-integral wall kernels, the cfd rates, the
delta-SPH rate, the binormal and the sensors by what the mathematics fixes:
the divergence theorem behind the boundary-integral operators, exact
reproduction of linear fields, the Laplacian of manufactured quadratic
fields, the momentum and continuity equations, hydrostatic rest, and frame
orientation.  Each check is paired with the plausible slip it catches (a
flipped normal, a missing area or density factor, a wrong Taylor sign, an
unapplied Shepard division), and those slips are planted in the inputs or
in the oracle's composition on the CPU to show that the check fails on
them.

The fluid operators are already pinned (test_oracle_cpu.py); here they are
combined with the wall terms.  Sign conventions follow the reference:
wall normals point out of the fluid (cfd/Boundary/BI/Interactions.cl:112),
grad_p holds grad(p)/rho, lap_u holds lap(u)/rho for the fluid and lap(u)
at the wall elements, div_u holds rho div(u) (cfd/Rates.cl:40-50).
"""
import ctypes

import numpy as np
import pytest

from oracle import pipeline as op
from oracle.pipeline import _p

import orc

D_ = ctypes.c_double
U_ = ctypes.c_uint
RHO0 = 1000.0


class Slab:
    """Fluid lattice z = (k + 1/2) dr above a flat wall of boundary-integral
    elements at z = 0 (the last axis is z), element spacing dr, area
    dr^(dims-1), outward normal -e_z, tangent e_x, binormal n x t."""

    def __init__(self, dims, n_lat, n_z, dr=0.01, hfac=2.0, wall=True):
        self.dims, self.dr = dims, dr
        self.h = hfac * dr
        vs = 2 if dims == 2 else 4
        ax = (np.arange(n_lat) + 0.5) * dr
        az = (np.arange(n_z) + 0.5) * dr
        grids = np.meshgrid(*([ax] * (dims - 1) + [az]), indexing="ij")
        fl = np.stack([g.ravel() for g in grids], 1)
        if wall:
            eg = np.meshgrid(*([ax] * (dims - 1)), indexing="ij")
            el = np.stack([g.ravel() for g in eg] + [np.zeros(eg[0].size)], 1)
        else:
            el = np.zeros((0, dims))
        nf, ne = fl.shape[0], el.shape[0]
        N = nf + ne
        r = np.zeros((N, vs))
        r[:nf, :dims] = fl
        r[nf:, :dims] = el
        self.normal = np.zeros((N, vs))
        self.tangent = np.zeros((N, vs))
        self.binormal = np.zeros((N, vs))
        self.normal[nf:, dims - 1] = -1.0
        self.tangent[nf:, 0] = 1.0
        if dims == 3:
            self.binormal[nf:, :3] = np.cross(self.normal[nf:, :3],
                                              self.tangent[nf:, :3])
        self.imove = np.ones(N, np.int32)
        self.imove[nf:] = -3
        self.m = np.full(N, RHO0 * dr ** dims)
        self.m[nf:] = dr ** (dims - 1)
        self.ll = orc.link(dims, r, self.h)
        perm = self.ll["perm"]
        for k in ("normal", "tangent", "binormal", "imove", "m"):
            setattr(self, k, getattr(self, k)[perm].copy())
        self.r = self.ll["r"]
        self.x = self.r[:, :dims]
        self.N, self.vs = N, vs
        self.fluid = self.imove == 1
        self.elem = self.imove == -3
        self.H, self.CONW, self.CONF = orc.defs(self.h, dims)
        L = n_lat * dr
        lat = np.all((self.x[:, :dims - 1] > 3 * self.h) &
                     (self.x[:, :dims - 1] < L - 3 * self.h), axis=1)
        z = self.x[:, dims - 1]
        # fluid particles whose support reaches the wall, and interior ones
        self.near = self.fluid & lat & (z < 2 * self.h)
        self.deep = self.fluid & lat & (z > 2 * self.h) & \
            (z < n_z * dr - 2 * self.h)
        self.welem = self.elem & lat

    def vec(self, f=None):
        a = np.zeros((self.N, self.vs))
        if f is not None:
            a[:, :self.dims] = f
        return a

    def call(self, name, *args):
        return orc.call(self.ll, name, *args)

    def lla(self):
        return orc.ll_args(self.ll)

    # ---- the oracle kernels on this slab ---------------------------------
    def gamma(self, normal=None, area=None):
        """Boundary-integral Shepard factor (cfd/Boundary/BI/Shepard.cl)."""
        n = self.normal if normal is None else normal
        m = self.m if area is None else area
        s = np.zeros(self.N)
        self.call("bi_shepard_compute", _p(self.imove), _p(self.r), _p(n),
                  _p(self.tangent), _p(self.binormal), _p(m), _p(s),
                  U_(self.N), *self.lla(), D_(self.H), D_(self.CONW),
                  D_(self.dims))
        return s

    def fluid_shepard(self, imove=None):
        """Discrete kernel mass of the fluid, sum_j W_ij m_j / rho_j
        (basic/Shepard.cl with the cfd exclusion)."""
        im = self.imove if imove is None else imove
        s = np.zeros(self.N)
        rho = np.full(self.N, RHO0)
        self.call("shepard", _p(im), _p(self.r), _p(rho), _p(self.m), _p(s),
                  U_(self.N), *self.lla(), D_(self.H), D_(self.CONW),
                  ctypes.c_int(1))
        return s

    def interactions(self, u, rho, p, lap_form=1):
        gp, lu, du = self.vec(), self.vec(), np.zeros(self.N)
        self.call("interactions", _p(self.imove), _p(self.r), _p(u), _p(rho),
                  _p(self.m), _p(p), _p(gp), _p(lu), _p(du), U_(self.N),
                  *self.lla(), D_(self.H), D_(self.CONF),
                  ctypes.c_int(lap_form))
        return gp, lu, du

    def bi_interactions(self, u, rho, p, gp, du, normal=None, area=None):
        n = self.normal if normal is None else normal
        m = self.m if area is None else area
        self.call("bi_interactions", _p(self.imove), _p(self.r), _p(n),
                  _p(u), _p(rho), _p(m), _p(p), _p(gp), _p(du), U_(self.N),
                  *self.lla(), D_(self.H), D_(self.CONW))

    def apply(self, shep, gp, lu, du):
        self.call("bi_shepard_apply", _p(self.imove), _p(shep), _p(gp),
                  _p(lu), _p(du), U_(self.N))

    def lapu(self, u, rho, lap_form):
        lu = self.vec()
        self.call("bi_lapu", _p(self.imove), _p(self.r), _p(u), _p(rho),
                  _p(self.m), _p(lu), U_(self.N), *self.lla(), D_(self.H),
                  D_(self.CONF), ctypes.c_int(lap_form))
        return lu

    def interp(self, rho, gp, p):
        self.call("bi_interp", _p(self.imove), _p(self.r), _p(self.m),
                  _p(rho), _p(gp), _p(p), U_(self.N), *self.lla(),
                  D_(self.H), D_(self.CONW))


def _mutations_fail(check, mutations):
    """Each planted slip must make `check` fail."""
    for name, kw in mutations:
        with pytest.raises(AssertionError):
            check(**kw)
            raise RuntimeError("slip not caught: " + name)


# ---- boundary-integral interactions and the Shepard division --------------
@pytest.mark.parametrize("dims", [2, 3])
def test_bi_uniform_pressure_has_no_gradient(dims):
    """Divergence theorem: for a uniform pressure the truncated fluid sum
    2p/rho sum_j grad W V_j near the wall is cancelled by the wall integral
    2p/rho int W n dS (Interactions.cl:51-133 with outward normals and the
    element area in m).  A flipped normal doubles the residual instead, a
    missing area leaves the fluid term alone."""
    s = Slab(dims, 40 if dims == 2 else 22, 16 if dims == 2 else 10)
    P = 2.0e4
    p = np.full(s.N, P)
    rho = np.full(s.N, RHO0)
    u = s.vec()

    def check(normal=None, area=None):
        gp, lu, du = s.interactions(u, rho, p)
        fluid_only = np.abs(gp[s.near, dims - 1]).max()
        s.bi_interactions(u, rho, p, gp, du, normal, area)
        scale = 2 * P / RHO0 / s.h
        assert fluid_only > 0.2 * scale  # the truncated sum is not small
        assert np.abs(gp[s.near, :dims]).max() < 0.03 * scale
        assert np.abs(gp[s.deep, :dims]).max() < 1e-9 * scale

    check()
    _mutations_fail(check, [("flipped normal", dict(normal=-s.normal)),
                            ("area 1", dict(area=np.where(s.elem, 1.0, s.m))),
                            ("half area", dict(area=np.where(s.elem, 0.5 * s.m,
                                                             s.m)))])


@pytest.mark.parametrize("dims", [2, 3])
def test_bi_linear_fields_near_the_wall(dims):
    """(fluid sum + wall integral) / gamma equals grad(p)/rho for a linear
    pressure and rho div(u) for a linear velocity at every particle whose
    support reaches the wall (integration by parts of the SPH operator;
    Interactions.cl:51-133 then Shepard.cl:173-195 apply).  The pressure
    operator is linear in p with the symmetric weight p_i + p_j, so it is
    p_i times its value for p = 1 (the uniform-pressure residual, bounded
    above) plus the difference form sum (p_j - p_i) ..., and the latter is
    the one the divergence theorem makes exact.  Without the Shepard
    division, with a flipped normal, half the element area, or a wall that
    moves against the flow, the first rows are off by 7 to 50 percent."""
    s = Slab(dims, 40 if dims == 2 else 22, 16 if dims == 2 else 10)
    a = np.array([3.0e3, -2.0e3, 1.5e3])[:dims]
    G = np.array([[0.2, -0.1, 0.3], [0.05, 0.4, -0.2],
                  [0.1, 0.0, -0.3]])[:dims, :dims]
    p = 5.0e3 + s.x @ a
    rho = np.full(s.N, RHO0)
    u = s.vec(s.x @ G.T)
    gam = s.gamma()
    one = np.ones(s.N)

    def check(apply=True, wall_u=None, normal=None, area=None):
        gp, lu, du = s.interactions(u, rho, p)
        uu = u if wall_u is None else wall_u
        s.bi_interactions(uu, rho, p, gp, du, normal, area)
        g1, _, d1 = s.interactions(s.vec(), rho, one)
        s.bi_interactions(s.vec(), rho, one, g1, d1, normal, area)
        gp -= p[:, None] * g1
        if apply:
            s.apply(gam, gp, lu, du)
        sel = s.near
        np.testing.assert_allclose(gp[sel, :dims] * RHO0,
                                   np.tile(a, (sel.sum(), 1)),
                                   atol=0.04 * np.abs(a).max())
        np.testing.assert_allclose(du[sel] / RHO0, np.trace(G),
                                   atol=0.04 * np.abs(G).max())

    check()
    u_rev = u.copy()
    u_rev[s.elem] = -u[s.elem]  # a wall moving against the flow
    _mutations_fail(check, [("no Shepard division", dict(apply=False)),
                            ("wall velocity reversed", dict(wall_u=u_rev)),
                            ("flipped normal", dict(normal=-s.normal)),
                            ("half area",
                             dict(area=np.where(s.elem, 0.5 * s.m, s.m)))])
    # the apply tool divides exactly the three fields, fluid only
    gp, lu, du = s.vec(s.x @ G.T), s.vec(s.x[:, ::-1]), s.x[:, 0].copy()
    g0, l0, d0 = gp.copy(), lu.copy(), du.copy()
    s.apply(gam, gp, lu, du)
    f, e = s.fluid, s.elem
    np.testing.assert_array_equal(gp[f], g0[f] / gam[f, None])
    np.testing.assert_array_equal(lu[f], l0[f] / gam[f, None])
    np.testing.assert_array_equal(du[f], d0[f] / gam[f])
    np.testing.assert_array_equal(gp[e], g0[e])
    np.testing.assert_array_equal(du[e], d0[e])


# ---- wall pressure: interpolation and its Shepard / reverse-EOS step -------
@pytest.mark.parametrize("dims", [2, 3])
def test_bi_interpolation_of_a_linear_pressure_is_exact(dims):
    """The Taylor-corrected wall interpolation (Interpolation.cl:54-107)
    sum_j (p_j - rho_i grad_p_i . r_ij) W_ij V_j reproduces a linear
    (hydrostatic) pressure at the element exactly, up to the fluid kernel
    mass; divided by that mass it is exact to rounding, and divided by the
    element's boundary-integral gamma = 1/2 it is within the lattice's
    kernel-mass error.  A reversed Taylor sign or a dropped rho factor
    breaks it."""
    s = Slab(dims, 40 if dims == 2 else 22, 16 if dims == 2 else 10)
    g = np.zeros(dims)
    g[-1] = -9.81
    gx = np.array([0.7, -0.4, 0.0])[:dims]  # a sloping free surface
    grad = RHO0 * (g + gx)  # grad p of a hydrostatic field in a tilted frame
    p_true = 3.0e3 + s.x @ grad
    rho = np.full(s.N, RHO0)
    rho[s.elem] = 2.0 * RHO0  # only rho_i of the element enters, as a factor
    gp_el = s.vec(np.tile(grad / (2.0 * RHO0), (s.N, 1)))
    mass = s.fluid_shepard()

    def check(gp=gp_el, rho_=rho):
        p = np.where(s.fluid, p_true, 0.0)
        s.interp(rho_, gp, p)
        w = s.welem
        np.testing.assert_allclose(p[w] / mass[w], p_true[w], rtol=1e-12,
                                   atol=1e-9 * np.abs(p_true).max())

    check()
    gp_rev = gp_el.copy()
    gp_rev[:, :dims] *= -1
    _mutations_fail(check, [("Taylor sign", dict(gp=gp_rev)),
                            ("rho_j for rho_i",
                             dict(rho_=np.full(s.N, RHO0)))])
    # the interpolation Shepard step with gamma from the boundary integral:
    # the element gamma is 1/2 (the self term), the fluid mass is 1/2 to the
    # lattice's accuracy, and the reverse EOS gives back p through the EOS
    gam = s.gamma()
    np.testing.assert_allclose(gam[s.elem], 0.5, rtol=0, atol=1e-15)
    p = np.where(s.fluid, p_true, 0.0)
    s.interp(rho, gp_el, p)
    iset = np.zeros(s.N, np.uint32)
    refd = np.array([RHO0])
    cs, p0 = 25.0, 1.0e2
    rho_el = rho.copy()
    s.call("bi_interp_shepard", _p(iset), _p(s.imove), _p(gam), _p(rho_el),
           _p(p), _p(refd), U_(s.N), D_(cs), D_(p0))
    w = s.welem
    np.testing.assert_allclose(p[w], p_true[w], rtol=0.02)
    # EOS(rho) = p: the reverse EOS is the inverse of basic/EOS.cl
    pe = np.zeros(s.N)
    all_eos = np.full(s.N, -1, np.int32)
    s.call("eos", _p(iset), _p(all_eos), _p(rho_el),
           _p(pe), _p(refd), U_(s.N), D_(cs), D_(p0))
    np.testing.assert_allclose(pe[w], p[w], rtol=1e-12)
    # fluid untouched; an element with no fluid support keeps gamma -> 1
    np.testing.assert_array_equal(rho_el[s.fluid], rho[s.fluid])
    one = np.zeros(s.N)
    pp = np.full(s.N, 7.0)
    rr = np.zeros(s.N)
    s.call("bi_interp_shepard", _p(iset), _p(s.imove), _p(one), _p(rr),
           _p(pp), _p(refd), U_(s.N), D_(cs), D_(p0))
    np.testing.assert_array_equal(pp[s.elem], 7.0)
    np.testing.assert_allclose(rr[s.elem], RHO0 + (7.0 - p0) / cs ** 2)


# ---- wall Laplacian of the velocity and the wall pressure gradient --------
@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("lap_form", [1, 2])
def test_bi_lapu_and_gradp_of_a_quadratic_shear(dims, lap_form):
    """u = (a z^2, 0[, 0]) vanishes on the wall and has lap(u) = 2a e_x, div
    u = 0.  The element sum of LapU.cl:58-123 (both Laplacian forms) covers
    half the kernel, so lap_u / gamma (gamma = 1/2) is lap(u) to the
    lattice's accuracy.  GradP.cl:55-83 then states the wall's momentum
    balance grad(p)/rho = nu lap(u) + g - du/dt; with mu/rho0, a wall
    acceleration and g the expected value follows from the field."""
    s = Slab(dims, 40 if dims == 2 else 22, 16 if dims == 2 else 10)
    acoef = 50.0
    u = s.vec()
    u[:, 0] = acoef * s.x[:, dims - 1] ** 2
    rho = np.full(s.N, RHO0)
    rho[s.elem] = 3.0 * RHO0  # the element's own rho must not enter
    lap_exact = np.zeros(dims)
    lap_exact[0] = 2 * acoef
    gam = s.gamma()
    w = s.welem

    def check(rho_=rho):
        lu = s.lapu(u, rho_, lap_form)
        np.testing.assert_allclose(lu[w, :dims] / gam[w, None],
                                   np.tile(lap_exact, (w.sum(), 1)),
                                   atol=0.05 * 2 * acoef)
        return lu

    lu = check()
    _mutations_fail(check, [("rho_i for rho_j",
                             dict(rho_=np.where(s.elem, RHO0, 3.0 * RHO0)))])
    # the pressure gradient from the wall momentum balance
    mu = 2.5
    g = np.zeros(s.vs)
    g[dims - 1] = -9.81
    acc = s.vec(np.tile(np.array([0.3, -0.2, 0.1])[:dims], (s.N, 1)))
    iset = np.zeros(s.N, np.uint32)
    gp = s.vec()
    mu_a, refd = np.array([mu]), np.array([RHO0])
    s.call("bi_gradp", _p(iset), _p(s.imove), _p(gam), _p(rho), _p(lu),
           _p(acc), _p(gp), _p(mu_a), _p(refd), U_(s.N), _p(g))
    expect = mu / RHO0 * lap_exact + g[:dims] - acc[0, :dims]
    np.testing.assert_allclose(gp[w, :dims], np.tile(expect, (w.sum(), 1)),
                               atol=0.05 * mu / RHO0 * 2 * acoef + 1e-12)
    assert np.all(gp[s.fluid] == 0)
    # a still wall with no fluid support: gamma -> 1, grad p / rho = g
    gp0 = s.vec()
    zero, zv = np.zeros(s.N), s.vec()
    s.call("bi_gradp", _p(iset), _p(s.imove), _p(zero), _p(rho), _p(zv),
           _p(zv), _p(gp0), _p(mu_a), _p(refd), U_(s.N), _p(g))
    np.testing.assert_array_equal(gp0[s.elem], np.tile(g, (s.elem.sum(), 1)))


# ---- cfd rates: the momentum and continuity equations ----------------------
@pytest.mark.parametrize("dims", [2, 3])
def test_rates_are_the_navier_stokes_right_hand_side(dims):
    """With u = G x + (a z^2) e_x and p = p0 + b.x on an interior lattice,
    the interaction sums are grad(p)/rho, lap(u)/rho and rho div(u); the
    rates (cfd/Rates.cl:55-77) must give the momentum equation
    du/dt = -grad(p)/rho + mu lap(u)/rho + g = -b/rho + 2 mu a/rho e_x + g
    and continuity drho/dt = -rho tr(G).  Non-fluid particles keep their
    rates.  A reversed pressure sign, nu for mu, or a dropped g fails."""
    s = Slab(dims, 26 if dims == 2 else 14, 26 if dims == 2 else 14,
             wall=False)
    b = np.array([400.0, -300.0, 250.0])[:dims]
    G = np.array([[0.2, -0.1, 0.3], [0.05, 0.4, -0.2],
                  [0.1, 0.0, -0.3]])[:dims, :dims]
    acoef = 3.0
    u = s.vec(s.x @ G.T)
    u[:, 0] += acoef * s.x[:, dims - 1] ** 2
    p = 1.0e3 + s.x @ b
    rho = np.full(s.N, RHO0)
    mu = 200.0
    g = np.zeros(s.vs)
    g[dims - 1] = -9.81
    gp, lu, du = s.interactions(u, rho, p, 1)
    c = s.x.mean(0)
    L = s.x.max(0) - s.x.min(0)
    inner = np.all(np.abs(s.x - c) < L / 2 - 2.5 * s.h, axis=1)
    assert inner.sum() > 10
    iset = np.zeros(s.N, np.uint32)
    imove = s.imove.copy()
    imove[::7] = -3  # non-fluid particles keep what they hold
    expect = -b / RHO0 + g[:dims]
    expect[0] += mu * 2 * acoef / RHO0  # 1.2 against a lattice error < 0.07

    def check(visc=mu, grav=g, gsign=1.0):
        dudt = s.vec(np.full((s.N, dims), 123.0))
        drhodt = np.full(s.N, -7.0)
        mu_a, g_a, gp_a = np.array([visc]), grav.copy(), gsign * gp
        s.call("rates", _p(iset), _p(imove), _p(rho), _p(gp_a), _p(lu),
               _p(du), _p(dudt), _p(drhodt), _p(mu_a), U_(s.N), _p(g_a))
        sel = inner & (imove == 1)
        np.testing.assert_allclose(dudt[sel, :dims],
                                   np.tile(expect, (sel.sum(), 1)),
                                   atol=0.08)
        np.testing.assert_allclose(drhodt[sel], -RHO0 * np.trace(G),
                                   atol=0.02 * RHO0 * np.abs(G).max())
        assert np.all(dudt[imove != 1, :dims] == 123.0)
        assert np.all(drhodt[imove != 1] == -7.0)

    check()
    _mutations_fail(check, [("nu for mu", dict(visc=mu / RHO0)),
                            ("no gravity", dict(grav=0 * g)),
                            ("pressure sign", dict(gsign=-1.0))])


# ---- delta-SPH rate ----------------------------------------------------------
@pytest.mark.parametrize("dims", [2, 3])
def test_delta_rate_diffuses_density(dims):
    """basic/deltaSPH.cl:330-350 adds delta dt (rho_i/rho0) lap_p to drhodt,
    where lap_p = sum_j (p_j - p_i) F_ij V_j = lap(p)/2 in the interior.
    Closed forms: a quadratic p = c |x|^2 gives lap_p = dims c, so the added
    rate is delta dt (rho/rho0) dims c (rho != rho0 tells rho_i/rho0 from
    its inverse); the term is a diffusion (it lowers a pressure maximum) and
    with uniform rho and volumes it conserves sum V_i drhodt_i exactly over
    a closed set.  Only fluid particles change (cfd/deltaSPH.cl:29)."""
    s = Slab(dims, 24 if dims == 2 else 14, 24 if dims == 2 else 14,
             wall=False)
    cc = 2.0e5
    xc = s.x - s.x.mean(0)
    rho_f = 1.2 * RHO0
    rho = np.full(s.N, rho_f)
    m = np.full(s.N, rho_f * s.dr ** dims)  # volumes dr^dims
    delta, dt = 0.1, 2.0e-4
    iset = np.zeros(s.N, np.uint32)
    refd = np.array([RHO0])
    delta_a = np.array([delta])

    def rate(p, imove):
        lp = np.zeros(s.N)
        s.call("lapp", _p(imove), _p(s.r), _p(rho), _p(m), _p(p), _p(lp),
               U_(s.N), *s.lla(), D_(s.H), D_(s.CONF))
        dr = np.full(s.N, 0.5)
        s.call("delta_rate", _p(iset), _p(imove), _p(rho), _p(lp), _p(dr),
               _p(refd), _p(delta_a), U_(s.N), D_(dt))
        return dr - 0.5

    imove = s.imove.copy()
    add = rate(cc * (xc ** 2).sum(1), imove)
    L = s.x.max(0) - s.x.min(0)
    inner = np.all(np.abs(xc) < L / 2 - 2.5 * s.h, axis=1)
    np.testing.assert_allclose(add[inner],
                               delta * dt * rho_f / RHO0 * dims * cc,
                               rtol=0.02)
    # a Gaussian pressure bump: the maximum's density falls, the total
    # volume-weighted change over the closed set is zero
    pb = 1.0e4 * np.exp(-(xc ** 2).sum(1) / (0.15 * L.min()) ** 2)
    add = rate(pb, imove)
    assert add[np.argmax(pb)] < 0
    assert abs(add.sum()) < 1e-12 * np.abs(add).sum()
    # non-fluid particles are left alone
    im2 = imove.copy()
    im2[::5] = 0
    im2[1::5] = -3
    add = rate(pb, im2)
    assert np.all(add[im2 != 1] == 0)


# ---- binormal ------------------------------------------------------------------
@pytest.mark.parametrize("dims", [2, 3])
def test_binormal_frame(dims):
    """basic/Binormal.cl:37-53.  3D: b = n x t is a unit vector orthogonal to
    n and t with det[n, t, b] = +1 (right-handed); a transposed cross
    product gives -1.  2D: the tangent is n turned clockwise (t . n = 0,
    |t| = |n|, n_x t_y - n_y t_x = -|n|^2) and the binormal is zero."""
    rng = np.random.default_rng(11)
    N = 200
    vs = 2 if dims == 2 else 4
    n = np.zeros((N, vs))
    t = np.zeros((N, vs))
    b = np.full((N, vs), 9.0)
    q = rng.normal(size=(N, dims))
    n[:, :dims] = q / np.linalg.norm(q, axis=1)[:, None]
    if dims == 3:
        v = rng.normal(size=(N, 3))
        v -= (v * n[:, :3]).sum(1)[:, None] * n[:, :3]
        t[:, :3] = v / np.linalg.norm(v, axis=1)[:, None]
    t0 = t.copy()
    getattr(op.lib(), "orc_binormal_f64_%dd" % dims)(_p(n), _p(t), _p(b),
                                                    U_(N))
    if dims == 3:
        np.testing.assert_array_equal(t, t0)
        np.testing.assert_allclose((b[:, :3] * n[:, :3]).sum(1), 0, atol=1e-14)
        np.testing.assert_allclose((b[:, :3] * t[:, :3]).sum(1), 0, atol=1e-14)
        np.testing.assert_allclose(np.linalg.norm(b[:, :3], axis=1), 1,
                                   rtol=1e-14)
        det = np.linalg.det(np.stack([n[:, :3], t[:, :3], b[:, :3]], 1))
        np.testing.assert_allclose(det, 1.0, rtol=1e-12)
        assert np.all(b[:, 3] == 0)
        # a bottom wall: n = -e_z, t = e_x -> b = -e_y
        n1 = np.array([[0, 0, -1.0, 0]])
        t1 = np.array([[1.0, 0, 0, 0]])
        b1 = np.zeros((1, 4))
        op.lib().orc_binormal_f64_3d(_p(n1), _p(t1), _p(b1), U_(1))
        np.testing.assert_array_equal(b1[0, :3], [0, -1.0, 0])
    else:
        np.testing.assert_array_equal(b, 0)
        np.testing.assert_allclose((t * n).sum(1), 0, atol=1e-15)
        np.testing.assert_allclose(n[:, 0] * t[:, 1] - n[:, 1] * t[:, 0], -1,
                                   rtol=1e-14)


# ---- sensors --------------------------------------------------------------------
@pytest.mark.parametrize("dims", [2, 3])
def test_sensors_interpolate_fields(dims):
    """cfd/Sensors.cl:57-130 + SensorsRenormalization.cl:42-67 at sensors
    (imove = 0) placed off the lattice: with the fluid kernel mass as the
    Shepard factor a constant field is reproduced to rounding and a linear
    one to the first-moment error of the lattice; near a wall, gamma from
    the boundary integral (its sensor branch) recovers the field as well.
    A sensor never reads itself or another sensor, and fluid particles are
    unchanged."""
    s0 = Slab(dims, 40 if dims == 2 else 20, 16 if dims == 2 else 10)
    rng = np.random.default_rng(5)
    ns = 30
    L = (40 if dims == 2 else 20) * s0.dr
    sens = np.zeros((ns, dims))
    sens[:, :dims - 1] = rng.uniform(3 * s0.h, L - 3 * s0.h, (ns, dims - 1))
    sens[:, dims - 1] = rng.uniform(0.1 * s0.h, 2.5 * s0.h, ns)
    # rebuild the slab with the sensors appended
    r = np.zeros((s0.N + ns, s0.vs))
    r[:s0.N] = s0.r
    r[s0.N:, :dims] = sens
    imove = np.concatenate([s0.imove, np.zeros(ns, np.int32)])
    m = np.concatenate([s0.m, np.full(ns, 1.0)])
    normal = np.concatenate([s0.normal, np.zeros((ns, s0.vs))])
    tangent = np.concatenate([s0.tangent, np.zeros((ns, s0.vs))])
    binormal = np.concatenate([s0.binormal, np.zeros((ns, s0.vs))])
    ll = orc.link(dims, r, s0.h)
    perm = ll["perm"]
    imove, m = imove[perm].copy(), m[perm].copy()
    normal, tangent = normal[perm].copy(), tangent[perm].copy()
    binormal = binormal[perm].copy()
    r = ll["r"]
    x = r[:, :dims]
    N = r.shape[0]
    H, CONW, _ = orc.defs(s0.h, dims)
    sen = imove == 0
    far = sen & (x[:, dims - 1] > 2 * s0.h)
    b = np.array([2.0, -1.0, 0.5])[:dims]
    u_true = np.zeros((N, s0.vs))
    u_true[:, :dims] = 1.5 + x @ np.diag([1.0, -2.0, 3.0][:dims])
    p_true = 1.0e3 + 1.0e3 * (x @ b)
    rho_true = RHO0 + 0.1 * (x @ b)

    def run(shep):
        u = np.where(sen[:, None], 0.0, u_true)
        rho = np.where(sen, 0.0, rho_true)
        p = np.where(sen, 0.0, p_true)
        u[sen] = 99.0  # sensors' old values must not be read
        orc.call(ll, "sensors", _p(imove), _p(r), _p(m), _p(u), _p(rho),
                 _p(p), U_(N), *orc.ll_args(ll), D_(H), D_(CONW))
        orc.call(ll, "sensors_renorm", _p(imove), _p(shep), _p(u), _p(rho),
                 _p(p), U_(N))
        return u, rho, p

    # Shepard factor = fluid kernel mass sum_j W V_j with V_j = m_j / rho_j
    mass = np.zeros(N)
    orc.call(ll, "shepard", _p(imove), _p(r), _p(rho_true), _p(m), _p(mass),
             U_(N), *orc.ll_args(ll), D_(H), D_(CONW), ctypes.c_int(1))
    u, rho, p = run(mass)
    fl = imove == 1
    np.testing.assert_array_equal(u[fl], u_true[fl])
    np.testing.assert_array_equal(p[fl], p_true[fl])
    # constant fields exactly, linear ones to the lattice's first moment
    u_c = np.where(sen[:, None], 0.0, np.full((N, s0.vs), 2.5))
    rho_c = np.where(sen, 0.0, RHO0)
    p_c = np.where(sen, 0.0, 4.0e3)
    orc.call(ll, "sensors", _p(imove), _p(r), _p(m), _p(u_c), _p(rho_c),
             _p(p_c), U_(N), *orc.ll_args(ll), D_(H), D_(CONW))
    # the mass sum used rho_true; with uniform rho the factor differs by
    # the volume ratio, so recompute it for the uniform density
    mass_c = np.zeros(N)
    rho_u = np.full(N, RHO0)
    orc.call(ll, "shepard", _p(imove), _p(r), _p(rho_u), _p(m),
             _p(mass_c), U_(N), *orc.ll_args(ll), D_(H), D_(CONW),
             ctypes.c_int(1))
    orc.call(ll, "sensors_renorm", _p(imove), _p(mass_c), _p(u_c), _p(rho_c),
             _p(p_c), U_(N))
    np.testing.assert_allclose(u_c[sen, :dims], 2.5, rtol=1e-13)
    np.testing.assert_allclose(rho_c[sen], RHO0, rtol=1e-13)
    np.testing.assert_allclose(p_c[sen], 4.0e3, rtol=1e-13)
    # linear fields, sensors away from the wall
    # (measured errors: 5e-6 on u, 6e-3 on p, 6e-7 on rho in 2D)
    np.testing.assert_allclose(u[far, :dims], u_true[far, :dims], atol=3e-5)
    np.testing.assert_allclose(p[far], p_true[far], atol=0.03)
    np.testing.assert_allclose(rho[far], rho_true[far], atol=3e-6)
    # near the wall the boundary-integral gamma (sensor branch of
    # Shepard.cl:61-150) stands in for the kernel mass
    gam = np.zeros(N)
    orc.call(ll, "bi_shepard_compute", _p(imove), _p(r), _p(normal),
             _p(tangent), _p(binormal), _p(m), _p(gam), U_(N),
             *orc.ll_args(ll), D_(H), D_(CONW), D_(dims))
    near = sen & ~far
    assert near.sum() > 3
    np.testing.assert_allclose(gam[sen], mass[sen], atol=0.015)
    u, rho, p = run(gam)
    np.testing.assert_allclose(p[sen], p_true[sen], rtol=0.02)


# ---- hydrostatic rest in a boundary-integral tank ---------------------------
@pytest.mark.parametrize("dims", [2, 3])
def test_bi_still_tank_stays_at_rest(dims):
    """Water at hydrostatic rest filling a boundary-integral tank: one full
    oracle step (EOS, BI Shepard, interactions, wall interpolation,
    LapU / GradP, wall interactions, Shepard apply, rates, delta-SPH) must
    leave du/dt = 0 up to the method's own wall accuracy at h = 2 dr: the
    zeroth-order residual of the truncated kernel times the hydrostatic p_i
    (bounded in test_bi_uniform_pressure_has_no_gradient) gives at most
    0.8 g (2D) / 0.35 g (3D) on the first row, falling to a few percent of g
    from the fourth row on.  With the wall normals flipped the first rows
    feel 30 to 100 g."""
    from aquagpusph_amd import cases
    dr = 0.005 if dims == 2 else 0.01
    Lx, Hd, W = (0.3, 0.15, 0.16) if dims == 2 else (0.16, 0.1, 0.16)
    case = cases.still_tank(dims, dr, L=Lx, H=Hd, W=W, bc="BI")
    h = 2 * dr
    g = 9.81

    def check(flip=False):
        c = dict(case)
        if flip:
            c["normal"] = -case["normal"]
        o = op.OracleSim(c, dtype=np.float64)
        o.step()
        f = o.f
        fl = f["imove"] == 1
        x = f["r"][:, :dims]
        a = np.abs(f["dudt"][:, :dims]).max(1) / g
        z = x[:, dims - 1]
        lat = np.all((x[:, :dims - 1] > 2 * h) &
                     (x[:, :dims - 1] < np.array([Lx, W])[:dims - 1] - 2 * h),
                     axis=1)
        first = 0.8 if dims == 2 else 0.35
        for k, bound in enumerate([first, 0.35, 0.35, 0.05, 0.05, 0.05]):
            row = fl & lat & (np.abs(z - (k + 0.5) * dr) < 0.1 * dr)
            assert row.sum() > 4
            assert a[row].max() < bound, (k, a[row].max())
        deep = fl & lat & (z > 2 * h) & (z < Hd - 2 * h)
        assert a[deep].max() < 0.05
        side = fl & ~lat & (z > 2 * h) & (z < Hd - 2 * h)
        assert a[side].max() < 0.7

    check()
    _mutations_fail(check, [("flipped normals", dict(flip=True))])
