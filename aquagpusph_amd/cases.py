"""Particle-case generators (the role of the reference's examples/*/Create.py).

Each generator returns a plain dict:
    dims, N, r, normal, tangent, u, dudt, rho, drhodt, m, imove, iset
    params: h, cs, courant, g, refd/visc_dyn/delta (per set), dt_min, dt_Ma,
            p0, domain_min, domain_max, dr
Positions are float64 lattices; consumers narrow them to float32.

Geometry of the dam break follows the reference's 2D example
(examples/2D/lobovsky_etal_2014/src/Create.py:41-140): a fluid column of
width B and height H at the left end of a tank of length L and height D,
cell-centred lattice of spacing dr, one layer of boundary-integral elements
(imove = -3, area stored in m, outward normals) on every tank face.
Physical set-up: rho0 = 998, c0 = 10 sqrt(g H), h = hfac dr, hydrostatic
initial pressure with the density from the linear EOS, artificial viscosity
alpha folded into visc_dyn = max(alpha/8 rho0 h c0, mu) in 2D and alpha/10 in
3D (examples/2D/spheric_testcase5_dambreak/src/Create.py:63,
examples/3D/spheric_testcase2_dambreak/src/Create.py:73).
"""
import math

import numpy as np

G = 9.81


def _lattice(*axes):
    grids = np.meshgrid(*axes, indexing="ij")
    return np.stack([g.ravel() for g in grids], axis=1)


def dam_break(dims, dr, L=None, D=0.6, B=0.6, H=0.3, W=0.6, hfac=2.0,
              refd=998.0, alpha=0.02, mu=8.9e-4, delta=0.1, courant=0.1,
              walls=("bottom", "top", "left", "right", "front", "back"),
              bc="BI", tait_gamma=None):
    """Dam break with one layer of wall elements.

    2D: tank [-B, L-B] x [0, D], column [-B, 0] x [0, H] (g along -y).
    3D: tank [-B, L-B] x [0, W] x [0, D], column [-B, 0] x [0, W] x [0, H]
        (g along -z).
    bc="BI": boundary-integral elements (imove = -3).
    bc="GP": mirroring elements (imove = -2) with ghost particles
    (imove = -1) behind each, at 0.5 dr, 1.5 dr, ... < 2h along the outward
    normal, associated with their element (the layout of the reference's
    examples/2D/spheric_testcase6_movingsquare/src/Create.py:85-146);
    "associations" holds N for every other particle.
    tait_gamma: the initial density from the hydrostatic pressure through
    the Tait law of Presets/basic/tait.xml instead of the linear EOS.
    """
    if L is None:
        L = 1.61 if dims == 2 else 1.6
    nx = int(round(B / dr))
    nh = int(round(H / dr))
    Nx = int(round(L / dr))
    Nd = int(round(D / dr))
    A = Nx * dr - B
    Dt = Nd * dr
    h = hfac * dr
    cs = 10.0 * math.sqrt(G * H)
    vs = 2 if dims == 2 else 4
    vert = dims - 1
    xs = -B + (np.arange(nx) + 0.5) * dr
    hs = (np.arange(nh) + 0.5) * dr
    if dims == 2:
        fluid = _lattice(xs, hs)
    else:
        ny = int(round(W / dr))
        Wt = ny * dr
        ys = (np.arange(ny) + 0.5) * dr
        fluid = _lattice(xs, ys, hs)
    nf = fluid.shape[0]

    # wall elements: (positions, normal, tangent)
    els = []
    xe = -B + (np.arange(Nx) + 0.5) * dr
    de = (np.arange(Nd) + 0.5) * dr
    if dims == 2:
        if "bottom" in walls:
            els.append((_lattice(xe, [0.0]), (0, -1), (0, 0)))
        if "top" in walls:
            els.append((_lattice(xe, [Dt]), (0, 1), (0, 0)))
        if "left" in walls:
            els.append((_lattice([-B], de), (-1, 0), (0, 0)))
        if "right" in walls:
            els.append((_lattice([A], de), (1, 0), (0, 0)))
    else:
        ye = (np.arange(ny) + 0.5) * dr
        if "bottom" in walls:
            els.append((_lattice(xe, ye, [0.0]), (0, 0, -1), (1, 0, 0)))
        if "top" in walls:
            els.append((_lattice(xe, ye, [Dt]), (0, 0, 1), (1, 0, 0)))
        if "left" in walls:
            els.append((_lattice([-B], ye, de), (-1, 0, 0), (0, 1, 0)))
        if "right" in walls:
            els.append((_lattice([A], ye, de), (1, 0, 0), (0, 1, 0)))
        if "front" in walls:
            els.append((_lattice(xe, [0.0], de), (0, -1, 0), (1, 0, 0)))
        if "back" in walls:
            els.append((_lattice(xe, [Wt], de), (0, 1, 0), (1, 0, 0)))
    ne = sum(e[0].shape[0] for e in els)
    ghosts = []  # (positions, normal, element index offset in els order)
    if bc == "GP":
        nl = int(math.ceil(2.0 * h / dr - 0.5))  # layers with d < 2h
        layers = [(k + 0.5) * dr for k in range(nl) if (k + 0.5) * dr < 2.0 * h]
        o = nf
        for pos, n, t in els:
            k = pos.shape[0]
            nn = np.asarray(n[:dims], float)
            for d in layers:
                ghosts.append((pos + d * nn, n, t, np.arange(o, o + k)))
            o += k
    ng = sum(g[0].shape[0] for g in ghosts)
    N = nf + ne + ng

    r = np.zeros((N, vs))
    normal = np.zeros((N, vs))
    tangent = np.zeros((N, vs))
    r[:nf, :dims] = fluid
    o = nf
    for pos, n, t in els:
        k = pos.shape[0]
        r[o:o + k, :dims] = pos
        normal[o:o + k, :dims] = n[:dims]
        tangent[o:o + k, :dims] = t[:dims]
        o += k
    assoc = np.full(N, N, np.uint32)
    for pos, n, t, ref in ghosts:
        k = pos.shape[0]
        r[o:o + k, :dims] = pos
        normal[o:o + k, :dims] = n[:dims]
        tangent[o:o + k, :dims] = t[:dims]
        assoc[o:o + k] = ref
        o += k
    imove = np.full(N, -3 if bc == "BI" else -2, np.int32)
    imove[:nf] = 1
    imove[nf + ne:] = -1
    press = np.zeros(N)
    press[:nf] = refd * G * (H - fluid[:, vert])
    if tait_gamma is None:
        rho = refd + press / cs ** 2
    else:
        B = refd * cs ** 2 / tait_gamma
        rho = refd * (press / B + 1.0) ** (1.0 / tait_gamma)
    rho[nf:] = refd
    m = np.empty(N)
    m[:nf] = refd * dr ** dims
    m[nf:nf + ne] = dr ** (dims - 1)
    m[nf + ne:] = refd * dr ** dims

    visc = max(alpha / (8.0 if dims == 2 else 10.0) * refd * h * cs, mu)
    g = [0.0] * dims
    g[vert] = -G
    lo = [-B - 0.25 * L] + [-0.25 * D] * (dims - 1)
    hi = [A + 0.25 * L] + [1.5 * D] * (dims - 1)
    if dims == 3:
        lo[1], hi[1] = -0.25 * W, 1.25 * W
    params = dict(h=h, cs=cs, courant=courant, g=g, refd=[refd],
                  visc_dyn=[visc], delta=[delta], p0=0.0,
                  dt_min=0.25 * courant * h / cs, dt_Ma=0.1,
                  domain_min=lo, domain_max=hi, dr=dr, hfac=hfac)
    out = dict(dims=dims, N=N, n_fluid=nf, r=r, normal=normal,
               tangent=tangent, u=np.zeros((N, vs)), dudt=np.zeros((N, vs)),
               rho=rho, drhodt=np.zeros(N), m=m, imove=imove,
               iset=np.zeros(N, np.uint32), params=params, bc=bc)
    if bc == "GP":
        out["associations"] = assoc
    return out


def still_tank(dims, dr, L=0.3, H=0.15, W=0.2, hfac=2.0, refd=998.0,
               alpha=0.02, mu=8.9e-4, delta=0.1, courant=0.1, bc="GP",
               u0=0.0):
    """Water at rest filling a tank [0, L] x [0, H] (x [0, W] in 3D) up to
    its open top: walls on the bottom and on the sides up to the free
    surface, hydrostatic pressure.  The fluid stays at rest, which checks
    the wall pressure extension (dp/dn = rho g.n).  u0 gives the fluid a
    uniform horizontal velocity (a tangential wall flow)."""
    case = dam_break(dims, dr, L=L + 0.0, D=H, B=L, H=H, W=W, hfac=hfac,
                     refd=refd, alpha=alpha, mu=mu, delta=delta,
                     courant=courant,
                     walls=("bottom", "left", "right", "front", "back"),
                     bc=bc)
    # shift so that the tank is [0, L]: dam_break puts the column at [-B, 0]
    vs = 2 if dims == 2 else 4
    case["r"][:, 0] += L
    case["params"]["domain_min"][0] += L
    case["params"]["domain_max"][0] += L
    if u0:
        fl = case["imove"] == 1
        case["u"][fl, 0] = u0
    case["params"]["still_tank"] = True
    return case


def decaying_box(n, seed=0, k0=4.0, urms=1.0, nu=1e-4, cs=20.0, refd=1.0,
                 courant=0.1, delta=0.1, p0_factor=3.0):
    """Decaying 2D turbulence in a closed unit box: n x n fluid particles,
    boundary-integral walls on the four sides, no gravity, and a seeded
    solenoidal velocity field with energy spectrum E(k) ~ k^4
    exp(-2 (k/k0)^2), k in units of 2 pi (the spectrum of the
    usual decaying-turbulence cube benchmark, here in 2D), scaled to the RMS speed urms.
    The field is the curl of a stream function sum a_mn sin(m pi x)
    sin(n pi y), so it is divergence-free and does not cross the walls.
    Kinematic viscosity nu (no artificial viscosity), background pressure
    p0 = p0_factor rho0 urms^2, delta-SPH delta."""
    dr = 1.0 / n
    case = dam_break(2, dr, L=1.0, D=1.0, B=1.0, H=1.0, refd=refd,
                     alpha=0.0, mu=refd * nu, delta=delta, courant=courant,
                     walls=("bottom", "top", "left", "right"))
    case["r"][:, 0] += 1.0  # dam_break puts the tank at [-1, 0]
    nf = case["n_fluid"]
    x, y = case["r"][:nf, 0], case["r"][:nf, 1]
    rng = np.random.default_rng(seed)
    K = int(np.ceil(4.0 * k0)) + 1
    kk = 2.0 * math.pi * k0
    u = np.zeros(nf)
    v = np.zeros(nf)
    for mm in range(1, 2 * K + 1):
        sx, cx = np.sin(mm * math.pi * x), np.cos(mm * math.pi * x)
        for nn in range(1, 2 * K + 1):
            k = math.pi * math.hypot(mm, nn)
            E = k ** 4 * math.exp(-2.0 * (k / kk) ** 2)
            if E < 1e-12 * kk ** 4:
                continue
            a = rng.normal() * math.sqrt(E) / k
            sy, cy = np.sin(nn * math.pi * y), np.cos(nn * math.pi * y)
            u += a * nn * math.pi * sx * cy    # d psi / dy
            v -= a * mm * math.pi * cx * sy    # -d psi / dx
    s = urms / math.sqrt(np.mean(u * u + v * v))
    case["u"][:nf, 0] = u * s
    case["u"][:nf, 1] = v * s
    case["rho"][:] = refd
    p = case["params"]
    h = 2.0 * dr
    p.update(cs=cs, g=[0.0, 0.0], p0=p0_factor * refd * urms ** 2,
             visc_dyn=[refd * nu], dt_min=0.0, dt=0.0,
             domain_min=[-0.5, -0.5], domain_max=[1.5, 1.5], nu=nu)
    case["params"] = p
    return case


def solenoidal_field(dims, n, seed=0, k0=4.0, urms=1.0, kmax=None):
    """A seeded random divergence-free velocity on the n^dims periodic
    lattice of a unit box, with the energy spectrum of the decaying
    turbulence benchmark, E(k) ~ k^4 exp(-2 (k/k0)^2) (k in units of
    2 pi), scaled to the RMS speed urms.  Built in Fourier space over the
    wavenumbers |k_i| <= kmax (default ceil(4 k0), where E has fallen by
    e^-32 from its peak): random complex amplitudes of modulus
    sqrt(E(k) / (shell measure)), drawn in a fixed order, projected onto
    the plane normal to k (the spectral form of div u = 0); the field is
    the real part of their Fourier series.  The draws do not depend on n,
    so lattices with n > 2 kmax sample one and the same field (a coarse
    and a fine run start from the same flow); smaller lattices keep the
    modes they resolve (|k_i| < n / 2).  Returns an (n^dims, dims) array in
    the lattice order of np.meshgrid(..., indexing="ij")."""
    rng = np.random.default_rng(seed)
    K = int(math.ceil(4.0 * k0)) if kmax is None else int(kmax)
    ks = np.arange(-K, K + 1)
    Kv = np.stack(np.meshgrid(*([ks] * dims), indexing="ij"), axis=0)
    kk = np.sqrt(np.sum(Kv ** 2, axis=0))
    shell = 2.0 * math.pi * kk if dims == 2 else 4.0 * math.pi * kk ** 2
    E = kk ** 4 * np.exp(-2.0 * (kk / k0) ** 2)
    amp = np.zeros_like(kk)
    nz = kk > 0
    amp[nz] = np.sqrt(E[nz] / shell[nz])
    U = (rng.normal(size=(dims,) + kk.shape) +
         1j * rng.normal(size=(dims,) + kk.shape)) * amp
    kd = np.where(nz, kk, 1.0)
    U = U - Kv * (np.sum(U * Kv, axis=0) / kd ** 2)
    # at the cell-centred lattice points x = (i + 1/2) / n: each mode's
    # phase at the half-spacing offset
    U = U * np.exp(1j * math.pi * np.sum(Kv, axis=0) / n)
    # onto the lattice's FFT grid: the resolved modes (no Nyquist mode)
    keep = np.all(np.abs(Kv) < n / 2, axis=0)
    G = np.zeros((dims,) + (n,) * dims, complex)
    idx = tuple(np.where(keep, Kv[a] % n, 0) for a in range(dims))
    for a in range(dims):
        np.add.at(G[a], tuple(ix[keep] for ix in idx), U[a][keep])
    u = np.stack([np.real(np.fft.ifftn(G[a])) * n ** dims
                  for a in range(dims)], -1).reshape(-1, dims)
    u -= u.mean(axis=0)
    return u * (urms / math.sqrt(np.mean(np.sum(u ** 2, axis=1))))


def periodic_box(dims, n, seed=0, k0=4.0, urms=1.0, nu=1e-4, cs=10.0,
                 refd=1.0, courant=0.1, delta=0.1, hfac=2.0, p0_factor=1.0,
                 L=1.0):
    """Decaying turbulence in a box of side L periodic along every axis
    (Presets/cfd/periodic.xml): n^dims fluid particles on the
    cell-centred lattice, no walls, no gravity, the velocity of
    solenoidal_field (E(k) ~ k^4 exp(-2 (k/k0)^2), k0 = 4, RMS urms: the
    reference's configuration of the learned-closure study, triply periodic
    in 3D).  Kinematic viscosity nu, background pressure p0 = p0_factor
    rho0 urms^2, delta-SPH delta.  params["periodic"] holds the box."""
    dr = L / n
    xs = (np.arange(n) + 0.5) * dr
    pos = _lattice(*([xs] * dims))
    N = pos.shape[0]
    vs = 2 if dims == 2 else 4
    r = np.zeros((N, vs))
    r[:, :dims] = pos
    u = np.zeros((N, vs))
    u[:, :dims] = solenoidal_field(dims, n, seed, k0, urms)
    h = hfac * dr
    params = dict(h=h, cs=cs, courant=courant, g=[0.0] * dims, refd=[refd],
                  visc_dyn=[refd * nu], delta=[delta],
                  p0=p0_factor * refd * urms ** 2, dt_min=0.0, dt_Ma=0.1,
                  dr=dr, hfac=hfac, nu=nu,
                  periodic={"min": [0.0] * dims, "L": [L] * dims})
    return dict(dims=dims, N=N, n_fluid=N, r=r, normal=np.zeros((N, vs)),
                tangent=np.zeros((N, vs)), u=u, dudt=np.zeros((N, vs)),
                rho=np.full(N, refd), drhodt=np.zeros(N),
                m=np.full(N, refd * dr ** dims), imove=np.ones(N, np.int32),
                iset=np.zeros(N, np.uint32), params=params, bc="none")


def taylor_green(n=32, L=2.0 * math.pi, U=1.0, hfac=4.0, cs=50.0,
                 courant=0.1, refd=1.0, Re=2000.0, alpha=0.0, delta=0.0,
                 drift=(0.0, 0.0)):
    """Periodic 2D Taylor-Green vortices (the reference's
    examples/2D/taylor_green/src/Create.py:44-117): n x n fluid particles on
    the cell-centred lattice of [-L/2, L/2]^2, u = U sin(kx) cos(ky),
    v = -U cos(kx) sin(ky), p = rho0 U^2 (cos 2kx + cos 2ky) / 4 with
    k = 2 pi / L, rho = rho0 + p / cs^2, background pressure p0 = 3 rho0 U^2,
    mu = max(alpha/8 rho0 h cs, rho0 U L / Re).  No walls: the box is closed
    by the eight plane portals of taylor_green_portals.  `drift` adds a
    uniform velocity (the vortices then travel through the portals)."""
    dr = L / n
    h = hfac * dr
    k = 2.0 * math.pi / L
    xs = -0.5 * L + (np.arange(n) + 0.5) * dr
    pos = _lattice(xs, xs)
    N = pos.shape[0]
    x, y = pos[:, 0], pos[:, 1]
    r = np.zeros((N, 2))
    r[:, :] = pos
    u = np.zeros((N, 2))
    u[:, 0] = U * np.sin(k * x) * np.cos(k * y)
    u[:, 1] = -U * np.cos(k * x) * np.sin(k * y)
    u += np.asarray(drift, float)[None, :]
    press = refd * U ** 2 * (np.cos(2.0 * k * x) + np.cos(2.0 * k * y)) / 4.0
    rho = refd + press / cs ** 2
    m = np.full(N, refd * dr ** 2)
    visc = max(alpha / 8.0 * refd * h * cs, refd * U * L / Re)
    params = dict(h=h, cs=cs, courant=courant, g=[0.0, 0.0], refd=[refd],
                  visc_dyn=[visc], delta=[delta], p0=3.0 * refd * U ** 2,
                  dt_min=0.0, dt_Ma=0.1, domain_min=[-L, -L],
                  domain_max=[L, L], dr=dr, hfac=hfac, L=L,
                  E_kin=float(np.sum(0.5 * m * (u[:, 0] ** 2 + u[:, 1] ** 2))))
    return dict(dims=2, N=N, n_fluid=N, r=r, normal=np.zeros((N, 2)),
                tangent=np.zeros((N, 2)), u=u, dudt=np.zeros((N, 2)),
                rho=rho, drhodt=np.zeros(N), m=m,
                imove=np.ones(N, np.int32), iset=np.zeros(N, np.uint32),
                params=params, bc="portal")


def taylor_green_portals(L):
    """The eight portals that make [-L/2, L/2]^2 periodic, in the order of
    examples/2D/taylor_green/src/templates/BCs.xml:21-54: (prefix, out_r,
    in_r, n).  The diagonal pairs carry the corner neighbourhoods."""
    a = 0.5 * L
    s = 1.0 / math.sqrt(2.0)
    return [("right_", (a, 0.0), (-a, 0.0), (1.0, 0.0)),
            ("left_", (-a, 0.0), (a, 0.0), (-1.0, 0.0)),
            ("top_", (0.0, a), (0.0, -a), (0.0, 1.0)),
            ("bottom_", (0.0, -a), (0.0, a), (0.0, -1.0)),
            ("topright_", (a, a), (-a, -a), (s, s)),
            ("topleft_", (-a, a), (a, -a), (-s, s)),
            ("bottomright_", (a, -a), (-a, a), (s, -s)),
            ("bottomleft_", (-a, -a), (a, a), (-s, -s))]


def periodic_slab(dims=3, n=(16, 8, 8), dr=0.05, U=1.0, hfac=2.0, cs=20.0,
                  courant=0.1, refd=1000.0, mu=1e-3, delta=0.1, wall=False):
    """A fluid block [-Lx/2, Lx/2) x [0, Ly) (x [0, Lz)) periodic along x
    through one pair of plane portals, free surfaces elsewhere, with the
    sheared flow u = U sin(2 pi z / Lz) along x (z the last axis; n gives
    (nx, nz) in 2D).  wall=True adds a row of boundary-integral elements
    (imove -3, area in m, normal -z) under the block and gravity along -z:
    a periodic channel."""
    nx = n[0]
    Lx = nx * dr
    axes = [-0.5 * Lx + (np.arange(nx) + 0.5) * dr]
    axes += [(np.arange(n[k]) + 0.5) * dr for k in range(1, dims)]
    Lz = n[dims - 1] * dr
    fl = _lattice(*axes)
    nf = fl.shape[0]
    walls = np.zeros((0, dims))
    if wall:
        walls = _lattice(*(axes[:-1] + [[0.0]]))
    N = nf + walls.shape[0]
    vs = 2 if dims == 2 else 4
    h = hfac * dr
    r = np.zeros((N, vs))
    r[:nf, :dims] = fl
    r[nf:, :dims] = walls
    normal = np.zeros((N, vs))
    normal[nf:, dims - 1] = -1.0
    u = np.zeros((N, vs))
    u[:nf, 0] = U * np.sin(2.0 * math.pi * fl[:, dims - 1] / Lz)
    imove = np.ones(N, np.int32)
    imove[nf:] = -3
    m = np.full(N, refd * dr ** dims)
    m[nf:] = dr ** (dims - 1)
    g = [0.0] * dims
    if wall:
        g[dims - 1] = -G
    hi = [Lx] + [2.0 * n[k] * dr for k in range(1, dims)]
    params = dict(h=h, cs=cs, courant=courant, g=g, refd=[refd],
                  visc_dyn=[mu], delta=[delta], p0=0.0,
                  dt_min=0.0, dt_Ma=0.1, domain_min=[-Lx] * dims,
                  domain_max=hi, dr=dr, hfac=hfac, L=Lx)
    return dict(dims=dims, N=N, n_fluid=nf, r=r, normal=normal,
                tangent=np.zeros((N, vs)), u=u, dudt=np.zeros((N, vs)),
                rho=np.full(N, refd), drhodt=np.zeros(N), m=m, imove=imove,
                iset=np.zeros(N, np.uint32), params=params, bc="portal")


def slab_portals(L, dims=3):
    """Portals of periodic_slab: x = +L/2 -> -L/2 and back."""
    a = 0.5 * L
    z = (0.0,) * (dims - 1)
    return [("right_", (a,) + z, (-a,) + z, (1.0,) + z),
            ("left_", (-a,) + z, (a,) + z, (-1.0,) + z)]


def sloshing_tank(width, fill, res=200, refd=998.0, alpha=0.02, mu=8.9e-4,
                  delta=0.1, courant=0.1, amp=0.03, freq_ratio=1.0):
    """A closed square sloshing tank of side `width` filled to fill * width,
    in the tank's own units (lengths / width, time unchanged: the Froude
    scaling keeps every tank at dr = 1/res, h = 2 dr): fluid lattice, one
    layer of boundary-integral elements on the four walls, hydrostatic
    pressure.  The tank moves horizontally as amp * width * sin(w t) with w
    = freq_ratio times the first sloshing mode sqrt(g k tanh(k d)), k = pi /
    width.  In tank units: g' = g / width, c0' = 10 sqrt(g' fill),
    excitation amplitude amp, mu' = mu / width^2.  Returns a case with the
    per-set values in params["set"]."""
    dr = 1.0 / res
    case = dam_break(2, dr, L=1.0, D=1.0, B=1.0, H=fill, hfac=2.0,
                     refd=refd, alpha=alpha, mu=mu, delta=delta,
                     courant=courant,
                     walls=("bottom", "top", "left", "right"))
    gs = G / width
    cs = 10.0 * math.sqrt(gs * fill)
    h = 2.0 * dr
    nf = case["n_fluid"]
    y = case["r"][:nf, 1]
    press = refd * gs * (fill - y)
    case["rho"][:nf] = refd + press / cs ** 2
    visc = max(alpha / 8.0 * refd * h * cs, mu / width ** 2)
    k = math.pi  # the tank is 1 wide in its units
    w1 = math.sqrt(gs * k * math.tanh(k * fill))
    p = case["params"]
    p.update(cs=cs, g=[0.0, -gs], visc_dyn=[visc], dt_min=0.0,
             width=width, fill=fill)
    p["set"] = dict(refd=refd, visc_dyn=visc, delta=delta, cs_set=cs,
                    g_set=(0.0, -gs), exc_amp=amp, exc_omega=freq_ratio * w1,
                    dt_set=0.0, t_set=0.0)
    return case


def sloshing_batch(n_tanks, seed=0, res=200, gap=None, **kw):
    """n_tanks independent sloshing tanks in one case, one particle set per
    tank (the batched-ensemble workload): widths uniform in 0.5-1.5 m, fill
    ratios 0.2-0.6, excitation amplitudes 1-5 % of the width at 0.8-1.2
    times the first mode (seeded).  Each tank in its own units (side 1,
    dr = 1/res); the tanks sit on a square grid, `gap` (default 4 h) apart
    so that no pair interaction crosses from one to another.  The
    concatenated arrays keep each tank's particles contiguous, in tank
    order; case["sets"] lists (n, per-set scalars)."""
    rng = np.random.default_rng(seed)
    widths = rng.uniform(0.5, 1.5, n_tanks)
    fills = rng.uniform(0.2, 0.6, n_tanks)
    amps = rng.uniform(0.01, 0.05, n_tanks)
    ratios = rng.uniform(0.8, 1.2, n_tanks)
    h = 2.0 / res
    gap = 4.0 * h if gap is None else gap
    cols = int(math.ceil(math.sqrt(n_tanks)))
    tanks = []
    for k in range(n_tanks):
        t = sloshing_tank(widths[k], fills[k], res=res, amp=amps[k],
                          freq_ratio=ratios[k], **kw)
        # dam_break puts the tank at [-1, 0] x [0, 1]
        t["r"][:, 0] += 1.0 + (k % cols) * (1.0 + gap)
        t["r"][:, 1] += (k // cols) * (1.0 + gap)
        tanks.append(t)
    out = {"dims": 2, "N": sum(t["N"] for t in tanks),
           "n_fluid": sum(t["n_fluid"] for t in tanks)}
    for key in ("r", "normal", "tangent", "u", "dudt"):
        out[key] = np.concatenate([t[key] for t in tanks])
    for key in ("rho", "drhodt", "m", "imove"):
        out[key] = np.concatenate([t[key] for t in tanks])
    out["iset"] = np.concatenate([np.full(t["N"], k, np.uint32)
                                  for k, t in enumerate(tanks)])
    out["sets"] = [(t["N"], t["params"]["set"]) for t in tanks]
    lo = out["r"].min(axis=0) - 1.0
    hi = out["r"].max(axis=0) + 1.0
    p0 = tanks[0]["params"]
    out["params"] = dict(h=h, cs=max(t["params"]["cs"] for t in tanks),
                         courant=p0["courant"], g=[0.0, -G],
                         refd=[s["refd"] for _, s in out["sets"]],
                         visc_dyn=[s["visc_dyn"] for _, s in out["sets"]],
                         delta=[s["delta"] for _, s in out["sets"]], p0=0.0,
                         dt_min=0.0, dt_Ma=0.1, domain_min=list(lo),
                         domain_max=list(hi), dr=1.0 / res, hfac=2.0)
    out["tanks"] = tanks
    out["bc"] = "BI"
    return out


def batch_member(batch, k):
    """Tank k of a sloshing_batch as a stand-alone case at the same place:
    its scalars become the global ones (cs, g, refd, visc_dyn, delta) and
    params["excitation"] = (exc_amp, exc_omega, g) for a time-dependent
    gravity."""
    o = sum(n for n, _ in batch["sets"][:k])
    n, sc = batch["sets"][k]
    out = {"dims": 2, "N": n}
    for key in ("r", "normal", "tangent", "u", "dudt", "rho", "drhodt", "m",
                "imove"):
        out[key] = batch[key][o:o + n].copy()
    out["n_fluid"] = int((out["imove"] == 1).sum())
    out["iset"] = np.zeros(n, np.uint32)
    p = dict(batch["params"])
    p.update(cs=sc["cs_set"], g=list(sc["g_set"]), refd=[sc["refd"]],
             visc_dyn=[sc["visc_dyn"]], delta=[sc["delta"]],
             excitation=(sc["exc_amp"], sc["exc_omega"], list(sc["g_set"])))
    out["params"] = p
    out["bc"] = "BI"
    return out


def excitation_xml(case):
    """The set_scalar tool that makes a stand-alone case's gravity follow
    its excitation: g = g0 + exc_amp exc_omega^2 sin(exc_omega t) along x,
    updated at the start of the interactions stage (as the batched kernels
    do per set)."""
    a, w, g0 = case["params"]["excitation"]
    f = lambda x: "%.9g" % float(np.float32(x))
    return ("    <Variables>\n"
            '        <Variable name="exc_amp" type="float" value="%s" />\n'
            '        <Variable name="exc_omega" type="float" value="%s" />\n'
            "    </Variables>\n    <Tools>\n"
            '        <Tool action="insert" after="Sort" type="set_scalar" '
            'name="excitation" in="g" value="%s + exc_amp*exc_omega*exc_omega'
            '*sin(exc_omega*t), %s"/>\n    </Tools>'
            % (f(a), f(w), f(g0[0]), f(g0[1])))


def with_sensors(case, pos, normal, tangent=None):
    """The case plus a second particle set of pressure sensors: fixed
    massless particles (imove = 0, m = 0, rho = refd, u = 0) at `pos` with
    the given normals/tangents, set 1 with the same refd / visc_dyn / delta
    (the sensor set of examples/3D/spheric_testcase2_dambreak: Create.py
    writes Sensors.dat with mass 0 and imove 0, templates/Sensors.xml loads
    it as its own ParticlesSet)."""
    dims = case["dims"]
    vs = 2 if dims == 2 else 4
    pos = np.asarray(pos, float).reshape(-1, dims)
    ns = pos.shape[0]

    def vec(a):
        out = np.zeros((ns, vs))
        if a is not None:
            out[:, :dims] = np.broadcast_to(np.asarray(a, float)[..., :dims],
                                            (ns, dims))
        return out

    p = dict(case["params"])
    sc = {"refd": p["refd"][0], "visc_dyn": p["visc_dyn"][0],
          "delta": p["delta"][0]}
    out = dict(case)
    N = int(case["N"])
    out["N"] = N + ns
    out["r"] = np.concatenate([case["r"], vec(pos)])
    out["normal"] = np.concatenate([case["normal"], vec(normal)])
    out["tangent"] = np.concatenate([case["tangent"], vec(tangent)])
    for k in ("u", "dudt"):
        out[k] = np.concatenate([case[k], np.zeros((ns, vs))])
    out["rho"] = np.concatenate([case["rho"], np.full(ns, sc["refd"])])
    out["drhodt"] = np.concatenate([case["drhodt"], np.zeros(ns)])
    out["m"] = np.concatenate([case["m"], np.zeros(ns)])
    out["imove"] = np.concatenate([case["imove"], np.zeros(ns, np.int32)])
    out["iset"] = np.concatenate([case.get("iset", np.zeros(N, np.uint32)),
                                  np.ones(ns, np.uint32)]).astype(np.uint32)
    out["sets"] = [(N, sc), (ns, dict(sc))]
    for k in ("refd", "visc_dyn", "delta"):
        p[k] = [p[k][0], p[k][0]]
    out["params"] = p
    out["n_sensors"] = ns
    return out


def dam_break_sensors(dims, B=0.6, H=0.3, W=0.6):
    """Sensor points for the dam break: on the left wall inside the column
    (normal +x, into the fluid), on the floor under it (normal +z), inside
    the column, and one on the far wall that the water reaches later."""
    v = dims - 1
    pts, nrm = [], []

    def add(x, z, n):
        q = [0.0] * dims
        q[0], q[v] = x, z
        if dims == 3:
            q[1] = 0.5 * W
        pts.append(q)
        nn = [0.0] * dims
        nn[n[0]] = n[1]
        nrm.append(nn)

    for k in range(4):
        add(-B, (0.07 + 0.2 * k) * H, (0, 1.0))
    for k in range(3):
        add(-B + (0.15 + 0.3 * k) * B, 0.0, (v, 1.0))
    for k in range(3):
        add(-B + (0.25 + 0.25 * k) * B, (0.3 + 0.2 * k) * H, (v, 1.0))
    add(0.9, 0.05 * H, (0, -1.0))
    return np.array(pts), np.array(nrm)


def dam_break_dr_for(n_fluid, dims=3, B=0.6, H=0.3, W=0.6):
    """Lattice spacing that gives about n_fluid particles in the column."""
    vol = B * H * (W if dims == 3 else 1.0)
    return (vol / n_fluid) ** (1.0 / dims)


def random_box(dims, n, seed=0, lo=0.0, hi=1.0, dr=None):
    """Seeded random fluid particles in a box (interaction unit tests)."""
    rng = np.random.default_rng(seed)
    vs = 2 if dims == 2 else 4
    r = np.zeros((n, vs))
    r[:, :dims] = rng.uniform(lo, hi, (n, dims))
    if dr is None:
        dr = (hi - lo) / n ** (1.0 / dims)
    u = np.zeros((n, vs))
    u[:, :dims] = rng.normal(0.0, 0.1, (n, dims))
    rho = 1000.0 + rng.normal(0.0, 1.0, n)
    return dict(dims=dims, N=n, r=r, u=u, rho=rho,
                m=np.full(n, 1000.0 * dr ** dims),
                imove=np.ones(n, np.int32), dr=dr)
