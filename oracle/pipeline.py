"""CPU oracle pipeline: one AQUAgpusph time step replayed on the CPU.

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may use this module, never the product library.

The kernels are the C restatements in aqua_oracle.hpp; this module replays
them in the tool order that the reference presets resolve to
(resources/Presets/src/basic.xml:139-175 with the insert semantics of
aquagpusph/InputOutput/State.cpp:803-1033), for the preset stack

    basic + improved_euler + [domain] + cfd + [deltaSPH] + [BI | GP]
          + [variableTimeStep]

(the ghost-particle order of resources/Presets/src/cfd/GP.xml:50-80 when the
case carries "associations").

Scalars follow the reference's host semantics: set_scalar expressions are
evaluated in double and narrowed to the variable type; preset definitions
(H, CONW, CONF, DIMS) are evaluated, narrowed to float and re-parsed from
their "%#G" text (aquagpusph/CalcServer/CalcServer.cpp:241-260).
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    """Load (building if needed) the oracle shared library."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_aqua.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
    return _LIB


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def definition(value):
    """Round a preset definition like the reference: float, then %#G."""
    return float(np.float32(float("%#G" % float(np.float32(value)))))


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class OracleSim:
    """Holds every pipeline variable as a numpy array and steps it."""

    VEC_FIELDS = ("r", "normal", "tangent", "binormal", "u", "dudt")
    SCA_FIELDS = ("rho", "drhodt", "m", "p")

    def __init__(self, case, dtype=np.float32, domain=True, bi=None,
                 delta=True, variable_dt=True, lap_form=1, gp_model=2,
                 freeslip_iset=None, bie=False, dr_factor=0.5,
                 min_bound_dist=0.0, delta_corr=None, energy=False,
                 midpoint=None, portals=None, precision=False,
                 excitation=None, les=None, elastic_bounce=None,
                 tait_gamma=None, periodic=None):
        self.dims = case["dims"]
        self.vs = 2 if self.dims == 2 else 4
        self.dtype = np.dtype(dtype)
        self.sfx = ("f32" if self.dtype == np.float32 else "f64") + \
            "_%dd" % self.dims
        gp = "associations" in case
        if bi is None:
            bi = not gp and not bie
        self.opts = dict(domain=domain, bi=bi, gp=gp, delta=delta,
                         variable_dt=variable_dt, lap_form=lap_form,
                         gp_model=gp_model, freeslip_iset=freeslip_iset,
                         bie=bie, dr_factor=dr_factor,
                         min_bound_dist=min_bound_dist,
                         delta_corr=delta_corr, energy=energy,
                         midpoint=midpoint)
        if portals and midpoint is not None:
            raise ValueError("portals are modelled with improved Euler only")
        # Presets/precision.xml: compensated walk sums, double positions
        # (r + dr64), double mass and energy sums; a float32 model
        self.precision = bool(precision)
        # (amp, omega, g0): the gravity follows a horizontal harmonic motion
        # of the frame, g = g0 + amp omega^2 sin(omega t) along x, set (in
        # double, narrowed) at the start of the interactions stage
        self.excitation = excitation
        # cfd/LES.xml (les = {"params": P}) or cfd/LES-smagorinsky.xml
        # (les = {"Cs": C}): the eddy viscosity after the delta-SPH rate
        self.les = les
        # cfd/elasticBounce.xml: {"dr": lattice spacing, and optionally
        # "dr_factor" (1.5), "min_bound_dist" (0.3), "elastic_factor" (0)}
        if elastic_bounce is not None:
            eb = dict(dr_factor=1.5, min_bound_dist=0.3, elastic_factor=0.0)
            eb.update(elastic_bounce)
            elastic_bounce = eb
        self.elastic_bounce = elastic_bounce
        # basic/tait.xml: the Tait EOS with this gamma (and its inverse for
        # the boundary-integral wall density)
        self.tait_gamma = tait_gamma
        # Presets/cfd/periodic.xml: {"min": box origin, "L": lengths (0 for
        # an axis that is not periodic)}; walls are not modelled with it
        self.periodic = None
        if periodic is not None:
            if portals or bie or midpoint is not None:
                raise ValueError("the periodic box is modelled with improved "
                                 "Euler and no walls or portals")
            vs = 2 if case["dims"] == 2 else 4
            pm = np.zeros(vs)
            pl = np.zeros(vs)
            pm[:case["dims"]] = periodic["min"][:case["dims"]]
            pl[:case["dims"]] = periodic["L"][:case["dims"]]
            # kept in the working precision (the kernels take pointers)
            self.periodic = (pm.astype(self.dtype), pl.astype(self.dtype))
        if precision and (midpoint is not None or dtype != np.float32):
            raise ValueError("the precision mode is a float32 improved-Euler "
                             "model")
        if midpoint is not None:
            # basic/time_scheme/midpoint.xml (+ midpoint/{r,autostop,
            # autorelax}.xml): iter_max, relax, r, Residual_max, mindec, inc
            mp = dict(iter_max=50, relax=0.0, r=False, residual_max=None,
                      mindec=None, inc=0.25)
            mp.update(midpoint)
            self.mp = mp
            self.relax = np.float32(mp["relax"])
            self.relax0 = self.relax
            self.residual = np.float32(0.0)
            self.residual_prev = np.float32(1.0)
            self.iter_midpoint = 0
            self.mp_iters = []  # inner iterations of each step
            self.residuals = []  # Residual_midpoint of each inner iteration
        N = self.N = int(case["N"])
        R = self.dtype
        f = {}
        for k in self.VEC_FIELDS:
            v = case.get(k)
            f[k] = np.zeros((N, self.vs), R) if v is None else \
                np.ascontiguousarray(v, dtype=R).reshape(N, self.vs).copy()
        for k in self.SCA_FIELDS:
            v = case.get(k)
            f[k] = np.zeros(N, R) if v is None else \
                np.ascontiguousarray(v, dtype=R).copy()
        f["imove"] = np.ascontiguousarray(case["imove"], np.int32).copy()
        f["iset"] = np.ascontiguousarray(
            case.get("iset", np.zeros(N)), np.uint32).copy()
        f["id"] = np.arange(N, dtype=np.uint32)
        for k in ("r", "u", "dudt", "normal", "tangent"):
            f[k + "_in"] = np.zeros_like(f[k])
        for k in ("rho", "drhodt", "m"):
            f[k + "_in"] = np.zeros_like(f[k])
        for k in ("id", "iset", "imove"):
            f[k + "_in"] = np.zeros_like(f[k])
        f["grad_p"] = np.zeros((N, self.vs), R)
        f["lap_u"] = np.zeros((N, self.vs), R)
        for k in ("div_u", "shepard", "lap_p", "dt_var"):
            f[k] = np.zeros(N, R)
        for k in ("icell", "id_sorted", "id_unsorted"):
            f[k] = np.zeros(N, np.uint32)
        if delta_corr:
            f["lap_p_corr"] = np.zeros((N, self.vs), R)
        if delta_corr == "full":
            ms = 4 if self.dims == 2 else 16
            for k in ("mls", "mls_fluid"):
                f[k] = np.zeros((N, ms), R)
        if energy:
            # cfd/energy_kin.xml and cfd/energy.xml arrays and scalars
            for k in ("energy_kin", "energy_dekdt", "energy_depdt",
                      "energy_decdt", "energy_ek", "energy_ep", "energy_ec"):
                f[k] = np.zeros(N, R)
            self.energy = {k: R.type(0) for k in (
                "energy_Kin", "energy_dEkdt", "energy_dEpdt", "energy_dEcdt",
                "energy_Ek", "energy_Ep", "energy_Ec", "energy_Ek_ref",
                "energy_Ep_ref", "energy_Ec_ref")}
        if bie:
            f["grad_w_bi"] = np.zeros((N, self.vs), R)
            f["div_u_bi"] = np.zeros(N, R)
        if gp:
            f["associations"] = np.ascontiguousarray(
                case["associations"], np.uint32).copy()
            f["associations_in"] = np.zeros(N, np.uint32)
            f["gp_icell"] = np.zeros(N, np.uint32)
            f["gp_rho"] = np.zeros(N, R)
            f["gp_p"] = np.zeros(N, R)
            for k in ("gp_u", "gp_r_in", "gp_u_in"):
                f[k] = np.zeros((N, self.vs), R)
        if les is not None:
            f["grad_u"] = np.zeros((N, 4 if self.dims == 2 else 16), R)
            f["nu_t"] = np.zeros(N, R)
            f["les_c"] = np.zeros(N, R)
        if precision:
            f["dr64"] = np.zeros((N, self.vs), np.float64)
            f["dr64_in"] = np.zeros((N, self.vs), np.float64)
            self.mass_fp64 = 0.0
        if portals:
            # cfd/portal.xml once per (prefix, out_r, in_r, n), in include
            # order; the set_scalar tools before each mirror/teleport tool
            # give every stage its own portal's planes
            f["imirrored"] = np.zeros(N, np.int32)
            f["icell_backup"] = np.zeros(N, np.uint32)
            self.portals = []
            for _, out_r, in_r, n in portals:
                pv = {}
                for key, v in (("out_r", out_r), ("in_r", in_r), ("n", n)):
                    a = np.zeros(self.vs, R)
                    a[:len(v)] = np.asarray(v, np.float32)[:self.vs]
                    pv[key] = a
                self.portals.append(pv)
        else:
            self.portals = []
        self.f = f
        self.ihoc = np.zeros(0, np.uint32)
        self.n_cells = np.zeros(4, np.uint32)
        self.r_min = np.zeros(self.vs, R)
        self.r_max = np.zeros(self.vs, R)

        p = case["params"]
        # inputs are float variables in the reference: narrow them to float
        # first so that the FP64 run starts from the same values
        f32 = np.float32
        self.sets = {k: np.asarray(p[k], f32).astype(R)
                     for k in ("refd", "visc_dyn", "delta")}
        g = np.zeros(self.vs, R)
        g[:self.dims] = np.asarray(p["g"][:self.dims], f32)
        self.g = g
        self.h = np.float32(p["h"])
        self.cs = np.float32(p["cs"])
        self.p0 = np.float32(p.get("p0", 0.0))
        self.courant = np.float32(p["courant"])
        self.dt_Ma = np.float32(p.get("dt_Ma", 0.1))
        self.dt_min = np.float32(p.get("dt_min", 0.0))
        self.support = np.float32(2.0)
        dmin = np.zeros(self.vs, R)
        dmax = np.zeros(self.vs, R)
        if domain:
            dmin[:self.dims] = np.asarray(p["domain_min"][:self.dims], f32)
            dmax[:self.dims] = np.asarray(p["domain_max"][:self.dims], f32)
        self.domain_min, self.domain_max = dmin, dmax
        # preset definitions (basic.xml Definitions), float -> %#G -> float
        self.H = definition(self.h)
        self.CONW = definition(1.0 / float(self.h) ** self.dims)
        self.CONF = definition(1.0 / float(self.h) ** (self.dims + 2))
        self.DIMS = definition(self.dims)
        self.t = self.dtype.type(0.0)
        self.dt = self.dtype.type(p.get("dt", 0.0))
        self.iter = 0
        self.cell_length = np.float32(self.support * self.h)
        L = lib()
        self._k = lambda name: getattr(L, "orc_%s_%s" % (name, self.sfx))

    # ---- helpers -----------------------------------------------------
    def _sum(self, a):
        """A sum reduction of the oracle's real type (accumulated in
        float64; the GPU's tree order differs, so sums agree to a tolerance
        only)."""
        return self.dtype.type(np.sum(a, dtype=np.float64))

    def _esum(self, a):
        """An energy sum: float (as _sum), or double in the precision
        mode (its energy variables are double)."""
        if self.precision:
            return np.float64(np.sum(a, dtype=np.float64))
        return self._sum(a)

    def _R(self, x):
        return self.dtype.type(x)

    def _r(self, x):
        """ctypes scalar of the oracle's real type."""
        return (ctypes.c_float if self.dtype == np.float32 else
                ctypes.c_double)(float(x))

    def _ll(self):
        f = self.f
        return (_p(f["icell"]), _p(self.ihoc), _p(self.n_cells))

    def _portal_stage(self, stage, ll):
        """The tools one cfd/portal.xml include puts at one place, for each
        portal in turn: icell backup, mirror, the stage's kernels, unmirror,
        icell restore (cfd/portal.xml stages 1-3).  stage: "mls" (before
        imove1_MLS), "forces" (before Interactions) or "lapp_corr" (before
        LapP Correction)."""
        f, o = self.f, self.opts
        k, R = self._k, self._r
        n = ctypes.c_uint(self.N)
        H, CONW, CONF = R(self.H), R(self.CONW), R(self.CONF)
        for pv in self.portals:
            f["icell_backup"][...] = f["icell"]
            k("portal_mirror")(_p(f["r"]), _p(f["imirrored"]), _p(f["icell"]),
                               n, _p(pv["in_r"]), _p(pv["out_r"]),
                               _p(pv["n"]), _p(self.r_min), _p(self.n_cells),
                               H)
            if stage == "mls":
                k("portal_mls")(_p(f["imove"]), _p(f["imirrored"]),
                                _p(f["r"]), _p(f["rho"]), _p(f["m"]),
                                _p(f["mls"]), n, ctypes.c_uint(1), *ll, H,
                                CONF)
            elif stage == "forces":
                k("portal_shepard")(_p(f["imove"]), _p(f["imirrored"]),
                                    _p(f["r"]), _p(f["rho"]), _p(f["m"]),
                                    _p(f["shepard"]), n, *ll, H, CONW)
                k("portal_interactions")(
                    _p(f["imove"]), _p(f["imirrored"]), _p(f["r"]),
                    _p(f["u"]), _p(f["rho"]), _p(f["m"]), _p(f["p"]),
                    _p(f["grad_p"]), _p(f["lap_u"]), _p(f["div_u"]), n, *ll,
                    H, CONF, ctypes.c_int(o["lap_form"]))
                if o["bi"]:
                    k("portal_bi")(_p(f["imove"]), _p(f["imirrored"]),
                                   _p(f["r"]), _p(f["normal"]), _p(f["u"]),
                                   _p(f["rho"]), _p(f["m"]), _p(f["p"]),
                                   _p(f["grad_p"]), _p(f["div_u"]), n, *ll,
                                   H, CONW)
                if o["delta"]:
                    k("portal_lapp")(_p(f["imove"]), _p(f["imirrored"]),
                                     _p(f["r"]), _p(f["rho"]), _p(f["m"]),
                                     _p(f["p"]), _p(f["lap_p"]), None, n,
                                     *ll, H, CONF)
                if o["delta_corr"] == "full":
                    k("portal_lapp")(_p(f["imove"]), _p(f["imirrored"]),
                                     _p(f["r"]), _p(f["rho"]), _p(f["m"]),
                                     _p(f["p"]), None, _p(f["lap_p_corr"]),
                                     n, *ll, H, CONF)
            else:
                k("portal_lapp_corr")(_p(f["imove"]), _p(f["imirrored"]),
                                      _p(f["r"]), _p(f["rho"]), _p(f["m"]),
                                      _p(f["lap_p_corr"]), _p(f["lap_p"]), n,
                                      *ll, H, CONF)
            k("portal_unmirror")(_p(f["r"]), _p(f["imirrored"]), n,
                                 _p(pv["in_r"]), _p(pv["out_r"]))
            f["icell"][...] = f["icell_backup"]

    def _les(self, ll):
        """cfd/LES*.xml, before Rates: velocity gradient, (features and
        network,) eddy viscosity, eddy-viscosity term."""
        from oracle import les as olq
        from oracle import mlp as omlp
        f, N, k, R = self.f, self.N, self._k, self._r
        n = ctypes.c_uint(N)
        H, CONF = R(self.H), R(self.CONF)
        k("les_gradu")(_p(f["imove"]), _p(f["r"]), _p(f["u"]), _p(f["rho"]),
                       _p(f["m"]), _p(f["grad_u"]), n, *ll, H, CONF)
        if "params" in self.les:
            x = olq.features(f["grad_u"], self.dims,
                             float(np.float32(self.les.get("hdr", 2.0))))
            x = x.astype(np.float32)
            c, _ = omlp.forward(self.les["params"], x, bf=True)
            f["les_c"][...] = c
            nu = olq.eddy_viscosity(f["imove"], f["grad_u"], self.dims,
                                    float(self.H), c=f["les_c"],
                                    cmin=float(np.float32(
                                        self.les.get("cmin", 0.0))))
        else:
            nu = olq.eddy_viscosity(f["imove"], f["grad_u"], self.dims,
                                    float(self.H),
                                    Cs=float(np.float32(self.les["Cs"])))
        f["nu_t"][...] = nu
        k("les_interactions")(_p(f["imove"]), _p(f["r"]), _p(f["u"]),
                              _p(f["rho"]), _p(f["m"]), _p(f["nu_t"]),
                              _p(f["dudt"]), n, *ll, H, CONF)

    def _portal_teleport(self):
        """"cfd portal mirror_teleport" of each portal (before Corrector,
        i.e. after the corrector kernel)."""
        for pv in self.portals:
            self._k("portal_teleport")(_p(self.f["r"]), ctypes.c_uint(self.N),
                                       _p(pv["in_r"]), _p(pv["out_r"]),
                                       _p(pv["n"]))

    # ---- link-list (LinkList.cpp:326-494) ----------------------------
    def link_list(self, r_name="r_in"):
        f, L, N = self.f, lib(), self.N
        self._k("minmax_vec")(_p(f[r_name]), ctypes.c_uint(N),
                              _p(self.r_min), _p(self.r_max))
        if self.periodic is not None:
            pm, pl = self.periodic
            rmin = self.r_min.copy()
            ok = self._k("n_cells_per")(_p(self.r_min), _p(self.r_max),
                                        _p(pm), _p(pl),
                                        self._r(self.cell_length),
                                        _p(self.n_cells), _p(rmin))
            if not ok:
                raise ValueError("periodic box under 3 cells of support * h")
            self.r_min[...] = rmin
            nw = int(self.n_cells[3])
            if self.ihoc.size < nw:
                self.ihoc = np.zeros(nw, np.uint32)
            self._k("icell_per")(_p(f["icell"]), _p(f[r_name]),
                                 ctypes.c_uint(N), _p(self.r_min),
                                 self._r(self.support), self._r(self.h),
                                 _p(self.n_cells), _p(pm), _p(pl))
        else:
            self._k("n_cells")(_p(self.r_min), _p(self.r_max),
                               self._r(self.cell_length), _p(self.n_cells))
            nw = int(self.n_cells[3])
            if self.ihoc.size < nw:
                self.ihoc = np.zeros(nw, np.uint32)
            self._k("icell")(_p(f["icell"]), _p(f[r_name]), ctypes.c_uint(N),
                             _p(self.r_min), self._r(self.support),
                             self._r(self.h), _p(self.n_cells))
        bits = L.orc_radix_key_bits(ctypes.c_ulonglong(nw))
        L.orc_stable_sort(_p(f["icell"]), _p(f["id_unsorted"]),
                          _p(f["id_sorted"]), ctypes.c_uint(N),
                          ctypes.c_uint(bits))
        L.orc_ihoc(_p(f["icell"]), _p(self.ihoc), ctypes.c_uint(N),
                   _p(self.n_cells))

    def reorder(self):
        """Backup copies + sort stage1/stage2 (basic.xml:141-148)."""
        f = self.f
        for k in ("id", "iset", "imove", "normal", "tangent", "m"):
            f[k + "_in"][...] = f[k]
        ids = f["id_sorted"]

        def sc(src, dst):
            dst[ids] = src
        sc(f["id_in"], f["id"])
        sc(f["iset_in"], f["iset"])
        sc(f["imove_in"], f["imove"])
        sc(f["r_in"], f["r"])
        sc(f["normal_in"], f["normal"])
        sc(f["tangent_in"], f["tangent"])
        sc(f["rho_in"], f["rho"])
        sc(f["m_in"], f["m"])
        sc(f["u_in"], f["u"])
        sc(f["dudt"], f["dudt_in"])
        sc(f["drhodt"], f["drhodt_in"])
        if self.precision:
            sc(f["dr64_in"], f["dr64"])

    # ---- one time step ----------------------------------------------
    def _forces(self, ll, reinit_lap_p=True):
        """Interactions stage: the "Sort" to "Rates" tools (reinits, Shepard,
        interactions, sensors, lap p, wall terms, rates, delta-SPH)."""
        f, N, o = self.f, self.N, self.opts
        k, R = self._k, self._r
        n = ctypes.c_uint(N)
        H, CONW, CONF = R(self.H), R(self.CONW), R(self.CONF)
        if o["delta_corr"]:
            f["lap_p_corr"][...] = 0
        if o["delta_corr"] == "full":
            # MLS.xml (prefix imove1_): matrices of the fluid particles
            k("mls_entry")(_p(f["imove"]), _p(f["r"]), _p(f["rho"]),
                           _p(f["m"]), _p(f["mls"]), n, ctypes.c_uint(1), *ll,
                           H, CONF)
            if self.portals:
                self._portal_stage("mls", ll)
            k("mls_inv")(_p(f["imove"]), _p(f["mls"]), n, ctypes.c_uint(1))
            f["mls_fluid"][...] = f["mls"]
        for key in ("lap_p", "lap_u", "div_u", "grad_p", "shepard"):
            if key != "lap_p" or reinit_lap_p:
                f[key][...] = 0
        if o["bi"]:
            k("bi_shepard_compute")(
                _p(f["imove"]), _p(f["r"]), _p(f["normal"]), _p(f["tangent"]),
                _p(f["binormal"]), _p(f["m"]), _p(f["shepard"]), n, *ll,
                H, CONW, R(self.DIMS))
        else:
            k("shepard")(_p(f["imove"]), _p(f["r"]), _p(f["rho"]), _p(f["m"]),
                         _p(f["shepard"]), n, *ll, H, CONW, ctypes.c_int(1))
        if o["gp"]:
            # mirror, interpolate, renormalise, unmirror (GP.xml:58-70)
            f["gp_r_in"][...] = f["r"]
            k("gp_mirror")(_p(f["imove"]), _p(f["associations"]),
                           _p(f["normal"]), _p(f["r"]), _p(f["gp_icell"]), n,
                           _p(self.r_min), _p(self.n_cells), H,
                           R(self.support))
            for key in ("gp_rho", "gp_p", "gp_u"):
                f[key][...] = 0
            k("gp_interp")(_p(f["imove"]), _p(f["r"]), _p(f["m"]),
                           _p(f["rho"]), _p(f["p"]), _p(f["u"]),
                           _p(f["gp_rho"]), _p(f["gp_p"]), _p(f["gp_u"]),
                           _p(f["shepard"]), n, _p(f["gp_icell"]), *ll, H,
                           CONW)
            k("gp_renorm")(_p(f["imove"]), _p(f["shepard"]), _p(f["gp_rho"]),
                           _p(f["gp_p"]), _p(f["gp_u"]), n)
            f["r"][...] = f["gp_r_in"]
            f["gp_u_in"][...] = f["u"]
        kh = "_kahan" if self.precision else ""
        k("interactions" + kh)(_p(f["imove"]), _p(f["r"]), _p(f["u"]),
                               _p(f["rho"]), _p(f["m"]), _p(f["p"]),
                               _p(f["grad_p"]), _p(f["lap_u"]),
                               _p(f["div_u"]), n, *ll, H, CONF,
                               ctypes.c_int(o["lap_form"]))
        k("sensors")(_p(f["imove"]), _p(f["r"]), _p(f["m"]), _p(f["u"]),
                     _p(f["rho"]), _p(f["p"]), n, *ll, H, CONW)
        k("sensors_renorm")(_p(f["imove"]), _p(f["shepard"]), _p(f["u"]),
                            _p(f["rho"]), _p(f["p"]), n)
        if o["delta_corr"] == "full":
            f["mls"][...] = f["mls_fluid"]
            k("lapp_full")(_p(f["imove"]), _p(f["r"]), _p(f["rho"]),
                           _p(f["m"]), _p(f["p"]), _p(f["lap_p_corr"]), n,
                           *ll, H, CONF)
        elif o["delta_corr"] == "simple":
            k("lapp_simple")(_p(f["iset"]), _p(f["imove"]),
                             _p(f["lap_p_corr"]), _p(self.sets["refd"]), n,
                             _p(self.g))
        if o["delta"]:
            k("lapp" + kh)(_p(f["imove"]), _p(f["r"]), _p(f["rho"]),
                           _p(f["m"]), _p(f["p"]), _p(f["lap_p"]), n, *ll, H,
                           CONF)
        if o["bie"]:
            # BIe.xml: boundary terms of the conservative integrals
            f["grad_w_bi"][...] = 0
            f["div_u_bi"][...] = 0
            k("bie_interactions")(_p(f["imove"]), _p(f["r"]), _p(f["normal"]),
                                  _p(f["u"]), _p(f["m"]), _p(f["grad_w_bi"]),
                                  _p(f["div_u_bi"]), n, *ll, H, CONW)
            k("bie_rates")(_p(f["imove"]), _p(f["rho"]), _p(f["p"]),
                           _p(f["u"]), _p(f["grad_w_bi"]), _p(f["div_u_bi"]),
                           _p(f["grad_p"]), _p(f["div_u"]), n)
        if o["gp"]:
            k("gp_prelapu")(_p(f["imove"]), _p(f["associations"]),
                            _p(f["normal"]), _p(f["gp_rho"]), _p(f["gp_u"]),
                            _p(f["rho"]), _p(f["u"]), n,
                            ctypes.c_int(o["gp_model"]))
            if o["freeslip_iset"] is not None:
                k("gp_freeslip")(_p(f["iset"]), _p(f["imove"]),
                                 _p(f["associations"]), _p(f["gp_u"]),
                                 _p(f["u"]), n,
                                 ctypes.c_uint(o["freeslip_iset"]))
            k("gp_lapu")(_p(f["imove"]), _p(f["r"]), _p(f["u"]), _p(f["rho"]),
                         _p(f["m"]), _p(f["lap_u"]), n, *ll, H, CONF,
                         ctypes.c_int(o["lap_form"]))
            k("gp_preinteractions")(
                _p(f["iset"]), _p(f["imove"]), _p(f["associations"]),
                _p(f["r"]), _p(f["normal"]), _p(f["dudt"]), _p(f["lap_u"]),
                _p(f["gp_p"]), _p(f["gp_u"]), _p(f["rho"]), _p(f["p"]),
                _p(f["u"]), _p(self.sets["visc_dyn"]), n, _p(self.g))
            k("gp_interactions")(_p(f["imove"]), _p(f["r"]), _p(f["u"]),
                                 _p(f["rho"]), _p(f["m"]), _p(f["p"]),
                                 _p(f["grad_p"]), _p(f["div_u"]), n, *ll, H,
                                 CONF)
            f["u"][...] = f["gp_u_in"]
        if o["bi"]:
            k("bi_lapu")(_p(f["imove"]), _p(f["r"]), _p(f["u"]), _p(f["rho"]),
                         _p(f["m"]), _p(f["lap_u"]), n, *ll, H, CONF,
                         ctypes.c_int(o["lap_form"]))
            k("bi_gradp")(_p(f["iset"]), _p(f["imove"]), _p(f["shepard"]),
                          _p(f["rho"]), _p(f["lap_u"]), _p(f["dudt"]),
                          _p(f["grad_p"]), _p(self.sets["visc_dyn"]),
                          _p(self.sets["refd"]), n, _p(self.g))
            k("bi_interp")(_p(f["imove"]), _p(f["r"]), _p(f["m"]),
                           _p(f["rho"]), _p(f["grad_p"]), _p(f["p"]), n, *ll,
                           H, CONW)
            if self.tait_gamma is None:
                k("bi_interp_shepard")(_p(f["iset"]), _p(f["imove"]),
                                       _p(f["shepard"]), _p(f["rho"]),
                                       _p(f["p"]), _p(self.sets["refd"]), n,
                                       R(self.cs), R(self.p0))
            else:
                k("bi_interp_shepard_tait")(
                    _p(f["iset"]), _p(f["imove"]), _p(f["shepard"]),
                    _p(f["rho"]), _p(f["p"]), _p(self.sets["refd"]), n,
                    R(self.cs), R(self.p0), R(self.tait_gamma))
            k("bi_interactions")(_p(f["imove"]), _p(f["r"]), _p(f["normal"]),
                                 _p(f["u"]), _p(f["rho"]), _p(f["m"]),
                                 _p(f["p"]), _p(f["grad_p"]), _p(f["div_u"]),
                                 n, *ll, H, CONW)
            k("bi_shepard_apply")(_p(f["imove"]), _p(f["shepard"]),
                                  _p(f["grad_p"]), _p(f["lap_u"]),
                                  _p(f["div_u"]), n)
        if self.portals:
            self._portal_stage("forces", ll)
        if o["delta_corr"] == "full":
            k("lapp_full_mls")(_p(f["imove"]), _p(f["mls"]),
                               _p(f["lap_p_corr"]), n)
        if o["delta_corr"]:
            k("lapp_corr")(_p(f["imove"]), _p(f["r"]), _p(f["rho"]),
                           _p(f["m"]), _p(f["lap_p_corr"]), _p(f["lap_p"]), n,
                           *ll, H, CONF)
            if self.portals:
                self._portal_stage("lapp_corr", ll)
        k("rates")(_p(f["iset"]), _p(f["imove"]), _p(f["rho"]),
                   _p(f["grad_p"]), _p(f["lap_u"]), _p(f["div_u"]),
                   _p(f["dudt"]), _p(f["drhodt"]), _p(self.sets["visc_dyn"]),
                   n, _p(self.g))
        if o["delta"]:
            k("delta_rate")(_p(f["iset"]), _p(f["imove"]), _p(f["rho"]),
                            _p(f["lap_p"]), _p(f["drhodt"]),
                            _p(self.sets["refd"]), _p(self.sets["delta"]), n,
                            R(self.dt))
        if self.les is not None:
            self._les(ll)
        if o["energy"]:
            # "cfd Energy kin" + its sum, after "Rates"
            k("energy_kin")(_p(f["energy_kin"]), _p(f["imove"]), _p(f["u"]),
                            _p(f["m"]), n)
            E = self.energy
            E["energy_Kin"] = self._esum(f["energy_kin"])

    def _step_midpoint(self):
        """One step of the preset stack basic + domain + cfd + midpoint
        (+ r / autostop / autorelax) + deltaSPH + BI + variableTimeStep, in
        the tool order the insert rules give it: the delta-SPH lap p reinit
        sits before the inner loop, the other reinits inside it."""
        f, N, o, mp = self.f, self.N, self.opts, self.mp
        k, R = self._k, self._r
        n = ctypes.c_uint(N)
        f32 = np.float32
        k("midpoint_predictor")(_p(f["r"]), _p(f["u"]), _p(f["dudt"]),
                                _p(f["rho"]), _p(f["drhodt"]), _p(f["r_in"]),
                                _p(f["u_in"]), _p(f["dudt_in"]),
                                _p(f["rho_in"]), _p(f["drhodt_in"]), n)
        if o["domain"]:
            k("domain")(_p(f["imove"]), _p(f["r_in"]), _p(f["u_in"]),
                        _p(f["dudt_in"]), _p(f["m"]), n,
                        _p(self.domain_min), _p(self.domain_max))
        self.link_list("r_in")
        self.reorder()
        self._eos()
        k("binormal")(_p(f["normal"]), _p(f["tangent"]), _p(f["binormal"]), n)
        ll = self._ll()
        if o["delta"]:
            f["lap_p"][...] = 0
        self.iter_midpoint = 0
        f["drhodt"][...] = f["drhodt_in"]
        f["dudt"][...] = f["dudt_in"]
        f["r_in"][...] = f["r"]
        f["u_in"][...] = f["u"]
        f["rho_in"][...] = f["rho"]
        iters = 0
        while self.iter_midpoint < mp["iter_max"]:
            f["dudt_in"][...] = f["dudt"]
            f["drhodt_in"][...] = f["drhodt"]
            k("midpoint_advance")(_p(f["imove"]), _p(f["u_in"]), _p(f["u"]),
                                  _p(f["dudt"]), _p(f["rho_in"]),
                                  _p(f["rho"]), _p(f["drhodt"]), n,
                                  R(self.dt))
            if mp["r"]:
                k("midpoint_r")(_p(f["imove"]), _p(f["r_in"]), _p(f["r"]),
                                _p(f["u"]), n, R(self.dt))
            self._eos()
            self._forces(ll, reinit_lap_p=False)
            if mp["residual_max"] is not None:
                res = np.zeros(N, self.dtype)
                k("midpoint_residuals")(_p(f["imove"]), _p(f["m"]),
                                        _p(f["u"]), _p(f["dudt_in"]),
                                        _p(f["dudt"]), _p(f["rho"]),
                                        _p(f["p"]), _p(f["drhodt_in"]),
                                        _p(f["drhodt"]), _p(res), n)
                self.residual = f32(self._sum(res))
                self.residuals.append(float(self.residual))
                # set_scalar tools: double arithmetic of float inputs
                if mp["mindec"] is not None:
                    with np.errstate(divide="ignore", invalid="ignore"):
                        ratio = np.float64(self.residual) / \
                            np.float64(self.residual_prev)
                    if ratio > float(f32(mp["mindec"])):
                        self.relax = f32(float(f32(mp["inc"])) +
                                         (1.0 - float(f32(mp["inc"]))) *
                                         float(self.relax))
                    if self.iter_midpoint == 0:
                        self.relax = self.relax0
                if float(self.residual) < float(f32(mp["residual_max"])):
                    self.iter_midpoint = mp["iter_max"]
                if mp["mindec"] is not None:
                    if self.iter_midpoint == mp["iter_max"]:
                        self.relax = f32(0.0)
                    self.residual_prev = self.residual
            k("midpoint_relax")(_p(f["imove"]), _p(f["dudt_in"]),
                                _p(f["dudt"]), _p(f["drhodt_in"]),
                                _p(f["drhodt"]), n, R(self.relax))
            self.iter_midpoint += 1
            iters += 1
        self.mp_iters.append(iters)
        k("midpoint_corrector")(_p(f["imove"]), _p(f["r_in"]), _p(f["r"]),
                                _p(f["u_in"]), _p(f["u"]), _p(f["dudt"]),
                                _p(f["rho_in"]), _p(f["rho"]),
                                _p(f["drhodt"]), n, R(self.dt))
        self._time_step()

    def _time_step(self):
        """fixed time step, TimeStep.cl + min reduction + assert, t, iter
        (basic.xml:164-169, cfd/variableTimeStep.xml)."""
        f, o = self.f, self.opts
        k, R = self._k, self._r
        n = ctypes.c_uint(self.N)
        self.dt = self.dtype.type(float(self.courant) * float(self.h) /
                                  float(self.cs))
        if o["variable_dt"]:
            k("timestep")(_p(f["imove"]), _p(f["u"]), _p(f["dt_var"]), n,
                          R(self.dt), R(self.dt_min), R(self.courant),
                          R(self.dt_Ma), R(self.h))
            self.dt = self.dtype.type(np.min(f["dt_var"]))
            assert self.dt > 0.0, "dt > 0.0 assertion failed"
        self.t = self.dtype.type(float(self.t) + float(self.dt))
        self.iter += 1

    def _eos(self):
        """basic/EOS.cl (or the Tait law of basic/tait.xml) on the sorted
        fields."""
        f, n, k, R = self.f, ctypes.c_uint(self.N), self._k, self._r
        args = (_p(f["iset"]), _p(f["imove"]), _p(f["rho"]), _p(f["p"]),
                _p(self.sets["refd"]), n, R(self.cs), R(self.p0))
        if self.tait_gamma is None:
            k("eos")(*args)
        else:
            k("tait_eos")(*args, R(self.tait_gamma))

    def step(self):
        if self.opts["midpoint"] is not None:
            return self._step_midpoint()
        if self.periodic is None:
            return self._step()
        L = self.periodic[1]
        lib().orc_set_periodic(ctypes.c_double(float(L[0])),
                               ctypes.c_double(float(L[1])),
                               ctypes.c_double(float(L[2]) if L.size > 2
                                               else 0.0))
        try:
            return self._step()
        finally:
            lib().orc_set_periodic(ctypes.c_double(0.0), ctypes.c_double(0.0),
                                   ctypes.c_double(0.0))

    def _step(self):
        f, N, o = self.f, self.N, self.opts
        k, R, ll = self._k, self._r, None
        n = ctypes.c_uint(N)
        if self.precision:
            getattr(lib(), "orc_predictor64_%dd" % self.dims)(
                _p(f["imove"]), _p(f["r"]), _p(f["dr64"]), _p(f["u"]),
                _p(f["dudt"]), _p(f["rho"]), _p(f["drhodt"]), _p(f["r_in"]),
                _p(f["dr64_in"]), _p(f["u_in"]), _p(f["dudt_in"]),
                _p(f["rho_in"]), _p(f["drhodt_in"]), n,
                ctypes.c_float(float(self.dt)))
        else:
            k("predictor")(_p(f["imove"]), _p(f["r"]), _p(f["u"]),
                           _p(f["dudt"]), _p(f["rho"]), _p(f["drhodt"]),
                           _p(f["r_in"]), _p(f["u_in"]), _p(f["dudt_in"]),
                           _p(f["rho_in"]), _p(f["drhodt_in"]), n, R(self.dt))
        if o["domain"]:
            k("domain")(_p(f["imove"]), _p(f["r_in"]), _p(f["u_in"]),
                        _p(f["dudt_in"]), _p(f["m"]), n,
                        _p(self.domain_min), _p(self.domain_max))
        if self.periodic is not None:
            # "cfd periodic wrap", before the link-list
            k("periodic_wrap")(_p(f["r_in"]), n, _p(self.periodic[0]),
                               _p(self.periodic[1]))
        self.link_list("r_in")
        self.reorder()
        self._eos()
        k("binormal")(_p(f["normal"]), _p(f["tangent"]), _p(f["binormal"]), n)
        if o["gp"]:
            # "Backup associations" + "Sort associations" (GP/Sort.cl)
            f["associations_in"][...] = f["associations"]
            lib().orc_gp_sort(_p(f["associations_in"]), _p(f["associations"]),
                              _p(f["id_sorted"]), n)
        ll = self._ll()
        if self.excitation is not None:
            a, w, g0 = (np.float32(self.excitation[0]),
                        np.float32(self.excitation[1]),
                        np.asarray(self.excitation[2], np.float32))
            a, w, t = float(a), float(w), float(np.float32(self.t))
            self.g[0] = self.dtype.type(np.float32(
                float(g0[0]) + a * w * w * math.sin(w * t)))
            self.g[1] = self.dtype.type(g0[1])
        self._forces(ll)
        if self.elastic_bounce is not None:
            eb = self.elastic_bounce
            k("elastic_bounce")(_p(f["imove"]), _p(f["r"]), _p(f["normal"]),
                                _p(f["u"]), _p(f["dudt"]), n, R(eb["dr"]),
                                R(self.dt), *ll, R(eb["dr_factor"]),
                                R(eb["min_bound_dist"]),
                                R(eb["elastic_factor"]))
        if o["bie"]:
            k("bie_elastic_bounce")(_p(f["imove"]), _p(f["r_in"]),
                                    _p(f["normal"]), _p(f["m"]),
                                    _p(f["u_in"]), _p(f["dudt"]), n,
                                    R(self.dt), *ll, R(o["dr_factor"]),
                                    R(o["min_bound_dist"]))
        if o["energy"]:
            # "cfd Power", its sums and the time integration (set_scalar in
            # double, narrowed), before the corrector
            k("energy_power")(_p(f["energy_dekdt"]), _p(f["energy_depdt"]),
                              _p(f["energy_decdt"]), _p(f["imove"]),
                              _p(f["u"]), _p(f["rho"]), _p(f["m"]),
                              _p(f["p"]), _p(f["dudt"]), _p(f["drhodt"]), n,
                              _p(self.g))
            E = self.energy
            for c in ("k", "p", "c"):
                E["energy_dE%sdt" % c] = self._esum(f["energy_de%sdt" % c])
                e = float(E["energy_E%s" % c]) + \
                    float(self.dt) * float(E["energy_dE%sdt" % c])
                E["energy_E%s" % c] = np.float64(e) if self.precision else \
                    self.dtype.type(e)
        if self.precision:
            getattr(lib(), "orc_corrector64_%dd" % self.dims)(
                _p(f["imove"]), _p(f["r"]), _p(f["dr64"]), _p(f["u"]),
                _p(f["dudt"]), _p(f["rho"]), _p(f["drhodt"]),
                _p(f["dudt_in"]), _p(f["drhodt_in"]), n,
                ctypes.c_float(float(self.dt)))
            # "fluid mass" + "total fluid mass" (double sum)
            self.mass_fp64 = float(np.sum(
                f["m"][f["imove"] == 1].astype(np.float64)))
        else:
            k("corrector")(_p(f["imove"]), _p(f["r"]), _p(f["u"]),
                           _p(f["dudt"]), _p(f["rho"]), _p(f["drhodt"]),
                           _p(f["dudt_in"]), _p(f["drhodt_in"]), n,
                           R(self.dt))
        if o["energy"]:
            # "cfd Energy" and its sums, after the corrector
            k("energy_parts")(_p(f["energy_ek"]), _p(f["energy_ep"]),
                              _p(f["energy_ec"]), _p(f["iset"]),
                              _p(f["imove"]), _p(f["r"]), _p(f["u"]),
                              _p(f["rho"]), _p(f["m"]),
                              _p(self.sets["refd"]), n, _p(self.g),
                              R(self.cs))
            E = self.energy
            for c in ("k", "p", "c"):
                E["energy_E%s_ref" % c] = self._esum(f["energy_e%s" % c])
        if o["bie"]:
            k("bie_pst")(_p(f["imove"]), _p(f["r"]), _p(f["normal"]),
                         _p(f["m"]), _p(f["rho"]), n, *ll, R(self.DIMS),
                         R(o["dr_factor"]))
        if self.portals:
            self._portal_teleport()
        # "fixed time step" set_scalar (basic.xml:164), double -> float
        self._time_step()

    def run(self, nsteps):
        for _ in range(nsteps):
            self.step()
