"""Test harness: the per-kernel C ABI driven step by step on the GPU.

Replays exactly the tool order of oracle/pipeline.py (the resolved
basic + improved_euler + domain + cfd + deltaSPH + BI + variableTimeStep
preset stack), calling every kernel through the C ABI with device buffers
held in torch tensors.  Used by the parity tests to compare each kernel and
whole steps against the oracle.
"""
import ctypes

import numpy as np
import torch

from aquagpusph_amd import _lib
from oracle.pipeline import definition

P = ctypes.c_void_p


def dp(t):
    return P(t.data_ptr())


class GpuSim:
    def __init__(self, case, domain=True, bi=True, delta=True,
                 variable_dt=True, lap_form=1, capacity=None):
        dev = torch.device("cuda")
        self.dev = dev
        self.dims = case["dims"]
        self.vs = 2 if self.dims == 2 else 4
        self.opts = dict(domain=domain, bi=bi, delta=delta,
                         variable_dt=variable_dt, lap_form=lap_form)
        N = self.N = int(case["N"])
        f = {}

        def vec(v):
            a = np.zeros((N, self.vs), np.float32) if v is None else \
                np.ascontiguousarray(v, np.float32).reshape(N, self.vs)
            return torch.from_numpy(a.copy()).to(dev)

        def sca(v):
            a = np.zeros(N, np.float32) if v is None else \
                np.ascontiguousarray(v, np.float32)
            return torch.from_numpy(a.copy()).to(dev)
        for k in ("r", "normal", "tangent", "binormal", "u", "dudt"):
            f[k] = vec(case.get(k))
        for k in ("rho", "drhodt", "m", "p"):
            f[k] = sca(case.get(k))
        f["imove"] = torch.from_numpy(
            np.ascontiguousarray(case["imove"], np.int32).copy()).to(dev)
        f["iset"] = torch.from_numpy(np.ascontiguousarray(
            case.get("iset", np.zeros(N)), np.uint32).view(np.int32).copy()
        ).to(dev)
        f["id"] = torch.arange(N, dtype=torch.int32, device=dev)
        for k in ("r", "u", "dudt", "normal", "tangent"):
            f[k + "_in"] = torch.zeros_like(f[k])
        for k in ("rho", "drhodt", "m", "id", "iset", "imove"):
            f[k + "_in"] = torch.zeros_like(f[k])
        f["grad_p"] = torch.zeros((N, self.vs), device=dev)
        f["lap_u"] = torch.zeros((N, self.vs), device=dev)
        for k in ("div_u", "shepard", "lap_p", "dt_var"):
            f[k] = torch.zeros(N, device=dev)
        for k in ("icell", "id_sorted", "id_unsorted"):
            f[k] = torch.zeros(N, dtype=torch.int32, device=dev)
        self.f = f
        p = case["params"]

        def dev_f(vals):
            return torch.tensor(np.asarray(vals, np.float32), device=dev)
        self.sets = {k: dev_f(p[k]) for k in ("refd", "visc_dyn", "delta")}
        g = np.zeros(self.vs, np.float32)
        g[:self.dims] = p["g"][:self.dims]
        self.g = dev_f(g)
        self.h = np.float32(p["h"])
        self.cs_h = np.float32(p["cs"])
        self.courant_h = np.float32(p["courant"])
        self.s = {k: dev_f([v]) for k, v in dict(
            h=self.h, cs=p["cs"], p0=p.get("p0", 0.0), courant=p["courant"],
            dt_Ma=p.get("dt_Ma", 0.1), dt_min=p.get("dt_min", 0.0),
            support=2.0, dt=p.get("dt", 0.0), t=0.0).items()}
        dmin = np.zeros(self.vs, np.float32)
        dmax = np.zeros(self.vs, np.float32)
        if domain:
            dmin[:self.dims] = p["domain_min"][:self.dims]
            dmax[:self.dims] = p["domain_max"][:self.dims]
        self.dmin, self.dmax = dev_f(dmin), dev_f(dmax)
        self.defs = _lib.Defs(definition(self.h),
                              definition(1.0 / float(self.h) ** self.dims),
                              definition(1.0 / float(self.h) **
                                         (self.dims + 2)),
                              definition(self.dims), lap_form)
        self.r_min = torch.zeros(self.vs, device=dev)
        self.r_max = torch.zeros(self.vs, device=dev)
        self.n_cells = torch.zeros(4, dtype=torch.int32, device=dev)
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)
        if capacity is None:
            capacity = self._capacity(case)
        self.capacity = int(capacity)
        self.ihoc = torch.zeros(self.capacity, dtype=torch.int32, device=dev)
        L = _lib.lib()
        self.ws = torch.empty(int(L.aq_linklist_workspace(N)),
                              dtype=torch.uint8, device=dev)
        self.rws = torch.empty(int(L.aq_reduce_workspace(N, 1)),
                               dtype=torch.uint8, device=dev)
        self.iter = 0
        self.stream = P(torch.cuda.current_stream().cuda_stream)

    def _capacity(self, case):
        """ihoc capacity: cells of the domain box (or positions) + margin."""
        p = case["params"]
        cl = float(np.float32(2.0) * np.float32(p["h"]))
        if self.opts["domain"]:
            lo = np.asarray(p["domain_min"][:self.dims], np.float64)
            hi = np.asarray(p["domain_max"][:self.dims], np.float64)
        else:
            r = np.asarray(case["r"])[:, :self.dims]
            lo, hi = r.min(0), r.max(0)
            ext = hi - lo
            lo, hi = lo - ext, hi + ext
        n = np.floor((hi - lo) / cl) + 8
        return int(np.prod(n))

    def ll_args(self):
        f = self.f
        return (dp(f["icell"]), dp(self.ihoc), dp(self.n_cells))

    def link_list(self, r_name="r_in"):
        f = self.f
        _lib.call("aq_linklist", self.dims, dp(f[r_name]), self.N,
                  dp(self.r_min), dp(self.r_max), dp(self.s["support"]),
                  dp(self.s["h"]), dp(self.n_cells), dp(f["icell"]),
                  dp(f["id_unsorted"]), dp(f["id_sorted"]), dp(self.ihoc),
                  self.capacity, dp(self.status), dp(self.ws), self.stream)

    def reorder(self):
        f = self.f
        for k in ("id", "iset", "imove", "normal", "tangent", "m"):
            f[k + "_in"].copy_(f[k])
        perm = dp(f["id_unsorted"])

        def pm(src, dst):
            esz = src.element_size() * (src.shape[1] if src.dim() > 1 else 1)
            _lib.call("aq_permute", dp(src), dp(dst), esz, perm, self.N,
                      self.stream)
        pm(f["id_in"], f["id"])
        pm(f["iset_in"], f["iset"])
        pm(f["imove_in"], f["imove"])
        pm(f["r_in"], f["r"])
        pm(f["normal_in"], f["normal"])
        pm(f["tangent_in"], f["tangent"])
        pm(f["rho_in"], f["rho"])
        pm(f["m_in"], f["m"])
        pm(f["u_in"], f["u"])
        pm(f["dudt"], f["dudt_in"])
        pm(f["drhodt"], f["drhodt_in"])

    def step(self):
        f, N, o, d, st = self.f, self.N, self.opts, self.dims, self.stream
        c = _lib.call
        s = self.s
        c("aq_predictor", d, dp(f["imove"]), dp(f["r"]), dp(f["u"]),
          dp(f["dudt"]), dp(f["rho"]), dp(f["drhodt"]), dp(f["r_in"]),
          dp(f["u_in"]), dp(f["dudt_in"]), dp(f["rho_in"]),
          dp(f["drhodt_in"]), N, dp(s["dt"]), st)
        if o["domain"]:
            c("aq_domain", d, dp(f["imove"]), dp(f["r_in"]), dp(f["u_in"]),
              dp(f["dudt_in"]), dp(f["m"]), N, dp(self.dmin), dp(self.dmax),
              st)
        self.link_list("r_in")
        self.reorder()
        c("aq_eos", dp(f["iset"]), dp(f["imove"]), dp(f["rho"]), dp(f["p"]),
          dp(self.sets["refd"]), N, dp(s["cs"]), dp(s["p0"]), st)
        c("aq_binormal", d, dp(f["normal"]), dp(f["tangent"]),
          dp(f["binormal"]), N, st)
        for k in ("lap_p", "lap_u", "div_u", "grad_p", "shepard"):
            f[k].zero_()
        ll = self.ll_args()
        if o["bi"]:
            c("aq_bi_shepard_compute", d, dp(f["imove"]), dp(f["r"]),
              dp(f["normal"]), dp(f["tangent"]), dp(f["binormal"]),
              dp(f["m"]), dp(f["shepard"]), N, *ll, self.defs, st)
        else:
            c("aq_shepard", d, dp(f["imove"]), dp(f["r"]), dp(f["rho"]),
              dp(f["m"]), dp(f["shepard"]), N, *ll, self.defs, 1, st)
        c("aq_interactions", d, dp(f["imove"]), dp(f["r"]), dp(f["u"]),
          dp(f["rho"]), dp(f["m"]), dp(f["p"]), dp(f["grad_p"]),
          dp(f["lap_u"]), dp(f["div_u"]), N, *ll, self.defs, st)
        c("aq_sensors", d, dp(f["imove"]), dp(f["r"]), dp(f["m"]),
          dp(f["u"]), dp(f["rho"]), dp(f["p"]), N, *ll, self.defs, st)
        c("aq_sensors_renorm", d, dp(f["imove"]), dp(f["shepard"]),
          dp(f["u"]), dp(f["rho"]), dp(f["p"]), N, st)
        if o["delta"]:
            c("aq_lapp", d, dp(f["imove"]), dp(f["r"]), dp(f["rho"]),
              dp(f["m"]), dp(f["p"]), dp(f["lap_p"]), N, *ll, self.defs, st)
        if o["bi"]:
            c("aq_bi_lapu", d, dp(f["imove"]), dp(f["r"]), dp(f["u"]),
              dp(f["rho"]), dp(f["m"]), dp(f["lap_u"]), N, *ll, self.defs,
              st)
            c("aq_bi_gradp", d, dp(f["iset"]), dp(f["imove"]),
              dp(f["shepard"]), dp(f["rho"]), dp(f["lap_u"]), dp(f["dudt"]),
              dp(f["grad_p"]), dp(self.sets["visc_dyn"]),
              dp(self.sets["refd"]), N, dp(self.g), st)
            c("aq_bi_interp", d, dp(f["imove"]), dp(f["r"]), dp(f["m"]),
              dp(f["rho"]), dp(f["grad_p"]), dp(f["p"]), N, *ll, self.defs,
              st)
            c("aq_bi_interp_shepard", dp(f["iset"]), dp(f["imove"]),
              dp(f["shepard"]), dp(f["rho"]), dp(f["p"]),
              dp(self.sets["refd"]), N, dp(s["cs"]), dp(s["p0"]), st)
            c("aq_bi_interactions", d, dp(f["imove"]), dp(f["r"]),
              dp(f["normal"]), dp(f["u"]), dp(f["rho"]), dp(f["m"]),
              dp(f["p"]), dp(f["grad_p"]), dp(f["div_u"]), N, *ll,
              self.defs, st)
            c("aq_bi_shepard_apply", d, dp(f["imove"]), dp(f["shepard"]),
              dp(f["grad_p"]), dp(f["lap_u"]), dp(f["div_u"]), N, st)
        c("aq_rates", d, dp(f["iset"]), dp(f["imove"]), dp(f["rho"]),
          dp(f["grad_p"]), dp(f["lap_u"]), dp(f["div_u"]), dp(f["dudt"]),
          dp(f["drhodt"]), dp(self.sets["visc_dyn"]), N, dp(self.g), st)
        if o["delta"]:
            c("aq_delta_rate", dp(f["iset"]), dp(f["imove"]), dp(f["rho"]),
              dp(f["lap_p"]), dp(f["drhodt"]), dp(self.sets["refd"]),
              dp(self.sets["delta"]), N, dp(s["dt"]), st)
        c("aq_corrector", d, dp(f["imove"]), dp(f["r"]), dp(f["u"]),
          dp(f["dudt"]), dp(f["rho"]), dp(f["drhodt"]), dp(f["dudt_in"]),
          dp(f["drhodt_in"]), N, dp(s["dt"]), st)
        s["dt"].fill_(float(np.float32(float(self.courant_h) *
                                       float(self.h) / float(self.cs_h))))
        if o["variable_dt"]:
            c("aq_timestep", d, dp(f["imove"]), dp(f["u"]), dp(f["dt_var"]),
              N, dp(s["dt"]), dp(s["dt_min"]), dp(s["courant"]),
              dp(s["dt_Ma"]), dp(s["h"]), st)
            ident = (ctypes.c_float * 1)(float("inf"))
            c("aq_reduce", 0, 0, 1, dp(f["dt_var"]), N, dp(s["dt"]),
              ident, dp(self.rws), st)
        s["t"].copy_((s["t"].double() + s["dt"].double()).float())
        self.iter += 1

    def run(self, n):
        for _ in range(n):
            self.step()

    def host(self, name):
        """Field as numpy in the oracle's layout/dtypes."""
        t = self.f[name].cpu().numpy()
        if name in ("icell", "id_sorted", "id_unsorted", "id", "iset",
                    "id_in", "iset_in"):
            t = t.view(np.uint32)
        return t
