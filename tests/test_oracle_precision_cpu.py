"""Pins the precision mode of the oracle (Presets/precision.xml):
compensated walk sums and double positions.

- The compensated sum (Kahan-Babuska / Neumaier) s + c of float32 terms is
  the exact sum (math.fsum) to within u|S| plus the rounding of the
  compensation's own plain sum, n u sum|lo_i| <= n u^2 sum|t_i| (t_i the
  partial sums), on random data and on sums where the plain float32 sum
  loses every small term.
- The compensated interactions and lap p keep the float32 pair terms of
  the plain kernels: against the float64 oracle their error is the terms'
  rounding only, no larger than the plain kernels' anywhere and smaller
  overall.
- Double positions: with a constant acceleration the improved-Euler
  position update is exact in real arithmetic; the double position r + dr64
  reproduces the float64 oracle's free fall to double rounding while float
  positions drift by their float rounding per step."""
import ctypes
import math

import numpy as np
import pytest

from oracle import pipeline as op
from oracle.pipeline import _p

from aquagpusph_amd import cases

U = 2.0 ** -24


def _nsum(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(2, np.float32)
    op.lib().orc_nsum_f32(_p(x), ctypes.c_uint(len(x)), _p(out))
    return float(out[0]) + float(out[1]), out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_compensated_sum_bound(seed):
    rng = np.random.default_rng(seed)
    for x in (rng.normal(0, 1, 5000),
              rng.normal(0, 1, 5000) * 10.0 ** rng.integers(-6, 6, 5000),
              np.concatenate([[1e8], rng.uniform(0, 1, 4000), [-1e8]]),
              np.array([1.0] + [1e-8] * 20000)):
        x = x.astype(np.float32)
        exact = math.fsum(float(v) for v in x)
        got, _ = _nsum(x)
        partial = np.abs(np.cumsum(x.astype(np.float64))).sum()
        bound = U * abs(exact) + 2 * len(x) * U * U * partial
        assert abs(got - exact) <= bound + 1e-45, (got - exact, bound)
        # the plain float32 sum, in the same order, for comparison
        plain = np.float32(0)
        for v in x:
            plain = np.float32(plain + v)
        assert abs(got - exact) <= abs(float(plain) - exact) + 1e-45
    # 1 + 20000 * 1e-8: the plain sum keeps 1, the compensated one 1.0002
    x = np.array([1.0] + [1e-8] * 20000, np.float32)
    got, _ = _nsum(x)
    assert abs(got - 1.0002) < 1e-6


@pytest.mark.parametrize("dims", [2, 3])
def test_compensated_walk_sums(dims):
    """Plain and compensated float32 kernels and the float64 kernel on the
    same (float32) state: the dam break after one step, with a random
    velocity field."""
    case = cases.dam_break(dims, 0.3 / (24 if dims == 2 else 8))
    o = op.OracleSim(case)
    o.step()
    f, N = o.f, ctypes.c_uint(o.N)
    # a sheared velocity field (the free fall alone gives lap u = 0)
    rng = np.random.default_rng(7)
    f["u"] = (f["u"] + rng.normal(0, 0.1, f["u"].shape)).astype(np.float32)
    lib = op.lib()
    sfx = "_%dd" % dims
    H, CONF = float(o.H), float(o.CONF)
    out = {}
    for name, dt, kern in (("plain", np.float32, "f32"),
                           ("kahan", np.float32, "f32"),
                           ("f64", np.float64, "f64")):
        g = {k: np.ascontiguousarray(f[k], dt) for k in
             ("r", "u", "rho", "m", "p")}
        acc = {k: np.zeros_like(np.ascontiguousarray(f[k], dt)) for k in
               ("grad_p", "lap_u", "div_u", "lap_p")}
        R = ctypes.c_float if dt == np.float32 else ctypes.c_double
        kh = "_kahan" if name == "kahan" else ""
        getattr(lib, "orc_interactions%s_%s%s" % (kh, kern, sfx))(
            _p(f["imove"]), _p(g["r"]), _p(g["u"]), _p(g["rho"]), _p(g["m"]),
            _p(g["p"]), _p(acc["grad_p"]), _p(acc["lap_u"]),
            _p(acc["div_u"]), N, *o._ll(), R(H), R(CONF), ctypes.c_int(1))
        getattr(lib, "orc_lapp%s_%s%s" % (kh, kern, sfx))(
            _p(f["imove"]), _p(g["r"]), _p(g["rho"]), _p(g["m"]), _p(g["p"]),
            _p(acc["lap_p"]), N, *o._ll(), R(H), R(CONF))
        out[name] = acc
    for key in ("grad_p", "lap_u", "div_u", "lap_p"):
        ref = out["f64"][key]
        e_plain = np.abs(out["plain"][key] - ref)
        e_kahan = np.abs(out["kahan"][key] - ref)
        scale = np.abs(ref).max()
        # the terms' own float32 rounding remains; the summation error goes
        assert e_kahan.sum() < e_plain.sum(), key
        assert e_kahan.max() <= e_plain.max() + 1e-6 * scale, key
        assert e_kahan.max() <= 1e-5 * scale, key
        # and the plain kernel is the very same sum without compensation
        assert not np.array_equal(out["plain"][key], out["kahan"][key])


@pytest.mark.parametrize("dims", [2, 3])
def test_double_positions_track_float64_free_fall(dims):
    """Free fall (no interactions: a lattice spaced 2.5 supports apart,
    100 m from the origin so that float positions round coarsely) under g
    over many small steps."""
    side = 4 if dims == 3 else 8
    g1 = np.arange(side) * 0.05 + 100.0
    pts = np.stack([a.ravel() for a in
                    np.meshgrid(*([g1] * dims), indexing="ij")], axis=1)
    n = pts.shape[0]
    vs = 2 if dims == 2 else 4
    r = np.zeros((n, vs))
    r[:, :dims] = pts
    r = r.astype(np.float32).astype(np.float64)  # the same start for all
    case = dict(dims=dims, N=n, r=r, u=np.zeros((n, vs)),
                rho=np.full(n, 1000.0), m=np.full(n, 1.0),
                imove=np.ones(n, np.int32),
                params=dict(h=0.01, cs=10.0, courant=0.1,
                            g=[0.0] * (dims - 1) + [-9.81], refd=[1000.0],
                            visc_dyn=[1e-3], delta=[0.0], p0=0.0))
    kw = dict(domain=False, bi=False, delta=False, variable_dt=False)
    sims = {"plain": op.OracleSim(case, **kw),
            "prec": op.OracleSim(case, precision=True, **kw),
            "f64": op.OracleSim(case, dtype=np.float64, **kw)}
    for o in sims.values():
        o.run(200)

    def pos(k):
        o = sims[k]
        a = o.f["r"].astype(np.float64)
        if k == "prec":
            a = a + o.f["dr64"]
        out = np.zeros_like(a)
        out[o.f["id"].astype(np.int64)] = a
        return out

    ref = pos("f64")
    e_plain = np.abs(pos("plain") - ref).max()
    e_prec = np.abs(pos("prec") - ref).max()
    assert e_plain > 1e-6
    assert e_prec < 1e-3 * e_plain, (e_prec, e_plain)
