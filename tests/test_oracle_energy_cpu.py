"""Pins the energy restatement of the oracle (cfd/Energy/EnergyKin.cl,
cfd/Energy/Energy.cl power/energy) with closed forms and with the calculus
that ties each rate to its energy: the power kernel's rates are the time
derivatives of the energy kernel's parts along u, du/dt and drho/dt (with
the linear EOS p = cs^2 (rho - rho0))."""
import ctypes

import numpy as np
import pytest

from oracle import pipeline as op
from oracle.pipeline import _p

from aquagpusph_amd import cases


def _k(name, dims, dt="f64"):
    return getattr(op.lib(), "orc_%s_%s_%dd" % (name, dt, dims))


def _fields(dims, n=64, seed=3):
    rng = np.random.default_rng(seed)
    vs = 2 if dims == 2 else 4
    f = dict(r=np.zeros((n, vs)), u=np.zeros((n, vs)), dudt=np.zeros((n, vs)))
    for k in f:
        f[k][:, :dims] = rng.normal(0, 1, (n, dims))
    f["rho"] = 1000.0 * (1 + rng.uniform(-0.02, 0.02, n))
    f["drhodt"] = rng.normal(0, 5.0, n)
    f["m"] = rng.uniform(0.5, 1.5, n)
    f["imove"] = np.ones(n, np.int32)
    f["imove"][::7] = -3
    f["iset"] = np.zeros(n, np.uint32)
    g = np.zeros(vs)
    g[dims - 1] = -9.81
    f["g"] = g
    return f


def _parts(dims, f, cs, refd):
    n = len(f["m"])
    ek, ep, ec = np.zeros(n), np.zeros(n), np.zeros(n)
    _k("energy_parts", dims)(_p(ek), _p(ep), _p(ec), _p(f["iset"]),
                             _p(f["imove"]), _p(f["r"]), _p(f["u"]),
                             _p(f["rho"]), _p(f["m"]), _p(refd),
                             ctypes.c_uint(n), _p(f["g"]),
                             ctypes.c_double(cs))
    return ek, ep, ec


@pytest.mark.parametrize("dims", [2, 3])
def test_energy_closed_forms(dims):
    f = _fields(dims)
    n = len(f["m"])
    fl = f["imove"] == 1
    ek = np.full(n, -1.0)
    _k("energy_kin", dims)(_p(ek), _p(f["imove"]), _p(f["u"]), _p(f["m"]),
                           ctypes.c_uint(n))
    u = f["u"][:, :dims]
    np.testing.assert_allclose(ek[fl], 0.5 * f["m"][fl] * (u[fl] ** 2).sum(1),
                               rtol=1e-14)
    assert np.all(ek[~fl] == 0)
    cs, refd = 15.0, np.array([1000.0])
    ek2, ep, ec = _parts(dims, f, cs, refd)
    np.testing.assert_array_equal(ek2, ek)
    z = f["r"][:, dims - 1]
    np.testing.assert_allclose(ep[fl], f["m"][fl] * 9.81 * z[fl], rtol=1e-14)
    # compressibility energy: zero at rho0, m cs^2 (x - 1 - log x) >= 0 with
    # x = rho0 / rho, ~ m cs^2 d^2 / 2 for rho = rho0 (1 + d)
    x = refd[0] / f["rho"]
    np.testing.assert_allclose(ec[fl], f["m"][fl] * cs ** 2 *
                               (x[fl] - 1 - np.log(x[fl])), rtol=1e-6)
    assert np.all(ec[fl] >= 0) and np.all(ec[~fl] == 0)
    f["rho"][:] = 1000.0
    assert np.all(_parts(dims, f, cs, refd)[2] == 0)
    d = 1e-3
    f["rho"][:] = 1000.0 * (1 + d)
    ec = _parts(dims, f, cs, refd)[2]
    np.testing.assert_allclose(ec[fl], f["m"][fl] * cs ** 2 * d ** 2 / 2,
                               rtol=2e-3)


@pytest.mark.parametrize("dims", [2, 3])
def test_power_is_the_rate_of_the_energy(dims):
    """dEk/dt, dEp/dt, dEc/dt of the power kernel against centred finite
    differences of the energy kernel along (r, u, rho) + eps (u, du/dt,
    drho/dt), with p from the linear EOS."""
    f = _fields(dims)
    n = len(f["m"])
    cs, refd = 15.0, np.array([1000.0])
    f["p"] = cs ** 2 * (f["rho"] - refd[0])
    dek, dep, dec = np.zeros(n), np.zeros(n), np.zeros(n)
    _k("energy_power", dims)(_p(dek), _p(dep), _p(dec), _p(f["imove"]),
                             _p(f["u"]), _p(f["rho"]), _p(f["m"]),
                             _p(f["p"]), _p(f["dudt"]), _p(f["drhodt"]),
                             ctypes.c_uint(n), _p(f["g"]))
    eps = 1e-6

    def at(s):
        g = dict(f)
        g["r"] = f["r"] + s * f["u"]
        g["u"] = f["u"] + s * f["dudt"]
        g["rho"] = f["rho"] + s * f["drhodt"]
        return _parts(dims, g, cs, refd)

    hi, lo = at(eps), at(-eps)
    for rate, a, b in zip((dek, dep, dec), hi, lo):
        fd = (a - b) / (2 * eps)
        np.testing.assert_allclose(rate, fd, rtol=1e-6,
                                   atol=1e-6 * np.abs(fd).max())
    assert np.all(dek[f["imove"] != 1] == 0)


def test_integrated_energy_tracks_the_energy():
    """Over a dam-break run (FP64), the time integral of each rate follows
    the change of the matching energy sum.  The rates are taken at the
    predicted state and integrated with the step's dt, one stage off the
    energies, so over the first 20 steps of a flow that starts from rest
    they agree to ~10 %, not to round-off."""
    case = cases.dam_break(2, 0.3 / 12)
    o = op.OracleSim(case, dtype=np.float64, energy=True)
    o.step()  # dt = 0 on the first step
    E0 = {c: float(o.energy["energy_E%s_ref" % c]) - float(
        o.energy["energy_E%s" % c]) for c in "kpc"}
    for _ in range(20):
        o.step()
    E = o.energy
    for c in "kp":
        integ = float(E["energy_E%s" % c])
        diff = float(E["energy_E%s_ref" % c]) - E0[c]
        assert abs(diff - integ) < 0.15 * abs(diff), (c, diff, integ)
    # the column collapses: potential energy down, kinetic energy up, and
    # the kinetic sum after the rates equals the one after the corrector to
    # the scheme's accuracy
    assert float(E["energy_Ep"]) < 0 < float(E["energy_Ek"])
    assert abs(float(E["energy_Kin"]) - float(E["energy_Ek_ref"])) < \
        0.2 * float(E["energy_Ek_ref"])
