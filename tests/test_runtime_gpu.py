"""GPU tests of the XML-driven runtime (the product path): case files built
from the presets run through aq_sim_* and are compared with the oracle."""
import os

import numpy as np
import pytest

from aquagpusph_amd import _lib, casefiles, cases
from aquagpusph_amd.sim import Simulation
from oracle import pipeline as op

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def relerr(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s else float(np.abs(a).max())


FIELDS = ("icell", "id_sorted", "id_unsorted", "id", "imove", "r", "u", "rho",
          "p", "dudt", "drhodt", "shepard", "grad_p", "lap_u", "div_u",
          "lap_p", "dt_var")


def _compare(sim, o, tol, fields=FIELDS):
    errs = {}
    for k in fields:
        a = sim.get(k)
        b = o.f[k]
        if a.dtype.kind in "iu":
            np.testing.assert_array_equal(a.view(b.dtype), b, err_msg=k)
            errs[k] = 0.0
        else:
            errs[k] = relerr(a, b)
    assert max(errs.values()) <= tol, errs
    return errs


@pytest.mark.parametrize("dims,nl,steps", [(2, 25, 2), (3, 8, 2),
                                           (2, 25, 30), (3, 8, 15)])
def test_xml_case_matches_oracle(cuda, tmp_path, dims, nl, steps):
    case = cases.dam_break(dims, 0.3 / nl)
    xml = casefiles.write_case(case, str(tmp_path))
    sim = Simulation(xml, dims)
    o = op.OracleSim(case)
    for _ in range(steps):
        o.step()
    sim.step(steps)
    errs = _compare(sim, o, 1e-5 if steps <= 2 else 1e-4)
    print(dims, steps, errs)
    assert abs(float(sim.get("t")) - float(o.t)) <= 1e-6 * float(o.t)
    assert int(sim.get("iter")) == steps
    assert float(sim.get("dt")) == pytest.approx(float(o.dt), rel=1e-5)
    sim.close()


RELAXED = ('    <Settings><Device platform="0" device="0" '
           'compile_flags="-cl-fast-relaxed-math" /></Settings>')


@pytest.mark.parametrize("dims,nl,steps", [(2, 25, 2), (3, 8, 2),
                                           (2, 25, 30), (3, 8, 15)])
def test_relaxed_math_case_matches_oracle(cuda, tmp_path, dims, nl, steps):
    """A case whose Device compile_flags ask for relaxed math runs the
    fused walk with reciprocal division, hardware square root and
    contracted multiply-adds: not bit-identical to the IEEE kernels, within
    the same tolerance of the FP32 oracle (1e-5 of each field's largest
    magnitude after two steps, 1e-4 after 15 or 30)."""
    case = cases.dam_break(dims, 0.3 / nl)
    xml = casefiles.write_case(case, str(tmp_path / "relaxed"),
                               extra_xml=RELAXED)
    xml_ieee = casefiles.write_case(case, str(tmp_path / "ieee"))
    sim = Simulation(xml, dims)
    ieee = Simulation(xml_ieee, dims)
    o = op.OracleSim(case)
    for _ in range(steps):
        o.step()
    sim.step(steps)
    ieee.step(steps)
    errs = _compare(sim, o, 1e-5 if steps <= 2 else 1e-4)
    print(dims, steps, errs)
    assert any(not np.array_equal(sim.get(f), ieee.get(f))
               for f in ("grad_p", "lap_u", "div_u", "lap_p"))
    sim.close()
    ieee.close()


ENERGY_PRESETS = casefiles.PBI_PRESETS + (
    "resources/Presets/cfd/energy_kin.xml", "resources/Presets/cfd/energy.xml")


@pytest.mark.parametrize("dims,nl,steps", [(2, 25, 2), (3, 8, 2),
                                           (2, 25, 20), (3, 8, 10)])
def test_energy_presets_match_oracle(cuda, tmp_path, dims, nl, steps):
    """cfd/energy_kin.xml and cfd/energy.xml on the dam break: the
    per-particle energies and rates, their sums and the time-integrated
    energies against the oracle (1e-5 of each array's largest magnitude
    after two steps, 1e-4 after 10-20; the sums are taken in a different
    order, and the GPU log is ocml's).  The compressibility energy
    m cs^2 (rho0/rho + log(rho/rho0) - 1) cancels to O(d^2) for
    rho = rho0 (1 + d) in FP32 (as in the reference kernel): it is held to
    8 ulp of m cs^2 per particle instead."""
    case = cases.dam_break(dims, 0.3 / nl)
    xml = casefiles.write_case(case, str(tmp_path), presets=ENERGY_PRESETS)
    sim = Simulation(xml, dims)
    o = op.OracleSim(case, energy=True)
    for _ in range(steps):
        o.step()
    sim.step(steps)
    tol = 1e-5 if steps <= 2 else 1e-4
    _compare(sim, o, tol)
    for k in ("energy_kin", "energy_dekdt", "energy_depdt", "energy_decdt",
              "energy_ek", "energy_ep"):
        assert relerr(sim.get(k), o.f[k]) <= tol, k
    mcs2 = o.f["m"].astype(np.float64) * float(o.cs) ** 2
    ec_abs = 8 * 2.0 ** -24 * mcs2
    assert np.all(np.abs(sim.get("energy_ec").astype(np.float64) -
                         o.f["energy_ec"]) <= ec_abs)
    # sums: to tol of the sum of magnitudes they add up; integrated
    # energies: to tol of the energy they integrate
    src = {"energy_Kin": "energy_kin", "energy_dEkdt": "energy_dekdt",
           "energy_dEpdt": "energy_depdt", "energy_dEcdt": "energy_decdt",
           "energy_Ek_ref": "energy_ek", "energy_Ep_ref": "energy_ep",
           "energy_Ec_ref": "energy_ec"}
    for k, v in o.energy.items():
        a, b = float(sim.get(k)), float(v)
        if k == "energy_Ec_ref":
            assert abs(a - b) <= ec_abs.sum(), (k, a, b)
            continue
        if k in src:
            scale = float(np.abs(o.f[src[k]].astype(np.float64)).sum())
        else:
            scale = abs(b) + abs(float(o.energy[k + "_ref"]))
        assert abs(a - b) <= tol * scale + 1e-30, (k, a, b)
    sim.close()


MIDPOINT_PRESETS = (
    "resources/Presets/basic.xml", "resources/Presets/basic/domain.xml",
    "resources/Presets/cfd.xml",
    "resources/Presets/basic/time_scheme/midpoint.xml",
    "resources/Presets/basic/time_scheme/midpoint/autostop.xml",
    "resources/Presets/basic/time_scheme/midpoint/autorelax.xml",
    "resources/Presets/cfd/deltaSPH.xml", "resources/Presets/cfd/BI.xml",
    "resources/Presets/cfd/variableTimeStep.xml")
MIDPOINT_R = "resources/Presets/basic/time_scheme/midpoint/r.xml"


@pytest.mark.parametrize("dims,with_r", [(2, False), (2, True), (3, False)])
def test_midpoint_scheme_matches_oracle(cuda, tmp_path, dims, with_r):
    """basic/time_scheme/midpoint.xml with autostop and autorelax: a
    while/end inner loop with device-evaluated conditions read back by the
    host.  Fields, the inner iteration count, the residual and the adapted
    relaxation factor against the oracle after 4 steps (1e-4 of each
    field's largest magnitude; the residual sum to 1e-3); graph replay
    (one graph per run of tools between control tools) equals eager
    execution bit for bit."""
    case = cases.dam_break(dims, 0.3 / (20 if dims == 2 else 6))
    presets = MIDPOINT_PRESETS
    if with_r:
        i = presets.index("resources/Presets/basic/time_scheme/midpoint.xml")
        presets = presets[:i + 1] + (MIDPOINT_R,) + presets[i + 1:]
    extra = ('    <Variables>\n'
             '        <Variable name="iter_midpoint_max" type="unsigned int" '
             'value="4" />\n'
             '        <Variable name="relax_midpoint" type="float" '
             'value="0.1" />\n'
             '    </Variables>')
    xml = casefiles.write_case(case, str(tmp_path), presets=presets,
                               extra_xml=extra)
    sim = Simulation(xml, dims)
    gsim = Simulation(xml, dims)
    o = op.OracleSim(case, midpoint=dict(iter_max=4, relax=0.1, r=with_r,
                                         residual_max=1e-8, mindec=0.2,
                                         inc=0.25))
    steps = 4
    for _ in range(steps):
        o.step()
    sim.step(steps)
    gsim.step(steps, graph=True)
    _compare(sim, o, 1e-4)
    assert int(sim.get("iter_midpoint")) == o.iter_midpoint
    assert float(sim.get("relax_midpoint")) == pytest.approx(
        float(o.relax), rel=1e-5, abs=1e-7)
    assert float(sim.get("Residual_midpoint")) == pytest.approx(
        float(o.residual), rel=1e-3)
    for f in ("r", "u", "rho", "p", "dudt", "drhodt", "lap_p"):
        np.testing.assert_array_equal(sim.get(f), gsim.get(f), err_msg=f)
    sim.close()
    gsim.close()


GP_FIELDS = FIELDS + ("associations", "gp_icell", "gp_rho", "gp_p", "gp_u")


@pytest.mark.parametrize("dims,nl,steps", [(2, 15, 2), (3, 6, 2),
                                           (2, 15, 25), (3, 6, 10)])
def test_gp_case_matches_oracle(cuda, tmp_path, dims, nl, steps):
    """Ghost-particle walls (cfd/GP.xml) on water at rest in a tank: every
    field, the ghost fields and the associations against the oracle."""
    case = cases.still_tank(dims, 0.15 / nl, bc="GP")
    xml = casefiles.write_case(case, str(tmp_path),
                               presets=casefiles.PGP_PRESETS)
    sim = Simulation(xml, dims)
    o = op.OracleSim(case)
    for _ in range(steps):
        o.step()
    sim.step(steps)
    errs = {}
    for k in GP_FIELDS:
        a = sim.get(k)
        b = o.f[k]
        if a.dtype.kind in "iu":
            np.testing.assert_array_equal(a.view(b.dtype), b, err_msg=k)
        else:
            errs[k] = relerr(a, b)
    print(dims, steps, errs)
    assert max(errs.values()) <= (1e-5 if steps <= 2 else 1e-4), errs
    sim.close()


@pytest.mark.parametrize("dims,nl,steps", [(2, 20, 2), (3, 6, 2),
                                           (2, 20, 25), (3, 6, 10)])
def test_bie_case_matches_oracle(cuda, tmp_path, dims, nl, steps):
    """Conservative boundary integrals (cfd/BIe.xml: wall terms, elastic
    bounce, particle shifting) on the dam break against the oracle."""
    case = cases.dam_break(dims, 0.3 / nl)
    xml = casefiles.write_case(case, str(tmp_path),
                               presets=casefiles.PBIE_PRESETS)
    sim = Simulation(xml, dims)
    o = op.OracleSim(case, bie=True)
    for _ in range(steps):
        o.step()
    sim.step(steps)
    errs = {}
    for k in FIELDS + ("grad_w_bi", "div_u_bi"):
        a = sim.get(k)
        b = o.f[k]
        if a.dtype.kind in "iu":
            np.testing.assert_array_equal(a.view(b.dtype), b, err_msg=k)
        else:
            errs[k] = relerr(a, b)
    print(dims, steps, errs)
    assert max(errs.values()) <= (1e-5 if steps <= 2 else 1e-4), errs
    sim.close()


@pytest.mark.parametrize("dims,corr,steps", [(2, "full", 2), (3, "full", 2),
                                            (2, "simple", 2), (2, "full", 20),
                                            (3, "simple", 10)])
def test_delta_sph_corrections_match_oracle(cuda, tmp_path, dims, corr,
                                            steps):
    """cfd/deltaSPH-full.xml (MLS matrices, full and MLS-renormalised
    corrections) and cfd/deltaSPH-simple.xml against the oracle."""
    case = cases.dam_break(dims, 0.3 / (20 if dims == 2 else 6))
    pres = list(casefiles.PBI_PRESETS)
    pres.insert(5, "resources/Presets/cfd/deltaSPH-%s.xml" % corr)
    xml = casefiles.write_case(case, str(tmp_path), presets=tuple(pres))
    sim = Simulation(xml, dims)
    o = op.OracleSim(case, delta_corr=corr)
    for _ in range(steps):
        o.step()
    sim.step(steps)
    keys = FIELDS + ("lap_p_corr",) + (("mls_fluid",) if corr == "full"
                                         else ())
    errs = {}
    for k in keys:
        a = sim.get(k)
        b = o.f[k]
        if a.dtype.kind in "iu":
            np.testing.assert_array_equal(a.view(b.dtype), b, err_msg=k)
        else:
            errs[k] = relerr(a, b)
    print(dims, corr, steps, errs)
    assert max(errs.values()) <= (1e-5 if steps <= 2 else 1e-4), errs
    sim.close()


PRECISION = "resources/Presets/precision.xml"


@pytest.mark.parametrize("dims,energy,steps", [(2, False, 2), (3, False, 2),
                                               (2, True, 20), (3, True, 10)])
def test_precision_preset_matches_oracle(cuda, tmp_path, dims, energy, steps):
    """Presets/precision.xml on the dam break: compensated walk sums,
    double positions r + dr64, double mass and energy sums, against the
    oracle's precision mode (the tolerances of the plain pipeline; the
    double positions' displacement likewise; the double mass exactly)."""
    case = cases.dam_break(dims, 0.3 / (25 if dims == 2 else 8))
    pres = (ENERGY_PRESETS if energy else casefiles.PBI_PRESETS) + \
        (PRECISION,)
    xml = casefiles.write_case(case, str(tmp_path), presets=pres)
    sim = Simulation(xml, dims)
    o = op.OracleSim(case, energy=energy, precision=True)
    for _ in range(steps):
        o.step()
    sim.step(steps)
    tol = 1e-5 if steps <= 2 else 1e-4
    errs = _compare(sim, o, tol)
    dr = sim.get("dr64")
    assert dr.dtype == np.float64
    assert np.abs(o.f["dr64"]).max() > 0
    # the double positions r + dr64 (dr64 alone is a rounding residual,
    # which flips with the last bit of r): their displacement from the
    # start, to tol
    ids = o.f["id"].astype(np.int64)
    r0 = case["r"].astype(np.float32).astype(np.float64)[ids]
    x_gpu = sim.get("r").astype(np.float64) + dr - r0
    x_orc = o.f["r"].astype(np.float64) + o.f["dr64"] - r0
    errs["r64"] = relerr(x_gpu, x_orc)
    assert errs["r64"] <= tol, errs
    m = float(sim.get("mass_fp64"))
    assert m == pytest.approx(o.mass_fp64, rel=1e-12)
    if energy:
        for k, v in o.energy.items():
            assert isinstance(sim.get(k), np.float64), k
        # (the compressibility part is held as in the plain energy test)
        for k in ("energy_Kin", "energy_Ek_ref", "energy_Ep_ref",
                  "energy_Ek", "energy_Ep"):
            a, b = float(sim.get(k)), float(o.energy[k])
            assert a == pytest.approx(b, rel=tol, abs=1e-9), k
    print(dims, energy, steps, errs)
    sim.close()


def test_precision_walk_differs_from_plain(cuda, tmp_path):
    """The compensated walk is in use: one step of the same case with and
    without the preset gives different (and oracle-matching) rates."""
    case = cases.dam_break(2, 0.3 / 25)
    a = Simulation(casefiles.write_case(case, str(tmp_path / "a")), 2)
    b = Simulation(casefiles.write_case(
        case, str(tmp_path / "b"),
        presets=casefiles.PBI_PRESETS + (PRECISION,)), 2)
    a.step(1)
    b.step(1)
    assert not np.array_equal(a.get("grad_p"), b.get("grad_p"))
    a.close()
    b.close()


BATCH = "resources/Presets/batch.xml"


@pytest.mark.parametrize("steps", [2, 20])
def test_sloshing_tank_matches_oracle(cuda, tmp_path, steps):
    """One excited sloshing tank (tank units, BI walls) with its gravity
    set each step by a set_scalar tool, against the oracle's excitation."""
    b = cases.sloshing_batch(2, res=40, seed=3)
    case = cases.batch_member(b, 1)
    xml = casefiles.write_case(case, str(tmp_path),
                               extra_xml=cases.excitation_xml(case))
    sim = Simulation(xml, 2)
    o = op.OracleSim(case, excitation=case["params"]["excitation"])
    for _ in range(steps):
        o.step()
    sim.step(steps)
    errs = _compare(sim, o, 1e-5 if steps <= 2 else 1e-4)
    np.testing.assert_allclose(sim.get("g"), o.g, rtol=1e-6)
    print(steps, errs)
    sim.close()


@pytest.mark.parametrize("steps", [2, 10])
def test_batch_matches_member_runs(cuda, tmp_path, steps):
    """Four tanks in one batched run (Presets/batch.xml: one link-list, one
    walk, per-set cs, g, excitation, dt, t) against each tank run alone at
    the same place: every field of every tank, its dt and its time."""
    b = cases.sloshing_batch(4, res=40, seed=5)
    xml = casefiles.write_case(b, str(tmp_path / "batch"),
                               presets=casefiles.PBI_PRESETS + (BATCH,))
    sim = Simulation(xml, 2)
    sim.step(steps)
    ids = sim.get("id").astype(np.int64)
    dt_set, t_set = sim.get("dt_set"), sim.get("t_set")
    o = 0
    errs = {}
    for k, (n, sc) in enumerate(b["sets"]):
        case = cases.batch_member(b, k)
        one = Simulation(casefiles.write_case(
            case, str(tmp_path / ("tank%d" % k)),
            extra_xml=cases.excitation_xml(case)), 2)
        one.step(steps)
        oid = one.get("id").astype(np.int64)
        mine = (ids >= o) & (ids < o + n)
        order = np.argsort(ids[mine])
        order1 = np.argsort(oid)
        for f in ("r", "u", "rho", "p", "dudt", "drhodt", "imove"):
            a = sim.get(f)[mine][order]
            c = one.get(f)[order1]
            if a.dtype.kind in "iu":
                np.testing.assert_array_equal(a, c, err_msg=f)
            else:
                errs[(k, f)] = relerr(a, c)
        assert float(dt_set[k]) == pytest.approx(float(one.get("dt")),
                                                 rel=1e-5)
        assert float(t_set[k]) == pytest.approx(float(one.get("t")),
                                                rel=1e-5)
        one.close()
        o += n
    print(steps, max(errs.values()))
    assert max(errs.values()) <= (1e-5 if steps <= 2 else 1e-4), errs
    # the tanks advance with different steps
    assert len(set(np.asarray(dt_set).tolist())) > 1
    sim.close()


@pytest.mark.parametrize("dims", [2, 3])
def test_autotune_keeps_state_and_results(cuda, tmp_path, dims):
    """The walk autotuner times its configurations on the live case but
    restores the state after each; every configuration walks in the same
    order, so the tuned run equals the untuned one bit for bit."""
    case = cases.dam_break(dims, 0.3 / (20 if dims == 2 else 6))
    xml = casefiles.write_case(case, str(tmp_path))
    a = Simulation(xml, dims)
    b = Simulation(xml, dims)
    a.step(2, graph=True)
    b.step(2, graph=True)
    rep = a.autotune(2)
    assert len(rep) == 9 and rep[-1].startswith("chosen:"), rep
    for f in ("r", "u", "rho", "p", "id", "t"):
        np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)
    a.step(3, graph=True)
    b.step(3, graph=True)
    for f in ("r", "u", "rho", "p", "id", "shepard", "t", "dt"):
        np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)
    a.close()
    b.close()


def _sheared(case):
    """A smooth sheared velocity on the fluid of a dam break."""
    dims = case["dims"]
    fl = case["imove"] == 1
    x, y = case["r"][fl, 0], case["r"][fl, dims - 1]
    case["u"][fl, 0] = 0.3 * np.sin(2 * np.pi * y / 0.3)
    case["u"][fl, dims - 1] = 0.3 * np.sin(2 * np.pi * x / 0.6)
    return case


@pytest.mark.parametrize("dims,nl,steps", [(2, 25, 2), (3, 8, 2),
                                           (2, 25, 10), (3, 8, 6)])
def test_moving_flow_matches_oracle(cuda, tmp_path, dims, nl, steps):
    """The P-BI dam break with the fluid moving from the start (a sheared
    velocity field), so that the boundary-integral wall terms, the
    delta-SPH diffusion and the variable time step all see a flow: every
    field within 1e-5 of its largest magnitude after 2 and 6-10 steps.
    The wall terms' atan / acospi / pow run the same float sequence on the
    device and in the oracle, and every sum is taken in the reference's
    order, so the only differences left are none: the fields are
    bit-identical."""
    case = _sheared(cases.dam_break(dims, 0.3 / nl))
    xml = casefiles.write_case(case, str(tmp_path))
    sim = Simulation(xml, dims)
    o = op.OracleSim(case)
    for _ in range(steps):
        o.step()
    sim.step(steps)
    errs = _compare(sim, o, 1e-5)
    print(dims, steps, errs)
    assert max(errs.values()) == 0.0, errs
    sim.close()


@pytest.mark.parametrize("dims,nl,steps", [(2, 25, 2), (3, 8, 2),
                                           (2, 25, 20), (3, 8, 10)])
def test_sensors_match_oracle(cuda, tmp_path, dims, nl, steps):
    """Pressure sensors as a second particle set (imove = 0, m = 0), laid
    out as in the reference's 3D dam break: on the wall inside the column
    (the boundary-integral Shepard's on-wall branch), on the floor, inside
    the fluid and on a dry wall.  cfd/Sensors.cl and
    SensorsRenormalization.cl run every step: u, rho, p and shepard at the
    sensors against the oracle (1e-5 of each field's largest sensor value
    after two steps, 1e-4 after 10-20), every other field as in the plain
    case; the wall sensors read the hydrostatic pressure to 12 % (the
    deepest one, 0.07 H above the floor, sits within h of the floor where
    the fluid-only interpolation is truncated: 8.5 % in 3D at H / dr = 8)."""
    pos, nrm = cases.dam_break_sensors(dims)
    case = cases.with_sensors(cases.dam_break(dims, 0.3 / nl), pos, nrm)
    xml = casefiles.write_case(case, str(tmp_path))
    sim = Simulation(xml, dims)
    o = op.OracleSim(case)
    for _ in range(steps):
        o.step()
    sim.step(steps)
    tol = 1e-5 if steps <= 2 else 1e-4
    _compare(sim, o, tol)
    sen = o.f["imove"] == 0
    assert sen.sum() == case["n_sensors"]
    np.testing.assert_array_equal(sim.get("imove") == 0, sen)
    errs = {k: relerr(sim.get(k)[sen], o.f[k][sen])
            for k in ("u", "rho", "p", "shepard")}
    print(dims, steps, errs)
    assert max(errs.values()) <= tol, errs
    # wall sensors (the first four by id) read about rho g (H - z)
    ids = o.f["id"][sen].astype(np.int64)
    first = np.argsort(ids)[:4]
    z = o.f["r"][sen][first, dims - 1].astype(np.float64)
    hyd = 998.0 * 9.81 * (0.3 - z)
    p = sim.get("p")[sen][first].astype(np.float64)
    assert np.all(np.abs(p - hyd) <= 0.12 * hyd), (p, hyd)
    sim.close()


BOUNCE_PRESETS = casefiles.PBI_PRESETS + (
    "resources/Presets/cfd/elasticBounce.xml",)


@pytest.mark.parametrize("dims,nl,steps,mbd", [(2, 25, 2, None),
                                               (3, 8, 2, None),
                                               (2, 25, 2, 0.6),
                                               (3, 8, 2, 0.6),
                                               (2, 25, 12, 0.6),
                                               (3, 8, 8, 0.6)])
def test_elastic_bounce_matches_oracle(cuda, tmp_path, dims, nl, steps, mbd):
    """cfd/elasticBounce.xml (cfd/Boundary/ElasticBounce.cl) on the sheared
    dam break, with the reference's definitions (nothing reaches the wall
    in two steps) and with __MIN_BOUND_DIST__ = 0.6 (every wall particle
    approaching an element bounces): every field against the oracle to 1e-5
    after two steps and 1e-4 after 8-12, and with the larger distance the
    bounce is seen to act (u differs from the run without it)."""
    case = _sheared(cases.dam_break(dims, 0.3 / nl))
    dr = case["params"]["dr"]
    extra = ('    <Variables><Variable name="dr" type="float" value="%s" />'
             '</Variables>' % casefiles._f(dr))
    eb = {"dr": float(np.float32(dr))}
    if mbd is not None:
        extra += ('\n    <Definitions><Define name="__MIN_BOUND_DIST__" '
                  'value="%s" evaluate="false"/></Definitions>' % mbd)
        eb["min_bound_dist"] = float(np.float32(mbd))
    xml = casefiles.write_case(case, str(tmp_path), presets=BOUNCE_PRESETS,
                               extra_xml=extra)
    sim = Simulation(xml, dims)
    o = op.OracleSim(case, elastic_bounce=eb)
    for _ in range(steps):
        o.step()
    sim.step(steps)
    errs = _compare(sim, o, 1e-5 if steps <= 2 else 1e-4)
    print(dims, steps, mbd, errs)
    if mbd is not None:
        plain = op.OracleSim(case)
        for _ in range(steps):
            plain.step()
        assert not np.array_equal(plain.f["u"], o.f["u"])
    sim.close()


TAIT_PRESETS = casefiles.PBI_PRESETS + ("resources/Presets/basic/tait.xml",)


@pytest.mark.parametrize("dims,nl,steps", [(2, 25, 2), (3, 8, 2),
                                           (2, 25, 12), (3, 8, 8)])
def test_tait_eos_matches_oracle(cuda, tmp_path, dims, nl, steps):
    """basic/tait.xml (gamma = 7) on the sheared P-BI dam break (initial
    density from the hydrostatic pressure through the same law): the fluid
    pressure by the same binary powering on both sides, the wall density
    by the Tait inverse (powf against the host pow: a few ulp); every field
    within 1e-5 of its largest magnitude after two steps and 1e-4 after
    8-12, and the pressure differs from the linear EOS run."""
    case = _sheared(cases.dam_break(dims, 0.3 / nl, tait_gamma=7.0))
    xml = casefiles.write_case(case, str(tmp_path), presets=TAIT_PRESETS)
    sim = Simulation(xml, dims)
    o = op.OracleSim(case, tait_gamma=7.0)
    for _ in range(steps):
        o.step()
    sim.step(steps)
    errs = _compare(sim, o, 1e-5 if steps <= 2 else 1e-4)
    print(dims, steps, errs)
    lin = op.OracleSim(case)
    for _ in range(steps):
        lin.step()
    assert not np.array_equal(lin.f["p"], o.f["p"])
    sim.close()


@pytest.mark.parametrize("dims,mode,steps", [(2, "smag", 2), (3, "smag", 2),
                                           (2, "mlp", 2), (3, "mlp", 2),
                                           (2, "mlp", 10)])
def test_les_presets_match_oracle(cuda, tmp_path, dims, mode, steps):
    """cfd/LES-smagorinsky.xml and cfd/LES.xml (the BF16 network with given
    parameters) on the dam break with a smooth sheared initial velocity:
    velocity gradient and eddy viscosity against the oracle to 1e-5 (the
    network's to 2e-3: its FP32 accumulation order differs from the
    oracle's), every field to 1e-5 after two steps and 1e-4 after ten (the
    Smagorinsky closure is bit-identical; the network's eddy viscosity
    carries its accumulation-order differences into u, dudt and div_u at
    the 1e-7 level)."""
    from oracle import mlp as omlp
    case = _sheared(cases.dam_break(dims, 0.3 / (20 if dims == 2 else 6)))
    pres = casefiles.PBI_PRESETS + (
        "resources/Presets/cfd/LES%s.xml"
        % ("" if mode == "mlp" else "-smagorinsky"),)
    xml = casefiles.write_case(case, str(tmp_path), presets=pres)
    sim = Simulation(xml, dims)
    if mode == "mlp":
        P = omlp.init(5)
        P[omlp.OFF_W3:omlp.OFF_B3] *= 0.1
        P[omlp.OFF_B3] = 0.05
        sim.set("les_params", P)
        les = {"params": P, "hdr": case["params"]["hfac"]}
    else:
        les = {"Cs": 0.15}
    o = op.OracleSim(case, les=les)
    for _ in range(steps):
        o.step()
    sim.step(steps)
    errs = _compare(sim, o, 1e-5 if steps <= 2 else 1e-4)
    for k in ("grad_u", "nu_t"):
        errs[k] = relerr(sim.get(k), o.f[k])
    print(dims, mode, steps, errs)
    tol = 1e-5 if steps <= 2 else 1e-4
    assert errs["nu_t"] <= (2e-3 if mode == "mlp" else tol), errs
    assert errs["grad_u"] <= tol, errs
    assert np.abs(o.f["nu_t"]).max() > 0
    sim.close()


@pytest.mark.parametrize("dims,n,steps", [(2, 32, 2), (3, 16, 2),
                                          (2, 32, 20), (3, 16, 10)])
def test_periodic_box_matches_oracle(cuda, tmp_path, dims, n, steps):
    """Presets/cfd/periodic.xml: decaying turbulence in a box periodic
    along every axis (no walls), the positions wrapped before the
    link-list, the fused cell walk over wrapped cells at shifted positions;
    every field bit-identical to the oracle's periodic walk (pinned against
    explicit periodic images in test_oracle_periodic_cpu.py) after 2 and
    10-20 steps, with particles crossing the box faces."""
    case = cases.periodic_box(dims, n)
    # a uniform drift carries particles through the faces: the lattice is
    # moved to within 1e-4 of the faces the drift crosses
    drift = np.array([3.0, -2.0, 1.0][:dims])
    off = np.sign(drift) * (0.5 / n - 1e-4)
    case["r"][:, :dims] = (case["r"][:, :dims] + off) % 1.0
    case["u"][:, :dims] += drift
    xml = casefiles.write_case(case, str(tmp_path),
                               presets=casefiles.PPER_PRESETS)
    sim = Simulation(xml, dims)
    o = op.OracleSim(case, bi=False, domain=False,
                     periodic=case["params"]["periodic"])
    for _ in range(steps):
        o.step()
    sim.step(steps)
    errs = _compare(sim, o, 1e-5 if steps <= 2 else 1e-4)
    print(dims, n, steps, errs)
    assert max(errs.values()) == 0.0, errs
    x = sim.get("r")[:, :dims]
    assert np.all((x >= 0.0) & (x <= 1.0))
    ids = o.f["id"].astype(np.int64)
    moved = np.abs(o.f["r"][:, :dims] - case["r"][ids, :dims]) > 0.5
    assert moved.any(), float(o.t)
    sim.close()


@pytest.mark.parametrize("dims,mode,steps", [(3, "smag", 2), (3, "mlp", 2),
                                             (3, "mlp", 8), (2, "smag", 6)])
def test_periodic_les_matches_oracle(cuda, tmp_path, dims, mode, steps):
    """The LES closure (Presets/cfd/LES.xml with a given network, or
    LES-smagorinsky.xml) in the periodic box: the cell-parallel LES
    gradient and viscous term over the wrapped cells, against the oracle:
    with Smagorinsky every field bit-identical; with the network (its eddy
    viscosity to 2e-3, its accumulation order differing from the oracle's)
    the velocity gradient and the state to 1e-5 after two steps and 1e-4
    after 8, the rates to 1e-4 (the pressure Laplacian to 1e-3)."""
    from oracle import mlp as omlp
    case = cases.periodic_box(dims, 16 if dims == 3 else 32)
    pres = casefiles.PPER_PRESETS + (
        "resources/Presets/cfd/LES%s.xml"
        % ("" if mode == "mlp" else "-smagorinsky"),)
    xml = casefiles.write_case(case, str(tmp_path), presets=pres,
                               extra_xml=MORRIS)
    sim = Simulation(xml, dims)
    if mode == "mlp":
        P = omlp.init(5)
        P[omlp.OFF_W3:omlp.OFF_B3] *= 0.1
        P[omlp.OFF_B3] = 0.05
        sim.set("les_params", P)
        les = {"params": P, "hdr": case["params"]["hfac"]}
    else:
        les = {"Cs": 0.15}
    o = op.OracleSim(case, bi=False, domain=False, lap_form=2, les=les,
                     periodic=case["params"]["periodic"])
    for _ in range(steps):
        o.step()
    sim.step(steps)
    tol = 1e-5 if steps <= 2 else 1e-4
    # the network's eddy viscosity is not bit-exact (its accumulation
    # order): the rates it feeds (dudt, and through the flow div u) are held
    # to 1e-4, the state (r, u, rho, p) to tol; lap_p sums the differences
    # p_j - p_i of a pressure with the background p0 in it, which amplify
    # the pressure's relative error by p / |p_j - p_i|: held to 1e-3
    if mode == "mlp":
        errs = _compare(sim, o, 1e-4,
                        fields=tuple(f for f in FIELDS if f != "lap_p"))
        errs["lap_p"] = relerr(sim.get("lap_p"), o.f["lap_p"])
        assert errs["lap_p"] <= 1e-3, errs
    else:
        errs = _compare(sim, o, tol)
    for k in ("r", "u", "rho", "p"):
        assert errs[k] <= tol, errs
    for k in ("grad_u", "nu_t"):
        errs[k] = relerr(sim.get(k), o.f[k])
    print(dims, mode, steps, errs)
    assert errs["nu_t"] <= (2e-3 if mode == "mlp" else tol), errs
    assert errs["grad_u"] <= tol, errs
    if mode == "smag":
        assert max(errs.values()) == 0.0, errs
    sim.close()


def test_periodic_box_too_small_is_refused(cuda, tmp_path):
    """A periodic axis shorter than 3 cells of support * h cannot be walked
    (wrapped neighbour cells would repeat): the link-list flags it and the
    step reports it."""
    case = cases.periodic_box(3, 8)  # 2 cells of 2h per axis
    xml = casefiles.write_case(case, str(tmp_path),
                               presets=casefiles.PPER_PRESETS)
    sim = Simulation(xml, 3)
    with pytest.raises(_lib.AquaError, match="periodic"):
        sim.step(1)
    sim.close()


PORTAL_BASE = ("resources/Presets/basic.xml",
               "resources/Presets/basic/time_scheme/improved_euler.xml",
               "resources/Presets/basic/domain.xml",
               "resources/Presets/cfd.xml",
               "resources/Presets/cfd/deltaSPH.xml")
MORRIS = ('    <Definitions><Define name="__LAP_FORMULATION__" '
          'value="__LAP_MORRIS__" evaluate="false"/></Definitions>')


@pytest.mark.parametrize("steps", [2, 20])
def test_taylor_green_portals_match_oracle(cuda, tmp_path, steps):
    """The periodic Taylor-Green vortices of the reference's
    examples/2D/taylor_green: eight cfd/portal.xml includes (sides and
    corners) with their set_scalar plane tools, Morris Laplacian,
    delta-SPH.  A uniform drift carries particles through the portals
    within the 20 steps (teleports)."""
    case = cases.taylor_green(n=32, hfac=4.0, drift=(5.0, 4.0))
    ports = cases.taylor_green_portals(case["params"]["L"])
    xml = casefiles.write_case(case, str(tmp_path), presets=PORTAL_BASE,
                               extra_xml=MORRIS, portals=ports)
    sim = Simulation(xml, 2)
    o = op.OracleSim(case, bi=False, variable_dt=False, lap_form=2,
                     portals=ports)
    for _ in range(steps):
        o.step()
    sim.step(steps)
    # (fixed time step: no dt_var)
    errs = _compare(sim, o, 1e-5 if steps <= 2 else 1e-4,
                    [k for k in FIELDS if k != "dt_var"])
    print(steps, errs)
    np.testing.assert_array_equal(sim.get("imirrored"), o.f["imirrored"])
    ids = o.f["id"].astype(np.int64)
    jumped = np.abs(o.f["r"][:, 0] - case["r"][ids, 0]) > \
        0.5 * case["params"]["L"]
    assert jumped.any() == (steps > 2)
    sim.close()


@pytest.mark.parametrize("dims,wall,corr,steps", [
    (2, True, None, 2), (3, True, None, 2), (2, False, "full", 2),
    (3, False, "full", 2), (2, True, "full", 10), (3, True, "simple", 4)])
def test_periodic_slab_portals_match_oracle(cuda, tmp_path, dims, wall, corr,
                                            steps):
    """One portal pair along x over a sheared block: with a boundary-integral
    wall (cfd portal BI), and with cfd/deltaSPH-full.xml (the portal MLS,
    full and lapp_corr stages) or deltaSPH-simple."""
    n = (16, 8) if dims == 2 else (12, 6, 6)
    case = cases.periodic_slab(dims, n, wall=wall)
    ports = cases.slab_portals(case["params"]["L"], dims)
    pres = list(PORTAL_BASE)
    if corr:
        pres.append("resources/Presets/cfd/deltaSPH-%s.xml" % corr)
    if wall:
        pres.append("resources/Presets/cfd/BI.xml")
    pres.append("resources/Presets/cfd/variableTimeStep.xml")
    xml = casefiles.write_case(case, str(tmp_path), presets=tuple(pres),
                               portals=ports)
    sim = Simulation(xml, dims)
    o = op.OracleSim(case, bi=wall, delta_corr=corr, portals=ports)
    for _ in range(steps):
        o.step()
    sim.step(steps)
    keys = FIELDS + (("lap_p_corr",) if corr else ()) + \
        (("mls_fluid",) if corr == "full" else ())
    errs = {}
    for k in keys:
        a, b = sim.get(k), o.f[k]
        if a.dtype.kind in "iu":
            np.testing.assert_array_equal(a.view(b.dtype), b, err_msg=k)
        else:
            errs[k] = relerr(a, b)
    print(dims, wall, corr, steps, errs)
    assert max(errs.values()) <= (1e-5 if steps <= 2 else 1e-4), errs
    sim.close()


@pytest.mark.parametrize("dims", [2, 3])
def test_graph_replay_equals_eager(cuda, tmp_path, dims):
    """A captured step replayed from a HIP graph gives bit-identical state,
    replay after replay."""
    case = cases.dam_break(dims, 0.3 / (20 if dims == 2 else 6))
    xml = casefiles.write_case(case, str(tmp_path))
    a = Simulation(xml, dims)
    b = Simulation(xml, dims)
    for k in range(3):
        a.step(4)
        b.step(4, graph=True)
        for f in ("r", "u", "rho", "p", "id", "icell", "shepard"):
            np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)
        assert float(a.get("t")) == float(b.get("t"))
    a.close()
    b.close()


@pytest.mark.parametrize("dims,bi,walk,bi_walk,no_fusion", [
    (2, True, "thread", "thread", "0"),
    (3, True, "thread", "thread", "0"),
    (3, False, "thread", "thread", "0"),
    (2, True, "cells", "cells", "0"),
    (3, True, "cells", "cells", "0"),
    (3, True, "cells", "cells", "1"),
    (2, True, "cells", "thread", "0"),
    (3, True, "cells", "thread", "0"),
    (3, True, "cells", "thread", "1"),
    (2, True, "cells", "list", "0"),
    (3, True, "cells", "list", "0"),
    (3, True, "cells", "list", "1")])
def test_walks_are_bit_identical(cuda, tmp_path, dims, bi, walk, bi_walk,
                                 no_fusion, monkeypatch):
    """The fused neighbour walk (Shepard + interactions + lap p + near-wall
    flags), the cell-parallel pair kernels (fused walk, boundary-integral
    element kernels) and the per-particle kernels run over compacted particle
    lists with cell-kind skipping reproduce the separate per-particle
    kernels, which follow the reference walk order, bit for bit."""
    case = cases.dam_break(dims, 0.3 / (20 if dims == 2 else 6))
    presets = casefiles.PBI_PRESETS if bi else tuple(
        p for p in casefiles.PBI_PRESETS if not p.endswith("BI.xml"))
    xml = casefiles.write_case(case, str(tmp_path), presets=presets)
    monkeypatch.setenv("AQ_NO_FUSION", "1")
    monkeypatch.setenv("AQ_WALK", "thread")
    monkeypatch.setenv("AQ_BI_WALK", "thread")
    a = Simulation(xml, dims)
    monkeypatch.setenv("AQ_NO_FUSION", no_fusion)
    monkeypatch.setenv("AQ_WALK", walk)
    monkeypatch.setenv("AQ_BI_WALK", bi_walk)
    b = Simulation(xml, dims)
    for _ in range(3):
        a.step(3)
        b.step(3, graph=True)
        for f in ("r", "u", "rho", "p", "shepard", "grad_p", "lap_u",
                  "div_u", "lap_p", "dudt", "drhodt", "id", "iset", "imove",
                  "m", "normal", "tangent", "binormal", "dudt_in",
                  "drhodt_in"):
            np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)


@pytest.mark.parametrize("dims,div", [(2, 60), (3, 16)])
def test_owner_cell_list_matches_all_cells(cuda, tmp_path, dims, div,
                                           monkeypatch):
    """The boundary-integral element kernels walking the owner-cell list
    (cells with wall elements or sensors; default) equal the same kernels
    visiting every non-empty cell (AQ_BI_WALK=cells) bit for bit on a dam
    break with sensors, whose wall cells give candidate lists of several
    segments."""
    pos, nrm = cases.dam_break_sensors(dims)
    case = cases.with_sensors(cases.dam_break(dims, 0.3 / div), pos, nrm)
    xml = casefiles.write_case(case, str(tmp_path))
    monkeypatch.setenv("AQ_BI_WALK", "cells")
    a = Simulation(xml, dims)
    monkeypatch.setenv("AQ_BI_WALK", "list")
    b = Simulation(xml, dims)
    for _ in range(2):
        a.step(3)
        b.step(3, graph=True)
        for f in ("r", "u", "rho", "p", "shepard", "grad_p", "lap_u",
                  "div_u", "dudt", "drhodt"):
            np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)
    a.close()
    b.close()


@pytest.mark.parametrize("dims", [2, 3])
def test_bi_chain_beside_walk_is_bit_identical(cuda, tmp_path, dims,
                                               monkeypatch):
    """The boundary-integral element chain run on a side stream beside the
    fused walk (AQ_BI_OVERLAP=1, joined before BI Interactions) gives the
    serial pipeline's results bit for bit, eagerly and replayed from a
    graph."""
    pos, nrm = cases.dam_break_sensors(dims)
    case = cases.with_sensors(
        cases.dam_break(dims, 0.3 / (40 if dims == 2 else 12)), pos, nrm)
    xml = casefiles.write_case(case, str(tmp_path))
    monkeypatch.delenv("AQ_BI_OVERLAP", raising=False)
    a = Simulation(xml, dims)
    monkeypatch.setenv("AQ_BI_OVERLAP", "1")
    b = Simulation(xml, dims)
    for graph in (False, True, True):
        a.step(3)
        b.step(3, graph=graph)
        for f in ("r", "u", "rho", "p", "shepard", "grad_p", "lap_u",
                  "div_u", "dudt", "drhodt"):
            np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)
    a.close()
    b.close()


def _dense_case(dims, n, seed):
    """Random fluid cloud with ~150 particles per link-list cell (more than
    a wavefront per cell, several candidate segments per neighbourhood)."""
    rng = np.random.default_rng(seed)
    vs = 2 if dims == 2 else 4
    r = np.zeros((n, vs))
    r[:, :dims] = rng.uniform(0, 1, (n, dims))
    per_cell = 150.0
    h = 0.5 * (per_cell / n) ** (1.0 / dims)
    u = np.zeros((n, vs))
    u[:, :dims] = rng.normal(0, 0.05, (n, dims))
    g = [0.0] * dims
    g[-1] = -9.81
    params = dict(h=h, cs=20.0, courant=0.1, g=g, refd=[1000.0],
                  visc_dyn=[1e-3], delta=[0.1], dt_min=0.0, dt_Ma=0.1,
                  domain_min=[-1.0] * dims, domain_max=[2.0] * dims)
    return dict(dims=dims, N=n, r=r, u=u,
                rho=1000.0 + rng.normal(0, 2.0, n),
                m=np.full(n, 1000.0 / n), imove=np.ones(n, np.int32),
                params=params)


@pytest.mark.parametrize("dims,n", [(2, 4000), (3, 12000)])
def test_step_head_and_tail_passes_are_bit_identical(cuda, tmp_path, dims, n,
                                                     monkeypatch):
    """The predictor + Domain pass, the reorder + EOS + Binormal pass and
    the Rates + delta-SPH rate + corrector pass give the state of their
    separate launches (AQ_NO_FUSION=1) bit for bit, with particles leaving
    the domain (their imove, mass and backed-up fields set by the Domain
    part)."""
    case = _dense_case(dims, n, 5)
    case["params"]["domain_max"] = [0.97] * dims
    presets = tuple(p for p in casefiles.PBI_PRESETS
                    if not p.endswith("BI.xml"))
    xml = casefiles.write_case(case, str(tmp_path), presets=presets)
    monkeypatch.setenv("AQ_NO_FUSION", "1")
    a = Simulation(xml, dims)
    monkeypatch.setenv("AQ_NO_FUSION", "0")
    b = Simulation(xml, dims)
    for graph in (False, True):
        a.step(2)
        b.step(2, graph=graph)
        for f in ("imove", "m", "r", "u", "rho", "r_in", "u_in", "dudt_in",
                  "rho_in", "drhodt_in", "dudt", "drhodt", "grad_p",
                  "lap_u", "div_u", "lap_p", "id", "iset", "p", "normal",
                  "tangent", "binormal", "id_in", "imove_in", "m_in"):
            np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)
    assert np.count_nonzero(b.get("imove") == -256) > 0
    a.close()
    b.close()


@pytest.mark.parametrize("dims,n", [(2, 6000), (3, 20000)])
def test_cell_walk_dense_cells(cuda, tmp_path, dims, n, monkeypatch):
    """Cells holding more particles than a wavefront and neighbourhoods of
    several candidate segments: the cell-parallel walk, the per-particle
    fused walk and the separate kernels agree bit for bit."""
    case = _dense_case(dims, n, 11)
    presets = tuple(p for p in casefiles.PBI_PRESETS
                    if not p.endswith("BI.xml"))
    xml = casefiles.write_case(case, str(tmp_path), presets=presets)
    sims = []
    for fuse, walk, segw in (("1", "cells", "16"), ("0", "cells", "16"),
                             ("0", "cells", "8"), ("0", "thread", "16")):
        monkeypatch.setenv("AQ_NO_FUSION", fuse)
        monkeypatch.setenv("AQ_WALK", walk)
        monkeypatch.setenv("AQ_WALK_SEGW", segw)
        sims.append(Simulation(xml, dims))
    for s in sims:
        s.step(2)
    for f in ("shepard", "grad_p", "lap_u", "div_u", "lap_p", "r", "u"):
        for k in (1, 2, 3):
            np.testing.assert_array_equal(sims[0].get(f), sims[k].get(f), f)


def test_upload_path_equals_file_path(cuda, tmp_path):
    """ParticlesSet without Load + Simulation.set == FastASCII Load."""
    case = cases.dam_break(3, 0.3 / 6)
    x1 = casefiles.write_case(case, str(tmp_path / "a"))
    x2 = casefiles.write_case(case, str(tmp_path / "b"), load=False)
    a = Simulation(x1, 3)
    b = Simulation(x2, 3)
    casefiles.upload(b, case)
    for f in ("r", "normal", "u", "rho", "m", "imove"):
        np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)
    a.step(3)
    b.step(3)
    np.testing.assert_array_equal(a.get("r"), b.get("r"))


def _fixture_case(tmp_path, body, dims, data, fields):
    np.savetxt(tmp_path / "particles.dat", data, fmt="%.17g")
    (tmp_path / "main.xml").write_text("""<?xml version="1.0" ?>
<sphInput>
%s
    <ParticlesSet n="%d">
        <Load fields="%s" file="particles.dat" format="FastASCII"/>
    </ParticlesSet>
</sphInput>""" % (body, data.shape[0], fields))
    return str(tmp_path / "main.xml")


def test_reference_radixsort_case(cuda, tmp_path):
    """The reference tests/3D/RadixSort main.xml tool set on its fixture."""
    d = np.loadtxt(os.path.join(GOLDEN, "radixsort_3d_particles.dat"))
    xml = _fixture_case(tmp_path, """
    <Variables>
        <Variable length="N" name="f" type="size_t*" />
        <Variable length="N" name="f_orig" type="size_t*" />
        <Variable length="N" name="perms" type="size_t*" />
        <Variable length="N" name="inv_perms" type="size_t*" />
        <Variable length="1" name="h" type="float" value="0.1"/>
    </Variables>
    <Tools>
        <Tool action="add" name="copy f" once="true" type="copy" in="f" out="f_orig" />
        <Tool action="add" name="sort f" once="true" type="radix-sort" in="f" perm="perms" inv_perm="inv_perms" />
        <Tool action="add" name="iter += 1" type="set_scalar" in="iter" value="iter + 1"/>
    </Tools>""", 3, d, "r, f")
    sim = Simulation(xml, 3)
    sim.step(2)  # once-tools run only in the first step
    f, f_orig = sim.get("f"), sim.get("f_orig")
    perms, inv = sim.get("perms"), sim.get("inv_perms")
    assert list(f) == sorted(f_orig)
    assert list(f_orig) != sorted(f_orig)
    np.testing.assert_array_equal(f, f_orig[perms])
    np.testing.assert_array_equal(f[inv], f_orig)
    assert int(sim.get("iter")) == 2


def test_batched_set_and_copy_runs(cuda, tmp_path):
    """Consecutive set tools, consecutive copy tools and consecutive
    set_scalar tools (dummies between them aside) run as one launch each;
    a run ends where a tool meets an array the run writes (a copy of a
    copy, a second set of one array), so every result is the one the tools
    give in order."""
    d = np.loadtxt(os.path.join(GOLDEN, "radixsort_3d_particles.dat"))
    xml = _fixture_case(tmp_path, """
    <Variables>
        <Variable length="N" name="f" type="size_t*" />
        <Variable length="N" name="a" type="size_t*" />
        <Variable length="N" name="b" type="size_t*" />
        <Variable length="N" name="c" type="float*" />
        <Variable length="N" name="e" type="vec*" />
        <Variable length="N" name="r1" type="vec*" />
        <Variable length="N" name="r2" type="vec*" />
        <Variable length="1" name="h" type="float" value="0.1"/>
    </Variables>
    <Tools>
        <Tool action="add" name="copy f a" type="copy" in="f" out="a" />
        <Tool action="add" name="copy a b" type="copy" in="a" out="b" />
        <Tool action="add" name="copy r r1" type="copy" in="r" out="r1" />
        <Tool action="add" name="set c" type="set" in="c" value="7.5" />
        <Tool action="add" name="set e one" type="set" in="e" value="VEC_ONE" />
        <Tool action="add" name="set e zero" type="set" in="e" value="VEC_ZERO" />
        <Tool action="add" name="set r2" type="set" in="r2" value="VEC_ONE" />
        <Tool action="add" name="copy r1 r2" type="copy" in="r1" out="r2" />
        <Tool action="add" name="iter += 1" type="set_scalar" in="iter" value="iter + 1"/>
        <Tool action="add" name="marker" type="dummy"/>
        <Tool action="add" name="h = 2 h" type="set_scalar" in="h" value="2 * h"/>
        <Tool action="add" name="h = h + iter" type="set_scalar" in="h" value="h + iter"/>
    </Tools>""", 3, d, "r, f")
    sim = Simulation(xml, 3)
    for graph in (False, True):
        sim.step(1, graph=graph)
        f = sim.get("f")
        np.testing.assert_array_equal(sim.get("a"), f)
        np.testing.assert_array_equal(sim.get("b"), f)
        np.testing.assert_array_equal(sim.get("r1"), sim.get("r"))
        np.testing.assert_array_equal(sim.get("r2"), sim.get("r"))
        assert np.all(sim.get("c") == np.float32(7.5))
        assert np.all(sim.get("e") == 0.0)
    # three set_scalar tools (one run, a dummy between), each reading what
    # the one before stored: h = 2 (2 0.1 + 1) + 2 after two steps
    assert int(sim.get("iter")) == 2
    np.testing.assert_allclose(float(sim.get("h")), 2 * (2 * 0.1 + 1) + 2,
                               rtol=1e-6)
    sim.close()


@pytest.mark.parametrize("dims", [2, 3])
def test_batched_launches_are_bit_identical(cuda, tmp_path, dims,
                                            monkeypatch):
    """The P-BI dam break with its set, copy and set_scalar runs batched and
    the sort stages' gathers in one pass each gives the state of one launch
    per tool and field (AQ_NO_BATCH=1) bit for bit, the sort's backup and
    unsorted fields and the time scalars included."""
    pos, nrm = cases.dam_break_sensors(dims)
    case = cases.with_sensors(
        cases.dam_break(dims, 0.3 / (40 if dims == 2 else 12)), pos, nrm)
    xml = casefiles.write_case(case, str(tmp_path))
    monkeypatch.setenv("AQ_NO_BATCH", "1")
    a = Simulation(xml, dims)
    monkeypatch.delenv("AQ_NO_BATCH")
    b = Simulation(xml, dims)
    for graph in (False, True):
        a.step(3)
        b.step(3, graph=graph)
        for f in ("r", "u", "rho", "p", "m", "id", "iset", "imove", "normal",
                  "tangent", "id_in", "iset_in", "imove_in", "normal_in",
                  "tangent_in", "m_in", "dudt", "dudt_in", "drhodt",
                  "drhodt_in", "shepard", "grad_p", "lap_u", "div_u",
                  "lap_p", "t", "dt", "iter"):
            np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)
    a.close()
    b.close()


def test_reference_reduction_case(cuda, tmp_path):
    """The reference tests/3D/Reduction tool set: vec sum (tol 1e-2) and max
    (exact) of its fixture."""
    d = np.loadtxt(os.path.join(GOLDEN, "reduction_3d_particles.dat"))
    xml = _fixture_case(tmp_path, """
    <Variables>
        <Variable length="1" name="r_sum" type="vec" />
        <Variable length="1" name="r_max" type="vec" />
        <Variable length="1" name="h" type="float" value="0.1"/>
    </Variables>
    <Tools>
        <Tool action="add" name="sum up" once="true" type="reduction" in="r" out="r_sum" null="VEC_ZERO">
            c = a + b;
        </Tool>
        <Tool action="add" name="get max" once="true" type="reduction" in="r" out="r_max" null="VEC_NEG_INFINITY">
            c = max(a, b);
        </Tool>
    </Tools>""", 3, d, "r")
    sim = Simulation(xml, 3)
    sim.step(1)
    r32 = d.astype(np.float32)
    np.testing.assert_allclose(sim.get("r_sum")[:3], d[:, :3].sum(0),
                               atol=1e-2)
    np.testing.assert_array_equal(sim.get("r_max")[:3], r32[:, :3].max(0))


def test_assert_tool_failure_is_reported(cuda, tmp_path):
    case = cases.dam_break(2, 0.3 / 10)
    xml = casefiles.write_case(case, str(tmp_path), extra_xml="""
    <Tools>
        <Tool action="insert" before="End" name="never" type="assert" condition="iter &lt; 2"/>
    </Tools>""")
    sim = Simulation(xml, 2)
    sim.step(1)
    with pytest.raises(_lib.AquaError, match="never"):
        sim.step(2)


def test_neighbour_count_preset(cuda, tmp_path):
    """basic/neighs_limit.xml: n_neighs bit-exact vs the oracle count, and the
    guard trips when the limit is too small."""
    import ctypes
    case = cases.dam_break(3, 0.3 / 6)
    # neighs_limit is evaluated from hfac: include the preset after the
    # case variables are registered
    xml = casefiles.write_case(case, str(tmp_path), extra_xml=(
        '    <Include file="resources/Presets/basic/neighs_limit.xml" '
        'when="end"/>'))
    sim = Simulation(xml, 3)
    sim.step(1)
    o = op.OracleSim(case)
    o.step()
    n = np.zeros(o.N, np.uint32)
    lim = int(sim.get("neighs_limit"))
    op.lib().orc_neighs_3d(op._p(o.f["imove"]), op._p(n), ctypes.c_uint(lim),
                           ctypes.c_uint(o.N), *o._ll())
    np.testing.assert_array_equal(sim.get("n_neighs"), n)
    assert int(sim.get("max_neighs")) == n.max()
    sim.set("neighs_limit", np.uint32(n.max()))
    with pytest.raises(_lib.AquaError, match="check number of neighbours"):
        sim.step(1)


@pytest.mark.parametrize("graph", [False, True])
def test_failure_names_its_step_and_freezes_the_state(cuda, tmp_path, graph):
    """A failure flagged in the middle of step(10) (here an assert that
    fails once iter reaches 5) is reported with the step it happened in,
    and the predictor / corrector leave the particles as they were at the
    end of that step for the remaining steps of the call."""
    case = cases.dam_break(2, 0.3 / 10)
    xml = casefiles.write_case(case, str(tmp_path), extra_xml="""
    <Tools>
        <Tool action="insert" before="End" name="stop at five" type="assert" condition="iter &lt; 5"/>
    </Tools>""")
    sim = Simulation(xml, 2)
    with pytest.raises(_lib.AquaError, match="iter = 5"):
        sim.step(10, graph=graph)
    ref = Simulation(casefiles.write_case(case, str(tmp_path / "ref")), 2)
    ref.step(5)

    def by_id(s, f):
        i = s.get("id").astype(np.int64)
        v = s.get(f)
        o = np.zeros_like(v)
        o[i] = v
        return o

    # the wall elements' density is not integrated: the boundary-integral
    # interpolation recomputes it every step from the (kept) fluid state,
    # so only the fluid's is held to the state at the failed step
    fluid = by_id(ref, "imove") == 1
    for f in ("r", "u"):
        np.testing.assert_array_equal(by_id(sim, f), by_id(ref, f),
                                      err_msg=f)
    np.testing.assert_array_equal(by_id(sim, "rho")[fluid],
                                  by_id(ref, "rho")[fluid])
    sim.close()
    ref.close()
