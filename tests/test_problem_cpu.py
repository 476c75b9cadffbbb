"""The tool list that case files resolve to, parsed on the host
(aq_xml_resolve), against the include/insert rules of the reference's
State.cpp: prefixed includes, insert before a glob (every match), try_insert
that finds nothing, ifdef.  The order checked here is the one the oracle
pipeline replays."""
import pytest

from aquagpusph_amd import casefiles, cases
from aquagpusph_amd.sim import problem_tools

BASE = ("resources/Presets/basic.xml",
        "resources/Presets/basic/time_scheme/improved_euler.xml",
        "resources/Presets/basic/domain.xml",
        "resources/Presets/cfd.xml",
        "resources/Presets/cfd/deltaSPH.xml")

SETS = ["portal_out_r", "portal_in_r", "portal_n"]


def _names(xml, dims):
    return [n for n, _ in problem_tools(xml, dims)]


def _block(names, first, count):
    i = names.index(first)
    return names[i:i + count]


def test_taylor_green_portal_tool_order(tmp_path):
    case = cases.taylor_green(n=16)
    ports = cases.taylor_green_portals(case["params"]["L"])
    xml = casefiles.write_case(case, str(tmp_path), presets=BASE,
                               portals=ports)
    names = _names(xml, 2)
    # the forces stage of every portal, in include order, right before the
    # Interactions dummy and after the ordinary delta-SPH Laplacian
    stage = ["cfd portal icell backup"] + [None] * 3 + [
        "cfd portal mirror", "cfd portal Shepard", "cfd portal interactions",
        "cfd portal delta-SPH", "cfd portal unmirror",
        "cfd portal icell restore"]
    at = names.index("cfd lap p") + 1
    for px, _, _, _ in ports:
        want = [px + s if s else None for s in stage]
        want[1:4] = [px + s for s in SETS]
        assert names[at:at + len(want)] == want, (px, names[at:at + 12])
        at += len(want)
    assert names[at] == "Interactions"
    # no MLS / LapP Correction in this stack: the try_inserts add nothing,
    # no BI: no portal BI, no delta-SPH-full: no portal full
    assert not any("portal_mls" in n or "portal_deltaSPH" in n or
                   "portal BI" in n or "delta-SPH full" in n for n in names)
    # teleports after the corrector kernel, each after its planes
    at = names.index("corrector") + 1
    for px, _, _, _ in ports:
        want = [px + s for s in SETS] + [px + "cfd portal mirror_teleport"]
        assert names[at:at + 4] == want
        at += 4
    assert names[at] == "Corrector"


@pytest.mark.parametrize("dims", [2, 3])
def test_slab_portal_stages_with_mls_and_bi(tmp_path, dims):
    n = (16, 8) if dims == 2 else (12, 6, 6)
    case = cases.periodic_slab(dims, n, wall=True)
    ports = cases.slab_portals(case["params"]["L"], dims)
    pres = BASE + ("resources/Presets/cfd/deltaSPH-full.xml",
                   "resources/Presets/cfd/BI.xml")
    xml = casefiles.write_case(case, str(tmp_path), presets=pres,
                               portals=ports)
    names = _names(xml, dims)
    # MLS stage between the MLS sums and their inversion
    a, b = names.index("imove1_MLS interactions"), names.index("imove1_MLS")
    mls = [px + s for px, _, _, _ in ports for s in
           ["cfd portal_mls icell backup"] + SETS +
           ["cfd portal_mls mirror", "cfd portal_mls MLS",
            "cfd portal_mls unmirror", "cfd portal_mls icell restore"]]
    assert names[a + 1:b] == mls
    # forces stage after the BI tools, with portal BI and the full term
    i = names.index("cfd BI Shepard") + 1
    assert names[i:i + 12] == ["right_cfd portal icell backup"] + [
        "right_" + s for s in SETS] + [
        "right_cfd portal " + s for s in
        ("mirror", "Shepard", "interactions", "BI", "delta-SPH",
         "delta-SPH full", "unmirror", "icell restore")]
    # lap_p correction stage after the ordinary correction
    i = names.index("cfd lap p apply correction") + 1
    j = names.index("LapP Correction")
    assert names[i:j] == [px + s for px, _, _, _ in ports for s in
                          ["cfd portal_deltaSPH icell backup"] + SETS +
                          ["cfd portal_deltaSPH mirror",
                           "cfd portal_deltaSPH lapp_corr",
                           "cfd portal_deltaSPH unmirror",
                           "cfd portal_deltaSPH icell restore"]]


def test_batch_preset_tool_order(tmp_path):
    """Presets/batch.xml swaps the scalar-parameter kernels for their
    per-set forms in place and adds the per-set time step and time after
    the global ones; one particle set per tank with its scalars."""
    b = cases.sloshing_batch(3, res=20)
    pres = casefiles.PBI_PRESETS + ("resources/Presets/batch.xml",)
    xml = casefiles.write_case(b, str(tmp_path), presets=pres)
    tools = problem_tools(xml, 2)
    names = [n for n, _ in tools]
    base = _names(casefiles.write_case(b, str(tmp_path / "plain")), 2)
    added = ["batch reset time steps", "batch time steps", "batch times"]
    assert [n for n in names if n not in added] == base
    i = names.index("cfd minimum time step")
    assert names[i + 1:i + 3] == added[:2]
    assert names[names.index("t = t + dt") + 1] == "batch times"
    text = open(xml).read()
    assert text.count("<ParticlesSet") == 3
    assert text.count('<Scalar name="exc_omega"') == 3


def test_reports_run_after_every_tool(tmp_path):
    """<Reports> are appended after all tools (CalcServer.cpp:472-520), as
    report_<type> tools; timing.report.xml holds a screen and a file
    report, performance.report.xml a performance report."""
    case = cases.dam_break(2, 0.3 / 10)
    presets = casefiles.PBI_PRESETS + (
        "resources/Presets/basic/timing.report.xml",
        "resources/Presets/basic/performance.report.xml")
    xml = casefiles.write_case(case, str(tmp_path), presets=presets)
    tl = problem_tools(xml, 2)
    assert tl[-3:] == [("Time control", "report_screen"),
                       ("Time control file", "report_file"),
                       ("Performance", "report_performance")]
    names = [n for n, _ in tl]
    assert names.index("iter += 1") < names.index("Time control")
    # a file report without a path is rejected at parse time
    bad = tmp_path / "bad.xml"
    bad.write_text('<?xml version="1.0" ?><sphInput><Reports>'
                   '<Report type="file" name="x" fields="t"/></Reports>'
                   '</sphInput>')
    with pytest.raises(Exception, match="path"):
        problem_tools(str(bad), 2)


# the include list of examples/2D/spheric_testcase5_dambreak/src/templates/
# Main.xml:22-32 (presets only; the case's own Settings/SPH/Time/Fluids
# files are replaced by this case's variables and particle set)
TESTCASE5 = ("resources/Presets/basic.xml",
             "resources/Presets/basic/time_scheme/improved_euler.xml",
             "resources/Presets/basic/domain.xml",
             "resources/Presets/cfd.xml",
             "resources/Presets/cfd/deltaSPH.xml",
             "resources/Presets/cfd/deltaSPH-full.xml",
             "resources/Presets/cfd/BI.xml",
             "resources/Presets/cfd/elasticBounce.xml",
             "resources/Presets/cfd/variableTimeStep.xml",
             "resources/Presets/basic/timing.report.xml",
             "resources/Presets/basic/performance.report.xml")


def test_spheric_testcase5_presets_resolve(tmp_path):
    """The preset stack of the reference's 2D dam-break example resolves,
    every kernel tool of it has a HIP implementation in the registry, and
    the elastic bounce runs right after the Rates marker (cfd/
    elasticBounce.xml inserts after="Rates"), i.e. after the cfd rates and
    the delta-SPH rate and before the corrector."""
    from aquagpusph_amd.sim import registered_kernels
    case = cases.dam_break(2, 0.3 / 10)
    xml = casefiles.write_case(case, str(tmp_path), presets=TESTCASE5,
                               extra_xml='    <Variables><Variable name="dr" '
                               'type="float" value="0.03" /></Variables>')
    tl = problem_tools(xml, 2, keys=True)
    reg = set(registered_kernels())
    missing = [(n, k) for n, t, k in tl if t == "kernel" and k not in reg]
    assert not missing, missing
    names = [n for n, _, _ in tl]
    i = names.index("Rates")
    assert names[i + 1] == "cfd elastic bounce"
    assert names.index("cfd rates") < i
    assert names.index("cfd delta-SPH") < i
    assert names.index("Corrector") > i + 1
    keys = dict((n, k) for n, _, k in tl)
    assert keys["cfd elastic bounce"] == "cfd/Boundary/ElasticBounce.cl:entry"
