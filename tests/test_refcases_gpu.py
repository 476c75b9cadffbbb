"""The reference's own tool-test cases run unchanged through this runtime:
tests/{2D,3D}/{Assert,Copy,RadixSort,Reduction,Set,SetScalar,Sort}/cMake/
main.xml and tests/ExternalTool/cMake/main.xml, stored as fixtures by
tests/golden/make_refcases.py.  As the reference's CMake configure step
does, only @RESOURCES_DIR@ and @EXTTOOL_LIB_NAME@ are substituted; each case
runs its Timing loop (one step) in a fresh directory and its output files are
checked with the assertions of the case's own check.py."""
import json
import os
import shutil

import numpy as np
import pytest

import aquagpusph_amd
from aquagpusph_amd import _lib
from aquagpusph_amd.sim import Simulation

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
PLUGIN = os.path.join(os.path.dirname(__file__), "plugin",
                      "libaq_demo_tool.so")
MANIFEST = json.load(open(os.path.join(GOLDEN, "refcases", "manifest.json")))
RESOURCES = os.path.join(os.path.dirname(aquagpusph_amd.__file__),
                         "resources")


def _setup(name, tmp_path, monkeypatch):
    m = MANIFEST[name]
    xml = open(os.path.join(GOLDEN, "refcases", name + ".xml")).read()
    xml = xml.replace("@RESOURCES_DIR@", RESOURCES)
    xml = xml.replace("@EXTTOOL_LIB_NAME@", os.path.basename(PLUGIN))
    (tmp_path / "main.xml").write_text(xml)
    shutil.copy(os.path.join(GOLDEN, m["particles"]),
                tmp_path / "particles.dat")
    if "EXTTOOL" in open(os.path.join(GOLDEN, "refcases",
                                      name + ".xml")).read():
        shutil.copy(PLUGIN, tmp_path / os.path.basename(PLUGIN))
    monkeypatch.chdir(tmp_path)
    return str(tmp_path / "main.xml"), m["dims"]


def _rows(path, kind):
    """The data lines after the 18 header lines (check.py's readlines()[18:]),
    ',' and ' ' both separating."""
    out = []
    for line in open(path).readlines()[18:]:
        line = line.replace(",", " ").strip()
        if line:
            out.append([kind(x) for x in line.split(" ")])
    return out


@pytest.mark.parametrize("dims", [2, 3])
def test_reference_copy_case(cuda, tmp_path, monkeypatch, dims):
    xml, d = _setup("%dD_Copy" % dims, tmp_path, monkeypatch)
    sim = Simulation(xml, d)
    assert sim.run() == 1
    sim.close()
    for row in _rows("set0.00000.dat", float):
        n = len(row) // 2
        assert row[:n] == row[n:]


@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("tool", ["RadixSort", "Sort"])
def test_reference_sort_cases(cuda, tmp_path, monkeypatch, dims, tool):
    xml, d = _setup("%dD_%s" % (dims, tool), tmp_path, monkeypatch)
    sim = Simulation(xml, d)
    assert sim.run() == 1
    sim.close()
    rows = _rows("set0.00000.dat", int)
    f, f_orig, perms, inv_perms = (list(c) for c in zip(*rows))
    assert len(f) == 500
    assert f == sorted(f)
    assert f == sorted(f_orig)
    assert f_orig != sorted(f_orig)
    for i in range(len(f)):
        assert f[i] == f_orig[perms[i]]
        assert f[inv_perms[i]] == f_orig[i]


@pytest.mark.parametrize("dims", [2, 3])
def test_reference_reduction_case(cuda, tmp_path, monkeypatch, dims):
    xml, d = _setup("%dD_Reduction" % dims, tmp_path, monkeypatch)
    sim = Simulation(xml, d)
    assert sim.run() == 1
    sim.close()
    r = np.array(_rows("set0.00000.dat", float))
    line = open("vars.out").readlines()[1]
    line = line.replace(",", " ").replace("(", "").replace(")", "").strip()
    fields = [float(x) for x in line.split(" ")]
    n = len(fields) // 2
    assert n == r.shape[1]
    for i in range(n):
        assert abs(r[:, i].sum() - fields[i]) < 1e-2
        assert abs(r[:, i].max() - fields[i + n]) < 1e-5
    assert os.path.exists("performance.json")


@pytest.mark.parametrize("dims", [2, 3])
def test_reference_set_case(cuda, tmp_path, monkeypatch, dims):
    xml, d = _setup("%dD_Set" % dims, tmp_path, monkeypatch)
    sim = Simulation(xml, d)
    assert sim.run() == 1
    sim.close()
    values = (0, 0, 0, 0, 0.2) if dims == 3 else (0, 0, 0.2)
    for row in _rows("set0.00000.dat", float):
        assert tuple(np.float32(row)) == tuple(np.float32(values))


@pytest.mark.parametrize("dims", [2, 3])
def test_reference_setscalar_case(cuda, tmp_path, monkeypatch, dims):
    xml, d = _setup("%dD_SetScalar" % dims, tmp_path, monkeypatch)
    sim = Simulation(xml, d)
    assert sim.run() == 1
    sim.close()
    line = open("vars.out").readlines()[1].strip()
    assert line == ("-1 -4 (-16,-0.75,5,6)" if dims == 3 else
                    "-1 -4 (-16,-0.75)")


@pytest.mark.parametrize("dims", [2, 3])
def test_reference_assert_case(cuda, tmp_path, monkeypatch, dims):
    """The bad assertion stops the run with an error naming it, and the tool
    after it in the step ("skipped tool", which would flip h back) is not
    run: h stays at the value the assertion rejected."""
    xml, d = _setup("%dD_Assert" % dims, tmp_path, monkeypatch)
    sim = Simulation(xml, d)
    with pytest.raises(_lib.AquaError, match="Bad assertion"):
        sim.run(graph=False)
    assert float(sim.get("h")) == -2.0
    assert int(sim.get("iter")) == 0
    sim.close()


def test_reference_installable_tool_case(cuda, tmp_path, monkeypatch,
                                         capfd):
    """tests/ExternalTool/cMake/main.xml: the installable tool is loaded
    from the plugin library named in the case and runs in the step (its
    check.py looks for the tool's log line and a clean finish)."""
    xml, d = _setup("ExternalTool", tmp_path, monkeypatch)
    sim = Simulation(xml, d)
    assert sim.run() == 1
    assert int(sim.get("iter")) == 1
    sim.close()
    out = capfd.readouterr().out
    assert "Executing Installed demo tool!" in out


@pytest.mark.parametrize("graph", [False, True])
def test_installable_tool_acts_on_variables(cuda, tmp_path, graph):
    """The plugin's tool with in/out attributes copies one variable into
    another on the runtime's stream after a set tool has cleared the
    target, every step, in eager steps and replayed from a captured graph;
    a plugin whose setup fails (a variable it asks for is
    missing) is reported with the tool's name."""
    shutil.copy(PLUGIN, tmp_path / "plugin.so")
    N = 300
    r = np.random.default_rng(3).uniform(-1, 1, (N, 2)).astype(np.float32)
    np.savetxt(tmp_path / "p.dat", r)
    (tmp_path / "main.xml").write_text("""<?xml version="1.0" ?>
<sphInput>
    <Variables>
        <Variable name="r" type="vec*" length="N" />
        <Variable name="r_copy" type="vec*" length="N" />
        <Variable name="h" type="float" value="1.0" />
    </Variables>
    <ParticlesSet n="%d">
        <Load fields="r" file="p.dat" format="FastASCII"/>
    </ParticlesSet>
    <Tools>
        <Tool action="add" name="clear copy" type="set" in="r_copy" value="VEC_ZERO"/>
        <Tool action="add" name="plugin copy" type="installable" path="./plugin.so" in="r" out="r_copy"/>
        <Tool action="add" name="iter += 1" type="set_scalar" in="iter" value="iter + 1"/>
    </Tools>
</sphInput>
""" % N)
    sim = Simulation(str(tmp_path / "main.xml"), 2)
    sim.step(3, graph=graph)
    np.testing.assert_array_equal(sim.get("r_copy"), r)
    np.testing.assert_array_equal(sim.get("r"), r)
    assert int(sim.get("iter")) == 3
    sim.close()
    bad = (tmp_path / "main.xml").read_text().replace('out="r_copy"',
                                                      'out="nothere"')
    (tmp_path / "bad.xml").write_text(bad)
    with pytest.raises(_lib.AquaError, match="plugin copy"):
        Simulation(str(tmp_path / "bad.xml"), 2)
