"""CPU tests of the drop-in boundary: the C ABI library and the XML
tool/variable surface (no GPU work)."""
import ctypes
import os
import re

import numpy as np
import pytest

from aquagpusph_amd import _lib, casefiles, cases


def test_library_exports_every_header_symbol():
    L = _lib.lib()
    syms = _lib.header_symbols()
    assert len(syms) > 40
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert b"gfx950" in L.aq_version()


def test_error_reporting_without_gpu():
    """Host-side argument checks fail with a code and a message."""
    L = _lib.lib()
    rc = L.aq_predictor(5, None, None, None, None, None, None, None, None,
                        None, None, None, 10, None, None)
    assert rc == -1
    assert b"dims" in L.aq_last_error()
    assert L.aq_strerror(-6) == b"assertion failed"


def _resolve(xml, dims):
    L = _lib.lib()
    L.aq_xml_resolve.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                 ctypes.c_char_p, ctypes.c_size_t,
                                 ctypes.POINTER(ctypes.c_size_t)]
    need = ctypes.c_size_t()
    buf = ctypes.create_string_buffer(1 << 20)
    _lib.check(L.aq_xml_resolve(xml.encode(), dims, buf, len(buf),
                                ctypes.byref(need)))
    tools, var, defs, N = [], [], {}, None
    for line in buf.value.decode().splitlines():
        if line.startswith("#var "):
            var.append(line[5:].split("\t"))
        elif line.startswith("#def "):
            k, v, e = line[5:].split("\t")
            defs[k] = (v, e)
        elif line.startswith("#N "):
            N = int(line[3:])
        else:
            tools.append(tuple(line.split("\t")))
    return tools, var, defs, N


# tool order of the reference preset stack basic + improved_euler + domain +
# cfd + deltaSPH + BI + variableTimeStep (State.cpp insert semantics)
PBI_ORDER = [
    ("predictor", "basic/time_scheme/improved_euler.cl:predictor"),
    ("Predictor", None), ("Domain", "basic/Domain.cl:entry"),
    ("link-list", None), ("Link-List", None), ("Backup id", None),
    ("Backup iset", None), ("Backup imove", None), ("Backup normal", None),
    ("Backup tangent", None), ("Backup m", None),
    ("sort stage1", "basic/Sort.cl:stage1"),
    ("sort stage2", "basic/Sort.cl:stage2"), ("EOS", "basic/EOS.cl:entry"),
    ("Binormal", "basic/Binormal.cl:entry"), ("Sort", None),
    ("cfd reinit lap_p", None), ("cfd Reinit lap_u", None),
    ("cfd Reinit div_u", None), ("cfd Reinit grad_p", None),
    ("cfd Reinit shepard", None),
    ("cfd Shepard", "cfd/Boundary/BI/Shepard.cl:compute"),
    ("cfd interactions", "cfd/Interactions.cl:entry"),
    ("cfd sensors", "cfd/Sensors.cl:entry"),
    ("cfd sensors renormalization", "cfd/SensorsRenormalization.cl:entry"),
    ("cfd lap p", "cfd/deltaSPH.cl:lapp"),
    ("cfd BI lap_u", "cfd/Boundary/BI/LapU.cl:freeslip"),
    ("cfd BI grad_p", "cfd/Boundary/BI/GradP.cl:freeslip"),
    ("cfd BI interpolation", "cfd/Boundary/BI/Interpolation.cl:entry"),
    ("cfd BI interpolation Shepard",
     "cfd/Boundary/BI/InterpolationShepard.cl:entry"),
    ("cfd BI interactions", "cfd/Boundary/BI/Interactions.cl:entry"),
    ("cfd BI Shepard", "cfd/Boundary/BI/Shepard.cl:apply"),
    ("Interactions", None), ("cfd rates", "cfd/Rates.cl:entry"),
    ("cfd delta-SPH", "cfd/deltaSPH.cl:deltaSPH"), ("Rates", None),
    ("corrector", "basic/time_scheme/improved_euler.cl:corrector"),
    ("Corrector", None), ("fixed time step", None),
    ("cfd variable time step", "cfd/TimeStep.cl:entry"),
    ("cfd minimum time step", None), ("cfd check time step", None),
    ("TimeStep", None), ("t = t + dt", None), ("iter += 1", None),
    ("End", None)]


def test_pbi_preset_stack_resolves_in_reference_order(tmp_path):
    case = cases.dam_break(2, 0.3 / 10)
    xml = casefiles.write_case(case, str(tmp_path))
    tools, var, defs, N = _resolve(xml, 2)
    assert [t[0] for t in tools] == [n for n, _ in PBI_ORDER]
    for (name, typ, key), (_, want) in zip(tools, PBI_ORDER):
        if want:
            assert typ == "kernel" and key == want, name
    assert N == case["N"]
    assert defs["H"] == ("h", "1")
    assert defs["__LAP_FORMULATION__"][0] == "__LAP_MONAGHAN__"
    # later registrations override earlier ones (courant from the case)
    assert [v for v in var if v[0] == "courant"][-1][3].startswith("0.1")


def test_tool_actions(tmp_path):
    """insert before/after (glob and comma lists), prefixes, replace, remove,
    try_* and ifdef, as State.cpp:739-1033."""
    (tmp_path / "sub.xml").write_text("""<?xml version="1.0" ?>
<sphInput>
    <Tools>
        <Tool action="add" name="A" type="dummy"/>
        <Tool action="insert" after_prefix="A" name="B" type="dummy"/>
    </Tools>
</sphInput>""")
    (tmp_path / "main.xml").write_text("""<?xml version="1.0" ?>
<sphInput>
    <Include file="sub.xml" prefix="p_"/>
    <Definitions><Define name="FLAG"/></Definitions>
    <Tools>
        <Tool action="add" name="one" type="dummy"/>
        <Tool action="add" name="two" type="dummy"/>
        <Tool action="insert" before="p_*" name="zero" type="dummy"/>
        <Tool action="insert" after="one,two" name="after12" type="dummy"/>
        <Tool action="insert" before="one,two" name="before12" type="dummy"/>
        <Tool action="replace" name="two" type="set_scalar" in="t" value="1"/>
        <Tool action="remove" name="p_B" type="dummy"/>
        <Tool action="try_remove" name="nope" type="dummy"/>
        <Tool action="add" name="flagged" type="dummy" ifdef="FLAG"/>
        <Tool action="add" name="skipped" type="dummy" ifndef="FLAG"/>
        <Tool action="insert" at="0" name="first" type="dummy"/>
    </Tools>
    <ParticlesSet n="3"/>
</sphInput>""")
    tools, _, _, N = _resolve(str(tmp_path / "main.xml"), 3)
    names = [t[0] for t in tools]
    # "zero" is inserted before every match of p_* (p_A and p_B): the same
    # tool at two places; removing p_B leaves both copies of "zero"
    assert names == ["first", "zero", "p_A", "zero", "before12", "one", "two",
                     "after12", "flagged"]
    assert dict((t[0], t[1]) for t in tools)["two"] == "set_scalar"
    assert N == 3


def test_missing_insert_anchor_is_an_error(tmp_path):
    (tmp_path / "m.xml").write_text("""<sphInput><Tools>
        <Tool action="insert" before="nothing" name="x" type="dummy"/>
    </Tools></sphInput>""")
    with pytest.raises(_lib.AquaError, match="cannot be found"):
        _resolve(str(tmp_path / "m.xml"), 2)


def test_xml_entities_and_reduction_text(tmp_path):
    case = cases.dam_break(3, 0.3 / 4)
    xml = casefiles.write_case(
        case, str(tmp_path),
        presets=casefiles.PBI_PRESETS + (
            "resources/Presets/basic/neighs_limit.xml",))
    tools, _, _, _ = _resolve(xml, 3)
    names = [t[0] for t in tools]
    i = names.index("CountNeighbours")
    assert names[i - 3:i] == ["number of neighbours",
                              "max number of neighbours",
                              "check number of neighbours"]
    assert names.index("Sort") < i


def test_pgp_preset_stack_order(tmp_path):
    """cfd/GP.xml inserted into the basic + cfd stack: the associations are
    re-sorted before "Sort", the mirror/interpolate/unmirror chain follows
    "cfd Shepard", the extensions and ghost interactions close the
    interactions stage (resources/Presets/src/cfd/GP.xml:50-80)."""
    case = cases.still_tank(2, 0.15 / 8, bc="GP")
    xml = casefiles.write_case(case, str(tmp_path),
                               presets=casefiles.PGP_PRESETS, load=False)
    names = [t[0] for t in _resolve(xml, 2)[0]]
    idx = {n: k for k, n in enumerate(names)}
    chain = ["Backup associations", "Sort associations", "Sort",
             "cfd Shepard", "cfd GP backup r", "cfd GP mirror",
             "cfd reinit gp_rho", "cfd reinit gp_p", "cfd reinit gp_u",
             "cfd GP interpolation", "cfd GP renormalization",
             "cfd GP unmirror", "cfd GP backup u", "cfd interactions",
             "cfd GP preLapU", "cfd GP LapU", "cfd GP preInteractions",
             "cfd GP interactions", "cfd GP restore u", "Interactions",
             "cfd rates"]
    pos = [idx[n] for n in chain]
    assert pos == sorted(pos), list(zip(chain, pos))
    assert idx["sort stage2"] < idx["Backup associations"]
