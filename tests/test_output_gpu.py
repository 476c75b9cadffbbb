"""Output of a run on the GPU: UnSort (aquagpusph/CalcServer/UnSort.cl.in)
bit for bit, the AQUAgpusph main loop with Timing end / output criteria
(TimeManager.cpp:103-150), the ASCII particle files in the original particle
order, and a run resumed from the state file AQUAgpusph.save.N.xml."""
import ctypes
import os

import numpy as np
import pytest
import torch

from aquagpusph_amd import _lib, casefiles, cases
from aquagpusph_amd.sim import Simulation

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("words", [1, 2, 4, 16])
def test_unsort_is_the_inverse_permutation(cuda, words):
    rng = np.random.default_rng(words)
    n = 100003
    ids = rng.permutation(n).astype(np.uint32)
    src = rng.integers(0, 2 ** 32, (n, words), dtype=np.uint64).astype(
        np.uint32)
    d_id = torch.from_numpy(ids.view(np.int32)).to(cuda)
    d_in = torch.from_numpy(src.view(np.int32)).to(cuda)
    d_out = torch.zeros_like(d_in)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(_lib.lib().aq_unsort(ctypes.c_void_p(d_id.data_ptr()),
                                    ctypes.c_void_p(d_in.data_ptr()),
                                    ctypes.c_void_p(d_out.data_ptr()),
                                    ctypes.c_uint(n), ctypes.c_uint(4 * words),
                                    s))
    torch.cuda.synchronize()
    want = np.zeros_like(src)
    want[ids] = src
    np.testing.assert_array_equal(d_out.cpu().numpy().view(np.uint32), want)


def _read_ascii(path, widths):
    rows = []
    for line in open(path):
        line = line.split("#")[0].strip()
        if not line:
            continue
        rows.append([float(x) for x in line.replace(",", " ").split()])
    a = np.array(rows)
    out, c = [], 0
    for w in widths:
        out.append(a[:, c:c + w])
        c += w
    return out


def test_run_saves_unsorted_files_and_resumes(cuda, tmp_path, monkeypatch):
    """The BI preset is left out: its wall-element pressure gradient reads
    du/dt at the element's slot of the sorted arrays, which the sort does
    not carry along (basic/Sort.cl:119-123 moves only du/dt_in), so with BI
    walls a resumed run differs from the continued one by that slot history
    (in the reference as well).  Without them the particle files, written
    with 9 significant digits, read back to the same float bits, and the
    resumed run equals the continued one bit for bit as long as no two
    particles that share a cell reached it from different cells before the
    resume (the stable sort keeps such ties in the order of the previous
    step, a resume in the order of the particle ids), which holds over the
    first steps of the dam break."""
    monkeypatch.chdir(tmp_path)
    dims = 2
    case = cases.dam_break(dims, 0.3 / 20)
    fields = casefiles.FIELDS
    presets = tuple(p for p in casefiles.PBI_PRESETS
                    if not p.endswith("BI.xml"))
    xml = casefiles.write_case(
        case, str(tmp_path / "case"), presets=presets,
        save=("ASCII", "Fluid.{index}.dat", fields),
        timing=[("End", "Steps", 8), ("Output", "IPF", 4)])
    sim = Simulation(xml, dims)
    assert sim.run() == 8
    # outputs at the first step, after 4 more (frames 1, 2) and at the end
    names = sorted(f for f in os.listdir(tmp_path) if f.startswith("Fluid."))
    assert names == ["Fluid.00000.dat", "Fluid.00001.dat", "Fluid.00002.dat"]
    assert os.path.exists(tmp_path / "AQUAgpusph.save.0.xml")
    assert int(sim.get("frame")) == 2 and int(sim.get("iter")) == 8
    # the last file holds the state in the original particle order, exactly
    ids = sim.get("id").astype(np.int64)
    widths = []
    for f in fields:
        v = sim.get(f)
        widths.append(1 if v.ndim == 1 else v.shape[1])
    cols = _read_ascii(tmp_path / "Fluid.00002.dat", widths)
    for f, col in zip(fields, cols):
        v = sim.get(f).astype(np.float64).reshape(len(ids), -1)
        orig = np.zeros_like(v)
        orig[ids] = v
        np.testing.assert_array_equal(col.astype(np.float32),
                                      orig.astype(np.float32), err_msg=f)
    # resume from the state file and compare with the continued run
    res = Simulation(str(tmp_path / "AQUAgpusph.save.0.xml"), dims)
    for k in ("t", "dt", "iter", "frame"):
        assert float(res.get(k)) == pytest.approx(float(sim.get(k)),
                                                  rel=1e-6), k
    sim.step(4)
    res.step(4)

    def orig(s, f):
        i = s.get("id").astype(np.int64)
        v = s.get(f).astype(np.float64).reshape(len(i), -1)
        o = np.zeros_like(v)
        o[i] = v
        return o

    for f in ("r", "u", "rho", "p"):
        np.testing.assert_array_equal(orig(res, f), orig(sim, f), err_msg=f)
    sim.close()
    res.close()


@pytest.mark.parametrize("graph", [False, True])
def test_time_control_reports(cuda, tmp_path, monkeypatch, capfd, graph):
    """basic/timing.report.xml: the file report writes "iter t dt" after
    every step (Reports/TabFile.cpp), in eager and in graph-replayed steps
    (a pinned copy plus a host node), and the screen report prints the
    same values; unbuilt report types are refused."""
    monkeypatch.chdir(tmp_path)
    case = cases.dam_break(2, 0.3 / 10)
    presets = casefiles.PBI_PRESETS + (
        "resources/Presets/basic/timing.report.xml",)
    xml = casefiles.write_case(case, str(tmp_path / "case"), presets=presets)
    sim = Simulation(xml, 2)
    rows = []
    for k in range(3):
        sim.step(2, graph=graph)
        rows.append((int(sim.get("iter")), float(np.float32(sim.get("t"))),
                     float(np.float32(sim.get("dt")))))
    sim.close()
    lines = [l.split() for l in open(tmp_path / "Timing.dat")]
    assert lines[0][0] == "#" and lines[0][1:] == ["iter", "t", "dt"]
    vals = [(int(a), float(b), float(c)) for a, b, c in lines[1:]]
    assert len(vals) == 6
    for k, (it, t, dt) in enumerate(rows):
        assert vals[2 * k + 1][0] == it
        assert np.float32(vals[2 * k + 1][1]) == np.float32(t)
    assert [v[0] for v in vals] == list(range(1, 7))
    out = capfd.readouterr().out
    assert out.count("Time control:") == 6
    assert "iter=6 " in out
    bad = tmp_path / "bad"
    xml = casefiles.write_case(case, str(bad), extra_xml=(
        '<Reports><Report type="dump" name="d" fields="r" path="x.dat"/>'
        '</Reports>'))
    with pytest.raises(Exception, match="not supported"):
        Simulation(xml, 2)
