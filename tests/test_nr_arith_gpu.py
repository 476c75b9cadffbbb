"""The fused walk's division and square root sequences without range
scaling (div_nr, sqrt_nr): equal bit for bit to the compiler's IEEE a / b
and sqrtf over their whole operand window on the hardware
(aq_check_nr_arith), and the walk that uses them equal bit for bit to the
walk on the IEEE sequences (AQ_DIV=ieee) and to the separate per-particle
kernels, including the first step from rest (zero numerators of both
signs), lattice neighbours at q == 2 (zero and near-zero kernel
derivatives) and a random cloud."""
import ctypes

import numpy as np
import pytest

from aquagpusph_amd import _lib, casefiles, cases
from aquagpusph_amd.sim import Simulation

pytestmark = pytest.mark.gpu

FIELDS = ("shepard", "grad_p", "lap_u", "div_u", "lap_p", "r", "u", "rho",
          "p")


def test_nr_sequences_match_ieee(cuda):
    L = _lib.lib()
    L.aq_check_nr_arith.argtypes = [ctypes.c_uint64, ctypes.c_uint64,
                                    ctypes.POINTER(ctypes.c_uint64)]
    for seed in (1, 12345):
        out = (ctypes.c_uint64 * 2)()
        _lib.check(L.aq_check_nr_arith(1 << 26, seed, out))
        assert (out[0], out[1]) == (0, 0), (seed, out[0], out[1])


def test_clmath_sequences_match_oracle(cuda):
    """The device atan / acospi of the boundary-integral terms equal the
    CPU oracle's float sequence bit for bit (the same Cephes-method
    operations; the oracle's ulp bounds are checked on the CPU): atan over a
    stride of every finite float of both signs, acospi over [0, 1]."""
    import torch
    from oracle import pipeline as op
    L = _lib.lib()
    one = int(np.float32(1.0).view(np.uint32))
    cases_ = [(0, np.arange(0, 0x7F800000, 127, dtype=np.uint32)),
              (0, np.arange(0x80000000, 0xFF800000, 131, dtype=np.uint32)),
              (1, np.arange(0, one + 1, 61, dtype=np.uint32))]
    for opc, bits in cases_:
        x = bits.view(np.float32)
        want = np.zeros_like(x)
        op.lib().orc_clmath_f32(ctypes.c_int(opc),
                                ctypes.c_void_p(x.ctypes.data),
                                ctypes.c_void_p(want.ctypes.data),
                                ctypes.c_uint(x.size))
        xd = torch.from_numpy(x.copy()).cuda()
        yd = torch.empty_like(xd)
        _lib.call("aq_clmath_eval", ctypes.c_int(opc),
                  ctypes.c_void_p(xd.data_ptr()),
                  ctypes.c_void_p(yd.data_ptr()), ctypes.c_uint32(x.size),
                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        got = yd.cpu().numpy()
        bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
        assert bad.size == 0, (opc, bad.size, x[bad[:4]], got[bad[:4]],
                               want[bad[:4]])


def _pair(xml, dims, monkeypatch, steps, **env):
    sims = []
    for div in ("nr", "ieee"):
        monkeypatch.setenv("AQ_DIV", div)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        sims.append(Simulation(xml, dims))
    for _ in range(steps):
        for s in sims:
            s.step(2, graph=True)
        for f in FIELDS:
            np.testing.assert_array_equal(sims[0].get(f), sims[1].get(f),
                                          err_msg=f)
    return sims


@pytest.mark.parametrize("dims", [2, 3])
def test_walk_nr_equals_ieee_dam_break(cuda, tmp_path, dims, monkeypatch):
    case = cases.dam_break(dims, 0.3 / (20 if dims == 2 else 6))
    xml = casefiles.write_case(case, str(tmp_path))
    _pair(xml, dims, monkeypatch, 3)


def test_walk_nr_equals_separate_kernels_dense(cuda, tmp_path, monkeypatch):
    from test_runtime_gpu import _dense_case
    case = _dense_case(3, 20000, 5)
    presets = tuple(p for p in casefiles.PBI_PRESETS
                    if not p.endswith("BI.xml"))
    xml = casefiles.write_case(case, str(tmp_path), presets=presets)
    monkeypatch.setenv("AQ_NO_FUSION", "1")
    monkeypatch.setenv("AQ_WALK", "thread")
    a = Simulation(xml, 3)
    monkeypatch.setenv("AQ_NO_FUSION", "0")
    monkeypatch.setenv("AQ_WALK", "cells")
    monkeypatch.setenv("AQ_DIV", "nr")
    b = Simulation(xml, 3)
    for _ in range(2):
        a.step(2)
        b.step(2)
        for f in FIELDS:
            np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)


@pytest.mark.parametrize("dims,res", [(2, 60), (3, 14)])
def test_walk_nr_equals_ieee_long_dam_break(cuda, tmp_path, dims, res,
                                            monkeypatch):
    """6000 steps of a collapsing column, from rest to a flow faster than
    1 m/s (checked): the walk on div_nr / sqrt_nr and the walk on the IEEE
    sequences stay bit-identical at every 1000th step."""
    case = cases.dam_break(dims, 0.3 / res)
    xml = casefiles.write_case(case, str(tmp_path))
    sims = []
    for div in ("nr", "ieee"):
        monkeypatch.setenv("AQ_DIV", div)
        sims.append(Simulation(xml, dims))
    for _ in range(6):
        for s in sims:
            s.step(1000, graph=True)
        for f in FIELDS:
            np.testing.assert_array_equal(sims[0].get(f), sims[1].get(f),
                                          err_msg=f)
    assert float(np.abs(sims[0].get("u")).max()) > 1.0
