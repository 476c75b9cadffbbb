"""The precision mode against the float64 oracle on the 2D dam break: after
300 steps from the same float32 start, the run with Presets/precision.xml
is closer to the float64 run than the plain float32 run, in positions and
velocities (the full 2000-step study is tests/precision_study.py)."""
import pytest

from precision_study import run_study

pytestmark = pytest.mark.gpu


def test_precision_mode_tracks_float64(cuda):
    res = run_study(steps=300, every=150)
    f = res["final"]
    print({k: {q: f[k][q] for q in ("pos_rms", "vel_rms", "front_err",
                                     "e_mech_err")}
           for k in ("plain", "precision")})
    assert f["precision"]["pos_rms"] < f["plain"]["pos_rms"]
    assert f["precision"]["vel_rms"] <= f["plain"]["vel_rms"] * 1.1
    assert f["precision"]["mass_fp64"] == pytest.approx(f["fp64"]["mass"],
                                                        rel=1e-12)
