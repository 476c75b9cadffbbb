"""A small run of the LES study (tests/les_study.py): fine 64 x 64 and
coarse 32 x 32 decaying turbulence; the network trained on the GPU from
the filtered fine run lowers its loss, and the coarse runs (plain,
learned closure, Smagorinsky) run and lose kinetic energy."""
import numpy as np
import pytest

from les_study import study

pytestmark = pytest.mark.gpu


def test_les_study_small(cuda):
    res = study(n=32, t_end=0.1, samples=4, snapshots=2, iters=60)
    tr = res["training"]
    print(tr["loss"], {k: v["ke_err"] for k, v in res["runs"].items()
                       if "ke_err" in v})
    assert tr["samples"] > 100
    assert tr["loss"][-1][1] < tr["loss"][0][1]
    for name in ("plain", "learned", "smagorinsky", "fine"):
        ke = np.asarray(res["runs"][name]["ke"])
        assert np.all(np.isfinite(ke)) and ke[-1] < ke[0], name
    # the closures only add dissipation
    assert res["runs"]["smagorinsky"]["ke"][-1] <= \
        res["runs"]["plain"]["ke"][-1]
