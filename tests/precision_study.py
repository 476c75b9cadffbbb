"""Precision study: the 2D dam break (tank 1.61 x 0.6 m, water column
0.6 x 0.3 m, dr = 3 mm, 20,000 fluid particles) or the 3D one (tank
1.6 x 0.6 x 0.6 m, column 0.6 x 0.6 x 0.3 m, --dims 3 --dr DR), BI walls,
delta-SPH, improved Euler with the variable time step, run three ways from
the same float32 start:

  plain      the GPU pipeline in float32;
  precision  the GPU pipeline with Presets/precision.xml (compensated walk
             sums, double positions r + dr64, double mass / energy sums);
  fp64       the oracle's float64 instantiation on the host (the reference
             for both; there is no float64 GPU pipeline).

Every `every` steps it records, per run: the time, the surge front (largest
x of the fluid), the fluid mass, the mechanical energy sum m (|u|^2/2 + g y)
and, against fp64, the RMS and largest position and velocity differences.

    python tests/precision_study.py [--dims 2|3] [--dr DR] [--steps 2000]
                                    [--every 100] [--out F]

prints one JSON line (and writes it to --out).  tests/test_precision_gpu.py
runs a short version and checks that the precision run stays closer to
fp64 than the plain one.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from aquagpusph_amd import casefiles, cases  # noqa: E402
from oracle import pipeline as op  # noqa: E402

PRECISION = "resources/Presets/precision.xml"


def _state(r, u, m, imove, ids, n, g):
    """Per-particle arrays in the original order and the diagnostics."""
    R = np.zeros((n, r.shape[1]))
    V = np.zeros((n, u.shape[1]))
    M = np.zeros(n)
    F = np.zeros(n, bool)
    R[ids] = r
    V[ids] = u
    M[ids] = m
    F[ids] = imove == 1
    d = r.shape[1] if r.shape[1] == 2 else 3
    ek = 0.5 * np.sum(M[F] * np.sum(V[F, :d] ** 2, axis=1))
    ep = -np.sum(M[F] * (R[F, :d] @ np.asarray(g[:d], float)))
    return dict(r=R, u=V, fluid=F, front=float(R[F, 0].max()),
                mass=float(M[F].sum()), e_mech=float(ek + ep),
                e_kin=float(ek))


def run_study(steps=2000, every=100, dr=0.003, log=None, dims=2):
    from aquagpusph_amd.sim import Simulation
    case = cases.dam_break(dims, dr)
    # every run starts from the same float32 values
    for k in ("r", "normal", "tangent", "u", "dudt", "rho", "drhodt", "m"):
        case[k] = np.asarray(case[k]).astype(np.float32).astype(np.float64)
    n = int(case["N"])
    g = np.array(case["params"]["g"], float)
    tmp = tempfile.mkdtemp(prefix="aq_prec_")
    sims = {
        "plain": Simulation(casefiles.write_case(
            case, os.path.join(tmp, "plain")), dims),
        "precision": Simulation(casefiles.write_case(
            case, os.path.join(tmp, "prec"),
            presets=casefiles.PBI_PRESETS + (PRECISION,)), dims),
    }
    ref = op.OracleSim(case, dtype=np.float64)
    rows = []
    done = 0
    t0 = time.time()
    while done < steps:
        k = min(every, steps - done)
        for s in sims.values():
            s.step(k, graph=True)
        ref.run(k)
        done += k
        f = ref.f
        st = {"fp64": _state(f["r"], f["u"], f["m"], f["imove"],
                             f["id"].astype(np.int64), n, g)}
        st["fp64"]["t"] = float(ref.t)
        for name, s in sims.items():
            r = s.get("r").astype(np.float64)
            if name == "precision":
                r = r + s.get("dr64")
            st[name] = _state(r, s.get("u").astype(np.float64),
                              s.get("m").astype(np.float64), s.get("imove"),
                              s.get("id").astype(np.int64), n, g)
            st[name]["t"] = float(s.get("t"))
            if name == "precision":
                st[name]["mass_fp64"] = float(s.get("mass_fp64"))
        row = {"step": done}
        b = st["fp64"]
        for name in ("plain", "precision", "fp64"):
            a = st[name]
            e = {"t": a["t"], "front": a["front"], "mass": a["mass"],
                 "e_mech": a["e_mech"]}
            if "mass_fp64" in a:
                e["mass_fp64"] = a["mass_fp64"]
            if name != "fp64":
                F = b["fluid"] & a["fluid"]
                dx = np.linalg.norm(a["r"][F] - b["r"][F], axis=1)
                du = np.linalg.norm(a["u"][F] - b["u"][F], axis=1)
                e.update(pos_rms=float(np.sqrt(np.mean(dx ** 2))),
                         pos_max=float(dx.max()),
                         vel_rms=float(np.sqrt(np.mean(du ** 2))),
                         front_err=a["front"] - b["front"],
                         e_mech_err=(a["e_mech"] - b["e_mech"]) /
                         max(b["e_kin"], 1e-30))
            row[name] = e
        rows.append(row)
        if log:
            log("step %d (%.0f s): pos_rms plain %.3g precision %.3g" %
                (done, time.time() - t0, row["plain"]["pos_rms"],
                 row["precision"]["pos_rms"]))
    for s in sims.values():
        s.close()
    e0 = rows[0]["fp64"]["e_mech"]
    return {
        "study": "precision",
        "workload": ("2D dam break, tank 1.61x0.6 m, column 0.6x0.3 m, "
                     if dims == 2 else
                     "3D dam break, tank 1.6x0.6x0.6 m, column 0.6x0.6x0.3 "
                     "m, ") + "dr=%g m, BI walls, delta-SPH, improved Euler, "
                    "variable dt" % dr,
        "n_particles": n, "n_fluid": int(case["n_fluid"]), "steps": steps,
        "reference": "oracle float64 instantiation on the host (no float64 "
                     "GPU pipeline)",
        "host_threads": os.environ.get("OMP_NUM_THREADS"),
        "e_mech_fp64_first_sample": e0,
        "final": rows[-1], "series": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--out", default=None)
    ap.add_argument("--dims", type=int, default=2)
    ap.add_argument("--dr", type=float, default=0.003)
    a = ap.parse_args()
    res = run_study(a.steps, a.every, dr=a.dr, dims=a.dims,
                    log=lambda s: print(s, file=sys.stderr, flush=True))
    line = json.dumps(res)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")
    print(line)


if __name__ == "__main__":
    main()
