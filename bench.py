#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the per-timestep SPH pipeline on one GPU.

Workload (default): the 3D dam break of the reference's configuration 1 --
tank 1.6 x 0.6 x 0.6 m, water column 0.6 x 0.6 x 0.3 m, dr = 3 mm, h = 2 dr,
Wendland C2, linear EOS with c0 = 10 sqrt(g H), delta-SPH (delta = 0.1),
artificial viscosity alpha = 0.02 folded into visc_dyn, CFL 0.1, one layer of
boundary-integral wall elements, improved-Euler time integration with the
variable time step -- the preset stack basic + improved_euler + domain + cfd
+ deltaSPH + BI + variableTimeStep, built from the XML presets and run by
the C++/HIP runtime.  One "step" is one full time step of that tool list
(link-list + sort + every interaction + integration + dt reduction),
replayed from a captured HIP graph.  Inputs are synthetic (the lattice of
the case generator) and resident in HBM before timing starts.

    python bench.py [--steps K] [--warmup W] [--dr DR] [--no-cpu]

Prints one JSON line.  --gpus N > 1 is refused: this build targets one GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3


def build_sim(dr, dims=3):
    from aquagpusph_amd import casefiles, cases
    from aquagpusph_amd.sim import Simulation
    case = cases.dam_break(dims, dr)
    d = tempfile.mkdtemp(prefix="aq_bench_")
    xml = casefiles.write_case(case, d, load=False)
    sim = Simulation(xml, dims)
    casefiles.upload(sim, case)
    return sim, case


def cpu_baseline(case, budget_s=25.0):
    """FP32 oracle (OpenMP C++ restatement) on the same case: a bounded
    number of full steps, reported as particle-steps/s."""
    from oracle import pipeline as op
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sim = op.OracleSim(case)
    sim.step()  # the first step has dt = 0; untimed
    t0 = time.perf_counter()
    n = 0
    while True:
        sim.step()
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 3:
            break
    return dict(value=case["N"] * n / el, unit="particle-steps/s",
                cores=threads, kind="port",
                sample="%d full steps of the same %d-particle case (after one "
                       "untimed step), FP32 oracle, %d OpenMP threads, "
                       "%.1f s" % (n, case["N"], threads, el))


# compulsory (algorithmic) HBM bytes per particle for the cfd interactions
# kernel in 3D: read imove 4, r 16, u 16, rho 4, m 4, p 4, icell 4 and write
# grad_p 16, lap_u 16, div_u 4 (read-modify-write of the zeroed outputs
# counts once), i.e. 88 B/particle (92 with the ihoc reads amortised)
INTERACTIONS_BYTES = 88


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--dr", type=float, default=0.003)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=3)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not "
              "supported (see DESIGN.md for how the path would shard)"
              % a.gpus, file=sys.stderr)
        sys.exit(2)

    sim, case = build_sim(a.dr)
    N = sim.N
    nf = int(case["n_fluid"])
    # warmup (the first call runs one eager step, then captures the graph)
    sim.step(max(1, a.warmup), graph=True)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(a.steps, graph=True)
    sim.sync()
    el = time.perf_counter() - t0
    ms = 1e3 * el / a.steps
    value = N * a.steps / el

    # per-tool GPU time (HIP events on the runtime's stream)
    prof = sim.profile(a.profile_steps)
    by_name = {n: (t, m) for n, t, m in prof}
    t_int = by_name["cfd interactions"][1] * 1e-3
    step_gpu_ms = sum(m for _, _, m in prof)
    achieved = INTERACTIONS_BYTES * N / t_int / 1e9
    # FP32 work of the interactions launch (accepted pairs ~ 60 flop,
    # rejected candidates ~ 10 flop; 268 / 1728 per particle on the lattice)
    flops = N * (268 * 60 + (1728 - 268) * 10)
    top = sorted(prof, key=lambda x: -x[2])[:8]

    out = {
        "metric": "particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "3D dam break, tank 1.6x0.6x0.6 m, column "
                        "0.6x0.6x0.3 m, dr=%g m, h=2dr, BI walls, delta-SPH, "
                        "improved Euler, variable dt" % a.dr,
            "n_particles": N, "n_fluid": nf, "dr": a.dr,
            "presets": "basic+improved_euler+domain+cfd+deltaSPH+BI+"
                       "variableTimeStep",
            "parallelism": "single GPU, HIP graph per step"},
        "fluid_particle_steps_per_s": nf * a.steps / el,
        "roofline": {
            "kernel": "cfd interactions (cfd/Interactions.cl:entry)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": None,
            "avg_launch_ms": t_int * 1e3,
            "bytes_per_particle": INTERACTIONS_BYTES,
            "fp32_tflops_est": flops / t_int / 1e12,
            "fp32_peak_tflops": PEAK_FP32_TFLOPS,
        },
        "tools_ms": {n: round(m, 4) for n, _, m in top},
        "step_gpu_ms_sum": step_gpu_ms,
    }
    sim.close()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
