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
# MI355X FP32 vector peak as published (it counts packed FP32: two FMAs per
# lane and cycle); the walk is built without packed FP32, so one FMA per
# lane and cycle, half of it, is what its instructions can reach
PEAK_FP32_TFLOPS = 157.3
PEAK_FP32_NONPACKED_TFLOPS = PEAK_FP32_TFLOPS / 2


def build_sim(dr, dims=3, precision=False):
    from aquagpusph_amd import casefiles, cases
    from aquagpusph_amd.sim import Simulation
    case = cases.dam_break(dims, dr)
    d = tempfile.mkdtemp(prefix="aq_bench_")
    pres = casefiles.PBI_PRESETS + (
        ("resources/Presets/precision.xml",) if precision else ())
    xml = casefiles.write_case(case, d, load=False, presets=pres)
    sim = Simulation(xml, dims)
    casefiles.upload(sim, case)
    return sim, case


# water-column resolutions giving ~1M and ~16M fluid particles: the sizes
# the particle-steps/s metric is quoted at
METRIC_SIZES = ((0.00476, "1M"), (0.00189, "16M"))


def size_point(dr, steps=10, warmup=2, profile_steps=2):
    """particle-steps/s of the same pipeline at another resolution (HIP
    graph steps, inputs resident; the same timing as the main line), with
    the fused walk's roofline at that size and the share of the step the
    walk and the other tools take."""
    sim, case = build_sim(dr)
    sim.step(warmup, graph=True)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(steps, graph=True)
    sim.sync()
    el = time.perf_counter() - t0
    ms = 1e3 * el / steps
    out = {"dr": dr, "n_particles": sim.N, "n_fluid": int(case["n_fluid"]),
           "steps": steps, "ms_per_step": ms,
           "particle_steps_per_s": sim.N * steps / el,
           "fluid_particle_steps_per_s": int(case["n_fluid"]) * steps / el}
    prof = sim.profile(profile_steps)
    walk = sim.walk_ms()
    gpu_ms = sum(m for _, _, m in prof)
    sim.close()
    if walk:
        pc = pair_counts(case)
        flops = (pc["cand"] * FLOP_CAND + pc["pairs"] * FLOP_PAIR
                 + pc["elem_pairs"] * FLOP_ELEM)
        ach = flops / (walk * 1e-3) / 1e12
        out["roofline"] = {
            "kernel": "fused fluid walk", "bound": "valu", "achieved": ach,
            "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / PEAK_FP32_TFLOPS,
            "frac_non_packed": ach / PEAK_FP32_NONPACKED_TFLOPS,
            "avg_launch_ms": walk,
            "hbm_achieved_gbs": WALK_BYTES * out["n_particles"]
            / (walk * 1e-3) / 1e9}
    # where the rest of the step goes: eager per-tool GPU time (summed
    # event intervals) against the graph-replayed step
    out["walk_ms"] = walk
    out["tools_gpu_ms"] = gpu_ms
    out["launch_gaps_ms"] = max(0.0, ms - gpu_ms)
    return out


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform string."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


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
                cores=threads, kind="port", cpu=cpu_model(),
                sample="%d full steps of the same %d-particle case (after one "
                       "untimed step), FP32 oracle, %d OpenMP threads, "
                       "%.1f s" % (n, case["N"], threads, el))


# Algorithmic figures of the fused fluid walk (one launch per step, at the
# position of the Shepard tool; it evaluates Shepard, the cfd interactions,
# lap p and the near-wall flags, see FusionPlan in csrc/runtime/sim.hpp).
# Compulsory HBM bytes per particle (3D, float4 vec): reads imove 4, r 16,
# u 16, rho 4, m 4, p 4, icell 4, grad_p/lap_u/div_u/lap_p 40 (the zeroed
# accumulators), normal/tangent/binormal 48 of the wall elements (counted
# for every particle: an upper bound), writes shepard 4, grad_p 16, lap_u 16,
# div_u 4, lap_p 4, near 1.
WALK_BYTES = 4 + 16 + 16 + 4 + 4 + 4 + 4 + 40 + 48 + 4 + 16 + 16 + 4 + 4 + 1
# FP32 operations (sqrt and divide counted as one) per candidate distance
# test (r_ij 3, |r_ij|^2 5, compare 1) and per accepted fluid pair (q 2,
# F(q) 6, u_ij 3, u.r 5, grad_p 10, Monaghan lap_u 15, div_u 4, lap_p 4);
# a wall element in the support costs about 100 (S_P, S_D with 4 acospi).
FLOP_CAND = 9
FLOP_PAIR = 49
FLOP_ELEM = 100

PROFILE_DIR = os.path.join(HERE, "profiles")


def pair_counts(case, sample=200000):
    """Candidates (particles in the 3^d cells around a fluid particle's cell,
    what the walk tests) exactly from the cell occupancy, and accepted
    fluid-fluid pairs (q < 2, j != i) and fluid-element pairs from a
    strided sample of the fluid particles (k-d tree counts)."""
    from scipy.spatial import cKDTree
    dims = case["dims"]
    r = np.asarray(case["r"], np.float64)[:, :dims]
    im = np.asarray(case["imove"])
    h = float(case["params"]["h"])
    cell = 2.0 * h
    rmin = r.min(0)
    ijk = np.floor((r - rmin) / cell).astype(np.int64) + 1
    dim = ijk.max(0) + 2
    flat = np.ravel_multi_index(ijk.T, dim)
    occ = np.bincount(flat, minlength=int(np.prod(dim)))
    fl = np.nonzero(im == 1)[0]
    cand = np.zeros(len(fl), np.int64)
    offs = np.stack(np.meshgrid(*[[-1, 0, 1]] * dims, indexing="ij"),
                    -1).reshape(-1, dims)
    for o in offs:
        cand += occ[np.ravel_multi_index((ijk[fl] + o).T, dim)]
    step = max(1, len(fl) // sample)
    q = r[fl[::step]]
    rad = 2.0 * h * (1 - 1e-6)
    t_f = cKDTree(r[im == 1])
    n_f = t_f.query_ball_point(q, rad, return_length=True, workers=-1) - 1
    n_e = np.zeros_like(n_f)
    if (im == -3).any():
        t_e = cKDTree(r[im == -3])
        n_e = t_e.query_ball_point(q, rad, return_length=True, workers=-1)
    return dict(n_fluid=len(fl), cand=int(cand.sum()),
                pairs=float(n_f.mean()) * len(fl),
                elem_pairs=float(n_e.mean()) * len(fl),
                sample_stride=step)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor),
    or None when no summary is committed."""
    path = os.path.join(PROFILE_DIR, kernel + "_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def streaming_counters(dr):
    """The committed HBM counters of the per-particle passes, the sort and
    the link-list kernels on the 3D dam break at dr = 3 mm (rocprofv3 --pmc
    FETCH_SIZE and WRITE_SIZE runs, copy-kernel scaled): {kernel: {MB per
    launch, us, TB/s}}, or None at another size or without the summary."""
    path = os.path.join(PROFILE_DIR, "streaming_kernels_dambreak3d_pmc.md")
    if abs(dr - 0.003) > 1e-9 or not os.path.exists(path):
        return None
    out = {}
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) != 9 or not cells[0].startswith("`"):
            continue
        name = cells[0].strip("`").replace("void ", "")
        try:
            out[name] = {"traffic_mb": float(cells[4]),
                         "avg_us": float(cells[6]),
                         "tbps": float(cells[7])}
        except ValueError:
            continue
    return out or None


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain device-to-device copy of known size, for scaling the
    FETCH_SIZE / WRITE_SIZE counters (run under rocprofv3 --pmc)."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                      "reps": reps}))


def batch_bench(n_tanks=256, res=200, steps=20, warmup=3, sample=8,
                tune=True):
    """The batched-ensemble workload: n_tanks independent excited 2D
    sloshing tanks (widths 0.5-1.5 m, fill 0.2-0.6, dr = width/200, BI
    walls, delta-SPH, per-tank dt) in one case (Presets/batch.xml: one
    link-list, one walk launch for every tank), after the walk autotuner;
    beside it, `sample` of the tanks run one at a time, scaled to all."""
    from aquagpusph_amd import casefiles, cases
    from aquagpusph_amd.sim import Simulation
    b = cases.sloshing_batch(n_tanks, res=res)
    d = tempfile.mkdtemp(prefix="aq_batch_")
    xml = casefiles.write_case(
        b, d, load=False,
        presets=casefiles.PBI_PRESETS + ("resources/Presets/batch.xml",))
    sim = Simulation(xml, 2)
    casefiles.upload(sim, b)
    tuning = sim.autotune(2) if tune else []
    sim.step(max(1, warmup), graph=True)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(steps, graph=True)
    sim.sync()
    el = time.perf_counter() - t0
    N = sim.N
    dts = np.asarray(sim.get("dt_set"), np.float64)
    sim.close()
    # the same tanks one case at a time (a sample, scaled to all)
    alone = []
    for k in range(min(sample, n_tanks)):
        case = cases.batch_member(b, k)
        dk = tempfile.mkdtemp(prefix="aq_tank_")
        one = Simulation(casefiles.write_case(
            case, dk, load=False, extra_xml=cases.excitation_xml(case)), 2)
        casefiles.upload(one, case)
        one.step(max(1, warmup), graph=True)
        one.sync()
        t1 = time.perf_counter()
        one.step(steps, graph=True)
        one.sync()
        alone.append((case["N"], 1e3 * (time.perf_counter() - t1) / steps))
        one.close()
    ms_alone = np.mean([m for _, m in alone]) * n_tanks
    return {
        "metric": "particle-steps/s",
        "value": N * steps / el,
        "unit": "particle-steps/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * el / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "%d 2D sloshing tanks, widths 0.5-1.5 m, "
                               "fill 0.2-0.6, dr=width/%d, BI walls, "
                               "delta-SPH, harmonic excitation, per-tank dt"
                               % (n_tanks, res),
                   "n_particles": N,
                   "tank_particles_min_max": [min(n for n, _ in b["sets"]),
                                              max(n for n, _ in b["sets"])],
                   "presets": "basic+improved_euler+domain+cfd+deltaSPH+"
                              "BI+variableTimeStep+batch",
                   "parallelism": "single GPU, one launch per tool for "
                                  "every tank, HIP graph per step"},
        "dt_set_min_max": [float(dts.min()), float(dts.max())],
        "autotune": tuning,
        "one_at_a_time": {"sampled_tanks": len(alone),
                          "ms_per_step_each": [round(m, 4)
                                               for _, m in alone],
                          "ms_per_step_all_tanks_est": ms_alone,
                          "batched_speedup": ms_alone / (1e3 * el / steps)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--dr", type=float, default=0.003)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--no-sizes", action="store_true",
                    help="skip the ~1M and ~16M fluid-particle points")
    ap.add_argument("--precision", action="store_true",
                    help="add Presets/precision.xml (compensated walk sums, "
                         "double positions) to the workload")
    ap.add_argument("--batch", type=int, default=0, metavar="TANKS",
                    help="run the batched sloshing-tank ensemble of TANKS "
                         "tanks instead (prints its own line)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB copy kernel (for the --pmc "
                         "counter scale) and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not "
              "supported (one rank per GPU, each owning a slab of cells "
              "with halo exchange, is the intended sharding)"
              % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration()
        return
    if a.batch:
        print(json.dumps(batch_bench(a.batch, steps=a.steps,
                                     warmup=a.warmup)))
        return

    sim, case = build_sim(a.dr, precision=a.precision)
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

    # per-tool GPU time: HIP events recorded on the runtime's stream around
    # every tool, over eager steps run right after the timed region
    prof = sim.profile(a.profile_steps)
    step_gpu_ms = sum(m for _, _, m in prof)
    top = sorted(prof, key=lambda x: -x[2])
    dom_name, _, dom_ms = top[0]
    by_name = {n: m for n, _, m in prof}
    fused = (dom_name == "cfd Shepard"
             and by_name.get("cfd interactions", 0.0) < 0.01 * dom_ms)
    pc = pair_counts(case)
    walk_ms = sim.walk_ms()
    t = (walk_ms if (fused and walk_ms) else dom_ms) * 1e-3
    if fused:
        kname = "fused fluid walk (cfd Shepard + cfd interactions + lap p)"
        flops = (pc["cand"] * FLOP_CAND + pc["pairs"] * FLOP_PAIR
                 + pc["elem_pairs"] * FLOP_ELEM)
        hbm = WALK_BYTES * N
    else:
        kname = dom_name
        flops = None
        hbm = None
    traffic = measured_traffic("fused_walk_dambreak3d") if fused else None
    achieved = flops / t / 1e12 if flops else None
    roof = {
        "kernel": kname,
        "bound": "valu",
        "achieved": achieved,
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP32_TFLOPS if flops else None,
        "traffic": traffic,
        "traffic_source": "profiles/fused_walk_dambreak3d_pmc.md "
                          "(rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE of this "
                          "kernel on this workload, copy-kernel scaled)",
        "peak_non_packed": PEAK_FP32_NONPACKED_TFLOPS,
        "frac_non_packed": (achieved / PEAK_FP32_NONPACKED_TFLOPS
                            if flops else None),
        "note": "FP32 vector ALU work, no matrix instructions (bound: the "
                "VALU). peak is the published MI355X FP32 vector figure, "
                "which counts packed FP32; this build emits none, so "
                "peak_non_packed (one FMA per lane and cycle) is the "
                "attainable roof. The walk is compute-bound: %.0f FLOP per "
                "%d compulsory bytes, IEEE division and square root"
                % (flops / N if flops else 0, WALK_BYTES),
        "avg_launch_ms": t * 1e3,
        "tool_ms": dom_ms,
        "flop_per_launch": flops,
        "candidates_per_fluid_particle": pc["cand"] / pc["n_fluid"],
        "pairs_per_fluid_particle": pc["pairs"] / pc["n_fluid"],
        "elements_per_fluid_particle": pc["elem_pairs"] / pc["n_fluid"],
        "hbm_algorithmic_bytes": hbm,
        "hbm_achieved_gbs": hbm / t / 1e9 if hbm else None,
        "hbm_frac": hbm / t / 1e9 / PEAK_HBM_GBS if hbm else None,
        "pair_interactions_per_s": pc["pairs"] / t,
    }

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
        "dtype": "f32" if not a.precision else "f32 (f64 positions)",
        "data": "synthetic",
        "config": {
            "workload": "3D dam break, tank 1.6x0.6x0.6 m, column "
                        "0.6x0.6x0.3 m, dr=%g m, h=2dr, BI walls, delta-SPH, "
                        "improved Euler, variable dt" % a.dr,
            "n_particles": N, "n_fluid": nf, "dr": a.dr,
            "presets": "basic+improved_euler+domain+cfd+deltaSPH+BI+"
                       "variableTimeStep" + ("+precision" if a.precision
                                             else ""),
            "parallelism": "single GPU, HIP graph per step"},
        "fluid_particle_steps_per_s": nf * a.steps / el,
        "roofline": roof,
        "tools_ms": {n: round(m, 4) for n, _, m in top[:8]},
        "step_gpu_ms_sum": step_gpu_ms,
        "hbm_counters": streaming_counters(a.dr),
        "hbm_counters_source": "profiles/streaming_kernels_dambreak3d_pmc.md "
                               "(rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of "
                               "each kernel on this workload, copy-kernel "
                               "scaled; not measured in this run)",
    }
    sim.close()
    if not a.no_sizes:
        out["metric_sizes"] = {name: size_point(dr)
                               for dr, name in METRIC_SIZES}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
