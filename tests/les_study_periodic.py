"""LES study on decaying turbulence in a triply periodic box (the learned
closure's configuration: a cube of side 1 periodic along every axis,
Presets/cfd/periodic.xml, no walls; a seeded solenoidal field with
E(k) ~ k^4 exp(-2 (k/k0)^2), k0 = 4, u_rms = 1, the same continuous field
sampled by the coarse and the fine lattice, cases.solenoidal_field).

  1. training flow (seed train_seed): a fine run (2n per axis, plain
     SPH), the reference;
  2. the shape of the closure, a priori: at `snapshots` times the fine
     velocities are filtered by the coarse SPH kernel (Wendland, support
     2 H) at the coarse lattice points: u_bar and
     tau = (u u)_bar - u_bar u_bar; the SPH gradient of u_bar there
     (computed on the GPU by the closure's own operator, one step of a
     coarse run with u = u_bar and dt = 0) gives the network inputs and
     S_bar; the target coefficient is
     c* = -(tau_d : S_bar) / (2 S_bar:S_bar) / (H^2 |S_bar|), tau_d the
     deviatoric part, clipped to [0, 2], each sample weighted by
     4 H^4 |S_bar|^2 S_bar:S_bar (the fit of the eddy-viscosity stress to
     tau_d in stress space), with the least-squares constant c_lsq of that
     fit; training on the GPU with the library's BF16 network kernels;
  3. the size of the closure, a posteriori: the coarse particles carry
     motion below the kernel scale that the filtered field does not, so
     the a priori coefficient overstates what the coarse run lacks; each
     closure's output is scaled (golden-section search over the scale, one
     coarse run of the training flow per evaluation) to bring the coarse
     run's kinetic energy closest to the fine run's, and so is the
     Smagorinsky coefficient (C_s^2);
  4. evaluation flow (seed eval_seed, not seen in training): a fine run and
     the coarse runs: plain SPH, the network unscaled (a priori) and
     scaled (cfd/LES.xml), Smagorinsky (cfd/LES-smagorinsky.xml)
     at C_s = 0.15 and at the fitted coefficient;
  5. the kinetic energy of every run at the same times, its error against
     the fine run (mean over the samples of |E - E_fine| / E_fine(0)), and
     the GPU time of every run.

    python tests/les_study_periodic.py [--dims 3] [--n 128] [--time 0.2]
                [--train-seed 0] [--eval-seed 1] [--out F]

prints one JSON line.  tests/test_les_gpu.py runs a small version.
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
sys.path.insert(0, HERE)

from aquagpusph_amd import casefiles, cases  # noqa: E402
from oracle import les as olq  # noqa: E402

from les_study import train  # noqa: E402

MORRIS = ('    <Definitions><Define name="__LAP_FORMULATION__" '
          'value="__LAP_MORRIS__" evaluate="false"/></Definitions>')
LES = "resources/Presets/cfd/LES.xml"
# the periodic pipeline at the Courant step courant h / cs (no velocity
# limited step: u_max / cs < 0.4 here, and the step is the coarse runs' and
# the fine run's alike)
PRESETS = tuple(p for p in casefiles.PPER_PRESETS
                if not p.endswith("variableTimeStep.xml"))
SMAG = "resources/Presets/cfd/LES-smagorinsky.xml"


def _sim(case, extra=(), params=None, cmin=None, cs=None):
    from aquagpusph_amd.sim import Simulation
    d = tempfile.mkdtemp(prefix="aq_lesp_")
    xml = casefiles.write_case(case, d,
                               presets=PRESETS + tuple(extra),
                               load=False, extra_xml=MORRIS)
    sim = Simulation(xml, case["dims"])
    casefiles.upload(sim, case)
    if params is not None:
        sim.set("les_params", params)
    if cmin is not None:
        sim.set("les_cmin", np.array([cmin], np.float32))
    if cs is not None:
        sim.set("les_cs", np.array([cs], np.float32))
    return sim


def _ke(sim):
    u = sim.get("u").astype(np.float64)
    m = sim.get("m").astype(np.float64)
    return float(0.5 * np.sum(m * np.sum(u ** 2, axis=1)))


def run(case, samples, steps_per_sample, extra=(), params=None, snap=None,
        cmin=None, log=None, cs=None):
    """KE at every sample, GPU seconds of the stepping, and (snap: the
    sample indices) particle snapshots (r, u)."""
    sim = _sim(case, extra, params, cmin, cs)
    ke = [_ke(sim)]
    gpu_s = 0.0
    snaps = []
    for k in range(samples):
        t0 = time.perf_counter()
        sim.step(steps_per_sample, graph=True)
        sim.sync()
        gpu_s += time.perf_counter() - t0
        ke.append(_ke(sim))
        if log:
            log("  sample %d/%d: t = %.4f, KE = %.6g (%.1f s)"
                % (k + 1, samples, float(sim.get("t")), ke[-1], gpu_s))
        if snap is not None and (k + 1) in snap:
            snaps.append({f: sim.get(f).copy() for f in ("r", "u")})
    t = float(sim.get("t"))
    sim.close()
    return np.array(ke), gpu_s, snaps, t


def wendland(q, dims):
    """The Wendland kernel of the pipeline (support 2 h) up to its
    normalisation: (1 + 2q)(1 - q/2)^4 for q < 2."""
    return np.where(q < 2.0, (1.0 + 2.0 * q) * (1.0 - 0.5 * q) ** 4, 0.0)


def kernel_filter(snap, n, dims, hc):
    """u_bar and tau at the n^dims coarse lattice points (cell centres of
    the unit box, x slowest) filtered from the fine particles by the coarse
    SPH kernel (Wendland, support 2 hc), Shepard-normalised over the
    particles (equal volumes):
        u_bar(x) = sum_j W(x - x_j) u_j / sum_j W(x - x_j),
        tau = (u u)_bar - u_bar u_bar.
    The particles are deposited (cloud in cell: multilinear weights on the
    2^dims surrounding nodes) on a periodic grid of spacing 1/(2n), whose
    odd nodes are the coarse points, and the kernel sums are periodic
    convolutions done by FFT.  Returns (u_bar, tau, the kernel-weighted
    particles per grid node: 1 on the initial lattice)."""
    import scipy.fft as sfft
    M = 2 * n
    d = 1.0 / M
    r = snap["r"][:, :dims].astype(np.float64)
    u = snap["u"][:, :dims].astype(np.float64)
    x = r / d
    base = np.floor(x).astype(np.int64)
    frac = x - base
    corners = []
    for k in range(2 ** dims):
        bits = [(k >> a) & 1 for a in range(dims)]
        node = tuple((base[:, a] + bits[a]) % M for a in range(dims))
        wc = np.prod([frac[:, a] if bits[a] else 1.0 - frac[:, a]
                      for a in range(dims)], axis=0)
        corners.append((np.ravel_multi_index(node, (M,) * dims), wc))
    K = int(np.ceil(2.0 * hc / d))
    off = np.arange(-K, K + 1)
    grids = np.meshgrid(*([off * d] * dims), indexing="ij")
    q = np.sqrt(sum(g ** 2 for g in grids)) / hc
    wk = np.zeros((M,) * dims)
    idx = np.ix_(*([off % M] * dims))
    wk[idx] = wendland(q, dims)
    wk /= wk.sum()
    Wf = sfft.rfftn(wk, workers=-1)
    pick = (slice(1, None, 2),) * dims

    def conv(vals):
        g = sum(np.bincount(f, vals * wc, minlength=M ** dims)
                for f, wc in corners).reshape((M,) * dims)
        c = sfft.irfftn(sfft.rfftn(g, workers=-1) * Wf, s=(M,) * dims,
                        workers=-1)
        return c[pick].reshape(-1)

    w = conv(np.ones(len(r)))
    ub = np.stack([conv(u[:, a]) for a in range(dims)], 1) / w[:, None]
    tau = np.zeros((len(w), dims, dims))
    for a in range(dims):
        for b in range(a, dims):
            t = conv(u[:, a] * u[:, b]) / w - ub[:, a] * ub[:, b]
            tau[:, a, b] = t
            tau[:, b, a] = t
    return ub, tau, w


def coarse_gradient(coarse, u, r=None):
    """The SPH velocity gradient of the field u at the coarse lattice
    points (or at the positions r): one step of a coarse run (its first
    step has dt = 0: nothing moves) with the closure's gradient operator;
    rows in the case order."""
    dims = coarse["dims"]
    c = dict(coarse)
    c["u"] = coarse["u"].copy()
    c["u"][:, :dims] = u[:, :dims]
    if r is not None:
        c["r"] = coarse["r"].copy()
        c["r"][:, :dims] = r[:, :dims]
    sim = _sim(c, (SMAG,))
    sim.step(1)
    G = sim.get("grad_u").astype(np.float64)
    ids = sim.get("id").astype(np.int64)
    sim.close()
    out = np.zeros_like(G)
    out[ids] = G
    return out


def _strain(G, dims):
    Gm = olq.matrices(G, dims)
    S = 0.5 * (Gm + np.transpose(Gm, (0, 2, 1)))
    return S, np.einsum("nab,nab->n", S, S)


def training_samples(snaps, coarse, per_snapshot=100000, seed=0):
    """Network inputs [M, 8], target c* [M] and the stress-space weights
    [M] from the fine snapshots.  The eddy-viscosity stress
    -2 c H^2 |S| S closest to the deviatoric subgrid stress tau_d,
    |tau_d + 2 c H^2 |S| S|^2 = w (c - c*)^2 + const, has the weight
    w = 4 H^4 |S|^2 S:S and c* = -(tau_d : S) / (2 H^2 |S| S:S); also the
    single least-squares constant c_lsq = sum w c* / sum w, and that of
    every snapshot."""
    dims = coarse["dims"]
    n = int(round(coarse["N"] ** (1.0 / dims)))
    hc = coarse["params"]["h"]
    rng = np.random.default_rng(seed)
    X, Y, Wt, per = [], [], [], []
    num = den = 0.0
    for s in snaps:
        ub, tau, cnt = kernel_filter(s, n, dims, hc)
        G = coarse_gradient(coarse, ub)
        S, ss = _strain(G, dims)
        tr = np.trace(tau, axis1=1, axis2=2)
        td = tau - tr[:, None, None] * np.eye(dims)[None] / dims
        smag = np.sqrt(2.0 * ss)
        tds = np.einsum("nab,nab->n", td, S)
        c = -tds / (2.0 * hc * hc * smag * ss + 1e-300)
        w = 4.0 * hc ** 4 * smag ** 2 * ss
        nk = float(np.sum(-2.0 * hc * hc * smag * tds))
        dk = float(np.sum(w))
        per.append(nk / dk)
        num += nk
        den += dk
        idx = np.nonzero(smag > 1e-6 * smag.max())[0]
        if len(idx) > per_snapshot:
            idx = rng.choice(idx, per_snapshot, replace=False)
        X.append(olq.features(G[idx], dims, coarse["params"]["hfac"]))
        Y.append(c[idx])
        Wt.append(w[idx])
    Y = np.concatenate(Y)
    Wt = np.concatenate(Wt)
    # a handful of the strongest strains would carry the whole weight
    Wt = np.minimum(Wt, np.quantile(Wt, 0.999))
    return np.concatenate(X).astype(np.float32), Y, Wt, num / den, per


def _scaled(P, a):
    """The network with its output multiplied by a (the output layer is
    linear)."""
    from oracle import mlp as omlp
    Q = np.array(P, np.float32)
    Q[omlp.OFF_W3:omlp.OFF_B3 + 1] *= np.float32(a)
    return Q


def fit_scale(objective, lo, hi, evals, log=None):
    """Golden-section search of objective(a) over log10 a in [lo, hi]
    (objective(0) is the plain run's error, the search keeps it as the
    bound); the best a and the (a, error) pairs evaluated."""
    g = (np.sqrt(5.0) - 1.0) / 2.0
    hist = []

    def f(x):
        e = objective(10.0 ** x)
        hist.append((10.0 ** x, e))
        if log:
            log("    scale %.4g: error %.4f" % (10.0 ** x, e))
        return e

    a, b = lo, hi
    c, d = b - g * (b - a), a + g * (b - a)
    fc, fd = f(c), f(d)
    for _ in range(evals - 2):
        if fc < fd:
            b, d, fd = d, c, fc
            c = b - g * (b - a)
            fc = f(c)
        else:
            a, c, fc = c, d, fd
            d = a + g * (b - a)
            fd = f(d)
    best = min(hist, key=lambda t: t[1])
    return best[0], hist


def study(dims=3, n=128, t_end=0.2, samples=10, snapshots=4, iters=400,
          train_seed=0, eval_seed=1, evals=6, log=None):
    """Training on the flow of train_seed (fine run, plain coarse run),
    evaluation on the flow of eval_seed (a fresh fine run and the coarse
    runs): the kinetic-energy errors are those of a realisation the
    closures have not seen."""
    say = log or (lambda s: None)
    coarse = cases.periodic_box(dims, n, seed=train_seed)
    hc = coarse["params"]["h"]
    dtc = coarse["params"]["courant"] * hc / coarse["params"]["cs"]
    spc = max(1, int(round(t_end / dtc / samples)))
    snap_at = sorted(set(int(k) for k in np.linspace(1, samples, snapshots)
                         .round()))
    times = np.arange(samples + 1) * spc * dtc
    fine = cases.periodic_box(dims, 2 * n, seed=train_seed)
    say("fine run (training flow): %d particles, %d steps"
        % (fine["N"], 2 * spc * samples))
    ke_ft, t_ft, snaps, _ = run(fine, samples, 2 * spc, snap=set(snap_at),
                                log=say)
    del fine
    say("training samples")
    X, Y, Wt, c_lsq, c_lsq_per = training_samples(snaps, coarse)
    del snaps
    say("a priori least-squares c = %.5g (per snapshot %s)"
        % (c_lsq, ", ".join("%.4g" % c for c in c_lsq_per)))
    Yp = np.clip(Y, 0.0, 2.0).astype(np.float32)
    say("training on %d samples" % len(Y))
    P, losses, t_train = train(X, Yp, iters=iters, weights=Wt)

    def err_on_training_flow(extra, params=None, cs=None):
        ke = run(coarse, samples, spc, extra, params, cs=cs)[0]
        return float(np.mean(np.abs(ke - ke_ft)) / ke_ft[0])

    # the size of each closure, a posteriori: the scale of its coefficient
    # that brings the coarse run's kinetic energy closest to the fine
    # run's on the training flow
    say("scale of the learned closure")
    a_l, h_l = fit_scale(lambda a: err_on_training_flow(
        (LES,), _scaled(P, a)), -3.0, 0.0, evals, say)
    say("Smagorinsky coefficient")
    c_s, h_s = fit_scale(lambda c: err_on_training_flow(
        (SMAG,), cs=float(np.sqrt(c))), -4.0, -1.0, evals, say)
    if eval_seed == train_seed:
        ke_f, t_f = ke_ft, t_ft
        ev = coarse
    else:
        fine = cases.periodic_box(dims, 2 * n, seed=eval_seed)
        say("fine run (evaluation flow)")
        ke_f, t_f, _, _ = run(fine, samples, 2 * spc, log=say)
        del fine
        ev = cases.periodic_box(dims, n, seed=eval_seed)
    out = {}
    for name, extra, params, cm, cs in (
            ("plain", (), None, None, None),
            ("learned_a_priori", (LES,), P, None, None),
            ("learned", (LES,), _scaled(P, a_l), None, None),
            ("smagorinsky", (SMAG,), None, None, None),
            ("smagorinsky_fitted", (SMAG,), None, None,
             float(np.sqrt(c_s)))):
        say("coarse run: %s" % name)
        ke, t, _, tc = run(ev, samples, spc, extra, params, cmin=cm,
                           log=say, cs=cs)
        out[name] = dict(ke=ke.tolist(), gpu_s=t, t_end=tc,
                         ke_err=float(np.mean(np.abs(ke - ke_f)) / ke_f[0]))
        say("  %s: KE(end) %.5f (fine %.5f), error %.4f, %.2f GPU s"
            % (name, ke[-1], ke_f[-1], out[name]["ke_err"], t))
    out["fine"] = dict(ke=list(map(float, ke_f)), gpu_s=t_f)
    return {
        "study": "les_periodic",
        "workload": "decaying turbulence in a %dD box periodic along every "
                    "axis, side 1, E(k) ~ k^4 exp(-2 (k/k0)^2), k0 = 4, "
                    "u_rms = 1, nu = 1e-4, coarse %d^%d, fine %d^%d, "
                    "t = %.3g s, trained on seed %d, evaluated on seed %d"
                    % (dims, n, dims, 2 * n, dims, t_end, train_seed,
                       eval_seed),
        "samples": samples, "steps_per_sample_coarse": spc,
        "times": times.tolist(),
        "training": {"samples": int(len(Y)), "snapshots": snap_at,
                     "ke_fine": list(map(float, ke_ft)),
                     "c_lsq": c_lsq, "c_lsq_per_snapshot": c_lsq_per,
                     "target_weighted_mean": float(np.dot(Wt, Yp) /
                                                   Wt.sum()),
                     "target_negative_weight": float(Wt[Y < 0].sum() /
                                                     Wt.sum()),
                     "iters": iters, "loss": losses, "gpu_s": t_train,
                     "scale_learned": a_l, "scale_search_learned": h_l,
                     "c_smagorinsky_fitted": c_s,
                     "search_smagorinsky": h_s},
        "runs": out,
        "time_ratio_fine_over_learned": t_f / out["learned"]["gpu_s"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--dims", type=int, default=3)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--time", type=float, default=0.2)
    ap.add_argument("--samples", type=int, default=10)
    ap.add_argument("--snapshots", type=int, default=4)
    ap.add_argument("--iters", type=int, default=400)
    ap.add_argument("--train-seed", type=int, default=0)
    ap.add_argument("--eval-seed", type=int, default=1)
    ap.add_argument("--evals", type=int, default=6)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    import torch
    torch.zeros(1, device="cuda")  # torch's HIP context before the library's
    res = study(a.dims, a.n, a.time, a.samples, a.snapshots, iters=a.iters,
                train_seed=a.train_seed, eval_seed=a.eval_seed,
                evals=a.evals,
                log=lambda s: print(s, file=sys.stderr, flush=True))
    line = json.dumps(res)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")
    print(line)


if __name__ == "__main__":
    main()
