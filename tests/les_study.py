"""LES study: the learned eddy-viscosity closure against plain SPH on
decaying 2D turbulence (cases.decaying_box: unit box with BI walls, seeded
solenoidal field with E(k) ~ k^4 exp(-2 (k/k0)^2), k0 = 4, u_rms = 1).

  1. fine run (2n x 2n particles, plain SPH): the reference, and the
     training data: at `snapshots` times its velocity is filtered onto the
     coarse lattice points (Shepard-normalised Wendland filter of the coarse
     smoothing length); the filtered field's SPH gradient on the coarse
     lattice gives the network inputs (the same operator the closure uses),
     the filtered subgrid stress tau = <uu> - <u><u> gives the target
     coefficient c* = -(tau_d : S) / (2 S:S) / (H^2 |S|) (interior points
     only), clipped to [0, 2] for the eddy-viscosity network and to
     [c_min, 2] (c_min < 0) for the one allowed to return energy
     (backscatter, les_cmin = c_min);
  2. training on the GPU: the BF16 network kernels (aq_mlp_grad, aq_mlp_adam)
     on minibatches of 10k samples;
  3. coarse runs (n x n): plain SPH, the learned closures (cfd/LES.xml)
     and Smagorinsky (cfd/LES-smagorinsky.xml, C_s = 0.15);
  4. the kinetic energy of every run at the same times, its error against
     the fine run (mean over samples of |E - E_fine| / E_fine(0)), and the
     GPU time of every run.

    python tests/les_study.py [--n 128] [--time 1.0] [--out F]

prints one JSON line.  tests/test_les_gpu.py runs a small version.
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
sys.path.insert(0, os.path.dirname(HERE))

from aquagpusph_amd import _lib, casefiles, cases  # noqa: E402
from oracle import les as olq  # noqa: E402
from oracle import mlp as omlp  # noqa: E402
from oracle import pipeline as op  # noqa: E402
from oracle.pipeline import _p  # noqa: E402

BASE = ("resources/Presets/basic.xml",
        "resources/Presets/basic/time_scheme/improved_euler.xml",
        "resources/Presets/basic/domain.xml",
        "resources/Presets/cfd.xml",
        "resources/Presets/cfd/deltaSPH.xml",
        "resources/Presets/cfd/BI.xml")
MORRIS = ('    <Definitions><Define name="__LAP_FORMULATION__" '
          'value="__LAP_MORRIS__" evaluate="false"/></Definitions>')


def _sim(case, extra=(), params=None, cmin=None):
    from aquagpusph_amd.sim import Simulation
    d = tempfile.mkdtemp(prefix="aq_les_")
    xml = casefiles.write_case(case, d, presets=BASE + tuple(extra),
                               load=False, extra_xml=MORRIS)
    sim = Simulation(xml, 2)
    casefiles.upload(sim, case)
    if params is not None:
        sim.set("les_params", params)
    if cmin is not None:
        sim.set("les_cmin", np.array([cmin], np.float32))
    return sim


def _ke(sim):
    u = sim.get("u").astype(np.float64)
    m = sim.get("m").astype(np.float64)
    fl = sim.get("imove") == 1
    return float(0.5 * np.sum(m[fl] * np.sum(u[fl] ** 2, axis=1)))


def run(case, samples, steps_per_sample, extra=(), params=None, snap=None,
        cmin=None):
    """KE at every sample, GPU seconds of the stepping, and (snap: the
    sample indices) particle snapshots."""
    sim = _sim(case, extra, params, cmin)
    ke = [_ke(sim)]
    gpu_s = 0.0
    snaps = []
    for k in range(samples):
        t0 = time.perf_counter()
        sim.step(steps_per_sample, graph=True)
        sim.sync()
        gpu_s += time.perf_counter() - t0
        ke.append(_ke(sim))
        if snap is not None and (k + 1) in snap:
            snaps.append({f: sim.get(f).copy() for f in
                          ("r", "u", "m", "rho", "imove")})
    t = float(sim.get("t"))
    sim.close()
    return np.array(ke), gpu_s, snaps, t


def _w(q):
    return np.where(q < 2.0, (1 + 2 * q) * (2 - q) ** 4, 0.0)


def training_samples(fine_snaps, coarse, hc):
    """(inputs [M, 8], target c* [M]) from the fine snapshots."""
    from scipy.spatial import cKDTree
    nf = coarse["n_fluid"]
    xc = coarse["r"][:nf, :2]
    X, Y = [], []
    for s in fine_snaps:
        fl = s["imove"] == 1
        rf = s["r"][fl, :2].astype(np.float64)
        uf = s["u"][fl, :2].astype(np.float64)
        vf = (s["m"][fl] / s["rho"][fl]).astype(np.float64)
        D = cKDTree(rf).sparse_distance_matrix(cKDTree(xc), 2.0 * hc,
                                               output_type="coo_matrix")
        w = vf[D.row] * _w(D.data / hc)
        nrm = np.bincount(D.col, w, minlength=nf)
        ub = np.stack([np.bincount(D.col, w * uf[D.row, a], minlength=nf)
                       for a in range(2)], axis=1) / nrm[:, None]
        uu = np.stack([[np.bincount(D.col, w * uf[D.row, a] * uf[D.row, b],
                                    minlength=nf) for b in range(2)]
                       for a in range(2)], axis=0) / nrm[None, None, :]
        tau = np.transpose(uu, (2, 0, 1)) - ub[:, :, None] * ub[:, None, :]
        # the filtered field's SPH gradient on the coarse lattice (the
        # closure's own operator)
        cc = dict(coarse)
        cc["u"] = coarse["u"].copy()
        cc["u"][:nf, :2] = ub
        o = op.OracleSim(cc, dtype=np.float64, domain=False,
                         delta=False, les={"Cs": 0.15})
        f = o.f
        for k in ("r", "u", "rho"):
            f[k + "_in"][...] = f[k]
        o.link_list("r_in")
        o.reorder()
        o._k("les_gradu")(_p(f["imove"]), _p(f["r"]), _p(f["u"]),
                          _p(f["rho"]), _p(f["m"]), _p(f["grad_u"]),
                          ctypes.c_uint(o.N), *o._ll(), o._r(o.H),
                          o._r(o.CONF))
        ids = f["id"].astype(np.int64)
        G = np.zeros((o.N, 4))
        G[ids] = f["grad_u"]
        G = G[:nf]
        Gm = olq.matrices(G, 2)
        S = 0.5 * (Gm + np.transpose(Gm, (0, 2, 1)))
        ss = np.einsum("nab,nab->n", S, S)
        tr = tau[:, 0, 0] + tau[:, 1, 1]
        td = tau.copy()
        td[:, 0, 0] -= 0.5 * tr
        td[:, 1, 1] -= 0.5 * tr
        nu_star = -np.einsum("nab,nab->n", td, S) / (2.0 * ss + 1e-30)
        smag = np.sqrt(2.0 * ss)
        c = nu_star / (hc * hc * smag + 1e-30)
        inner = np.all((xc > 2.5 * hc) & (xc < 1.0 - 2.5 * hc), axis=1) & \
            (smag > 1e-3 * smag.max())
        X.append(olq.features(G, 2, 2.0)[inner])
        Y.append(np.clip(c[inner], -2.0, 2.0))
    return np.concatenate(X).astype(np.float32), \
        np.concatenate(Y).astype(np.float32)


def train(X, Y, iters=400, batch=10000, lr=2e-3, seed=0, weights=None):
    """The network trained on the GPU by the library's kernels.  weights:
    minibatches drawn with probability proportional to them (with
    replacement), so that the mean squared error is the weighted one."""
    import torch
    L = _lib.lib()
    L.aq_mlp_workspace.restype = ctypes.c_size_t
    L.aq_mlp_workspace.argtypes = [ctypes.c_uint32]
    L.aq_mlp_grad.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_uint32] + \
        [ctypes.c_void_p] * 4
    L.aq_mlp_adam.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int,
                                                      ctypes.c_float,
                                                      ctypes.c_void_p]
    rng = np.random.default_rng(seed)
    P = omlp.init(seed)
    P[omlp.OFF_W3:omlp.OFF_B3] *= 0.1
    pw = None
    if weights is not None:
        pw = np.asarray(weights, np.float64)
        pw = pw / pw.sum()
    P[omlp.OFF_B3] = float(np.mean(Y) if pw is None else np.dot(pw, Y))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
    Pd, md, vd = dev(P), dev(np.zeros_like(P)), dev(np.zeros_like(P))
    gd = torch.zeros(omlp.NPARAM + 1, device="cuda")
    b = min(batch, len(Y))
    ws = torch.zeros(L.aq_mlp_workspace(b) // 4 + 1, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    losses = []
    t0 = time.perf_counter()
    for k in range(1, iters + 1):
        idx = (rng.choice(len(Y), b, replace=False) if pw is None
               else rng.choice(len(Y), b, replace=True, p=pw))
        Xb, Yb = dev(X[idx]), dev(Y[idx])
        _lib.check(L.aq_mlp_grad(_p_t(Xb), _p_t(Yb), b, _p_t(Pd), _p_t(gd),
                                 _p_t(ws), s))
        _lib.check(L.aq_mlp_adam(_p_t(Pd), _p_t(gd), _p_t(md), _p_t(vd), k,
                                 lr, s))
        if k == 1 or k % 50 == 0:
            losses.append((k, float(gd[-1].item())))
    torch.cuda.synchronize()
    return Pd.cpu().numpy(), losses, time.perf_counter() - t0


def _p_t(t):
    return ctypes.c_void_p(t.data_ptr())


def study(n=128, t_end=1.0, samples=20, snapshots=8, iters=400, log=None,
          cmin=-0.1):
    coarse = cases.decaying_box(n)
    fine = cases.decaying_box(2 * n)
    hc = coarse["params"]["h"]
    dtc = coarse["params"]["courant"] * hc / coarse["params"]["cs"]
    spc = max(1, int(round(t_end / dtc / samples)))
    snap = set(np.linspace(1, samples, snapshots).round().astype(int))
    say = log or (lambda s: None)
    say("fine run: %d particles, %d steps" % (fine["N"], 2 * spc * samples))
    ke_f, t_f, snaps, tf = run(fine, samples, 2 * spc, snap=snap)
    say("training samples")
    X, Y = training_samples(snaps, coarse, hc)
    say("training on %d samples" % len(Y))
    P, losses, t_train = train(X, np.clip(Y, 0.0, 2.0), iters=iters)
    Pb, losses_b, _ = train(X, np.clip(Y, cmin, 2.0), iters=iters)
    les = ("resources/Presets/cfd/LES.xml",)
    out = {}
    for name, extra, params, cm in (
            ("plain", (), None, None),
            ("learned", les, P, None),
            ("learned_backscatter", les, Pb, cmin),
            ("smagorinsky", ("resources/Presets/cfd/LES-smagorinsky.xml",),
             None, None)):
        say("coarse run: %s" % name)
        ke, t, _, tc = run(coarse, samples, spc, extra, params, cmin=cm)
        out[name] = dict(ke=ke.tolist(), gpu_s=t, t_end=tc,
                         ke_err=float(np.mean(np.abs(ke - ke_f)) / ke_f[0]))
    out["fine"] = dict(ke=ke_f.tolist(), gpu_s=t_f, t_end=tf)
    return {
        "study": "les",
        "workload": "decaying 2D turbulence in a closed unit box, BI walls, "
                    "E(k) ~ k^4 exp(-2 (k/k0)^2), k0 = 4, u_rms = 1, "
                    "nu = 1e-4, coarse %dx%d, fine %dx%d, t = %.3g s"
                    % (n, n, 2 * n, 2 * n, t_end),
        "samples": samples, "steps_per_sample_coarse": spc,
        "training": {"samples": int(len(Y)),
                     "target_mean": float(np.clip(Y, 0.0, 2.0).mean()),
                     "target_mean_signed": float(np.clip(Y, cmin, 2.0).mean()),
                     "target_negative_fraction": float((Y < 0).mean()),
                     "cmin": cmin, "iters": iters, "loss": losses,
                     "loss_backscatter": losses_b, "gpu_s": t_train},
        "runs": out,
        "time_ratio_fine_over_learned": t_f / out["learned"]["gpu_s"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--time", type=float, default=1.0)
    ap.add_argument("--samples", type=int, default=20)
    ap.add_argument("--iters", type=int, default=400)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    import torch
    torch.zeros(1, device="cuda")  # torch's HIP context before the library's
    res = study(a.n, a.time, a.samples, iters=a.iters,
                log=lambda s: print(s, file=sys.stderr, flush=True))
    line = json.dumps(res)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")
    print(line)


if __name__ == "__main__":
    main()
