"""The BF16-MFMA network kernels (aq_mlp_forward / aq_mlp_grad /
aq_mlp_adam) against the oracle's BF16 mode (oracle/mlp.py, itself pinned
by finite differences): outputs, every block of the gradient and the loss,
determinism, and Adam steps.  Differences come from FP32 accumulation
order and FP32 GELU against float64 (a BF16 rounding of a hidden value can
flip), hence the 2e-3-of-largest tolerance."""
import ctypes

import numpy as np
import pytest
import torch

from aquagpusph_amd import _lib
from oracle import mlp

pytestmark = pytest.mark.gpu


def _L():
    L = _lib.lib()
    L.aq_mlp_nparam.restype = ctypes.c_int
    L.aq_mlp_workspace.restype = ctypes.c_size_t
    L.aq_mlp_workspace.argtypes = [ctypes.c_uint32]
    L.aq_mlp_forward.argtypes = [ctypes.c_void_p, ctypes.c_uint32,
                                 ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p]
    L.aq_mlp_grad.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_uint32] + \
        [ctypes.c_void_p] * 4
    L.aq_mlp_adam.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int,
                                                      ctypes.c_float,
                                                      ctypes.c_void_p]
    return L


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _s():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(a, dt=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(
        "cuda", dt)


def _data(m, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (m, 8)).astype(np.float32)
    t = (np.tanh(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2]).astype(np.float32)
    return X, t


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("m", [1, 16, 37, 1000, 10000])
def test_forward_matches_oracle(cuda, m):
    L = _L()
    assert L.aq_mlp_nparam() == mlp.NPARAM
    P = mlp.init(1)
    X, _ = _data(m, 2)
    y = torch.zeros(m, device="cuda")
    Xd, Pd = _dev(X), _dev(P)  # held: _p keeps only the address
    _lib.check(L.aq_mlp_forward(_p(Xd), m, _p(Pd), _p(y), _s()))
    torch.cuda.synchronize()
    want, _ = mlp.forward(P, X, bf=True)
    assert rel(y.cpu().numpy(), want) <= 2e-3


@pytest.mark.parametrize("m", [16, 37, 1000, 10000])
def test_gradient_matches_oracle(cuda, m):
    L = _L()
    P = mlp.init(2)
    X, t = _data(m, 3)
    ws = torch.zeros(L.aq_mlp_workspace(m) // 4 + 1, device="cuda")
    g1 = torch.zeros(mlp.NPARAM + 1, device="cuda")
    g2 = torch.zeros(mlp.NPARAM + 1, device="cuda")
    Xd, td, Pd = _dev(X), _dev(t), _dev(P)  # held: _p keeps the address
    args = (_p(Xd), _p(td), m, _p(Pd))
    _lib.check(L.aq_mlp_grad(*args, _p(g1), _p(ws), _s()))
    _lib.check(L.aq_mlp_grad(*args, _p(g2), _p(ws), _s()))
    torch.cuda.synchronize()
    a, b = g1.cpu().numpy(), g2.cpu().numpy()
    np.testing.assert_array_equal(a, b)  # deterministic
    want, loss = mlp.grad(P, X, t, bf=True)
    assert a[-1] == pytest.approx(loss, rel=1e-4)
    blocks = {"W1": (mlp.OFF_W1, mlp.OFF_B1), "b1": (mlp.OFF_B1, mlp.OFF_W2),
              "W2": (mlp.OFF_W2, mlp.OFF_B2), "b2": (mlp.OFF_B2, mlp.OFF_W3),
              "w3": (mlp.OFF_W3, mlp.OFF_B3), "b3": (mlp.OFF_B3, mlp.NPARAM)}
    errs = {k: rel(a[i:j], want[i:j]) for k, (i, j) in blocks.items()}
    print(m, errs)
    assert max(errs.values()) <= 2e-3, errs
    W1g = a[mlp.OFF_W1:mlp.OFF_B1].reshape(mlp.NH, mlp.KIN)
    assert np.all(W1g[:, mlp.NIN:] == 0)


def test_adam_training_matches_oracle(cuda):
    L = _L()
    m = 4096
    P = mlp.init(7)
    X, t = _data(m, 4)
    Pd = _dev(P)
    md = torch.zeros(mlp.NPARAM, device="cuda")
    vd = torch.zeros(mlp.NPARAM, device="cuda")
    gd = torch.zeros(mlp.NPARAM + 1, device="cuda")
    ws = torch.zeros(L.aq_mlp_workspace(m) // 4 + 1, device="cuda")
    Xd, td = _dev(X), _dev(t)
    Po, mo, vo = P.copy(), np.zeros_like(P), np.zeros_like(P)
    losses = []
    for k in range(1, 31):
        _lib.check(L.aq_mlp_grad(_p(Xd), _p(td), m, _p(Pd), _p(gd), _p(ws),
                                 _s()))
        _lib.check(L.aq_mlp_adam(_p(Pd), _p(gd), _p(md), _p(vd), k, 3e-3,
                                 _s()))
        losses.append(float(gd[-1].item()))
        if k <= 3:
            go, _ = mlp.grad(Po, X, t, bf=True)
            Po, mo, vo = mlp.adam(Po, go.astype(np.float32), mo, vo, k, 3e-3)
            assert rel(Pd.cpu().numpy(), Po) <= 1e-3, k
    assert losses[-1] < 0.5 * losses[0], losses
