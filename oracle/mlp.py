"""CPU restatement of the LES eddy-viscosity network (csrc/mlp.hip).

TEST INFRASTRUCTURE ONLY: tests/ may use this module, never the product.

A 7 -> 64 -> 64 -> 1 perceptron with tanh-form GELU hidden units and the
mean squared error, forward and backward in float64.  With bf16=True every
operand of a matrix product is first rounded to BF16 (round to nearest
even) exactly where the GPU kernels feed a BF16 matrix instruction: the
inputs and W1 (layer 1), the hidden activations and W2 (layer 2), dZ2 and
H1 (dW2), dZ2 and W2 (dH1), dZ1 and the inputs (dW1); the output layer, the
biases and the loss stay unrounded.  The network is not in the reference
(a north-star addition); it is pinned by finite differences of the float64
mode (tests/test_oracle_mlp_cpu.py).

Parameter layout (one flat array of NPARAM): W1[64][16] (inputs padded
from 7 to 16), b1[64], W2[64][64] (out, in), b2[64], w3[64], b3.
"""
import numpy as np

NIN, KIN, NH = 7, 16, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * KIN
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH
NPARAM = OFF_B3 + 1


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    b = a.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16
    return (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(P):
    P = np.asarray(P, np.float64)
    return dict(W1=P[OFF_W1:OFF_B1].reshape(NH, KIN), b1=P[OFF_B1:OFF_W2],
                W2=P[OFF_W2:OFF_B2].reshape(NH, NH), b2=P[OFF_B2:OFF_W3],
                w3=P[OFF_W3:OFF_B3], b3=P[OFF_B3])


def init(seed=0):
    """He-style random parameters (float32), padded columns zero."""
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAM, np.float32)
    W1 = np.zeros((NH, KIN))
    W1[:, :NIN] = rng.normal(0, np.sqrt(2.0 / NIN), (NH, NIN))
    P[OFF_W1:OFF_B1] = W1.ravel()
    P[OFF_W2:OFF_B2] = rng.normal(0, np.sqrt(2.0 / NH), NH * NH)
    P[OFF_W3:OFF_B3] = rng.normal(0, np.sqrt(1.0 / NH), NH)
    return P


def gelu(x):
    c = np.sqrt(2.0 / np.pi)
    return 0.5 * x * (1.0 + np.tanh(c * (x + 0.044715 * x ** 3)))


def gelu_d(x):
    c = np.sqrt(2.0 / np.pi)
    t = np.tanh(c * (x + 0.044715 * x ** 3))
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * c * (1 + 3 * 0.044715 * x * x)


def _inputs(X):
    X = np.asarray(X, np.float64)
    Xp = np.zeros((X.shape[0], KIN))
    Xp[:, :NIN] = X[:, :NIN]
    return Xp


def forward(P, X, bf=True):
    """y[M] and the intermediates."""
    q = unpack(P)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    Xp = _inputs(X)
    Z1 = r(Xp) @ r(q["W1"]).T + q["b1"]
    H1 = gelu(Z1)
    Z2 = r(H1) @ r(q["W2"]).T + q["b2"]
    H2 = gelu(Z2)
    y = H2 @ q["w3"] + q["b3"]
    return y, dict(Xp=Xp, Z1=Z1, H1=H1, Z2=Z2, H2=H2)


def grad(P, X, target, bf=True):
    """(gradient of the mean squared error [NPARAM], loss)."""
    q = unpack(P)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    y, c = forward(P, X, bf)
    M = len(y)
    e = y - np.asarray(target, np.float64)
    loss = float(np.sum(e * e) / M)
    dy = 2.0 * e / M
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = dy @ c["H2"]
    g[OFF_B3] = dy.sum()
    dZ2 = np.outer(dy, q["w3"]) * gelu_d(c["Z2"])
    g[OFF_B2:OFF_W3] = dZ2.sum(axis=0)
    g[OFF_W2:OFF_B2] = (r(dZ2).T @ r(c["H1"])).ravel()
    dH1 = r(dZ2) @ r(q["W2"])
    dZ1 = dH1 * gelu_d(c["Z1"])
    g[OFF_B1:OFF_W2] = dZ1.sum(axis=0)
    g[OFF_W1:OFF_B1] = (r(dZ1).T @ r(c["Xp"])).ravel()
    return g, loss


def adam(P, G, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step in float32 (returns new P, m, v)."""
    P, G, m, v = (np.asarray(a, np.float32) for a in (P, G, m, v))
    m = np.float32(beta1) * m + np.float32(1 - beta1) * G
    v = np.float32(beta2) * v + np.float32(1 - beta2) * G * G
    c1 = np.float32(1 - beta1 ** t)
    c2 = np.float32(1 - beta2 ** t)
    P = P - np.float32(lr) * (m / c1) / (np.sqrt(v / c2) + np.float32(eps))
    return P, m, v
