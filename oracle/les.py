"""CPU restatement of the LES closure's per-particle parts (csrc/les.hip):
the network inputs from the velocity gradient and the eddy viscosity.

TEST INFRASTRUCTURE ONLY: tests/ may use this module, never the product.
The pair sums (velocity gradient, eddy-viscosity term) are the C
restatements orc_les_gradu / orc_les_interactions; the network is
oracle/mlp.py.  Pinned by tests/test_oracle_les_cpu.py.
"""
import numpy as np


def matrices(grad_u, dims):
    """(N, D, D) gradient tensors from the matrix-layout array."""
    n = 2 if dims == 2 else 4
    G = np.asarray(grad_u, np.float64).reshape(-1, n, n)[:, :dims, :dims]
    return G


def features(grad_u, dims, hdr):
    """The 8 network inputs per particle (the last is padding)."""
    G = matrices(grad_u, dims)
    S = 0.5 * (G + np.transpose(G, (0, 2, 1)))
    W = 0.5 * (G - np.transpose(G, (0, 2, 1)))
    s2 = np.einsum("nab,nab->n", S, S)
    w2 = np.einsum("nab,nab->n", W, W)
    nn = s2 + w2 + 1e-30
    S2 = S @ S
    W2 = W @ W
    tr = lambda A: np.trace(A, axis1=1, axis2=2)
    x = np.zeros((G.shape[0], 8))
    x[:, 0] = s2 / nn
    x[:, 1] = w2 / nn
    x[:, 2] = tr(S2 @ S) / nn ** 1.5
    x[:, 3] = tr(W2 @ S) / nn ** 1.5
    x[:, 4] = tr(W2 @ S2) / nn ** 2
    x[:, 5] = tr(W2 @ S @ W @ S2) / nn ** 3
    x[:, 6] = hdr
    return x


def strain_rate(grad_u, dims):
    """|S| = sqrt(2 S:S)."""
    G = matrices(grad_u, dims)
    S = 0.5 * (G + np.transpose(G, (0, 2, 1)))
    return np.sqrt(2.0 * np.einsum("nab,nab->n", S, S))


def eddy_viscosity(imove, grad_u, dims, H, c=None, Cs=None, cmin=0.0):
    """max(c, cmin) H^2 |S| (network) or (Cs H)^2 |S| (Smagorinsky); 0 off
    the fluid.  A float32 grad_u is evaluated in float32 with the kernel's
    operation order (S:S summed over a then b, ((k H) H) |S|)."""
    if np.asarray(grad_u).dtype == np.float32:
        f = np.float32
        n = 2 if dims == 2 else 4
        G = np.asarray(grad_u, f).reshape(-1, n, n)
        s2 = np.zeros(G.shape[0], f)
        for a in range(dims):
            for b in range(dims):
                sab = f(0.5) * (G[:, a, b] + G[:, b, a])
                s2 = s2 + sab * sab
        smag = np.sqrt(f(2) * s2)
        k = (np.maximum(np.asarray(c, f), f(cmin)) if Cs is None
             else np.full(G.shape[0], f(Cs) * f(Cs), f))
        nu = ((k * f(H)) * f(H)) * smag
        return np.where(np.asarray(imove) == 1, nu, f(0))
    k = np.maximum(np.asarray(c, np.float64), cmin) if Cs is None else Cs * Cs
    nu = k * H * H * strain_rate(grad_u, dims)
    return np.where(np.asarray(imove) == 1, nu, 0.0)
