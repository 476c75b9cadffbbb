"""Helpers to drive single oracle kernels on arbitrary particle sets."""
import ctypes

import numpy as np

from oracle import pipeline as op
from oracle.pipeline import _p


def link(dims, r, h, dtype=np.float64):
    """Oracle link-list of positions r (N x vs): returns dict with the
    sorted order applied (positions permuted like the reference sort)."""
    L = op.lib()
    R = np.dtype(dtype)
    sfx = ("f32" if R == np.float32 else "f64") + "_%dd" % dims
    vs = 2 if dims == 2 else 4
    r = np.ascontiguousarray(r, R).reshape(-1, vs)
    N = r.shape[0]
    rmin = np.zeros(vs, R)
    rmax = np.zeros(vs, R)
    getattr(L, "orc_minmax_vec_" + sfx)(_p(r), ctypes.c_uint(N), _p(rmin),
                                        _p(rmax))
    nc = np.zeros(4, np.uint32)
    creal = ctypes.c_float if R == np.float32 else ctypes.c_double
    cl = R.type(np.float32(2.0) * np.float32(h))
    getattr(L, "orc_n_cells_" + sfx)(_p(rmin), _p(rmax), creal(cl), _p(nc))
    ic = np.zeros(N, np.uint32)
    getattr(L, "orc_icell_" + sfx)(_p(ic), _p(r), ctypes.c_uint(N), _p(rmin),
                                   creal(2.0), creal(float(np.float32(h))),
                                   _p(nc))
    perm = np.zeros(N, np.uint32)
    inv = np.zeros(N, np.uint32)
    L.orc_stable_sort(_p(ic), _p(perm), _p(inv), ctypes.c_uint(N),
                      ctypes.c_uint(L.orc_radix_key_bits(
                          ctypes.c_ulonglong(int(nc[3])))))
    hoc = np.zeros(int(nc[3]), np.uint32)
    L.orc_ihoc(_p(ic), _p(hoc), ctypes.c_uint(N), _p(nc))
    return dict(r=r[perm].copy(), perm=perm, inv=inv, icell=ic, ihoc=hoc,
                n_cells=nc, rmin=rmin, rmax=rmax, N=N, sfx=sfx, R=R,
                creal=creal)


def call(ll, name, *args):
    return getattr(op.lib(), "orc_%s_%s" % (name, ll["sfx"]))(*args)


def ll_args(ll):
    return (_p(ll["icell"]), _p(ll["ihoc"]), _p(ll["n_cells"]))


def defs(h, dims):
    return (op.definition(h), op.definition(1.0 / h ** dims),
            op.definition(1.0 / h ** (dims + 2)))
