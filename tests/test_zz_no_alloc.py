"""aq_linklist_noalloc (the link-list with its workspace in the library's
fixed device scratch, no allocation) against aq_linklist with a caller
workspace: grid, r_min / r_max, icell, the sort permutations and ihoc bit
for bit, on two calls in a row (the second smaller, so it runs over the
scratch the first one left), and the size limit is refused on the host."""
import ctypes

import numpy as np
import pytest
import torch

from aquagpusph_amd import _lib

pytestmark = pytest.mark.gpu

P = ctypes.c_void_p


def dp(t):
    return P(t.data_ptr())


def _inputs(dims, N, seed):
    rng = np.random.default_rng(seed)
    vs = 2 if dims == 2 else 4
    r = np.zeros((N, vs), np.float32)
    r[:, :dims] = rng.uniform(-1.0, 1.5, (N, dims)).astype(np.float32)
    # a dense clump so that cells hold many particles (long equal-key runs)
    k = N // 4
    r[:k, :dims] = rng.normal(0.2, 0.05, (k, dims)).astype(np.float32)
    return r


def _run(dims, r, noalloc, cap, h=0.02):
    dev = torch.device("cuda")
    N, vs = r.shape
    rr = torch.from_numpy(r).to(dev)
    out = dict(r_min=torch.zeros(vs, device=dev),
               r_max=torch.zeros(vs, device=dev),
               n_cells=torch.zeros(4, dtype=torch.int32, device=dev),
               icell=torch.zeros(N, dtype=torch.int32, device=dev),
               id_unsorted=torch.zeros(N, dtype=torch.int32, device=dev),
               id_sorted=torch.zeros(N, dtype=torch.int32, device=dev),
               ihoc=torch.zeros(cap, dtype=torch.int32, device=dev),
               status=torch.zeros(1, dtype=torch.int32, device=dev))
    support = torch.tensor([2.0], device=dev)
    hh = torch.tensor([h], device=dev)
    stream = P(torch.cuda.current_stream().cuda_stream)
    common = (dims, dp(rr), N, dp(out["r_min"]), dp(out["r_max"]),
              dp(support), dp(hh), dp(out["n_cells"]), dp(out["icell"]),
              dp(out["id_unsorted"]), dp(out["id_sorted"]), dp(out["ihoc"]),
              cap, dp(out["status"]))
    if noalloc:
        _lib.call("aq_linklist_noalloc", *common, stream)
    else:
        L = _lib.lib()
        L.aq_linklist_workspace.restype = ctypes.c_size_t
        ws = torch.empty(int(L.aq_linklist_workspace(N)), dtype=torch.uint8,
                         device=dev)
        _lib.call("aq_linklist", *common, dp(ws), stream)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


@pytest.mark.parametrize("dims", [2, 3])
def test_noalloc_linklist_equals_workspace_linklist(cuda, dims):
    cap = 1 << 22
    for N, seed in ((1000000, 1), (123457, 2)):
        print("aq_linklist_noalloc dims=%d N=%d seed=%d" % (dims, N, seed),
              flush=True)
        r = _inputs(dims, N, seed)
        a = _run(dims, r, False, cap)
        b = _run(dims, r, True, cap)
        assert a["status"][0] == 0 and b["status"][0] == 0
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)
        nc = int(a["n_cells"][3])
        assert 0 < nc <= cap
        # the permutation sorts icell, and id_sorted inverts it
        ic = a["icell"].astype(np.int64)
        assert np.all(np.diff(ic) >= 0)
        inv = a["id_sorted"].astype(np.int64)
        perm = a["id_unsorted"].astype(np.int64)
        np.testing.assert_array_equal(inv[perm], np.arange(N))


def test_noalloc_linklist_refuses_sizes_beyond_its_scratch(cuda):
    N = (1 << 20) + 1
    dev = torch.device("cuda")
    z = torch.zeros(4, device=dev)
    zi = torch.zeros(4, dtype=torch.int32, device=dev)
    with pytest.raises(_lib.AquaError, match="no-allocation limit"):
        _lib.call("aq_linklist_noalloc", 3, dp(z), N, dp(z), dp(z), dp(z),
                  dp(z), dp(zi), dp(zi), dp(zi), dp(zi), dp(zi), 4, dp(zi),
                  P(torch.cuda.current_stream().cuda_stream))
