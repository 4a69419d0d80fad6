"""serbf_knn_noalloc, the kNN search that allocates nothing (its working arrays are a fixed
device array per GPU), against serbf_knn + serbf_get_stencils on the same nodes: identical
indices and distances on two calls in a row, the second on a smaller node set (so stale
counts from the first call would show), and the host-side refusals (too many nodes, host
pointers) that run no GPU work."""
import ctypes as C

import numpy as np
import pytest
import torch

from se_rbf_projects_amd import engine as E

pytestmark = pytest.mark.gpu


def _engine_knn(xyz, n):
    eng = E.Engine(0)
    try:
        eng.set_nodes(xyz)
        eng.knn(n)
        return eng.stencils()
    finally:
        eng.close()


def _noalloc_knn(xyz, n):
    N = len(xyz)
    dev = torch.device("cuda:0")
    x, y, z = (torch.tensor(np.ascontiguousarray(xyz[:, k]), dtype=torch.float64, device=dev) for k in range(3))
    idx = torch.full((N, n), -1, dtype=torch.int32, device=dev)
    D = torch.full((N, n), -1.0, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    print(f"serbf_knn_noalloc N={N} n={n}", flush=True)
    rc = E.lib().serbf_knn_noalloc(0, N, x.data_ptr(), y.data_ptr(), z.data_ptr(), n, idx.data_ptr(), D.data_ptr())
    assert rc == 0, E.lib().serbf_last_error().decode()
    return idx.cpu().numpy(), D.cpu().numpy()


@pytest.mark.parametrize("level_n", [((5, 55), (3, 31)), ((4, 10), (2, 7)), ((7, 4), (5, 55))])
def test_noalloc_knn_equals_engine_knn_twice(level_n):
    for level, n in level_n:                   # the second call is on a smaller node set
        xyz = E.icos_nodes(level)
        xyz = xyz[E.spatial_order(xyz)]
        ref_idx, ref_D = _engine_knn(xyz, n)
        idx, D = _noalloc_knn(xyz, n)
        assert np.array_equal(idx, ref_idx)
        assert np.array_equal(D, ref_D)
        assert np.all(idx[:, 0] == np.arange(len(xyz)))


def test_noalloc_knn_jittered_nodes():
    xyz = E.jitter_nodes(E.icos_nodes(6), 0.3, seed=7)
    ref_idx, ref_D = _engine_knn(xyz, 40)
    idx, D = _noalloc_knn(xyz, 40)
    assert np.array_equal(idx, ref_idx) and np.array_equal(D, ref_D)


def test_noalloc_knn_refusals():
    L = E.lib()
    dev = torch.device("cuda:0")
    big = (1 << 20) + 1
    t = torch.zeros(16, dtype=torch.float64, device=dev)
    ti = torch.zeros(16, dtype=torch.int32, device=dev)
    p = t.data_ptr()
    assert L.serbf_knn_noalloc(0, big, p, p, p, 4, ti.data_ptr(), p) == 1
    assert b"1048576" in L.serbf_last_error()
    h = np.zeros(16)
    hp = h.ctypes.data
    assert L.serbf_knn_noalloc(0, 16, hp, hp, hp, 4, ti.data_ptr(), p) == 1
    assert b"device memory" in L.serbf_last_error()
    assert L.serbf_knn_noalloc(0, 16, p, p, p, 17, ti.data_ptr(), p) == 1
