"""The bench's accounting: algorithmic bytes of a stage launch and the matrix-core work of the
tile kernel, checked against direct counts on small inputs."""
import numpy as np

import bench


def test_stage_bytes_counts():
    # one node, one tracer, one level, n = 1: 4.25 state values, n (3 weights + 4-byte index),
    # 16 basis values
    assert bench.stage_bytes(1, 1, 1, 1, 4, 4) == 4.25 * 4 + (3 * 4 + 4) + 16 * 4
    # linear in N; the state term scales with ntr * Nv and the value size
    b1 = bench.stage_bytes(10, 55, 60, 4, 4, 4)
    assert bench.stage_bytes(20, 55, 60, 4, 4, 4) == 2 * b1
    assert bench.stage_bytes(10, 55, 60, 4, 8, 4) - b1 == 10 * 4 * 60 * 4.25 * 4


def test_tile_flops_against_brute_force():
    rng = np.random.default_rng(0)
    N, n = 37, 5                       # three tiles, the last one partial (5 nodes)
    idx = rng.integers(0, 60, size=(N, n)).astype(np.int32)
    dense, useful, umean = bench.tile_flops(idx, S=64, ntr=4)
    unions = [len(np.unique(idx[t * 16:(t + 1) * 16])) for t in range(3)]
    u16 = [(u + 15) // 16 * 16 for u in unions]
    assert abs(umean - np.mean(unions)) < 1e-12
    assert dense == sum(u16) * 16 * 64 * 4 * 2 * 2
    assert useful == N * n * 64 * 4 * 2 * 2
    d1, u1, _ = bench.tile_flops(idx, S=64, ntr=4, hv=False)
    assert d1 == dense / 2 and u1 == useful / 2


def test_more_than_one_gpu_is_refused():
    """The bench measures one GPU; asked for more it exits with status 2 and says why, before
    it loads the library or touches a device."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=root)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""
