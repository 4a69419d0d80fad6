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


def test_bench_householder_and_lapack_residuals(golden):
    """The bench's numpy frame maps each node to the pole like the oracle's; its LAPACK check
    of the config-mode gradient systems finds the oracle's weights (its own LU) within
    rounding, and finds a perturbed weight set."""
    import oracle_lib as O
    from se_rbf_projects_amd.ctypes_api import config_deform, variable_stencil_sizes

    xyz = golden.icos(642)[:, :3]
    for p in (xyz[5], xyz[100], np.array([0.0, 0.0, -1.0]), np.array([0.0, 0.6, 0.8])):
        H = bench.householder(p)
        assert np.allclose(H @ p, [0, 0, 1], atol=1e-15) and np.allclose(H @ H.T, np.eye(3), atol=1e-15)
        assert np.allclose(H, O.householder(p, fix_pole=True), atol=1e-15)
    cfg = config_deform(642, Nv=10, n=40, deg=4, nv=5)
    idx, D = O.knn(xyz, 40)
    nn, deg = variable_stencil_sizes(D, n_min=31, n_max=40)
    Dx, Dy, L, st = O.weights_var(cfg, xyz, idx, nn, deg)
    assert not st.any()
    r = bench.lapack_residuals(xyz, idx, nn, deg, Dx, Dy, samples=40)
    assert r["sampled_nodes"] == 40 and r["max_weight_rel_diff"] < 1e-8 and r["max_scaled_residual"] < 1e-13
    Dx2 = Dx.copy()
    Dx2[::7, 3] *= 1.001
    r2 = bench.lapack_residuals(xyz, idx, nn, deg, Dx2, Dy, samples=40)
    assert r2["max_weight_rel_diff"] > 1e-5
