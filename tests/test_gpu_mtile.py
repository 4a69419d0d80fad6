"""The matrix-core stage kernels (16-node tiles, dense W x Q products on
v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64; version 4 stages the state in LDS, version 5
reads it directly as the matrix operand, version 6 is version 5 as a persistent pipeline of
matrix and update wavefronts) against the column-vector kernel and the FP64 path, with and without
hyperviscosity and with the learned closure's per-point coefficients.  10242 nodes leave a
partial last tile."""
import numpy as np
import pytest

from se_rbf_projects_amd import closure as CL
from se_rbf_projects_amd import engine as E
from se_rbf_projects_amd.ctypes_api import F32, F64, config_deform, courant_dt

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def deform60():
    xyz = E.icos_nodes(5)
    xyz = xyz[E.spatial_order(xyz)]
    N = len(xyz)
    cfg = config_deform(N, Nv=60, n=55, deg=4, nv=7)
    cfg.ntracers = 4
    eng = E.Engine(0)
    eng.set_nodes(xyz)
    eng.knn(cfg.n)
    eng.configure(cfg)
    eng.weights()
    q0 = np.stack([E.initial_condition(k, cfg, xyz, seed=3) for k in CL.TRACERS], axis=1)
    yield dict(eng=eng, cfg=cfg, q0=q0, dt=courant_dt(cfg, N))
    eng.set_kernel_version(2)
    eng.close()


def _run(eng, q0, prec, kver, steps, dt):
    eng.set_kernel_version(kver)
    eng.set_state(q0)
    eng.run(prec, steps, dt)
    return eng.state(prec)


@pytest.mark.parametrize("kver", [4, 5, 6])
def test_matrix_core_stage_matches_column_kernel(deform60, kver):
    eng, q0, dt = deform60["eng"], deform60["q0"], deform60["dt"]
    q64 = _run(eng, q0, F64, 2, 20, dt)
    q2 = _run(eng, q0, F32, 2, 20, dt)
    q4 = _run(eng, q0, F32, kver, 20, dt)
    d24 = np.linalg.norm(q4 - q2) / np.linalg.norm(q2)
    d64 = np.linalg.norm(q4 - q64) / np.linalg.norm(q64)
    print(f"kver {kver} vs kver 2 (FP32): {d24:.2e}; vs FP64: {d64:.2e}; kver 2 vs FP64: "
          f"{np.linalg.norm(q2 - q64) / np.linalg.norm(q64):.2e}")
    assert d24 < 1e-5
    assert d64 < 1e-4
    # the constant tracer stays constant up to FP32 rounding
    assert np.max(np.abs(q4[:, 3] - 1.0)) < 1e-5


@pytest.mark.parametrize("kver", [5, 6])
def test_matrix_core_stage_fp64(deform60, kver):
    """Versions 5 and 6 in FP64 (v_mfma_f64_16x16x4f64) against the FP64 column kernel: the
    same sums in another order, so they agree to a few units of FP64 rounding per step."""
    eng, q0, dt = deform60["eng"], deform60["q0"], deform60["dt"]
    q2 = _run(eng, q0, F64, 2, 20, dt)
    q5 = _run(eng, q0, F64, kver, 20, dt)
    d = np.linalg.norm(q5 - q2) / np.linalg.norm(q2)
    print(f"kver {kver} vs kver 2 (FP64): {d:.2e}")
    assert d < 1e-12
    assert np.max(np.abs(q5[:, 3] - 1.0)) < 1e-12


@pytest.mark.parametrize("kver", [4, 5, 6])
def test_matrix_core_stage_without_hyperviscosity(deform60, kver):
    eng, cfg, q0, dt = deform60["eng"], deform60["cfg"], deform60["q0"], deform60["dt"]
    coef = cfg.hv_coef
    cfg.hv_coef = 0.0
    eng.configure(cfg)
    eng.weights()
    try:
        q64 = _run(eng, q0, F64, 2, 10, dt)
        q2 = _run(eng, q0, F32, 2, 10, dt)
        q4 = _run(eng, q0, F32, kver, 10, dt)
        q64k = _run(eng, q0, F64, kver, 10, dt)
    finally:
        cfg.hv_coef = coef
        eng.configure(cfg)
        eng.weights()
    assert np.linalg.norm(q4 - q2) / np.linalg.norm(q2) < 1e-5
    assert np.linalg.norm(q4 - q64) / np.linalg.norm(q64) < 1e-4
    assert np.linalg.norm(q64k - q64) / np.linalg.norm(q64) < 1e-12


@pytest.mark.parametrize("kver", [4, 5, 6])
def test_matrix_core_stage_with_closure(deform60, kver):
    eng, q0, dt = deform60["eng"], deform60["q0"], deform60["dt"]
    rng = np.random.default_rng(4)
    params = CL.init_params(4, out_scale=0.5)
    norm = np.concatenate([np.zeros(16), np.ones(16)]).astype(np.float32)
    cl = E.Closure(0)
    cl.set_params(params, norm)
    eng.set_state(q0)
    eng.attach_closure(cl, every=4, s_max=16.0)
    q2 = _run(eng, q0, F32, 2, 12, dt)
    q4 = _run(eng, q0, F32, kver, 12, dt)
    eng.attach_closure(None)
    cl.close()
    assert np.linalg.norm(q4 - q2) / np.linalg.norm(q2) < 1e-5
