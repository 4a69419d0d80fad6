"""The matrix-core stage kernel (version 5: 16-node tiles, dense W x Q products on
v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64 with the state read directly as the B
operand) at the mixed-precision headline's exact shape -- n = 55, degree 4, vertical
n = 7, 4 tracers x 60 levels (padded to 64), Delta^4 hyperviscosity in every stage,
deformational flow -- on 10242 Hilbert-ordered nodes (a partial last tile), against the
oracle's RK4 on the same stencils and weights, in FP64, in FP32, and in FP32 with the
closure's per-point multiplier handed over in FP16."""
import numpy as np
import pytest

import fp32bars
import oracle_lib as O
from se_rbf_projects_amd import closure as CL
from se_rbf_projects_amd import engine as E
from se_rbf_projects_amd.ctypes_api import F32, F64, config_deform, courant_dt

pytestmark = pytest.mark.gpu

STEPS = 8


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
    idx, _ = eng.stencils()
    Dx, Dy, L, st = eng.get_weights()
    assert not st.any()
    q0 = np.stack([E.initial_condition(k, cfg, xyz, seed=3) for k in CL.TRACERS], axis=1)
    dt = courant_dt(cfg, N)
    P = O.Problem(cfg, xyz, idx, Dx, Dy, L)
    oq = P.rk4(q0, STEPS, dt)
    yield dict(eng=eng, cfg=cfg, q0=q0, dt=dt, P=P, oq=oq, xyz=xyz)
    eng.set_kernel_version(7)
    eng.close()


def _run(eng, q0, prec, kver, steps, dt):
    eng.set_kernel_version(kver)
    eng.set_state(q0)
    eng.run(prec, steps, dt)
    return eng.state(prec)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_tile_plan_covers_config_D():
    """The shape takes the tile kernel: every tile's union fits (at most 320 columns)."""
    xyz = E.icos_nodes(5)
    xyz = xyz[E.spatial_order(xyz)]
    eng = E.Engine(0)
    try:
        eng.set_nodes(xyz)
        eng.knn(55)
        cfg = config_deform(len(xyz), Nv=60, n=55, deg=4, nv=7)
        cfg.ntracers = 4
        eng.configure(cfg)
        eng.weights()
        eng.set_state(np.zeros((len(xyz), 4, 60)))
        eng.run(F32, 1, 1.0)
        ntiles, umax = eng.tile_info()
    finally:
        eng.close()
    assert ntiles == (len(xyz) + 15) // 16 and 0 < umax <= 320


def test_config_D_shape_fp64_matches_oracle(deform60):
    d = deform60
    q = _run(d["eng"], d["q0"], F64, 5, STEPS, d["dt"])
    err = rel_l2(q, d["oq"])
    print(f"config D shape, kernel 5 FP64 vs oracle after {STEPS} steps: {err:.2e}")
    assert err < 1e-11
    assert np.max(np.abs(q[:, 3] - 1.0)) < 1e-12      # the constant tracer


def test_config_D_shape_fp32_matches_oracle(deform60):
    d = deform60
    q = _run(d["eng"], d["q0"], F32, 5, STEPS, d["dt"])
    qc = _run(d["eng"], d["q0"], F32, 2, STEPS, d["dt"])
    fp32bars.check("config_D_shape_v5", rel_l2(q, d["oq"]))
    fp32bars.check("config_D_shape_v2", rel_l2(qc, d["oq"]))
    assert not np.array_equal(q, qc)                     # the tile kernel ran
    assert np.max(np.abs(q[:, 3] - 1.0)) < 1e-5


def test_config_D_shape_fp32_vector_kernel_matches_oracle(deform60):
    """Kernel version 7 (FP32 vector tiles: a wavefront per tracer, float4 state accesses, the
    vertical stencil from the own columns by row shifts) at the headline's exact shape."""
    d = deform60
    q = _run(d["eng"], d["q0"], F32, 7, STEPS, d["dt"])
    q5 = _run(d["eng"], d["q0"], F32, 5, STEPS, d["dt"])
    fp32bars.check("config_D_shape_v7", rel_l2(q, d["oq"]))
    fp32bars.check("config_D_shape_v7_vs_v5", rel_l2(q, q5))
    assert not np.array_equal(q, q5)                     # the vector kernel ran
    assert np.max(np.abs(q[:, 3] - 1.0)) < 1e-5
    # FP64 under version 7 takes the version-5 kernel
    q64 = _run(d["eng"], d["q0"], F64, 7, STEPS, d["dt"])
    assert rel_l2(q64, d["oq"]) < 1e-11


def test_partial_last_tile_with_fields_there(deform60):
    """10242 nodes leave a last tile of 2 nodes whose union positions 2..15 are other nodes'
    columns.  A smooth field that is not constant on those nodes from the first step (the
    configured tracers are zero there at first): one FP32 RK4 step of the vector kernel (7) and the
    matrix-core kernel (5) against the oracle, at the last tile's nodes and everywhere."""
    d = deform60
    eng, xyz, dt = d["eng"], d["xyz"], d["dt"]
    N = len(xyz)
    assert N % 16 == 2
    lev = (np.arange(60) + 0.5) / 60
    q0 = np.stack([np.sin(3 * xyz[:, 0:1] + t) * np.cos(2 * xyz[:, 1:2]) + 0.1 * xyz[:, 2:3] * lev[None, :]
                   for t in range(4)], axis=1)
    oq = d["P"].rk4(q0, 1, dt)
    last = slice(N - N % 16, N)
    dq = np.max(np.abs(oq - q0))
    for kver in (7, 5):
        q = _run(eng, q0, F32, kver, 1, dt)
        err_last = np.max(np.abs(q[last] - oq[last])) / dq
        print(f"kernel {kver}: last tile error / step change {err_last:.2e}, rel l2 {rel_l2(q, oq):.2e}")
        assert err_last < 1e-3
        assert rel_l2(q, oq) < 1e-6
    eng.set_kernel_version(7)


def test_config_D_shape_without_hyperviscosity(deform60):
    d = deform60
    eng, cfg = d["eng"], d["cfg"]
    coef = cfg.hv_coef
    cfg.hv_coef = 0.0
    eng.configure(cfg)
    eng.weights()
    try:
        q64 = _run(eng, d["q0"], F64, 5, 4, d["dt"])
        q32 = _run(eng, d["q0"], F32, 5, 4, d["dt"])
        q32v = _run(eng, d["q0"], F32, 7, 4, d["dt"])
        idx, _ = eng.stencils()
        Dx, Dy, L, _ = eng.get_weights()
    finally:
        cfg.hv_coef = coef
        eng.configure(cfg)
        eng.weights()
    oq = O.Problem(_no_hv(cfg), d["xyz"], idx, Dx, Dy, L).rk4(d["q0"], 4, d["dt"])
    assert rel_l2(q64, oq) < 1e-11
    fp32bars.check("config_D_shape_v5_no_hv", rel_l2(q32, oq))
    fp32bars.check("config_D_shape_v7_no_hv", rel_l2(q32v, oq))


def _no_hv(cfg):
    import ctypes as C

    c = type(cfg)()
    C.memmove(C.addressof(c), C.addressof(cfg), C.sizeof(c))
    c.hv_coef = 0.0
    return c


@pytest.mark.parametrize("kver", [5, 7])
def test_config_D_closure_fp16_multiplier_matches_oracle(deform60, kver):
    """The headline leg's path: the closure's multiplier s handed to the FP32 tile kernel in
    FP16, which forms c = hv_coef * s in every stage.  Against the oracle's RK4 with the same
    per-point coefficient (the FP32 evaluation, rounded through FP16 as the kernel sees it),
    evaluated at the first step and held, as every = 16 > steps does on the GPU."""
    d = deform60
    eng, cfg, q0, dt = d["eng"], d["cfg"], d["q0"], d["dt"]
    cl = E.Closure(0)
    cl.set_params(CL.init_params(4, out_scale=0.5), np.concatenate([np.zeros(16), np.ones(16)]).astype(np.float32))
    eng.set_kernel_version(kver)
    try:
        eng.set_state(q0)
        eng.attach_closure(cl, every=16, s_max=16.0)
        g32 = eng.closure_gamma(F32, 0.0, dt)
        hv = np.float32(cfg.hv_coef)
        s16 = (g32.astype(np.float32) / hv).astype(np.float16)
        gamma = (hv * s16.astype(np.float32)).astype(np.float64)
        assert np.std(s16.astype(np.float64)) > 1e-3 and s16.max() <= 16.0
        eng.set_state(q0)
        eng.run(F32, STEPS, dt)
        q32 = eng.state(F32)
        eng.set_state(q0)
        g64 = eng.closure_gamma(F64, 0.0, dt)
        eng.set_state(q0)
        eng.run(F64, STEPS, dt)
        q64 = eng.state(F64)
    finally:
        eng.attach_closure(None)
        eng.set_kernel_version(7)
        cl.close()
    P = d["P"]
    qo = q0.copy()
    for s in range(STEPS):
        t = s * dt
        f1 = P.rhs(qo, t, gamma)
        f2 = P.rhs(qo + dt / 2 * f1, t + dt / 2, gamma)
        f3 = P.rhs(qo + dt / 2 * f2, t + dt / 2, gamma)
        f4 = P.rhs(qo + dt * f3, t + dt, gamma)
        qo = qo + dt / 6 * (((f1 + 2 * f2) + 2 * f3) + f4)
    fp32bars.check("config_D_closure_fp16" + ("" if kver == 5 else f"_v{kver}"), rel_l2(q32, qo))
    # FP64 with its own (FP64) coefficient against the oracle with that coefficient
    qo = q0.copy()
    for s in range(STEPS):
        t = s * dt
        f1 = P.rhs(qo, t, g64)
        f2 = P.rhs(qo + dt / 2 * f1, t + dt / 2, g64)
        f3 = P.rhs(qo + dt / 2 * f2, t + dt / 2, g64)
        f4 = P.rhs(qo + dt * f3, t + dt, g64)
        qo = qo + dt / 6 * (((f1 + 2 * f2) + 2 * f3) + f4)
    assert rel_l2(q64, qo) < 1e-11


@pytest.mark.parametrize("scale", [1.0, 1.0e-6, 3.0e5])
def test_closure_tile_evaluation_matches_oracle(deform60, scale):
    """The FP32 closure evaluation on the tile plan (the five feature operators as split-FP16
    matrix-core products, the vertical derivative by row shifts, the network per 64 points) at the
    headline's shape: its multipliers against the oracle's features and BF16 network on 20000
    sampled points, with the per-node stability bound, and against the per-point FP64 evaluation
    everywhere.  The state scaled by 1e-6 (FP16 subnormals without the tile's power-of-two scale)
    and by 3e5 (past the FP16 range) checks that scale.  At 3e5 the features themselves are
    ill-conditioned in any FP32 arithmetic: they are asinh of K h^k D q, so on the constant plateaus
    of the slotted cylinders, where D q is exactly 0, FP32 rounding of the sums (2^-24 of 3e5 times
    the row sums of h^8 Delta^4, up to ~2e3) becomes arguments of order 100, an error of up to ~14
    in the normalised input where plain FP32 products are emulated in numpy.  There the test
    asks for finite multipliers within the clamps and agreement at 95% of the points."""
    d = deform60
    eng, cfg, dt = d["eng"], d["cfg"], d["dt"]
    q0 = d["q0"] * scale
    params, norm, rep = CL.load()
    cl = E.Closure(0)
    cl.set_params(params, norm)
    t = 0.3 * cfg.T
    try:
        eng.set_state(q0)
        eng.attach_closure(cl, every=16, s_max=rep["s_max"])
        bound = eng.closure_bound(rep["kappa"])
        g32 = eng.closure_gamma(F32, t, dt) / cfg.hv_coef
        g64 = eng.closure_gamma(F64, t, dt) / cfg.hv_coef
        idx, _ = eng.stencils()
        Dx, Dy, L, _ = eng.get_weights()
        L1, L2 = eng.lap_weights()
        vstart, vW = eng.vertical()
    finally:
        eng.attach_closure(None)
        cl.close()
    feat, _ = O.closure_features(cfg, d["xyz"], idx, Dx, Dy, L, L1, L2, vstart, vW, q0, t, dt)
    rng = np.random.default_rng(5)
    pick = rng.choice(len(feat), 20000, replace=False)
    o = O.mlp_forward(params, norm, feat[pick], emulate=True)
    sref = np.minimum(np.where(o > 20.0, o, np.log1p(np.exp(np.minimum(o, 20.0)))), rep["s_max"])
    node = pick // (4 * cfg.Nv)
    sref = np.minimum(sref, bound[node].astype(np.float64))
    got = g32.reshape(-1)[pick]
    # the network's BF16 operands leave relative differences of a few 1e-3 in s (up to ~60 here)
    rel = np.abs(got - sref) / np.maximum(1.0, sref)
    rel64 = np.abs(g32 - g64) / np.maximum(1.0, g64)
    print(f"scale {scale:g}: tile evaluation vs oracle: share rel < 4e-3: {np.mean(rel < 4e-3):.4f}, max {rel.max():.2e}; "
          f"vs FP64 per-point: share rel < 4e-3: {np.mean(rel64 < 4e-3):.4f}, max {rel64.max():.2e}")
    assert np.all(np.isfinite(g32)) and np.all(g32 >= 0.0)
    assert np.all(g32 <= rep["s_max"]) and np.all(g32 <= bound[:, None, None] + 1e-3)
    if scale > 1.0:
        assert np.mean(rel < 4e-3) > 0.95 and np.mean(rel64 < 4e-3) > 0.95
        return
    assert np.mean(rel < 4e-3) > 0.99 and rel.max() < 2e-2
    # FP32 and FP64 feature sums can flip a BF16 rounding of the network input at a few points
    assert np.mean(rel64 < 4e-3) > 0.99 and rel64.max() < 0.1


def test_closure_tile_evaluation_non_finite_state(deform60):
    """A NaN at one level of one node's tracer-0 column: every multiplier stays finite; tracer 0
    falls back to the constant coefficient (s = 1) at that level of every node whose stencil holds
    that node (the levels are separate columns of the products; on the tile plan the NaN reaches
    every row of the tiles whose union holds it, since 0 x NaN is NaN, and the 12-level vertical
    window of the node itself) and every point it changes is at s = 1 (the network alone would
    drop the NaN in a ReLU and return an arbitrary finite value); tracers 1..3 are bit for bit
    unchanged (their products, features and scales never see tracer 0)."""
    d = deform60
    eng, cfg, dt = d["eng"], d["cfg"], d["dt"]
    q0 = d["q0"]
    params, norm, rep = CL.load()
    cl = E.Closure(0)
    cl.set_params(params, norm)
    t = 0.3 * cfg.T
    p = 5000
    try:
        eng.set_state(q0)
        eng.attach_closure(cl, every=16, s_max=rep["s_max"])
        bound = eng.closure_bound(rep["kappa"])
        ref = eng.closure_gamma(F32, t, dt) / cfg.hv_coef
        qn = q0.copy()
        qn[p, 0, 10] = np.nan
        eng.set_state(qn)
        got = eng.closure_gamma(F32, t, dt) / cfg.hv_coef
        idx, _ = eng.stencils()
    finally:
        eng.attach_closure(None)
        cl.close()
        eng.set_state(q0)
    assert np.all(np.isfinite(got))
    assert np.array_equal(got[:, 1:], ref[:, 1:])
    fallback = np.minimum(1.0, bound.astype(np.float64))
    hit = np.nonzero((idx == p).any(axis=1))[0]
    assert len(hit) > 0
    assert np.allclose(got[hit, 0, 10], fallback[hit], rtol=1e-6, atol=0)
    cn, cv = np.nonzero(got[:, 0] != ref[:, 0])
    assert np.allclose(got[cn, 0, cv], fallback[cn], rtol=1e-6, atol=0)
    print(f"NaN at node {p}, level 10: {len(hit)} stencils hold it; {len(cn)} points fell back to s = 1, "
          f"levels {sorted(set(cv.tolist()))}")
