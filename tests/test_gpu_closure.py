"""The learned closure on the GPU against the oracle: the BF16 matrix-core forward and
backward passes, Adam, the features, the coefficients inside the RK4, and what a trained
network gains over the constant coefficient."""
import numpy as np
import pytest

import fp32bars
import oracle_lib as O
from se_rbf_projects_amd import closure as CL
from se_rbf_projects_amd import engine as E
from se_rbf_projects_amd.ctypes_api import F32, F64, config_deform, courant_dt

pytestmark = pytest.mark.gpu


def _problem(seed, m=2048, out_scale=0.5):
    rng = np.random.default_rng(seed)
    params = CL.init_params(seed, out_scale=out_scale)
    params[CL.O_B1:CL.O_W2] = rng.normal(0, 0.3, CL.NH)
    params[CL.O_B2:CL.O_W3] = rng.normal(0, 0.3, CL.NH)
    norm = np.zeros(32, np.float32)
    norm[:10] = rng.normal(0, 0.2, 10)
    norm[16:26] = rng.uniform(0.5, 1.5, 10)
    feat = np.zeros((m, 16), np.float32)
    feat[:, :10] = rng.normal(0, 1, (m, 10))
    A = rng.normal(0, 1, m).astype(np.float32)
    R = rng.normal(0, 1, m).astype(np.float32)
    return params, norm, feat, A, R


def test_mfma_forward_matches_oracle():
    params, norm, feat, A, R = _problem(11, m=1024)
    cl = E.Closure(0)
    cl.set_params(params, norm)
    cl.set_data(feat, A, R)
    o = cl.eval(0, 1024)
    ref = O.mlp_forward(params, norm, feat, emulate=True)
    exact = O.mlp_forward(params, norm, feat, emulate=False)
    print("forward: max |o - oracle(bf16)|", np.max(np.abs(o - ref)), " vs FP64", np.max(np.abs(o - exact)))
    assert np.max(np.abs(o - ref)) < 2e-3 * max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(o - exact)) < 5e-2 * max(1.0, np.max(np.abs(exact)))
    # a sub-range starts at its own first sample
    assert np.array_equal(cl.eval(512, 256), o[512:768])


def test_backward_matches_oracle():
    params, norm, feat, A, R = _problem(12, m=2048)
    cl = E.Closure(0)
    cl.set_params(params, norm)
    cl.set_data(feat, A, R)
    g, loss = cl.grad(0, 2048)
    ref, lref = O.mlp_grad(params, norm, feat, A, R, emulate=True)
    assert loss == pytest.approx(lref, rel=1e-4)
    blocks = {"W1": (0, CL.O_B1), "b1": (CL.O_B1, CL.O_W2), "W2": (CL.O_W2, CL.O_B2), "b2": (CL.O_B2, CL.O_W3),
              "w3": (CL.O_W3, CL.O_B3), "b3": (CL.O_B3, CL.NPARAMS)}
    for name, (a, b) in blocks.items():
        err = np.linalg.norm(g[a:b] - ref[a:b]) / np.linalg.norm(ref[a:b])
        print(f"grad {name}: rel err {err:.2e}")
        assert err < 3e-3, name
    # a second range (the tail) is independent of the first
    g2, l2 = cl.grad(1024, 1024)
    ref2, lref2 = O.mlp_grad(params, norm, feat[1024:], A[1024:], R[1024:], emulate=True)
    assert np.linalg.norm(g2 - ref2) / np.linalg.norm(ref2) < 3e-3
    assert l2 == pytest.approx(lref2, rel=1e-4)


def test_adam_steps_match_oracle():
    params, norm, feat, A, R = _problem(13, m=512)
    cl = E.Closure(0)
    cl.set_params(params, norm)
    cl.set_data(feat, A, R)
    p = params.copy()
    m = np.zeros(CL.NPARAMS)
    v = np.zeros(CL.NPARAMS)
    lr = 3e-3
    for step in range(1, 4):
        g, _ = cl.grad(0, 512)          # the gradient the next Adam step will use
        cl.train(1, 512, lr)
        p, m, v = O.adam(p, m, v, g.astype(np.float64), step, lr)
        got = cl.params()
        assert np.max(np.abs(got - p)) < 2e-6, step
    # the BF16 copies follow the parameters: the forward pass sees the updated network
    o = cl.eval(0, 512)
    assert np.max(np.abs(o - O.mlp_forward(p, norm, feat[:512], emulate=True))) < 2e-3


@pytest.fixture(scope="module")
def small(golden):
    N = 642
    xyz = golden.icos(N)[:, :3]
    cfg = config_deform(N, Nv=10, n=31, deg=4, nv=5)
    cfg.ntracers = 4
    eng = E.Engine(0)
    eng.set_nodes(xyz)
    eng.knn(cfg.n)
    eng.configure(cfg)
    eng.weights()
    q0 = np.stack([E.initial_condition(k, cfg, xyz, seed=5) for k in CL.TRACERS], axis=1)
    eng.set_state(q0)
    idx, _ = eng.stencils()
    Dx, Dy, L, _ = eng.get_weights()
    L1, L2 = eng.lap_weights()
    vstart, vW = eng.vertical()
    yield dict(eng=eng, cfg=cfg, xyz=xyz, q0=q0, idx=idx, W=(Dx, Dy, L, L1, L2), vstart=vstart, vW=vW,
               dt=courant_dt(cfg, N))
    eng.close()


def test_lap_weights_match_oracle(small):
    _, _, _, L1, L2, st = O.weights_lap(small["cfg"], small["xyz"], small["idx"])
    assert not st.any()
    for got, ref in zip(small["W"][3:], (L1, L2)):
        assert np.max(np.abs(got - ref)) < 1e-9 * np.max(np.abs(ref))


def test_features_match_oracle(small):
    eng, cfg = small["eng"], small["cfg"]
    eng.set_state(small["q0"])
    t = 2.5e5
    feat, lap = eng.closure_features(t, small["dt"])
    ref, lref = O.closure_features(cfg, small["xyz"], small["idx"], *small["W"], small["vstart"], small["vW"],
                                   small["q0"], t, small["dt"])
    d = np.abs(feat - ref)
    print("features: max abs diff per feature", d.max(axis=0)[:10])
    assert np.all(d <= 1e-5 * np.maximum(1.0, np.abs(ref)))
    assert np.max(np.abs(lap - lref)) < 1e-10 * np.max(np.abs(lref))


def _oracle_gamma(small, params, norm, t, q, s_max=8.0):
    cfg = small["cfg"]
    feat, _ = O.closure_features(cfg, small["xyz"], small["idx"], *small["W"], small["vstart"], small["vW"], q, t,
                                 small["dt"])
    o = O.mlp_forward(params, norm, feat, emulate=True)
    s = np.log1p(np.exp(np.minimum(o, 30.0)))
    s = np.where(np.isfinite(s), np.minimum(s, s_max), 1.0)
    return cfg.hv_coef * s.reshape(q.shape), feat


def test_coefficients_match_oracle_and_fall_back(small):
    eng, cfg = small["eng"], small["cfg"]
    params, norm, *_ = _problem(21, m=16, out_scale=0.6)
    cl = E.Closure(0)
    cl.set_params(params, norm)
    eng.set_state(small["q0"])
    eng.attach_closure(cl, every=8, s_max=8.0)
    t = 1.0e5
    g64 = eng.closure_gamma(F64, t, small["dt"])
    ref, _ = _oracle_gamma(small, params, norm, t, small["q0"])
    rel = np.abs(g64 - ref) / abs(cfg.hv_coef)
    print("gamma FP64: max |ds|", rel.max(), " s range", (ref / cfg.hv_coef).min(), (ref / cfg.hv_coef).max())
    assert np.mean(rel < 2e-3) > 0.999 and rel.max() < 5e-2
    g32 = eng.closure_gamma(F32, t, small["dt"])
    rel32 = np.abs(g32 - ref) / abs(cfg.hv_coef)
    assert np.mean(rel32 < 1e-2) > 0.99 and rel32.max() < 1e-1
    # an output far above s_max is clamped to s_max everywhere
    big = params.copy()
    big[CL.O_W3:CL.O_B3] = 0.0
    big[CL.O_B3] = 50.0
    cl.set_params(big, norm)
    gb = eng.closure_gamma(F64, t, small["dt"])
    assert np.all(gb == 8.0 * cfg.hv_coef)
    # a network that returns NaN falls back to the constant coefficient everywhere
    bad = params.copy()
    bad[CL.O_B3] = np.nan
    cl.set_params(bad, norm)
    gn = eng.closure_gamma(F64, t, small["dt"])
    assert np.all(gn == cfg.hv_coef)
    eng.attach_closure(None)
    cl.close()


def test_rk4_with_closure_matches_oracle(golden):
    """Eight FP64 steps with the closure re-evaluated at the first step (hyperviscosity in every
    stage) against the oracle RK4 with the same per-point coefficients; FP32 beside it."""
    N = 642
    xyz = golden.icos(N)[:, :3]
    cfg = config_deform(N, Nv=10, n=31, deg=4, nv=5)
    eng = E.Engine(0)
    eng.set_nodes(xyz)
    eng.knn(cfg.n)
    eng.configure(cfg)
    eng.weights()
    q0 = E.initial_condition(E.IC_SLOTTED, cfg, xyz, seed=1)
    eng.set_state(q0)
    dt = courant_dt(cfg, N)
    params, norm, *_ = _problem(31, m=16, out_scale=0.6)
    cl = E.Closure(0)
    cl.set_params(params, norm)
    eng.attach_closure(cl, every=8)
    gamma = eng.closure_gamma(F64, 0.0, dt)
    assert np.std(gamma / cfg.hv_coef) > 1e-3                     # a non-trivial coefficient field
    eng.set_state(q0)
    eng.run(F64, 8, dt)
    q64 = eng.state(F64)
    idx, _ = eng.stencils()
    Dx, Dy, L, _ = eng.get_weights()
    P = O.Problem(cfg, xyz, idx, Dx, Dy, L)
    qo = q0.copy()
    for s in range(8):
        t = s * dt
        f1 = P.rhs(qo, t, gamma)
        f2 = P.rhs(qo + dt / 2 * f1, t + dt / 2, gamma)
        f3 = P.rhs(qo + dt / 2 * f2, t + dt / 2, gamma)
        f4 = P.rhs(qo + dt * f3, t + dt, gamma)
        qo = qo + dt / 6 * (((f1 + 2 * f2) + 2 * f3) + f4)
    err = np.linalg.norm(q64 - qo) / np.linalg.norm(qo)
    print("RK4 + closure, FP64 vs oracle:", err)
    assert err < 1e-11
    eng.set_state(q0)
    eng.run(F32, 8, dt)
    fp32bars.check("closure_rk4_642x10_f32", np.linalg.norm(eng.state(F32) - q64) / np.linalg.norm(q64))
    eng.attach_closure(None)
    cl.close()
    eng.close()


def test_trained_closure_gains_over_constant_coefficient():
    """Train on coarse (642 nodes) vs restricted fine (10242 nodes) runs of the deformational
    flow.  On held-out samples the trained network must lower the M-step error the constant
    coefficient leaves (normalised to 1), and do at least as well as the best constant multiplier
    the stability clamp admits."""
    feat, A, R, meta = CL.make_dataset(fine_level=5, coarse_level=3, Nv=10, n=31, deg=4, nv=5, snapshots=6)
    params, norm, rep = CL.train(feat, A, R, steps=1500, batch=2048, lr=3e-3, log_every=500)
    print({k: v for k, v in rep.items() if k.startswith("loss") or k.startswith("best") or k.startswith("s_")})
    assert rep["loss_const_holdout"] == pytest.approx(1.0, rel=0.2)
    assert rep["loss_trained_holdout"] < 0.995 * rep["loss_const_holdout"]
    assert rep["loss_trained_holdout"] <= 1.002 * rep["loss_best_admissible_constant_holdout"]
