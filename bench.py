"""Benchmark of the RBF-FD RK4 hot path on one MI355X.

Workload (mixed-precision RK4 configuration): icosahedral level-8 nodes (655362)
x 60 levels, kNN stencils n = 55, PHS r^7 + degree-4 tangent-plane polynomials,
vertical 1D RBF-FD n = 7, Delta^4 hyperviscosity, the deformational flow, four
tracers (Gaussian hills, slotted cylinders, cosine bells, constant).  Stencils
and weights are built on the GPU (kNN + batched FP64 weight solves) before the
timed region; inputs are resident in HBM when it starts.

One step = one classic RK4 step (4 fused stage launches) over the whole state.
One point-update = one (node, level) advanced one RK4 step, all four tracers.
Timed legs on the same inputs:
  f64          the plain method in FP64 (hyperviscosity in every stage)
  f32+closure  FP32 state and SpMV, hyperviscosity in every stage with the learned
               closure's per-point coefficients (re-evaluated every 128 steps)  -> value
  f32          the same with the constant coefficient (its stage kernel gives the roofline)
Both precisions use 16-node tiles on the matrix cores: FP32 the vector tile kernel (version 7),
FP64 the matrix-core tile kernel (version 5).
After the timed legs, the f64 and f32+closure legs run over the full 12-day period of the
reversing flow (about 7000 steps each, untimed for `value`) and the line states their error
against the exact solution q(T) = q(0) (--no-period skips this; --period-sweep adds FP32 at
constant coefficients and at the closure's per-node stability bound alone, for comparison).
Secondary legs (reported beside the headline, not in it): the uneven batched weight solves at
the same level with a LAPACK residual on sampled nodes (config C), and the plain FP64 method
over a full period one level lower (163842 nodes x 60, one field of two Gaussian hills;
config B).  The CPU baseline is the oracle built here at -O3 -march=native, timed on all host
cores over a bounded sample of the workload, plus one thread on the CPU-runnable case (2562
nodes x 10 levels, 100 steps).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-period] [--period-sweep]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_F32_PEAK_TFLOPS = 157.3   # dense FP32 matrix peak of MI355X
TILE = 16                       # nodes per matrix-core tile


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def stage_bytes(N, n, Nv, ntr, bq, bw, basis=16):
    """Algorithmic HBM bytes of one fused stage launch, averaged over the 4 stages.

    State: per point and stage 4.25 values on average (stage 1: gather q_T, write F
    and q_B = 3; stages 2-3: gather, read q_T, read+write F, write q = 5; stage 4:
    gather, read q_T and F, write q_T = 4), each gathered column read once.
    Stencil: n indices (4 B) + 3 weight sets (Dx, Dy, L) per node; velocity basis
    (16 values) per node."""
    return N * (ntr * Nv * 4.25 * bq + n * (3 * bw + 4) + basis * bw)


def tile_flops(idx, S, ntr, hv=True):
    """Matrix-core work of one stage launch of the tile kernel: per 16-node tile the union of
    the stencils, rounded up to 16 columns, times 16 nodes x S levels x ntr tracers x 2 flops,
    once for the advection operator and once for Delta^k; and the useful part of it (n stencil
    entries per node)."""
    N, n = idx.shape
    nt = (N + TILE - 1) // TILE
    pad = np.full((nt * TILE - N, n), -1, idx.dtype)
    keys = np.sort(np.concatenate([idx, pad]).reshape(nt, TILE * n), axis=1)
    u = 1 + np.count_nonzero(np.diff(keys, axis=1), axis=1) - (keys[:, 0] < 0)
    u16 = (u + 15) // 16 * 16
    ops = 2 if hv else 1
    dense = float(np.sum(u16)) * TILE * S * ntr * 2 * ops
    useful = float(N) * n * S * ntr * 2 * ops
    return dense, useful, float(np.mean(u))


def load_traffic(kernel_key):
    """Measured HBM bytes per launch of the stage kernel from the committed rocprofv3 counter
    summary (profiles/stage_kernel_traffic.json: L2 read requests by size plus WRITE_SIZE), or
    None."""
    path = os.path.join(ROOT, "profiles", "stage_kernel_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d["kernels"][kernel_key]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def error_vs(q, qref):
    """Relative l2 / linf over the three transported tracers, and l2 per tracer."""
    d = q[:, :3] - qref[:, :3]
    return {"rel_l2": float(np.linalg.norm(d) / np.linalg.norm(qref[:, :3])),
            "rel_linf": float(np.max(np.abs(d)) / np.max(np.abs(qref[:, :3]))),
            "rel_l2_per_tracer": [float(np.linalg.norm(d[:, i]) / np.linalg.norm(qref[:, i])) for i in range(3)]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=128, help="timed steps (a multiple of the closure cadence)")
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--nv", type=int, default=60)
    ap.add_argument("--n", type=int, default=55)
    ap.add_argument("--cpu-steps", type=int, default=2, help="RK4 steps of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-mb", type=int, default=1024, help="size of the stream-copy bandwidth reference")
    ap.add_argument("--no-period", action="store_true", help="skip the full-period error legs")
    ap.add_argument("--period-sweep", action="store_true",
                    help="with the period legs, also FP32 at constant coefficients (1, 28, 30 and the multipliers "
                         "given by --sweep-multipliers) and at the closure's stability bound alone")
    ap.add_argument("--sweep-multipliers", type=str, default="", help="comma-separated extra constant multipliers")
    ap.add_argument("--closure-every", type=int, default=128, help="steps between closure re-evaluations")
    ap.add_argument("--no-config-c", action="store_true", help="skip the uneven batched weight-solve leg")
    ap.add_argument("--no-config-b", action="store_true", help="skip the FP64 full-period leg at one level lower")
    ap.add_argument("--kernel-version", type=int, default=7, choices=(2, 5, 7),
                    help="stage kernel: 7 FP32 vector tiles + FP64 matrix-core tiles (default), 5 matrix-core "
                         "tiles in both precisions, 2 column kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU path is not part of "
              "this build", file=sys.stderr)
        sys.exit(2)

    from se_rbf_projects_amd import closure as CL
    from se_rbf_projects_amd import engine as E
    from se_rbf_projects_amd.ctypes_api import F32, F64, config_deform, courant_dt

    t0 = time.time()
    xyz = E.icos_nodes(args.level)
    xyz = xyz[E.spatial_order(xyz)]
    N = len(xyz)
    ntr = 4
    cfg = config_deform(N, Nv=args.nv, n=args.n, deg=4, nv=7)
    cfg.ntracers = ntr
    log(f"nodes N={N} generated in {time.time() - t0:.1f}s")

    eng = E.Engine(0)
    eng.set_kernel_version(args.kernel_version)
    eng.set_nodes(xyz)
    t = time.time()
    eng.knn(cfg.n)
    knn_s = time.time() - t
    log(f"kNN n={cfg.n}: {knn_s:.3f}s")
    eng.configure(cfg)
    t = time.time()
    eng.weights()
    w_s = time.time() - t
    log(f"weights: {w_s:.3f}s ({N / w_s:.3e} nodes/s, 2 systems per node)")

    q0 = np.stack([E.initial_condition(k, cfg, xyz, seed=2024) for k in CL.TRACERS], axis=1)
    eng.set_state(q0)
    nsteps_T = int(np.ceil(cfg.T / courant_dt(cfg, N)))
    dt = cfg.T / nsteps_T

    if os.path.exists(CL.PARAMS_PATH):
        params, norm, rep = CL.load()
        closure_kind = "trained"
    else:
        params, norm, rep = CL.init_params(0), np.concatenate([np.zeros(16), np.ones(16)]).astype(np.float32), {}
        closure_kind = "untrained (constant multiplier 1; no closure_params.npz)"
    s_max = float(rep.get("s_max", CL.S_MAX))
    kappa = rep.get("kappa")    # the per-node stability bound the network was trained under
    closure = E.Closure(0)
    closure.set_params(params, norm)

    def attach(use_cl):
        eng.attach_closure(closure if use_cl else None, every=args.closure_every, s_max=s_max)
        if use_cl and kappa:
            eng.closure_bound(kappa)

    legs = (("f64", F64, False, False), ("f32+closure", F32, False, True), ("f32", F32, False, False))
    res, final = {}, {}
    for name, prec, split, use_cl in legs:
        eng.set_hv_split(split)
        attach(use_cl)
        eng.set_state(q0)
        eng.run(prec, args.warmup, dt)
        eng.set_state(q0)
        ms, stage_ms = eng.run_timed(prec, args.steps, dt)
        res[name] = (ms, stage_ms)
        final[name] = eng.state(prec)
        log(f"{name}: {args.steps} steps {ms:.2f} ms, {ms / args.steps:.3f} ms/step, "
            f"{ms / (4 * args.steps) * 1e3:.1f} us per stage (closure included)")
    q64 = final["f64"]
    short = {name: error_vs(final[name], q64) for name in ("f32+closure", "f32")}
    log(f"after {args.steps} steps, difference from the f64 leg: {short}")
    pts = N * cfg.Nv
    value = {name: pts * args.steps / (res[name][0] / 1e3) for name in res}
    b32 = stage_bytes(N, cfg.n, cfg.Nv, ntr, 4, 4)
    achieved = b32 / (res["f32"][1] / 1e3) / 1e9
    b64 = stage_bytes(N, cfg.n, cfg.Nv, ntr, 8, 8)
    achieved64 = b64 / (res["f64"][1] / 1e3) / 1e9

    period = None
    if not args.no_period:
        period = {"steps": nsteps_T, "dt_s": dt, "reference": "exact solution q(T) = q(0) of the reversing flow",
                  "norms": "relative l2 / linf over the three transported tracers"}
        for name, prec, split, use_cl in legs[:2]:
            eng.set_hv_split(split)
            attach(use_cl)
            eng.set_state(q0)
            t = time.time()
            eng.run(prec, nsteps_T, dt)
            q = eng.state(prec)
            period[name] = error_vs(q, q0)
            period[name]["mass_const_tracer"] = float(np.mean(q[:, 3]) - 1.0)
            period[name]["seconds"] = time.time() - t
            log(f"period {name}: {period[name]}")
        period["f32+closure_vs_f64_l2"] = period["f32+closure"]["rel_l2"] / period["f64"]["rel_l2"]
        if args.period_sweep:
            # the closure against the constant coefficient at the largest stable multipliers and
            # against its stability bound alone (s = kappa G_ref / G_p at every point)
            eng.set_hv_split(False)
            eng.attach_closure(None)
            mults = [1.0, 28.0, 30.0] + [float(m) for m in args.sweep_multipliers.split(",") if m.strip()]
            fields = [(f"f32 constant coefficient x {m:g}", np.full(N, m)) for m in mults]
            if kappa:
                b = eng.closure_bound(kappa)
                eng.closure_bound(0.0)
                fields.append((f"f32 stability bound alone (kappa {kappa:g})", np.minimum(b.astype(np.float64), s_max)))
            for key, sfield in fields:
                eng.set_state(q0)
                eng.set_gamma(np.repeat(cfg.hv_coef * sfield[:, None], ntr * cfg.Nv, axis=1))
                eng.set_state(q0)
                eng.run(F32, nsteps_T, dt)
                period[key] = error_vs(eng.state(F32), q0)
                eng.set_gamma(None)
                log(f"period {key}: {period[key]}")
            eng.set_state(q0)
    eng.attach_closure(None)
    eng.set_hv_split(False)

    copy_ms = eng.stream_copy(args.copy_mb << 20, reps=10)
    copy_gbs = 2 * (args.copy_mb << 20) / (copy_ms / 1e3) / 1e9
    log(f"stream copy {args.copy_mb} MiB: {copy_ms:.3f} ms, {copy_gbs:.0f} GB/s (read + write)")

    weights_c = None if args.no_config_c else config_c_weights(E, args.level)
    period_b = None if args.no_config_b else config_b_period(E, args.level - 1, args.kernel_version)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(eng, cfg, xyz, q0, dt, args.cpu_steps)

    kern = "fp32 step (Delta^4 in every stage)"
    # the committed counter summary is of the default stage kernel
    traffic = load_traffic(kern) if args.kernel_version == 7 else None
    idx, _ = eng.stencils()
    dense_fl, useful_fl, umean = tile_flops(idx, 64, ntr)
    mfma_tf = dense_fl / (res["f32"][1] / 1e3) / 1e12
    out = {
        "metric": "RK4 point-updates/s @655362x60, FP32+closure vs FP64 plain; l2 err",
        "value": value["f32+closure"],
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["f32+closure"][0] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (generated icosahedral nodes, analytic flow and tracers, seed 2024)",
        "config": {"workload": f"mixed-precision RK4 + learned closure, icosahedral {N} nodes x {cfg.Nv} levels x "
                               f"{ntr} tracers, deformational flow, n={cfg.n} deg 4, vertical n=7, Delta^4 "
                               "hyperviscosity",
                   "nodes": N, "levels": cfg.Nv, "tracers": ntr, "stencil": cfg.n, "dt_s": dt,
                   "point_update": "one (node, level) advanced one RK4 step (all 4 tracers)",
                   "closure": closure_kind, "closure_every_steps": args.closure_every, "closure_s_max": s_max,
                   "closure_stability_bound_kappa": kappa,
                   "stage_kernel": {2: "column kernel (version 2)",
                                    5: "matrix-core tiles (version 5)",
                                    7: "FP32 vector tiles (version 7), FP64 matrix-core tiles (version 5)"}[args.kernel_version]
},
        "fp64_plain": {"value": value["f64"], "ms_per_step": res["f64"][0] / args.steps,
                       "stage_us": res["f64"][1] * 1e3, "roofline_achieved_GBs": achieved64,
                       "roofline_frac": achieved64 / HBM_PEAK_GBS},
        "f32_no_closure": {"value": value["f32"], "ms_per_step": res["f32"][0] / args.steps},
        "speedup_vs_fp64_plain": value["f32+closure"] / value["f64"],
        "tracer_point_updates_per_s": value["f32+closure"] * ntr,
        "error": {"after_steps": args.steps, "vs_f64_leg": short, "period": period},
        "stream_copy_GBs": copy_gbs,
        "stage_frac_of_copy": achieved / copy_gbs,
        "setup": {"knn_s": knn_s, "weights_s": w_s, "weight_solves_per_s": 2 * N / w_s},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "frac_of_tuned_copy": achieved / copy_gbs,
                     "kernel": {7: "mvec_kernel<STAGE, 1> (FP32 vector tiles, constant coefficient)",
                                5: "mtile2_kernel<float, STAGE, 1, 4, 2> (matrix-core tiles)",
                                2: "stage2_kernel (column kernel)"}[args.kernel_version]
                               + ", average of the 4 launches of a step",
                     "bytes_per_launch": b32,
                     "note": "average FP32 stage launch of the f32 leg (HIP events); algorithmic bytes model "
                             "bench.stage_bytes; traffic = L2 read requests x size (128/64/32 B) + WRITE_SIZE per launch",
                     "mfma_view": {"dense_flops_per_launch": dense_fl, "useful_flops_per_launch": useful_fl,
                                   "mean_tile_union": umean, "achieved_TFLOPs": mfma_tf,
                                   "peak_TFLOPs": MFMA_F32_PEAK_TFLOPS, "frac": mfma_tf / MFMA_F32_PEAK_TFLOPS}
                     if args.kernel_version in (5, 7) else None},
        "cpu_baseline": cpu,
        "config_c_weights": weights_c,
        "config_b_fp64_period": period_b,
    }
    print(json.dumps(out), flush=True)
    closure.close()
    eng.close()


def config_c_weights(E, level):
    """Batched uneven weight solves: jittered icosahedral nodes (tangential U(0, 0.3 h)
    displacements, seed 7), stencil sizes 31..84 from local density and degrees 3..5, all
    (n + M)^2 FP64 saddle systems in one launch with the autotuned schedule."""
    from se_rbf_projects_amd.ctypes_api import config_deform, variable_stencil_sizes
    xyz = E.icos_nodes(level)
    xyz = E.jitter_nodes(xyz, 0.3, 7)
    xyz = xyz[E.spatial_order(xyz)]
    N = len(xyz)
    cfg = config_deform(N, Nv=60, n=84, deg=4, nv=7)
    eng = E.Engine(0)
    eng.set_nodes(xyz)
    eng.knn(84)
    _, D = eng.stencils()
    nn, deg = variable_stencil_sizes(D)
    eng.configure(cfg)
    eng.set_stencil_sizes(nn, deg)
    t = time.time()
    rep = eng.weights_autotune(sample=32768)
    tune_s = time.time() - t
    eng.weights()
    ms, systems, v = eng.weights_stats()
    Dx, Dy, _, st = eng.get_weights()
    lapack = lapack_residuals(xyz, eng.stencils()[0], nn, deg, Dx, Dy, samples=256)
    M = (deg + 1) * (deg + 2) // 2
    m = (nn + M).astype(np.float64)
    flops = float(np.sum(2.0 / 3.0 * m ** 3 + 2 * m ** 2 * 4 + 4 * m ** 2 * 4)
                  + np.sum(2.0 / 3.0 * (nn + 1.0) ** 3 + 6 * (nn + 1.0) ** 2))
    eng.close()
    out = {"nodes": N, "stencil_sizes": [int(nn.min()), int(nn.max())], "degrees": sorted(set(deg.tolist())),
           "systems": systems, "kernel_ms": ms, "solves_per_s": systems / (ms * 1e-3), "variant": v,
           "autotune": [{"variant": a, "ms": b, "solves_per_s": c} for a, b, c in rep], "autotune_s": tune_s,
           "fp64_tflops": flops / (ms * 1e-3) / 1e12, "failed_systems": int(np.count_nonzero(st)),
           "vs_lapack": lapack}
    log(f"config C vs LAPACK on {lapack['sampled_nodes']} sampled nodes: max weight difference "
        f"{lapack['max_weight_rel_diff']:.2e}, max scaled residual {lapack['max_scaled_residual']:.2e}")
    log(f"config C weights: {systems:.0f} systems in {ms:.1f} ms ({out['solves_per_s']:.3e}/s, variant {v}, "
        f"{out['fp64_tflops']:.2f} TFLOP/s FP64)")
    return out


def config_b_period(E, level, kernel_version):
    """The plain FP64 method at scale: icosahedral level 7 (163842 nodes) x 60 levels, n = 55,
    degree-4 tangent polynomials, vertical n = 7, Delta^4, the deformational flow, one field of
    two Gaussian hills (seed 2024), run over the full 12-day period and compared with its
    initial state (the exact solution of the reversing flow)."""
    from se_rbf_projects_amd.ctypes_api import F64, config_deform, courant_dt
    xyz = E.icos_nodes(level)
    xyz = xyz[E.spatial_order(xyz)]
    N = len(xyz)
    cfg = config_deform(N, Nv=60, n=55, deg=4, nv=7)
    eng = E.Engine(0)
    eng.set_kernel_version(kernel_version)
    eng.set_nodes(xyz)
    eng.knn(cfg.n)
    eng.configure(cfg)
    eng.weights()
    q0 = E.initial_condition(E.IC_HILLS, cfg, xyz, seed=2024)
    nT = int(np.ceil(cfg.T / courant_dt(cfg, N)))
    dt = cfg.T / nT
    eng.set_state(q0)
    eng.run(F64, 8, dt)
    eng.set_state(q0)
    ms, stage_ms = eng.run_timed(F64, nT, dt)
    q = eng.state(F64)
    eng.close()
    d = q - q0
    out = {"nodes": N, "levels": cfg.Nv, "tracers": 1, "steps": nT, "dt_s": dt, "ms_per_step": ms / nT,
           "stage_us": stage_ms * 1e3, "point_updates_per_s": N * cfg.Nv * nT / (ms * 1e-3),
           "period_rel_l2": float(np.linalg.norm(d) / np.linalg.norm(q0)),
           "period_rel_linf": float(np.max(np.abs(d)) / np.max(np.abs(q0))),
           "mass_rel": float((np.sum(q) - np.sum(q0)) / np.sum(q0))}
    log(f"config B: {N} x {cfg.Nv} FP64, {nT} steps in {ms / 1e3:.2f} s ({ms / nT:.3f} ms/step, "
        f"{out['point_updates_per_s']:.3e} point-updates/s), period rel l2 {out['period_rel_l2']:.4f}, "
        f"linf {out['period_rel_linf']:.4f}")
    return out


def householder(p):
    """The reflector H = I - beta v v^T with H p = e_z (the stable branch for z > 0), and the
    frame diag(1, -1, -1) at the south pole, as the weight solves use in config mode."""
    x, y, z = p
    lon = x * x + y * y
    if lon == 0.0:
        return np.eye(3) if z > 0 else np.diag([1.0, -1.0, -1.0])
    v2 = -lon / (z + 1.0) if z > 0.0 else z - 1.0
    beta = 2.0 * v2 * v2 / (lon + v2 * v2)
    v = np.array([x / v2, y / v2, 1.0])
    return np.eye(3) - beta * np.outer(v, v)


def lapack_residuals(xyz, idx, nn, deg, Dx, Dy, samples=256, phs=7):
    """The config-mode gradient systems of sampled nodes (PHS r^phs + degree-d tangent
    monomials on the node's rotated, centred stencil scaled by its radius) assembled in numpy
    and solved with numpy.linalg.solve (LAPACK dgesv): the largest difference of the GPU
    weights from LAPACK's, and the largest scaled residual ||A w - b|| / (||A|| ||w||) of the
    GPU weights in the same systems (polynomial part completed by least squares)."""
    N = len(Dx)
    worst_w, worst_r, count = 0.0, 0.0, 0
    for p in np.linspace(0, N - 1, samples).astype(int):
        n, d = int(nn[p]), int(deg[p])
        X = (householder(xyz[idx[p, 0]]) @ xyz[idx[p, :n]].T).T
        X = X - X[0]
        s = np.max(np.linalg.norm(X, axis=1))
        X = X / s
        r = np.linalg.norm(X[:, None, :] - X[None, :, :], axis=2)
        mons = [(q - j, j) for q in range(d + 1) for j in range(q + 1)]
        P = np.stack([X[:, 0] ** a * X[:, 1] ** b for a, b in mons], axis=1)
        M = len(mons)
        A = np.zeros((n + M, n + M))
        A[:n, :n] = r ** phs
        A[:n, n:] = P
        A[n:, :n] = P.T
        r0 = np.linalg.norm(X, axis=1)
        b = np.zeros((n + M, 2))
        b[:n, 0] = phs * (-X[:, 0]) * r0 ** (phs - 2)
        b[:n, 1] = phs * (-X[:, 1]) * r0 ** (phs - 2)
        b[n + mons.index((1, 0)), 0] = 1.0
        b[n + mons.index((0, 1)), 1] = 1.0
        ref = np.linalg.solve(A, b)
        got = np.zeros_like(ref)
        got[:n, 0] = Dx[p, :n] * s
        got[:n, 1] = Dy[p, :n] * s
        worst_w = max(worst_w, float(np.max(np.abs(got[:n] - ref[:n])) / np.max(np.abs(ref[:n]))))
        got[n:] = np.linalg.lstsq(A[:, n:], b - A[:, :n] @ got[:n], rcond=None)[0]
        worst_r = max(worst_r, float(np.linalg.norm(A @ got - b) / (np.linalg.norm(A) * np.linalg.norm(got))))
        count += 1
    return {"sampled_nodes": count, "max_weight_rel_diff": worst_w, "max_scaled_residual": worst_r,
            "systems": "gradient systems (PHS r^7 + tangent monomials), numpy.linalg.solve"}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(eng, cfg, xyz, q0, dt, steps):
    """The oracle's FP64 RK4 (C + OpenMP over nodes), compiled here at -O3 -march=native,
    on the same stencils and weights for a bounded number of steps on all host cores, and on
    one core over the CPU-runnable case (2562 nodes x 10 levels, n = 31, 100 steps)."""
    import subprocess
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native"], check=True,
                       capture_output=True, text=True)
        native = os.path.join(ROOT, "oracle", "_ref", "liboracle_native.so")
        build = "gcc -O3 -march=native -fopenmp -ffp-contract=off"
        import oracle_lib as O
        Ol = O.lib(native)
    except Exception as ex:  # pragma: no cover
        log(f"native oracle build unavailable ({ex}); timing the shipped build")
        native, build = None, "gcc -O3 -fopenmp -ffp-contract=off"
        import oracle_lib as O
        Ol = O.lib()
    from se_rbf_projects_amd.ctypes_api import F64, config_A, courant_dt
    cores = len(os.sched_getaffinity(0))
    Ol.orc_set_threads(cores)
    threads = Ol.orc_max_threads()
    idx, _ = eng.stencils()
    Dx, Dy, L, _ = eng.get_weights()
    P = O.Problem(cfg, xyz, idx, Dx, Dy, L)
    t = time.time()
    q_cpu = P.rk4(q0, steps, dt)
    s = time.time() - t
    eng.set_hv_split(False)
    eng.set_state(q0)
    eng.run(F64, steps, dt)
    parity = float(np.linalg.norm(eng.state(F64) - q_cpu) / np.linalg.norm(q_cpu))
    log(f"full-scale parity, GPU FP64 vs oracle after {steps} steps: rel l2 {parity:.3e}")
    pts = len(xyz) * cfg.Nv
    log(f"cpu baseline: {steps} steps in {s:.2f}s on {threads} threads")
    # one thread on the CPU-runnable case
    from se_rbf_projects_amd import engine as E
    xa = E.icos_nodes(4)
    ca = config_A(len(xa))
    ia, _ = O.knn(xa, ca.n)
    wa = O.weights(ca, xa, ia)
    qa = O.initial(E.IC_BELL, ca, xa)
    dta = courant_dt(ca, len(xa))
    Ol.orc_set_threads(1)
    t = time.time()
    O.Problem(ca, xa, ia, wa[0], wa[1], wa[2]).rk4(qa, 100, dta)
    s1 = time.time() - t
    Ol.orc_set_threads(threads)
    log(f"cpu baseline, 1 thread, 2562 x 10 x 100 steps: {s1:.2f}s")
    return {"value": pts * steps / s, "unit": "point-updates/s", "cores": threads,
            "kind": "port", "sample": f"{steps} RK4 steps of the full workload (FP64 oracle, OpenMP over nodes, "
                                      f"{threads} threads)",
            "cpu": cpu_model(), "build": build,
            "one_thread_config_A": {"value": len(xa) * ca.Nv * 100 / s1, "unit": "point-updates/s",
                                    "sample": f"{len(xa)} nodes x {ca.Nv} levels, n={ca.n}, 100 RK4 steps, "
                                              "FP64, 1 thread", "seconds": s1},
            "gpu_f64_vs_oracle_rel_l2": parity}


if __name__ == "__main__":
    main()
