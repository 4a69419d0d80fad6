"""Benchmark of the RBF-FD RK4 hot path on one MI355X.

Workload (mixed-precision RK4 configuration): icosahedral level-8 nodes (655362)
x 60 levels, kNN stencils n = 55, PHS r^7 + degree-4 tangent-plane polynomials,
vertical 1D RBF-FD n = 7, Delta^4 hyperviscosity, the deformational flow, four
tracers (Gaussian hills, slotted cylinders, cosine bells, constant).  Stencils
and weights are built on the GPU (kNN + batched FP64 weight solves) before the
timed region; inputs are resident in HBM when it starts.

One step = one classic RK4 step (4 fused stage launches) over the whole state.
value = FP32 tracer point-updates/s (N * Nv * tracers * steps / s); the FP64
method on the same inputs is timed beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--nv", type=int, default=60)
    ap.add_argument("--n", type=int, default=55)
    ap.add_argument("--tracers", type=int, default=4)
    ap.add_argument("--cpu-steps", type=int, default=2, help="RK4 steps of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-mb", type=int, default=1024, help="size of the stream-copy bandwidth reference")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU path is not part of "
              "this build", file=sys.stderr)
        sys.exit(2)

    from se_rbf_projects_amd import engine as E
    from se_rbf_projects_amd.ctypes_api import F32, F64, config_deform, courant_dt

    t0 = time.time()
    xyz = E.icos_nodes(args.level)
    xyz = xyz[E.spatial_order(xyz)]
    N = len(xyz)
    cfg = config_deform(N, Nv=args.nv, n=args.n, deg=4, nv=7)
    cfg.ntracers = args.tracers
    ntr = args.tracers
    log(f"nodes N={N} generated in {time.time() - t0:.1f}s")

    eng = E.Engine(0)
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

    kinds = [E.IC_HILLS, E.IC_SLOTTED, E.IC_BELLS2, E.IC_CONST][:ntr]
    q0 = np.stack([E.initial_condition(k, cfg, xyz, seed=2024) for k in kinds], axis=1)
    eng.set_state(q0)
    dt = courant_dt(cfg, N)

    res, final = {}, {}
    for prec, name, split in ((F64, "f64", False), (F32, "f32", True)):
        eng.set_hv_split(split)
        eng.set_state(q0)
        eng.run(prec, args.warmup, dt)
        eng.set_state(q0)
        ms, stage_ms = eng.run_timed(prec, args.steps, dt)
        res[name] = (ms, stage_ms)
        final[name] = eng.state(prec)
        log(f"{name}: {args.steps} steps {ms:.2f} ms, {ms / args.steps:.3f} ms/step, stage {stage_ms * 1e3:.1f} us")
    q64, q32 = final["f64"], final["f32"]
    err32 = float(np.linalg.norm(q32 - q64) / np.linalg.norm(q64))
    linf32 = float(np.max(np.abs(q32 - q64)) / np.max(np.abs(q64)))
    drift = {name: float(np.linalg.norm(q - q0) / np.linalg.norm(q0)) for name, q in (("f64", q64), ("f32", q32))}
    log(f"f32 vs f64 rel l2 {err32:.3e}, rel linf {linf32:.3e}; change from q0: {drift}")
    pts = N * cfg.Nv * ntr
    v32 = pts * args.steps / (res["f32"][0] / 1e3)
    v64 = pts * args.steps / (res["f64"][0] / 1e3)
    b32 = stage_bytes(N, cfg.n, cfg.Nv, ntr, 4, 4)
    achieved = b32 / (res["f32"][1] / 1e3) / 1e9
    b64 = stage_bytes(N, cfg.n, cfg.Nv, ntr, 8, 8)
    achieved64 = b64 / (res["f64"][1] / 1e3) / 1e9

    copy_ms = eng.stream_copy(args.copy_mb << 20, reps=10)
    copy_gbs = 2 * (args.copy_mb << 20) / (copy_ms / 1e3) / 1e9
    log(f"stream copy {args.copy_mb} MiB: {copy_ms:.3f} ms, {copy_gbs:.0f} GB/s (read + write)")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(eng, cfg, xyz, q0, dt, args.cpu_steps)

    out = {
        "metric": "RK4 point-updates/s @655362x60, FP32+closure vs FP64 plain; l2 err",
        "value": v32,
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["f32"][0] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (generated icosahedral nodes, analytic flow and tracers)",
        "config": {"workload": f"mixed-precision RK4, icosahedral {N} nodes x {cfg.Nv} levels x {ntr} tracers, "
                               f"deformational flow, n={cfg.n} deg 4, vertical n=7, Delta^4 hyperviscosity",
                   "nodes": N, "levels": cfg.Nv, "tracers": ntr, "stencil": cfg.n, "dt_s": dt,
                   "point_update": "one (node, level, tracer) advanced one RK4 step"},
        "fp64": {"value": v64, "ms_per_step": res["f64"][0] / args.steps, "stage_us": res["f64"][1] * 1e3,
                 "roofline_achieved_GBs": achieved64, "roofline_frac": achieved64 / HBM_PEAK_GBS},
        "speedup_f32_vs_f64": v32 / v64,
        "error": {"f32_vs_f64_rel_l2": err32, "f32_vs_f64_rel_linf": linf32, "after_steps": args.steps,
                  "rel_l2_change_from_q0": drift},
        "stream_copy_GBs": copy_gbs,
        "setup": {"knn_s": knn_s, "weights_s": w_s, "weight_solves_per_s": 2 * N / w_s},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "stage_kernel<float,64,*,4,true>", "bytes_per_launch": b32},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    eng.close()


def cpu_baseline(eng, cfg, xyz, q0, dt, steps):
    """The oracle's FP64 RK4 (C + OpenMP over nodes) on the same stencils and
    weights, for a bounded number of steps on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        import oracle_lib as O
    except Exception as ex:  # pragma: no cover
        log(f"cpu baseline unavailable: {ex}")
        return None
    idx, _ = eng.stencils()
    Dx, Dy, L, _ = eng.get_weights()
    P = O.Problem(cfg, xyz, idx, Dx, Dy, L)
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    t = time.time()
    q_cpu = P.rk4(q0, steps, dt)
    s = time.time() - t
    from se_rbf_projects_amd.ctypes_api import F64
    eng.set_hv_split(False)
    eng.set_state(q0)
    eng.run(F64, steps, dt)
    parity = float(np.linalg.norm(eng.state(F64) - q_cpu) / np.linalg.norm(q_cpu))
    log(f"full-scale parity, GPU FP64 vs oracle after {steps} steps: rel l2 {parity:.3e}")
    pts = len(xyz) * cfg.Nv * max(1, cfg.ntracers)
    log(f"cpu baseline: {steps} steps in {s:.2f}s on {cores} cores")
    return {"value": pts * steps / s, "unit": "point-updates/s", "cores": int(os.environ.get("OMP_NUM_THREADS", cores)),
            "kind": "port", "sample": f"{steps} RK4 steps of the full workload (FP64 oracle, OpenMP over nodes)",
            "cpu": platform.processor() or platform.machine(), "gpu_f64_vs_oracle_rel_l2": parity}


if __name__ == "__main__":
    main()
