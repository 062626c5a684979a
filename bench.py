"""Benchmark: mobility solves/sec to GMRES 1e-8 on one MI355X.

A step is one full mobility solve (SphereSTKMobMat::solveMob,
BIE/SphereSTKMobMat.cpp:146-286): RHS traction apply, GMRES(30) to
||r||/||b|| <= 1e-8 with the explicit residual check, the final single-layer
apply and the rigid-body averages, from a cold start (x0 = 0) with the inputs
resident in device memory.  Both forms are timed on the same workload: the FP64
plain form and the mixed form (FP32 far field inside FP64 iterative
refinement); `value` is the mixed form's solves/sec, the FP64 figure, their
ratio and the mixed form's velocity error against the FP64 form are reported
beside it.  The far-field traction kernel's roofline comes from HIP events
recorded by the library on its own stream.  The CPU baseline is the oracle's
FP64 far field (the dominant cost) on a bounded sample of targets, scaled to one
solve with the GPU-measured iteration count.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_VALU_PEAK = 78.6  # TFLOP/s, MI355X FP64 vector (datasheet)
FP32_VALU_PEAK = 157.3  # TFLOP/s, MI355X FP32 vector (datasheet; counts packed FP32, which this build does not emit)
FLOPS_TRAC = 30  # algorithmic flops per far traction pair
FLOPS_SL = 29


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def run_form(S, c, precision, steps, warmup):
    t0 = time.perf_counter()
    M = S.SphereSTKMobMat(c["pos"], c["radius"], c["quat"], c["order"], counts=c["counts"], precision=precision,
                          timing=True)
    setup = time.perf_counter() - t0
    zero = np.zeros(M.op.n_dof)
    info = None
    for _ in range(warmup):
        M.set_unknown(zero)
        v, info = M.solveMob(c["force"])
    M.op.reset_stats()
    times = []
    for _ in range(steps):
        M.set_unknown(zero)
        t0 = time.perf_counter()
        v, info = M.solveMob(c["force"])
        times.append(time.perf_counter() - t0)
    stats = {(k, p): M.op.kernel_stats(kind=k, prec=p) for k in range(2) for p in range(3)}
    return M, v, info, times, stats, setup


def pmc_traffic(c, pairs):
    """HBM bytes per launch of the FP64 traction kernel from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE passes (profiles/), when they cover this workload."""
    path = os.path.join(ROOT, "profiles", "far_traction_f64_cfg2_pmc.json")
    if not os.path.exists(path) or not c["name"].startswith("cfg2"):
        return None
    with open(path) as f:
        return json.load(f)["traffic_bytes"]


def cpu_baseline(c, iters, budget_s=15.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    op = O.Operator(c["pos"], c["radius"], c["quat"], c["order"])
    x = np.random.default_rng(0).standard_normal(op.dof)
    n = op.npts
    k = max(64, min(n, 4096))
    t0 = time.perf_counter()
    op.far_sums_range(x, 0, k, cTrac=1.0)
    dt = time.perf_counter() - t0
    # grow the sample to fill roughly budget_s of CPU work
    k2 = int(min(n, max(k, k * budget_s / max(dt, 1e-6) * 0.5)))
    if k2 > k:
        t0 = time.perf_counter()
        op.far_sums_range(x, 0, k2, cTrac=1.0)
        dt = time.perf_counter() - t0
        k = k2
    t_apply = dt * n / k
    applies = iters + 3  # RHS + Krylov iterations + explicit residual + final single layer
    return dict(value=1.0 / (t_apply * applies), unit="solves/s", cores=O.num_threads(), kind="port",
                sample="oracle FP64 far-field traction sum for %d of %d targets (%.1f s), scaled to %d applies "
                       "per solve (direct sum in place of STKFMM)" % (k, n, dt, applies))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp64,mixed")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (requested --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import spheresimulator_amd as S
    log("building config %d" % args.config)
    c = S.configs.CONFIGS[args.config]()
    res = {}
    for form in args.forms.split(","):
        prec = S.PREC_FP64 if form == "fp64" else S.PREC_MIXED
        log("timing %s form: %d warmup + %d steps" % (form, args.warmup, args.steps))
        M, v, info, times, stats, setup = run_form(S, c, prec, args.steps, args.warmup)
        res[form] = dict(v=v, info=info, times=times, stats=stats, n_dof=M.op.n_dof, near=M.op.n_near_pairs,
                         setup=setup)
        log("%s: setup %.3f s, %.3f s/solve, iters=%d outer=%d relres=%.2e" % (
            form, setup, np.mean(times), info["iters"], info["outer"], info["relres"]))
        del M
    main_form = "mixed" if "mixed" in res else "fp64"
    r = res[main_form]
    ms = 1000.0 * float(np.mean(r["times"]))
    out = {"metric": "mobility solves/sec to GMRES 1e-8 (mixed FP32-far-field + FP64 refinement vs FP64), 1 GPU",
           "value": 1000.0 / ms, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if main_form == "fp64" else "f32+f64", "data": "synthetic",
           "config": {"workload": c["name"], "spheres": int(len(c["radius"])), "sh_order": int(c["order"].max()),
                      "dof": int(r["n_dof"]), "near_pairs": int(r["near"]), "gmres_tol": 1e-8, "restart": 30,
                      "form": main_form}}
    # one-time set-up (geometry, near list, work lists, level grids; the
    # reference rebuilds its operator every time step) and the rate with it
    out["setup_s"] = {k: float(v["setup"]) for k, v in res.items()}
    out["solves_per_s_with_setup"] = {k: 1.0 / (float(v["setup"]) + float(np.mean(v["times"])))
                                      for k, v in res.items()}
    if "fp64" in res:
        f = res["fp64"]
        out["fp64_solves_per_s"] = 1.0 / float(np.mean(f["times"]))
        out["fp64_iters"] = f["info"]["iters"]
        st = f["stats"][(0, 0)]  # traction, FP64
        if st["launches"]:
            avg_s = st["ms"] / st["launches"] / 1000.0
            ach = FLOPS_TRAC * st["pairs"] / avg_s / 1e12
            out["roofline"] = {"bound": "valu", "kernel": "k_far_f64<traction>", "achieved": ach,
                               "peak": FP64_VALU_PEAK, "unit": "TFLOP/s", "frac": ach / FP64_VALU_PEAK,
                               "traffic": pmc_traffic(c, st["pairs"]), "avg_launch_ms": st["ms"] / st["launches"],
                               "pairs_per_launch": st["pairs"], "flops_per_pair": FLOPS_TRAC}
    if "mixed" in res and "fp64" in res:
        vm = res["mixed"]["v"].reshape(-1, 6)
        vf = res["fp64"]["v"].reshape(-1, 6)
        out["mixed_vs_fp64"] = {
            "speedup": float(np.mean(res["fp64"]["times"]) / np.mean(res["mixed"]["times"])),
            "rel_err_V": float(np.linalg.norm(vm[:, :3] - vf[:, :3]) / np.linalg.norm(vf[:, :3])),
            "rel_err_Omega": float(np.linalg.norm(vm[:, 3:] - vf[:, 3:]) / np.linalg.norm(vf[:, 3:])),
            "mixed_iters": res["mixed"]["info"]["iters"], "mixed_outer": res["mixed"]["info"]["outer"],
            "mixed_fp32_applies": res["mixed"]["info"]["applies_fp32"],
            "mixed_fp64_applies": res["mixed"]["info"]["applies_fp64"]}
        st = res["mixed"]["stats"][(0, 1)]  # traction, FP32 inner operator (all levels of one apply)
        if st["launches"]:
            avg_s = st["ms"] / st["launches"] / 1000.0
            ach = FLOPS_TRAC * st["pairs"] / avg_s / 1e12
            out["mixed_vs_fp64"]["fp32_far"] = {
                "bound": "valu", "achieved": ach, "peak": FP32_VALU_PEAK, "unit": "TFLOP/s",
                "frac": ach / FP32_VALU_PEAK, "avg_apply_ms": st["ms"] / st["launches"], "pairs_per_apply": st["pairs"]}
        st = res["mixed"]["stats"][(0, 2)]  # traction, FP64 multi-level outer applies
        if st["launches"]:
            avg_s = st["ms"] / st["launches"] / 1000.0
            ach = FLOPS_TRAC * st["pairs"] / avg_s / 1e12
            out["mixed_vs_fp64"]["fp64_outer_far"] = {
                "bound": "valu", "achieved": ach, "peak": FP64_VALU_PEAK, "unit": "TFLOP/s",
                "frac": ach / FP64_VALU_PEAK, "avg_apply_ms": st["ms"] / st["launches"], "pairs_per_apply": st["pairs"]}
    if not args.no_cpu:
        it = res.get("fp64", r)["info"]["iters"]
        log("CPU baseline (oracle, bounded sample)")
        out["cpu_baseline"] = cpu_baseline(c, it)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
