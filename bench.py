"""Benchmark: mobility solves/sec to GMRES 1e-8 on one MI355X.

Workload: the 16384-sphere, p=12 suspension (lognormal radii, vf 0.2, chains
of 8 with close gaps; configs.config5, seed 5), the size the target is quoted
at.  A step is one full mobility solve (SphereSTKMobMat::solveMob,
BIE/SphereSTKMobMat.cpp:146-286) from a cold start (x0 = 0) with the inputs
resident in device memory: the right-hand side, GMRES to ||r||/||b|| <= 1e-8,
the final single layer and the rigid-body averages.

* value: the mixed form's solves/sec (FP32 multi-level inner operator inside
  FP64 iterative refinement).
* FP64 plain form: one full FP64 operator apply is timed here; its solve time
  is that apply time x the applies of a completed FP64 solve of this workload
  (tests/golden/cfg5_fp64_reference.npz, made by
  tests/golden/make_cfg5_fp64_reference.py in two GPU jobs), labelled
  "extrapolated"; the fixture's own two-job wall-clock is reported beside it.
* Error: the mixed velocities against that FP64 solution.
* roofline: the mixed form's pair-sum kernel with the most time per timed
  solve, from HIP events the library records on its own stream around the
  launches of each apply: k_far_f64_lvl (the FP64 residual apply: mid range
  and six levels) or k_far_f32 (per apply seven lists: mid range, four
  levels, two source-cell lists; the coarse operator's applies and the FP32
  operator's); the other is reported under "mixed".
* cpu_baseline: the oracle's full operator apply (far, self and near terms)
  on a bounded sample of target spheres, scaled to a full FP64 solve with the
  FP64 apply count (kind "port", all host threads), labelled "extrapolated".
* The 1024-sphere and batched workloads follow as extra keys when time allows.

The driver runs `--steps 20 --warmup 5` under a time limit: when the first
solve shows that the requested steps would not fit `--budget` seconds, the
warmup is cut first and then the steps, and the line says so (`requested`).
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
REF_FIXTURE = os.path.join(ROOT, "tests", "golden", "cfg5_fp64_reference.npz")
T0 = time.perf_counter()


def log(msg):
    print("[bench %6.1f s] %s" % (time.perf_counter() - T0, msg), file=sys.stderr, flush=True)


def rel_err(v, ref):
    v, ref = v.reshape(-1, 6), ref.reshape(-1, 6)
    return (float(np.linalg.norm(v[:, :3] - ref[:, :3]) / np.linalg.norm(ref[:, :3])),
            float(np.linalg.norm(v[:, 3:] - ref[:, 3:]) / np.linalg.norm(ref[:, 3:])))


def solve_loop(M, force, steps, warmup, budget_end=None, fixed_after=0.0):
    """warmup untimed solves, then steps timed ones, each from x0 = 0.  With
    budget_end, the counts are cut (warmup first, at least one timed step) once
    the first solve's time shows they would not finish fixed_after seconds
    before it."""
    zero = np.zeros(M.n_dof)
    req_steps, req_warm = steps, warmup
    count = [0]

    def one():
        M.set_unknown(zero)
        t0 = time.perf_counter()
        v, info = M.solveMob(force)
        t = time.perf_counter() - t0
        count[0] += 1
        log("solve %d: %.3f s" % (count[0], t))  # a progress line per solve
        return t, v, info

    times = []
    if warmup == 0:
        M.op.reset_stats()
    t1, v, info = one()
    warm_done = 1 if warmup > 0 else 0
    if warmup == 0:
        times.append(t1)
    if budget_end is not None:
        fit = int((budget_end - time.perf_counter() - fixed_after) // (1.02 * t1))
        rw, rs = warmup - warm_done, steps - len(times)
        if rw + rs > fit:
            rw = max(0, min(rw, fit - rs))
            rs = max(0 if times else 1, min(rs, fit - rw))
            warmup, steps = warm_done + rw, len(times) + rs
    for _ in range(warmup - warm_done):
        one()
    if warm_done or warmup:
        M.op.reset_stats()
    while len(times) < steps:
        t, v, info = one()
        times.append(t)
    return v, info, times, dict(steps=steps, warmup=warmup, requested_steps=req_steps, requested_warmup=req_warm)


def far_roofline(st, peak, kernel, launches_per_apply):
    """Achieved rate of the far-field traction launches of one operator apply:
    algorithmic flops (FLOPS_TRAC per source-target pair, all lists of the
    apply) over the time between the HIP events the library records on its
    own stream before the first and after the last of those launches."""
    if not st["launches"]:
        return None
    apply_ms = st["ms"] / st["launches"]
    ach = FLOPS_TRAC * st["pairs"] / (apply_ms / 1000.0) / 1e12
    return {"bound": "valu", "kernel": kernel, "achieved": ach, "peak": peak, "unit": "TFLOP/s",
            "frac": ach / peak, "traffic": None, "avg_launch_ms": apply_ms / launches_per_apply,
            "launches_per_apply": launches_per_apply, "avg_apply_ms": apply_ms,
            "pairs_per_apply": st["pairs"], "flops_per_pair": FLOPS_TRAC, "applies_timed": st["launches"]}


def dominant(a, b, st_a, st_b, nsolves):
    """(the roofline object with more kernel time per solve, the other); each
    gains ms_per_solve."""
    for r, st in ((a, st_a), (b, st_b)):
        if r is not None:
            r["ms_per_solve"] = st["ms"] / max(1, nsolves)
    if b is not None and (a is None or b["ms_per_solve"] > a["ms_per_solve"]):
        return b, a
    return a, b


def cpu_baseline(c, applies, budget_s=8.0):
    """The oracle's full operator apply (far + self + near) on a sample of
    target spheres, scaled to all spheres and to `applies` applies per solve."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    ref = O.Operator(c["pos"], c["radius"], c["quat"], c["order"], 0.5, 0.0, 1.0, 1.0)
    x = np.random.default_rng(0).standard_normal(ref.dof)
    out = np.zeros(ref.dof)
    n = len(c["radius"])
    t0 = time.perf_counter()
    k = 0
    stride = max(1, n // 997)
    while k < n and (time.perf_counter() - t0 < budget_s or k == 0):
        i = (k * stride) % n
        ref.apply_p2p_spheres(x, i, i + 1, out, 0.5, 0.0, 1.0)
        k += 1
    dt = time.perf_counter() - t0
    t_apply = dt / k * n
    return dict(value=1.0 / (t_apply * applies), unit="solves/s", cores=O.num_threads(), kind="port",
                sample="extrapolated: the oracle's full FP64 apply (direct far sum in place of STKFMM, self and "
                       "near terms) on %d of %d target spheres (%.1f s), scaled to all spheres (%.1f s per apply) "
                       "and %d applies per FP64 solve" % (k, n, dt, t_apply, applies))


def fp64_apply_time(S, c, budget_left):
    """Set-up and one timed FP64 apply A64 x (direct far sums) on the workload."""
    t0 = time.perf_counter()
    op = S.SphereSTKSHOperator(c["pos"], c["radius"], c["quat"], c["order"], 0.5, 0.0, 1.0, 1.0, timing=True)
    setup = time.perf_counter() - t0
    x = np.random.default_rng(1).standard_normal(op.n_dof)
    t0 = time.perf_counter()
    op.apply(x)
    t_apply = time.perf_counter() - t0
    st = op.kernel_stats(0, 0)
    return setup, t_apply, st


def extra_workload(S, name, c, deadline, mixed_steps=2, fp64_steps=1):
    """Mixed (one warmup solve) and FP64 (no warmup) solves of a smaller
    workload, timed, not extrapolated."""
    res = {}
    for form, prec, steps, warm in (("mixed", S.PREC_MIXED, mixed_steps, 1), ("fp64", S.PREC_FP64, fp64_steps, 0)):
        if time.perf_counter() > deadline:
            res[form] = "skipped: time budget"
            continue
        t0 = time.perf_counter()
        M = S.SphereSTKMobMat(c["pos"], c["radius"], c["quat"], c["order"], counts=c["counts"], precision=prec)
        setup = time.perf_counter() - t0
        v, info, times, cnt = solve_loop(M, c["force"], steps, warm)
        res[form] = dict(solves_per_s=1.0 / float(np.mean(times)), s_per_solve=float(np.mean(times)),
                         setup_s=setup, iters=info["iters"], outer=info["outer"], relres=info["relres"],
                         steps=cnt["steps"], v=v)
        log("%s %s: setup %.2f s, %.3f s/solve, iters %d" % (name, form, setup, np.mean(times), info["iters"]))
        del M
    if isinstance(res.get("mixed"), dict) and isinstance(res.get("fp64"), dict):
        eV, eW = rel_err(res["mixed"]["v"], res["fp64"]["v"])
        res["mixed_vs_fp64"] = dict(speedup=res["mixed"]["solves_per_s"] / res["fp64"]["solves_per_s"],
                                    rel_err_V=eV, rel_err_Omega=eW)
    for f in ("mixed", "fp64"):
        if isinstance(res.get(f), dict):
            res[f].pop("v")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--budget", type=float, default=570.0, help="seconds the whole run may take")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (requested --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    end = T0 + args.budget
    import spheresimulator_amd as S
    log("building the 16384-sphere workload")
    c = S.configs.config5()
    ref = np.load(REF_FIXTURE, allow_pickle=False)
    ref_meta = json.loads(str(ref["meta"]))
    fp64_applies = int(ref_meta["fp64_applies"])

    # mixed form: the headline
    t0 = time.perf_counter()
    M = S.SphereSTKMobMat(c["pos"], c["radius"], c["quat"], c["order"], precision=S.PREC_MIXED, timing=True)
    setup_mixed = time.perf_counter() - t0
    log("mixed set-up %.2f s; solving (%d warmup + %d steps requested)" % (setup_mixed, args.warmup, args.steps))
    # reserve time for the FP64 apply (~24 s), the CPU sample, the smaller
    # workloads and writing the line
    reserve = 32.0 + (0.0 if args.no_cpu else 12.0) + (0.0 if args.no_extras else 42.0)
    v, info, times, cnt = solve_loop(M, c["force"], args.steps, args.warmup, budget_end=end, fixed_after=reserve)
    ms = 1000.0 * float(np.mean(times))
    log("mixed: %.3f s/solve over %d steps (warmup %d), iters %d outer %d relres %.2e" % (
        ms / 1000, cnt["steps"], cnt["warmup"], info["iters"], info["outer"], info["relres"]))
    st32 = M.op.kernel_stats(0, 1)
    stc = M.op.kernel_stats(0, 3)
    st64ml = M.op.kernel_stats(0, 2)
    n_dof, n_near = M.n_dof, M.op.n_near_pairs
    del M
    eV, eW = rel_err(v, ref["vel"])

    log("FP64 form: one timed apply")
    setup64, t_apply64, st64 = fp64_apply_time(S, c, end - time.perf_counter())
    t_fp64 = t_apply64 * fp64_applies
    log("FP64 apply %.2f s -> %.1f s per solve (%d applies)" % (t_apply64, t_fp64, fp64_applies))

    out = {"metric": "mobility solves/sec to GMRES 1e-8 (mixed FP32-far-field + FP64 refinement vs FP64), 1 GPU",
           "value": 1000.0 / ms, "unit": "solves/s", "n_gpus": 1, "steps": cnt["steps"], "warmup": cnt["warmup"],
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": "cfg5: 16384 spheres p=12, lognormal radii (sigma 0.2), vf 0.2, chains of 8",
                      "spheres": int(len(c["radius"])), "sh_order": 12, "dof": int(n_dof),
                      "near_pairs": int(n_near), "gmres_tol": 1e-8, "restart": 30, "form": "mixed"}}
    if (cnt["steps"], cnt["warmup"]) != (args.steps, args.warmup):
        out["requested"] = {"steps": args.steps, "warmup": args.warmup,
                            "note": "cut to fit the %.0f s budget after the first solve took %.1f s" %
                                    (args.budget, times[0] if times else ms / 1000)}
    out["mixed"] = {"iters": info["iters"], "outer": info["outer"], "relres": info["relres"],
                    "applies_fp32": info["applies_fp32"], "applies_fp64": info["applies_fp64"],
                    "setup_s": setup_mixed, "solves_per_s_with_setup": 1.0 / (setup_mixed + ms / 1000.0),
                    "rel_err_V_vs_fp64": eV, "rel_err_Omega_vs_fp64": eW}
    out["fp64"] = {"solves_per_s": 1.0 / t_fp64, "s_per_solve": t_fp64, "basis": "extrapolated",
                   "apply_s": t_apply64, "applies_per_solve": fp64_applies, "iters": ref_meta["fp64_iters"],
                   "setup_s": setup64, "two_job_solve_s": ref_meta["fp64_solve_s"],
                   "far_traction": far_roofline(st64, FP64_VALU_PEAK, "k_far_f64<traction>", 1)}
    out["mixed_vs_fp64"] = {"speedup": t_fp64 / (ms / 1000.0), "rel_err_V": eV, "rel_err_Omega": eW}
    # the mixed solve's pair-sum kernels: k_far_f32 (the coarse operator's lists
    # every Krylov iteration, the FP32 operator's for its residuals) and
    # k_far_f64_lvl (the FP64 residual apply); `roofline` is the one with more
    # time per timed solve, the other goes under "mixed"
    both = {"ms": stc["ms"] + st32["ms"], "launches": stc["launches"] + st32["launches"],
            "pairs": (stc["pairs"] * stc["launches"] + st32["pairs"] * st32["launches"]) /
                     max(1, stc["launches"] + st32["launches"])}
    f32 = far_roofline(both, FP32_VALU_PEAK, "k_far_f32<traction>: all FP32 pair-sum launches of a solve (coarse "
                       "operator: mid range, 4 levels, 2 cell lists per apply; FP32 operator: likewise)", 7)
    if f32 is not None:
        f32["coarse"] = far_roofline(stc, FP32_VALU_PEAK, "k_far_f32<traction>, coarse operator", 7)
        f32["inner"] = far_roofline(st32, FP32_VALU_PEAK, "k_far_f32<traction>, FP32 operator", 7)
    f64 = far_roofline(st64ml, FP64_VALU_PEAK, "k_far_f64_lvl<traction>: the mid-range list and 6 level lists of "
                       "one FP64 residual apply", 7)
    out["roofline"], other = dominant(f32, f64, both, st64ml, cnt["steps"])
    if other is not None:
        out["mixed"]["fp32_far" if other is f32 else "fp64_outer_far"] = other
    if not args.no_cpu:
        log("CPU baseline (oracle, bounded sample)")
        out["cpu_baseline"] = cpu_baseline(c, fp64_applies)
    if not args.no_extras:
        deadline = end - 5.0
        if time.perf_counter() + 10 < deadline:
            out["cfg2"] = extra_workload(S, "cfg2", S.configs.config2(), deadline, 3, 2)
        if time.perf_counter() + 30 < deadline:
            out["cfg4"] = extra_workload(S, "cfg4", S.configs.config4(), deadline, 2, 1)
    out["wall_s"] = time.perf_counter() - T0
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
