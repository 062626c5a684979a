"""Writes tests/golden/cfg5_fp64_reference.npz: the FP64 plain form's solution
of the dense 16384-sphere workload (p=12, vf 0.2, chains; configs.config5()),
the reference against which the mixed form's velocity error is reported there.

A full FP64 solve of that workload (GMRES(30) with IMGS to 1e-8, direct far
sums: about 60 applies of 24 s) is longer than one GPU job, so it runs in two
jobs split at the first GMRES restart, where the whole solver state is the
iterate x:

  phase cycle1  cold FP64 solve capped at 30 iterations (one restart cycle);
                stores x30 - x_mixed, quantised per sphere to 16 bits
                (x_mixed, the mixed form's solution of the same system, is
                recomputed bit for bit by the second job), which moves the
                iterate by about 1e-5 * 2^-15 of its size: far below the 1e-8
                tolerance, so the second cycle starts where an uninterrupted
                run's would;
  phase finish  the FP64 solve warm-started from that iterate to 1e-8; writes
                the velocities, the iteration count (30 + the second job's)
                and the timing of both jobs.

Run on the GPU box from the repository root, in two jobs:
  python3 tests/golden/make_cfg5_fp64_reference.py --phase cycle1 --work gpurun_out/cfg5_ref
  python3 tests/golden/make_cfg5_fp64_reference.py --phase finish --work gpurun_out/cfg5_ref
The second job needs the first job's files in --work (gpurun_out/ comes back
to the caller, who carries it into the second job); it writes
cfg5_fp64_reference.npz there, which is committed as tests/golden/.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def log(msg):
    print("[cfg5-ref] " + msg, flush=True)


def mixed_solution(S, c):
    M = S.SphereSTKMobMat(c["pos"], c["radius"], c["quat"], c["order"], precision=S.PREC_MIXED)
    t0 = time.perf_counter()
    v, info = M.solveMob(c["force"])
    log("mixed solve %.1f s: outer %d inner %d relres %.3e" % (time.perf_counter() - t0, info["outer"], info["iters"],
                                                                info["relres"]))
    return M.get_unknown(), v, info


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--phase", choices=("cycle1", "finish"), required=True)
    ap.add_argument("--work", required=True)
    a = ap.parse_args()
    os.makedirs(a.work, exist_ok=True)
    import spheresimulator_amd as S
    t0 = time.perf_counter()
    c = S.configs.config5()
    log("inputs %.1f s" % (time.perf_counter() - t0))
    xm, vm, im = mixed_solution(S, c)
    if a.phase == "cycle1":
        M = S.SphereSTKMobMat(c["pos"], c["radius"], c["quat"], c["order"], precision=S.PREC_FP64,
                              params=S.solver_params(max_iters=30), timing=True)
        t0 = time.perf_counter()
        v, info = M.solveMob(c["force"])
        dt = time.perf_counter() - t0
        log("FP64 cycle 1: %.1f s, iters %d relres %.3e applies %d" % (dt, info["iters"], info["relres"],
                                                                      info["applies_fp64"]))
        x30 = M.get_unknown()
        off = np.concatenate([[0], np.cumsum(3 * (c["order"].astype(np.int64) + 1) * (2 * c["order"] + 2))])
        dx = x30 - xm
        scale = np.array([np.abs(dx[off[i]:off[i + 1]]).max() for i in range(len(c["radius"]))]) / 32767.0
        scale[scale == 0] = 1.0
        q = np.rint(dx / np.repeat(scale, np.diff(off))).astype(np.int16)
        recon = xm + q * np.repeat(scale, np.diff(off))
        log("quantised x30: relative change %.3e" % (np.linalg.norm(recon - x30) / np.linalg.norm(x30)))
        np.save(os.path.join(a.work, "dq.npy"), q)
        np.save(os.path.join(a.work, "dscale.npy"), scale)
        st = M.op.kernel_stats(0, 0)
        meta = dict(cycle1_s=dt, cycle1_iters=info["iters"], cycle1_relres=info["relres"],
                    cycle1_applies=info["applies_fp64"], xm_norm=float(np.linalg.norm(xm)),
                    xm_sum=float(xm.sum()), trac_ms_per_launch=st["ms"] / max(st["launches"], 1),
                    trac_pairs=st["pairs"])
        json.dump(meta, open(os.path.join(a.work, "cycle1.json"), "w"), indent=1)
        return
    meta = json.load(open(os.path.join(a.work, "cycle1.json")))
    q = np.load(os.path.join(a.work, "dq.npy"))
    scale = np.load(os.path.join(a.work, "dscale.npy"))
    log("x_mixed here vs cycle 1: norm %.17g / %.17g" % (np.linalg.norm(xm), meta["xm_norm"]))
    off = np.concatenate([[0], np.cumsum(3 * (c["order"].astype(np.int64) + 1) * (2 * c["order"] + 2))])
    x30 = xm + q * np.repeat(scale, np.diff(off))
    M = S.SphereSTKMobMat(c["pos"], c["radius"], c["quat"], c["order"], precision=S.PREC_FP64, timing=True)
    M.set_unknown(x30)
    t0 = time.perf_counter()
    v, info = M.solveMob(c["force"])
    dt = time.perf_counter() - t0
    log("FP64 finish: %.1f s, iters %d relres %.3e converged %d applies %d" % (
        dt, info["iters"], info["relres"], info["converged"], info["applies_fp64"]))
    if not info["converged"]:
        raise SystemExit("FP64 solve did not converge")
    vr = v.reshape(-1, 6)
    vmr = vm.reshape(-1, 6)
    eV = np.linalg.norm(vmr[:, :3] - vr[:, :3]) / np.linalg.norm(vr[:, :3])
    eW = np.linalg.norm(vmr[:, 3:] - vr[:, 3:]) / np.linalg.norm(vr[:, 3:])
    log("mixed vs FP64: rel err V %.3e Omega %.3e" % (eV, eW))
    st = M.op.kernel_stats(0, 0)
    info_all = dict(
        workload="16384 spheres p=12 lognormal R vf=0.2 chains (configs.config5, seed 5)",
        fp64_iters=int(meta["cycle1_iters"] + info["iters"]), fp64_relres=float(info["relres"]),
        # an uninterrupted solve does not repeat the right-hand side, the
        # residual at the restart and the final single layer of the two jobs
        fp64_applies=int(meta["cycle1_applies"] + info["applies_fp64"] - 3),
        fp64_solve_s=float(meta["cycle1_s"] + dt - 3 * st["ms"] / max(st["launches"], 1) / 1000.0),
        cycle1_s=meta["cycle1_s"], finish_s=dt,
        trac_ms_per_launch=[meta["trac_ms_per_launch"], st["ms"] / max(st["launches"], 1)],
        trac_pairs_per_launch=st["pairs"], mixed_rel_err_V=float(eV), mixed_rel_err_Omega=float(eW),
        mixed_iters=int(im["iters"]), mixed_outer=int(im["outer"]))
    log(json.dumps(info_all))
    np.savez(os.path.join(a.work, "cfg5_fp64_reference.npz"), vel=v, meta=json.dumps(info_all))
    json.dump(info_all, open(os.path.join(a.work, "finish.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
