"""Profiling driver for the far-field kernels' counters.

Runs three FP64 traction far-field applies on the 1024-sphere p=12 workload
(with --inner also three applies of the mixed form's FP32 inner operator) and
then one 512 MiB device-to-device copy of known size, whose FETCH_SIZE /
WRITE_SIZE readings calibrate the traffic counters.  Meant to run under
`rocprofv3 --pmc <counters>`, one pass per counter group:
    python3 -m spheresimulator_amd.profile_far [--inner]
"""
import sys

import numpy as np


def main():
    import torch

    import spheresimulator_amd as S
    c = S.configs.config2()
    op = S.SphereSTKSHOperator(c["pos"], c["radius"], c["quat"], c["order"], 0.5, 0.0, 1.0, 1.0)
    x = np.random.default_rng(0).standard_normal(op.n_dof)
    out = np.zeros(op.n_dof)
    for _ in range(3):
        op.applyP2POP(x, out, 0.0, 0.0, 1.0)
    if "--inner" in sys.argv[1:]:
        for _ in range(3):
            op.apply_inner(x)
    n = 64 * 1024 * 1024
    a = torch.randn(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    print("done", flush=True)


if __name__ == "__main__":
    main()
