"""Times the batched dense solve in its two layouts on GRI-3.0 Newton matrices.

    python lu_bench.py [--matrices N] [--repeat R]

Layout 0 is one wave per matrix (the BDF kernel's register LU, FP64 and the FP32 inverse);
layout 1 is one 256-thread workgroup per matrix with the augmented matrix in LDS (FP64). The
matrices are M = I - gamma J for GRI states of the plain workload's generator, gamma log-uniform
in [1e-9, 1e-6] s (the range the BDF steps take), tiled to N. Prints one JSON line: device
milliseconds per launch (HIP events) and matrices per second for each layout.
"""
import argparse
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--matrices", type=int, default=65536)
    ap.add_argument("--states", type=int, default=2048, help="distinct Jacobians, tiled to --matrices")
    ap.add_argument("--repeat", type=int, default=3)
    a = ap.parse_args()
    import hpc_ode_amd as H
    from hpc_ode_amd import workloads as WL

    mech = H.Mechanism(os.path.join(HERE, "hpc_ode_amd", "mechanisms", "gri30.yaml"))
    rng = np.random.default_rng(5)
    y, P, _ = WL.gri_plain(mech.species, mech.W, a.states, seed=5)
    J = mech.jac(y, P)
    gam = 10.0 ** rng.uniform(-9, -6, a.states)
    M0 = np.eye(mech.n_state)[None] - gam[:, None, None] * J
    reps = -(-a.matrices // a.states)
    M = np.tile(M0, (reps, 1, 1))[: a.matrices]
    b = rng.standard_normal((a.matrices, mech.n_state))
    out = {"matrices": a.matrices, "n": mech.n_state, "repeat": a.repeat, "layouts": {}}
    for name, fp32, layout in (("wave_fp64", False, 0), ("wave_fp32_inverse", True, 0), ("block_fp64", False, 1)):
        times = []
        for _ in range(a.repeat):
            x, ok, ms = H.lu_solve(M, b, fp32=fp32, layout=layout, timing=True)
            times.append(ms)
        ms = float(np.median(times))
        out["layouts"][name] = {"kernel_ms": ms, "matrices_per_s": a.matrices / (ms / 1e3),
                                "all_ok": bool(ok.all())}
        print(f"[lu_bench] {name}: {ms:.3f} ms", flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
