"""Benchmark: GRI-3.0 cells integrated per second at dt = 1e-4 s on one MI355X.

Workload: 1,048,576 GRI-3.0 cells (T0 U[1200,2000] K, P U[1,20] atm, CH4/air phi U[0.6,1.6],
1e-8 radical seeding), one BDF substep of dt = 1e-4 s per cell, rtol 1e-6, atol 1e-10.
A "step" is one pass of the integrator over all cells. The measured line is the
FP32-LU (+FP64 residual refinement) configuration; the plain FP64 configuration runs on
the same cells and is reported beside it. Inputs are resident in HBM before timing starts.

    python bench.py [--steps K] [--warmup W] [--cells N] [--gpus 1]

Prints one JSON line (see the roofline / cpu_baseline objects for the method).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MECH = os.path.join(ROOT, "hpc_ode_amd", "mechanisms", "gri30.yaml")
PEAK_FP64_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=1 << 20)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--refine-sweeps", type=int, default=1)
    ap.add_argument("--waves-per-block", type=int, default=0)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--sort", type=int, default=1)
    ap.add_argument("--lu-layout", type=int, default=0, choices=(0, 1),
                    help="FP32 Newton matrix: 0 explicit inverse (Gauss-Jordan), 1 in-place LU")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--accuracy-cells", type=int, default=4096,
                    help="cells of the error statement against an rtol-1e-12 solution (0 = skip)")
    ap.add_argument("--profile", action="store_true", help="per-phase cycle counters (stderr)")
    ap.add_argument("--workload", choices=("plain", "uneven"), default="plain",
                    help="plain: T0 U[1200,2000] K, dt 1e-4 s; uneven: T0 U[300,2500] K, phi log-U[0.3,3], "
                         "dt log-U[1e-6,1e-3] s, launch shape and queue order from the autotuner")
    ap.add_argument("--autotune", action="store_true", help="autotune the plain workload too")
    ap.add_argument("--tune-sample", type=int, default=32768)
    ap.add_argument("--serve-seconds", type=float, default=16.0,
                    help="duration of the serving-loop measurement, split over the offered loads (0 = skip)")
    ap.add_argument("--cu-share", type=int, default=0,
                    help="serving: largest share of the device one launch may take, in percent (0 = 100)")
    ap.add_argument("--serve-loads", default="0.3,0.7",
                    help="offered loads as fractions of the measured serving capacity")
    ap.add_argument("--serve-substeps", type=int, default=2000,
                    help="serving: each load runs until it has offered at least this many substeps")
    ap.add_argument("--serve-slots", type=int, default=8, help="serving: launch slots (1..16)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def measured_traffic(workload, ncells, lu_precision):
    """HBM bytes per launch of the BDF kernel from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the same counters on a copy of known size), when
    that summary was taken on this workload; None otherwise."""
    path = os.path.join(ROOT, "profiles", "bdf_kernel_gri_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("workload_kind") != workload or rec.get("cells") != ncells or rec.get("lu_precision") != lu_precision:
        return None
    return rec.get("traffic_bytes")


def serving_measurement(H, gri, seconds, fractions, cu_share=0, min_substeps=2000, slots=8):
    """p50/p99 substep latency of the serving loop at offered loads stated as fractions of its
    measured capacity (requests: 70 % H2/air, 30 % GRI-3.0, log-uniform [1, 65536] cells,
    dt U[1e-5, 1e-4] s, two substeps per request stream, Poisson arrivals), each load offered
    for its share of `seconds` or until it has offered min_substeps substeps. The arrival rate
    is set from the mean size of the requests the load will draw (its first min_substeps / 2),
    so that the offered cells per second are the stated fraction of the capacity."""
    from hpc_ode_amd import serving as SV
    h2 = H.Mechanism(os.path.join(ROOT, "hpc_ode_amd", "mechanisms", "h2o2.yaml"))
    mechs = [h2, gri]
    srv = H.Server(mechs, n_slots=slots, cu_share_pct=cu_share)
    seed, substeps = 11, 2
    try:
        cap = SV.capacity(srv, mechs)
        drawn, _, _ = SV.request_mix(-(-min_substeps // substeps), seed + 1)
        mc = float(drawn.mean())
        loads = []
        for f in fractions:
            rate = f * cap / (mc * substeps)
            r = SV.run_load(srv, mechs, rate, seconds / len(fractions), seed=seed, substeps=substeps,
                            log=lambda m: print(m, file=sys.stderr, flush=True), min_substeps=min_substeps)
            r["fraction_of_capacity"] = f
            loads.append(r)
        return {"capacity_cells_per_s": cap, "mean_cells_per_request": mc, "slots": slots,
                "cu_share_pct": cu_share or 100,
                "mix": "70% H2/air (h2o2.yaml), 30% GRI-3.0; cells log-U[1,65536]; dt U[1e-5,1e-4] s; "
                       "2 substeps per stream with state resident between them; Poisson arrivals",
                "precision": "the server's default: FP32 Newton matrix, one FP64 refinement sweep, FP64 fallback",
                "loads": loads}
    finally:
        srv.close()


def workload_name(kind, n):
    cells = f"{n:,}"
    if kind == "uneven":
        return (f"GRI-3.0 {cells}-cell uneven substep (T0 U[300,2500] K, phi log-U[0.3,3], dt log-U[1e-6,1e-3] s), "
                "FP32-LU + FP64 refinement, autotuned persistent kernel")
    return f"GRI-3.0 {cells}-cell dt=1e-4 s substep, FP32-LU + FP64 refinement"


def _errors(y, truth):
    """Max relative error in T and in mass fractions above 1e-10 (and above 1e-6)."""
    eT = np.abs(y[:, 0] - truth[:, 0]) / truth[:, 0]
    Yt, Y = truth[:, 1:], y[:, 1:]
    rel = np.abs(Y - Yt) / np.maximum(np.abs(Yt), 1e-300)
    return {"T_max_rel": float(eT.max()), "T_p99_rel": float(np.quantile(eT, 0.99)),
            "Y_max_rel_above_1e-10": float(rel[np.abs(Yt) > 1e-10].max()),
            "Y_max_rel_above_1e-6": float(rel[np.abs(Yt) > 1e-6].max())}


def _agreement(y, ref, rtol=1e-6, atol=1e-10):
    """GPU against the CPU BDF at the same tolerances, cell by cell: the strict match criterion
    (relative rtol on T and on every mass fraction above 1e-10) and CVODE's own per-component
    weight (|dy| <= rtol |y| + atol)."""
    dT = np.abs(y[:, 0] - ref[:, 0]) / ref[:, 0]
    Yr, dY = ref[:, 1:], np.abs(y[:, 1:] - ref[:, 1:])
    strict = (dT <= rtol) & np.all((np.abs(Yr) <= 1e-10) | (dY <= rtol * np.abs(Yr)), axis=1)
    w = dY / (rtol * np.abs(Yr) + atol)
    weighted = (dT <= rtol) & np.all(w <= 1.0, axis=1)
    return {"T_max_rel": float(dT.max()), "T_median_rel": float(np.median(dT)),
            "cells_within_rtol_T_and_Y_above_1e-10": float(strict.mean()),
            "cells_within_rtol_y_plus_atol": float(weighted.mean()),
            "max_dy_over_rtol_y_plus_atol": float(w.max())}


def accuracy_statement(orc, y0, P, dt, y32, st32, nst32, y64, st64, nst64, ncells, nthreads):
    """On an evenly spaced subset of the workload:

    * errors of the GPU results and of the CPU BDF at the production tolerance against the same
      BDF at rtol 1e-12 / atol 1e-16;
    * the GPU against the CPU BDF at the production tolerance, cell by cell, beside the
      round-off floor of that comparison: the CPU BDF against itself when T0, P or Y_O2 moves
      by one ulp (oracle/floor.py). A result that is not bit-identical to the CPU's cannot
      match on more cells than the floor does."""
    from oracle import floor as F
    idx = np.linspace(0, len(y0) - 1, ncells).astype(np.int64)
    t0 = time.perf_counter()
    truth, st_t, _ = orc.integrate(P[idx], dt[idx], y0[idx], rtol=1e-12, atol=1e-16, maxsteps=10 ** 6,
                                   nthreads=nthreads)
    fl = F.roundoff_floor(orc, P[idx], dt[idx], y0[idx], nthreads=nthreads, perturbations=("T+", "P+", "O2+"))
    took = time.perf_counter() - t0
    y6, st6 = fl["y"], fl["status"]
    ok = st_t == 0
    out = {"cells": int(ncells), "truth": "CPU BDF (same algorithm) rtol 1e-12 atol 1e-16",
           "truth_failed": int((~ok).sum()), "cpu_seconds": took,
           "fp32lu_gpu": _errors(y32[idx][ok], truth[ok]),
           "cpu_rtol1e-6": _errors(y6[ok], truth[ok]),
           "roundoff_floor": F.summary(fl)}
    ok6 = st6 == 0
    out["fp32lu_gpu_vs_cpu_rtol1e-6"] = _agreement(y32[idx][ok6], y6[ok6])
    out["fp32lu_gpu_vs_cpu_rtol1e-6"].update(F.compare(y32[idx], st32[idx], {"nst": nst32[idx]}, fl))
    if y64 is not None:
        out["fp64_gpu"] = _errors(y64[idx][ok], truth[ok])
        out["fp64_gpu_vs_cpu_rtol1e-6"] = _agreement(y64[idx][ok6], y6[ok6])
        out["fp64_gpu_vs_cpu_rtol1e-6"].update(F.compare(y64[idx], st64[idx], {"nst": nst64[idx]}, fl))
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(orc, FL, counts, y0, P, dt, seconds, H2_PATH):
    """The oracle's OpenMP cell loop (the shape of CVodeIntegratorOMP, schedule(runtime) over
    cells) on this GPU's share of the host cores, in the same run as the GPU:

    * a uniform random sample of the measured workload for about `seconds` (cells/s is the
      headline figure; cells are independent, so the rate extrapolates linearly);
    * the same on one thread;
    * the plain GPU configuration (32,768 GRI cells, dt 1e-4 s, FP64) and the plain CPU
      configuration (100 H2/air cells, dt 1e-3 s) timed in full."""
    import oracle as O
    from hpc_ode_amd import workloads as WL
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(len(os.sched_getaffinity(0)), 16)
    omp = O.omp_info()
    n = len(y0)
    rng = np.random.default_rng(7)

    def sample_rate(nth, budget):
        t_cpu, done, sample = 0.0, 0, []
        chunk = 4 * nth
        while t_cpu < budget and done < n:
            idx = rng.choice(n, size=chunk, replace=False)
            tt = time.perf_counter()
            _, so, sto = orc.integrate(P[idx], dt[idx], y0[idx], nthreads=nth)
            t_cpu += time.perf_counter() - tt
            done += chunk
            sample.append(sto)
        return done, t_cpu, {k: np.concatenate([s[k] for s in sample]) for k in sample[0]}

    done, t_cpu, merged = sample_rate(nthreads, seconds)
    done1, t1, _ = sample_rate(1, max(seconds / 2, 1.0))
    full = {}
    yg, Pg, dtg = WL.gri_plain(orc.mech.species, np.array(orc.mech.W), 32768, seed=3)
    tt = time.perf_counter()
    _, sg, _ = orc.integrate(Pg, dtg, yg, nthreads=nthreads)
    tg = time.perf_counter() - tt
    full["gri_32768_dt1e-4"] = {"cells": 32768, "seconds": tg, "cells_per_s": 32768 / tg,
                                "failed": int((sg != 0).sum())}
    o2 = O.Oracle(H2_PATH)
    yh, Ph, dth = WL.h2_air(o2.mech.species, np.array(o2.mech.W), 100, seed=0)
    tt = time.perf_counter()
    _, sh, _ = o2.integrate(Ph, dth, yh, nthreads=nthreads)
    th = time.perf_counter() - tt
    full["h2_air_100_dt1e-3"] = {"cells": 100, "seconds": th, "cells_per_s": 100 / th,
                                 "failed": int((sh != 0).sum())}
    cpu = {"value": done / t_cpu, "unit": "cells/s", "cores": nthreads, "kind": "port",
           "sample": f"{done} cells drawn uniformly from the {n}-cell workload, "
                     f"OpenMP schedule(runtime) over cells, {t_cpu:.1f} s",
           "one_thread": {"value": done1 / t1, "cells": done1, "seconds": t1},
           "host": {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
                    "cpu_model": _cpu_model(), "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
                    "threads_used": nthreads, "schedule": omp["schedule"], "chunk": omp["chunk"],
                    "note": "threads = OMP_NUM_THREADS, this GPU's share of the host's cores"},
           "full_configs": full}
    return cpu, FL.mean_cell_flops(counts, merged), nthreads


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (cells shard with no "
                                   "exchange across GPUs, see the sharding note in the design document)"}))
        return 2
    import torch
    import hpc_ode_amd as H
    from hpc_ode_amd import flops as FL
    from hpc_ode_amd import workloads as WL

    mech = H.Mechanism(MECH)
    if a.workload == "uneven":
        y0, P, dt = WL.gri_uneven(mech.species, mech.W, a.cells, seed=a.seed)
    else:
        y0, P, dt = WL.gri_plain(mech.species, mech.W, a.cells, seed=a.seed)
    tune = a.autotune or a.workload == "uneven"
    n, ns = y0.shape
    dev = torch.device("cuda:0")
    d_y0 = torch.from_numpy(y0).to(dev)
    d_y = torch.empty_like(d_y0)
    d_P = torch.from_numpy(P).to(dev)
    d_dt = torch.from_numpy(dt).to(dev)
    d_st = torch.empty(n, dtype=torch.int32, device=dev)
    d_stats = torch.empty((n, H.NSTATS), dtype=torch.int32, device=dev)
    d_prof = torch.zeros((n, H.NPROF), dtype=torch.int64, device=dev) if a.profile else None
    d_order = None
    if a.sort and not tune:
        order = np.argsort(-y0[:, 0], kind="stable").astype(np.int32)
        d_order = torch.from_numpy(order).to(dev)
    stream = torch.cuda.current_stream(dev)
    if a.calibrate_copy > 0:
        nb = a.calibrate_copy << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] calibration copy of {nb} bytes done", file=sys.stderr, flush=True)
        del src, dst
    torch.cuda.synchronize()

    import ctypes as C

    def run(o, K, W):
        ker = []
        for i in range(W + K):
            if i == W:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
            t = H.Timing()
            rc = H.lib().hpo_integrate_device(
                mech.handle, n, C.c_void_p(d_y0.data_ptr()), C.c_void_p(d_y.data_ptr()), ns,
                C.c_void_p(d_P.data_ptr()), C.c_void_p(d_dt.data_ptr()), C.c_void_p(d_st.data_ptr()),
                C.c_void_p(d_stats.data_ptr()),
                C.c_void_p(d_order.data_ptr()) if d_order is not None else None,
                C.c_void_p(d_prof.data_ptr()) if d_prof is not None else None,
                C.byref(o), C.c_void_p(stream.cuda_stream), C.byref(t))
            if rc != 0:
                raise RuntimeError(H.lib().hpo_last_error().decode())
            if i >= W:
                ker.append(t.kernel_ms)
            print(f"[bench] {'warmup' if i < W else 'step'} {i} kernel {t.kernel_ms:.1f} ms grid {t.grid} "
                  f"block {t.block}", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        st = d_st.cpu().numpy()
        stats = d_stats.cpu().numpy()
        if d_prof is not None:
            pr = d_prof.cpu().numpy().astype(np.float64)
            tot = pr[:, 4].sum()
            parts = ", ".join(f"{f} {100 * pr[:, i].sum() / tot:.1f}%" for i, f in enumerate(H.PROF_FIELDS[:4]))
            ncall = stats.mean(axis=0)
            per = [pr[:, 0].sum() / stats[:, 1].sum(), pr[:, 1].sum() / max(stats[:, 2].sum(), 1),
                   pr[:, 2].sum() / max(stats[:, 3].sum(), 1), pr[:, 3].sum() / max(stats[:, 4].sum(), 1)]
            print(f"[bench] profile (lu_precision={o.lu_precision}): {parts}; cycles/cell {pr[:, 4].mean():.3g}; "
                  f"cycles per RHS {per[0]:.0f}, per Jacobian {per[1]:.0f}, per LU {per[2]:.0f}, per solve {per[3]:.0f}; "
                  f"mean nst {ncall[0]:.1f} nfe {ncall[1]:.1f} nje {ncall[2]:.2f} nsetups {ncall[3]:.1f} nni {ncall[4]:.1f}",
                  file=sys.stderr, flush=True)
        return wall / K * 1e3, float(np.mean(ker)), st, stats, t

    o32 = H.opts(lu_precision=1, refine_sweeps=a.refine_sweeps, waves_per_block=a.waves_per_block,
                 blocks_per_cu=a.blocks_per_cu, lu_layout=a.lu_layout)
    tuned = None
    if tune:
        d_order = torch.empty(n, dtype=torch.int32, device=dev)
        rep = H.TuneReport()
        rc = H.lib().hpo_autotune(mech.handle, n, C.c_void_p(d_y0.data_ptr()), ns, C.c_void_p(d_P.data_ptr()),
                                  C.c_void_p(d_dt.data_ptr()), a.tune_sample, C.byref(o32),
                                  C.c_void_p(d_order.data_ptr()), C.byref(rep), C.c_void_p(stream.cuda_stream))
        if rc != 0:
            raise RuntimeError(H.lib().hpo_last_error().decode())
        best = rep.candidates()[rep.best]
        tuned = {"sample_cells": rep.sample_cells, "tune_ms": rep.tune_ms, "chosen": best,
                 "candidates": rep.candidates()}
        print(f"[bench] autotune: {best} in {rep.tune_ms:.0f} ms", file=sys.stderr, flush=True)
    ms32, ker32, st32, stats32, t32 = run(o32, a.steps, a.warmup)
    y32 = d_y.cpu().numpy()
    fp64 = None
    y64 = None
    if not a.skip_fp64:
        # the FP64 leg runs the same cells with at most 3 timed steps and 1 warmup: it is reported
        # beside the measured line and its steps take 1.7x as long
        o64 = H.opts(lu_precision=0, waves_per_block=o32.waves_per_block, blocks_per_cu=o32.blocks_per_cu)
        k64, w64 = min(a.steps, 3), min(a.warmup, 1)
        ms64, ker64, st64, stats64, _ = run(o64, k64, w64)
        y64 = d_y.cpu().numpy()
        fp64 = dict(ms_per_step=ms64, kernel_ms=ker64, cells_per_s=n / (ms64 / 1e3),
                    failed=int((st64 != 0).sum()), steps=k64, warmup=w64)

    # ---- CPU baseline and algorithmic work: the oracle on a bounded sample ----
    cpu = None
    counts = None
    work = None
    import oracle as O
    orc = O.Oracle(MECH)
    counts = FL.mechanism_counts(orc.t)
    nthreads = min(len(os.sched_getaffinity(0)), 16)
    if not a.no_cpu:
        cpu, work, nthreads = cpu_baseline(orc, FL, counts, y0, P, dt, a.cpu_seconds,
                                           os.path.join(ROOT, "hpc_ode_amd", "mechanisms", "h2o2.yaml"))

    accuracy = None
    if a.accuracy_cells > 0 and not a.no_cpu:
        ncheck = min(a.accuracy_cells, n, 1024 if a.workload == "uneven" else n)
        accuracy = accuracy_statement(orc, y0, P, dt, y32, st32, stats32[:, 0], y64,
                                      st64 if y64 is not None else None,
                                      stats64[:, 0] if y64 is not None else None, ncheck, nthreads)

    serving = None
    if a.serve_seconds > 0:
        serving = serving_measurement(H, mech, a.serve_seconds, [float(x) for x in a.serve_loads.split(",")],
                                      a.cu_share, a.serve_substeps, a.serve_slots)

    value = n / (ms32 / 1e3)
    roof = None
    if work is not None:
        flop_launch = work * n
        achieved = flop_launch / (ker32 / 1e3) / 1e12
        roof = {"bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_TFLOPS, "traffic": measured_traffic(a.workload, n, 1),
                "flop_per_cell": work, "kernel_ms": ker32}
    out = {
        "metric": "GRI-3.0 cells integrated/sec at dt=1e-4 s, FP32-LU vs FP64 plain; p99 latency",
        "value": value, "unit": "cells/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms32, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (FP32 Newton-matrix LU)", "data": "synthetic",
        "config": {"workload": workload_name(a.workload, n),
                   "autotune": tuned,
                   "cells": n, "mechanism": "gri30.yaml (53 species, 325 reactions)",
                   "rtol": 1e-6, "atol": 1e-10, "refine_sweeps": a.refine_sweeps,
                   "lu_layout": H.LU_LAYOUTS[o32.lu_layout],
                   "grid": t32.grid, "block": t32.block, "lds_bytes": t32.lds_bytes},
        "failed_cells": int((st32 != 0).sum()),
        "fp64_plain": fp64,
        "speedup_fp32lu_vs_fp64": (fp64["ms_per_step"] / ms32) if fp64 else None,
        "roofline": roof,
        "cpu_baseline": cpu,
        "accuracy": accuracy,
        "serving": serving,
        "work_model": {k: counts[k] for k in ("F_rhs", "T_rhs", "F_jac", "F_LU", "F_solve")} if counts else None,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
