"""CPU checks of bench.py's measurement helpers: the one-GPU refusal, the HBM-traffic lookup
behind roofline.traffic, and the accuracy/agreement statistics on inputs with known answers."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_traffic_comes_from_the_profile_of_the_same_workload_only():
    with open(os.path.join(ROOT, "profiles", "bdf_kernel_gri_traffic.json")) as f:
        rec = json.load(f)
    t = bench.measured_traffic(rec["workload_kind"], rec["cells"], rec["lu_precision"])
    assert t == rec["traffic_bytes"]
    # the counters scale each read/write by what a copy of known size reads/writes
    fetch_scale = rec["copy_bytes"] / (rec["copy_fetch_kib"] * 1024)
    write_scale = rec["copy_bytes"] / (rec["copy_write_kib"] * 1024)
    expect = rec["fetch_kib"] * 1024 * fetch_scale + rec["write_kib"] * 1024 * write_scale
    assert abs(t - expect) <= 1e-6 * expect
    assert bench.measured_traffic(rec["workload_kind"], rec["cells"] // 2, rec["lu_precision"]) is None
    assert bench.measured_traffic("uneven" if rec["workload_kind"] == "plain" else "plain", rec["cells"],
                                  rec["lu_precision"]) is None


def test_error_and_agreement_statistics():
    rng = np.random.default_rng(0)
    truth = np.concatenate([rng.uniform(1000, 2000, (50, 1)), rng.uniform(0, 0.2, (50, 9))], axis=1)
    truth[:, 5] = 1e-12                          # below the 1e-10 floor: not counted
    truth[3, 2] = 0.1
    e = bench._errors(truth.copy(), truth)
    assert e["T_max_rel"] == 0 and e["Y_max_rel_above_1e-10"] == 0
    y = truth.copy()
    y[7, 0] *= 1 + 3e-6                          # one cell off by 3e-6 in T
    y[3, 2] *= 1 + 2e-6                          # one cell off by 2e-6 in a species
    y[:, 5] = 2e-12                              # tiny species: 100 % off, ignored by the strict test
    e = bench._errors(y, truth)
    assert abs(e["T_max_rel"] - 3e-6) < 1e-12
    assert abs(e["Y_max_rel_above_1e-10"] - 2e-6) < 1e-12
    a = bench._agreement(y, truth)
    assert a["cells_within_rtol_T_and_Y_above_1e-10"] == 48 / 50
    # CVODE's weight: |dY| <= 1e-6 |Y| + 1e-10 fails for the species cell, passes for 1e-12 values
    assert a["cells_within_rtol_y_plus_atol"] == 48 / 50
    assert a["max_dy_over_rtol_y_plus_atol"] > 1.0
