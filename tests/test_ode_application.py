"""ODEApplication: the reference application's entry point (app/main.cpp:17-88) with the GPU
integrator. It takes a config.yaml with the reference's keys (app/config.yaml:1-23), creates the
timestamped execution folder under outFileService.outFolder, copies config.yaml there, and writes
the results CSV, performance.yaml and out.log (FileLogger) into it: the files the reference's
Python runner collects (app/main.py:107-130, 157-168).

The CPU test checks the refusals (a Cantera integrator type, a missing argument) and the folder
and log written before them; the GPU test runs the reference's own E2E case (res_gri_100.csv,
dt 1e-3 s, 101325.15 Pa) with its config keys and compares the CSV with the reference's serial
CVODE results file."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "hpc_ode_amd", "bin", "ODEApplication")
GOLD = os.path.join(ROOT, "tests", "golden")

CONFIG = """outFileService:
    outFolder: {out} # if null, default is ./out
reader:
    type: csvReader
    filename: {inp} # Only for file readers
writer:
    type: csvWriter
    filename: result.csv
integrator:
    type: {itype}
    omp:
        cpus: 10
        schedule:
            type: static
            chunk: 1
    mechanism: {mech}
    reltol: 1.0e-6
    abstol: 1.0e-10
    pressure: 101325.15
    dt: 1.0e-3
logger:
    type: FileLogger
    logLevel: DEBUG
"""


def _exe():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (python -m hpc_ode_amd.build)")
    return EXE


def _config(tmp_path, itype, gri_path):
    out = tmp_path / "out"
    cfg = tmp_path / "config.yaml"
    cfg.write_text(CONFIG.format(out=out, inp=os.path.join(GOLD, "res_gri_100.csv"), itype=itype, mech=gri_path))
    return cfg, out


def _exec_folder(out):
    dirs = [d for d in out.iterdir() if d.is_dir()]
    assert len(dirs) == 1
    return dirs[0]


def test_usage_and_refusals(tmp_path, gri_path):
    r = subprocess.run([_exe()], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "No argument provided" in r.stdout
    r = subprocess.run([_exe(), "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "<path to config.yaml>" in r.stdout
    cfg, out = _config(tmp_path, "CanteraIntegratorOMP", gri_path)
    r = subprocess.run([_exe(), str(cfg)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "Integrator type is not valid" in r.stderr
    ex = _exec_folder(out)
    import re
    assert re.fullmatch(r"\d{4}-\d\d-\d\d_\d\d:\d\d:\d\d", ex.name)      # %F_%T
    assert (ex / "config.yaml").read_text() == cfg.read_text()
    log = (ex / "out.log").read_text()
    assert "Starting ODEApplication..." in log and "Copying config.yaml to results directory" in log


@pytest.mark.gpu
@pytest.mark.parametrize("itype", ["CVodeIntegrator", "BDFIntegratorHIP"])
def test_reference_e2e_case(tmp_path, gri_path, itype):
    import yaml
    cfg, out = _config(tmp_path, itype, gri_path)
    r = subprocess.run([_exe(), str(cfg)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    ex = _exec_folder(out)
    assert sorted(p.name for p in ex.iterdir()) == ["config.yaml", "out.log", "performance.yaml", "result.csv"]
    p = yaml.safe_load((ex / "performance.yaml").read_text())
    assert set(p) == {"readTime", "integrationTime", "writeTime", "totalTime"}
    assert abs(p["totalTime"] - (p["readTime"] + p["integrationTime"] + p["writeTime"])) < 1e-9
    log = (ex / "out.log").read_text()
    assert "End of ODEApplication execution" in log
    with open(ex / "result.csv") as f:
        head = f.readline().strip().split(",")
    data = np.loadtxt(ex / "result.csv", delimiter=",", skiprows=1, ndmin=2)
    ref = np.loadtxt(os.path.join(GOLD, "results_serial100.csv"), delimiter=",", skiprows=1, ndmin=2)
    assert head[0] == "0CO0" and head[53:] == ["ENTHA", "TEMPE", "Points:0", "Points:1", "Points:2"]
    assert data.shape == ref.shape
    # the files print 6 significant digits: T, and every mass fraction above 1e-10, agree to that
    assert np.all(np.abs(data[:, 54] - ref[:, 54]) <= 5e-6 * ref[:, 54])
    m = np.abs(ref[:, :53]) > 1e-10
    assert np.all(np.abs(data[:, :53] - ref[:, :53])[m] <= 5e-6 * np.abs(ref[:, :53])[m])
    assert np.array_equal(data[:, 55:], ref[:, 55:])      # coordinates pass through
