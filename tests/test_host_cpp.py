"""The C++ layer above the C ABI: Mesh / CSV reader / CSV writer / BDFIntegratorHIP, which
mirror the reference's Integrator interface (include/ODEIntegrator/Integrators/Integrator.hpp:22-40)
and its mesh file formats (src/InputOutput/Reader/csvReader.cpp, Writer/csvWriter.cpp).

CPU tests run the C++ unit tests (tests/cpp/test_host.cpp) and pin the writer's number format
against C++ iostreams. GPU tests run the reference's own 100-cell E2E case through the drop-in
integrator and compare with the reference's serial CVODE results file, and integrate H2/air
cells (the plain CPU configuration: 100 cells, 1e-3 s) through the CSV path against the oracle.
"""
import os
import subprocess
import textwrap

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hpc_ode_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")


def _exe(name):
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        pytest.fail(f"{p} is not built (python -m hpc_ode_amd.build)")
    return p


def _read_csv(path):
    with open(path) as f:
        head = f.readline().strip().split(",")
    data = np.loadtxt(path, delimiter=",", skiprows=1, ndmin=2)
    return head, data


def _read_raw(path, ncells, nsp):
    raw = np.fromfile(path, dtype=np.uint8)
    nd = ncells * (nsp + 1) * 8
    state = raw[:nd].view(np.float64).reshape(ncells, nsp + 1)
    status = raw[nd:nd + 4 * ncells].view(np.int32)
    return state, status


def test_cpp_host_unit_tests(tmp_path):
    r = subprocess.run([_exe("test_host"), "cpu", GOLD, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


def test_csv_writer_number_format_matches_iostreams(tmp_path):
    """The writer's text equals what std::scientific / std::defaultfloat streams print
    (csvWriter.cpp:61-85), checked against a C++ iostream program on the same values."""
    subprocess.run([_exe("test_host"), "cpu", GOLD, str(tmp_path)], check=True, capture_output=True, timeout=120)
    written = (tmp_path / "gri_10_results.csv").read_text().splitlines()
    vals = [float(x) for line in written[1:] for x in line.split(",")]
    vals += [0.0, -0.0, 1e-300, 123456789.0, 1523.4567, 300.0, 0.001204, -2.322e6]
    src = tmp_path / "fmt.cpp"
    src.write_text(textwrap.dedent("""
        #include <iostream>
        #include <cstdio>
        int main() { double v; while (std::scanf("%lf", &v) == 1)
          std::cout << std::scientific << v << ' ' << std::defaultfloat << v << '\\n'; }
    """))
    exe = tmp_path / "fmt"
    subprocess.run(["g++", "-O1", "-o", str(exe), str(src)], check=True, timeout=120)
    inp = "\n".join(repr(v) for v in vals)
    out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    for v, line in zip(vals, out):
        sci, dflt = line.split(" ")
        assert sci == "%e" % v and dflt == "%g" % v, (v, line)
    # and the written species columns are %e, the temperature column %g
    head = written[0].split(",")
    ti = head.index("TEMPE")
    for line in written[1:]:
        f = line.split(",")
        assert all(s == "%e" % float(s) for s in f[:ti])
        assert f[ti] == "%g" % float(f[ti])


@pytest.mark.gpu
def test_cpp_host_integrator_on_gpu(tmp_path):
    r = subprocess.run([_exe("test_host"), "gpu", GOLD, str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("lu", ["fp64", "fp32"])
def test_reference_e2e_100_cells(tmp_path, lu, gri_path):
    """E2Etests/100: the reference's serial CVODE output for tests/Integrators/data/res_gri_100.csv
    at dt = 1e-3 s, 101325.15 Pa (app/config.yaml:16-19). Its file holds 5 significant digits:
    T and every mass fraction above 1e-10 must agree to that precision (the cells are 300 K
    mixtures, so the printed state is unchanged by the step)."""
    out, raw, perf = tmp_path / "result.csv", tmp_path / "state.bin", tmp_path / "performance.yaml"
    r = subprocess.run([_exe("hpo_integrate_csv"), "--input", os.path.join(GOLD, "res_gri_100.csv"),
                        "--output", str(out), "--mechanism", gri_path, "--dt", "1e-3",
                        "--pressure", "101325.15", "--lu", lu, "--raw", str(raw), "--quiet",
                        "--performance-yaml", str(perf)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    # the reference application's performance.yaml keys (app/main.cpp:76-79), read as its runner does
    import yaml
    p = yaml.safe_load(perf.read_text())
    assert set(p) == {"readTime", "integrationTime", "writeTime", "totalTime"}
    assert abs(p["totalTime"] - (p["readTime"] + p["integrationTime"] + p["writeTime"])) < 1e-9
    _, ref = _read_csv(os.path.join(GOLD, "results_serial100.csv"))
    n = ref.shape[0]
    state, status = _read_raw(str(raw), n, 53)
    assert (status == 0).all()
    T_ref, Y_ref = ref[:, 54], ref[:, :53]
    assert np.all(np.abs(state[:, 0] - T_ref) <= 5e-5 * T_ref)
    Y = state[:, 1:]
    m = np.abs(Y_ref) > 1e-10
    assert np.all(np.abs(Y - Y_ref)[m] <= 5e-5 * np.abs(Y_ref)[m] + 1e-15)
    # the written CSV is the reference writer's format: 53 species, ENTHA, TEMPE, Points:0-2
    head, data = _read_csv(str(out))
    assert head[0] == "0CO0" and head[52] == "CO52" and head[53:] == ["ENTHA", "TEMPE", "Points:0", "Points:1",
                                                                     "Points:2"]
    assert data.shape == (n, 58)


@pytest.mark.gpu
def test_h2_cells_through_csv_path_match_oracle(tmp_path, h2_path):
    """The plain CPU configuration (H2/air, 100 cells, T0 U[1000,1500] K, phi U[0.5,2],
    101325 Pa, dt 1e-3 s) read from a CSV, integrated by BDFIntegratorHIP and compared with the
    oracle in full precision."""
    import oracle as O
    from hpc_ode_amd import workloads as WL
    orc = O.Oracle(h2_path)
    sp, W = orc.mech.species, orc.mech.W
    y0, P, dt = WL.h2_air(sp, W, 100, seed=21)
    K = len(sp)
    Y = np.concatenate([y0[:, 1:], 1.0 - y0[:, 1:].sum(axis=1, keepdims=True)], axis=1)
    inp = tmp_path / "h2.csv"
    with open(inp, "w") as f:
        f.write(",".join(f"CO{k}" for k in range(K)) + ",TEMPE\n")
        for i in range(len(y0)):
            f.write(",".join(repr(float(v)) for v in Y[i]) + f",{float(y0[i, 0])!r}\n")
    raw = tmp_path / "h2.bin"
    r = subprocess.run([_exe("hpo_integrate_csv"), "--input", str(inp), "--output", str(tmp_path / "o.csv"),
                        "--mechanism", h2_path, "--dt", "1e-3", "--pressure", "101325", "--raw", str(raw),
                        "--quiet"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    state, status = _read_raw(str(raw), len(y0), K)
    assert (status == 0).all()
    # the CSV carries the eliminated species; the state the solver sees is [T, Y_0..Y_{K-2}]
    y_in_csv = np.concatenate([y0[:, :1], Y[:, :-1]], axis=1)
    yo, so, _ = orc.integrate(P, dt, y_in_csv, nthreads=8)
    assert (so == 0).all()
    yg = state[:, :K]
    dT = np.abs(yg[:, 0] - yo[:, 0]) / yo[:, 0]
    assert dT.max() < 1e-5
    w = np.abs(yg - yo) / (1e-6 * np.abs(yo) + 1e-10)
    assert w.max() < 100.0
    # last species closes the mass balance, as CVodeIntegrator.cpp:90-96
    assert np.allclose(state[:, 1:].sum(axis=1), 1.0, rtol=0, atol=1e-13)


@pytest.mark.gpu
def test_reference_h2_states_with_xml_mechanism(tmp_path, h2li_path):
    """500 of the 10,000 hot H2/air states of the reference's ref_data/res_h2.csv (every 20th
    row), with their CTML mechanism h2_li_2004.xml (ref_data/README.md:15), read by the CSV
    reader, integrated for 1e-4 s at 101325 Pa through BDFIntegratorHIP and compared with the
    oracle in full precision."""
    import oracle as O
    src = os.path.join(GOLD, "res_h2_sample.csv")
    raw = tmp_path / "h2li.bin"
    r = subprocess.run([_exe("hpo_integrate_csv"), "--input", src, "--output", str(tmp_path / "o.csv"),
                        "--mechanism", h2li_path, "--dt", "1e-4", "--pressure", "101325", "--raw", str(raw),
                        "--quiet"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    _, d = _read_csv(src)
    n, K = d.shape[0], 9
    state, status = _read_raw(str(raw), n, K)
    assert (status == 0).all()
    y0 = np.concatenate([d[:, -1:], d[:, :K - 1]], axis=1)
    orc = O.Oracle(h2li_path)
    yo, so, _ = orc.integrate(np.full(n, 101325.0), np.full(n, 1e-4), y0, nthreads=8)
    assert (so == 0).all()
    yg = state[:, :K]
    assert np.max(np.abs(yg[:, 0] - yo[:, 0]) / yo[:, 0]) < 1e-5
    assert np.max(np.abs(yg - yo) / (1e-6 * np.abs(yo) + 1e-10)) < 100.0
    # the written file keeps the input's columns; ENTHA is h(T, Y) of the integrated state (the
    # file prints 7 significant digits), and since the reactor holds the pressure, that is the
    # input's Cantera enthalpy to the integration tolerance (T rises by up to 1,450 K here)
    import hpc_ode_amd as H
    head, out = _read_csv(str(tmp_path / "o.csv"))
    assert head[K:] == ["ENTHA", "TEMPE"]
    Yf = np.concatenate([state[:, 1:K], 1.0 - state[:, 1:K].sum(axis=1, keepdims=True)], axis=1)
    h = H.Mechanism(h2li_path).enthalpy_mass(state[:, 0], Yf)
    assert np.all(np.abs(out[:, K] - h) <= 6e-7 * np.abs(h))
    assert np.all(np.abs(out[:, K] - d[:, K]) <= 2e-6 * np.abs(d[:, K]))
