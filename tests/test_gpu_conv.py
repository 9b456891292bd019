"""The constant-volume reactor on the GPU (hpo_opts::reactor = 1, P = density), against the
oracle's constant-volume reactor (pinned by conservation of internal energy, central
differences and the constant-pressure rates in tests/test_oracle_conv.py): RHS and Jacobian of
both cell layouts, integrated cells judged at the stated bar against the round-off floor of the
oracle's constant-volume integration, internal energy kept by the GPU result, and the CSV
path (hpo_integrate_csv --reactor CONV) with ENTHA = h(T, Y) of the result."""
import os
import subprocess

import numpy as np
import pytest

import hpc_ode_amd as H
from hpc_ode_amd import workloads as WL

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
R = 8314.4621


@pytest.fixture(scope="module")
def gri(gri_path):
    import oracle as O
    return H.Mechanism(gri_path), O.Oracle(gri_path, reactor="CONV")


def _states(m, o, n, seed):
    """GRI cells at their constant-pressure start and after 1e-4 s at constant volume."""
    y0, P, dt = WL.gri_plain(m.species, m.W, n, seed=seed)
    rho = m.density(y0, P)
    y1, s1, _ = o.integrate(rho, dt, y0, nthreads=8)
    assert (s1 == 0).all()
    return np.concatenate([y0, y1]), np.concatenate([rho, rho])


@pytest.mark.parametrize("layout", [1, 2])
def test_conv_rhs_and_jacobian(gri, layout):
    """Same bars as the constant-pressure component tests (tests/rhs_bars.py; the Jacobian
    column-scaled to 1e-9)."""
    m, o = gri
    y, rho = _states(m, o, 16, 5)
    from rhs_bars import rhs_error
    f = m.rhs(y, rho, cell_layout=layout, reactor="CONV")
    assert np.all(np.isfinite(f))
    err, ref = rhs_error(o, rho, y, f)
    assert err.max() < 1.0, (err.max(), np.unravel_index(err.argmax(), err.shape))
    J = m.jac(y, rho, cell_layout=layout, reactor="CONV")
    for i in range(len(y)):
        Jr = o.jac(rho[i], y[i])
        e = np.abs(J[i] - Jr) / (np.abs(Jr).max(axis=0, keepdims=True) + 1e-300)
        assert e.max() < 1e-9, (i, e.max(), np.unravel_index(e.argmax(), e.shape))
    # the constant-pressure kernel on the same arguments is a different function
    fp = m.rhs(y[:4], rho[:4], cell_layout=layout)
    assert not np.allclose(fp[:, 0], f[:4, 0], rtol=1e-3)


def _u(o, y):
    W = np.asarray(o.mech.W)
    out = np.empty(len(y))
    scale = np.empty(len(y))
    for i, yy in enumerate(y):
        T = yy[0]
        a = np.where((T <= np.asarray(o.mech.tmid))[:, None], o.mech.nasa_lo, o.mech.nasa_hi)
        hRT = a[:, 0] + T * (a[:, 1] / 2 + T * (a[:, 2] / 3 + T * (a[:, 3] / 4 + T * a[:, 4] / 5))) + a[:, 5] / T
        Yf = np.append(yy[1:], 1.0 - yy[1:].sum())
        U = (hRT - 1.0) * R * T / W
        out[i], scale[i] = (Yf * U).sum(), np.abs(Yf * U).sum()
    return out, scale


@pytest.mark.parametrize("lu_precision,layout", [(0, 1), (1, 1), (1, 2)])
def test_conv_integration_matches_oracle(gri, lu_precision, layout):
    from oracle import floor as F
    m, o = gri
    y0, P, dt = WL.gri_plain(m.species, m.W, 256, seed=11)
    rho = m.density(y0, P)
    fl = F.roundoff_floor(o, rho, dt, y0, nthreads=8)
    assert (fl["status"] == 0).all()
    yg, sg, stg = m.integrate(y0, rho, dt, lu_precision=lu_precision, refine_sweeps=1, cell_layout=layout,
                              reactor="CONV")
    assert (sg == 0).all()
    print("floor", F.summary(fl), "gpu", F.compare(yg, sg, stg, fl), flush=True)
    F.check_against_floor(yg, sg, stg, fl)
    F.check_spread(yg, fl)
    u0, sc = _u(o, y0)
    u1, _ = _u(o, yg)
    assert np.all(np.abs(u1 - u0) <= 2e-6 * sc)


def test_conv_csv_path(tmp_path, h2li_path):
    """500 H2/air states of the reference's res_h2.csv through hpo_integrate_csv --reactor CONV
    at 101325 Pa for 1e-4 s: the oracle's constant-volume result at each cell's starting density,
    and ENTHA = h(T, Y) of the result (which moves: at constant volume u is kept, not h)."""
    import oracle as O
    exe = os.path.join(ROOT, "hpc_ode_amd", "bin", "hpo_integrate_csv")
    src = os.path.join(GOLD, "res_h2_sample.csv")
    raw, outp = tmp_path / "s.bin", tmp_path / "o.csv"
    r = subprocess.run([exe, "--input", src, "--output", str(outp), "--mechanism", h2li_path, "--dt", "1e-4",
                        "--pressure", "101325", "--reactor", "CONV", "--raw", str(raw), "--quiet"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    d = np.loadtxt(src, delimiter=",", skiprows=1)
    n, K = d.shape[0], 9
    buf = np.fromfile(str(raw), dtype=np.uint8)
    state = buf[:n * (K + 1) * 8].view(np.float64).reshape(n, K + 1)
    status = buf[n * (K + 1) * 8:].view(np.int32)
    assert (status == 0).all()
    y0 = np.concatenate([d[:, -1:], d[:, :K - 1]], axis=1)
    o = O.Oracle(h2li_path, reactor="CONV")
    rho = o.density(np.full(n, 101325.0), y0)
    yo, so, _ = o.integrate(rho, np.full(n, 1e-4), y0, nthreads=8)
    assert (so == 0).all()
    yg = state[:, :K]
    assert np.max(np.abs(yg[:, 0] - yo[:, 0]) / yo[:, 0]) < 1e-5
    assert np.max(np.abs(yg - yo) / (1e-6 * np.abs(yo) + 1e-10)) < 300.0
    out = np.loadtxt(str(outp), delimiter=",", skiprows=1)
    h = H.Mechanism(h2li_path).enthalpy_mass(state[:, 0], state[:, 1:K + 1])
    assert np.all(np.abs(out[:, K] - h) <= 6e-7 * np.abs(h))
    assert np.abs(out[:, K] - d[:, K]).max() > 1e3   # h moved by the p dv work
