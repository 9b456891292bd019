"""The block-per-cell layout: one 256-thread workgroup per cell (hpo_opts::cell_layout = 2),
the Newton matrix inverted across the workgroup. It runs mechanisms whose state has more
variables than a wave has lanes, here nDodecane_Reitz (the reference's
resources/mechanisms/nDodecane_Reitz.yaml: 100 species, 553 reactions, reactions with four
products), and GRI-3.0 as the autotuner's alternative to one wave per cell.

Each result is judged at the stated bar (T and every mass fraction above 1e-10 within a
relative 1e-6 of the oracle) against the oracle's round-off floor on the same cells
(oracle/floor.py), plus element conservation (the oracle's Jacobian of nDodecane is pinned by
central differences in tests/test_oracle_chemistry.py)."""
import os

import numpy as np
import pytest
import yaml

import hpc_ode_amd as H
from hpc_ode_amd import workloads as WL

pytestmark = pytest.mark.gpu

DODECANE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hpc_ode_amd", "mechanisms",
                        "nDodecane_Reitz.yaml")


def _elements(path, species, W, y):
    from oracle import mech as OM
    with open(path) as f:
        doc = yaml.load(f, Loader=OM._Loader)
    comp = {s["name"]: s["composition"] for s in doc["species"]}
    els = sorted({e for c in comp.values() for e in c})
    E = np.array([[comp[sp].get(e, 0) for e in els] for sp in species], float) / np.asarray(W)[:, None]
    Yf = np.concatenate([y[:, 1:], 1.0 - y[:, 1:].sum(axis=1, keepdims=True)], axis=1)
    return Yf @ E


def _check(m, o, y0, P, dt, **kw):
    from oracle import floor as F
    fl = F.roundoff_floor(o, P, dt, y0, nthreads=8)
    assert (fl["status"] == 0).all()
    print(f"cells {len(y0)} {kw}", flush=True)
    yg, sg, stg = m.integrate(y0, P, dt, **kw)
    assert (sg == 0).all(), np.unique(sg, return_counts=True)
    print("floor", F.summary(fl), "gpu", F.compare(yg, sg, stg, fl), flush=True)
    F.check_against_floor(yg, sg, stg, fl)
    F.check_spread(yg, fl)
    return yg


def _states(mech, o, y0, P, dt):
    """Initial states, and the oracle's states after the substep (radicals, products)."""
    y1, s1, _ = o.integrate(P, dt, y0, nthreads=8)
    assert (s1 == 0).all()
    return np.concatenate([y0, y1]), np.concatenate([P, P])


def _component_check(m, o, y, P):
    """RHS: within 1e-9 of the oracle's value, 1e-13 of the component's term magnitude, or
    twice the spread an 8-ulp change of the eliminated species' mass fraction causes
    (tests/rhs_bars.py).
    Jacobian: within 1e-9 of the largest entry of its column (the wave kernel's bar)."""
    from rhs_bars import rhs_error
    f = m.rhs(y, P, cell_layout=2)
    assert np.all(np.isfinite(f))
    err, ref = rhs_error(o, P, y, f)
    i, k = np.unravel_index(err.argmax(), err.shape)
    print(f"RHS: worst {err.max():.3e} of the bar at state {i} component {k}: gpu {f[i, k]:.17g} oracle "
          f"{ref[i, k]:.17g} y {y[i, k]:.6e} T {y[i, 0]:.2f}", flush=True)
    rhs_ok = err.max() < 1.0
    if m.n_state < 64:
        fw = m.rhs(y, P, cell_layout=1)
        ew, _ = rhs_error(o, P, y, fw)
        print(f"RHS of the wave kernel on the same states: worst {ew.max():.3e} of the bar", flush=True)
    J = m.jac(y, P, cell_layout=2)
    for i in range(len(y)):
        Jr = o.jac(P[i], y[i])
        e = np.abs(J[i] - Jr) / (np.abs(Jr).max(axis=0, keepdims=True) + 1e-300)
        assert e.max() < 1e-9, f"state {i}: max column-scaled J error {e.max():.3e} at {np.unravel_index(e.argmax(), e.shape)}"
    assert rhs_ok, f"RHS error {err.max():.3e} of the bar"


def test_block_components_gri(gri_path):
    """The block kernel's RHS and Jacobian on GRI-3.0 against the oracle (same bar as the wave
    kernel's component tests: 1e-9 relative, column-scaled for J)."""
    import oracle as O
    m, o = H.Mechanism(gri_path), O.Oracle(gri_path)
    y0, P, dt = WL.gri_plain(m.species, m.W, 16, seed=5)
    _component_check(m, o, *_states(m, o, y0, P, dt))


def test_block_components_ndodecane():
    """nDodecane_Reitz (100 species, 553 reactions, up to four products): RHS and Jacobian of
    the block kernel against the oracle on fresh fuel/air mixtures and on their states after
    the substep."""
    import oracle as O
    m, o = H.Mechanism(DODECANE), O.Oracle(DODECANE)
    y0, P, dt = WL.dodecane_air(m.species, m.W, 16, seed=6)
    _component_check(m, o, *_states(m, o, y0, P, dt))


@pytest.mark.parametrize("ns", [53, 100])
@pytest.mark.parametrize("fp32", [False, True])
def test_block_gauss_jordan(ns, fp32):
    """The block kernel's Newton-matrix code (hpo_lu_solve_layout layout 2): Gauss-Jordan
    inverse across the workgroup, applied to b, against LAPACK (numpy.linalg.solve) on random
    dense matrices that need row exchanges; a singular matrix is reported."""
    rng = np.random.default_rng(ns + fp32)
    M = rng.standard_normal((8, ns, ns)) + 4.0 * np.eye(ns)[None] * rng.uniform(0, 1, (8, 1, 1))
    M[3] = np.triu(M[3])[::-1]
    M[5, :, 7] = 0.0
    b = rng.standard_normal((8, ns))
    x, ok = H.lu_solve(M, b, fp32=fp32, layout=2)
    assert ok.tolist() == [1, 1, 1, 1, 1, 0, 1, 1]
    eps = np.finfo(np.float32 if fp32 else np.float64).eps
    for i in range(8):
        if i == 5:
            continue
        ref = np.linalg.solve(M[i], b[i])
        err = np.abs(x[i] - ref).max() / np.abs(ref).max()
        assert err <= 50 * np.linalg.cond(M[i]) * eps, (i, err)


@pytest.mark.parametrize("lu_precision", [0, 1])
def test_block_layout_gri_matches_oracle(gri_path, lu_precision):
    import oracle as O
    m, o = H.Mechanism(gri_path), O.Oracle(gri_path)
    y0, P, dt = WL.gri_plain(m.species, m.W, 256, seed=11)
    _check(m, o, y0, P, dt, lu_precision=lu_precision, refine_sweeps=1, cell_layout=2)


@pytest.mark.parametrize("lu_precision", [0, 1])
def test_ndodecane_matches_oracle(lu_precision):
    import oracle as O
    m, o = H.Mechanism(DODECANE), O.Oracle(DODECANE)
    assert m.n_state == 100
    y0, P, dt = WL.dodecane_air(m.species, m.W, 256, seed=3)
    yg = _check(m, o, y0, P, dt, lu_precision=lu_precision, refine_sweeps=1)   # cell_layout 0 picks the block
    drift = np.abs(_elements(DODECANE, m.species, m.W, yg) - _elements(DODECANE, m.species, m.W, y0)).max()
    assert drift < 1e-11, drift


def test_wave_layout_refused_for_ndodecane():
    m = H.Mechanism(DODECANE)
    y0, P, dt = WL.dodecane_air(m.species, m.W, 4, seed=1)
    with pytest.raises(H.HpoError):
        m.integrate(y0, P, dt, cell_layout=1)
