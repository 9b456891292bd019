"""The GRI-3.0 1,048,576-cell substeps with the FP32 Newton matrix and one FP64 refinement
sweep (the measured configuration at dt = 1e-4 s, and the wide-spread one), checked at their
full size through properties that do not depend on it:

* every cell returns success;
* element conservation: the reactor is closed, so each cell's elemental mass fractions
  (C, H, O, N, Ar) stay at their initial values. The BDF corrections are combinations of
  right-hand sides and Newton steps, both orthogonal to the element vectors, so the drift is
  rounding only (the oracle's is below 4e-16);
* the cells integrated alone give bit-identical rows to the same cells inside the full batch
  (one wave per cell runs the same instructions whatever else is in the launch);
* 256 evenly spaced cells agree with the oracle at the stated bar (T and every mass fraction
  above 1e-10 within a relative 1e-6) as often as the oracle agrees with itself under one-ulp
  input changes (oracle/floor.py)."""
import numpy as np
import pytest
import yaml

import hpc_ode_amd as H
import oracle as O
from oracle import floor as F
from hpc_ode_amd import workloads as WL
from oracle import mech as OM

pytestmark = pytest.mark.gpu

N = 1 << 20


def _element_matrix(path, species, W):
    """Mass of element e per kilogram of species k per unit mass fraction: a_ek / W_k."""
    with open(path) as f:
        doc = yaml.load(f, Loader=OM._Loader)
    comp = {s["name"]: s["composition"] for s in doc["species"]}
    els = sorted({e for c in comp.values() for e in c})
    A = np.array([[comp[sp].get(e, 0) for e in els] for sp in species], float)
    return els, A / np.asarray(W)[:, None]


def _elements(y, E):
    Yf = np.concatenate([y[:, 1:], 1.0 - y[:, 1:].sum(axis=1, keepdims=True)], axis=1)
    return Yf @ E


@pytest.mark.parametrize("workload,n,lu_precision", [("plain", N, 1), ("uneven", N, 1), ("plain", 32768, 0)])
def test_full_size_substep_properties(gri_path, workload, n, lu_precision):
    """plain: T0 U[1200,2000] K, dt 1e-4 s. uneven: T0 U[300,2500] K, phi log-U[0.3,3], dt
    log-U[1e-6,1e-3] s, per-cell step counts spread over more than 10x. The 32,768-cell case is
    the plain GPU method: FP64 everywhere, the 53x53 Newton-matrix LU included (state
    [T, Y_0..Y_51])."""
    m = H.Mechanism(gri_path)
    gen = WL.gri_plain if workload == "plain" else WL.gri_uneven
    y0, P, dt = gen(m.species, m.W, n, seed=3)
    print(f"integrating {n} GRI cells ({workload}), lu_precision {lu_precision}", flush=True)
    y, st, stats = m.integrate(y0, P, dt, lu_precision=lu_precision, refine_sweeps=1)
    assert (st == 0).all(), np.unique(st, return_counts=True)
    assert np.isfinite(y).all()
    assert (y[:, 0] > 250.0).all() and (y[:, 0] < 4000.0).all()

    els, E = _element_matrix(gri_path, m.species, m.W)
    drift = np.abs(_elements(y, E) - _elements(y0, E)).max(axis=0)
    print("element drift max:", dict(zip(els, drift.tolist())), flush=True)
    # measured on MI355X: 2.2e-15 (plain) and 3.1e-13 (uneven, up to ~900 steps), both in H
    assert drift.max() < 1e-11

    k = 4096
    ya, sa, _ = m.integrate(y0[:k], P[:k], dt[:k], lu_precision=lu_precision, refine_sweeps=1)
    assert np.array_equal(ya, y[:k]) and np.array_equal(sa, st[:k])

    idx = np.linspace(0, n - 1, 256).astype(np.int64)
    fl = F.roundoff_floor(O.Oracle(gri_path), P[idx], dt[idx], y0[idx], nthreads=8)
    assert (fl["status"] == 0).all()
    yo = fl["y"]
    dT = np.abs(y[idx, 0] - yo[:, 0]) / yo[:, 0]
    w = np.abs(y[idx] - yo) / (1e-6 * np.abs(yo) + 1e-10)
    print(f"oracle sample: T max rel {dT.max():.2e}, median {np.median(dT):.2e}, max weighted {w.max():.1f}")
    print("floor", F.summary(fl), "gpu", F.compare(y[idx], st[idx], {k: v[idx] for k, v in stats.items()}, fl))
    F.check_against_floor(y[idx], st[idx], {k: v[idx] for k, v in stats.items()}, fl)
    assert dT.max() < 1e-5 and w.max() < 300.0
