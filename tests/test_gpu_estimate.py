"""ESTiMatE-Mech on the GPU (the reference's resources/mechanisms/ESTiMatE-Mech_04042022.xml: 214
species, 1,558 reactions, 22 PLOG rates, irreversible lumped reactions with five products and
half-integer coefficients): the block-per-cell kernel with its tables partly in global memory
and the Newton matrix in the workgroup's scratch.

* RHS and Jacobian against the oracle (pinned by the PLOG closed forms, central differences
  and element conservation in tests/test_oracle_plog.py) at pressures below, between and above
  the PLOG levels (0.01 to 100 atm), at the component tests' bars;
* integrated n-dodecane/air cells (the states ref_data/generate.py tabulates for this
  mechanism) at the stated bar against the oracle's round-off floor (oracle/floor.py), in both
  precisions and at several pressures, with element conservation;
* the constant-volume reactor is refused (the pressure of such a cell moves with its state)."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import hpc_ode_amd as H
from hpc_ode_amd import workloads as WL

pytestmark = pytest.mark.gpu

EST = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hpc_ode_amd", "mechanisms",
                   "ESTiMatE-Mech_04042022.xml")
ATM = 101325.0


@pytest.fixture(scope="module")
def est():
    import oracle as O
    m = H.Mechanism(EST)
    assert (m.n_species, m.n_reactions) == (214, 1558)
    return m, O.Oracle(EST)


def _elements(species, W, y):
    root = ET.parse(EST).getroot()
    comp = {}
    for s in root.iter("species"):
        a = s.find("atomArray")
        if a is not None:
            comp[s.get("name")] = {t.split(":")[0]: float(t.split(":")[1]) for t in a.text.split()}
    els = sorted({e for c in comp.values() for e in c})
    E = np.array([[comp[sp].get(e, 0.0) for e in els] for sp in species]) / np.asarray(W)[:, None]
    Yf = np.concatenate([y[:, 1:], 1.0 - y[:, 1:].sum(axis=1, keepdims=True)], axis=1)
    return Yf @ E


def test_estimate_components(est):
    """RHS within the component bar (1e-9 relative, 1e-13 of the terms' magnitude, the
    eliminated species' spread: tests/rhs_bars.py) and the Jacobian within 1e-9 of its
    column's largest entry, on fresh mixtures with radicals and on reacted states."""
    from rhs_bars import rhs_error
    m, o = est
    y, _, _ = WL.kerosene_air(m.species, m.W, 10, seed=8)
    rng = np.random.default_rng(3)
    for i in range(5):
        for s in ("H", "O", "OH", "HO2", "CH3", "H2O", "CO", "C2H4", "CH2O"):
            y[i, 1 + m.species.index(s)] = rng.uniform(1e-5, 1e-3)
    P = np.array([0.01, 0.03, 0.3, 1.0, 2.5, 7.0, 10.0, 30.0, 100.0, 1.0]) * ATM
    print("states: 10, pressures 0.01..100 atm", flush=True)
    f = m.rhs(y, P)
    assert np.all(np.isfinite(f))
    err, ref = rhs_error(o, P, y, f)
    i, k = np.unravel_index(err.argmax(), err.shape)
    print(f"RHS worst {err.max():.3e} of the bar at state {i} component {k}", flush=True)
    assert err.max() < 1.0
    J = m.jac(y[:4], P[:4])
    for i in range(4):
        Jr = o.jac(P[i], y[i])
        e = np.abs(J[i] - Jr) / (np.abs(Jr).max(axis=0, keepdims=True) + 1e-300)
        assert e.max() < 1e-9, (i, e.max(), np.unravel_index(e.argmax(), e.shape))


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
    drift = np.abs(_elements(m.species, m.W, yg) - _elements(m.species, m.W, y0)).max()
    assert drift < 1e-11, drift
    return yg


@pytest.mark.parametrize("lu_precision", [0, 1])
def test_estimate_matches_oracle(est, lu_precision):
    m, o = est
    y0, P, dt = WL.kerosene_air(m.species, m.W, 96, seed=5)
    _check(m, o, y0, P, dt, lu_precision=lu_precision, refine_sweeps=1)


def test_estimate_pressures(est):
    """The same mixtures at 0.05 to 60 atm: the PLOG rates at, between and beyond their levels."""
    m, o = est
    y0, _, dt = WL.kerosene_air(m.species, m.W, 64, seed=6)
    P = np.exp(np.random.default_rng(1).uniform(np.log(0.05), np.log(60.0), 64)) * ATM
    _check(m, o, y0, P, dt, lu_precision=1, refine_sweeps=1)


def test_estimate_refuses_constant_volume(est):
    m, _ = est
    y0, P, dt = WL.kerosene_air(m.species, m.W, 2, seed=1)
    with pytest.raises(H.HpoError, match="constant pressure"):
        m.integrate(y0, m.density(y0, P), dt, reactor="CONV")
    with pytest.raises(H.HpoError, match="constant pressure"):
        m.rhs(y0, m.density(y0, P), reactor="CONV")
