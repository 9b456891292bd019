"""Pressure-dependent Arrhenius rates (PLOG, Cantera's PlogRate) in the oracle, and the
mechanism that needs them: ESTiMatE-Mech (the reference's
resources/mechanisms/ESTiMatE-Mech_04042022.xml, 214 species, 1,558 reactions, 22 PLOG
reactions, copied unchanged into hpc_ode_amd/mechanisms/).

What fixes the PLOG rate independently of the oracle's code: at a level's pressure the rate is
that level's Arrhenius expression; between two levels ln k is linear in ln P (so at the
geometric mean of two level pressures it is the mean of the two ln k); outside the range it is
the end level's rate. Checked on a copy of h2o2.yaml whose H + O2 <=> O + OH reaction is made a
three-level PLOG, through the forward rate the oracle reports, plus central differences of the
RHS against the analytic Jacobian (the d ln k / dT of the interpolated rate)."""
import os

import numpy as np
import pytest
import yaml

import oracle as O
from oracle.mech import _Loader
from hpc_ode_amd import workloads as WL

MECH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hpc_ode_amd", "mechanisms")
EST = os.path.join(MECH, "ESTiMatE-Mech_04042022.xml")
R = 8314.4621
ATM = 101325.0
LEVELS = [(0.1, 1.0e16, -0.5, 16000.0), (1.0, 2.65e16, -0.6707, 17041.0), (10.0, 9.0e16, -0.8, 18500.0)]


@pytest.fixture(scope="module")
def plog_h2(tmp_path_factory):
    with open(os.path.join(MECH, "h2o2.yaml")) as f:
        doc = yaml.load(f, Loader=_Loader)
    r = next(i for i, rd in enumerate(doc["reactions"]) if rd["equation"].startswith("H + O2 <=> O + OH"))
    doc["reactions"][r] = {"equation": "H + O2 <=> O + OH", "type": "pressure-dependent-Arrhenius",
                           "rate-constants": [{"P": f"{p} atm", "A": A, "b": b, "Ea": Ea} for p, A, b, Ea in LEVELS]}
    path = tmp_path_factory.mktemp("plog") / "h2o2_plog.yaml"
    path.write_text(yaml.safe_dump(doc, sort_keys=False))
    return O.Oracle(str(path)), r, str(path)


def _kf(o, r, P, y):
    rt = o.rates(P, y)
    C = rt["conc"]
    sp = o.mech.species
    return rt["fwd"][r] / (C[sp.index("H")] * C[sp.index("O2")])


def _level(i, T):
    # the file's units: cm, mol, cal/mol -> A * 1e-3 (bimolecular), theta = Ea * 4184 / R
    p, A, b, Ea = LEVELS[i]
    return A * 1e-3 * T ** b * np.exp(-Ea * 4184.0 / R / T)


def test_plog_rate_at_between_and_outside_levels(plog_h2):
    o, r, _ = plog_h2
    y, _, _ = WL.h2_air(o.mech.species, o.mech.W, 1, seed=3)
    y = y[0]
    y[1 + o.mech.species.index("H")] = 1e-3
    for T in (900.0, 1400.0, 2100.0):
        y[0] = T
        for i, (p, *_rest) in enumerate(LEVELS):
            assert _kf(o, r, p * ATM, y) == pytest.approx(_level(i, T), rel=1e-13)
        for i in range(2):
            pm = np.sqrt(LEVELS[i][0] * LEVELS[i + 1][0]) * ATM
            assert _kf(o, r, pm, y) == pytest.approx(np.sqrt(_level(i, T) * _level(i + 1, T)), rel=1e-13)
        w = np.log(4.0 / 1.0) / np.log(10.0 / 1.0)
        assert _kf(o, r, 4.0 * ATM, y) == pytest.approx(_level(1, T) ** (1 - w) * _level(2, T) ** w, rel=1e-13)
        assert _kf(o, r, 0.01 * ATM, y) == pytest.approx(_level(0, T), rel=1e-13)
        assert _kf(o, r, 100.0 * ATM, y) == pytest.approx(_level(2, T), rel=1e-13)


def _fd_err(o, P, y):
    n = len(y)
    J = o.jac(P, y)
    F = np.zeros((n, n))
    for c in range(n):
        h = 1e-4 * (y[0] if c == 0 else max(abs(y[c]), 1e-3))
        yp, ym = y.copy(), y.copy()
        yp[c] += h
        ym[c] -= h
        F[:, c] = (o.rhs(P, yp) - o.rhs(P, ym)) / (2 * h)
    scale = np.abs(F).max(axis=0) + np.abs(J).max(axis=0) + 1e-300
    return (np.abs(J - F) / scale).max()


@pytest.mark.parametrize("P_atm", [0.3, 1.0, 5.0])
def test_plog_jacobian_central_differences(plog_h2, P_atm):
    o, _, _ = plog_h2
    y, _, _ = WL.h2_air(o.mech.species, o.mech.W, 1, seed=7)
    y = y[0]
    y[0] = 1300.0
    for s in ("H", "O", "OH", "HO2"):
        y[1 + o.mech.species.index(s)] = 1e-3
    assert _fd_err(o, P_atm * ATM, y) < 1e-5


def test_plog_refused_at_constant_volume(plog_h2):
    _, _, path = plog_h2
    with pytest.raises(ValueError, match="constant pressure"):
        O.Oracle(path, reactor="CONV")


@pytest.fixture(scope="module")
def est():
    return O.Oracle(EST)


def test_estimate_tables(est):
    """214 species, 1,558 reactions, 22 of them PLOG with 106 pressure levels; up to five
    species on a side and half-integer product coefficients on irreversible lumped reactions."""
    t = est.t
    assert (est.K, est.R) == (214, 1558)
    assert int((np.diff(t.plog_ptr) > 0).sum()) == 22 and len(t.plog) == 106
    assert t.nprod.max() == 5
    half = t.prod_nu[t.prod_nu != np.floor(t.prod_nu)]
    assert len(half) == 11 and set(half.tolist()) == {0.5, 1.5}
    assert (t.rev[t.prod_nu.max(axis=1) != np.floor(t.prod_nu.max(axis=1))] == 0).all()


def test_estimate_plog_rate_at_a_level(est):
    """CH3 + OH <=> CH2(S) + H2O (reaction 0154 of the file): at 1 atm the rate is the file's
    1-atm level, A = 5.282e14 (SI-kmol), b = -1.518, E = 890.45 K."""
    sp = est.mech.species
    r = next(i for i, rx in enumerate(est.mech.reactions) if rx.equation.replace(" ", "") == "CH3+OH[=]CH2(S)+H2O")
    y, P, _ = WL.kerosene_air(sp, est.mech.W, 1, seed=1)
    y = y[0]
    y[1 + sp.index("CH3")] = 1e-4
    y[1 + sp.index("OH")] = 1e-4
    for T in (1000.0, 1700.0):
        y[0] = T
        rt = est.rates(ATM, y)
        C = rt["conc"]
        kf = rt["fwd"][r] / (C[sp.index("CH3")] * C[sp.index("OH")])
        assert kf == pytest.approx(5.282e14 * T ** -1.518 * np.exp(-890.45 / T), rel=1e-13)


def test_estimate_jacobian_and_conservation(est):
    import xml.etree.ElementTree as ET
    m = est.mech
    root = ET.parse(EST).getroot()
    comp = {}
    for s in root.iter("species"):
        a = s.find("atomArray")
        if s.get("name") in m.species and a is not None:
            comp[s.get("name")] = {t.split(":")[0]: float(t.split(":")[1]) for t in a.text.split()}
    els = sorted({e for c in comp.values() for e in c})
    E = np.array([[comp[s].get(e, 0.0) for e in els] for s in m.species])
    y, P, _ = WL.kerosene_air(m.species, m.W, 2, seed=4)
    rng = np.random.default_rng(2)
    for i in range(2):
        yy = y[i].copy()
        yy[0] = 1350.0 + 200 * i
        for s in ("H", "O", "OH", "HO2", "CH3", "H2O", "CO", "C2H4", "CH2O"):
            yy[1 + m.species.index(s)] = rng.uniform(1e-4, 1e-2)
        if min(abs(yy[0] - tm) for tm in set(m.tmid.tolist())) > 2.0:
            assert _fd_err(est, P[i], yy) < 1e-5
        w = est.rates(P[i], yy)["wdot"]
        assert np.abs(w @ E).max() <= 1e-10 * np.abs(w[:, None] * E).max()
