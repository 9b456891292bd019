"""Pins the oracle's chemistry (RHS and Jacobian) against the reference's pyJac golden
vectors, central finite differences and conservation laws. CPU only."""
import os

import numpy as np
import pytest

import oracle as O
from hpc_ode_amd import workloads as WL

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "pyjac_gri30.npz")


@pytest.fixture(scope="module")
def gri(gri_path):
    return O.Oracle(gri_path)


@pytest.fixture(scope="module")
def h2(h2_path):
    return O.Oracle(h2_path)


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


def test_rhs_matches_pyjac_golden(gri, golden):
    """dydt of the table-driven oracle == pyJac's generated dydt (dydt.c:7-133) at 30 states
    spanning 300-3000 K, every NASA7 mid-point, 0.5-40 atm, zero and trace species."""
    for i in range(len(golden["y"])):
        f = gri.rhs(golden["P"][i], golden["y"][i])
        ref = golden["f"][i]
        scale = np.abs(ref).max()
        err = np.abs(f - ref) / (np.abs(ref) + 1e-9 * scale)
        assert err.max() < 1e-10, (i, err.max())


def test_jacobian_matches_pyjac_golden(gri, golden):
    """J of the oracle == pyJac's eval_jacob (jacob.c:3-371) except where pyJac is known to be
    inexact: (a) d(dT/dt)/dT, where pyJac deviates by up to 1e-4 relative while the oracle
    matches central differences (test_jacobian_central_differences); (b) at 300 K pyJac carries
    absolute noise of ~2e-12 in near-zero entries, so that state is compared norm-wise."""
    for j, i in enumerate(golden["jidx"]):
        P, y = golden["P"][i], golden["y"][i]
        J = gri.jac(P, y)
        ref = golden["J"][j]
        if y[0] < 500.0:
            assert np.abs(J - ref).max() <= 1e-7 * np.abs(ref).max()
            continue
        colscale = np.abs(ref).max(axis=0) + 1e-300
        err = np.abs(J - ref) / colscale
        err[0, 0] = 0.0
        assert err.max() < 1e-8, (i, err.max(), np.unravel_index(err.argmax(), err.shape))
        assert abs(J[0, 0] - ref[0, 0]) <= 1e-3 * abs(ref[0, 0])


def _fd_jac(o, P, y, rel=1e-4):
    n = len(y)
    J = np.zeros((n, n))
    for c in range(n):
        h = rel * (y[0] if c == 0 else max(abs(y[c]), 1e-3))
        yp, ym = y.copy(), y.copy()
        yp[c] += h
        ym[c] -= h
        J[:, c] = (o.rhs(P, yp) - o.rhs(P, ym)) / (2 * h)
    return J


def _check_fd(o, P, y, tol):
    J = o.jac(P, y)
    F = _fd_jac(o, P, y)
    colscale = np.abs(F).max(axis=0) + np.abs(J).max(axis=0) + 1e-300
    err = np.abs(J - F) / colscale
    return err.max()


@pytest.mark.parametrize("T", [900.0, 1250.0, 1700.0, 2300.0])
def test_jacobian_central_differences_gri(gri, T):
    rng = np.random.default_rng(int(T))
    y, P, _ = WL.gri_plain(gri.mech.species, gri.mech.W, 1, seed=int(T))
    y = y[0]
    y[0] = T
    for s in ("H", "O", "OH", "HO2", "CH3", "H2O", "CO"):
        y[1 + gri.mech.species.index(s)] = rng.uniform(1e-4, 1e-2)
    assert _check_fd(gri, P[0], y, 1e-5) < 1e-5


def test_jacobian_central_differences_h2(h2):
    y, P, _ = WL.h2_air(h2.mech.species, h2.mech.W, 3, seed=5)
    for i in range(3):
        yy = y[i].copy()
        for s in ("H", "O", "OH", "HO2"):
            yy[1 + h2.mech.species.index(s)] = 1e-3
        assert _check_fd(h2, P[i], yy, 1e-5) < 1e-5


def test_fd_check_detects_planted_jacobian_mistakes(gri):
    """The finite-difference check above has teeth: a transposed Jacobian or one column with a
    flipped sign fails it."""
    y, P, _ = WL.gri_plain(gri.mech.species, gri.mech.W, 1, seed=9)
    y = y[0]
    y[0] = 1700.0
    for s in ("H", "O", "OH"):
        y[1 + gri.mech.species.index(s)] = 1e-3
    J = gri.jac(P[0], y)
    F = _fd_jac(gri, P[0], y)
    cs = np.abs(F).max(axis=0) + np.abs(J).max(axis=0) + 1e-300
    assert (np.abs(J - F) / cs).max() < 1e-5
    assert (np.abs(J.T - F) / cs).max() > 1e-2
    Jb = J.copy()
    Jb[:, 4] *= -1
    assert (np.abs(Jb - F) / cs).max() > 1e-2


def test_mass_and_element_conservation(gri):
    """sum_k W_k wdot_k = 0 and every element is conserved by the species production rates."""
    m = gri.mech
    import yaml
    from oracle.mech import _Loader
    with open(os.path.join(os.path.dirname(os.path.dirname(__file__)), "hpc_ode_amd", "mechanisms",
                           "gri30.yaml")) as f:
        doc = yaml.load(f, Loader=_Loader)
    comp = {s["name"]: s["composition"] for s in doc["species"]}
    elements = ["O", "H", "C", "N", "Ar"]
    E = np.array([[comp[s].get(e, 0) for e in elements] for s in m.species], float)
    y, P, _ = WL.gri_plain(m.species, m.W, 4, seed=3)
    for i in range(4):
        yy = y[i].copy()
        for s in ("H", "O", "OH", "HO2", "CH3", "CO", "NO", "HCN"):
            yy[1 + m.species.index(s)] = 1e-3
        r = gri.rates(P[i], yy)
        w = r["wdot"]
        scale = np.abs(w * m.W).max()
        assert abs((w * m.W).sum()) < 1e-11 * scale
        assert np.all(np.abs(E.T @ w) < 1e-11 * np.abs(w).max() * 10)
        assert r["wdot"][m.species.index("AR")] == 0.0   # spec_rates.c:3110


def test_detailed_balance_and_closed_form_rate(gri):
    """k_f by hand for O + H2 <=> H + OH (A 3.87e4 cm3/mol/s, b 2.7, Ea 6260 cal/mol), and
    fwd/rev ratio equal to Kc from NASA7 Gibbs energies computed independently here."""
    m = gri.mech
    T, P = 1500.0, 101325.0
    y, _, _ = WL.gri_plain(m.species, m.W, 1, seed=1)
    y = y[0]
    y[0] = T
    for s_ in ("H2", "H", "O", "OH"):
        y[1 + m.species.index(s_)] = 1e-3
    r = gri.rates(P, y)
    idx = [i for i, rx in enumerate(m.reactions) if rx.equation.startswith("O + H2 <=> H + OH")][0]
    C = r["conc"]
    kf = 3.87e4 * 1e-3 * T ** 2.7 * np.exp(-6260.0 * 4184.0 / 8314.4621 / T)
    sp = m.species
    assert np.isclose(r["fwd"][idx], kf * C[sp.index("O")] * C[sp.index("H2")], rtol=1e-12)

    def g_RT(name):
        k = sp.index(name)
        a = m.nasa_lo[k] if T <= m.tmid[k] else m.nasa_hi[k]
        h = a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + a[5] / T
        s = a[0] * np.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4 + a[6]
        return h - s
    Kc = np.exp(-(g_RT("H") + g_RT("OH") - g_RT("O") - g_RT("H2")))   # dnu = 0
    ratio = (r["fwd"][idx] / (C[sp.index("O")] * C[sp.index("H2")])) / (
        r["rev"][idx] / (C[sp.index("H")] * C[sp.index("OH")]))
    assert np.isclose(ratio, Kc, rtol=1e-12)


def test_h2o2_reactions_identical_to_gri(gri, h2):
    """h2o2.yaml is extracted from GRI-Mech 3.0: shared reactions have the same rate parameters."""
    def key(rx, sp):
        return (tuple((sp[k], c) for k, c in rx.reactants), tuple((sp[k], c) for k, c in rx.products))
    g = {}
    for rx in gri.mech.reactions:   # duplicate reactions share a key
        g.setdefault(key(rx, gri.mech.species), []).append((rx.A, rx.b, rx.theta))
    shared = 0
    for rx in h2.mech.reactions:
        k = key(rx, h2.mech.species)
        if k in g and not rx.efficiencies:
            assert (rx.A, rx.b, rx.theta) in g[k], rx.equation
            shared += 1
    assert shared >= 15


def test_equilibrium_has_zero_temperature_rate(gri):
    """Integrated far to equilibrium, dT/dt and every wdot vanish relative to their scale."""
    y, P, _ = WL.gri_plain(gri.mech.species, gri.mech.W, 1, seed=4)
    y0 = y[0].copy()
    y0[0] = 2000.0
    ye, st, _ = gri.integrate(P[0], 1.0, y0[None], rtol=1e-10, atol=1e-16, maxsteps=200000)
    assert st[0] == 0
    f = gri.rhs(P[0], ye[0])
    f0 = gri.rhs(P[0], y0)
    assert abs(f[0]) < 1e-6 * abs(f0[0])


def test_h2li_xml_mechanism_jacobian_and_conservation(h2li_path):
    """The CTML mechanism of the reference's res_h2.csv (h2_li_2004.xml, 9 species, 21
    reactions): the analytic Jacobian matches central differences on the reference's own
    states, and production rates conserve mass."""
    o = O.Oracle(h2li_path)
    d = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "res_h2_sample.csv"), delimiter=",",
                   skiprows=1)
    tmid = set(o.mech.tmid.tolist())
    for row in d[::50]:
        y = np.concatenate([[row[-1]], row[:8]])          # [T, Y_0..Y_7]; N2 (last) eliminated
        if min(abs(y[0] - t) for t in tmid) < 1.0:
            continue   # central differences would straddle the NASA7 range switch
        yy = y.copy()
        for s in ("H", "O", "OH", "HO2"):
            yy[1 + o.mech.species.index(s)] = max(yy[1 + o.mech.species.index(s)], 1e-4)
        assert _check_fd(o, 101325.0, yy, 1e-5) < 1e-5
        w = o.rates(101325.0, y)["wdot"]
        assert abs((w * o.mech.W).sum()) <= 1e-11 * np.abs(w * o.mech.W).max() + 1e-300


@pytest.fixture(scope="module")
def dodecane():
    return O.Oracle(os.path.join(os.path.dirname(os.path.dirname(__file__)), "hpc_ode_amd", "mechanisms",
                                 "nDodecane_Reitz.yaml"))


def test_ndodecane_tables(dodecane):
    """nDodecane_Reitz (the reference's resources/mechanisms, 100 species, 553 reactions):
    reactions with four product species and the eliminated last species A4."""
    t = dodecane.t
    assert (dodecane.K, dodecane.R) == (100, 553)
    assert t.nprod.max() == 4 and t.nreac.max() <= 4
    assert dodecane.mech.species[-1] == "A4"


@pytest.mark.parametrize("T", [800.0, 1100.0, 1500.0, 2200.0])
def test_ndodecane_jacobian_central_differences(dodecane, T):
    """The analytic Jacobian of the 100-species mechanism (what the block-per-cell kernel
    reproduces) matches central differences of the RHS."""
    m = dodecane.mech
    rng = np.random.default_rng(int(T))
    y, P, _ = WL.dodecane_air(m.species, m.W, 1, seed=int(T))
    y = y[0]
    y[0] = T
    for s in ("h", "o", "oh", "ho2", "ch3", "h2o", "co", "c2h4", "ch2o"):
        y[1 + m.species.index(s)] = rng.uniform(1e-4, 1e-2)
    if min(abs(T - tm) for tm in set(m.tmid.tolist())) < 2.0:
        pytest.skip("central differences would straddle a NASA7 range switch")
    assert _check_fd(dodecane, P[0], y, 1e-5) < 1e-5


def test_ndodecane_mass_and_element_conservation(dodecane):
    m = dodecane.mech
    import yaml
    from oracle.mech import _Loader
    with open(os.path.join(os.path.dirname(os.path.dirname(__file__)), "hpc_ode_amd", "mechanisms",
                           "nDodecane_Reitz.yaml")) as f:
        doc = yaml.load(f, Loader=_Loader)
    comp = {s["name"]: s["composition"] for s in doc["species"]}
    elements = sorted({e for c in comp.values() for e in c})
    E = np.array([[comp[s].get(e, 0) for e in elements] for s in m.species], float)
    y, P, _ = WL.dodecane_air(m.species, m.W, 3, seed=5)
    for i in range(3):
        yy = y[i].copy()
        yy[0] = 1400.0
        for s in ("h", "o", "oh", "ho2", "ch3", "co", "c2h4"):
            yy[1 + m.species.index(s)] = 1e-3
        w = dodecane.rates(P[i], yy)["wdot"]
        scale = np.abs(w * m.W).max()
        assert abs((w * m.W).sum()) < 1e-11 * scale
        assert np.all(np.abs(E.T @ w) < 1e-11 * np.abs(w).max() * 10)
