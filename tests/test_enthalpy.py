"""hpo_mech_enthalpy_mass, the ENTHA the host integrator writes, against the enthalpies the
reference's data files hold: ref_data/res_h2.csv (every 20th row, tests/golden/res_h2_sample.csv)
and res_gri_100.csv carry Cantera's enthalpy_mass at each row's (T, Y) (ref_data/generate.py
writes gas.enthalpy_mass beside gas.T and gas.Y). The files' gas constant is CODATA 2018's
8314.46261815 J/(kmol K); the library uses 8314.4621, a relative 6.2e-8 below it."""
import os

import numpy as np
import pytest

import hpc_ode_amd as H

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MECH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hpc_ode_amd", "mechanisms")


def _csv(name):
    head = open(os.path.join(GOLD, name)).readline().strip().split(",")
    d = np.loadtxt(os.path.join(GOLD, name), delimiter=",", skiprows=1, ndmin=2)
    return head, d


def test_enthalpy_matches_cantera_h2_states():
    """500 H2/air states at 1000-2100 K, written at full precision: agreement to the gas
    constant's ratio (the R difference alone is 6.2e-8)."""
    head, d = _csv("res_h2_sample.csv")
    m = H.Mechanism(os.path.join(MECH, "h2_li_2004.xml"))
    K = m.n_species
    assert head[K] == "ENTHA" and head[K + 1] == "TEMPE"
    h = m.enthalpy_mass(d[:, K + 1], d[:, :K])
    rel = np.abs(h - d[:, K]) / np.abs(d[:, K])
    assert rel.max() < 1e-7, rel.max()
    assert np.allclose(h * (8314.46261815324 / 8314.4621), d[:, K], rtol=2e-9, atol=0)


def test_enthalpy_matches_cantera_gri_states():
    """The GRI-3.0 triple-flame mixtures of res_gri_100.csv (printed to five significant
    digits)."""
    head, d = _csv("res_gri_100.csv")
    m = H.Mechanism(os.path.join(MECH, "gri30.yaml"))
    K = m.n_species
    ih = [i for i, c in enumerate(head) if "ENTHA" in c][0]
    it = [i for i, c in enumerate(head) if "TEMPE" in c][0]
    h = m.enthalpy_mass(d[:, it], d[:, :K])
    assert np.all(np.abs(h - d[:, ih]) <= 5e-5 * np.abs(d[:, ih]))


def test_enthalpy_closed_forms():
    """Pure species: h equals the NASA7 enthalpy over W; a mixture is the mass-weighted sum;
    at 298.15 K pure O2 and N2 (elements in their reference state) have h close to 0 (the
    polynomial fits leave about 50 J/kg)."""
    m = H.Mechanism(os.path.join(MECH, "gri30.yaml"))
    K = m.n_species
    t = m.tables()
    for name in ("O2", "N2"):
        Y = np.zeros((1, K))
        Y[0, m.species.index(name)] = 1.0
        assert abs(m.enthalpy_mass(np.array([298.15]), Y)[0]) < 200.0   # J/kg (fit residual), against 7e5 at 1000 K
    rng = np.random.default_rng(1)
    Y = rng.dirichlet(np.ones(K), 4)
    T = np.array([300.0, 999.0, 1001.0, 2500.0])
    h = m.enthalpy_mass(T, Y)
    for i in range(4):
        hk = np.empty(K)
        for k in range(K):
            a = t["nasa"][k, :7] if T[i] <= t["tmid"][k] else t["nasa"][k, 7:]
            x = T[i]
            hk[k] = (a[0] + x * (a[1] / 2 + x * (a[2] / 3 + x * (a[3] / 4 + x * a[4] / 5))) + a[5] / x) * 8314.4621 * x
        assert np.isclose(h[i], (Y[i] * hk / m.W).sum(), rtol=1e-13)


def test_enthalpy_argument_checks():
    m = H.Mechanism(os.path.join(MECH, "h2o2.yaml"))
    with pytest.raises(ValueError):
        m.enthalpy_mass(np.ones(2), np.ones((2, 3)))
    import ctypes as C
    assert H.lib().hpo_mech_enthalpy_mass(m.handle, 1, None, None, m.n_species, None) == -1
