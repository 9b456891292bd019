"""Seeded synthetic cell states for the benchmark configurations.

Each generator returns (y, P, dt): y is [ncells, n_state] with rows
y = [T, Y_0 .. Y_{K-2}] (the layout CVodeIntegrator builds from the Mesh,
src/Integrators/CVodeIntegrator.cpp:64-77), P in Pa and dt in s per cell.
Mixtures are drawn in mole fractions and converted with the mechanism's
molecular weights.
"""
from __future__ import annotations

import numpy as np

ATM = 101325.0


def _to_state(X: np.ndarray, W: np.ndarray, T: np.ndarray) -> np.ndarray:
    Y = X * W[None, :]
    Y /= Y.sum(axis=1, keepdims=True)
    return np.concatenate([T[:, None], Y[:, :-1]], axis=1)


def h2_air(species, W, n, seed=0, T_range=(1000.0, 1500.0), phi_range=(0.5, 2.0), dt=1e-3, P=ATM):
    """H2/air cells: T0 ~ U(T_range), X_H2 : X_O2 : X_N2 = 2 phi : 1 : 3.76, phi ~ U(phi_range)."""
    rng = np.random.default_rng(seed)
    K = len(species)
    T = rng.uniform(*T_range, n)
    phi = rng.uniform(*phi_range, n)
    X = np.zeros((n, K))
    X[:, species.index("H2")] = 2.0 * phi
    X[:, species.index("O2")] = 1.0
    X[:, species.index("N2")] = 3.76
    X /= X.sum(axis=1, keepdims=True)
    y = _to_state(X, np.asarray(W), T)
    return y, np.full(n, P), np.full(n, dt)


def ch4_air(species, W, n, seed=0, T_range=(1200.0, 2000.0), P_atm_range=(1.0, 20.0),
            phi_range=(0.6, 1.6), phi_log=False, dt=1e-4, dt_log_range=None, radical_seed=1e-8):
    """CH4/air cells for GRI-3.0: X_CH4 : X_O2 : X_N2 = phi : 2 : 7.52, then mass fractions
    of OH, H and O set to radical_seed and the composition renormalised."""
    rng = np.random.default_rng(seed)
    K = len(species)
    W = np.asarray(W)
    T = rng.uniform(*T_range, n)
    P = rng.uniform(*P_atm_range, n) * ATM
    if phi_log:
        phi = np.exp(rng.uniform(np.log(phi_range[0]), np.log(phi_range[1]), n))
    else:
        phi = rng.uniform(*phi_range, n)
    X = np.zeros((n, K))
    X[:, species.index("CH4")] = phi
    X[:, species.index("O2")] = 2.0
    X[:, species.index("N2")] = 7.52
    X /= X.sum(axis=1, keepdims=True)
    Y = X * W[None, :]
    Y /= Y.sum(axis=1, keepdims=True)
    if radical_seed:
        for s in ("OH", "H", "O"):
            Y[:, species.index(s)] = radical_seed
        Y /= Y.sum(axis=1, keepdims=True)
    y = np.concatenate([T[:, None], Y[:, :-1]], axis=1)
    if dt_log_range is not None:
        dtv = np.exp(rng.uniform(np.log(dt_log_range[0]), np.log(dt_log_range[1]), n))
    else:
        dtv = np.full(n, dt)
    return y, P, dtv


def gri_plain(species, W, n, seed=0):
    """GRI-3.0 plain-method cells: T0 U[1200,2000] K, P U[1,20] atm, phi U[0.6,1.6], dt 1e-4 s."""
    return ch4_air(species, W, n, seed=seed)


def gri_uneven(species, W, n, seed=0):
    """Wide stiffness spread: T0 U[300,2500] K, phi log-U[0.3,3], dt log-U[1e-6,1e-3] s."""
    return ch4_air(species, W, n, seed=seed, T_range=(300.0, 2500.0), phi_range=(0.3, 3.0),
                   phi_log=True, dt_log_range=(1e-6, 1e-3))


def dodecane_air(species, W, n, seed=0, T_range=(1000.0, 1600.0), P_atm_range=(1.0, 20.0),
                 phi_range=(0.6, 1.6), dt=1e-4):
    """n-dodecane/air cells for nDodecane_Reitz (100 species): X_C12H26 : X_O2 : X_N2 =
    phi : 18.5 : 18.5 * 3.76, T0 ~ U(T_range), P ~ U(P_atm_range) atm, phi ~ U(phi_range)."""
    rng = np.random.default_rng(seed)
    K = len(species)
    low = [s.lower() for s in species]
    W = np.asarray(W)
    T = rng.uniform(*T_range, n)
    P = rng.uniform(*P_atm_range, n) * ATM
    phi = rng.uniform(*phi_range, n)
    X = np.zeros((n, K))
    X[:, low.index("c12h26")] = phi
    X[:, low.index("o2")] = 18.5
    X[:, low.index("n2")] = 18.5 * 3.76
    X /= X.sum(axis=1, keepdims=True)
    y = _to_state(X, W, T)
    return y, P, np.full(n, dt)


def kerosene_air(species, W, n, seed=0, T_range=(1000.0, 2100.0), phi_range=(0.5, 3.5), P_atm=1.0, dt=1e-4):
    """n-dodecane/air cells for ESTiMatE-Mech (214 species), the states the reference's
    ref_data/generate.py tabulates for that mechanism: fuel N-C12H26 in air (O2:N2 = 1:3.76),
    X_fuel : X_O2 = phi : 18.5, T0 ~ U(T_range), phi ~ U(phi_range), at P_atm atmospheres."""
    rng = np.random.default_rng(seed)
    K = len(species)
    W = np.asarray(W)
    T = rng.uniform(*T_range, n)
    phi = rng.uniform(*phi_range, n)
    X = np.zeros((n, K))
    X[:, species.index("N-C12H26")] = phi
    X[:, species.index("O2")] = 18.5
    X[:, species.index("N2")] = 18.5 * 3.76
    X /= X.sum(axis=1, keepdims=True)
    y = _to_state(X, W, T)
    return y, np.full(n, P_atm * ATM), np.full(n, dt)
