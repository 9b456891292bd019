"""Writes tests/golden/pyjac_gri30.npz: the reference's own pyJac GRI-3.0 RHS and Jacobian
(src/Mechanism/CPU/dydt.c:7-133 and jacob.c:3-371) evaluated at fixed states.

The library is oracle/_ref/libpyjac_gri30.so, compiled from the reference's C sources where
they lie by oracle/Makefile ('make -C oracle ref'). Run from the repository root:

    python tests/golden/make_pyjac_golden.py

States (y = [T, Y_0..Y_51], P in Pa): lean/rich CH4-air mixtures with and without radical
pools, temperatures on both sides of every NASA7 mid-point (1000, 1368, 1382, 1478 K and the
points themselves), cold (300 K) and very hot (3000 K) states, species at exactly zero, and
pressures from 0.5 to 40 atm.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from oracle import mech as M  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "pyjac_gri30.npz")
MECH = os.path.join(ROOT, "hpc_ode_amd", "mechanisms", "gri30.yaml")


def states():
    m = M.load(MECH)
    sp, W = m.species, m.W
    rng = np.random.default_rng(20240601)
    Ts = [300.0, 800.0, 999.0, 1000.0, 1001.0, 1200.0, 1368.0, 1375.0, 1382.0, 1478.0, 1479.0,
          1600.0, 2000.0, 2400.0, 3000.0]
    rows, Ps = [], []
    for i, T in enumerate(Ts):
        for variant in range(2):
            X = np.zeros(len(sp))
            phi = rng.uniform(0.4, 2.5)
            X[sp.index("CH4")] = phi
            X[sp.index("O2")] = 2.0
            X[sp.index("N2")] = 7.52
            X[sp.index("AR")] = 0.09
            if variant == 1:
                # radical pool and products; every species present at a random level
                X += 10.0 ** rng.uniform(-9, -3, len(sp))
                for s in ("H2O", "CO2", "CO", "OH", "H", "O", "HO2", "CH3", "CH2O", "NO"):
                    X[sp.index(s)] += 10.0 ** rng.uniform(-3, -0.5)
            Y = X * W
            Y /= Y.sum()
            rows.append(np.concatenate([[T], Y[:-1]]))
            Ps.append(101325.0 * (0.5 if i % 3 == 0 else (1.0 if i % 3 == 1 else 40.0)))
    return np.array(rows), np.array(Ps)


def main():
    if not O.pyjac_available():
        raise SystemExit("build the reference library first: make -C oracle ref")
    y, P = states()
    f = np.stack([O.pyjac_rhs(P[i], y[i]) for i in range(len(y))])
    # full Jacobians for a subset keep the fixture small
    jidx = np.array([0, 5, 7, 13, 17, 21, 25, 29])
    J = np.stack([O.pyjac_jac(P[i], y[i]) for i in jidx])
    np.savez_compressed(OUT, y=y, P=P, f=f, jidx=jidx, J=J)
    print(f"wrote {OUT}: {len(y)} RHS vectors, {len(jidx)} Jacobians")


if __name__ == "__main__":
    main()
