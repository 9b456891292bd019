"""Mechanisms read from the reference's CTI files (hpc_ode_amd/mechanisms/*.cti, copied unchanged
from resources/mechanisms/) drive the kernels exactly as the same mechanisms read from YAML:
the RHS, the analytic Jacobian and an integration are bit-identical between the two readers on
the same states, which covers every table the device uses (stoichiometry, third-body
efficiencies, falloff and Troe parameters) beyond what tests/test_cti_reader.py compares on the
host. GRI-3.0 runs the wave layout, nDodecane_Reitz the block layout."""
import os

import numpy as np
import pytest

import hpc_ode_amd as H
from hpc_ode_amd import workloads as WL

pytestmark = pytest.mark.gpu

MECH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hpc_ode_amd", "mechanisms")


def _states(name, m, n):
    if name == "gri30":
        return WL.gri_plain(m.species, m.W, n, seed=21)
    return WL.dodecane_air(m.species, m.W, n, seed=21)


@pytest.mark.parametrize("name", ["gri30", "nDodecane_Reitz"])
def test_cti_mechanism_matches_yaml_on_gpu(name):
    c = H.Mechanism(os.path.join(MECH, name + ".cti"))
    y = H.Mechanism(os.path.join(MECH, name + ".yaml"))
    n = 64 if name == "nDodecane_Reitz" else 256
    y0, P, dt = _states(name, y, n)
    print(f"{name}: {n} cells", flush=True)
    assert np.array_equal(c.rhs(y0, P), y.rhs(y0, P))
    assert np.array_equal(c.jac(y0[:16], P[:16]), y.jac(y0[:16], P[:16]))
    yc, sc, stc = c.integrate(y0, P, dt, lu_precision=1)
    yy, sy, sty = y.integrate(y0, P, dt, lu_precision=1)
    assert (sc == 0).all() and (sy == 0).all()
    assert np.array_equal(yc, yy)
    assert np.array_equal(stc["nst"], sty["nst"])
