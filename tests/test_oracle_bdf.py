"""Pins the oracle's dense LU and CVODE-style BDF against LAPACK, closed forms and
high-accuracy solutions. CPU only."""
import numpy as np
import pytest
import scipy.linalg as sla
from scipy.integrate import solve_ivp

import oracle as O


def test_lu_matches_lapack_getrf():
    """Same partial pivoting as LAPACK dgetrf: identical pivot sequence and factors."""
    rng = np.random.default_rng(1)
    for n in (1, 5, 10, 53):
        A = rng.standard_normal((n, n))
        lu, piv, ier = O.getrf(A)
        assert ier == 0
        lu_ref, piv_ref = sla.lu_factor(A)
        assert np.array_equal(piv, piv_ref)
        assert np.allclose(lu.reshape(n, n).T, lu_ref, rtol=1e-12, atol=1e-12)
        b = rng.standard_normal(n)
        x = O.getrs(lu, piv, b)
        assert np.allclose(A @ x, b, rtol=1e-10, atol=1e-10)


def test_lu_singular_reports_step():
    A = np.eye(4)
    A[2, 2] = 0.0
    _, _, ier = O.getrf(A)
    assert ier == 3        # SUNDIALS convention: 1-based index of the zero pivot


def test_lu_tie_break_first_max():
    """Ties in |pivot| pick the first row, like LAPACK's idamax."""
    A = np.array([[1.0, 2.0], [-1.0, 3.0]])
    _, piv, _ = O.getrf(A)
    assert piv[0] == 0


def test_bdf_scalar_decay_closed_form():
    lam = -1000.0
    for rtol in (1e-4, 1e-6, 1e-8):
        y, rc, st = O.cvode_generic(lambda t, y: lam * y, lambda t, y: np.array([[lam]]),
                                    [1.0], 0.0, 0.01, rtol=rtol, atol=1e-14)
        assert rc == 0
        exact = np.exp(lam * 0.01)
        assert abs(y[0] - exact) < 50 * rtol * 1.0


def test_bdf_error_scales_with_tolerance():
    """Global error of a stiff linear system decreases with rtol (BDF error control works)."""
    A = np.array([[-1.0, 1.0], [0.0, -1000.0]])
    y0 = np.array([1.0, 1.0])
    exact = sla.expm(A * 2.0) @ y0
    errs = []
    for rtol in (1e-3, 1e-5, 1e-7, 1e-9):
        y, rc, _ = O.cvode_generic(lambda t, y: A @ y, lambda t, y: A, y0, 0.0, 2.0, rtol=rtol, atol=1e-14)
        assert rc == 0
        errs.append(np.abs(y - exact).max())
    assert errs[0] > errs[1] > errs[2] > errs[3]
    assert errs[3] < 1e-7


def test_bdf_robertson_vs_radau():
    """Robertson's stiff kinetics (the classic BDF test) against scipy Radau at rtol 1e-12."""
    def f(t, y):
        return np.array([-0.04 * y[0] + 1e4 * y[1] * y[2],
                         0.04 * y[0] - 1e4 * y[1] * y[2] - 3e7 * y[1] ** 2,
                         3e7 * y[1] ** 2])

    def jac(t, y):
        return np.array([[-0.04, 1e4 * y[2], 1e4 * y[1]],
                         [0.04, -1e4 * y[2] - 6e7 * y[1], -1e4 * y[1]],
                         [0.0, 6e7 * y[1], 0.0]])
    y0 = [1.0, 0.0, 0.0]
    ref = solve_ivp(f, (0, 40.0), y0, method="Radau", rtol=1e-12, atol=1e-16, jac=jac).y[:, -1]
    y, rc, st = O.cvode_generic(f, jac, y0, 0.0, 40.0, rtol=1e-8, atol=1e-14)
    assert rc == 0
    assert np.allclose(y, ref, rtol=1e-5, atol=1e-12)
    assert st["nje"] < st["nsetups"] <= st["nni"]     # Jacobian reuse across setups


def test_bdf_maxsteps_and_bad_input():
    lam = -1.0
    y, rc, st = O.cvode_generic(lambda t, y: lam * y, lambda t, y: np.array([[lam]]), [1.0], 0.0, 100.0,
                                maxsteps=3)
    assert rc == -1                                      # CV_TOO_MUCH_WORK
    assert st["nst"] == 3
    y, rc, _ = O.cvode_generic(lambda t, y: lam * y, lambda t, y: np.array([[lam]]), [1.0], 0.0, 1.0,
                               rtol=-1.0, atol=0.0)
    assert rc == -22                                     # CV_ILL_INPUT


def test_bdf_fp32_lu_matches_fp64():
    """FP32 Newton-matrix LU with FP64 refinement converges to the same answer at tolerance."""
    A = np.array([[-1.0, 1.0, 0.0], [0.0, -1e4, 1.0], [1e3, 0.0, -1e5]])
    y0 = np.array([1.0, 0.5, 0.1])
    y64, rc64, _ = O.cvode_generic(lambda t, y: A @ y, lambda t, y: A, y0, 0.0, 1.0, rtol=1e-8, atol=1e-14)
    for sw in (0, 1, 2):
        y32, rc32, _ = O.cvode_generic(lambda t, y: A @ y, lambda t, y: A, y0, 0.0, 1.0, rtol=1e-8,
                                       atol=1e-14, lu_fp32=True, refine_sweeps=sw)
        assert rc32 == 0
        assert np.allclose(y32, y64, rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("T0", [1300.0, 1800.0])
def test_gri_cell_vs_radau(gri_path, T0):
    """One GRI-3.0 cell: the rtol-1e-12 BDF and scipy Radau agree, and the rtol-1e-6 run is
    within its tolerance's reach of both."""
    from hpc_ode_amd import workloads as WL
    g = O.Oracle(gri_path)
    y, P, _ = WL.gri_plain(g.mech.species, g.mech.W, 1, seed=2)
    y0 = y[0]
    y0[0] = T0
    dt = 1e-4
    yt, st, _ = g.integrate(P[0], dt, y0[None], rtol=1e-12, atol=1e-20, maxsteps=10 ** 6)
    ys = solve_ivp(lambda t, yy: g.rhs(P[0], yy), (0, dt), y0, method="Radau", rtol=1e-12, atol=1e-18,
                   jac=lambda t, yy: g.jac(P[0], yy)).y[:, -1]
    assert st[0] == 0
    assert abs(yt[0, 0] - ys[0]) / ys[0] < 1e-9
    m = np.abs(ys) > 1e-10
    assert np.max(np.abs(yt[0] - ys)[m] / np.abs(ys)[m]) < 1e-5
    y6, st6, _ = g.integrate(P[0], dt, y0[None])
    assert abs(y6[0, 0] - ys[0]) / ys[0] < 1e-4


def test_tight_reference_settings_are_converged(gri_path):
    """The rtol-1e-12 / atol-1e-16 run used as the accuracy reference agrees with an
    atol-1e-20 run to far below the rtol-1e-6 errors it is used to measure."""
    from hpc_ode_amd import workloads as WL
    g = O.Oracle(gri_path)
    y0, P, dt = WL.gri_plain(g.mech.species, g.mech.W, 8, seed=3)
    a, sa, _ = g.integrate(P, dt, y0, rtol=1e-12, atol=1e-16, maxsteps=10 ** 6, nthreads=8)
    b, sb, _ = g.integrate(P, dt, y0, rtol=1e-12, atol=1e-20, maxsteps=10 ** 6, nthreads=8)
    assert (sa == 0).all() and (sb == 0).all()
    assert np.max(np.abs(a[:, 0] - b[:, 0]) / b[:, 0]) < 1e-9
    m = np.abs(b[:, 1:]) > 1e-10
    assert np.max(np.abs(a[:, 1:] - b[:, 1:])[m] / np.abs(b[:, 1:])[m]) < 1e-6
