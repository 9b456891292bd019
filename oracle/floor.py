"""The round-off floor of the integration match criterion.

TEST INFRASTRUCTURE ONLY (the checker's side): tests/, the smoke check and bench.py's
accuracy statement use it to judge a GPU result against the oracle.

The criterion. A cell matches when its temperature and every mass fraction above 1e-10
agree with the oracle's result to a relative 1e-6 (the integration rtol). Both runs use
CVODE's BDF (src/Integrators/CVodeIntegrator.cpp:98-155) at rtol 1e-6 / atol 1e-10.

Why a floor. CVODE's step-size, order and Jacobian-reuse decisions are threshold tests
(error norm <= 1, convergence rate, eta >= 1.5, ...). On igniting and near-equilibrium
cells, the chemistry amplifies a difference of one unit in the last place of an input by
many orders of magnitude within the substep. One flipped decision then moves the answer by
the method's global error, which is tens of times rtol on small mass fractions. So two runs
of the same CPU BDF whose inputs differ by one ulp disagree at this bar on a fixed share of
cells. An implementation that is not bit-identical to the oracle cannot do better than that
share. ``roundoff_floor`` measures it on the same cells as the GPU run, with one-ulp
perturbations of T, P and the O2 mass fraction.

How a GPU result is judged (``check_against_floor``):
* its match fraction is at least the floor's, less a statistical margin;
* on the "robust" cells, those where every perturbed CPU run matches, the GPU matches too,
  up to the rate at which one more perturbation would break such a cell;
* the fraction of cells with the oracle's step count is at least the floor's, less a margin.
"""
from __future__ import annotations

import math

import numpy as np

RTOL = 1e-6
Y_FLOOR = 1e-10
PERTURBATIONS = ("T+", "T-", "P+", "P-", "O2+")


def strict_match(y: np.ndarray, ref: np.ndarray, rtol: float = RTOL, y_floor: float = Y_FLOOR) -> np.ndarray:
    """Per-cell boolean: |T - T_ref| <= rtol |T_ref| and, for every mass fraction of the
    reference above y_floor, |Y - Y_ref| <= rtol |Y_ref|."""
    dT = np.abs(y[:, 0] - ref[:, 0]) <= rtol * np.abs(ref[:, 0])
    Yr = ref[:, 1:]
    dY = np.abs(y[:, 1:] - Yr)
    return dT & np.all((np.abs(Yr) <= y_floor) | (dY <= rtol * np.abs(Yr)), axis=1)


def deviation(y: np.ndarray, ref: np.ndarray, rtol: float = RTOL, atol: float = Y_FLOOR):
    """Per cell: |T - T_ref| / T_ref, and per cell the largest |y - y_ref| / (rtol |y_ref| + atol)
    over the state's components."""
    dT = np.abs(y[:, 0] - ref[:, 0]) / np.abs(ref[:, 0])
    w = (np.abs(y - ref) / (rtol * np.abs(ref) + atol)).max(axis=1)
    return dT, w


def _perturb(name, y0, P, o2_col):
    y, p = y0.copy(), P.copy()
    if name == "T+":
        y[:, 0] = np.nextafter(y[:, 0], np.inf)
    elif name == "T-":
        y[:, 0] = np.nextafter(y[:, 0], -np.inf)
    elif name == "P+":
        p = np.nextafter(p, np.inf)
    elif name == "P-":
        p = np.nextafter(p, -np.inf)
    elif name == "O2+":
        y[:, o2_col] = np.nextafter(y[:, o2_col], np.inf)
    else:
        raise ValueError(name)
    return y, p


def roundoff_floor(orc, P, dt, y0, nthreads=0, perturbations=PERTURBATIONS, **kw) -> dict:
    """Runs the oracle on the cells and on their one-ulp perturbations.

    Returns the reference result (``y``, ``status``, ``stats``) and, per perturbation, the
    per-cell match at the strict bar and whether the step count equals the reference's."""
    y0 = np.ascontiguousarray(y0, np.float64)
    P = np.ascontiguousarray(np.broadcast_to(np.asarray(P, np.float64), (len(y0),)))
    dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float64), (len(y0),)))
    o2_col = 1 + [sp.upper() for sp in orc.mech.species].index("O2")
    yref, sref, stref = orc.integrate(P, dt, y0, nthreads=nthreads, **kw)
    match, same_nst = {}, {}
    spread_T, spread_w = 0.0, 0.0
    for name in perturbations:
        yp, pp = _perturb(name, y0, P, o2_col)
        y, s, st = orc.integrate(pp, dt, yp, nthreads=nthreads, **kw)
        match[name] = strict_match(y, yref) & (s == sref)
        same_nst[name] = st["nst"] == stref["nst"]
        dT, w = deviation(y, yref)
        spread_T, spread_w = max(spread_T, float(dT.max())), max(spread_w, float(w.max()))
    M = np.array([match[k] for k in perturbations])
    S = np.array([same_nst[k] for k in perturbations])
    robust = M.all(axis=0)
    # leave-one-out: the rate at which a further perturbation breaks a cell the others keep
    loo = []
    for i in range(len(perturbations)):
        others = np.delete(M, i, axis=0).all(axis=0)
        if others.any():
            loo.append(M[i][others].mean())
    return {
        "y": yref, "status": sref, "stats": stref,
        "match": match, "same_nst": same_nst,
        "fraction": float(M.mean()), "fraction_min": float(M.mean(axis=1).min()),
        "same_nst_fraction": float(S.mean()),
        "robust": robust, "robust_fraction": float(robust.mean()),
        "robust_keep_rate": float(min(loo)) if loo else 1.0,
        "spread_T": spread_T, "spread_w": spread_w,
        "cells": int(len(y0)),
    }


def summary(fl: dict) -> dict:
    """The JSON-friendly part of roundoff_floor's result."""
    return {
        "cells": fl["cells"],
        "perturbations": "one ulp of T0, of P and of Y_O2 (" + ", ".join(fl["match"]) + ")",
        "cells_within_rtol_T_and_Y_above_1e-10": {k: float(v.mean()) for k, v in fl["match"].items()},
        "mean": fl["fraction"],
        "same_nst": fl["same_nst_fraction"],
        "robust_cells": fl["robust_fraction"],
        "robust_keep_rate": fl["robust_keep_rate"],
        "spread_T": fl["spread_T"], "spread_w": fl["spread_w"],
    }


def compare(y, status, stats, fl: dict) -> dict:
    """A result (GPU) judged against the floor measured on the same cells."""
    m = strict_match(y, fl["y"]) & (status == fl["status"])
    rob = fl["robust"]
    return {
        "cells_within_rtol_T_and_Y_above_1e-10": float(m.mean()),
        "floor": fl["fraction"],
        "robust_cells_matched": float(m[rob].mean()) if rob.any() else 1.0,
        "same_nst": float((stats["nst"] == fl["stats"]["nst"]).mean()) if stats is not None else None,
        "floor_same_nst": fl["same_nst_fraction"],
        "max_dT": float(deviation(y, fl["y"])[0].max()),
        "max_w": float(deviation(y, fl["y"])[1].max()),
    }


def check_against_floor(y, status, stats, fl: dict, sigmas: float = 4.0) -> dict:
    """Asserts the three conditions of the module docstring; returns the comparison."""
    c = compare(y, status, stats, fl)
    n = fl["cells"]
    f = fl["fraction"]
    margin = sigmas * math.sqrt(max(f * (1 - f), 0.01) / n)
    assert c["cells_within_rtol_T_and_Y_above_1e-10"] >= f - margin, (
        f"match fraction {c['cells_within_rtol_T_and_Y_above_1e-10']:.3f} below the round-off floor "
        f"{f:.3f} - {margin:.3f}")
    nrob = int(fl["robust"].sum())
    if nrob:
        keep = fl["robust_keep_rate"]
        rmargin = sigmas * math.sqrt(max(keep * (1 - keep), 0.01) / nrob)
        assert c["robust_cells_matched"] >= keep - rmargin, (
            f"robust cells matched {c['robust_cells_matched']:.3f} < {keep:.3f} - {rmargin:.3f} "
            f"({nrob} robust cells)")
    if stats is not None:
        s = fl["same_nst_fraction"]
        smargin = sigmas * math.sqrt(max(s * (1 - s), 0.01) / n)
        assert c["same_nst"] >= s - smargin, f"same step count {c['same_nst']:.3f} < {s:.3f} - {smargin:.3f}"
    return c


def check_spread(y, fl: dict, factor: float = 4.0, T_min: float = 1e-5, w_min: float = 300.0):
    """Every cell, matched or not: |dT| / T and the largest |dy| / (rtol |y| + atol) stay within
    `factor` times the largest deviation the one-ulp runs of the floor showed (and never below
    T_min, w_min). A sanity bound on the cells the strict criterion does not judge."""
    dT, w = deviation(y, fl["y"])
    bT, bw = max(T_min, factor * fl["spread_T"]), max(w_min, factor * fl["spread_w"])
    assert dT.max() <= bT, f"T max rel diff {dT.max():.3e} > {bT:.3e} (floor spread {fl['spread_T']:.3e})"
    assert w.max() <= bw, f"max |dy|/(rtol|y|+atol) {w.max():.3e} > {bw:.3e} (floor spread {fl['spread_w']:.3e})"
    return dT, w
