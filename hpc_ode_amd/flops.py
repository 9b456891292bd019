"""Algorithmic work model of one BDF substep of one cell.

W_cell = nfe*F_rhs + nje*F_jac + nsetups*(F_form + F_LU) + nni*F_solve + nst*F_step

The per-call counts F_* are static operation counts of the table-driven evaluator for a
given mechanism (add/sub/mul/div = 1 flop, fused multiply-add = 2); transcendental calls
(exp, log, log10, pow) are counted separately and add no flops. The counters
(nfe, nje, nsetups, nni, nst) are taken from the CPU oracle run on the same cells.
Dense LU / solve counts are the textbook ones for an n x n system.
"""
from __future__ import annotations

import numpy as np


def lu_flops(n: int) -> float:
    return 2.0 / 3.0 * n ** 3 - 0.5 * n ** 2 - n / 6.0


def solve_flops(n: int) -> float:
    return 2.0 * n * n - n


def form_flops(n: int) -> float:
    return float(n * n + n)


def rhs_counts(kind, rev, nreac_nu, nprod_nu, n_eff, K):
    """(flops, transcendentals) of one RHS evaluation.

    kind/rev: per reaction; nreac_nu/nprod_nu: sum of stoichiometric coefficients per side;
    n_eff: efficiency entries per reaction."""
    fl = 0.0
    tr = 0
    # species: Y, 1/W sums, density, NASA7 cp/h/s (Horner: 4+5+5 fma), g, C
    fl += K * (2 + 2 * 4 + 2 * 5 + 2 * 5 + 1 + 2 + 2)
    tr += 2  # log T, log(P_atm/(R T))
    for r in range(len(kind)):
        fl += 4                           # lnA + b lnT - theta/T
        tr += 1                           # exp
        fl += nreac_nu[r]                 # product of concentrations
        if rev[r]:
            fl += 2 * (nreac_nu[r] + nprod_nu[r]) + 3 + nprod_nu[r]  # dG, dnu, Kc exponent, reverse product
            tr += 1
        fl += 2                           # (fwd - rev) * pm
        if kind[r] >= 1:
            fl += 1 + 2 * n_eff[r]        # third-body sum
        if kind[r] >= 2:
            fl += 4 + 3                   # k0, Pr, Pr/(1+Pr)
            tr += 1
        if kind[r] == 3:
            fl += 25                      # Fcent, A, B, F
            tr += 6
    return fl, tr


def jac_counts(kind, rev, nreac_nu, nprod_nu, n_eff, K, n_terms):
    fl, tr = rhs_counts(kind, rev, nreac_nu, nprod_nu, n_eff, K)
    for r in range(len(kind)):
        ns = nreac_nu[r] + (nprod_nu[r] if rev[r] else 0)
        fl += 3 * ns + 10 + 2 * n_eff[r]  # slot derivatives, d/dT, d/dm
        if kind[r] >= 2:
            fl += 20
        if kind[r] == 3:
            fl += 30
    fl += 2 * n_terms                     # assembly of d wdot / dC
    fl += K * (2 * K + 12 * K)            # sigma and the chain rule to (T, Y)
    return fl, tr


def mechanism_counts(oracle_tables) -> dict:
    """F_rhs, F_jac and transcendental counts from the oracle's flat tables."""
    t = oracle_tables
    K = int(t.K)
    kind = t.kind
    rev = t.rev
    nreac_nu = [float(t.reac_nu[r, : t.nreac[r]].sum()) for r in range(t.R)]
    nprod_nu = [float(t.prod_nu[r, : t.nprod[r]].sum()) for r in range(t.R)]
    n_eff = [int(t.eff_ptr[r + 1] - t.eff_ptr[r]) for r in range(t.R)]
    # Jacobian assembly terms: (species row or T row) x (slots of the reaction)
    n_terms = 0
    for r in range(t.R):
        net = {}
        for i in range(t.nreac[r]):
            net[t.reac_sp[r, i]] = net.get(t.reac_sp[r, i], 0.0) - t.reac_nu[r, i]
        for i in range(t.nprod[r]):
            net[t.prod_sp[r, i]] = net.get(t.prod_sp[r, i], 0.0) + t.prod_nu[r, i]
        rows = sum(1 for v in net.values() if v != 0.0) + 1
        slots = t.nreac[r] + (t.nprod[r] if rev[r] else 0) + n_eff[r]
        n_terms += rows * slots
    frhs, trhs = rhs_counts(kind, rev, nreac_nu, nprod_nu, n_eff, K)
    fjac, tjac = jac_counts(kind, rev, nreac_nu, nprod_nu, n_eff, K, n_terms)
    n = K
    return dict(F_rhs=frhs, T_rhs=trhs, F_jac=fjac, T_jac=tjac, F_LU=lu_flops(n),
                F_solve=solve_flops(n), F_form=form_flops(n), F_step=float(40 * n), n=n)


def cell_flops(counts: dict, nfe, nje, nsetups, nni, nst) -> float:
    c = counts
    return (nfe * c["F_rhs"] + nje * c["F_jac"] + nsetups * (c["F_form"] + c["F_LU"]) +
            nni * c["F_solve"] + nst * c["F_step"])


def mean_cell_flops(counts: dict, stats: dict) -> float:
    return float(np.mean(cell_flops(counts, stats["nfe"], stats["nje"], stats["nsetups"],
                                    stats["nni"], stats["nst"])))
