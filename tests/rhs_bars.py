"""Error bars for comparing a kernel's chemistry RHS with the oracle's at floating-point
noise, shared by the GPU component tests.

A species rate sums nu_kr (fwd_r - rev_r) pm_r over reactions whose terms cancel near
equilibrium, so rounding in any summation order moves it by a few ulps of the sum of the
terms' magnitudes (``rhs_magnitude``), not of the result. The state carries T and the first
K-1 mass fractions; the eliminated one, 1 - sum(Y), is formed by the kernels in another
order than the oracle's, and where that species is a minor one (CH3CHO in GRI-3.0, A4 in
nDodecane_Reitz) its reactions make the RHS sensitive to those last bits
(``last_species_spread``)."""
import numpy as np

R = 8314.4621


def rhs_magnitude(o, P, y):
    """Per RHS component, the sum of the magnitudes of the terms it adds up (constant pressure
    or constant volume, as the oracle `o` is built)."""
    M = o.mech
    K, Rn = len(M.species), len(M.reactions)
    N = np.zeros((K, Rn))
    for r, rx in enumerate(M.reactions):
        for k, nu in list(rx.reactants) + list(rx.products):
            N[k, r] += abs(nu)
    rates = o.rates(P, y)
    mag_w = N @ ((np.abs(rates["fwd"]) + np.abs(rates["rev"])) * np.abs(rates["presmod"]))
    W = np.asarray(M.W)
    rho = float(rates["conc"] @ W)
    T = y[0]
    a = np.where((T <= np.asarray(M.tmid))[:, None], M.nasa_lo, M.nasa_hi)
    hRT = a[:, 0] + T * (a[:, 1] / 2 + T * (a[:, 2] / 3 + T * (a[:, 3] / 4 + T * a[:, 4] / 5))) + a[:, 5] / T
    cpR = a[:, 0] + T * (a[:, 1] + T * (a[:, 2] + T * (a[:, 3] + T * a[:, 4])))
    cv = 1.0 if getattr(o, "reactor", "CONP") == "CONV" else 0.0
    Yf = np.append(y[1:], 1.0 - y[1:].sum())
    c = (Yf * (cpR - cv) / W).sum() * R
    mag = np.empty(K)
    mag[0] = (np.abs(hRT - cv) * R * T * mag_w).sum() / (rho * c)
    mag[1:] = mag_w[:-1] * W[:-1] / rho
    return mag


def last_species_spread(o, P, y, ulps=8):
    """Per RHS component, how far the oracle's RHS moves when the eliminated species' mass
    fraction moves by `ulps` units of 1.0 (the N2 mass fraction, about 0.7, is moved instead:
    one ulp of it is one ulp of 1.0 in the difference)."""
    n2 = 1 + [sp.upper() for sp in o.mech.species].index("N2")
    f0 = o.rhs(P, y)
    spread = np.zeros_like(f0)
    for d in (np.inf, -np.inf):
        yy = y.copy()
        for _ in range(ulps):
            yy[n2] = np.nextafter(yy[n2], d)
        spread = np.maximum(spread, np.abs(o.rhs(P, yy) - f0))
    return spread


def rhs_error(o, P, y, f):
    """|f - f_oracle| over the bar 1e-9 |f_oracle| + 1e-13 magnitude + 2 spread, per component
    (states y [n, K], parameters P [n]); below 1 everywhere means agreement at noise level.
    Returns (error ratios, oracle RHS)."""
    ref = np.stack([o.rhs(P[i], y[i]) for i in range(len(y))])
    mag = np.stack([rhs_magnitude(o, P[i], y[i]) for i in range(len(y))])
    last = np.stack([last_species_spread(o, P[i], y[i]) for i in range(len(y))])
    return np.abs(f - ref) / (1e-9 * np.abs(ref) + 1e-13 * mag + 2.0 * last + 1e-300), ref
