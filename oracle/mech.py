"""Oracle-side mechanism reader: Cantera YAML or CTML (XML) -> flat numpy tables.

TEST INFRASTRUCTURE ONLY. This module is part of the CPU checker: only tests,
the smoke check and the CPU-baseline leg of bench.py may import it. The product
(hpc_ode_amd) has its own C++ reader; a test asserts both produce equal tables.

What it restates: the physics constants and unit conventions that pyJac baked
into the reference's generated C (src/Mechanism/CPU):
  * R = 8314.4621 J/(kmol K)                    (chem_utils.c:42, rxn_rates_pres_mod.c:20)
  * A in SI-kmol, i.e. (cm^3/mol)^(n-1) -> (m^3/kmol)^(n-1): factor 1e-3^(n-1), n = reaction
    order counting M for three-body and low-pressure limits  (rxn_rates.c:8-9, rxn_rates_pres_mod.c:27-29)
  * theta = Ea / R with Ea in cal/mol converted by 4184 J/kcal  (rxn_rates.c:33)
  * NASA7 low range used for T <= T_mid (chem_utils.c:197, rxn_rates.c:2803)
  * the last species of the phase is the eliminated one (mechanism.h:5)
"""
from __future__ import annotations

import dataclasses
import re
from typing import Dict, List, Tuple

import numpy as np
import yaml

R_GAS = 8314.4621          # J/(kmol K)
P_ATM = 101325.0           # Pa
CAL = 4184.0               # J/kcal == J per (cal/mol) * (mol/kmol)

# atomic weights as Cantera 2.x tabulates them (kg/kmol)
ATOMIC_WEIGHTS = {
    "H": 1.008, "D": 2.014102, "He": 4.002602, "C": 12.011, "N": 14.007, "O": 15.999,
    "F": 18.998403163, "Ne": 20.1797, "Si": 28.085, "S": 32.06, "Cl": 35.45,
    "Ar": 39.95, "Kr": 83.798, "Xe": 131.293, "E": 0.000548579909,
}

RXN_ELEMENTARY, RXN_THREE_BODY, RXN_LINDEMANN, RXN_TROE = 0, 1, 2, 3


@dataclasses.dataclass
class Reaction:
    equation: str
    kind: int
    reversible: bool
    reactants: List[Tuple[int, float]]
    products: List[Tuple[int, float]]
    A: float
    b: float
    theta: float
    A0: float = 0.0
    b0: float = 0.0
    theta0: float = 0.0
    efficiencies: Dict[int, float] = dataclasses.field(default_factory=dict)
    default_eff: float = 1.0
    troe: Tuple[float, float, float, float] = (0.0, 0.0, 0.0, 0.0)
    troe_has_t2: bool = False
    # pressure-dependent Arrhenius (PLOG): (P in Pa, A, b, theta) per level, ascending in P;
    # empty for every other reaction. The reaction is elementary otherwise.
    plog: List[Tuple[float, float, float, float]] = dataclasses.field(default_factory=list)


@dataclasses.dataclass
class Mechanism:
    species: List[str]
    W: np.ndarray              # [K] kg/kmol
    tmid: np.ndarray           # [K]
    nasa_lo: np.ndarray        # [K,7]
    nasa_hi: np.ndarray        # [K,7]
    reactions: List[Reaction]

    @property
    def n_species(self) -> int:
        return len(self.species)

    @property
    def n_state(self) -> int:
        return len(self.species)


def _length_factor(units: dict) -> float:
    """m per length unit; concentrations scale with length^3/quantity."""
    return {"cm": 1e-2, "m": 1.0, "mm": 1e-3}[units.get("length", "m")]


def _quantity_factor(units: dict) -> float:
    """kmol per quantity unit."""
    return {"mol": 1e-3, "kmol": 1.0, "molec": 1.0 / 6.02214076e26}[units.get("quantity", "kmol")]


def _ea_factor(units: dict) -> float:
    """J/kmol per activation-energy unit."""
    u = units.get("activation-energy", "J/kmol")
    return {"cal/mol": CAL, "kcal/mol": CAL * 1e3, "J/mol": 1e3, "kJ/mol": 1e6,
            "J/kmol": 1.0, "K": R_GAS}[u]


_TERM = re.compile(r"^\s*(?:(\d+(?:\.\d+)?)\s+)?(\S+)\s*$")


PRESSURE_UNITS = {"Pa": 1.0, "kPa": 1e3, "MPa": 1e6, "bar": 1e5, "atm": P_ATM, "torr": P_ATM / 760.0}


def _plog_levels(levels):
    """Sorted (P, A, b, theta) levels; a pressure given twice (a sum of rates at one level) and a
    non-positive A are refused."""
    levels = sorted(levels)
    ps = [lv[0] for lv in levels]
    if len(set(ps)) != len(ps):
        raise ValueError("PLOG with several rate expressions at one pressure unsupported")
    if any(lv[1] <= 0 for lv in levels):
        raise ValueError("PLOG with a non-positive pre-exponential factor unsupported")
    return levels


def _parse_side(side: str) -> Tuple[List[Tuple[str, float]], str | None, bool]:
    """Returns (species terms, falloff collider or None, has_plus_M)."""
    collider = None
    m = re.search(r"\(\+\s*([^)]+?)\s*\)", side)
    if m:
        collider = m.group(1)
        side = side[: m.start()] + side[m.end():]
    terms = []
    has_m = False
    for tok in side.split(" + "):
        tok = tok.strip()
        if not tok:
            continue
        mm = _TERM.match(tok)
        if mm is None:
            raise ValueError(f"cannot parse term {tok!r}")
        coef = float(mm.group(1)) if mm.group(1) else 1.0
        name = mm.group(2)
        if name == "M":
            has_m = True
            continue
        terms.append((name, coef))
    return terms, collider, has_m


def _merge(terms, index):
    out: Dict[int, float] = {}
    for name, c in terms:
        k = index[name]
        out[k] = out.get(k, 0.0) + c
    return sorted(out.items())


class _Loader(yaml.SafeLoader):
    """YAML 1.2 booleans: species such as NO / N / Y stay strings (YAML 1.1 would make them bools)."""


_Loader.yaml_implicit_resolvers = {
    k: [(tag, rx) for tag, rx in v if tag != "tag:yaml.org,2002:bool"]
    for k, v in yaml.SafeLoader.yaml_implicit_resolvers.items()
}
_Loader.add_implicit_resolver("tag:yaml.org,2002:bool", re.compile(r"^(?:true|false|True|False|TRUE|FALSE)$"),
                              list("tTfF"))


def load(path: str) -> Mechanism:
    """Reads a Cantera YAML file, or a CTML file when the path ends in .xml / .ctml."""
    if path.endswith((".xml", ".ctml")):
        return load_xml(path)
    return load_yaml(path)


def load_xml(path: str) -> Mechanism:
    """Cantera CTML: rate constants already in SI-kmol units, activation energies with a units
    attribute (cal/mol when absent)."""
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    phase = root.find(".//phase")
    names = phase.find("speciesArray").text.split()
    index = {n: i for i, n in enumerate(names)}
    sref = phase.find("speciesArray").get("datasrc", "#").lstrip("#")
    sdata = next(e for e in root.iter("speciesData") if not sref or e.get("id") == sref)
    K = len(names)
    W = np.zeros(K)
    tmid = np.zeros(K)
    lo = np.zeros((K, 7))
    hi = np.zeros((K, 7))
    for sp in sdata.findall("species"):
        n = sp.get("name")
        if n not in index:
            continue
        k = index[n]
        W[k] = sum(ATOMIC_WEIGHTS[t.split(":")[0]] * float(t.split(":")[1])
                   for t in sp.find("atomArray").text.split())
        blocks = sorted(sp.find("thermo").findall("NASA"), key=lambda b: float(b.get("Tmin")))
        coef = [np.array([float(v) for v in b.find("floatArray").text.replace(",", " ").split()]) for b in blocks]
        tmid[k] = float(blocks[0].get("Tmax"))
        lo[k] = coef[0]
        hi[k] = coef[-1]

    def ea(e):
        if e is None:
            return 0.0
        u = e.get("units", "cal/mol")
        return float(e.text) * {"cal/mol": CAL, "kcal/mol": CAL * 1e3, "J/mol": 1e3, "kJ/mol": 1e6,
                                "J/kmol": 1.0, "K": R_GAS}[u] / R_GAS

    def arr(a):
        b = a.find("b")
        return float(a.find("A").text), float(b.text) if b is not None else 0.0, ea(a.find("E"))

    def pairs(e):
        if e is None or not e.text:
            return []
        return [(t.rsplit(":", 1)[0], float(t.rsplit(":", 1)[1])) for t in e.text.split()]

    rref = phase.find("reactionArray").get("datasrc", "#").lstrip("#")
    rdata = next(e for e in root.iter("reactionData") if not rref or e.get("id") == rref)
    rxns: List[Reaction] = []
    for x in rdata.findall("reaction"):
        eq = x.find("equation").text
        rev = x.get("reversible", "yes") == "yes"
        if x.find("order") is not None:
            raise ValueError("non-standard reaction orders unsupported")
        reac = _merge(pairs(x.find("reactants")), index)
        prod = _merge(pairs(x.find("products")), index)
        rc = x.find("rateCoeff")
        hiA = [a for a in rc.findall("Arrhenius") if a.get("name") != "k0"][0]
        A, b, th = arr(hiA)
        effs, deff = {}, 1.0
        e = rc.find("efficiencies")
        if e is not None:
            deff = float(e.get("default", "1.0"))
            effs = {index[n]: v for n, v in pairs(e) if n in index}
        typ = x.get("type", "elementary")
        if typ == "elementary":
            rx = Reaction(eq, RXN_ELEMENTARY, rev, reac, prod, A, b, th)
        elif typ == "plog":
            levels = []
            for a in rc.findall("Arrhenius"):
                pe = a.find("P")
                levels.append((float(pe.text) * PRESSURE_UNITS[pe.get("units", "Pa")], *arr(a)))
            rx = Reaction(eq, RXN_ELEMENTARY, rev, reac, prod, A, b, th, plog=_plog_levels(levels))
        elif typ == "threeBody":
            rx = Reaction(eq, RXN_THREE_BODY, rev, reac, prod, A, b, th, efficiencies=effs, default_eff=deff)
        elif typ == "falloff":
            A0, b0, th0 = arr([a for a in rc.findall("Arrhenius") if a.get("name") == "k0"][0])
            f = rc.find("falloff")
            ftype = f.get("type", "Lindemann") if f is not None else "Lindemann"
            if ftype == "Troe":
                t = [float(v) for v in f.text.split()]
                kind, troe, has_t2 = RXN_TROE, (t[0], t[1], t[2], t[3] if len(t) > 3 else 0.0), len(t) > 3
            elif ftype == "Lindemann":
                kind, troe, has_t2 = RXN_LINDEMANN, (0.0, 0.0, 0.0, 0.0), False
            else:
                raise ValueError(f"falloff type {ftype} unsupported")
            rx = Reaction(eq, kind, rev, reac, prod, A, b, th, A0, b0, th0, efficiencies=effs, default_eff=deff,
                          troe=troe, troe_has_t2=has_t2)
        else:
            raise ValueError(f"reaction type {typ} unsupported")
        rxns.append(rx)
    return Mechanism(names, W, tmid, lo, hi, rxns)


def load_yaml(path: str) -> Mechanism:
    with open(path) as f:
        doc = yaml.load(f, Loader=_Loader)
    units = doc.get("units", {})
    phase = doc["phases"][0]
    names = list(phase["species"])
    index = {n: i for i, n in enumerate(names)}
    spdb = {s["name"]: s for s in doc["species"]}
    K = len(names)
    W = np.zeros(K)
    tmid = np.zeros(K)
    lo = np.zeros((K, 7))
    hi = np.zeros((K, 7))
    for k, n in enumerate(names):
        s = spdb[n]
        W[k] = sum(ATOMIC_WEIGHTS[e] * c for e, c in s["composition"].items())
        th = s["thermo"]
        if th["model"] != "NASA7":
            raise ValueError(f"{n}: thermo model {th['model']} unsupported")
        tr = th["temperature-ranges"]
        data = th["data"]
        if len(data) == 1:
            tmid[k] = tr[-1]
            lo[k] = data[0]
            hi[k] = data[0]
        else:
            tmid[k] = tr[1]
            lo[k] = data[0]
            hi[k] = data[1]

    lf = _length_factor(units)
    qf = _quantity_factor(units)
    ef = _ea_factor(units)
    conc_factor = lf ** 3 / qf    # (length^3/quantity) -> m^3/kmol

    def arr(rc, order):
        A = float(rc["A"]) * conc_factor ** (order - 1)
        return A, float(rc["b"]), float(rc["Ea"]) * ef / R_GAS

    rxns: List[Reaction] = []
    for rd in doc["reactions"]:
        eq = rd["equation"]
        if "<=>" in eq:
            lhs, rhs = eq.split("<=>")
            rev = True
        elif "=>" in eq:
            lhs, rhs = eq.split("=>")
            rev = False
        else:
            lhs, rhs = eq.split("=")
            rev = True
        if "orders" in rd:
            raise ValueError("non-standard reaction orders unsupported")
        rt, rcol, rm = _parse_side(lhs)
        pt, pcol, pm = _parse_side(rhs)
        reac = _merge(rt, index)
        prod = _merge(pt, index)
        order = sum(c for _, c in reac)
        typ = rd.get("type", "elementary")
        effs = {index[k]: float(v) for k, v in rd.get("efficiencies", {}).items() if k in index}
        deff = float(rd.get("default-efficiency", 1.0))
        if typ == "three-body" or (typ == "elementary" and rm):
            A, b, th = arr(rd["rate-constant"], order + 1)
            rx = Reaction(eq, RXN_THREE_BODY, rev, reac, prod, A, b, th,
                          efficiencies=effs, default_eff=deff)
        elif typ == "falloff":
            A, b, th = arr(rd["high-P-rate-constant"], order)
            A0, b0, th0 = arr(rd["low-P-rate-constant"], order + 1)
            if rcol is not None and rcol != "M":
                effs = {index[rcol]: 1.0}
                deff = 0.0
            if "Troe" in rd:
                t = rd["Troe"]
                kind = RXN_TROE
                troe = (float(t["A"]), float(t["T3"]), float(t["T1"]), float(t.get("T2", 0.0)))
                has_t2 = "T2" in t
            elif "SRI" in rd:
                raise ValueError("SRI falloff unsupported")
            else:
                kind = RXN_LINDEMANN
                troe = (0.0, 0.0, 0.0, 0.0)
                has_t2 = False
            rx = Reaction(eq, kind, rev, reac, prod, A, b, th, A0, b0, th0,
                          efficiencies=effs, default_eff=deff, troe=troe, troe_has_t2=has_t2)
        elif typ == "elementary":
            A, b, th = arr(rd["rate-constant"], order)
            rx = Reaction(eq, RXN_ELEMENTARY, rev, reac, prod, A, b, th)
        elif typ == "pressure-dependent-Arrhenius":
            levels = []
            for rc in rd["rate-constants"]:
                pv = rc["P"]
                if isinstance(pv, str):
                    v, u = pv.split()
                    pv = float(v) * PRESSURE_UNITS[u]
                levels.append((float(pv), *arr(rc, order)))
            levels = _plog_levels(levels)
            rx = Reaction(eq, RXN_ELEMENTARY, rev, reac, prod, *levels[0][1:], plog=levels)
        else:
            raise ValueError(f"reaction type {typ} unsupported")
        rxns.append(rx)
    return Mechanism(names, W, tmid, lo, hi, rxns)


# ----------------------------------------------------------------------------
# Flat tables for the C oracle (oracle.h: struct or_mech)
# ----------------------------------------------------------------------------
MAX_SLOTS = 6   # reactant / product species per reaction side


@dataclasses.dataclass
class Tables:
    K: int
    R: int
    W: np.ndarray
    tmid: np.ndarray
    nasa: np.ndarray          # [K, 14]: low 7 then high 7
    kind: np.ndarray          # [R] int32
    rev: np.ndarray           # [R] int32
    nreac: np.ndarray         # [R] int32
    nprod: np.ndarray         # [R] int32
    reac_sp: np.ndarray       # [R, MAX_SLOTS] int32
    reac_nu: np.ndarray       # [R, MAX_SLOTS] float64
    prod_sp: np.ndarray
    prod_nu: np.ndarray
    arr: np.ndarray           # [R, 8]: lnA, b, theta, A, lnA0, b0, theta0, A0
    troe: np.ndarray          # [R, 5]: a, T3, T1, T2, has_T2
    eff_ptr: np.ndarray       # [R+1] int32
    eff_sp: np.ndarray        # [E] int32
    eff_val: np.ndarray       # [E] float64  (efficiency - 1)
    eff_base: np.ndarray      # [R] float64 default efficiency
    plog_ptr: np.ndarray      # [R+1] int32: levels of reaction r are plog[plog_ptr[r]:plog_ptr[r+1]]
    plog: np.ndarray          # [L, 4]: ln P (Pa), lnA, b, theta


def tables(m: Mechanism) -> Tables:
    K = m.n_species
    R = len(m.reactions)
    kind = np.zeros(R, np.int32)
    rev = np.zeros(R, np.int32)
    nreac = np.zeros(R, np.int32)
    nprod = np.zeros(R, np.int32)
    reac_sp = np.zeros((R, MAX_SLOTS), np.int32)
    reac_nu = np.zeros((R, MAX_SLOTS))
    prod_sp = np.zeros((R, MAX_SLOTS), np.int32)
    prod_nu = np.zeros((R, MAX_SLOTS))
    arr = np.zeros((R, 8))
    troe = np.zeros((R, 5))
    eff_ptr = [0]
    eff_sp: List[int] = []
    eff_val: List[float] = []
    eff_base = np.ones(R)
    plog_ptr = [0]
    plog: List[Tuple[float, float, float, float]] = []
    for r, rx in enumerate(m.reactions):
        kind[r] = rx.kind
        rev[r] = int(rx.reversible)
        if len(rx.reactants) > MAX_SLOTS or len(rx.products) > MAX_SLOTS:
            raise ValueError("too many species on one side")
        nreac[r] = len(rx.reactants)
        nprod[r] = len(rx.products)
        for s, (k, c) in enumerate(rx.reactants):
            reac_sp[r, s] = k
            reac_nu[r, s] = c
        for s, (k, c) in enumerate(rx.products):
            prod_sp[r, s] = k
            prod_nu[r, s] = c
        arr[r] = (np.log(rx.A) if rx.A > 0 else -np.inf, rx.b, rx.theta, rx.A,
                  np.log(rx.A0) if rx.A0 > 0 else -np.inf, rx.b0, rx.theta0, rx.A0)
        troe[r] = (*rx.troe, float(rx.troe_has_t2))
        if rx.kind != RXN_ELEMENTARY:
            eff_base[r] = rx.default_eff
            for k in sorted(rx.efficiencies):
                eff_sp.append(k)
                eff_val.append(rx.efficiencies[k] - rx.default_eff)
        eff_ptr.append(len(eff_sp))
        for P, A, b, th in rx.plog:
            plog.append((np.log(P), np.log(A), b, th))
        plog_ptr.append(len(plog))
    nasa = np.concatenate([m.nasa_lo, m.nasa_hi], axis=1)
    return Tables(K, R, m.W.copy(), m.tmid.copy(), np.ascontiguousarray(nasa), kind, rev,
                  nreac, nprod, reac_sp, reac_nu, prod_sp, prod_nu, arr, troe,
                  np.asarray(eff_ptr, np.int32), np.asarray(eff_sp, np.int32),
                  np.asarray(eff_val, np.float64), eff_base,
                  np.asarray(plog_ptr, np.int32), np.asarray(plog, np.float64).reshape(-1, 4))
