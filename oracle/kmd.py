"""Oracle-side mechanism reader — TEST INFRASTRUCTURE ONLY.

Builds the reference's KineticModelData tables from an NCAR-format YAML file
with PyYAML (BaseLoader, so every scalar stays a string and a species named
``NO`` is not turned into a boolean).  It restates
``KineticModelData::initChemNCAR`` (src/core/TChem_KineticModelData.cpp:80-1017)
independently of the product's C++ loader, so the two can be compared table by
table, and feeds the C oracle (oracle/oracle.c).

Only the tables the gas-phase box model (KForwardAerosol + ReactionRatesAerosol)
uses are built: species, stoichiometry, ARRHENIUS / TROE / CMAQ_H2O2
parameters, EMISSION sources and convExponent.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import yaml

CONV_PPM = 7.24296358205307e10  # src/core/TChem_Util.hpp:183
KBOLT = 1.3806504e-23  # src/core/TChem_Util.hpp:195
LENGTHOFSPECNAME = 18  # src/core/TChem_Util.hpp:133


@dataclass
class OracleKMD:
    nSpec: int
    nConstSpec: int
    nReac: int
    maxSpecInReac: int
    names: list
    reacNreac: np.ndarray
    reacNprod: np.ndarray
    reacSidx: np.ndarray  # [nReac][maxSpecInReac]
    reacNuki: np.ndarray
    convExponent: np.ndarray
    arrIdx: np.ndarray
    arrPar: np.ndarray  # [nArr][5] A B C D E
    troeIdx: np.ndarray
    troePar: np.ndarray  # [nTroe][8]
    cmaqIdx: np.ndarray
    cmaqPar: np.ndarray  # [nCmaq][6]
    emisSpecies: np.ndarray
    emisRate: np.ndarray
    photoReac: np.ndarray
    types: list = field(default_factory=list)

    @property
    def m(self) -> int:
        return self.nSpec - self.nConstSpec


def _f(v) -> float:
    return float(v)


def load(path: str) -> OracleKMD:
    with open(path) as fh:
        root = yaml.load(fh, Loader=yaml.BaseLoader)
    if "NCAR-version" not in root:
        raise ValueError("not an NCAR mechanism file")
    species = root["species"]
    const = root.get("constant_species") or []
    reactions = root.get("reactions") or []

    # species: active first, constant appended; first name wins (.cpp:107-131)
    index: dict = {}
    names = []
    for i, sp in enumerate(list(species) + list(const)):
        nm = sp["name"].strip()
        index.setdefault(nm, i)
        names.append(nm[:LENGTHOFSPECNAME])
    if "M" not in index:
        raise ValueError("M species is required")
    nSpec = len(species) + len(const)
    nR = len(reactions)

    nreac = np.zeros(nR, np.int32)
    nprod = np.zeros(nR, np.int32)
    rinfo, pinfo, types = [], [], []
    maxspec = 0
    for r, rx in enumerate(reactions):
        reac = rx["reactants"] or {}
        prod = rx["products"] or {}
        nreac[r] = len(reac)
        nprod[r] = len(prod)
        rd = {}
        for k, v in reac.items():
            if k not in index:
                raise ValueError(f"unknown reactant {k} in reaction {r}")
            rd.setdefault(index[k], _f(v))
        pd = {}
        for k, v in prod.items():
            if k in index:
                pd.setdefault(index[k], _f(v))
        rinfo.append(dict(sorted(rd.items())))
        pinfo.append(dict(sorted(pd.items())))
        maxspec = max(maxspec, len(reac), len(prod))
        types.append(rx["type"])
    maxspec *= 2

    sidx = np.zeros((nR, maxspec), np.int32)
    nuki = np.zeros((nR, maxspec), np.float64)
    for r in range(nR):
        for c, (s, v) in enumerate(rinfo[r].items()):
            sidx[r, c] = s
            nuki[r, c] = -v
        for c, (s, v) in enumerate(pinfo[r].items()):
            sidx[r, maxspec // 2 + c] = s
            nuki[r, maxspec // 2 + c] = v

    arr_idx, arr_par, troe_idx, troe_par, cmaq_idx, cmaq_par, photo = [], [], [], [], [], [], []
    for r, rx in enumerate(reactions):
        co = rx.get("coefficients") or {}
        per_min = rx.get("time_unit") == "min"
        t = rx["type"]
        if t == "ARRHENIUS":
            A = _f(co["A"]) / (60.0 if per_min else 1.0) if "A" in co else 1.0
            D = _f(co.get("D", 300.0))
            B = _f(co.get("B", 0.0))
            if "Ea" in co:
                if "C" in co:
                    raise ValueError("Ea and C both present")
                C = -_f(co["Ea"]) / KBOLT
            else:
                C = _f(co.get("C", 0.0))
            E = _f(co.get("E", 0.0))
            arr_idx.append(r)
            arr_par.append([A, B, C, D, E])
            if str(rx.get("note", "")).startswith("PHOTOLYSIS"):
                photo.append(r)
        elif t == "TROE":
            A = _f(co["k0_A"]) / (60.0 if per_min else 1.0) if "k0_A" in co else 1.0
            kinf = _f(co["kinf_A"]) / (60.0 if per_min else 1.0) if "kinf_A" in co else 1.0
            troe_idx.append(r)
            troe_par.append([A * 1e6, _f(co.get("k0_B", 0.0)), _f(co.get("k0_C", 0.0)), kinf,
                             _f(co.get("kinf_B", 0.0)), _f(co.get("kinf_C", 0.0)),
                             _f(co.get("Fc", 0.6)), _f(co.get("N", 1.0))])
        elif t == "CMAQ_H2O2":
            k1 = _f(co["k1_A"]) / (60.0 if per_min else 1.0) if "k1_A" in co else 1.0
            k2 = _f(co["k2_A"]) / (60.0 if per_min else 1.0) if "k2_A" in co else 1.0
            cmaq_idx.append(r)
            cmaq_par.append([k1, _f(co.get("k1_B", 0.0)), _f(co.get("k1_C", 0.0)), k2 * 1e6,
                             _f(co.get("k2_B", 0.0)), _f(co.get("k2_C", 0.0))])
        elif t in ("JPL", "R_JPL_ARRHENIUS", "PHOTO_RATE"):
            pass  # E3SM kinetics: not evaluated by the gas-phase box model
        else:
            raise ValueError(f"unknown reaction type {t}")

    emis_sp, emis_rate = [], []
    for src in root.get("sources") or []:
        if src["type"] != "EMISSION":
            raise ValueError("unknown source type")
        emis_sp.append(index[src["species"]])
        emis_rate.append(_f(src["coefficients"]["emission_rate"]))

    conv = np.array([sum(abs(int(nuki[r, j])) for j in range(nreac[r])) - 1 for r in range(nR)], np.int32)

    def a(x, dt, shape=None):
        arr = np.array(x, dtype=dt)
        if shape is not None and arr.size == 0:
            arr = arr.reshape(shape)
        return np.ascontiguousarray(arr)

    return OracleKMD(
        nSpec=nSpec, nConstSpec=len(const), nReac=nR, maxSpecInReac=maxspec, names=names,
        reacNreac=nreac, reacNprod=nprod, reacSidx=sidx, reacNuki=nuki, convExponent=conv,
        arrIdx=a(arr_idx, np.int32), arrPar=a(arr_par, np.float64, (0, 5)),
        troeIdx=a(troe_idx, np.int32), troePar=a(troe_par, np.float64, (0, 8)),
        cmaqIdx=a(cmaq_idx, np.int32), cmaqPar=a(cmaq_par, np.float64, (0, 6)),
        emisSpecies=a(emis_sp, np.int32), emisRate=a(emis_rate, np.float64),
        photoReac=a(photo, np.int32), types=types,
    )


def elimination_order(k: OracleKMD) -> np.ndarray:
    """Fill-reducing order of the active species for the dense Newton solves:
    greedy minimum Markowitz count on the structure of I - J (entry (i, s)
    where species i has a nonzero net coefficient in a reaction with active
    reactant s; the diagonal always).  Each step takes the remaining species
    with the smallest (row count - 1) * (column count - 1) over the remaining
    species (ties: lowest index) and adds its fill.  The product's mechanism
    compiler uses the same rule (tests/test_oracle.py checks the two orders
    agree); the order only changes the rounding of the linear solves."""
    m = k.m
    half = k.maxSpecInReac // 2
    pat = np.eye(m, dtype=bool)
    for r in range(k.nReac):
        net = {}
        for j in range(k.reacNreac[r]):
            s = int(k.reacSidx[r, j])
            if s < m:
                net[s] = net.get(s, 0.0) + float(k.reacNuki[r, j])
        for j in range(k.reacNprod[r]):
            s, nu = int(k.reacSidx[r, half + j]), float(k.reacNuki[r, half + j])
            if s < m and nu != 0.0:
                net[s] = net.get(s, 0.0) + nu
        rows = [e for e, v in net.items() if v != 0.0]
        for j in range(k.reacNreac[r]):
            s = int(k.reacSidx[r, j])
            if s < m:
                pat[rows, s] = True
    done = np.zeros(m, bool)
    order = []
    for _ in range(m):
        live = ~done
        rc = pat[:, live].sum(axis=1).astype(np.int64)
        cc = pat[live, :].sum(axis=0).astype(np.int64)
        score = np.where(live, (rc - 1) * (cc - 1), np.iinfo(np.int64).max)
        b = int(np.argmin(score))  # first minimum: lowest index on ties
        order.append(b)
        done[b] = True
        live = ~done
        rows = np.nonzero(live & pat[:, b])[0]
        cols = np.nonzero(live & pat[b, :])[0]
        if rows.size and cols.size:
            pat[np.ix_(rows, cols)] = True
    return np.array(order, np.int32)


def load_scenario(path: str, k: OracleKMD) -> np.ndarray:
    """State rows [nBatch][nSpec+3] from a scenario/mechanism YAML, restating
    AtmChemistry::setScenarioConditions (src/core/TChem_AtmosphericChemistry.cpp:429-504)."""
    with open(path) as fh:
        root = yaml.load(fh, Loader=yaml.BaseLoader)
    env = root["environmental_conditions"]
    pres = [float(v) for v in env["pressure"]["initial_value"]]
    temp = [float(v) for v in env["temperature"]["initial_value"]]
    init = root["initial_state"]
    n = len(pres)
    state = np.zeros((n, k.nSpec + 3))
    for name, spec in init.items():
        idx = next(i for i, nm in enumerate(k.names) if nm[:LENGTHOFSPECNAME] == name[:LENGTHOFSPECNAME])
        vals = [float(v) for v in spec["initial_value"]]
        units = spec["units"]
        for i in range(n):
            conv = CONV_PPM * pres[i] / temp[i]
            if units == "molecules m-3":
                v = vals[i] / conv
            elif units == "molecules cm-3":
                v = vals[i] / conv / 1e6
            elif units == "mol m-3":
                v = vals[i]
            else:
                raise ValueError(f"unit type does not exist {units}")
            state[i, idx + 3] = v
    state[:, 1] = pres
    state[:, 2] = temp
    return state
