"""Oracle-side aerosol model reader — TEST INFRASTRUCTURE ONLY.

Restates ``AerosolModelData::initChem`` and ``scenarioConditionParticles``
(src/core/TChem_AerosolModelData.cpp:60-249, 263-300) with PyYAML, independently
of the product's C++ loader, and feeds the C oracle's single-particle SIMPOL.1
functions (oracle/oracle.c).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import yaml


@dataclass
class OracleAMD:
    nSpec: int
    nParticles: int
    names: list
    mw: np.ndarray
    density: np.ndarray
    kappa: np.ndarray
    gas: np.ndarray   # [nSimpol] kinetic-model species index
    aero: np.ndarray  # [nSimpol] aerosol species index within a particle
    par: np.ndarray   # [nSimpol][8] B1 B2 B3 B4 diffusion mw Nstar computeAlpha

    @property
    def nSimpol(self) -> int:
        return len(self.gas)


def load(path: str, gas_names: list) -> OracleAMD:
    """Aerosol tables of `path`; gas_names = the kinetic model's species names."""
    with open(path) as f:
        root = yaml.load(f, Loader=yaml.BaseLoader)
    items = root["NCAR-version"]
    names, mw, dens, kap = [], [], [], []
    gas_info = {}
    n_part = -1
    for it in items:
        if not isinstance(it, dict) or "type" not in it:
            continue
        if it["type"] == "CHEM_SPEC":
            if "phase" in it:
                if it["phase"] == "AEROSOL":
                    names.append(it["name"])
                    mw.append(float(it["molecular weight [kg mol-1]"]))
                    dens.append(float(it["density [kg m-3]"]))
                    kap.append(float(it["kappa"]))
            else:
                gas_info.setdefault(it["name"], it)
        elif it["type"] == "AERO_REP_SINGLE_PARTICLE":
            n_part = int(it["maximum computational particles"])
    gas, aero, par = [], [], []
    for it in items:
        if not isinstance(it, dict) or it.get("type") != "MECHANISM":
            continue
        for r in it.get("reactions", []):
            if r["type"] != "SIMPOL_PHASE_TRANSFER":
                continue
            g = r["gas-phase species"]
            info = gas_info[g]
            nstar = info.get("N star")
            gas.append(gas_names.index(g))
            aero.append(names.index(r["aerosol-phase species"]))
            par.append([float(b) for b in r["B"]] + [
                float(info["diffusion coeff [m2 s-1]"]), float(info["molecular weight [kg mol-1]"]),
                999.0 if nstar is None else float(nstar), 0.0 if nstar is None else 1.0])
    return OracleAMD(len(names), n_part, names, np.array(mw), np.array(dens), np.array(kap),
                     np.array(gas, np.int32), np.array(aero, np.int32),
                     np.array(par, np.float64).reshape(-1, 8))


def load_particles(path: str, a: OracleAMD, n_gas_species: int, state: np.ndarray):
    """Particle block of a scenario file: returns number concentrations
    [nBatch][nParticles] and writes initial_value / N into `state` rows."""
    with open(path) as f:
        root = yaml.load(f, Loader=yaml.BaseLoader)
    p = root["particles"]
    n = state.shape[0]
    nc = np.array([[float(p["num_concentration"]["initial_value"][i])] * a.nParticles for i in range(n)])
    base = 3 + n_gas_species
    for name, spec in p.get("initial_state", {}).items():
        idx = a.names.index(name)
        for i in range(n):
            v = float(spec["initial_value"][i])
            for ip in range(a.nParticles):
                state[i, base + a.nSpec * ip + idx] = v / nc[i, ip]
    return nc
