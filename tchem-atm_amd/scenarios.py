"""Seeded synthetic cell states for the CB05 box-model workloads.

All draws use numpy's Philox4x32 counter-based generator keyed by the seed,
one array per quantity, in this order: pressure [n], temperature [n],
active-species mixing ratios [n][m] (ppb, log-uniform), photolysis rates
[n][nPhoto], and for the day/night workload the night flags [n].

State rows follow the reference's layout [density, pressure(Pa),
temperature(K), species...] (src/core/TChem_Util.hpp:635-676) in its ppm
convention (rate constants carry conv^(order-1) with conv = CONV_PPM p / T,
src/core/impl/TChem_Impl_KForwardAerosol.hpp:57), so a mixing ratio of x ppb
is stored as x * 1e-3.  Constant species take the mechanism file's
initial_state values (CH4, M, O2, H2O, H2, N2 for CB05CL_AE5).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

CONV_PPM = 7.24296358205307e10


@dataclass
class Workload:
    name: str
    state: np.ndarray          # [n][nSpec+3]
    photo: np.ndarray          # [n][nPhoto]
    tbeg: float
    tend: float
    dt_in: float
    dtmin: float
    dtmax: float
    atol_time: float
    rtol_time: float
    atol_newton: float = 1e-18
    rtol_newton: float = 1e-8
    max_newton: int = 20
    iters_per_interval: int = 10
    jacobian_interval: int = 1

    @property
    def n(self) -> int:
        return self.state.shape[0]

    @property
    def cell_hours(self) -> float:
        return self.n * (self.tend - self.tbeg) / 3600.0


def _const_species(base_state_row: np.ndarray, m: int, n_spec: int) -> np.ndarray:
    return base_state_row[3 + m: 3 + n_spec]


def cb05_cells(n: int, seed: int, m: int, n_spec: int, n_photo: int, base_state_row: np.ndarray,
               t_range=(270.0, 310.0), night_fraction: float = 0.0) -> tuple[np.ndarray, np.ndarray]:
    """State rows and photolysis rates for n independent cells."""
    g = np.random.Generator(np.random.Philox(key=seed))
    p = g.uniform(5.0e4, 1.013e5, n)
    T = g.uniform(t_range[0], t_range[1], n)
    ppb = 10.0 ** g.uniform(-3.0, 2.0, (n, m))
    photo = g.uniform(0.0, 1.0e-3, (n, n_photo))
    if night_fraction > 0.0:
        night = g.uniform(0.0, 1.0, n) < night_fraction
        photo[night, :] = 0.0
    state = np.zeros((n, n_spec + 3))
    state[:, 1] = p
    state[:, 2] = T
    state[:, 3:3 + m] = ppb * 1e-3
    state[:, 3 + m:3 + n_spec] = _const_species(base_state_row, m, n_spec)
    return state, photo


def config0(kmd, base_state_row, n: int = 1024, seed: int = 0) -> Workload:
    """1024 cells, T ~ U[270,310] K, 3600 s, dtmax 60 s, atol 1e-12, rtol 1e-6."""
    m = kmd.nSpec - kmd.nConstSpec
    st, ph = cb05_cells(n, seed, m, kmd.nSpec, kmd.nPhoto, base_state_row)
    return Workload("cb05_config0", st, ph, 0.0, 3600.0, 1e-20, 1e-20, 60.0, 1e-12, 1e-6)


def config1(kmd, base_state_row, n: int = 131072, seed: int = 1) -> Workload:
    """131072 cells (1820 columns x 72 levels), T ~ U[220,320] K, half the cells at night."""
    m = kmd.nSpec - kmd.nConstSpec
    st, ph = cb05_cells(n, seed, m, kmd.nSpec, kmd.nPhoto, base_state_row, t_range=(220.0, 320.0),
                        night_fraction=0.5)
    return Workload("cb05_config1", st, ph, 0.0, 3600.0, 1e-20, 1e-20, 60.0, 1e-12, 1e-6)


def tol_arrays(w: Workload, m: int):
    tol_newton = np.array([w.atol_newton, w.rtol_newton])
    tol_time = np.tile([w.atol_time, w.rtol_time], (m, 1))
    return tol_newton, tol_time
