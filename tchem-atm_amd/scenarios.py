"""Seeded synthetic cell states for the CB05 box-model workloads.

Every random number is keyed by (seed, cell, slot): slot s of cell c is the
Philox4x32-10 block of counter (s, c, 0, 0) under key (seed, 0), turned into
a double in [0, 1) from its first two words (27 + 26 bits).  A cell's state
therefore depends only on the seed and its own index, so any subset of a
workload (the first cells, a strided sample, a list of indices) holds exactly
the rows the full workload holds.  Slots per cell: 0 pressure, 1 temperature,
2 .. 2+m-1 active-species mixing ratios (ppb, log-uniform, mechanism order),
then nPhoto photolysis rates, then the night flag.

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


_PHILOX_M0, _PHILOX_M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_PHILOX_W0, _PHILOX_W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
_MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32(c0, c1, c2, c3, k0, k1, rounds=10):
    """Philox4x32 (Salmon et al., SC'11) on uint64 arrays holding 32-bit
    words; returns the four output words."""
    c0, c1, c2, c3 = (np.asarray(x, np.uint64) & _MASK32 for x in (c0, c1, c2, c3))
    k0, k1 = np.uint64(k0) & _MASK32, np.uint64(k1) & _MASK32
    for _ in range(rounds):
        p0 = _PHILOX_M0 * c0
        p1 = _PHILOX_M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK32
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
        k0 = (k0 + _PHILOX_W0) & _MASK32
        k1 = (k1 + _PHILOX_W1) & _MASK32
    return c0, c1, c2, c3


def keyed_uniform(seed: int, cells, slots) -> np.ndarray:
    """U[0,1) doubles [len(cells)][len(slots)] keyed by (seed, cell, slot)."""
    cells = np.asarray(cells, np.uint64).reshape(-1, 1)
    slots = np.asarray(slots, np.uint64).reshape(1, -1)
    w0, w1, _, _ = philox4x32(np.broadcast_to(slots, (cells.shape[0], slots.shape[1])), cells, 0, 0, seed, 0)
    return ((w0 >> np.uint64(5)).astype(np.float64) * 67108864.0 + (w1 >> np.uint64(6)).astype(np.float64)) \
        / 9007199254740992.0


def _cell_index(n: int, cells) -> np.ndarray:
    return np.arange(n, dtype=np.int64) if cells is None else np.asarray(cells, np.int64)


def cb05_cells(n: int, seed: int, m: int, n_spec: int, n_photo: int, base_state_row: np.ndarray,
               t_range=(270.0, 310.0), night_fraction: float = 0.0, cells=None) -> tuple[np.ndarray, np.ndarray]:
    """State rows and photolysis rates of cells 0..n-1 (or of the given cell
    indices): p ~ U[5e4, 1.013e5] Pa, T ~ U[t_range] K, active species
    10^U[-3, 2] ppb, photolysis U[0, 1e-3] s^-1, all photolysis 0 where the
    night flag draw is below night_fraction."""
    idx = _cell_index(n, cells)
    u = keyed_uniform(seed, idx, np.arange(3 + m + n_photo))
    p = 5.0e4 + (1.013e5 - 5.0e4) * u[:, 0]
    T = t_range[0] + (t_range[1] - t_range[0]) * u[:, 1]
    ppb = 10.0 ** (-3.0 + 5.0 * u[:, 2:2 + m])
    photo = 1.0e-3 * u[:, 2 + m:2 + m + n_photo]
    if night_fraction > 0.0:
        night = u[:, 2 + m + n_photo] < night_fraction
        photo[night, :] = 0.0
    state = np.zeros((idx.size, n_spec + 3))
    state[:, 1] = p
    state[:, 2] = T
    state[:, 3:3 + m] = ppb * 1e-3
    state[:, 3 + m:3 + n_spec] = _const_species(base_state_row, m, n_spec)
    return state, photo


def night_flags(n: int, seed: int, m: int, n_photo: int, night_fraction: float = 0.5, cells=None) -> np.ndarray:
    """The night flags cb05_cells draws for these cells."""
    idx = _cell_index(n, cells)
    return keyed_uniform(seed, idx, [2 + m + n_photo])[:, 0] < night_fraction


def config0(kmd, base_state_row, n: int = 1024, seed: int = 0, cells=None) -> Workload:
    """1024 cells, T ~ U[270,310] K, 3600 s, dtmax 60 s, atol 1e-12, rtol 1e-6
    (cells: a subset of cell indices instead of 0..n-1)."""
    m = kmd.nSpec - kmd.nConstSpec
    st, ph = cb05_cells(n, seed, m, kmd.nSpec, kmd.nPhoto, base_state_row, cells=cells)
    return Workload("cb05_config0", st, ph, 0.0, 3600.0, 1e-20, 1e-20, 60.0, 1e-12, 1e-6)


def config1(kmd, base_state_row, n: int = 131072, seed: int = 1, cells=None) -> Workload:
    """131072 cells (1820 columns x 72 levels), T ~ U[220,320] K, half the
    cells at night (cells: a subset of cell indices instead of 0..n-1)."""
    m = kmd.nSpec - kmd.nConstSpec
    st, ph = cb05_cells(n, seed, m, kmd.nSpec, kmd.nPhoto, base_state_row, t_range=(220.0, 320.0),
                        night_fraction=0.5, cells=cells)
    return Workload("cb05_config1", st, ph, 0.0, 3600.0, 1e-20, 1e-20, 60.0, 1e-12, 1e-6)


def aerosol_soa(kmd, base_state_row, n_aero: int, n_particles: int, n: int = 16384, seed: int = 4):
    """CB05CL_AE5 with the two-product SOA scheme and single-particle
    SIMPOL.1 partitioning: n cells with config-0 draws of pressure,
    temperature and gas mixing ratios (Philox keyed by `seed`), then
    particles per m3 N ~ 10^U[6,9] and primary organic aerosol (aerosol
    species 0) 10^U[-9,-7] kg m-3 shared by the particles; other particle
    species start at zero.  3600 s, dtmax 60 s, time tolerances (1e-12, 1e-6).
    Returns the workload (state rows [n][nSpec + 3 + n_aero * n_particles])
    and number concentrations [n][n_particles]."""
    m = kmd.nSpec - kmd.nConstSpec
    gas, _ = cb05_cells(n, seed, m, kmd.nSpec, 0, base_state_row)
    g = np.random.Generator(np.random.Philox(key=seed + 1000))
    nc = np.repeat((10.0 ** g.uniform(6.0, 9.0, n))[:, None], n_particles, axis=1)
    poa = 10.0 ** g.uniform(-9.0, -7.0, n)
    state = np.zeros((n, kmd.nSpec + 3 + n_aero * n_particles))
    state[:, :kmd.nSpec + 3] = gas
    for ip in range(n_particles):
        state[:, kmd.nSpec + 3 + ip * n_aero] = poa / nc[:, ip]
    w = Workload("cb05_soa_simpol", state, np.zeros((n, 0)), 0.0, 3600.0, 1e-20, 1e-20, 60.0, 1e-12, 1e-6)
    return w, nc


def tol_arrays(w: Workload, m: int):
    tol_newton = np.array([w.atol_newton, w.rtol_newton])
    tol_time = np.tile([w.atol_time, w.rtol_time], (m, 1))
    return tol_newton, tol_time


def config4_batches(seed: int = 3, n_batches: int = 48, lo: int = 256, hi: int = 65536,
                    size_scale: float = 1.0) -> np.ndarray:
    """Cell counts of the uneven batches: log-uniform in [lo, hi], times
    size_scale (at least one cell)."""
    g = np.random.Generator(np.random.Philox(key=seed + 1000))
    sizes = 10.0 ** g.uniform(np.log10(lo), np.log10(hi), n_batches)
    return np.maximum(1, np.round(sizes * size_scale)).astype(np.int64)


def config4(kmd, base_state_row, seed: int = 3, n_batches: int = 48, size_scale: float = 1.0):
    """Uneven batches of the synthetic mechanism, states drawn as in config 1
    (T ~ U[220,320] K; the synthetic mechanism has no photolysis), all
    batches concatenated into one workload for one persistent launch.
    Returns (workload, batch sizes)."""
    sizes = config4_batches(seed, n_batches, size_scale=size_scale)
    m = kmd.nSpec - kmd.nConstSpec
    st, ph = cb05_cells(int(sizes.sum()), seed, m, kmd.nSpec, kmd.nPhoto, base_state_row, t_range=(220.0, 320.0),
                        night_fraction=0.5)
    return Workload("synthetic250_config4", st, ph, 0.0, 3600.0, 1e-20, 1e-20, 60.0, 1e-12, 1e-6), sizes


def synthetic_mechanism_yaml(path: str, seed: int = 3, n_species: int = 250, n_reactions: int = 750,
                             n_radicals: int = 12, window: int = 40, coupling: str = "uniform") -> str:
    """Writes a seeded synthetic mechanism in the NCAR YAML format the loader
    reads: n_species active species S000.. plus the constant third body M and
    n_reactions reactions with 1-2 reactants (integer stoichiometry, a
    repeated reactant becomes nu = 2) and 1-4 distinct products with yields
    ~ U[0.1, 1].  Rate parameters: first-order A ~ 10^U[-6,-1] s^-1,
    bimolecular A ~ 10^U[-16,-10] cm^3 molecule^-1 s^-1, C = -Ea/R with
    Ea/R ~ U[-500, 3000] K; 10 % of the reactions are TROE with k0_A and
    kinf_A drawn from the bimolecular range, Fc = 0.6, N = 1.

    coupling="uniform" (the default): reactants (two with probability 0.8)
    and products drawn uniformly from all species.  The LU factor of such a
    mechanism is about half dense in the fill-reducing order (30312 of
    62500 entries at the default sizes): the dense MFMA form's case.

    coupling="cascade" (a structured variant, as in atmospheric
    mechanisms): species 0..n_radicals-1 are radicals, the rest organics
    in generation order.  60 % of the reactions are organic + radical, 20 %
    organic photolysis-like first order, 20 % radical + radical; organic
    products come from the next `window` organics, radical products from
    the radicals.  The Jacobian is then a cascade plus dense radical rows
    and columns, which a fill-reducing order factors with little fill
    (6965 entries): the sparse form's case."""
    g = np.random.Generator(np.random.Philox(key=seed))
    names = [f"S{i:03d}" for i in range(n_species)]
    nr = n_radicals
    out = ["NCAR-version: v1.0",
           "environmental_conditions:",
           "  pressure:", "    evolving: 'false'", "    initial_value: [101325]", "    units: Pa",
           "  temperature:", "    evolving: 'false'", "    initial_value: [298.0]", "    units: K",
           "model_info:",
           "  chemistry_time_step:", "    units: min", "    value: 1.0",
           "  evolving_conditions:", "    exist: 'false'",
           "  output_time_step:", "    units: min", "    value: 1.0",
           "  simulation_length:", "    units: hr", "    value: 1.0",
           "initial_state:"]
    for nm in names:
        out += [f"  {nm}:", "    initial_value: [0.001]", "    units: mol m-3"]
    out += ["  M:", "    initial_value: [1000000.0]", "    units: mol m-3"]
    out += ["constant_species:", "- description: third body", "  name: M", "species:"]
    for i, nm in enumerate(names):
        out += [f"- description: synthetic {'radical' if i < nr else 'organic'}", f"  name: {nm}"]
    out += ["reactions:"]

    def later_organics(i, k):
        lo = max(i + 1, nr)
        hi = min(n_species, lo + window)
        if hi - lo < k:
            lo = max(nr, hi - max(k, window))
        return [int(x) for x in g.choice(np.arange(lo, hi), k, replace=False)]

    for r in range(n_reactions):
        u = g.uniform()
        nprod = 1 + int(g.integers(0, 4))
        if coupling == "uniform":
            reac = [int(x) for x in g.integers(0, n_species, 1 + int(u < 0.8))]
            prods = [int(x) for x in g.choice(n_species, nprod, replace=False)]
        elif u < 0.6:    # organic + radical
            org = int(g.integers(nr, n_species - 1))
            reac = [org, int(g.integers(0, nr))]
            n_org = max(1, nprod - int(g.integers(0, 2)))
            prods = later_organics(org, min(n_org, nprod))
            prods += [int(x) for x in g.choice(nr, nprod - len(prods), replace=False)]
        elif u < 0.8:  # organic first order
            org = int(g.integers(nr, n_species - 1))
            reac = [org]
            prods = later_organics(org, max(1, nprod - 1))
            prods += [int(x) for x in g.choice(nr, nprod - len(prods), replace=False)]
        else:          # radical + radical
            reac = [int(x) for x in g.integers(0, nr, 2)]
            prods = [int(x) for x in g.choice(np.arange(0, nr + window), nprod, replace=False)]
        nu = {}
        for x in reac:
            nu[x] = nu.get(x, 0) + 1
        prods = list(dict.fromkeys(prods))
        yields = g.uniform(0.1, 1.0, len(prods))
        out.append(f"- MUSICA_name: R{r + 1}")
        out.append("  reactants:")
        for x, v in nu.items():
            out.append(f"    {names[x]}: {v}")
        out.append("  products:")
        for x, y in zip(prods, yields):
            out.append(f"    {names[x]}: {float(y)!r}")
        ea_r = float(g.uniform(-500.0, 3000.0))
        bimolecular = sum(nu.values()) == 2
        if bimolecular and g.uniform() < 0.125:  # 10 % of all reactions (80 % are bimolecular)
            k0 = 10.0 ** float(g.uniform(-16.0, -10.0))
            kinf = 10.0 ** float(g.uniform(-16.0, -10.0))
            out += ["  type: TROE", "  coefficients:", f"    k0_A: {k0!r}", "    k0_B: 0.0", f"    k0_C: {-ea_r!r}",
                    f"    kinf_A: {kinf!r}", "    kinf_B: 0.0", f"    kinf_C: {-ea_r!r}", "    Fc: 0.6", "    N: 1.0"]
        else:
            A = 10.0 ** float(g.uniform(-16.0, -10.0) if bimolecular else g.uniform(-6.0, -1.0))
            out += ["  type: ARRHENIUS", "  coefficients:", f"    A: {A!r}", "    B: 0.0", f"    C: {-ea_r!r}"]
    text = "\n".join(out) + "\n"
    with open(path, "w") as f:
        f.write(text)
    return path
