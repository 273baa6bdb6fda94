"""Pins of the CPU oracle (oracle/) against the reference's goldens and against
what the mathematics fixes.  CPU only.

Each test names what it pins:
  * CB05CL_AE5 golden (src/unit-tests/CB05CL_AE5): source term, rate
    constants with conv factors, TR stage/BDF2 stage at tiny steps, doubling
    step growth from dtmin;
  * TChem's Emissions.dat output: emission source term, dtmax clamp;
  * CAMP analytic solutions of the Arrhenius/Troe/CMAQ_H2O2 test mechanisms
    (an independent implementation): every rate-constant form and unit rule;
  * closed forms: TR-BDF2 one-step amplification on y' = -k y, LU against
    scipy, analytic Jacobian against central differences.
"""
import numpy as np
import pytest
import scipy.linalg

from conftest import golden


def _tadv(oracle_mod, n, tend, dt, dtmin, dtmax, iters, newton=20, jint=1, tbeg=0.0):
    ta = np.zeros(n, oracle_mod.TADV_DTYPE)
    ta["tbeg"], ta["tend"], ta["dt"], ta["dtmin"], ta["dtmax"] = tbeg, tend, dt, dtmin, dtmax
    ta["max_num_newton_iterations"] = newton
    ta["num_time_iterations_per_interval"] = iters
    ta["jacobian_interval"] = jint
    return ta


def run_intervals(o, state, tn, tt, ta, n_intervals, photo=None, mode=0):
    """The reference driver loop (src/examples/TChem_AtmosphericChemistry.cpp:312-352);
    returns the rows it would write to the .dat file."""
    n = state.shape[0]
    t = np.full(n, ta["tbeg"][0])
    dt = np.full(n, ta["dtmax"][0])
    rows = [np.concatenate([np.stack([np.full(n, -1.0), t, dt], 1), state], 1)]
    for it in range(n_intervals):
        o.run_batch(tn, tt, ta, state, t, dt, photo=photo, jacobian_mode=mode)
        rows.append(np.concatenate([np.stack([np.full(n, float(it)), t, dt], 1), state], 1))
        ta["tbeg"] = t
        ta["dt"] = dt
        if t.mean() > ta["tend"][0] * 0.9999:
            break
    return np.array(rows)


def compare_tchem_norms(test, ref):
    """src/unit-tests/scripts/compare_tchem.py:54-82: per-column L1/L2/Linf of
    the difference over the column L2 of the reference (floor 1e-23)."""
    diff = test - ref
    l2 = np.linalg.norm(ref, 2, 0)
    mx = np.where(l2 > 1e-23, l2, 1.0)
    return max((np.linalg.norm(diff, o, 0) / mx).max() for o in (1, 2, np.inf))


def test_cb05cl_ae5_golden(oracle_mod, cb05_oracle):
    o = cb05_oracle
    state = oracle_mod.kmd.load_scenario(golden("cb05cl_ae5.yaml"), o.k)
    m = o.m
    ta = _tadv(oracle_mod, 1, 600.0, 1e-20, 1e-20, 10.0, 10)
    rows = run_intervals(o, state, np.array([1e-18, 1e-8]), np.tile([1e-12, 1e-4], (m, 1)), ta, 2)
    ref = np.load(golden("cb05cl_ae5_full_gas.npy"))
    test = rows[:, 0, :]
    assert test.shape == ref.shape
    assert compare_tchem_norms(test, ref) < 1e-8  # the reference's ctest threshold
    # stricter: every species above the comparator's floor to 1e-9 relative
    big = np.linalg.norm(ref, 2, 0) > 1e-23
    rel = np.abs(test - ref)[:, big] / np.maximum(np.abs(ref[:, big]), 1e-300)
    assert rel[1:].max() < 1e-9
    # time column: doubling from dtmin; dt_out is the next proposal
    np.testing.assert_allclose(test[1:, 1], [1e-20 * (2**10 - 1), 1e-20 * (2**20 - 1)], rtol=1e-15)
    np.testing.assert_allclose(test[1:, 2], [1e-20 * 2**10, 1e-20 * 2**20], rtol=1e-15)


def test_emission_output_of_the_reference(oracle_mod):
    """src/examples/runs/atmospheric_chemistry/Emission/Emissions.dat is
    TChem's own output for config_emissions.yaml (runModel.sh settings)."""
    o = oracle_mod.Oracle(golden("emission.yaml"))
    state = oracle_mod.kmd.load_scenario(golden("emission.yaml"), o.k)
    ta = _tadv(oracle_mod, 1, 100.0, 1e-20, 1e-20, 0.1, 10)
    ref = np.loadtxt(golden("emission_Emissions.dat"), skiprows=1)
    rows = run_intervals(o, state, np.array([1e-18, 1e-8]), np.tile([1e-12, 1e-6], (o.m, 1)), ta, 200)[:, 0, :]
    n = min(len(rows), len(ref))
    assert n == len(ref)
    np.testing.assert_allclose(rows[:n, 1], ref[:n, 1], rtol=1e-10, atol=1e-25)  # t (10 digits printed)
    # dt: the file's last row (dt = 0 after the clamped final step) comes from an
    # older end rule; the reference's active UCI golden (uci_col/full_gas.dat:
    # dt_out = 2 x the clamped step) fixes the current one, tested below
    np.testing.assert_allclose(rows[:n - 1, 2], ref[:n - 1, 2], rtol=1e-10)
    np.testing.assert_allclose(rows[:n, 6:8], ref[:n, 6:8], rtol=1e-10)  # A, B (the file predates the M column)


def _chain(ka, kb, t):
    A = np.exp(-ka * t)
    B = ka / (kb - ka) * (np.exp(-ka * t) - np.exp(-kb * t))
    return A, B, 1.0 - A - B


@pytest.mark.parametrize("mech,results", [("troe.yaml", "troe_results.txt"),
                                          ("cmaq_h2o2.yaml", "CMAQ_H2O2_results.txt"),
                                          ("arrhenius.yaml", "rxn_arrhenius_results.txt")])
def test_rate_constants_against_camp_analytic_solution(oracle_mod, mech, results):
    """CAMP's analytic columns (A_true, B_true, C_true) of A -> B -> C fix the
    first-order rate constants; ours come from the restated KForwardAerosol."""
    o = oracle_mod.Oracle(golden(mech))
    st = oracle_mod.kmd.load_scenario(golden(mech), o.k)
    kf = o.kforward(st[0, 2], st[0, 1])
    if mech == "arrhenius.yaml":  # A + D -> B, A -> B, B -> C + D with D constant
        ka, kb = kf[0] * st[0, 3 + o.k.names.index("D")] + kf[1], kf[2]
    else:
        ka, kb = kf[0], kf[1]
    R = np.loadtxt(golden(results))
    A, B, C = _chain(ka, kb, R[:, 0])
    # A_true = exp(-ka t) fixes ka directly
    np.testing.assert_allclose(ka, -np.log(R[1, 2]) / R[1, 0], rtol=1e-7)
    np.testing.assert_allclose(A[1:4], R[1:4, 2], rtol=1e-6)
    np.testing.assert_allclose(B[1:], R[1:, 4], rtol=1e-7)
    np.testing.assert_allclose(C[1:], R[1:, 6], rtol=1e-7)


def test_integrator_against_camp_analytic_troe(oracle_mod):
    """runModelTroe.sh settings (tol-time 1e-8, dtmax 0.1, 50 iterations per
    interval) against the analytic A -> B -> C solution."""
    o = oracle_mod.Oracle(golden("troe.yaml"))
    st = oracle_mod.kmd.load_scenario(golden("troe.yaml"), o.k)
    R = np.loadtxt(golden("troe_results.txt"))
    tn, tt = np.array([1e-18, 1e-8]), np.tile([1e-12, 1e-8], (o.m, 1))
    for tend in (1.0, 5.0, 20.0):
        s = st.copy()
        ta = _tadv(oracle_mod, 1, tend, 1e-20, 1e-20, 0.1, 50)
        run_intervals(o, s, tn, tt, ta, 200)
        row = R[int(tend)]
        got = s[0, 3:6]
        np.testing.assert_allclose(got, row[[2, 4, 6]], rtol=1e-5, atol=1e-9)  # global error ~ steps x tol


def test_trbdf2_one_step_amplification(oracle_mod, tmp_path):
    """For y' = -k y one TR-BDF2 step multiplies y by
    R(z) = (a_g (1 + g z/2)/(1 - g z/2) - a_n) / (1 - c z), z = -k h."""
    k = 0.7
    f = tmp_path / "decay.yaml"
    f.write_text("NCAR-version: v1.0\nspecies:\n- name: A\n- name: B\nconstant_species:\n- name: M\n"
                 f"reactions:\n- reactants:\n    A: 1\n  products:\n    B: 1\n  type: ARRHENIUS\n"
                 f"  coefficients:\n    A: {k}\n")
    o = oracle_mod.Oracle(str(f))
    g, c_tr, a_g, a_n, c_bdf, _ = oracle_mod.trbdf2_constants()
    assert abs(g - (2 - np.sqrt(2))) < 1e-16 and abs(c_bdf - c_tr) < 4e-16  # both stages share I - (g/2) h J
    for h in (1e-3, 0.1, 1.0, 10.0):
        st = np.array([[0.0, 1e5, 300.0, 1.0, 0.0, 0.0]])
        ta = _tadv(oracle_mod, 1, h, h, h, h, 1)
        t, dt = np.zeros(1), np.zeros(1)
        o.run_batch(np.array([1e-30, 1e-14]), np.tile([1.0, 1.0], (2, 1)), ta, st, t, dt)
        z = -k * h
        R = (a_g * (1 + g * z / 2) / (1 - g * z / 2) - a_n) / (1 - c_bdf * z)
        assert t[0] == h
        np.testing.assert_allclose(st[0, 3], R, rtol=1e-13)
        np.testing.assert_allclose(st[0, 3] + st[0, 4], 1.0, rtol=1e-14)  # mass balance


def test_analytic_jacobian_against_central_differences(cb05_oracle):
    o = cb05_oracle
    rng = np.random.default_rng(5)
    state = oracle_base_state(o)
    for trial in range(3):
        c = state.copy()
        c[:o.m] = 10 ** rng.uniform(-6, -1, o.m)
        kf = o.kforward(rng.uniform(250, 310), rng.uniform(5e4, 1e5), rng.uniform(0, 1e-3, len(o.k.photoReac)))
        J = o.jacobian(kf, c)
        Jfd = np.zeros_like(J)
        for j in range(o.m):
            # omega is at most quadratic in any single species (reactant powers
            # <= 2), so central differences are exact up to rounding: use a wide h
            h = 1e-2 * c[j]
            cp, cm = c.copy(), c.copy()
            cp[j] += h
            cm[j] -= h
            Jfd[:, j] = (o.omega(kf, cp) - o.omega(kf, cm)) / (2 * h)
        scale = np.abs(J).max(axis=1, keepdims=True) + 1e-300
        assert (np.abs(J - Jfd) / scale).max() < 1e-9
        # structural zeros agree
        assert np.all((np.abs(Jfd) > 1e-12 * scale) <= (J != 0))


def oracle_base_state(o):
    import oracle
    return oracle.kmd.load_scenario(golden("cb05cl_ae5.yaml"), o.k)[0, 3:]


def test_source_term_invariants(cb05_oracle):
    """omega of species no reaction touches is zero; linear in every
    first-order rate constant; constant species never change."""
    o = cb05_oracle
    c = oracle_base_state(o)
    kf = o.kforward(290.0, 9e4)
    w = o.omega(kf, c)
    w2 = o.omega(2.0 * kf, c)
    np.testing.assert_allclose(w2, 2.0 * w, rtol=1e-14, atol=0)
    assert w.shape == (o.m,)
    assert np.all(o.omega(np.zeros_like(kf), c) == 0.0)


def test_lu_against_scipy(oracle_mod, cb05_oracle):
    rng = np.random.default_rng(1)
    mats = [rng.standard_normal((n, n)) for n in (1, 5, 33, 74)]
    o = cb05_oracle
    kf = o.kforward(280.0, 8e4, np.full(len(o.k.photoReac), 5e-4))
    J = o.jacobian(kf, oracle_base_state(o))
    mats += [np.eye(o.m) - 0.29 * h * J for h in (1e-3, 1.0, 60.0)]
    for A in mats:
        LU, piv, info = oracle_mod.lu(A)
        assert info == 0
        lu_s, piv_s = scipy.linalg.lu_factor(A)
        np.testing.assert_array_equal(piv, piv_s)
        np.testing.assert_allclose(LU, lu_s, rtol=1e-12, atol=1e-14 * np.abs(A).max())
        b = rng.standard_normal(A.shape[0])
        x = oracle_mod.lu_solve(LU, piv, b)
        assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(A, np.inf) * np.linalg.norm(x) + 1e-13
    _, _, info = oracle_mod.lu(np.zeros((3, 3)))
    assert info == 1


def test_step_size_rules(oracle_mod):
    """Clamp to t_end before the step; dt_out is the next proposal (twice the
    clamped step when the error is small), capped by dtmax."""
    o = oracle_mod.Oracle(golden("emission.yaml"))
    st = oracle_mod.kmd.load_scenario(golden("emission.yaml"), o.k)
    tn, tt = np.array([1e-18, 1e-8]), np.tile([1e-12, 1e-6], (o.m, 1))
    ta = _tadv(oracle_mod, 1, 0.3, 0.25, 1e-20, 10.0, 10)
    t, dt = np.zeros(1), np.zeros(1)
    stats = np.zeros(1, oracle_mod.STATS_DTYPE)
    o.run_batch(tn, tt, ta, st, t, dt, stats=stats)
    assert t[0] == 0.3
    assert dt[0] == pytest.approx(2 * 0.05, rel=1e-12)  # steps 0.25 then 0.05 (clamped)
    assert stats["accepted"][0] == 2
    ta = _tadv(oracle_mod, 1, 100.0, 3.0, 1e-20, 4.0, 1)
    o.run_batch(tn, tt, ta, st, t, dt)
    assert dt[0] == 4.0  # 2 * 3 capped by dtmax


def test_error_rejection_and_newton_failure_paths(oracle_mod, cb05_oracle):
    o = cb05_oracle
    m = o.m
    base = oracle_mod.kmd.load_scenario(golden("cb05cl_ae5.yaml"), o.k)
    # a 60 s first step from the YAML state is rejected (dt -> h/4) before progress
    st = base.copy()
    ta = _tadv(oracle_mod, 1, 3600.0, 60.0, 1e-20, 60.0, 5)
    t, dt = np.zeros(1), np.zeros(1)
    stats = np.zeros(1, oracle_mod.STATS_DTYPE)
    o.run_batch(np.array([1e-18, 1e-8]), np.tile([1e-12, 1e-6], (m, 1)), ta, st, t, dt, stats=stats)
    assert stats["rejected"][0] >= 1 and stats["failed"][0] == 0
    # Newton limited to one iteration with a large minimum step: hard failure
    st = base.copy()
    ta = _tadv(oracle_mod, 1, 3600.0, 100.0, 100.0, 100.0, 5, newton=1)
    t, dt = np.zeros(1), np.zeros(1)
    stats = np.zeros(1, oracle_mod.STATS_DTYPE)
    o.run_batch(np.array([1e-30, 1e-14]), np.tile([1e-12, 1e-6], (m, 1)), ta, st, t, dt, stats=stats)
    assert stats["failed"][0] == 1 and stats["newton_failures"][0] == 4
    assert t[0] == 3600.0 and dt[0] == -1.0
    assert np.all(st[0, 3:3 + m] == 0.0)
    assert np.all(st[0, 3 + m:] == base[0, 3 + m:])  # constant species untouched
    # a failed cell (dt_in < 0) and a finished cell (t_out >= tend) are skipped
    t2, dt2 = np.array([0.0]), np.array([5.0])
    ta["dt"] = -1.0
    o.run_batch(np.array([1e-18, 1e-8]), np.tile([1e-12, 1e-6], (m, 1)), ta, st, t2, dt2)
    assert t2[0] == 0.0 and dt2[0] == 5.0


def test_reference_algorithm_mode_agrees(cb05_oracle, oracle_mod):
    """The CPU baseline form (forward-difference Jacobian, rate constants in every
    source-term call) converges to the same solution as the analytic form."""
    o = cb05_oracle
    m = o.m
    base = oracle_mod.kmd.load_scenario(golden("cb05cl_ae5.yaml"), o.k)
    out = []
    for mode in (0, 1):
        st = base.copy()
        ta = _tadv(oracle_mod, 1, 30.0, 1e-20, 1e-20, 10.0, 5000)
        t, dt = np.zeros(1), np.zeros(1)
        o.run_batch(np.array([1e-18, 1e-8]), np.tile([1e-12, 1e-6], (m, 1)), ta, st, t, dt, jacobian_mode=mode)
        assert t[0] == 30.0
        out.append(st[0, 3:3 + m])
    rel = np.abs(out[0] - out[1]) / np.maximum(np.abs(out[0]), 1e-12)
    assert rel.max() < 1e-4


@pytest.mark.parametrize("mech", ["cb05cl_ae5.yaml", "carbonbond05.yaml", "cb05_soa_gas.yaml"])
def test_elimination_order_matches_the_library(pkg, oracle_mod, mech):
    """The oracle factors the Newton systems in the fill-reducing order it
    computes itself (oracle/kmd.py: elimination_order); it is the order the
    library's mechanism compiler picks (queried on the host)."""
    k = oracle_mod.kmd_mod.load(golden(mech))
    order = oracle_mod.kmd_mod.elimination_order(k)
    assert sorted(order.tolist()) == list(range(k.m))
    kmd = pkg.KineticModelData(golden(mech))
    np.testing.assert_array_equal(order, kmd.eliminationOrder())


def test_elimination_order_matches_on_the_synthetic_mechanism(pkg, oracle_mod, tmp_path):
    import importlib
    S = importlib.import_module("tchem-atm_amd.scenarios")
    path = S.synthetic_mechanism_yaml(str(tmp_path / "synthetic250.yaml"), coupling="cascade")
    k = oracle_mod.kmd_mod.load(path)
    np.testing.assert_array_equal(oracle_mod.kmd_mod.elimination_order(k),
                                  pkg.KineticModelData(path).eliminationOrder())


def test_dense_lu_order_changes_only_rounding(cb05_oracle, oracle_mod):
    """The same Newton system factored in species order and in the
    elimination order: the solutions agree to rounding."""
    k = cb05_oracle.k
    rng = np.random.default_rng(5)
    c = np.concatenate([10.0 ** rng.uniform(-6, -1, k.m), np.ones(k.nConstSpec)])
    kf = cb05_oracle.kforward(290.0, 9e4, rng.uniform(0, 1e-3, len(k.photoReac)))
    J = cb05_oracle.jacobian(kf, c)
    A = np.eye(k.m) - 10.0 * J
    b = rng.standard_normal(k.m)
    order = cb05_oracle.lu_order
    LU, piv, info = oracle_mod.lu(A)
    x0 = oracle_mod.lu_solve(LU, piv, b)
    LUp, pivp, infop = oracle_mod.lu(A[np.ix_(order, order)])
    xp = np.empty(k.m)
    xp[order] = oracle_mod.lu_solve(LUp, pivp, b[order])
    assert info == 0 and infop == 0
    assert np.abs(x0 - xp).max() <= 1e-12 * np.abs(x0).max()


def test_carbonbond05_long_trajectory_golden(oracle_mod):
    """src/examples/runs/atmospheric_chemistry/CarbonBond05/full_gas.dat: one
    cell, 12 intervals of 100 TR-BDF2 iterations to 10 800 s (runModel.sh:
    dtmin 1e-20, dtmax 10 s, rtol 1e-6, Newton 1e-18 / 1e-8, 20 iterations;
    100 iterations per interval reproduce the file's interval times).

    The file comes from an older mechanism file (its species are in another
    order and M has other units), so it pins the step sequence: from the
    file's own initial row, 70 doublings from 1e-20 s, then the 10 s cap
    with no rejection, every interval ending where the file's does, and the
    last step clamped to 10 800 s.  Its species columns agree only to the
    older mechanism's difference (compare_tchem norm 0.17); the bar below
    guards against gross errors."""
    mech = golden("carbonbond05.yaml")
    o = oracle_mod.Oracle(mech)
    k = o.k
    path = golden("carbonbond05_full_gas.dat")
    names = open(path).readline().split()[6:]
    gold = np.loadtxt(path, skiprows=1)
    col = np.array([names.index(nm) for nm in k.names])
    G = np.concatenate([gold[:, :6], gold[:, 6:][:, col]], axis=1)
    st = np.ascontiguousarray(G[:1, 3:])
    tn, tt = np.array([1e-18, 1e-8]), np.tile([1e-12, 1e-6], (k.m, 1))
    ta = _tadv(oracle_mod, 1, 10800.0, 1e-20, 1e-20, 10.0, 100)
    t, dt = np.zeros(1), np.zeros(1)
    rows = []
    for it in range(12):
        s = np.zeros(1, oracle_mod.STATS_DTYPE)
        o.run_batch(tn, tt, ta, st, t, dt, stats=s)
        assert s["rejected"][0] == 0 and s["failed"][0] == 0
        rows.append(np.concatenate([[it, t[0], dt[0]], st[0]]))
        ta["tbeg"], ta["dt"] = t, dt
    rows = np.array(rows)
    np.testing.assert_allclose(rows[:, 1], G[1:, 1], rtol=1e-9)  # the file prints 11 digits
    assert np.all(rows[:-1, 2] == 10.0) and rows[0, 1] == pytest.approx(1e-20 * (2.0 ** 70 - 1) + 300.0, rel=1e-15)
    assert rows[-1, 1] == 10800.0
    assert compare_tchem_norms(rows[:, 3:], G[1:, 3:]) < 0.25


_TINY_MECH = """NCAR-version: v1.0
species:
- name: A
- name: B
- name: C
- name: D
constant_species:
- name: M
sources:
- coefficients:
    emission_rate: 0.25
  species: B
  type: EMISSION
reactions:
- reactants: {A: 1, B: 1}
  products: {C: 1}
  type: ARRHENIUS
  coefficients: {A: 1.0}
- reactants: {A: 2}
  products: {B: 1, D: 0.5}
  type: ARRHENIUS
  coefficients: {A: 1.0}
- reactants: {C: 1, M: 1}
  products: {A: 1, M: 1}
  type: ARRHENIUS
  coefficients: {A: 1.0}
- reactants: {D: 1}
  products: {D: 1, B: 2.0}
  type: ARRHENIUS
  coefficients: {A: 1.0}
- reactants: {A: 1, B: 1, C: 1}
  products: {C: 0.3, A: 1.7}
  type: ARRHENIUS
  coefficients: {A: 1.0}
"""
# (reactant powers, product yields) per reaction over species A B C D M
_TINY_R = [({0: 1, 1: 1}, {2: 1.0}), ({0: 2}, {1: 1.0, 3: 0.5}), ({2: 1, 4: 1}, {0: 1.0, 4: 1.0}),
           ({3: 1}, {3: 1.0, 1: 2.0}), ({0: 1, 1: 1, 2: 1}, {2: 0.3, 0: 1.7})]


def test_source_term_and_jacobian_by_brute_force(oracle_mod, tmp_path):
    """A hand-written mechanism with a second-order reactant (2 A), a
    constant species as reactant and product, a species that is both
    reactant and product (catalyst, and net yields 0.3 - 1 and 1.7 - 1), a
    three-body reaction and an emission: the oracle's source term and
    analytic Jacobian against the mass-action formulas evaluated term by term
    in extended precision (long double), at random concentrations and rate
    constants."""
    f = tmp_path / "tiny.yaml"
    f.write_text(_TINY_MECH)
    o = oracle_mod.Oracle(str(f))
    assert o.k.nSpec == 5 and o.m == 4
    rng = np.random.default_rng(11)
    ld = np.longdouble
    for _ in range(50):
        c = 10.0 ** rng.uniform(-3, 1, 5)
        kf = 10.0 ** rng.uniform(-2, 2, 5)
        cl, kl = c.astype(ld), kf.astype(ld)
        om = np.zeros(4, ld)
        om[1] += ld(0.25)
        J = np.zeros((4, 4), ld)
        mag = np.zeros(4, ld)
        for r, (reac, prod) in enumerate(_TINY_R):
            rop = kl[r] * np.prod([cl[s] ** n for s, n in reac.items()])
            net = {}
            for s, n in reac.items():
                net[s] = net.get(s, ld(0)) - n
            for s, y in prod.items():
                net[s] = net.get(s, ld(0)) + ld(y)
            for s, nu in net.items():
                if s < 4:
                    om[s] += nu * rop
            # the oracle adds reactant and product entries one by one
            for s, nu in [(s, -n) for s, n in reac.items()] + list(prod.items()):
                if s < 4:
                    mag[s] += abs(nu * rop)
            for j, nj in reac.items():
                if j >= 4:
                    continue
                d = kl[r] * nj * cl[j] ** (nj - 1) * np.prod([cl[s] ** n for s, n in reac.items() if s != j])
                for s, nu in net.items():
                    if s < 4:
                        J[s, j] += nu * d
        got = o.omega(kf, c)
        assert np.all(np.abs(got - om.astype(float)) <= 1e-14 * (mag + 0.25).astype(float))
        Jg = o.jacobian(kf, c)
        np.testing.assert_allclose(Jg, J.astype(float), rtol=1e-13, atol=1e-13 * float(np.abs(J).max()))


@pytest.mark.parametrize("case", ["linear", "quadratic"])
def test_trbdf2_error_estimate_is_asymptotically_exact(oracle_mod, tmp_path, case):
    """The step controller accepts a step when max |est| / (atol + rtol |u|)
    <= 1.  On y' = -k y (A -> B) and on y' = -2 k y^2 (2 A -> B), whose exact
    solutions are closed forms, the largest accepted step h* found by
    bisection on the accept / reject decision equals, to 1 %, the step at
    which the TRUE one-step error |y_1 - y(h)| reaches the tolerance: the
    estimate 2 k_err h (f_n / g - f_g / (g (1 - g)) + f_1 / (1 - g)) is the
    local error to leading order (a dropped term or a wrong constant moves
    h* by 25 % or more).  Also pinned: the growth rule (dt_out = 2 h when
    the estimate is below 1/8 of the tolerance, h otherwise)."""
    k = 0.7
    reac = "A: 1" if case == "linear" else "A: 2"
    f = tmp_path / f"{case}.yaml"
    f.write_text("NCAR-version: v1.0\nspecies:\n- name: A\n- name: B\nconstant_species:\n- name: M\n"
                 f"reactions:\n- reactants:\n    {reac}\n  products:\n    B: 1\n  type: ARRHENIUS\n"
                 f"  coefficients:\n    A: {k}\n")
    o = oracle_mod.Oracle(str(f))
    y0, rtol = 1.0, 1e-9
    p0 = 300.0 / oracle_mod.kmd_mod.CONV_PPM  # ppm conversion factor 1 at 300 K
    k = o.kforward(300.0, p0)[0]
    if case == "linear":  # A -> B: B = y0 - A
        exact = lambda h: np.array([y0 * np.exp(-k * h), y0 - y0 * np.exp(-k * h)])
    else:  # 2 A -> B: B = (y0 - A) / 2
        exact = lambda h: np.array([y0 / (1.0 + 2.0 * k * y0 * h), 0.5 * (y0 - y0 / (1.0 + 2.0 * k * y0 * h))])
    tn = np.array([1e-30, 1e-15])

    def step(h, rt):
        st = np.array([[0.0, p0, 300.0, y0, 0.0, 0.0]])
        ta = _tadv(oracle_mod, 1, 1e30, h, 1e-30, 1e30, 1)
        t, dt = np.zeros(1), np.zeros(1)
        o.run_batch(tn, np.tile([0.0, rt], (2, 1)), ta, st, t, dt)
        return t[0] > 0.0, st[0, 3:5], dt[0]

    def bisect(pred, lo, hi):
        for _ in range(60):
            mid = np.sqrt(lo * hi)
            lo, hi = (mid, hi) if pred(mid) else (lo, mid)
        return lo

    h_acc = bisect(lambda h: step(h, rtol)[0], 1e-6, 1.0)
    def true_norm(h):  # the controller's norm applied to the true one-step error
        y1 = step(h, 1e30)[1]
        return np.max(np.abs(y1 - exact(h)) / (rtol * np.abs(y1)))
    h_true = bisect(lambda h: true_norm(h) <= 1.0, 1e-6, 1.0)
    assert abs(h_acc / h_true - 1.0) < 1e-2, (h_acc, h_true)
    # growth: the proposal doubles while the estimate is at most 1/8 of the
    # tolerance, and stays h above
    h_grow = bisect(lambda h: step(h, rtol)[2] == 2.0 * h, 1e-6, h_acc)
    h_grow_true = bisect(lambda h: true_norm(h) <= 0.125, 1e-6, 1.0)
    assert abs(h_grow / h_grow_true - 1.0) < 1e-2, (h_grow, h_grow_true)
    assert step(1.01 * h_grow, rtol)[2] == 1.01 * h_grow
