"""Pins of the oracle's derivatives of the discrete TR-BDF2 map (the
reference has no derivative path; these are checked against what the
mathematics fixes).  CPU only.

  * y' = -k y: the tangent linear of a fixed step sequence equals the
    closed-form derivative of the product of TR-BDF2 amplification factors,
    in the rate log-scale and in the initial value;
  * CB05CL_AE5: the tangent linear agrees with central differences (h = 1e-6
    in the log-scale, relative 1e-6 in the initial values; the oracle's own
    fixed-step map) for random reactions and initial species, on step
    sequences the oracle's integrator chose, to 1e-7 of each column's scale
    plus the difference quotient's rounding level.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from conftest import golden
from test_oracle import _tadv


def _decay(oracle_mod, tmp_path, k):
    f = tmp_path / "decay.yaml"
    f.write_text("NCAR-version: v1.0\nspecies:\n- name: A\n- name: B\nconstant_species:\n- name: M\n"
                 f"reactions:\n- reactants:\n    A: 1\n  products:\n    B: 1\n  type: ARRHENIUS\n"
                 f"  coefficients:\n    A: {k}\n")
    return oracle_mod.Oracle(str(f))


def test_tlm_closed_form_decay(oracle_mod, tmp_path):
    k = 0.7
    o = _decay(oracle_mod, tmp_path, k)
    _, c, a_g, a_n, _, _ = oracle_mod.trbdf2_constants()
    h = np.array([1e-3, 0.1, 0.5, 1.0, 2.0, 10.0, 0.3])
    theta = 0.25
    row = np.array([0.0, 1e5, 300.0, 1.3, 0.2, 0.0])
    u, S = o.tlm_fixed_steps([1e-30, 1e-14], 20, 1, row, h, sens=[0, -1, -2], lscale=[theta])
    ke = k * np.exp(theta)
    z = -ke * h
    P = (1 + c * z) / (1 - c * z)
    R = (a_g * P - a_n) / (1 - c * z)
    dR = a_g * (2 * c / (1 - c * z) ** 2) / (1 - c * z) + (a_g * P - a_n) * c / (1 - c * z) ** 2
    prodR = np.prod(R)
    yA = row[3] * prodR
    dyA_dtheta = yA * np.sum(dR * z / R)  # dz/dtheta = z
    np.testing.assert_allclose(u, [yA, row[3] + row[4] - yA], rtol=1e-13)
    np.testing.assert_allclose(S[:, 0], [dyA_dtheta, -dyA_dtheta], rtol=1e-12)
    np.testing.assert_allclose(S[:, 1], [prodR, 1 - prodR], rtol=1e-12)   # d/d A0
    np.testing.assert_allclose(S[:, 2], [0.0, 1.0], atol=1e-15)            # d/d B0


def _cb05_cells(oracle_mod, o, n, seed):
    g = np.random.Generator(np.random.Philox(key=seed))
    base = oracle_mod.kmd.load_scenario(golden("cb05cl_ae5.yaml"), o.k)[0]
    m = o.m
    rows = np.repeat(base[None], n, 0)
    rows[:, 1] = g.uniform(5.0e4, 1.013e5, n)
    rows[:, 2] = g.uniform(270.0, 310.0, n)
    rows[:, 3:3 + m] = 10.0 ** g.uniform(-3.0, 2.0, (n, m)) * 1e-3
    photo = g.uniform(0.0, 1.0e-3, (n, len(o.k.photoReac)))
    return rows, photo


def test_tlm_against_central_differences_cb05(oracle_mod, cb05_oracle):
    o = cb05_oracle
    m, nR = o.m, o.k.nReac
    rows, photo = _cb05_cells(oracle_mod, o, 2, seed=7)
    tn = np.array([1e-18, 1e-8])
    tt = np.tile([1e-12, 1e-6], (m, 1))
    ta = _tadv(oracle_mod, 1, 300.0, 1e-20, 1e-20, 60.0, 100000)
    g = np.random.default_rng(11)
    reacs = g.choice(nR, 8, replace=False)
    specs = g.choice(m, 4, replace=False)
    sens = list(reacs) + [-(s + 1) for s in specs]
    eps = 1e-6
    jobs = []  # (cell, column, perturbed row, lscale, divisor)
    recs = []
    for c in range(rows.shape[0]):
        _, h = o.record_steps(tn, tt, ta, rows[c], photo[c])
        assert h is not None and len(h) > 20
        recs.append(h)
        for j, s in enumerate(sens):
            for sign in (1.0, -1.0):
                r2, ls = rows[c].copy(), np.zeros(nR)
                if s >= 0:
                    ls[s], d = sign * eps, eps
                else:
                    d = eps * rows[c, 3 + (-s - 1)]
                    r2[3 + (-s - 1)] += sign * d
                jobs.append((c, j, sign, r2, ls, d))
    # ctypes releases the GIL: the fixed-step runs go to a thread pool
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        outs = list(ex.map(lambda jb: o.tlm_fixed_steps(tn, 20, 1, jb[3], recs[jb[0]], (), photo=photo[jb[0]],
                                                        lscale=jb[4])[0], jobs))
        tlm = list(ex.map(lambda c: o.tlm_fixed_steps(tn, 20, 1, rows[c], recs[c], sens, photo=photo[c]),
                          range(rows.shape[0])))
    fdv = {}
    for jb, out in zip(jobs, outs):
        c, j, sign, _, _, d = jb
        fdv[(c, j)] = fdv.get((c, j), 0.0) + sign * out / (2 * d)
    for c in range(rows.shape[0]):
        u, S = tlm[c]
        for j, s in enumerate(sens):
            d = eps if s >= 0 else eps * rows[c, 3 + (-s - 1)]
            fd = fdv[(c, j)]
            # componentwise: 1e-7 of the column's scale plus the rounding
            # level of the difference quotient, ~sqrt(steps) eps_mach |u_i| / d
            scale = np.abs(S[:, j]).max()
            tol = 1e-7 * scale + 4.0 * np.sqrt(len(recs[c])) * 1.1e-16 * np.abs(u) / d
            err = np.abs(fd - S[:, j])
            assert (err <= tol).all(), (c, s, (err / np.maximum(tol, 1e-300)).max())
