"""Discrete adjoint on the GPU (needs an MI355X), CB05CL_AE5 with the
differentiation workload's distributions (config 0 draws, seed 2, 1800 s)
on a few cells.

Loss L = sum_cells sum_k w_k log c_k(t_end), w ~ N(0, 1) seeded, over the
species at or above the absolute tolerance 1e-12 ppm at t_end in the base
run (below it a species' value is set by the Newton tolerance, atol 1e-18,
not by the rate parameters: two correct integrations differ there by orders
of magnitude, and so do their derivatives), lambda = dL/dc = w / c.  Checked:
  * against the oracle's discrete tangent linear (forward sensitivity) on the
    step sizes the GPU recorded: dL/dtheta_r for 32 seeded-random reactions
    and dL/dc0 for 8 initial species, normwise 1e-6;
  * against central finite differences of the GPU forward itself, h = 1e-6
    in theta (and relative 1e-6 in c0), all perturbations in one launch with
    per-cell theta; pairs whose perturbed run took a different step sequence
    are excluded (the derivative is that of the fixed-step map);
  * the autograd wrapper, linearity and repeatability of the backward
    sweep, and the NaN marking of overflowed or mismatched records.
"""
import importlib

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

TEND = 1800.0
MAX_STEPS = 8192
FLOOR = 1e-12


@pytest.fixture(scope="module")
def cb05(gpu, pkg):
    kmd = pkg.KineticModelData(golden("cb05cl_ae5.yaml"))
    kmcd = kmd.createConstData()
    base = pkg.setScenarioConditions(golden("cb05cl_ae5.yaml"), kmd)
    return kmd, kmcd, base


@pytest.fixture(scope="module")
def work(cb05):
    scen = importlib.import_module("tchem-atm_amd.scenarios")
    kmd, _, base = cb05
    return scen.config0(kmd, base[0], n=4, seed=2)


def _forward(pkg, torch, kmcd, w, state, photo, theta=None, max_steps=MAX_STEPS):
    n, m = state.shape[0], kmcd.nActive
    st = torch.tensor(state, device="cuda")
    ph = torch.tensor(photo, device="cuda")
    tn = torch.tensor([w.atol_newton, w.rtol_newton], dtype=torch.float64, device="cuda")
    tt = torch.tensor(np.tile([w.atol_time, w.rtol_time], (m, 1)), device="cuda")
    tadv = pkg.time_advance_tensor(n, 0.0, TEND, w.dt_in, w.dtmin, w.dtmax, w.max_newton, 1 << 30, 1)
    t = torch.zeros(n, dtype=torch.float64, device="cuda")
    dt = torch.zeros(n, dtype=torch.float64, device="cuda")
    ws = pkg.Adjoint.workspace(kmcd, n, max_steps)
    th = None if theta is None else torch.tensor(theta, device="cuda")
    pkg.Adjoint.forward(tn, tt, tadv, st, t, dt, ws, max_steps, kmcd, photo_rates=ph, log_rate_scale=th)
    ns, h = pkg.Adjoint.steps(ws, n, max_steps, kmcd)
    torch.cuda.synchronize()
    return st, ws, ns.cpu().numpy(), h.cpu().numpy(), t.cpu().numpy()


def _weights(final, m):
    """[n][m] loss weights: w ~ N(0, 1) per species, 0 where c(t_end) < FLOOR."""
    w = np.random.Generator(np.random.Philox(key=2)).standard_normal(m)
    return np.where(final[:, 3:3 + m] >= FLOOR, w, 0.0)


def _loss_terms(final, m, wts):
    c = final[:, 3:3 + m]
    return np.where(wts != 0.0, wts * np.log(np.where(wts != 0.0, c, 1.0)), 0.0).sum(1)


def _lambda(final, m, nspec, wts):
    lam = np.zeros((final.shape[0], nspec))
    c = final[:, 3:3 + m]
    lam[:, :m] = np.where(wts != 0.0, wts / np.where(wts != 0.0, c, 1.0), 0.0)
    return lam


@pytest.fixture(scope="module")
def adj(cb05, work, gpu, pkg):
    """Forward + backward of the base cells."""
    torch = gpu
    kmd, kmcd, _ = cb05
    m, nspec = kmd.nActive, kmd.nSpec
    st, ws, ns, h, t = _forward(pkg, torch, kmcd, work, work.state, work.photo)
    final = st.cpu().numpy()
    wts = _weights(final, m)
    lam = torch.tensor(_lambda(final, m, nspec, wts), device="cuda")
    lam_out = torch.empty_like(lam)
    grad = torch.empty((work.n, kmd.nReac), dtype=torch.float64, device="cuda")
    pkg.Adjoint.backward(lam, lam_out, grad, ws, MAX_STEPS, kmcd)
    torch.cuda.synchronize()
    return dict(final=final, ns=ns, h=h, t=t, ws=ws, lam=lam, lam_out=lam_out.cpu().numpy(),
                grad=grad.cpu().numpy(), wts=wts)


def _pick(nreac, m):
    g = np.random.default_rng(2)
    return g.choice(nreac, 32, replace=False), g.choice(m, 8, replace=False)


def test_forward_record(adj, work, cb05):
    kmd = cb05[0]
    assert (adj["ns"] > 100).all() and (adj["ns"] < MAX_STEPS).all()
    np.testing.assert_array_equal(adj["t"], TEND)
    for c in range(work.n):
        h = adj["h"][c, :adj["ns"][c]]
        assert (h > 0).all()
        assert abs(h.sum() - TEND) <= 1e-9 * TEND  # the accepted steps tile [0, tend]
    assert np.isfinite(adj["grad"]).all() and np.isfinite(adj["lam_out"][:, :kmd.nActive]).all()
    np.testing.assert_array_equal(adj["lam_out"][:, kmd.nActive:], 0.0)


def test_adjoint_matches_oracle_tangent_linear(adj, work, cb05, cb05_oracle):
    from concurrent.futures import ThreadPoolExecutor
    kmd = cb05[0]
    o = cb05_oracle
    m, nR = kmd.nActive, kmd.nReac
    reacs, specs = _pick(nR, m)
    sens = list(reacs) + [-(s + 1) for s in specs]
    tn = np.array([work.atol_newton, work.rtol_newton])

    def run(c):
        h = adj["h"][c, :adj["ns"][c]]
        return o.tlm_fixed_steps(tn, work.max_newton, 1, work.state[c], h, sens, photo=work.photo[c])

    with ThreadPoolExecutor(max_workers=4) as ex:
        res = list(ex.map(run, range(work.n)))
    for c, (u, S) in enumerate(res):
        # the oracle's trajectory on the GPU's steps ends where the GPU did
        assert np.abs(u - adj["final"][c, 3:3 + m]).max() <= 1e-6 * np.abs(u).max()
        lam = np.where(adj["wts"][c] != 0.0, adj["wts"][c] / np.where(adj["wts"][c] != 0.0, u, 1.0), 0.0)
        ref_r = lam @ S[:, :32]
        ref_s = lam @ S[:, 32:]
        got_r = adj["grad"][c, reacs]
        got_s = adj["lam_out"][c, specs]
        assert np.abs(got_r - ref_r).max() <= 1e-6 * np.abs(ref_r).max(), (c, np.abs(got_r - ref_r).max())
        assert np.abs(got_s - ref_s).max() <= 1e-6 * np.abs(ref_s).max(), (c, np.abs(got_s - ref_s).max())


def test_adjoint_matches_central_differences(adj, work, cb05, gpu, pkg):
    torch = gpu
    kmd, kmcd, _ = cb05
    m, nR, n = kmd.nActive, kmd.nReac, work.n
    reacs, specs = _pick(nR, m)
    eps = 1e-6
    rows, photo, theta, meta = [], [], [], []
    for c in range(n):
        for r in reacs:
            for sgn in (1.0, -1.0):
                th = np.zeros(nR)
                th[r] = sgn * eps
                rows.append(work.state[c])
                photo.append(work.photo[c])
                theta.append(th)
                meta.append((c, "r", r, sgn, eps))
        for s in specs:
            for sgn in (1.0, -1.0):
                row = work.state[c].copy()
                d = eps * row[3 + s]
                row[3 + s] += sgn * d
                rows.append(row)
                photo.append(work.photo[c])
                theta.append(np.zeros(nR))
                meta.append((c, "s", s, sgn, d))
    st, _, ns, h, _ = _forward(pkg, torch, kmcd, work, np.array(rows), np.array(photo), np.array(theta))
    final = st.cpu().numpy()
    cells = np.array([mt[0] for mt in meta])
    L = _loss_terms(final, m, adj["wts"][cells])
    fd, same = {}, {}
    for i, (c, kind, idx, sgn, d) in enumerate(meta):
        key = (c, kind, idx)
        fd[key] = fd.get(key, 0.0) + sgn * L[i] / (2 * d)
        ok = ns[i] == adj["ns"][c] and np.array_equal(h[i, :ns[i]], adj["h"][c, :ns[i]])
        same[key] = same.get(key, True) and ok
    valid = [k for k in fd if same[k]]
    assert len(valid) >= 0.75 * len(fd), f"{len(fd) - len(valid)} of {len(fd)} perturbations changed the steps"
    worst = 0.0
    for c in range(n):
        gr = adj["grad"][c]
        lo = adj["lam_out"][c]
        scale_r = np.abs(gr[reacs]).max()
        scale_s = np.abs(lo[specs]).max()
        nsq = np.sqrt(adj["ns"][c])
        for key in valid:
            if key[0] != c:
                continue
            _, kind, idx = key
            got = gr[idx] if kind == "r" else lo[idx]
            d = eps if kind == "r" else eps * work.state[c, 3 + idx]
            # 1e-5 of the gradient's scale plus the rounding level of the
            # difference quotient of L (sqrt(steps) eps_mach sum |w| / d)
            tol = 1e-5 * (scale_r if kind == "r" else scale_s) + 4 * nsq * 1.1e-16 * np.abs(adj["wts"][c]).sum() / d
            worst = max(worst, abs(fd[key] - got) / tol)
            assert abs(fd[key] - got) <= tol, (key, fd[key], got, tol)
    print(f"central differences: {len(valid)} of {len(fd)} pairs compared, worst error / tolerance {worst:.3f}")


def test_autograd_function(adj, work, cb05, gpu, pkg):
    torch = gpu
    A = importlib.import_module("tchem-atm_amd.adjoint")
    kmd, kmcd, _ = cb05
    m = kmd.nActive
    tn = torch.tensor([work.atol_newton, work.rtol_newton], dtype=torch.float64, device="cuda")
    tt = torch.tensor(np.tile([work.atol_time, work.rtol_time], (m, 1)), device="cuda")
    prob = A.AdjointProblem(kmcd, tn, tt, 0.0, TEND, work.dt_in, work.dtmin, work.dtmax, work.max_newton,
                            max_steps=MAX_STEPS, photo_rates=torch.tensor(work.photo, device="cuda"))
    theta = torch.zeros(kmd.nReac, dtype=torch.float64, device="cuda", requires_grad=True)
    state = torch.tensor(work.state, device="cuda", requires_grad=True)
    final = A.KineticsAdjoint.apply(theta, state, prob)
    w = torch.tensor(adj["wts"], device="cuda")
    loss = (w * torch.log(torch.clamp(final[:, 3:3 + m], min=FLOOR))).sum()
    loss.backward()
    np.testing.assert_array_equal(final.detach().cpu().numpy(), adj["final"])
    np.testing.assert_allclose(theta.grad.cpu().numpy(), adj["grad"].sum(0), rtol=1e-12, atol=0)
    np.testing.assert_allclose(state.grad[:, 3:].cpu().numpy(), adj["lam_out"], rtol=1e-12, atol=0)
    np.testing.assert_array_equal(state.grad[:, :3].cpu().numpy(), 0.0)


def test_backward_linear_and_repeatable(adj, work, cb05, gpu, pkg):
    torch = gpu
    kmd, kmcd, _ = cb05
    lam2 = 2.0 * adj["lam"]
    lo = torch.empty_like(lam2)
    g = torch.empty((work.n, kmd.nReac), dtype=torch.float64, device="cuda")
    pkg.Adjoint.backward(lam2, lo, g, adj["ws"], MAX_STEPS, kmcd)
    torch.cuda.synchronize()
    np.testing.assert_allclose(g.cpu().numpy(), 2.0 * adj["grad"], rtol=1e-13, atol=0)
    np.testing.assert_allclose(lo.cpu().numpy(), 2.0 * adj["lam_out"], rtol=1e-13, atol=0)
    lam_in = adj["lam"].clone()
    pkg.Adjoint.backward(lam_in, lam_in, g, adj["ws"], MAX_STEPS, kmcd)  # lambda_out may alias lambda
    torch.cuda.synchronize()
    np.testing.assert_array_equal(lam_in.cpu().numpy(), adj["lam_out"])
    np.testing.assert_array_equal(g.cpu().numpy(), adj["grad"])


def test_overflow_and_mismatch_give_nan(work, cb05, gpu, pkg):
    torch = gpu
    kmd, kmcd, _ = cb05
    n = work.n
    st, ws, ns, _, _ = _forward(pkg, torch, kmcd, work, work.state, work.photo, max_steps=64)
    assert (ns == -1).all()  # every cell takes more than 64 accepted steps
    lam = torch.ones((n, kmd.nSpec), dtype=torch.float64, device="cuda")
    lo = torch.empty_like(lam)
    g = torch.empty((n, kmd.nReac), dtype=torch.float64, device="cuda")
    pkg.Adjoint.backward(lam, lo, g, ws, 64, kmcd)
    torch.cuda.synchronize()
    assert torch.isnan(g).all() and torch.isnan(lo).all()
    # a record made for n cells read as n - 1 cells
    _, ws2, ns2, _, _ = _forward(pkg, torch, kmcd, work, work.state, work.photo)
    assert (ns2 > 0).all()
    pkg.Adjoint.backward(lam[:n - 1], lo[:n - 1], g[:n - 1], ws2, MAX_STEPS, kmcd)
    torch.cuda.synchronize()
    assert torch.isnan(g[:n - 1]).all()
    with pytest.raises(pkg.TChemError):
        pkg.Adjoint.workspaceBytes(kmcd, n, 0)
