"""E3SM/UCI kinetics batch functions on the GPU against the reference's
own unit-test goldens, compared as text, byte for byte, the way its ctests
do (TChem::Test::compareFiles):

  ReactionRates       uci.yaml, 10 cloned cells     src/unit-tests/TChem_Test_ReactionRates.hpp
  RateofProgress      uci_v2.yaml, 10 cloned cells  src/unit-tests/TChem_Test_RateofProgress.hpp
  NetProductionRates  uci_v2.yaml, 10 cloned cells  src/unit-tests/TChem_Test_NetProductionRates.hpp:30-50
  NetProductionRates  uci_v2_test3.yaml with the 80-cell column input (photolysis rates and
                      external forcing per cell)   src/unit-tests/TChem_Test_NetProductionRates.hpp:55-80

The files are written as the reference's examples write them
(writeReactionRates / write2DMatrix, src/core/TChem_Util.hpp:1036-1118:
std::scientific, six digits).
"""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def write_reaction_rates(n_spec, omega):
    """TChem::Test::writeReactionRates: header, then 'i   value' lines."""
    lines = [f"## nSpec({n_spec})    Omega({len(omega)})"]
    lines += [f"{i}   {v:e}" for i, v in enumerate(omega)]
    return "\n".join(lines) + "\n"


def write_2d_matrix(n_row, n_col, mat):
    """TChem::Test::write2DMatrix: header, then rows of 'value ' and a newline."""
    out = [f"## nRows({n_row})    ({mat.shape[0]})", f"## nCols({n_col})    ({mat.shape[1]})"]
    out += ["".join(f"{v:e} " for v in row) for row in mat]
    return "\n".join(out) + "\n"


def _setup(pkg, torch, mech, inputfile=None, n_clone=10):
    kmd = pkg.KineticModelData(golden(mech))
    kmcd = kmd.createConstData()
    inp = golden(inputfile or mech)
    state = pkg.setScenarioConditions(inp, kmd)
    photo = pkg.setScenarioConditionsPhotolysisReactions(inp, kmd, state.shape[0])
    ext = pkg.setScenarioConditionsExternalForcing(inp, kmd, state.shape[0])
    if n_clone and state.shape[0] == 1:  # --use_cloned_samples=true --batch_size=10
        state = np.repeat(state, n_clone, axis=0)
        photo = np.repeat(photo, n_clone, axis=0)
        ext = np.repeat(ext, n_clone, axis=0)
    d = "cuda"
    return (kmd, kmcd, torch.tensor(state, device=d),
            torch.tensor(photo, device=d) if photo.shape[1] else None, torch.tensor(ext, device=d))


def test_reaction_rates_golden(gpu, pkg):
    torch = gpu
    kmd, kmcd, state, _, _ = _setup(pkg, torch, "uci.yaml")
    kfwd = torch.full((state.shape[0], kmd.nReac), np.nan, dtype=torch.float64, device="cuda")
    pkg.ReactionRates.runDeviceBatch(state, kfwd, kmcd)
    got = write_reaction_rates(kmd.nReac, kfwd.cpu().numpy()[0])
    assert got == open(golden("e3sm_reaction_rates.dat")).read()
    assert torch.equal(kfwd, kfwd[:1].expand_as(kfwd))  # cloned cells


def test_rate_of_progress_golden(gpu, pkg):
    torch = gpu
    kmd, kmcd, state, photo, _ = _setup(pkg, torch, "uci_v2.yaml")
    rop = torch.full((state.shape[0], kmd.nReac), np.nan, dtype=torch.float64, device="cuda")
    pkg.RateofProgress.runDeviceBatch(state, photo, rop, kmcd)
    got = write_reaction_rates(kmd.nReac, rop.cpu().numpy()[0])
    assert got == open(golden("e3sm_rate_of_progress.dat")).read()


def test_net_production_rates_golden(gpu, pkg):
    torch = gpu
    kmd, kmcd, state, photo, ext = _setup(pkg, torch, "uci_v2.yaml")
    omega = torch.full((state.shape[0], kmd.nSpec), np.nan, dtype=torch.float64, device="cuda")
    pkg.NetProductionRates.runDeviceBatch(state, photo, ext, omega, kmcd)
    got = write_reaction_rates(kmd.nReac, omega.cpu().numpy()[0])
    assert got == open(golden("e3sm_net_production_rates.dat")).read()


def test_net_production_rates_column_golden(gpu, pkg):
    """80 cells of a column, per-cell photolysis rates (22 PHOTO_RATE
    reactions) and external forcing."""
    torch = gpu
    kmd, kmcd, state, photo, ext = _setup(pkg, torch, "uci_v2_test3.yaml", "uci_col_input_conditions.yaml")
    assert state.shape[0] == 80 and photo.shape == (80, 22)
    omega = torch.full((80, kmd.nSpec), np.nan, dtype=torch.float64, device="cuda")
    pkg.NetProductionRates.runDeviceBatch(state, photo, ext, omega, kmcd)
    got = write_2d_matrix(80, kmd.nSpec, omega.cpu().numpy())
    ref = open(golden("uci_col_net_production_rates.dat")).read()
    if got != ref:  # report the first differing value
        g, r = got.split(), ref.split()
        bad = [(i, a, b) for i, (a, b) in enumerate(zip(g, r)) if a != b]
        pytest.fail(f"{len(bad)} values differ, first {bad[:3]}")


def test_box_model_entry_points_reject_e3sm_mechanisms(gpu, pkg):
    """The UCI mechanisms (JPL, PHOTO_RATE) run the kinetics entry points only."""
    torch = gpu
    kmd = pkg.KineticModelData(golden("uci_v2_test3.yaml"))
    kmcd = kmd.createConstData()
    st = torch.zeros((1, kmd.nSpec + 3), dtype=torch.float64, device="cuda")
    with pytest.raises(pkg.TChemError, match="sourceterm"):
        pkg.SourceTerm.runDeviceBatch(st, None, None, torch.zeros((1, kmd.nActive), dtype=torch.float64,
                                                                  device="cuda"), kmcd)
