"""The box-model executable (TChem_AtmosphericChemistry.x on the C ABI) run
with the reference's CB05CL_AE5 ctest command line
(src/unit-tests/CB05CL_AE5/CMakeLists.txt:23-35) and checked the way that
ctest checks it: the written full_gas.dat against the reference's
full_gas.hdf5 (kept as tests/golden/cb05cl_ae5_full_gas.npy) with the norm
criterion of src/unit-tests/scripts/compare_tchem.py at 1e-8.  Also checks
the other files the reference writes (kmod.echo, wall_times.json)."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden
from test_oracle import compare_tchem_norms

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tchem-atm_amd", "TChem_AtmosphericChemistry.x")


def _run(tmp_path, *extra):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (make -C tchem-atm_amd/csrc)")
    args = [EXE, f"--chemfile={golden('cb05cl_ae5.yaml')}", "--outputfile=full_gas.dat",
            "--outputfile_times=AtmosphericChemistry_wall_times.json", "--time-iterations-per-interval=10",
            "--dtmin=1e-20", "--dtmax=10", "--tend=600", "--atol-newton=1e-18", "--rtol-newton=1e-8",
            "--max-newton-iterations=20", "--max-time-iterations=2", *extra]
    print("case:", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r


@pytest.mark.parametrize("lu", ["dense", "sparse"])
def test_cb05cl_ae5_ctest_through_the_executable(tmp_path, lu, gpu):
    r = _run(tmp_path, f"--lu={lu}")
    assert "Number of Species 81" in r.stdout and "Number of Reactions 187" in r.stdout
    test = np.loadtxt(tmp_path / "full_gas.dat", skiprows=1)
    ref = np.load(golden("cb05cl_ae5_full_gas.npy"))
    assert test.shape == ref.shape
    assert compare_tchem_norms(test, ref) < 1e-8
    header = (tmp_path / "full_gas.dat").read_text().splitlines()[0].split("\t")
    assert [h.strip() for h in header[:6]] == ["iter", "t", "dt", "Density[kg/m3]", "Pressure[Pascal]",
                                                "Temperature[K]"]
    assert (tmp_path / "kmod.echo").stat().st_size > 1000
    body = json.loads((tmp_path / "AtmosphericChemistry_wall_times.json").read_text())
    assert body["Aerosol Chemistry"]["number_of_time_iters"] == 2
    assert body["Aerosol Chemistry"]["number_of_samples"] == 1


def test_cloned_samples_and_mixed_precision(tmp_path, gpu):
    _run(tmp_path, "--batch_size=64", "--use_cloned_samples=true", "--precision=mixed", "--schedule=persistent")
    test = np.loadtxt(tmp_path / "full_gas.dat", skiprows=1)
    ref = np.load(golden("cb05cl_ae5_full_gas.npy"))
    assert test.shape == (3 * 64, ref.shape[1])
    final = test[test[:, 0] == 1]
    assert np.array_equal(final, np.repeat(final[:1], 64, axis=0))  # every clone identical
    big = np.abs(ref[-1, 6:]) > 1e-23
    rel = np.abs(final[0, 6:] - ref[-1, 6:])[big] / np.abs(ref[-1, 6:][big])
    assert rel.max() < 1e-6


EXE_E3SM = os.path.join(ROOT, "tchem-atm_amd", "TChem_AtmosphericChemistryE3SM.x")


def test_uci_column_e3sm_golden(tmp_path, gpu):
    """The reference's AtmosphericChemistry_uci ctest
    (src/unit-tests/TChem_Test_AtmosphericChemistryE3SM.hpp:57-93): UCI
    mechanism, 80 cells of a column with per-cell photolysis rates and
    external forcing, TR-BDF2 to 1800 s, against its full_gas.dat.

    The reference's step-size controller lives in Tines, which the reference
    does not vendor.  In the cells where both controllers take the same
    steps (pure doubling from 1e-20 s, then the step clamped to t_end) the
    final states agree to 1e-8; elsewhere the accepted-step sequences differ
    and the states agree to the integration tolerance."""
    args = [EXE_E3SM, f"--chemfile={golden('uci_v2_test3.yaml')}", "--outputfile=full_gas.dat",
            f"--inputfile={golden('uci_col_input_conditions.yaml')}", "--time-iterations-per-interval=100",
            "--tol-time=1e-6", "--dtmin=1e-20", "--dtmax=3600", "--tend=1800", "--atol-newton=1e-18",
            "--rtol-newton=1e-8", "--max-newton-iterations=20", "--max-time-iterations=20000", "--verbose=true"]
    print("case:", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    test = np.loadtxt(tmp_path / "full_gas.dat", skiprows=1)
    ref = np.loadtxt(golden("uci_col_full_gas.dat"), skiprows=1)
    assert test.shape == ref.shape == (160, 88)
    np.testing.assert_array_equal(test[:80], ref[:80])  # initial rows: scenario reader, units
    ft, fr = test[80:], ref[80:]
    np.testing.assert_array_equal(ft[:, 1], 1800.0)
    np.testing.assert_array_equal(ft[:, 3:6], fr[:, 3:6])
    d = np.abs(ft[:, 6:] - fr[:, 6:])
    exact = (d / np.maximum(np.abs(fr[:, 6:]), 1e-30) < 1e-8).all(axis=1)
    doubling = np.abs(fr[:, 2] - 2 * (1800 - 1e-20 * (2 ** 77 - 1))) < 1e-6
    print(f"cells identical to the golden: {exact.sum()} of 80; golden cells with pure doubling: {doubling.sum()}")
    assert exact.sum() >= 40 and exact[exact].size == (exact & doubling).sum()
    assert (d <= 10 * (1e-12 + 1e-6 * np.abs(fr[:, 6:]))).all()  # within 10x the integration tolerance


EXE_AERO = os.path.join(ROOT, "tchem-atm_amd", "TChem_AerosolChemistry.x")


@pytest.mark.parametrize("ref_name,chem,aero,part,extra", [
    ("simpol_full_gas.npy", "simpol_gas.yaml", "simpol_aero.yaml", "simpol_particles.yaml", ()),
    ("simpol_batched_full_gas.npy", "simpol_gas.yaml", "simpol_aero.yaml", "simpol_particles.yaml",
     ("--use_cloned_samples=true", "--batch_size=3")),
    ("cb05_soa_batched_full_gas.npy", "cb05_soa_gas.yaml", "cb05_soa_aero.yaml", "cb05_soa_particles.yaml",
     ("--use_cloned_samples=true", "--batch_size=3")),
])
def test_tines_ode_aerosol_ctest_through_the_executable(tmp_path, gpu, ref_name, chem, aero, part, extra):
    """The reference's TinesODE ctests (src/unit-tests/TinesODE/CMakeLists.txt)
    on TChem_AerosolChemistry.x, checked as compare_tchem.py checks them."""
    args = [EXE_AERO, f"--chemfile={golden(chem)}", f"--aerofile={golden(aero)}",
            f"--inputfile_particles={golden(part)}", "--outputfile=full_gas.dat",
            "--outputfile_times=AerosolChemistry_wall_times.json", "--dtmin=1e-20", "--dtmax=10", "--tend=600",
            "--atol-newton=1e-18", "--rtol-newton=1e-8", "--max-time-iterations=2", *extra]
    print("case:", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    test = np.loadtxt(tmp_path / "full_gas.dat", skiprows=1)
    ref = np.load(golden(ref_name))
    assert test.shape == ref.shape
    assert compare_tchem_norms(test, ref) < 1e-8
    header = [h.strip() for h in (tmp_path / "full_gas.dat").read_text().splitlines()[0].split("\t") if h.strip()]
    assert len(header) == ref.shape[1]
    assert header[-1].endswith("_p0")
    body = json.loads((tmp_path / "AerosolChemistry_wall_times.json").read_text())
    assert body["Aerosol Chemistry"]["number_of_time_iters"] == 2
