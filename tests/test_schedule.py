"""The compiled sparse programs (tchem-atm_amd/csrc/schedule.cpp), checked on
the CPU by tests/native/schedule_check.cpp: bounds, race freedom of every
round, padding isolation, and agreement with direct formulas and a
long-double dense solve.  The checker is built from the library's host
sources with g++."""
import os
import subprocess

import pytest

from conftest import ROOT, golden

CSRC = os.path.join(ROOT, "tchem-atm_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("native") / "schedule_check")
    srcs = [os.path.join(ROOT, "tests", "native", "schedule_check.cpp")] + [
        os.path.join(CSRC, f) for f in ("yaml_lite.cpp", "kinetic_model.cpp", "mechanism.cpp", "schedule.cpp",
                                        "aerosol_model.cpp")]
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + CSRC, *srcs, "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("mech", ["cb05cl_ae5.yaml", "carbonbond05.yaml", "troe.yaml", "photolysis.yaml",
                                  "emission.yaml"])
def test_sparse_programs(checker, mech):
    r = subprocess.run([checker, golden(mech)], capture_output=True, text=True)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-3000:]


@pytest.mark.parametrize("mech", ["uci.yaml", "uci_v2.yaml", "uci_v2_test3.yaml"])
def test_sparse_programs_e3sm_mechanisms(checker, mech):
    r = subprocess.run([checker, golden(mech)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-3000:]


@pytest.mark.parametrize("seed,n_species,n_reactions,n_radicals,window",
                         [(5, 40, 120, 6, 10), (6, 90, 270, 10, 25), (7, 120, 400, 16, 60), (8, 80, 300, 4, 60)])
def test_sparse_programs_synthetic_patterns(checker, tmp_path, seed, n_species, n_reactions, n_radicals, window):
    """The list-scheduled LU program on factor patterns of other shapes and
    fill: seeded synthetic mechanisms with narrow and wide coupling windows
    (every round race free, the solves against a long-double dense solve)."""
    import importlib
    scen = importlib.import_module("tchem-atm_amd.scenarios")
    path = str(tmp_path / f"synthetic_{seed}.yaml")
    scen.synthetic_mechanism_yaml(path, seed=seed, n_species=n_species, n_reactions=n_reactions,
                                  n_radicals=n_radicals, window=window, coupling="cascade")
    r = subprocess.run([checker, path], capture_output=True, text=True)
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stdout[-3000:]
    assert "rounds lu" in r.stdout


@pytest.mark.parametrize("chem,aero", [("cb05_soa_gas.yaml", "cb05_soa_aero.yaml"),
                                       ("simpol_gas.yaml", "simpol_aero.yaml")])
def test_aerosol_lu_program(checker, chem, aero):
    """The sparse aerosol form's LU program over gas + particle unknowns."""
    r = subprocess.run([checker, golden(chem), golden(aero)], capture_output=True, text=True)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-3000:]
    assert "aerosol LU program" in r.stdout
