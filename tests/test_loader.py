"""Mechanism loader (KineticModelData) and the C ABI surface — CPU only."""
import ctypes
import os
import filecmp
import subprocess

import numpy as np
import pytest

from conftest import golden

ECHO_GOLDENS = [
    # (mechanism, kmod.echo written by the reference for it)
    ("uci.yaml", "kmod_echo_reaction_rates.txt"),
    ("uci_v2.yaml", "kmod_echo_net_production_rates.txt"),
    ("uci_v2.yaml", "kmod_echo_rate_of_progress.txt"),
    ("uci_v2_test3.yaml", "kmod_echo_uci_col.txt"),
    ("uci_v2_test3.yaml", "kmod_echo_atmchem_e3sm_uci_col.txt"),
]


@pytest.mark.parametrize("mech,echo", ECHO_GOLDENS)
def test_echo_matches_reference_byte_for_byte(pkg, tmp_path, mech, echo):
    """src/unit-tests/TChem_Test_{ReactionRates,NetProductionRates,RateofProgress,
    AtmosphericChemistryE3SM}.hpp compare kmod.echo byte-exactly."""
    out = tmp_path / "kmod.echo"
    pkg.KineticModelData(golden(mech), str(out))
    assert filecmp.cmp(out, golden(echo), shallow=False)


def test_cb05_structure(pkg):
    k = pkg.KineticModelData(golden("cb05cl_ae5.yaml"))
    assert (k.nSpec, k.nConstSpec, k.nReac, k.nPhoto) == (81, 7, 187, 25)
    assert k.nActive == 74
    assert k.speciesNames[74:] == ["CH4", "M", "O2", "H2O", "H2", "M", "N2"]
    assert list(k.photolysisReactions) == [0, 7, 8, 13, 14, 24, 34, 49, 50, 51, 60, 62, 69, 72, 73, 84, 88, 94,
                                           99, 103, 133, 138, 146, 155, 184]


def test_cb05_echo_stoichiometry_matches_oracle_tables(pkg, tmp_path, cb05_oracle):
    """The product loader's echo lists every stoichiometric entry; rebuild it from
    the oracle's independent (PyYAML) tables and compare."""
    out = tmp_path / "kmod.echo"
    pkg.KineticModelData(golden("cb05cl_ae5.yaml"), str(out))
    lines = out.read_text().splitlines()
    k = cb05_oracle.k
    start = lines.index("Reaction data : species and Arrhenius pars") + 1
    for r in range(k.nReac):
        parts = [f"{r + 1:<5d}\t1\t{k.reacNreac[r]:2d}\t{k.reacNprod[r]:2d} | "]
        for j in range(k.reacNreac[r]):
            parts.append(f"{k.reacNuki[r, j]:f}*{k.names[k.reacSidx[r, j]]} | ")
        half = k.maxSpecInReac // 2
        for j in range(k.reacNprod[r]):
            parts.append(f"{k.reacNuki[r, half + j]:f}*{k.names[k.reacSidx[r, half + j]]} | ")
        assert lines[start + r] == "".join(parts)


@pytest.mark.parametrize("mech", ["cb05cl_ae5.yaml", "carbonbond05.yaml", "arrhenius.yaml", "troe.yaml",
                                  "photolysis.yaml"])
def test_scenario_conditions_match_oracle(pkg, oracle_mod, mech):
    k = pkg.KineticModelData(golden(mech))
    st = pkg.setScenarioConditions(golden(mech), k)
    ok = oracle_mod.kmd.load(golden(mech))
    assert k.speciesNames == ok.names
    np.testing.assert_array_equal(st, oracle_mod.kmd.load_scenario(golden(mech), ok))


def test_loader_errors(pkg, tmp_path):
    bad = tmp_path / "nom.yaml"
    bad.write_text("NCAR-version: v1.0\nspecies:\n- name: A\nreactions:\n- reactants:\n    A: 1\n"
                   "  products:\n    A: 1\n  type: ARRHENIUS\n")
    with pytest.raises(pkg.TChemError, match="M species"):
        pkg.KineticModelData(str(bad))
    bad.write_text("NCAR-version: v1.0\nspecies:\n- name: A\nconstant_species:\n- name: M\nreactions:\n"
                   "- reactants:\n    B: 1\n  products:\n    A: 1\n  type: ARRHENIUS\n")
    with pytest.raises(pkg.TChemError, match="species does not exist B"):
        pkg.KineticModelData(str(bad))
    bad.write_text("NCAR-version: v1.0\nspecies:\n- name: A\nconstant_species:\n- name: M\nreactions:\n"
                   "- reactants:\n    A: 1\n  products:\n    A: 1\n  type: FOO\n")
    with pytest.raises(pkg.TChemError, match="type"):
        pkg.KineticModelData(str(bad))
    with pytest.raises(pkg.TChemError):
        pkg.KineticModelData(str(tmp_path / "missing.yaml"))


def test_yaml_reader_features(pkg, tmp_path):
    """Anchors/aliases, compact sequences, flow maps, quoted and multi-line scalars."""
    f = tmp_path / "m.yaml"
    f.write_text(
        "NCAR-version: v1.0\n"
        "species:\n- description: a long\n    description\n  name: 'NO'\n- name: \"B\"\n"
        "constant_species:\n- name: M  # comment\n"
        "reactions:\n"
        "- reactants: &r1\n    NO: 1\n  products: {B: 2.0}\n  type: ARRHENIUS\n"
        "  coefficients: {A: 1.5e-3, C: -100}\n"
        "- reactants: *r1\n  products: {}\n  type: TROE\n  adjust_reaction:\n  - B\n")
    k = pkg.KineticModelData(str(f), str(tmp_path / "echo"))
    assert k.speciesNames == ["NO", "B", "M"]
    echo = (tmp_path / "echo").read_text()
    assert "1    \t1\t 1\t 1 | -1.000000*NO | 2.000000*B | " in echo
    assert "2    \t1\t 1\t 0 | -1.000000*NO | " in echo
    assert "# of adjustment pairs                            : 1" in echo


def test_c_abi_exports_every_declared_symbol(pkg):
    header = open(golden("../../include/tchem_atm.h")).read()
    import re
    declared = set(re.findall(r"^\w[\w\s\*]*?\b(tchem_\w+)\(", header, re.M))
    assert declared == set(pkg.EXPORTED_SYMBOLS)
    out = subprocess.run(["nm", "-D", "--defined-only", pkg.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert declared <= exported
    lib = ctypes.CDLL(pkg.LIB_PATH)
    for s in declared:
        assert hasattr(lib, s)


def test_driver_library_exports_every_declared_symbol(pkg):
    """libtchem_driver.so (include/tchem_driver.h): the reference's
    TChem_Driver extern "C" names, loadable without a device."""
    import re
    header = open(golden("../../include/tchem_driver.h")).read()
    declared = set(re.findall(r"^\w[\w\s\*]*?\b(initialize|finalize|TChem_\w+)\(", header, re.M))
    assert {"initialize", "finalize", "TChem_doTimestep", "TChem_getStateVector", "TChem_setStateVector",
            "TChem_getAerosolSpeciesKappa", "TChem_setNumberConcentrationVector"} <= declared
    path = os.path.join(os.path.dirname(pkg.LIB_PATH), "libtchem_driver.so")
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert declared <= exported
    lib = ctypes.CDLL(path)
    lib.TChem_getLengthOfStateVector.restype = ctypes.c_int
    assert lib.TChem_getLengthOfStateVector() == -1  # not initialized: no device call


def test_c_abi_argument_errors_without_device(pkg):
    lib = pkg.lib()
    assert lib.tchem_kmd_load_yaml(None, None, None) == -1
    assert lib.tchem_kmd_info(None, None, None, None, None) == -1
    assert "null" in lib.tchem_last_error().decode()
    # batch calls with a null handle fail cleanly before touching a device
    assert lib.tchem_sourceterm_run_device_batch(4, None, None, None, None, None, None) == -1
    assert lib.tchem_jacobian_run_device_batch(4, None, None, None, None, None) == -1
    assert lib.tchem_atmchem_update_time_advance(-1, None, None, None, None) == -1


def test_scenario_photolysis_rates_and_external_forcing(pkg):
    """setScenarioConditionsPhotolysisReactions / ExternalForcing
    (src/core/TChem_AtmosphericChemistry.cpp:511-631) on the reference's
    80-cell column input, against a direct reading of the file (the species
    names read as strings: YAML 1.1 would turn NO into a boolean)."""
    import yaml
    f = golden("uci_col_input_conditions.yaml")
    kmd = pkg.KineticModelData(golden("uci_v2_test3.yaml"))
    photo = pkg.setScenarioConditionsPhotolysisReactions(f, kmd, 80)
    ext = pkg.setScenarioConditionsExternalForcing(f, kmd, 80)
    raw = yaml.load(open(f), Loader=yaml.BaseLoader)
    assert photo.shape == (80, len(raw["photolysis_rates"]))
    for j in range(photo.shape[1]):
        np.testing.assert_array_equal(photo[:, j], [float(v) for v in raw["photolysis_rates"][f"r_{j}"]["value"]])
    expect = np.zeros((80, kmd.nActive))
    for e in raw["external_forcing"].values():
        expect[:, kmd.speciesNames.index(e["species_name"])] = [float(v) for v in e["value"]]
    np.testing.assert_array_equal(ext, expect)
    assert pkg.setScenarioConditionsPhotolysisReactions(golden("cb05cl_ae5.yaml"),
                                                        pkg.KineticModelData(golden("cb05cl_ae5.yaml")), 1).shape == (1, 0)


def test_aerosol_loader_matches_oracle_reader(pkg, oracle_mod):
    """AerosolModelData (C++ loader, host only) against the oracle's
    independent PyYAML reader: species order, properties, particle count,
    transfer parameters; the particle scenario reader against the oracle's."""
    from oracle import amd as amd_mod
    for chem, aero, part in (("simpol_gas.yaml", "simpol_aero.yaml", "simpol_particles.yaml"),
                             ("cb05_soa_gas.yaml", "cb05_soa_aero.yaml", "cb05_soa_particles.yaml")):
        kmd = pkg.KineticModelData(golden(chem))
        amd = pkg.AerosolModelData(golden(aero), kmd)
        ref = amd_mod.load(golden(aero), list(kmd.speciesNames))
        assert amd.speciesNames == ref.names and amd.nParticles == ref.nParticles and amd.nSimpol == len(ref.gas)
        np.testing.assert_array_equal(amd.density, ref.density)
        np.testing.assert_array_equal(amd.molecularWeights, ref.mw)
        np.testing.assert_array_equal(amd.kappa, ref.kappa)
        state = np.zeros((1, amd.stateDim))
        nc = amd.scenarioConditionParticles(golden(part), state)
        state_ref = np.zeros((1, amd.stateDim))
        nc_ref = amd_mod.load_particles(golden(part), ref, kmd.nSpec, state_ref)
        np.testing.assert_array_equal(nc, nc_ref)
        np.testing.assert_array_equal(state, state_ref)
    amd.setNumberofParticles(3)
    assert amd.nParticles == 3 and amd.stateDim == 3 + kmd.nSpec + 3 * amd.nSpec
    # no particle block: nothing written
    st = np.zeros((1, amd.stateDim))
    assert amd.scenarioConditionParticles(golden("cb05_soa_gas.yaml"), st) is None and not st.any()


def test_aerosol_loader_errors(pkg, tmp_path):
    kmd = pkg.KineticModelData(golden("simpol_gas.yaml"))
    bad = open(golden("simpol_aero.yaml")).read().replace("gas-phase species: ethanol", "gas-phase species: nope")
    (tmp_path / "bad.yaml").write_text(bad)
    with pytest.raises(pkg.TChemError, match="nope"):
        pkg.AerosolModelData(str(tmp_path / "bad.yaml"), kmd)
    with pytest.raises(pkg.TChemError):
        pkg.AerosolModelData(str(tmp_path / "missing.yaml"), kmd)
