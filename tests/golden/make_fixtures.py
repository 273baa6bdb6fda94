"""Copies the reference's test data into tests/golden/ (data files only).

Run from the repository root with the reference checked out at
/root/reference:  python tests/golden/make_fixtures.py

Every fixture keeps its origin path in ORIGINS.txt.  The HDF5 golden is
converted to .npy with the h5py that ships in /opt/conda (data read only).
"""
import os
import shutil
import subprocess
import sys

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

FILES = {
    # mechanisms (inputs)
    "cb05cl_ae5.yaml": "src/examples/runs/atmospheric_chemistry/CB05CL_AE5/config_full_gas.yaml",
    "carbonbond05.yaml": "src/examples/runs/atmospheric_chemistry/CarbonBond05/config_full_gas.yaml",
    "arrhenius.yaml": "src/examples/runs/atmospheric_chemistry/Arrhenius/config_arrhenius/config_arrhenius.yaml",
    "troe.yaml": "src/examples/runs/atmospheric_chemistry/Troe/config_troe/config_troe.yaml",
    "photolysis.yaml": "src/examples/runs/atmospheric_chemistry/Photolysis/config_photolysis.yaml",
    "cmaq_h2o2.yaml": "src/examples/runs/atmospheric_chemistry/CMAQ_H2O2/config_cmaq_h2o2.yaml",
    "emission.yaml": "src/examples/runs/atmospheric_chemistry/Emission/config_emissions.yaml",
    "uci.yaml": "src/examples/runs/atmospheric_chemistry/uci_e3sm_v3/uci.yaml",
    "uci_v2.yaml": "src/examples/runs/atmospheric_chemistry/uci_e3sm_v3/uci_v2.yaml",
    "uci_v2_test3.yaml": "src/examples/runs/atmospheric_chemistry/uci_col/uci_v2_test3.yaml",
    # expected outputs
    "carbonbond05_full_gas.dat": "src/examples/runs/atmospheric_chemistry/CarbonBond05/full_gas.dat",
    "kmod_echo_reaction_rates.txt": "src/unit-tests/references/reaction_rates/kmod.echo",
    "kmod_echo_net_production_rates.txt": "src/unit-tests/references/net_production_rates/kmod.echo",
    "kmod_echo_rate_of_progress.txt": "src/unit-tests/references/rate_of_progress/kmod.echo",
    "kmod_echo_uci_col.txt": "src/unit-tests/references/net_production_rates/uci_col/kmod.echo",
    "kmod_echo_atmchem_e3sm_uci_col.txt": "src/unit-tests/references/atmospheric_chemistry_e3sm/uci_col/kmod.echo",
    "uci_col_full_gas.dat": "src/unit-tests/references/atmospheric_chemistry_e3sm/uci_col/full_gas.dat",
    "photolysis_results.txt": "src/examples/runs/atmospheric_chemistry/Photolysis/photolysis_results.txt",
    "troe_results.txt": "src/examples/runs/atmospheric_chemistry/Troe/troe_results.txt",
    "rxn_arrhenius_results.txt": "src/examples/runs/atmospheric_chemistry/Arrhenius/rxn_arrhenius_results.txt",
    "CMAQ_H2O2_results.txt": "src/examples/runs/atmospheric_chemistry/CMAQ_H2O2/CMAQ_H2O2_results.txt",
    "emission_results.txt": "src/examples/runs/atmospheric_chemistry/Emission/emission_results.txt",
    "emission_Emissions.dat": "src/examples/runs/atmospheric_chemistry/Emission/Emissions.dat",
    "e3sm_reaction_rates.dat": "src/unit-tests/references/reaction_rates/reaction-rates.dat",
    "e3sm_rate_of_progress.dat": "src/unit-tests/references/rate_of_progress/rate-of-progress.dat",
    "e3sm_net_production_rates.dat": "src/unit-tests/references/net_production_rates/net-production-rates.dat",
    "uci_col_net_production_rates.dat": "src/unit-tests/references/net_production_rates/uci_col/net_production_rates.dat",
    "uci_col_input_conditions.yaml": "src/examples/runs/atmospheric_chemistry/uci_col/input_conditions_col.yaml",
    # single-particle SIMPOL.1 aerosol (inputs and expected outputs)
    "simpol_gas.yaml": "src/examples/runs/atmospheric_chemistry/Simpol/config_gas.yaml",
    "simpol_aero.yaml": "src/examples/runs/atmospheric_chemistry/Simpol/test_SIMPOL_phase_transfer.yaml",
    "simpol_particles.yaml": "src/examples/runs/atmospheric_chemistry/Simpol/scenario_conditions_particle.yaml",
    "simpol_rhs_host.txt": "src/unit-tests/references/partmc_integration/simpol/Simpol_RHS_HOST.txt",
    "simpol_camp_results.txt": "src/examples/runs/atmospheric_chemistry/Simpol/camp_out/SIMPOL_phase_transfer_results.txt",
    "cb05_soa_gas.yaml": "src/examples/runs/atmospheric_chemistry/CB05CL_AE5_w_simpolSOA/config_full_gas.yaml",
    "cb05_soa_aero.yaml": "src/examples/runs/atmospheric_chemistry/CB05CL_AE5_w_simpolSOA/mechanism_aero.yaml",
    "cb05_soa_particles.yaml": "src/examples/runs/atmospheric_chemistry/CB05CL_AE5_w_simpolSOA/scenario_conditions_particle.yaml",
    "driver_config_solver.yaml": "src/unit-tests/TChem_Driver/config_solver.yaml",
    "driver_output_ref.txt": "src/unit-tests/TChem_Driver/references/output_ref.txt",
    "driver_gas_species_names_ref.txt": "src/unit-tests/TChem_Driver/references/gas_species_names_ref.txt",
    "driver_aero_species_props_ref.txt": "src/unit-tests/TChem_Driver/references/aero_species_props_ref.txt",
}

H5 = {"cb05cl_ae5_full_gas.npy": "src/unit-tests/CB05CL_AE5/references/full_gas.hdf5",
      "simpol_full_gas.npy": "src/unit-tests/TinesODE/references/full_gas.hdf5",
      "simpol_batched_full_gas.npy": "src/unit-tests/TinesODE/references/batched_full_gas.hdf5",
      "cb05_soa_batched_full_gas.npy": "src/unit-tests/TinesODE/references/batched_cb_full_gas.hdf5"}


def main():
    origins = []
    for dst, src in FILES.items():
        shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, dst))
        origins.append(f"{dst}\t{src}")
    for dst, src in H5.items():
        code = ("import h5py, numpy as np, sys; "
                "np.save(sys.argv[2], np.array(h5py.File(sys.argv[1], 'r')['ref']))")
        subprocess.run(["/opt/conda/bin/python3.9", "-c", code, os.path.join(REF, src), os.path.join(HERE, dst)],
                       check=True)
        origins.append(f"{dst}\t{src} (dataset 'ref', converted to .npy)")
    with open(os.path.join(HERE, "ORIGINS.txt"), "w") as fh:
        fh.write("# fixture\torigin under sandialabs/TChem-atm\n")
        fh.write("\n".join(origins) + "\n")


if __name__ == "__main__":
    sys.exit(main())
