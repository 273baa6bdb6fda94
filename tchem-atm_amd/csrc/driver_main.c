/* TChem_Driver.x: exercises the coupling interface the way the reference's
 * DriverInterface ctest does (src/unit-tests/TChem_Driver/TChem_Driver.c and
 * CMakeLists.txt): one batch entry at 81060 Pa and 270.5 K, a few gas
 * species and one aerosol species per particle set, 1e4 particles m-3, one
 * 30 s step; writes gas_species_names.txt, aero_species_props.txt and
 * output.txt in the working directory (the files the ctest compares).
 *
 * usage: TChem_Driver.x gas.yaml aerosol.yaml solver.yaml */
#include <stdio.h>
#include <stdlib.h>

#include "../../include/tchem_driver.h"

int main(int argc, char* argv[]) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s gas.yaml aerosol.yaml solver.yaml\n", argv[0]);
    return 2;
  }
  initialize(argv[1], argv[2], argv[3], 1);
  const int len = TChem_getLengthOfStateVector();
  const int nPart = TChem_getNumberConcentrationVectorSize();
  const int nGas = TChem_getNumberOfSpecies();
  const int nAero = TChem_getNumberOfAeroSpecies();
  if (len < 0 || nPart < 0) {
    fprintf(stderr, "initialize failed: %s\n", TChem_getLastError());
    return 1;
  }
  double* state = (double*)calloc((size_t)len, sizeof(double));
  double* nc = (double*)calloc((size_t)(nPart > 0 ? nPart : 1), sizeof(double));
  char name[100];

  FILE* f = fopen("gas_species_names.txt", "w");
  for (int i = 0; i < nGas; ++i) {
    TChem_getSpeciesName(&i, name, sizeof name);
    fprintf(f, "%s\n", name);
  }
  fclose(f);

  /* density, pressure, temperature; then the entries the ctest sets */
  state[0] = 0.0;
  state[1] = 81060.0;
  state[2] = 270.5;
  if (len > 81) {
    state[50] = 0.06;
    state[51] = 1e-06;
    state[36] = 0.001;
    state[81] = 1e8;
  }
  for (int p = 0; p < nPart; ++p) {
    nc[p] = 10000.0;
    state[3 + nGas + nAero * p] = 1e-08; /* first aerosol species of particle p */
  }
  TChem_setStateVector(state, 0);
  TChem_setNumberConcentrationVector(nc, 0);

  f = fopen("aero_species_props.txt", "w");
  for (int i = 0; i < nAero; ++i) {
    TChem_getAerosolSpeciesName(&i, name, sizeof name);
    fprintf(f, "%s %f %f %f\n", name, TChem_getAerosolSpeciesDensity(&i), TChem_getAerosolSpeciesMW(&i),
            TChem_getAerosolSpeciesKappa(&i));
  }
  fclose(f);

  const double del_t = 30.0;
  TChem_doTimestep(&del_t);
  if (TChem_getLastError()[0]) {
    fprintf(stderr, "doTimestep failed: %s\n", TChem_getLastError());
    return 1;
  }
  TChem_getStateVector(state, 0);
  f = fopen("output.txt", "w");
  for (int i = 0; i < len; ++i) fprintf(f, "%e \n", state[i]);
  fclose(f);

  finalize();
  free(state);
  free(nc);
  return 0;
}
