// Host-side aerosol model data: the NCAR-format aerosol file reader for
// single-particle SIMPOL.1 gas-particle partitioning.
//
// Mirrors AerosolModelData (src/core/TChem_AerosolModelData.cpp:60-249):
//   * aerosol species: CHEM_SPEC items with `phase: AEROSOL`, in file order,
//     with molecular weight, density and kappa (.cpp:85-104);
//   * gas species properties (diffusion coefficient, N star, molecular
//     weight) come from the CHEM_SPEC items without a phase (.cpp:105-110,
//     177-202);
//   * the particle count from AERO_REP_SINGLE_PARTICLE (.cpp:113-116);
//   * SIMPOL_PHASE_TRANSFER reactions of MECHANISM items (.cpp:121-166);
//     other reaction types are skipped.
// Unknown vector of a cell (src/core/problems/TChem_Impl_AerosolChemistry_Problem.hpp:80-85):
//   [active gas species (kinetic model order), particle 0 species, particle 1
//   species, ...]; the state row is [density, pressure, temperature, gas
//   species (active then constant), particle species].
#pragma once

#include <array>
#include <map>
#include <string>
#include <vector>

#include "kinetic_model.hpp"
#include "mechanism.hpp"

namespace tchem_amd {

struct SimpolParam {
  int gas;   // kinetic-model species index of the gas-phase species
  int aero;  // aerosol species index within a particle
  double B1, B2, B3, B4;
  double diffusion;  // m2 s-1
  double mw;         // gas-phase molecular weight, kg mol-1
  double Nstar;      // 999 when absent
  int computeAlpha;  // 1 when N star is given
};

struct AerosolModel {
  int nSpec = 0;         // aerosol species per particle
  int nParticles = -1;
  int nGasActive = 0, nGasConst = 0;
  std::vector<std::string> names;  // aerosol species, index order
  std::map<std::string, int> index;
  std::vector<double> mw, density, kappa;
  std::vector<SimpolParam> simpol;

  int nUnknowns() const { return nGasActive + nSpec * nParticles; }
  // state row length: density, pressure, temperature, gas species, particle species
  int stateDim(const KineticModel& km) const { return 3 + km.nSpec + nSpec * nParticles; }
};

// Parses the aerosol file against the kinetic model `km`; errors that make
// the reference abort throw std::runtime_error.
AerosolModel loadAerosolModel(const std::string& path, const KineticModel& km);

// `particles` block of a scenario file
// (AerosolModelData::scenarioConditionParticles, .cpp:263-300): fills
// numberConc [nBatch][nParticles] with each cell's number concentration and
// writes initial_value / number concentration into every particle's slot of
// the listed species in `state` rows [nBatch][stateDim].  Returns false when
// the file has no `particles` block (nothing written).
bool loadScenarioParticles(const std::string& path, const AerosolModel& am, const KineticModel& km, int nBatch,
                           std::vector<double>& state, std::vector<double>& numberConc);

// Elimination order of the aerosol problem's unknowns (gas species in the
// mechanism's internal order 0..m-1, then particle species m..mt-1): the
// particle species are eliminated first, then the gas species in their
// fill-reducing order.  Eliminating the gas species first would fill the
// particle rows with every particle's columns (all particles couple through
// the same gas species); this way the fill stays among the gas species that
// share a particle.
inline int aerosolPivot(int u, int m, int nParticleUnknowns) { return u < m ? u + nParticleUnknowns : u - m; }

// Mechanism tables of the aerosol problem for the sparse programs: `t` with
// m = mt = t.m + nSpec * nParticles and the Jacobian pattern, in
// elimination order (aerosolPivot), extended by every transfer derivative
// (entries without gas-phase contributions).  `entries` receives the
// (row, col) of each transfer derivative, in elimination order, in the order
// the integrator adds them: per (particle, transfer) pair (g,g), (a,g),
// (g, species i of the particle) for every i, (a, species i) for every i,
// where g is the gas species and a the particle species of the transfer.
MechTables aerosolTables(const MechTables& t, const AerosolModel& am, std::vector<std::array<int, 2>>* entries);

}  // namespace tchem_amd
