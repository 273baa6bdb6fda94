// Host-side kinetic model data: the NCAR-format mechanism reader.
//
// Mirrors the tables that KineticModelData::initChemNCAR builds
// (src/core/TChem_KineticModelData.cpp:80-1017) with the same index
// conventions, so the loader echo (kmod.echo) and every downstream table have
// the reference's layout:
//   * species: active species first, constant species appended
//     (.cpp:110-131); names truncated to 18 characters (TChem_Util.hpp:133);
//   * reactant entries ordered by species index, stored with negated
//     coefficients in columns [0, maxSpecInReac/2); product entries at column
//     offset maxSpecInReac/2 (.cpp:314-338);
//   * per-type parameter records (.cpp:391-783), emission sources
//     (.cpp:891-937) and convExponent = sum|nu_reactant| - 1 (.cpp:939-947).
#pragma once

#include <array>
#include <map>
#include <ostream>
#include <string>
#include <vector>

namespace tchem_amd {

// Constants from src/core/TChem_Util.hpp:183,195
constexpr double kConvPPM = 7.24296358205307E+10;
constexpr double kBoltzmann = 1.3806504E-23;
constexpr int kLengthOfSpecName = 18;

struct ArrheniusParam {
  int reaction;
  double A, B, C, D, E;
};
struct CmaqH2O2Param {
  int reaction;
  double A1, B1, C1, A2, B2, C2;
};
struct JplParam {
  int reaction;
  double k0_A, k0_B, k0_C, kinf_A, kinf_B, kinf_C, Fc;
};
struct RJplArrheniusParam {
  int reaction;
  double k0_A, k0_B, k0_C, kinf_A, kinf_B, kinf_C, Fc, A, B, C;
};
struct EmissionParam {
  int species;
  double rate;
};
struct AdjustReactionParam {
  int reaction;
  int species;
};
struct ProdO1DParam {
  int reactions[3];
  int photolysis_reaction;
  int species1, species2, species3;
  double A1, C1, A2, C2, A3, C3;
};

enum ReactionType : int {
  kArrhenius = 0,
  kTroe = 1,
  kCmaqH2O2 = 2,
  kJpl = 3,
  kRJplArrhenius = 4,
  kPhotoRate = 5,
};

struct KineticModel {
  // sizes
  int nSpec = 0;       // active + constant species
  int nConstSpec = 0;  // constant (tracer) species appended at the end
  int nReac = 0;
  int maxSpecInReac = 0;
  int nFallReac = 0;
  int M_index = -1;
  double CONV_PPM = kConvPPM;

  std::vector<std::string> speciesNames;  // truncated to 18 chars
  std::map<std::string, int> speciesIndex;

  // per reaction
  std::vector<int> reactionType;  // ReactionType
  std::vector<int> isRev, reacNrp, reacNreac, reacNprod, convExponent;
  std::vector<int> reacSidx;       // [nReac][maxSpecInReac]
  std::vector<double> reacNuki;    // [nReac][maxSpecInReac]
  std::vector<std::string> note;   // the reaction's `note` field ("" if absent)

  std::vector<ArrheniusParam> arrhenius;
  std::vector<CmaqH2O2Param> cmaqH2O2;
  std::vector<JplParam> jpl;
  std::vector<RJplArrheniusParam> rJplArrhenius;
  std::vector<int> reacPfal;                    // [nFallReac]
  std::vector<std::array<double, 8>> reacPpar;  // k0_A*1e6, k0_B, k0_C, kinf_A, kinf_B, kinf_C, Fc, N
  std::vector<EmissionParam> emission;
  std::vector<AdjustReactionParam> adjustReaction;
  std::vector<ProdO1DParam> prodO1D;
  int countPhotoRate = 0;

  // Reactions whose rate constant is a photolysis rate: ARRHENIUS reactions
  // whose note starts with "PHOTOLYSIS" (the CB05 convention: "A := photolysis
  // rate"), in reaction order.  Per-cell photolysis rates replace their A.
  std::vector<int> photolysisReactions;

  int nActive() const { return nSpec - nConstSpec; }
  int sidx(int r, int j) const { return reacSidx[size_t(r) * maxSpecInReac + j]; }
  double nuki(int r, int j) const { return reacNuki[size_t(r) * maxSpecInReac + j]; }
};

// Parse `path`; writes the loader echo to `echo` when non-null.  Errors that
// make the reference call exit(1) throw std::runtime_error here.
KineticModel loadKineticModel(const std::string& path, std::ostream* echo);

// Scenario reader: pressure/temperature per cell and the `initial_state`
// block, with the reference's unit handling
// (src/core/TChem_AtmosphericChemistry.cpp:429-504).  Returns state rows
// [nBatch][nSpec+3] = [density, pressure, temperature, species...].
std::vector<double> loadScenarioState(const std::string& path, const KineticModel& km, int* nBatch);

// `photolysis_rates` block (r_0, r_1, ... each with one value per cell) of a
// scenario file (src/core/TChem_AtmosphericChemistry.cpp:511-560): returns
// [nBatch][count] (count 0 when the block is absent).
std::vector<double> loadScenarioPhotolysisRates(const std::string& path, int nBatch, int* count);

// `external_forcing` block (species_name + one value per cell) of a scenario
// file (src/core/TChem_AtmosphericChemistry.cpp:562-631): returns
// [nBatch][nActive], zero for species without forcing.
std::vector<double> loadScenarioExternalForcing(const std::string& path, const KineticModel& km, int nBatch,
                                                int* count);

}  // namespace tchem_amd
