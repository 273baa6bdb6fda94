// Mechanism compiler: turns a KineticModel into the flat, index-based tables
// the HIP kernels read.  All tables are shared by every cell; per-cell data
// (state rows, photolysis rates) never enter here.
//
// Layout choices (the reference scatters with atomics, one team per cell,
// src/core/impl/TChem_Impl_ReactionRatesAerosol.hpp:116-138):
//   * rate constants: one record per reaction (type, 8 parameters, conv
//     exponent, photolysis slot) so a lane evaluates any reaction without
//     branching on per-type index lists;
//   * rate of progress: reactant species/power table [nReac][maxReactants];
//   * source term: species-major CSR of (reaction, net coefficient) — a
//     deterministic gather instead of atomic scatter;
//   * Jacobian: per-reaction list of derivative "pairs" (reaction, active
//     reactant) and a CSR over the structural nonzeros of J holding
//     (pair, coefficient) contributions.  Every diagonal entry is a
//     structural nonzero so the iteration matrix I - c h J is written in one
//     pass.
#pragma once

#include <string>
#include <vector>

#include "kinetic_model.hpp"

namespace tchem_amd {

// kRateExternal: JPL / ratio-JPL / PHOTO_RATE reactions, whose rate constants
// only the E3SM kinetics kernel computes (the box model then takes them from it)
enum RateKind : int { kRateArrhenius = 0, kRateTroe = 1, kRateCmaqH2O2 = 2, kRateExternal = 3 };

struct MechTables {
  int nSpec = 0, nConst = 0, m = 0, nReac = 0, nPhoto = 0;
  int maxReactants = 0;
  double convPPM = kConvPPM;

  // Active species are renumbered in a fill-reducing elimination order
  // (greedy minimum Markowitz count on the structure of I - J).  Every table
  // below uses the internal numbering; perm[a] is the reference's species
  // index of internal species a, iperm its inverse.  Constant species keep
  // their indices (>= m).
  std::vector<int> perm, iperm;  // [m]

  // rate constants, per reaction
  bool externalRates = false;    // some rate constants need the E3SM kinetics (kRateExternal)
  std::vector<int> rateKind;     // RateKind
  std::vector<double> ratePar;   // [nReac][8]
  std::vector<int> convExp;      // [nReac]
  std::vector<int> photoSlot;    // [nReac], -1 or index into the per-cell photolysis rates
  std::vector<int> photoReac;    // [nPhoto] reaction index of each photolysis slot

  // rate of progress
  std::vector<int> reacSpec;  // [nReac][maxReactants], -1 padded
  std::vector<int> reacPow;   // [nReac][maxReactants]

  // source term gather (species-major), emission per active species.  Rows
  // are cut into segments of at most kOmegaSegment entries so long rows
  // (radicals in ~100 reactions) spread over several lanes; segment sums are
  // added per row in segment order.
  static constexpr int kOmegaSegment = 16;
  std::vector<int> omegaPtr;     // [m+1]
  std::vector<int> omegaReac;    // [nnzOmega]
  std::vector<double> omegaCoef; // [nnzOmega]
  std::vector<int> segBegin, segEnd;  // [nSeg] entry ranges
  std::vector<int> rowSegPtr;         // [m+1] segments of row a
  std::vector<double> emission;  // [m]

  // Jacobian
  std::vector<int> pairPtr;      // [nReac+1] pairs of reaction r
  std::vector<int> pairSlot;     // [nPairs] reactant slot j within the reaction
  std::vector<int> jnzRow, jnzCol;   // [nnzJ]
  std::vector<int> jnzPtr;       // [nnzJ+1]
  std::vector<int> jcPair;       // [nContrib]
  std::vector<double> jcCoef;    // [nContrib]

  // Serial forms of the source-term gather and of the Jacobian for the
  // dense LU form: every reactant and product entry of the reference's
  // tables, uncombined (zero yields included), in the order the reference's
  // loops add them (reaction order; within a reaction the reactants, then
  // the products; for J, the reactant slot of the derivative before the
  // target entry), src/core/impl/TChem_Impl_ReactionRatesAerosol.hpp:116-138.
  // Summed one entry at a time, they give the reference's serial sums bit
  // for bit.  Rows and columns in the internal numbering.
  std::vector<int> serOmegaPtr;     // [m+1]
  std::vector<int> serOmegaReac;    // [nSerOmega]
  std::vector<double> serOmegaCoef; // [nSerOmega]
  std::vector<int> serJRow, serJCol; // [nSerJ] structural entries, row-major
  std::vector<int> serJPtr;         // [nSerJ+1]
  std::vector<int> serJPair;        // [nSerContrib]
  std::vector<double> serJCoef;     // [nSerContrib]

  int nPairs() const { return int(pairSlot.size()); }
  int nSerJ() const { return int(serJRow.size()); }
  int nSeg() const { return int(segBegin.size()); }
  int nnzJ() const { return int(jnzRow.size()); }
};

// Throws std::runtime_error for mechanisms the gas-phase box model cannot
// evaluate (JPL / R_JPL_ARRHENIUS / PHOTO_RATE reactions belong to the E3SM
// kinetics path).
MechTables compileMechanism(const KineticModel& km);

// Greedy minimum-Markowitz elimination order of the m x m structure `pattern`
// (row-major booleans, diagonal included): at each step the remaining index
// with the smallest (row count - 1) * (column count - 1) is eliminated (ties:
// lowest index) and its fill is added.  Returns the order (position -> index).
std::vector<int> markowitzOrder(int m, std::vector<char> pattern);

}  // namespace tchem_amd
