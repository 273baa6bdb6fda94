// Compiled per-wavefront schedules for the sparse form of the per-cell
// linear algebra: rates of progress, source term, Jacobian assembly, LU
// factorisation of I - c h J in the fill-reducing species order with the
// pivot sequence fixed at compile time (no row interchanges), and the
// triangular solves.
//
// Every phase is a list of "rounds"; in a round each of the 64 lanes of the
// wavefront executes at most one record (a reaction, a segment, a matrix
// entry ...).  Records carry every index the lane needs, laid out
// [slot][lane] so the loads of a round are one coalesced access with no
// pointer chasing.  Chains of work on one target (a sum over contributions,
// the updates one LU entry receives within a level) are padded to the
// longest chain of their round with entries that read a zero slot.
//
// Matrix storage (LDS, per cell): the nnzF entries of the symbolic LU
// pattern F (structure of I - J plus fill) in row-major order, then the m
// right-hand-side entries, then one zero entry.  Positions fit in 16 bits,
// so a pair of positions packs into one int (lo | hi << 16).
#pragma once

#include <vector>

#include "mechanism.hpp"

namespace tchem_amd {

struct SparseSchedule {
  int m = 0, nnzF = 0;  // matrix entries; RHS at [nnzF, nnzF+m); zero at nnzF+m
  int zeroSlot = 0;     // = nnzF + m
  int nLevels = 0;      // factorisation levels
  int nFwdLevels = 0;   // forward-substitution levels (factors reused)
  int nBackLevels = 0;  // back-substitution levels

  // structure (internal order), for tests and diagnostics
  std::vector<int> rowPtr, colIdx;  // CSR of F
  std::vector<int> diagIdx;         // [m] position of (k,k)
  int nMultipliers = 0, nUpdates = 0;

  // rate of progress: one int4 per (round, lane):
  //   x = reaction (nReac on padding lanes: rate constant 0), y/z/w = species | power << 16 (0xffff: none)
  int ropRounds = 0;
  std::vector<int> ropRec;  // [ropRounds*64*4]

  // chained phases share one layout: per round g an int2 (offset, length)
  // in units of 64-entry slots, a target per lane ([g*64 + lane], -1 none)
  // and chain entries at [(offset + k) * 64 + lane].

  // source-term segments: target = segment id; entries (reaction or the
  // rate-of-progress zero slot nReac, coefficient)
  int segRounds = 0;
  std::vector<int> segRound, segTarget, segReac;
  std::vector<double> segCoef;

  // Jacobian derivative pairs: int4 per (round, lane): x = reaction | slot << 24, y/z/w as ropRec;
  // lane l of round g writes derivative slot g*64 + l (padding lanes write 0);
  // pairSlots = derivative slots a cell needs (zero slot nPairs included)
  int pairRounds = 0, pairSlots = 0;
  std::vector<int> pairRec;

  // Jacobian entries: target = matrix position | diag << 30 (-1 none); every
  // position of F is a target (fill positions with an empty chain, so the
  // assembly also clears them); entries (pair or the derivative zero slot
  // nPairs, coefficient)
  int jacRounds = 0;
  std::vector<int> jacRound, jacTarget, jacPair;
  std::vector<double> jacCoef;

  // factorisation, per level: multiplier rounds, then update rounds.
  // multiplier records, int2 per (round, lane) = (dst, src):
  //   dst >= 0: A[dst] *= 1 / A[src]; dst < 0: pivot record, rdiag[-dst-1] = 1 / A[src]; src = -1: none
  std::vector<int> levMul, levUpd;  // int2 (first round, end round) per level
  std::vector<int> mulRec;
  // updates: target position (matrix entry or RHS entry), entries packed (l | u << 16): target -= A[l] * A[u]
  std::vector<int> updRound, updTarget, updPair;

  // forward substitution with stored factors (RHS target rhs0 + i, entries (L_ik | b_k << 16))
  std::vector<int> fwdLev, fwdRound, fwdTarget, fwdPair;
  // back substitution: target row i, entries (U_ij | x_j << 16); x_i = (b_i - sum) * rdiag_i
  std::vector<int> backLev, backRound, backTarget, backPair;
};

// Builds the schedules; requires at most three reactants per reaction and
// nnzF + m + 1 <= 65535.  Throws std::runtime_error otherwise.
SparseSchedule buildSparseSchedule(const MechTables& t);

}  // namespace tchem_amd
