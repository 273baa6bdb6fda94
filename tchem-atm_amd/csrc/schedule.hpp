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

  // ---- LU / substitution program (copied into LDS, shared by the cells of
  // a workgroup).  A sequence of rounds; header word per round = len | kind << 16:
  //   kMul: 64 words (dst | src << 16): dst < 0x8000: A[dst] *= 1 / A[src];
  //         dst = 0x8000 | k: pivot record, rdiag[k] = 1 / A[src] (zero or
  //         non-finite pivot: factorisation fails);
  //   kUpd: 64 target words, then len x 64 entry words (lo | hi << 16):
  //         A[target] -= sum A[lo] * A[hi];
  //   kBack: as kUpd with entries (U position | column k << 16):
  //         A[target] -= sum A[lo] * (A[rhs0 + k] * rdiag[k]).
  // Entry words of one lane are at [(round body + 64 * (1 + j)) + lane].
  // Padding is branch free: idle lanes name the trash slot (zeroSlot + 1)
  // as target (multiplier padding: dst = trash, src = zero slot) and idle
  // chain entries read the zero slot (column m for kBack, whose RHS entry and
  // reciprocal pivot are zero).  Sections, in order: factorisation with the RHS
  // eliminated alongside (LU), column-oriented back substitution leaving
  // b_i = u_ii x_i (BACK: x_i = b_i * rdiag_i afterwards), and forward
  // substitution with stored factors (FWD, kUpd rounds).
  enum : int { kMul = 0, kUpd = 1, kBack = 2 };
  static constexpr int kMaxChain = 4;  // entries per lane per round
  int trashSlot = 0;                   // zeroSlot + 1
  std::vector<unsigned> progHdr, progBody;  // progBody ends with (1 + kMaxChain) x 64 words of slack
  int luRounds = 0, backRounds = 0, fwdRounds = 0;  // rounds per section
  int luBody = 0, backBody = 0, fwdBody = 0;        // body word offset of each section
};

// Builds the schedules; requires at most three reactants per reaction and
// nnzF + m + 2 < 32768.  Throws std::runtime_error otherwise.
SparseSchedule buildSparseSchedule(const MechTables& t);

}  // namespace tchem_amd
