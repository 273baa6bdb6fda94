// Compiled per-wavefront programs for the sparse form of the per-cell work:
// rates of progress, derivative pairs, source-term segment sums, Jacobian
// assembly (the evaluation program), and the LU factorisation of I - c h J in
// the fill-reducing species order with the pivot sequence fixed at compile
// time, back and forward substitution (the LU program).
//
// A program is a sequence of "rounds"; in a round each of the 64 lanes of
// the wavefront executes one record.  A round is a header word
// (len | kind << 16, read by the whole wavefront) and a body of 64-word
// slots, slot s of lane l at body[64 * s + l], so every load of a round is
// one conflict-free LDS access.  Both programs are copied into LDS once per
// workgroup and read by every wavefront (cell) of the workgroup.
//
// Per-cell matrix storage: the nnzF entries of the symbolic LU pattern F
// (structure of I - J plus fill) in row-major order, the m right-hand-side
// entries, a zero slot and a trash slot.  Positions fit in 15 bits, so two
// indices pack into one word (lo | hi << 16).
//
// Padding is branch free: idle lanes write trash slots, idle chain entries
// read zero slots (a zero coefficient, a zero rate, a zero matrix entry).
#pragma once

#include <vector>

#include "mechanism.hpp"

namespace tchem_amd {

struct SparseSchedule {
  enum : int { kMul = 0, kUpd = 1, kBack = 2, kRop = 3, kPair = 4, kSeg = 5, kJac = 6, kNuT = 7, kJtv = 8,
               kMulUpd = 9 };
  static constexpr int kMaxChain = 4;            // chain entries per lane per round
  static constexpr int kMaxMulUpdChain = kMaxChain - 2;  // chain entries of a kMulUpd lane
  static constexpr int kSlack = 1 + kMaxChain;   // slots read past a section's last round
  static constexpr unsigned kCont = 0x80000000u;  // target flag: continues a chain split over rounds
  static constexpr unsigned kNone = 0xffffu;      // packed reactant field: no reactant

  int m = 0, nnzF = 0;
  int zeroSlot = 0, trashSlot = 0;  // nnzF + m, nnzF + m + 1
  int nLevels = 0, nBackLevels = 0;
  int nMultipliers = 0, nUpdates = 0;
  std::vector<int> rowPtr, colIdx;  // CSR of F (internal order)
  std::vector<int> diagIdx;         // [m] position of (k,k)

  // ---- evaluation program
  //   kRop:  2 words: (r | fa << 16), (fb | fc << 16); rop[r] = kf[r] * fa * fb * fc,
  //          field f = species | power << 12 (kNone: factor 1); padding r = nReac (kf = 0)
  //   kPair: 2 words: (r | slot << 14 | fa << 16), (fb | fc << 16); derivative of
  //          rop[r] by the reactant in `slot`, written to dpr[64 * round + lane]
  //   kSeg:  target (segment | kCont), len entry words (reaction | coef << 16):
  //          seg[t] = (cont ? seg[t] : 0) + sum coefTable[coef] * rop[reaction]
  //   kJac:  target (position | diag << 30 | kCont), entries (pair | coef << 16):
  //          A[t] = (cont ? A[t] : diag) + scale * sum coefTable[coef] * dpr[pair]
  // Sections in order: rop, pair, seg, jac.  Padding: seg target nSeg (trash),
  // jac target trashSlot, entries (zero slot | coefficient 0).
  // Transposed programs (adjoint) assemble and factor I - c h J^T instead
  // and append two sections:
  //   kNuT:  target (reaction | kCont), entries (species | coef << 16):
  //          w[t] = (cont ? w[t] : 0) + sum coefTable[coef] * mu[species]   (nu_r . mu)
  //   kJtv:  target (species | kCont), entries (pair | reaction << 16):
  //          v[t] = (cont ? v[t] : 0) + sum dpr[pair] * w[reaction]         (J^T mu)
  //          padding: w target nReac + 1, v target m; entries read the zero
  //          slots mu[m], dpr[nPairs], w[nReac]
  bool transposed = false;
  std::vector<double> coefTable;            // [0] = 0
  std::vector<unsigned> evalHdr, evalBody;  // evalBody ends with kSlack x 64 words of slack
  int ropRounds = 0, pairRounds = 0, segRounds = 0, jacRounds = 0, nutRounds = 0, jtvRounds = 0;
  int pairBody = 0, segBody = 0, jacBody = 0, nutBody = 0, jtvBody = 0;  // body word offsets of the sections
  int pairSlots = 0;                           // derivative slots of a cell (zero slot nPairs included)

  // ---- LU program
  //   kMul:  1 word (dst | src << 16): dst < 0x8000: A[dst] *= 1 / A[src];
  //          dst = 0x8000 | k: pivot record, rdiag[k] = 1 / A[src] (a zero or
  //          non-finite pivot fails the factorisation); padding (trash | zero)
  //   kUpd:  target, len entries (lo | hi << 16): A[t] -= sum A[lo] * A[hi]
  //   kMulUpd (len 2 + c, c <= kMaxMulUpdChain): target t, (ik | kj << 16),
  //          (kk | mul << 15 | d << 16), then c entries (lo | hi << 16):
  //          l = A[ik] * (1 / A[kk]) (0 when A[ik] = 0); multiplier lanes
  //          (mul) store A[t = ik] = l, update lanes
  //          A[t] -= sum A[lo] * A[hi] + l * A[kj]; rdiag[d] = 1 / A[kk]
  //          (d = k on one lane per factored pivot, m + 1 on the others).
  //          Update lanes read A[ik] before the multiplier lane's store (all
  //          lanes of a round read before any writes).  A lane with chain
  //          entries only has ik = kj = kk = zero.  Padding (trash, zero |
  //          zero, zero | (m + 1) << 16, zero | zero).
  //   kBack: target, len entries (U position | k << 16):
  //          A[t] -= sum A[lo] * (A[rhs0 + k] * rdiag[k]); padding (zero | m << 16)
  // Sections in order: factorisation with the RHS eliminated alongside (LU),
  // column-oriented back substitution leaving b_i = u_ii x_i (BACK; then
  // x_i = b_i * rdiag_i), forward substitution with stored factors (FWD).
  std::vector<unsigned> progHdr, progBody;  // progBody ends with kSlack x 64 words of slack
  int luRounds = 0, backRounds = 0, fwdRounds = 0;
  int backBody = 0, fwdBody = 0;

  // LDS slot of each factor position (CSR order of F): the positions are
  // renumbered so that the lanes of a round touch distinct LDS banks
  // (bankAwareSlots); every program word holds slots.  RHS, zero and trash
  // slots keep their numbers.
  std::vector<int> slotOf;
  // extra LDS passes (bank conflicts, 32 banks of 4 bytes per half
  // wavefront) summed over the matrix accesses of one pass through the LU
  // program and the Jacobian rounds, before and after the renumbering
  long long bankPassesBefore = 0, bankPassesAfter = 0;
};

// Builds both programs.  Requires at most three reactants per reaction,
// nSpec < 4095, reactant powers < 16, nReac < 16384, at most 65535 distinct
// stoichiometric coefficients and nnzF + m + 2 < 32768; throws
// std::runtime_error otherwise.
SparseSchedule buildSparseSchedule(const MechTables& t, bool transposed = false);

}  // namespace tchem_amd
