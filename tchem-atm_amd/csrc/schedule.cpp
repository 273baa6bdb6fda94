// Schedule compiler for the sparse per-cell linear algebra (schedule.hpp).
//
// Symbolic LU: elimination without interchanges in the internal (minimum
// Markowitz) species order; F = structure of I - J plus fill.  Pivot k
// depends on every earlier pivot k' < k whose row or column it touches
// (F(k,k') or F(k',k) nonzero); its level is one more than the highest
// level among those, so pivots of one level touch disjoint entries: the
// multipliers of a level read only final pivots, and its updates read only
// final multipliers and final rows of U.
#include "schedule.hpp"

#include <algorithm>
#include <map>
#include <stdexcept>

namespace tchem_amd {

namespace {

constexpr int kLanes = 64;

int packSpecies(int s, int pw) { return s < 0 ? 0xffff : (s | (pw << 16)); }
int pack16(int lo, int hi) { return lo | (hi << 16); }

struct Chain {
  int target = -1;
  std::vector<int> a;      // packed or single index per entry
  std::vector<double> c;   // coefficients (phases that have them)
};

// Lays chains into rounds of 64 lanes, longest chains first; appends to the
// round (offset, length) list, the per-lane targets and the entries.  The
// offset counts 64-entry slots from the start of `entries`.
void layRounds(std::vector<Chain> chains, std::vector<int>& round, std::vector<int>& target, std::vector<int>& entries,
               std::vector<double>* coef, int pad) {
  std::stable_sort(chains.begin(), chains.end(),
                   [](const Chain& x, const Chain& y) { return x.a.size() > y.a.size(); });
  for (size_t base = 0; base < chains.size(); base += kLanes) {
    const size_t n = std::min(size_t(kLanes), chains.size() - base);
    int len = 0;
    for (size_t l = 0; l < n; ++l) len = std::max(len, int(chains[base + l].a.size()));
    round.push_back(int(entries.size() / kLanes));
    round.push_back(len);
    for (int l = 0; l < kLanes; ++l) target.push_back(size_t(l) < n ? chains[base + l].target : -1);
    for (int k = 0; k < len; ++k)
      for (int l = 0; l < kLanes; ++l) {
        const Chain* ch = size_t(l) < n ? &chains[base + l] : nullptr;
        const bool live = ch && size_t(k) < ch->a.size();
        entries.push_back(live ? ch->a[k] : pad);
        if (coef) coef->push_back(live ? ch->c[k] : 0.0);
      }
  }
}

int rounds(const std::vector<int>& round) { return int(round.size() / 2); }

}  // namespace

SparseSchedule buildSparseSchedule(const MechTables& t) {
  SparseSchedule s;
  const int m = t.m;
  if (t.maxReactants > 3) throw std::runtime_error("sparse schedule: more than three reactants in a reaction");
  if (t.nSpec >= 0xffff || t.nReac >= (1 << 24)) throw std::runtime_error("sparse schedule: too many species or reactions");
  s.m = m;

  // ---- symbolic LU of I - J, no interchanges
  std::vector<char> F(size_t(m) * m, 0);
  for (int i = 0; i < m; ++i) F[size_t(i) * m + i] = 1;
  for (size_t n = 0; n < t.jnzRow.size(); ++n) F[size_t(t.jnzRow[n]) * m + t.jnzCol[n]] = 1;
  for (int k = 0; k < m; ++k)
    for (int i = k + 1; i < m; ++i) {
      if (!F[size_t(i) * m + k]) continue;
      for (int j = k + 1; j < m; ++j)
        if (F[size_t(k) * m + j]) F[size_t(i) * m + j] = 1;
    }
  std::vector<int> pos(size_t(m) * m, -1);
  s.rowPtr.assign(m + 1, 0);
  for (int i = 0; i < m; ++i) {
    for (int j = 0; j < m; ++j)
      if (F[size_t(i) * m + j]) {
        pos[size_t(i) * m + j] = int(s.colIdx.size());
        s.colIdx.push_back(j);
      }
    s.rowPtr[i + 1] = int(s.colIdx.size());
  }
  s.nnzF = int(s.colIdx.size());
  s.zeroSlot = s.nnzF + m;
  if (s.zeroSlot >= 0xffff) throw std::runtime_error("sparse schedule: factor pattern too large for 16-bit positions");
  s.diagIdx.resize(m);
  for (int k = 0; k < m; ++k) s.diagIdx[k] = pos[size_t(k) * m + k];
  auto Fz = [&](int i, int j) { return F[size_t(i) * m + j] != 0; };
  auto P = [&](int i, int j) { return pos[size_t(i) * m + j]; };
  const int rhs0 = s.nnzF;
  const int Z = s.zeroSlot;

  // ---- rates of progress
  const int R = t.maxReactants;
  auto reactantField = [&](int r, int j) {
    if (j >= R) return 0xffff;
    return packSpecies(t.reacSpec[size_t(r) * R + j], t.reacPow[size_t(r) * R + j]);
  };
  s.ropRounds = (t.nReac + kLanes - 1) / kLanes;
  // padding lanes: reaction nReac (rate constant 0, no reactants), which
  // writes 0 into the zero slot rop[nReac]
  s.ropRec.assign(size_t(s.ropRounds) * kLanes * 4, 0xffff);
  for (int r = 0; r < s.ropRounds * kLanes; ++r) {
    int* rec = &s.ropRec[size_t(r) * 4];
    rec[0] = r < t.nReac ? r : t.nReac;
    for (int j = 0; j < 3; ++j) rec[1 + j] = r < t.nReac ? reactantField(r, j) : 0xffff;
  }

  // ---- source-term segments
  {
    std::vector<Chain> ch;
    for (int g = 0; g < t.nSeg(); ++g) {
      Chain c;
      c.target = g;
      for (int e = t.segBegin[g]; e < t.segEnd[g]; ++e) {
        c.a.push_back(t.omegaReac[e]);
        c.c.push_back(t.omegaCoef[e]);
      }
      ch.push_back(c);
    }
    layRounds(ch, s.segRound, s.segTarget, s.segReac, &s.segCoef, t.nReac);
    s.segRounds = rounds(s.segRound);
  }

  // ---- derivative pairs
  {
    std::vector<int> pairReac(t.nPairs());
    for (int r = 0; r < t.nReac; ++r)
      for (int p = t.pairPtr[r]; p < t.pairPtr[r + 1]; ++p) pairReac[p] = r;
    s.pairRounds = (t.nPairs() + kLanes - 1) / kLanes;
    s.pairSlots = std::max(t.nPairs() + 1, s.pairRounds * kLanes);
    s.pairRec.assign(size_t(s.pairRounds) * kLanes * 4, 0xffff);
    for (int p = 0; p < s.pairRounds * kLanes; ++p) {
      int* rec = &s.pairRec[size_t(p) * 4];
      if (p >= t.nPairs()) {  // padding: d = kf[nReac] * 1 * c_0^0 = 0
        rec[0] = t.nReac;
        rec[1] = packSpecies(0, 1);
        continue;
      }
      const int r = pairReac[p];
      rec[0] = r | (t.pairSlot[p] << 24);
      for (int j = 0; j < 3; ++j) rec[1 + j] = reactantField(r, j);
    }
  }

  // ---- Jacobian entries: every position of F is a target
  {
    std::vector<Chain> ch;
    std::vector<int> jn(size_t(m) * m, -1);
    for (int n = 0; n < t.nnzJ(); ++n) jn[size_t(t.jnzRow[n]) * m + t.jnzCol[n]] = n;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) {
        if (!Fz(i, j)) continue;
        Chain c;
        c.target = P(i, j) | ((i == j ? 1 : 0) << 30);
        const int n = jn[size_t(i) * m + j];
        if (n >= 0)
          for (int e = t.jnzPtr[n]; e < t.jnzPtr[n + 1]; ++e) {
            c.a.push_back(t.jcPair[e]);
            c.c.push_back(t.jcCoef[e]);
          }
        ch.push_back(c);
      }
    layRounds(ch, s.jacRound, s.jacTarget, s.jacPair, &s.jacCoef, t.nPairs());
    s.jacRounds = rounds(s.jacRound);
  }

  // ---- factorisation levels
  std::vector<int> lev(m, 0);
  for (int k = 0; k < m; ++k) {
    int L = 0;
    for (int kp = 0; kp < k; ++kp)
      if (Fz(k, kp) || Fz(kp, k)) L = std::max(L, lev[kp]);
    lev[k] = L + 1;
  }
  s.nLevels = m ? *std::max_element(lev.begin(), lev.end()) : 0;
  const int T = Z + 1;
  s.trashSlot = T;
  if (T >= 0x8000) throw std::runtime_error("sparse schedule: factor pattern too large for 15-bit positions");

  // Packs chains (target, entries) into rounds of 64 lanes with at most
  // kMaxChain entries per lane; a long chain continues in a later round of
  // the same group (its partial sum is in the target entry).
  auto emitChains = [&](std::vector<Chain> chains, int kind, unsigned padEntry) {
    int emitted = 0;
    for (auto& c : chains) std::reverse(c.a.begin(), c.a.end());  // pop_back yields the original order
    while (true) {
      std::vector<Chain*> live;
      for (auto& c : chains)
        if (!c.a.empty()) live.push_back(&c);
      if (live.empty()) break;
      std::stable_sort(live.begin(), live.end(),
                       [](const Chain* x, const Chain* y) { return x->a.size() > y->a.size(); });
      if (live.size() > size_t(kLanes)) live.resize(kLanes);
      const int len = std::min<int>(SparseSchedule::kMaxChain, int(live.front()->a.size()));
      std::vector<std::vector<int>> part(kLanes);
      for (size_t l = 0; l < live.size(); ++l)
        for (int j = 0; j < len && !live[l]->a.empty(); ++j) {
          part[l].push_back(live[l]->a.back());
          live[l]->a.pop_back();
        }
      s.progHdr.push_back(unsigned(len) | (unsigned(kind) << 16));
      for (int l = 0; l < kLanes; ++l)
        s.progBody.push_back(size_t(l) < live.size() ? unsigned(live[l]->target) : unsigned(T));
      for (int j = 0; j < len; ++j)
        for (int l = 0; l < kLanes; ++l)
          s.progBody.push_back(size_t(j) < part[l].size() ? unsigned(part[l][size_t(j)]) : padEntry);
      ++emitted;
    }
    return emitted;
  };

  const unsigned zz = unsigned(Z) | (unsigned(Z) << 16);
  s.luBody = 0;
  for (int L = 1; L <= s.nLevels; ++L) {
    // multiplier and pivot records
    std::vector<unsigned> recs;
    for (int k = 0; k < m; ++k) {
      if (lev[k] != L) continue;
      recs.push_back((0x8000u | unsigned(k)) | (unsigned(P(k, k)) << 16));
      for (int i = k + 1; i < m; ++i)
        if (Fz(i, k)) {
          recs.push_back(unsigned(P(i, k)) | (unsigned(P(k, k)) << 16));
          ++s.nMultipliers;
        }
    }
    for (size_t base = 0; base < recs.size(); base += kLanes) {
      s.progHdr.push_back(unsigned(SparseSchedule::kMul) << 16);
      for (int l = 0; l < kLanes; ++l)
        s.progBody.push_back(base + l < recs.size() ? recs[base + l] : (unsigned(T) | (unsigned(Z) << 16)));
      ++s.luRounds;
    }
    // updates grouped by target, contributions in pivot order; the RHS rides along as column m
    std::map<int, Chain> byTarget;
    for (int k = 0; k < m; ++k) {
      if (lev[k] != L) continue;
      for (int i = k + 1; i < m; ++i) {
        if (!Fz(i, k)) continue;
        for (int j = k + 1; j < m; ++j) {
          if (!Fz(k, j)) continue;
          Chain& c = byTarget[P(i, j)];
          c.target = P(i, j);
          c.a.push_back(pack16(P(i, k), P(k, j)));
          ++s.nUpdates;
        }
        Chain& c = byTarget[rhs0 + i];
        c.target = rhs0 + i;
        c.a.push_back(pack16(P(i, k), rhs0 + k));
      }
    }
    std::vector<Chain> ch;
    for (auto& e : byTarget) ch.push_back(e.second);
    s.luRounds += emitChains(ch, SparseSchedule::kUpd, zz);
  }

  // ---- back substitution, column oriented: when column k is final
  // (x_k = b_k * rdiag_k), every row i < k with U_ik != 0 gets
  // b_i -= U_ik * x_k.  Columns of one level touch disjoint rows of later levels.
  std::vector<int> bl(m, 0);
  for (int i = m - 1; i >= 0; --i) {
    int L = 0;
    for (int j = i + 1; j < m; ++j)
      if (Fz(i, j)) L = std::max(L, bl[j]);
    bl[i] = L + 1;
  }
  s.nBackLevels = m ? *std::max_element(bl.begin(), bl.end()) : 0;
  s.backBody = int(s.progBody.size());
  const unsigned zcol = unsigned(Z) | (unsigned(m) << 16);  // U = 0, column m: b = 0, rdiag = 0
  for (int L = 1; L <= s.nBackLevels; ++L) {
    std::map<int, Chain> byTarget;
    for (int k = 0; k < m; ++k) {
      if (bl[k] != L) continue;
      for (int i = 0; i < k; ++i)
        if (Fz(i, k)) {
          Chain& c = byTarget[rhs0 + i];
          c.target = rhs0 + i;
          c.a.push_back(pack16(P(i, k), k));
        }
    }
    std::vector<Chain> ch;
    for (auto& e : byTarget) ch.push_back(e.second);
    s.backRounds += emitChains(ch, SparseSchedule::kBack, zcol);
  }

  // ---- forward substitution with stored factors (unit lower), column oriented by LU level
  s.fwdBody = int(s.progBody.size());
  for (int L = 1; L <= s.nLevels; ++L) {
    std::map<int, Chain> byTarget;
    for (int k = 0; k < m; ++k) {
      if (lev[k] != L) continue;
      for (int i = k + 1; i < m; ++i)
        if (Fz(i, k)) {
          Chain& c = byTarget[rhs0 + i];
          c.target = rhs0 + i;
          c.a.push_back(pack16(P(i, k), rhs0 + k));
        }
    }
    std::vector<Chain> ch;
    for (auto& e : byTarget) ch.push_back(e.second);
    s.fwdRounds += emitChains(ch, SparseSchedule::kUpd, zz);
  }
  s.nFwdLevels = s.nLevels;
  // slack after the last round: the interpreter reads the next round's words
  // (up to 1 + kMaxChain slots) before it knows there is none
  s.progBody.insert(s.progBody.end(), size_t(kLanes) * (1 + SparseSchedule::kMaxChain), 0u);
  return s;
}

}  // namespace tchem_amd
