// Program compiler for the sparse per-cell work (schedule.hpp).
//
// Symbolic LU: elimination without interchanges in the internal (minimum
// Markowitz) species order; F = structure of I - J plus fill.  Pivot k
// depends on every earlier pivot k' < k whose row or column it touches
// (F(k,k') or F(k',k) nonzero); its level is one more than the highest
// level among those, so pivots of one level touch disjoint entries.  The
// forward-substitution section runs level by level; the factorisation is
// list-scheduled (entries become final as their contributions are applied,
// rounds take whatever is ready), with the levels as its urgency measure.
#include "schedule.hpp"

#include <algorithm>
#include <array>
#include <map>
#include <stdexcept>

namespace tchem_amd {

namespace {

constexpr int kLanes = 64;

unsigned pack16(unsigned lo, unsigned hi) { return lo | (hi << 16); }

// A chain of entries accumulated into one target.
struct Chain {
  unsigned target = 0;
  std::vector<unsigned> a;  // entry words in accumulation order
};

// Appends rounds of `kind` to (hdr, body): up to 64 chains per round with
// at most kMaxChain entries each, longest remaining chains first; a chain
// longer than that continues in later rounds, its target word then carries
// kCont when `markCont`.  Idle lanes get `padTarget`, idle entries
// `padEntry`.  Returns the number of rounds.
int emitChains(std::vector<Chain> chains, int kind, unsigned padTarget, unsigned padEntry, bool markCont,
               std::vector<unsigned>& hdr, std::vector<unsigned>& body) {
  int emitted = 0;
  std::vector<size_t> done(chains.size(), 0);  // entries of each chain already emitted
  while (true) {
    std::vector<size_t> live;
    for (size_t c = 0; c < chains.size(); ++c)
      if (done[c] < chains[c].a.size()) live.push_back(c);
    if (live.empty()) break;
    std::stable_sort(live.begin(), live.end(), [&](size_t x, size_t y) {
      return chains[x].a.size() - done[x] > chains[y].a.size() - done[y];
    });
    if (live.size() > size_t(kLanes)) live.resize(kLanes);
    const int len = int(std::min<size_t>(SparseSchedule::kMaxChain, chains[live.front()].a.size() - done[live.front()]));
    hdr.push_back(unsigned(len) | (unsigned(kind) << 16));
    for (int l = 0; l < kLanes; ++l) {
      if (size_t(l) >= live.size()) {
        body.push_back(padTarget);
        continue;
      }
      const size_t c = live[size_t(l)];
      body.push_back(chains[c].target | ((markCont && done[c] > 0) ? SparseSchedule::kCont : 0u));
    }
    for (int j = 0; j < len; ++j)
      for (int l = 0; l < kLanes; ++l) {
        unsigned w = padEntry;
        if (size_t(l) < live.size()) {
          const size_t c = live[size_t(l)];
          if (done[c] + size_t(j) < chains[c].a.size()) w = chains[c].a[done[c] + size_t(j)];
        }
        body.push_back(w);
      }
    for (size_t l = 0; l < live.size(); ++l) {
      const size_t c = live[l];
      done[c] = std::min(chains[c].a.size(), done[c] + size_t(len));
    }
    ++emitted;
  }
  return emitted;
}

}  // namespace

SparseSchedule buildSparseSchedule(const MechTables& t, bool transposed) {
  SparseSchedule s;
  s.transposed = transposed;
  const int m = t.m;
  if (t.maxReactants > 3) throw std::runtime_error("sparse schedule: more than three reactants in a reaction");
  if (t.nSpec >= 4095 || t.nReac >= 16384) throw std::runtime_error("sparse schedule: too many species or reactions");
  for (int p : t.reacPow)
    if (p >= 16) throw std::runtime_error("sparse schedule: reactant power above 15");
  s.m = m;

  // ---- symbolic LU of I - J, no interchanges
  std::vector<char> F(size_t(m) * m, 0);
  for (int i = 0; i < m; ++i) F[size_t(i) * m + i] = 1;
  for (size_t n = 0; n < t.jnzRow.size(); ++n) {
    const int i = transposed ? t.jnzCol[n] : t.jnzRow[n], j = transposed ? t.jnzRow[n] : t.jnzCol[n];
    F[size_t(i) * m + j] = 1;
  }
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
  s.trashSlot = s.zeroSlot + 1;
  if (s.trashSlot >= 0x8000) throw std::runtime_error("sparse schedule: factor pattern too large for 15-bit positions");
  s.diagIdx.resize(m);
  for (int k = 0; k < m; ++k) s.diagIdx[k] = pos[size_t(k) * m + k];
  auto Fz = [&](int i, int j) { return F[size_t(i) * m + j] != 0; };
  auto P = [&](int i, int j) { return unsigned(pos[size_t(i) * m + j]); };
  const unsigned rhs0 = unsigned(s.nnzF);
  const unsigned Z = unsigned(s.zeroSlot), T = unsigned(s.trashSlot);
  const unsigned slack = unsigned(kLanes * SparseSchedule::kSlack);

  // ======== evaluation program ========
  const int R = t.maxReactants;
  auto field = [&](int r, int j) -> unsigned {
    if (j >= R || t.reacSpec[size_t(r) * R + j] < 0) return SparseSchedule::kNone;
    return unsigned(t.reacSpec[size_t(r) * R + j]) | (unsigned(t.reacPow[size_t(r) * R + j]) << 12);
  };
  // coefficient table: distinct stoichiometric coefficients, [0] = 0
  std::map<double, unsigned> coefIndex;
  s.coefTable.push_back(0.0);
  coefIndex[0.0] = 0;
  auto coef = [&](double c) {
    auto it = coefIndex.find(c);
    if (it != coefIndex.end()) return it->second;
    const unsigned i = unsigned(s.coefTable.size());
    if (i > 0xffff) throw std::runtime_error("sparse schedule: too many distinct coefficients");
    s.coefTable.push_back(c);
    coefIndex[c] = i;
    return i;
  };

  // rates of progress: reaction r on lane r % 64 of round r / 64
  s.ropRounds = (t.nReac + kLanes - 1) / kLanes;
  for (int g = 0; g < s.ropRounds; ++g) {
    s.evalHdr.push_back(unsigned(SparseSchedule::kRop) << 16);
    std::vector<unsigned> w0(kLanes), w1(kLanes);
    for (int l = 0; l < kLanes; ++l) {
      const int r = g * kLanes + l;
      if (r < t.nReac) {
        w0[l] = unsigned(r) | (field(r, 0) << 16);
        w1[l] = field(r, 1) | (field(r, 2) << 16);
      } else {
        w0[l] = unsigned(t.nReac) | (SparseSchedule::kNone << 16);
        w1[l] = SparseSchedule::kNone | (SparseSchedule::kNone << 16);
      }
    }
    s.evalBody.insert(s.evalBody.end(), w0.begin(), w0.end());
    s.evalBody.insert(s.evalBody.end(), w1.begin(), w1.end());
  }
  // derivative pairs: pair p on lane p % 64 of round p / 64
  s.pairBody = int(s.evalBody.size());
  {
    std::vector<int> pairReac(t.nPairs());
    for (int r = 0; r < t.nReac; ++r)
      for (int p = t.pairPtr[r]; p < t.pairPtr[r + 1]; ++p) pairReac[p] = r;
    s.pairRounds = (t.nPairs() + kLanes - 1) / kLanes;
    s.pairSlots = std::max(t.nPairs() + 1, s.pairRounds * kLanes);
    for (int g = 0; g < s.pairRounds; ++g) {
      s.evalHdr.push_back(unsigned(SparseSchedule::kPair) << 16);
      std::vector<unsigned> w0(kLanes), w1(kLanes);
      for (int l = 0; l < kLanes; ++l) {
        const int p = g * kLanes + l;
        if (p < t.nPairs()) {
          const int r = pairReac[p];
          w0[l] = unsigned(r) | (unsigned(t.pairSlot[p]) << 14) | (field(r, 0) << 16);
          w1[l] = field(r, 1) | (field(r, 2) << 16);
        } else {  // d = kf[nReac] * 1 * c_0^0 = 0
          w0[l] = unsigned(t.nReac) | ((0u | (1u << 12)) << 16);
          w1[l] = SparseSchedule::kNone | (SparseSchedule::kNone << 16);
        }
      }
      s.evalBody.insert(s.evalBody.end(), w0.begin(), w0.end());
      s.evalBody.insert(s.evalBody.end(), w1.begin(), w1.end());
    }
  }
  // segment sums
  s.segBody = int(s.evalBody.size());
  {
    std::vector<Chain> ch;
    for (int g = 0; g < t.nSeg(); ++g) {
      Chain c;
      c.target = unsigned(g);
      for (int e = t.segBegin[g]; e < t.segEnd[g]; ++e) c.a.push_back(pack16(unsigned(t.omegaReac[e]), coef(t.omegaCoef[e])));
      ch.push_back(c);
    }
    s.segRounds = emitChains(ch, SparseSchedule::kSeg, unsigned(t.nSeg()), pack16(unsigned(t.nReac), 0), true,
                             s.evalHdr, s.evalBody);
  }
  // Jacobian entries: every position of F is a target (fill positions get 0)
  s.jacBody = int(s.evalBody.size());
  {
    std::vector<int> jn(size_t(m) * m, -1);  // structural nonzero of the assembled matrix at (i, j)
    for (int n = 0; n < t.nnzJ(); ++n) {
      const int i = transposed ? t.jnzCol[n] : t.jnzRow[n], j = transposed ? t.jnzRow[n] : t.jnzCol[n];
      jn[size_t(i) * m + j] = n;
    }
    std::vector<Chain> ch;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) {
        if (!Fz(i, j)) continue;
        Chain c;
        c.target = P(i, j) | ((i == j ? 1u : 0u) << 30);
        const int n = jn[size_t(i) * m + j];
        if (n >= 0)
          for (int e = t.jnzPtr[n]; e < t.jnzPtr[n + 1]; ++e)
            c.a.push_back(pack16(unsigned(t.jcPair[e]), coef(t.jcCoef[e])));
        if (c.a.empty()) c.a.push_back(pack16(unsigned(t.nPairs()), 0));  // a round still writes the entry
        ch.push_back(c);
      }
    s.jacRounds = emitChains(ch, SparseSchedule::kJac, T, pack16(unsigned(t.nPairs()), 0), true, s.evalHdr,
                             s.evalBody);
  }
  if (transposed) {
    // nu_r . mu: reaction-major stoichiometry (net coefficients of active species)
    s.nutBody = int(s.evalBody.size());
    std::vector<Chain> nu(t.nReac);
    for (int r = 0; r < t.nReac; ++r) nu[r].target = unsigned(r);
    for (int a = 0; a < m; ++a)
      for (int e = t.omegaPtr[a]; e < t.omegaPtr[a + 1]; ++e)
        nu[t.omegaReac[e]].a.push_back(pack16(unsigned(a), coef(t.omegaCoef[e])));
    for (auto& c : nu)  // every w[r] is written (0 when reaction r touches no active species)
      if (c.a.empty()) c.a.push_back(pack16(unsigned(m), 0));
    s.nutRounds = emitChains(nu, SparseSchedule::kNuT, unsigned(t.nReac + 1), pack16(unsigned(m), 0), true,
                             s.evalHdr, s.evalBody);
    // J^T mu = sum over derivative pairs of species j of dpr[p] * w[reaction(p)]
    s.jtvBody = int(s.evalBody.size());
    std::vector<Chain> jt(m);
    for (int j = 0; j < m; ++j) jt[j].target = unsigned(j);
    for (int r = 0; r < t.nReac; ++r)
      for (int p = t.pairPtr[r]; p < t.pairPtr[r + 1]; ++p)
        jt[t.reacSpec[size_t(r) * R + t.pairSlot[p]]].a.push_back(pack16(unsigned(p), unsigned(r)));
    for (auto& c : jt)  // every v[j] is written
      if (c.a.empty()) c.a.push_back(pack16(unsigned(t.nPairs()), unsigned(t.nReac)));
    s.jtvRounds = emitChains(jt, SparseSchedule::kJtv, unsigned(m), pack16(unsigned(t.nPairs()), unsigned(t.nReac)),
                             true, s.evalHdr, s.evalBody);
  }
  s.evalBody.insert(s.evalBody.end(), slack, 0u);

  // ======== LU program ========
  std::vector<int> lev(m, 0);
  for (int k = 0; k < m; ++k) {
    int L = 0;
    for (int kp = 0; kp < k; ++kp)
      if (Fz(k, kp) || Fz(kp, k)) L = std::max(L, lev[kp]);
    lev[k] = L + 1;
  }
  s.nLevels = m ? *std::max_element(lev.begin(), lev.end()) : 0;
  const unsigned zz = pack16(Z, Z);
  // Rounds by readiness (list scheduling).  Pivot k factors (multipliers
  // l_ik = a_ik / a_kk, reciprocal pivot) once its diagonal and column are
  // final; the contribution of pivot p to (i, j) -- a_ij -= l_ip u_pj, the
  // RHS riding along as column m -- is ready once p has factored and u_pj is
  // final.  Each round takes what is ready, most urgent targets first (the
  // level of the pivot that reads them first), so the trailing updates of a
  // dense block wait and then go in chains of up to kMaxChain contributions
  // per lane instead of one contribution per lane and round.  Pivots that
  // are ready factor in a kMulUpd round when their multiplier and pivot
  // records fit; its other lanes each carry one contribution of those
  // pivots (scaling a_ip themselves: they read the unscaled value) and up to
  // kMaxMulUpdChain contributions of pivots factored earlier.  Otherwise
  // they factor in kMul rounds.  Later rounds read the stored multipliers.
  // Readiness is taken from the state before each round, so no round reads
  // an entry it writes.
  {
    const unsigned trashDiag = unsigned(m + 1) << 16;
    const int nT = s.nnzF + m;  // targets: factor positions, then the RHS entry of each row
    struct Contribution {
      int pivot;
      unsigned ik, kj;
      bool done;
    };
    std::vector<std::vector<Contribution>> contrib(static_cast<size_t>(nT));
    std::vector<int> rem(size_t(nT), 0);
    for (int k = 0; k < m; ++k)
      for (int i = k + 1; i < m; ++i) {
        if (!Fz(i, k)) continue;
        for (int j = k + 1; j <= m; ++j) {
          if (j < m && !Fz(k, j)) continue;
          const unsigned tgt = j == m ? rhs0 + unsigned(i) : P(i, j);
          contrib[tgt].push_back({k, P(i, k), j == m ? rhs0 + unsigned(k) : P(k, j), false});
          ++rem[tgt];
          if (j < m) ++s.nUpdates;
        }
      }
    std::vector<int> urg(size_t(nT), 0);
    for (int i = 0; i < m; ++i) {
      for (int q = s.rowPtr[i]; q < s.rowPtr[i + 1]; ++q) urg[size_t(q)] = lev[std::min(i, s.colIdx[size_t(q)])];
      urg[rhs0 + unsigned(i)] = lev[i];
    }
    std::vector<char> factored(m, 0);
    int nFactored = 0;
    long remaining = 0;
    for (int v : rem) remaining += v;
    auto isFinal = [&](unsigned e) { return e == Z || rem[e] == 0; };
    // most urgent first: the level that reads the target first, brought
    // forward by the contributions still to apply (a long backlog must
    // start early: a lane applies at most kMaxChain per round)
    auto byUrgency = [&](std::vector<unsigned>& tg) {
      std::stable_sort(tg.begin(), tg.end(), [&](unsigned x, unsigned y) {
        const int kx = 3 * urg[x] - rem[x], ky = 3 * urg[y] - rem[y];
        if (kx != ky) return kx < ky;
        return rem[x] > rem[y];
      });
    };
    while (nFactored < m || remaining > 0) {
      std::vector<int> R;
      for (int k = 0; k < m; ++k) {
        if (factored[k] || !isFinal(P(k, k))) continue;
        bool ok = true;
        for (int i = k + 1; i < m && ok; ++i)
          if (Fz(i, k) && !isFinal(P(i, k))) ok = false;
        if (ok) R.push_back(k);
      }
      if (!R.empty()) {
        constexpr int kC = SparseSchedule::kMaxMulUpdChain;
        // (target, ik | kj << 16, kk | mul << 15 | rdiag index << 16, chain entries)
        std::vector<std::array<unsigned, 3 + kC>> lanes;
        auto lane = [&](unsigned t, unsigned w1, unsigned w2) {
          std::array<unsigned, 3 + kC> x;
          x.fill(zz);
          x[0] = t;
          x[1] = w1;
          x[2] = w2;
          lanes.push_back(x);
        };
        for (int k : R) {
          const unsigned kk = P(k, k);
          bool first = true;
          for (int i = k + 1; i < m; ++i)
            if (Fz(i, k)) {
              lane(P(i, k), pack16(P(i, k), Z), kk | 0x8000u | (first ? unsigned(k) << 16 : trashDiag));
              first = false;
              ++s.nMultipliers;
            }
          if (first) lane(T, zz, kk | (unsigned(k) << 16));  // pivot record only
        }
        if (lanes.size() <= size_t(kLanes)) {
          std::vector<char> inR(m, 0);
          for (int k : R) inR[k] = 1;
          // from the state before the round: one contribution of a pivot of
          // R (scaled in the lane) and up to kC of factored pivots per target
          std::vector<int> head(static_cast<size_t>(nT), -1);
          std::vector<std::vector<size_t>> chainOf(static_cast<size_t>(nT));
          std::vector<unsigned> tg;
          for (int e = 0; e < nT; ++e) {
            if (rem[e] == 0) continue;
            for (size_t c = 0; c < contrib[e].size(); ++c) {
              const Contribution& x = contrib[e][c];
              if (x.done || !isFinal(x.kj)) continue;
              if (inR[x.pivot]) {
                if (head[e] < 0) head[e] = int(c);
              } else if (factored[x.pivot] && chainOf[e].size() < size_t(kC)) {
                chainOf[e].push_back(c);
              }
            }
            if (head[e] >= 0 || !chainOf[e].empty()) tg.push_back(unsigned(e));
          }
          byUrgency(tg);
          int chain = 0;
          for (unsigned e : tg) {
            if (lanes.size() >= size_t(kLanes)) break;
            unsigned w1 = zz, w2 = Z | trashDiag;
            if (head[e] >= 0) {
              Contribution& x = contrib[e][size_t(head[e])];
              w1 = pack16(x.ik, x.kj);
              w2 = P(x.pivot, x.pivot) | trashDiag;
              x.done = true;
              --rem[e];
              --remaining;
            }
            lane(e, w1, w2);
            for (size_t j = 0; j < chainOf[e].size(); ++j) {
              Contribution& x = contrib[e][chainOf[e][j]];
              lanes.back()[3 + j] = pack16(x.ik, x.kj);
              x.done = true;
              --rem[e];
              --remaining;
            }
            chain = std::max(chain, int(chainOf[e].size()));
          }
          s.progHdr.push_back(unsigned(2 + chain) | (unsigned(SparseSchedule::kMulUpd) << 16));
          for (int slot = 0; slot < 3 + chain; ++slot)
            for (int l = 0; l < kLanes; ++l) {
              if (size_t(l) < lanes.size()) s.progBody.push_back(lanes[l][slot]);
              else s.progBody.push_back(slot == 0 ? T : (slot == 2 ? (Z | trashDiag) : zz));
            }
          ++s.luRounds;
        } else {
          std::vector<unsigned> recs;
          for (int k : R) {
            recs.push_back((0x8000u | unsigned(k)) | (P(k, k) << 16));
            for (int i = k + 1; i < m; ++i)
              if (Fz(i, k)) recs.push_back(P(i, k) | (P(k, k) << 16));
          }
          for (size_t base = 0; base < recs.size(); base += kLanes) {
            s.progHdr.push_back(unsigned(SparseSchedule::kMul) << 16);
            for (int l = 0; l < kLanes; ++l) s.progBody.push_back(base + l < recs.size() ? recs[base + l] : pack16(T, Z));
            ++s.luRounds;
          }
        }
        for (int k : R) factored[k] = 1;
        nFactored += int(R.size());
        continue;
      }
      // no pivot ready: chained updates with stored multipliers
      // readiness from the state before the round: an entry finished in
      // this round is not read in it
      std::vector<unsigned> tg;
      std::vector<std::vector<size_t>> readyOf(static_cast<size_t>(nT));
      for (int e = 0; e < nT; ++e) {
        if (rem[e] == 0) continue;
        for (size_t c = 0; c < contrib[e].size(); ++c) {
          const Contribution& x = contrib[e][c];
          if (!x.done && factored[x.pivot] && isFinal(x.kj)) readyOf[e].push_back(c);
        }
        if (!readyOf[e].empty()) tg.push_back(unsigned(e));
      }
      if (tg.empty()) throw std::runtime_error("sparse schedule: LU program stalled");
      byUrgency(tg);
      if (tg.size() > size_t(kLanes)) tg.resize(kLanes);
      int len = 0;
      for (unsigned e : tg) len = std::max(len, std::min(int(readyOf[e].size()), SparseSchedule::kMaxChain));
      std::vector<std::vector<unsigned>> words(size_t(len), std::vector<unsigned>(kLanes, zz));
      for (size_t l = 0; l < tg.size(); ++l) {
        const unsigned e = tg[l];
        int taken = 0;
        for (size_t c : readyOf[e]) {
          if (taken == len) break;
          Contribution& x = contrib[e][c];
          words[size_t(taken)][l] = pack16(x.ik, x.kj);
          x.done = true;
          ++taken;
        }
        rem[e] -= taken;
        remaining -= taken;
      }
      s.progHdr.push_back(unsigned(len) | (unsigned(SparseSchedule::kUpd) << 16));
      for (int l = 0; l < kLanes; ++l) s.progBody.push_back(size_t(l) < tg.size() ? tg[l] : T);
      for (int j = 0; j < len; ++j)
        for (int l = 0; l < kLanes; ++l) s.progBody.push_back(words[size_t(j)][size_t(l)]);
      ++s.luRounds;
    }
  }

  // back substitution, column oriented: when column k is final
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
  for (int L = 1; L <= s.nBackLevels; ++L) {
    std::map<unsigned, Chain> byTarget;
    for (int k = 0; k < m; ++k) {
      if (bl[k] != L) continue;
      for (int i = 0; i < k; ++i)
        if (Fz(i, k)) {
          Chain& c = byTarget[rhs0 + unsigned(i)];
          c.target = rhs0 + unsigned(i);
          c.a.push_back(pack16(P(i, k), unsigned(k)));
        }
    }
    std::vector<Chain> ch;
    for (auto& e : byTarget) ch.push_back(e.second);
    s.backRounds += emitChains(ch, SparseSchedule::kBack, T, pack16(Z, unsigned(m)), false, s.progHdr, s.progBody);
  }

  // forward substitution with stored factors (unit lower), column oriented by LU level
  s.fwdBody = int(s.progBody.size());
  for (int L = 1; L <= s.nLevels; ++L) {
    std::map<unsigned, Chain> byTarget;
    for (int k = 0; k < m; ++k) {
      if (lev[k] != L) continue;
      for (int i = k + 1; i < m; ++i)
        if (Fz(i, k)) {
          Chain& c = byTarget[rhs0 + unsigned(i)];
          c.target = rhs0 + unsigned(i);
          c.a.push_back(pack16(P(i, k), rhs0 + unsigned(k)));
        }
    }
    std::vector<Chain> ch;
    for (auto& e : byTarget) ch.push_back(e.second);
    s.fwdRounds += emitChains(ch, SparseSchedule::kUpd, T, zz, false, s.progHdr, s.progBody);
  }
  s.progBody.insert(s.progBody.end(), slack, 0u);
  return s;
}

}  // namespace tchem_amd
