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
#include <random>
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

// Bank-aware numbering of the factor positions.  Every matrix access of a
// round (one operand field read or written by the 64 lanes) is an LDS
// instruction; on 32 banks of 4 bytes per half wavefront it takes as many
// passes as the most-used bank holds distinct addresses.  The positions
// [0, nnzF) are permuted (RHS, zero and trash slots fixed) to minimise the
// passes summed over the LU, BACK, FWD and Jacobian rounds: local search
// over swaps from the row-major numbering, deterministic.  The program
// words are rewritten in place.
void bankAwareSlots(SparseSchedule& s) {
  using SS = SparseSchedule;
  const int P = s.nnzF;
  constexpr int kBanks = 32, kHalf = 32;
  s.slotOf.resize(size_t(P));
  for (int p = 0; p < P; ++p) s.slotOf[size_t(p)] = p;
  // an access: for one (round, field, half wavefront), the distinct movable
  // positions and the bank counts of the fixed addresses; weight 2 for a
  // target that is read and written
  struct Access {
    std::vector<int> pos;
    std::array<int, kBanks> fixedCnt{};
    int weight = 1;
  };
  std::vector<Access> acc;
  // field extractors over a section: calls f(words of lane l) per round
  auto addField = [&](const std::vector<std::vector<unsigned>>& vals, int weight) {
    // vals[l] = addresses of lane l for this field (0 or 1 address)
    for (int h = 0; h < 2; ++h) {
      Access a;
      a.weight = weight;
      std::vector<int> seen;
      for (int l = h * kHalf; l < (h + 1) * kHalf; ++l)
        for (unsigned v : vals[size_t(l)]) seen.push_back(int(v));
      std::sort(seen.begin(), seen.end());
      seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
      for (int v : seen) {
        if (v < P) a.pos.push_back(v);
        else a.fixedCnt[size_t(v % kBanks)]++;
      }
      if (a.pos.size() + size_t(std::count_if(a.fixedCnt.begin(), a.fixedCnt.end(), [](int c) { return c > 0; })) > 1)
        acc.push_back(std::move(a));
    }
  };
  // walk a section of the LU program: callback per round with (kind, len, word pointer)
  auto walkLu = [&](int r0, int nr, size_t body0, auto&& fn) {
    size_t b = body0;
    for (int r = r0; r < r0 + nr; ++r) {
      const int kind = int(s.progHdr[size_t(r)] >> 16), len = int(s.progHdr[size_t(r)] & 0xffff);
      fn(kind, len, b);
      b += size_t(kLanes) * (kind == SS::kMul ? 1 : size_t(1 + len));
    }
  };
  const unsigned rhs0 = unsigned(P);
  auto collect = [&](int kind, int len, size_t b) {
    std::vector<std::vector<unsigned>> f(kLanes);
    auto word = [&](int j, int l) { return s.progBody[b + size_t(kLanes) * size_t(j) + size_t(l)]; };
    auto flush = [&](int weight) {
      addField(f, weight);
      for (auto& x : f) x.clear();
    };
    if (kind == SS::kMul) {
      for (int l = 0; l < kLanes; ++l) f[size_t(l)].push_back(word(0, l) >> 16);
      flush(1);
      for (int l = 0; l < kLanes; ++l) {
        const unsigned dst = word(0, l) & 0xffff;
        f[size_t(l)].push_back((dst & 0x8000) ? unsigned(s.trashSlot) : dst);
      }
      flush(2);
    } else if (kind == SS::kUpd || kind == SS::kBack) {
      for (int l = 0; l < kLanes; ++l) f[size_t(l)].push_back(word(0, l));
      flush(2);
      for (int j = 1; j <= len; ++j) {
        for (int l = 0; l < kLanes; ++l) f[size_t(l)].push_back(word(j, l) & 0xffff);
        flush(1);
        if (kind == SS::kUpd) {
          for (int l = 0; l < kLanes; ++l) f[size_t(l)].push_back(word(j, l) >> 16);
          flush(1);
        } else {
          for (int l = 0; l < kLanes; ++l) f[size_t(l)].push_back(rhs0 + (word(j, l) >> 16));
          flush(1);
        }
      }
    } else if (kind == SS::kMulUpd) {
      for (int l = 0; l < kLanes; ++l) f[size_t(l)].push_back(word(0, l));
      flush(2);
      for (int l = 0; l < kLanes; ++l) f[size_t(l)].push_back(word(1, l) & 0xffff);
      flush(1);
      for (int l = 0; l < kLanes; ++l) f[size_t(l)].push_back(word(1, l) >> 16);
      flush(1);
      for (int l = 0; l < kLanes; ++l) f[size_t(l)].push_back(word(2, l) & 0x7fff);
      flush(1);
      for (int j = 3; j <= len; ++j) {
        for (int l = 0; l < kLanes; ++l) f[size_t(l)].push_back(word(j, l) & 0xffff);
        flush(1);
        for (int l = 0; l < kLanes; ++l) f[size_t(l)].push_back(word(j, l) >> 16);
        flush(1);
      }
    }
  };
  walkLu(0, s.luRounds, 0, collect);
  walkLu(s.luRounds, s.backRounds, size_t(s.backBody), collect);
  walkLu(s.luRounds + s.backRounds, s.fwdRounds, size_t(s.fwdBody), collect);
  // Jacobian rounds: targets (position in the low 30 bits)
  {
    size_t b = size_t(s.jacBody);
    const int r0 = s.ropRounds + s.pairRounds + s.segRounds;
    for (int r = r0; r < r0 + s.jacRounds; ++r) {
      const int len = int(s.evalHdr[size_t(r)] & 0xffff);
      std::vector<std::vector<unsigned>> f(kLanes);
      for (int l = 0; l < kLanes; ++l) f[size_t(l)].push_back(s.evalBody[b + size_t(l)] & 0x3fffffffu);
      addField(f, 2);
      b += size_t(kLanes) * size_t(1 + len);
    }
  }
  // per position: the accesses it is in
  std::vector<std::vector<int>> inAcc(static_cast<size_t>(P));
  for (size_t g = 0; g < acc.size(); ++g)
    for (int p : acc[g].pos) inAcc[size_t(p)].push_back(int(g));
  std::vector<std::array<int, kBanks>> cnt(acc.size());
  auto passes = [&](const std::array<int, kBanks>& c) { return *std::max_element(c.begin(), c.end()); };
  long long total = 0;
  for (size_t g = 0; g < acc.size(); ++g) {
    cnt[g] = acc[g].fixedCnt;
    for (int p : acc[g].pos) cnt[g][size_t(s.slotOf[size_t(p)] % kBanks)]++;
    total += (long long)acc[g].weight * (passes(cnt[g]) - 1);
  }
  s.bankPassesBefore = total;
  // local search: swap the slots of two positions in different banks when
  // that does not add passes
  std::mt19937_64 rng(0x5eedULL);
  // budget: 400 tries per position, and at most ~3e8 access updates (a
  // deterministic work bound that keeps kmcd creation fast on the largest
  // mechanisms, whose sparse form is not the chosen one)
  const long long iters = std::min<long long>(400LL * P, 1500000LL);
  long long work = 0;
  const long long maxWork = 300000000LL;
  std::vector<int> mark(acc.size(), 0);
  int stamp = 0;
  for (long long it = 0; it < iters && P > 1 && work < maxWork; ++it) {
    const int a = int(rng() % unsigned(P)), b = int(rng() % unsigned(P));
    work += (long long)(inAcc[size_t(a)].size() + inAcc[size_t(b)].size()) * 2;
    const int sa = s.slotOf[size_t(a)], sb = s.slotOf[size_t(b)];
    const int ra = sa % kBanks, rb = sb % kBanks;
    if (a == b || ra == rb) continue;
    ++stamp;
    long long delta = 0;
    // accesses holding both keep their counts
    for (int g : inAcc[size_t(b)]) mark[size_t(g)] = stamp;
    auto moveCost = [&](int g, int from, int to) {
      std::array<int, kBanks>& c = cnt[size_t(g)];
      const int before = passes(c);
      c[size_t(from)]--;
      c[size_t(to)]++;
      const int after = passes(c);
      c[size_t(to)]--;
      c[size_t(from)]++;
      return (long long)acc[size_t(g)].weight * (after - before);
    };
    std::vector<int> onlyA, onlyB;
    for (int g : inAcc[size_t(a)])
      if (mark[size_t(g)] != stamp) onlyA.push_back(g);
      else mark[size_t(g)] = -stamp;  // both
    for (int g : inAcc[size_t(b)])
      if (mark[size_t(g)] == stamp) onlyB.push_back(g);
    for (int g : onlyA) delta += moveCost(g, ra, rb);
    for (int g : onlyB) delta += moveCost(g, rb, ra);
    if (delta <= 0) {
      for (int g : onlyA) {
        cnt[size_t(g)][size_t(ra)]--;
        cnt[size_t(g)][size_t(rb)]++;
      }
      for (int g : onlyB) {
        cnt[size_t(g)][size_t(rb)]--;
        cnt[size_t(g)][size_t(ra)]++;
      }
      s.slotOf[size_t(a)] = sb;
      s.slotOf[size_t(b)] = sa;
      total += delta;
    }
  }
  s.bankPassesAfter = total;
  // rewrite the program words: every position field below nnzF
  auto mapPos = [&](unsigned v) { return v < unsigned(P) ? unsigned(s.slotOf[v]) : v; };
  auto remapLu = [&](int kind, int len, size_t b) {
    auto wref = [&](int j, int l) -> unsigned& { return s.progBody[b + size_t(kLanes) * size_t(j) + size_t(l)]; };
    for (int l = 0; l < kLanes; ++l) {
      if (kind == SS::kMul) {
        unsigned& w = wref(0, l);
        const unsigned dst = w & 0xffff, src = w >> 16;
        w = ((dst & 0x8000) ? dst : mapPos(dst)) | (mapPos(src) << 16);
      } else if (kind == SS::kUpd) {
        wref(0, l) = mapPos(wref(0, l));
        for (int j = 1; j <= len; ++j) {
          unsigned& w = wref(j, l);
          w = mapPos(w & 0xffff) | (mapPos(w >> 16) << 16);
        }
      } else if (kind == SS::kBack) {
        wref(0, l) = mapPos(wref(0, l));
        for (int j = 1; j <= len; ++j) {
          unsigned& w = wref(j, l);
          w = mapPos(w & 0xffff) | (w & 0xffff0000u);
        }
      } else if (kind == SS::kMulUpd) {
        wref(0, l) = mapPos(wref(0, l));
        unsigned& w1 = wref(1, l);
        w1 = mapPos(w1 & 0xffff) | (mapPos(w1 >> 16) << 16);
        unsigned& w2 = wref(2, l);
        w2 = mapPos(w2 & 0x7fff) | (w2 & 0xffff8000u);
        for (int j = 3; j <= len; ++j) {
          unsigned& w = wref(j, l);
          w = mapPos(w & 0xffff) | (mapPos(w >> 16) << 16);
        }
      }
    }
  };
  walkLu(0, s.luRounds, 0, remapLu);
  walkLu(s.luRounds, s.backRounds, size_t(s.backBody), remapLu);
  walkLu(s.luRounds + s.backRounds, s.fwdRounds, size_t(s.fwdBody), remapLu);
  {
    size_t b = size_t(s.jacBody);
    const int r0 = s.ropRounds + s.pairRounds + s.segRounds;
    for (int r = r0; r < r0 + s.jacRounds; ++r) {
      const int len = int(s.evalHdr[size_t(r)] & 0xffff);
      for (int l = 0; l < kLanes; ++l) {
        unsigned& w = s.evalBody[b + size_t(l)];
        w = mapPos(w & 0x3fffffffu) | (w & 0xc0000000u);
      }
      b += size_t(kLanes) * size_t(1 + len);
    }
  }
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
  bankAwareSlots(s);
  return s;
}

}  // namespace tchem_amd
