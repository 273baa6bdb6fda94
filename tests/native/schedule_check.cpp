// Host-side checker of the compiled sparse programs (tchem-atm_amd/csrc/schedule.cpp).
//
// Emulates the rounds the way the device interpreter runs them (every lane of
// a round reads before any lane writes, rounds in order) and checks: every
// index stays inside its buffer, no lane of a round reads an entry another
// lane of the same round writes, padding never touches real entries, and the
// results agree with direct formulas -- rates of progress bit for bit, the
// source term and I - c h J to 1e-14 relative, and the LU program's solves
// (fresh factorisation, and forward + back substitution with stored factors)
// with a long-double dense partial-pivoting solve to 1e-6 relative.
//
// usage: schedule_check <mechanism.yaml> [<aerosol.yaml>]   (exit status 0
// when every check passes).  With an aerosol file, the LU program of the
// aerosol problem (gas + particle unknowns, aerosolTables) is checked too.
#include <cmath>
#include <cstdio>
#include <map>
#include <random>
#include <vector>
#include "aerosol_model.hpp"
#include "kinetic_model.hpp"
#include "mechanism.hpp"
#include "schedule.hpp"
using namespace tchem_amd;

static int fails = 0;
#define CHECK(c, ...)                                                  \
  do {                                                                 \
    if (!(c)) {                                                        \
      ++fails;                                                         \
      if (fails < 30) {                                                \
        std::printf("FAIL %s:%d ", __FILE__, __LINE__);                \
        std::printf(__VA_ARGS__);                                      \
        std::printf("\n");                                             \
      }                                                                \
    }                                                                  \
  } while (0)

static double ipow(double x, int n) {
  double r = n >= 1 ? x : 1.0;
  for (int i = 1; i < n; ++i) r *= x;
  return r;
}

struct Emu {
  const SparseSchedule& s; int m; std::vector<double> A, rd; bool bad = false;
  Emu(const SparseSchedule& s_, int m_) : s(s_), m(m_) {}
  // run rounds [r0, r1) starting at body word offset off; returns end offset
  size_t run(int r0, int r1, size_t off) {
    const int Z = s.zeroSlot, T = s.trashSlot, rhs0 = s.nnzF;
    for (int r = r0; r < r1; ++r) {
      const unsigned h = s.progHdr[r]; const int len = h & 0xffff, kind = h >> 16;
      std::map<long, int> wr; std::vector<std::pair<long, int>> rdset;
      std::vector<std::pair<long, double>> outA, outR;
      if (kind == SparseSchedule::kMul) {
        CHECK(len == 0, "mul len");
        for (int l = 0; l < 64; ++l) {
          unsigned w = s.progBody[off + l]; int dst = w & 0xffff, src = w >> 16;
          CHECK(src <= Z, "mul src %d", src);
          double piv = A[src], inv = 1.0 / piv;
          rdset.push_back({src, l});
          if (dst & 0x8000) { int k = dst & 0x7fff; CHECK(k < m, "rdiag k"); outR.push_back({k, inv}); if (!(piv != 0 && std::isfinite(piv) && std::isfinite(inv))) bad = true; wr[-1 - k] = l; }
          else { CHECK(dst < Z || dst == T, "mul dst %d", dst); if (dst != T) { CHECK(!wr.count(dst), "dup mul dst"); wr[dst] = l; rdset.push_back({dst, l}); } outA.push_back({dst, A[dst] * inv}); }
        }
        off += 64;
      } else if (kind == SparseSchedule::kMulUpd) {
        CHECK(len >= 2 && len <= 2 + SparseSchedule::kMaxMulUpdChain, "mulupd len");
        std::map<long, int> mulw;  // multiplier slots written by mul lanes: update lanes may read them (old value)
        for (int l = 0; l < 64; ++l) {
          const unsigned w2 = s.progBody[off + 128 + l];
          if (w2 & 0x8000u) mulw[long(s.progBody[off + l])] = l;
        }
        for (int l = 0; l < 64; ++l) {
          const int t = int(s.progBody[off + l]);
          const unsigned w1 = s.progBody[off + 64 + l], w2 = s.progBody[off + 128 + l];
          const int ik = w1 & 0xffff, kj = w1 >> 16, kk = w2 & 0x7fff, ri = w2 >> 16;
          const bool mul = (w2 & 0x8000u) != 0;
          CHECK(t < Z || t == T, "mulupd target %d", t);
          CHECK(ik <= Z && kj <= Z && kk <= Z, "mulupd index");
          CHECK(ri < m || ri == m + 1, "mulupd rdiag %d", ri);
          if (mul) CHECK(t == ik && kj == Z, "mul lane layout");
          const double p = A[kk], inv = 1.0 / p, lv = A[ik] == 0.0 ? 0.0 : A[ik] * inv;
          double acc = A[t];
          for (int j = 0; j < len - 2; ++j) {
            const unsigned e = s.progBody[off + 64 * (3 + j) + l];
            const int lo = e & 0xffff, hi = e >> 16;
            CHECK(lo <= Z && hi <= Z, "mulupd chain index");
            if (t != T) CHECK(lo != T && hi != T, "reads trash");
            CHECK(!mul || (lo == Z && hi == Z), "chain on a multiplier lane");
            CHECK(!mulw.count(lo) && !mulw.count(hi), "chain reads a multiplier stored in the same round");
            acc -= A[lo] * A[hi];
            rdset.push_back({lo, l});
            rdset.push_back({hi, l});
          }
          const double val = mul ? lv : acc - lv * A[kj];
          if (ri != m + 1) {
            CHECK(!wr.count(-1 - ri), "dup rdiag");
            wr[-1 - ri] = l;
            outR.push_back({ri, inv});
            if (!(p != 0 && std::isfinite(p) && std::isfinite(inv))) bad = true;
          }
          rdset.push_back({kk, l});
          if (!mul) {
            auto it = mulw.find(ik);
            CHECK(it == mulw.end() || t != ik, "update lane writes a multiplier");
            if (it == mulw.end()) rdset.push_back({ik, l});  // reading the unscaled value is intended
            rdset.push_back({kj, l});
          }
          if (t != T) {
            CHECK(!wr.count(t), "dup target %d", t);
            wr[t] = l;
            if (!mul) rdset.push_back({t, l});
          }
          outA.push_back({t, val});
        }
        off += size_t(64) * (1 + len);
      } else {
        for (int l = 0; l < 64; ++l) {
          int t = int(s.progBody[off + l]);
          CHECK(t < Z || t == T, "target %d", t);
          double acc = A[t];
          for (int j = 0; j < len; ++j) {
            unsigned e = s.progBody[off + 64 * (1 + j) + l]; int lo = e & 0xffff, hi = e >> 16;
            CHECK(lo <= Z, "lo %d", lo);
            if (t != T) CHECK(lo != T, "reads trash");
            if (kind == SparseSchedule::kUpd) { CHECK(hi <= Z, "hi"); acc -= A[lo] * A[hi]; rdset.push_back({lo, l}); rdset.push_back({hi, l}); }
            else { CHECK(hi <= m, "back col %d", hi); acc -= A[lo] * (A[rhs0 + hi] * rd[hi]); rdset.push_back({lo, l}); rdset.push_back({rhs0 + hi, l}); rdset.push_back({-1 - hi, l}); }
          }
          if (t != T) { CHECK(!wr.count(t), "dup target %d", t); wr[t] = l; rdset.push_back({t, l}); }
          outA.push_back({t, acc});
        }
        off += 64 * (1 + len);
      }
      for (auto& p : rdset) { auto it = wr.find(p.first); if (it != wr.end() && it->second != p.second) CHECK(false, "race on %ld round %d kind %d", p.first, r, kind); }
      for (auto& o : outA) A[o.first] = o.second;
      for (auto& o : outR) rd[o.first] = o.second;
      CHECK(A[Z] == 0.0, "zero slot modified");
    }
    return off;
  }
};

static std::vector<long double> dense_solve(std::vector<long double> M, std::vector<long double> b, int m) {
  for (int k = 0; k < m; ++k) { int p = k; for (int i = k + 1; i < m; ++i) if (fabsl(M[size_t(i) * m + k]) > fabsl(M[size_t(p) * m + k])) p = i;
    if (p != k) { for (int j = 0; j < m; ++j) std::swap(M[size_t(k) * m + j], M[size_t(p) * m + j]); std::swap(b[k], b[p]); }
    for (int i = k + 1; i < m; ++i) { long double l = M[size_t(i) * m + k] / M[size_t(k) * m + k]; if (l == 0) continue; for (int j = k; j < m; ++j) M[size_t(i) * m + j] -= l * M[size_t(k) * m + j]; b[i] -= l * b[k]; } }
  for (int k = m - 1; k >= 0; --k) { long double v = b[k]; for (int j = k + 1; j < m; ++j) v -= M[size_t(k) * m + j] * b[j]; b[k] = v / M[size_t(k) * m + k]; }
  return b;
}

static void check_lu_program(const MechTables& t, const SparseSchedule& s) {

  const int m = t.m;
  size_t bodyBytes = s.progBody.size() * 4, hdrBytes = s.progHdr.size() * 4;
  int slots[3] = {0, 0, 0};
  for (size_t r = 0; r < s.progHdr.size(); ++r) slots[r < size_t(s.luRounds) ? 0 : (r < size_t(s.luRounds + s.backRounds) ? 1 : 2)] += 1 + (s.progHdr[r] & 0xffff);
  std::printf("m %d nnzF %d levels %d back levels %d | rounds lu %d back %d fwd %d | sequential slots lu %d back %d fwd %d | program %.1f KB + hdr %.1f KB\n",
              m, s.nnzF, s.nLevels, s.nBackLevels, s.luRounds, s.backRounds, s.fwdRounds, slots[0], slots[1], slots[2], bodyBytes / 1024.0, hdrBytes / 1024.0);
  std::printf("bank passes (extra LDS passes per program pass) before %lld after %lld\n", s.bankPassesBefore, s.bankPassesAfter);
  {  // the renumbering is a permutation of [0, nnzF)
    std::vector<char> seen(size_t(s.nnzF), 0);
    for (int p = 0; p < s.nnzF; ++p) {
      const int q = s.slotOf[size_t(p)];
      CHECK(q >= 0 && q < s.nnzF && !seen[size_t(q)], "slotOf is not a permutation at %d", p);
      seen[size_t(q)] = 1;
    }
  }
  CHECK((int)s.progHdr.size() == s.luRounds + s.backRounds + s.fwdRounds, "round count");
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  for (int trial = 0; trial < 6; ++trial) {
    // random matrix on the pattern of F, diagonally weighted like I - ch J
    std::vector<long double> D(size_t(m) * m, 0.0L);
    Emu e(s, m);
    e.A.assign(s.trashSlot + 1, 0.0); e.rd.assign(m + 2, 0.0);
    std::vector<char> isJ(size_t(m) * m, 0);
    for (size_t n = 0; n < t.jnzRow.size(); ++n) {
      const int i = s.transposed ? t.jnzCol[n] : t.jnzRow[n], j = s.transposed ? t.jnzRow[n] : t.jnzCol[n];
      isJ[size_t(i) * m + j] = 1;
    }
    const double scale = std::pow(10.0, trial - 3);
    for (int i = 0; i < m; ++i) for (int p = s.rowPtr[i]; p < s.rowPtr[i + 1]; ++p) {
      int j = s.colIdx[p]; double v = 0;
      if (isJ[size_t(i) * m + j]) v = scale * (U(rng) - 0.5);
      if (i == j) v += 1.0 + scale * U(rng);
      e.A[size_t(s.slotOf[size_t(p)])] = v; D[size_t(i) * m + j] = v;
    }
    e.A[s.trashSlot] = std::nan("");
    std::vector<double> b(m); std::vector<long double> bl(m);
    for (int i = 0; i < m; ++i) { b[i] = U(rng) - 0.5; e.A[s.nnzF + i] = b[i]; bl[i] = b[i]; }
    size_t off = e.run(0, s.luRounds, 0);
    CHECK(off == size_t(s.backBody), "lu body end %zu vs %d", off, s.backBody);
    CHECK(!e.bad, "bad pivot");
    std::vector<double> F = e.A, R = e.rd;
    off = e.run(s.luRounds, s.luRounds + s.backRounds, off);
    CHECK(off == size_t(s.fwdBody), "back body end");
    auto x = dense_solve(D, bl, m);
    double err = 0, xm = 0;
    for (int i = 0; i < m; ++i) { double xi = e.A[s.nnzF + i] * e.rd[i]; err = std::max(err, std::fabs(xi - (double)x[i])); xm = std::max(xm, std::fabs((double)x[i])); }
    // reuse: fwd + back with new rhs
    e.A = F; e.rd = R;
    std::vector<double> b2(m); std::vector<long double> b2l(m);
    for (int i = 0; i < m; ++i) { b2[i] = U(rng) - 0.5; e.A[s.nnzF + i] = b2[i]; b2l[i] = b2[i]; }
    off = e.run(s.luRounds + s.backRounds, s.luRounds + s.backRounds + s.fwdRounds, s.fwdBody);
    CHECK(off + 64 * (1 + SparseSchedule::kMaxChain) == s.progBody.size(), "fwd body end");
    e.run(s.luRounds, s.luRounds + s.backRounds, s.backBody);
    auto x2 = dense_solve(D, b2l, m);
    double err2 = 0;
    for (int i = 0; i < m; ++i) err2 = std::max(err2, std::fabs(e.A[s.nnzF + i] * e.rd[i] - (double)x2[i]));
    std::printf("trial %d scale %.0e: solve err %.3g, reuse err %.3g (|x| %.3g)\n", trial, scale, err, err2, xm);
    // static pivots on random matrices lose some digits at large c h; a wrong
    // schedule gives O(1) errors
    CHECK(err <= 1e-6 * (xm + 1) && err2 <= 1e-6 * (xm + 1), "solve accuracy");
  }
}

// kNuT and kJtv sections of a transposed schedule: w = nu^T mu and v = J^T mu
// against direct sums, starting at round rr / body word off.
static void check_adjoint_sections(const MechTables& t, const SparseSchedule& s, const std::vector<double>& c,
                                   const std::vector<double>& kf, const std::vector<double>& rop,
                                   const std::vector<double>& dpr, int rr, size_t off) {
  const int m = t.m, nR = t.nReac;
  (void)c; (void)kf; (void)rop;
  std::mt19937_64 rng(3);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  std::vector<double> mu(m + 1, 0.0), w(nR + 2, 0.0), v(m + 1, 0.0);
  for (int i = 0; i < m; ++i) mu[i] = U(rng);
  w[nR + 1] = std::nan("");
  CHECK(off == size_t(s.nutBody), "nut body");
  std::vector<int> wr(nR, 0), vr(m, 0);
  const int nutEnd = rr + s.nutRounds;
  for (; rr < (int)s.evalHdr.size(); ++rr) {
    const int kind = s.evalHdr[rr] >> 16, len = s.evalHdr[rr] & 0xffff;
    CHECK(kind == (rr < nutEnd ? SparseSchedule::kNuT : SparseSchedule::kJtv), "adjoint section kind");
    if (rr == nutEnd) CHECK(off == size_t(s.jtvBody), "jtv body");
    std::map<unsigned, int> tg;
    for (int l = 0; l < 64; ++l) {
      const unsigned tw = s.progBody.empty() ? 0 : s.evalBody[off + l];
      const unsigned tt = tw & ~SparseSchedule::kCont;
      double acc = 0.0;
      if (kind == SparseSchedule::kNuT) {
        CHECK((int)tt <= nR + 1, "w target");
        if ((int)tt < nR) { CHECK(!tg.count(tt), "dup w target"); tg[tt] = l; if (!(tw & SparseSchedule::kCont)) wr[tt]++; }
        acc = (tw & SparseSchedule::kCont) ? w[tt] : 0.0;
        for (int j = 0; j < len; ++j) {
          const unsigned e = s.evalBody[off + 64 * (1 + j) + l];
          CHECK((int)(e & 0xffff) <= m && (e >> 16) < s.coefTable.size(), "nut entry");
          acc += s.coefTable[e >> 16] * mu[e & 0xffff];
        }
        if ((int)tt <= nR + 1) w[tt] = acc;
      } else {
        CHECK((int)tt <= m, "v target");
        if ((int)tt < m) { CHECK(!tg.count(tt), "dup v target"); tg[tt] = l; if (!(tw & SparseSchedule::kCont)) vr[tt]++; }
        acc = (tw & SparseSchedule::kCont) ? v[tt] : 0.0;
        for (int j = 0; j < len; ++j) {
          const unsigned e = s.evalBody[off + 64 * (1 + j) + l];
          CHECK((int)(e & 0xffff) < s.pairSlots && (int)(e >> 16) <= nR, "jtv entry");
          acc += dpr[e & 0xffff] * w[e >> 16];
        }
        if ((int)tt <= m) v[tt] = acc;
      }
    }
    off += 64 * (1 + len);
  }
  CHECK(off + 64 * SparseSchedule::kSlack == s.evalBody.size(), "adjoint body end");
  for (int r = 0; r < nR; ++r) CHECK(wr[r] == 1, "w[%d] initialised %d times", r, wr[r]);
  for (int j = 0; j < m; ++j) CHECK(vr[j] == 1, "v[%d] initialised %d times", j, vr[j]);
  // direct: w_r = sum_a nu_ar mu_a; v = J^T mu with J = sum_p nu_r(p) dpr_p e_s(p)^T
  std::vector<long double> wd(nR, 0.0L), vd(m, 0.0L), vs(m, 0.0L);
  for (int a = 0; a < m; ++a)
    for (int e = t.omegaPtr[a]; e < t.omegaPtr[a + 1]; ++e) wd[t.omegaReac[e]] += (long double)t.omegaCoef[e] * mu[a];
  double we = 0, ve = 0;
  for (int r = 0; r < nR; ++r) we = std::max(we, (double)fabsl(wd[r] - w[r]));
  for (int n = 0; n < t.nnzJ(); ++n)
    for (int e = t.jnzPtr[n]; e < t.jnzPtr[n + 1]; ++e) {
      const long double jv = (long double)t.jcCoef[e] * dpr[t.jcPair[e]];
      vd[t.jnzCol[n]] += jv * mu[t.jnzRow[n]];
      vs[t.jnzCol[n]] += fabsl(jv * mu[t.jnzRow[n]]);
    }
  for (int j = 0; j < m; ++j) ve = std::max(ve, (double)(fabsl(vd[j] - v[j]) / (vs[j] + 1e-300)));
  std::printf("adjoint sections: nu^T mu abs err %.2e, J^T mu rel err %.2e\n", we, ve);
  CHECK(we < 1e-12 && ve < 1e-13, "adjoint sections");
}

static void check_eval_program(const MechTables& t, const SparseSchedule& s) {

  const int m = t.m, nR = t.nReac, nS = t.nSpec;
  std::printf("eval: rop %d pair %d seg %d jac %d rounds; body %.1f KB; coef %zu; lu program %.1f KB\n", s.ropRounds, s.pairRounds,
              s.segRounds, s.jacRounds, s.evalBody.size() * 4 / 1024.0, s.coefTable.size(), s.progBody.size() * 4 / 1024.0);
  int segSlots = 0, jacSlots = 0;
  for (int r = s.ropRounds + s.pairRounds; r < (int)s.evalHdr.size(); ++r) { int len = s.evalHdr[r] & 0xffff, kind = s.evalHdr[r] >> 16; (kind == SparseSchedule::kSeg ? segSlots : jacSlots) += 1 + len; }
  std::printf("slots: seg %d jac %d\n", segSlots, jacSlots);
  CHECK(s.coefTable[0] == 0.0, "coef 0");
  std::mt19937_64 rng(11);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  auto fac = [&](const std::vector<double>& c, unsigned f) { if (f == SparseSchedule::kNone) return 1.0; CHECK((int)(f & 0xfff) < nS, "species idx"); return ipow(c[f & 0xfff], int(f >> 12)); };
  for (int trial = 0; trial < 3; ++trial) {
    std::vector<double> c(nS), kf(nR + 1, 0.0), rop(nR + 1, 0.0), dpr(s.pairSlots, 0.0), seg(t.nSeg() + 1, 0.0);
    for (auto& x : c) x = std::pow(10.0, -3 + 5 * U(rng));
    for (int r = 0; r < nR; ++r) kf[r] = std::pow(10.0, -6 + 6 * U(rng));
    size_t off = 0; int rr = 0;
    for (int g = 0; g < s.ropRounds; ++g, ++rr) {
      CHECK((int)(s.evalHdr[rr] >> 16) == SparseSchedule::kRop, "rop kind");
      for (int l = 0; l < 64; ++l) { unsigned w0 = s.evalBody[off + l], w1 = s.evalBody[off + 64 + l]; int r = w0 & 0xffff; CHECK(r <= nR, "rop r");
        double v = kf[r] * fac(c, w0 >> 16) * fac(c, w1 & 0xffff) * fac(c, w1 >> 16); if (r == nR) CHECK(v == 0.0, "pad rop"); rop[r] = v; }
      off += 128;
    }
    CHECK(off == (size_t)s.pairBody, "pair body");
    for (int g = 0; g < s.pairRounds; ++g, ++rr) {
      CHECK((int)(s.evalHdr[rr] >> 16) == SparseSchedule::kPair, "pair kind");
      for (int l = 0; l < 64; ++l) { unsigned w0 = s.evalBody[off + l], w1 = s.evalBody[off + 64 + l];
        int r = w0 & 0x3fff, js = (w0 >> 14) & 3; unsigned f[3] = {w0 >> 16, w1 & 0xffff, w1 >> 16}; CHECK(r <= nR && js < 3 && f[js] != SparseSchedule::kNone, "pair rec");
        int ns = f[js] >> 12; double d = kf[r] * ns * ipow(c[f[js] & 0xfff], ns - 1);
        for (int j = 0; j < 3; ++j) if (j != js) d *= fac(c, f[j]);
        CHECK(g * 64 + l < s.pairSlots, "dpr slot"); dpr[g * 64 + l] = d; }
      off += 128;
    }
    CHECK(off == (size_t)s.segBody, "seg body");
    CHECK(dpr[t.nPairs()] == 0.0 && rop[nR] == 0.0, "zero slots");
    // direct rop / dpr
    for (int r = 0; r < nR; ++r) { double v = kf[r]; for (int j = 0; j < t.maxReactants; ++j) { int sp = t.reacSpec[r * t.maxReactants + j]; if (sp < 0) break; v *= ipow(c[sp], t.reacPow[r * t.maxReactants + j]); } CHECK(v == rop[r], "rop %d", r); }
    // seg + jac
    const double ch = 0.1 * (trial + 1);
    std::vector<double> A(s.trashSlot + 1, std::nan("")); A[s.zeroSlot] = 0;
    std::vector<int> jwrites(s.nnzF, 0);
    const int rEnd = s.ropRounds + s.pairRounds + s.segRounds + s.jacRounds;
    for (; rr < rEnd; ++rr) {
      int kind = s.evalHdr[rr] >> 16, len = s.evalHdr[rr] & 0xffff;
      CHECK(len >= 1 && len <= SparseSchedule::kMaxChain, "len");
      std::map<unsigned, int> tg;
      for (int l = 0; l < 64; ++l) {
        unsigned tw = s.evalBody[off + l];
        if (kind == SparseSchedule::kSeg) {
          unsigned tt = tw & ~SparseSchedule::kCont; CHECK((int)tt <= t.nSeg(), "seg target");
          if ((int)tt < t.nSeg()) { CHECK(!tg.count(tt), "dup seg target"); tg[tt] = l; }
          double acc = (tw & SparseSchedule::kCont) ? seg[tt] : 0.0;
          for (int j = 0; j < len; ++j) { unsigned e = s.evalBody[off + 64 * (1 + j) + l]; CHECK((e & 0xffff) <= (unsigned)nR && (e >> 16) < s.coefTable.size(), "seg entry"); acc += s.coefTable[e >> 16] * rop[e & 0xffff]; }
          seg[tt] = acc;
        } else {
          CHECK(kind == SparseSchedule::kJac, "kind");
          unsigned tt = tw & 0x3fffffffu; CHECK((int)tt < s.zeroSlot || (int)tt == s.trashSlot, "jac target %u", tt);
          if ((int)tt != s.trashSlot) { CHECK(!tg.count(tt), "dup jac target"); tg[tt] = l; if (!(tw & SparseSchedule::kCont)) jwrites[tt]++; }
          double init = (tw & SparseSchedule::kCont) ? A[tt] : (((tw >> 30) & 1) ? 1.0 : 0.0);
          double sum = 0; for (int j = 0; j < len; ++j) { unsigned e = s.evalBody[off + 64 * (1 + j) + l]; CHECK((int)(e & 0xffff) < s.pairSlots && (e >> 16) < s.coefTable.size(), "jac entry"); sum += s.coefTable[e >> 16] * dpr[e & 0xffff]; }
          A[tt] = init + (-ch) * sum;
        }
      }
      off += 64 * (1 + len);
    }
    if (s.transposed) check_adjoint_sections(t, s, c, kf, rop, dpr, rr, off);
    else CHECK(off + 64 * SparseSchedule::kSlack == s.evalBody.size(), "eval body end");
    for (int p = 0; p < s.nnzF; ++p) CHECK(jwrites[p] == 1, "jac position %d initialised %d times", p, jwrites[p]);
    // omega vs direct
    double maxe = 0;
    for (int a = 0; a < m; ++a) {
      double f = t.emission[a]; for (int g = t.rowSegPtr[a]; g < t.rowSegPtr[a + 1]; ++g) f += seg[g];
      long double v = t.emission[a]; for (int e = t.omegaPtr[a]; e < t.omegaPtr[a + 1]; ++e) v += (long double)t.omegaCoef[e] * rop[t.omegaReac[e]];
      long double scale = 0; for (int e = t.omegaPtr[a]; e < t.omegaPtr[a + 1]; ++e) scale += fabsl((long double)t.omegaCoef[e] * rop[t.omegaReac[e]]);
      maxe = std::max(maxe, (double)(fabsl(v - f) / (scale + 1e-300)));
    }
    CHECK(maxe < 1e-14, "omega err %g", maxe);
    // J vs direct on pattern
    std::vector<long double> D(size_t(m) * m, 0.0L);
    for (int i = 0; i < m; ++i) D[size_t(i) * m + i] = 1.0L;
    std::vector<long double> scl(size_t(m) * m, 0.0L);
    for (int n = 0; n < t.nnzJ(); ++n) {
      const size_t at = s.transposed ? size_t(t.jnzCol[n]) * m + t.jnzRow[n] : size_t(t.jnzRow[n]) * m + t.jnzCol[n];
      for (int e = t.jnzPtr[n]; e < t.jnzPtr[n + 1]; ++e) {
        long double v = (long double)t.jcCoef[e] * dpr[t.jcPair[e]];
        D[at] += -ch * v;
        scl[at] += fabsl(ch * v);
      }
    }
    double je = 0;
    for (int i = 0; i < m; ++i) for (int p = s.rowPtr[i]; p < s.rowPtr[i + 1]; ++p) { int j = s.colIdx[p]; long double sc = scl[size_t(i) * m + j] + (i == j); je = std::max(je, (double)(fabsl(D[size_t(i) * m + j] - A[size_t(s.slotOf[size_t(p)])]) / (sc + 1e-300))); }
    std::printf("trial %d: omega rel err %.2e, J rel err %.2e\n", trial, maxe, je);
    CHECK(je < 1e-14, "J err");
  }
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  KineticModel km = loadKineticModel(argv[1], nullptr);
  MechTables t = compileMechanism(km);
  for (bool tr : {false, true}) {
    SparseSchedule s = buildSparseSchedule(t, tr);
    std::printf("-- %s programs\n", tr ? "transposed" : "forward");
    check_eval_program(t, s);
    check_lu_program(t, s);
  }
  if (argc > 2) {
    AerosolModel am = loadAerosolModel(argv[2], km);
    std::vector<std::array<int, 2>> ent;
    MechTables ta = aerosolTables(t, am, &ent);
    SparseSchedule s = buildSparseSchedule(ta, false);
    std::printf("-- aerosol LU program (%d transfer derivatives)\n", int(ent.size()));
    for (const auto& e : ent) {
      bool found = false;
      for (int q = s.rowPtr[e[0]]; q < s.rowPtr[e[0] + 1]; ++q) found = found || s.colIdx[q] == e[1];
      CHECK(found, "transfer derivative (%d, %d) outside the factor pattern", e[0], e[1]);
    }
    check_lu_program(ta, s);
  }
  std::printf("%s (%d failures)\n", fails ? "FAILED" : "OK", fails);
  return fails ? 1 : 0;
}
