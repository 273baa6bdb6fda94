#include "mechanism.hpp"

#include <algorithm>
#include <cstdio>
#include <map>
#include <stdexcept>

namespace tchem_amd {

MechTables compileMechanism(const KineticModel& km) {
  MechTables t;
  t.nSpec = km.nSpec;
  t.nConst = km.nConstSpec;
  t.m = km.nActive();
  t.nReac = km.nReac;
  t.convPPM = km.CONV_PPM;
  const int nR = km.nReac, m = t.m;

  // rate-constant records
  t.rateKind.assign(nR, kRateArrhenius);
  for (int r = 0; r < nR; ++r) {
    const int ty = km.reactionType[r];
    if (ty != kArrhenius && ty != kTroe && ty != kCmaqH2O2) {
      t.rateKind[r] = kRateExternal;
      t.externalRates = true;
    }
  }
  t.ratePar.assign(size_t(nR) * 8, 0.0);
  t.convExp = km.convExponent;
  t.photoSlot.assign(nR, -1);
  for (const auto& a : km.arrhenius) {
    t.rateKind[a.reaction] = kRateArrhenius;
    double* p = &t.ratePar[size_t(a.reaction) * 8];
    p[0] = a.A; p[1] = a.B; p[2] = a.C; p[3] = a.D; p[4] = a.E;
  }
  for (size_t i = 0; i < km.reacPfal.size(); ++i) {
    const int r = km.reacPfal[i];
    t.rateKind[r] = kRateTroe;
    for (int j = 0; j < 8; ++j) t.ratePar[size_t(r) * 8 + j] = km.reacPpar[i][j];
  }
  for (const auto& c : km.cmaqH2O2) {
    t.rateKind[c.reaction] = kRateCmaqH2O2;
    double* p = &t.ratePar[size_t(c.reaction) * 8];
    p[0] = c.A1; p[1] = c.B1; p[2] = c.C1; p[3] = c.A2; p[4] = c.B2; p[5] = c.C2;
  }
  t.photoReac = km.photolysisReactions;
  t.nPhoto = int(t.photoReac.size());
  for (int q = 0; q < t.nPhoto; ++q) t.photoSlot[t.photoReac[q]] = q;

  // net coefficient of species i in reaction r (active species only);
  // reactant and product entries of the same species are combined
  std::vector<std::map<int, double>> net(nR);  // r -> (species -> nu), reference numbering
  for (int r = 0; r < nR; ++r) {
    for (int j = 0; j < km.reacNreac[r]; ++j) {
      const int s = km.sidx(r, j);
      if (s < m) net[r][s] += km.nuki(r, j);
    }
    const int joff = km.maxSpecInReac / 2;
    for (int j = 0; j < km.reacNprod[r]; ++j) {
      const int s = km.sidx(r, j + joff);
      const double nu = km.nuki(r, j + joff);
      if (s < m && nu != 0.0) net[r][s] += nu;  // unknown products keep (0, 0.0)
    }
  }

  // structure of I - J in the reference numbering -> elimination order
  {
    std::vector<char> pat(size_t(m) * m, 0);
    for (int i = 0; i < m; ++i) pat[size_t(i) * m + i] = 1;
    for (int r = 0; r < nR; ++r)
      for (int j = 0; j < km.reacNreac[r]; ++j) {
        const int s = km.sidx(r, j);
        if (s >= m) continue;
        for (const auto& e : net[r])
          if (e.second != 0.0) pat[size_t(e.first) * m + s] = 1;
      }
    t.perm = markowitzOrder(m, pat);
    t.iperm.assign(m, 0);
    for (int a = 0; a < m; ++a) t.iperm[t.perm[a]] = a;
  }
  auto in = [&](int s) { return s < m ? t.iperm[s] : s; };  // reference -> internal index

  // reactants (ordered by species index, as the reference's std::map)
  int maxR = 0;
  for (int r = 0; r < nR; ++r) maxR = std::max(maxR, km.reacNreac[r]);
  t.maxReactants = std::max(maxR, 1);
  t.reacSpec.assign(size_t(nR) * t.maxReactants, -1);
  t.reacPow.assign(size_t(nR) * t.maxReactants, 0);
  for (int r = 0; r < nR; ++r)
    for (int j = 0; j < km.reacNreac[r]; ++j) {
      t.reacSpec[size_t(r) * t.maxReactants + j] = in(km.sidx(r, j));
      t.reacPow[size_t(r) * t.maxReactants + j] = std::abs(int(km.nuki(r, j)));
    }

  // source term CSR (internal rows, entries in reaction order) and segments
  std::vector<std::vector<std::pair<int, double>>> rows(m);
  for (int r = 0; r < nR; ++r)
    for (const auto& e : net[r])
      if (e.second != 0.0) rows[t.iperm[e.first]].push_back({r, e.second});
  t.omegaPtr.assign(m + 1, 0);
  t.rowSegPtr.assign(m + 1, 0);
  for (int a = 0; a < m; ++a) {
    t.omegaPtr[a + 1] = t.omegaPtr[a] + int(rows[a].size());
    for (const auto& e : rows[a]) {
      t.omegaReac.push_back(e.first);
      t.omegaCoef.push_back(e.second);
    }
    for (int b = t.omegaPtr[a]; b < t.omegaPtr[a + 1]; b += MechTables::kOmegaSegment) {
      t.segBegin.push_back(b);
      t.segEnd.push_back(std::min(b + MechTables::kOmegaSegment, t.omegaPtr[a + 1]));
    }
    t.rowSegPtr[a + 1] = int(t.segBegin.size());
  }
  t.emission.assign(m, 0.0);
  for (const auto& e : km.emission) {
    if (e.species < m) t.emission[t.iperm[e.species]] = e.rate;  // assignment: the last entry wins
    else std::fprintf(stderr, "tchem: emission for constant species %d ignored\n", e.species);
  }

  // Jacobian pairs and contributions
  t.pairPtr.assign(nR + 1, 0);
  std::vector<int> pairReac, pairSpec;
  for (int r = 0; r < nR; ++r) {
    for (int j = 0; j < km.reacNreac[r]; ++j) {
      const int s = km.sidx(r, j);
      if (s < m) {
        t.pairSlot.push_back(j);
        pairReac.push_back(r);
        pairSpec.push_back(t.iperm[s]);
      }
    }
    t.pairPtr[r + 1] = int(t.pairSlot.size());
  }
  std::map<std::pair<int, int>, std::vector<std::pair<int, double>>> jmap;  // (row, col) -> contributions
  for (int i = 0; i < m; ++i) jmap[{i, i}];                                // structural diagonal
  for (int p = 0; p < t.nPairs(); ++p) {
    const int r = pairReac[p], s = pairSpec[p];
    for (const auto& e : net[r])
      if (e.second != 0.0) jmap[{t.iperm[e.first], s}].push_back({p, e.second});
  }
  t.jnzPtr.push_back(0);
  for (const auto& e : jmap) {
    t.jnzRow.push_back(e.first.first);
    t.jnzCol.push_back(e.first.second);
    for (const auto& c : e.second) {
      t.jcPair.push_back(c.first);
      t.jcCoef.push_back(c.second);
    }
    t.jnzPtr.push_back(int(t.jcPair.size()));
  }

  // serial gathers (dense form): uncombined entries in the reference's order
  {
    const int joff = km.maxSpecInReac / 2;
    std::vector<std::vector<std::pair<int, double>>> srows(m);
    for (int r = 0; r < nR; ++r) {
      for (int j = 0; j < km.reacNreac[r]; ++j)
        if (km.sidx(r, j) < m) srows[t.iperm[km.sidx(r, j)]].push_back({r, km.nuki(r, j)});
      for (int j = 0; j < km.reacNprod[r]; ++j)
        if (km.sidx(r, j + joff) < m) srows[t.iperm[km.sidx(r, j + joff)]].push_back({r, km.nuki(r, j + joff)});
    }
    t.serOmegaPtr.assign(m + 1, 0);
    for (int a = 0; a < m; ++a) {
      t.serOmegaPtr[a + 1] = t.serOmegaPtr[a] + int(srows[a].size());
      for (const auto& e : srows[a]) {
        t.serOmegaReac.push_back(e.first);
        t.serOmegaCoef.push_back(e.second);
      }
    }
    std::map<std::pair<int, int>, std::vector<std::pair<int, double>>> smap;
    for (int i = 0; i < m; ++i) smap[{i, i}];
    int p = 0;  // pairs are numbered in (reaction, active reactant slot) order
    for (int r = 0; r < nR; ++r)
      for (int js = 0; js < km.reacNreac[r]; ++js) {
        const int ks = km.sidx(r, js);
        if (ks >= m) continue;
        const int col = t.iperm[ks];
        for (int j = 0; j < km.reacNreac[r]; ++j)
          if (km.sidx(r, j) < m) smap[{t.iperm[km.sidx(r, j)], col}].push_back({p, km.nuki(r, j)});
        for (int j = 0; j < km.reacNprod[r]; ++j)
          if (km.sidx(r, j + joff) < m) smap[{t.iperm[km.sidx(r, j + joff)], col}].push_back({p, km.nuki(r, j + joff)});
        ++p;
      }
    t.serJPtr.push_back(0);
    for (const auto& e : smap) {
      t.serJRow.push_back(e.first.first);
      t.serJCol.push_back(e.first.second);
      for (const auto& c : e.second) {
        t.serJPair.push_back(c.first);
        t.serJCoef.push_back(c.second);
      }
      t.serJPtr.push_back(int(t.serJPair.size()));
    }
  }
  return t;
}

std::vector<int> markowitzOrder(int m, std::vector<char> pat) {
  std::vector<int> order;
  std::vector<char> done(m, 0);
  order.reserve(m);
  for (int step = 0; step < m; ++step) {
    long best = -1;
    int bi = -1;
    for (int i = 0; i < m; ++i) {
      if (done[i]) continue;
      long rc = 0, cc = 0;
      for (int j = 0; j < m; ++j)
        if (!done[j]) {
          rc += pat[size_t(i) * m + j];
          cc += pat[size_t(j) * m + i];
        }
      const long score = (rc - 1) * (cc - 1);
      if (bi < 0 || score < best) { best = score; bi = i; }
    }
    order.push_back(bi);
    done[bi] = 1;
    for (int r = 0; r < m; ++r) {
      if (done[r] || !pat[size_t(r) * m + bi]) continue;
      for (int c = 0; c < m; ++c)
        if (!done[c] && pat[size_t(bi) * m + c]) pat[size_t(r) * m + c] = 1;
    }
  }
  return order;
}

}  // namespace tchem_amd
