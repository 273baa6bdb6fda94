#include "aerosol_model.hpp"

#include <stdexcept>

#include "yaml_lite.hpp"

namespace tchem_amd {

using yaml::Node;

namespace {
[[noreturn]] void fail(const std::string& msg) {
  throw std::runtime_error("Yaml : Error when interpreting aerosol model: " + msg);
}
double number(const Node& n, const std::string& key, const std::string& who) {
  const Node* v = n.find(key);
  if (!v) fail(who + " has no '" + key + "'");
  return v->asDouble();
}
}  // namespace

AerosolModel loadAerosolModel(const std::string& path, const KineticModel& km) {
  const Node root = yaml::parseFile(path);
  const Node* items = root.find("NCAR-version");
  if (!items || !items->isSeq()) throw std::runtime_error("we do not have a parser for this file: " + path);
  AerosolModel am;
  am.nGasActive = km.nActive();
  am.nGasConst = km.nConstSpec;

  std::map<std::string, const Node*> gasInfo;  // CHEM_SPEC items without a phase, by name
  for (const Node& item : items->items) {
    if (!item.isMap() || !item.has("type")) continue;
    const std::string type = item["type"].str();
    if (type == "CHEM_SPEC") {
      const std::string name = item["name"].str();
      const Node* phase = item.find("phase");
      if (phase) {
        if (phase->str() == "AEROSOL") {
          am.index[name] = int(am.names.size());
          am.names.push_back(name);
          am.mw.push_back(number(item, "molecular weight [kg mol-1]", name));
          am.density.push_back(number(item, "density [kg m-3]", name));
          am.kappa.push_back(number(item, "kappa", name));
        }
      } else {
        gasInfo.insert({name, &item});  // first definition wins (std::map::insert)
      }
    } else if (type == "AERO_REP_SINGLE_PARTICLE") {
      am.nParticles = item["maximum computational particles"].asInt();
    }
  }
  am.nSpec = int(am.names.size());

  for (const Node& item : items->items) {
    if (!item.isMap() || !item.has("type") || item["type"].str() != "MECHANISM") continue;
    const Node* reactions = item.find("reactions");
    if (!reactions) continue;
    for (const Node& r : reactions->items) {
      if (r["type"].str() != "SIMPOL_PHASE_TRANSFER") continue;
      SimpolParam s{};
      const Node& B = r["B"];
      if (B.size() < 4) fail("SIMPOL_PHASE_TRANSFER needs four B parameters");
      s.B1 = B.items[0].asDouble();
      s.B2 = B.items[1].asDouble();
      s.B3 = B.items[2].asDouble();
      s.B4 = B.items[3].asDouble();
      const std::string aero = r["aerosol-phase species"].str();
      auto ia = am.index.find(aero);
      if (ia == am.index.end()) fail("species does not exist " + aero + " in SIMPOL_PHASE_TRANSFER");
      s.aero = ia->second;
      const std::string gas = r["gas-phase species"].str();
      auto ig = km.speciesIndex.find(gas);
      if (ig == km.speciesIndex.end()) fail("species does not exist " + gas + " in SIMPOL_PHASE_TRANSFER");
      s.gas = ig->second;
      if (s.gas >= am.nGasActive) fail("gas-phase species " + gas + " of SIMPOL_PHASE_TRANSFER is a constant species");
      auto gi = gasInfo.find(gas);
      if (gi == gasInfo.end()) fail("gas-phase species " + gas + " has no CHEM_SPEC entry in the aerosol file");
      const Node& g = *gi->second;
      s.diffusion = number(g, "diffusion coeff [m2 s-1]", gas);
      if (const Node* ns = g.find("N star")) {
        s.Nstar = ns->asDouble();
        s.computeAlpha = 1;
      } else {
        s.Nstar = 999.0;
        s.computeAlpha = 0;
      }
      s.mw = number(g, "molecular weight [kg mol-1]", gas);
      am.simpol.push_back(s);
    }
  }
  if (am.nParticles < 0) fail("no AERO_REP_SINGLE_PARTICLE with 'maximum computational particles'");
  return am;
}

bool loadScenarioParticles(const std::string& path, const AerosolModel& am, const KineticModel& km, int nBatch,
                           std::vector<double>& state, std::vector<double>& numberConc) {
  const Node root = yaml::parseFile(path);
  const Node* particles = root.find("particles");
  if (!particles) return false;
  const int dim = am.stateDim(km);
  if (state.size() < size_t(nBatch) * size_t(dim)) throw std::runtime_error("particles: state rows too short");
  const Node& nc = (*particles)["num_concentration"]["initial_value"];
  if (int(nc.size()) < nBatch) throw std::runtime_error("particles: too few num_concentration values");
  numberConc.assign(size_t(nBatch) * size_t(am.nParticles), 0.0);
  for (int i = 0; i < nBatch; ++i)
    for (int p = 0; p < am.nParticles; ++p) numberConc[size_t(i) * am.nParticles + p] = nc.items[i].asDouble();
  const int base = 3 + km.nSpec;
  if (const Node* init = particles->find("initial_state")) {
    for (const auto& e : init->entries) {
      auto it = am.index.find(e.first);
      if (it == am.index.end()) fail("species does not exist " + e.first + " in the particle initial state");
      const Node& vals = e.second["initial_value"];
      if (int(vals.size()) < nBatch) throw std::runtime_error("particles: too few initial values for " + e.first);
      for (int i = 0; i < nBatch; ++i) {
        const double v = vals.items[i].asDouble();
        for (int p = 0; p < am.nParticles; ++p)
          state[size_t(i) * dim + base + am.nSpec * p + it->second] = v / numberConc[size_t(i) * am.nParticles + p];
      }
    }
  }
  return true;
}

MechTables aerosolTables(const MechTables& t, const AerosolModel& am, std::vector<std::array<int, 2>>* entries) {
  const int np = am.nSpec * am.nParticles, mt = t.m + np;
  auto piv = [&](int u) { return aerosolPivot(u, t.m, np); };
  MechTables ta = t;
  ta.m = mt;
  std::vector<char> have(size_t(mt) * mt, 0);
  for (size_t q = 0; q < t.jnzRow.size(); ++q) {
    ta.jnzRow[q] = piv(t.jnzRow[q]);
    ta.jnzCol[q] = piv(t.jnzCol[q]);
    have[size_t(ta.jnzRow[q]) * mt + ta.jnzCol[q]] = 1;
  }
  std::vector<std::array<int, 2>> ent;
  for (int ip = 0; ip < am.nParticles; ++ip)
    for (const SimpolParam& sp : am.simpol) {
      const int g = piv(t.iperm[sp.gas]), base = piv(t.m + ip * am.nSpec), ai = base + sp.aero;
      ent.push_back({g, g});
      ent.push_back({ai, g});
      for (int i = 0; i < am.nSpec; ++i) ent.push_back({g, base + i});
      for (int i = 0; i < am.nSpec; ++i) ent.push_back({ai, base + i});
    }
  for (const auto& e : ent) {
    if (have[size_t(e[0]) * mt + e[1]]) continue;
    have[size_t(e[0]) * mt + e[1]] = 1;
    ta.jnzRow.push_back(e[0]);
    ta.jnzCol.push_back(e[1]);
    ta.jnzPtr.push_back(ta.jnzPtr.back());  // no gas-phase contributions
  }
  if (entries) *entries = ent;
  return ta;
}

}  // namespace tchem_amd
