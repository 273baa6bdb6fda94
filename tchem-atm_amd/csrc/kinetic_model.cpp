#include "kinetic_model.hpp"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <stdexcept>

#include "yaml_lite.hpp"

namespace tchem_amd {

namespace {

using yaml::Node;

void echoPrint(std::ostream* os, const char* fmt, ...) {
  if (!os) return;
  char buf[1024];
  va_list args;
  va_start(args, fmt);
  const int n = std::vsnprintf(buf, sizeof(buf), fmt, args);
  va_end(args);
  if (n < int(sizeof(buf))) {
    *os << buf;
    return;
  }
  std::vector<char> big(size_t(n) + 1);
  va_start(args, fmt);
  std::vsnprintf(big.data(), big.size(), fmt, args);
  va_end(args);
  *os << big.data();
}

std::string truncName(const std::string& s) {
  return s.size() > size_t(kLengthOfSpecName) ? s.substr(0, kLengthOfSpecName) : s;
}

// Optional coefficient lookup with the reference's default values.
double coef(const Node* coefs, const char* key, double dflt, bool* present = nullptr) {
  const Node* n = coefs ? coefs->find(key) : nullptr;
  if (present) *present = (n != nullptr);
  return n ? n->asDouble() : dflt;
}

bool timeUnitIsMinutes(const Node& reaction) {
  const Node* tu = reaction.find("time_unit");
  return tu && tu->isScalar() && tu->str() == "min";
}

[[noreturn]] void fail(const std::string& msg) { throw std::runtime_error("Yaml : Error when interpreting kinetic model: " + msg); }

}  // namespace

KineticModel loadKineticModel(const std::string& path, std::ostream* echo) {
  const Node root = yaml::parseFile(path);
  if (!root.has("NCAR-version")) throw std::runtime_error("we do not have a parser for this file: " + path);

  KineticModel km;
  const Node& species = root["species"];
  static const Node kEmpty;
  const Node* reactionsNode = root.find("reactions");  // a mechanism may have sources only
  const Node& reactions = reactionsNode ? *reactionsNode : kEmpty;
  const Node* constSpecies = root.find("constant_species");

  // species (src/core/TChem_KineticModelData.cpp:88-131)
  km.nConstSpec = constSpecies ? int(constSpecies->size()) : 0;
  km.nSpec = int(species.size()) + km.nConstSpec;
  km.nReac = int(reactions.size());
  int sp_i = 0;
  auto addSpecies = [&](const Node& sp) {
    const std::string name = sp["name"].str();
    km.speciesIndex.insert({name, sp_i});  // first occurrence wins (std::map::insert)
    km.speciesNames.push_back(truncName(name));
    ++sp_i;
  };
  for (const Node& sp : species.items) addSpecies(sp);
  if (constSpecies)
    for (const Node& sp : constSpecies->items) addSpecies(sp);

  // M is mandatory (.cpp:134-146)
  {
    auto it = km.speciesIndex.find("M");
    if (it == km.speciesIndex.end()) fail("M species does not exist; the name of the M species is hard-coded to M");
    km.M_index = it->second;
  }

  echoPrint(echo, "kmod.list : # of species                                      : %d\n", km.nSpec);
  echoPrint(echo, "kmod.list : # of species with constant concentration          : %d\n", km.nConstSpec);
  echoPrint(echo, "kmod.list : # of reactions                                    : %d\n", km.nReac);
  echoPrint(echo, "No. \t Species\n");
  for (int i = 0; i < km.nSpec; i++) echoPrint(echo, "%-3d\t%-32s\n", i + 1, km.speciesNames[i].c_str());
  echoPrint(echo, "-------------------------------------------------------------------------\n");

  // reactions, pass 1: stoichiometry and counts (.cpp:190-282)
  const int nR = km.nReac;
  km.isRev.assign(nR, 1);
  km.reacNrp.assign(nR, 0);
  km.reacNreac.assign(nR, 0);
  km.reacNprod.assign(nR, 0);
  km.convExponent.assign(nR, 0);
  km.reactionType.assign(nR, 0);
  km.note.assign(nR, std::string());

  std::vector<std::map<int, double>> reactantsInfo(nR), productsInfo(nR);
  int countArrhenius = 0, countCmaq = 0, countJpl = 0, countRJpl = 0, countAdjust = 0;
  km.nFallReac = 0;
  km.maxSpecInReac = 0;
  for (int r = 0; r < nR; ++r) {
    const Node& reaction = reactions.items[r];
    const Node* reactants = reaction.find("reactants");
    if (!reactants) fail("reaction No " + std::to_string(r) + " does not have reactants");
    km.reacNreac[r] = int(reactants->size());
    for (const auto& e : reactants->entries) {
      auto it = km.speciesIndex.find(e.first);
      if (it == km.speciesIndex.end())
        fail("species does not exist " + e.first + " in reactants of reaction No " + std::to_string(r));
      reactantsInfo[r].insert({it->second, e.second.asDouble()});
    }
    const Node* products = reaction.find("products");
    if (!products) fail("reaction No " + std::to_string(r) + " does not have products");
    km.reacNprod[r] = int(products->size());
    for (const auto& e : products->entries) {
      auto it = km.speciesIndex.find(e.first);
      if (it != km.speciesIndex.end()) {
        productsInfo[r].insert({it->second, e.second.asDouble()});
      } else {
        // the reference only warns: the slot keeps index 0 and coefficient 0
        std::fprintf(stderr, "Yaml : Warning: species does not exist %s in products of reaction No %d\n",
                     e.first.c_str(), r);
      }
    }
    km.reacNrp[r] = km.reacNreac[r] + km.reacNprod[r];
    km.maxSpecInReac = std::max(km.maxSpecInReac, km.reacNreac[r]);
    km.maxSpecInReac = std::max(km.maxSpecInReac, km.reacNprod[r]);

    const std::string type = reaction["type"].str();
    if (type == "ARRHENIUS") { km.reactionType[r] = kArrhenius; ++countArrhenius; }
    else if (type == "TROE") { km.reactionType[r] = kTroe; ++km.nFallReac; }
    else if (type == "CMAQ_H2O2") { km.reactionType[r] = kCmaqH2O2; ++countCmaq; }
    else if (type == "JPL") { km.reactionType[r] = kJpl; ++countJpl; }
    else if (type == "R_JPL_ARRHENIUS") { km.reactionType[r] = kRJplArrhenius; ++countRJpl; }
    else if (type == "PHOTO_RATE") { km.reactionType[r] = kPhotoRate; ++km.countPhotoRate; }
    else fail("reaction type does not exist: reaction No " + std::to_string(r) + " type " + type);

    if (const Node* adj = reaction.find("adjust_reaction")) countAdjust += int(adj->size());
    if (const Node* note = reaction.find("note"))
      if (note->isScalar()) km.note[r] = note->str();
  }
  km.maxSpecInReac *= 2;

  echoPrint(echo, "kmod.list : Max # of species in a reaction                    : %d\n", km.maxSpecInReac);
  echoPrint(echo, "kmod.list :  # of arrhenius type reactions                    : %d\n", countArrhenius);
  echoPrint(echo, "kmod.list :  # of CMAQ_H2O2 type reactions                    : %d\n", countCmaq);
  echoPrint(echo, "kmod.list :  # of Troe type reactions                         : %d\n", km.nFallReac);
  echoPrint(echo, "kmod.list :  # of JPL-Troe type reactions                     : %d\n", countJpl);
  echoPrint(echo, "kmod.list :  # of Ratio JPL arrhenius type reactions          : %d\n", countRJpl);
  echoPrint(echo, "kmod.list :  # of adjustment pairs                            : %d\n", countAdjust);
  echoPrint(echo, "kmod.list :  # of Photolysis rates                            : %d\n", km.countPhotoRate);

  // stoichiometric tables (.cpp:302-338); unused slots stay zero
  const int ms = km.maxSpecInReac;
  km.reacSidx.assign(size_t(nR) * ms, 0);
  km.reacNuki.assign(size_t(nR) * ms, 0.0);
  for (int r = 0; r < nR; ++r) {
    int count = 0;
    for (const auto& e : reactantsInfo[r]) {
      km.reacSidx[size_t(r) * ms + count] = e.first;
      km.reacNuki[size_t(r) * ms + count] = -e.second;
      ++count;
    }
    count = ms / 2;
    for (const auto& e : productsInfo[r]) {
      km.reacSidx[size_t(r) * ms + count] = e.first;
      km.reacNuki[size_t(r) * ms + count] = e.second;
      ++count;
    }
  }

  // reactions, pass 2: per-type parameters (.cpp:391-783)
  for (int r = 0; r < nR; ++r) {
    const Node& reaction = reactions.items[r];
    const Node* c = reaction.find("coefficients");
    if (c && !c->isMap()) c = nullptr;  // "coefficients:" with no entries
    const bool perMin = timeUnitIsMinutes(reaction);
    switch (km.reactionType[r]) {
      case kArrhenius: {
        bool hasA = false;
        double A = coef(c, "A", 1.0, &hasA);
        if (hasA && perMin) A /= 60.0;
        const double D = coef(c, "D", 300.0);
        const double B = coef(c, "B", 0.0);
        bool hasEa = false, hasC = false;
        const double Ea = coef(c, "Ea", 0.0, &hasEa);
        const double Cc = coef(c, "C", 0.0, &hasC);
        double C = 0.0;
        if (hasEa) {
          C = -Ea / kBoltzmann;
          if (hasC) fail("Ea and C are both present for reaction No " + std::to_string(r));
        } else if (hasC) {
          C = Cc;
        }
        const double E = coef(c, "E", 0.0);
        km.arrhenius.push_back(ArrheniusParam{r, A, B, C, D, E});
        break;
      }
      case kCmaqH2O2: {
        bool has = false;
        double k1A = coef(c, "k1_A", 1.0, &has);
        if (has && perMin) k1A /= 60.0;
        double k2A = coef(c, "k2_A", 1.0, &has);
        if (has && perMin) k2A /= 60.0;
        km.cmaqH2O2.push_back(CmaqH2O2Param{r, k1A, coef(c, "k1_B", 0.0), coef(c, "k1_C", 0.0), k2A * 1e6,
                                            coef(c, "k2_B", 0.0), coef(c, "k2_C", 0.0)});
        break;
      }
      case kJpl:
      case kRJplArrhenius: {
        bool has = false;
        double k0A = coef(c, "k0_A", 1.0, &has);
        if (has && perMin) k0A /= 60.0;
        double kinfA = coef(c, "kinf_A", 1.0, &has);
        if (has && perMin) kinfA /= 60.0;
        const double k0B = coef(c, "k0_B", 0.0), k0C = coef(c, "k0_C", 0.0);
        const double kinfB = coef(c, "kinf_B", 0.0), kinfC = coef(c, "kinf_C", 0.0);
        const double Fc = coef(c, "Fc", 0.6);
        if (km.reactionType[r] == kJpl) {
          km.jpl.push_back(JplParam{r, k0A, k0B, k0C, kinfA, kinfB, kinfC, Fc});
        } else {
          double A = coef(c, "A", 1.0, &has);
          if (has && perMin) A /= 60.0;
          km.rJplArrhenius.push_back(
              RJplArrheniusParam{r, k0A, k0B, k0C, kinfA, kinfB, kinfC, Fc, A, coef(c, "B", 0.0), coef(c, "C", 0.0)});
        }
        break;
      }
      case kTroe: {
        bool has = false;
        double A = coef(c, "k0_A", 1.0, &has);
        if (has && perMin) A /= 60.0;
        double kinfA = coef(c, "kinf_A", 1.0, &has);
        if (has && perMin) kinfA /= 60.0;
        km.reacPfal.push_back(r);
        // k0_A carries [M] in CAMP's convention: k0_A * 1e6 (.cpp:743-746)
        km.reacPpar.push_back({A * 1e6, coef(c, "k0_B", 0.0), coef(c, "k0_C", 0.0), kinfA, coef(c, "kinf_B", 0.0),
                               coef(c, "kinf_C", 0.0), coef(c, "Fc", 0.6), coef(c, "N", 1.0)});
        break;
      }
      default:
        break;
    }
    if (const Node* adj = reaction.find("adjust_reaction")) {
      for (const Node& s : adj->items) {
        auto it = km.speciesIndex.find(s.str());
        if (it == km.speciesIndex.end())
          fail("species does not exist " + s.str() + " in adjust_reaction for reaction No " + std::to_string(r));
        km.adjustReaction.push_back(AdjustReactionParam{r, it->second});
      }
    }
  }

  // modifier_prod_O1D (.cpp:785-889)
  if (const Node* mod = root.find("modifier_prod_O1D")) {
    ProdO1DParam p{};
    const Node* c = mod->find("coefficients");
    p.A1 = coef(c, "A1", 1.0);
    p.C1 = coef(c, "C1", 0.0);
    p.A2 = coef(c, "A2", 1.0);
    p.C2 = coef(c, "C2", 0.0);
    p.A3 = coef(c, "A3", 1.0);
    p.C3 = coef(c, "C3", 0.0);
    auto spIdx = [&](const char* key) {
      const std::string n = (*mod)[key].str();
      auto it = km.speciesIndex.find(n);
      if (it == km.speciesIndex.end()) fail("species does not exist " + n + " in modifier_prod_O1D");
      return it->second;
    };
    p.species1 = spIdx("species_name_1");
    p.species2 = spIdx("species_name_2");
    p.species3 = spIdx("species_name_3");
    p.photolysis_reaction = (*mod)["photolysis_reaction_index"].asInt();
    int count = 0;
    for (const Node& r : (*mod)["reaction_list"].items) {
      if (count >= 3) fail("modifier_prod_O1D supports three reactions");
      p.reactions[count++] = r.asInt();
    }
    km.prodO1D.push_back(p);
  }

  // emission sources (.cpp:891-937)
  if (const Node* sources = root.find("sources")) {
    for (const Node& src : sources->items) {
      const std::string type = src["type"].str();
      if (type != "EMISSION") fail("source type does not exist: " + type);
    }
    for (const Node& src : sources->items) {
      const Node* rate = src.find("coefficients") ? src["coefficients"].find("emission_rate") : nullptr;
      if (!rate) fail("EMISSION, emission_rate is not provided");
      const Node* sp = src.find("species");
      if (!sp) fail("EMISSION, species name is not provided");
      auto it = km.speciesIndex.find(sp->str());
      if (it == km.speciesIndex.end()) fail("species does not exist " + sp->str());
      km.emission.push_back(EmissionParam{it->second, rate->asDouble()});
    }
  }

  // convExponent = sum |int(nu_reactant)| - 1 (.cpp:939-947)
  for (int r = 0; r < nR; ++r) {
    int nu_sum = 0;
    for (int j = 0; j < km.reacNreac[r]; ++j) nu_sum += std::abs(int(km.nuki(r, j)));
    km.convExponent[r] = nu_sum - 1;
  }

  echoPrint(echo, "kmod.list :  # sources e.g., EMISSION                         : %d\n", int(km.emission.size()));
  echoPrint(echo, "Reaction data : species and Arrhenius pars\n");
  for (int r = 0; r < nR; r++) {
    echoPrint(echo, "%-5d\t%1d\t%2d\t%2d | ", r + 1, km.isRev[r], km.reacNreac[r], km.reacNprod[r]);
    for (int j = 0; j < km.reacNreac[r]; j++)
      echoPrint(echo, "%f*%s | ", km.nuki(r, j), km.speciesNames[km.sidx(r, j)].c_str());
    const int joff = ms / 2;
    for (int j = 0; j < km.reacNprod[r]; j++)
      echoPrint(echo, "%f*%s | ", km.nuki(r, j + joff), km.speciesNames[km.sidx(r, j + joff)].c_str());
    echoPrint(echo, "\n");
  }

  // photolysis reactions (CB05 convention)
  for (int r = 0; r < nR; ++r)
    if (km.reactionType[r] == kArrhenius && km.note[r].rfind("PHOTOLYSIS", 0) == 0) km.photolysisReactions.push_back(r);

  return km;
}

std::vector<double> loadScenarioState(const std::string& path, const KineticModel& km, int* nBatchOut) {
  const Node root = yaml::parseFile(path);
  const Node& pressure = root["environmental_conditions"]["pressure"]["initial_value"];
  const Node& temperature = root["environmental_conditions"]["temperature"]["initial_value"];
  const Node& initial = root["initial_state"];
  const int nBatch = int(pressure.size());
  if (int(temperature.size()) != nBatch)
    throw std::runtime_error("scenario: pressure and temperature lists differ in length");
  const int dim = km.nSpec + 3;
  std::vector<double> state(size_t(nBatch) * dim, 0.0);

  // species index by name (first of the 18-character-truncated names)
  std::vector<int> idx;
  for (const auto& e : initial.entries) {
    int found = -1;
    for (int i = 0; i < km.nSpec; ++i)
      if (std::strncmp(km.speciesNames[i].c_str(), e.first.c_str(), kLengthOfSpecName) == 0) {
        found = i;
        break;
      }
    if (found < 0) throw std::runtime_error("scenario: species " + e.first + " is not in the mechanism");
    idx.push_back(found);
  }
  for (int i = 0; i < nBatch; ++i) {
    const double p = pressure.items[i].asDouble();
    const double T = temperature.items[i].asDouble();
    state[size_t(i) * dim + 1] = p;
    state[size_t(i) * dim + 2] = T;
    int j = 0;
    for (const auto& e : initial.entries) {
      const std::string units = e.second["units"].str();
      const Node& vals = e.second["initial_value"];
      if (int(vals.size()) <= i) throw std::runtime_error("scenario: too few initial values for " + e.first);
      const double v = vals.items[i].asDouble();
      double out;
      const double conv = kConvPPM * p / T;
      if (units == "molecules m-3") out = v / conv;
      else if (units == "molecules cm-3") out = v / conv / 1e6;
      else if (units == "mol m-3") out = v;
      else throw std::runtime_error("unit type does not exist " + units);
      state[size_t(i) * dim + idx[j] + 3] = out;
      ++j;
    }
  }
  *nBatchOut = nBatch;
  return state;
}

std::vector<double> loadScenarioPhotolysisRates(const std::string& path, int nBatch, int* countOut) {
  const Node root = yaml::parseFile(path);
  const Node* pr = root.find("photolysis_rates");
  int count = pr ? int(pr->size()) : 0;
  std::vector<double> out(size_t(nBatch) * size_t(count), 0.0);
  for (int j = 0; j < count; ++j) {
    const Node& vals = (*pr)["r_" + std::to_string(j)]["value"];
    if (int(vals.size()) != nBatch)
      throw std::runtime_error("scenario: number of values of photolysis rate r_" + std::to_string(j) +
                               " differs from the number of conditions");
    for (int i = 0; i < nBatch; ++i) out[size_t(i) * count + j] = vals.items[i].asDouble();
  }
  *countOut = count;
  return out;
}

std::vector<double> loadScenarioExternalForcing(const std::string& path, const KineticModel& km, int nBatch,
                                                int* countOut) {
  const Node root = yaml::parseFile(path);
  const int m = km.nActive();
  std::vector<double> out(size_t(nBatch) * size_t(m), 0.0);
  int count = 0;
  if (const Node* ef = root.find("external_forcing")) {
    for (const auto& e : ef->entries) {
      const std::string name = e.second["species_name"].str();
      int idx = -1;
      for (int i = 0; i < m; ++i)  // active species are listed first
        if (std::strncmp(km.speciesNames[i].c_str(), name.c_str(), kLengthOfSpecName) == 0) {
          idx = i;
          break;
        }
      if (idx < 0) throw std::runtime_error("scenario: species " + name + " does not exist in the species list");
      const Node& vals = e.second["value"];
      if (int(vals.size()) != nBatch)
        throw std::runtime_error("scenario: number of values of external forcing differs from the number of conditions");
      for (int i = 0; i < nBatch; ++i) out[size_t(i) * m + idx] = vals.items[i].asDouble();
      ++count;
    }
  }
  *countOut = count;
  return out;
}

}  // namespace tchem_amd
