// TChem_Driver coupling interface (include/tchem_driver.h) on the C ABI of
// libtchem_atm.so: one process-wide aerosol box model whose host state
// vectors a coupled model reads and writes, integrated on the GPU by
// tchem_aerosolchem_run_device_batch.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/tchem_atm.h"
#include "../../include/tchem_driver.h"
#include "yaml_lite.hpp"

namespace {

std::string g_error;

void fail(const std::string& msg) {
  g_error = msg;
  std::fprintf(stderr, "TChem_Driver: %s\n", msg.c_str());
}

struct DeviceBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DeviceBuf() {
    if (p) (void)hipFree(p);
  }
  bool reserve(size_t b) {
    if (b <= bytes) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (hipMalloc(&p, b) != hipSuccess) return false;
    bytes = b;
    return true;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Solver settings of the numerics file (TChem_Driver.cpp:155-183): keys and
// defaults as the reference reads them.
struct Numerics {
  double atol_newton = 1e-10, rtol_newton = 1e-6, dtmin = 1e-8, atol_time = 1e-12, rtol_time = 1e-4;
  int max_newton = 100, max_time_iterations = 1000;
  bool verbose = false;
};

struct Driver {
  tchem_kmd_t kmd = nullptr;
  tchem_kmcd_t kmcd = nullptr;
  tchem_amd_t amd = nullptr;
  tchem_amcd_t amcd = nullptr;
  int nBatch = 0, nSpec = 0, nConst = 0, nAero = 0, nPart = 0, len = 0, mt = 0;
  std::vector<std::string> gasNames, aeroNames;
  std::vector<double> density, mw, kappa;
  std::vector<double> state;  // [nBatch][len]
  std::vector<double> nc;     // [nBatch][nPart]
  Numerics num;
  DeviceBuf dState, dNc, dTadv, dT, dDt, dTolN, dTolT, dStats;
  ~Driver() {
    if (amcd) tchem_amcd_destroy(amcd);
    if (amd) tchem_amd_destroy(amd);
    if (kmcd) tchem_kmcd_destroy(kmcd);
    if (kmd) tchem_kmd_destroy(kmd);
  }
};

std::unique_ptr<Driver> g;

Numerics readNumerics(const char* path) {
  Numerics n;
  const tchem_amd::yaml::Node root = tchem_amd::yaml::parseFile(path);
  const tchem_amd::yaml::Node* s = root.find("solver_info");
  if (!s) return n;
  auto num = [&](const char* key, double& v) {
    if (const tchem_amd::yaml::Node* x = s->find(key)) v = x->asDouble();
  };
  auto integer = [&](const char* key, int& v) {
    if (const tchem_amd::yaml::Node* x = s->find(key)) v = int(x->asDouble());
  };
  num("atol_newton", n.atol_newton);
  num("rtol_newton", n.rtol_newton);
  num("dtmin", n.dtmin);
  num("atol_time", n.atol_time);
  num("rtol_time", n.rtol_time);
  integer("max_newton_iterations", n.max_newton);
  integer("max_num_time_iterations", n.max_time_iterations);
  if (const tchem_amd::yaml::Node* x = s->find("verbose")) n.verbose = x->str() == "true" || x->str() == "True";
  return n;
}

bool ok(int rc, const char* what) {
  if (rc == TCHEM_OK) return true;
  fail(std::string(what) + ": " + tchem_last_error());
  return false;
}

int copyName(const std::string& s, char* result, size_t buffer_size) {
  if (!result || buffer_size == 0) return -1;
  const size_t n = s.size() < buffer_size - 1 ? s.size() : buffer_size - 1;
  std::memcpy(result, s.data(), n);
  result[n] = '\0';
  return int(s.size());
}

}  // namespace

extern "C" {

void initialize(const char* gasFile, const char* aeroFile, const char* numericsFile, const int nBatch) {
  g_error.clear();
  g.reset();
  if (!gasFile || !aeroFile || !numericsFile || nBatch < 1) {
    fail("initialize: bad argument");
    return;
  }
  std::unique_ptr<Driver> d(new Driver);
  d->nBatch = nBatch;
  if (!ok(tchem_kmd_load_yaml(gasFile, nullptr, &d->kmd), "gas model")) return;
  int nReac = 0, nPhoto = 0;
  if (!ok(tchem_kmd_info(d->kmd, &d->nSpec, &d->nConst, &nReac, &nPhoto), "gas model")) return;
  char buf[256];
  for (int i = 0; i < d->nSpec; ++i) {
    if (!ok(tchem_kmd_species_name(d->kmd, i, buf, sizeof buf), "gas model")) return;
    d->gasNames.push_back(buf);
  }
  if (!ok(tchem_kmcd_create(d->kmd, &d->kmcd), "gas model constant data")) return;
  if (!ok(tchem_amd_load_yaml(aeroFile, d->kmd, &d->amd), "aerosol model")) return;
  int nSimpol = 0;
  if (!ok(tchem_amd_info(d->amd, &d->nAero, &d->nPart, &nSimpol), "aerosol model")) return;
  for (int i = 0; i < d->nAero; ++i) {
    double r = 0, w = 0, k = 0;
    if (!ok(tchem_amd_species_name(d->amd, i, buf, sizeof buf), "aerosol model") ||
        !ok(tchem_amd_species_props(d->amd, i, &r, &w, &k), "aerosol model"))
      return;
    d->aeroNames.push_back(buf);
    d->density.push_back(r);
    d->mw.push_back(w);
    d->kappa.push_back(k);
  }
  if (!ok(tchem_amcd_create(d->amd, d->kmcd, &d->amcd), "aerosol model constant data")) return;
  try {
    d->num = readNumerics(numericsFile);
  } catch (const std::exception& e) {
    fail(std::string("numerics file: ") + e.what());
    return;
  }
  d->len = 3 + d->nSpec + d->nAero * d->nPart;
  d->mt = d->nSpec - d->nConst + d->nAero * d->nPart;
  d->state.assign(size_t(nBatch) * d->len, 0.0);
  d->nc.assign(size_t(nBatch) * (d->nPart > 0 ? d->nPart : 1), 0.0);
  g = std::move(d);
}

void finalize(void) { g.reset(); }

int TChem_getNumberOfSpecies(void) { return g ? g->nSpec : -1; }
int TChem_getLengthOfStateVector(void) { return g ? g->len : -1; }
int TChem_getStateVectorSize(void) { return TChem_getLengthOfStateVector(); }
int TChem_getNumberConcentrationVectorSize(void) { return g ? g->nPart : -1; }
int TChem_getNumberOfAeroSpecies(void) { return g ? g->nAero : -1; }
const char* TChem_getLastError(void) { return g_error.c_str(); }

void TChem_getStateVector(double* array, const int iBatch) {
  if (!g || !array || iBatch < 0 || iBatch >= g->nBatch) return fail("TChem_getStateVector: bad argument");
  std::memcpy(array, g->state.data() + size_t(iBatch) * g->len, sizeof(double) * g->len);
}

void TChem_setStateVector(double* array, const int iBatch) {
  if (!g || !array || iBatch < 0 || iBatch >= g->nBatch) return fail("TChem_setStateVector: bad argument");
  std::memcpy(g->state.data() + size_t(iBatch) * g->len, array, sizeof(double) * g->len);
}

void TChem_getAllStateVectorHost(double* view) {
  if (!g || !view) return fail("TChem_getAllStateVectorHost: bad argument");
  std::memcpy(view, g->state.data(), sizeof(double) * g->state.size());
}

void TChem_setNumberConcentrationVector(double* array, const int iBatch) {
  if (!g || !array || iBatch < 0 || iBatch >= g->nBatch)
    return fail("TChem_setNumberConcentrationVector: bad argument");
  std::memcpy(g->nc.data() + size_t(iBatch) * g->nPart, array, sizeof(double) * g->nPart);
}

int TChem_getSpeciesName(int* index, char* result, const size_t buffer_size) {
  if (!g || !index || *index < 0 || *index >= g->nSpec) return -1;
  return copyName(g->gasNames[*index], result, buffer_size);
}

int TChem_getAerosolSpeciesName(int* index, char* result, const size_t buffer_size) {
  if (!g || !index || *index < 0 || *index >= g->nAero) return -1;
  return copyName(g->aeroNames[*index], result, buffer_size);
}

double TChem_getAerosolSpeciesDensity(int* index) {
  return (g && index && *index >= 0 && *index < g->nAero) ? g->density[*index] : -1.0;
}
double TChem_getAerosolSpeciesMW(int* index) {
  return (g && index && *index >= 0 && *index < g->nAero) ? g->mw[*index] : -1.0;
}
double TChem_getAerosolSpeciesKappa(int* index) {
  return (g && index && *index >= 0 && *index < g->nAero) ? g->kappa[*index] : -1.0;
}

// Every batch entry from 0 to del_t: intervals of max_num_time_iterations
// TR-BDF2 steps (dt_in = dtmin, dtmax = del_t) until every entry reaches
// del_t, at most max_num_time_iterations intervals (the reference's output
// loop bound, TChem_Driver.cpp:596-614); species of every entry copied back.
void TChem_doTimestep(const double& del_t) {
  if (!g) return fail("TChem_doTimestep: not initialized");
  Driver& d = *g;
  if (!(del_t > 0.0)) return;
  const int n = d.nBatch;
  const size_t stBytes = sizeof(double) * d.state.size();
  if (!d.dState.reserve(stBytes) || !d.dNc.reserve(sizeof(double) * d.nc.size()) ||
      !d.dTadv.reserve(sizeof(tchem_time_advance_t) * n) || !d.dT.reserve(sizeof(double) * n) ||
      !d.dDt.reserve(sizeof(double) * n) || !d.dTolN.reserve(2 * sizeof(double)) ||
      !d.dTolT.reserve(2 * sizeof(double) * d.mt) || !d.dStats.reserve(sizeof(tchem_cell_stats_t) * n))
    return fail("TChem_doTimestep: device allocation failed");
  const int steps = d.num.max_time_iterations > 0 ? d.num.max_time_iterations : 1000;
  std::vector<tchem_time_advance_t> tadv(n);
  for (auto& ta : tadv) {
    ta = tchem_time_advance_t{};
    ta._tbeg = 0.0;
    ta._tend = del_t;
    ta._dt = d.num.dtmin;
    ta._dtmin = d.num.dtmin;
    ta._dtmax = del_t;
    ta._max_num_newton_iterations = d.num.max_newton;
    ta._num_time_iterations_per_interval = steps;
    ta._jacobian_interval = 1;
  }
  const double tolN[2] = {d.num.atol_newton, d.num.rtol_newton};
  std::vector<double> tolT(2 * size_t(d.mt));
  for (int i = 0; i < d.mt; ++i) {
    tolT[2 * i] = d.num.atol_time;
    tolT[2 * i + 1] = d.num.rtol_time;
  }
  std::vector<double> zeros(n, 0.0), t(n, 0.0);
  hipError_t e = hipSuccess;
  if ((e = hipMemcpy(d.dState.p, d.state.data(), stBytes, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d.dNc.p, d.nc.data(), sizeof(double) * d.nc.size(), hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d.dTadv.p, tadv.data(), sizeof(tchem_time_advance_t) * n, hipMemcpyHostToDevice)) !=
          hipSuccess ||
      (e = hipMemcpy(d.dT.p, zeros.data(), sizeof(double) * n, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d.dDt.p, zeros.data(), sizeof(double) * n, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d.dTolN.p, tolN, sizeof tolN, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d.dTolT.p, tolT.data(), sizeof(double) * tolT.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(std::string("TChem_doTimestep: ") + hipGetErrorString(e));
  std::vector<tchem_cell_stats_t> st(n), total(n);
  for (int it = 0; it < steps; ++it) {
    if (!ok(tchem_aerosolchem_run_device_batch(nullptr, n, d.dTolN.as<double>(), d.dTolT.as<double>(), nullptr,
                                               d.dTadv.as<tchem_time_advance_t>(), d.dState.as<double>(),
                                               d.dNc.as<double>(), d.dT.as<double>(), d.dDt.as<double>(),
                                               d.dState.as<double>(), d.dStats.as<tchem_cell_stats_t>(), d.amcd,
                                               nullptr),
            "TChem_doTimestep") ||
        !ok(tchem_atmchem_update_time_advance(n, d.dTadv.as<tchem_time_advance_t>(), d.dT.as<double>(),
                                              d.dDt.as<double>(), nullptr),
            "TChem_doTimestep"))
      return;
    if ((e = hipMemcpy(t.data(), d.dT.p, sizeof(double) * n, hipMemcpyDeviceToHost)) != hipSuccess ||
        (e = hipMemcpy(st.data(), d.dStats.p, sizeof(tchem_cell_stats_t) * n, hipMemcpyDeviceToHost)) != hipSuccess)
      return fail(std::string("TChem_doTimestep: ") + hipGetErrorString(e));
    bool done = true;
    for (int c = 0; c < n; ++c) {
      total[c].accepted += st[c].accepted;
      total[c].rejected += st[c].rejected;
      total[c].newton_iterations += st[c].newton_iterations;
      total[c].failed |= st[c].failed;
      done = done && t[c] >= del_t;
    }
    if (done) break;
  }
  std::vector<double> out(d.state.size());
  if ((e = hipMemcpy(out.data(), d.dState.p, stBytes, hipMemcpyDeviceToHost)) != hipSuccess)
    return fail(std::string("TChem_doTimestep: ") + hipGetErrorString(e));
  const int mg = d.nSpec - d.nConst;
  for (int c = 0; c < n; ++c) {
    double* row = d.state.data() + size_t(c) * d.len;
    const double* src = out.data() + size_t(c) * d.len;
    for (int j = 0; j < mg; ++j) row[3 + j] = src[3 + j];
    for (int j = 3 + d.nSpec; j < d.len; ++j) row[j] = src[j];
    if (total[c].failed || t[c] < del_t)
      std::fprintf(stderr, "TChem_Driver: batch entry %d did not reach %g s (t = %g)\n", c, del_t, t[c]);
  }
  if (d.num.verbose)
    for (int c = 0; c < n; ++c)
      std::printf("TChem_Driver: entry %d: %d steps, %d rejected, %d Newton iterations\n", c, total[c].accepted,
                  total[c].rejected, total[c].newton_iterations);
}

}  // extern "C"
