// Box-model driver with the command line and output files of the
// reference's TChem_AtmosphericChemistry.x
// (src/examples/TChem_AtmosphericChemistry.cpp:46-352) and, built with
// TCHEM_E3SM_DRIVER, of TChem_AtmosphericChemistryE3SM.x
// (src/examples/TChem_AtmosphericChemistryE3SM.cpp: E3SM kinetics, scenario
// from --inputfile with per-cell photolysis rates and external forcing,
// --atol-time) and, built with TCHEM_AEROSOL_DRIVER, of
// TChem_AerosolChemistry.x (src/examples/TChem_AerosolChemistry.cpp:46-440:
// --aerofile single-particle SIMPOL.1 aerosol model, particle scenario from
// --inputfile_particles, --number_of_particles, FP64): load the mechanism
// (writing kmod.echo), read the scenario, relaunch the batched integrator
// interval by interval until the mean cell time exceeds 0.9999 tend, write
// the state of every cell after every interval (iter, t, dt, state row) and
// the wall time of every interval (wall_times.json).  All compute goes
// through the C ABI of libtchem_atm.so.
//
// Options are given as --name=value (booleans: true/false/1/0), as in the
// reference's CommandLineParser (src/core/TChem_CommandLineParser.hpp).
// team_thread_size / vector_thread_size / inputsPath are accepted and
// ignored.  Additional options of this driver: --precision=fp64|mixed,
// --lu=auto|dense|sparse, --schedule=grid|persistent.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/tchem_atm.h"

namespace {

#ifdef TCHEM_E3SM_DRIVER
constexpr bool kE3sm = true;
#else
constexpr bool kE3sm = false;
#endif
#ifdef TCHEM_AEROSOL_DRIVER
constexpr bool kAero = true;
#else
constexpr bool kAero = false;
#endif

struct Options {
  std::string chemfile = "chem.yaml", inputfile = "None", times = "wall_times.json";
  std::string aerofile = "aero.yaml", particles = "None";
  int nParticles = -1;
  std::string outputfile = kE3sm ? "AtmosphericChemistryE3SM.dat"
                                 : (kAero ? "AerosolChemistry.dat" : "AtmosphericChemistry.dat");
  double tbeg = 0, tend = 1, dtmin = 1e-8, dtmax = 1e-1, rtol_time = 1e-4, atol_newton = 1e-10, rtol_newton = 1e-6;
  double atol_time = 1e-12;
  int per_interval = 10, max_iters = 1000, max_newton = 100, jac_interval = 1, batch = 1;
  bool verbose = false, write_profiles = true, cloned = false;
  int precision = TCHEM_PRECISION_FP64, lu = TCHEM_LU_AUTO, schedule = TCHEM_SCHEDULE_GRID;
};

bool parseBool(const std::string& v) { return v == "true" || v == "1" || v == "True" || v == "TRUE"; }

void usage(const char* argv0) {
  std::printf(
      "Usage: %s [options]\n  options (--name=value):\n"
      "  --chemfile --outputfile --outputfile_times --inputsPath (E3SM: --inputfile --atol-time)\n"
      "  (aerosol: --aerofile --inputfile --inputfile_particles --number_of_particles --atol-time)\n"
      "  --tbeg --tend --dtmin --dtmax --tol-time --atol-newton --rtol-newton\n"
      "  --time-iterations-per-interval --max-time-iterations --max-newton-iterations --jacobian-interval\n"
      "  --batch_size --use_cloned_samples --write-time-profiles --verbose\n"
      "  --team_thread_size --vector_thread_size (ignored)\n"
      "  --precision=fp64|mixed --lu=auto|dense|sparse --schedule=grid|persistent\n",
      argv0);
}

// returns 0 to continue, 1 when help was printed, -1 on a bad option
int parse(int argc, char** argv, Options& o) {
  for (int i = 1; i < argc; ++i) {
    const std::string s = argv[i];
    if (s == "--help") {
      usage(argv[0]);
      return 1;
    }
    if (s.rfind("--", 0) != 0) continue;
    const size_t eq = s.find('=');
    const std::string key = s.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
    const std::string val = eq == std::string::npos ? "true" : s.substr(eq + 1);
    try {
      if (key == "chemfile") o.chemfile = val;
      else if ((kE3sm || kAero) && key == "inputfile") o.inputfile = val;
      else if ((kE3sm || kAero) && key == "atol-time") o.atol_time = std::stod(val);
      else if (kAero && key == "aerofile") o.aerofile = val;
      else if (kAero && key == "inputfile_particles") o.particles = val;
      else if (kAero && key == "number_of_particles") o.nParticles = std::stoi(val);
      else if (kAero && key == "use-cvode") ;
      else if (key == "outputfile") o.outputfile = val;
      else if (key == "outputfile_times") o.times = val;
      else if (key == "tbeg") o.tbeg = std::stod(val);
      else if (key == "tend") o.tend = std::stod(val);
      else if (key == "dtmin") o.dtmin = std::stod(val);
      else if (key == "dtmax") o.dtmax = std::stod(val);
      else if (key == "tol-time") o.rtol_time = std::stod(val);
      else if (key == "atol-newton") o.atol_newton = std::stod(val);
      else if (key == "rtol-newton") o.rtol_newton = std::stod(val);
      else if (key == "time-iterations-per-interval") o.per_interval = std::stoi(val);
      else if (key == "max-time-iterations") o.max_iters = std::stoi(val);
      else if (key == "max-newton-iterations") o.max_newton = std::stoi(val);
      else if (key == "jacobian-interval") o.jac_interval = std::stoi(val);
      else if (key == "batch_size") o.batch = std::stoi(val);
      else if (key == "use_cloned_samples") o.cloned = parseBool(val);
      else if (key == "write-time-profiles") o.write_profiles = parseBool(val);
      else if (key == "verbose") o.verbose = parseBool(val);
      else if (key == "precision") o.precision = val == "mixed" ? TCHEM_PRECISION_MIXED : TCHEM_PRECISION_FP64;
      else if (key == "lu") o.lu = val == "dense" ? TCHEM_LU_DENSE : (val == "sparse" ? TCHEM_LU_SPARSE : TCHEM_LU_AUTO);
      else if (key == "schedule")
        o.schedule = val == "persistent" ? TCHEM_SCHEDULE_PERSISTENT : TCHEM_SCHEDULE_GRID;
      else if (key == "team_thread_size" || key == "vector_thread_size" || key == "inputsPath" ||
               key == "echo-command-line")
        ;
      else {
        std::fprintf(stderr, "unknown option --%s\n", key.c_str());
        return -1;
      }
    } catch (const std::exception&) {
      std::fprintf(stderr, "bad value for --%s: %s\n", key.c_str(), val.c_str());
      return -1;
    }
  }
  return 0;
}

#define CHECK_T(call)                                                                  \
  do {                                                                                 \
    const int rc_ = (call);                                                            \
    if (rc_) {                                                                         \
      std::fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, tchem_last_error());    \
      return 1;                                                                        \
    }                                                                                  \
  } while (0)
#define CHECK_H(call)                                                                  \
  do {                                                                                 \
    const hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                            \
      std::fprintf(stderr, "%s failed: %s\n", #call, hipGetErrorString(e_));           \
      return 1;                                                                        \
    }                                                                                  \
  } while (0)

template <typename T>
struct DeviceArray {
  T* p = nullptr;
  ~DeviceArray() {
    if (p) (void)hipFree(p);
  }
};

void writeState(FILE* f, int iter, const std::vector<double>& t, const std::vector<double>& dt,
                const std::vector<double>& state, int n, int dim) {
  for (int i = 0; i < n; ++i) {
    std::fprintf(f, "%d \t %15.10e \t  %15.10e \t ", iter, t[i], dt[i]);
    for (int k = 0; k < dim; ++k) std::fprintf(f, "%15.10e \t", state[size_t(i) * dim + k]);
    std::fprintf(f, "\n");
  }
}

}  // namespace

int main(int argc, char** argv) {
  Options o;
  const int pr = parse(argc, argv, o);
  if (pr) return pr > 0 ? 0 : 2;

  tchem_kmd_t kmd = nullptr;
  CHECK_T(tchem_kmd_load_yaml(o.chemfile.c_str(), "kmod.echo", &kmd));
  int nSpec = 0, nConst = 0, nReac = 0, nPhoto = 0;
  CHECK_T(tchem_kmd_info(kmd, &nSpec, &nConst, &nReac, &nPhoto));
  std::printf("Number of Species %d \n", nSpec);
  std::printf("Number of Reactions %d \n", nReac);
  const int gasDim = nSpec + 3;
  std::vector<std::string> names(nSpec);
  for (int k = 0; k < nSpec; ++k) {
    char buf[64];
    CHECK_T(tchem_kmd_species_name(kmd, k, buf, sizeof(buf)));
    names[k] = buf;
  }
  // aerosol model: particle species appended to the state rows and unknowns
  tchem_amd_t amd = nullptr;
  int nAero = 0, nPart = 0;
  if (kAero) {
    CHECK_T(tchem_amd_load_yaml(o.aerofile.c_str(), kmd, &amd));
    if (o.nParticles > 0) CHECK_T(tchem_amd_set_number_of_particles(amd, o.nParticles));
    int nSimpol = 0;
    CHECK_T(tchem_amd_info(amd, &nAero, &nPart, &nSimpol));
    std::printf("Number of Aerosol Species %d \n", nAero);
    std::printf("Number of Aerosol Particles %d \n", nPart);
    for (int ip = 0; ip < nPart; ++ip)
      for (int k = 0; k < nAero; ++k) {
        char buf[128];
        CHECK_T(tchem_amd_species_name(amd, k, buf, sizeof(buf)));
        names.push_back(std::string(buf) + "_p" + std::to_string(ip));
      }
  }
  const int dim = gasDim + nAero * nPart, m = nSpec - nConst + nAero * nPart;

  const std::string input = ((kE3sm || kAero) && o.inputfile != "None") ? o.inputfile : o.chemfile;
  int nFiles = 0;
  CHECK_T(tchem_scenario_load(kmd, input.c_str(), nullptr, 0, &nFiles));
  std::vector<double> scen(size_t(nFiles) * dim);
  {
    std::vector<double> gas(size_t(nFiles) * gasDim);
    CHECK_T(tchem_scenario_load(kmd, input.c_str(), gas.data(), nFiles, &nFiles));
    for (int i = 0; i < nFiles; ++i)
      std::copy(gas.begin() + size_t(i) * gasDim, gas.begin() + size_t(i + 1) * gasDim, scen.begin() + size_t(i) * dim);
  }
  std::vector<double> ncScen(size_t(nFiles) * (nPart > 0 ? nPart : 1), 0.0);
  if (kAero && o.particles != "None") {
    int found = 0;
    CHECK_T(tchem_scenario_particles(amd, o.particles.c_str(), nFiles, dim, scen.data(), ncScen.data(), &found));
  }
  // E3SM: per-cell photolysis rates and external forcing of the scenario
  int nPhotoCols = 0;
  std::vector<double> photoScen, extScen;
  if (kE3sm) {
    CHECK_T(tchem_scenario_photolysis_rates(kmd, input.c_str(), nFiles, nullptr, 0, &nPhotoCols));
    photoScen.resize(size_t(nFiles) * nPhotoCols);
    if (nPhotoCols)
      CHECK_T(tchem_scenario_photolysis_rates(kmd, input.c_str(), nFiles, photoScen.data(), nPhotoCols, &nPhotoCols));
    extScen.resize(size_t(nFiles) * m);
    int nExt = 0;
    CHECK_T(tchem_scenario_external_forcing(kmd, input.c_str(), nFiles, extScen.data(), &nExt));
  }
  int n = nFiles;
  std::vector<double> state, photo, ext, nc;
  auto clone = [](const std::vector<double>& row0, int width, int count) {
    std::vector<double> v(size_t(count) * width);
    for (int i = 0; i < count; ++i) std::copy(row0.begin(), row0.begin() + width, v.begin() + size_t(i) * width);
    return v;
  };
  if (nFiles == 1 && o.cloned && o.batch > 1) {
    std::printf("Using cloned samples ... only for numerical experiments\n");
    n = o.batch;
    state = clone(scen, dim, n);
    if (kAero) nc = clone(ncScen, nPart > 0 ? nPart : 1, n);
    if (kE3sm) {
      photo = clone(photoScen, nPhotoCols, n);
      ext = clone(extScen, m, n);
    }
  } else {
    state = scen;
    photo = photoScen;
    ext = extScen;
    nc = ncScen;
  }

  tchem_kmcd_t kmcd = nullptr;
  CHECK_T(tchem_kmcd_create(kmd, &kmcd));
  tchem_amcd_t amcd = nullptr;
  if (kAero) CHECK_T(tchem_amcd_create(amd, kmcd, &amcd));

  DeviceArray<double> d_state, d_t, d_dt, d_tn, d_tt, d_photo, d_ext, d_nc;
  DeviceArray<tchem_time_advance_t> d_tadv;
  CHECK_H(hipMalloc(&d_state.p, sizeof(double) * state.size()));
  CHECK_H(hipMalloc(&d_t.p, sizeof(double) * n));
  CHECK_H(hipMalloc(&d_dt.p, sizeof(double) * n));
  CHECK_H(hipMalloc(&d_tn.p, sizeof(double) * 2));
  CHECK_H(hipMalloc(&d_tt.p, sizeof(double) * 2 * (m > 0 ? m : 1)));
  CHECK_H(hipMalloc(&d_tadv.p, sizeof(tchem_time_advance_t) * n));
  // the gas driver fills dt with dtmax (examples/TChem_AtmosphericChemistry.cpp:227); the
  // aerosol driver leaves it zero (examples/TChem_AerosolChemistry.cpp:299-302)
  std::vector<double> t(n, o.tbeg), dt(n, kAero ? 0.0 : o.dtmax);
  const double tn[2] = {o.atol_newton, o.rtol_newton};
  std::vector<double> tt(size_t(2) * (m > 0 ? m : 1));
  for (int i = 0; i < m; ++i) {
    tt[2 * i] = o.atol_time;  // 1e-12 (fixed in TChem_AtmosphericChemistry.x, --atol-time in the E3SM/aerosol drivers)
    tt[2 * i + 1] = o.rtol_time;
  }
  if (kE3sm) {
    if (nPhotoCols) {
      CHECK_H(hipMalloc(&d_photo.p, sizeof(double) * photo.size()));
      CHECK_H(hipMemcpy(d_photo.p, photo.data(), sizeof(double) * photo.size(), hipMemcpyHostToDevice));
    }
    CHECK_H(hipMalloc(&d_ext.p, sizeof(double) * (ext.empty() ? 1 : ext.size())));
    if (!ext.empty()) CHECK_H(hipMemcpy(d_ext.p, ext.data(), sizeof(double) * ext.size(), hipMemcpyHostToDevice));
  }
  if (kAero && nPart > 0) {
    CHECK_H(hipMalloc(&d_nc.p, sizeof(double) * nc.size()));
    CHECK_H(hipMemcpy(d_nc.p, nc.data(), sizeof(double) * nc.size(), hipMemcpyHostToDevice));
  }
  tchem_time_advance_t ta{};
  ta._tbeg = o.tbeg;
  ta._tend = o.tend;
  ta._dt = o.dtmin;
  ta._dtmin = o.dtmin;
  ta._dtmax = o.dtmax;
  ta._max_num_newton_iterations = o.max_newton;
  ta._num_time_iterations_per_interval = o.per_interval;
  ta._jacobian_interval = o.jac_interval;
  std::vector<tchem_time_advance_t> tadv(n, ta);
  CHECK_H(hipMemcpy(d_state.p, state.data(), sizeof(double) * state.size(), hipMemcpyHostToDevice));
  CHECK_H(hipMemcpy(d_t.p, t.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  CHECK_H(hipMemcpy(d_dt.p, dt.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  CHECK_H(hipMemcpy(d_tn.p, tn, sizeof(tn), hipMemcpyHostToDevice));
  CHECK_H(hipMemcpy(d_tt.p, tt.data(), sizeof(double) * tt.size(), hipMemcpyHostToDevice));
  CHECK_H(hipMemcpy(d_tadv.p, tadv.data(), sizeof(tchem_time_advance_t) * n, hipMemcpyHostToDevice));

  FILE* fout = nullptr;
  if (o.write_profiles) {
    fout = std::fopen(o.outputfile.c_str(), "w");
    if (!fout) {
      std::fprintf(stderr, "cannot write %s\n", o.outputfile.c_str());
      return 1;
    }
    std::fprintf(fout, "%s \t %s \t %s \t ", "iter", "t", "dt");
    std::fprintf(fout, "%s \t %s \t %s \t", "Density[kg/m3]", "Pressure[Pascal]", "Temperature[K]");
    for (const std::string& nm : names) std::fprintf(fout, "%s \t", nm.c_str());
    std::fprintf(fout, "\n");
    writeState(fout, -1, t, dt, state, n, dim);
  }
  FILE* ftimes = std::fopen(o.times.c_str(), "w");
  if (!ftimes) {
    std::fprintf(stderr, "cannot write %s\n", o.times.c_str());
    return 1;
  }
  std::fprintf(ftimes, "{\n");
  std::fprintf(ftimes, kE3sm ? " \"Atmospheric Chemistry E3SM\": \n {\n" : " \"Aerosol Chemistry\": \n {\n");

  tchem_atmchem_options_t opts{};
  opts.precision = o.precision;
  opts.lu = o.lu;
  opts.schedule = o.schedule;
  int iter = 0;
  double tsum = 0.0;
  for (; iter < o.max_iters && tsum <= o.tend * 0.9999; ++iter) {
    const auto w0 = std::chrono::steady_clock::now();
    if (kAero)
      CHECK_T(tchem_aerosolchem_run_device_batch(nullptr, n, d_tn.p, d_tt.p, nullptr, d_tadv.p, d_state.p, d_nc.p,
                                                 d_t.p, d_dt.p, d_state.p, nullptr, amcd, nullptr));
    else if (kE3sm)
      CHECK_T(tchem_atmchem_e3sm_run_device_batch(&opts, n, d_tn.p, d_tt.p, nullptr, d_tadv.p, d_state.p, d_photo.p,
                                                  nPhotoCols, m > 0 ? d_ext.p : nullptr, d_t.p, d_dt.p, d_state.p,
                                                  nullptr, kmcd, nullptr));
    else
      CHECK_T(tchem_atmchem_run_device_batch(&opts, n, d_tn.p, d_tt.p, nullptr, d_tadv.p, d_state.p, nullptr, d_t.p,
                                             d_dt.p, d_state.p, nullptr, kmcd, nullptr));
    CHECK_H(hipDeviceSynchronize());
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
    std::fprintf(ftimes, "\"%s%d\": %20.14e, \n", "wall_time_iter_", iter, wall);
    CHECK_H(hipMemcpy(t.data(), d_t.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    CHECK_H(hipMemcpy(dt.data(), d_dt.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (o.write_profiles) {
      CHECK_H(hipMemcpy(state.data(), d_state.p, sizeof(double) * state.size(), hipMemcpyDeviceToHost));
      writeState(fout, iter, t, dt, state, n, dim);
    }
    if (o.verbose) {
      if (!o.write_profiles)
        CHECK_H(hipMemcpy(state.data(), d_state.p, sizeof(double) * dim, hipMemcpyDeviceToHost));
      std::printf("%e %e %e %e", t[0], state[0], state[1], state[2]);
      for (int k = 3; k < dim; ++k) std::printf(" %e", state[k]);
      std::printf("\n");
    }
    CHECK_T(tchem_atmchem_update_time_advance(n, d_tadv.p, d_t.p, d_dt.p, nullptr));
    tsum = 0.0;
    for (int i = 0; i < n; ++i) tsum += t[i];
    tsum /= n;
  }
  std::fprintf(ftimes, "%s: %d, \n", "\"number_of_time_iters\"", iter);
  std::fprintf(ftimes, "%s: %d \n", "\"number_of_samples\"", n);
  std::fprintf(ftimes, "} \n ");
  std::fprintf(ftimes, "}\n ");
  std::fclose(ftimes);
  if (fout) std::fclose(fout);
  if (amcd) tchem_amcd_destroy(amcd);
  if (amd) tchem_amd_destroy(amd);
  tchem_kmcd_destroy(kmcd);
  tchem_kmd_destroy(kmd);
  return 0;
}
