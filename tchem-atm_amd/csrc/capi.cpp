// C ABI (include/tchem_atm.h): handles, validation, device table upload and
// kernel launches.  No exception crosses this boundary; every entry point
// returns a status code and records a message for tchem_last_error().
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <array>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/tchem_atm.h"
#include "aerosol_model.hpp"
#include "device_mech.hpp"
#include "kinetic_model.hpp"
#include "mechanism.hpp"
#include "schedule.hpp"

namespace tchem_amd {
hipError_t launchIntegrator(const IntegratorArgs& a, int precision, int grid, size_t lds, hipStream_t stream);
hipError_t integratorMaxBlocksPerCU(int precision, int sparse, int m, int threads, size_t lds, int* blocks, bool rec);
int sparseMaxCellsPerGroup(int precision, int m);
size_t integratorStaticLds(int precision, int sparse, int m);
hipError_t launchResetCounter(int* counter, hipStream_t stream);
hipError_t launchAutotuneSample(int n, long long stride, int nb, int np, int iters, const double* state,
                                const double* photo, const tchem_time_advance_t* tadv, double* st0, double* ph,
                                tchem_time_advance_t* ta, double* tt0, int* live, hipStream_t stream);
hipError_t launchCopyWords(void* dst, const void* src, size_t words, hipStream_t stream);
size_t mfmaCellLdsBytes(const DevMech& M);
hipError_t launchDenseMfmaSolve(int n, int nBatch, const double* A, const double* b, double* x, hipStream_t stream);
hipError_t noAllocScratch(void** ptr, size_t* bytes);
hipError_t launchRates(const DevMech& M, int nBatch, const double* state, const double* photo, const double* lscale,
                       int lstride, double* kf, const int* skip, hipStream_t stream, int rowDim = 0);
int aerosolCellsPerGroup(const DevMech& M, const DevAerosol& X, size_t maxLds);
hipError_t launchAerosol(const IntegratorArgs& a, int grid, size_t lds, hipStream_t stream);
hipError_t launchAerosolRhs(const IntegratorArgs& a, size_t lds, hipStream_t stream);
hipError_t launchLoopControl(int n, const double* t_out, const tchem_time_advance_t* tadv, double fraction, int* ctl,
                             bool reset, hipStream_t stream);
size_t adjointCellBytes(const DevMech& M, const DevSparse& S);
hipError_t adjointMaxBlocksPerCU(int m, int threads, size_t lds, int* blocks);
hipError_t launchAdjoint(const AdjointArgs& a, int grid, size_t lds, hipStream_t stream);
hipError_t launchE3smKinetics(const DevE3sm& E, int mode, int nBatch, const double* state, const double* photo,
                              int nPhoto, const double* ext, double* out, int ldOut, hipStream_t stream);
hipError_t launchAdjointSetup(int nBatch, int maxSteps, const tchem_time_advance_t* tadv, const double* tol_newton,
                              int* header, double* tol, int* newton, hipStream_t stream);
hipError_t launchUpdateTadv(int n, tchem_time_advance_t* tadv, const double* t_out, const double* dt_out, int* ctl,
                            hipStream_t stream);
hipError_t launchSourceTerm(const DevMech& M, int nBatch, const double* state, const double* photo, const double* ext,
                            double* omega, hipStream_t stream);
hipError_t launchJacobian(const DevMech& M, int nBatch, const double* state, const double* photo, double* jac,
                          hipStream_t stream);
}  // namespace tchem_amd

using namespace tchem_amd;

namespace {
// roctx ranges around the batch calls: the reference's Kokkos::Profiling
// regions (src/core/TChem_AtmosphericChemistry.cpp:61,166,193,207), seen by
// rocprofv3 --marker-trace.  The ROCm roctx library is not linked: loading a
// second ROCm's profiler registration next to the HIP runtime a host
// framework (torch) brings makes that runtime find no device.  It is used
// when the process already holds it (a profiler preloaded it) or when
// TCHEM_ROCTX=1 asks for it; otherwise a range costs one null check.
struct RoctxApi {
  int (*push)(const char*) = nullptr;
  int (*pop)() = nullptr;
};
const RoctxApi& roctxApi() {
  static const RoctxApi api = [] {
    RoctxApi a;
    const char* lib = "librocprofiler-sdk-roctx.so.1";
    void* h = dlopen(lib, RTLD_NOW | RTLD_NOLOAD);
    const char* want = std::getenv("TCHEM_ROCTX");
    if (!h && want && want[0] == '1') h = dlopen(lib, RTLD_NOW | RTLD_LOCAL);
    if (h) {
      a.push = reinterpret_cast<int (*)(const char*)>(dlsym(h, "roctxRangePushA"));
      a.pop = reinterpret_cast<int (*)()>(dlsym(h, "roctxRangePop"));
      if (!a.push || !a.pop) a = RoctxApi{};
    }
    return a;
  }();
  return api;
}
struct ProfileRange {
  explicit ProfileRange(const char* name) : on(roctxApi().push != nullptr) {
    if (on) roctxApi().push(name);
  }
  ~ProfileRange() {
    if (on) roctxApi().pop();
  }
  ProfileRange(const ProfileRange&) = delete;
  ProfileRange& operator=(const ProfileRange&) = delete;
  bool on;
};
}  // namespace

struct tchem_kmd_s {
  KineticModel km;
};

struct tchem_kmcd_s {
  MechTables tables;
  bool boxModel = true;     // the mechanism compiles for the box-model integrator
  std::string boxError;     // why not
  bool cb05Kinetics = true;  // ... with the CB05 rate constants (runDeviceBatch, source term, Jacobian, adjoint)
  std::string e3smBoxError;  // empty when the E3SM box model (AtmosphericChemistryE3SM) can run
  DevE3sm e3{};             // E3SM kinetics tables (every mechanism the loader reads)
  SparseSchedule sched, schedT;  // forward and transposed (adjoint) programs
  DevMech dev{};
  DevSparse sparse{}, sparseT{};
  std::vector<void*> allocations;
  int device = 0;
  int numCUs = 0;
  size_t maxLds = 0;
  int* queue = nullptr;
  double* kf = nullptr;  // [capacity][nReac] rate constants of the last batch
  size_t kfCapacity = 0;  // cells
  int graphs = 0;         // live captured graphs that read this kmcd's tables
  ~tchem_kmcd_s() {
    if (kf) allocations.push_back(kf);
    int cur = 0;
    if (hipGetDevice(&cur) == hipSuccess && cur != device) (void)hipSetDevice(device);
    for (void* p : allocations) (void)hipFree(p);
    if (cur != device) (void)hipSetDevice(cur);
  }
};

struct tchem_amd_s {
  AerosolModel am;
  KineticModel km;  // the gas model it was read against (names, state layout)
};

struct tchem_amcd_s {
  tchem_kmcd_s* kmcd = nullptr;  // not owned
  DevAerosol X{};
  DevSparse S{};                 // programs over the mt unknowns (S.enabled = 0: dense form only)
  SparseSchedule sched;
  int stateDim = 0;
  std::vector<void*> allocations;
  int device = 0;
  ~tchem_amcd_s() {
    int cur = 0;
    if (hipGetDevice(&cur) == hipSuccess && cur != device) (void)hipSetDevice(device);
    for (void* p : allocations) (void)hipFree(p);
    if (cur != device) (void)hipSetDevice(cur);
  }
};

// A captured interval loop owns the buffers its kernels write (rate
// constants, work-queue counter, loop control), so later calls on the same
// kmcd with other batch sizes cannot move them; it reads the kmcd's
// mechanism tables, so the kmcd refuses destruction while graphs exist.
struct tchem_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int* ctl = nullptr;    // [0] stopped, [1] intervals run (device); [64] work-queue counter
  double* kf = nullptr;  // [nBatch][nReac] rate constants of the captured launches
  tchem_kmcd_s* kmcd = nullptr;
  int device = 0;
  int intervals = 0;
  ~tchem_graph_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ctl) (void)hipFree(ctl);
    if (kf) (void)hipFree(kf);
    if (kmcd) --kmcd->graphs;
  }
};

namespace {

thread_local std::string g_error;

int setError(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

int hipFail(hipError_t e, const char* what) {
  return setError(TCHEM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
int upload(tchem_kmcd_s* k, const std::vector<T>& v, const T** out) {
  if (v.empty()) {
    *out = nullptr;
    return TCHEM_OK;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, sizeof(T) * v.size());
  if (e != hipSuccess) return hipFail(e, "hipMalloc");
  k->allocations.push_back(p);
  e = hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hipFail(e, "hipMemcpy");
  *out = static_cast<const T*>(p);
  return TCHEM_OK;
}

// Options after validation, with the LU form resolved.
int resolveOptions(const tchem_atmchem_options_t* opts, const tchem_kmcd_s* k, tchem_atmchem_options_t* o,
                   const char* who) {
  *o = tchem_atmchem_options_t{};
  if (opts) *o = *opts;
  for (int r : o->reserved)
    if (r != 0) return setError(TCHEM_ERR_INVALID, std::string(who) + ": reserved option fields must be zero");
  if (o->precision != TCHEM_PRECISION_FP64 && o->precision != TCHEM_PRECISION_MIXED)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": unknown precision");
  if (o->schedule != TCHEM_SCHEDULE_GRID && o->schedule != TCHEM_SCHEDULE_PERSISTENT)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": unknown schedule");
  if (o->blocks_per_cu < 0) return setError(TCHEM_ERR_INVALID, std::string(who) + ": negative blocks_per_cu");
  if (o->lu != TCHEM_LU_AUTO && o->lu != TCHEM_LU_DENSE && o->lu != TCHEM_LU_SPARSE && o->lu != TCHEM_LU_MFMA)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": unknown lu");
  // AUTO in mixed precision: the MFMA form when the factor is dense (at
  // least 30 % of the m^2 entries in the fill-reducing order) and large
  // enough to fill its 256 x 256 tiles (m >= 128)
  if (o->lu == TCHEM_LU_AUTO && o->precision == TCHEM_PRECISION_MIXED && o->cells_per_block <= 1 &&
      k->dev.m >= 128 && k->dev.m <= 255 && k->dev.oMfjTermW != 0 && mfmaCellLdsBytes(k->dev) <= k->maxLds &&
      (!k->sparse.enabled || double(k->sparse.nnzF) >= 0.3 * double(k->dev.m) * double(k->dev.m)))
    o->lu = TCHEM_LU_MFMA;
  if (o->lu == TCHEM_LU_MFMA) {
    if (o->precision != TCHEM_PRECISION_MIXED)
      return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": the MFMA LU form is mixed precision (FP32 factors)");
    if (k->dev.m > 255 || k->dev.oMfjTermW == 0)
      return setError(TCHEM_ERR_UNSUPPORTED,
                      std::string(who) + ": the MFMA LU form takes at most 255 species and 16383 derivative pairs");
    if (o->cells_per_block > 1)
      return setError(TCHEM_ERR_INVALID, std::string(who) + ": the MFMA LU form runs one cell per workgroup");
    o->cells_per_block = 1;
    if (mfmaCellLdsBytes(k->dev) > k->maxLds)
      return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": mechanism too large for the MFMA form's LDS");
    return TCHEM_OK;
  }
  if (o->cells_per_block < 0) return setError(TCHEM_ERR_INVALID, std::string(who) + ": negative cells_per_block");
  if (o->programs < TCHEM_PROGRAMS_AUTO || o->programs > TCHEM_PROGRAMS_MEMORY)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": unknown programs placement");
  const size_t bytesA = o->precision == TCHEM_PRECISION_FP64 ? sizeof(double) : sizeof(float);
  const size_t staticLds = integratorStaticLds(o->precision, 1, k->dev.m);
  const size_t maxLds = k->maxLds > staticLds ? k->maxLds - staticLds : 0;
  const size_t cellB = sparseCellBytes(k->dev, k->sparse, bytesA);
  // AUTO: programs staged into LDS when that still leaves room for as many
  // cells per workgroup as reading them from memory would (the kernel is
  // latency bound: resident cells set its throughput); memory otherwise
  if (o->programs == TCHEM_PROGRAMS_AUTO) {
    const int cap = sparseMaxCellsPerGroup(o->precision, k->dev.m);
    const size_t prog = sparseProgramBytes(k->sparse);
    const size_t coefB = sparseCoefBytes(k->sparse);
    const size_t inMem = coefB < maxLds ? std::min<size_t>(cap, cellB ? (maxLds - coefB) / cellB : 0) : 0;
    const size_t inLds = prog < maxLds ? std::min<size_t>(cap, cellB ? (maxLds - prog) / cellB : 0) : 0;
    o->programs = (inLds >= 1 && inLds >= inMem) ? TCHEM_PROGRAMS_LDS : TCHEM_PROGRAMS_MEMORY;
  }
  const bool staged = o->programs == TCHEM_PROGRAMS_LDS;
  auto sparseFits = [&](int cells) {
    return (staged ? sparseProgramBytes(k->sparse) : sparseCoefBytes(k->sparse)) + size_t(cells) * cellB <= maxLds;
  };
  if (o->lu == TCHEM_LU_AUTO)
    o->lu = (k->sparse.enabled && sparseFits(1)) ? TCHEM_LU_SPARSE : TCHEM_LU_DENSE;
  if (o->lu == TCHEM_LU_SPARSE) {
    if (!k->sparse.enabled)
      return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": mechanism has no sparse schedule");
    const int cap = sparseMaxCellsPerGroup(o->precision, k->dev.m);
    if (o->cells_per_block == 0) {
      int c = cap;
      while (c > 1 && !sparseFits(c)) --c;
      o->cells_per_block = c;
    }
    if (o->cells_per_block > cap)
      return setError(TCHEM_ERR_INVALID, std::string(who) + ": cells_per_block above " + std::to_string(cap));
    if (!sparseFits(o->cells_per_block))
      return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": " + std::to_string(o->cells_per_block) +
                                                 " cells per block do not fit in LDS");
  } else {
    if (o->cells_per_block > 1)
      return setError(TCHEM_ERR_INVALID, std::string(who) + ": the dense form runs one cell per block");
    o->cells_per_block = 1;
    if (integratorLdsBytes(k->dev, bytesA) > maxLds)
      return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": the dense form's matrix does not fit in LDS");
  }
  return TCHEM_OK;
}

// whether the sparse programs are staged into LDS (resolved options)
bool sparseStaged(const tchem_atmchem_options_t& o) { return o.programs == TCHEM_PROGRAMS_LDS; }

// LDS bytes of one workgroup (resolved options)
size_t ldsBytes(const tchem_kmcd_s* k, const tchem_atmchem_options_t& o) {
  if (o.lu == TCHEM_LU_MFMA) return mfmaCellLdsBytes(k->dev);
  const size_t bytesA = o.precision == TCHEM_PRECISION_FP64 ? sizeof(double) : sizeof(float);
  if (o.lu == TCHEM_LU_SPARSE)
    return (sparseStaged(o) ? sparseProgramBytes(k->sparse) : sparseCoefBytes(k->sparse)) +
           size_t(o.cells_per_block) * sparseCellBytes(k->dev, k->sparse, bytesA);
  return integratorLdsBytes(k->dev, bytesA);
}

// Grows the per-cell rate-constant buffer to nBatch cells (allocation only
// when a larger batch arrives; the previous buffer may still be read by work
// queued on s, so s is synchronised first).
int ensureKf(tchem_kmcd_s* kmcd, int nBatch, hipStream_t s) {
  if (size_t(nBatch) <= kmcd->kfCapacity) return TCHEM_OK;
  hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hipFail(e, "hipStreamSynchronize");
  if (kmcd->kf) (void)hipFree(kmcd->kf);
  kmcd->kf = nullptr;
  kmcd->kfCapacity = 0;
  void* p = nullptr;
  e = hipMalloc(&p, sizeof(double) * size_t(nBatch) * size_t(kmcd->dev.nReac > 0 ? kmcd->dev.nReac : 1));
  if (e != hipSuccess) return hipFail(e, "hipMalloc (rate constants)");
  kmcd->kf = static_cast<double*>(p);
  kmcd->kfCapacity = size_t(nBatch);
  return TCHEM_OK;
}

// Resident workgroups per CU of the persistent schedule (options value, or
// the occupancy limit when 0).
int resolveBlocksPerCU(const tchem_kmcd_s* kmcd, const tchem_atmchem_options_t& o, bool rec, int* bpc) {
  *bpc = o.blocks_per_cu;
  if (*bpc > 0) return TCHEM_OK;
  const int sparse = o.lu == TCHEM_LU_MFMA ? 3 : (o.lu == TCHEM_LU_SPARSE ? (sparseStaged(o) ? 1 : 2) : 0);
  hipError_t e = integratorMaxBlocksPerCU(o.precision, sparse, kmcd->dev.m, 64 * o.cells_per_block,
                                          ldsBytes(kmcd, o), bpc, rec);
  if (e != hipSuccess) return hipFail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  if (*bpc < 1) *bpc = 1;
  return TCHEM_OK;
}

// Rate constants (optionally log-scaled) into a.kf, then the integrator
// launch for the resolved options; a carries the per-call arrays.
int runIntegrator(tchem_kmcd_s* kmcd, const tchem_atmchem_options_t& o, const double* lscale, int lstride,
                  IntegratorArgs& a, hipStream_t s, const char* who, bool ratesReady = false) {
  const size_t lds = ldsBytes(kmcd, o);
  if (lds > kmcd->maxLds)
    return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": mechanism needs " + std::to_string(lds) +
                                               " bytes of LDS per workgroup, device limit " +
                                               std::to_string(kmcd->maxLds));
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess || cur != kmcd->device)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": kmcd belongs to another device");
  a.M = kmcd->dev;
  a.S = kmcd->sparse;
  a.sparse = o.lu == TCHEM_LU_MFMA ? 3 : (o.lu == TCHEM_LU_SPARSE ? (sparseStaged(o) ? 1 : 2) : 0);
  if (!a.queue) a.queue = kmcd->queue;
  a.persistent = (o.schedule == TCHEM_SCHEDULE_PERSISTENT) ? 1 : 0;
  a.cellsPerGroup = o.cells_per_block;
  a.cellBytes = (a.sparse == 1 || a.sparse == 2)
                    ? int(sparseCellBytes(kmcd->dev, kmcd->sparse,
                                          o.precision == TCHEM_PRECISION_FP64 ? sizeof(double) : sizeof(float)))
                    : 0;
  if (!ratesReady) {
    hipError_t e = launchRates(kmcd->dev, a.nBatch, a.state, a.photo, lscale, lstride, const_cast<double*>(a.kf), a.skip,
                               s);
    if (e != hipSuccess) return hipFail(e, "rates_kernel launch");
  }
  const bool rec = a.rec.h != nullptr;
  const int cpg = a.cellsPerGroup;
  const int groups = (a.nBatch + cpg - 1) / cpg;
  int grid = groups;
  if (a.persistent) {
    int bpc = 0;
    const int rc = resolveBlocksPerCU(kmcd, o, rec, &bpc);
    if (rc) return rc;
    const long long g = (long long)bpc * kmcd->numCUs;
    grid = int(g < groups ? g : groups);
    hipError_t e = launchResetCounter(a.queue, s);
    if (e != hipSuccess) return hipFail(e, "reset_counter_kernel");
  }
  hipError_t e = launchIntegrator(a, o.precision, grid, lds, s);
  if (e != hipSuccess) return hipFail(e, "trbdf2_kernel launch");
  return TCHEM_OK;
}

// Byte offsets of the adjoint workspace sections (256-byte aligned).
struct AdjointLayout {
  size_t header, tol, newton, nsteps, h, ckpt, consts, kf, traj, total;
};
AdjointLayout adjointLayout(const DevMech& M, int nBatch, int maxSteps) {
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t n = size_t(nBatch), nCkpt = size_t(maxSteps / kAdjointSegment + 1);
  AdjointLayout L{};
  size_t o = 0;
  L.header = o; o += al(4 * sizeof(int));
  L.tol = o; o += al(2 * sizeof(double));
  L.newton = o; o += al(2 * sizeof(int) * n);
  L.nsteps = o; o += al(sizeof(int) * n);
  L.h = o; o += al(sizeof(double) * n * size_t(maxSteps));
  L.ckpt = o; o += al(sizeof(double) * n * nCkpt * size_t(M.m));
  L.consts = o; o += al(sizeof(double) * n * size_t(M.nSpec - M.m));
  L.kf = o; o += al(sizeof(double) * n * size_t(M.nReac));
  L.traj = o; o += al(sizeof(double) * n * size_t(2 * kAdjointSegment + 1) * size_t(M.m));
  L.total = o;
  return L;
}

int adjointCheck(const tchem_kmcd_s* k, int nBatch, int maxSteps, const char* who) {
  if (!k) return setError(TCHEM_ERR_INVALID, std::string(who) + ": null kmcd");
  if (!k->cb05Kinetics) return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": " + k->boxError);
  if (nBatch < 0) return setError(TCHEM_ERR_INVALID, std::string(who) + ": negative nBatch");
  if (maxSteps < 1 || maxSteps > (1 << 24))
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": maxSteps must be in [1, 2^24]");
  if (!k->sparse.enabled || !k->sparseT.enabled)
    return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": mechanism has no sparse (transposed) schedule");
  return TCHEM_OK;
}

}  // namespace

extern "C" {

const char* tchem_last_error(void) { return g_error.c_str(); }

int tchem_kmd_load_yaml(const char* path, const char* echo_path, tchem_kmd_t* out) {
  if (!path || !out) return setError(TCHEM_ERR_INVALID, "tchem_kmd_load_yaml: null argument");
  *out = nullptr;
  try {
    std::unique_ptr<tchem_kmd_s> h(new tchem_kmd_s);
    if (echo_path) {
      std::ofstream echo(echo_path);
      if (!echo) return setError(TCHEM_ERR_INVALID, std::string("cannot write echo file ") + echo_path);
      h->km = loadKineticModel(path, &echo);
    } else {
      h->km = loadKineticModel(path, nullptr);
    }
    *out = h.release();
    return TCHEM_OK;
  } catch (const std::exception& e) {
    return setError(TCHEM_ERR_PARSE, e.what());
  }
}

int tchem_kmd_destroy(tchem_kmd_t kmd) {
  delete kmd;
  return TCHEM_OK;
}

int tchem_kmd_info(tchem_kmd_t kmd, int* nSpec, int* nConstSpec, int* nReac, int* nPhoto) {
  if (!kmd) return setError(TCHEM_ERR_INVALID, "tchem_kmd_info: null handle");
  if (nSpec) *nSpec = kmd->km.nSpec;
  if (nConstSpec) *nConstSpec = kmd->km.nConstSpec;
  if (nReac) *nReac = kmd->km.nReac;
  if (nPhoto) *nPhoto = int(kmd->km.photolysisReactions.size());
  return TCHEM_OK;
}

int tchem_kmd_species_name(tchem_kmd_t kmd, int i, char* buf, size_t len) {
  if (!kmd || !buf || len == 0) return setError(TCHEM_ERR_INVALID, "tchem_kmd_species_name: bad argument");
  if (i < 0 || i >= kmd->km.nSpec) return setError(TCHEM_ERR_INVALID, "tchem_kmd_species_name: index out of range");
  std::strncpy(buf, kmd->km.speciesNames[i].c_str(), len - 1);
  buf[len - 1] = '\0';
  return TCHEM_OK;
}

int tchem_kmd_photolysis_reactions(tchem_kmd_t kmd, int* reactions, int capacity) {
  if (!kmd || (!reactions && capacity > 0)) return setError(TCHEM_ERR_INVALID, "tchem_kmd_photolysis_reactions: bad argument");
  const auto& v = kmd->km.photolysisReactions;
  if (capacity < int(v.size())) return setError(TCHEM_ERR_INVALID, "tchem_kmd_photolysis_reactions: capacity too small");
  for (size_t q = 0; q < v.size(); ++q) reactions[q] = v[q];
  return TCHEM_OK;
}

int tchem_kmd_elimination_order(tchem_kmd_t kmd, int* order, int capacity) {
  if (!kmd || !order) return setError(TCHEM_ERR_INVALID, "tchem_kmd_elimination_order: bad argument");
  try {
    const MechTables t = compileMechanism(kmd->km);
    if (capacity < t.m) return setError(TCHEM_ERR_INVALID, "tchem_kmd_elimination_order: capacity too small");
    for (int a = 0; a < t.m; ++a) order[a] = t.perm[a];
    return TCHEM_OK;
  } catch (const std::exception& e) {
    return setError(TCHEM_ERR_UNSUPPORTED, std::string("tchem_kmd_elimination_order: ") + e.what());
  }
}

int tchem_scenario_load(tchem_kmd_t kmd, const char* path, double* state_host, int capacity_rows, int* nBatch) {
  if (!kmd || !path || !nBatch) return setError(TCHEM_ERR_INVALID, "tchem_scenario_load: null argument");
  try {
    int n = 0;
    std::vector<double> s = loadScenarioState(path, kmd->km, &n);
    *nBatch = n;
    if (state_host) {
      if (capacity_rows < n) return setError(TCHEM_ERR_INVALID, "tchem_scenario_load: capacity too small");
      std::memcpy(state_host, s.data(), sizeof(double) * s.size());
    }
    return TCHEM_OK;
  } catch (const std::exception& e) {
    return setError(TCHEM_ERR_PARSE, e.what());
  }
}

int tchem_kmcd_create(tchem_kmd_t kmd, tchem_kmcd_t* out) {
  if (!kmd || !out) return setError(TCHEM_ERR_INVALID, "tchem_kmcd_create: null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return setError(TCHEM_ERR_NO_DEVICE, "tchem_kmcd_create: no HIP device");
  std::unique_ptr<tchem_kmcd_s> k(new tchem_kmcd_s);
  try {
    k->tables = compileMechanism(kmd->km);
  } catch (const std::exception& e) {
    // e.g. JPL / PHOTO_RATE reactions: only the E3SM kinetics entry points work
    k->boxModel = false;
    k->boxError = e.what();
    k->tables = MechTables{};
  }
  if (k->boxModel && k->tables.m > 256) {
    k->boxModel = false;
    k->boxError = "more than 256 active species";
    k->tables = MechTables{};
  }
  const MechTables& t = k->tables;
  hipError_t e = hipGetDevice(&k->device);
  if (e != hipSuccess) return hipFail(e, "hipGetDevice");
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, k->device);
  if (e != hipSuccess) return hipFail(e, "hipGetDeviceProperties");
  k->numCUs = prop.multiProcessorCount;
  k->maxLds = prop.sharedMemPerBlock;

  DevMech& d = k->dev;
  d.nSpec = t.nSpec;
  d.nConst = t.nConst;
  d.m = t.m;
  d.nReac = t.nReac;
  d.nPhoto = t.nPhoto;
  d.maxR = t.maxReactants;
  d.nPairs = t.nPairs();
  d.nnzJ = t.nnzJ();
  d.nSeg = t.nSeg();
  d.nSerJ = t.nSerJ();
  d.convPPM = t.convPPM;
  std::vector<int> rowcol(t.jnzRow.size());
  for (size_t n = 0; n < rowcol.size(); ++n) rowcol[n] = (t.jnzRow[n] << 16) | t.jnzCol[n];
  std::vector<int> serRowcol(t.serJRow.size());
  for (size_t n = 0; n < serRowcol.size(); ++n) serRowcol[n] = (t.serJRow[n] << 16) | t.serJCol[n];
  // pack every table into one int and one double buffer
  std::vector<int> ib;
  std::vector<double> db;
  auto addI = [&](const std::vector<int>& v) {
    const int off = int(ib.size());
    ib.insert(ib.end(), v.begin(), v.end());
    while (ib.size() % 4) ib.push_back(0);
    return off;
  };
  auto addD = [&](const std::vector<double>& v) {
    const int off = int(db.size());
    db.insert(db.end(), v.begin(), v.end());
    while (db.size() % 2) db.push_back(0.0);
    return off;
  };
  d.oPerm = addI(t.perm);
  d.oRateKind = addI(t.rateKind);
  d.oConvExp = addI(t.convExp);
  d.oPhotoSlot = addI(t.photoSlot);
  d.oReacSpec = addI(t.reacSpec);
  d.oReacPow = addI(t.reacPow);
  d.oOmegaReac = addI(t.omegaReac);
  d.oSegBegin = addI(t.segBegin);
  d.oSegEnd = addI(t.segEnd);
  d.oRowSegPtr = addI(t.rowSegPtr);
  d.oPairPtr = addI(t.pairPtr);
  d.oPairSlot = addI(t.pairSlot);
  d.oJnzRowCol = addI(rowcol);
  d.oJnzPtr = addI(t.jnzPtr);
  d.oJcPair = addI(t.jcPair);
  d.oRatePar = addD(t.ratePar);
  d.oOmegaCoef = addD(t.omegaCoef);
  d.oEmission = addD(t.emission);
  d.oJcCoef = addD(t.jcCoef);
  d.oSerOmegaPtr = addI(t.serOmegaPtr);
  d.oSerOmegaReac = addI(t.serOmegaReac);
  d.oSerOmegaCoef = addD(t.serOmegaCoef);
  d.oSerJRowCol = addI(serRowcol);
  d.oSerJPtr = addI(t.serJPtr);
  d.oSerJPair = addI(t.serJPair);
  d.oSerJCoef = addD(t.serJCoef);
  d.oMfjQuarter = d.oMfjTermW = d.oMfjTermC = 0;
  if (t.m <= 255 && t.nPairs() < (1 << 14)) {
    // the MFMA form's assembly lists: the structural entries of each column
    // quarter dealt to the 256 threads by term count (largest first, to the
    // least loaded thread), each thread's terms in entry order, padded to the
    // quarter's longest run, stored [term row][thread]
    constexpr int kThreads = 256;
    std::vector<int> quarter(5, 0);
    std::vector<int> termW, termC;
    for (int q = 0; q < 4; ++q) {
      quarter[size_t(q)] = int(termW.size() / kThreads);
      std::vector<size_t> ent;
      for (size_t n = 0; n < t.serJRow.size(); ++n)
        if ((t.serJCol[n] >> 6) == q) ent.push_back(n);
      std::stable_sort(ent.begin(), ent.end(), [&](size_t x, size_t y) {
        return t.serJPtr[x + 1] - t.serJPtr[x] > t.serJPtr[y + 1] - t.serJPtr[y];
      });
      std::vector<std::vector<size_t>> own(kThreads);
      std::vector<int> load(kThreads, 0);
      for (size_t n : ent) {
        const int th = int(std::min_element(load.begin(), load.end()) - load.begin());
        own[size_t(th)].push_back(n);
        load[size_t(th)] += std::max(1, t.serJPtr[n + 1] - t.serJPtr[n]);
      }
      const int rows = *std::max_element(load.begin(), load.end());
      const size_t base = termW.size();
      termW.resize(base + size_t(rows) * kThreads, 0);
      termC.resize(base + size_t(rows) * kThreads, 0);
      for (int th = 0; th < kThreads; ++th) {
        int j = 0;
        for (size_t n : own[size_t(th)]) {
          const int row = t.serJRow[n], col = t.serJCol[n];
          const unsigned dest = unsigned(row * 65 + (col & 63));
          const int e0 = t.serJPtr[n], e1 = t.serJPtr[n + 1];
          for (int e = e0; e < std::max(e1, e0 + 1); ++e, ++j) {
            unsigned w = e < e1 ? unsigned(t.serJPair[size_t(e)]) : 0u;
            float c = e < e1 ? float(t.serJCoef[size_t(e)]) : 0.0f;
            if (e + 1 >= e1) w |= (dest << 14) | (row == col ? (1u << 29) : 0u) | (1u << 31);
            int bits;
            std::memcpy(&bits, &c, sizeof(bits));
            termW[base + size_t(j) * kThreads + size_t(th)] = int(w);
            termC[base + size_t(j) * kThreads + size_t(th)] = bits;
          }
        }
      }
    }
    quarter[4] = int(termW.size() / kThreads);
    d.oMfjQuarter = addI(quarter);
    d.oMfjTermW = addI(termW);
    d.oMfjTermC = addI(termC);
  }
  // sparse schedules (mechanisms they cannot express run the dense form only)
  auto words = [](const std::vector<unsigned>& v) { return std::vector<int>(v.begin(), v.end()); };
  auto pack = [&](const SparseSchedule& sc, DevSparse& sp) {
    sp.enabled = 1;
    sp.nnzF = sc.nnzF;
    sp.zeroSlot = sc.zeroSlot;
    sp.trashSlot = sc.trashSlot;
    sp.pairSlots = sc.pairSlots;
    sp.oEvalHdr = addI(words(sc.evalHdr));
    sp.oEvalBody = addI(words(sc.evalBody));
    sp.evalBodyWords = int(sc.evalBody.size());
    sp.oCoef = addD(sc.coefTable);
    sp.nCoef = int(sc.coefTable.size());
    sp.ropRounds = sc.ropRounds;
    sp.pairRounds = sc.pairRounds;
    sp.segRounds = sc.segRounds;
    sp.jacRounds = sc.jacRounds;
    sp.pairBody = sc.pairBody;
    sp.segBody = sc.segBody;
    sp.jacBody = sc.jacBody;
    sp.nutRounds = sc.nutRounds;
    sp.jtvRounds = sc.jtvRounds;
    sp.nutBody = sc.nutBody;
    sp.jtvBody = sc.jtvBody;
    sp.oProgHdr = addI(words(sc.progHdr));
    sp.oProgBody = addI(words(sc.progBody));
    sp.luRounds = sc.luRounds;
    sp.backRounds = sc.backRounds;
    sp.fwdRounds = sc.fwdRounds;
    sp.backBody = sc.backBody;
    sp.fwdBody = sc.fwdBody;
  };
  try {
    if (!k->boxModel) throw std::runtime_error("no box model");
    k->sched = buildSparseSchedule(t);
    pack(k->sched, k->sparse);
  } catch (const std::exception&) {
    k->sparse = DevSparse{};
  }
  if (k->sparse.enabled) {
    // the adjoint shares the cell layout of the forward programs: the
    // transposed factor pattern has the same size
    try {
      k->schedT = buildSparseSchedule(t, true);
      if (k->schedT.nnzF == k->sched.nnzF && k->schedT.pairSlots == k->sched.pairSlots)
        pack(k->schedT, k->sparseT);
    } catch (const std::exception&) {
      k->sparseT = DevSparse{};
    }
  }
  {
    // E3SM kinetics tables straight from the loader's data
    const KineticModel& km = kmd->km;
    DevE3sm& E = k->e3;
    E.nSpec = km.nSpec;
    E.nConst = km.nConstSpec;
    E.nReac = km.nReac;
    E.maxSpec = km.maxSpecInReac;
    E.M_index = km.M_index;
    E.convPPM = km.CONV_PPM;
    std::vector<int> ai, ti, ci, ji, ri, adj, pI;
    std::vector<double> ap, tp, cp, jp, rp, pD;
    for (const auto& a : km.arrhenius) {
      ai.push_back(a.reaction);
      ap.insert(ap.end(), {a.A, a.B, a.C, a.D, a.E});
    }
    for (size_t i = 0; i < km.reacPfal.size(); ++i) {
      ti.push_back(km.reacPfal[i]);
      tp.insert(tp.end(), km.reacPpar[i].begin(), km.reacPpar[i].end());
    }
    for (const auto& c : km.cmaqH2O2) {
      ci.push_back(c.reaction);
      cp.insert(cp.end(), {c.A1, c.B1, c.C1, c.A2, c.B2, c.C2});
    }
    for (const auto& j : km.jpl) {
      ji.push_back(j.reaction);
      jp.insert(jp.end(), {j.k0_A, j.k0_B, j.k0_C, j.kinf_A, j.kinf_B, j.kinf_C, j.Fc});
    }
    for (const auto& j : km.rJplArrhenius) {
      ri.push_back(j.reaction);
      rp.insert(rp.end(), {j.k0_A, j.k0_B, j.k0_C, j.kinf_A, j.kinf_B, j.kinf_C, j.Fc, j.A, j.B, j.C});
    }
    for (const auto& a : km.adjustReaction) adj.insert(adj.end(), {a.reaction, a.species});
    for (const auto& o : km.prodO1D) {
      pI.insert(pI.end(), {o.reactions[0], o.reactions[1], o.reactions[2], o.photolysis_reaction, o.species1,
                           o.species2, o.species3});
      pD.insert(pD.end(), {o.A1, o.C1, o.A2, o.C2, o.A3, o.C3});
    }
    E.nArr = int(ai.size());
    E.nTroe = int(ti.size());
    E.nCmaq = int(ci.size());
    E.nJpl = int(ji.size());
    E.nRJpl = int(ri.size());
    E.nAdjust = int(km.adjustReaction.size());
    E.nProd = int(km.prodO1D.size());
    E.oArrIdx = addI(ai);
    E.oArrPar = addD(ap);
    E.oTroeIdx = addI(ti);
    E.oTroePar = addD(tp);
    E.oCmaqIdx = addI(ci);
    E.oCmaqPar = addD(cp);
    E.oJplIdx = addI(ji);
    E.oJplPar = addD(jp);
    E.oRJplIdx = addI(ri);
    E.oRJplPar = addD(rp);
    E.oAdjust = addI(adj);
    E.oProdI = addI(pI);
    E.oProdD = addD(pD);
    E.oConvExp = addI(km.convExponent);
    E.oNreac = addI(km.reacNreac);
    E.oNprod = addI(km.reacNprod);
    E.oSidx = addI(km.reacSidx);
    E.oNuki = addD(km.reacNuki);
  }
  ib.push_back(0);
  db.push_back(0.0);
  int rc = TCHEM_OK;
  if ((rc = upload(k.get(), ib, &d.ib)) || (rc = upload(k.get(), db, &d.db))) return rc;
  k->e3.ib = d.ib;
  k->e3.db = d.db;
  k->cb05Kinetics = k->boxModel && !t.externalRates;
  if (!k->cb05Kinetics && k->boxModel)
    k->boxError = "the mechanism has JPL / ratio-JPL / PHOTO_RATE reactions (E3SM kinetics): use the E3SM entry points";
  {
    // the E3SM box model takes its rate constants once per cell: every species
    // they depend on ([M], adjusted reactions, prod_O1D) must be constant
    const KineticModel& km = kmd->km;
    const int m = km.nActive();
    std::string why;
    if (!k->boxModel) why = k->boxError;
    if (why.empty() && km.M_index >= 0 && km.M_index < m && !km.jpl.empty()) why = "[M] is an active species";
    for (const auto& a : km.adjustReaction)
      if (why.empty() && a.species < m) why = "a reaction is adjusted by an active species";
    for (const auto& o : km.prodO1D)
      if (why.empty() && (o.species1 < m || o.species2 < m || o.species3 < m))
        why = "prod_O1D depends on an active species";
    k->e3smBoxError = why;
  }
  void* q = nullptr;
  e = hipMalloc(&q, 256);
  if (e != hipSuccess) return hipFail(e, "hipMalloc");
  k->allocations.push_back(q);
  k->queue = static_cast<int*>(q);
  *out = k.release();
  return TCHEM_OK;
}

int tchem_kmcd_destroy(tchem_kmcd_t kmcd) {
  if (kmcd && kmcd->graphs > 0)
    return setError(TCHEM_ERR_INVALID, "tchem_kmcd_destroy: " + std::to_string(kmcd->graphs) +
                                           " captured graph(s) still read this kmcd; destroy them first");
  delete kmcd;
  return TCHEM_OK;
}

int tchem_atmchem_workspace_bytes(tchem_kmcd_t kmcd, const tchem_atmchem_options_t* opts, size_t* bytes) {
  if (kmcd && !kmcd->cb05Kinetics)
    return setError(TCHEM_ERR_UNSUPPORTED, std::string("tchem_atmchem_workspace_bytes: ") + kmcd->boxError);
  if (!kmcd || !bytes) return setError(TCHEM_ERR_INVALID, "tchem_atmchem_workspace_bytes: null argument");
  tchem_atmchem_options_t o;
  const int rc = resolveOptions(opts, kmcd, &o, "tchem_atmchem_workspace_bytes");
  if (rc) return rc;
  *bytes = ldsBytes(kmcd, o);
  return TCHEM_OK;
}

int tchem_atmchem_run_device_batch(const tchem_atmchem_options_t* opts, int nBatch, const double* tol_newton,
                                   const double* tol_time, double* fac, const tchem_time_advance_t* tadv,
                                   double* state, const double* photo_rates, double* t_out, double* dt_out,
                                   double* state_out, tchem_cell_stats_t* stats, tchem_kmcd_t kmcd, void* stream) {
  ProfileRange profile_range("TChem::AtmosphericChemistry::runDeviceBatch::kmcd");
  if (kmcd && !kmcd->cb05Kinetics)
    return setError(TCHEM_ERR_UNSUPPORTED, std::string("run_device_batch: ") + kmcd->boxError);
  (void)fac;
  if (!kmcd) return setError(TCHEM_ERR_INVALID, "run_device_batch: null kmcd");
  if (nBatch < 0) return setError(TCHEM_ERR_INVALID, "run_device_batch: negative nBatch");
  tchem_atmchem_options_t o;
  const int rc = resolveOptions(opts, kmcd, &o, "run_device_batch");
  if (rc) return rc;
  if (nBatch == 0) return TCHEM_OK;
  if (!tol_newton || !tol_time || !tadv || !state || !t_out || !dt_out || !state_out)
    return setError(TCHEM_ERR_INVALID, "run_device_batch: null device array");
  if (kmcd->dev.nPhoto == 0) photo_rates = nullptr;
  hipStream_t s = static_cast<hipStream_t>(stream);
  {
    const int rk = ensureKf(kmcd, nBatch, s);
    if (rk) return rk;
  }
  IntegratorArgs a{};
  a.nBatch = nBatch;
  a.tol_newton = tol_newton;
  a.tol_time = tol_time;
  a.tadv = tadv;
  a.state = state;
  a.photo = photo_rates;
  a.t_out = t_out;
  a.dt_out = dt_out;
  a.state_out = state_out;
  a.stats = stats;
  a.kf = kmcd->kf;
  return runIntegrator(kmcd, o, nullptr, 0, a, s, "run_device_batch");
}

int tchem_atmchem_run_device_batch_noalloc(const tchem_atmchem_options_t* opts, int nBatch,
                                           const double* tol_newton, const double* tol_time, double* fac,
                                           const tchem_time_advance_t* tadv, double* state,
                                           const double* photo_rates, double* t_out, double* dt_out,
                                           double* state_out, tchem_cell_stats_t* stats, tchem_kmcd_t kmcd,
                                           void* stream) {
  ProfileRange profile_range("TChem::AtmosphericChemistry::runDeviceBatch::noalloc");
  const char* who = "run_device_batch_noalloc";
  if (kmcd && !kmcd->cb05Kinetics) return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": " + kmcd->boxError);
  (void)fac;
  if (!kmcd) return setError(TCHEM_ERR_INVALID, std::string(who) + ": null kmcd");
  if (nBatch < 0) return setError(TCHEM_ERR_INVALID, std::string(who) + ": negative nBatch");
  tchem_atmchem_options_t o;
  const int rc = resolveOptions(opts, kmcd, &o, who);
  if (rc) return rc;
  if (nBatch == 0) return TCHEM_OK;
  if (!tol_newton || !tol_time || !tadv || !state || !t_out || !dt_out || !state_out)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": null device array");
  if (kmcd->dev.nPhoto == 0) photo_rates = nullptr;
  void* scr = nullptr;
  size_t cap = 0;
  hipError_t e = noAllocScratch(&scr, &cap);
  if (e != hipSuccess) return hipFail(e, "hipGetSymbolAddress (scratch)");
  const size_t perCell = sizeof(double) * size_t(kmcd->dev.nReac > 0 ? kmcd->dev.nReac : 1);
  const size_t kfBytes = (perCell * size_t(nBatch) + 255) & ~size_t(255);
  if (kfBytes + 256 > cap)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": " + std::to_string(nBatch) + " cells need " +
                                           std::to_string(kfBytes + 256) + " bytes of scratch; it holds " +
                                           std::to_string(cap) + " (at most " + std::to_string((cap - 256) / perCell) +
                                           " cells of this mechanism per call)");
  IntegratorArgs a{};
  a.nBatch = nBatch;
  a.tol_newton = tol_newton;
  a.tol_time = tol_time;
  a.tadv = tadv;
  a.state = state;
  a.photo = photo_rates;
  a.t_out = t_out;
  a.dt_out = dt_out;
  a.state_out = state_out;
  a.stats = stats;
  a.kf = static_cast<double*>(scr);
  a.queue = reinterpret_cast<int*>(static_cast<char*>(scr) + kfBytes);
  return runIntegrator(kmcd, o, nullptr, 0, a, static_cast<hipStream_t>(stream), who);
}

int tchem_dense_lu_solve_mfma(int n, int nBatch, const double* A, const double* b, double* x, void* stream) {
  ProfileRange profile_range("TChem::DenseLuSolveMfma");
  if (n < 1 || n > 255 || nBatch < 0) return setError(TCHEM_ERR_INVALID, "tchem_dense_lu_solve_mfma: bad size");
  if (nBatch == 0) return TCHEM_OK;
  if (!A || !b || !x) return setError(TCHEM_ERR_INVALID, "tchem_dense_lu_solve_mfma: null device array");
  hipError_t e = launchDenseMfmaSolve(n, nBatch, A, b, x, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hipFail(e, "dense_mfma_solve_kernel launch");
  return TCHEM_OK;
}

int tchem_atmchem_e3sm_run_device_batch(const tchem_atmchem_options_t* opts, int nBatch, const double* tol_newton,
                                        const double* tol_time, double* fac, const tchem_time_advance_t* tadv,
                                        double* state, const double* photo_rates, int nPhotoRates,
                                        const double* external_sources, double* t_out, double* dt_out,
                                        double* state_out, tchem_cell_stats_t* stats, tchem_kmcd_t kmcd,
                                        void* stream) {
  ProfileRange profile_range("TChem::AtmosphericChemistryE3SM::runDeviceBatch");
  (void)fac;
  const char* who = "e3sm_run_device_batch";
  if (!kmcd) return setError(TCHEM_ERR_INVALID, std::string(who) + ": null kmcd");
  if (!kmcd->e3smBoxError.empty())
    return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": " + kmcd->e3smBoxError);
  if (nBatch < 0 || nPhotoRates < 0 || nPhotoRates > kmcd->e3.nReac)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": bad size");
  tchem_atmchem_options_t o;
  int rc = resolveOptions(opts, kmcd, &o, who);
  if (rc) return rc;
  if (nBatch == 0) return TCHEM_OK;
  if (!tol_newton || !tol_time || !tadv || !state || !t_out || !dt_out || !state_out ||
      (nPhotoRates > 0 && !photo_rates))
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": null device array");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if ((rc = ensureKf(kmcd, nBatch, s))) return rc;
  // rate constants from the E3SM kinetics (constant over the interval: they
  // depend on T, p, the photolysis rates and constant species only)
  hipError_t e = launchE3smKinetics(kmcd->e3, 3, nBatch, state, nPhotoRates ? photo_rates : nullptr, nPhotoRates,
                                    nullptr, kmcd->kf, kmcd->e3.nReac, s);
  if (e != hipSuccess) return hipFail(e, "e3sm_kinetics_kernel launch");
  IntegratorArgs a{};
  a.nBatch = nBatch;
  a.tol_newton = tol_newton;
  a.tol_time = tol_time;
  a.tadv = tadv;
  a.state = state;
  a.t_out = t_out;
  a.dt_out = dt_out;
  a.state_out = state_out;
  a.stats = stats;
  a.kf = kmcd->kf;
  a.ext = external_sources;
  return runIntegrator(kmcd, o, nullptr, 0, a, s, who, true);
}

int tchem_atmchem_update_time_advance(int nBatch, tchem_time_advance_t* tadv, const double* t_out,
                                      const double* dt_out, void* stream) {
  if (nBatch < 0 || (nBatch > 0 && (!tadv || !t_out || !dt_out)))
    return setError(TCHEM_ERR_INVALID, "update_time_advance: bad argument");
  hipError_t e = launchUpdateTadv(nBatch, tadv, t_out, dt_out, nullptr, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hipFail(e, "update_tadv_kernel");
  return TCHEM_OK;
}

int tchem_sourceterm_run_device_batch(int nBatch, const double* state, const double* photo_rates,
                                      const double* external_sources, double* omega, tchem_kmcd_t kmcd,
                                      void* stream) {
  ProfileRange profile_range("TChem::SourceTerm::runDeviceBatch");
  if (kmcd && !kmcd->cb05Kinetics)
    return setError(TCHEM_ERR_UNSUPPORTED, std::string("sourceterm: ") + kmcd->boxError);
  if (!kmcd || nBatch < 0) return setError(TCHEM_ERR_INVALID, "sourceterm: bad argument");
  if (nBatch == 0) return TCHEM_OK;
  if (!state || !omega) return setError(TCHEM_ERR_INVALID, "sourceterm: null device array");
  if (kmcd->dev.nPhoto == 0) photo_rates = nullptr;
  hipError_t e = launchSourceTerm(kmcd->dev, nBatch, state, photo_rates, external_sources, omega,
                                  static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hipFail(e, "sourceterm_kernel launch");
  return TCHEM_OK;
}

int tchem_jacobian_run_device_batch(int nBatch, const double* state, const double* photo_rates, double* jac,
                                    tchem_kmcd_t kmcd, void* stream) {
  ProfileRange profile_range("TChem::Jacobian::runDeviceBatch");
  if (kmcd && !kmcd->cb05Kinetics)
    return setError(TCHEM_ERR_UNSUPPORTED, std::string("jacobian: ") + kmcd->boxError);
  if (!kmcd || nBatch < 0) return setError(TCHEM_ERR_INVALID, "jacobian: bad argument");
  if (nBatch == 0) return TCHEM_OK;
  if (!state || !jac) return setError(TCHEM_ERR_INVALID, "jacobian: null device array");
  if (kmcd->dev.nPhoto == 0) photo_rates = nullptr;
  const size_t lds = integratorLdsBytes(kmcd->dev, sizeof(double));
  if (lds > kmcd->maxLds) return setError(TCHEM_ERR_UNSUPPORTED, "jacobian: mechanism too large for LDS");
  hipError_t e = launchJacobian(kmcd->dev, nBatch, state, photo_rates, jac, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hipFail(e, "jacobian_kernel launch");
  return TCHEM_OK;
}

int tchem_adjoint_workspace_bytes(tchem_kmcd_t kmcd, int nBatch, int maxSteps, size_t* bytes) {
  int rc = adjointCheck(kmcd, nBatch, maxSteps, "tchem_adjoint_workspace_bytes");
  if (rc) return rc;
  if (!bytes) return setError(TCHEM_ERR_INVALID, "tchem_adjoint_workspace_bytes: null bytes");
  *bytes = adjointLayout(kmcd->dev, nBatch, maxSteps).total;
  return TCHEM_OK;
}

int tchem_adjoint_forward(int nBatch, const double* tol_newton, const double* tol_time,
                          const tchem_time_advance_t* tadv, double* state, const double* photo_rates,
                          const double* log_rate_scale, int log_rate_scale_stride, double* t_out, double* dt_out,
                          tchem_cell_stats_t* stats, void* workspace, int maxSteps, tchem_kmcd_t kmcd,
                          void* stream) {
  ProfileRange profile_range("TChem::Adjoint::forward");
  const char* who = "tchem_adjoint_forward";
  int rc = adjointCheck(kmcd, nBatch, maxSteps, who);
  if (rc) return rc;
  if (log_rate_scale_stride != 0 && log_rate_scale_stride != kmcd->dev.nReac)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": log_rate_scale_stride must be 0 or nReac");
  if (nBatch == 0) return TCHEM_OK;
  if (!tol_newton || !tol_time || !tadv || !state || !t_out || !dt_out || !workspace)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": null device array");
  if (reinterpret_cast<uintptr_t>(workspace) % 256)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": workspace must be 256-byte aligned");
  if (kmcd->dev.nPhoto == 0) photo_rates = nullptr;
  tchem_atmchem_options_t req{};
  req.precision = TCHEM_PRECISION_FP64;
  req.schedule = TCHEM_SCHEDULE_PERSISTENT;
  req.lu = TCHEM_LU_SPARSE;
  tchem_atmchem_options_t o;
  rc = resolveOptions(&req, kmcd, &o, who);
  if (rc) return rc;
  const AdjointLayout W = adjointLayout(kmcd->dev, nBatch, maxSteps);
  char* ws = static_cast<char*>(workspace);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = launchAdjointSetup(nBatch, maxSteps, tadv, tol_newton, reinterpret_cast<int*>(ws + W.header),
                                    reinterpret_cast<double*>(ws + W.tol), reinterpret_cast<int*>(ws + W.newton), s);
  if (e != hipSuccess) return hipFail(e, "adjoint_setup_kernel");
  IntegratorArgs a{};
  a.nBatch = nBatch;
  a.tol_newton = tol_newton;
  a.tol_time = tol_time;
  a.tadv = tadv;
  a.state = state;
  a.photo = photo_rates;
  a.t_out = t_out;
  a.dt_out = dt_out;
  a.state_out = state;
  a.stats = stats;
  a.kf = reinterpret_cast<double*>(ws + W.kf);
  a.rec.nsteps = reinterpret_cast<int*>(ws + W.nsteps);
  a.rec.h = reinterpret_cast<double*>(ws + W.h);
  a.rec.ckpt = reinterpret_cast<double*>(ws + W.ckpt);
  a.rec.consts = reinterpret_cast<double*>(ws + W.consts);
  a.rec.maxSteps = maxSteps;
  return runIntegrator(kmcd, o, log_rate_scale, log_rate_scale_stride, a, s, who);
}

int tchem_adjoint_backward(int nBatch, const double* lambda, double* lambda_out, double* grad, void* workspace,
                           int maxSteps, tchem_kmcd_t kmcd, void* stream) {
  ProfileRange profile_range("TChem::Adjoint::backward");
  const char* who = "tchem_adjoint_backward";
  int rc = adjointCheck(kmcd, nBatch, maxSteps, who);
  if (rc) return rc;
  if (nBatch == 0) return TCHEM_OK;
  if (!lambda || !lambda_out || !grad || !workspace)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": null device array");
  if (reinterpret_cast<uintptr_t>(workspace) % 256)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": workspace must be 256-byte aligned");
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess || cur != kmcd->device)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": kmcd belongs to another device");
  const DevMech& M = kmcd->dev;
  const size_t cellB = adjointCellBytes(M, kmcd->sparse);
  int cpg = 8;
  while (cpg > 1 && size_t(cpg) * cellB > kmcd->maxLds) --cpg;
  if (cellB > kmcd->maxLds)
    return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": one cell needs " + std::to_string(cellB) +
                                               " bytes of LDS, device limit " + std::to_string(kmcd->maxLds));
  const size_t lds = size_t(cpg) * cellB;
  const AdjointLayout W = adjointLayout(M, nBatch, maxSteps);
  char* ws = static_cast<char*>(workspace);
  AdjointArgs a{};
  a.fw.M = M;
  a.fw.S = kmcd->sparse;
  a.fw.nBatch = nBatch;
  a.fw.kf = reinterpret_cast<const double*>(ws + W.kf);
  a.fw.queue = kmcd->queue;
  a.fw.persistent = 1;
  a.fw.sparse = 2;
  a.fw.cellsPerGroup = cpg;
  a.fw.cellBytes = int(cellB);
  a.fw.rec.nsteps = reinterpret_cast<int*>(ws + W.nsteps);
  a.fw.rec.h = reinterpret_cast<double*>(ws + W.h);
  a.fw.rec.ckpt = reinterpret_cast<double*>(ws + W.ckpt);
  a.fw.rec.consts = reinterpret_cast<double*>(ws + W.consts);
  a.fw.rec.maxSteps = maxSteps;
  a.ST = kmcd->sparseT;
  a.header = reinterpret_cast<const int*>(ws + W.header);
  a.tol = reinterpret_cast<const double*>(ws + W.tol);
  a.newton = reinterpret_cast<const int*>(ws + W.newton);
  a.traj = reinterpret_cast<double*>(ws + W.traj);
  a.lambda_in = lambda;
  a.lambda_out = lambda_out;
  a.grad = grad;
  a.cellsPerGroup = cpg;
  a.cellBytes = int(cellB);
  int bpc = 0;
  hipError_t e = adjointMaxBlocksPerCU(M.m, 64 * cpg, lds, &bpc);
  if (e != hipSuccess) return hipFail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  if (bpc < 1) bpc = 1;
  const int groups = (nBatch + cpg - 1) / cpg;
  const long long g = (long long)bpc * kmcd->numCUs;
  const int grid = int(g < groups ? g : groups);
  hipStream_t s = static_cast<hipStream_t>(stream);
  e = launchResetCounter(kmcd->queue, s);
  if (e != hipSuccess) return hipFail(e, "reset_counter_kernel");
  e = launchAdjoint(a, grid, lds, s);
  if (e != hipSuccess) return hipFail(e, "adjoint_kernel launch");
  return TCHEM_OK;
}

int tchem_adjoint_steps(int nBatch, const void* workspace, int maxSteps, int* nsteps, double* h, tchem_kmcd_t kmcd,
                        void* stream) {
  const char* who = "tchem_adjoint_steps";
  int rc = adjointCheck(kmcd, nBatch, maxSteps, who);
  if (rc) return rc;
  if (nBatch == 0) return TCHEM_OK;
  if (!workspace || !nsteps) return setError(TCHEM_ERR_INVALID, std::string(who) + ": null device array");
  const AdjointLayout W = adjointLayout(kmcd->dev, nBatch, maxSteps);
  const char* ws = static_cast<const char*>(workspace);
  hipStream_t s = static_cast<hipStream_t>(stream);
  // device copies: both arrays are the caller's
  hipError_t e = launchCopyWords(nsteps, ws + W.nsteps, size_t(nBatch), s);
  if (e == hipSuccess && h) e = launchCopyWords(h, ws + W.h, 2 * size_t(nBatch) * size_t(maxSteps), s);
  if (e != hipSuccess) return hipFail(e, "copy launch");
  return TCHEM_OK;
}

int tchem_atmchem_graph_create(const tchem_atmchem_options_t* opts, int nBatch, const double* tol_newton,
                               const double* tol_time, tchem_time_advance_t* tadv, double* state,
                               const double* photo_rates, double* t_out, double* dt_out, double* state_out,
                               tchem_cell_stats_t* stats, int nIntervals, double stopFraction, tchem_kmcd_t kmcd,
                               tchem_graph_t* out) {
  if (kmcd && !kmcd->cb05Kinetics)
    return setError(TCHEM_ERR_UNSUPPORTED, std::string("tchem_atmchem_graph_create: ") + kmcd->boxError);
  const char* who = "tchem_atmchem_graph_create";
  if (!kmcd || !out) return setError(TCHEM_ERR_INVALID, std::string(who) + ": null argument");
  *out = nullptr;
  if (nBatch < 1 || nIntervals < 1) return setError(TCHEM_ERR_INVALID, std::string(who) + ": nBatch and nIntervals must be >= 1");
  if (!tol_newton || !tol_time || !tadv || !state || !t_out || !dt_out || !state_out)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": null device array");
  tchem_atmchem_options_t o;
  int rc = resolveOptions(opts, kmcd, &o, who);
  if (rc) return rc;
  if (kmcd->dev.nPhoto == 0) photo_rates = nullptr;
  // everything that queries or allocates happens before the capture
  if (o.schedule == TCHEM_SCHEDULE_PERSISTENT) {
    int bpc = 0;
    if ((rc = resolveBlocksPerCU(kmcd, o, false, &bpc))) return rc;
    o.blocks_per_cu = bpc;
  }
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hipFail(e, "hipDeviceSynchronize");
  std::unique_ptr<tchem_graph_s> g(new tchem_graph_s);
  g->device = kmcd->device;
  g->intervals = nIntervals;
  {
    void* p = nullptr;
    e = hipMalloc(&p, 512);
    if (e != hipSuccess) return hipFail(e, "hipMalloc");
    g->ctl = static_cast<int*>(p);
    e = hipMalloc(&p, sizeof(double) * size_t(nBatch) * size_t(kmcd->dev.nReac > 0 ? kmcd->dev.nReac : 1));
    if (e != hipSuccess) return hipFail(e, "hipMalloc (rate constants)");
    g->kf = static_cast<double*>(p);
    e = hipMemset(g->ctl, 0, 512);  // outside the capture
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return hipFail(e, "hipMemset");
  }
  hipStream_t cs = nullptr;
  e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e != hipSuccess) return hipFail(e, "hipStreamCreateWithFlags");
  const bool stop = stopFraction > 0.0;
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(cs);
    return hipFail(e, "hipStreamBeginCapture");
  }
  e = launchLoopControl(nBatch, t_out, tadv, stopFraction, g->ctl, true, cs);
  if (e != hipSuccess) rc = hipFail(e, "loop_reset_kernel");
  for (int it = 0; it < nIntervals && rc == TCHEM_OK; ++it) {
    IntegratorArgs a{};
    a.skip = g->ctl;
    a.nBatch = nBatch;
    a.tol_newton = tol_newton;
    a.tol_time = tol_time;
    a.tadv = tadv;
    a.state = state;
    a.photo = photo_rates;
    a.t_out = t_out;
    a.dt_out = dt_out;
    a.state_out = state_out;
    a.stats = stats;
    a.kf = g->kf;
    a.queue = g->ctl + 64;
    rc = runIntegrator(kmcd, o, nullptr, 0, a, cs, who);
    if (rc == TCHEM_OK) {
      e = launchUpdateTadv(nBatch, tadv, t_out, dt_out, g->ctl, cs);
      if (e != hipSuccess) rc = hipFail(e, "update_tadv_kernel");
    }
    if (rc == TCHEM_OK && stop) {
      e = launchLoopControl(nBatch, t_out, tadv, stopFraction, g->ctl, false, cs);
      if (e != hipSuccess) rc = hipFail(e, "stop_check_kernel");
    }
  }
  hipError_t ee = hipStreamEndCapture(cs, &g->graph);
  (void)hipStreamDestroy(cs);
  if (rc) return rc;
  if (ee != hipSuccess) return hipFail(ee, "hipStreamEndCapture");
  e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) return hipFail(e, "hipGraphInstantiate");
  g->kmcd = kmcd;
  ++kmcd->graphs;
  *out = g.release();
  return TCHEM_OK;
}

int tchem_graph_launch(tchem_graph_t graph, void* stream) {
  ProfileRange profile_range("TChem::AtmosphericChemistry::graphLaunch");
  if (!graph || !graph->exec) return setError(TCHEM_ERR_INVALID, "tchem_graph_launch: null graph");
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess || cur != graph->device)
    return setError(TCHEM_ERR_INVALID, "tchem_graph_launch: graph belongs to another device");
  hipError_t e = hipGraphLaunch(graph->exec, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hipFail(e, "hipGraphLaunch");
  return TCHEM_OK;
}

int tchem_graph_status(tchem_graph_t graph, int* intervals_run, int* stopped) {
  if (!graph || !graph->ctl) return setError(TCHEM_ERR_INVALID, "tchem_graph_status: null graph");
  int h[2] = {0, 0};
  hipError_t e = hipDeviceSynchronize();  // the launch may be on any stream
  if (e == hipSuccess) e = hipMemcpy(h, graph->ctl, sizeof(h), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hipFail(e, "hipMemcpy");
  if (stopped) *stopped = h[0];
  if (intervals_run) *intervals_run = h[1];
  return TCHEM_OK;
}

int tchem_graph_destroy(tchem_graph_t graph) {
  delete graph;
  return TCHEM_OK;
}

int tchem_atmchem_autotune(tchem_atmchem_options_t* opts, int nBatch, const double* tol_newton,
                           const double* tol_time, const tchem_time_advance_t* tadv, const double* state,
                           const double* photo_rates, int sampleCells, int sampleIterations, double* best_seconds,
                           tchem_kmcd_t kmcd, void* stream) {
  ProfileRange profile_range("TChem::AtmosphericChemistry::autotune");
  if (kmcd && !kmcd->cb05Kinetics)
    return setError(TCHEM_ERR_UNSUPPORTED, std::string("tchem_atmchem_autotune: ") + kmcd->boxError);
  const char* who = "tchem_atmchem_autotune";
  if (!kmcd || !opts) return setError(TCHEM_ERR_INVALID, std::string(who) + ": null argument");
  if (nBatch < 1 || sampleCells < 0 || sampleIterations < 0)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": bad size");
  if (!tol_newton || !tol_time || !tadv || !state)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": null device array");
  if (kmcd->dev.nPhoto == 0) photo_rates = nullptr;
  tchem_atmchem_options_t base = *opts;
  base.cells_per_block = 0;
  base.blocks_per_cu = 0;
  const int luIn = base.lu;
  // candidate LU forms
  std::vector<int> lus;
  if (luIn == TCHEM_LU_AUTO || luIn == TCHEM_LU_SPARSE) {
    tchem_atmchem_options_t t = base;
    t.lu = TCHEM_LU_SPARSE;
    tchem_atmchem_options_t r;
    if (resolveOptions(&t, kmcd, &r, who) == TCHEM_OK) lus.push_back(TCHEM_LU_SPARSE);
    else if (luIn == TCHEM_LU_SPARSE) return TCHEM_ERR_UNSUPPORTED;
  }
  for (int lu : {TCHEM_LU_DENSE, TCHEM_LU_MFMA}) {
    if (luIn != TCHEM_LU_AUTO && luIn != lu) continue;
    if (luIn == TCHEM_LU_AUTO && lu == TCHEM_LU_MFMA && kmcd->dev.m < 128) continue;  // tiles mostly padding
    tchem_atmchem_options_t t = base;
    t.lu = lu;
    tchem_atmchem_options_t r;
    if (resolveOptions(&t, kmcd, &r, who) == TCHEM_OK) lus.push_back(lu);
    else if (luIn == lu) return TCHEM_ERR_UNSUPPORTED;
  }
  {
    tchem_atmchem_options_t r;
    int rc = resolveOptions(&base, kmcd, &r, who);  // validates the remaining fields
    if (rc) return rc;
  }
  const int nb = kmcd->dev.nSpec + 3, np = kmcd->dev.nPhoto;
  const int n = sampleCells > 0 ? std::min(sampleCells, nBatch) : std::min(nBatch, kmcd->numCUs * 48);
  const int iters = sampleIterations > 0 ? sampleIterations : 24;
  const size_t stride = size_t(nBatch / n);  // sample rows 0, stride, 2 stride, ...
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc = ensureKf(kmcd, n, s);
  if (rc) return rc;
  // scratch: pristine sample rows, working rows, tadv, t/dt
  std::vector<void*> bufs;
  auto alloc = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return nullptr;
    bufs.push_back(p);
    return p;
  };
  auto freeAll = [&]() {
    (void)hipStreamSynchronize(s);
    for (void* p : bufs) (void)hipFree(p);
  };
  double* st0 = static_cast<double*>(alloc(sizeof(double) * size_t(n) * nb));
  double* st = static_cast<double*>(alloc(sizeof(double) * size_t(n) * nb));
  double* ph = (np && photo_rates) ? static_cast<double*>(alloc(sizeof(double) * size_t(n) * np)) : nullptr;
  tchem_time_advance_t* ta = static_cast<tchem_time_advance_t*>(alloc(sizeof(tchem_time_advance_t) * size_t(n)));
  double* tt = static_cast<double*>(alloc(sizeof(double) * size_t(n) * 2));
  if (!st0 || !st || (np && photo_rates && !ph) || !ta || !tt) {
    freeAll();
    return setError(TCHEM_ERR_HIP, std::string(who) + ": scratch allocation failed");
  }
  // the sample, extracted on the device (tt0: t_out = tbeg, dt_out = 0)
  double* tt0 = static_cast<double*>(alloc(sizeof(double) * size_t(n) * 2));
  int* live = static_cast<int*>(alloc(sizeof(int)));
  if (!tt0 || !live) {
    freeAll();
    return setError(TCHEM_ERR_HIP, std::string(who) + ": scratch allocation failed");
  }
  int hlive = 0;
  hipError_t e = hipMemcpy(live, &hlive, sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess)
    e = launchAutotuneSample(n, (long long)stride, nb, ph ? np : 0, iters, state, ph ? photo_rates : nullptr, tadv, st0,
                             ph, ta, tt0, live, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = hipMemcpy(&hlive, live, sizeof(int), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    freeAll();
    return hipFail(e, "sample copy");
  }
  if (hlive == 0) {
    freeAll();
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": no cell of the sample has time left to integrate");
  }
  auto trial = [&](const tchem_atmchem_options_t& o, float* ms) -> int {
    hipError_t ex = hipMemcpyAsync(st, st0, sizeof(double) * size_t(n) * nb, hipMemcpyDeviceToDevice, s);
    if (ex == hipSuccess) ex = hipMemcpyAsync(tt, tt0, sizeof(double) * size_t(n) * 2, hipMemcpyDeviceToDevice, s);
    if (ex != hipSuccess) return hipFail(ex, "trial reset");
    IntegratorArgs a{};
    a.nBatch = n;
    a.tol_newton = tol_newton;
    a.tol_time = tol_time;
    a.tadv = ta;
    a.state = st;
    a.photo = ph;
    a.t_out = tt;
    a.dt_out = tt + n;
    a.state_out = st;
    a.kf = kmcd->kf;
    // host clock around a synchronised launch: robust where events are not
    // (under some profilers they complete early); a trial runs for ms
    if ((ex = hipStreamSynchronize(s)) != hipSuccess) return hipFail(ex, "hipStreamSynchronize");
    const auto t0 = std::chrono::steady_clock::now();
    const int r = runIntegrator(kmcd, o, nullptr, 0, a, s, who);
    if (r) return r;
    if ((ex = hipStreamSynchronize(s)) != hipSuccess) return hipFail(ex, "hipStreamSynchronize");
    *ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return TCHEM_OK;
  };
  float best = -1.0f;
  tchem_atmchem_options_t bestO{};
  const char* venv = std::getenv("TCHEM_AUTOTUNE_VERBOSE");  // trial timings on stderr
  const bool verbose = venv && *venv && *venv != '0';
  if (verbose) {
    std::fprintf(stderr, "autotune: %d sample cells, %d iterations, candidate LU forms:", n, iters);
    for (int lu : lus) std::fprintf(stderr, " %d", lu);
    std::fprintf(stderr, "\n");
  }
  // (lu, programs placement) candidates; placement only matters for the sparse form
  std::vector<std::pair<int, int>> forms;
  for (int lu : lus) {
    if (lu == TCHEM_LU_SPARSE && base.programs == TCHEM_PROGRAMS_AUTO) {
      forms.push_back({lu, TCHEM_PROGRAMS_LDS});
      forms.push_back({lu, TCHEM_PROGRAMS_MEMORY});
    } else {
      forms.push_back({lu, base.programs});
    }
  }
  for (const auto& form : forms) {
    tchem_atmchem_options_t t = base;
    t.lu = form.first;
    t.programs = form.second;
    tchem_atmchem_options_t r;
    if (resolveOptions(&t, kmcd, &r, who) != TCHEM_OK) continue;  // e.g. programs too large for LDS
    const int cmax = r.cells_per_block;  // as many as fit
    for (int c = cmax; c >= 1 && rc == TCHEM_OK; --c) {
      tchem_atmchem_options_t oc = r;
      oc.cells_per_block = c;
      int bmax = 1;
      if (oc.schedule == TCHEM_SCHEDULE_PERSISTENT && (rc = resolveBlocksPerCU(kmcd, oc, false, &bmax))) break;
      for (int b = bmax; b >= 1 && b >= bmax - 1; --b) {
        oc.blocks_per_cu = oc.schedule == TCHEM_SCHEDULE_PERSISTENT ? b : 0;
        float ms = 0.0f;
        if ((rc = trial(oc, &ms)) || (rc = trial(oc, &ms))) break;  // first run warms the kernel instance
        if (verbose)
          std::fprintf(stderr, "autotune: lu %d programs %d cells %d blocks/CU %d: %.3f ms\n", oc.lu, oc.programs,
                       oc.cells_per_block, oc.blocks_per_cu, double(ms));
        if (best < 0.0f || ms < best) {  // ms >= 0: host clock
          best = ms;
          bestO = oc;
        }
        if (oc.schedule != TCHEM_SCHEDULE_PERSISTENT) break;
      }
    }
    if (rc) break;
  }
  freeAll();
  if (rc) return rc;
  if (best < 0.0f) return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": no launch configuration fits");
  *opts = bestO;
  if (best_seconds) *best_seconds = double(best) * 1e-3;
  return TCHEM_OK;
}

int tchem_scenario_photolysis_rates(tchem_kmd_t kmd, const char* path, int nBatch, double* rates_host,
                                    int capacity_cols, int* count) {
  if (!kmd || !path || !count || nBatch < 1)
    return setError(TCHEM_ERR_INVALID, "tchem_scenario_photolysis_rates: bad argument");
  try {
    int n = 0;
    const std::vector<double> r = loadScenarioPhotolysisRates(path, nBatch, &n);
    *count = n;
    if (rates_host) {
      if (capacity_cols < n)
        return setError(TCHEM_ERR_INVALID, "tchem_scenario_photolysis_rates: capacity below " + std::to_string(n));
      for (int i = 0; i < nBatch; ++i)
        for (int j = 0; j < n; ++j) rates_host[size_t(i) * capacity_cols + j] = r[size_t(i) * n + j];
    }
  } catch (const std::exception& e) {
    return setError(TCHEM_ERR_PARSE, e.what());
  }
  return TCHEM_OK;
}

int tchem_scenario_external_forcing(tchem_kmd_t kmd, const char* path, int nBatch, double* forcing_host, int* count) {
  if (!kmd || !path || !forcing_host || nBatch < 1)
    return setError(TCHEM_ERR_INVALID, "tchem_scenario_external_forcing: bad argument");
  try {
    int n = 0;
    const std::vector<double> f = loadScenarioExternalForcing(path, kmd->km, nBatch, &n);
    std::copy(f.begin(), f.end(), forcing_host);
    if (count) *count = n;
  } catch (const std::exception& e) {
    return setError(TCHEM_ERR_PARSE, e.what());
  }
  return TCHEM_OK;
}

namespace {
int runE3sm(const char* who, int mode, int nBatch, const double* state, const double* photo, int nPhoto,
            const double* ext, double* out, int ldOut, tchem_kmcd_t kmcd, void* stream) {
  if (!kmcd || nBatch < 0 || nPhoto < 0) return setError(TCHEM_ERR_INVALID, std::string(who) + ": bad argument");
  if (nBatch == 0) return TCHEM_OK;
  if (!state || !out || (nPhoto > 0 && !photo))
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": null device array");
  const DevE3sm& E = kmcd->e3;
  if (nPhoto > E.nReac) return setError(TCHEM_ERR_INVALID, std::string(who) + ": more photolysis rates than reactions");
  const size_t lds = sizeof(double) * (size_t(E.nSpec) + 2 * size_t(E.nReac) + size_t(E.nSpec - E.nConst));
  if (lds > kmcd->maxLds) return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": mechanism too large for LDS");
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess || cur != kmcd->device)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": kmcd belongs to another device");
  hipError_t e = launchE3smKinetics(E, mode, nBatch, state, nPhoto ? photo : nullptr, nPhoto, ext, out, ldOut,
                                    static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hipFail(e, "e3sm_kinetics_kernel launch");
  return TCHEM_OK;
}
}  // namespace

int tchem_reactionrates_run_device_batch(int nBatch, const double* state, double* kfwd, tchem_kmcd_t kmcd,
                                         void* stream) {
  return runE3sm("reactionrates", 0, nBatch, state, nullptr, 0, nullptr, kfwd, kmcd ? kmcd->e3.nReac : 0, kmcd,
                 stream);
}

int tchem_rateofprogress_run_device_batch(int nBatch, const double* state, const double* photo_rates,
                                          int nPhotoRates, double* rop, tchem_kmcd_t kmcd, void* stream) {
  return runE3sm("rateofprogress", 1, nBatch, state, photo_rates, nPhotoRates, nullptr, rop,
                 kmcd ? kmcd->e3.nReac : 0, kmcd, stream);
}

int tchem_netproductionrates_run_device_batch(int nBatch, const double* state, const double* photo_rates,
                                              int nPhotoRates, const double* external_sources, double* omega,
                                              int ldOmega, tchem_kmcd_t kmcd, void* stream) {
  if (kmcd && ldOmega < kmcd->e3.nSpec - kmcd->e3.nConst)
    return setError(TCHEM_ERR_INVALID, "netproductionrates: ldOmega below the number of active species");
  return runE3sm("netproductionrates", 2, nBatch, state, photo_rates, nPhotoRates, external_sources, omega, ldOmega,
                 kmcd, stream);
}

}  // extern "C"

// ---- single-particle SIMPOL.1 aerosol model ----------------------------------

int tchem_amd_load_yaml(const char* path, tchem_kmd_t kmd, tchem_amd_t* out) {
  if (!path || !kmd || !out) return setError(TCHEM_ERR_INVALID, "tchem_amd_load_yaml: null argument");
  *out = nullptr;
  try {
    std::unique_ptr<tchem_amd_s> h(new tchem_amd_s);
    h->am = loadAerosolModel(path, kmd->km);
    h->km = kmd->km;
    *out = h.release();
    return TCHEM_OK;
  } catch (const std::exception& e) {
    return setError(TCHEM_ERR_PARSE, e.what());
  }
}

int tchem_amd_destroy(tchem_amd_t amd) {
  delete amd;
  return TCHEM_OK;
}

int tchem_amd_info(tchem_amd_t amd, int* nSpec, int* nParticles, int* nSimpol) {
  if (!amd) return setError(TCHEM_ERR_INVALID, "tchem_amd_info: null handle");
  if (nSpec) *nSpec = amd->am.nSpec;
  if (nParticles) *nParticles = amd->am.nParticles;
  if (nSimpol) *nSimpol = int(amd->am.simpol.size());
  return TCHEM_OK;
}

int tchem_amd_set_number_of_particles(tchem_amd_t amd, int nParticles) {
  if (!amd || nParticles < 0) return setError(TCHEM_ERR_INVALID, "tchem_amd_set_number_of_particles: bad argument");
  amd->am.nParticles = nParticles;
  return TCHEM_OK;
}

int tchem_amd_species_name(tchem_amd_t amd, int i, char* buf, size_t len) {
  if (!amd || !buf || len == 0) return setError(TCHEM_ERR_INVALID, "tchem_amd_species_name: bad argument");
  if (i < 0 || i >= amd->am.nSpec) return setError(TCHEM_ERR_INVALID, "tchem_amd_species_name: index out of range");
  std::strncpy(buf, amd->am.names[i].c_str(), len - 1);
  buf[len - 1] = '\0';
  return TCHEM_OK;
}

int tchem_amd_species_props(tchem_amd_t amd, int i, double* density, double* mw, double* kappa) {
  if (!amd) return setError(TCHEM_ERR_INVALID, "tchem_amd_species_props: null handle");
  if (i < 0 || i >= amd->am.nSpec) return setError(TCHEM_ERR_INVALID, "tchem_amd_species_props: index out of range");
  if (density) *density = amd->am.density[i];
  if (mw) *mw = amd->am.mw[i];
  if (kappa) *kappa = amd->am.kappa[i];
  return TCHEM_OK;
}

int tchem_scenario_particles(tchem_amd_t amd, const char* path, int nBatch, int stateDim, double* state_host,
                             double* number_conc_host, int* found) {
  const char* who = "tchem_scenario_particles";
  if (!amd || !path || nBatch < 0 || !found) return setError(TCHEM_ERR_INVALID, std::string(who) + ": bad argument");
  if (stateDim != amd->am.stateDim(amd->km))
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": stateDim differs from the model's state length");
  if (nBatch > 0 && (!state_host || (amd->am.nParticles > 0 && !number_conc_host)))
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": null host array");
  try {
    std::vector<double> st(state_host, state_host + size_t(nBatch) * size_t(stateDim));
    std::vector<double> nc;
    *found = loadScenarioParticles(path, amd->am, amd->km, nBatch, st, nc) ? 1 : 0;
    if (*found) {
      std::copy(st.begin(), st.end(), state_host);
      std::copy(nc.begin(), nc.end(), number_conc_host);
    }
    return TCHEM_OK;
  } catch (const std::exception& e) {
    return setError(TCHEM_ERR_PARSE, e.what());
  }
}

int tchem_amcd_create(tchem_amd_t amd, tchem_kmcd_t kmcd, tchem_amcd_t* out) {
  const char* who = "tchem_amcd_create";
  if (!amd || !kmcd || !out) return setError(TCHEM_ERR_INVALID, std::string(who) + ": null argument");
  *out = nullptr;
  if (!kmcd->cb05Kinetics)
    return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": gas mechanism: " + kmcd->boxError);
  const AerosolModel& am = amd->am;
  const MechTables& t = kmcd->tables;
  if (am.nGasActive != t.m || amd->km.nSpec != t.nSpec)
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": the aerosol model was read against another gas model");
  const int mt = t.m + am.nSpec * am.nParticles;
  if (mt > 256) return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": more than 256 unknowns per cell");
  std::unique_ptr<tchem_amcd_s> h(new tchem_amcd_s);
  h->kmcd = kmcd;
  h->device = kmcd->device;
  h->stateDim = am.stateDim(amd->km);
  DevAerosol& X = h->X;
  X.nAero = am.nSpec;
  X.nPart = am.nParticles;
  X.nSimpol = int(am.simpol.size());
  X.mt = mt;
  std::vector<int> ib;
  std::vector<double> db(am.mw.begin(), am.mw.end());
  X.oMw = 0;
  X.oDens = int(db.size());
  db.insert(db.end(), am.density.begin(), am.density.end());
  X.oPar = int(db.size());
  for (const SimpolParam& s : am.simpol) {
    ib.push_back(t.iperm[s.gas]);  // gas unknowns are in the internal order
    const double par[8] = {s.B1, s.B2, s.B3, s.B4, s.diffusion, s.mw, s.Nstar, double(s.computeAlpha)};
    db.insert(db.end(), par, par + 8);
  }
  X.oGas = 0;
  X.oAero = int(ib.size());
  for (const SimpolParam& s : am.simpol) ib.push_back(s.aero);
  // sparse form: the gas Jacobian pattern plus every transfer derivative
  // over mt unknowns, particle unknowns after the gas species
  std::vector<int> pib;
  std::vector<double> pdb;
  try {
    std::vector<std::array<int, 2>> ent;  // (row, col) of every transfer derivative, table order
    const MechTables ta = aerosolTables(t, am, &ent);
    h->sched = buildSparseSchedule(ta, false);
    const SparseSchedule& sc = h->sched;
    auto position = [&](int r, int c) {  // the LDS slot of factor entry (r, c)
      for (int q = sc.rowPtr[r]; q < sc.rowPtr[r + 1]; ++q)
        if (sc.colIdx[q] == c) return sc.slotOf[size_t(q)];
      throw std::runtime_error("transfer derivative outside the factor pattern");
    };
    X.oJpos = int(ib.size());
    for (const auto& e : ent) ib.push_back(position(e[0], e[1]));
    auto addI = [&](const std::vector<unsigned>& v) {
      const int off = int(pib.size());
      pib.insert(pib.end(), v.begin(), v.end());
      while (pib.size() % 4) pib.push_back(0);
      return off;
    };
    DevSparse& sp = h->S;
    sp.enabled = 1;
    sp.nnzF = sc.nnzF;
    sp.zeroSlot = sc.zeroSlot;
    sp.trashSlot = sc.trashSlot;
    sp.pairSlots = sc.pairSlots;
    sp.oEvalHdr = addI(sc.evalHdr);
    sp.oEvalBody = addI(sc.evalBody);
    sp.evalBodyWords = int(sc.evalBody.size());
    sp.oCoef = 0;
    pdb = sc.coefTable;
    sp.nCoef = int(sc.coefTable.size());
    sp.ropRounds = sc.ropRounds;
    sp.pairRounds = sc.pairRounds;
    sp.segRounds = sc.segRounds;
    sp.jacRounds = sc.jacRounds;
    sp.pairBody = sc.pairBody;
    sp.segBody = sc.segBody;
    sp.jacBody = sc.jacBody;
    sp.oProgHdr = addI(sc.progHdr);
    sp.oProgBody = addI(sc.progBody);
    sp.luRounds = sc.luRounds;
    sp.backRounds = sc.backRounds;
    sp.fwdRounds = sc.fwdRounds;
    sp.backBody = sc.backBody;
    sp.fwdBody = sc.fwdBody;
  } catch (const std::exception&) {
    h->S = DevSparse{};  // the dense form still runs
    pib.clear();
    pdb.clear();
  }
  ib.push_back(0);
  db.push_back(0.0);
  void* pi = nullptr;
  void* pd = nullptr;
  hipError_t e = hipMalloc(&pi, sizeof(int) * ib.size());
  if (e != hipSuccess) return hipFail(e, "hipMalloc");
  h->allocations.push_back(pi);
  e = hipMalloc(&pd, sizeof(double) * db.size());
  if (e != hipSuccess) return hipFail(e, "hipMalloc");
  h->allocations.push_back(pd);
  if ((e = hipMemcpy(pi, ib.data(), sizeof(int) * ib.size(), hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(pd, db.data(), sizeof(double) * db.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return hipFail(e, "hipMemcpy");
  X.ib = static_cast<const int*>(pi);
  X.db = static_cast<const double*>(pd);
  if (h->S.enabled) {
    void* qi = nullptr;
    void* qd = nullptr;
    if ((e = hipMalloc(&qi, sizeof(int) * pib.size())) != hipSuccess) return hipFail(e, "hipMalloc");
    h->allocations.push_back(qi);
    if ((e = hipMalloc(&qd, sizeof(double) * pdb.size())) != hipSuccess) return hipFail(e, "hipMalloc");
    h->allocations.push_back(qd);
    if ((e = hipMemcpy(qi, pib.data(), sizeof(int) * pib.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(qd, pdb.data(), sizeof(double) * pdb.size(), hipMemcpyHostToDevice)) != hipSuccess)
      return hipFail(e, "hipMemcpy");
    X.pib = static_cast<const int*>(qi);
    X.pdb = static_cast<const double*>(qd);
  }
  const bool denseFits = aerosolCellsPerGroup(kmcd->dev, X, kmcd->maxLds) >= 1;
  const bool sparseFits = h->S.enabled && aerosolSparseCellBytes(kmcd->dev, h->S, X) <= kmcd->maxLds;
  if (!sparseFits) h->S = DevSparse{};
  if (!denseFits && !sparseFits)
    return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": one cell does not fit the LDS");
  *out = h.release();
  return TCHEM_OK;
}

int tchem_amcd_destroy(tchem_amcd_t amcd) {
  delete amcd;
  return TCHEM_OK;
}

namespace {
IntegratorArgs aerosolArgs(tchem_amcd_s* h, int nBatch) {
  IntegratorArgs a{};
  a.M = h->kmcd->dev;
  a.S = h->kmcd->sparse;
  a.X = h->X;
  a.nBatch = nBatch;
  a.cellBytes = int((aerosolLdsBytes(a.M, a.X) + 15) & ~size_t(15));
  return a;
}
}  // namespace

int tchem_aerosol_rhs_run_device_batch(int nBatch, const double* state, const double* number_conc, double* omega,
                                       tchem_amcd_t amcd, void* stream) {
  ProfileRange profile_range("TChem::AerosolRHS::runDeviceBatch");
  const char* who = "aerosol_rhs_run_device_batch";
  if (!amcd || nBatch < 0) return setError(TCHEM_ERR_INVALID, std::string(who) + ": bad argument");
  if (nBatch == 0) return TCHEM_OK;
  if (!state || !omega || (amcd->X.nPart > 0 && !number_conc))
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": null device array");
  IntegratorArgs a = aerosolArgs(amcd, nBatch);
  a.state = const_cast<double*>(state);
  a.numberConc = number_conc;
  a.state_out = omega;  // the kernel's output rows [nBatch][mt]
  const size_t lds = (aerosolLdsBytes(a.M, a.X, false) + 15) & ~size_t(15);
  hipError_t e = launchAerosolRhs(a, lds, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hipFail(e, "aerosol_rhs_kernel launch");
  return TCHEM_OK;
}

int tchem_aerosolchem_run_device_batch(const tchem_atmchem_options_t* opts, int nBatch, const double* tol_newton,
                                       const double* tol_time, double* fac, const tchem_time_advance_t* tadv,
                                       double* state, const double* number_conc, double* t_out, double* dt_out,
                                       double* state_out, tchem_cell_stats_t* stats, tchem_amcd_t amcd,
                                       void* stream) {
  ProfileRange profile_range("TChem::AerosolChemistry::runDeviceBatch");
  (void)fac;
  const char* who = "aerosolchem_run_device_batch";
  if (!amcd || nBatch < 0) return setError(TCHEM_ERR_INVALID, std::string(who) + ": bad argument");
  if (opts && opts->precision != TCHEM_PRECISION_FP64)
    return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": the aerosol integrator is FP64 only");
  if (opts && opts->lu == TCHEM_LU_SPARSE && !amcd->S.enabled)
    return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": no sparse programs for this model");
  const bool sparse = amcd->S.enabled && !(opts && opts->lu == TCHEM_LU_DENSE);
  if (!sparse && aerosolCellsPerGroup(amcd->kmcd->dev, amcd->X, amcd->kmcd->maxLds) < 1)
    return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": the dense form needs at most 128 unknowns");
  if (nBatch == 0) return TCHEM_OK;
  if (!tol_newton || !tol_time || !tadv || !state || !t_out || !dt_out || !state_out ||
      (amcd->X.nPart > 0 && !number_conc))
    return setError(TCHEM_ERR_INVALID, std::string(who) + ": null device array");
  tchem_kmcd_s* k = amcd->kmcd;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc = ensureKf(k, nBatch, s);
  if (rc) return rc;
  hipError_t e = launchRates(k->dev, nBatch, state, nullptr, nullptr, 0, k->kf, nullptr, s, amcd->stateDim);
  if (e != hipSuccess) return hipFail(e, "rates_kernel launch");
  IntegratorArgs a = aerosolArgs(amcd, nBatch);
  a.tol_newton = tol_newton;
  a.tol_time = tol_time;
  a.tadv = tadv;
  a.state = state;
  a.t_out = t_out;
  a.dt_out = dt_out;
  a.state_out = state_out;
  a.stats = stats;
  a.kf = k->kf;
  a.numberConc = number_conc;
  int cells = aerosolCellsPerGroup(a.M, a.X, k->maxLds);
  if (sparse) {
    a.sparse = 1;
    a.S = amcd->S;
    a.cellBytes = int(aerosolSparseCellBytes(a.M, a.S, a.X));
    cells = int(k->maxLds / size_t(a.cellBytes));
    if (cells > 8) cells = 8;
  }
  if (cells < 1) return setError(TCHEM_ERR_UNSUPPORTED, std::string(who) + ": one cell does not fit the LDS");
  const int perGroup = cells < nBatch ? cells : nBatch;
  a.cellsPerGroup = perGroup;
  const int grid = (nBatch + perGroup - 1) / perGroup;
  e = launchAerosol(a, grid, size_t(perGroup) * size_t(a.cellBytes), s);
  if (e != hipSuccess) return hipFail(e, "aerosol integrator launch");
  return TCHEM_OK;
}
