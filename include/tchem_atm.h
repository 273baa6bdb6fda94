/* C ABI of the MI355X-native batched gas-phase kinetics library.
 *
 * Every entry point replaces one piece of the reference's C++ surface; the
 * reference interface is cited beside each declaration (paths relative to
 * sandialabs/TChem-atm).  Conventions:
 *   - plain pointers and sizes only; `stream` is a hipStream_t passed as
 *     void* (NULL = the default stream);
 *   - device pointers are marked [device]; all device arrays are row-major
 *     FP64 unless stated (the reference's LayoutRight views);
 *   - every function returns TCHEM_OK (0) or a negative TCHEM_ERR_* code,
 *     never throws and never exits; tchem_last_error() gives the message;
 *   - batch calls enqueue on `stream` and return without synchronising
 *     (the reference's Kokkos::parallel_for is asynchronous too, and its
 *     driver fences: src/examples/TChem_AtmosphericChemistry.cpp:317).
 */
#ifndef TCHEM_ATM_H
#define TCHEM_ATM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TCHEM_OK 0
#define TCHEM_ERR_INVALID -1  /* bad argument or size */
#define TCHEM_ERR_PARSE -2    /* mechanism / scenario file could not be read */
#define TCHEM_ERR_UNSUPPORTED -3
#define TCHEM_ERR_HIP -4      /* a HIP runtime call failed */
#define TCHEM_ERR_NO_DEVICE -5

typedef struct tchem_kmd_s* tchem_kmd_t;   /* host kinetic model data */
typedef struct tchem_kmcd_s* tchem_kmcd_t; /* device-resident constant data */
typedef struct tchem_graph_s* tchem_graph_t; /* captured launch sequence (HIP graph) */
typedef struct tchem_amd_s* tchem_amd_t;   /* host aerosol model data */
typedef struct tchem_amcd_s* tchem_amcd_t; /* device-resident aerosol constant data */

/* Byte layout of TChem::TimeAdvance (src/core/TChem_Util.hpp:469-476). */
typedef struct {
  double _tbeg, _tend;
  double _dt, _dtmin, _dtmax;
  int _max_num_newton_iterations;
  int _num_time_iterations_per_interval;
  int _num_outer_time_iterations_per_interval;
  int _jacobian_interval;
} tchem_time_advance_t;

/* Per-cell counters (no counterpart in the reference; the reference only
 * prints a warning for failed cells, src/core/problems/TChem_Impl_AtmosphericChemistry.hpp:213-221). */
typedef struct {
  int accepted, rejected, newton_failures, newton_iterations, factorizations, failed;
  unsigned long long t_start, t_end; /* device real-time counter (100 MHz) when the cell started and ended in
                                        this launch (0 for skipped cells) */
} tchem_cell_stats_t;

/* Kernel form and launch shape of tchem_atmchem_run_device_batch. */
#define TCHEM_PRECISION_FP64 0  /* all-FP64 ("plain") */
#define TCHEM_PRECISION_MIXED 1 /* FP32 iteration matrix + LU, FP64 state/residual/error/step control */
#define TCHEM_SCHEDULE_GRID 0       /* one workgroup per cell, grid = nBatch */
#define TCHEM_SCHEDULE_PERSISTENT 1 /* resident workgroups pull cells from a device work queue */
#define TCHEM_LU_AUTO 0   /* sparse when the mechanism has a sparse schedule, dense otherwise */
#define TCHEM_LU_DENSE 1  /* dense LU with partial pivoting (the oracle's arithmetic) */
#define TCHEM_LU_SPARSE 2 /* level-scheduled sparse LU, pivots fixed by the fill-reducing order */
#define TCHEM_LU_MFMA 3   /* dense FP32 LU on the matrix cores, pivots fixed by the fill-reducing order: one cell
                             per workgroup of 4 wavefronts; mixed precision, m <= 255 */
#define TCHEM_PROGRAMS_AUTO 0   /* LDS when they leave room for 4 cells, device memory otherwise */
#define TCHEM_PROGRAMS_LDS 1    /* staged into LDS once per workgroup, shared by its cells */
#define TCHEM_PROGRAMS_MEMORY 2 /* read from device memory (L2-resident); the LDS holds the cells and the
                                   coefficient table */
typedef struct {
  int precision;         /* TCHEM_PRECISION_* */
  int schedule;          /* TCHEM_SCHEDULE_* */
  int blocks_per_cu;     /* persistent schedule: resident workgroups per CU (0 = from occupancy) */
  int lu;                /* TCHEM_LU_* */
  int cells_per_block;   /* sparse form: cells (wavefronts) per workgroup sharing the LDS-resident
                            mechanism program (0 = as many as fit); the dense form runs 1 */
  int programs;          /* sparse form: where the compiled mechanism programs live, TCHEM_PROGRAMS_* */
  int reserved[2];       /* must be zero */
} tchem_atmchem_options_t;

/* Last error message of the calling thread ("" when none). */
const char* tchem_last_error(void);

/* ---- kinetic model (KineticModelData) ----------------------------------- */

/* Replaces KineticModelData::KineticModelData(mechfile[, echofile])
 * (src/core/TChem_KineticModelData.cpp:53-78, initChemNCAR :80-1017).
 * echo_path: nullable; when given, the loader echo ("kmod.echo") is written there. */
int tchem_kmd_load_yaml(const char* path, const char* echo_path, tchem_kmd_t* out);
int tchem_kmd_destroy(tchem_kmd_t kmd);

/* Sizes: kmcd.nSpec, kmcd.nConstSpec, kmcd.nReac (src/core/TChem_KineticModelNCAR_ConstData.hpp:30-113)
 * and the number of photolysis reactions (ARRHENIUS reactions whose note
 * starts with "PHOTOLYSIS"). Any output pointer may be NULL. */
int tchem_kmd_info(tchem_kmd_t kmd, int* nSpec, int* nConstSpec, int* nReac, int* nPhoto);

/* Species name i (the reference's kmcd.speciesNames, 18 characters max). */
int tchem_kmd_species_name(tchem_kmd_t kmd, int i, char* buf, size_t len);

/* Reaction indices of the photolysis slots, in slot order (nPhoto entries). */
int tchem_kmd_photolysis_reactions(tchem_kmd_t kmd, int* reactions, int capacity);

/* No counterpart in the reference (introspection of the compiled tables):
 * the elimination order the device LU forms factor the Newton systems in,
 * order[position] = active species index (m = nSpec - nConstSpec entries),
 * a greedy minimum-Markowitz order of the structure of I - J.  Host only. */
int tchem_kmd_elimination_order(tchem_kmd_t kmd, int* order, int capacity);

/* Replaces AtmChemistry::setScenarioConditions (src/core/TChem_AtmosphericChemistry.cpp:429-504):
 * reads a scenario YAML into host state rows [nBatch][nSpec+3].  Call with
 * state_host == NULL to query nBatch. */
int tchem_scenario_load(tchem_kmd_t kmd, const char* path, double* state_host, int capacity_rows, int* nBatch);

/* ---- device constant data (KineticModelNCAR_ConstData) ------------------ */

/* Replaces createNCAR_KineticModelConstData<device>(kmd)
 * (src/core/TChem_KineticModelNCAR_ConstData.hpp:115-148): compiles the
 * mechanism into device tables on the current HIP device. The kmcd is
 * independent of the kmd afterwards.  tchem_kmcd_destroy returns
 * TCHEM_ERR_INVALID (and destroys nothing) while captured graphs made from
 * the kmcd exist: they read its tables. */
int tchem_kmcd_create(tchem_kmd_t kmd, tchem_kmcd_t* out);
int tchem_kmcd_destroy(tchem_kmcd_t kmcd);

/* Replaces AtmosphericChemistry::getWorkSpaceSize (src/core/TChem_AtmosphericChemistry.hpp:47-81):
 * LDS bytes of one workgroup in the form opts selects (NULL: defaults): its cells_per_block cell regions plus
 * what the workgroup stages once for its cells (the sparse form's programs, or with the programs in device
 * memory its coefficient table). */
int tchem_atmchem_workspace_bytes(tchem_kmcd_t kmcd, const tchem_atmchem_options_t* opts, size_t* bytes);

/* ---- batched kernels ------------------------------------------------------ */

/* Replaces AtmosphericChemistry::runDeviceBatch(policy, tol_newton, tol_time,
 * fac, tadv, state, t_out, dt_out, state_out, kmcd)
 * (src/core/TChem_AtmosphericChemistry.hpp:83-97, .cpp:210-271) and its
 * per-cell-kmcd overload (.hpp:141-155, .cpp:319-378) through photo_rates.
 *   tol_newton [device] 2: (atol, rtol)
 *   tol_time   [device] [m][2]: (atol, rtol) per species
 *   fac        unused (FD-Jacobian scratch in the reference); may be NULL
 *   tadv       [device] [nBatch] tchem_time_advance_t
 *   state      [device] [nBatch][nSpec+3] in/out: active species integrated in place
 *   photo_rates [device] [nBatch][nPhoto] or NULL: per-cell A of the photolysis reactions
 *   t_out, dt_out [device] [nBatch] in/out: cells with t_out >= tend are skipped
 *   state_out  [device] [nBatch][nSpec+3]: density, pressure, temperature copied;
 *              species copied too when state_out != state (may alias state)
 *   stats      [device] [nBatch] tchem_cell_stats_t or NULL (overwritten)
 * Per-cell failure follows the reference: species zeroed, t_out = tend, dt_out = -1. */
int tchem_atmchem_run_device_batch(const tchem_atmchem_options_t* opts, int nBatch, const double* tol_newton,
                                   const double* tol_time, double* fac, const tchem_time_advance_t* tadv,
                                   double* state, const double* photo_rates, double* t_out, double* dt_out,
                                   double* state_out, tchem_cell_stats_t* stats, tchem_kmcd_t kmcd, void* stream);

/* The driver's between-interval update (src/examples/TChem_AtmosphericChemistry.cpp:340-350):
 * tadv[i]._tbeg = t_out[i], tadv[i]._dt = dt_out[i]. */
int tchem_atmchem_update_time_advance(int nBatch, tchem_time_advance_t* tadv, const double* t_out,
                                      const double* dt_out, void* stream);

/* Replaces AtmosphericChemistryE3SM::runDeviceBatch(policy, tol_newton, tol_time, fac, tadv,
 * state, photo_rates, external_sources, t_out, dt_out, state_out, kmcd)
 * (src/core/TChem_AtmosphericChemistryE3SM.hpp, .cpp:36-232; problem:
 * src/core/problems/TChem_Impl_AtmosphericChemistryE3SM{,_Problem}.hpp): the
 * same TR-BDF2 box model with the E3SM kinetics' net production rates
 * (rate constants of the E3SM kinetics with per-cell photolysis rates, plus
 * per-cell external sources).  Arguments as tchem_atmchem_run_device_batch,
 * plus photo_rates [device] [nBatch][nPhotoRates] or NULL and
 * external_sources [device] [nBatch][m] or NULL.  Requires every species the
 * rate constants depend on ([M], adjusted reactions, prod_O1D) to be a
 * constant species, as in the UCI mechanisms. */
int tchem_atmchem_e3sm_run_device_batch(const tchem_atmchem_options_t* opts, int nBatch, const double* tol_newton,
                                        const double* tol_time, double* fac, const tchem_time_advance_t* tadv,
                                        double* state, const double* photo_rates, int nPhotoRates,
                                        const double* external_sources, double* t_out, double* dt_out,
                                        double* state_out, tchem_cell_stats_t* stats, tchem_kmcd_t kmcd,
                                        void* stream);

/* No counterpart in the reference: the FP32 dense factorisation of the
 * MFMA form (TCHEM_LU_MFMA) on its own, for dense systems of the caller.
 * Solves A_i x_i = b_i for nBatch matrices A [nBatch][n][n], b, x [nBatch][n]
 * (FP64 in and out, [device]) with n <= 255, factored in FP32 on the matrix
 * cores without row interchanges (2 x 2 diagonal pivot blocks; a singular
 * pivot block gives NaN for that system).
 * One workgroup of 256 threads per system. */
int tchem_dense_lu_solve_mfma(int n, int nBatch, const double* A, const double* b, double* x, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): tchem_atmchem_run_device_batch with the same arguments and
 * results, whose scratch (the batch's rate constants and the work-queue
 * counter) lives in a file-scope device array of 16 MiB per GPU instead of
 * library allocations; the call allocates nothing.  Limit: nBatch * nReac *
 * 8 + 256 bytes <= 16 MiB (11 214 CB05 cells per call); larger batches get
 * TCHEM_ERR_INVALID.  Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (other streams, other threads) race. */
int tchem_atmchem_run_device_batch_noalloc(const tchem_atmchem_options_t* opts, int nBatch,
                                           const double* tol_newton, const double* tol_time, double* fac,
                                           const tchem_time_advance_t* tadv, double* state,
                                           const double* photo_rates, double* t_out, double* dt_out,
                                           double* state_out, tchem_cell_stats_t* stats, tchem_kmcd_t kmcd,
                                           void* stream);

/* The driver's interval loop (src/examples/TChem_AtmosphericChemistry.cpp:312-352)
 * captured into a HIP graph, bound to the given arrays by pointer: one
 * launch replays nIntervals x (tchem_atmchem_run_device_batch, then
 * tchem_atmchem_update_time_advance) on their current contents.  With
 * stopFraction > 0 the driver's stop test runs on the device after every
 * interval (mean t_out > stopFraction * tend, tend of cell 0; the driver uses
 * 0.9999) and the remaining intervals of the launch do nothing once it holds,
 * so a launch of nIntervals reproduces the eager loop exactly; stopFraction 0
 * runs all nIntervals.  Options are resolved once at capture (blocks_per_cu 0
 * -> the occupancy limit).  graph_create is a setup call (synchronises the
 * device, allocates); tchem_graph_status reads the intervals run and the stop
 * flag of the last launch (synchronous).  Lifetime: the graph owns its rate
 * constants, work-queue counter and loop control, so later calls on the same
 * kmcd (any batch size) leave it valid; it reads the kmcd's tables and the
 * caller's arrays, which must outlive it (the kmcd refuses destruction until
 * its graphs are destroyed).  Launches of one graph must not overlap. */
int tchem_atmchem_graph_create(const tchem_atmchem_options_t* opts, int nBatch, const double* tol_newton,
                               const double* tol_time, tchem_time_advance_t* tadv, double* state,
                               const double* photo_rates, double* t_out, double* dt_out, double* state_out,
                               tchem_cell_stats_t* stats, int nIntervals, double stopFraction, tchem_kmcd_t kmcd,
                               tchem_graph_t* out);
int tchem_graph_launch(tchem_graph_t graph, void* stream);
int tchem_graph_status(tchem_graph_t graph, int* intervals_run, int* stopped);
int tchem_graph_destroy(tchem_graph_t graph);

/* Autotuner (no counterpart in the reference, whose launch shape is fixed):
 * times tchem_atmchem_run_device_batch on a strided sample of the batch
 * (sampleCells, 0 = 48 per CU; each cell limited to sampleIterations TR-BDF2
 * iterations from its tadv, 0 = 24) for every LU form opts->lu allows (AUTO:
 * sparse and dense), every program placement opts->programs allows (AUTO:
 * LDS and device memory), every cells_per_block that fits, and the two
 * largest blocks_per_cu of the persistent schedule, and writes the fastest
 * into *opts (precision and schedule are kept).  The caller's arrays are not
 * modified.  Synchronous; allocates scratch. */
int tchem_atmchem_autotune(tchem_atmchem_options_t* opts, int nBatch, const double* tol_newton,
                           const double* tol_time, const tchem_time_advance_t* tadv, const double* state,
                           const double* photo_rates, int sampleCells, int sampleIterations, double* best_seconds,
                           tchem_kmcd_t kmcd, void* stream);

/* Source term omega[nBatch][m] at state rows [nBatch][nSpec+3]: the batched
 * computeFunction of src/examples/TChem_GasChemistry_RHSs.cpp:184-236
 * (AtmosphericChemistry_Problem::computeFunction,
 * src/core/problems/TChem_Impl_AtmosphericChemistry_Problem.hpp:127-143),
 * argument order as NetProductionRates::runDeviceBatch (src/core/TChem_NetProductionRates.hpp:61-69).
 *   photo_rates [device] [nBatch][nPhoto] or NULL
 *   external_sources [device] [nBatch][m] or NULL: added to omega */
int tchem_sourceterm_run_device_batch(int nBatch, const double* state, const double* photo_rates,
                                      const double* external_sources, double* omega, tchem_kmcd_t kmcd,
                                      void* stream);

/* Jacobian jac[nBatch][m][m] = d omega_i / d x_j at state rows: the batched
 * computeJacobian of src/examples/TChem_GasChemistry_RHSs.cpp:254-309
 * (analytic here; forward differences or Sacado AD in the reference,
 * src/core/problems/TChem_Impl_AtmosphericChemistry_Problem.hpp:166-241). */
int tchem_jacobian_run_device_batch(int nBatch, const double* state, const double* photo_rates, double* jac,
                                    tchem_kmcd_t kmcd, void* stream);

/* ---- E3SM / UCI kinetics batch functions ----------------------------------- */
/* The reference's E3SM kinetics (Arrhenius and Troe without the ppm
 * conversion factor, CMAQ_H2O2, JPL and ratio-JPL falloff with [M], adjusted
 * reactions and prod_O1D, photolysis rates given per cell).  State rows as
 * above; species concentrations are the row's species block.  These work
 * for every mechanism the loader reads, including those the box-model
 * integrator rejects.
 *
 * Replaces ReactionRates::runDeviceBatch(policy, state, kfwd, kmcd)
 * (src/core/TChem_ReactionRates.cpp:118-134; Impl: TChem_Impl_ReactionRates.hpp).
 *   kfwd [device] [nBatch][nReac] */
int tchem_reactionrates_run_device_batch(int nBatch, const double* state, double* kfwd, tchem_kmcd_t kmcd,
                                         void* stream);

/* Replaces RateofProgress::runDeviceBatch(policy, state, photo_rates, rate_of_progress, kmcd)
 * (src/core/TChem_RateofProgress.cpp; Impl: TChem_Impl_RateofProgress.hpp:52-102).
 *   photo_rates [device] [nBatch][nPhotoRates] or NULL: rate constants of
 *               reactions 0..nPhotoRates-1 (the PHOTO_RATE reactions come first)
 *   rop [device] [nBatch][nReac] */
int tchem_rateofprogress_run_device_batch(int nBatch, const double* state, const double* photo_rates,
                                          int nPhotoRates, double* rop, tchem_kmcd_t kmcd, void* stream);

/* Replaces NetProductionRates::runDeviceBatch(policy, state, photo_rates, external_sources,
 * net_production_rates, kmcd) (src/core/TChem_NetProductionRates.hpp:61-78;
 * Impl: TChem_Impl_NetProductionRates.hpp:52-112).
 *   external_sources [device] [nBatch][m] or NULL
 *   omega [device] [nBatch][ldOmega], ldOmega >= m: active species' net
 *         production rates, the remaining ldOmega - m entries of a row set to 0
 *         (the reference's callers pass [nBatch][nSpec]) */
int tchem_netproductionrates_run_device_batch(int nBatch, const double* state, const double* photo_rates,
                                              int nPhotoRates, const double* external_sources, double* omega,
                                              int ldOmega, tchem_kmcd_t kmcd, void* stream);

/* Replaces AtmChemistry::setScenarioConditionsPhotolysisReactions
 * (src/core/TChem_AtmosphericChemistry.cpp:511-560): the scenario file's
 * photolysis_rates block as host rows [nBatch][capacity_cols] (count columns
 * used; rates_host NULL queries count; count 0 without the block). */
int tchem_scenario_photolysis_rates(tchem_kmd_t kmd, const char* path, int nBatch, double* rates_host,
                                    int capacity_cols, int* count);

/* Replaces AtmChemistry::setScenarioConditionsExternalForcing
 * (src/core/TChem_AtmosphericChemistry.cpp:562-631): host rows [nBatch][m],
 * zero for species without forcing; count = forced species (nullable). */
int tchem_scenario_external_forcing(tchem_kmd_t kmd, const char* path, int nBatch, double* forcing_host, int* count);

/* ---- discrete adjoint ----------------------------------------------------- */
/* No counterpart in the reference, which has no derivative path: gradients
 * of a scalar loss on the final concentrations of a recorded TR-BDF2 run
 * with respect to the initial active species and to per-reaction log-scale
 * parameters theta_r (kf_r = k_r exp(theta_r); for photolysis reactions a
 * uniform log-scale on the per-cell rate).  The forward run is the FP64
 * sparse persistent integrator of tchem_atmchem_run_device_batch with its
 * accepted step sizes and a checkpoint of the state every 16 accepted steps
 * recorded into a caller-owned device workspace; the backward sweep
 * replays each 16-step segment from its checkpoint and runs the transposed
 * stage solves in reverse.  Step sizes are held fixed (the derivative of the
 * discrete map along the accepted steps).
 *
 * Workspace bytes for nBatch cells and at most maxSteps accepted steps per
 * cell (1 <= maxSteps <= 2^24). */
int tchem_adjoint_workspace_bytes(tchem_kmcd_t kmcd, int nBatch, int maxSteps, size_t* bytes);

/* Forward run with recording, from tadv[i]._tbeg to ._tend in one launch
 * (at most _num_time_iterations_per_interval iterations).  Arguments as
 * tchem_atmchem_run_device_batch, plus
 *   log_rate_scale [device] theta, NULL for 0: [nReac] shared by every cell
 *                  (log_rate_scale_stride 0) or [nBatch][nReac] (stride nReac)
 *   workspace      [device] tchem_adjoint_workspace_bytes(kmcd, nBatch, maxSteps) bytes
 * A cell that fails, or takes more than maxSteps accepted steps, is marked
 * in the record; its backward results are NaN. */
int tchem_adjoint_forward(int nBatch, const double* tol_newton, const double* tol_time,
                          const tchem_time_advance_t* tadv, double* state, const double* photo_rates,
                          const double* log_rate_scale, int log_rate_scale_stride, double* t_out, double* dt_out,
                          tchem_cell_stats_t* stats, void* workspace, int maxSteps, tchem_kmcd_t kmcd,
                          void* stream);

/* Backward sweep over the record of the last tchem_adjoint_forward on this
 * workspace (same nBatch and maxSteps; a mismatch gives NaN results).
 *   lambda     [device] [nBatch][nSpec] dL/d(final species), reference order
 *              (constant species ignored)
 *   lambda_out [device] [nBatch][nSpec] dL/d(initial species); 0 for constant
 *              species; may alias lambda
 *   grad       [device] [nBatch][nReac] dL/dtheta_r of each cell
 * The workspace's replay scratch is overwritten; the record is kept, so the
 * sweep may be repeated with another lambda. */
int tchem_adjoint_backward(int nBatch, const double* lambda, double* lambda_out, double* grad, void* workspace,
                           int maxSteps, tchem_kmcd_t kmcd, void* stream);

/* Copies the record's accepted-step counts nsteps[nBatch] (-1: failed or
 * overflowed) and, when h is not NULL, step sizes h[nBatch][maxSteps]. */
int tchem_adjoint_steps(int nBatch, const void* workspace, int maxSteps, int* nsteps, double* h, tchem_kmcd_t kmcd,
                        void* stream);


/* ---- single-particle SIMPOL.1 aerosol model ------------------------------ */
/* Unknowns of a cell: the active gas species, then nParticles particles of
 * nSpec aerosol species each.  State rows [nSpec_gas + 3 + nSpec * nParticles]:
 * density, pressure, temperature, gas species (active then constant),
 * particle species (particle-major), kg m-3 per particle. */

/* Replaces AerosolModelData(mechfile, kmd) (src/core/TChem_AerosolModelData.cpp:30-249):
 * aerosol species, particle count and SIMPOL_PHASE_TRANSFER reactions of an
 * NCAR-format aerosol file, against the gas model kmd. */
int tchem_amd_load_yaml(const char* path, tchem_kmd_t kmd, tchem_amd_t* out);
int tchem_amd_destroy(tchem_amd_t amd);
/* nSpec_, nParticles_, nSimpol_tran_ (AerosolModelData.hpp:31-41); any may be NULL */
int tchem_amd_info(tchem_amd_t amd, int* nSpec, int* nParticles, int* nSimpol);
/* Replaces AerosolModelData::setNumberofParticles (.cpp:251-260) */
int tchem_amd_set_number_of_particles(tchem_amd_t amd, int nParticles);
/* aerosol species name and (density, molecular weight, kappa) of index i
 * (TChem_Driver.cpp: getAerosolSpeciesName / Density / MW / Kappa) */
int tchem_amd_species_name(tchem_amd_t amd, int i, char* buf, size_t len);
int tchem_amd_species_props(tchem_amd_t amd, int i, double* density, double* mw, double* kappa);
/* Replaces AerosolModelData::scenarioConditionParticles (.cpp:263-300):
 * number_conc_host [nBatch][nParticles] written; particle species of
 * state_host rows [nBatch][stateDim] set to initial_value / number
 * concentration.  *found = 0 (nothing written) when the file has no
 * `particles` block. */
int tchem_scenario_particles(tchem_amd_t amd, const char* path, int nBatch, int stateDim, double* state_host,
                             double* number_conc_host, int* found);

/* Replaces create_AerosolModelConstData<device>(amd) (AerosolModelData.hpp:94-108)
 * for use with the gas model's kmcd, which must outlive it.  At most 256
 * unknowns per cell (the dense form: 128). */
int tchem_amcd_create(tchem_amd_t amd, tchem_kmcd_t kmcd, tchem_amcd_t* out);
int tchem_amcd_destroy(tchem_amcd_t amcd);

/* Batched aerosol source term (src/core/impl/TChem_Impl_Aerosol_RHS.hpp:64-103;
 * the reference's AerosolChemistry_RHSs example), FP64:
 *   state [device] [nBatch][stateDim], number_conc [device] [nBatch][nParticles]
 *   omega [device] [nBatch][mt] in the unknown order above. */
int tchem_aerosol_rhs_run_device_batch(int nBatch, const double* state, const double* number_conc, double* omega,
                                       tchem_amcd_t amcd, void* stream);

/* Replaces AerosolChemistry::runDeviceBatch(policy, tol_newton, tol_time, fac,
 * tadv, state, number_conc, t_out, dt_out, state_out, kmcd, amcd)
 * (src/core/TChem_AerosolChemistry.hpp, .cpp:31-258): TR-BDF2 over the gas
 * and particle unknowns of every cell, FP64, analytic Jacobian; opts->lu
 * selects the sparse form (AUTO, SPARSE: the mechanism's compiled programs
 * over all unknowns) or DENSE (LU with partial pivoting, <= 128 unknowns).
 * Arguments as tchem_atmchem_run_device_batch with tol_time [mt][2] and
 * state rows of stateDim; opts NULL or FP64 precision (other fields
 * ignored). */
int tchem_aerosolchem_run_device_batch(const tchem_atmchem_options_t* opts, int nBatch, const double* tol_newton,
                                       const double* tol_time, double* fac, const tchem_time_advance_t* tadv,
                                       double* state, const double* number_conc, double* t_out, double* dt_out,
                                       double* state_out, tchem_cell_stats_t* stats, tchem_amcd_t amcd,
                                       void* stream);

#ifdef __cplusplus
}
#endif

#endif
