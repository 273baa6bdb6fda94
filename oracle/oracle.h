/* FP64 CPU oracle for the batched gas-phase kinetics path.
 *
 * TEST INFRASTRUCTURE ONLY.  This restatement is the checker for the HIP
 * implementation: only tests, the smoke check and the benchmark's CPU
 * baseline leg may load it.  The product path never calls into it.
 *
 * It restates, in plain C, the reference's arithmetic for
 *   - rate constants   src/core/impl/TChem_Impl_KForwardAerosol.hpp:43-145
 *   - source term      src/core/impl/TChem_Impl_ReactionRatesAerosol.hpp:56-143
 *   - the box-model integration loop
 *                      src/core/TChem_AtmosphericChemistry.cpp:74-165,
 *                      src/core/problems/TChem_Impl_AtmosphericChemistry.hpp:74-223
 * on the reference's own table layout (KineticModelData, see
 * src/core/TChem_KineticModelData.cpp:302-338).  The TR-BDF2 integrator and
 * Newton solver live in Tines, which the reference does not vendor; the
 * stepper below is the published TR-BDF2 method (Bank et al. 1985, Hosea &
 * Shampine 1996) with the step-size rules the reference goldens fix
 * (doubling growth, dtmax clamp, clamp to t_end before the step, dt_out =
 * next proposal) and the retry/failure pattern of
 * src/core/impl/TChem_Impl_TimeIntegratorImplicitEuler.hpp:242-306.
 */
#ifndef TCHEM_ORACLE_H
#define TCHEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nSpec, nConstSpec, nReac, maxSpecInReac;
  const int* reacNreac;    /* [nReac] */
  const int* reacNprod;    /* [nReac] */
  const int* reacSidx;     /* [nReac][maxSpecInReac] */
  const double* reacNuki;  /* [nReac][maxSpecInReac] */
  const int* convExponent; /* [nReac] */
  int nArr;
  const int* arrIdx;       /* [nArr] reaction index */
  const double* arrPar;    /* [nArr][5] A, B, C, D, E */
  int nTroe;
  const int* troeIdx;      /* [nTroe] */
  const double* troePar;   /* [nTroe][8] k0_A(*1e6), k0_B, k0_C, kinf_A, kinf_B, kinf_C, Fc, N */
  int nCmaq;
  const int* cmaqIdx;      /* [nCmaq] */
  const double* cmaqPar;   /* [nCmaq][6] A1, B1, C1, A2(*1e6), B2, C2 */
  int nEmis;
  const int* emisSpecies;  /* [nEmis] */
  const double* emisRate;  /* [nEmis] */
  int nPhoto;
  const int* photoReac;    /* [nPhoto] reactions whose A is a per-cell photolysis rate */
  double CONV_PPM;
  /* nullable [m]: elimination order of the dense Newton solves (position ->
   * active species), a fill-reducing order; NULL = species order */
  const int* lu_order;
} oracle_kmd;

typedef struct { /* byte layout of TChem::TimeAdvance (src/core/TChem_Util.hpp:469-476) */
  double tbeg, tend, dt, dtmin, dtmax;
  int max_num_newton_iterations;
  int num_time_iterations_per_interval;
  int num_outer_time_iterations_per_interval;
  int jacobian_interval;
} oracle_time_advance;

/* per-cell counters written by the integrator */
typedef struct {
  int accepted, rejected, newton_failures, newton_iterations, factorizations, failed;
} oracle_stats;

/* kfor[nReac]; photo (nullable) holds nPhoto per-cell photolysis rates that
 * replace A of the listed reactions (the reference's per-cell kmcd array) */
void oracle_kforward(const oracle_kmd* k, double T, double p, const double* photo, double* kfor);

/* omega[m] at concentrations c[nSpec] (active then constant species) */
void oracle_omega(const oracle_kmd* k, const double* kfor, const double* c, double* omega);

/* analytic Jacobian J[m][m] = d omega_i / d c_j over active species */
void oracle_jacobian(const oracle_kmd* k, const double* kfor, const double* c, double* J);

/* LU with partial pivoting, row-major A[n][lda]; returns 0 or k+1 for a zero pivot */
int oracle_lu(int n, double* A, int lda, int* piv);
void oracle_lu_solve(int n, const double* A, int lda, const int* piv, double* b);

/* Batched source term / Jacobian over state rows [nBatch][nSpec+3]. */
void oracle_sourceterm_batch(const oracle_kmd* k, int nBatch, const double* state, const double* photo,
                             double* omega);
void oracle_jacobian_batch(const oracle_kmd* k, int nBatch, const double* state, const double* photo,
                           double* jac);

/* One runDeviceBatch interval over nBatch cells (OpenMP, static schedule).
 *   jacobian_mode: 0 = analytic Jacobian, rate constants computed once per
 *                      cell (the oracle form the GPU path is checked against)
 *                  1 = the reference algorithm: forward-difference Jacobian
 *                      with m+1 source-term evaluations, all rate constants
 *                      recomputed in every source-term evaluation
 *   state: [nBatch][nSpec+3], species advanced in place (as the reference).
 *   t_out/dt_out: in/out per cell; cells with t_out >= tend are skipped.
 *   stats: nullable [nBatch] counters (accumulated).
 *   nthreads: 0 = OpenMP default.
 */
void oracle_run_batch(const oracle_kmd* k, int nBatch, const double* tol_newton, const double* tol_time,
                      const oracle_time_advance* tadv, double* state, const double* photo, double* t_out,
                      double* dt_out, oracle_stats* stats, int jacobian_mode, int nthreads);

/* One cell with rate constants kf_r * exp(lscale_r) (lscale nullable),
 * recording accepted step sizes h_out[max_steps]; row integrated in place.
 * Returns the number of accepted steps, -1 on failure or overflow. */
int oracle_record_steps(const oracle_kmd* k, const double* tol_newton, const double* tol_time,
                        const oracle_time_advance* ta, double* row, const double* photo, const double* lscale,
                        double* h_out, int max_steps);

/* Fixed step sequence h[nsteps] with the discrete tangent linear of nsens
 * parameters (sens[c] >= 0: log-scale of reaction sens[c]; < 0: initial
 * value of active species -sens[c]-1).  u_out [m], S_out [m][nsens]. */
int oracle_tlm_fixed_steps(const oracle_kmd* k, const double* tol_newton, int max_newton, int jint,
                           const double* row, const double* photo, const double* lscale, const double* h, int nsteps,
                           int nsens, const int* sens, double* u_out, double* S_out);

/* ---- single-particle SIMPOL.1 aerosol model (AerosolModelData) ----
 * Unknowns of a cell: [active gas species (mg), particle 0 species (nSpec),
 * particle 1 species, ...]; state rows [nSpec_gas + 3 + nSpec * nParticles]:
 * density, pressure, temperature, gas species, particle species. */
typedef struct {
  int nSpec, nParticles, nSimpol;
  const double* mw;       /* [nSpec] kg mol-1 */
  const double* density;  /* [nSpec] kg m-3 */
  const int* gas;         /* [nSimpol] gas species (kinetic model index, active) */
  const int* aero;        /* [nSimpol] aerosol species index within a particle */
  const double* par;      /* [nSimpol][8] B1 B2 B3 B4 diffusion mw Nstar computeAlpha */
} oracle_amd;

/* alpha, mean free path [m], kg m-3 -> ppm factor and equilibrium constant
 * of SIMPOL transfer s at T, p */
void oracle_simpol_constants(const double* par8, double T, double p, double* out4);

/* omega[mt] = gas-phase source term + SIMPOL transfer at unknowns x[mt];
 * cconst: constant gas species; nc: particle number concentrations */
void oracle_aerosol_omega(const oracle_kmd* k, const oracle_amd* a, double T, double p, const double* kfor,
                          const double* cconst, const double* nc, const double* x, double* omega);
/* analytic J[mt][mt] = d omega / d x */
void oracle_aerosol_jacobian(const oracle_kmd* k, const oracle_amd* a, double T, double p, const double* kfor,
                             const double* cconst, const double* nc, const double* x, double* J);
/* batched omega over state rows; nc [nBatch][nParticles]; omega [nBatch][mt] */
void oracle_aerosol_rhs_batch(const oracle_kmd* k, const oracle_amd* a, int nBatch, const double* state,
                              const double* nc, double* omega);
/* One AerosolChemistry::runDeviceBatch interval (analytic Jacobian);
 * tol_time [mt][2]; other arguments as oracle_run_batch. */
void oracle_aerosol_run_batch(const oracle_kmd* k, const oracle_amd* a, int nBatch, const double* tol_newton,
                              const double* tol_time, const oracle_time_advance* tadv, double* state,
                              const double* nc, double* t_out, double* dt_out, oracle_stats* stats, int nthreads);

/* exp / log / log10 / pow of the rate constants over arrays (for tests) */
void oracle_detmath_batch(int fn, int n, const double* x, const double* y, double* out);

/* TR-BDF2 constants used by the oracle (for tests) */
void oracle_trbdf2_constants(double* out6);

#ifdef __cplusplus
}
#endif

#endif
