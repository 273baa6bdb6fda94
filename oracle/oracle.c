/* FP64 CPU oracle — TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Each function cites the reference code it restates.  Parity status of each
 * part is stated in the header comment of the function; pins live in
 * tests/test_oracle.py, tests/test_oracle_adjoint.py and
 * tests/test_oracle_aerosol.py.
 */
#include "oracle.h"

#include <math.h>

/* exp / log / pow from IEEE operations only (the rate constants are then the
 * same bits on the host and on the GPU); pinned against the C library in
 * tests/test_detmath.py. */
#include "../tchem-atm_amd/csrc/detmath.h"
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* TR-BDF2 with gamma = 2 - sqrt(2) (Bank et al. 1985). */
#define TRB_GAMMA (2.0 - 1.41421356237309504880)

/* pow(double, int) of the reference (ats<real_type>::pow(c, niup)) for the
 * small integer exponents of mechanisms: repeated multiplication, which for
 * n <= 2 is the correctly rounded power. */
static double pow_int(double x, int n) { return tchem_ipow(x, n); }

/* Rate constants: src/core/impl/TChem_Impl_KForwardAerosol.hpp:54-131.
 * Same expression order as the reference; photolysis overrides replace A of
 * the listed Arrhenius reactions (the per-cell kmcd "model variation" form,
 * src/core/TChem_AtmosphericChemistry.cpp:79).
 * Pinned by: the CAMP analytic solutions' rate constants at fixed T, p and
 * the reference's CB05 goldens (tests/test_oracle.py:
 * test_rate_constants_against_camp_analytic_solution, test_cb05cl_ae5_golden). */
void oracle_kforward(const oracle_kmd* k, double t, double p, const double* photo, double* kfor) {
  const double t_1 = 1.0 / t;
  const double one = 1.0;
  const double conv = k->CONV_PPM * p * t_1;
  for (int i = 0; i < k->nArr; ++i) {
    const double* P = k->arrPar + 5 * i;
    const int r = k->arrIdx[i];
    double A = P[0];
    if (photo)
      for (int q = 0; q < k->nPhoto; ++q)
        if (k->photoReac[q] == r) A = photo[q];
    const double B = P[1], C = P[2], D = P[3], E = P[4];
    kfor[r] = A * tchem_exp(C * t_1) * (B == 0.0 ? one : tchem_pow(t / D, B)) * (E == 0.0 ? one : (one + E * p)) *
              pow_int(conv, k->convExponent[r]);
  }
  for (int i = 0; i < k->nTroe; ++i) {
    const double* rp = k->troePar + 8 * i;
    const int r = k->troeIdx[i];
    const double k0_A = rp[0], k0_B = rp[1], k0_C = rp[2];
    const double kinf_A = rp[3], kinf_B = rp[4], kinf_C = rp[5], Fc = rp[6], N = rp[7];
    const double k0 =
        k0_A * (k0_C == 0.0 ? one : tchem_exp(k0_C * t_1)) * (k0_B == 0.0 ? one : tchem_pow(t / 300.0, k0_B)) * conv;
    const double k0_kinf =
        k0 / (kinf_A * (kinf_C == 0.0 ? one : tchem_exp(kinf_C * t_1)) *
              (kinf_B == 0.0 ? one : tchem_pow(t / 300.0, kinf_B)));
    kfor[r] = k0 / (1.0 + k0_kinf) * tchem_pow(Fc, 1.0 / (1.0 + pow_int(tchem_log10(k0_kinf) / N, 2))) *
              pow_int(conv, k->convExponent[r]);
  }
  for (int i = 0; i < k->nCmaq; ++i) {
    const double* P = k->cmaqPar + 6 * i;
    const int r = k->cmaqIdx[i];
    kfor[r] = (P[0] * tchem_exp(P[2] * t_1) * (P[1] == 0.0 ? one : tchem_pow(t / 300.0, P[1])) +
               P[3] * tchem_exp(P[5] * t_1) * (P[4] == 0.0 ? one : tchem_pow(t / 300.0, P[4])) * conv) *
              pow_int(conv, k->convExponent[r]);
  }
}

/* Source term: src/core/impl/TChem_Impl_ReactionRatesAerosol.hpp:78-138.
 * omega := 0; omega[sp] := emission rate; rop_r = kfor_r * prod c^|int nu|;
 * omega[k] += nu * rop for active k, reactants then products, in the
 * reference's loop order (the serial order of its atomic adds).
 * Pinned by: the CB05 golden trajectory and brute-force/invariant checks. */
void oracle_omega(const oracle_kmd* k, const double* kfor, const double* c, double* omega) {
  const int m = k->nSpec - k->nConstSpec;
  const int ms = k->maxSpecInReac;
  for (int i = 0; i < m; ++i) omega[i] = 0.0;
  for (int i = 0; i < k->nEmis; ++i)
    if (k->emisSpecies[i] < m) omega[k->emisSpecies[i]] = k->emisRate[i];
  for (int r = 0; r < k->nReac; ++r) {
    double rop = kfor[r];
    for (int j = 0; j < k->reacNreac[r]; ++j) {
      const int kspec = k->reacSidx[r * ms + j];
      const int niup = abs((int)k->reacNuki[r * ms + j]);
      rop *= pow_int(c[kspec], niup);
    }
    for (int j = 0; j < k->reacNreac[r]; ++j) {
      const int kspec = k->reacSidx[r * ms + j];
      if (kspec < m) omega[kspec] += k->reacNuki[r * ms + j] * rop;
    }
    const int joff = ms / 2;
    for (int j = 0; j < k->reacNprod[r]; ++j) {
      const int kspec = k->reacSidx[r * ms + j + joff];
      if (kspec < m) omega[kspec] += k->reacNuki[r * ms + j + joff] * rop;
    }
  }
}

/* Analytic Jacobian of oracle_omega with respect to the active species.
 * The reference differentiates the same function by forward differences or
 * Sacado forward AD (src/core/problems/TChem_Impl_AtmosphericChemistry_Problem.hpp:166-241).
 * Pinned by: central differences on CB05 and a long-double brute force on
 * small hand-written mechanisms (tests/test_oracle.py:
 * test_analytic_jacobian_against_central_differences,
 * test_source_term_and_jacobian_by_brute_force). */
void oracle_jacobian(const oracle_kmd* k, const double* kfor, const double* c, double* J) {
  const int m = k->nSpec - k->nConstSpec;
  const int ms = k->maxSpecInReac;
  memset(J, 0, sizeof(double) * (size_t)m * m);
  for (int r = 0; r < k->nReac; ++r) {
    const int nre = k->reacNreac[r];
    for (int s = 0; s < nre; ++s) {
      const int ks = k->reacSidx[r * ms + s];
      if (ks >= m) continue; /* constant species: no derivative */
      /* d rop / d c_ks = kfor * n_s c_ks^(n_s-1) * prod_{j != s} c_j^n_j */
      const int ns = abs((int)k->reacNuki[r * ms + s]);
      double d = kfor[r] * (double)ns * pow_int(c[ks], ns - 1);
      for (int j = 0; j < nre; ++j) {
        if (j == s) continue;
        d *= pow_int(c[k->reacSidx[r * ms + j]], abs((int)k->reacNuki[r * ms + j]));
      }
      for (int j = 0; j < nre; ++j) {
        const int kt = k->reacSidx[r * ms + j];
        if (kt < m) J[(size_t)kt * m + ks] += k->reacNuki[r * ms + j] * d;
      }
      const int joff = ms / 2;
      for (int j = 0; j < k->reacNprod[r]; ++j) {
        const int kt = k->reacSidx[r * ms + j + joff];
        if (kt < m) J[(size_t)kt * m + ks] += k->reacNuki[r * ms + j + joff] * d;
      }
    }
  }
}

/* Dense LU with partial pivoting (row interchanges), Doolittle form;
 * multipliers use the reciprocal of the pivot, l = a_ik * (1 / a_kk); every
 * update a_ij - l u_kj is one fused multiply-add (one rounding), as are the
 * substitution updates of oracle_lu_solve.
 * The reference factors with Tines' UTV; both solve the same Newton system.
 * Pinned by: scipy.linalg.lu_factor and the backward error on random and
 * CB05 iteration matrices. */
int oracle_lu(int n, double* A, int lda, int* piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double amax = fabs(A[(size_t)k * lda + k]);
    for (int i = k + 1; i < n; ++i) {
      const double a = fabs(A[(size_t)i * lda + k]);
      if (a > amax) { amax = a; p = i; }
    }
    piv[k] = p;
    if (amax == 0.0) return k + 1;
    if (p != k)
      for (int j = 0; j < n; ++j) {
        const double tmp = A[(size_t)k * lda + j];
        A[(size_t)k * lda + j] = A[(size_t)p * lda + j];
        A[(size_t)p * lda + j] = tmp;
      }
    const double inv = 1.0 / A[(size_t)k * lda + k];
    for (int i = k + 1; i < n; ++i) {
      const double l = A[(size_t)i * lda + k] * inv;
      A[(size_t)i * lda + k] = l;
      if (l != 0.0)
        for (int j = k + 1; j < n; ++j) A[(size_t)i * lda + j] = fma(-l, A[(size_t)k * lda + j], A[(size_t)i * lda + j]);
    }
  }
  return 0;
}

void oracle_lu_solve(int n, const double* A, int lda, const int* piv, double* b) {
  for (int k = 0; k < n; ++k) {
    const int p = piv[k];
    if (p != k) { const double t = b[k]; b[k] = b[p]; b[p] = t; }
  }
  for (int k = 0; k < n; ++k) {
    const double yk = b[k];
    if (yk != 0.0)
      for (int i = k + 1; i < n; ++i) b[i] = fma(-A[(size_t)i * lda + k], yk, b[i]);
  }
  for (int k = n - 1; k >= 0; --k) {
    const double xk = b[k] * (1.0 / A[(size_t)k * lda + k]);
    b[k] = xk;
    if (xk != 0.0)
      for (int i = 0; i < k; ++i) b[i] = fma(-A[(size_t)i * lda + k], xk, b[i]);
  }
}

/* the portable elementary functions of the rate constants, for tests:
 * fn 0 exp(x), 1 log(x), 2 log10(x), 3 pow(x, y) */
void oracle_detmath_batch(int fn, int n, const double* x, const double* y, double* out) {
  for (int i = 0; i < n; ++i)
    out[i] = fn == 0 ? tchem_exp(x[i]) : fn == 1 ? tchem_log(x[i]) : fn == 2 ? tchem_log10(x[i]) : tchem_pow(x[i], y[i]);
}

void oracle_trbdf2_constants(double* out6) {
  const double g = TRB_GAMMA;
  out6[0] = g;
  out6[1] = 0.5 * g;                                /* TR stage: u - un - (g/2) h (f + fn) */
  out6[2] = 1.0 / (g * (2.0 - g));                  /* BDF2: coefficient of u_gamma */
  out6[3] = (1.0 - g) * (1.0 - g) / (g * (2.0 - g));/* BDF2: coefficient of u_n */
  out6[4] = (1.0 - g) / (2.0 - g);                  /* BDF2: coefficient of h f(u) (== g/2) */
  out6[5] = (-3.0 * g * g + 4.0 * g - 2.0) / (12.0 * (2.0 - g)); /* local error constant k */
}

/* ---------------------------------------------------------------------------
 * Per-cell integrator state
 */
typedef struct {
  const oracle_kmd* k;
  const oracle_amd* a; /* NULL: gas-phase box model */
  const double* nc;    /* [nParticles] number concentrations (aerosol) */
  int m, mg, mode;     /* unknowns, active gas species, Jacobian mode */
  double T, p;
  const double* photo;
  const double* lscale; /* [nReac] log-scale of the rate constants, or NULL */
  double* kfor;  /* [nReac] */
  double* c;     /* [nSpec]: active part is the evaluation point */
  double* omega; /* [m] */
  double* J;     /* [m*m] */
  double* A;     /* [m*m] */
  int* piv;      /* [m] */
  double* fh;    /* [m] FD scratch */
} cell_ctx;

/* rate constants of the cell, kf_r * exp(lscale_r) when lscale is given */
static void cell_kforward(cell_ctx* x) {
  oracle_kforward(x->k, x->T, x->p, x->photo, x->kfor);
  if (x->lscale)
    for (int r = 0; r < x->k->nReac; ++r) x->kfor[r] *= tchem_exp(x->lscale[r]);
}

/* source term at u (active part); the constant part of c is fixed per cell */
static void rhs(cell_ctx* x, const double* u, double* f) {
  memcpy(x->c, u, sizeof(double) * x->mg);
  if (x->mode == 1) cell_kforward(x); /* reference: every call */
  if (x->a)
    oracle_aerosol_omega(x->k, x->a, x->T, x->p, x->kfor, x->c + x->mg, x->nc, u, f);
  else
    oracle_omega(x->k, x->kfor, x->c, f);
}

/* Forward-difference Jacobian with m+1 source-term evaluations — the
 * reference's default (NumericalJacobianForwardDifference,
 * src/core/problems/TChem_Impl_AtmosphericChemistry_Problem.hpp:217-241).
 * The increment rule is Tines' and is not vendored; this uses
 * h_j = sqrt(eps) * max(|u_j|, 1e-10). */
static void jac_fd(cell_ctx* x, const double* u, double* J) {
  const int m = x->m;
  double* u2 = x->A; /* borrow: A is rebuilt right after */
  double* f0 = x->omega;
  rhs(x, u, f0);
  for (int j = 0; j < m; ++j) u2[j] = u[j];
  for (int j = 0; j < m; ++j) {
    const double h = 1.4901161193847656e-08 * (fabs(u[j]) > 1e-10 ? fabs(u[j]) : 1e-10);
    u2[j] = u[j] + h;
    rhs(x, u2, x->fh);
    u2[j] = u[j];
    for (int i = 0; i < m; ++i) J[(size_t)i * m + j] = (x->fh[i] - f0[i]) / h;
  }
}

static void jac(cell_ctx* x, const double* u, double* J) {
  if (x->mode == 1) {
    jac_fd(x, u, J);
  } else if (x->a) {
    memcpy(x->c, u, sizeof(double) * x->mg);
    oracle_aerosol_jacobian(x->k, x->a, x->T, x->p, x->kfor, x->c + x->mg, x->nc, u, J);
  } else {
    memcpy(x->c, u, sizeof(double) * x->m);
    oracle_jacobian(x->k, x->kfor, x->c, J);
  }
}

/* Newton iteration for one TR-BDF2 stage (Tines NewtonSolver, restated):
 * G(u) = u - b - ch * f(u), iteration matrix I - ch*J, factorised in the
 * kinetic model's lu_order when it has one (a symmetric permutation of the
 * same system: the order changes only the rounding of the solve).  The Jacobian and its
 * factorisation are refreshed when the iteration count within the stage is a
 * multiple of jacobian_interval (reference default 1: every iteration).
 * Converged when max_i |dx_i| / (atol + rtol |u_i|) <= 1. */
static int newton(cell_ctx* x, double* u, const double* b, double ch, const double* tol_newton, int max_it,
                  int jac_interval, double* f, double* dx, oracle_stats* st) {
  const int m = x->m;
  /* elimination order of the gas-phase solves (NULL: species order) */
  const int* ord = x->a ? NULL : x->k->lu_order;
  if (jac_interval < 1) jac_interval = 1;
  for (int it = 0; it < max_it; ++it) {
    if (it % jac_interval == 0) {
      jac(x, u, x->J);
      for (int i = 0; i < m; ++i) {
        const int ri = ord ? ord[i] : i;
        for (int j = 0; j < m; ++j) {
          const int cj = ord ? ord[j] : j;
          x->A[(size_t)i * m + j] = (ri == cj ? 1.0 : 0.0) - ch * x->J[(size_t)ri * m + cj];
        }
      }
      if (oracle_lu(m, x->A, m, x->piv) != 0) return 0;
      if (st) st->factorizations++;
    }
    rhs(x, u, f);
    for (int i = 0; i < m; ++i) dx[i] = -(u[i] - b[i] - ch * f[i]);
    if (ord) {
      double* y = x->fh; /* FD scratch, free once the Jacobian is built */
      for (int i = 0; i < m; ++i) y[i] = dx[ord[i]];
      oracle_lu_solve(m, x->A, m, x->piv, y);
      for (int i = 0; i < m; ++i) dx[ord[i]] = y[i];
    } else {
      oracle_lu_solve(m, x->A, m, x->piv, dx);
    }
    double nrm = 0.0;
    for (int i = 0; i < m; ++i) {
      u[i] += dx[i];
      const double w = fabs(dx[i]) / (tol_newton[0] + tol_newton[1] * fabs(u[i]));
      if (!(w <= nrm)) nrm = w; /* NaN propagates as non-convergence */
    }
    if (st) st->newton_iterations++;
    if (nrm <= 1.0) return 1;
  }
  return 0;
}

/* One cell of AtmosphericChemistry::runDeviceBatch
 * (src/core/TChem_AtmosphericChemistry.cpp:78-163): integrate the active
 * species of `row` in place from tbeg for at most
 * num_time_iterations_per_interval TR-BDF2 iterations.
 *
 * Step control (Tines' TimeIntegratorTrBDF2 is not vendored; the rules below
 * are the ones the reference goldens fix, thresholds are this project's):
 *   h = min(max(dt, dtmin), tend - t)           clamp to t_end before the step
 *   Newton failure in either stage -> dt = h/2, up to 4 attempts, then hard
 *     failure (species zeroed, t_out = tend, dt_out = -1)
 *   e = max_i |est_i| / (atol_i + rtol_i |u_i|), est = LTE estimate
 *   e > 1 and h > dtmin -> reject, dt = h/4 (the iteration is consumed)
 *   otherwise accept; next dt = 2h if e <= 1/8 else h, clamped to dtmax.
 */
static void integrate_cell(const oracle_kmd* k, const double* tol_newton, const double* tol_time,
                           const oracle_time_advance* ta, double* row, const double* photo, double* t_out,
                           double* dt_out, oracle_stats* st, int mode, double* work, int* iwork,
                           const double* lscale, double* h_rec, int max_rec, int* n_rec, const oracle_amd* am,
                           const double* nc) {
  const int nSpec = k->nSpec, mg = nSpec - k->nConstSpec;
  const int m = am ? mg + am->nSpec * am->nParticles : mg; /* unknowns */
#define ROWCOL(i) ((i) < mg ? 3 + (i) : 3 + nSpec + ((i) - mg))
  const double tend = ta->tend;
  if (*t_out >= tend) return; /* .cpp:83 */
  const double dt_in = ta->dt;
  if (dt_in < 0.0) return; /* a failed cell stays failed */

  double g6[6];
  oracle_trbdf2_constants(g6);
  const double g = g6[0], c_tr = g6[1], a_g = g6[2], a_n = g6[3], c_bdf = g6[4], k_err = g6[5];

  cell_ctx x;
  memset(&x, 0, sizeof x);
  x.k = k;
  x.a = am;
  x.nc = nc;
  x.mg = mg;
  x.m = m;
  x.mode = mode;
  x.p = row[1];
  x.T = row[2];
  x.photo = photo;
  x.lscale = lscale;
  double* w = work;
  x.kfor = w; w += k->nReac;
  x.c = w; w += nSpec;
  x.omega = w; w += m;
  x.J = w; w += (size_t)m * m;
  x.A = w; w += (size_t)m * m;
  x.fh = w; w += m;
  x.piv = iwork;
  double* un = w; w += m;
  double* u = w; w += m;
  double* ug = w; w += m;
  double* fn = w; w += m;
  double* fg = w; w += m;
  double* f1 = w; w += m;
  double* f = w; w += m;
  double* dx = w; w += m;
  double* b = w; w += m;

  for (int i = 0; i < nSpec; ++i) x.c[i] = row[3 + i];
  for (int i = 0; i < m; ++i) un[i] = row[ROWCOL(i)];
  if (mode == 0) cell_kforward(&x); /* constant per cell */
  int nrec = 0;

  const int max_it = ta->num_time_iterations_per_interval;
  const int max_newton = ta->max_num_newton_iterations;
  const int jint = ta->jacobian_interval;
  double t = ta->tbeg, dt = dt_in;
  int fn_valid = 0, failed = 0;

  for (int iter = 0; iter < max_it && t < tend; ++iter) {
    int conv = 0;
    double h = 0.0;
    for (int attempt = 0; attempt < 4 && !conv; ++attempt) {
      h = dt > ta->dtmin ? dt : ta->dtmin;
      if (h > tend - t) h = tend - t;
      if (!(t + h > t)) break; /* no representable progress: hard failure */
      if (!fn_valid) { rhs(&x, un, fn); fn_valid = 1; }
      /* TR stage */
      for (int i = 0; i < m; ++i) { u[i] = un[i]; b[i] = un[i] + c_tr * h * fn[i]; }
      conv = newton(&x, u, b, c_tr * h, tol_newton, max_newton, jint, f, dx, st);
      if (conv) {
        /* BDF2 stage */
        for (int i = 0; i < m; ++i) { ug[i] = u[i]; b[i] = a_g * ug[i] - a_n * un[i]; }
        conv = newton(&x, u, b, c_bdf * h, tol_newton, max_newton, jint, f, dx, st);
      }
      if (!conv) {
        if (st) st->newton_failures++;
        dt = 0.5 * h;
      }
    }
    if (!conv) { failed = 1; break; }

    rhs(&x, ug, fg);
    rhs(&x, u, f1);
    double e = 0.0;
    for (int i = 0; i < m; ++i) {
      const double est = 2.0 * k_err * h * (fn[i] / g - fg[i] / (g * (1.0 - g)) + f1[i] / (1.0 - g));
      const double wi = fabs(est) / (tol_time[2 * i] + tol_time[2 * i + 1] * fabs(u[i]));
      if (!(wi <= e)) e = wi;
    }
    if (e > 1.0 && h > ta->dtmin) {
      if (st) st->rejected++;
      dt = 0.25 * h;
      continue;
    }
    t = (h == tend - t) ? tend : t + h;
    for (int i = 0; i < m; ++i) { un[i] = u[i]; fn[i] = f1[i]; }
    dt = (e <= 0.125) ? 2.0 * h : h;
    if (dt > ta->dtmax) dt = ta->dtmax;
    if (st) st->accepted++;
    if (h_rec) {
      if (nrec < max_rec) h_rec[nrec] = h;
      ++nrec;
    }
  }
  if (n_rec) *n_rec = (failed || nrec > max_rec) ? -1 : nrec;

  if (failed) {
    for (int i = 0; i < m; ++i) row[ROWCOL(i)] = 0.0;
    *t_out = tend;
    *dt_out = -1.0;
    if (st) st->failed = 1;
  } else {
    for (int i = 0; i < m; ++i) row[ROWCOL(i)] = un[i];
    *t_out = t;
    *dt_out = dt;
  }
}
#undef ROWCOL

void oracle_run_batch(const oracle_kmd* k, int nBatch, const double* tol_newton, const double* tol_time,
                      const oracle_time_advance* tadv, double* state, const double* photo, double* t_out,
                      double* dt_out, oracle_stats* stats, int jacobian_mode, int nthreads) {
  const int nSpec = k->nSpec, m = nSpec - k->nConstSpec;
  const size_t wlen = (size_t)k->nReac + nSpec + 2 * (size_t)m * m + 12 * (size_t)m;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    double* work = (double*)malloc(sizeof(double) * wlen);
    int* iwork = (int*)malloc(sizeof(int) * (size_t)m);
#pragma omp for schedule(static)
    for (int c = 0; c < nBatch; ++c) {
      integrate_cell(k, tol_newton, tol_time, &tadv[c], state + (size_t)c * (nSpec + 3),
                     photo ? photo + (size_t)c * k->nPhoto : NULL, &t_out[c], &dt_out[c],
                     stats ? &stats[c] : NULL, jacobian_mode, work, iwork, NULL, NULL, 0, NULL, NULL, NULL);
    }
    free(work);
    free(iwork);
  }
}

void oracle_sourceterm_batch(const oracle_kmd* k, int nBatch, const double* state, const double* photo,
                             double* omega) {
  const int nSpec = k->nSpec, m = nSpec - k->nConstSpec;
#pragma omp parallel
  {
    double* kfor = (double*)malloc(sizeof(double) * (size_t)k->nReac);
#pragma omp for schedule(static)
    for (int c = 0; c < nBatch; ++c) {
      const double* row = state + (size_t)c * (nSpec + 3);
      oracle_kforward(k, row[2], row[1], photo ? photo + (size_t)c * k->nPhoto : NULL, kfor);
      oracle_omega(k, kfor, row + 3, omega + (size_t)c * m);
    }
    free(kfor);
  }
}

void oracle_jacobian_batch(const oracle_kmd* k, int nBatch, const double* state, const double* photo,
                           double* jacv) {
  const int nSpec = k->nSpec, m = nSpec - k->nConstSpec;
#pragma omp parallel
  {
    double* kfor = (double*)malloc(sizeof(double) * (size_t)k->nReac);
#pragma omp for schedule(static)
    for (int c = 0; c < nBatch; ++c) {
      const double* row = state + (size_t)c * (nSpec + 3);
      oracle_kforward(k, row[2], row[1], photo ? photo + (size_t)c * k->nPhoto : NULL, kfor);
      oracle_jacobian(k, kfor, row + 3, jacv + (size_t)c * m * m);
    }
    free(kfor);
  }
}

/* ---------------------------------------------------------------------------
 * Derivatives of the discrete map (no counterpart in the reference).
 */

static size_t cell_work_len(const oracle_kmd* k) {
  const int m = k->nSpec - k->nConstSpec;
  return (size_t)k->nReac + k->nSpec + 2 * (size_t)m * m + 12 * (size_t)m;
}

/* One integrate_cell run (analytic Jacobian, rate constants scaled by
 * exp(lscale)) recording the accepted step sizes: returns the number of
 * accepted steps, or -1 when the cell failed or took more than max_steps.
 * row is integrated in place.
 * Pinned by: the integrator's own pins (the run is oracle_run_batch's). */
int oracle_record_steps(const oracle_kmd* k, const double* tol_newton, const double* tol_time,
                        const oracle_time_advance* ta, double* row, const double* photo, const double* lscale,
                        double* h_out, int max_steps) {
  const int m = k->nSpec - k->nConstSpec;
  double* work = (double*)malloc(sizeof(double) * cell_work_len(k));
  int* iwork = (int*)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
  double t_out = ta->tbeg, dt_out = ta->dt;
  int n = -1;
  integrate_cell(k, tol_newton, tol_time, ta, row, photo, &t_out, &dt_out, NULL, 0, work, iwork, lscale, h_out,
                 max_steps, &n, NULL, NULL);
  free(work);
  free(iwork);
  return n;
}

/* v += scale * nu_r * rop_r(c) over active species (reactants then products) */
static void add_nu_rop(const oracle_kmd* k, const double* kfor, const double* c, int r, double scale, double* v,
                       int stride) {
  const int m = k->nSpec - k->nConstSpec;
  const int ms = k->maxSpecInReac;
  double rop = kfor[r];
  for (int j = 0; j < k->reacNreac[r]; ++j) rop *= pow_int(c[k->reacSidx[r * ms + j]], abs((int)k->reacNuki[r * ms + j]));
  for (int j = 0; j < k->reacNreac[r]; ++j) {
    const int ks = k->reacSidx[r * ms + j];
    if (ks < m) v[(size_t)ks * stride] += scale * k->reacNuki[r * ms + j] * rop;
  }
  const int joff = ms / 2;
  for (int j = 0; j < k->reacNprod[r]; ++j) {
    const int ks = k->reacSidx[r * ms + j + joff];
    if (ks < m) v[(size_t)ks * stride] += scale * k->reacNuki[r * ms + j + joff] * rop;
  }
}

/* Integrates one cell over a given sequence of step sizes h[0..nsteps) with
 * the oracle's TR-BDF2 stages and Newton solver (no error control), and
 * propagates the discrete tangent linear S = du/dp of nsens parameters:
 * sens[c] >= 0 is the log-scale theta_r of reaction r (kf_r * exp(theta_r),
 * d f / d theta_r = nu_r rop_r), sens[c] < 0 the initial value of active
 * species -sens[c]-1.  Per step, with J0, Jg, J1 the Jacobians at u_n, u_g,
 * u_{n+1} and F(u) the parameter columns of d f / d p,
 *   (I - c_tr h Jg) Sg = (I + c_tr h J0) S + c_tr h (F(u_g) + F(u_n))
 *   (I - c_bdf h J1) S' = a_g Sg - a_n S + c_bdf h F(u_{n+1}),
 * the derivative of the stage equations solved exactly.
 *   row: [nSpec+3] initial state (not modified); u_out [m]; S_out [m][nsens]
 *   (nullable).  Returns 0, or -1 when a Newton solve failed.
 * Pinned by: central differences of this same function in the parameters
 * (tests/test_oracle_adjoint.py), and on y' = lambda y by the closed-form
 * derivative of the TR-BDF2 amplification factor. */
int oracle_tlm_fixed_steps(const oracle_kmd* k, const double* tol_newton, int max_newton, int jint,
                           const double* row, const double* photo, const double* lscale, const double* h, int nsteps,
                           int nsens, const int* sens, double* u_out, double* S_out) {
  const int nSpec = k->nSpec, m = nSpec - k->nConstSpec;
  double g6[6];
  oracle_trbdf2_constants(g6);
  const double c_tr = g6[1], a_g = g6[2], a_n = g6[3], c_bdf = g6[4];
  cell_ctx x;
  memset(&x, 0, sizeof x); /* gas-phase cell: no aerosol model */
  x.k = k;
  x.mg = m;
  x.m = m;
  x.mode = 0;
  x.p = row[1];
  x.T = row[2];
  x.photo = photo;
  x.lscale = lscale;
  double* work = (double*)malloc(sizeof(double) * cell_work_len(k));
  int* iwork = (int*)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
  const size_t mm = (size_t)m * m, ms = (size_t)m * (nsens > 0 ? nsens : 1);
  double* big = (double*)calloc(3 * mm + 5 * ms + 8 * (size_t)m, sizeof(double));
  double* w = work;
  x.kfor = w; w += k->nReac;
  x.c = w; w += nSpec;
  x.omega = w; w += m;
  x.J = w; w += mm;
  x.A = w; w += mm;
  x.fh = w; w += m;
  x.piv = iwork;
  double* J0 = big;
  double* Jg = J0 + mm;
  double* M = Jg + mm;
  double* S = M + mm;
  double* Sg = S + ms;
  double* R = Sg + ms;
  double* col = R + ms;
  double* tmp = col + ms;
  double* un = tmp + ms;
  double* u = un + m;
  double* ug = u + m;
  double* fn = ug + m;
  double* f = fn + m;
  double* dx = f + m;
  double* b = dx + m;
  int* piv = (int*)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
  for (int i = 0; i < nSpec; ++i) x.c[i] = row[3 + i];
  for (int i = 0; i < m; ++i) un[i] = row[3 + i];
  cell_kforward(&x);
  for (int c = 0; c < nsens; ++c)
    for (int i = 0; i < m; ++i) S[(size_t)i * nsens + c] = (sens[c] < 0 && i == -sens[c] - 1) ? 1.0 : 0.0;
  int rc = 0;
  rhs(&x, un, fn);
  /* F(u) columns added with a scale into R ([m][nsens]) */
  #define ADD_F(uvec, scale)                                                   \
    do {                                                                       \
      memcpy(x.c, (uvec), sizeof(double) * m);                                 \
      for (int cc = 0; cc < nsens; ++cc)                                       \
        if (sens[cc] >= 0) add_nu_rop(k, x.kfor, x.c, sens[cc], (scale), R + cc, nsens); \
    } while (0)
  for (int s = 0; s < nsteps && rc == 0; ++s) {
    const double hh = h[s];
    for (int i = 0; i < m; ++i) { u[i] = un[i]; b[i] = un[i] + c_tr * hh * fn[i]; }
    if (!newton(&x, u, b, c_tr * hh, tol_newton, max_newton, jint, f, dx, NULL)) { rc = -1; break; }
    for (int i = 0; i < m; ++i) { ug[i] = u[i]; b[i] = a_g * ug[i] - a_n * un[i]; }
    if (!newton(&x, u, b, c_bdf * hh, tol_newton, max_newton, jint, f, dx, NULL)) { rc = -1; break; }
    if (S_out && nsens > 0) {
      /* R = (I + c_tr h J0) S + c_tr h (F(u_g) + F(u_n)) */
      memcpy(x.c, un, sizeof(double) * m);
      oracle_jacobian(k, x.kfor, x.c, J0);
      for (int i = 0; i < m; ++i)
        for (int c = 0; c < nsens; ++c) {
          double acc = S[(size_t)i * nsens + c];
          for (int j = 0; j < m; ++j) acc += c_tr * hh * J0[(size_t)i * m + j] * S[(size_t)j * nsens + c];
          R[(size_t)i * nsens + c] = acc;
        }
      ADD_F(ug, c_tr * hh);
      ADD_F(un, c_tr * hh);
      memcpy(x.c, ug, sizeof(double) * m);
      oracle_jacobian(k, x.kfor, x.c, Jg);
      for (size_t e = 0; e < mm; ++e) M[e] = -c_tr * hh * Jg[e];
      for (int i = 0; i < m; ++i) M[(size_t)i * m + i] += 1.0;
      if (oracle_lu(m, M, m, piv) != 0) { rc = -1; break; }
      for (int c = 0; c < nsens; ++c) {
        for (int i = 0; i < m; ++i) col[i] = R[(size_t)i * nsens + c];
        oracle_lu_solve(m, M, m, piv, col);
        for (int i = 0; i < m; ++i) Sg[(size_t)i * nsens + c] = col[i];
      }
      /* R = a_g Sg - a_n S + c_bdf h F(u_{n+1}) */
      for (size_t e = 0; e < ms; ++e) R[e] = a_g * Sg[e] - a_n * S[e];
      ADD_F(u, c_bdf * hh);
      memcpy(x.c, u, sizeof(double) * m);
      oracle_jacobian(k, x.kfor, x.c, Jg);
      for (size_t e = 0; e < mm; ++e) M[e] = -c_bdf * hh * Jg[e];
      for (int i = 0; i < m; ++i) M[(size_t)i * m + i] += 1.0;
      if (oracle_lu(m, M, m, piv) != 0) { rc = -1; break; }
      for (int c = 0; c < nsens; ++c) {
        for (int i = 0; i < m; ++i) col[i] = R[(size_t)i * nsens + c];
        oracle_lu_solve(m, M, m, piv, col);
        for (int i = 0; i < m; ++i) S[(size_t)i * nsens + c] = col[i];
      }
    }
    for (int i = 0; i < m; ++i) un[i] = u[i];
    rhs(&x, un, fn);
  }
  #undef ADD_F
  (void)tmp;
  if (u_out) memcpy(u_out, un, sizeof(double) * m);
  if (S_out && nsens > 0) memcpy(S_out, S, sizeof(double) * ms);
  free(piv);
  free(big);
  free(work);
  free(iwork);
  return rc;
}

/* ---------------------------------------------------------------------------
 * Single-particle SIMPOL.1 gas-particle partitioning
 */
#define SP_RUNIV 8.314472          /* src/core/TChem_Util.hpp:176 */
#define SP_PI 3.141592653589793    /* TChem_Util.hpp:212 */
#define SP_ATMPA 101325.0          /* TChem_Util.hpp:227-229 */
/* src/core/impl/TChem_Impl_SingleParticleUtils.hpp:25-30 (long double
 * literals there; double here, a difference below 1e-16 of a 1e-25 term) */
#define SP_MIN_MASS 1.0e-25
#define SP_MIN_MW 0.1
#define SP_MIN_DENSITY 1800.0

/* src/core/impl/TChem_Impl_SIMPOL_constant.hpp:60-128: mass accommodation
 * (Ervens et al. 2003) when N* is given, else 0.1; mean free path
 * 3 D / sqrt(8 R T / (pi MW)) (TChem_Util.hpp:1211-1238); SIMPOL.1 vapour
 * pressure; kg m-3 -> ppm factor; partitioning constant.  Same expression
 * order as the reference.  Pinned by the SIMPOL RHS golden
 * (Simpol_RHS_HOST.txt) through oracle_aerosol_omega. */
void oracle_simpol_constants(const double* par, double T, double p, double* out4) {
  const double B1 = par[0], B2 = par[1], B3 = par[2], B4 = par[3], D = par[4], MW = par[5], Nstar = par[6];
  double alpha = 0.1;
  if (par[7] != 0.0) {
    const double N23 = pow(Nstar, 2.0 / 3.0);
    double dh = -10.0 * (Nstar - 1.0) + 7.53 * (N23 - 1.0) - 1.0;
    double ds = -32.0 * (Nstar - 1.0) + 9.21 * (N23 - 1.0) - 1.3;
    dh = dh * 4184.0;
    ds = ds * 4.184;
    const double dG = dh - T * ds;
    alpha = exp(-dG / (SP_RUNIV * T));
    alpha = alpha / (1.0 + alpha);
  }
  const double mfp = 3.0 * D / sqrt(8.0 * SP_RUNIV * T / (SP_PI * MW));
  double vp = B1 / T + B2 + B3 * T + B4 * log(T);
  vp = SP_ATMPA * pow(10, vp);
  const double conv = SP_RUNIV / MW * 1.0e6;
  out4[0] = alpha;
  out4[1] = mfp;
  out4[2] = conv * T / p;
  out4[3] = vp / p / MW * 1.0e6;
}

/* per-particle sums of SingleParticleUtils.hpp:62-111: volume -> effective
 * radius, total mass and moles */
static void particle_sums(const oracle_amd* a, const double* y, double* radius, double* volume, double* mass,
                          double* moles) {
  double v = SP_MIN_MASS / SP_MIN_DENSITY;
  for (int i = 0; i < a->nSpec; ++i) v += y[i] / a->density[i];
  *volume = v;
  *radius = pow((v * 3.0 / 4.0 / SP_PI), 1.0 / 3.0);
  double ms = SP_MIN_MASS, mo = SP_MIN_MASS / SP_MIN_MW;
  for (int i = 0; i < a->nSpec; ++i) {
    ms += y[i];
    mo += y[i] / a->mw[i];
  }
  *mass = ms;
  *moles = mo;
}

/* Fuchs-Sutugin transition-regime factor (SingleParticleUtils.hpp:128-135) */
static double fuchs_sutugin(double Kn, double alpha) {
  return (0.75 * alpha * (1.0 + Kn)) / (Kn * Kn + (1.0 + 0.283 * alpha) * Kn + 0.75 * alpha);
}

/* src/core/impl/TChem_Impl_Aerosol_RHS.hpp:64-103: omega := 0, gas-phase
 * source term over c = [x_gas, const] (ReactionRatesAerosol), then for each
 * particle and each SIMPOL transfer (SIMPOL_phase_transfer.hpp:61-150):
 *   k_c = 4 pi r D f(lambda / r, alpha),  k_e = k_c (K_eq MW_avg / mass)
 *   diff = k_c x_gas - k_e y_aero
 *   omega_gas -= N_p diff,  omega_aero += diff / (kg m-3 -> ppm)
 * accumulated serially in (particle, transfer) order.
 * Pinned by: the SIMPOL RHS golden, a mass-balance invariant and the
 * equilibrium zero (tests/test_oracle_aerosol.py). */
void oracle_aerosol_omega(const oracle_kmd* k, const oracle_amd* a, double T, double p, const double* kfor,
                          const double* cconst, const double* nc, const double* x, double* omega) {
  const int mg = k->nSpec - k->nConstSpec, mt = mg + a->nSpec * a->nParticles;
  double* c = (double*)malloc(sizeof(double) * (size_t)(k->nSpec > 0 ? k->nSpec : 1));
  for (int i = 0; i < mg; ++i) c[i] = x[i];
  for (int i = mg; i < k->nSpec; ++i) c[i] = cconst[i - mg];
  for (int i = 0; i < mt; ++i) omega[i] = 0.0;
  oracle_omega(k, kfor, c, omega);
  free(c);
  for (int ip = 0; ip < a->nParticles; ++ip) {
    const double* y = x + mg + (size_t)ip * a->nSpec;
    double radius, volume, mass, moles;
    particle_sums(a, y, &radius, &volume, &mass, &moles);
    const double avgMW = mass / moles;
    for (int s = 0; s < a->nSimpol; ++s) {
      double c4[4];
      oracle_simpol_constants(a->par + 8 * s, T, p, c4);
      const double D = a->par[8 * s + 4];
      const double Kn = c4[1] / radius;
      double cond = 4.0 * SP_PI * radius * D * fuchs_sutugin(Kn, c4[0]);
      double evap = cond * 1.0 * (c4[3] * avgMW / mass);
      const int g = a->gas[s], ai = mg + ip * a->nSpec + a->aero[s];
      cond *= x[g];
      evap *= x[ai];
      const double diff = -evap + cond;
      omega[g] += -nc[ip] * diff;
      omega[ai] += diff / c4[2];
    }
  }
}

/* Analytic Jacobian of oracle_aerosol_omega.  The reference uses a
 * forward-difference Jacobian for this problem
 * (TChem_Impl_AerosolChemistry_Problem.hpp:211-239); both solve the same
 * Newton system.  With r = (3V / 4 pi)^(1/3), V = V0 + sum y_i / rho_i,
 * Kn = lambda / r:
 *   d k_c / d y_i = 4 pi D (f - Kn f'(Kn)) (r / 3V) / rho_i
 *   k_e = k_c K_eq / moles, d k_e / d y_i = K_eq (d k_c / d y_i) / moles
 *                                           - k_e / (moles MW_i)
 *   d diff / d x_gas = k_c,  d diff / d y_i = x_gas d k_c - y_aero d k_e - [i = aero] k_e
 * Pinned by: central differences of oracle_aerosol_omega
 * (tests/test_oracle_aerosol.py). */
void oracle_aerosol_jacobian(const oracle_kmd* k, const oracle_amd* a, double T, double p, const double* kfor,
                             const double* cconst, const double* nc, const double* x, double* J) {
  const int mg = k->nSpec - k->nConstSpec, mt = mg + a->nSpec * a->nParticles;
  memset(J, 0, sizeof(double) * (size_t)mt * mt);
  double* c = (double*)malloc(sizeof(double) * ((size_t)(k->nSpec > 0 ? k->nSpec : 1) + (size_t)mg * mg + 2 * (size_t)a->nSpec + 1));
  double* Jg = c + (k->nSpec > 0 ? k->nSpec : 1);
  double* dkc = Jg + (size_t)mg * mg;
  double* dke = dkc + a->nSpec;
  for (int i = 0; i < mg; ++i) c[i] = x[i];
  for (int i = mg; i < k->nSpec; ++i) c[i] = cconst[i - mg];
  if (mg > 0) oracle_jacobian(k, kfor, c, Jg);
  for (int i = 0; i < mg; ++i)
    for (int j = 0; j < mg; ++j) J[(size_t)i * mt + j] = Jg[(size_t)i * mg + j];
  for (int ip = 0; ip < a->nParticles; ++ip) {
    const int base = mg + ip * a->nSpec;
    const double* y = x + base;
    double radius, volume, mass, moles;
    particle_sums(a, y, &radius, &volume, &mass, &moles);
    const double avgMW = mass / moles;
    for (int s = 0; s < a->nSimpol; ++s) {
      double c4[4];
      oracle_simpol_constants(a->par + 8 * s, T, p, c4);
      const double D = a->par[8 * s + 4], alpha = c4[0], Keq = c4[3];
      const double Kn = c4[1] / radius;
      const double den = Kn * Kn + (1.0 + 0.283 * alpha) * Kn + 0.75 * alpha;
      const double f = fuchs_sutugin(Kn, alpha);
      const double fp = 0.75 * alpha * (den - (1.0 + Kn) * (2.0 * Kn + 1.0 + 0.283 * alpha)) / (den * den);
      const double kc = 4.0 * SP_PI * radius * D * f;
      const double ke = kc * 1.0 * (Keq * avgMW / mass);
      const double dkc_dV = 4.0 * SP_PI * D * (f - Kn * fp) * (radius / (3.0 * volume));
      for (int i = 0; i < a->nSpec; ++i) {
        dkc[i] = dkc_dV / a->density[i];
        dke[i] = Keq * dkc[i] / moles - ke / (moles * a->mw[i]);
      }
      const int g = a->gas[s], ai = base + a->aero[s];
      const double xg = x[g], ya = x[ai];
      /* d diff / d x_g = kc */
      J[(size_t)g * mt + g] += -nc[ip] * kc;
      J[(size_t)ai * mt + g] += kc / c4[2];
      for (int i = 0; i < a->nSpec; ++i) {
        double dd = xg * dkc[i] - ya * dke[i];
        if (base + i == ai) dd -= ke;
        J[(size_t)g * mt + base + i] += -nc[ip] * dd;
        J[(size_t)ai * mt + base + i] += dd / c4[2];
      }
    }
  }
  free(c);
}

static void aerosol_split_row(const oracle_kmd* k, const oracle_amd* a, const double* row, double* x) {
  const int mg = k->nSpec - k->nConstSpec;
  for (int i = 0; i < mg; ++i) x[i] = row[3 + i];
  for (int i = 0; i < a->nSpec * a->nParticles; ++i) x[mg + i] = row[3 + k->nSpec + i];
}

void oracle_aerosol_rhs_batch(const oracle_kmd* k, const oracle_amd* a, int nBatch, const double* state,
                              const double* nc, double* omega) {
  const int mg = k->nSpec - k->nConstSpec, mt = mg + a->nSpec * a->nParticles;
  const int dim = 3 + k->nSpec + a->nSpec * a->nParticles;
  double* kfor = (double*)malloc(sizeof(double) * ((size_t)k->nReac + 1));
  double* x = (double*)malloc(sizeof(double) * ((size_t)mt + 1));
  for (int c = 0; c < nBatch; ++c) {
    const double* row = state + (size_t)c * dim;
    oracle_kforward(k, row[2], row[1], NULL, kfor);
    aerosol_split_row(k, a, row, x);
    oracle_aerosol_omega(k, a, row[2], row[1], kfor, row + 3 + mg, nc + (size_t)c * a->nParticles, x,
                         omega + (size_t)c * mt);
  }
  free(kfor);
  free(x);
}

void oracle_aerosol_run_batch(const oracle_kmd* k, const oracle_amd* a, int nBatch, const double* tol_newton,
                              const double* tol_time, const oracle_time_advance* tadv, double* state,
                              const double* nc, double* t_out, double* dt_out, oracle_stats* stats, int nthreads) {
  const int mg = k->nSpec - k->nConstSpec, mt = mg + a->nSpec * a->nParticles;
  const int dim = 3 + k->nSpec + a->nSpec * a->nParticles;
  const size_t wlen = (size_t)k->nReac + k->nSpec + 2 * (size_t)mt * mt + 12 * (size_t)mt + 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    double* work = (double*)malloc(sizeof(double) * wlen);
    int* iwork = (int*)malloc(sizeof(int) * ((size_t)mt + 1));
#pragma omp for schedule(static)
    for (int c = 0; c < nBatch; ++c) {
      integrate_cell(k, tol_newton, tol_time, &tadv[c], state + (size_t)c * dim, NULL, &t_out[c], &dt_out[c],
                     stats ? &stats[c] : NULL, 0, work, iwork, NULL, NULL, 0, NULL, a,
                     nc + (size_t)c * a->nParticles);
    }
    free(work);
    free(iwork);
  }
}
