// HIP kernels for CDNA4 (gfx950): rate constants, source term, analytic
// Jacobian and the TR-BDF2 box-model integrator.
//
// Mapping: one cell per 64-lane wavefront (one wavefront per workgroup).
//   * per-species vectors (state, stage values, residuals, Newton updates)
//     live in registers, species i on lane i % 64, register slot i / 64;
//   * data other lanes must read goes through LDS: concentrations, rate
//     constants, rate-of-progress / derivative scratch, and the dense
//     iteration matrix I - c h J with its LU factors;
//   * the LU (partial pivoting) maps lanes to the columns of the trailing
//     submatrix; the pivot search is a wavefront arg-max over shuffles; a
//     row whose multiplier is exactly zero is skipped by the whole wavefront
//     (ballot), so the structural sparsity of the chemistry Jacobian costs
//     nothing without fixing a pivot order.
// The iteration matrix and its factorisation are in TA precision (double for
// the all-FP64 form, float for the mixed form); everything else is FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "device_mech.hpp"

namespace tchem_amd {

namespace {

constexpr int kWave = 64;

__device__ __forceinline__ void wave_sync() { __syncthreads(); }  // workgroup == one wavefront

__device__ __forceinline__ double readlane(double v, int l) {
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(x & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane(int(x >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ float readlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
  return v;
}

__device__ __forceinline__ double pow_small(double x, int n) {
  return n == 1 ? x : (n == 2 ? x * x : pow(x, double(n)));
}

// Rate constant of reaction r: src/core/impl/TChem_Impl_KForwardAerosol.hpp:54-131.
__device__ double rate_constant(const DevMech& M, int r, double t, double p, double t_1, double conv,
                                const double* photo) {
  const double* P = M.ratePar + 8 * r;
  const double one = 1.0;
  const int kind = M.rateKind[r];
  double k;
  if (kind == 0) {  // Arrhenius
    double A = P[0];
    const int slot = M.photoSlot[r];
    if (photo != nullptr && slot >= 0) A = photo[slot];
    k = A * exp(P[2] * t_1) * (P[1] == 0.0 ? one : pow(t / P[3], P[1])) * (P[4] == 0.0 ? one : (one + P[4] * p));
  } else if (kind == 1) {  // Troe
    const double k0 = P[0] * (P[2] == 0.0 ? one : exp(P[2] * t_1)) * (P[1] == 0.0 ? one : pow(t / 300.0, P[1])) * conv;
    const double k0_kinf =
        k0 / (P[3] * (P[5] == 0.0 ? one : exp(P[5] * t_1)) * (P[4] == 0.0 ? one : pow(t / 300.0, P[4])));
    const double lg = log10(k0_kinf) / P[7];
    k = k0 / (1.0 + k0_kinf) * pow(P[6], 1.0 / (1.0 + lg * lg));
  } else {  // CMAQ_H2O2
    k = P[0] * exp(P[2] * t_1) * (P[1] == 0.0 ? one : pow(t / 300.0, P[1])) +
        P[3] * exp(P[5] * t_1) * (P[4] == 0.0 ? one : pow(t / 300.0, P[4])) * conv;
  }
  return k * pow(conv, double(M.convExp[r]));
}

template <typename TA>
struct CellLds {
  TA* A;       // [m][m] iteration matrix / LU factors (row-major, lda = m)
  double* c;   // [nSpec] concentrations: active (evaluation point) then constant
  double* kf;  // [nReac] rate constants
  double* scr; // [max(nReac, nPairs)] rate of progress or derivative pairs
  int* piv;    // [m]
};

template <typename TA>
__device__ CellLds<TA> carve(char* base, const DevMech& M, bool withMatrix = true) {
  CellLds<TA> L;
  size_t off = withMatrix ? sizeof(TA) * size_t(M.m) * size_t(M.m) : 0;
  off = (off + 15) & ~size_t(15);
  L.A = reinterpret_cast<TA*>(base);
  L.c = reinterpret_cast<double*>(base + off);
  L.kf = L.c + M.nSpec;
  L.scr = L.kf + M.nReac;
  const int nScr = M.nReac > M.nPairs ? M.nReac : M.nPairs;
  L.piv = reinterpret_cast<int*>(L.scr + nScr);
  return L;
}

// rate constants of one cell into LDS
template <typename TA>
__device__ void load_rate_constants(const DevMech& M, const CellLds<TA>& L, int lane, double T, double p,
                                    const double* photo) {
  const double t_1 = 1.0 / T;
  const double conv = M.convPPM * p * t_1;
  for (int r = lane; r < M.nReac; r += kWave) L.kf[r] = rate_constant(M, r, T, p, t_1, conv, photo);
}

// rate of progress of reaction r at the concentrations in LDS
__device__ __forceinline__ double rate_of_progress(const DevMech& M, const double* c, const double* kf, int r) {
  double v = kf[r];
  const int* sp = M.reacSpec + r * M.maxR;
  const int* pw = M.reacPow + r * M.maxR;
  for (int j = 0; j < M.maxR; ++j) {
    const int s = sp[j];
    if (s < 0) break;
    v *= pow_small(c[s], pw[j]);
  }
  return v;
}

// omega at the concentrations already in LDS (c); result in registers,
// species i = lane + 64 q.  src/core/impl/TChem_Impl_ReactionRatesAerosol.hpp:78-138
template <int QM, typename TA>
__device__ void source_term_from_lds(const DevMech& M, const CellLds<TA>& L, int lane, double (&f)[QM]) {
  for (int r = lane; r < M.nReac; r += kWave) L.scr[r] = rate_of_progress(M, L.c, L.kf, r);
  wave_sync();
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int i = lane + kWave * q;
    double s = 0.0;
    if (i < M.m) {
      s = M.emission[i];
      const int e1 = M.omegaPtr[i + 1];
      for (int e = M.omegaPtr[i]; e < e1; ++e) s += M.omegaCoef[e] * L.scr[M.omegaReac[e]];
    }
    f[q] = s;
  }
  wave_sync();
}

template <int QM, typename TA>
__device__ void eval_rhs(const DevMech& M, const CellLds<TA>& L, int lane, const double (&u)[QM], double (&f)[QM]) {
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int i = lane + kWave * q;
    if (i < M.m) L.c[i] = u[q];
  }
  wave_sync();
  source_term_from_lds<QM>(M, L, lane, f);
}

// Assemble A = diag * I + scale * J at the concentrations in LDS.
template <typename TA>
__device__ void assemble_jacobian_from_lds(const DevMech& M, const CellLds<TA>& L, int lane, double diag,
                                           double scale) {
  const int m = M.m;
  const int total = m * m;
  for (int idx = lane; idx < total; idx += kWave) L.A[idx] = TA(0);
  // derivative pairs: d rop_r / d c_s = k * n_s c_s^(n_s-1) * prod_{j != s} c_j^n_j
  for (int r = lane; r < M.nReac; r += kWave) {
    const int p0 = M.pairPtr[r], p1 = M.pairPtr[r + 1];
    if (p0 == p1) continue;
    const int* sp = M.reacSpec + r * M.maxR;
    const int* pw = M.reacPow + r * M.maxR;
    for (int p = p0; p < p1; ++p) {
      const int js = M.pairSlot[p];
      const int ns = pw[js];
      double d = L.kf[r] * double(ns) * pow_small(L.c[sp[js]], ns - 1);
      for (int j = 0; j < M.maxR; ++j) {
        const int s = sp[j];
        if (s < 0) break;
        if (j == js) continue;
        d *= pow_small(L.c[s], pw[j]);
      }
      L.scr[p] = d;
    }
  }
  wave_sync();
  for (int n = lane; n < M.nnzJ; n += kWave) {
    double s = 0.0;
    const int e1 = M.jnzPtr[n + 1];
    for (int e = M.jnzPtr[n]; e < e1; ++e) s += M.jcCoef[e] * L.scr[M.jcPair[e]];
    const int rc = M.jnzRowCol[n];
    const int row = rc >> 16, col = rc & 0xffff;
    L.A[row * m + col] = TA((row == col ? diag : 0.0) + scale * s);
  }
  wave_sync();
}

template <int QM, typename TA>
__device__ void eval_iteration_matrix(const DevMech& M, const CellLds<TA>& L, int lane, const double (&u)[QM],
                                      double ch) {
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int i = lane + kWave * q;
    if (i < M.m) L.c[i] = u[q];
  }
  wave_sync();
  assemble_jacobian_from_lds(M, L, lane, 1.0, -ch);
}

// In-place LU with partial pivoting of the m x m matrix in LDS.  Returns
// false for a zero (or non-finite) pivot.  Mirrors oracle_lu's arithmetic:
// multipliers l = a_ik / a_kk, trailing update a_ij -= l * a_kj.
template <int QM, typename TA>
__device__ bool lu_factor(const CellLds<TA>& L, int m, int lane) {
  TA* A = L.A;
  bool ok = true;
  for (int k = 0; k < m; ++k) {
    // pivot search: smallest index among the largest |a_ik|, i >= k
    TA best = TA(-1);
    int bi = m;
    for (int i = k + lane; i < m; i += kWave) {
      const TA v = fabs(A[i * m + k]);
      if (v > best) { best = v; bi = i; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const TA ob = __shfl_xor(best, off, kWave);
      const int oi = __shfl_xor(bi, off, kWave);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    const int p = __builtin_amdgcn_readfirstlane(bi);
    if (p >= m || !(best > TA(0))) { ok = false; break; }
    if (lane == 0) L.piv[k] = p;
    if (p != k)
      for (int j = lane; j < m; j += kWave) {
        const TA t = A[k * m + j];
        A[k * m + j] = A[p * m + j];
        A[p * m + j] = t;
      }
    wave_sync();
    const TA akk = A[k * m + k];
    TA l[QM], uk[QM];
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = k + 1 + lane + kWave * q;
      l[q] = TA(0);
      uk[q] = TA(0);
      if (i < m) {
        l[q] = A[i * m + k] / akk;
        A[i * m + k] = l[q];
        uk[q] = A[k * m + i];  // pivot row, column i (same offset pattern)
      }
    }
    // rank-1 update of the trailing submatrix, one row at a time for the
    // rows with a nonzero multiplier; lanes own columns k+1+lane+64c
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      unsigned long long mask = __ballot(l[q] != TA(0));
      while (mask) {
        const int b = __builtin_ctzll(mask);
        mask &= mask - 1;
        const TA li = readlane(l[q], b);
        const int i = k + 1 + kWave * q + b;
#pragma unroll
        for (int c = 0; c < QM; ++c) {
          const int j = k + 1 + lane + kWave * c;
          if (j < m) A[i * m + j] -= li * uk[c];
        }
      }
    }
    wave_sync();
  }
  wave_sync();
  return ok;
}

template <int QM, typename TA>
__device__ __forceinline__ TA bcast(const TA (&v)[QM], int idx) {
  const int q = idx / kWave, l = idx % kWave;
  TA out = TA(0);
#pragma unroll
  for (int s = 0; s < QM; ++s)
    if (s == q) out = readlane(v[s], l);
  return out;
}

// Solve (LU) x = b with b distributed over lanes (row i = lane + 64 q).
template <int QM, typename TA>
__device__ void lu_solve(const CellLds<TA>& L, int m, int lane, TA (&b)[QM]) {
  const TA* A = L.A;
  for (int k = 0; k < m; ++k) {
    const int p = L.piv[k];
    if (p != k) {
      const TA bk = bcast<QM>(b, k), bp = bcast<QM>(b, p);
#pragma unroll
      for (int q = 0; q < QM; ++q) {
        const int i = lane + kWave * q;
        if (i == k) b[q] = bp;
        else if (i == p) b[q] = bk;
      }
    }
  }
  for (int k = 0; k < m - 1; ++k) {
    const TA yk = bcast<QM>(b, k);
    if (yk == TA(0)) continue;
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      if (i > k && i < m) b[q] -= A[i * m + k] * yk;
    }
  }
  for (int k = m - 1; k >= 0; --k) {
    const TA xk = bcast<QM>(b, k) / A[k * m + k];
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      if (i == k) b[q] = xk;
    }
    if (xk == TA(0)) continue;
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      if (i < k) b[q] -= A[i * m + k] * xk;
    }
  }
}

struct CellCounters {
  int accepted = 0, rejected = 0, newton_failures = 0, newton_iterations = 0, factorizations = 0, failed = 0;
};

// Newton iteration for one TR-BDF2 stage: G(u) = u - b - ch f(u), iteration
// matrix I - ch J refreshed every jacobian_interval iterations; converged when
// max_i |dx_i| / (atol + rtol |u_i|) <= 1 (oracle/oracle.c: newton).
template <int QM, typename TA>
__device__ bool newton(const DevMech& M, const CellLds<TA>& L, int lane, double (&u)[QM], const double (&bc)[QM],
                       double ch, double atol, double rtol, int max_it, int jint, CellCounters& cnt) {
  const int m = M.m;
  if (jint < 1) jint = 1;
  for (int it = 0; it < max_it; ++it) {
    if (it % jint == 0) {
      eval_iteration_matrix<QM>(M, L, lane, u, ch);
      if (!lu_factor<QM>(L, m, lane)) return false;
      ++cnt.factorizations;
    }
    double f[QM];
    eval_rhs<QM>(M, L, lane, u, f);
    TA x[QM];
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      x[q] = (i < m) ? TA(-(u[q] - bc[q] - ch * f[q])) : TA(0);
    }
    lu_solve<QM>(L, m, lane, x);
    double nrm = 0.0;
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      if (i < m) {
        const double dx = double(x[q]);
        u[q] += dx;
        double w = fabs(dx) / (atol + rtol * fabs(u[q]));
        if (!(w == w)) w = INFINITY;
        nrm = fmax(nrm, w);
      }
    }
    nrm = wave_max(nrm);
    ++cnt.newton_iterations;
    if (nrm <= 1.0) return true;
  }
  return false;
}

// One cell of AtmosphericChemistry::runDeviceBatch
// (src/core/TChem_AtmosphericChemistry.cpp:78-163) with the TR-BDF2 step
// control of oracle/oracle.c: integrate_cell.
template <typename TA, int QM>
__device__ void integrate_cell(const IntegratorArgs& a, const CellLds<TA>& L, int cell, int lane) {
  const DevMech& M = a.M;
  const int m = M.m, nSpec = M.nSpec, dim = nSpec + 3;
  const tchem_time_advance_t ta = a.tadv[cell];
  const double tend = ta._tend;
  double* row = a.state + size_t(cell) * dim;
  CellCounters cnt;

  const double t_in = a.t_out[cell];
  if (t_in >= tend || ta._dt < 0.0) {  // skipped (.cpp:83) or already failed
    if (a.stats && lane == 0) a.stats[cell] = tchem_cell_stats_t{0, 0, 0, 0, 0, ta._dt < 0.0 ? 1 : 0};
    return;
  }

  const double gam = 2.0 - 1.41421356237309504880;
  const double c_tr = 0.5 * gam;
  const double a_g = 1.0 / (gam * (2.0 - gam));
  const double a_n = (1.0 - gam) * (1.0 - gam) / (gam * (2.0 - gam));
  const double c_bdf = (1.0 - gam) / (2.0 - gam);
  const double k_err = (-3.0 * gam * gam + 4.0 * gam - 2.0) / (12.0 * (2.0 - gam));

  const double p = row[1], T = row[2];
  load_rate_constants(M, L, lane, T, p, a.photo ? a.photo + size_t(cell) * M.nPhoto : nullptr);
  for (int i = m + lane; i < nSpec; i += kWave) L.c[i] = row[3 + i];

  double un[QM], u[QM], ug[QM], fn[QM], fg[QM], f1[QM], bc[QM];
  double atl[QM], rtl[QM];
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int i = lane + kWave * q;
    un[q] = (i < m) ? row[3 + i] : 0.0;
    atl[q] = (i < m) ? a.tol_time[2 * i] : 1.0;
    rtl[q] = (i < m) ? a.tol_time[2 * i + 1] : 0.0;
    fn[q] = 0.0;
  }
  const double atol_n = a.tol_newton[0], rtol_n = a.tol_newton[1];
  wave_sync();

  const int max_it = ta._num_time_iterations_per_interval;
  const int max_newton = ta._max_num_newton_iterations;
  const int jint = ta._jacobian_interval;
  double t = ta._tbeg, dt = ta._dt;
  bool fn_valid = false, failed = false;

  for (int iter = 0; iter < max_it && t < tend; ++iter) {
    bool conv = false;
    double h = 0.0;
    for (int attempt = 0; attempt < 4 && !conv; ++attempt) {
      h = dt > ta._dtmin ? dt : ta._dtmin;
      if (h > tend - t) h = tend - t;
      if (!(t + h > t)) break;  // no representable progress: hard failure
      if (!fn_valid) {
        eval_rhs<QM>(M, L, lane, un, fn);
        fn_valid = true;
      }
#pragma unroll
      for (int q = 0; q < QM; ++q) {
        u[q] = un[q];
        bc[q] = un[q] + c_tr * h * fn[q];
      }
      conv = newton<QM>(M, L, lane, u, bc, c_tr * h, atol_n, rtol_n, max_newton, jint, cnt);
      if (conv) {
#pragma unroll
        for (int q = 0; q < QM; ++q) {
          ug[q] = u[q];
          bc[q] = a_g * ug[q] - a_n * un[q];
        }
        conv = newton<QM>(M, L, lane, u, bc, c_bdf * h, atol_n, rtol_n, max_newton, jint, cnt);
      }
      if (!conv) {
        ++cnt.newton_failures;
        dt = 0.5 * h;
      }
    }
    if (!conv) {
      failed = true;
      break;
    }
    eval_rhs<QM>(M, L, lane, ug, fg);
    eval_rhs<QM>(M, L, lane, u, f1);
    double e = 0.0;
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      if (i < m) {
        const double est = 2.0 * k_err * h * (fn[q] / gam - fg[q] / (gam * (1.0 - gam)) + f1[q] / (1.0 - gam));
        double w = fabs(est) / (atl[q] + rtl[q] * fabs(u[q]));
        if (!(w == w)) w = INFINITY;
        e = fmax(e, w);
      }
    }
    e = wave_max(e);
    if (e > 1.0 && h > ta._dtmin) {
      ++cnt.rejected;
      dt = 0.25 * h;
      continue;
    }
    t = (h == tend - t) ? tend : t + h;
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      un[q] = u[q];
      fn[q] = f1[q];
    }
    dt = (e <= 0.125) ? 2.0 * h : h;
    if (dt > ta._dtmax) dt = ta._dtmax;
    ++cnt.accepted;
  }

  double* rowo = a.state_out + size_t(cell) * dim;
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int i = lane + kWave * q;
    if (i < m) {
      const double v = failed ? 0.0 : un[q];
      row[3 + i] = v;
      if (rowo != row) rowo[3 + i] = v;
    }
  }
  if (rowo != row)
    for (int i = m + lane; i < nSpec; i += kWave) rowo[3 + i] = row[3 + i];
  if (lane == 0) {
    rowo[0] = row[0];
    rowo[1] = row[1];
    rowo[2] = row[2];
    a.t_out[cell] = failed ? tend : t;
    a.dt_out[cell] = failed ? -1.0 : dt;
    if (a.stats) {
      cnt.failed = failed ? 1 : 0;
      a.stats[cell] = tchem_cell_stats_t{cnt.accepted, cnt.rejected, cnt.newton_failures, cnt.newton_iterations,
                                         cnt.factorizations, cnt.failed};
    }
  }
}

template <typename TA, int QM>
__global__ __launch_bounds__(64) void trbdf2_kernel(IntegratorArgs a) {
  extern __shared__ __align__(16) char lds_raw[];
  const CellLds<TA> L = carve<TA>(lds_raw, a.M);
  const int lane = threadIdx.x;
  if (!a.persistent) {
    integrate_cell<TA, QM>(a, L, blockIdx.x, lane);
    return;
  }
  for (;;) {
    int cell = 0;
    if (lane == 0) cell = atomicAdd(a.queue, 1);
    cell = __builtin_amdgcn_readfirstlane(__shfl(cell, 0, kWave));
    if (cell >= a.nBatch) break;
    integrate_cell<TA, QM>(a, L, cell, lane);
    wave_sync();
  }
}

__global__ void reset_counter_kernel(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

__global__ void update_tadv_kernel(int n, tchem_time_advance_t* tadv, const double* t_out, const double* dt_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    tadv[i]._tbeg = t_out[i];
    tadv[i]._dt = dt_out[i];
  }
}

// Source term per cell: one wavefront per cell.
template <int QM>
__global__ __launch_bounds__(64) void sourceterm_kernel(DevMech M, int nBatch, const double* state,
                                                        const double* photo, const double* ext, double* omega) {
  extern __shared__ __align__(16) char lds_raw[];
  const CellLds<double> L = carve<double>(lds_raw, M, false);  // no matrix region
  const int lane = threadIdx.x, cell = blockIdx.x;
  if (cell >= nBatch) return;
  const double* row = state + size_t(cell) * (M.nSpec + 3);
  load_rate_constants(M, L, lane, row[2], row[1], photo ? photo + size_t(cell) * M.nPhoto : nullptr);
  for (int i = lane; i < M.nSpec; i += kWave) L.c[i] = row[3 + i];
  wave_sync();
  double f[QM];
  source_term_from_lds<QM>(M, L, lane, f);
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int i = lane + kWave * q;
    if (i < M.m) omega[size_t(cell) * M.m + i] = f[q] + (ext ? ext[size_t(cell) * M.m + i] : 0.0);
  }
}

// Jacobian per cell: assembled in LDS, written out coalesced.
__global__ __launch_bounds__(64) void jacobian_kernel(DevMech M, int nBatch, const double* state, const double* photo,
                                                      double* jac) {
  extern __shared__ __align__(16) char lds_raw[];
  const CellLds<double> L = carve<double>(lds_raw, M);
  const int lane = threadIdx.x, cell = blockIdx.x;
  if (cell >= nBatch) return;
  const double* row = state + size_t(cell) * (M.nSpec + 3);
  load_rate_constants(M, L, lane, row[2], row[1], photo ? photo + size_t(cell) * M.nPhoto : nullptr);
  for (int i = lane; i < M.nSpec; i += kWave) L.c[i] = row[3 + i];
  wave_sync();
  assemble_jacobian_from_lds(M, L, lane, 0.0, 1.0);
  const int total = M.m * M.m;
  double* out = jac + size_t(cell) * total;
  for (int idx = lane; idx < total; idx += kWave) out[idx] = L.A[idx];
}

}  // namespace

// ---- launchers (called from the C ABI) --------------------------------------

static int qmOf(int m) { return m <= 64 ? 1 : (m <= 128 ? 2 : (m <= 256 ? 4 : 0)); }

hipError_t launchIntegrator(const IntegratorArgs& a, int precision, int grid, size_t lds, hipStream_t stream) {
  const int qm = qmOf(a.M.m);
  if (qm == 0) return hipErrorInvalidValue;
  dim3 blk(64), grd(grid);
#define TCHEM_LAUNCH(TA, QM) hipLaunchKernelGGL((trbdf2_kernel<TA, QM>), grd, blk, lds, stream, a)
  if (precision == TCHEM_PRECISION_FP64) {
    if (qm == 1) TCHEM_LAUNCH(double, 1);
    else if (qm == 2) TCHEM_LAUNCH(double, 2);
    else TCHEM_LAUNCH(double, 4);
  } else {
    if (qm == 1) TCHEM_LAUNCH(float, 1);
    else if (qm == 2) TCHEM_LAUNCH(float, 2);
    else TCHEM_LAUNCH(float, 4);
  }
#undef TCHEM_LAUNCH
  return hipGetLastError();
}

hipError_t integratorMaxBlocksPerCU(int precision, int m, size_t lds, int* blocks) {
  const int qm = qmOf(m);
  if (qm == 0) return hipErrorInvalidValue;
  const void* fn = nullptr;
  if (precision == TCHEM_PRECISION_FP64)
    fn = qm == 1 ? (const void*)trbdf2_kernel<double, 1> : qm == 2 ? (const void*)trbdf2_kernel<double, 2>
                                                                     : (const void*)trbdf2_kernel<double, 4>;
  else
    fn = qm == 1 ? (const void*)trbdf2_kernel<float, 1> : qm == 2 ? (const void*)trbdf2_kernel<float, 2>
                                                                    : (const void*)trbdf2_kernel<float, 4>;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, fn, 64, lds);
}

hipError_t launchResetCounter(int* counter, hipStream_t stream) {
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream, counter);
  return hipGetLastError();
}

hipError_t launchUpdateTadv(int n, tchem_time_advance_t* tadv, const double* t_out, const double* dt_out,
                            hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(update_tadv_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, tadv, t_out, dt_out);
  return hipGetLastError();
}

hipError_t launchSourceTerm(const DevMech& M, int nBatch, const double* state, const double* photo, const double* ext,
                            double* omega, hipStream_t stream) {
  const int qm = qmOf(M.m);
  if (qm == 0) return hipErrorInvalidValue;
  if (nBatch <= 0) return hipSuccess;
  const size_t lds = integratorLdsBytes(M, 0);
  if (qm == 1) hipLaunchKernelGGL(sourceterm_kernel<1>, dim3(nBatch), dim3(64), lds, stream, M, nBatch, state, photo, ext, omega);
  else if (qm == 2) hipLaunchKernelGGL(sourceterm_kernel<2>, dim3(nBatch), dim3(64), lds, stream, M, nBatch, state, photo, ext, omega);
  else hipLaunchKernelGGL(sourceterm_kernel<4>, dim3(nBatch), dim3(64), lds, stream, M, nBatch, state, photo, ext, omega);
  return hipGetLastError();
}

hipError_t launchJacobian(const DevMech& M, int nBatch, const double* state, const double* photo, double* jac,
                          hipStream_t stream) {
  if (nBatch <= 0) return hipSuccess;
  const size_t lds = integratorLdsBytes(M, sizeof(double));
  hipLaunchKernelGGL(jacobian_kernel, dim3(nBatch), dim3(64), lds, stream, M, nBatch, state, photo, jac);
  return hipGetLastError();
}

}  // namespace tchem_amd
