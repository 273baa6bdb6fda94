// HIP kernels for CDNA4 (gfx950): rate constants, source term, analytic
// Jacobian and the TR-BDF2 box-model integrator.
//
// Mapping: one cell per 64-lane wavefront (one wavefront per workgroup).
//   * per-species vectors (state, stage values, residuals, Newton updates)
//     live in registers: internal species a on lane a % 64, slot a / 64;
//   * data other lanes read goes through LDS: concentrations, rate
//     constants, rate-of-progress / derivative / segment scratch, and the
//     dense iteration matrix I - c h J with its LU factors;
//   * species are numbered in a fill-reducing elimination order compiled
//     from the mechanism, so the dense LU meets few nonzero multipliers.
// LU with partial pivoting: the pivot row is found with one ballot when the
// diagonal already has the largest magnitude in its column (the common
// case) and with a wavefront arg-max otherwise; rows whose multiplier is
// exactly zero are skipped by the whole wavefront; lanes own the columns of
// the trailing update; the Newton right-hand side is eliminated alongside
// the matrix so only back substitution remains.
// The iteration matrix and its factors are in TA precision (double for the
// all-FP64 form, float for the mixed form); everything else is FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>

#include "detmath.h"
#include "device_mech.hpp"
#include "schedule.hpp"

namespace tchem_amd {

// Optional per-phase cycle counters for profiling builds (-DTCHEM_PHASE_TIMERS):
// wavefront-summed shader-clock cycles of the source term, Jacobian assembly,
// factorisation, back substitution and whole cells.
#ifdef TCHEM_PHASE_TIMERS
__device__ unsigned long long g_phase_cycles[16];
__shared__ unsigned long long tchem_pt_lds[16][16];  // [wavefront][slot], flushed per cell
#define TCHEM_PT_BEGIN(nm) const unsigned long long tchem_pt_##nm = __builtin_readcyclecounter();
#define TCHEM_PT_END(nm, slot)          \
  if ((threadIdx.x & 63) == 0)          \
    ::tchem_amd::tchem_pt_lds[threadIdx.x >> 6][slot] += __builtin_readcyclecounter() - tchem_pt_##nm;
#define TCHEM_PT_RESET() \
  if ((threadIdx.x & 63) < 16) tchem_pt_lds[threadIdx.x >> 6][threadIdx.x & 63] = 0;
#define TCHEM_PT_FLUSH()                                                                          \
  if ((threadIdx.x & 63) < 16 && tchem_pt_lds[threadIdx.x >> 6][threadIdx.x & 63])                \
    atomicAdd(&g_phase_cycles[threadIdx.x & 63], tchem_pt_lds[threadIdx.x >> 6][threadIdx.x & 63]); \
  if ((threadIdx.x & 63) < 16) tchem_pt_lds[threadIdx.x >> 6][threadIdx.x & 63] = 0;
#else
#define TCHEM_PT_BEGIN(nm)
#define TCHEM_PT_END(nm, slot)
#define TCHEM_PT_RESET()
#define TCHEM_PT_FLUSH()
#endif

}  // namespace tchem_amd

#include "mfma_lu.hpp"

namespace tchem_amd {

namespace {

constexpr int kWave = 64;

// Wavefront-scope barrier for data exchanged through LDS between the lanes
// of one wavefront: the LDS executes a wavefront's operations in order, so
// only compiler reordering has to be prevented (no s_waitcnt, no s_barrier).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double readlane(double v, int l) {
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(x & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane(int(x >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ float readlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ unsigned long long ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
  return v;
}

// value of slot `idx` (lane idx % 64, register slot idx / 64) on every lane
template <int QM, typename TA>
__device__ __forceinline__ TA bcast(const TA (&v)[QM], int idx) {
  const int q = idx >> 6, l = idx & 63;
  TA out = TA(0);
#pragma unroll
  for (int s = 0; s < QM; ++s)
    if (s == q) out = readlane(v[s], l);
  return out;
}

// pow(double, int) of the reference for the small integer powers of
// mechanisms: repeated multiplication (x^2 is the correctly rounded square)
__device__ __forceinline__ double pow_small(double x, int n) { return tchem_ipow(x, n); }

// Rate constant of reaction r: src/core/impl/TChem_Impl_KForwardAerosol.hpp:54-131,
// in the reference's expression order, with the portable exp / log / pow of
// detmath.h and no fused operations, so every device and the host give the
// same bits.
__device__ double rate_constant(const DevMech& M, int r, double t, double p, double t_1, double conv,
                                const double* photo) {
#pragma clang fp contract(off)
  const double* P = M.ratePar() + 8 * r;
  const double one = 1.0;
  const int kind = M.rateKind()[r];
  double k;
  if (kind == 0) {  // Arrhenius
    double A = P[0];
    const int slot = M.photoSlot()[r];
    if (photo != nullptr && slot >= 0) A = photo[slot];
    k = A * tchem_exp(P[2] * t_1) * (P[1] == 0.0 ? one : tchem_pow(t / P[3], P[1])) *
        (P[4] == 0.0 ? one : (one + P[4] * p));
  } else if (kind == 1) {  // Troe
    const double k0 = P[0] * (P[2] == 0.0 ? one : tchem_exp(P[2] * t_1)) *
                      (P[1] == 0.0 ? one : tchem_pow(t / 300.0, P[1])) * conv;
    const double k0_kinf = k0 / (P[3] * (P[5] == 0.0 ? one : tchem_exp(P[5] * t_1)) *
                                 (P[4] == 0.0 ? one : tchem_pow(t / 300.0, P[4])));
    const double lg = tchem_log10(k0_kinf) / P[7];
    k = k0 / (1.0 + k0_kinf) * tchem_pow(P[6], 1.0 / (1.0 + lg * lg));
  } else {  // CMAQ_H2O2
    k = P[0] * tchem_exp(P[2] * t_1) * (P[1] == 0.0 ? one : tchem_pow(t / 300.0, P[1])) +
        P[3] * tchem_exp(P[5] * t_1) * (P[4] == 0.0 ? one : tchem_pow(t / 300.0, P[4])) * conv;
  }
  return k * tchem_ipow(conv, M.convExp()[r]);
}

template <typename TA>
struct CellLds {
  TA* A;        // [m][m] iteration matrix / LU factors (row-major, lda = m)
  double* c;    // [nSpec] concentrations: active (internal order) then constant
  double* kf;   // [nReac] rate constants
  double* scr;  // [max(nReac, nPairs)] rate of progress or derivative pairs
  double* seg;  // [nSeg] source-term segment sums
  TA* rdiag;    // [m] reciprocal pivots 1 / u_kk
  int* swp;     // [2m] row interchanges (k, p) of the factorisation
  const double* src;  // nullable: this cell's external sources [m], reference order
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
  L.seg = L.scr + nScr;
  L.rdiag = reinterpret_cast<TA*>(L.seg + M.nSeg);
  L.swp = reinterpret_cast<int*>(L.seg + M.nSeg + M.m);
  L.src = nullptr;
  return L;
}

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
  const int* sp = M.reacSpec() + r * M.maxR;
  const int* pw = M.reacPow() + r * M.maxR;
  for (int j = 0; j < M.maxR; ++j) {
    const int s = sp[j];
    if (s < 0) break;
    v *= pow_small(c[s], pw[j]);
  }
  return v;
}

// omega (internal order) at the concentrations in LDS; result in registers.
// src/core/impl/TChem_Impl_ReactionRatesAerosol.hpp:78-138 as a gather: one
// lane per species adds the emission rate, then nu * rop of every reactant
// and product entry of that species one at a time, in the order of the
// reference's loops (the serial order of its atomic adds), each product
// rounded before the add: the same sum, bit for bit.
template <int QM, typename TA>
__device__ void source_term_from_lds(const DevMech& M, const CellLds<TA>& L, int lane, double (&f)[QM]) {
#pragma clang fp contract(off)
#pragma unroll 2
  for (int r = lane; r < M.nReac; r += kWave) L.scr[r] = rate_of_progress(M, L.c, L.kf, r);
  wave_sync();
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int a = lane + kWave * q;
    double s = 0.0;
    if (a < M.m) {
      s = M.emission()[a];
      if (L.src) s += L.src[M.perm()[a]];
      const int e1 = M.serOmegaPtr()[a + 1];
      for (int e = M.serOmegaPtr()[a]; e < e1; ++e) {
        const double v = M.serOmegaCoef()[e] * L.scr[M.serOmegaReac()[e]];
        s = s + v;
      }
    }
    f[q] = s;
  }
  wave_sync();
}

template <int QM, typename TA>
__device__ void eval_rhs(const DevMech& M, const CellLds<TA>& L, int lane, const double (&u)[QM], double (&f)[QM]) {
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int a = lane + kWave * q;
    if (a < M.m) L.c[a] = u[q];
  }
  wave_sync();
  source_term_from_lds<QM>(M, L, lane, f);
}

// A = diag * I + scale * J at the concentrations in LDS.  J is the analytic
// Jacobian of source_term_from_lds: derivative pairs d rop_r / d c_s, then
// per structural entry the contributions nu * d one at a time in the
// reference loop order (reaction, reactant slot of the derivative, then the
// reactants and products of the reaction), each product rounded before the
// add.  With external_order the entries are placed at the reference's
// species indices.
template <typename TA>
__device__ void assemble_jacobian_from_lds(const DevMech& M, const CellLds<TA>& L, int lane, double diag, double scale,
                                           bool external_order) {
#pragma clang fp contract(off)
  const int m = M.m;
  {
    const size_t bytes = sizeof(TA) * size_t(m) * size_t(m);
    const int n16 = int(bytes / 16);
    uint4* A16 = reinterpret_cast<uint4*>(L.A);
    for (int idx = lane; idx < n16; idx += kWave) A16[idx] = make_uint4(0, 0, 0, 0);
    for (int idx = n16 * int(16 / sizeof(TA)) + lane; idx < m * m; idx += kWave) L.A[idx] = TA(0);
  }
  // derivative pairs: d rop_r / d c_s = k * n_s c_s^(n_s-1) * prod_{j != s} c_j^n_j
  for (int r = lane; r < M.nReac; r += kWave) {
    const int p0 = M.pairPtr()[r], p1 = M.pairPtr()[r + 1];
    if (p0 == p1) continue;
    const int* sp = M.reacSpec() + r * M.maxR;
    const int* pw = M.reacPow() + r * M.maxR;
    for (int p = p0; p < p1; ++p) {
      const int js = M.pairSlot()[p];
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
  for (int n = lane; n < M.nSerJ; n += kWave) {
    double s = 0.0;
    const int e1 = M.serJPtr()[n + 1];
    for (int e = M.serJPtr()[n]; e < e1; ++e) {
      const double v = M.serJCoef()[e] * L.scr[M.serJPair()[e]];
      s = s + v;
    }
    const int rc = M.serJRowCol()[n];
    int row = rc >> 16, col = rc & 0xffff;
    const double v = (row == col ? diag : 0.0) + scale * s;
    if (external_order) {
      row = M.perm()[row];
      col = M.perm()[col];
    }
    L.A[row * m + col] = TA(v);
  }
  wave_sync();
}

// In-place LU with partial pivoting of the m x m matrix in LDS (row
// interchanges recorded in L.swp, count in nswp), eliminating the right-hand
// side b (row a on lane a % 64, slot a / 64) alongside.  Same arithmetic as
// oracle_lu + the forward half of oracle_lu_solve: pivot = first row of
// largest |a_ik|, l = a_ik * (1 / a_kk), a_ij = fma(-l, a_kj, a_ij),
// b_i = fma(-l, b_k, b_i).  The reciprocal pivots are kept for the back
// substitution.  Returns false for a zero or non-finite pivot.
template <int QM, typename TA>
__device__ bool lu_factor_fused(const CellLds<TA>& L, int m, int lane, TA (&b)[QM], int& nswp) {
#pragma clang fp contract(off)
  TA* A = L.A;
  nswp = 0;
  for (int k = 0; k < m; ++k) {
    TA a[QM], uk[QM];
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      a[q] = (i >= k && i < m) ? A[i * m + k] : TA(0);
      const int j = k + 1 + lane + kWave * q;
      uk[q] = (j < m) ? A[k * m + j] : TA(0);
    }
    const TA akk = bcast<QM>(a, k);
    const TA dabs = fabs(akk);
    unsigned long long gt = 0;
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      gt |= ballot(i > k && i < m && fabs(a[q]) > dabs);
    }
    int p = k;
    TA apiv = akk;
    if (gt != 0 || !(dabs > TA(0))) {
      // arg-max over rows i >= k; ties go to the smallest row index
      TA best = TA(-1);
      int bi = m;
#pragma unroll
      for (int q = 0; q < QM; ++q) {
        const int i = lane + kWave * q;
        if (i >= k && i < m) {
          const TA v = fabs(a[q]);
          if (v > best) { best = v; bi = i; }
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const TA ob = __shfl_xor(best, off, kWave);
        const int oi = __shfl_xor(bi, off, kWave);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
      }
      p = __builtin_amdgcn_readfirstlane(bi);
      if (p >= m || !(best > TA(0))) return false;
      apiv = bcast<QM>(a, p);
    }
    if (!isfinite(double(apiv))) return false;
    if (p != k) {
      for (int j = lane; j < m; j += kWave) {
        const TA t = A[k * m + j];
        A[k * m + j] = A[p * m + j];
        A[p * m + j] = t;
      }
      const TA bk = bcast<QM>(b, k), bp = bcast<QM>(b, p);
#pragma unroll
      for (int q = 0; q < QM; ++q) {
        const int i = lane + kWave * q;
        if (i == k) { a[q] = apiv; b[q] = bp; }
        else if (i == p) { a[q] = akk; b[q] = bk; }
      }
      if (lane == 0) {
        L.swp[2 * nswp] = k;
        L.swp[2 * nswp + 1] = p;
      }
      ++nswp;
      wave_sync();
#pragma unroll
      for (int q = 0; q < QM; ++q) {
        const int j = k + 1 + lane + kWave * q;
        uk[q] = (j < m) ? A[k * m + j] : TA(0);
      }
    }
    const TA bk = bcast<QM>(b, k);
    const TA inv = TA(1) / apiv;
    if (lane == 0) L.rdiag[k] = inv;
    TA l[QM];
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      l[q] = TA(0);
      if (i > k && i < m && a[q] != TA(0)) {
        l[q] = a[q] * inv;
        A[i * m + k] = l[q];
        b[q] = fma(-l[q], bk, b[q]);
      }
    }
    // trailing update: rows with a nonzero multiplier, two at a time
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      unsigned long long mask = ballot(l[q] != TA(0));
      while (mask) {
        const int b0 = __builtin_ctzll(mask);
        mask &= mask - 1;
        const TA l0 = readlane(l[q], b0);
        const int i0 = kWave * q + b0;
        if (mask) {
          const int b1 = __builtin_ctzll(mask);
          mask &= mask - 1;
          const TA l1 = readlane(l[q], b1);
          const int i1 = kWave * q + b1;
#pragma unroll
          for (int c = 0; c < QM; ++c) {
            const int j = k + 1 + lane + kWave * c;
            if (j < m && uk[c] != TA(0)) {
              const TA x0 = A[i0 * m + j];
              const TA x1 = A[i1 * m + j];
              A[i0 * m + j] = fma(-l0, uk[c], x0);
              A[i1 * m + j] = fma(-l1, uk[c], x1);
            }
          }
        } else {
#pragma unroll
          for (int c = 0; c < QM; ++c) {
            const int j = k + 1 + lane + kWave * c;
            if (j < m && uk[c] != TA(0)) A[i0 * m + j] = fma(-l0, uk[c], A[i0 * m + j]);
          }
        }
      }
    }
    wave_sync();
  }
  return true;
}

// Forward elimination with stored factors: interchanges, then unit-lower L.
template <int QM, typename TA>
__device__ void lu_forward(const CellLds<TA>& L, int m, int lane, TA (&b)[QM], int nswp) {
#pragma clang fp contract(off)
  const TA* A = L.A;
  for (int s = 0; s < nswp; ++s) {
    const int k = L.swp[2 * s], p = L.swp[2 * s + 1];
    const TA bk = bcast<QM>(b, k), bp = bcast<QM>(b, p);
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      if (i == k) b[q] = bp;
      else if (i == p) b[q] = bk;
    }
  }
  TA col[QM];
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int i = lane + kWave * q;
    col[q] = (i > 0 && i < m) ? A[i * m] : TA(0);
  }
  for (int k = 0; k < m - 1; ++k) {
    TA nxt[QM];
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      nxt[q] = (i > k + 1 && i < m) ? A[i * m + k + 1] : TA(0);
    }
    const TA yk = bcast<QM>(b, k);
    if (yk != TA(0)) {
#pragma unroll
      for (int q = 0; q < QM; ++q) b[q] = fma(-col[q], yk, b[q]);
    }
#pragma unroll
    for (int q = 0; q < QM; ++q) col[q] = nxt[q];
  }
}

// Back substitution with the upper factor (column order, as oracle_lu_solve);
// the next column is fetched from LDS while the current one is applied.
template <int QM, typename TA>
__device__ void lu_back(const CellLds<TA>& L, int m, int lane, TA (&b)[QM]) {
#pragma clang fp contract(off)
  const TA* A = L.A;
  TA col[QM];
  TA ukk = L.rdiag[m - 1];
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int i = lane + kWave * q;
    col[q] = (i < m - 1) ? A[i * m + m - 1] : TA(0);
  }
  for (int k = m - 1; k >= 0; --k) {
    TA nxt[QM];
    TA nkk = TA(1);
#pragma unroll
    for (int q = 0; q < QM; ++q) nxt[q] = TA(0);
    if (k > 0) {
      nkk = L.rdiag[k - 1];
#pragma unroll
      for (int q = 0; q < QM; ++q) {
        const int i = lane + kWave * q;
        nxt[q] = (i < k - 1) ? A[i * m + k - 1] : TA(0);
      }
    }
    const TA xk = bcast<QM>(b, k) * ukk;  // ukk holds 1 / u_kk
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      if (i == k) b[q] = xk;
    }
    if (xk != TA(0)) {
#pragma unroll
      for (int q = 0; q < QM; ++q) b[q] = fma(-col[q], xk, b[q]);
    }
    ukk = nkk;
#pragma unroll
    for (int q = 0; q < QM; ++q) col[q] = nxt[q];
  }
}

// ---- sparse form ------------------------------------------------------------
// Compiled schedules (schedule.hpp): the iteration matrix is held as the
// nonzeros of its symbolic LU pattern, factorised without interchanges in
// the fill-reducing order by list-scheduled rounds; the RHS is stored after
// the matrix entries and eliminated alongside.

template <typename TA>
struct SpLds {
  const unsigned* hdr;    // LU program headers (LU + BACK sections), shared by the workgroup
  const unsigned* body;   // LU program bodies
  const unsigned* ehdr;   // evaluation program headers
  const unsigned* ebody;  // evaluation program bodies
  const double* coef;     // stoichiometric coefficient table
  TA* A;        // [trashSlot+1] factor entries, RHS, zero slot, trash slot
  TA* rdiag;    // [m+2] reciprocal pivots, rdiag[m] = 0, rdiag[m+1] = trash
  TA* dpr;      // [pairSlots] Jacobian derivative pairs, dpr[nPairs] = 0
  double* c;    // [nSpec]
  double* kf;   // [nReac+1], kf[nReac] = 0
  double* rop;  // [nReac+1], rop[nReac] = 0
  double* seg;  // [nSeg+1], seg[nSeg] = trash
  const double* src;  // nullable: this cell's external sources [m], reference order
  // discrete adjoint only (nullptr otherwise)
  double* mu;   // [m+1] adjoint vector, mu[m] = 0
  double* w;    // [nReac+2] nu_r . mu, w[nReac] = 0, w[nReac+1] = trash
  double* v;    // [m+1] J^T mu, v[m] = trash
};

// STAGED: programs at the start of the workgroup's LDS, then one cell region
// per wavefront; otherwise the programs are read from device memory (they
// stay resident in L2) and the cell regions start at 0, or, with COEF, after
// the coefficient table staged at the start of the LDS (sparseCoefBytes).
template <typename TA, bool STAGED, bool COEF = false>
__device__ SpLds<TA> carve_sparse(char* base, const DevMech& M, const DevSparse& S, int cellBytes, int wave) {
  auto al = [](size_t x) { return (x + 15) & ~size_t(15); };
  SpLds<TA> L;
  char* cell;
  if (STAGED) {
    const unsigned* w = reinterpret_cast<const unsigned*>(base);
    L.hdr = w;
    L.body = w + sparseLuHdrWords(S);
    L.ehdr = w + sparseEvalOffsetWords(S);
    L.ebody = L.ehdr + sparseEvalHdrWords(S);
    L.coef = reinterpret_cast<const double*>(w + sparseCoefOffsetWords(S));
    cell = base + sparseProgramBytes(S) + size_t(wave) * size_t(cellBytes);
  } else {
    L.hdr = reinterpret_cast<const unsigned*>(M.ib + S.oProgHdr);
    L.body = reinterpret_cast<const unsigned*>(M.ib + S.oProgBody);
    L.ehdr = reinterpret_cast<const unsigned*>(M.ib + S.oEvalHdr);
    L.ebody = reinterpret_cast<const unsigned*>(M.ib + S.oEvalBody);
    L.coef = COEF ? reinterpret_cast<const double*>(base) : M.db + S.oCoef;
    cell = base + (COEF ? sparseCoefBytes(S) : 0) + size_t(wave) * size_t(cellBytes);
  }
  size_t off = 0;
  L.A = reinterpret_cast<TA*>(cell);
  off += al(sizeof(TA) * size_t(S.trashSlot + 1));
  L.rdiag = reinterpret_cast<TA*>(cell + off);
  off += al(sizeof(TA) * size_t(M.m + 2));
  L.dpr = reinterpret_cast<TA*>(cell + off);
  off += al(sizeof(TA) * size_t(S.pairSlots));
  L.c = reinterpret_cast<double*>(cell + off);
  L.kf = L.c + M.nSpec;
  L.rop = L.kf + M.nReac + 1;
  L.seg = L.rop + M.nReac + 1;
  L.mu = L.w = L.v = nullptr;
  L.src = nullptr;
  return L;
}

// x^n for small n >= 0 by repeated multiplication (exact for n <= 2; at
// most n-1 roundings above); powers up to 3 without a loop (selects, the
// same products in the same order)
__device__ __forceinline__ double ipow(double x, int n) {
  double r = n >= 1 ? x : 1.0;
  r = n >= 2 ? r * x : r;
  r = n >= 3 ? r * x : r;
  for (int i = 3; i < n; ++i) r *= x;
  return r;
}
// packed reactant field: species | power << 12, kNone for no reactant (a
// factor 1: power 0 of an in-range slot, no branch)
__device__ __forceinline__ double reactant_factor(const double* c, unsigned f) {
  const bool none = f == SparseSchedule::kNone;
  return ipow(c[none ? 0u : (f & 0xfff)], none ? 0 : int(f >> 12));
}

// Rates of progress (kRop rounds) and, with pairs, the derivatives
// d rop_r / d c_s (kPair rounds) at the concentrations in LDS: the same
// products in the same order as the dense form.  Rounds are independent;
// the next round's two words are loaded while a round computes.
template <typename TA>
__device__ __forceinline__ void sp_rates(const DevSparse& S, const SpLds<TA>& L, int lane, bool pairs) {
  {
    const unsigned* b = L.ebody;
    unsigned w0 = b[lane], w1 = b[kWave + lane];
    for (int g = 0; g < S.ropRounds; ++g) {
      b += 2 * kWave;
      const unsigned n0 = b[lane], n1 = b[kWave + lane];
      const int r = int(w0 & 0xffff);
      L.rop[r] = L.kf[r] * reactant_factor(L.c, w0 >> 16) * reactant_factor(L.c, w1 & 0xffff) *
                 reactant_factor(L.c, w1 >> 16);
      w0 = n0;
      w1 = n1;
    }
  }
  if (!pairs) return;
  const unsigned* b = L.ebody + S.pairBody;
  unsigned w0 = b[lane], w1 = b[kWave + lane];
  for (int g = 0; g < S.pairRounds; ++g) {
    b += 2 * kWave;
    const unsigned n0 = b[lane], n1 = b[kWave + lane];
    const int r = int(w0 & 0x3fff), js = int((w0 >> 14) & 3);
    const unsigned f0 = w0 >> 16, f1 = w1 & 0xffff, f2 = w1 >> 16;
    const unsigned fs = js == 0 ? f0 : (js == 1 ? f1 : f2);
    const int ns = int(fs >> 12);
    double d = L.kf[r] * double(ns) * ipow(L.c[fs & 0xfff], ns - 1);
    if (js != 0) d *= reactant_factor(L.c, f0);
    if (js != 1) d *= reactant_factor(L.c, f1);
    if (js != 2) d *= reactant_factor(L.c, f2);
    L.dpr[g * kWave + lane] = TA(d);
    w0 = n0;
    w1 = n1;
  }
}

// one segment-sum round, LEN entries per lane (w[0] = target, w[1..] = (reaction | coef << 16))
template <int LEN, typename TA>
__device__ __forceinline__ void seg_round(const unsigned (&w)[1 + SparseSchedule::kMaxChain], const SpLds<TA>& L) {
  const unsigned t = w[0] & ~SparseSchedule::kCont;
  double acc = (w[0] & SparseSchedule::kCont) ? L.seg[t] : 0.0;
  double cf[LEN], rv[LEN];
#pragma unroll
  for (int j = 0; j < LEN; ++j) {
    cf[j] = L.coef[w[1 + j] >> 16];
    rv[j] = L.rop[w[1 + j] & 0xffff];
  }
#pragma unroll
  for (int j = 0; j < LEN; ++j) acc += cf[j] * rv[j];
  L.seg[t] = acc;
}
// one Jacobian round: A[t] = (cont ? A[t] : diag) + scale * sum coef * dpr, in TA precision
template <int LEN, typename TA>
__device__ __forceinline__ void jac_round(const unsigned (&w)[1 + SparseSchedule::kMaxChain], const SpLds<TA>& L,
                                          TA tdiag, TA tscale) {
  const unsigned t = w[0] & 0x3fffffffu;
  const TA init = (w[0] & SparseSchedule::kCont) ? L.A[t] : (((w[0] >> 30) & 1u) ? tdiag : TA(0));
  TA cf[LEN], dv[LEN];
#pragma unroll
  for (int j = 0; j < LEN; ++j) {
    cf[j] = TA(L.coef[w[1 + j] >> 16]);
    dv[j] = L.dpr[w[1 + j] & 0xffff];
  }
  TA sum = TA(0);
#pragma unroll
  for (int j = 0; j < LEN; ++j) sum += cf[j] * dv[j];
  L.A[t] = init + tscale * sum;
}

// nu_r . mu (kNuT): w[t] = (cont ? w[t] : 0) + sum coef * mu[species]
template <int LEN, typename TA>
__device__ __forceinline__ void nut_round(const unsigned (&w)[1 + SparseSchedule::kMaxChain], const SpLds<TA>& L) {
  const unsigned t = w[0] & ~SparseSchedule::kCont;
  double acc = (w[0] & SparseSchedule::kCont) ? L.w[t] : 0.0;
  double cf[LEN], mv[LEN];
#pragma unroll
  for (int j = 0; j < LEN; ++j) {
    cf[j] = L.coef[w[1 + j] >> 16];
    mv[j] = L.mu[w[1 + j] & 0xffff];
  }
#pragma unroll
  for (int j = 0; j < LEN; ++j) acc += cf[j] * mv[j];
  L.w[t] = acc;
}
// J^T mu (kJtv): v[t] = (cont ? v[t] : 0) + sum dpr[pair] * w[reaction]
template <int LEN, typename TA>
__device__ __forceinline__ void jtv_round(const unsigned (&w)[1 + SparseSchedule::kMaxChain], const SpLds<TA>& L) {
  const unsigned t = w[0] & ~SparseSchedule::kCont;
  double acc = (w[0] & SparseSchedule::kCont) ? L.v[t] : 0.0;
  double dv[LEN], wv[LEN];
#pragma unroll
  for (int j = 0; j < LEN; ++j) {
    dv[j] = double(L.dpr[w[1 + j] & 0xffff]);
    wv[j] = L.w[w[1 + j] >> 16];
  }
#pragma unroll
  for (int j = 0; j < LEN; ++j) acc += dv[j] * wv[j];
  L.v[t] = acc;
}

// Round driver for programs read from device memory, where a round's words
// take longer to arrive than a round takes to run: the words of round r + 2
// are loaded while round r runs, into three register sets used in rotation
// (no copies, so no wait until a set's round comes up), and the headers come
// through scalar loads.  exec(h, w) runs one round with header h and this
// lane's words w; size(h) is the round's body length in words.  The loads
// stop at the last round, so nothing past the program is read.
template <class Size, class Exec>
__device__ __forceinline__ void pipelined_rounds(const unsigned* hdr, const unsigned* body, int nr, int lane,
                                                 Size size, Exec exec) {
  constexpr int kW = 1 + SparseSchedule::kMaxChain;
  if (nr <= 0) return;
  using cptr = const __attribute__((address_space(4))) unsigned*;
  const cptr sh = (cptr)hdr;
  unsigned wa[kW], wb[kW], wc[kW];
  auto load = [&](unsigned (&w)[kW], const unsigned* b) {
#pragma unroll
    for (int j = 0; j < kW; ++j) w[j] = b[kWave * j + lane];
  };
  unsigned h0 = sh[0];
  unsigned h1 = nr > 1 ? sh[1] : 0u;
  const unsigned* b1 = body + size(h0);  // body of round r + 1
  load(wa, body);
  if (nr > 1) load(wb, b1);
  int r = 0;
  // one round: run round r from `cur`, load round r + 2 into `nxt2`
  auto step = [&](unsigned (&cur)[kW], unsigned (&nxt2)[kW]) -> bool {
    const unsigned* b2 = b1 + size(h1);
    unsigned h2 = 0u;
    if (r + 2 < nr) {
      h2 = sh[r + 2];
      load(nxt2, b2);
    }
    exec(h0, cur);
    if (++r >= nr) return false;
    h0 = h1;
    h1 = h2;
    b1 = b2;
    return true;
  };
  while (step(wa, wc) && step(wb, wa) && step(wc, wb)) {
  }
}

// Runs nr chained rounds of the evaluation program starting at header
// index r0 and body word offset body0: kSeg / kJac rounds, and with ADJ
// also kNuT / kJtv.  Rounds run in order (a chain split over rounds
// continues in a later one); the next round's words are loaded while a
// round computes.
template <bool ADJ, typename TA>
__device__ __forceinline__ void eval_rounds(const SpLds<TA>& L, int lane, int r0, int nr, int body0, double diag, double scale) {
  constexpr int kW = 1 + SparseSchedule::kMaxChain;
  const unsigned* hdr = L.ehdr + r0;
  const unsigned* b = L.ebody + body0;
  const TA tdiag = TA(diag), tscale = TA(scale);
  unsigned w[kW];
#pragma unroll
  for (int j = 0; j < kW; ++j) w[j] = b[kWave * j + lane];
  unsigned h = nr > 0 ? __builtin_amdgcn_readfirstlane(hdr[0]) : 0u;
  for (int r = 0; r < nr; ++r) {
    const unsigned hn = r + 1 < nr ? hdr[r + 1] : 0u;
    const int kind = int(h >> 16), len = int(h & 0xffff);
    b += kWave * (1 + len);
    unsigned wn[kW];
#pragma unroll
    for (int j = 0; j < kW; ++j) wn[j] = b[kWave * j + lane];
    if (kind == SparseSchedule::kSeg) {
      switch (len) {
        case 1: seg_round<1>(w, L); break;
        case 2: seg_round<2>(w, L); break;
        case 3: seg_round<3>(w, L); break;
        default: seg_round<4>(w, L); break;
      }
    } else if (!ADJ || kind == SparseSchedule::kJac) {
      switch (len) {
        case 1: jac_round<1>(w, L, tdiag, tscale); break;
        case 2: jac_round<2>(w, L, tdiag, tscale); break;
        case 3: jac_round<3>(w, L, tdiag, tscale); break;
        default: jac_round<4>(w, L, tdiag, tscale); break;
      }
    } else if (kind == SparseSchedule::kNuT) {
      switch (len) {
        case 1: nut_round<1>(w, L); break;
        case 2: nut_round<2>(w, L); break;
        case 3: nut_round<3>(w, L); break;
        default: nut_round<4>(w, L); break;
      }
    } else {
      switch (len) {
        case 1: jtv_round<1>(w, L); break;
        case 2: jtv_round<2>(w, L); break;
        case 3: jtv_round<3>(w, L); break;
        default: jtv_round<4>(w, L); break;
      }
    }
    h = __builtin_amdgcn_readfirstlane(hn);
#pragma unroll
    for (int j = 0; j < kW; ++j) w[j] = wn[j];
  }
}

// eval_rounds (kSeg / kJac only) for programs in device memory
template <typename TA>
__device__ __forceinline__ void eval_rounds_pf(const SpLds<TA>& L, int lane, int r0, int nr, int body0, double diag,
                                               double scale) {
  constexpr int kW = 1 + SparseSchedule::kMaxChain;
  const TA tdiag = TA(diag), tscale = TA(scale);
  pipelined_rounds(
      L.ehdr + r0, L.ebody + body0, nr, lane, [](unsigned h) { return kWave * (1 + int(h & 0xffff)); },
      [&](unsigned h, const unsigned (&w)[kW]) {
        const int kind = int(h >> 16), len = int(h & 0xffff);
        if (kind == SparseSchedule::kSeg) {
          switch (len) {
            case 1: seg_round<1>(w, L); break;
            case 2: seg_round<2>(w, L); break;
            case 3: seg_round<3>(w, L); break;
            default: seg_round<4>(w, L); break;
          }
        } else {
          switch (len) {
            case 1: jac_round<1>(w, L, tdiag, tscale); break;
            case 2: jac_round<2>(w, L, tdiag, tscale); break;
            case 3: jac_round<3>(w, L, tdiag, tscale); break;
            default: jac_round<4>(w, L, tdiag, tscale); break;
          }
        }
      });
}

// Source-term segment sums (kSeg rounds) and, with jac, A = diag * I +
// scale * J on the factor pattern (kJac rounds; fill positions cleared).
template <typename TA, bool PF = false>
__device__ void sp_sums(const DevSparse& S, const SpLds<TA>& L, int lane, bool jac, double diag, double scale) {
  if constexpr (PF)
    eval_rounds_pf(L, lane, S.ropRounds + S.pairRounds, S.segRounds + (jac ? S.jacRounds : 0), S.segBody, diag,
                   scale);
  else
    eval_rounds<false>(L, lane, S.ropRounds + S.pairRounds, S.segRounds + (jac ? S.jacRounds : 0), S.segBody, diag,
                       scale);
}

// per-species sums of the segment sums plus emission and external sources
// (register layout)
template <int QM>
__device__ __forceinline__ void row_sums(const DevMech& M, const double* seg, const double* src, int lane,
                                         double (&f)[QM]) {
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int a = lane + kWave * q;
    double s = 0.0;
    if (a < M.m) {
      s = M.emission()[a];
      if (src) s += src[M.perm()[a]];
      const int s1 = M.rowSegPtr()[a + 1];
      for (int g = M.rowSegPtr()[a]; g < s1; ++g) s += seg[g];
    }
    f[q] = s;
  }
}

// One chained-update round with LEN entries per lane (w[0] = target,
// w[1..LEN] = entries): every load is issued before the first use.
template <int LEN, typename TA>
__device__ __forceinline__ void upd_round(const unsigned (&w)[1 + SparseSchedule::kMaxChain], TA* A) {
  TA acc = A[w[0]];
  TA l[LEN], u[LEN];
#pragma unroll
  for (int j = 0; j < LEN; ++j) {
    l[j] = A[w[1 + j] & 0xffff];
    u[j] = A[w[1 + j] >> 16];
  }
#pragma unroll
  for (int j = 0; j < LEN; ++j) acc -= l[j] * u[j];
  A[w[0]] = acc;
}
// Reciprocal of a pivot: the correctly rounded quotient in FP64; in FP32 the
// hardware reciprocal (within 1 ulp: one instruction instead of the eleven of
// a correctly rounded division, well inside the FP32 factorisation's own
// rounding).  Zero and non-finite pivots give non-finite reciprocals in both.
template <typename TA>
__device__ __forceinline__ TA pivot_recip(TA p) {
  if constexpr (sizeof(TA) == sizeof(float))
    return __builtin_amdgcn_rcpf(p);
  else
    return TA(1) / p;
}
// A kMulUpd round with LEN chain entries per lane (w[3..3+LEN)); returns
// true for a zero or non-finite pivot on a lane that stores one.
template <int LEN, typename TA>
__device__ __forceinline__ bool mulupd_round(const unsigned (&w)[1 + SparseSchedule::kMaxChain], TA* A, TA* rdiag,
                                             int m) {
  const int t = int(w[0]);
  const int ik = int(w[1] & 0xffff), kj = int(w[1] >> 16);
  const int kk = int(w[2] & 0x7fff), ri = int(w[2] >> 16);
  const bool mul = (w[2] & 0x8000u) != 0;
  const TA p = A[kk];
  const TA a = A[ik];
  const TA b = A[kj];
  TA acc = A[t];
  TA lo[LEN > 0 ? LEN : 1], hi[LEN > 0 ? LEN : 1];
#pragma unroll
  for (int j = 0; j < LEN; ++j) {
    lo[j] = A[w[3 + j] & 0xffff];
    hi[j] = A[w[3 + j] >> 16];
  }
  const TA inv = pivot_recip(p);
  const TA l = a == TA(0) ? TA(0) : a * inv;
#pragma unroll
  for (int j = 0; j < LEN; ++j) acc -= lo[j] * hi[j];
  A[t] = mul ? l : acc - l * b;
  rdiag[ri] = inv;
  return ri != m + 1 && !(p != TA(0) && isfinite(p) && isfinite(inv));
}
template <int LEN, typename TA>
__device__ __forceinline__ void back_round(const unsigned (&w)[1 + SparseSchedule::kMaxChain], TA* A,
                                           const TA* rdiag, int rhs0) {
  TA acc = A[w[0]];
  TA U[LEN], b[LEN], r[LEN];
#pragma unroll
  for (int j = 0; j < LEN; ++j) {
    const int k = int(w[1 + j] >> 16);
    U[j] = A[w[1 + j] & 0xffff];
    b[j] = A[rhs0 + k];
    r[j] = rdiag[k];
  }
#pragma unroll
  for (int j = 0; j < LEN; ++j) acc -= U[j] * (b[j] * r[j]);
  A[w[0]] = acc;
}

// Runs rounds of the LU / substitution program (schedule.hpp): multiplier
// and pivot records (kMul), chained updates A[t] -= sum A[lo] * A[hi] (kUpd),
// multipliers fused with updates that scale A[ik] themselves, plus chained
// updates (kMulUpd)
// and column-oriented back-substitution updates
// A[t] -= sum A[lo] * (A[rhs0 + k] * rdiag[k]) (kBack).  Lanes of a round
// touch disjoint targets; rounds run in order.  Branch free per lane
// (padding reads zero slots, writes trash slots); the header two rounds
// ahead and the next round's words are loaded while a round runs, so a
// round waits only for its matrix data.  Returns false when a pivot is
// zero or not finite.
template <typename TA>
__device__ bool run_program(const unsigned* hdr, const unsigned* body, int nRounds, TA* A, TA* rdiag, int m,
                            int rhs0, int trash, int lane) {
  constexpr int kW = 1 + SparseSchedule::kMaxChain;
  bool bad = false;
  if (nRounds <= 0) return true;
  unsigned h0 = __builtin_amdgcn_readfirstlane(hdr[0]);
  unsigned h1 = nRounds > 1 ? hdr[1] : 0u;
  unsigned w[kW];
#pragma unroll
  for (int j = 0; j < kW; ++j) w[j] = body[kWave * j + lane];
  for (int r = 0; r < nRounds; ++r) {
    const unsigned h2 = r + 2 < nRounds ? hdr[r + 2] : 0u;
    const int kind = int(h0 >> 16), len = int(h0 & 0xffff);
    const unsigned* bnext = body + (kind == SparseSchedule::kMul ? kWave : kWave * (1 + len));
    unsigned wn[kW];
#pragma unroll
    for (int j = 0; j < kW; ++j) wn[j] = bnext[kWave * j + lane];
    if (kind == SparseSchedule::kMul) {
      const int dst = int(w[0] & 0xffff), src = int(w[0] >> 16);
      const bool piv = (dst & 0x8000) != 0;
      const int ai = piv ? trash : dst;
      const int ri = piv ? (dst & 0x7fff) : m + 1;
      const TA p = A[src];
      const TA a = A[ai];
      const TA inv = pivot_recip(p);
      A[ai] = a * inv;
      rdiag[ri] = inv;
      bad |= piv && !(p != TA(0) && isfinite(p) && isfinite(inv));
    } else if (kind == SparseSchedule::kUpd) {
      switch (len) {
        case 1: upd_round<1>(w, A); break;
        case 2: upd_round<2>(w, A); break;
        case 3: upd_round<3>(w, A); break;
        default: upd_round<4>(w, A); break;
      }
    } else if (kind == SparseSchedule::kMulUpd) {
      static_assert(SparseSchedule::kMaxMulUpdChain == 2, "kMulUpd chain cases");
      switch (len) {
        case 2: bad |= mulupd_round<0>(w, A, rdiag, m); break;
        case 3: bad |= mulupd_round<1>(w, A, rdiag, m); break;
        default: bad |= mulupd_round<2>(w, A, rdiag, m); break;
      }
    } else {
      switch (len) {
        case 1: back_round<1>(w, A, rdiag, rhs0); break;
        case 2: back_round<2>(w, A, rdiag, rhs0); break;
        case 3: back_round<3>(w, A, rdiag, rhs0); break;
        default: back_round<4>(w, A, rdiag, rhs0); break;
      }
    }
    wave_sync();
    body = bnext;
    h0 = __builtin_amdgcn_readfirstlane(h1);
    h1 = h2;
#pragma unroll
    for (int j = 0; j < kW; ++j) w[j] = wn[j];
  }
  return ballot(bad) == 0;
}

// run_program for programs in device memory (pipelined_rounds)
template <typename TA>
__device__ bool run_program_pf(const unsigned* hdr, const unsigned* body, int nRounds, TA* A, TA* rdiag, int m,
                               int rhs0, int trash, int lane) {
  constexpr int kW = 1 + SparseSchedule::kMaxChain;
  bool bad = false;
  pipelined_rounds(
      hdr, body, nRounds, lane,
      [](unsigned h) { return int(h >> 16) == SparseSchedule::kMul ? kWave : kWave * (1 + int(h & 0xffff)); },
      [&](unsigned h, const unsigned (&w)[kW]) {
        const int kind = int(h >> 16), len = int(h & 0xffff);
        if (kind == SparseSchedule::kMul) {
          const int dst = int(w[0] & 0xffff), src = int(w[0] >> 16);
          const bool piv = (dst & 0x8000) != 0;
          const int ai = piv ? trash : dst;
          const int ri = piv ? (dst & 0x7fff) : m + 1;
          const TA p = A[src];
          const TA a = A[ai];
          const TA inv = pivot_recip(p);
          A[ai] = a * inv;
          rdiag[ri] = inv;
          bad |= piv && !(p != TA(0) && isfinite(p) && isfinite(inv));
        } else if (kind == SparseSchedule::kUpd) {
          switch (len) {
            case 1: upd_round<1>(w, A); break;
            case 2: upd_round<2>(w, A); break;
            case 3: upd_round<3>(w, A); break;
            default: upd_round<4>(w, A); break;
          }
        } else if (kind == SparseSchedule::kMulUpd) {
          switch (len) {
            case 2: bad |= mulupd_round<0>(w, A, rdiag, m); break;
            case 3: bad |= mulupd_round<1>(w, A, rdiag, m); break;
            default: bad |= mulupd_round<2>(w, A, rdiag, m); break;
          }
        } else {
          switch (len) {
            case 1: back_round<1>(w, A, rdiag, rhs0); break;
            case 2: back_round<2>(w, A, rdiag, rhs0); break;
            case 3: back_round<3>(w, A, rdiag, rhs0); break;
            default: back_round<4>(w, A, rdiag, rhs0); break;
          }
        }
        wave_sync();
      });
  return ballot(bad) == 0;
}

// the LU / substitution interpreter for programs staged in LDS or read from
// device memory
template <bool STAGED, typename TA>
__device__ __forceinline__ bool program(const unsigned* hdr, const unsigned* body, int nRounds, TA* A, TA* rdiag,
                                        int m, int rhs0, int trash, int lane) {
  if constexpr (STAGED)
    return run_program(hdr, body, nRounds, A, rdiag, m, rhs0, trash, lane);
  else
    return run_program_pf(hdr, body, nRounds, A, rdiag, m, rhs0, trash, lane);
}

// ---- backends -----------------------------------------------------------------
// Both provide: Lds (per-cell LDS view), Fact (state of the stored factors),
// carve, init, evaluate (f at u; for a Newton iteration also the residual
// and, when refreshing, the iteration matrix), factor_solve (factorise +
// solve; false on a bad pivot) and solve (stored factors).

template <typename TA_, int QM>
struct DenseBackend {
  using TA = TA_;
  static constexpr bool kCollective = false;  // factorisation by the whole workgroup (collective_solve)
  using Lds = CellLds<TA>;
  struct Fact { int nswp = 0; };
  // unknowns, state row length, state column and tolerance row of unknown ai
  static __device__ __forceinline__ int unknowns(const IntegratorArgs& a) { return a.M.m; }
  static __device__ __forceinline__ int rowDim(const IntegratorArgs& a) { return a.M.nSpec + 3; }
  static __device__ __forceinline__ int column(const IntegratorArgs& a, int ai) { return 3 + a.M.perm()[ai]; }
  static __device__ __forceinline__ int tolIndex(const IntegratorArgs& a, int ai) { return a.M.perm()[ai]; }
  // per-cell bindings: this cell's external sources
  static __device__ __forceinline__ void bind(const IntegratorArgs& a, Lds& L, int cell) {
    L.src = a.ext ? a.ext + size_t(cell) * a.M.m : nullptr;
  }
  static __device__ void stage(const IntegratorArgs&, char*) {}
  static __device__ Lds carve(char* base, const IntegratorArgs& a, int) { return tchem_amd::carve<TA>(base, a.M); }
  static __device__ void init(const IntegratorArgs&, const Lds&, int) {}
  // f at u; for a Newton iteration also x = -(u - bc - ch f) and, when
  // refreshing, the iteration matrix I - ch J at the same u
  static __device__ void evaluate(const IntegratorArgs& a, const Lds& L, int lane, const double (&u)[QM],
                                  const double (&bc)[QM], double ch, bool newton, bool refresh, double (&f)[QM],
                                  TA (&x)[QM]) {
#pragma clang fp contract(off)
    eval_rhs<QM>(a.M, L, lane, u, f);
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      x[q] = (newton && i < a.M.m) ? TA(-(u[q] - bc[q] - ch * f[q])) : TA(0);
    }
    if (refresh) assemble_jacobian_from_lds(a.M, L, lane, 1.0, -ch, false);
  }
  static __device__ bool factor_solve(const IntegratorArgs& a, const Lds& L, int lane, TA (&x)[QM], Fact& fs) {
    if (!lu_factor_fused<QM>(L, a.M.m, lane, x, fs.nswp)) return false;
    lu_back<QM>(L, a.M.m, lane, x);
    return true;
  }
  static __device__ void solve(const IntegratorArgs& a, const Lds& L, int lane, TA (&x)[QM], const Fact& fs) {
    lu_forward<QM>(L, a.M.m, lane, x, fs.nswp);
    lu_back<QM>(L, a.M.m, lane, x);
  }
};

template <typename TA_, int QM, bool STAGED>
struct SparseBackend {
  using TA = TA_;
  static constexpr bool kCollective = false;  // factorisation by the whole workgroup (collective_solve)
  using Lds = SpLds<TA>;
  struct Fact {};
  static __device__ __forceinline__ int unknowns(const IntegratorArgs& a) { return a.M.m; }
  static __device__ __forceinline__ int rowDim(const IntegratorArgs& a) { return a.M.nSpec + 3; }
  static __device__ __forceinline__ int column(const IntegratorArgs& a, int ai) { return 3 + a.M.perm()[ai]; }
  static __device__ __forceinline__ int tolIndex(const IntegratorArgs& a, int ai) { return a.M.perm()[ai]; }
  static __device__ __forceinline__ void bind(const IntegratorArgs& a, Lds& L, int cell) {
    L.src = a.ext ? a.ext + size_t(cell) * a.M.m : nullptr;
  }
  // the mechanism program (LU and BACK sections) is copied into the
  // workgroup's LDS once; every wavefront of the workgroup reads it.  With
  // the programs in device memory the coefficient table alone is copied.
  static __device__ void stage(const IntegratorArgs& a, char* base) {
    const DevSparse& S = a.S;
    if (!STAGED) {
      double* cd = reinterpret_cast<double*>(base);
      for (int i = threadIdx.x; i < S.nCoef; i += blockDim.x) cd[i] = a.M.db[S.oCoef + i];
      __syncthreads();
      return;
    }
    unsigned* dst = reinterpret_cast<unsigned*>(base);
    auto copy = [&](size_t at, const int* src, int n) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) dst[at + size_t(i)] = unsigned(src[i]);
    };
    copy(0, a.M.ib + S.oProgHdr, S.luRounds + S.backRounds);
    copy(sparseLuHdrWords(S), a.M.ib + S.oProgBody, S.fwdBody + 64 * SparseSchedule::kSlack);
    const size_t e0 = sparseEvalOffsetWords(S);
    copy(e0, a.M.ib + S.oEvalHdr, S.ropRounds + S.pairRounds + S.segRounds + S.jacRounds);
    copy(e0 + sparseEvalHdrWords(S), a.M.ib + S.oEvalBody, S.evalBodyWords);
    double* cd = reinterpret_cast<double*>(dst + sparseCoefOffsetWords(S));
    for (int i = threadIdx.x; i < S.nCoef; i += blockDim.x) cd[i] = a.M.db[S.oCoef + i];
    __syncthreads();
  }
  static __device__ Lds carve(char* base, const IntegratorArgs& a, int wave) {
    return carve_sparse<TA, STAGED, !STAGED>(base, a.M, a.S, a.cellBytes, wave);
  }
  // zero slots read by padded records
  static __device__ void init(const IntegratorArgs& a, const Lds& L, int lane) {
    if (lane == 0) {
      L.kf[a.M.nReac] = 0.0;
      L.rop[a.M.nReac] = 0.0;
      L.dpr[a.M.nPairs] = TA(0);
      L.A[a.S.zeroSlot] = TA(0);
      L.rdiag[a.M.m] = TA(0);
    }
    wave_sync();
  }
  static __device__ __forceinline__ void put_c(const IntegratorArgs& a, const Lds& L, int lane, const double (&u)[QM]) {
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      if (i < a.M.m) L.c[i] = u[q];
    }
    wave_sync();
  }
  static __device__ __forceinline__ void evaluate(const IntegratorArgs& a, const Lds& L, int lane,
                                                  const double (&u)[QM], const double (&bc)[QM], double ch,
                                                  bool newton, bool refresh, double (&f)[QM], TA (&x)[QM]) {
    TCHEM_PT_BEGIN(rates)
    put_c(a, L, lane, u);
    sp_rates(a.S, L, lane, refresh);
    wave_sync();
    TCHEM_PT_END(rates, 11)
    TCHEM_PT_BEGIN(sums)
    sp_sums<TA, !STAGED>(a.S, L, lane, refresh, 1.0, -ch);
    wave_sync();
    TCHEM_PT_END(sums, 12)
    TCHEM_PT_BEGIN(rows)
    row_sums<QM>(a.M, L.seg, L.src, lane, f);
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      x[q] = (newton && i < a.M.m) ? TA(-(u[q] - bc[q] - ch * f[q])) : TA(0);
      if (newton && i < a.M.m) L.A[a.S.nnzF + i] = x[q];
    }
    if (newton) wave_sync();
    TCHEM_PT_END(rows, 13)
  }
  // after the BACK section the RHS holds u_ii x_i
  static __device__ __forceinline__ void get_x(const IntegratorArgs& a, const Lds& L, int lane, TA (&x)[QM]) {
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      x[q] = (i < a.M.m) ? L.A[a.S.nnzF + i] * L.rdiag[i] : TA(0);
    }
  }
  static __device__ void back(const IntegratorArgs& a, const Lds& L, int lane) {
    TCHEM_PT_BEGIN(back)
    program<STAGED>(L.hdr + a.S.luRounds, L.body + a.S.backBody, a.S.backRounds, L.A, L.rdiag, a.M.m, a.S.nnzF,
                    a.S.trashSlot, lane);
    TCHEM_PT_END(back, 10)
  }
  static __device__ __forceinline__ bool factor_solve(const IntegratorArgs& a, const Lds& L, int lane, TA (&x)[QM],
                                                      Fact&) {
    TCHEM_PT_BEGIN(lu)
    const bool ok = program<STAGED>(L.hdr, L.body, a.S.luRounds, L.A, L.rdiag, a.M.m, a.S.nnzF, a.S.trashSlot, lane);
    TCHEM_PT_END(lu, 9)
    if (!ok) return false;
    back(a, L, lane);
    get_x(a, L, lane, x);
    return true;
  }
  static __device__ __forceinline__ void solve(const IntegratorArgs& a, const Lds& L, int lane, TA (&x)[QM],
                                               const Fact&) {
    const unsigned* fh = reinterpret_cast<const unsigned*>(a.M.ib + a.S.oProgHdr) + a.S.luRounds + a.S.backRounds;
    const unsigned* fb = reinterpret_cast<const unsigned*>(a.M.ib + a.S.oProgBody) + a.S.fwdBody;
    program<false>(fh, fb, a.S.fwdRounds, L.A, L.rdiag, a.M.m, a.S.nnzF, a.S.trashSlot, lane);  // always in memory
    back(a, L, lane);
    get_x(a, L, lane, x);
  }
};

struct CellCounters {
  int accepted = 0, rejected = 0, newton_failures = 0, newton_iterations = 0, factorizations = 0, failed = 0;
};

// wavefront-uniform copy of a value every lane holds (lets the compiler branch on it with scalar code)
__device__ __forceinline__ double uniform_d(double v) {
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane(int(x & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane(int(x >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// One cell of AtmosphericChemistry::runDeviceBatch
// (src/core/TChem_AtmosphericChemistry.cpp:78-163) with the TR-BDF2 step
// control of oracle/oracle.c: integrate_cell, written as a state machine so
// the function evaluation and the linear solve each have a single call site:
//   STEP   start or retry a step: h = min(max(dt, dtmin), tend - t); fail when
//          t + h == t; fetch f(u_n) first if it is not known;
//   FN     f(u_n);
//   NEWTON one Newton iteration of the current stage: residual
//          G(u) = u - bc - ch f(u) and, every jacobian_interval iterations,
//          the iteration matrix I - ch J at the same u; converged when
//          max_i |dx_i| / (atol + rtol |u_i|) <= 1; after max_newton
//          iterations without convergence, or on a zero or non-finite
//          pivot, dt = h/2 and the step is retried (at most 4 attempts);
//   STAGE  f at the converged stage value: after the TR stage set up BDF2;
//          after BDF2 estimate the local error and accept or reject.
// REC: also record the accepted steps for the discrete adjoint (a.rec).
template <class B, int QM, bool REC>
__device__ void integrate_cell(const IntegratorArgs& a, const typename B::Lds& Lw, int cell, int lane) {
  // no implicit fused multiply-adds: the stage, residual and error
  // arithmetic rounds as oracle_run_batch does
#pragma clang fp contract(off)
  using TA = typename B::TA;
  const DevMech& M = a.M;
  const int m = B::unknowns(a), mg = M.m, nSpec = M.nSpec, dim = B::rowDim(a);
  typename B::Lds L = Lw;
  B::bind(a, L, cell);
  const tchem_time_advance_t ta = a.tadv[cell];
  const double tend = ta._tend;
  double* row = a.state + size_t(cell) * dim;
  CellCounters cnt;

  const double t_in = a.t_out[cell];
  if (t_in >= tend || ta._dt < 0.0) {  // skipped (.cpp:83) or already failed
    if (a.stats && lane == 0) a.stats[cell] = tchem_cell_stats_t{0, 0, 0, 0, 0, ta._dt < 0.0 ? 1 : 0, 0, 0};
    if (REC) {
      const int nCkpt0 = a.rec.maxSteps / kAdjointSegment + 1;
      for (int ai = lane; ai < m; ai += kWave) a.rec.ckpt[size_t(cell) * nCkpt0 * m + ai] = row[3 + M.perm()[ai]];
      for (int i = m + lane; i < nSpec; i += kWave) a.rec.consts[size_t(cell) * (nSpec - m) + (i - m)] = row[3 + i];
      if (lane == 0) a.rec.nsteps[cell] = ta._dt < 0.0 ? -1 : 0;
    }
    return;
  }

  const unsigned long long t_start = a.stats ? (unsigned long long)__builtin_amdgcn_s_memrealtime() : 0ull;
  const double gam = 2.0 - 1.41421356237309504880;
  const double c_tr = 0.5 * gam;
  const double a_g = 1.0 / (gam * (2.0 - gam));
  const double a_n = (1.0 - gam) * (1.0 - gam) / (gam * (2.0 - gam));
  const double c_bdf = (1.0 - gam) / (2.0 - gam);
  const double k_err = (-3.0 * gam * gam + 4.0 * gam - 2.0) / (12.0 * (2.0 - gam));

  TCHEM_PT_BEGIN(cell)
  {
    const double* kf = a.kf + size_t(cell) * M.nReac;
    for (int r = lane; r < M.nReac; r += kWave) L.kf[r] = kf[r];
  }
  for (int i = mg + lane; i < nSpec; i += kWave) L.c[i] = row[3 + i];

  // un (state) and fn (f at un) live across steps; u, bc, est per step
  double un[QM], u[QM], fn[QM], bc[QM], est[QM];
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int ai = lane + kWave * q;
    un[q] = (ai < m) ? row[B::column(a, ai)] : 0.0;
    u[q] = un[q];
    fn[q] = bc[q] = est[q] = 0.0;
  }
  const double atol_n = a.tol_newton[0], rtol_n = a.tol_newton[1];
  const int nCkpt = REC ? a.rec.maxSteps / kAdjointSegment + 1 : 0;
  bool overflow = false;
  if (REC) {
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int ai = lane + kWave * q;
      if (ai < m) a.rec.ckpt[size_t(cell) * nCkpt * m + ai] = un[q];
    }
    for (int i = m + lane; i < nSpec; i += kWave) a.rec.consts[size_t(cell) * (nSpec - m) + (i - m)] = row[3 + i];
  }
  wave_sync();

  const int max_it = ta._num_time_iterations_per_interval;
  const int max_newton = ta._max_num_newton_iterations;
  const int jint = ta._jacobian_interval < 1 ? 1 : ta._jacobian_interval;
  double t = ta._tbeg, dt = ta._dt, h = 0.0, ch = 0.0;
  bool fn_valid = false, failed = false;
  int iter = 0, attempt = 0, stage = 0, nit = 0;
  enum : int { kStep, kFn, kNewton, kStage, kDone };
  int phase = kStep;
  typename B::Fact fs;

  while (phase != kDone) {
    if (phase == kStep) {
      if (iter >= max_it || !(t < tend)) {
        phase = kDone;
        continue;
      }
      h = dt > ta._dtmin ? dt : ta._dtmin;
      if (h > tend - t) h = tend - t;
      if (!(t + h > t)) {  // no representable progress: hard failure
        failed = true;
        phase = kDone;
        continue;
      }
      if (!fn_valid) {
#pragma unroll
        for (int q = 0; q < QM; ++q) u[q] = un[q];
        phase = kFn;
        continue;
      }
#pragma unroll
      for (int q = 0; q < QM; ++q) {
        u[q] = un[q];
        bc[q] = un[q] + c_tr * h * fn[q];
      }
      ch = c_tr * h;
      stage = 0;
      nit = 0;
      phase = kNewton;
      continue;
    }

    if (phase == kNewton && nit >= max_newton) {  // Newton failed: halve the step and retry
      ++cnt.newton_failures;
      dt = 0.5 * h;
      if (++attempt >= 4) {
        failed = true;
        phase = kDone;
      } else {
        phase = kStep;
      }
      continue;
    }
    const bool nw = phase == kNewton;
    const bool refresh = nw && (nit % jint) == 0;
    double f[QM];
    TA x[QM];
    TCHEM_PT_BEGIN(ev)
    B::evaluate(a, L, lane, u, bc, ch, nw, refresh, f, x);
    TCHEM_PT_END(ev, 0)

    if (nw) {
      TCHEM_PT_BEGIN(lin)
      bool ok = true;
      if constexpr (B::kCollective) {  // one call site for the workgroup's factorisation code
        ok = B::collective_solve(a, L, lane, x, refresh);
        if (ok && refresh) ++cnt.factorizations;
      } else if (refresh) {
        ok = B::factor_solve(a, L, lane, x, fs);
        if (ok) ++cnt.factorizations;
      } else {
        B::solve(a, L, lane, x, fs);
      }
      TCHEM_PT_END(lin, refresh ? 2 : 3)
      if (!ok) {  // zero or non-finite pivot
        nit = max_newton;
        continue;
      }
      {
        double nrm = 0.0;
#pragma unroll
        for (int q = 0; q < QM; ++q) {
          const int i = lane + kWave * q;
          if (i < m) {
            const double dx = double(x[q]);
            u[q] += dx;
            double w = fabs(dx) / (atol_n + rtol_n * fabs(u[q]));
            if (!(w == w)) w = INFINITY;
            nrm = fmax(nrm, w);
          }
        }
        nrm = uniform_d(wave_max(nrm));
        ++cnt.newton_iterations;
        ++nit;
        if (nrm <= 1.0) phase = kStage;
      }
      continue;
    }

    if (phase == kFn) {
#pragma unroll
      for (int q = 0; q < QM; ++q) fn[q] = f[q];
      fn_valid = true;
      phase = kStep;
      continue;
    }

    // kStage
    if (stage == 0) {
      // f at the TR-stage value enters the error estimate; the stage value
      // itself only enters the BDF2 constant term
#pragma unroll
      for (int q = 0; q < QM; ++q) {
        est[q] = fn[q] / gam - f[q] / (gam * (1.0 - gam));
        bc[q] = a_g * u[q] - a_n * un[q];
      }
      ch = c_bdf * h;
      stage = 1;
      nit = 0;
      phase = kNewton;
      continue;
    }
    double e = 0.0;
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int ai = lane + kWave * q;
      if (ai < m) {
        const int xs = B::tolIndex(a, ai);
        const double lte = 2.0 * k_err * h * (est[q] + f[q] / (1.0 - gam));
        double w = fabs(lte) / (a.tol_time[2 * xs] + a.tol_time[2 * xs + 1] * fabs(u[q]));
        if (!(w == w)) w = INFINITY;
        e = fmax(e, w);
      }
    }
    e = uniform_d(wave_max(e));
    ++iter;
    attempt = 0;
    phase = kStep;
    if (e > 1.0 && h > ta._dtmin) {
      ++cnt.rejected;
      dt = 0.25 * h;
      continue;
    }
    t = (h == tend - t) ? tend : t + h;
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      un[q] = u[q];
      fn[q] = f[q];
    }
    dt = (e <= 0.125) ? 2.0 * h : h;
    if (dt > ta._dtmax) dt = ta._dtmax;
    if (REC) {
      const int k = cnt.accepted;  // index of this step
      if (k < a.rec.maxSteps) {
        if (lane == 0) a.rec.h[size_t(cell) * a.rec.maxSteps + k] = h;
        if ((k + 1) % kAdjointSegment == 0) {
#pragma unroll
          for (int q = 0; q < QM; ++q) {
            const int ai = lane + kWave * q;
            if (ai < m) a.rec.ckpt[(size_t(cell) * nCkpt + (k + 1) / kAdjointSegment) * m + ai] = un[q];
          }
        }
      } else {
        overflow = true;
      }
    }
    ++cnt.accepted;
  }

  TCHEM_PT_END(cell, 6)
  wave_sync();
  TCHEM_PT_FLUSH()
  double* rowo = a.state_out + size_t(cell) * dim;
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int ai = lane + kWave * q;
    if (ai < m) {
      const int col = B::column(a, ai);
      const double v = failed ? 0.0 : un[q];
      row[col] = v;
      if (rowo != row) rowo[col] = v;
    }
  }
  if (rowo != row)
    for (int i = mg + lane; i < nSpec; i += kWave) rowo[3 + i] = row[3 + i];
  if (lane == 0) {
    rowo[0] = row[0];
    rowo[1] = row[1];
    rowo[2] = row[2];
    a.t_out[cell] = failed ? tend : t;
    a.dt_out[cell] = failed ? -1.0 : dt;
    if (REC) a.rec.nsteps[cell] = (failed || overflow) ? -1 : cnt.accepted;
    if (a.stats) {
      cnt.failed = failed ? 1 : 0;
      a.stats[cell] = tchem_cell_stats_t{cnt.accepted, cnt.rejected, cnt.newton_failures, cnt.newton_iterations,
                                         cnt.factorizations, cnt.failed, t_start,
                                         (unsigned long long)__builtin_amdgcn_s_memrealtime()};
    }
  }
}


// ---- dense LU on the matrix cores (mfma_lu.hpp) ------------------------------
// One cell per workgroup of four wavefronts: wavefront 0 runs integrate_cell
// (FP64 state, residual, error estimate and step control) and hands two
// jobs to the whole workgroup: the FP64 source term (serial per-species
// sums), and at every Newton iteration the assembly of the FP32 iteration
// matrix, which the four wavefronts hold in their registers for the MFMA
// factorisation and solve (mfma::factorAndSolve).  Wavefronts 1-3 wait at
// the workgroup barrier for the next command (source term, factor and
// solve, or finish).

// d rop_r / d c_s of every derivative pair into out (the first part of
// assemble_jacobian_from_lds)
template <typename TO>
__device__ __forceinline__ void derivative_pairs(const DevMech& M, const double* c, const double* kf, TO* out,
                                                 int tid, int nthreads) {
  for (int r = tid; r < M.nReac; r += nthreads) {
    const int p0 = M.pairPtr()[r], p1 = M.pairPtr()[r + 1];
    if (p0 == p1) continue;
    const int* sp = M.reacSpec() + r * M.maxR;
    const int* pw = M.reacPow() + r * M.maxR;
    for (int p = p0; p < p1; ++p) {
      const int js = M.pairSlot()[p];
      const int ns = pw[js];
      double d = kf[r] * double(ns) * pow_small(c[sp[js]], ns - 1);
      for (int j = 0; j < M.maxR; ++j) {
        const int s = sp[j];
        if (s < 0) break;
        if (j == js) continue;
        d *= pow_small(c[s], pw[j]);
      }
      out[p] = TO(d);
    }
  }
}

// LDS bytes of one cell of the MFMA form: the dense cell region without a
// matrix, the FP32 derivative pairs, the source term, then the
// factorisation's region
inline size_t mfmaCellBytes(const DevMech& M) {
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  return al(integratorLdsBytes(M, 0)) + al(4 * size_t(M.nPairs + 1)) + al(8 * size_t(M.m)) + mfma::luLdsBytes();
}

// The source term by the whole workgroup (256 threads): rates of progress,
// then the serial per-species gathers of source_term_from_lds into fbuf.
template <class Lds>
__device__ void mfma_rhs_collective(const IntegratorArgs& a, const Lds& L) {
#pragma clang fp contract(off)
  const DevMech& M = a.M;
  const int tid = threadIdx.x;
  for (int r = tid; r < M.nReac; r += 256) L.scr[r] = rate_of_progress(M, L.c, L.kf, r);
  __syncthreads();
  for (int i = tid; i < M.m; i += 256) {
    double s = M.emission()[i];
    if (L.src) s += L.src[M.perm()[i]];
    const int e1 = M.serOmegaPtr()[i + 1];
    for (int e = M.serOmegaPtr()[i]; e < e1; ++e) {
      const double v = M.serOmegaCoef()[e] * L.scr[M.serOmegaReac()[e]];
      s = s + v;
    }
    L.fbuf[i] = s;
  }
  __syncthreads();
}

template <int QM>
struct MfmaBackend {
  using TA = float;
  static constexpr bool kCollective = true;  // factorisation by the whole workgroup (collective_solve)
  struct Lds : CellLds<float> {
    float* pairs;   // [nPairs] derivative pairs of the last Jacobian refresh (FP32)
    double* fbuf;   // [m] source term of the workgroup's evaluation
    mfma::Lu lu;
  };
  struct Fact {};
  static __device__ __forceinline__ int unknowns(const IntegratorArgs& a) { return a.M.m; }
  static __device__ __forceinline__ int rowDim(const IntegratorArgs& a) { return a.M.nSpec + 3; }
  static __device__ __forceinline__ int column(const IntegratorArgs& a, int ai) { return 3 + a.M.perm()[ai]; }
  static __device__ __forceinline__ int tolIndex(const IntegratorArgs& a, int ai) { return a.M.perm()[ai]; }
  static __device__ __forceinline__ void bind(const IntegratorArgs& a, Lds& L, int cell) {
    L.src = a.ext ? a.ext + size_t(cell) * a.M.m : nullptr;
  }
  static __device__ void stage(const IntegratorArgs&, char*) {}
  static __device__ Lds carve(char* base, const IntegratorArgs& a, int) {
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    Lds L;
    static_cast<CellLds<float>&>(L) = tchem_amd::carve<float>(base, a.M, false);
    size_t off = al(integratorLdsBytes(a.M, 0));
    L.pairs = reinterpret_cast<float*>(base + off);
    off += al(4 * size_t(a.M.nPairs + 1));
    L.fbuf = reinterpret_cast<double*>(base + off);
    off += al(8 * size_t(a.M.m));
    L.lu = mfma::carveLu(base + off);
    return L;
  }
  static __device__ void init(const IntegratorArgs&, const Lds&, int) {}
  // f (FP64, serial source term) and the Newton residual; on a refresh c h
  // of the iteration matrix (its derivative pairs and assembly are the
  // workgroup's, in factor_solve)
  static __device__ void evaluate(const IntegratorArgs& a, const Lds& L, int lane, const double (&u)[QM],
                                  const double (&bc)[QM], double ch, bool newton, bool refresh, double (&f)[QM],
                                  TA (&x)[QM]) {
#pragma clang fp contract(off)
    const int m = a.M.m;
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      if (i < m) L.c[i] = u[q];
    }
    if (lane == 0) L.lu.cmd[0] = 2;
    __syncthreads();  // hands the source term to the workgroup
    mfma_rhs_collective(a, L);
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      f[q] = i < m ? L.fbuf[i] : 0.0;
      x[q] = (newton && i < m) ? TA(-(u[q] - bc[q] - ch * f[q])) : TA(0);
    }
    if (refresh) {  // the workgroup computes the derivative pairs in collective_solve
      if (lane == 0) *L.lu.ch = ch;
      wave_sync();
    }
  }
  static __device__ bool factor_solve(const IntegratorArgs& a, const Lds& L, int lane, TA (&x)[QM], Fact&);
  // factors kept between refreshes: the kept derivative pairs and c h are
  // assembled and factored again (the same matrix, the same factors)
  static __device__ void solve(const IntegratorArgs& a, const Lds& L, int lane, TA (&x)[QM], const Fact&) {
    collective_solve(a, L, lane, x, false);
  }
  static __device__ bool collective_solve(const IntegratorArgs& a, const Lds& L, int lane, TA (&x)[QM],
                                          bool refresh);
};

// Assembly of I - c h J (FP32 sums of the serial Jacobian lists) plus the
// identity padding and the right-hand side (stashed in z) as column 255,
// by column quarters through LDS into the tiles of the four wavefronts,
// then the factorisation and back substitution.
template <int QM>
__device__ __forceinline__ void mfma_collective(const IntegratorArgs& a, const typename MfmaBackend<QM>::Lds& L,
                                                int wave, int lane) {
#pragma clang fp contract(off)
  using namespace mfma;
  const DevMech& M = a.M;
  const int m = M.m;
  const int tid = threadIdx.x;
  TCHEM_PT_BEGIN(dp)
  if (L.lu.cmd[2]) derivative_pairs(M, L.c, L.kf, L.pairs, tid, 256);  // refresh: pairs of the current iterate
  TCHEM_PT_END(dp, 1)
  const float scale = float(-*L.lu.ch);
  f32x16 acc[16];
  TCHEM_PT_BEGIN(asm)
  float4* stage4 = reinterpret_cast<float4*>(L.lu.stage);
#pragma unroll
  for (int q = 0; q < 4; ++q) {  // unrolled: the tiles of quarter q are known at compile time
    for (int idx = tid; idx < int(kStageWords / 4); idx += 256) stage4[idx] = float4{0.0f, 0.0f, 0.0f, 0.0f};
    __syncthreads();
    {  // this thread's run of terms: coalesced, independent loads
      const int j1 = M.mfjQuarter()[q + 1];
      float sum = 0.0f;
#pragma unroll 4
      for (int j = M.mfjQuarter()[q]; j < j1; ++j) {
        const unsigned w = M.mfjTermW()[size_t(j) * 256 + tid];
        sum = fmaf(M.mfjTermC()[size_t(j) * 256 + tid], L.pairs[w & 0x3fff], sum);
        if (w >> 31) {
          L.lu.stage[(w >> 14) & 0x7fff] = (((w >> 29) & 1u) ? 1.0f : 0.0f) + scale * sum;
          sum = 0.0f;
        }
      }
    }
    for (int p = m + tid; p < kN - 1; p += 256)
      if ((p >> 6) == q) L.lu.stage[p * 65 + (p - 64 * q)] = 1.0f;
    if (q == 3) {
      for (int i = tid; i < m; i += 256) L.lu.stage[i * 65 + 63] = L.lu.z[i];
      if (tid == 0) L.lu.stage[(kN - 1) * 65 + 63] = 1.0f;
    }
    __syncthreads();
    if (q == 0) loadQuarter<0>(acc, L.lu.stage, wave, lane);
    if (q == 1) loadQuarter<1>(acc, L.lu.stage, wave, lane);
    if (q == 2) loadQuarter<2>(acc, L.lu.stage, wave, lane);
    if (q == 3) loadQuarter<3>(acc, L.lu.stage, wave, lane);
    __syncthreads();
  }
  TCHEM_PT_END(asm, 4)
  factorAndSolve(acc, L.lu, wave, lane);
}

template <int QM>
__device__ bool MfmaBackend<QM>::factor_solve(const IntegratorArgs& a, const Lds& L, int lane, TA (&x)[QM], Fact&) {
  return collective_solve(a, L, lane, x, true);
}

template <int QM>
__device__ bool MfmaBackend<QM>::collective_solve(const IntegratorArgs& a, const Lds& L, int lane, TA (&x)[QM],
                                                  bool refresh) {
  const int m = a.M.m;
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int i = lane + kWave * q;
    if (i < m) L.lu.z[i] = x[q];
  }
  if (lane == 0) {
    L.lu.cmd[0] = 1;
    L.lu.cmd[1] = 0;
    L.lu.cmd[2] = refresh ? 1 : 0;
  }
  __syncthreads();  // hands the matrix to the workgroup
  mfma_collective<QM>(a, L, 0, lane);
  // x = -(U^-1 e_255)[0..m)
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int i = lane + kWave * q;
    x[q] = (i < m) ? -L.lu.z[i] : TA(0);
  }
  bool ok = L.lu.cmd[1] == 0;
#pragma unroll
  for (int q = 0; q < QM; ++q) ok = ok && isfinite(x[q]);
  return ok;
}

template <int QM>
__global__ __launch_bounds__(256, 1) void trbdf2_mfma_kernel(IntegratorArgs a) {
  extern __shared__ __align__(16) char lds_raw[];
  if (a.skip && __builtin_amdgcn_readfirstlane(*a.skip)) return;
  TCHEM_PT_RESET()
  using B = MfmaBackend<QM>;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);  // uniform: tile choices are scalar branches
  typename B::Lds L = B::carve(lds_raw, a, 0);
  if (wave != 0) {
    // helpers: one collective factorisation per command until told to finish
    for (;;) {
      __syncthreads();
      const int cmd = __builtin_amdgcn_readfirstlane(L.lu.cmd[0]);
      if (cmd == 0) break;
      if (cmd == 2) {
        mfma_rhs_collective(a, L);
      } else {
        mfma_collective<QM>(a, L, wave, lane);
      }
    }
    return;
  }
  // one call site, so integrate_cell (with the collective's register tiles) is inlined
  for (int first = 1;; first = 0) {
    int cell = a.nBatch;
    if (a.persistent) {
      if (lane == 0) cell = atomicAdd(a.queue, 1);
      cell = __builtin_amdgcn_readfirstlane(__shfl(cell, 0, kWave));
    } else if (first) {
      cell = blockIdx.x;
    }
    if (cell >= a.nBatch) break;
    integrate_cell<B, QM, false>(a, L, cell, lane);
    wave_sync();
  }
  if (lane == 0) L.lu.cmd[0] = 0;
  __syncthreads();  // releases the helpers
}

// Batched dense solve A x = b (n <= 255, FP64 in and out, FP32 MFMA
// factorisation without interchanges): the LU of the MFMA form on its own.
__global__ __launch_bounds__(256, 1) void dense_mfma_solve_kernel(int n, int nBatch, const double* A, const double* b,
                                                                  double* x) {
  extern __shared__ __align__(16) char lds_raw[];
  using namespace mfma;
  const int cell = blockIdx.x;
  if (cell >= nBatch) return;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);  // uniform: tile choices are scalar branches
  const Lu L = carveLu(lds_raw);
  const double* Ac = A + size_t(cell) * n * n;
  const double* bc = b + size_t(cell) * n;
  if (threadIdx.x == 0) L.cmd[1] = 0;
  f32x16 acc[16];
  // [A | b] padded with the identity, by column quarters through LDS
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    for (int idx = threadIdx.x; idx < kN * 64; idx += blockDim.x) {
      const int row = idx >> 6, col = 64 * q + (idx & 63);
      float v = (row == col) ? 1.0f : 0.0f;
      if (row < n && col < n) v = float(Ac[size_t(row) * n + col]);
      else if (col == kN - 1) v = row < n ? float(bc[row]) : (row == kN - 1 ? 1.0f : 0.0f);
      L.stage[row * 65 + (idx & 63)] = v;
    }
    __syncthreads();
    if (q == 0) loadQuarter<0>(acc, L.stage, wave, lane);
    if (q == 1) loadQuarter<1>(acc, L.stage, wave, lane);
    if (q == 2) loadQuarter<2>(acc, L.stage, wave, lane);
    if (q == 3) loadQuarter<3>(acc, L.stage, wave, lane);
    __syncthreads();
  }
  factorAndSolve(acc, L, wave, lane);
  for (int i = threadIdx.x; i < n; i += blockDim.x) x[size_t(cell) * n + i] = L.cmd[1] ? NAN : -double(L.z[i]);
}

// ---- single-particle SIMPOL.1 aerosol integrator ---------------------------
// AerosolChemistry::runDeviceBatch (src/core/TChem_AerosolChemistry.cpp:31-186):
// the gas-phase box model extended by gas-particle partitioning
// (src/core/impl/TChem_Impl_Aerosol_RHS.hpp:64-103,
// TChem_Impl_SIMPOL_phase_transfer.hpp:61-150).  One cell per wavefront,
// FP64 dense LU with partial pivoting over the mt gas + particle unknowns
// (the reference factors a dense Jacobian too); the Jacobian is analytic.

constexpr double kSpRuniv = 8.314472;          // src/core/TChem_Util.hpp:176
constexpr double kSpPi = 3.141592653589793;    // TChem_Util.hpp:212
constexpr double kSpAtmPa = 101325.0;          // TChem_Util.hpp:227-229
constexpr double kSpMinMass = 1.0e-25;         // TChem_Impl_SingleParticleUtils.hpp:25-30
constexpr double kSpMinMw = 0.1;
constexpr double kSpMinDensity = 1800.0;

struct AeroLds : CellLds<double> {
  double* x;    // [mt] unknowns (gas internal order, then particles)
  double* f;    // [mt] source term
  double* sc;   // [nSimpol][4] alpha, mean free path, kg m-3 -> ppm, partitioning constant
  double* ps;   // [nPart][4] effective radius, volume, mass, moles
  double* pd;   // [nPart * nSimpol][6] k_c, k_e, diff, d k_c / dV, (unused) x2
  const double* nc;  // this cell's particle number concentrations
  double T, p;
};

__device__ AeroLds carve_aero(char* base, const DevMech& M, const DevAerosol& X, bool withMatrix = true) {
  AeroLds L;
  const int mt = X.mt;
  size_t off = withMatrix ? sizeof(double) * size_t(mt) * size_t(mt) : 0;
  off = (off + 15) & ~size_t(15);
  L.A = reinterpret_cast<double*>(base);
  L.c = reinterpret_cast<double*>(base + off);
  L.kf = L.c + M.nSpec;
  L.scr = L.kf + M.nReac;
  const int nScr = M.nReac > M.nPairs ? M.nReac : M.nPairs;
  L.seg = L.scr + nScr;
  L.x = L.seg + M.nSeg;
  L.f = L.x + mt;
  L.rdiag = L.f + mt;
  L.sc = L.rdiag + mt;
  L.ps = L.sc + 4 * X.nSimpol;
  L.pd = L.ps + 4 * X.nPart;
  L.swp = reinterpret_cast<int*>(L.pd + 6 * X.nPart * X.nSimpol);
  L.src = nullptr;
  L.nc = nullptr;
  L.T = L.p = 0.0;
  return L;
}

// SIMPOL_constant.hpp:60-128 (same expression order as the reference)
__device__ void simpol_constants(const double* par, double T, double p, double* out) {
  const double B1 = par[0], B2 = par[1], B3 = par[2], B4 = par[3], D = par[4], MW = par[5], Nstar = par[6];
  double alpha = 0.1;
  if (par[7] != 0.0) {
    const double N23 = pow(Nstar, 2.0 / 3.0);
    double dh = -10.0 * (Nstar - 1.0) + 7.53 * (N23 - 1.0) - 1.0;
    double ds = -32.0 * (Nstar - 1.0) + 9.21 * (N23 - 1.0) - 1.3;
    dh = dh * 4184.0;
    ds = ds * 4.184;
    const double dG = dh - T * ds;
    alpha = exp(-dG / (kSpRuniv * T));
    alpha = alpha / (1.0 + alpha);
  }
  const double mfp = 3.0 * D / sqrt(8.0 * kSpRuniv * T / (kSpPi * MW));
  double vp = B1 / T + B2 + B3 * T + B4 * log(T);
  vp = kSpAtmPa * pow(10.0, vp);
  const double conv = kSpRuniv / MW * 1.0e6;
  out[0] = alpha;
  out[1] = mfp;
  out[2] = conv * T / p;
  out[3] = vp / p / MW * 1.0e6;
}

__device__ __forceinline__ double fuchs_sutugin(double Kn, double alpha) {
  return (0.75 * alpha * (1.0 + Kn)) / (Kn * Kn + (1.0 + 0.283 * alpha) * Kn + 0.75 * alpha);
}

// Transfer terms of every (particle, transfer) pair at the unknowns in L.x:
// per-particle sums (lanes over particles), then k_c, k_e, diff and
// d k_c / dV (lanes over pairs).  SingleParticleUtils.hpp:62-160.
template <class LdsT>
__device__ void simpol_terms(const IntegratorArgs& a, const LdsT& L, int lane) {
  const DevAerosol& X = a.X;
  const double* mw = X.db + X.oMw;
  const double* dens = X.db + X.oDens;
  for (int ip = lane; ip < X.nPart; ip += kWave) {
    const double* y = L.x + a.M.m + ip * X.nAero;
    double v = kSpMinMass / kSpMinDensity;
    for (int i = 0; i < X.nAero; ++i) v += y[i] / dens[i];
    double ms = kSpMinMass, mo = kSpMinMass / kSpMinMw;
    for (int i = 0; i < X.nAero; ++i) {
      ms += y[i];
      mo += y[i] / mw[i];
    }
    L.ps[4 * ip + 0] = pow((v * 3.0 / 4.0 / kSpPi), 1.0 / 3.0);
    L.ps[4 * ip + 1] = v;
    L.ps[4 * ip + 2] = ms;
    L.ps[4 * ip + 3] = mo;
  }
  wave_sync();
  const int nPair = X.nPart * X.nSimpol;
  for (int j = lane; j < nPair; j += kWave) {
    const int ip = j / X.nSimpol, s = j - ip * X.nSimpol;
    const double* c4 = L.sc + 4 * s;
    const double alpha = c4[0], D = X.db[X.oPar + 8 * s + 4];
    const double radius = L.ps[4 * ip], volume = L.ps[4 * ip + 1], mass = L.ps[4 * ip + 2];
    const double moles = L.ps[4 * ip + 3];
    const double Kn = c4[1] / radius;
    const double den = Kn * Kn + (1.0 + 0.283 * alpha) * Kn + 0.75 * alpha;
    const double fk = fuchs_sutugin(Kn, alpha);
    const double fp = 0.75 * alpha * (den - (1.0 + Kn) * (2.0 * Kn + 1.0 + 0.283 * alpha)) / (den * den);
    const double kc = 4.0 * kSpPi * radius * D * fk;
    const double ke = kc * 1.0 * (c4[3] * (mass / moles) / mass);
    const int g = X.ib[X.oGas + s], ai = a.M.m + ip * X.nAero + X.ib[X.oAero + s];
    double* o = L.pd + 6 * j;
    o[0] = kc;
    o[1] = ke;
    o[2] = -(ke * L.x[ai]) + kc * L.x[g];
    o[3] = 4.0 * kSpPi * D * (fk - Kn * fp) * (radius / (3.0 * volume));
  }
  wave_sync();
}

// unknown layout and per-cell bindings shared by the aerosol backends
struct AerosolLayout {
  static __device__ __forceinline__ int unknowns(const IntegratorArgs& a) { return a.X.mt; }
  static __device__ __forceinline__ int rowDim(const IntegratorArgs& a) {
    return a.M.nSpec + 3 + a.X.nAero * a.X.nPart;
  }
  static __device__ __forceinline__ int column(const IntegratorArgs& a, int ai) {
    return ai < a.M.m ? 3 + a.M.perm()[ai] : 3 + a.M.nSpec + (ai - a.M.m);
  }
  static __device__ __forceinline__ int tolIndex(const IntegratorArgs& a, int ai) {
    return ai < a.M.m ? a.M.perm()[ai] : ai;
  }
  // this cell's number concentrations, temperature and pressure, and the
  // SIMPOL constants (functions of T and p only)
  template <class LdsT>
  static __device__ void bind(const IntegratorArgs& a, LdsT& L, int cell) {
    const double* row = a.state + size_t(cell) * rowDim(a);
    L.nc = a.numberConc + size_t(cell) * a.X.nPart;
    L.T = row[2];
    L.p = row[1];
    const int lane = threadIdx.x & (kWave - 1);
    for (int s = lane; s < a.X.nSimpol; s += kWave) simpol_constants(a.X.db + a.X.oPar + 8 * s, L.T, L.p, L.sc + 4 * s);
    wave_sync();
  }
};

template <int QM>
struct AerosolBackend : AerosolLayout {
  using TA = double;
  static constexpr bool kCollective = false;  // factorisation by the whole workgroup (collective_solve)
  using Lds = AeroLds;
  struct Fact { int nswp = 0; };
  static __device__ void stage(const IntegratorArgs&, char*) {}
  static __device__ Lds carve(char* base, const IntegratorArgs& a, int wave) {
    return carve_aero(base + size_t(wave) * size_t(a.cellBytes), a.M, a.X);
  }
  static __device__ void init(const IntegratorArgs&, const Lds&, int) {}

  // f at u: gas source term, then every transfer in (particle, transfer)
  // order as the reference's serial accumulation; for a Newton iteration
  // also x = -(u - bc - ch f) and, when refreshing, I - ch J
  static __device__ void evaluate(const IntegratorArgs& a, const Lds& L, int lane, const double (&u)[QM],
                                  const double (&bc)[QM], double ch, bool newton, bool refresh, double (&f)[QM],
                                  double (&x)[QM]) {
    const DevMech& M = a.M;
    const DevAerosol& X = a.X;
    const int mt = X.mt;
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int ai = lane + kWave * q;
      if (ai < mt) L.x[ai] = u[q];
      if (ai < M.m) L.c[ai] = u[q];
    }
    wave_sync();
    double fg[QM];
    source_term_from_lds<QM>(M, L, lane, fg);
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int ai = lane + kWave * q;
      if (ai < mt) L.f[ai] = ai < M.m ? fg[q] : 0.0;
    }
    simpol_terms(a, L, lane);  // syncs first
    if (lane == 0) {
      for (int ip = 0; ip < X.nPart; ++ip)
        for (int s = 0; s < X.nSimpol; ++s) {
          const double diff = L.pd[6 * (ip * X.nSimpol + s) + 2];
          const int g = X.ib[X.oGas + s], ai = M.m + ip * X.nAero + X.ib[X.oAero + s];
          L.f[g] += -L.nc[ip] * diff;
          L.f[ai] += diff / L.sc[4 * s + 2];
        }
    }
    wave_sync();
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int ai = lane + kWave * q;
      f[q] = ai < mt ? L.f[ai] : 0.0;
      x[q] = (newton && ai < mt) ? -(u[q] - bc[q] - ch * f[q]) : 0.0;
    }
    if (refresh) assemble(a, L, lane, -ch);
  }

  // A = I + scale * J: gas-phase derivative pairs and Jacobian entries as
  // assemble_jacobian_from_lds (leading dimension mt), then the transfer
  // derivatives of SIMPOL_phase_transfer.hpp differentiated analytically:
  //   d diff / d x_gas = k_c
  //   d diff / d y_i = x_gas d k_c - y_aero d k_e - [i = aero] k_e
  //   d k_c = (d k_c / dV) / rho_i,  d k_e = K_eq d k_c / moles - k_e / (moles MW_i)
  static __device__ void assemble(const IntegratorArgs& a, const Lds& L, int lane, double scale) {
    const DevMech& M = a.M;
    const DevAerosol& X = a.X;
    const int mt = X.mt;
    for (int idx = lane; idx < mt * mt; idx += kWave) L.A[idx] = 0.0;
    wave_sync();
    for (int i = lane; i < mt; i += kWave) L.A[i * mt + i] = 1.0;
    for (int r = lane; r < M.nReac; r += kWave) {
      const int p0 = M.pairPtr()[r], p1 = M.pairPtr()[r + 1];
      if (p0 == p1) continue;
      const int* sp = M.reacSpec() + r * M.maxR;
      const int* pw = M.reacPow() + r * M.maxR;
      for (int pp = p0; pp < p1; ++pp) {
        const int js = M.pairSlot()[pp];
        const int ns = pw[js];
        double d = L.kf[r] * double(ns) * pow_small(L.c[sp[js]], ns - 1);
        for (int j = 0; j < M.maxR; ++j) {
          const int s2 = sp[j];
          if (s2 < 0) break;
          if (j == js) continue;
          d *= pow_small(L.c[s2], pw[j]);
        }
        L.scr[pp] = d;
      }
    }
    wave_sync();
    for (int n = lane; n < M.nnzJ; n += kWave) {
      double sum = 0.0;
      const int e1 = M.jnzPtr()[n + 1];
      for (int e = M.jnzPtr()[n]; e < e1; ++e) sum += M.jcCoef()[e] * L.scr[M.jcPair()[e]];
      const int rc = M.jnzRowCol()[n];
      const int row = rc >> 16, col = rc & 0xffff;
      L.A[row * mt + col] = (row == col ? 1.0 : 0.0) + scale * sum;
    }
    wave_sync();
    if (lane == 0) {
      const double* mw = X.db + X.oMw;
      const double* dens = X.db + X.oDens;
      for (int ip = 0; ip < X.nPart; ++ip) {
        const int base = M.m + ip * X.nAero;
        const double moles = L.ps[4 * ip + 3];
        for (int s = 0; s < X.nSimpol; ++s) {
          const double* o = L.pd + 6 * (ip * X.nSimpol + s);
          const double kc = o[0], ke = o[1], dkdv = o[3];
          const double keq = L.sc[4 * s + 3], kg = L.sc[4 * s + 2], n = L.nc[ip];
          const int g = X.ib[X.oGas + s], ai = base + X.ib[X.oAero + s];
          const double xg = L.x[g], ya = L.x[ai];
          L.A[g * mt + g] += scale * (-n * kc);
          L.A[ai * mt + g] += scale * (kc / kg);
          for (int i = 0; i < X.nAero; ++i) {
            const double dkc = dkdv / dens[i];
            const double dke = keq * dkc / moles - ke / (moles * mw[i]);
            double dd = xg * dkc - ya * dke;
            if (base + i == ai) dd -= ke;
            L.A[g * mt + base + i] += scale * (-n * dd);
            L.A[ai * mt + base + i] += scale * (dd / kg);
          }
        }
      }
    }
    wave_sync();
  }
  static __device__ bool factor_solve(const IntegratorArgs& a, const Lds& L, int lane, double (&x)[QM], Fact& fs) {
    if (!lu_factor_fused<QM>(L, a.X.mt, lane, x, fs.nswp)) return false;
    lu_back<QM>(L, a.X.mt, lane, x);
    return true;
  }
  static __device__ void solve(const IntegratorArgs& a, const Lds& L, int lane, double (&x)[QM], const Fact& fs) {
    lu_forward<QM>(L, a.X.mt, lane, x, fs.nswp);
    lu_back<QM>(L, a.X.mt, lane, x);
  }
};


// Sparse form of the aerosol integrator: the mechanism's compiled programs
// built over the mt unknowns (gas Jacobian pattern + the transfer
// derivatives' pattern, in elimination order: particle species first, then
// the gas species), read from device memory; the transfer terms and their derivatives are added to the
// source term and to the factor entries by one lane, in the dense form's
// order.
struct AeroSpLds : SpLds<double> {
  double* x;    // [mt] unknowns
  double* f;    // [mt] source term
  double* sc;   // [nSimpol][4]
  double* ps;   // [nPart][4]
  double* pd;   // [nPart * nSimpol][6]
  const double* nc;
  double T, p;
};

__device__ AeroSpLds carve_aero_sparse(char* base, const IntegratorArgs& a, int wave) {
  auto al = [](size_t x) { return (x + 15) & ~size_t(15); };
  const DevMech& M = a.M;
  const DevSparse& S = a.S;
  const DevAerosol& X = a.X;
  AeroSpLds L;
  L.hdr = reinterpret_cast<const unsigned*>(X.pib + S.oProgHdr);
  L.body = reinterpret_cast<const unsigned*>(X.pib + S.oProgBody);
  L.ehdr = reinterpret_cast<const unsigned*>(X.pib + S.oEvalHdr);
  L.ebody = reinterpret_cast<const unsigned*>(X.pib + S.oEvalBody);
  L.coef = X.pdb + S.oCoef;
  char* cell = base + size_t(wave) * size_t(a.cellBytes);
  size_t off = 0;
  L.A = reinterpret_cast<double*>(cell);
  off += al(8 * size_t(S.trashSlot + 1));
  L.rdiag = reinterpret_cast<double*>(cell + off);
  off += al(8 * size_t(X.mt + 2));
  L.dpr = reinterpret_cast<double*>(cell + off);
  off += al(8 * size_t(S.pairSlots));
  L.c = reinterpret_cast<double*>(cell + off);
  L.kf = L.c + M.nSpec;
  L.rop = L.kf + M.nReac + 1;
  L.seg = L.rop + M.nReac + 1;
  off += al(8 * (size_t(M.nSpec) + 2 * size_t(M.nReac + 1) + size_t(M.nSeg + 1)));
  L.x = reinterpret_cast<double*>(cell + off);
  L.f = L.x + X.mt;
  L.sc = L.f + X.mt;
  L.ps = L.sc + 4 * X.nSimpol;
  L.pd = L.ps + 4 * X.nPart;
  L.mu = L.w = L.v = nullptr;
  L.src = nullptr;
  L.nc = nullptr;
  L.T = L.p = 0.0;
  return L;
}

template <int QM>
struct AerosolSparseBackend : AerosolLayout {
  using TA = double;
  static constexpr bool kCollective = false;  // factorisation by the whole workgroup (collective_solve)
  using Lds = AeroSpLds;
  struct Fact {};
  static __device__ void stage(const IntegratorArgs&, char*) {}
  static __device__ Lds carve(char* base, const IntegratorArgs& a, int wave) { return carve_aero_sparse(base, a, wave); }
  static __device__ void init(const IntegratorArgs& a, const Lds& L, int lane) {
    if (lane == 0) {
      L.kf[a.M.nReac] = 0.0;
      L.rop[a.M.nReac] = 0.0;
      L.dpr[a.M.nPairs] = 0.0;
      L.A[a.S.zeroSlot] = 0.0;
      L.rdiag[a.X.mt] = 0.0;
    }
    wave_sync();
  }
  static __device__ void evaluate(const IntegratorArgs& a, const Lds& L, int lane, const double (&u)[QM],
                                  const double (&bc)[QM], double ch, bool newton, bool refresh, double (&f)[QM],
                                  double (&x)[QM]) {
    const DevMech& M = a.M;
    const DevAerosol& X = a.X;
    const int mt = X.mt;
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      if (i < mt) L.x[i] = u[q];
      if (i < M.m) L.c[i] = u[q];
    }
    wave_sync();
    sp_rates(a.S, L, lane, refresh);
    wave_sync();
    sp_sums(a.S, L, lane, refresh, 1.0, -ch);
    wave_sync();
    double fg[QM];
    row_sums<QM>(M, L.seg, nullptr, lane, fg);
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      if (i < mt) L.f[i] = i < M.m ? fg[q] : 0.0;
    }
    simpol_terms(a, L, lane);  // syncs first
    // Transfer terms and derivatives.  Targets in a particle's own rows or
    // columns are written by that particle's lane, the shared gas-species
    // targets (f_g and the (g, g) entries) by lane 0 afterwards; each target
    // receives its contributions in (particle, transfer) order either way.
    const int nj = 2 + 2 * X.nAero;
    const double scale = -ch;
    for (int ip = lane; ip < X.nPart; ip += kWave) {
      const double* mw = X.db + X.oMw;
      const double* dens = X.db + X.oDens;
      const int base = M.m + ip * X.nAero;
      const double moles = L.ps[4 * ip + 3];
      for (int s = 0; s < X.nSimpol; ++s) {
        const int j = ip * X.nSimpol + s;
        const double* o = L.pd + 6 * j;
        const double kc = o[0], ke = o[1], diff = o[2], dkdv = o[3];
        const double keq = L.sc[4 * s + 3], kg = L.sc[4 * s + 2], n = L.nc[ip];
        const int g = X.ib[X.oGas + s], ai = base + X.ib[X.oAero + s];
        L.f[ai] += diff / kg;
        if (refresh) {
          const int* jp = X.ib + X.oJpos + j * nj;
          const double xg = L.x[g], ya = L.x[ai];
          L.A[jp[1]] += scale * (kc / kg);
          for (int i = 0; i < X.nAero; ++i) {
            const double dkc = dkdv / dens[i];
            const double dke = keq * dkc / moles - ke / (moles * mw[i]);
            double dd = xg * dkc - ya * dke;
            if (base + i == ai) dd -= ke;
            L.A[jp[2 + i]] += scale * (-n * dd);
            L.A[jp[2 + X.nAero + i]] += scale * (dd / kg);
          }
        }
      }
    }
    wave_sync();
    if (lane == 0) {
      for (int ip = 0; ip < X.nPart; ++ip)
        for (int s = 0; s < X.nSimpol; ++s) {
          const int j = ip * X.nSimpol + s;
          const double n = L.nc[ip];
          L.f[X.ib[X.oGas + s]] += -n * L.pd[6 * j + 2];
          if (refresh) L.A[X.ib[X.oJpos + j * nj]] += scale * (-n * L.pd[6 * j]);
        }
    }
    wave_sync();
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      f[q] = i < mt ? L.f[i] : 0.0;
      x[q] = (newton && i < mt) ? -(u[q] - bc[q] - ch * f[q]) : 0.0;
      if (newton && i < mt) L.A[a.S.nnzF + pivot(a, i)] = x[q];
    }
    if (newton) wave_sync();
  }
  // elimination index of unknown i: particle species first, then the gas
  // species (aerosolPivot)
  static __device__ __forceinline__ int pivot(const IntegratorArgs& a, int i) {
    return i < a.M.m ? i + (a.X.mt - a.M.m) : i - a.M.m;
  }
  static __device__ __forceinline__ void get_x(const IntegratorArgs& a, const Lds& L, int lane, double (&x)[QM]) {
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      const int k = pivot(a, i);
      x[q] = (i < a.X.mt) ? L.A[a.S.nnzF + k] * L.rdiag[k] : 0.0;
    }
  }
  static __device__ void back(const IntegratorArgs& a, const Lds& L, int lane) {
    run_program(L.hdr + a.S.luRounds, L.body + a.S.backBody, a.S.backRounds, L.A, L.rdiag, a.X.mt, a.S.nnzF,
                a.S.trashSlot, lane);
  }
  static __device__ bool factor_solve(const IntegratorArgs& a, const Lds& L, int lane, double (&x)[QM], Fact&) {
    if (!run_program(L.hdr, L.body, a.S.luRounds, L.A, L.rdiag, a.X.mt, a.S.nnzF, a.S.trashSlot, lane)) return false;
    back(a, L, lane);
    get_x(a, L, lane, x);
    return true;
  }
  static __device__ void solve(const IntegratorArgs& a, const Lds& L, int lane, double (&x)[QM], const Fact&) {
    run_program(L.hdr + a.S.luRounds + a.S.backRounds, L.body + a.S.fwdBody, a.S.fwdRounds, L.A, L.rdiag, a.X.mt,
                a.S.nnzF, a.S.trashSlot, lane);
    back(a, L, lane);
    get_x(a, L, lane, x);
  }
};

// Batched aerosol source term (gas + transfer) at the state rows: one
// wavefront per cell; omega [nBatch][mt] in the reference's unknown order
// (active gas species, then particle species).
template <int QM>
__global__ __launch_bounds__(64) void aerosol_rhs_kernel(IntegratorArgs a) {
  extern __shared__ __align__(16) char lds_raw[];
  const int lane = threadIdx.x, cell = blockIdx.x;
  if (cell >= a.nBatch) return;
  AeroLds L = carve_aero(lds_raw, a.M, a.X, false);  // no iteration matrix
  AerosolBackend<QM>::bind(a, L, cell);
  const double* row = a.state + size_t(cell) * AerosolBackend<QM>::rowDim(a);
  const DevMech& M = a.M;
  const double t_1 = 1.0 / L.T;
  const double conv = M.convPPM * L.p * t_1;
  for (int r = lane; r < M.nReac; r += kWave) L.kf[r] = rate_constant(M, r, L.T, L.p, t_1, conv, nullptr);
  for (int i = M.m + lane; i < M.nSpec; i += kWave) L.c[i] = row[3 + i];
  double u[QM], bc[QM], f[QM], x[QM];
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int ai = lane + kWave * q;
    u[q] = ai < a.X.mt ? row[AerosolBackend<QM>::column(a, ai)] : 0.0;
    bc[q] = 0.0;
  }
  wave_sync();
  AerosolBackend<QM>::evaluate(a, L, lane, u, bc, 0.0, false, false, f, x);
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int ai = lane + kWave * q;
    if (ai < a.X.mt) a.state_out[size_t(cell) * a.X.mt + (ai < M.m ? M.perm()[ai] : ai)] = f[q];
  }
}

// Workgroup = cellsPerGroup wavefronts, one cell per wavefront at a time;
// after staging, wavefronts never synchronise with each other.
template <class B, int QM, int kThreads, bool REC>
__global__ __launch_bounds__(kThreads) void trbdf2_kernel(IntegratorArgs a) {
  extern __shared__ __align__(16) char lds_raw[];
  if (a.skip && __builtin_amdgcn_readfirstlane(*a.skip)) return;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  B::stage(a, lds_raw);
  const typename B::Lds L = B::carve(lds_raw, a, wave);
  B::init(a, L, lane);
  TCHEM_PT_RESET()
  if (!a.persistent) {
    const int cell = blockIdx.x * a.cellsPerGroup + wave;
    if (cell < a.nBatch) integrate_cell<B, QM, REC>(a, L, cell, lane);
    return;
  }
  for (;;) {
    int cell = 0;
    if (lane == 0) cell = atomicAdd(a.queue, 1);
    cell = __builtin_amdgcn_readfirstlane(__shfl(cell, 0, kWave));
    if (cell >= a.nBatch) break;
    integrate_cell<B, QM, REC>(a, L, cell, lane);
    wave_sync();
  }
}

// Rate constants of every (cell, reaction), one thread each; read by the
// integrator at the start of a cell.  Keeping the transcendental code out of
// the integrator keeps its register footprint small.
// lscale (nullable): kf_r *= exp(lscale[cell * lstride + r]), the log-scale
// parameters the discrete adjoint differentiates with respect to (lstride 0:
// one set for every cell).
// rowDim: state row length (0: nSpec + 3, the gas-phase box model's rows)
__global__ __launch_bounds__(256) void rates_kernel(DevMech M, int nBatch, const double* state, const double* photo,
                                                   const double* lscale, int lstride, double* kf, const int* skip,
                                                   int rowDim) {
  if (skip && *skip) return;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)nBatch * M.nReac) return;
  const int cell = int(idx / M.nReac), r = int(idx - (long long)cell * M.nReac);
  const double* row = state + size_t(cell) * (rowDim > 0 ? rowDim : M.nSpec + 3);
  const double p = row[1], T = row[2];
  const double t_1 = 1.0 / T;
  const double conv = M.convPPM * p * t_1;
  const double k = rate_constant(M, r, T, p, t_1, conv, photo ? photo + size_t(cell) * M.nPhoto : nullptr);
  kf[idx] = lscale ? k * tchem_exp(lscale[size_t(cell) * lstride + r]) : k;
}

// ---- E3SM / UCI kinetics batch functions --------------------------------------
// ReactionRates / RateofProgress / NetProductionRates of the reference's
// E3SM kinetics (src/core/impl/TChem_Impl_{KForward,KForwardJPL,AdjustReactions,
// RateofProgress,NetProductionRates}.hpp), one wavefront per cell:
//   mode 0: kf   (ReactionRates)
//   mode 1: rop  (RateofProgress; photolysis rates replace kf of the first nPhoto reactions)
//   mode 2: omega of the active species (NetProductionRates: external sources + nu . rop)
//   mode 3: kf with the photolysis rates applied (the box model's rate constants)
// Species concentrations are the state row's species block (active then constant).

__device__ __forceinline__ double t300_pow(double T, double B) { return B == 0.0 ? 1.0 : pow(T / 300.0, B); }
__device__ __forceinline__ double exp_or_one(double C, double t_1) { return C == 0.0 ? 1.0 : exp(C * t_1); }

__global__ __launch_bounds__(64) void e3sm_kinetics_kernel(DevE3sm E, int mode, int nBatch, const double* state,
                                                         const double* photo, int nPhoto, const double* ext,
                                                         double* out, int ldOut) {
  extern __shared__ __align__(16) char lds_raw[];
  const int cell = blockIdx.x, lane = threadIdx.x;
  if (cell >= nBatch) return;
  double* x = reinterpret_cast<double*>(lds_raw);  // [nSpec]
  double* kf = x + E.nSpec;                         // [nReac]
  double* rop = kf + E.nReac;                       // [nReac]
  double* om = rop + E.nReac;                       // [m]
  const int m = E.nSpec - E.nConst;
  const double* row = state + size_t(cell) * (E.nSpec + 3);
  const double p = row[1], T = row[2];
  const double t_1 = 1.0 / T;
  const double conv = E.convPPM * p * t_1;
  for (int i = lane; i < E.nSpec; i += kWave) x[i] = row[3 + i];
  for (int r = lane; r < E.nReac; r += kWave)
    kf[r] = (mode >= 1 && photo && r < nPhoto) ? photo[size_t(cell) * nPhoto + r] : 0.0;
  wave_sync();
  const int* ib = E.ib;
  const double* db = E.db;
  // KForward (TChem_Impl_KForward.hpp:52-141): Arrhenius and Troe without the
  // conv factor, CMAQ_H2O2 with it
  for (int i = lane; i < E.nArr; i += kWave) {
    const double* P = db + E.oArrPar + 5 * i;
    kf[ib[E.oArrIdx + i]] = P[0] * exp(P[2] * t_1) * (P[1] == 0.0 ? 1.0 : pow(T / P[3], P[1])) *
                            (P[4] == 0.0 ? 1.0 : (1.0 + P[4] * p));
  }
  for (int i = lane; i < E.nTroe; i += kWave) {
    const double* P = db + E.oTroePar + 8 * i;
    const double k0 = P[0] * exp_or_one(P[2], t_1) * t300_pow(T, P[1]);
    const double k0_kinf = k0 / (P[3] * exp_or_one(P[5], t_1) * t300_pow(T, P[4]));
    kf[ib[E.oTroeIdx + i]] = k0 / (1.0 + k0_kinf) * pow(P[6], 1.0 / (1.0 + pow(log10(k0_kinf) / P[7], 2.0)));
  }
  for (int i = lane; i < E.nCmaq; i += kWave) {
    const double* P = db + E.oCmaqPar + 6 * i;
    const int r = ib[E.oCmaqIdx + i];
    kf[r] = (P[0] * exp(P[2] * t_1) * t300_pow(T, P[1]) + P[3] * exp(P[5] * t_1) * t300_pow(T, P[4]) * conv) *
            pow(conv, double(ib[E.oConvExp + r]));
  }
  // KForwardJPL (TChem_Impl_KForwardJPL.hpp:44-143), m = [M]
  const double mM = E.M_index >= 0 ? x[E.M_index] : 0.0;
  for (int i = lane; i < E.nJpl + E.nRJpl; i += kWave) {
    const bool rj = i >= E.nJpl;
    const int j = rj ? i - E.nJpl : i;
    const double* P = db + (rj ? E.oRJplPar + 10 * j : E.oJplPar + 7 * j);
    const double k0 = P[0] * exp_or_one(P[2], t_1) * t300_pow(T, P[1]);
    const double kinf = P[3] * exp_or_one(P[5], t_1) * t300_pow(T, P[4]);
    const double xpo = k0 * mM / kinf;
    const double lx = log10(xpo);
    const double factor = P[6] == 1.0 ? 1.0 : pow(P[6], 1.0 / (1.0 + lx * lx));
    double k = k0 / (1.0 + xpo) * factor;
    if (rj) k /= P[7] * exp_or_one(P[9], t_1) * t300_pow(T, P[8]);
    kf[ib[(rj ? E.oRJplIdx : E.oJplIdx) + j]] = k;
  }
  wave_sync();
  // AdjustReactions (TChem_Impl_AdjustReactions.hpp:52-96), in list order
  if (lane == 0) {
    for (int i = 0; i < E.nAdjust; ++i) kf[ib[E.oAdjust + 2 * i]] *= x[ib[E.oAdjust + 2 * i + 1]];
    for (int i = 0; i < E.nProd; ++i) {
      const int* I = ib + E.oProdI + 7 * i;
      const double* P = db + E.oProdD + 6 * i;
      const double fc = P[0] * exp(P[1] * t_1) * x[I[4]] + P[2] * exp(P[3] * t_1) * x[I[5]] +
                        P[4] * exp(P[5] * t_1) * x[I[6]];
      for (int k = 0; k < 3; ++k) kf[I[k]] *= kf[I[3]] / fc;
    }
  }
  wave_sync();
  double* o = out + size_t(cell) * ldOut;
  if (mode == 0 || mode == 3) {
    for (int r = lane; r < E.nReac; r += kWave) o[r] = kf[r];
    return;
  }
  // RateofProgress (TChem_Impl_RateofProgress.hpp:73-87)
  for (int r = lane; r < E.nReac; r += kWave) {
    double v = kf[r];
    for (int j = 0; j < ib[E.oNreac + r]; ++j)
      v *= pow(x[ib[E.oSidx + r * E.maxSpec + j]], double(abs(int(db[E.oNuki + r * E.maxSpec + j]))));
    rop[r] = v;
  }
  wave_sync();
  if (mode == 1) {
    for (int r = lane; r < E.nReac; r += kWave) o[r] = rop[r];
    return;
  }
  // NetProductionRates (TChem_Impl_NetProductionRates.hpp:60-104): external
  // sources, then nu . rop over active species in reaction order
  for (int a = lane; a < m; a += kWave) om[a] = ext ? ext[size_t(cell) * m + a] : 0.0;
  wave_sync();
  if (lane == 0) {
    const int joff = E.maxSpec / 2;
    for (int r = 0; r < E.nReac; ++r) {
      for (int j = 0; j < ib[E.oNreac + r]; ++j) {
        const int k = ib[E.oSidx + r * E.maxSpec + j];
        if (k < m) om[k] += db[E.oNuki + r * E.maxSpec + j] * rop[r];
      }
      for (int j = 0; j < ib[E.oNprod + r]; ++j) {
        const int k = ib[E.oSidx + r * E.maxSpec + j + joff];
        if (k < m) om[k] += db[E.oNuki + r * E.maxSpec + j + joff] * rop[r];
      }
    }
  }
  wave_sync();
  for (int a = lane; a < ldOut; a += kWave) o[a] = a < m ? om[a] : 0.0;
}

// ---- discrete adjoint --------------------------------------------------------
// Reverse-mode derivative of the recorded TR-BDF2 trajectory (accepted steps
// only, step sizes held fixed) with respect to the initial active species
// and the log-scale parameters theta_r of the rate constants
// (kf_r = k_r exp(theta_r), so d f / d theta_r = nu_r rop_r).  For one step
// u_n -> u_g -> u_{n+1},
//   TR:   u_g - u_n - c_tr h (f(u_g) + f(u_n)) = 0
//   BDF2: u_{n+1} - a_g u_g + a_n u_n - c_bdf h f(u_{n+1}) = 0,
// lambda = dL/du_{n+1} propagates as
//   (I - c_bdf h J(u_{n+1}))^T mu2 = lambda
//   (I - c_tr h J(u_g))^T mu1     = a_g mu2
//   lambda_n = -a_n mu2 + mu1 + c_tr h J(u_n)^T mu1
//   dL/dtheta_r += c_bdf h rop_r(u_{n+1}) (nu_r . mu2)
//                + c_tr h (rop_r(u_g) + rop_r(u_n)) (nu_r . mu1).
// The transposed systems are assembled and factorised directly by the
// transposed programs (I - c h J^T on the transposed factor pattern).
// Checkpoints every kAdjointSegment accepted steps; each segment is replayed
// forward with the forward programs (the same arithmetic as the recording
// integrator, hence the same Newton iterates), its states kept in device
// scratch, then swept in reverse.

template <int QM>
__device__ __forceinline__ void load_vec(const double* src, int m, int lane, double (&u)[QM]) {
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int i = lane + kWave * q;
    u[q] = i < m ? src[i] : 0.0;
  }
}
template <int QM>
__device__ __forceinline__ void store_vec(double* dst, int m, int lane, const double (&u)[QM]) {
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int i = lane + kWave * q;
    if (i < m) dst[i] = u[q];
  }
}

// One Newton solve of a stage, as the integrator runs it (same refresh
// rule, convergence test and arithmetic).
template <int QM>
__device__ __forceinline__ void replay_newton(const IntegratorArgs& f, const SpLds<double>& L, int lane, double (&u)[QM],
                              const double (&bc)[QM], double ch, int max_newton, int jint, double atol, double rtol) {
  using B = SparseBackend<double, QM, false>;
  typename B::Fact fs;
  for (int nit = 0; nit < max_newton; ++nit) {
    const bool refresh = (nit % jint) == 0;
    double fv[QM], x[QM];
    B::evaluate(f, L, lane, u, bc, ch, true, refresh, fv, x);
    if (refresh) {
      if (!B::factor_solve(f, L, lane, x, fs)) return;
    } else {
      B::solve(f, L, lane, x, fs);
    }
    double nrm = 0.0;
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      const int i = lane + kWave * q;
      if (i < f.M.m) {
        u[q] += x[q];
        double w = fabs(x[q]) / (atol + rtol * fabs(u[q]));
        if (!(w == w)) w = INFINITY;
        nrm = fmax(nrm, w);
      }
    }
    if (uniform_d(wave_max(nrm)) <= 1.0) return;
  }
}

// Solves (I - ch J^T) mu = rhs at the Jacobian pairs in LDS; mu in registers
// and in L.mu, w = nu . mu.  Returns false on a zero or non-finite pivot.
template <int QM>
__device__ __forceinline__ bool adjoint_solve(const AdjointArgs& a, const SpLds<double>& LT, int lane, double ch,
                              const double (&rhs)[QM], double (&mu)[QM]) {
  const DevSparse& S = a.ST;
  const int m = a.fw.M.m;
  eval_rounds<true>(LT, lane, S.ropRounds + S.pairRounds + S.segRounds, S.jacRounds, S.jacBody, 1.0, -ch);
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int i = lane + kWave * q;
    if (i < m) LT.A[S.nnzF + i] = rhs[q];
  }
  wave_sync();
  const bool ok = run_program(LT.hdr, LT.body, S.luRounds, LT.A, LT.rdiag, m, S.nnzF, S.trashSlot, lane);
  run_program(LT.hdr + S.luRounds, LT.body + S.backBody, S.backRounds, LT.A, LT.rdiag, m, S.nnzF, S.trashSlot, lane);
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int i = lane + kWave * q;
    mu[q] = i < m ? LT.A[S.nnzF + i] * LT.rdiag[i] : 0.0;
    if (i < m) LT.mu[i] = mu[q];
  }
  wave_sync();
  eval_rounds<true>(LT, lane, S.ropRounds + S.pairRounds + S.segRounds + S.jacRounds, S.nutRounds, S.nutBody, 0.0,
                    0.0);
  wave_sync();
  return ok;
}

// grad_r += c rop_r w_r
__device__ __forceinline__ void accumulate_grad(const SpLds<double>& L, double* g, int nReac, int lane, double c) {
  for (int r = lane; r < nReac; r += kWave) g[r] += c * L.rop[r] * L.w[r];
}

template <int QM>
__device__ void adjoint_cell(const AdjointArgs& a, const SpLds<double>& L, const SpLds<double>& LT, double* g,
                             int cell, int lane) {
  const IntegratorArgs& f = a.fw;
  const DevMech& M = f.M;
  const int m = M.m, nSpec = M.nSpec, nReac = M.nReac;
  const int maxSteps = f.rec.maxSteps;
  const int nCkpt = maxSteps / kAdjointSegment + 1;
  constexpr int kTraj = 2 * kAdjointSegment + 1;
  double* lam_out = a.lambda_out + size_t(cell) * nSpec;
  double* gout = a.grad + size_t(cell) * nReac;
  // a record of other sizes: only its header (offset 0) is read
  const bool foreign = a.header[0] != f.nBatch || a.header[1] != maxSteps;
  const int n = foreign ? -1 : f.rec.nsteps[cell];
  if (n < 0) {  // failed, overflowed or foreign record
    for (int i = lane; i < nSpec; i += kWave) lam_out[i] = NAN;
    for (int r = lane; r < nReac; r += kWave) gout[r] = NAN;
    return;
  }
  const double gam = 2.0 - 1.41421356237309504880;
  const double c_tr = 0.5 * gam;
  const double a_g = 1.0 / (gam * (2.0 - gam));
  const double a_n = (1.0 - gam) * (1.0 - gam) / (gam * (2.0 - gam));
  const double c_bdf = (1.0 - gam) / (2.0 - gam);
  const double atol = a.tol[0], rtol = a.tol[1];
  const int max_newton = a.newton[2 * cell], jint = a.newton[2 * cell + 1];

  for (int r = lane; r < nReac; r += kWave) {
    L.kf[r] = f.kf[size_t(cell) * nReac + r];
    g[r] = 0.0;
  }
  for (int i = m + lane; i < nSpec; i += kWave) L.c[i] = f.rec.consts[size_t(cell) * (nSpec - m) + (i - m)];
  double lam[QM];
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int ai = lane + kWave * q;
    lam[q] = ai < m ? a.lambda_in[size_t(cell) * nSpec + M.perm()[ai]] : 0.0;
  }
  wave_sync();

  using B = SparseBackend<double, QM, false>;
  double* tr = a.traj + size_t(cell) * kTraj * m;
  const double* hrec = f.rec.h + size_t(cell) * maxSteps;
  bool ok = true;
  for (int e = (n + kAdjointSegment - 1) / kAdjointSegment - 1; e >= 0; --e) {
    const int k0 = e * kAdjointSegment;
    const int len = n - k0 < kAdjointSegment ? n - k0 : kAdjointSegment;
    // replay the segment from its checkpoint
    {
      double un[QM], u[QM], fn[QM], bc[QM], x[QM];
      load_vec<QM>(f.rec.ckpt + (size_t(cell) * nCkpt + e) * m, m, lane, un);
      store_vec<QM>(tr, m, lane, un);
#pragma unroll
      for (int q = 0; q < QM; ++q) bc[q] = 0.0;
      B::evaluate(f, L, lane, un, bc, 0.0, false, false, fn, x);
      for (int j = 0; j < len; ++j) {
        const double h = hrec[k0 + j];
#pragma unroll
        for (int q = 0; q < QM; ++q) {
          u[q] = un[q];
          bc[q] = un[q] + c_tr * h * fn[q];
        }
        replay_newton<QM>(f, L, lane, u, bc, c_tr * h, max_newton, jint, atol, rtol);
        store_vec<QM>(tr + size_t(kAdjointSegment + 1 + j) * m, m, lane, u);
#pragma unroll
        for (int q = 0; q < QM; ++q) bc[q] = a_g * u[q] - a_n * un[q];
        replay_newton<QM>(f, L, lane, u, bc, c_bdf * h, max_newton, jint, atol, rtol);
        store_vec<QM>(tr + size_t(j + 1) * m, m, lane, u);
#pragma unroll
        for (int q = 0; q < QM; ++q) un[q] = u[q];
        if (j + 1 < len) B::evaluate(f, L, lane, un, bc, 0.0, false, false, fn, x);
      }
    }
    // reverse sweep; rop and the Jacobian pairs at u_{n+1} are still in LDS
    // from the previous (later) step except for the segment's last step
    for (int j = len - 1; j >= 0; --j) {
      const double h = hrec[k0 + j];
      double u[QM], mu2[QM], mu1[QM], rhs[QM];
      if (j == len - 1) {
        load_vec<QM>(tr + size_t(j + 1) * m, m, lane, u);
        B::put_c(f, L, lane, u);
        sp_rates(a.ST, LT, lane, true);
        wave_sync();
      }
      ok &= adjoint_solve<QM>(a, LT, lane, c_bdf * h, lam, mu2);
      accumulate_grad(LT, g, nReac, lane, c_bdf * h);
      load_vec<QM>(tr + size_t(kAdjointSegment + 1 + j) * m, m, lane, u);
      B::put_c(f, L, lane, u);
      sp_rates(a.ST, LT, lane, true);
      wave_sync();
#pragma unroll
      for (int q = 0; q < QM; ++q) rhs[q] = a_g * mu2[q];
      ok &= adjoint_solve<QM>(a, LT, lane, c_tr * h, rhs, mu1);
      accumulate_grad(LT, g, nReac, lane, c_tr * h);
      load_vec<QM>(tr + size_t(j) * m, m, lane, u);
      B::put_c(f, L, lane, u);
      sp_rates(a.ST, LT, lane, true);
      wave_sync();
      accumulate_grad(LT, g, nReac, lane, c_tr * h);
      eval_rounds<true>(LT, lane, a.ST.ropRounds + a.ST.pairRounds + a.ST.segRounds + a.ST.jacRounds + a.ST.nutRounds,
                        a.ST.jtvRounds, a.ST.jtvBody, 0.0, 0.0);
      wave_sync();
#pragma unroll
      for (int q = 0; q < QM; ++q) {
        const int i = lane + kWave * q;
        lam[q] = i < m ? -a_n * mu2[q] + mu1[q] + c_tr * h * LT.v[i] : 0.0;
      }
      wave_sync();
    }
  }
  ok = ballot(!ok) == 0;
  for (int i = m + lane; i < nSpec; i += kWave) lam_out[i] = 0.0;
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int ai = lane + kWave * q;
    if (ai < m) lam_out[M.perm()[ai]] = ok ? lam[q] : NAN;
  }
  for (int r = lane; r < nReac; r += kWave) gout[r] = ok ? g[r] : NAN;
}

// One wavefront per cell, cells taken from a device work queue.
template <int QM>
__global__ __launch_bounds__(512) __attribute__((flatten)) void adjoint_kernel(AdjointArgs a) {
  extern __shared__ __align__(16) char lds_raw[];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const DevMech& M = a.fw.M;
  SpLds<double> L = carve_sparse<double, false>(lds_raw, M, a.fw.S, a.cellBytes, wave);
  auto al = [](size_t x) { return (x + 15) & ~size_t(15); };
  char* x = reinterpret_cast<char*>(L.A) + sparseCellBytes(M, a.fw.S, sizeof(double));
  L.mu = reinterpret_cast<double*>(x);
  x += al(8 * size_t(M.m + 1));
  L.w = reinterpret_cast<double*>(x);
  x += al(8 * size_t(M.nReac + 2));
  L.v = reinterpret_cast<double*>(x);
  x += al(8 * size_t(M.m + 1));
  double* g = reinterpret_cast<double*>(x);
  SpLds<double> LT = L;
  LT.hdr = reinterpret_cast<const unsigned*>(M.ib + a.ST.oProgHdr);
  LT.body = reinterpret_cast<const unsigned*>(M.ib + a.ST.oProgBody);
  LT.ehdr = reinterpret_cast<const unsigned*>(M.ib + a.ST.oEvalHdr);
  LT.ebody = reinterpret_cast<const unsigned*>(M.ib + a.ST.oEvalBody);
  LT.coef = M.db + a.ST.oCoef;
  if (lane == 0) {
    L.kf[M.nReac] = 0.0;
    L.rop[M.nReac] = 0.0;
    L.dpr[M.nPairs] = 0.0;
    L.A[a.fw.S.zeroSlot] = 0.0;
    L.rdiag[M.m] = 0.0;
    L.mu[M.m] = 0.0;
    L.w[M.nReac] = 0.0;
  }
  wave_sync();
  for (;;) {
    int cell = 0;
    if (lane == 0) cell = atomicAdd(a.fw.queue, 1);
    cell = __builtin_amdgcn_readfirstlane(__shfl(cell, 0, kWave));
    if (cell >= a.fw.nBatch) break;
    adjoint_cell<QM>(a, L, LT, g, cell, lane);
    wave_sync();
  }
}

// Per-cell Newton settings and the record header of an adjoint forward run.
__global__ void adjoint_setup_kernel(int nBatch, int maxSteps, const tchem_time_advance_t* tadv,
                                     const double* tol_newton, int* header, double* tol, int* newton) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    header[0] = nBatch;
    header[1] = maxSteps;
    tol[0] = tol_newton[0];
    tol[1] = tol_newton[1];
  }
  if (i < nBatch) {
    newton[2 * i] = tadv[i]._max_num_newton_iterations;
    newton[2 * i + 1] = tadv[i]._jacobian_interval < 1 ? 1 : tadv[i]._jacobian_interval;
  }
}

// Control words of a captured interval loop: ctl[0] = stopped, ctl[1] =
// intervals run.
__global__ void loop_reset_kernel(int* ctl) {
  if (threadIdx.x == 0) {
    ctl[0] = 0;
    ctl[1] = 0;
  }
}

// The driver's stop test after an interval (src/examples/TChem_AtmosphericChemistry.cpp:340-350):
// stop when mean_i t_out[i] > fraction * tend (tend of cell 0); a fixed
// reduction order, so the decision is reproducible.
__global__ __launch_bounds__(1024) void stop_check_kernel(int n, const double* t_out, const tchem_time_advance_t* tadv,
                                                         double fraction, int* ctl) {
  __shared__ double part[1024];
  if (ctl[0]) return;  // block-uniform
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) s += t_out[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && part[0] / double(n) > fraction * tadv[0]._tend) ctl[0] = 1;
}

// The autotuner's sample (rows 0, stride, 2 stride, ... of the caller's
// batch) copied by the device: the caller's arrays are never handed to a
// runtime copy (the runtime that owns them may not be the library's, e.g.
// under a profiler that loads its own), and the records get the trial's
// iteration cap, t_out = tbeg and dt_out = 0.  live counts cells with time
// left.
__global__ void autotune_sample_kernel(int n, long long stride, int nb, int np, int iters, const double* state,
                                       const double* photo, const tchem_time_advance_t* tadv, double* st0,
                                       double* ph, tchem_time_advance_t* ta, double* tt0, int* live) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long src = i * stride;
  for (int j = 0; j < nb; ++j) st0[size_t(i) * nb + j] = state[size_t(src) * nb + j];
  for (int j = 0; j < np; ++j) ph[size_t(i) * np + j] = photo[size_t(src) * np + j];
  tchem_time_advance_t r = tadv[src];
  r._num_time_iterations_per_interval = iters;
  ta[i] = r;
  tt0[i] = r._tbeg;
  tt0[n + i] = 0.0;
  if (r._tend > r._tbeg && r._dt > 0.0) atomicAdd(live, 1);
}

// dst[i] = src[i] for n 4-byte words (device copies of caller arrays)
__global__ void copy_words_kernel(unsigned* dst, const unsigned* src, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

__global__ void reset_counter_kernel(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

__global__ void update_tadv_kernel(int n, tchem_time_advance_t* tadv, const double* t_out, const double* dt_out,
                                   int* ctl) {
  if (ctl && ctl[0]) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (ctl && i == 0) ctl[1] += 1;  // intervals run
  if (i < n) {
    tadv[i]._tbeg = t_out[i];
    tadv[i]._dt = dt_out[i];
  }
}

template <typename TA>
__device__ void load_concentrations(const DevMech& M, const CellLds<TA>& L, int lane, const double* row) {
  for (int a = lane; a < M.m; a += kWave) L.c[a] = row[3 + M.perm()[a]];
  for (int i = M.m + lane; i < M.nSpec; i += kWave) L.c[i] = row[3 + i];
}

// Source term per cell (one wavefront per cell), reference species order.
template <int QM>
__global__ __launch_bounds__(64) void sourceterm_kernel(DevMech M, int nBatch, const double* state,
                                                        const double* photo, const double* ext, double* omega) {
  extern __shared__ __align__(16) char lds_raw[];
  const CellLds<double> L = carve<double>(lds_raw, M, false);  // no matrix region
  const int lane = threadIdx.x, cell = blockIdx.x;
  if (cell >= nBatch) return;
  const double* row = state + size_t(cell) * (M.nSpec + 3);
  load_rate_constants(M, L, lane, row[2], row[1], photo ? photo + size_t(cell) * M.nPhoto : nullptr);
  load_concentrations(M, L, lane, row);
  wave_sync();
  double f[QM];
  source_term_from_lds<QM>(M, L, lane, f);
#pragma unroll
  for (int q = 0; q < QM; ++q) {
    const int a = lane + kWave * q;
    if (a < M.m) {
      const size_t o = size_t(cell) * M.m + M.perm()[a];
      omega[o] = f[q] + (ext ? ext[o] : 0.0);
    }
  }
}

// Jacobian per cell: assembled in LDS in reference order, written out coalesced.
__global__ __launch_bounds__(64) void jacobian_kernel(DevMech M, int nBatch, const double* state, const double* photo,
                                                      double* jac) {
  extern __shared__ __align__(16) char lds_raw[];
  const CellLds<double> L = carve<double>(lds_raw, M);
  const int lane = threadIdx.x, cell = blockIdx.x;
  if (cell >= nBatch) return;
  const double* row = state + size_t(cell) * (M.nSpec + 3);
  load_rate_constants(M, L, lane, row[2], row[1], photo ? photo + size_t(cell) * M.nPhoto : nullptr);
  load_concentrations(M, L, lane, row);
  wave_sync();
  assemble_jacobian_from_lds(M, L, lane, 0.0, 1.0, true);
  const int total = M.m * M.m;
  double* out = jac + size_t(cell) * total;
  for (int idx = lane; idx < total; idx += kWave) out[idx] = L.A[idx];
}

}  // namespace

// ---- launchers (called from the C ABI) --------------------------------------

static int qmOf(int m) { return m <= 64 ? 1 : (m <= 128 ? 2 : (m <= 256 ? 4 : 0)); }

template <class B, int QM, int kThreads, bool REC = false>
static const void* integratorFn() {
  return reinterpret_cast<const void*>(&trbdf2_kernel<B, QM, kThreads, REC>);
}

// Largest workgroup of the sparse form (m <= 128): 16 wavefronts (cells) for
// the mixed form, 12 for FP64; 8 for larger mechanisms.  The occupancy of this
// latency-bound kernel sets its throughput, so each workgroup size gets the
// instance with the largest register budget that still lets its wavefronts
// be resident: FP64 up to 8 cells 512 threads (256 VGPRs), 9-12 cells 768
// (168); mixed up to 12 cells 768, 13-16 cells 1024 (128 VGPRs, 4 per SIMD,
// a few registers spilled).
int sparseMaxCellsPerGroup(int precision, int m) {
  if (m > 128) return 8;
  return precision == TCHEM_PRECISION_MIXED ? 16 : 12;
}

// kernel instance for (precision, sparse form: 0 dense, 1 staged programs,
// 2 programs in device memory; species slots per lane; recording for the
// adjoint, FP64 sparse forms only; cells per workgroup); nullptr when there
// is none
static const void* integratorKernel(int precision, int sparse, int qm, bool rec, int cells) {
  if (sparse == 3) {  // dense LU on the matrix cores: mixed precision only
    if (rec || precision != TCHEM_PRECISION_MIXED) return nullptr;
    return qm == 1 ? reinterpret_cast<const void*>(&trbdf2_mfma_kernel<1>)
                   : (qm == 2 ? reinterpret_cast<const void*>(&trbdf2_mfma_kernel<2>)
                              : reinterpret_cast<const void*>(&trbdf2_mfma_kernel<4>));
  }
#define TCHEM_PICK_D(TA) \
  return qm == 1 ? integratorFn<DenseBackend<TA, 1>, 1, 64>() \
                 : (qm == 2 ? integratorFn<DenseBackend<TA, 2>, 2, 64>() : integratorFn<DenseBackend<TA, 4>, 4, 64>())
#define TCHEM_PICK_S(TA, T12, ST, REC)                                                        \
  return qm == 1 ? integratorFn<SparseBackend<TA, 1, ST>, 1, T12, REC>()                      \
                 : (qm == 2 ? integratorFn<SparseBackend<TA, 2, ST>, 2, T12, REC>()           \
                            : integratorFn<SparseBackend<TA, 4, ST>, 4, 512, REC>())
  if (rec) {
    if (precision != TCHEM_PRECISION_FP64 || sparse == 0) return nullptr;
    if (cells > 8 && qm <= 2) {
      if (sparse == 1) TCHEM_PICK_S(double, 768, true, true);
      TCHEM_PICK_S(double, 768, false, true);
    }
    if (sparse == 1) TCHEM_PICK_S(double, 512, true, true);
    TCHEM_PICK_S(double, 512, false, true);
  }
  if (precision == TCHEM_PRECISION_FP64) {
    if (cells > 8 && qm <= 2) {
      if (sparse == 1) TCHEM_PICK_S(double, 768, true, false);
      if (sparse == 2) TCHEM_PICK_S(double, 768, false, false);
    }
    if (sparse == 1) TCHEM_PICK_S(double, 512, true, false);
    if (sparse == 2) TCHEM_PICK_S(double, 512, false, false);
    TCHEM_PICK_D(double);
  }
  if (cells > 12 && qm <= 2) {
    if (sparse == 1) TCHEM_PICK_S(float, 1024, true, false);
    if (sparse == 2) TCHEM_PICK_S(float, 1024, false, false);
  }
  if (sparse == 1) TCHEM_PICK_S(float, 768, true, false);
  if (sparse == 2) TCHEM_PICK_S(float, 768, false, false);
  TCHEM_PICK_D(float);
#undef TCHEM_PICK_D
#undef TCHEM_PICK_S
}

hipError_t launchIntegrator(const IntegratorArgs& a, int precision, int grid, size_t lds, hipStream_t stream) {
  const int qm = qmOf(a.M.m);
  if (qm == 0) return hipErrorInvalidValue;
  const void* fn = integratorKernel(precision, a.sparse, qm, a.rec.h != nullptr, a.cellsPerGroup);
  if (!fn) return hipErrorInvalidValue;
  hipFuncAttributes attr;
  const hipError_t ea = hipFuncGetAttributes(&attr, fn);
  if (ea != hipSuccess) return ea;
  const int threads = a.sparse == 3 ? 256 : 64 * a.cellsPerGroup;
  if (threads > attr.maxThreadsPerBlock) return hipErrorInvalidConfiguration;  // instance bound
  IntegratorArgs arg = a;
  void* params[] = {&arg};
  return hipLaunchKernel(fn, dim3(grid), dim3(threads), params, lds, stream);
}

// static LDS of the integrator instance (profiling builds declare some)
size_t integratorStaticLds(int precision, int sparse, int m) {
  const int qm = qmOf(m);
  if (qm == 0) return 0;
  hipFuncAttributes attr;
  if (hipFuncGetAttributes(&attr, integratorKernel(precision, sparse, qm, false, 1)) != hipSuccess) return 0;
  return attr.sharedSizeBytes;
}

hipError_t integratorMaxBlocksPerCU(int precision, int sparse, int m, int threads, size_t lds, int* blocks, bool rec) {
  const int qm = qmOf(m);
  if (qm == 0) return hipErrorInvalidValue;
  const void* fn = integratorKernel(precision, sparse, qm, rec, threads / 64);
  if (sparse == 3) threads = 256;
  if (!fn) return hipErrorInvalidValue;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, fn, threads, lds);
}

hipError_t launchRates(const DevMech& M, int nBatch, const double* state, const double* photo, const double* lscale,
                       int lstride, double* kf, const int* skip, hipStream_t stream, int rowDim) {
  const long long n = (long long)nBatch * M.nReac;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(rates_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, M, nBatch, state, photo,
                     lscale, lstride, kf, skip, rowDim);
  return hipGetLastError();
}

// ---- aerosol integrator launches (FP64 dense, mt <= 128 unknowns) ----
static const void* aerosolKernel(int mt, bool sparse) {
  if (sparse) {
    if (mt <= 64) return integratorFn<AerosolSparseBackend<1>, 1, 512>();
    if (mt <= 128) return integratorFn<AerosolSparseBackend<2>, 2, 512>();
    if (mt <= 256) return integratorFn<AerosolSparseBackend<4>, 4, 512>();
    return nullptr;
  }
  if (mt <= 64) return integratorFn<AerosolBackend<1>, 1, 256>();
  if (mt <= 128) return integratorFn<AerosolBackend<2>, 2, 256>();
  return nullptr;
}

// cells (wavefronts) per workgroup of the dense aerosol integrator: up to
// 4, as many as the LDS holds (0 when one cell does not fit or mt > 128)
int aerosolCellsPerGroup(const DevMech& M, const DevAerosol& X, size_t maxLds) {
  if (X.mt > 128) return 0;
  const size_t cb = (aerosolLdsBytes(M, X) + 15) & ~size_t(15);
  int c = int(maxLds / cb);
  return c > 4 ? 4 : c;
}

hipError_t launchAerosol(const IntegratorArgs& a, int grid, size_t lds, hipStream_t stream) {
  const bool sparse = a.sparse != 0;
  const void* fn = aerosolKernel(a.X.mt, sparse);
  if (!fn || a.cellsPerGroup < 1 || a.cellsPerGroup > (sparse ? 8 : 4)) return hipErrorInvalidValue;
  IntegratorArgs arg = a;
  void* params[] = {&arg};
  return hipLaunchKernel(fn, dim3(grid), dim3(64 * a.cellsPerGroup), params, lds, stream);
}

hipError_t launchAerosolRhs(const IntegratorArgs& a, size_t lds, hipStream_t stream) {
  if (a.nBatch <= 0) return hipSuccess;
  IntegratorArgs arg = a;
  void* params[] = {&arg};
  const void* fn = a.X.mt <= 64 ? reinterpret_cast<const void*>(&aerosol_rhs_kernel<1>)
                  : (a.X.mt <= 128 ? reinterpret_cast<const void*>(&aerosol_rhs_kernel<2>)
                                   : (a.X.mt <= 256 ? reinterpret_cast<const void*>(&aerosol_rhs_kernel<4>) : nullptr));
  if (!fn) return hipErrorInvalidValue;
  return hipLaunchKernel(fn, dim3(a.nBatch), dim3(64), params, lds, stream);
}

static const void* adjointKernel(int qm) {
  return qm == 1 ? reinterpret_cast<const void*>(&adjoint_kernel<1>)
                 : (qm == 2 ? reinterpret_cast<const void*>(&adjoint_kernel<2>)
                            : reinterpret_cast<const void*>(&adjoint_kernel<4>));
}

size_t adjointCellBytes(const DevMech& M, const DevSparse& S) {
  return sparseCellBytes(M, S, sizeof(double)) + adjointExtraBytes(M);
}

hipError_t adjointMaxBlocksPerCU(int m, int threads, size_t lds, int* blocks) {
  const int qm = qmOf(m);
  if (qm == 0) return hipErrorInvalidValue;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, adjointKernel(qm), threads, lds);
}

hipError_t launchAdjoint(const AdjointArgs& a, int grid, size_t lds, hipStream_t stream) {
  const int qm = qmOf(a.fw.M.m);
  if (qm == 0) return hipErrorInvalidValue;
  AdjointArgs arg = a;
  void* params[] = {&arg};
  return hipLaunchKernel(adjointKernel(qm), dim3(grid), dim3(64 * a.cellsPerGroup), params, lds, stream);
}

hipError_t launchAdjointSetup(int nBatch, int maxSteps, const tchem_time_advance_t* tadv, const double* tol_newton,
                              int* header, double* tol, int* newton, hipStream_t stream) {
  const int n = nBatch > 0 ? nBatch : 1;
  hipLaunchKernelGGL(adjoint_setup_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, nBatch, maxSteps, tadv,
                     tol_newton, header, tol, newton);
  return hipGetLastError();
}

hipError_t launchE3smKinetics(const DevE3sm& E, int mode, int nBatch, const double* state, const double* photo,
                              int nPhoto, const double* ext, double* out, int ldOut, hipStream_t stream) {
  if (nBatch <= 0) return hipSuccess;
  const size_t lds = sizeof(double) * (size_t(E.nSpec) + 2 * size_t(E.nReac) + size_t(E.nSpec - E.nConst));
  hipLaunchKernelGGL(e3sm_kinetics_kernel, dim3(nBatch), dim3(64), lds, stream, E, mode, nBatch, state, photo, nPhoto,
                     ext, out, ldOut);
  return hipGetLastError();
}

hipError_t launchAutotuneSample(int n, long long stride, int nb, int np, int iters, const double* state,
                                const double* photo, const tchem_time_advance_t* tadv, double* st0, double* ph,
                                tchem_time_advance_t* ta, double* tt0, int* live, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(autotune_sample_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, stride, nb, np, iters,
                     state, photo, tadv, st0, ph, ta, tt0, live);
  return hipGetLastError();
}

hipError_t launchCopyWords(void* dst, const void* src, size_t words, hipStream_t stream) {
  if (words == 0) return hipSuccess;
  const size_t blocks = std::min<size_t>((words + 255) / 256, 4096);
  hipLaunchKernelGGL(copy_words_kernel, dim3(unsigned(blocks)), dim3(256), 0, stream, static_cast<unsigned*>(dst),
                     static_cast<const unsigned*>(src), words);
  return hipGetLastError();
}

hipError_t launchResetCounter(int* counter, hipStream_t stream) {
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream, counter);
  return hipGetLastError();
}

hipError_t launchUpdateTadv(int n, tchem_time_advance_t* tadv, const double* t_out, const double* dt_out, int* ctl,
                            hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(update_tadv_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, tadv, t_out, dt_out, ctl);
  return hipGetLastError();
}

hipError_t launchLoopControl(int n, const double* t_out, const tchem_time_advance_t* tadv, double fraction, int* ctl,
                             bool reset, hipStream_t stream) {
  if (reset) hipLaunchKernelGGL(loop_reset_kernel, dim3(1), dim3(64), 0, stream, ctl);
  else hipLaunchKernelGGL(stop_check_kernel, dim3(1), dim3(1024), 0, stream, n, t_out, tadv, fraction, ctl);
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

// ---- scratch of the no-allocation entry point ---------------------------------
// Device memory the no-allocation variant of runDeviceBatch uses in place of
// library allocations: the batch's rate constants, then the work-queue
// counter.  Each GPU holds its own copy of the array, at its own address,
// for the life of the process; every call on a GPU shares that GPU's copy,
// so calls that overlap on one GPU race.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Device address of the current GPU's scratch, looked up once per device
// (hipGetSymbolAddress with that device current) and kept; host code uses
// this address, never the array's name.
hipError_t noAllocScratch(void** ptr, size_t* bytes) {
  static std::mutex mu;
  static void* cache[256] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 256) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!cache[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    cache[dev] = p;
  }
  *ptr = cache[dev];
  *bytes = kNoAllocScratchBytes;
  return hipSuccess;
}

size_t mfmaCellLdsBytes(const DevMech& M) { return mfmaCellBytes(M); }

hipError_t launchDenseMfmaSolve(int n, int nBatch, const double* A, const double* b, double* x, hipStream_t stream) {
  if (nBatch <= 0) return hipSuccess;
  hipLaunchKernelGGL(dense_mfma_solve_kernel, dim3(nBatch), dim3(256), mfma::luLdsBytes(), stream, n, nBatch, A, b,
                     x);
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

#ifdef TCHEM_PHASE_TIMERS
extern "C" int tchem_debug_phase_cycles(unsigned long long* out, int reset) {
  if (hipMemcpyFromSymbol(out, HIP_SYMBOL(tchem_amd::g_phase_cycles), sizeof(unsigned long long) * 16) != hipSuccess)
    return -1;
  if (reset) {
    const unsigned long long z[16] = {};
    if (hipMemcpyToSymbol(HIP_SYMBOL(tchem_amd::g_phase_cycles), z, sizeof(z)) != hipSuccess) return -1;
  }
  return 0;
}
#endif
