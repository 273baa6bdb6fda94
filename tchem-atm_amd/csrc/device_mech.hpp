// Device-side view of the compiled mechanism (see mechanism.hpp) and the
// argument blocks of the kernels.  Plain structs passed by value as kernel
// arguments; every pointer is a device pointer owned by the kmcd object.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/tchem_atm.h"

namespace tchem_amd {

struct DevMech {
  int nSpec, nConst, m, nReac, nPhoto, maxR, nPairs, nnzJ, nSeg;
  double convPPM;
  // Every table lives in one int buffer and one double buffer; the kernels
  // address them by 32-bit offsets so a wavefront keeps two base pointers
  // instead of twenty.
  const int* ib;
  const double* db;
  int oPerm, oRateKind, oConvExp, oPhotoSlot, oReacSpec, oReacPow, oOmegaReac, oSegBegin, oSegEnd, oRowSegPtr,
      oPairPtr, oPairSlot, oJnzRowCol, oJnzPtr, oJcPair;
  int oRatePar, oOmegaCoef, oEmission, oJcCoef;

  __host__ __device__ const int* perm() const { return ib + oPerm; }             // [m] internal -> reference species
  __host__ __device__ const int* rateKind() const { return ib + oRateKind; }     // [nReac]
  __host__ __device__ const int* convExp() const { return ib + oConvExp; }       // [nReac]
  __host__ __device__ const int* photoSlot() const { return ib + oPhotoSlot; }   // [nReac]
  __host__ __device__ const int* reacSpec() const { return ib + oReacSpec; }     // [nReac][maxR]
  __host__ __device__ const int* reacPow() const { return ib + oReacPow; }       // [nReac][maxR]
  __host__ __device__ const int* omegaReac() const { return ib + oOmegaReac; }   // [nnzOmega]
  __host__ __device__ const int* segBegin() const { return ib + oSegBegin; }     // [nSeg]
  __host__ __device__ const int* segEnd() const { return ib + oSegEnd; }         // [nSeg]
  __host__ __device__ const int* rowSegPtr() const { return ib + oRowSegPtr; }   // [m+1]
  __host__ __device__ const int* pairPtr() const { return ib + oPairPtr; }       // [nReac+1]
  __host__ __device__ const int* pairSlot() const { return ib + oPairSlot; }     // [nPairs]
  __host__ __device__ const int* jnzRowCol() const { return ib + oJnzRowCol; }   // [nnzJ] row * 65536 + col
  __host__ __device__ const int* jnzPtr() const { return ib + oJnzPtr; }         // [nnzJ+1]
  __host__ __device__ const int* jcPair() const { return ib + oJcPair; }         // [nContrib]
  __host__ __device__ const double* ratePar() const { return db + oRatePar; }    // [nReac][8]
  __host__ __device__ const double* omegaCoef() const { return db + oOmegaCoef; }// [nnzOmega]
  __host__ __device__ const double* emission() const { return db + oEmission; }  // [m]
  __host__ __device__ const double* jcCoef() const { return db + oJcCoef; }      // [nContrib]
};

// Device view of the compiled sparse schedule (schedule.hpp); arrays live in
// the mechanism's int / double buffers, addressed by offsets.
struct DevSparse {
  int enabled;  // 0: mechanism has no sparse schedule
  int nnzF, zeroSlot, nLevels, nFwdLevels, nBackLevels;
  int ropRounds, segRounds, pairRounds, pairSlots, jacRounds;
  int oRopRec, oSegRound, oSegTarget, oSegReac, oPairRec, oJacRound, oJacTarget, oJacPair;
  int oLevMul, oLevUpd, oMulRec, oUpdRound, oUpdTarget, oUpdPair;
  int oFwdLev, oFwdRound, oFwdTarget, oFwdPair;
  int oBackLev, oBackRound, oBackTarget, oBackPair;
  int oSegCoef, oJacCoef;  // double buffer
};

struct IntegratorArgs {
  DevMech M;
  DevSparse S;
  int nBatch;
  const double* tol_newton;          // [2]
  const double* tol_time;            // [m][2]
  const tchem_time_advance_t* tadv;  // [nBatch]
  double* state;                     // [nBatch][nSpec+3]
  const double* photo;               // [nBatch][nPhoto] or null
  const double* kf;                  // [nBatch][nReac] rate constants (rates_kernel)
  double* t_out;
  double* dt_out;
  double* state_out;
  tchem_cell_stats_t* stats;         // or null
  int* queue;                        // persistent schedule: next-cell counter
  int persistent;
  int sparse;                        // 1: sparse static-pivot linear algebra
};

// LDS bytes of one cell: the iteration matrix in TA precision (bytesA per
// entry, 0 for kernels without a matrix), then doubles for concentrations,
// rate constants, rate-of-progress / derivative scratch and source-term
// segment sums and reciprocal pivots, then the row-interchange list of the
// factorisation.
inline size_t integratorLdsBytes(const DevMech& M, size_t bytesA) {
  const size_t nScr = size_t(M.nReac > M.nPairs ? M.nReac : M.nPairs);
  size_t a = bytesA * size_t(M.m) * size_t(M.m);
  a = (a + 15) & ~size_t(15);
  size_t b = sizeof(double) * (size_t(M.nSpec) + size_t(M.nReac) + nScr + size_t(M.nSeg));
  return a + b + sizeof(double) * size_t(M.m) + sizeof(int) * 2 * size_t(M.m);
}


// LDS bytes of one cell for the sparse form: factor entries + RHS + zero
// slot and reciprocal pivots in TA precision, then doubles for
// concentrations, rate constants (+ a zero), rates of progress (+ zero
// slot), derivative pairs (+ zero slot, padded to whole rounds) and
// source-term segment sums.
inline size_t sparseLdsBytes(const DevMech& M, const DevSparse& S, size_t bytesA) {
  size_t a = bytesA * size_t(S.zeroSlot + 1);
  a = (a + 15) & ~size_t(15);
  size_t r = bytesA * size_t(M.m);
  r = (r + 15) & ~size_t(15);
  const size_t d = sizeof(double) * (size_t(M.nSpec) + size_t(M.nReac + 1) + size_t(M.nReac + 1) +
                                     size_t(S.pairSlots) + size_t(M.nSeg));
  return a + r + d;
}

}  // namespace tchem_amd
