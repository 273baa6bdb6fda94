// Device-side view of the compiled mechanism (see mechanism.hpp) and the
// argument blocks of the kernels.  Plain structs passed by value as kernel
// arguments; every pointer is a device pointer owned by the kmcd object.
#pragma once

#include "../../include/tchem_atm.h"

namespace tchem_amd {

struct DevMech {
  int nSpec, nConst, m, nReac, nPhoto, maxR, nPairs, nnzJ;
  double convPPM;
  const int* rateKind;      // [nReac]
  const double* ratePar;    // [nReac][8]
  const int* convExp;       // [nReac]
  const int* photoSlot;     // [nReac]
  const int* reacSpec;      // [nReac][maxR]
  const int* reacPow;       // [nReac][maxR]
  const int* omegaPtr;      // [m+1]
  const int* omegaReac;
  const double* omegaCoef;
  const double* emission;   // [m]
  const int* pairPtr;       // [nReac+1]
  const int* pairSlot;      // [nPairs]
  const int* jnzRowCol;     // [nnzJ] row * 65536 + col
  const int* jnzPtr;        // [nnzJ+1]
  const int* jcPair;
  const double* jcCoef;
};

struct IntegratorArgs {
  DevMech M;
  int nBatch;
  const double* tol_newton;          // [2]
  const double* tol_time;            // [m][2]
  const tchem_time_advance_t* tadv;  // [nBatch]
  double* state;                     // [nBatch][nSpec+3]
  const double* photo;               // [nBatch][nPhoto] or null
  double* t_out;
  double* dt_out;
  double* state_out;
  tchem_cell_stats_t* stats;         // or null
  int* queue;                        // persistent schedule: next-cell counter
  int persistent;
};

// LDS bytes of one cell in the TR-BDF2 kernel: iteration matrix in TA
// precision, then doubles for concentrations, rate constants and the
// rate-of-progress / derivative scratch, then pivots.
inline size_t integratorLdsBytes(const DevMech& M, size_t bytesA) {
  const size_t nScr = size_t(M.nReac > M.nPairs ? M.nReac : M.nPairs);
  size_t a = bytesA * size_t(M.m) * size_t(M.m);
  a = (a + 15) & ~size_t(15);
  size_t b = sizeof(double) * (size_t(M.nSpec) + size_t(M.nReac) + nScr);
  return a + b + sizeof(int) * size_t(M.m);
}

}  // namespace tchem_amd
