// Device-side view of the compiled mechanism (see mechanism.hpp) and the
// argument blocks of the kernels.  Plain structs passed by value as kernel
// arguments; every pointer is a device pointer owned by the kmcd object.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/tchem_atm.h"

namespace tchem_amd {

struct DevMech {
  int nSpec, nConst, m, nReac, nPhoto, maxR, nPairs, nnzJ, nSeg, nSerJ;
  double convPPM;
  // Every table lives in one int buffer and one double buffer; the kernels
  // address them by 32-bit offsets so a wavefront keeps two base pointers
  // instead of twenty.
  const int* ib;
  const double* db;
  int oPerm, oRateKind, oConvExp, oPhotoSlot, oReacSpec, oReacPow, oOmegaReac, oSegBegin, oSegEnd, oRowSegPtr,
      oPairPtr, oPairSlot, oJnzRowCol, oJnzPtr, oJcPair, oSerOmegaPtr, oSerOmegaReac, oSerJRowCol, oSerJPtr,
      oSerJPair;
  int oRatePar, oOmegaCoef, oEmission, oJcCoef, oSerOmegaCoef, oSerJCoef;
  // the serial Jacobian lists regrouped for the MFMA form's assembly
  // (m <= 255 only; offsets 0 otherwise): per column quarter, each of the
  // 256 threads of the workgroup walks its own run of terms, stored
  // [term row][thread] so that a row is one coalesced load
  int oMfjQuarter, oMfjTermW, oMfjTermC;

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
  // serial gathers of the dense form (mechanism.hpp)
  __host__ __device__ const int* serOmegaPtr() const { return ib + oSerOmegaPtr; }   // [m+1]
  __host__ __device__ const int* serOmegaReac() const { return ib + oSerOmegaReac; } // [nSerOmega]
  __host__ __device__ const double* serOmegaCoef() const { return db + oSerOmegaCoef; }
  __host__ __device__ const int* serJRowCol() const { return ib + oSerJRowCol; }     // [nSerJ] row * 65536 + col
  __host__ __device__ const int* serJPtr() const { return ib + oSerJPtr; }           // [nSerJ+1]
  __host__ __device__ const int* serJPair() const { return ib + oSerJPair; }         // [nSerContrib]
  __host__ __device__ const double* serJCoef() const { return db + oSerJCoef; }
  __host__ __device__ const int* mfjQuarter() const { return ib + oMfjQuarter; }  // [5] first term row of each quarter
  // [rows][256] term words: derivative pair (bits 0-13), stage index of the
  // entry (bits 14-28, on its last term), diagonal entry (bit 29), last
  // term of an entry (bit 31); padding terms are 0 with coefficient 0
  __host__ __device__ const unsigned* mfjTermW() const { return reinterpret_cast<const unsigned*>(ib + oMfjTermW); }
  __host__ __device__ const float* mfjTermC() const {  // [rows][256] FP32 coefficients
    return reinterpret_cast<const float*>(ib + oMfjTermC);
  }
};

// Device view of the compiled sparse programs (schedule.hpp); arrays live
// in the mechanism's int / double buffers, addressed by offsets.
struct DevSparse {
  int enabled;  // 0: mechanism has no sparse programs
  int nnzF, zeroSlot, trashSlot, pairSlots;
  // evaluation program (staged into LDS) and its coefficient table
  int oEvalHdr, oEvalBody, evalBodyWords, oCoef, nCoef;
  int ropRounds, pairRounds, segRounds, jacRounds;
  int pairBody, segBody, jacBody;
  int nutRounds, jtvRounds, nutBody, jtvBody;  // adjoint sections (transposed programs only)
  // LU program: the LU and BACK sections are staged into LDS, FWD (factor
  // reuse) is read from memory
  int oProgHdr, oProgBody;
  int luRounds, backRounds, fwdRounds;
  int backBody, fwdBody;
};

// LDS layout of a sparse-form workgroup: LU program headers and staged
// bodies, evaluation program headers and bodies, coefficient table, then one
// cell region per wavefront.  Word counts are rounded up to 4 (16 bytes).
__host__ __device__ inline size_t al4(size_t words) { return (words + 3) & ~size_t(3); }
__host__ __device__ inline size_t sparseLuHdrWords(const DevSparse& S) { return al4(size_t(S.luRounds + S.backRounds)); }
// LU and BACK bodies plus the slack the interpreter may read past the last round
__host__ __device__ inline size_t sparseLuBodyWords(const DevSparse& S) { return al4(size_t(S.fwdBody) + 64 * 5); }
__host__ __device__ inline size_t sparseEvalHdrWords(const DevSparse& S) {
  return al4(size_t(S.ropRounds + S.pairRounds + S.segRounds + S.jacRounds));
}
__host__ __device__ inline size_t sparseEvalOffsetWords(const DevSparse& S) {
  return sparseLuHdrWords(S) + sparseLuBodyWords(S);
}
__host__ __device__ inline size_t sparseCoefOffsetWords(const DevSparse& S) {
  return sparseEvalOffsetWords(S) + sparseEvalHdrWords(S) + al4(size_t(S.evalBodyWords));
}
__host__ __device__ inline size_t sparseProgramBytes(const DevSparse& S) {
  return 4 * sparseCoefOffsetWords(S) + ((8 * size_t(S.nCoef) + 15) & ~size_t(15));
}
// With the programs in device memory only the coefficient table is staged:
// it sits at the start of the workgroup's LDS, before the cell regions (its
// entries are read on the critical path of every segment and Jacobian round).
__host__ __device__ inline size_t sparseCoefBytes(const DevSparse& S) {
  return (8 * size_t(S.nCoef) + 15) & ~size_t(15);
}

// Forward record of the discrete adjoint: accepted step sizes and the state
// every kAdjointSegment accepted steps (checkpoint j = state after 16 j
// steps, internal species order).
constexpr int kAdjointSegment = 16;
struct AdjointRecord {
  int* nsteps;      // [nBatch] accepted steps, -1 when the cell failed or the record overflowed
  double* h;        // [nBatch][maxSteps]
  double* ckpt;     // [nBatch][maxSteps / 16 + 1][m]
  double* consts;   // [nBatch][nConst] constant-species concentrations
  int maxSteps;
};

// Single-particle SIMPOL.1 aerosol tables (aerosol_model.hpp).  Unknowns of
// a cell: the m active gas species (internal order), then nPart particles of
// nAero species each; state rows [nSpec + 3 + nAero * nPart].
struct DevAerosol {
  int nAero, nPart, nSimpol, mt;  // mt = m + nAero * nPart
  const int* ib;
  const double* db;
  int oMw, oDens;   // [nAero] kg mol-1, kg m-3
  int oGas, oAero;  // [nSimpol] gas unknown (internal order), aerosol species within a particle
  int oPar;         // [nSimpol][8] B1 B2 B3 B4 diffusion mw Nstar computeAlpha
  // sparse form: the mechanism's programs compiled over the mt unknowns
  // (IntegratorArgs::S) live in these buffers; oJpos indexes X.ib:
  // [nPart * nSimpol][2 + 2 nAero] factor positions of the transfer
  // derivatives (g,g), (a,g), (g, particle species i), (a, particle species i)
  const int* pib;
  const double* pdb;
  int oJpos;
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
  int sparse;                        // 0 dense LU; 1 sparse, programs staged in LDS; 2 sparse, programs in memory
  int cellsPerGroup;                 // wavefronts (cells) per workgroup
  int cellBytes;                     // LDS bytes of one cell region
  AdjointRecord rec;                 // h == nullptr: no record
  const int* skip;                   // nullable: the launch does nothing when *skip != 0 (captured interval loops)
  const double* ext;                 // nullable [nBatch][m]: per-cell external sources added to the source term
  DevAerosol X;                      // aerosol integrator only
  const double* numberConc;          // aerosol integrator only: [nBatch][nPart] particles m-3
};

// LDS bytes of one cell of the aerosol integrator: the gas cell region
// (integratorLdsBytes) with an mt x mt FP64 matrix, plus the unknown copy,
// the source term, reciprocal pivots and interchanges over mt, SIMPOL
// constants [nSimpol][4], particle sums [nPart][4] and transfer terms
// [nPart * nSimpol][6].
// LDS bytes of one cell of the sparse aerosol integrator: the sparse cell
// region (sparseCellBytes, reciprocal pivots over mt) plus the unknown copy,
// source term and SIMPOL scratch of aerosolLdsBytes.
inline size_t aerosolSparseCellBytes(const DevMech& M, const DevSparse& S, const DevAerosol& X) {
  auto al = [](size_t x) { return (x + 15) & ~size_t(15); };
  return al(8 * size_t(S.trashSlot + 1)) + al(8 * size_t(X.mt + 2)) + al(8 * size_t(S.pairSlots)) +
         al(8 * (size_t(M.nSpec) + 2 * size_t(M.nReac + 1) + size_t(M.nSeg + 1))) +
         al(8 * (2 * size_t(X.mt) + 4 * size_t(X.nSimpol) + 4 * size_t(X.nPart) +
                 6 * size_t(X.nPart) * size_t(X.nSimpol)));
}

inline size_t aerosolLdsBytes(const DevMech& M, const DevAerosol& X, bool withMatrix = true) {
  const size_t nScr = size_t(M.nReac > M.nPairs ? M.nReac : M.nPairs);
  size_t a = withMatrix ? sizeof(double) * size_t(X.mt) * size_t(X.mt) : 0;
  a = (a + 15) & ~size_t(15);
  const size_t d = size_t(M.nSpec) + size_t(M.nReac) + nScr + size_t(M.nSeg) + 3 * size_t(X.mt) +
                   4 * size_t(X.nSimpol) + 4 * size_t(X.nPart) + 6 * size_t(X.nPart) * size_t(X.nSimpol);
  return a + sizeof(double) * d + sizeof(int) * 2 * size_t(X.mt);
}

// LDS bytes of one cell: the iteration matrix in TA precision (bytesA per
// entry, 0 for kernels without a matrix), then doubles for concentrations,
// rate constants, rate-of-progress / derivative scratch and source-term
// segment sums and reciprocal pivots, then the row-interchange list of the
// factorisation.
__host__ __device__ inline size_t integratorLdsBytes(const DevMech& M, size_t bytesA) {
  const size_t nScr = size_t(M.nReac > M.nPairs ? M.nReac : M.nPairs);
  size_t a = bytesA * size_t(M.m) * size_t(M.m);
  a = (a + 15) & ~size_t(15);
  size_t b = sizeof(double) * (size_t(M.nSpec) + size_t(M.nReac) + nScr + size_t(M.nSeg));
  return a + b + sizeof(double) * size_t(M.m) + sizeof(int) * 2 * size_t(M.m);
}


// LDS bytes of one cell region of the sparse form: factor entries + RHS +
// zero and trash slots, reciprocal pivots (+ a zero for column m and a
// trash slot) and the Jacobian derivative pairs in TA precision, then
// doubles for concentrations, rate constants (+ a zero), rates of progress
// (+ zero slot) and source-term segment sums (+ trash slot).
__host__ __device__ inline size_t sparseCellBytes(const DevMech& M, const DevSparse& S, size_t bytesA) {
  auto al = [](size_t x) { return (x + 15) & ~size_t(15); };
  return al(bytesA * size_t(S.trashSlot + 1)) + al(bytesA * size_t(M.m + 2)) + al(bytesA * size_t(S.pairSlots)) +
         al(sizeof(double) * (size_t(M.nSpec) + size_t(M.nReac + 1) + size_t(M.nReac + 1) + size_t(M.nSeg + 1)));
}

// Tables of the E3SM/UCI kinetics (ReactionRates / RateofProgress /
// NetProductionRates of the reference): every rate-constant form the loader
// reads, in the reference's parameter layout, indexed into the kmcd's int
// and double buffers.
struct DevE3sm {
  const int* ib;
  const double* db;
  int nSpec, nConst, nReac, maxSpec, M_index;
  double convPPM;
  int nArr, nTroe, nCmaq, nJpl, nRJpl, nAdjust, nProd;
  int oArrIdx, oArrPar;      // [nArr], [nArr][5] A B C D E
  int oTroeIdx, oTroePar;    // [nTroe], [nTroe][8] k0_A(x1e6) k0_B k0_C kinf_A kinf_B kinf_C Fc N
  int oCmaqIdx, oCmaqPar;    // [nCmaq], [nCmaq][6] A1 B1 C1 A2(x1e6) B2 C2
  int oJplIdx, oJplPar;      // [nJpl], [nJpl][7] k0_A k0_B k0_C kinf_A kinf_B kinf_C Fc
  int oRJplIdx, oRJplPar;    // [nRJpl], [nRJpl][10] + A B C
  int oAdjust;               // [nAdjust][2] reaction, species
  int oProdI, oProdD;        // [nProd][7] reactions[3], photolysis reaction, species 1-3; [nProd][6] A1 C1 A2 C2 A3 C3
  int oConvExp, oNreac, oNprod, oSidx, oNuki;  // per reaction; [nReac][maxSpec]
};

// Arguments of the discrete adjoint kernel (backward sweep over a record).
struct AdjointArgs {
  IntegratorArgs fw;        // M, S (forward programs), nBatch, kf, rec, queue
  DevSparse ST;             // transposed programs
  const int* header;        // [4] nBatch, maxSteps of the record
  const double* tol;        // [2] Newton atol, rtol of the recording run
  const int* newton;        // [nBatch][2] max Newton iterations, Jacobian interval
  double* traj;             // [nBatch][2 kAdjointSegment + 1][m] replayed states
  const double* lambda_in;  // [nBatch][nSpec] dL/d(final species), reference order
  double* lambda_out;       // [nBatch][nSpec] dL/d(initial species); constant species 0
  double* grad;             // [nBatch][nReac] dL/dtheta per cell
  int cellsPerGroup;
  int cellBytes;            // LDS bytes of one cell region
};

// LDS of the adjoint beyond the sparse cell region: mu, w, v, grad
__host__ __device__ inline size_t adjointExtraBytes(const DevMech& M) {
  auto al = [](size_t x) { return (x + 15) & ~size_t(15); };
  return al(8 * size_t(M.m + 1)) + al(8 * size_t(M.nReac + 2)) + al(8 * size_t(M.m + 1)) + al(8 * size_t(M.nReac));
}

}  // namespace tchem_amd
