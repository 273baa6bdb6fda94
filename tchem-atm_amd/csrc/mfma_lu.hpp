// Dense LU on the matrix cores for mechanisms whose factor is dense (no
// counterpart in the reference, which factors a dense m x m matrix with
// Tines per Newton iteration, src/core/problems/TChem_Impl_AtmosphericChemistry.hpp:138-155):
// the FP32 iteration matrix I - c h J of one cell, padded to 256 x 256, lives
// in the registers of the four wavefronts of a workgroup (16 tiles of
// 32 x 32 per wavefront), and is factored in place without interchanges
// between blocks (the pivots of the fill-reducing order, as in the sparse
// form) by 32 x 32 blocks in Schur-complement form (factorAndSolve):
// the diagonal tile is inverted by one wavefront, and the panel products
// and the trailing update are 32 x 32 x 32 products of 16
// v_mfma_f32_32x32x2f32 instructions each.
// The right-hand side is column 255 (rows >= m padded with the identity), so
// the factorisation also does the forward substitution; the solution is
// x = -z[0..m), z = U^-1 e_255, by block back substitution with the kept
// inverses of the diagonal blocks.  Tile (I, J) belongs to wavefront
// 2 (I & 1) + (J & 1), so every trailing update spreads over the four
// wavefronts.
//
// MFMA register layouts (v_mfma_f32_32x32x2f32, D = A B + C):
//   A (32 x 2): lane l holds A[l % 32][l / 32];  B (2 x 32): lane l holds
//   B[l / 32][l % 32];  C/D (32 x 32): register r of lane l holds
//   C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
#pragma once

// phase timers of profiling builds (kernels.hip defines them first)
#ifndef TCHEM_PT_BEGIN
#define TCHEM_PT_BEGIN(nm)
#define TCHEM_PT_END(nm, slot)
#endif

namespace mfma {

using f32x16 = __attribute__((ext_vector_type(16))) float;

constexpr int kTile = 32;
constexpr int kNb = 8;             // tiles per side
constexpr int kN = kTile * kNb;    // padded order 256
constexpr int kLd = kTile + 1;     // LDS row stride of a tile (odd: conflict-free column reads)
constexpr int kTileWords = kTile * kLd;

__device__ __forceinline__ int crow(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// LDS of one cell of the MFMA form (floats unless noted), after the
// integrator's cell region
struct Lu {
  float* stage;    // [kN][64 + 1] column quarter of the matrix during assembly; reused for the panels
  float* lpanel;   // [kNb - 1][kTileWords] -P_ik of the current step (aliases stage)
  float* upanel;   // [kNb - 1][kTileWords] A_kj of the current step (aliases stage)
  float* diag;     // [kTileWords] diagonal tile, then D_k^-1 - I of the current step
  float* uinv;     // [kNb][kTileWords] D_k^-1 of every block (back substitution)
  float* z;        // [kN] back substitution unknowns
  float* ypart;    // [4][kTile] partial block sums of the four wavefronts
  int* cmd;        // [2] command word of the driver wavefront, zero-pivot flag
  double* ch;      // [1] c h of the current iteration matrix
};

constexpr size_t kStageWords = size_t(kN) * 65;
__host__ __device__ inline size_t luLdsBytes() {
  const size_t panels = size_t(2 * (kNb - 1)) * kTileWords;
  const size_t stage = kStageWords > panels ? kStageWords : panels;
  return 4 * (stage + kTileWords * (1 + kNb) + kN + 4 * kTile) + 16 + 16;
}

__device__ inline Lu carveLu(char* base) {
  Lu L;
  float* f = reinterpret_cast<float*>(base);
  const size_t panels = size_t(2 * (kNb - 1)) * kTileWords;
  const size_t stage = kStageWords > panels ? kStageWords : panels;
  L.stage = f;
  L.lpanel = f;
  L.upanel = f + size_t(kNb - 1) * kTileWords;
  L.diag = f + stage;
  L.uinv = L.diag + kTileWords;
  L.z = L.uinv + size_t(kNb) * kTileWords;
  L.ypart = L.z + kN;
  char* tail = reinterpret_cast<char*>(L.ypart + 4 * kTile);
  L.cmd = reinterpret_cast<int*>(tail);
  L.ch = reinterpret_cast<double*>(tail + 16);
  return L;
}

// tile t of wavefront w: (I, J) = (2 (t / 4) + w / 2, 2 (t % 4) + w % 2)
__device__ __forceinline__ int tileI(int t, int w) { return 2 * (t >> 2) + (w >> 1); }
__device__ __forceinline__ int tileJ(int t, int w) { return 2 * (t & 3) + (w & 1); }

// C-layout tile -> LDS row-major [32][kLd]
__device__ __forceinline__ void storeTile(const f32x16& c, float* dst, int lane) {
#pragma unroll
  for (int r = 0; r < 16; ++r) dst[crow(r, lane) * kLd + (lane & 31)] = c[r];
}
// -(C-layout tile) -> LDS row-major [32][kLd]
__device__ __forceinline__ void storeTileNeg(const f32x16& c, float* dst, int lane) {
#pragma unroll
  for (int r = 0; r < 16; ++r) dst[crow(r, lane) * kLd + (lane & 31)] = -c[r];
}
// LDS row-major [32][ld] -> C-layout tile
__device__ __forceinline__ f32x16 loadTile(const float* src, int ld, int lane) {
  f32x16 c;
#pragma unroll
  for (int r = 0; r < 16; ++r) c[r] = src[crow(r, lane) * ld + (lane & 31)];
  return c;
}

// acc += A B for 32 x 32 A, B in LDS (row-major, stride kLd): 16 MFMAs of K = 2
__device__ __forceinline__ f32x16 mma32(f32x16 acc, const float* A, const float* B, int lane) {
  const int i = lane & 31, kk = lane >> 5;
#pragma unroll
  for (int s = 0; s < kTile / 2; ++s) {
    const float a = A[i * kLd + 2 * s + kk];
    const float b = B[(2 * s + kk) * kLd + i];
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
  }
  return acc;
}

__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wavefront inverts a tile in its own registers (C layout), in place,
// by block Gauss-Jordan with 2 x 2 pivot blocks P = {p, p+1} and no
// interchanges (the pivots of the fill-reducing order, as in the sparse
// form).  With S = M_PP^-1, one step is the rank-2 update
//   M <- M + A B,  A = E_P - M_:P (32 x 2),
//   B = S M_P: with columns P replaced by S + I (2 x 32),
// which leaves S M_PQ in the pivot rows, -M_OP S in the pivot columns, S in
// the pivot block and M_OQ - M_OP S M_PQ elsewhere: after 16 steps the tile
// holds M^-1.  Each step is one v_mfma_f32_32x32x2f32; its operands are
// read from a column-major copy of the tile in LDS (buf, [32][36] floats,
// 16-byte aligned).  A zero or non-finite pivot block sets *bad.
constexpr int kCm = 36;  // column stride of the column-major copy
static_assert(kStageWords >= size_t(2 * (kNb - 1)) * kTileWords + size_t(kTile) * kCm,
              "the inversion buffer sits in the stage region past the panels");
__device__ __forceinline__ void invertTileInPlace(f32x16& acc, float* buf, int lane, bool* bad) {
#pragma clang fp contract(off)
  const int c = lane & 31, h = lane >> 5;
  bool ok = true;
  for (int p = 0; p < kTile; p += 2) {
    // registers 4j..4j+3 hold rows 8j + 4h .. 8j + 4h + 3 of column c
#pragma unroll
    for (int j = 0; j < 4; ++j)
      *reinterpret_cast<float4*>(buf + c * kCm + 8 * j + 4 * h) = float4{acc[4 * j], acc[4 * j + 1], acc[4 * j + 2],
                                                                           acc[4 * j + 3]};
    wsync();
    const float mcol = buf[(p + h) * kCm + c];  // M[c][p + h]
    const float r0 = buf[c * kCm + p], r1 = buf[c * kCm + p + 1];  // M[p][c], M[p + 1][c]
    const float m00 = buf[p * kCm + p], m10 = buf[p * kCm + p + 1];
    const float m01 = buf[(p + 1) * kCm + p], m11 = buf[(p + 1) * kCm + p + 1];
    wsync();  // every read done before the next step's copy
    const float det = m00 * m11 - m01 * m10;
    ok = ok && det != 0.0f && isfinite(det);
    const float rdet = __builtin_amdgcn_rcpf(det);
    const float s0 = h ? -m10 * rdet : m11 * rdet;  // row h of S
    const float s1 = h ? m00 * rdet : -m01 * rdet;
    const float a = (c == p + h ? 1.0f : 0.0f) - mcol;
    float b = s0 * r0 + s1 * r1;
    if (c == p) b = s0 + (h == 0 ? 1.0f : 0.0f);
    if (c == p + 1) b = s1 + (h == 1 ? 1.0f : 0.0f);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
  }
  if (!ok) *bad = true;
}

// After invertTileInPlace: D^-1 into dinvK (kept for the back substitution)
// and D^-1 - I into L.diag (the panel operand: with the identity taken out,
// P_ik = A_ik D^-1 = A_ik + A_ik (D^-1 - I) is an accumulation into the tile
// registers), both row-major with stride kLd.
__device__ __forceinline__ void storeInverse(const f32x16& inv, const Lu& L, float* dinvK, int lane) {
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = crow(r, lane), col = lane & 31;
    dinvK[row * kLd + col] = inv[r];
    L.diag[row * kLd + col] = row == col ? inv[r] - 1.0f : inv[r];
  }
}

// Block LU without interchanges between blocks, Schur-complement form:
//   step k: P_ik = A_ik D_k^-1 (D_k the diagonal tile, inverted by its
//   owner), A_ij -= P_ik A_kj for i, j > k.
// The block row A_kj (j > k) is left as it stands, so the factor is
// A = L U with L_ik = P_ik (unit diagonal blocks) and U_kj = A_kj, U_kk = D_k;
// the right-hand side is column 255, so the elimination also applies L^-1
// to it and the solution needs only the back substitution with U, whose
// diagonal blocks are inverted already.  Look-ahead: in step k's trailing
// update the owner of tile (k+1, k+1) updates it first and inverts it while
// the other wavefronts finish their tiles.
//
// All four wavefronts of the workgroup call this with their own tiles (wave
// = 0..3); it contains the workgroup barriers.  Every write to a tile is an
// MFMA accumulation into its own registers.  On return z[0..256) holds
// U^-1 e_255; cmd[1] is set when a pivot was zero or not finite.
__device__ __forceinline__ void factorAndSolve(f32x16 (&acc)[16], const Lu& L, int wave, int lane) {
  bool bad = false;
  float* invBuf = L.stage + size_t(2 * (kNb - 1)) * kTileWords;  // past the panels: free during the factorisation
  // step -1 only inverts D_0; step k < 7 inverts D_{k+1} in its trailing update
  for (int k = -1; k < kNb; ++k) {
    TCHEM_PT_BEGIN(pn)
    if (k >= 0) {
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const int I = tileI(t, wave), J = tileJ(t, wave);
        if (J == k && I > k) {  // P_ik, kept negated for the trailing update
          float* slot = L.lpanel + size_t(I - k - 1) * kTileWords;
          storeTile(acc[t], slot, lane);
          wsync();
          acc[t] = mma32(acc[t], slot, L.diag, lane);
          wsync();
          storeTileNeg(acc[t], slot, lane);
        }
        if (I == k && J > k) storeTile(acc[t], L.upanel + size_t(J - k - 1) * kTileWords, lane);  // U_kj
      }
      __syncthreads();
    }
    TCHEM_PT_END(pn, 7)
    TCHEM_PT_BEGIN(dg)
    const int n = k + 1;  // next diagonal tile, updated and inverted first by its owner
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int I = tileI(t, wave), J = tileJ(t, wave);
      if (I == n && J == n && n < kNb) {  // tile (n, n) belongs to wavefront 3 (n & 1)
        if (k >= 0)
          acc[t] = mma32(acc[t], L.lpanel + size_t(I - k - 1) * kTileWords, L.upanel + size_t(J - k - 1) * kTileWords,
                         lane);
        invertTileInPlace(acc[t], invBuf, lane, &bad);  // the tile is not needed after its inverse
        storeInverse(acc[t], L, L.uinv + size_t(n) * kTileWords, lane);
      }
    }
    TCHEM_PT_END(dg, 5)
    TCHEM_PT_BEGIN(tr)
    if (k >= 0) {
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const int I = tileI(t, wave), J = tileJ(t, wave);
        if (I > k && J > k && !(I == n && J == n))
          acc[t] = mma32(acc[t], L.lpanel + size_t(I - k - 1) * kTileWords, L.upanel + size_t(J - k - 1) * kTileWords,
                         lane);
      }
    }
    __syncthreads();
    TCHEM_PT_END(tr, 8)
  }
  if (bad && lane == 0) L.cmd[1] = 1;
  TCHEM_PT_BEGIN(bs)
  // z = U^-1 e_255: the last block is column 31 of D_7^-1, then blocks 6..0
  if (wave == 0 && lane < kTile) L.z[(kNb - 1) * kTile + lane] = L.uinv[size_t(kNb - 1) * kTileWords + lane * kLd + 31];
  __syncthreads();
  for (int I = kNb - 2; I >= 0; --I) {
    // the two wavefronts holding block row I: U_IJ z_J of their tiles, row
    // sums by a transposed store (the stage region is free by now)
    if ((wave >> 1) == (I & 1)) {
      float part[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) part[r] = 0.0f;
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const int J = tileJ(t, wave);
        if (tileI(t, wave) == I && J > I) {
          const float zj = L.z[J * kTile + (lane & 31)];
#pragma unroll
          for (int r = 0; r < 16; ++r) part[r] = fmaf(acc[t][r], zj, part[r]);
        }
      }
      float* yt = L.stage + (wave & 1) * kTileWords;
#pragma unroll
      for (int r = 0; r < 16; ++r) yt[crow(r, lane) * kLd + (lane & 31)] = part[r];
    }
    __syncthreads();
    if (wave == 0) {
      // lane i + 32 h sums row i of half h; y_I = -(both halves)
      const float* yt = L.stage + (lane >> 5) * kTileWords + (lane & 31) * kLd;
      float y = 0.0f;
#pragma unroll
      for (int c = 0; c < kTile; ++c) y += yt[c];
      y += __shfl_xor(y, 32, 64);
      if (lane < kTile) L.ypart[lane] = -y;
      wsync();
      // z_I = D_I^-1 y_I: lane i + 32 h takes columns 16 h .. 16 h + 15 of row i
      const float* ui = L.uinv + size_t(I) * kTileWords + (lane & 31) * kLd + 16 * (lane >> 5);
      const float* yv = L.ypart + 16 * (lane >> 5);
      float zi = 0.0f;
#pragma unroll
      for (int c = 0; c < 16; ++c) zi = fmaf(ui[c], yv[c], zi);
      zi += __shfl_xor(zi, 32, 64);
      if (lane < kTile) L.z[I * kTile + lane] = zi;
    }
    __syncthreads();
  }
  TCHEM_PT_END(bs, 14)
}

// Loads the tiles of wavefront `wave` whose column block pair is q (tiles
// t with t % 4 == q) from the staged column quarter [kN][65].
template <int Q>
__device__ __forceinline__ void loadQuarter(f32x16 (&acc)[16], const float* stage, int wave, int lane) {
#pragma unroll
  for (int t = Q; t < 16; t += 4)
    acc[t] = loadTile(stage + (tileI(t, wave) * kTile) * 65 + (wave & 1) * kTile, 65, lane);
}

}  // namespace mfma
