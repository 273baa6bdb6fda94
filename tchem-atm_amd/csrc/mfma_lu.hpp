// Dense LU on the matrix cores for mechanisms whose factor is dense (no
// counterpart in the reference, which factors a dense m x m matrix with
// Tines per Newton iteration, src/core/problems/TChem_Impl_AtmosphericChemistry.hpp:138-155):
// the FP32 iteration matrix I - c h J of one cell, padded to 256 x 256, lives
// in the accumulator registers of the four wavefronts of a workgroup (16
// tiles of 32 x 32 per wavefront, 256 registers per lane), and is factored
// in place without row interchanges (the pivots of the fill-reducing order,
// as the sparse form) by 32 x 32 blocks:
//   for k = 0..7: factor the diagonal tile (one wavefront, rows in
//   registers), invert its two triangles (L^-1, U^-1 in LDS), panels
//   L_ik = A_ik U_kk^-1 and U_kj = L_kk^-1 A_kj and the trailing update
//   A_ij -= L_ik U_kj, each a 32 x 32 x 32 product of 16
//   v_mfma_f32_32x32x2f32 instructions.
// The right-hand side is column 255 (rows >= m padded with the identity), so
// the factorisation also does the forward substitution; the solution is
// x = -z[0..m), z = U^-1 e_255, by block back substitution with the kept
// U_kk^-1.  Tile (I, J) belongs to wavefront 2 (I & 1) + (J & 1), so every
// trailing update spreads over the four wavefronts.
//
// MFMA register layouts (v_mfma_f32_32x32x2f32, D = A B + C):
//   A (32 x 2): lane l holds A[l % 32][l / 32];  B (2 x 32): lane l holds
//   B[l / 32][l % 32];  C/D (32 x 32): register r of lane l holds
//   C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
#pragma once

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
  float* lpanel;   // [kNb - 1][kTileWords] -L_ik of the current step (aliases stage)
  float* upanel;   // [kNb - 1][kTileWords] U_kj (aliases stage)
  float* diag;     // [kTileWords] diagonal tile, factored in place
  float* linv;     // [kTileWords] L_kk^-1
  float* uinv;     // [kNb][kTileWords] U_kk^-1 of every block (back substitution)
  float* z;        // [kN] back substitution unknowns
  float* ypart;    // [4][kTile] partial block sums of the four wavefronts
  int* cmd;        // [2] command word of the driver wavefront, zero-pivot flag
  double* ch;      // [1] c h of the current iteration matrix
};

constexpr size_t kStageWords = size_t(kN) * 65;
__host__ __device__ inline size_t luLdsBytes() {
  const size_t panels = size_t(2 * (kNb - 1)) * kTileWords;
  const size_t stage = kStageWords > panels ? kStageWords : panels;
  return 4 * (stage + kTileWords * (2 + kNb) + kN + 4 * kTile) + 16 + 16;
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
  L.linv = L.diag + kTileWords;
  L.uinv = L.linv + kTileWords;
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

// One wavefront factors the diagonal tile in LDS (no interchanges; row i on
// lane i, 32 lanes) and leaves L^-1 - I in linv, U^-1 in uinvK (kept for the
// back substitution) and U^-1 - I in diag: with the identity taken out, the
// panels L_ik = A_ik U^-1 = A_ik + A_ik (U^-1 - I) and U_kj = L^-1 A_kj =
// A_kj + (L^-1 - I) A_kj become accumulations into the tile registers, the
// same form as the trailing update.  A zero or non-finite pivot sets *bad.
__device__ inline void factorDiag(const Lu& L, float* uinvK, int lane, bool* bad) {
#pragma clang fp contract(off)
  float row[kTile];
  const int i = lane & 31;
#pragma unroll
  for (int j = 0; j < kTile; ++j) row[j] = L.diag[i * kLd + j];
  bool ok = true;
#pragma unroll
  for (int p = 0; p < kTile; ++p) {
    // row p of the partly factored tile, read by every lane from lane p
    float up[kTile];
#pragma unroll
    for (int j = p; j < kTile; ++j) up[j] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(row[j]), p));
    ok = ok && up[p] != 0.0f && isfinite(up[p]);
    if (i > p) {
      const float l = row[p] / up[p];
      row[p] = l;
#pragma unroll
      for (int j = p + 1; j < kTile; ++j) row[j] = fmaf(-l, up[j], row[j]);
    }
  }
  if (lane < 32) {
#pragma unroll
    for (int j = 0; j < kTile; ++j) L.diag[i * kLd + j] = row[j];
  }
  wsync();
  // lanes 0..31: column c of L^-1 (forward, unit diagonal); lanes 32..63:
  // column c of U^-1 (backward)
  const int c = lane & 31;
  float x[kTile];
  if (lane < 32) {
#pragma unroll
    for (int r = 0; r < kTile; ++r) {
      float s = (r == c) ? 1.0f : 0.0f;
#pragma unroll
      for (int p = 0; p < r; ++p) s = fmaf(-L.diag[r * kLd + p], x[p], s);
      x[r] = s;
    }
#pragma unroll
    for (int r = 0; r < kTile; ++r) L.linv[r * kLd + c] = (r == c) ? 0.0f : x[r];
  } else {
#pragma unroll
    for (int r = kTile - 1; r >= 0; --r) {
      float s = (r == c) ? 1.0f : 0.0f;
#pragma unroll
      for (int p = r + 1; p < kTile; ++p) s = fmaf(-L.diag[r * kLd + p], x[p], s);
      x[r] = s / L.diag[r * kLd + r];
    }
#pragma unroll
    for (int r = 0; r < kTile; ++r) uinvK[r * kLd + c] = x[r];
  }
  wsync();
  if (lane >= 32) {
#pragma unroll
    for (int r = 0; r < kTile; ++r) L.diag[r * kLd + c] = (r == c) ? x[r] - 1.0f : x[r];
  }
  if (!ok) *bad = true;
}

// Factorisation and back substitution of the assembled tiles.  All four
// wavefronts of the workgroup call it with their own tiles (wave = 0..3); it
// contains the workgroup barriers.  Every write to a tile is an MFMA
// accumulation into its own registers.  On return z[0..256) holds
// U^-1 e_255; cmd[1] is set when a pivot was zero or not finite.
__device__ __forceinline__ void factorAndSolve(f32x16 (&acc)[16], const Lu& L, int wave, int lane) {
  for (int k = 0; k < kNb; ++k) {
    bool bad = false;
    if (wave == 2 * (k & 1) + (k & 1)) {
#pragma unroll
      for (int t = 0; t < 16; ++t)
        if (tileI(t, wave) == k && tileJ(t, wave) == k) storeTile(acc[t], L.diag, lane);
      wsync();
      factorDiag(L, L.uinv + size_t(k) * kTileWords, lane, &bad);
    }
    if (bad && lane == 0) L.cmd[1] = 1;
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int I = tileI(t, wave), J = tileJ(t, wave);
      if (J == k && I > k) {  // L_ik, kept negated for the trailing update
        float* slot = L.lpanel + size_t(I - k - 1) * kTileWords;
        storeTile(acc[t], slot, lane);
        wsync();
        acc[t] = mma32(acc[t], slot, L.diag, lane);
        wsync();
        storeTileNeg(acc[t], slot, lane);
      }
      if (I == k && J > k) {  // U_kj, kept in the registers for the back substitution
        float* slot = L.upanel + size_t(J - k - 1) * kTileWords;
        storeTile(acc[t], slot, lane);
        wsync();
        acc[t] = mma32(acc[t], L.linv, slot, lane);
        wsync();
        storeTile(acc[t], slot, lane);
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int I = tileI(t, wave), J = tileJ(t, wave);
      if (I > k && J > k)
        acc[t] = mma32(acc[t], L.lpanel + size_t(I - k - 1) * kTileWords, L.upanel + size_t(J - k - 1) * kTileWords,
                       lane);
    }
    __syncthreads();
  }
  // z = U^-1 e_255: the last block is column 31 of U_77^-1, then blocks 6..0
  if (wave == 0 && lane < kTile) L.z[(kNb - 1) * kTile + lane] = L.uinv[size_t(kNb - 1) * kTileWords + lane * kLd + 31];
  __syncthreads();
  for (int I = kNb - 2; I >= 0; --I) {
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
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      float v = part[r];
#pragma unroll
      for (int off = 16; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      if ((lane & 31) == 0) L.ypart[wave * kTile + crow(r, lane)] = v;
    }
    __syncthreads();
    if (wave == 0) {
      float y = 0.0f;
      if (lane < kTile) y = -(L.ypart[lane] + L.ypart[kTile + lane] + L.ypart[2 * kTile + lane] + L.ypart[3 * kTile + lane]);
      wsync();
      if (lane < kTile) L.ypart[lane] = y;
      wsync();
      if (lane < kTile) {
        const float* ui = L.uinv + size_t(I) * kTileWords + lane * kLd;
        float zi = 0.0f;
#pragma unroll
        for (int c = 0; c < kTile; ++c) zi = fmaf(ui[c], L.ypart[c], zi);
        L.z[I * kTile + lane] = zi;
      }
    }
    __syncthreads();
  }
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
