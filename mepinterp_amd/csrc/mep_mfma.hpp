// mep_mfma.hpp -- small complex matrix products of one k-point on FP64 MFMA.
//
// C (n x n) = op(A) * B for n <= NM (NM = 16 or 32), every operand complex
// in LDS, column-major with leading dimension LD (NM + 1 keeps the strided
// fragment reads conflict-free).  The operands must be zero in rows and
// columns n .. NM-1: the loads are then unconditional.  The 4 waves of a
// 256-thread workgroup own the 2 x 2 output tiles of 16 x 16; each tile is
// accumulated with v_mfma_f64_16x16x4_f64, three real products per complex
// one: P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi), Re C = P1 - P2,
// Im C = P3 - P1 - P2.
// Fragment layout (measured on gfx950): A lane l = row l%16, k l/16;
// B lane l = col l%16, k l/16; accumulator element v of lane l = row
// l/16 + 4v, col l%16.
#pragma once
#include "mep_common.hpp"

typedef double mep_v4d __attribute__((ext_vector_type(4)));

// B element loaders: (k-row, col) -> complex
struct BPlain {
  const zd* B;
  int ld;
  __device__ inline zd operator()(int r, int c) const { return B[r + ld * c]; }
};

template <int NM, bool ADJ_A, class BLoad>
__device__ inline void zgemm_tiles(int n, const zd* __restrict__ A, int lda, const BLoad& bload,
                                   zd* __restrict__ C, int ldc) {
  static_assert(NM % 16 == 0, "tiles of 16");
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r0 = 16 * (w & 1), c0 = 16 * (w >> 1);
  if (r0 >= n || c0 >= n) return;  // wave-uniform
  const int ar = r0 + (lane & 15), bc = c0 + (lane & 15), kq = lane >> 4;
  mep_v4d p1 = {0, 0, 0, 0}, p2 = {0, 0, 0, 0}, p3 = {0, 0, 0, 0};
#pragma unroll
  for (int k0 = 0; k0 < NM; k0 += 4) {
    if (k0 >= n) break;
    const int m = k0 + kq;
    const zd a = ADJ_A ? zconj(A[m + lda * ar]) : A[ar + lda * m];
    const zd b = bload(m, bc);
    p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.re, p1, 0, 0, 0);
    p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.im, p2, 0, 0, 0);
    p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re + a.im, b.re + b.im, p3, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = r0 + kq + 4 * v, col = c0 + (lane & 15);
    if (row < n && col < n) C[row + ldc * col] = zd{p1[v] - p2[v], p3[v] - p1[v] - p2[v]};
  }
}
