// mep_mfma.hpp -- small complex matrix products of one k-point on FP64 MFMA.
//
// C (n x n) = op(A) * B for n <= 32, every operand complex in LDS,
// column-major with leading dimension LD (NMAX + 1 keeps the strided
// fragment reads conflict-free).  The 4 waves of a 256-thread workgroup own
// the 2 x 2 output tiles of 16 x 16; each tile is accumulated with
// v_mfma_f64_16x16x4_f64 as C_re += Ar Br - Ai Bi, C_im += Ar Bi + Ai Br,
// with two independent accumulator pairs (even / odd k-steps).
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

template <bool ADJ_A, class BLoad>
__device__ inline void zgemm_tiles(int n, const zd* __restrict__ A, int lda, const BLoad& bload,
                                   zd* __restrict__ C, int ldc) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r0 = 16 * (w & 1), c0 = 16 * (w >> 1);
  if (r0 >= n || c0 >= n) return;
  const int ar = r0 + (lane & 15), bc = c0 + (lane & 15), kq = lane >> 4;
  mep_v4d re0 = {0, 0, 0, 0}, im0 = {0, 0, 0, 0}, re1 = {0, 0, 0, 0}, im1 = {0, 0, 0, 0};
  for (int k0 = 0; k0 < n; k0 += 8) {
    const int ka = k0 + kq, kb = k0 + 4 + kq;
    zd a0 = zd{0.0, 0.0}, b0 = zd{0.0, 0.0}, a1 = zd{0.0, 0.0}, b1 = zd{0.0, 0.0};
    if (ar < n && ka < n) a0 = ADJ_A ? zconj(A[ka + lda * ar]) : A[ar + lda * ka];
    if (bc < n && ka < n) b0 = bload(ka, bc);
    if (ar < n && kb < n) a1 = ADJ_A ? zconj(A[kb + lda * ar]) : A[ar + lda * kb];
    if (bc < n && kb < n) b1 = bload(kb, bc);
    re0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.re, b0.re, re0, 0, 0, 0);
    im0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.re, b0.im, im0, 0, 0, 0);
    re1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.re, b1.re, re1, 0, 0, 0);
    im1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.re, b1.im, im1, 0, 0, 0);
    re0 = __builtin_amdgcn_mfma_f64_16x16x4f64(-a0.im, b0.im, re0, 0, 0, 0);
    im0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.im, b0.re, im0, 0, 0, 0);
    re1 = __builtin_amdgcn_mfma_f64_16x16x4f64(-a1.im, b1.im, re1, 0, 0, 0);
    im1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.im, b1.re, im1, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = r0 + kq + 4 * v, col = c0 + (lane & 15);
    if (row < n && col < n) C[row + ldc * col] = zd{re0[v] + re1[v], im0[v] + im1[v]};
  }
}
