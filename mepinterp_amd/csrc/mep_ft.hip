// mep_ft.hip -- Fourier sums H(k), dH(k), A(k), Omega(k) as one batched
// complex GEMM  C[k, j] = sum_R P[k, R] * X[R, j]  on gfx950 MFMA.
//
// Replaces the per-k R-loop of FT_R_to_k (wann_interp.f90:151-384).  The
// k-independent factors i*c_a(R) (velocity) and i*(c_a r_b - c_b r_a)
// (curvature) are folded into the columns of X(R) once, at upload; the
// tight-binding phase e^{2 pi i k.(tau_m - tau_n)} is a per-element epilogue
// applied by the consumers (mep_eigh.hip, mep_velo.hip).
//
// Layouts (complex interleaved, row-major over the GEMM):
//   P  [M = k-block][K = nR]     phase tile, written by phase_kernel
//   X  [K = nR][N = nmat*n*n]    column j = mat*n*n + row + n*col
//   C  [M][N]                    per k: nmat column-major n x n matrices
#include "mep_common.hpp"
#include "mep_kernels.hpp"

// --------------------------------------------------------------------------
// k-points of a Monkhorst-Pack block, in the reference's order
// (core.f90:165-167, k_space.f90:360-362): linear index ki (0-based)
// -> (kix fastest, kiy, kiz).  Bit-exact with the oracle.
// --------------------------------------------------------------------------
__global__ void mp_kpts_kernel(int64_t k0, int64_t nk, int n1, int n2, int n3,
                               double* __restrict__ kpts) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nk) return;
  int64_t ki = k0 + t;
  int64_t x = ki % n1, y = (ki / n1) % n2, z = ki / ((int64_t)n1 * n2);
  kpts[3 * t + 0] = (2.0 * (double)(x + 1) - (double)n1 - 1.0) / (2.0 * (double)n1);
  kpts[3 * t + 1] = (2.0 * (double)(y + 1) - (double)n2 - 1.0) / (2.0 * (double)n2);
  kpts[3 * t + 2] = (2.0 * (double)(z + 1) - (double)n3 - 1.0) / (2.0 * (double)n3);
}

// P[k][R] = exp(i * 2pi * (k . R))  (wann_interp.f90:258-260 angle order)
__global__ void phase_kernel(int64_t nk, int nR, const double* __restrict__ kpts,
                             const double* __restrict__ R, zd* __restrict__ P) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nk * nR) return;
  int64_t k = t / nR;
  int s = (int)(t - k * nR);
  const double two_pi = 2.0 * 3.14159265358979323846;
  double ang = two_pi * (kpts[3 * k] * R[3 * s] + kpts[3 * k + 1] * R[3 * s + 1] +
                         kpts[3 * k + 2] * R[3 * s + 2]);
  double sn, cs;
  sincos(ang, &sn, &cs);
  P[t] = zd{cs, sn};
}

// --------------------------------------------------------------------------
// FP64 complex GEMM on v_mfma_f64_16x16x4_f64.
// Block: 256 threads = 4 waves in a 2x2 grid, block tile 64 (k) x 64 (cols)
// complex; each wave owns 32x32 = 2x2 MFMA tiles, with separate real and
// imaginary accumulators (C_re += Ar Br - Ai Bi, C_im += Ar Bi + Ai Br).
// Operand fragments (16x4 A, 4x16 B): lane l holds row/col l%16 at k l/16;
// accumulators of the f64 instruction: element v of lane l is row l/16 + 4v,
// col l%16 (measured on gfx950; the f32/bf16 16x16 forms use 4*(l/16) + v).
// --------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));

#define FT_BM 64
#define FT_BN 64
#define FT_KC 16

__global__ __launch_bounds__(256) void ft_gemm_f64_kernel(
    int M, int N, int K, const zd* __restrict__ P, const zd* __restrict__ X,
    zd* __restrict__ C) {
  __shared__ double As_re[FT_KC][FT_BM], As_im[FT_KC][FT_BM];
  __shared__ double Bs_re[FT_KC][FT_BN], Bs_im[FT_KC][FT_BN];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t m0 = (int64_t)blockIdx.y * FT_BM;
  const int64_t n0 = (int64_t)blockIdx.x * FT_BN;

  v4d acc_re[2][2], acc_im[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      acc_re[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
      acc_im[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
    }

  for (int k0 = 0; k0 < K; k0 += FT_KC) {
    // stage A: 64 rows x 16 k  (consecutive threads -> consecutive R of a row)
#pragma unroll
    for (int it = 0; it < (FT_BM * FT_KC) / 256; ++it) {
      int idx = tid + 256 * it;
      int row = idx / FT_KC, kk = idx % FT_KC;
      zd v = zd{0.0, 0.0};
      if (m0 + row < M && k0 + kk < K) v = P[(m0 + row) * (int64_t)K + k0 + kk];
      As_re[kk][row] = v.re;
      As_im[kk][row] = v.im;
    }
    // stage B: 16 k x 64 cols (consecutive threads -> consecutive columns)
#pragma unroll
    for (int it = 0; it < (FT_BN * FT_KC) / 256; ++it) {
      int idx = tid + 256 * it;
      int kk = idx / FT_BN, col = idx % FT_BN;
      zd v = zd{0.0, 0.0};
      if (n0 + col < N && k0 + kk < K) v = X[(int64_t)(k0 + kk) * N + n0 + col];
      Bs_re[kk][col] = v.re;
      Bs_im[kk][col] = v.im;
    }
    __syncthreads();
#pragma unroll
    for (int kq = 0; kq < FT_KC / 4; ++kq) {
      const int kk = kq * 4 + lk;
      double ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        ar[a] = As_re[kk][32 * wr + 16 * a + lr];
        ai[a] = As_im[kk][32 * wr + 16 * a + lr];
        nai[a] = -ai[a];
        br[a] = Bs_re[kk][32 * wc + 16 * a + lr];
        bi[a] = Bs_im[kk][32 * wc + 16 * a + lr];
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          acc_re[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], acc_re[a][b], 0, 0, 0);
          acc_im[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], acc_im[a][b], 0, 0, 0);
          acc_re[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[a], bi[b], acc_re[a][b], 0, 0, 0);
          acc_im[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], acc_im[a][b], 0, 0, 0);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int64_t row = m0 + 32 * wr + 16 * a + lk + 4 * v;
        int64_t col = n0 + 32 * wc + 16 * b + lr;
        if (row < M && col < N) C[row * (int64_t)N + col] = zd{acc_re[a][b][v], acc_im[a][b][v]};
      }
}

// --------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------
void launch_mp_kpts(hipStream_t s, int64_t k0, int64_t nk, const int mp[3], double* kpts) {
  int64_t nb = (nk + 255) / 256;
  hipLaunchKernelGGL(mp_kpts_kernel, dim3((unsigned)nb), dim3(256), 0, s, k0, nk, mp[0], mp[1],
                     mp[2], kpts);
}

void launch_phase(hipStream_t s, int64_t nk, int nR, const double* kpts, const double* R, zd* P) {
  int64_t tot = nk * nR;
  int64_t nb = (tot + 255) / 256;
  hipLaunchKernelGGL(phase_kernel, dim3((unsigned)nb), dim3(256), 0, s, nk, nR, kpts, R, P);
}

void launch_ft_gemm_f64(hipStream_t s, int M, int N, int K, const zd* P, const zd* X, zd* C) {
  dim3 grid((N + FT_BN - 1) / FT_BN, (M + FT_BM - 1) / FT_BM);
  hipLaunchKernelGGL(ft_gemm_f64_kernel, grid, dim3(256), 0, s, M, N, K, P, X, C);
}
