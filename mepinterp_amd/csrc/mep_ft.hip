// mep_ft.hip -- Fourier sums H(k), dH(k), A(k), Omega(k) as one batched
// real GEMM on gfx950 FP64 MFMA.
//
// Replaces the per-k R-loop of FT_R_to_k (wann_interp.f90:151-384).  The
// k-independent factors i*c_a(R) (velocity) and i*(c_a r_b - c_b r_a)
// (curvature) are folded into the columns of X(R) once, at upload; the
// tight-binding phase e^{2 pi i k.(tau_m - tau_n)} is a per-element epilogue
// applied by the consumers (mep_eigh.hip, mep_velo.hip).
//
// R and -R are paired at upload:  e^{i phi} X_R + e^{-i phi} X_-R
//   = cos(phi) (X_R + X_-R) + sin(phi) i (X_R - X_-R),
// so the phase tile is real (cos rows, sin rows, and a row of ones for
// R = 0) and the sum is a real GEMM against the interleaved (re, im) columns
// of the combined coefficients: half the flops of the complex product, and
// the output is already the interleaved complex H(k).  Without a position
// operator only the upper triangle of each (Hermitian) matrix is produced.
//
// Layouts (doubles, row-major over the GEMM):
//   Pt [K][M]       phase tile, transposed (k-points contiguous)
//   X  [K][N]       N = 2 * nmat * np: (re, im) of matrix mat, element p
//                   (p = i + j(j+1)/2 for i <= j packed, i + n j full)
//   C  [M][N]       per k the nmat matrices, complex interleaved
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

// Pt[e][k] = cos(2 pi k.R_e), Pt[Kh + e][k] = sin(2 pi k.R_e) for the
// paired entries e < Kh (wann_interp.f90:258-260 angle); Pt[2 Kh][k] = 1
// when the model has R = 0.
__global__ void phase_real_kernel(int64_t nk, int Kh, int has_zero, const double* __restrict__ kpts,
                                  const double* __restrict__ Rh, double* __restrict__ Pt) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nk * (int64_t)(Kh + has_zero)) return;
  const int e = (int)(t / nk);
  const int64_t k = t - (int64_t)e * nk;
  if (e == Kh) {
    Pt[(int64_t)(2 * Kh) * nk + k] = 1.0;
    return;
  }
  const double two_pi = 2.0 * 3.14159265358979323846;
  const double ang = two_pi * (kpts[3 * k] * Rh[3 * e] + kpts[3 * k + 1] * Rh[3 * e + 1] +
                               kpts[3 * k + 2] * Rh[3 * e + 2]);
  double sn, cs;
  sincos(ang, &sn, &cs);
  Pt[(int64_t)e * nk + k] = cs;
  Pt[(int64_t)(Kh + e) * nk + k] = sn;
}

// the same tile rounded to FP32 (angles and sin/cos in FP64)
__global__ void phase_real_f32_kernel(int64_t nk, int Kh, int has_zero,
                                      const double* __restrict__ kpts,
                                      const double* __restrict__ Rh, float* __restrict__ Pt) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nk * (int64_t)(Kh + has_zero)) return;
  const int e = (int)(t / nk);
  const int64_t k = t - (int64_t)e * nk;
  if (e == Kh) {
    Pt[(int64_t)(2 * Kh) * nk + k] = 1.0f;
    return;
  }
  const double two_pi = 2.0 * 3.14159265358979323846;
  const double ang = two_pi * (kpts[3 * k] * Rh[3 * e] + kpts[3 * k + 1] * Rh[3 * e + 1] +
                               kpts[3 * k + 2] * Rh[3 * e + 2]);
  double sn, cs;
  sincos(ang, &sn, &cs);
  Pt[(int64_t)e * nk + k] = (float)cs;
  Pt[(int64_t)(Kh + e) * nk + k] = (float)sn;
}

// --------------------------------------------------------------------------
// FP64 real GEMM C = Pt^T X on v_mfma_f64_16x16x4_f64.
// Block: 256 threads = 4 waves in a 2x2 grid, block tile 128 (k-points) x
// 128 (columns), K staged 16 at a time through double-buffered LDS with the
// next chunk's global loads in flight during the MFMAs.  Each wave owns
// 64 x 64 = 4 x 4 MFMA tiles.  Operand fragments (16x4 A, 4x16 B): lane l
// holds row / col l%16 at k l/16; accumulator element v of lane l is row
// l/16 + 4v, col l%16 (measured on gfx950).
// Tiles are mapped XCD-aware: workgroups go round-robin over the 8 XCDs,
// so each XCD receives a contiguous range of tiles, i.e. all column tiles
// of a few k-row tiles, and the phase rows they share stay in its L2.
// --------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));

#define DG_BM 128
#define DG_BN 128
#define DG_KC 16
#define DG_LD (DG_BM + 16)  // row pitch: 4 k-rows of a fragment read hit distinct banks

__global__ __launch_bounds__(256, 2) void ft_gemm_f64_kernel(int M, int N, int K,
                                                             const double* __restrict__ Pt,
                                                             const double* __restrict__ X,
                                                             double* __restrict__ C, int tiles_n,
                                                             int tiles_total) {
  __shared__ double As[2][DG_KC][DG_LD], Bs[2][DG_KC][DG_LD];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;
  // XCD-aware tile index
  const int b = blockIdx.x;
  const int xcd = b & 7, loc = b >> 3;
  const int base = tiles_total >> 3, rem = tiles_total & 7;
  const int L = xcd * base + (xcd < rem ? xcd : rem) + loc;
  const int64_t m0 = (int64_t)(L / tiles_n) * DG_BM;
  const int64_t n0 = (int64_t)(L % tiles_n) * DG_BN;

  v4d acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  // staging map: thread -> (kk = tid/128 + 2 it, col = tid%128), 8 per operand
  const int sc = tid & 127, sk = tid >> 7;
  double ra[8], rb[8];
  auto gload = [&](int k0) {
#pragma unroll
    for (int it = 0; it < 8; ++it) {
      const int kk = k0 + sk + 2 * it;
      ra[it] = (kk < K && m0 + sc < M) ? Pt[(int64_t)kk * M + m0 + sc] : 0.0;
      rb[it] = (kk < K && n0 + sc < N) ? X[(int64_t)kk * N + n0 + sc] : 0.0;
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int it = 0; it < 8; ++it) {
      As[buf][sk + 2 * it][sc] = ra[it];
      Bs[buf][sk + 2 * it][sc] = rb[it];
    }
  };
  const int nch = (K + DG_KC - 1) / DG_KC;
  gload(0);
  lstore(0);
  __syncthreads();
  for (int ch = 0; ch < nch; ++ch) {
    const int buf = ch & 1;
    if (ch + 1 < nch) gload((ch + 1) * DG_KC);
#pragma unroll
    for (int kq = 0; kq < DG_KC / 4; ++kq) {
      const int kk = kq * 4 + lk;
      double a[4], bb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[i] = As[buf][kk][64 * wr + 16 * i + lr];
        bb[i] = Bs[buf][kk][64 * wc + 16 * i + lr];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], bb[j], acc[i][j], 0, 0, 0);
    }
    if (ch + 1 < nch) lstore(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t row = m0 + 64 * wr + 16 * i + lk + 4 * v;
        const int64_t col = n0 + 64 * wc + 16 * j + lr;
        if (row < M && col < N) C[row * (int64_t)N + col] = acc[i][j][v];
      }
}

// --------------------------------------------------------------------------
// FP32 real GEMM C = Pt^T X on v_mfma_f32_16x16x4_f32 (FP32 accumulation),
// same tiling and XCD mapping as the FP64 kernel.  Accumulator element v of
// lane l is row 4(l/16) + v, col l%16 (measured on gfx950).
// --------------------------------------------------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256, 2) void ft_gemm_f32_kernel(int M, int N, int K,
                                                             const float* __restrict__ Pt,
                                                             const float* __restrict__ X,
                                                             float* __restrict__ C, int tiles_n,
                                                             int tiles_total) {
  __shared__ float As[2][DG_KC][DG_LD], Bs[2][DG_KC][DG_LD];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;
  const int b = blockIdx.x;
  const int xcd = b & 7, loc = b >> 3;
  const int base = tiles_total >> 3, rem = tiles_total & 7;
  const int L = xcd * base + (xcd < rem ? xcd : rem) + loc;
  const int64_t m0 = (int64_t)(L / tiles_n) * DG_BM;
  const int64_t n0 = (int64_t)(L % tiles_n) * DG_BN;

  v4f acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.0f, 0.0f, 0.0f, 0.0f};

  const int sc = tid & 127, sk = tid >> 7;
  float ra[8], rb[8];
  auto gload = [&](int k0) {
#pragma unroll
    for (int it = 0; it < 8; ++it) {
      const int kk = k0 + sk + 2 * it;
      ra[it] = (kk < K && m0 + sc < M) ? Pt[(int64_t)kk * M + m0 + sc] : 0.0f;
      rb[it] = (kk < K && n0 + sc < N) ? X[(int64_t)kk * N + n0 + sc] : 0.0f;
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int it = 0; it < 8; ++it) {
      As[buf][sk + 2 * it][sc] = ra[it];
      Bs[buf][sk + 2 * it][sc] = rb[it];
    }
  };
  const int nch = (K + DG_KC - 1) / DG_KC;
  gload(0);
  lstore(0);
  __syncthreads();
  for (int ch = 0; ch < nch; ++ch) {
    const int buf = ch & 1;
    if (ch + 1 < nch) gload((ch + 1) * DG_KC);
#pragma unroll
    for (int kq = 0; kq < DG_KC / 4; ++kq) {
      const int kk = kq * 4 + lk;
      float a[4], bb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[i] = As[buf][kk][64 * wr + 16 * i + lr];
        bb[i] = Bs[buf][kk][64 * wc + 16 * i + lr];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], bb[j], acc[i][j], 0, 0, 0);
    }
    if (ch + 1 < nch) lstore(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t row = m0 + 64 * wr + 16 * i + 4 * lk + v;
        const int64_t col = n0 + 64 * wc + 16 * j + lr;
        if (row < M && col < N) C[row * (int64_t)N + col] = acc[i][j][v];
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

void launch_phase(hipStream_t s, int64_t nk, int Kh, int has_zero, const double* kpts,
                  const double* Rh, double* Pt) {
  int64_t tot = nk * (int64_t)(Kh + has_zero);
  int64_t nb = (tot + 255) / 256;
  hipLaunchKernelGGL(phase_real_kernel, dim3((unsigned)nb), dim3(256), 0, s, nk, Kh, has_zero, kpts,
                     Rh, Pt);
}

// --------------------------------------------------------------------------
// The FP32 Fourier GEMM C = Pt^T X in split bfloat16 ("BF16x3"): every FP32
// operand is x = hi + mid + lo with three round-to-nearest bf16 parts (8
// significand bits each, 24 together), and C accumulates in FP32 the six
// products whose size reaches 2^-24 of hi*hi (lo*hi, hi*lo, mid*mid,
// mid*hi, hi*mid, hi*hi, smallest first) on v_mfma_f32_16x16x16_bf16.  The
// dropped mid*lo, lo*mid, lo*lo lie below that, so the result matches the
// FP32 GEMM to its own rounding.  The split happens once per element, when
// the K-chunk is staged: LDS holds the three planes m-major with the chunk's
// 16 k contiguous (pitch 20 bf16 = 40 B: a 16-lane fragment read of 8 B per
// lane covers the 32 banks once).  Tiles, XCD order and double buffering
// as ft_gemm_f32_kernel.
// --------------------------------------------------------------------------
typedef short bf_v4 __attribute__((ext_vector_type(4)));
constexpr int B3_P = 20;  // bf16 per staged row: 16 k + 4 pad

__device__ inline unsigned short bf16_rn(float x) {
  const unsigned u = __float_as_uint(x);
  return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float bf16_f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__global__ __launch_bounds__(256, 2) void ft_gemm_bf16x3_kernel(int M, int N, int K,
                                                                const float* __restrict__ Pt,
                                                                const float* __restrict__ X,
                                                                float* __restrict__ C, int tiles_n,
                                                                int tiles_total) {
  // [buffer][plane hi / mid / lo][row m or n of the tile][k]
  __shared__ __attribute__((aligned(16))) unsigned short As[2][3][DG_BM * B3_P];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[2][3][DG_BN * B3_P];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;
  const int b = blockIdx.x;
  const int xcd = b & 7, loc = b >> 3;
  const int base = tiles_total >> 3, rem = tiles_total & 7;
  const int L = xcd * base + (xcd < rem ? xcd : rem) + loc;
  const int64_t m0 = (int64_t)(L / tiles_n) * DG_BM;
  const int64_t n0 = (int64_t)(L % tiles_n) * DG_BN;

  v4f acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.0f, 0.0f, 0.0f, 0.0f};

  // staging: thread (sc, sk) takes the eight k = 8 sk .. 8 sk + 7 of tile row sc
  const int sc = tid & 127, sk = tid >> 7;
  float ra[8], rb[8];
  auto gload = [&](int k0) {
#pragma unroll
    for (int it = 0; it < 8; ++it) {
      const int kk = k0 + 8 * sk + it;
      ra[it] = (kk < K && m0 + sc < M) ? Pt[(int64_t)kk * M + m0 + sc] : 0.0f;
      rb[it] = (kk < K && n0 + sc < N) ? X[(int64_t)kk * N + n0 + sc] : 0.0f;
    }
  };
  auto split_store = [&](const float* r, unsigned short* dst0, int buf_stride) {
    unsigned short h[3][8];
#pragma unroll
    for (int it = 0; it < 8; ++it) {
      const float x = r[it];
      const unsigned short a0 = bf16_rn(x);
      const float r1 = x - bf16_f(a0);
      const unsigned short a1 = bf16_rn(r1);
      const float r2 = r1 - bf16_f(a1);
      h[0][it] = a0;
      h[1][it] = a1;
      h[2][it] = bf16_rn(r2);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      // row sc, k 8 sk .. 8 sk + 7 at byte 40 sc + 16 sk: two 8-byte stores
      // (the offset is 8-byte aligned, not 16)
      uint2 v0, v1;
      v0.x = (unsigned)h[q][0] | ((unsigned)h[q][1] << 16);
      v0.y = (unsigned)h[q][2] | ((unsigned)h[q][3] << 16);
      v1.x = (unsigned)h[q][4] | ((unsigned)h[q][5] << 16);
      v1.y = (unsigned)h[q][6] | ((unsigned)h[q][7] << 16);
      unsigned short* d = dst0 + q * buf_stride + sc * B3_P + 8 * sk;
      *reinterpret_cast<uint2*>(d) = v0;
      *reinterpret_cast<uint2*>(d + 4) = v1;
    }
  };
  auto lstore = [&](int buf) {
    split_store(ra, &As[buf][0][0], DG_BM * B3_P);
    split_store(rb, &Bs[buf][0][0], DG_BN * B3_P);
  };
  const int nch = (K + DG_KC - 1) / DG_KC;
  gload(0);
  lstore(0);
  __syncthreads();
  for (int ch = 0; ch < nch; ++ch) {
    const int buf = ch & 1;
    if (ch + 1 < nch) gload((ch + 1) * DG_KC);
    // fragments: row 16 i + lr of the wave's 64, k = 4 lk .. 4 lk + 3
    bf_v4 a[3][4], bb[3][4];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[q][i] = *reinterpret_cast<const bf_v4*>(&As[buf][q][(64 * wr + 16 * i + lr) * B3_P + 4 * lk]);
        bb[q][i] = *reinterpret_cast<const bf_v4*>(&Bs[buf][q][(64 * wc + 16 * i + lr) * B3_P + 4 * lk]);
      }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        v4f c = acc[i][j];
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[2][i], bb[0][j], c, 0, 0, 0);  // lo hi
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0][i], bb[2][j], c, 0, 0, 0);  // hi lo
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[1][i], bb[1][j], c, 0, 0, 0);  // mid mid
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[1][i], bb[0][j], c, 0, 0, 0);  // mid hi
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0][i], bb[1][j], c, 0, 0, 0);  // hi mid
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0][i], bb[0][j], c, 0, 0, 0);  // hi hi
        acc[i][j] = c;
      }
    if (ch + 1 < nch) lstore(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t row = m0 + 64 * wr + 16 * i + 4 * lk + v;
        const int64_t col = n0 + 64 * wc + 16 * j + lr;
        if (row < M && col < N) C[row * (int64_t)N + col] = acc[i][j][v];
      }
}

void launch_ft_gemm_bf16x3(hipStream_t s, int M, int N, int K, const float* Pt, const float* X,
                           float* C) {
  const int tiles_m = (M + DG_BM - 1) / DG_BM, tiles_n = (N + DG_BN - 1) / DG_BN;
  const int tiles = tiles_m * tiles_n;
  hipLaunchKernelGGL(ft_gemm_bf16x3_kernel, dim3((unsigned)tiles), dim3(256), 0, s, M, N, K, Pt, X,
                     C, tiles_n, tiles);
}

void launch_ft_gemm_f64(hipStream_t s, int M, int N, int K, const double* Pt, const double* X,
                        double* C) {
  const int tiles_m = (M + DG_BM - 1) / DG_BM, tiles_n = (N + DG_BN - 1) / DG_BN;
  const int tiles = tiles_m * tiles_n;
  hipLaunchKernelGGL(ft_gemm_f64_kernel, dim3((unsigned)tiles), dim3(256), 0, s, M, N, K, Pt, X, C,
                     tiles_n, tiles);
}

void launch_phase_f32(hipStream_t s, int64_t nk, int Kh, int has_zero, const double* kpts,
                      const double* Rh, float* Pt) {
  int64_t tot = nk * (int64_t)(Kh + has_zero);
  int64_t nb = (tot + 255) / 256;
  hipLaunchKernelGGL(phase_real_f32_kernel, dim3((unsigned)nb), dim3(256), 0, s, nk, Kh, has_zero,
                     kpts, Rh, Pt);
}

void launch_ft_gemm_f32(hipStream_t s, int M, int N, int K, const float* Pt, const float* X,
                        float* C) {
  const int tiles_m = (M + DG_BM - 1) / DG_BM, tiles_n = (N + DG_BN - 1) / DG_BN;
  const int tiles = tiles_m * tiles_n;
  hipLaunchKernelGGL(ft_gemm_f32_kernel, dim3((unsigned)tiles), dim3(256), 0, s, M, N, K, Pt, X, C,
                     tiles_n, tiles);
}

// --------------------------------------------------------------------------
// Separable Fourier sums on a Monkhorst-Pack mesh.
//
// On the mesh the phase factorises, e^{2 pi i k.R} = ex(x, Rx) ey(y, Ry)
// ez(z, Rz), so with the coefficients X(R) scattered into a box
// Xb[ix][iy][iz][q] (R = origin + (ix, iy, iz); zero where no R)
//   Y1(z, ix, iy, q) = sum_iz Xb(ix, iy, iz, q) ez(z, iz)
//   Y2(y, z, ix, q)  = sum_iy Y1(z, ix, iy, q) ey(y, iy)
//   H(x, y, z, q)    = sum_ix Y2(y, z, ix, q) ex(x, ix)
// i.e. about bx + bx by / n1 complex products per k and column instead of
// NR: the stage is bound by writing H(k).  The first sum runs over the
// z-planes a k-block touches (sep_z_kernel); the other two are fused per
// mesh line (y, z) (sep_xy_kernel: Y2 in registers, one thread per column,
// the x-loop writes the line's k-points).  Phase tables, per mesh index:
// e*[index * BX + i] (ex padded with zeros to BX entries).
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void sep_z_kernel(int64_t NC, int bx, int by, int bz, int z0,
                                                    const zd* __restrict__ Xb,
                                                    const zd* __restrict__ ez,
                                                    zd* __restrict__ Y1) {
  const int64_t q = blockIdx.x * (int64_t)256 + threadIdx.x;
  const int c = blockIdx.y;  // (zi * bx + ix) * by + iy
  if (q >= NC) return;
  const int iy = c % by, ix = (c / by) % bx, zi = c / (bx * by);
  const zd* xs = Xb + (int64_t)(ix * by + iy) * bz * NC + q;
  const zd* ph = ez + (int64_t)(z0 + zi) * bz;
  double sr = 0.0, si = 0.0;
  for (int iz = 0; iz < bz; ++iz) {
    const zd x = xs[(int64_t)iz * NC], p = ph[iz];
    sr = fma(x.re, p.re, fma(-x.im, p.im, sr));
    si = fma(x.re, p.im, fma(x.im, p.re, si));
  }
  Y1[(int64_t)c * NC + q] = zd{sr, si};
}

// H(k) is written once and read by the next kernel long after the caches
// have turned over: non-temporal stores
__device__ inline void store_h(zd* p, double re, double im) {
  __builtin_nontemporal_store(re, &p->re);
  __builtin_nontemporal_store(im, &p->im);
}
__device__ inline void store_h(zf* p, double re, double im) {
  __builtin_nontemporal_store((float)re, &p->re);
  __builtin_nontemporal_store((float)im, &p->im);
}

template <int BX, class OT>
__global__ __launch_bounds__(256) void sep_xy_kernel(int64_t NC, int bx, int by, int n1, int n2,
                                                     int64_t k0, int64_t nk, int64_t line0,
                                                     int64_t nlines, int z0,
                                                     const zd* __restrict__ Y1,
                                                     const zd* __restrict__ ex,
                                                     const zd* __restrict__ ey,
                                                     OT* __restrict__ H) {
  const int64_t q = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (q >= NC) return;  // no barrier below
  for (int64_t li = blockIdx.y; li < nlines; li += gridDim.y) {
    const int64_t line = line0 + li;
    const int y = (int)(line % n2), zi = (int)(line / n2) - z0;
    // Y2(ix) = sum_iy Y1(zi, ix, iy, q) ey(y, iy)
    double yr[BX], yi[BX];
#pragma unroll
    for (int ix = 0; ix < BX; ++ix) {
      yr[ix] = 0.0;
      yi[ix] = 0.0;
      if (ix < bx) {
        const zd* src = Y1 + (int64_t)(zi * bx + ix) * by * NC + q;
        const zd* ph = ey + (int64_t)y * by;
        for (int iy = 0; iy < by; ++iy) {
          const zd v = src[(int64_t)iy * NC], p = ph[iy];
          yr[ix] = fma(v.re, p.re, fma(-v.im, p.im, yr[ix]));
          yi[ix] = fma(v.re, p.im, fma(v.im, p.re, yi[ix]));
        }
      }
    }
    // the line's k-points inside [k0, k0 + nk)
    const int64_t kb = line * n1;
    const int xa = (int)(k0 > kb ? k0 - kb : 0);
    const int xb = (int)(k0 + nk - kb < n1 ? k0 + nk - kb : n1);
    OT* out = H + (kb - k0) * NC + q;
    for (int x = xa; x < xb; ++x) {
      const zd* ph = ex + (int64_t)x * BX;
      double hr = 0.0, hi = 0.0;
#pragma unroll
      for (int ix = 0; ix < BX; ++ix) {
        const zd p = ph[ix];
        hr = fma(yr[ix], p.re, fma(-yi[ix], p.im, hr));
        hi = fma(yr[ix], p.im, fma(yi[ix], p.re, hi));
      }
      store_h(out + (int64_t)x * NC, hr, hi);
    }
  }
}

void launch_sep_ft(hipStream_t s, const SepFT& f, int64_t k0, int64_t nk, zd* H64, zf* H32) {
  const int64_t plane = (int64_t)f.n1 * f.n2;
  const int z0 = (int)(k0 / plane), z1 = (int)((k0 + nk - 1) / plane);
  const int nz = z1 - z0 + 1;
  const unsigned qb = (unsigned)((f.NC + 255) / 256);
  hipLaunchKernelGGL(sep_z_kernel, dim3(qb, (unsigned)(nz * f.bx * f.by)), dim3(256), 0, s, f.NC,
                     f.bx, f.by, f.bz, z0, f.Xb, f.ez, f.Y1);
  const int64_t line0 = k0 / f.n1, line1 = (k0 + nk - 1) / f.n1;
  const int64_t nlines = line1 - line0 + 1;
  const unsigned ly = (unsigned)(nlines < 65535 ? nlines : 65535);
#define MEP_SEP_XY(BX)                                                                          \
  if (H32)                                                                                      \
    hipLaunchKernelGGL((sep_xy_kernel<BX, zf>), dim3(qb, ly), dim3(256), 0, s, f.NC, f.bx, f.by, \
                       f.n1, f.n2, k0, nk, line0, nlines, z0, f.Y1, f.ex, f.ey, H32);           \
  else                                                                                          \
    hipLaunchKernelGGL((sep_xy_kernel<BX, zd>), dim3(qb, ly), dim3(256), 0, s, f.NC, f.bx, f.by, \
                       f.n1, f.n2, k0, nk, line0, nlines, z0, f.Y1, f.ex, f.ey, H64);
  if (f.bx <= 8) {
    MEP_SEP_XY(8)
  } else {
    MEP_SEP_XY(16)
  }
#undef MEP_SEP_XY
}
