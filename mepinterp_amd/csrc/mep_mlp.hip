// mep_mlp.hip -- the learned part of adaptive k refinement.
//
// A small MLP predicts log|Omega(k)| from sin/cos Fourier features of k:
//   features  x = [sin(2 pi m k_a), cos(2 pi m k_a)], m = 1..n_freq, a = x,y,z
//             (6 n_freq values, zero-padded to FP, a multiple of 16)
//   layers    h1 = tanh(x W1 + b1), h2 = tanh(h1 W2 + b2), h3 = tanh(h2 W3 + b3),
//             y = h3 W4 + b4 (W4 padded to 16 output columns; column 0 is y)
// FP32 weights; training (mean squared error, Adam) and inference run in the
// kernels below: an FP32 MFMA GEMM for every product (one wavefront per
// 16 x 16 tile, either operand transposed), bias/activation, their
// derivatives, column sums for the bias gradients, and the Adam update.
// Parameter layout (floats): W1[FP][H] b1[H] W2[H][H] b2[H] W3[H][H] b3[H]
// W4[H][16] b4[16], matrices row-major (input index slow).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/mepinterp.h"

typedef float mlp_v4f __attribute__((ext_vector_type(4)));

extern void mep_set_error(const std::string& msg);  // mep_capi.hip

namespace {

constexpr int OUTP = 16;  // padded output width

// C[M][N] (row-major, ldc) = op(A)[M][K] op(B)[K][N]; op(A)(r, c) =
// transA ? A[c*lda + r] : A[r*lda + c], likewise B.  M, N multiples of 16,
// K a multiple of 4.  One wavefront per 16 x 16 tile, four per block.
__global__ __launch_bounds__(256) void mlp_gemm_kernel(int M, int N, int K, const float* __restrict__ A,
                                                       int lda, int transA, const float* __restrict__ B,
                                                       int ldb, int transB, float* __restrict__ C,
                                                       int ldc) {
  const int lane = threadIdx.x & 63;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int tn = N / 16;
  if (tile >= (M / 16) * tn) return;
  const int r0 = 16 * (tile / tn), c0 = 16 * (tile % tn);
  const int lr = lane & 15, lk = lane >> 4;
  mlp_v4f acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 4) {
    const int kk = k0 + lk;
    const float a = transA ? A[(int64_t)kk * lda + r0 + lr] : A[(int64_t)(r0 + lr) * lda + kk];
    const float b = transB ? B[(int64_t)(c0 + lr) * ldb + kk] : B[(int64_t)kk * ldb + c0 + lr];
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) C[(int64_t)(r0 + 4 * lk + v) * ldc + c0 + lr] = acc[v];
}

// features of rows r < rows (k from kpts[idx[r]] or kpts[r]); zero rows after
__global__ void mlp_feat_kernel(int B, int rows, const double* __restrict__ kpts,
                                const int64_t* __restrict__ idx, int nfreq, int FP,
                                float* __restrict__ X) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B) return;
  float* xr = X + (int64_t)r * FP;
  for (int c = 0; c < FP; ++c) xr[c] = 0.f;
  if (r >= rows) return;
  const int64_t s = idx ? idx[r] : r;
  const double two_pi = 6.283185307179586476925;
  for (int a = 0; a < 3; ++a) {
    const double ka = kpts[3 * s + a];
    for (int m = 0; m < nfreq; ++m) {
      double sn, cs;
      sincos(two_pi * (m + 1) * ka, &sn, &cs);
      xr[a * 2 * nfreq + 2 * m] = (float)sn;
      xr[a * 2 * nfreq + 2 * m + 1] = (float)cs;
    }
  }
}

// H = act(Z + b) over B x N; act: 1 = tanh, 0 = identity
__global__ void mlp_bias_act_kernel(int64_t total, int N, const float* __restrict__ Z,
                                    const float* __restrict__ b, float* __restrict__ H, int act) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const float z = Z[i] + b[i % N];
  H[i] = act ? tanhf(z) : z;
}

// dZ = dH * (1 - H^2)
__global__ void mlp_act_back_kernel(int64_t total, const float* __restrict__ dH,
                                    const float* __restrict__ H, float* __restrict__ dZ) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const float h = H[i];
  dZ[i] = dH[i] * (1.f - h * h);
}

// db[c] = sum_r dZ[r][c], rows in order
__global__ void mlp_colsum_kernel(int B, int N, const float* __restrict__ dZ, float* __restrict__ db) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float s = 0.f;
  for (int r = 0; r < B; ++r) s += dZ[(int64_t)r * N + c];
  db[c] = s;
}

// mean squared error over the first `rows` rows of column 0 and its
// gradient dY (zero in padded rows and columns); one block, fixed order
__global__ __launch_bounds__(256) void mlp_loss_kernel(int B, int rows, const float* __restrict__ Y,
                                                       const float* __restrict__ t,
                                                       const int64_t* __restrict__ idx,
                                                       float* __restrict__ dY,
                                                       double* __restrict__ loss_out) {
  __shared__ double red[256];
  double s = 0.0;
  const float inv = 1.f / (float)rows;
  for (int r = threadIdx.x; r < B; r += 256) {
    float* d = dY + (int64_t)r * OUTP;
    for (int c = 0; c < OUTP; ++c) d[c] = 0.f;
    if (r < rows) {
      const float e = Y[(int64_t)r * OUTP] - t[idx ? idx[r] : r];
      d[0] = 2.f * e * inv;
      s += (double)e * (double)e;
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss_out = red[0] / rows;
}

__global__ void mlp_adam_kernel(int64_t P, float* __restrict__ p, const float* __restrict__ g,
                                float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                float b2, float eps, float c1, float c2) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= P) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

struct mep_mlp {
  int nfreq = 0, F = 0, FP = 0, H = 0;
  int64_t P = 0, oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0, oW4 = 0, ob4 = 0;
  float *dp = nullptr, *dg = nullptr, *dm = nullptr, *dv = nullptr;
  int64_t bcap = 0;
  float *dX = nullptr, *dZ = nullptr, *dH1 = nullptr, *dH2 = nullptr, *dH3 = nullptr;
  float *dY = nullptr, *ddY = nullptr, *ddH = nullptr, *ddZ = nullptr;
  double* dloss = nullptr;
  int64_t lcap = 0;
  int64_t dcap = 0;  // training / prediction data capacity (k-points)
  double* dk = nullptr;
  float* dt = nullptr;
  int64_t* didx = nullptr;
  float* dout = nullptr;
  hipStream_t st = nullptr;
  int64_t step = 0;
};

namespace {

#define MLP_HIP(x)                                                          \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                 \
      mep_set_error(std::string(#x) + ": " + hipGetErrorString(e_));        \
      return MEP_ERR_HIP;                                                   \
    }                                                                       \
  } while (0)

void gemm(mep_mlp* q, int M, int N, int K, const float* A, int lda, int tA, const float* B, int ldb,
          int tB, float* C, int ldc) {
  const int tiles = (M / 16) * (N / 16);
  hipLaunchKernelGGL(mlp_gemm_kernel, dim3(nblk(tiles, 4)), dim3(256), 0, q->st, M, N, K, A, lda, tA,
                     B, ldb, tB, C, ldc);
}

int ensure_batch(mep_mlp* q, int64_t B) {
  if (B <= q->bcap) return MEP_OK;
  hipFree(q->dX); hipFree(q->dZ); hipFree(q->dH1); hipFree(q->dH2); hipFree(q->dH3);
  hipFree(q->dY); hipFree(q->ddY); hipFree(q->ddH); hipFree(q->ddZ);
  const int64_t H = q->H;
  MLP_HIP(hipMalloc(&q->dX, sizeof(float) * B * q->FP));
  MLP_HIP(hipMalloc(&q->dZ, sizeof(float) * B * H));
  MLP_HIP(hipMalloc(&q->dH1, sizeof(float) * B * H));
  MLP_HIP(hipMalloc(&q->dH2, sizeof(float) * B * H));
  MLP_HIP(hipMalloc(&q->dH3, sizeof(float) * B * H));
  MLP_HIP(hipMalloc(&q->dY, sizeof(float) * B * OUTP));
  MLP_HIP(hipMalloc(&q->ddY, sizeof(float) * B * OUTP));
  MLP_HIP(hipMalloc(&q->ddH, sizeof(float) * B * H));
  MLP_HIP(hipMalloc(&q->ddZ, sizeof(float) * B * H));
  q->bcap = B;
  return MEP_OK;
}

int ensure_data(mep_mlp* q, int64_t n) {
  if (n <= q->dcap) return MEP_OK;
  hipFree(q->dk); hipFree(q->dt); hipFree(q->didx); hipFree(q->dout);
  MLP_HIP(hipMalloc(&q->dk, sizeof(double) * 3 * n));
  MLP_HIP(hipMalloc(&q->dt, sizeof(float) * n));
  MLP_HIP(hipMalloc(&q->didx, sizeof(int64_t) * n));
  MLP_HIP(hipMalloc(&q->dout, sizeof(float) * n));
  q->dcap = n;
  return MEP_OK;
}

// forward pass of B rows whose features are in dX; output in dY
void forward(mep_mlp* q, int B) {
  const int H = q->H;
  const int64_t tot = (int64_t)B * H;
  gemm(q, B, H, q->FP, q->dX, q->FP, 0, q->dp + q->oW1, H, 0, q->dZ, H);
  hipLaunchKernelGGL(mlp_bias_act_kernel, dim3(nblk(tot, 256)), dim3(256), 0, q->st, tot, H, q->dZ,
                     q->dp + q->ob1, q->dH1, 1);
  gemm(q, B, H, H, q->dH1, H, 0, q->dp + q->oW2, H, 0, q->dZ, H);
  hipLaunchKernelGGL(mlp_bias_act_kernel, dim3(nblk(tot, 256)), dim3(256), 0, q->st, tot, H, q->dZ,
                     q->dp + q->ob2, q->dH2, 1);
  gemm(q, B, H, H, q->dH2, H, 0, q->dp + q->oW3, H, 0, q->dZ, H);
  hipLaunchKernelGGL(mlp_bias_act_kernel, dim3(nblk(tot, 256)), dim3(256), 0, q->st, tot, H, q->dZ,
                     q->dp + q->ob3, q->dH3, 1);
  const int64_t to = (int64_t)B * OUTP;
  gemm(q, B, OUTP, H, q->dH3, H, 0, q->dp + q->oW4, OUTP, 0, q->dZ, OUTP);
  hipLaunchKernelGGL(mlp_bias_act_kernel, dim3(nblk(to, 256)), dim3(256), 0, q->st, to, OUTP, q->dZ,
                     q->dp + q->ob4, q->dY, 0);
}

// backward pass from ddY into dg (every parameter's gradient is written)
void backward(mep_mlp* q, int B) {
  const int H = q->H;
  const int64_t tot = (int64_t)B * H;
  float* g = q->dg;
  gemm(q, H, OUTP, B, q->dH3, H, 1, q->ddY, OUTP, 0, g + q->oW4, OUTP);
  hipLaunchKernelGGL(mlp_colsum_kernel, dim3(1), dim3(64), 0, q->st, B, OUTP, q->ddY, g + q->ob4);
  gemm(q, B, H, OUTP, q->ddY, OUTP, 0, q->dp + q->oW4, OUTP, 1, q->ddH, H);
  hipLaunchKernelGGL(mlp_act_back_kernel, dim3(nblk(tot, 256)), dim3(256), 0, q->st, tot, q->ddH,
                     q->dH3, q->ddZ);
  gemm(q, H, H, B, q->dH2, H, 1, q->ddZ, H, 0, g + q->oW3, H);
  hipLaunchKernelGGL(mlp_colsum_kernel, dim3(nblk(H, 64)), dim3(64), 0, q->st, B, H, q->ddZ, g + q->ob3);
  gemm(q, B, H, H, q->ddZ, H, 0, q->dp + q->oW3, H, 1, q->ddH, H);
  hipLaunchKernelGGL(mlp_act_back_kernel, dim3(nblk(tot, 256)), dim3(256), 0, q->st, tot, q->ddH,
                     q->dH2, q->ddZ);
  gemm(q, H, H, B, q->dH1, H, 1, q->ddZ, H, 0, g + q->oW2, H);
  hipLaunchKernelGGL(mlp_colsum_kernel, dim3(nblk(H, 64)), dim3(64), 0, q->st, B, H, q->ddZ, g + q->ob2);
  gemm(q, B, H, H, q->ddZ, H, 0, q->dp + q->oW2, H, 1, q->ddH, H);
  hipLaunchKernelGGL(mlp_act_back_kernel, dim3(nblk(tot, 256)), dim3(256), 0, q->st, tot, q->ddH,
                     q->dH1, q->ddZ);
  gemm(q, q->FP, H, B, q->dX, q->FP, 1, q->ddZ, H, 0, g + q->oW1, H);
  hipLaunchKernelGGL(mlp_colsum_kernel, dim3(nblk(H, 64)), dim3(64), 0, q->st, B, H, q->ddZ, g + q->ob1);
}

int64_t round_batch(int64_t b) { return ((std::max<int64_t>(b, 16) + 63) / 64) * 64; }

}  // namespace

extern "C" int mep_mlp_create(int32_t n_freq, int32_t hidden, uint64_t seed, mep_mlp** out) {
  if (!out || n_freq < 1 || n_freq > 64 || hidden < 16 || hidden > 1024 || hidden % 16) {
    mep_set_error("mep_mlp_create: n_freq in [1, 64], hidden a multiple of 16 in [16, 1024]");
    return MEP_ERR_ARG;
  }
  mep_mlp* q = new mep_mlp();
  q->nfreq = n_freq;
  q->F = 6 * n_freq;
  q->FP = ((q->F + 15) / 16) * 16;
  q->H = hidden;
  const int64_t H = hidden, FP = q->FP;
  int64_t o = 0;
  q->oW1 = o; o += FP * H;
  q->ob1 = o; o += H;
  q->oW2 = o; o += H * H;
  q->ob2 = o; o += H;
  q->oW3 = o; o += H * H;
  q->ob3 = o; o += H;
  q->oW4 = o; o += H * OUTP;
  q->ob4 = o; o += OUTP;
  q->P = o;
  // seeded uniform +-sqrt(6 / (fan_in + fan_out)) weights, zero biases;
  // padded feature rows and output columns stay zero
  std::vector<float> p((size_t)q->P, 0.f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int64_t off, int64_t fin, int64_t fout, int64_t cols, int64_t rows_used,
                  int64_t cols_used) {
    std::uniform_real_distribution<double> U(-std::sqrt(6.0 / (fin + fout)), std::sqrt(6.0 / (fin + fout)));
    for (int64_t r = 0; r < rows_used; ++r)
      for (int64_t c = 0; c < cols_used; ++c) p[(size_t)(off + r * cols + c)] = (float)U(rng);
  };
  fill(q->oW1, q->F, H, H, q->F, H);
  fill(q->oW2, H, H, H, H, H);
  fill(q->oW3, H, H, H, H, H);
  fill(q->oW4, H, 1, OUTP, H, 1);
  hipError_t e = hipStreamCreate(&q->st);
  if (e == hipSuccess) e = hipMalloc(&q->dp, sizeof(float) * q->P);
  if (e == hipSuccess) e = hipMalloc(&q->dg, sizeof(float) * q->P);
  if (e == hipSuccess) e = hipMalloc(&q->dm, sizeof(float) * q->P);
  if (e == hipSuccess) e = hipMalloc(&q->dv, sizeof(float) * q->P);
  if (e == hipSuccess) e = hipMemcpy(q->dp, p.data(), sizeof(float) * q->P, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(q->dm, 0, sizeof(float) * q->P);
  if (e == hipSuccess) e = hipMemset(q->dv, 0, sizeof(float) * q->P);
  if (e != hipSuccess) {
    mep_set_error(std::string("mep_mlp_create: ") + hipGetErrorString(e));
    mep_mlp_destroy(q);
    return MEP_ERR_HIP;
  }
  *out = q;
  return MEP_OK;
}

extern "C" int mep_mlp_destroy(mep_mlp* q) {
  if (!q) return MEP_OK;
  hipFree(q->dp); hipFree(q->dg); hipFree(q->dm); hipFree(q->dv);
  hipFree(q->dX); hipFree(q->dZ); hipFree(q->dH1); hipFree(q->dH2); hipFree(q->dH3);
  hipFree(q->dY); hipFree(q->ddY); hipFree(q->ddH); hipFree(q->ddZ); hipFree(q->dloss);
  hipFree(q->dk); hipFree(q->dt); hipFree(q->didx); hipFree(q->dout);
  if (q->st) hipStreamDestroy(q->st);
  delete q;
  return MEP_OK;
}

extern "C" int mep_mlp_num_params(const mep_mlp* q, int64_t* n) {
  if (!q || !n) {
    mep_set_error("mep_mlp_num_params: bad argument");
    return MEP_ERR_ARG;
  }
  *n = q->P;
  return MEP_OK;
}

extern "C" int mep_mlp_get_params(mep_mlp* q, float* params) {
  if (!q || !params) {
    mep_set_error("mep_mlp_get_params: bad argument");
    return MEP_ERR_ARG;
  }
  MLP_HIP(hipMemcpyAsync(params, q->dp, sizeof(float) * q->P, hipMemcpyDeviceToHost, q->st));
  MLP_HIP(hipStreamSynchronize(q->st));
  return MEP_OK;
}

extern "C" int mep_mlp_set_params(mep_mlp* q, const float* params) {
  if (!q || !params) {
    mep_set_error("mep_mlp_set_params: bad argument");
    return MEP_ERR_ARG;
  }
  MLP_HIP(hipMemcpyAsync(q->dp, params, sizeof(float) * q->P, hipMemcpyHostToDevice, q->st));
  MLP_HIP(hipStreamSynchronize(q->st));
  return MEP_OK;
}

extern "C" int mep_mlp_predict(mep_mlp* q, int64_t n, const double* kpts, float* out) {
  if (!q || n < 0 || (n > 0 && (!kpts || !out))) {
    mep_set_error("mep_mlp_predict: bad argument");
    return MEP_ERR_ARG;
  }
  if (n == 0) return MEP_OK;
  const int64_t B = 4096;
  if (int rc = ensure_batch(q, B)) return rc;
  if (int rc = ensure_data(q, n)) return rc;
  MLP_HIP(hipMemcpyAsync(q->dk, kpts, sizeof(double) * 3 * n, hipMemcpyHostToDevice, q->st));
  for (int64_t b0 = 0; b0 < n; b0 += B) {
    const int rows = (int)std::min<int64_t>(B, n - b0);
    hipLaunchKernelGGL(mlp_feat_kernel, dim3(nblk(B, 256)), dim3(256), 0, q->st, (int)B, rows,
                       q->dk + 3 * b0, (const int64_t*)nullptr, q->nfreq, q->FP, q->dX);
    forward(q, (int)B);
    // column 0 of Y: strided copy
    MLP_HIP(hipMemcpy2DAsync(q->dout + b0, sizeof(float), q->dY, sizeof(float) * OUTP, sizeof(float),
                             rows, hipMemcpyDeviceToDevice, q->st));
  }
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipMemcpyAsync(out, q->dout, sizeof(float) * n, hipMemcpyDeviceToHost, q->st));
  MLP_HIP(hipStreamSynchronize(q->st));
  return MEP_OK;
}

extern "C" int mep_mlp_grad(mep_mlp* q, int64_t n, const double* kpts, const float* target,
                            float* grad, float* loss) {
  if (!q || n < 1 || !kpts || !target || !grad) {
    mep_set_error("mep_mlp_grad: bad argument");
    return MEP_ERR_ARG;
  }
  const int64_t B = round_batch(n);
  if (int rc = ensure_batch(q, B)) return rc;
  if (int rc = ensure_data(q, n)) return rc;
  if (!q->dloss) MLP_HIP(hipMalloc(&q->dloss, sizeof(double) * 64));
  q->lcap = std::max<int64_t>(q->lcap, 64);
  MLP_HIP(hipMemcpyAsync(q->dk, kpts, sizeof(double) * 3 * n, hipMemcpyHostToDevice, q->st));
  MLP_HIP(hipMemcpyAsync(q->dt, target, sizeof(float) * n, hipMemcpyHostToDevice, q->st));
  hipLaunchKernelGGL(mlp_feat_kernel, dim3(nblk(B, 256)), dim3(256), 0, q->st, (int)B, (int)n, q->dk,
                     (const int64_t*)nullptr, q->nfreq, q->FP, q->dX);
  forward(q, (int)B);
  hipLaunchKernelGGL(mlp_loss_kernel, dim3(1), dim3(256), 0, q->st, (int)B, (int)n, q->dY, q->dt,
                     (const int64_t*)nullptr, q->ddY, q->dloss);
  backward(q, (int)B);
  MLP_HIP(hipGetLastError());
  double l = 0.0;
  MLP_HIP(hipMemcpyAsync(grad, q->dg, sizeof(float) * q->P, hipMemcpyDeviceToHost, q->st));
  MLP_HIP(hipMemcpyAsync(&l, q->dloss, sizeof(double), hipMemcpyDeviceToHost, q->st));
  MLP_HIP(hipStreamSynchronize(q->st));
  if (loss) *loss = (float)l;
  return MEP_OK;
}

extern "C" int mep_mlp_train(mep_mlp* q, int64_t n, const double* kpts, const float* target,
                             int32_t epochs, int32_t batch, float lr, uint64_t seed,
                             float* final_loss) {
  if (!q || n < 1 || !kpts || !target || epochs < 1 || batch < 1 || !(lr > 0.f)) {
    mep_set_error("mep_mlp_train: bad argument");
    return MEP_ERR_ARG;
  }
  const int64_t B = round_batch(std::min<int64_t>(batch, n));
  const int64_t nb = (n + B - 1) / B;
  if (int rc = ensure_batch(q, B)) return rc;
  if (int rc = ensure_data(q, n)) return rc;
  if (nb > q->lcap) {
    hipFree(q->dloss);
    q->dloss = nullptr;
    MLP_HIP(hipMalloc(&q->dloss, sizeof(double) * nb));
    q->lcap = nb;
  }
  MLP_HIP(hipMemcpyAsync(q->dk, kpts, sizeof(double) * 3 * n, hipMemcpyHostToDevice, q->st));
  MLP_HIP(hipMemcpyAsync(q->dt, target, sizeof(float) * n, hipMemcpyHostToDevice, q->st));
  std::vector<int64_t> perm((size_t)n);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs; ++ep) {
    std::iota(perm.begin(), perm.end(), (int64_t)0);
    std::mt19937_64 rng(seed + 0x9e3779b97f4a7c15ull * (uint64_t)(ep + 1));
    std::shuffle(perm.begin(), perm.end(), rng);
    // the upload must finish before the host reshuffles for the next epoch
    MLP_HIP(hipMemcpyAsync(q->didx, perm.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, q->st));
    for (int64_t bi = 0; bi < nb; ++bi) {
      const int rows = (int)std::min<int64_t>(B, n - bi * B);
      const int64_t* ix = q->didx + bi * B;
      hipLaunchKernelGGL(mlp_feat_kernel, dim3(nblk(B, 256)), dim3(256), 0, q->st, (int)B, rows, q->dk,
                         ix, q->nfreq, q->FP, q->dX);
      forward(q, (int)B);
      hipLaunchKernelGGL(mlp_loss_kernel, dim3(1), dim3(256), 0, q->st, (int)B, rows, q->dY, q->dt, ix,
                         q->ddY, q->dloss + bi);
      backward(q, (int)B);
      q->step += 1;
      const float c1 = 1.f - std::pow(b1, (float)q->step), c2 = 1.f - std::pow(b2, (float)q->step);
      hipLaunchKernelGGL(mlp_adam_kernel, dim3(nblk(q->P, 256)), dim3(256), 0, q->st, q->P, q->dp, q->dg,
                         q->dm, q->dv, lr, b1, b2, eps, c1, c2);
    }
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipStreamSynchronize(q->st));
  }
  std::vector<double> lb((size_t)nb);
  MLP_HIP(hipMemcpy(lb.data(), q->dloss, sizeof(double) * nb, hipMemcpyDeviceToHost));
  if (final_loss) {
    double s = 0.0, w = 0.0;
    for (int64_t bi = 0; bi < nb; ++bi) {
      const double rows = (double)std::min<int64_t>(B, n - bi * B);
      s += lb[(size_t)bi] * rows;
      w += rows;
    }
    *final_loss = (float)(s / w);
  }
  return MEP_OK;
}
