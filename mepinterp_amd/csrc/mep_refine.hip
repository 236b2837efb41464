// mep_refine.hip -- FP64 refinement of the FP32 eigenpairs of the
// mixed-precision path.
//
// The FP32 eigensolver (mep_eigh.hip, R = float) leaves eigenvalues with an
// absolute error of order n eps32 |H| and eigenvectors rotated by angles of
// order n eps32 |H| / gap (large inside near-degenerate groups).  Per
// k-point, with H(k) (the FP32 Fourier sum, tight-binding phase applied) and
// U staged in FP64, on FP64 MFMA:
//   1. one step of the Ogita-Aishima refinement of the eigenvector matrix
//      (RefSyEv, 2018): G = U^H U, S = U^H H U, lambda = diag(S)/diag(G),
//      U <- U (I + F) with F_ij = (S_ij - lambda_j G_ij)/(lambda_j - lambda_i)
//      for pairs further apart than dE_thr and -G_ij/2 otherwise;
//   2. Rayleigh quotients E_n = Re(u_n^H H u_n)/|u_n|^2;
//   3. groups of consecutive bands with E_{n+1} - E_n < dE_thr: Rayleigh-Ritz
//      in FP64 (M = U_g^H H U_g, cyclic Jacobi, U_g <- U_g Y);
//   4. normalisation and the eigensolver's gauge convention.
// U is written in FP64 (the velocity rotation reads it), E in FP64.
#include "mep_common.hpp"
#include "mep_kernels.hpp"
#include "mep_mfma.hpp"

// Hermitian Jacobi on the g x g matrix M (column-major, leading dimension
// ld) accumulating the unitary Y (g x g, ld); one thread.  Rotation of the
// classic cyclic method after the phase that makes M_pq real.
__device__ static void herm_jacobi(int g, zd* M, zd* Y, int ld) {
  for (int i = 0; i < g; ++i)
    for (int j = 0; j < g; ++j) Y[i + ld * j] = zd{i == j ? 1.0 : 0.0, 0.0};
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int q = 0; q < g; ++q) {
      diag += M[q + ld * q].re * M[q + ld * q].re;
      for (int p = 0; p < q; ++p) off += zabs2(M[p + ld * q]);
    }
    if (off <= 1e-32 * diag || off == 0.0) break;
    for (int q = 1; q < g; ++q)
      for (int p = 0; p < q; ++p) {
        const zd b = M[p + ld * q];
        const double ab = sqrt(zabs2(b));
        if (ab == 0.0) continue;
        const zd ph = zd{b.re / ab, b.im / ab};  // e^{i phi}
        const double a = M[p + ld * p].re, d = M[q + ld * q].re;
        const double th = (d - a) / (2.0 * ab);
        const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        // G = D R with D = diag(1 at p, e^{-i phi} at q), R = [[c, s], [-s, c]]
        // columns: x_p' = c x_p - s e^{-i phi} x_q,  x_q' = s x_p + c e^{-i phi} x_q
        const zd eq = zconj(ph);
        for (int i = 0; i < g; ++i) {  // M <- M G
          const zd xp = M[i + ld * p], xq = zmul(M[i + ld * q], eq);
          M[i + ld * p] = zsub(zscale(xp, c), zscale(xq, s));
          M[i + ld * q] = zadd(zscale(xp, s), zscale(xq, c));
        }
        for (int j = 0; j < g; ++j) {  // M <- G^H M
          const zd xp = M[p + ld * j], xq = zmul(M[q + ld * j], ph);
          M[p + ld * j] = zsub(zscale(xp, c), zscale(xq, s));
          M[q + ld * j] = zadd(zscale(xp, s), zscale(xq, c));
        }
        M[p + ld * q] = zd{0.0, 0.0};
        M[q + ld * p] = zd{0.0, 0.0};
        M[p + ld * p].im = 0.0;
        M[q + ld * q].im = 0.0;
        for (int i = 0; i < g; ++i) {  // Y <- Y G
          const zd xp = Y[i + ld * p], xq = zmul(Y[i + ld * q], eq);
          Y[i + ld * p] = zsub(zscale(xp, c), zscale(xq, s));
          Y[i + ld * q] = zadd(zscale(xp, s), zscale(xq, c));
        }
      }
  }
}

template <int NMAX>
__global__ __launch_bounds__(256) void refine_kernel(int n, int64_t nk, const zf* __restrict__ Hk,
                                                     int64_t strideHk, int packed,
                                                     const double* __restrict__ kpts,
                                                     const double* __restrict__ tau,
                                                     double dE_thr, const zf* __restrict__ U32,
                                                     double* __restrict__ E, zd* __restrict__ U) {
  constexpr int LD = NMAX + 1;
  __shared__ zd Hs[NMAX * LD], Us[NMAX * LD], Ts[NMAX * LD], Ys[NMAX * LD];
  __shared__ double Er[NMAX];
  __shared__ int gstart[NMAX + 1];
  __shared__ int ngrp;
  const int t = threadIdx.x;
  const int n2 = n * n;
  const double two_pi = 2.0 * 3.14159265358979323846;
  // rows and columns >= n of every matrix stay zero (the products read them)
  for (int c = t; c < NMAX * LD; c += 256) Hs[c] = Us[c] = Ts[c] = Ys[c] = zd{0.0, 0.0};
  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    const zf* Hb = Hk + k * strideHk;
    // H(k) in FP64 (Hermitian, tight-binding phase) into Hs
    auto load_h = [&]() {
      for (int idx = t; idx < n2; idx += 256) {
        const int i = idx % n, j = idx / n;
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        const int64_t o = packed ? lo + ((int64_t)hi * (hi + 1)) / 2 : lo + (int64_t)n * hi;
        const zf hf = Hb[o];
        zd h = zd{(double)hf.re, (double)hf.im};
        if (tau && lo != hi) {
          const double ang = two_pi * (kpts[3 * k] * (tau[3 * hi] - tau[3 * lo]) +
                                       kpts[3 * k + 1] * (tau[3 * hi + 1] - tau[3 * lo + 1]) +
                                       kpts[3 * k + 2] * (tau[3 * hi + 2] - tau[3 * lo + 2]));
          double sn, cs;
          sincos(ang, &sn, &cs);
          h = zmul(zd{cs, sn}, h);
        }
        if (lo == hi) h.im = 0.0;
        else if (i > j) h = zconj(h);
        Hs[i + LD * j] = h;
      }
    };
    __syncthreads();
    load_h();
    for (int idx = t; idx < n2; idx += 256) {
      const zf uf = U32[k * n2 + idx];
      Us[idx % n + LD * (idx / n)] = zd{(double)uf.re, (double)uf.im};
    }
    __syncthreads();
    // ---- one step of eigenvector refinement (X = U):
    //   G = X^H X, S = X^H H X, lambda_i = S_ii / G_ii,
    //   F_ij = (S_ij - lambda_j G_ij) / (lambda_j - lambda_i) for separated
    //   pairs, F_ij = -G_ij / 2 inside a group and on the diagonal
    //   (F_ii = (1 - G_ii) / 2);  X <- X (I + F)
    zgemm_tiles<NMAX, false>(n, Hs, LD, BPlain{Us, LD}, Ts, LD);  // T = H X
    zgemm_tiles<NMAX, true>(n, Us, LD, BPlain{Us, LD}, Ys, LD);   // G = X^H X
    __syncthreads();
    zgemm_tiles<NMAX, true>(n, Us, LD, BPlain{Ts, LD}, Hs, LD);   // S = X^H T (H is reloaded later)
    __syncthreads();
    if (t < n) Er[t] = Hs[t + LD * t].re / Ys[t + LD * t].re;
    __syncthreads();
    for (int idx = t; idx < n2; idx += 256) {
      const int i = idx % n, j = idx / n;
      const zd g = Ys[i + LD * j];
      zd f;
      if (i == j) {
        f = zd{0.5 * (1.0 - g.re), 0.0};
      } else {
        const double dl = Er[j] - Er[i];
        if (fabs(dl) > dE_thr) {
          const zd sij = Hs[i + LD * j];
          const double inv = 1.0 / dl;
          f = zd{(sij.re - Er[j] * g.re) * inv, (sij.im - Er[j] * g.im) * inv};
        } else {
          f = zd{-0.5 * g.re, -0.5 * g.im};
        }
      }
      Ts[i + LD * j] = f;
    }
    __syncthreads();
    zgemm_tiles<NMAX, false>(n, Us, LD, BPlain{Ts, LD}, Ys, LD);  // X F
    __syncthreads();
    for (int idx = t; idx < n2; idx += 256) {
      const int o = idx % n + LD * (idx / n);
      Us[o] = zadd(Us[o], Ys[o]);
    }
    load_h();
    __syncthreads();
    // ---- Rayleigh quotients of the refined vectors -------------------------
    zgemm_tiles<NMAX, false>(n, Hs, LD, BPlain{Us, LD}, Ts, LD);  // T = H X
    __syncthreads();
    if (t < n) {
      double num = 0.0, den = 0.0;
      for (int i = 0; i < n; ++i) {
        const zd u = Us[i + LD * t], tv = Ts[i + LD * t];
        num += u.re * tv.re + u.im * tv.im;
        den += zabs2(u);
      }
      Er[t] = num / den;
    }
    __syncthreads();
    // ---- near-degenerate groups: Rayleigh-Ritz in FP64 ----------------------
    if (t == 0) {
      int g = 0;
      gstart[0] = 0;
      for (int j = 1; j < n; ++j)
        if (!(Er[j] - Er[j - 1] < dE_thr)) gstart[++g] = j;
      gstart[++g] = n;
      ngrp = g;
    }
    __syncthreads();
    for (int gi = t; gi < ngrp; gi += 256) {
      const int a = gstart[gi], b = gstart[gi + 1], g = b - a;
      if (g < 2) continue;
      // M = U_g^H T_g into Hs columns a..b-1 (H is no longer needed), rows 0..g-1
      zd* M = Hs + LD * a;
      zd* Y = Ys + LD * a;
      for (int q = 0; q < g; ++q)
        for (int p = 0; p < g; ++p) {
          zd s = zd{0.0, 0.0};
          for (int i = 0; i < n; ++i) s = zadd(s, zcmul(Us[i + LD * (a + p)], Ts[i + LD * (a + q)]));
          M[p + LD * q] = s;
        }
      for (int q = 0; q < g; ++q) {  // exact Hermitian part
        for (int p = 0; p < q; ++p) {
          const zd x = M[p + LD * q], y = zconj(M[q + LD * p]);
          const zd h = zd{0.5 * (x.re + y.re), 0.5 * (x.im + y.im)};
          M[p + LD * q] = h;
          M[q + LD * p] = zconj(h);
        }
        M[q + LD * q].im = 0.0;
      }
      herm_jacobi(g, M, Y, LD);
      // ascending order of the Ritz values (selection sort on the columns)
      for (int q = 0; q < g; ++q) {
        int best = q;
        for (int r = q + 1; r < g; ++r)
          if (M[r + LD * r].re < M[best + LD * best].re) best = r;
        if (best != q) {
          const zd tmp = M[q + LD * q];
          M[q + LD * q] = M[best + LD * best];
          M[best + LD * best] = tmp;
          for (int i = 0; i < g; ++i) {
            const zd y = Y[i + LD * q];
            Y[i + LD * q] = Y[i + LD * best];
            Y[i + LD * best] = y;
          }
        }
        Er[a + q] = M[q + LD * q].re;
      }
      // U_g <- U_g Y through T's columns a..b-1 (T is no longer needed there)
      for (int q = 0; q < g; ++q)
        for (int i = 0; i < n; ++i) {
          zd s = zd{0.0, 0.0};
          for (int p = 0; p < g; ++p) s = zadd(s, zmul(Us[i + LD * (a + p)], Y[p + LD * q]));
          Ts[i + LD * (a + q)] = s;
        }
      for (int q = 0; q < g; ++q)
        for (int i = 0; i < n; ++i) Us[i + LD * (a + q)] = Ts[i + LD * (a + q)];
    }
    __syncthreads();
    // ---- normalise; gauge: largest-modulus component real and positive -----
    if (t < n) {
      double nrm = 0.0, best = -1.0;
      zd u = zd{1.0, 0.0};
      for (int i = 0; i < n; ++i) {
        const zd x = Us[i + LD * t];
        const double m2 = zabs2(x);
        nrm += m2;
        if (m2 > best) {
          best = m2;
          u = x;
        }
      }
      const double au = sqrt(zabs2(u));
      const double sc = 1.0 / sqrt(nrm);
      const zd phs = au > 0.0 ? zd{u.re / au * sc, -u.im / au * sc} : zd{sc, 0.0};
      for (int i = 0; i < n; ++i) Us[i + LD * t] = zmul(Us[i + LD * t], phs);
    }
    __syncthreads();
    for (int idx = t; idx < n2; idx += 256) U[k * n2 + idx] = Us[idx % n + LD * (idx / n)];
    if (t < n) E[k * n + t] = Er[t];
  }
}

void launch_refine(hipStream_t s, int n, int64_t nk, const zf* Hk, int64_t strideHk, int packed,
                   const KGeom& g, double dE_thr, const zf* U32, double* E, zd* U) {
  unsigned grid = (unsigned)(nk < 65536 ? nk : 65536);
  if (n <= 16)
    hipLaunchKernelGGL(refine_kernel<16>, dim3(grid), dim3(256), 0, s, n, nk, Hk, strideHk, packed,
                       g.kpts, g.tau, dE_thr, U32, E, U);
  else
    hipLaunchKernelGGL(refine_kernel<32>, dim3(grid), dim3(256), 0, s, n, nk, Hk, strideHk, packed,
                       g.kpts, g.tau, dE_thr, U32, E, U);
}
