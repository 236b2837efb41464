// mep_eigh.hip -- batched Hermitian eigensolver, one k-point per wavefront.
//
// Replaces zheevd('V','U') of zheevd_wrapper (matrix_math.f90:192-226) inside
// get_wann_interp (wann_interp.f90:95).  Contract kept: only the upper
// triangle of H(k) is read, eigenvalues ascending, eigenvectors in columns.
//
// Supported sizes: n <= 32 (the caller checks).
// Method (per k, matrix resident in LDS, lane i <-> row i):
//   1. Householder reduction A = Q T Q^H (rank-2 trailing updates); the
//      reflectors are kept in the annihilated columns of A.
//   2. Diagonal phases P make the subdiagonal real: T = P T_r P^H.
//   3. Implicit-shift QL on T_r (scalar recurrence evaluated redundantly by
//      every lane; each rotation updates one row of Z per lane).
//   4. U = Q P Z: lane j keeps column j of P Z in registers and applies the
//      reflectors in reverse order; columns are written in ascending order.
//   5. Gauge convention shared with the oracle: each eigenvector's
//      largest-modulus component is made real and positive.
#include "mep_common.hpp"
#include "mep_kernels.hpp"

template <int NMAX>
__global__ __launch_bounds__(64) void eigh_kernel(int n, int64_t nk, const zd* __restrict__ Hk,
                                                  int64_t strideHk, const double* __restrict__ kpts,
                                                  const double* __restrict__ tau,
                                                  double* __restrict__ E, zd* __restrict__ U,
                                                  int* __restrict__ info) {
  constexpr int LDA = NMAX;
  constexpr int LDZ = NMAX + 1;
  __shared__ zd a[NMAX * NMAX];
  __shared__ double Zt[NMAX * LDZ];  // Zt[j*LDZ + i] = Z(i, j)
  __shared__ zd vv[NMAX], ww[NMAX], sub[NMAX], ph[NMAX];
  __shared__ double tauv[NMAX], d[NMAX], e[NMAX];
  const int t = threadIdx.x;
  const double two_pi = 2.0 * 3.14159265358979323846;
  const double EPS = 2.220446049250313e-16;

  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    const zd* H = Hk + k * strideHk;
    const double kx = kpts[3 * k], ky = kpts[3 * k + 1], kz = kpts[3 * k + 2];
    // ---- load the upper triangle (tight-binding phase applied) ----------
    for (int idx = t; idx < n * n; idx += 64) {
      int i = idx % n, j = idx / n;
      if (i <= j) {
        zd h = H[i + (int64_t)n * j];
        if (tau && i != j) {
          double ang = two_pi * (kx * (tau[3 * j] - tau[3 * i]) + ky * (tau[3 * j + 1] - tau[3 * i + 1]) +
                                 kz * (tau[3 * j + 2] - tau[3 * i + 2]));
          double sn, cs;
          sincos(ang, &sn, &cs);
          h = zmul(zd{cs, sn}, h);
        }
        if (i == j) {
          a[i + LDA * i] = zd{h.re, 0.0};
        } else {
          a[i + LDA * j] = h;
          a[j + LDA * i] = zconj(h);
        }
      }
    }
    __syncthreads();

    // ---- 1. Householder tridiagonalisation ------------------------------
    for (int kc = 0; kc + 2 < n; ++kc) {
      const int m = n - kc - 1;
      zd x = (t < m) ? a[(kc + 1 + t) + LDA * kc] : zd{0.0, 0.0};
      double tail2 = wave_sum((t >= 1 && t < m) ? zabs2(x) : 0.0);
      zd x0 = zd{__shfl(x.re, 0, 64), __shfl(x.im, 0, 64)};
      double tk = 0.0;
      zd alpha = x0;
      zd v = zd{0.0, 0.0};
      if (tail2 > 0.0) {
        double ax02 = zabs2(x0);
        double xn = sqrt(tail2 + ax02);
        double ax0 = sqrt(ax02);
        alpha = (ax0 > 0.0) ? zscale(x0, -xn / ax0) : zd{-xn, 0.0};
        v = x;
        if (t == 0) v = zsub(x, alpha);
        double vn2 = wave_sum((t < m) ? zabs2(v) : 0.0);
        tk = 2.0 / vn2;
        if (t < NMAX) vv[t] = (t < m) ? v : zd{0.0, 0.0};
        __syncthreads();
        zd p = zd{0.0, 0.0};
        if (t < m) {
          const zd* row = &a[kc + 1 + t + LDA * (kc + 1)];
          for (int j = 0; j < m; ++j) p = zadd(p, zmul(row[LDA * j], vv[j]));
          p = zscale(p, tk);
        }
        double K = 0.5 * tk * wave_sum((t < m) ? (v.re * p.re + v.im * p.im) : 0.0);
        zd w = zsub(p, zscale(v, K));
        if (t < NMAX) ww[t] = (t < m) ? w : zd{0.0, 0.0};
        __syncthreads();
        if (t < m) {
          zd* row = &a[kc + 1 + t + LDA * (kc + 1)];
          for (int j = 0; j < m; ++j) {
            zd vj = vv[j], wj = ww[j];
            zd upd = zadd(zmul(v, zconj(wj)), zmul(w, zconj(vj)));
            row[LDA * j] = zsub(row[LDA * j], upd);
          }
        }
      }
      // keep the reflector in the annihilated column, the subdiagonal aside
      if (t < m) a[(kc + 1 + t) + LDA * kc] = (tail2 > 0.0) ? v : zd{0.0, 0.0};
      if (t == 0) {
        sub[kc] = alpha;
        tauv[kc] = tk;
      }
      __syncthreads();
    }
    if (t == 0 && n >= 2) {
      sub[n - 2] = a[(n - 1) + LDA * (n - 2)];
      tauv[n - 2] = 0.0;
    }
    if (t < n) d[t] = a[t + LDA * t].re;
    for (int idx = t; idx < NMAX * LDZ; idx += 64) {
      int j = idx / LDZ, i = idx % LDZ;
      Zt[idx] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    // ---- 2. real tridiagonal by diagonal phases -------------------------
    if (t == 0) {
      ph[0] = zd{1.0, 0.0};
      for (int i = 0; i + 1 < n; ++i) {
        zd s = sub[i];
        double as = sqrt(zabs2(s));
        e[i] = as;
        ph[i + 1] = (as > 0.0) ? zmul(ph[i], zscale(s, 1.0 / as)) : ph[i];
      }
      e[n - 1] = 0.0;
    }
    __syncthreads();
    // ---- 3. implicit QL; every lane evaluates the same scalar recurrence --
    int fail = 0;
    for (int l = 0; l < n && !fail; ++l) {
      int iter = 0;
      for (;;) {
        int m = l;
        for (; m < n - 1; ++m) {
          double dd = fabs(d[m]) + fabs(d[m + 1]);
          if (fabs(e[m]) <= EPS * dd) break;
        }
        if (m == l) break;
        if (++iter > 60) {
          fail = 1;
          break;
        }
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + copysign(r, g));
        double s = 1.0, c = 1.0, p = 0.0;
        bool underflow = false;
        for (int i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          r = hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            underflow = true;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          if (t < n) {
            double zf = Zt[(i + 1) * LDZ + t];
            double zi = Zt[i * LDZ + t];
            Zt[(i + 1) * LDZ + t] = s * zi + c * zf;
            Zt[i * LDZ + t] = c * zi - s * zf;
          }
        }
        if (underflow) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    }
    if (fail && t == 0) atomicOr(info, 1);
    __syncthreads();
    // ---- ascending rank of each eigenvalue ------------------------------
    int rnk = 0;
    double dj = 0.0;
    if (t < n) {
      dj = d[t];
      for (int i = 0; i < n; ++i) {
        double di = d[i];
        rnk += (di < dj) || (di == dj && i < t);
      }
    }
    // ---- 4. back-transformation U = Q P Z (lane j <-> column j) ----------
    zd W[NMAX];
#pragma unroll
    for (int i = 0; i < NMAX; ++i)
      W[i] = (i < n && t < n) ? zscale(ph[i], Zt[t * LDZ + i]) : zd{0.0, 0.0};
    for (int kc = n - 3; kc >= 0; --kc) {
      double tk = tauv[kc];
      if (tk == 0.0) continue;
      zd s = zd{0.0, 0.0};
#pragma unroll
      for (int i = 0; i < NMAX; ++i)
        if (i > kc && i < n) s = zadd(s, zcmul(a[i + LDA * kc], W[i]));
      s = zscale(s, tk);
#pragma unroll
      for (int i = 0; i < NMAX; ++i)
        if (i > kc && i < n) W[i] = zsub(W[i], zmul(a[i + LDA * kc], s));
    }
    // gauge convention: largest-modulus component (first on ties) real > 0
    {
      double best = -1.0;
      zd u = zd{1.0, 0.0};
#pragma unroll
      for (int i = 0; i < NMAX; ++i) {
        double m2 = zabs2(W[i]);
        if (i < n && m2 > best) {
          best = m2;
          u = W[i];
        }
      }
      double au = sqrt(zabs2(u));
      if (au > 0.0) {
        zd phs = zd{u.re / au, -u.im / au};
#pragma unroll
        for (int i = 0; i < NMAX; ++i) W[i] = zmul(W[i], phs);
      }
    }
    if (t < n) {
      zd* Uk = U + k * (int64_t)n * n + (int64_t)n * rnk;
#pragma unroll
      for (int i = 0; i < NMAX; ++i)
        if (i < n) Uk[i] = W[i];
      E[k * n + rnk] = dj;
    }
    __syncthreads();
  }
}

void launch_eigh(hipStream_t s, int n, int64_t nk, const zd* Hk, int64_t strideHk, const KGeom& g,
                 double* E, zd* U, int* info) {
  unsigned grid = (unsigned)(nk < 65536 ? nk : 65536);
  if (n <= 8)
    hipLaunchKernelGGL(eigh_kernel<8>, dim3(grid), dim3(64), 0, s, n, nk, Hk, strideHk, g.kpts,
                       g.tau, E, U, info);
  else if (n <= 16)
    hipLaunchKernelGGL(eigh_kernel<16>, dim3(grid), dim3(64), 0, s, n, nk, Hk, strideHk, g.kpts,
                       g.tau, E, U, info);
  else
    hipLaunchKernelGGL(eigh_kernel<32>, dim3(grid), dim3(64), 0, s, n, nk, Hk, strideHk, g.kpts,
                       g.tau, E, U, info);
}
