// mep_velo.hip -- velocity (and connection / curvature) in the Hamiltonian
// gauge, one k-point per wavefront.
//
// Replaces W_to_H_gaugeTRAFO + get_velo (wann_interp.f90:100-101, 406-658).
// Inputs per k: the Fourier GEMM output in the Wannier gauge, stacked as
//   mat 0: H, 1-3: dH_a, [4-6: A_a, 7-9: Omega_yz, Omega_zx, Omega_xy]
// plus U (eigenvectors) and E.  The tight-binding phase of the reference's
// convention (wann_interp.f90:192-241) is applied on load:
//   X'_nm = e^{2 pi i k.(tau_m - tau_n)} X_nm,
//   dH'_a = ph (dH_a + i c_a(tau_m - tau_n) H),
//   Om'_ab = ph (Om_ab + i c_a(dtau) A_b - i c_b(dtau) A_a).
// Output: V per k planar, V[k][a][i + n j]; with a position operator also
// A^(H) (3,n,n) and Omega^(H) (3,3,n,n) in the reference layout.
#include "mep_common.hpp"
#include "mep_kernels.hpp"
#include "mep_mfma.hpp"

struct TBPhase {
  double kx, ky, kz;
  const double* kpts;
  const double* tau;
  double cm[9];
  __device__ inline TBPhase at(int64_t k) const {
    TBPhase r = *this;
    r.kx = kpts[3 * k];
    r.ky = kpts[3 * k + 1];
    r.kz = kpts[3 * k + 2];
    return r;
  }
  __device__ inline zd phase(int n, int m, double dtau[3]) const {
    dtau[0] = tau[3 * m] - tau[3 * n];
    dtau[1] = tau[3 * m + 1] - tau[3 * n + 1];
    dtau[2] = tau[3 * m + 2] - tau[3 * n + 2];
    const double two_pi = 2.0 * 3.14159265358979323846;
    double ang = two_pi * (kx * dtau[0] + ky * dtau[1] + kz * dtau[2]);
    double sn, cs;
    sincos(ang, &sn, &cs);
    return zd{cs, sn};
  }
  // c_a(dt) = sum_b cm[a][b] dt_b, with a selected without indexing cm
  __device__ inline double c(int a, const double dt[3]) const {
    double c0 = cm[0] * dt[0] + cm[1] * dt[1] + cm[2] * dt[2];
    double c1 = cm[3] * dt[0] + cm[4] * dt[1] + cm[5] * dt[2];
    double c2 = cm[6] * dt[0] + cm[7] * dt[1] + cm[8] * dt[2];
    return a == 0 ? c0 : (a == 1 ? c1 : c2);
  }
};

// Wannier-gauge element (row, col) of matrix `mat`: full column-major
// matrices of N x N, or (packed) the upper triangles of Hermitian matrices
// with (i <= j) at i + j(j+1)/2 and the lower half as conjugates.  HT is the
// stored complex type (FP64, or FP32 from the mixed-precision GEMM).
__device__ inline zd up64(zd x) { return x; }
__device__ inline zd up64(zf x) { return zd{(double)x.re, (double)x.im}; }
template <class HT>
__device__ inline zd load_w(const HT* X, int N, int packed, int row, int col, int mat) {
  if (!packed) return up64(X[(int64_t)mat * N * N + row + (int64_t)N * col]);
  const int64_t NP = ((int64_t)N * (N + 1)) / 2;
  if (row <= col) return up64(X[mat * NP + row + ((int64_t)col * (col + 1)) / 2]);
  return zconj(up64(X[mat * NP + col + ((int64_t)row * (row + 1)) / 2]));
}

// element (row, col) of mat `mat` after the TB epilogue (mats 0..9 as above)
template <class HT>
__device__ inline zd load_tb(const HT* X, int packed, int N, int row, int col, int mat,
                             const TBPhase& tb) {
  zd x = load_w(X, N, packed, row, col, mat);
  if (!tb.tau) return x;
  double dt[3];
  zd ph = tb.phase(row, col, dt);
  if (mat >= 1 && mat <= 3) {
    zd h = load_w(X, N, packed, row, col, 0);
    x = zadd(x, zmuli(zscale(h, tb.c(mat - 1, dt))));
  } else if (mat >= 7) {
    // Omega_ab with (a,b) = (1,2), (2,0), (0,1)
    int a = (mat == 7) ? 1 : (mat == 8) ? 2 : 0;
    int b = (mat == 7) ? 2 : (mat == 8) ? 0 : 1;
    zd Ab = load_w(X, N, packed, row, col, 4 + b), Aa = load_w(X, N, packed, row, col, 4 + a);
    x = zadd(x, zmuli(zsub(zscale(Ab, tb.c(a, dt)), zscale(Aa, tb.c(b, dt)))));
  }
  return zmul(ph, x);
}

// --------------------------------------------------------------------------
// no position operator: V_a = U^H dH_a U for the three directions, or
// V_a = dH_a without the gauge transformation.  V is written planar:
// V[k][a][i + n*j].
//
// One k-point per 256-thread workgroup; the waves own the 2 x 2 output
// tiles of 16 x 16.  U and the three dH_a are staged in LDS (dH_a is
// Hermitian: its packed upper triangle is read once, coalesced, and both
// halves are written).  T_a = dH_a U is formed in registers for all three a
// (nine MFMA chains: three real products per complex one), written over
// dH_a, and V_a = U^H T_a goes from the accumulators straight to HBM.
// LDS rows and columns >= n stay zero, so no operand load is guarded.
// --------------------------------------------------------------------------
template <class HT>
__global__ __launch_bounds__(256, 2) void velo_mfma_kernel(int n, int64_t nk,
                                                           const HT* __restrict__ Hk,
                                                           int64_t strideHk, int packed,
                                                           TBPhase tb0, int gauge_trafo,
                                                           const zd* __restrict__ U,
                                                           zd* __restrict__ V) {
  constexpr int NM = 32, LD = NM + 1, PL = NM * LD;
  __shared__ zd Us[PL], Ds[3 * PL];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int n2 = n * n;
  const int np = packed ? n * (n + 1) / 2 : n2;
  for (int c = t; c < PL; c += 256) Us[c] = zd{0.0, 0.0};
  for (int c = t; c < 3 * PL; c += 256) Ds[c] = zd{0.0, 0.0};
  const int r0 = 16 * (w & 1), c0 = 16 * (w >> 1);
  const int ar = r0 + (lane & 15), bc = c0 + (lane & 15), kq = lane >> 4;
  const bool tile = r0 < n && c0 < n;  // wave-uniform
  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    const TBPhase tb = tb0.at(k);
    const HT* X = Hk + k * strideHk;
    zd* Vk = V + k * 3 * (int64_t)n2;
    __syncthreads();  // the previous k-point's readers are done
    // ---- stage: every global load of the k-point before the LDS stores
    {
      constexpr int NU = (NM * NM + 255) / 256, ND = (NM * (NM + 1) / 2 + 255) / 256;
      zd ut[NU], dt[3][ND], ht[ND];
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        const int idx = t + 256 * u;
        ut[u] = (gauge_trafo && idx < n2) ? U[k * (int64_t)n2 + idx] : zd{0.0, 0.0};
      }
      // element p of the packed upper triangle: (i, j), i <= j, p = i + j(j+1)/2
      int pi[ND], pj[ND];
#pragma unroll
      for (int u = 0; u < ND; ++u) {
        const int p = t + 256 * u;
        int j = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
        if ((j + 1) * (j + 2) / 2 <= p) ++j;
        if (j * (j + 1) / 2 > p) --j;
        pj[u] = j;
        pi[u] = p - j * (j + 1) / 2;
      }
      if (packed) {
        const int64_t NPk = np;
#pragma unroll
        for (int u = 0; u < ND; ++u) {
          const int p = t + 256 * u;
          const bool ok = p < np;
#pragma unroll
          for (int a = 0; a < 3; ++a) dt[a][u] = ok ? up64(X[(1 + a) * NPk + p]) : zd{0.0, 0.0};
          ht[u] = (ok && tb.tau) ? up64(X[p]) : zd{0.0, 0.0};
        }
      }
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        const int idx = t + 256 * u;
        if (idx < n2) Us[idx % n + LD * (idx / n)] = ut[u];
      }
      if (packed) {
#pragma unroll
        for (int u = 0; u < ND; ++u) {
          const int p = t + 256 * u;
          if (p >= np) continue;
          const int i = pi[u], j = pj[u];
          zd ph = zd{1.0, 0.0};
          double dtau[3] = {0.0, 0.0, 0.0};
          if (tb.tau) ph = tb.phase(i, j, dtau);
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            // dH'_a(i, j) = e^{2 pi i k.(tau_j - tau_i)} (dH_a + i c_a(tau_j - tau_i) H)
            zd x = dt[a][u];
            if (tb.tau) x = zmul(ph, zadd(x, zmuli(zscale(ht[u], tb.c(a, dtau)))));
            Ds[a * PL + j + LD * i] = zconj(x);
            Ds[a * PL + i + LD * j] = x;  // last: the diagonal keeps x as stored
          }
        }
      } else {
        for (int a = 0; a < 3; ++a)
          for (int idx = t; idx < n2; idx += 256) {
            const int i = idx % n, j = idx / n;
            Ds[a * PL + i + LD * j] = load_tb(X, 0, n, i, j, 1 + a, tb);
          }
      }
    }
    __syncthreads();
    if (!gauge_trafo) {
      for (int a = 0; a < 3; ++a)
        for (int idx = t; idx < n2; idx += 256)
          Vk[a * (int64_t)n2 + idx] = Ds[a * PL + idx % n + LD * (idx / n)];
      continue;
    }
    // ---- T_a = dH_a U (C = sum_m A(., m) B(m, .): P1 = Ar Br, P2 = Ai Bi,
    // P3 = (Ar + Ai)(Br + Bi); Re C = P1 - P2, Im C = P3 - P1 - P2)
    mep_v4d p1[3], p2[3], p3[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) p1[a] = p2[a] = p3[a] = mep_v4d{0, 0, 0, 0};
    if (tile)
#pragma unroll
      for (int k0 = 0; k0 < NM; k0 += 4) {
        if (k0 >= n) break;
        const int m = k0 + kq;
        const zd b = Us[m + LD * bc];
        const double bs = b.re + b.im;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const zd x = Ds[a * PL + ar + LD * m];
          p1[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.re, b.re, p1[a], 0, 0, 0);
          p2[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.im, b.im, p2[a], 0, 0, 0);
          p3[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.re + x.im, bs, p3[a], 0, 0, 0);
        }
      }
    __syncthreads();  // every wave has read dH
    if (tile)
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = r0 + kq + 4 * v;
          Ds[a * PL + row + LD * bc] = zd{p1[a][v] - p2[a][v], p3[a][v] - p1[a][v] - p2[a][v]};
        }
    __syncthreads();
    // ---- V_a = U^H T_a, A(row, m) = conj(U(m, row))
#pragma unroll
    for (int a = 0; a < 3; ++a) p1[a] = p2[a] = p3[a] = mep_v4d{0, 0, 0, 0};
    if (tile) {
#pragma unroll
      for (int k0 = 0; k0 < NM; k0 += 4) {
        if (k0 >= n) break;
        const int m = k0 + kq;
        const zd u = Us[m + LD * ar];
        const double us = u.re - u.im;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const zd y = Ds[a * PL + m + LD * bc];
          p1[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(u.re, y.re, p1[a], 0, 0, 0);
          p2[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(-u.im, y.im, p2[a], 0, 0, 0);
          p3[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(us, y.re + y.im, p3[a], 0, 0, 0);
        }
      }
      if (bc < n)
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = r0 + kq + 4 * v;
            if (row < n)
              Vk[a * (int64_t)n2 + row + (int64_t)n * bc] =
                  zd{p1[a][v] - p2[a][v], p3[a][v] - p1[a][v] - p2[a][v]};
          }
    }
  }
}

// --------------------------------------------------------------------------
// 32 < n <= 128, no position operator: V_a = U^H dH_a U on FP64 MFMA, three
// real products per complex one.  Pass 0 forms T_a = dH_a U for a = x, y, z
// in the workgroup's slot Tscr[blockIdx] (3 n^2 complex), pass 1 V_a =
// U^H T_a; the 3 nt^2 (component, 16 x 16 tile) items of a pass are spread
// over the four waves, with one barrier per pass.  Operands come from L2 /
// HBM in chunks of VW_CH four-deep steps: the chunk's loads are all issued
// before its MFMAs, so their latencies overlap.
// --------------------------------------------------------------------------
typedef double velo_v4d __attribute__((ext_vector_type(4)));
constexpr int VW_CH = 8;

// lower-triangle tile p (row-major: (0,0), (1,0), (1,1), (2,0), ...) of an
// edge of tiles: V_a = U^H T_a is Hermitian, so the second product forms
// the tiles on and below the diagonal and writes each off-diagonal one
// twice (V(c, r) = conj V(r, c))
__device__ inline void lower_tile(int p, int* tr, int* tc) {
  int r = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
  if ((r + 1) * (r + 2) / 2 <= p) ++r;
  if (r * (r + 1) / 2 > p) --r;
  *tr = r;
  *tc = p - r * (r + 1) / 2;
}

template <int PASS>
__device__ inline void velo_wide_pass(int n, int wv, int lane, const zd* __restrict__ X, int packed,
                                      const TBPhase& tb, const zd* __restrict__ Uk,
                                      zd* __restrict__ T, zd* __restrict__ Vk) {
  const int lr = lane & 15, kq = lane >> 4;
  const int n2 = n * n, nt = (n + 15) / 16;
  const int nit = PASS == 0 ? nt * nt : nt * (nt + 1) / 2;  // tiles per component
  for (int item = wv; item < 3 * nit; item += 4) {
    const int a = item / nit, ti = item % nit;
    int tr = ti / nt, tc = ti % nt;
    if (PASS == 1) lower_tile(ti, &tr, &tc);
    const int r0 = 16 * tr, c0 = 16 * tc;
    const int ar = r0 + lr, bc = c0 + lr;
    const zd* Ta = T + (int64_t)a * n2;
    velo_v4d p1 = {0, 0, 0, 0}, p2 = {0, 0, 0, 0}, p3 = {0, 0, 0, 0};
    for (int m0 = 0; m0 < n; m0 += 4 * VW_CH) {
      zd av[VW_CH], bv[VW_CH];
#pragma unroll
      for (int i = 0; i < VW_CH; ++i) {
        const int m = m0 + 4 * i + kq;
        av[i] = bv[i] = zd{0.0, 0.0};
        if (m < n) {
          if (ar < n)
            av[i] = PASS == 0 ? load_tb(X, packed, n, ar, m, 1 + a, tb) : zconj(Uk[m + (int64_t)n * ar]);
          if (bc < n) bv[i] = PASS == 0 ? Uk[m + (int64_t)n * bc] : Ta[m + (int64_t)n * bc];
        }
      }
#pragma unroll
      for (int i = 0; i < VW_CH; ++i) {
        if (m0 + 4 * i >= n) break;  // uniform
        p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i].re, bv[i].re, p1, 0, 0, 0);
        p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i].im, bv[i].im, p2, 0, 0, 0);
        p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i].re + av[i].im, bv[i].re + bv[i].im, p3, 0, 0, 0);
      }
    }
    zd* out = PASS == 0 ? T + (int64_t)a * n2 : Vk + (int64_t)a * n2;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = r0 + kq + 4 * v;
      if (row < n && bc < n) {
        const zd z = zd{p1[v] - p2[v], p3[v] - p1[v] - p2[v]};
        out[row + (int64_t)n * bc] = z;
        if (PASS == 1 && tr != tc) out[bc + (int64_t)n * row] = zconj(z);
      }
    }
  }
}

__global__ __launch_bounds__(256, 2) void velo_wide_kernel(int n, int64_t nk, const zd* __restrict__ Hk,
                                                           int64_t strideHk, int packed, TBPhase tb0,
                                                           int gauge_trafo, const zd* __restrict__ U,
                                                           zd* __restrict__ V, zd* __restrict__ Tscr) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int n2 = n * n;
  zd* T = Tscr + (int64_t)blockIdx.x * 3 * n2;
  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    const TBPhase tb = tb0.at(k);
    const zd* X = Hk + k * strideHk;
    const zd* Uk = U + k * (int64_t)n2;
    zd* Vk = V + k * 3 * (int64_t)n2;
    if (!gauge_trafo) {
      for (int idx = t; idx < 3 * n2; idx += 256)
        Vk[idx] = load_tb(X, packed, n, (idx % n2) % n, (idx % n2) / n, 1 + idx / n2, tb);
      continue;
    }
    velo_wide_pass<0>(n, wv, lane, X, packed, tb, Uk, T, Vk);
    __syncthreads();  // T complete
    velo_wide_pass<1>(n, wv, lane, X, packed, tb, Uk, T, Vk);
    __syncthreads();  // T read before the next k-point overwrites it
  }
}

// --------------------------------------------------------------------------
// 32 < n <= 64, no position operator: the same products with the operands
// in LDS.  Plane U holds U (pitch LV = n | 1, odd, so the 16-byte reads of
// a 16-lane column tile fall in distinct banks); plane D holds dH_a (both
// triangles, unpacked once per component), then T_a = dH_a U over it.  Per
// component: stage dH_a | T = D U into registers | T over D | V_a = U^H T
// to HBM.  The MFMA loop then issues two LDS reads and two adds per three
// products (the L2 version above spends about fourteen VALU instructions per
// product on addresses and guards, and VALU and FP64 MFMA issue serialise).
// --------------------------------------------------------------------------
constexpr int VL_MAXT = 4;  // 16 x 16 tiles per matrix edge (n <= 64)

__host__ __device__ inline int velo_lds_pitch(int n) { return n | 1; }

__global__ __launch_bounds__(256) void velo_lds_kernel(int n, int64_t nk, const zd* __restrict__ Hk,
                                                       int64_t strideHk, int packed, TBPhase tb0,
                                                       const zd* __restrict__ U,
                                                       zd* __restrict__ V) {
  extern __shared__ __attribute__((aligned(16))) double vl_smem[];
  const int LV = velo_lds_pitch(n), PLV = LV * n;
  zd* Us = reinterpret_cast<zd*>(vl_smem);
  zd* Ds = Us + PLV;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int lr = lane & 15, kq = lane >> 4;
  const int n2 = n * n, nt = (n + 15) / 16, ntile = nt * nt;
  // tile rows / columns n .. 16 nt - 1 lie outside the planes: their
  // operand loads read a clamped index and select zero
  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    const TBPhase tb = tb0.at(k);
    const zd* X = Hk + k * strideHk;
    const zd* Uk = U + k * (int64_t)n2;
    zd* Vk = V + k * 3 * (int64_t)n2;
    __syncthreads();  // the previous k-point's last reads of U and T are done
    for (int idx = t; idx < n2; idx += 256) stz(Us + idx % n + LV * (idx / n), Uk[idx]);
    for (int a = 0; a < 3; ++a) {
      for (int idx = t; idx < n2; idx += 256) {
        const int i = idx % n, j = idx / n;
        stz(Ds + i + LV * j, load_tb(X, packed, n, i, j, 1 + a, tb));
      }
      __syncthreads();  // U and dH_a staged
      // T = D U: up to VL_MAXT^2 / 4 tiles per wave, kept in registers
      velo_v4d q1[4], q2[4], q3[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) q1[u] = q2[u] = q3[u] = velo_v4d{0, 0, 0, 0};
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int tile = wv + 4 * u;
        if (tile >= ntile) break;  // uniform
        const int ar = 16 * (tile / nt) + lr, bc = 16 * (tile % nt) + lr;
        const bool aok = ar < n, bok = bc < n;
        const int arc = aok ? ar : 0, bcc = bok ? bc : 0;
        for (int m0 = 0; m0 < n; m0 += 4) {
          const int m = m0 + kq, mc = m < n ? m : 0;
          zd x = ldz(Ds + arc + LV * mc), y = ldz(Us + mc + LV * bcc);
          if (!(aok && m < n)) x = zd{0.0, 0.0};
          if (!(bok && m < n)) y = zd{0.0, 0.0};
          q1[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.re, y.re, q1[u], 0, 0, 0);
          q2[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.im, y.im, q2[u], 0, 0, 0);
          q3[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.re + x.im, y.re + y.im, q3[u], 0, 0, 0);
        }
      }
      __syncthreads();  // every wave has read dH_a
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int tile = wv + 4 * u;
        if (tile >= ntile) break;
        const int r0 = 16 * (tile / nt), bc = 16 * (tile % nt) + lr;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = r0 + kq + 4 * v;
          if (row < n && bc < n)
            stz(Ds + row + LV * bc, zd{q1[u][v] - q2[u][v], q3[u][v] - q1[u][v] - q2[u][v]});
        }
      }
      __syncthreads();  // T complete
      // V_a = U^H T, straight to HBM: the tiles on and below the diagonal
      // (V_a is Hermitian), each off-diagonal one written twice
      const int nlt = nt * (nt + 1) / 2;
#pragma unroll
      for (int u = 0; u < 3; ++u) {
        const int tile = wv + 4 * u;
        if (tile >= nlt) break;
        int tr, tc;
        lower_tile(tile, &tr, &tc);
        const int r0 = 16 * tr;
        const int ar = r0 + lr, bc = 16 * tc + lr;
        const bool aok = ar < n, bok = bc < n;
        const int arc = aok ? ar : 0, bcc = bok ? bc : 0;
        velo_v4d p1 = {0, 0, 0, 0}, p2 = {0, 0, 0, 0}, p3 = {0, 0, 0, 0};
        for (int m0 = 0; m0 < n; m0 += 4) {
          const int m = m0 + kq, mc = m < n ? m : 0;
          zd x = ldz(Us + mc + LV * arc), y = ldz(Ds + mc + LV * bcc);
          x.im = -x.im;  // A(row, m) = conj(U(m, row))
          if (!(aok && m < n)) x = zd{0.0, 0.0};
          if (!(bok && m < n)) y = zd{0.0, 0.0};
          p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(x.re, y.re, p1, 0, 0, 0);
          p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(x.im, y.im, p2, 0, 0, 0);
          p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(x.re + x.im, y.re + y.im, p3, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = r0 + kq + 4 * v;
          if (row < n && bc < n) {
            const zd z = zd{p1[v] - p2[v], p3[v] - p1[v] - p2[v]};
            Vk[a * (int64_t)n2 + row + (int64_t)n * bc] = z;
            if (tr != tc) Vk[a * (int64_t)n2 + bc + (int64_t)n * row] = zconj(z);
          }
        }
      }
      __syncthreads();  // T read before the next component's dH overwrites it
    }
  }
}

size_t velo_lds_bytes(int n) { return sizeof(zd) * 2 * (size_t)velo_lds_pitch(n) * n; }

// V (3,n,n) per k: planar [a][i + n j]  <->  reference (a fastest)
__global__ void v_planar_to_ref(int n, int64_t nk, const zd* __restrict__ Vp, zd* __restrict__ Vr) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t per = 3 * (int64_t)n * n;
  if (t >= nk * per) return;
  int64_t k = t / per, r = t % per;
  int a = (int)(r % 3);
  int64_t ij = r / 3;
  Vr[t] = Vp[k * per + a * (int64_t)n * n + ij];
}
__global__ void v_ref_to_planar(int n, int64_t nk, const zd* __restrict__ Vr, zd* __restrict__ Vp) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t per = 3 * (int64_t)n * n, n2 = (int64_t)n * n;
  if (t >= nk * per) return;
  int64_t k = t / per, r = t % per;
  int a = (int)(r / n2);
  int64_t ij = r % n2;
  Vp[t] = Vr[k * per + a + 3 * ij];
}

void launch_v_layout(hipStream_t s, int n, int64_t nk, const zd* src, zd* dst, bool to_ref) {
  int64_t tot = nk * 3 * (int64_t)n * n;
  unsigned nb = (unsigned)((tot + 255) / 256);
  if (to_ref)
    hipLaunchKernelGGL(v_planar_to_ref, dim3(nb), dim3(256), 0, s, n, nk, src, dst);
  else
    hipLaunchKernelGGL(v_ref_to_planar, dim3(nb), dim3(256), 0, s, n, nk, src, dst);
}

// --------------------------------------------------------------------------
// position operator present (n <= 16): PRB 74 195118 Eqs. 21-31 with the
// reference's evaluation order (rotate_gauge :502-550, get_gauge_covar_deriv
// :565-617, conn_gaugeTrafo :620-633, curv_gaugeTrafo :636-658, get_velo
// :406-444).  textbook != 0 selects the printed equations instead.
// --------------------------------------------------------------------------
#define PM 16
#define PLD (PM + 1)

// C = op(A) * B over the n x n leading blocks of LDS matrices (ld PLD)
__device__ inline void lds_mm(int n, const zd* A, bool adjA, const zd* B, zd* C) {
  for (int idx = threadIdx.x; idx < n * n; idx += 64) {
    int i = idx % n, j = idx / n;
    zd s = zd{0.0, 0.0};
    for (int l = 0; l < n; ++l) {
      zd aa = adjA ? zconj(A[l + PLD * i]) : A[i + PLD * l];
      s = zadd(s, zmul(aa, B[l + PLD * j]));
    }
    C[i + PLD * j] = s;
  }
}

__global__ __launch_bounds__(64) void velo_pos_kernel(int n, int64_t nk, const zd* __restrict__ Hk,
                                                      int64_t strideHk, TBPhase tb, int gauge_trafo,
                                                      int textbook, double kubo_tol,
                                                      const zd* __restrict__ U,
                                                      const double* __restrict__ E,
                                                      zd* __restrict__ V, zd* __restrict__ Aout,
                                                      zd* __restrict__ Omout) {
  __shared__ zd Us[PM * PLD], T1[PM * PLD], T2[PM * PLD];
  __shared__ zd Hb[3][PM * PLD], Ab[3][PM * PLD], Dm[3][PM * PLD], Om[3][PM * PLD];
  __shared__ zd Abar[3][PM * PLD];
  __shared__ double Es[PM];
  const int t = threadIdx.x;
  const int N = n;
  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    const TBPhase tk = tb.at(k);
    const zd* X = Hk + k * strideHk;
    const int64_t NN = (int64_t)n * n;
    for (int idx = t; idx < n * n; idx += 64) {
      int i = idx % n, j = idx / n;
      Us[i + PLD * j] = U[k * NN + idx];
      for (int a = 0; a < 3; ++a) {
        Hb[a][i + PLD * j] = load_tb(X, 0, N, i, j, 1 + a, tk);
        Ab[a][i + PLD * j] = load_tb(X, 0, N, i, j, 4 + a, tk);
        Om[a][i + PLD * j] = load_tb(X, 0, N, i, j, 7 + a, tk);
      }
    }
    if (t < n) Es[t] = E[k * n + t];
    __syncthreads();
    if (gauge_trafo) {
      for (int a = 0; a < 3; ++a) {
        // Hbar = U^H dH U ; tmp1 = dH U
        lds_mm(n, Hb[a], false, Us, T1);
        __syncthreads();
        lds_mm(n, Us, true, T1, Hb[a]);
        __syncthreads();
        if (textbook) {
          lds_mm(n, Ab[a], false, Us, T2);
          __syncthreads();
          lds_mm(n, Us, true, T2, Ab[a]);
        } else {
          // stale temporary: A = (U^H (dH U)) U
          lds_mm(n, Us, true, T1, T2);
          __syncthreads();
          lds_mm(n, T2, false, Us, Ab[a]);
        }
        __syncthreads();
        // Omega_bar = U^H Omega U
        lds_mm(n, Us, true, Om[a], T2);
        __syncthreads();
        lds_mm(n, T2, false, Us, Om[a]);
        __syncthreads();
      }
      // D_a,ij = Hbar_a,ij / (E_j - E_i) for i != j and |E_j - E_i| > kubo_tol
      for (int idx = t; idx < n * n; idx += 64) {
        int i = idx % n, j = idx / n;
        double de = Es[j] - Es[i];
        for (int a = 0; a < 3; ++a) {
          zd dv = zd{0.0, 0.0};
          if (i != j && fabs(de) > kubo_tol)
            dv = zd{Hb[a][i + PLD * j].re / de, Hb[a][i + PLD * j].im / de};
          Dm[a][i + PLD * j] = dv;
          Abar[a][i + PLD * j] = Ab[a][i + PLD * j];
        }
      }
      __syncthreads();
      // A^(H) = A + i D
      for (int idx = t; idx < n * n; idx += 64) {
        int i = idx % n, j = idx / n;
        for (int a = 0; a < 3; ++a)
          Ab[a][i + PLD * j] = zadd(Ab[a][i + PLD * j], zmuli(Dm[a][i + PLD * j]));
      }
      __syncthreads();
      // Omega^(H)_ab = Om_ab - [D_a, A_b] + [D_b, A_a] - i [D_a, D_b]
      // for (a,b) in {(1,2), (2,0), (0,1)}; A is A^(H) (reference) or Abar
      for (int c = 0; c < 3; ++c) {
        int a = (c == 0) ? 1 : (c == 1) ? 2 : 0;
        int b = (c == 0) ? 2 : (c == 1) ? 0 : 1;
        zd(*Asrc)[PM * PLD] = textbook ? Abar : Ab;
        for (int idx = t; idx < n * n; idx += 64) {
          int i = idx % n, j = idx / n;
          zd s = Om[c][i + PLD * j];
          zd c1 = zd{0, 0}, c2 = zd{0, 0}, c3 = zd{0, 0};
          for (int l = 0; l < n; ++l) {
            // [D_a, A_b]_ij
            c1 = zadd(c1, zsub(zmul(Dm[a][i + PLD * l], Asrc[b][l + PLD * j]),
                               zmul(Asrc[b][i + PLD * l], Dm[a][l + PLD * j])));
            // [D_b, A_a]_ij
            c2 = zadd(c2, zsub(zmul(Dm[b][i + PLD * l], Asrc[a][l + PLD * j]),
                               zmul(Asrc[a][i + PLD * l], Dm[b][l + PLD * j])));
            // [D_a, D_b]_ij
            c3 = zadd(c3, zsub(zmul(Dm[a][i + PLD * l], Dm[b][l + PLD * j]),
                               zmul(Dm[b][i + PLD * l], Dm[a][l + PLD * j])));
          }
          s = zsub(s, c1);
          s = zadd(s, c2);
          s = zsub(s, zmuli(c3));
          T1[i + PLD * j] = s;
        }
        __syncthreads();
        for (int idx = t; idx < n * n; idx += 64) {
          int i = idx % n, j = idx / n;
          Om[c][i + PLD * j] = T1[i + PLD * j];
        }
        __syncthreads();
      }
    }
    // velocity
    for (int idx = t; idx < n * n; idx += 64) {
      int i = idx % n, j = idx / n;
      for (int a = 0; a < 3; ++a) {
        zd v = Hb[a][i + PLD * j];
        if (i != j) {
          // V_nm = H_nm - i (E_m - E_n) A_nm
          const zd* Asel = (gauge_trafo && textbook) ? Abar[a] : Ab[a];
          zd ce = zd{0.0, Es[j] - Es[i]};
          v = zsub(v, zmul(ce, Asel[i + PLD * j]));
        }
        V[k * 3 * NN + a * NN + idx] = v;
        Aout[k * 3 * NN + a + 3 * (int64_t)idx] = Ab[a][i + PLD * j];
      }
      // Omega (3,3,n,n): antisymmetric, zero diagonal in (a,b)
      zd* O = Omout + k * 9 * NN + 9 * (int64_t)idx;
      zd oyz = Om[0][i + PLD * j], ozx = Om[1][i + PLD * j], oxy = Om[2][i + PLD * j];
      O[0] = zd{0, 0};
      O[1 + 3 * 1] = zd{0, 0};
      O[2 + 3 * 2] = zd{0, 0};
      O[1 + 3 * 2] = oyz;
      O[2 + 3 * 1] = zsub(zd{0, 0}, oyz);
      O[2 + 3 * 0] = ozx;
      O[0 + 3 * 2] = zsub(zd{0, 0}, ozx);
      O[0 + 3 * 1] = oxy;
      O[1 + 3 * 0] = zsub(zd{0, 0}, oxy);
    }
    __syncthreads();
  }
}

// C = op(A) * B for n x n matrices of pitch n in global memory (the
// workgroup's slot), 256 threads
__device__ inline void gmm(int n, const zd* A, bool adjA, const zd* B, zd* C) {
  for (int idx = threadIdx.x; idx < n * n; idx += 256) {
    const int i = idx % n, j = idx / n;
    zd s0 = zd{0.0, 0.0}, s1 = zd{0.0, 0.0};
    int l = 0;
    for (; l + 1 < n; l += 2) {
      const zd a0 = adjA ? zconj(A[l + (int64_t)n * i]) : A[i + (int64_t)n * l];
      const zd a1 = adjA ? zconj(A[l + 1 + (int64_t)n * i]) : A[i + (int64_t)n * (l + 1)];
      s0 = zadd(s0, zmul(a0, B[l + (int64_t)n * j]));
      s1 = zadd(s1, zmul(a1, B[l + 1 + (int64_t)n * j]));
    }
    if (l < n) {
      const zd a0 = adjA ? zconj(A[l + (int64_t)n * i]) : A[i + (int64_t)n * l];
      s0 = zadd(s0, zmul(a0, B[l + (int64_t)n * j]));
    }
    C[i + (int64_t)n * j] = zadd(s0, s1);
  }
}

// The position-operator path for 16 < n <= 128: velo_pos_kernel's steps in
// the same order (rotate_gauge :502-550, get_gauge_covar_deriv :565-617,
// conn_gaugeTrafo :620-633, curv_gaugeTrafo :636-658, get_velo :406-444),
// with its eighteen n x n matrices in the workgroup's slot of `slots`
// (18 n^2 complex each; the slots lie in the consumed eigensolver records)
// instead of LDS.
__global__ __launch_bounds__(256) void velo_pos_global_kernel(int n, int64_t nk,
                                                              const zd* __restrict__ Hk,
                                                              int64_t strideHk, TBPhase tb,
                                                              int gauge_trafo, int textbook,
                                                              double kubo_tol,
                                                              const zd* __restrict__ U,
                                                              const double* __restrict__ E,
                                                              zd* __restrict__ V, zd* __restrict__ Aout,
                                                              zd* __restrict__ Omout,
                                                              zd* __restrict__ slots) {
  __shared__ double Es[128];
  const int t = threadIdx.x;
  const int64_t NN = (int64_t)n * n;
  zd* base = slots + (int64_t)blockIdx.x * 18 * NN;
  zd *Us = base, *T1 = base + NN, *T2 = base + 2 * NN;
  zd *Hb[3], *Ab[3], *Dm[3], *Om[3], *Abar[3];
  for (int a = 0; a < 3; ++a) {
    Hb[a] = base + (3 + a) * NN;
    Ab[a] = base + (6 + a) * NN;
    Dm[a] = base + (9 + a) * NN;
    Om[a] = base + (12 + a) * NN;
    Abar[a] = base + (15 + a) * NN;
  }
  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    const TBPhase tk = tb.at(k);
    const zd* X = Hk + k * strideHk;
    __syncthreads();  // the previous k-point's reads of the slot are done
    for (int idx = t; idx < n * n; idx += 256) {
      const int i = idx % n, j = idx / n;
      Us[idx] = U[k * NN + idx];
      for (int a = 0; a < 3; ++a) {
        Hb[a][idx] = load_tb(X, 0, n, i, j, 1 + a, tk);
        Ab[a][idx] = load_tb(X, 0, n, i, j, 4 + a, tk);
        Om[a][idx] = load_tb(X, 0, n, i, j, 7 + a, tk);
      }
    }
    if (t < n) Es[t] = E[k * n + t];
    __syncthreads();
    if (gauge_trafo) {
      for (int a = 0; a < 3; ++a) {
        gmm(n, Hb[a], false, Us, T1);  // dH U
        __syncthreads();
        gmm(n, Us, true, T1, Hb[a]);    // Hbar = U^H dH U
        __syncthreads();
        if (textbook) {
          gmm(n, Ab[a], false, Us, T2);
          __syncthreads();
          gmm(n, Us, true, T2, Ab[a]);
        } else {
          // stale temporary: A = (U^H (dH U)) U
          gmm(n, Us, true, T1, T2);
          __syncthreads();
          gmm(n, T2, false, Us, Ab[a]);
        }
        __syncthreads();
        gmm(n, Us, true, Om[a], T2);  // Omega_bar = U^H Omega U
        __syncthreads();
        gmm(n, T2, false, Us, Om[a]);
        __syncthreads();
      }
      // D_a,ij = Hbar_a,ij / (E_j - E_i) for i != j and |E_j - E_i| > kubo_tol
      for (int idx = t; idx < n * n; idx += 256) {
        const int i = idx % n, j = idx / n;
        const double de = Es[j] - Es[i];
        for (int a = 0; a < 3; ++a) {
          zd dv = zd{0.0, 0.0};
          if (i != j && fabs(de) > kubo_tol) dv = zd{Hb[a][idx].re / de, Hb[a][idx].im / de};
          Dm[a][idx] = dv;
          Abar[a][idx] = Ab[a][idx];
        }
      }
      __syncthreads();
      // A^(H) = A + i D
      for (int idx = t; idx < n * n; idx += 256)
        for (int a = 0; a < 3; ++a) Ab[a][idx] = zadd(Ab[a][idx], zmuli(Dm[a][idx]));
      __syncthreads();
      // Omega^(H)_ab = Om_ab - [D_a, A_b] + [D_b, A_a] - i [D_a, D_b]
      // for (a,b) in {(1,2), (2,0), (0,1)}; A is A^(H) (reference) or Abar
      for (int c = 0; c < 3; ++c) {
        const int a = (c == 0) ? 1 : (c == 1) ? 2 : 0;
        const int b = (c == 0) ? 2 : (c == 1) ? 0 : 1;
        zd* const* Asrc = textbook ? Abar : Ab;
        for (int idx = t; idx < n * n; idx += 256) {
          const int i = idx % n, j = idx / n;
          zd s_ = Om[c][idx];
          zd c1 = zd{0, 0}, c2 = zd{0, 0}, c3 = zd{0, 0};
          for (int l = 0; l < n; ++l) {
            const int64_t il = i + (int64_t)n * l, lj = l + (int64_t)n * j;
            c1 = zadd(c1, zsub(zmul(Dm[a][il], Asrc[b][lj]), zmul(Asrc[b][il], Dm[a][lj])));
            c2 = zadd(c2, zsub(zmul(Dm[b][il], Asrc[a][lj]), zmul(Asrc[a][il], Dm[b][lj])));
            c3 = zadd(c3, zsub(zmul(Dm[a][il], Dm[b][lj]), zmul(Dm[b][il], Dm[a][lj])));
          }
          s_ = zsub(s_, c1);
          s_ = zadd(s_, c2);
          s_ = zsub(s_, zmuli(c3));
          T1[idx] = s_;
        }
        __syncthreads();
        for (int idx = t; idx < n * n; idx += 256) Om[c][idx] = T1[idx];
        __syncthreads();
      }
    }
    // velocity, connection and curvature in the reference's layouts
    for (int idx = t; idx < n * n; idx += 256) {
      const int i = idx % n, j = idx / n;
      for (int a = 0; a < 3; ++a) {
        zd v = Hb[a][idx];
        if (i != j) {
          // V_nm = H_nm - i (E_m - E_n) A_nm
          const zd* Asel = (gauge_trafo && textbook) ? Abar[a] : Ab[a];
          const zd ce = zd{0.0, Es[j] - Es[i]};
          v = zsub(v, zmul(ce, Asel[idx]));
        }
        V[k * 3 * NN + a * NN + idx] = v;
        Aout[k * 3 * NN + a + 3 * (int64_t)idx] = Ab[a][idx];
      }
      zd* O = Omout + k * 9 * NN + 9 * (int64_t)idx;
      const zd oyz = Om[0][idx], ozx = Om[1][idx], oxy = Om[2][idx];
      O[0] = zd{0, 0};
      O[1 + 3 * 1] = zd{0, 0};
      O[2 + 3 * 2] = zd{0, 0};
      O[1 + 3 * 2] = oyz;
      O[2 + 3 * 1] = zsub(zd{0, 0}, oyz);
      O[2 + 3 * 0] = ozx;
      O[0 + 3 * 2] = zsub(zd{0, 0}, ozx);
      O[0 + 3 * 1] = oxy;
      O[1 + 3 * 0] = zsub(zd{0, 0}, oxy);
    }
  }
}

bool velo_pos_scratch_ok(int n, size_t scratch_bytes) {
  return n <= PM || scratch_bytes >= 18 * (size_t)n * n * sizeof(zd);
}

void launch_velo(hipStream_t s, const VeloArgs& va, int64_t nk, const zd* Hk, int64_t strideHk,
                 const KGeom& g, const zd* U, const double* E, zd* V, zd* A, zd* Om) {
  unsigned grid = (unsigned)(nk < 65536 ? nk : 65536);
  TBPhase tb;
  tb.kx = tb.ky = tb.kz = 0.0;
  tb.kpts = g.kpts;
  tb.tau = g.tau;
  for (int i = 0; i < 9; ++i) tb.cm[i] = g.cmat[i];
  if (va.pos_op && va.n <= PM) {
    hipLaunchKernelGGL(velo_pos_kernel, dim3(grid), dim3(64), 0, s, va.n, nk, Hk, strideHk, tb,
                       va.gauge_trafo, va.textbook, va.kubo_tol, U, E, V, A, Om);
    return;
  }
  if (va.pos_op) {
    // one workgroup per slot of 18 n^2 complex in the scratch
    const int64_t slot = 18 * (int64_t)va.n * va.n * (int64_t)sizeof(zd);
    int64_t pg = (int64_t)(va.scratch_bytes / (size_t)slot);
    if (pg > nk) pg = nk;
    if (pg > 2048) pg = 2048;
    if (pg < 1) pg = 1;  // the caller guarantees one slot (velo_pos_scratch_ok)
    hipLaunchKernelGGL(velo_pos_global_kernel, dim3((unsigned)pg), dim3(256), 0, s, va.n, nk, Hk,
                       strideHk, tb, va.gauge_trafo, va.textbook, va.kubo_tol, U, E, V, A, Om,
                       va.scratch);
    return;
  }
  if (va.n > 32 && va.n <= 16 * VL_MAXT && va.gauge_trafo) {
    const size_t lds = velo_lds_bytes(va.n);
    (void)hipFuncSetAttribute((const void*)velo_lds_kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    unsigned lgrid = (unsigned)(nk < 8192 ? nk : 8192);
    hipLaunchKernelGGL(velo_lds_kernel, dim3(lgrid), dim3(256), lds, s, va.n, nk, Hk, strideHk,
                       va.packed, tb, U, V);
    return;
  }
  if (va.n > 32) {
    // T slots: 3 n^2 complex per workgroup (48 n^2 bytes), within the
    // eigensolver's rotation records that va.scratch reuses (64 n^2 bytes or
    // more per k-point of the block, and there are at least wgrid of them)
    unsigned wgrid = (unsigned)(nk < 2048 ? nk : 2048);
    hipLaunchKernelGGL(velo_wide_kernel, dim3(wgrid), dim3(256), 0, s, va.n, nk, Hk, strideHk,
                       va.packed, tb, va.gauge_trafo, U, V, va.scratch);
    return;
  }
  hipLaunchKernelGGL(velo_mfma_kernel<zd>, dim3(grid), dim3(256), 0, s, va.n, nk, Hk, strideHk,
                     va.packed, tb, va.gauge_trafo, U, V);
}

void launch_velo_h32(hipStream_t s, const VeloArgs& va, int64_t nk, const zf* Hk,
                     int64_t strideHk, const KGeom& g, const zd* U, zd* V) {
  unsigned grid = (unsigned)(nk < 65536 ? nk : 65536);
  TBPhase tb;
  tb.kx = tb.ky = tb.kz = 0.0;
  tb.kpts = g.kpts;
  tb.tau = g.tau;
  for (int i = 0; i < 9; ++i) tb.cm[i] = g.cmat[i];
  hipLaunchKernelGGL(velo_mfma_kernel<zf>, dim3(grid), dim3(256), 0, s, va.n, nk, Hk, strideHk,
                     va.packed, tb, va.gauge_trafo, U, V);
}
