// mep_bandsum.hip -- Fermi-Dirac weights and the Berry / Kubo / MEP band
// sums of one batch of k-points, accumulated into per-workgroup partial
// records (reduced once per run by reduce_parts_kernel).
//
// One k-point at a time per 256-thread workgroup (grid-stride over k; fixed
// k -> workgroup map and fixed-order reductions, so results are
// deterministic).  Rows of the reference it replaces:
//   d1_fd_stat, d1_fd_get_N_el          statistics.f90:65-124
//   kubo_ahc_tens, velo_ahc_tens,
//   kubo_ohc_tens                       kubo.f90:37-198
//   kubo_mep_CS / LC / IC               kubo_mep.f90:16-149
//   mep_niu_CS / LC / IC                mep_niu.f90:20-159
//
// Per k the workgroup stages in LDS the three velocity planes V_a (padded
// column-major) and the antisymmetric reciprocal table
// r(m, n) = 1 / (E_n - E_m) (0 on the diagonal), so no band sum divides.
//   pair phase    lane (n0, part) of every wave walks n = part, part + P,
//                 ...; wave 0 sums the Berry curvature of n0 (and finds the
//                 Niu exclusion windows), wave 1 the local (LC) MEP terms,
//                 waves 2-3 (single (hw, E_F)) the velo/ohc pair sums; the
//                 band sums are wavefront shuffle reductions (several
//                 (hw, E_F) items: hwsum_kernel, launched after this one);
//   itinerant     C_j = V_k B_l - V_l B_k, (j,k,l) cyclic, with
//                 B_l(m, n0) = V_l(m, n0) r(m, n0): three complex n x n
//                 products on FP64 MFMA (three real products per complex
//                 one); the IC tensor of band n0 is the
//                 diagonal -Re[A_i C_j](n0, n0) with
//                 A_i(n0, n) = V_i(n0, n) r(n, n0)^2, contracted straight
//                 from the accumulator registers.  mep_niu's guard
//                 |(E0-En)^2 (E0-Em)| > kubo_tol excludes, for fixed (n0, n),
//                 one contiguous window of m around n0 (energies are
//                 sorted); those terms are subtracted from C (or the kept
//                 ones summed directly when the window covers most bands).
#include "mep_common.hpp"
#include "mep_kernels.hpp"
#include "mep_lane.hpp"
#include "../../include/mepinterp.h"

typedef double mep_v4d __attribute__((ext_vector_type(4)));

AccLayout acc_layout(const BandArgs& b) {
  AccLayout L;
  int64_t o = 0;
  L.Ne = o;   o += b.n_ef;
  L.ahc = o;  o += 9 * (int64_t)b.n_ef;
  L.velo = o; o += 2 * 9 * (int64_t)b.n_hw * b.n_ef;
  L.ohc = o;  o += 2 * 9 * (int64_t)b.n_hw * b.n_ef;
  L.kic = o;  o += 9 * (int64_t)b.n_ef;
  L.klc = o;  o += 9 * (int64_t)b.n_ef;
  L.kcs = o;  o += 9 * (int64_t)b.n_ef;
  L.nic = o;  o += 9 * (int64_t)(b.n_val > 0 ? b.n_val : 1);
  L.nlc = o;  o += 9 * (int64_t)(b.n_val > 0 ? b.n_val : 1);
  L.ncs = o;  o += 9 * (int64_t)(b.n_val > 0 ? b.n_val : 1);
  L.total = o;
  return L;
}

constexpr int BS_THREADS = 256;

// LDS plan: V planes (zd) 3*PL | r table PL | E NMAX | fd nef*NMAX |
// F2 guarded NT*9*NMAX | IC tile sums 9*NT*NT*nef | red NMAX |
// per-band Berry and LC tensors 18*NMAX | pair sums 96 | acc total |
// Niu windows 2*NMAX*NMAX bytes              (doubles after the V planes)
template <int NMAX>
struct BsLds {
  static constexpr int LD = NMAX + 1;
  static constexpr int PL = NMAX * LD;
  static constexpr int NT = NMAX / 16;  // 16-row tiles per matrix
  static size_t doubles(int nef, int64_t total) {
    return (size_t)PL + NMAX + (size_t)nef * NMAX + (size_t)NT * 9 * NMAX +
           (size_t)9 * NT * NT * nef + NMAX + 18 * NMAX + 96 + (size_t)total;
  }
  static size_t bytes(int nef, int64_t total) {
    return sizeof(zd) * 3 * (size_t)PL + sizeof(double) * doubles(nef, total) +
           2 * (size_t)NMAX * NMAX;
  }
};


// FUSED: V is not read from HBM but formed in the V planes from the packed
// H(k) / dH_a(k) of the Fourier stage (Hk, per-k stride strideHk) and the
// eigenvectors U: V_a = U^H dH_a U on FP64 MFMA, with the same products and
// the same order as velo_mfma_kernel (so V, and every sum, is bit-identical
// to the two-kernel path).  n in 17..32, no position operator, Wannier
// convention.
template <int NMAX, bool FUSED>
__global__ __launch_bounds__(BS_THREADS, 2) void bandsum_kernel(BandArgs b, double Tsmear,
                                                                AccLayout L, int64_t gap,
                                                                int64_t nk,
                                                                const double* __restrict__ E,
                                                                const zd* __restrict__ V,
                                                                const zd* __restrict__ A,
                                                                const zd* __restrict__ Om,
                                                                const double* __restrict__ w,
                                                                double* __restrict__ part,
                                                                const zd* __restrict__ Hk,
                                                                int64_t strideHk,
                                                                const zd* __restrict__ U) {
  using S_ = BsLds<NMAX>;
  constexpr int LD = S_::LD, PL = S_::PL, NT = S_::NT;
  // U's plane (PL complex) lives over the per-k tables between the r table
  // and the accumulators while V is formed
  static_assert(!FUSED || (NMAX == 32 && PL + NMAX + NT * 9 * NMAX + NMAX + 18 * NMAX + 96 >= 2 * PL),
                "fused velocities: U's plane must fit below the accumulators");
  extern __shared__ __attribute__((aligned(16))) double smem[];
  const int n = b.n, nef = b.n_ef, nhw = b.n_hw, nval = b.n_val;
  zd* Vs = reinterpret_cast<zd*>(smem);
  double* Rs = reinterpret_cast<double*>(Vs + 3 * PL);
  double* Es = Rs + PL;
  double* fd = Es + NMAX;
  double* QF = fd + (int64_t)nef * NMAX;        // [NT][9][NMAX] guarded IC per band
  double* TS = QF + NT * 9 * NMAX;              // [3*NT*NT tiles][3][nef] Kubo IC tile sums
  double* red = TS + 9 * NT * NT * nef;
  double* OMA = red + NMAX;                     // [9][NMAX] Berry curvature per band
  double* F3A = OMA + 9 * NMAX;                 // [9][NMAX] Kubo LC tensor per band
  double* PW = F3A + 9 * NMAX;                  // velo / ohc pair sums: [4 waves][24]
                                                // (Hermitian V) or [36] (general V)
  double* acc = PW + 96;
  uint8_t* win = reinterpret_cast<uint8_t*>(acc + L.total);  // [n0][n] lo, hi
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int n2 = n * n;
  const bool finiteT = b.T_kelvin > 1e-2;
  const bool doAHC = (b.jobs & MEP_JOB_AHC) != 0;
  const bool doK = (b.jobs & MEP_JOB_KUBO) != 0;
  const bool doNjob = (b.jobs & MEP_JOB_MEP) != 0;
  const bool pmode = doAHC && (nhw * nef == 1);
  // V(n, m) = conj(V(m, n)) holds for V built here without a position
  // operator (with one, the reference's connection is not Hermitian; V from
  // a caller is taken as it is)
  const bool herm = b.v_herm != 0;
  const double tol = b.kubo_tol;
  for (int64_t c = t; c < L.total; c += BS_THREADS) acc[c] = 0.0;
  // V planes, r, E and fd are staged for indices < n only: their padding
  // (up to NMAX) stays zero for every k, so loads need no bounds branch
  for (int c = t; c < 7 * PL + (1 + nef) * NMAX; c += BS_THREADS) smem[c] = 0.0;

  const zd wq = zd{b.hw[0], b.eta};
  const zd w2 = zmul(wq, wq);

  // FUSED: U and the packed dH_a of a k-point, loaded into registers during
  // the last (light) phase of the previous k-point
  constexpr int NU = (NMAX * NMAX + BS_THREADS - 1) / BS_THREADS;
  constexpr int ND = (NMAX * (NMAX + 1) / 2 + BS_THREADS - 1) / BS_THREADS;
  const int np = n * (n + 1) / 2;
  zd ut[FUSED ? NU : 1], dt[3][FUSED ? ND : 1];
  auto fetch = [&](int64_t kk) {
    const int tt = threadIdx.x;
    const bool ok = kk < nk;
    const zd* Uk = U + (ok ? kk : 0) * (int64_t)n2;
    const zd* X = Hk + (ok ? kk : 0) * strideHk;
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int idx = tt + BS_THREADS * u;
      ut[u] = (ok && idx < n2) ? Uk[idx] : zd{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < ND; ++u) {
      const int p = tt + BS_THREADS * u;
#pragma unroll
      for (int a = 0; a < 3; ++a)
        dt[a][u] = (ok && p < np) ? X[(int64_t)(1 + a) * np + p] : zd{0.0, 0.0};
    }
  };
  if constexpr (FUSED) fetch(blockIdx.x);

  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    // the thread index is made opaque per k-point: otherwise the compiler
    // hoists the lane-dependent indexes and masks of every unrolled loop out
    // of this one and spills them (the wave index stays scalar)
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    const int lane = t & 63, wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const double wk = w ? w[k] : 1.0;
    __syncthreads();
    // ---- stage V (planar [a][i + n j]) and E ------------------------------
    if constexpr (FUSED) {
      // V_a = U^H (dH_a U): U over the tables' region, dH_a (both halves of
      // the packed upper triangle) in the V planes, T_a = dH_a U written over
      // dH_a, then V_a from the accumulators into the V planes.  Rows and
      // columns >= n stay zero in every plane, so no operand load is guarded.
      zd* Us = reinterpret_cast<zd*>(Rs);
      for (int c = t; c < PL; c += BS_THREADS) stz(Us + c, zd{0.0, 0.0});
      __syncthreads();  // U's plane is zero
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        const int idx = t + BS_THREADS * u;
        if (idx < n2) stz(Us + idx % n + LD * (idx / n), ut[u]);
      }
#pragma unroll
      for (int u = 0; u < ND; ++u) {
        const int p = t + BS_THREADS * u;
        if (p >= np) continue;
        // element p of the packed upper triangle: (i, j), i <= j, p = i + j(j+1)/2
        int j = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
        if ((j + 1) * (j + 2) / 2 <= p) ++j;
        if (j * (j + 1) / 2 > p) --j;
        const int i = p - j * (j + 1) / 2;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          stz(Vs + a * PL + j + LD * i, zconj(dt[a][u]));
          stz(Vs + a * PL + i + LD * j, dt[a][u]);  // last: the diagonal keeps the value as stored
        }
      }
      __syncthreads();
      // the 4 waves own the 2 x 2 output tiles of 16 x 16 (n > 16: all busy)
      const int r0 = 16 * (wv & 1), c0 = 16 * (wv >> 1);
      const int ar = r0 + (lane & 15), bc = c0 + (lane & 15), kq = lane >> 4;
      mep_v4d q1[3], q2[3], q3[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) q1[a] = q2[a] = q3[a] = mep_v4d{0, 0, 0, 0};
#pragma unroll
      for (int k0 = 0; k0 < NMAX; k0 += 4) {
        if (k0 >= n) break;
        const int m = k0 + kq;
        const zd bu = ldz(Us + m + LD * bc);
        const double bsum = bu.re + bu.im;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const zd x = ldz(Vs + a * PL + ar + LD * m);
          q1[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.re, bu.re, q1[a], 0, 0, 0);
          q2[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.im, bu.im, q2[a], 0, 0, 0);
          q3[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.re + x.im, bsum, q3[a], 0, 0, 0);
        }
      }
      __syncthreads();  // every wave has read dH
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = r0 + kq + 4 * v;
          stz(Vs + a * PL + row + LD * bc, zd{q1[a][v] - q2[a][v], q3[a][v] - q1[a][v] - q2[a][v]});
        }
      __syncthreads();
#pragma unroll
      for (int a = 0; a < 3; ++a) q1[a] = q2[a] = q3[a] = mep_v4d{0, 0, 0, 0};
#pragma unroll
      for (int k0 = 0; k0 < NMAX; k0 += 4) {
        if (k0 >= n) break;
        const int m = k0 + kq;
        const zd uu = ldz(Us + m + LD * ar);  // A(row, m) = conj(U(m, row))
        const double us = uu.re - uu.im;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const zd y = ldz(Vs + a * PL + m + LD * bc);
          q1[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(uu.re, y.re, q1[a], 0, 0, 0);
          q2[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(-uu.im, y.im, q2[a], 0, 0, 0);
          q3[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(us, y.re + y.im, q3[a], 0, 0, 0);
        }
      }
      __syncthreads();  // every wave has read T and U
      // V (zero in rows and columns >= n, as T and U are) into the planes
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = r0 + kq + 4 * v;
          stz(Vs + a * PL + row + LD * bc, zd{q1[a][v] - q2[a][v], q3[a][v] - q1[a][v] - q2[a][v]});
        }
      // the r table, E and the occupations are zero again, E in its slot
      const int zn = PL + NMAX + nef * NMAX;
      for (int c = t; c < zn; c += BS_THREADS) {
        const int e = c - PL;
        Rs[c] = (e >= 0 && e < n) ? E[k * n + e] : 0.0;
      }
    } else {
    const zd* Vk = V + k * 3 * (int64_t)n2;
    {
      // every load of the k-point is issued before the first LDS store
      constexpr int JS = BS_THREADS / NMAX, NJ = NMAX / JS;
      const int i = t % NMAX, jb = t / NMAX;
      zd vt[3][NJ];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj) {
          const int j = jb + JS * jj;
          vt[a][jj] = (i < n && j < n) ? Vk[a * n2 + i + n * j] : zd{0.0, 0.0};
        }
      const double et = (t < n) ? E[k * n + t] : 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj) {
          const int j = jb + JS * jj;
          if (i < n && j < n) stz(Vs + a * PL + i + LD * j, vt[a][jj]);
        }
      if (t < n) Es[t] = et;
    }
    }
    __syncthreads();
    // ---- r(m, c) = 1 / (E_c - E_m); occupations ---------------------------
    {
      // one division per pair m < c: r(c, m) = -r(m, c) exactly in IEEE
      // arithmetic; the diagonal stays 0 (zeroed once with the padding)
      const int npair = n * (n - 1) / 2;
      for (int pi = t; pi < npair; pi += BS_THREADS) {
        int c = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)pi)) * 0.5f);  // p = c(c-1)/2 + m
        if (c * (c - 1) / 2 > pi) --c;
        if ((c + 1) * c / 2 <= pi) ++c;
        const int m = pi - c * (c - 1) / 2;
        const double v = 1.0 / (Es[c] - Es[m]);
        Rs[m + LD * c] = v;
        Rs[c + LD * m] = -v;
      }
    }
    if (doNjob)
      for (int idx = t; idx < nval * NMAX; idx += BS_THREADS) {
        // mep_niu's guard |(E0-En)^2 (E0-Em)| > kubo_tol: energies are
        // sorted, so the excluded m of (n0, n) form one range around n0
        const int nn = idx % NMAX, n0 = idx / NMAX;
        if (nn >= n || nn == n0) continue;
        const double E0 = Es[n0], dn = E0 - Es[nn];
        const double dn2 = dn * dn;
        int lo = n0, hi = n0;
        while (lo > 0 && !(fabs(dn2 * (E0 - Es[lo - 1])) > tol)) --lo;
        while (hi < n - 1 && !(fabs(dn2 * (E0 - Es[hi + 1])) > tol)) ++hi;
        win[2 * (n0 * NMAX + nn)] = (uint8_t)lo;
        win[2 * (n0 * NMAX + nn) + 1] = (uint8_t)hi;
      }
    for (int idx = t; idx < nef * n; idx += BS_THREADS) {
      if (b.fd_in) {
        fd[idx] = b.fd_in[k * (int64_t)nef * n + idx];
      } else {
        const int ef = (nef == 1) ? 0 : idx % nef, nb = (nef == 1) ? idx : idx / nef;
        const double e = Es[nb], f = b.ef[ef];
        fd[idx] = finiteT ? 1.0 / (1.0 + exp((e - f) / Tsmear)) : ((e < f) ? 1.0 : 0.0);
      }
    }
    __syncthreads();

    if (!herm) {
      // ==== pair phase, general V (position operator: the reference's
      // connection makes V non-Hermitian, so V(n, n0) is read as stored).
      // Lane (p0, pp) of each wave walks n = pp, pp + P, ...: wave 0 the
      // Berry curvature of p0, wave 1 its local (LC) terms, waves 2-3 the
      // velo / ohc pair sums at the single (hw, E_F).
      constexpr int PARTS = 64 / NMAX;
      const int p0 = lane % NMAX, pp = lane / NMAX;
      if (wv == 0) {
        double om[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) om[q] = 0.0;
        if (doAHC)
          for (int nn = pp; nn < n; nn += PARTS) {
            zd x[3], y[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              x[a] = ldz(Vs + a * PL + p0 + LD * nn);  // V_a(n0, n)
              y[a] = ldz(Vs + a * PL + nn + LD * p0);  // V_a(n, n0)
            }
            const double r = Rs[nn + LD * p0];  // 1 / (E0 - En)
            const double r2 = r * r;
            // Omega_n0,ij = -2 Im(V_i(n0,n) V_j(n,n0)) / (En - E0)^2
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
              for (int i = 0; i < 3; ++i) om[i + 3 * j] -= 2.0 * zmul(x[i], y[j]).im * r2;
          }
#pragma unroll
        for (int q = 0; q < 9; ++q) om[q] = (NMAX == 32) ? swap32_sum(om[q], om[q]) : rows4_sum(om[q]);
        if (pp == 0)
#pragma unroll
          for (int q = 0; q < 9; ++q) OMA[q * NMAX + p0] = om[q];
      } else if (wv == 1) {
        double f3[9], f3m[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) f3[q] = f3m[q] = 0.0;
        const bool doN = doNjob && p0 < n && p0 < nval;
        if (doK || doNjob) {
          const double E0 = Es[p0];
          zd z0[3];
#pragma unroll
          for (int l = 0; l < 3; ++l) z0[l] = ldz(Vs + l * PL + p0 + LD * p0);
          for (int nn = pp; nn < n; nn += PARTS) {
            zd x[3], y[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              x[a] = ldz(Vs + a * PL + p0 + LD * nn);
              y[a] = ldz(Vs + a * PL + nn + LD * p0);
            }
            const double r = Rs[nn + LD * p0];  // 0 for nn == n0: adds nothing
            const double r3 = r * r * r;
            // F3_ij = eps_jkl Re(V_i(n0,n) V_k(n,n0) V_l(n0,n0)) / (E0-En)^3
            zd cr[3];
            cr[0] = zsub(zmul(y[1], z0[2]), zmul(y[2], z0[1]));
            cr[1] = zsub(zmul(y[2], z0[0]), zmul(y[0], z0[2]));
            cr[2] = zsub(zmul(y[0], z0[1]), zmul(y[1], z0[0]));
            const double dn = E0 - Es[nn];
            const bool keep3 = doN && (fabs(dn * dn * dn) > tol);
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
              for (int i = 0; i < 3; ++i) {
                const double v = zmul(x[i], cr[j]).re * r3;
                f3[i + 3 * j] += v;
                if (keep3) f3m[i + 3 * j] += v;
              }
          }
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) {
          f3[q] = (NMAX == 32) ? swap32_sum(f3[q], f3[q]) : rows4_sum(f3[q]);
          f3m[q] = (NMAX == 32) ? swap32_sum(f3m[q], f3m[q]) : rows4_sum(f3m[q]);
        }
        if (pp == 0) {
#pragma unroll
          for (int q = 0; q < 9; ++q) F3A[q * NMAX + p0] = f3[q];
          if (doN)
#pragma unroll
            for (int q = 0; q < 9; ++q) acc[L.nlc + q + 9 * p0] += wk * f3m[q];
        }
      } else if (pmode && wv == 2) {
        // components q = i + 3j:
        //   velo_q += (f0 - fn) Im(V_i(n0,n) V_j(n,n0)) / ((E0-En)^2 - (hw + i eta)^2)
        //   ohc_q  += i (f0 - fn) V_i(n,n0) V_j(n0,n) / ((E0-En)(E0-En - hw - i eta))
        double pv[36];
#pragma unroll
        for (int u = 0; u < 36; ++u) pv[u] = 0.0;
        const double E0 = Es[p0];
        for (int nn = pp; nn < n; nn += PARTS) {
          zd x[3], y[3];
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            x[a] = ldz(Vs + a * PL + p0 + LD * nn);
            y[a] = ldz(Vs + a * PL + nn + LD * p0);
          }
          const double dE = E0 - Es[nn];
          const double dfv = fd[p0] - fd[nn];
          const zd denv = zd{dE * dE - w2.re, -w2.im};
          const zd deno = zmul(zd{dE, 0.0}, zd{dE - wq.re, -wq.im});
          const double iv = 1.0 / zabs2(denv), io = 1.0 / zabs2(deno);
          const bool off = nn != p0;
          const zd rv = off ? zd{denv.re * iv * dfv, -denv.im * iv * dfv} : zd{0.0, 0.0};
          const zd ro = off ? zd{deno.re * io * dfv, -deno.im * io * dfv} : zd{0.0, 0.0};
#pragma unroll
          for (int q = 0; q < 9; ++q) {
            const int i = q % 3, j = q / 3;
            const double vim = zmul(x[i], y[j]).im;
            pv[2 * q] += vim * rv.re;
            pv[2 * q + 1] += vim * rv.im;
            const zd o = zmuli(zmul(zmul(y[i], x[j]), ro));
            pv[18 + 2 * q] += o.re;
            pv[19 + 2 * q] += o.im;
          }
        }
#pragma unroll
        for (int u = 0; u < 36; ++u) pv[u] = rows4_sum(row16_sum(pv[u]));
        if (lane == 0)
#pragma unroll
          for (int u = 0; u < 36; ++u) PW[u] = pv[u];
      }
    } else {
    // ==== pair phase: wave w owns the bands n0 = 8w .. 8w + 7; lane
    // (n0 = 8w + lane % 8, part = lane / 8) walks n = part, part + 8, ...
    // V is Hermitian, so V_a(n, n0) = conj(x_a) with x_a = V_a(n0, n), and
    // every pair sum is built from P_ij = x_i conj(x_j) (i < j) and |x_i|^2.
    // r = 1 / (E0 - En) is 0 for n = n0 and in the padding.
    if (wv * 8 < n) {
      const int n0 = wv * 8 + (lane & 7), part = lane >> 3;
      const bool doN = doNjob && n0 < nval && n0 < n;
      const double E0 = Es[n0];
      double om[3] = {0.0, 0.0, 0.0};        // sum r^2 Im P_ij
      double gd[3] = {0.0, 0.0, 0.0};        // sum r^3 |x_i|^2
      double gdm[3] = {0.0, 0.0, 0.0};       // the same over Niu-kept n
      zd go[3], gom[3];                      // sum r^3 P_ij (i < j), kept
      double ps[24];                         // velo / ohc pair sums
#pragma unroll
      for (int c = 0; c < 3; ++c) go[c] = gom[c] = zd{0.0, 0.0};
#pragma unroll
      for (int c = 0; c < 24; ++c) ps[c] = 0.0;
      const double f0 = pmode ? fd[n0] : 0.0;
      for (int nn = part; nn < n; nn += 8) {
        zd x[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) x[a] = ldz(Vs + a * PL + n0 + LD * nn);
        const double r = Rs[nn + LD * n0];
        const double r2 = r * r, r3 = r2 * r;
        zd P[3];
        double d[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) d[a] = x[a].re * x[a].re + x[a].im * x[a].im;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const zd xi = x[c == 2 ? 1 : 0], xj = x[c == 0 ? 1 : 2];
          P[c] = zd{xi.re * xj.re + xi.im * xj.im, xi.im * xj.re - xi.re * xj.im};
        }
        if (doAHC)
#pragma unroll
          for (int c = 0; c < 3; ++c) om[c] += r2 * P[c].im;
        if (doK || doNjob) {
          const double dn = E0 - Es[nn];
          const double r3m = (doN && fabs(dn * dn * dn) > tol) ? r3 : 0.0;
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            gd[a] += r3 * d[a];
            gdm[a] += r3m * d[a];
          }
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            go[c] = zd{go[c].re + r3 * P[c].re, go[c].im + r3 * P[c].im};
            gom[c] = zd{gom[c].re + r3m * P[c].re, gom[c].im + r3m * P[c].im};
          }
        }
        if (pmode) {
          const double dE = E0 - Es[nn];
          const double dfv = f0 - fd[nn];
          const zd denv = zd{dE * dE - w2.re, -w2.im};
          const zd deno = zmul(zd{dE, 0.0}, zd{dE - wq.re, -wq.im});
          // one reciprocal for both denominators
          const double av = zabs2(denv), ao = zabs2(deno);
          const double ia = rcp_nr(av * ao);
          const double iv = ia * ao, io = ia * av;
          // selects, not a branch, drop n == n0 (where 1/deno is not finite)
          const bool off = nn != n0;
          const zd rv = off ? zd{denv.re * iv * dfv, -denv.im * iv * dfv} : zd{0.0, 0.0};
          const zd ro = off ? zd{deno.re * io * dfv, -deno.im * io * dfv} : zd{0.0, 0.0};
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            ps[2 * c] += rv.re * P[c].im;
            ps[2 * c + 1] += rv.im * P[c].im;
            ps[6 + 2 * c] += ro.re * P[c].re;
            ps[7 + 2 * c] += ro.re * P[c].im;
            ps[12 + 2 * c] += ro.im * P[c].re;
            ps[13 + 2 * c] += ro.im * P[c].im;
            ps[18 + c] += ro.re * d[c];
            ps[21 + c] += ro.im * d[c];
          }
        }
      }
      // local terms of n0: F3_ij = Re(z_l G_ik - z_k G_il), (j, k, l) cyclic,
      // z = V(n0, n0), G_ik = sum_n r^3 x_i conj(x_k) (Hermitian in i, k)
      double v24[24];
      {
        zd z[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) z[a] = ldz(Vs + a * PL + n0 + LD * n0);
        auto G = [&](const double* gd_, const zd* go_, int i, int k) {
          if (i == k) return zd{gd_[i], 0.0};
          const zd g = go_[i + k - 1];
          return i < k ? g : zd{g.re, -g.im};
        };
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const int k = (j + 1) % 3, l = (j + 2) % 3;
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const zd a = G(gd, go, i, k), bb = G(gd, go, i, l);
            const zd am = G(gdm, gom, i, k), bm = G(gdm, gom, i, l);
            v24[i + 3 * j] = (z[l].re * a.re - z[l].im * a.im) - (z[k].re * bb.re - z[k].im * bb.im);
            v24[9 + i + 3 * j] =
                (z[l].re * am.re - z[l].im * am.im) - (z[k].re * bm.re - z[k].im * bm.im);
          }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          v24[18 + c] = om[c];
          v24[21 + c] = 0.0;
        }
      }
      // sums over the 8 parts (lane bits 3-5), scattered: lane (n0, g) ends
      // with the slots 12 g5 + 6 g4 + 3 g3 + s, s < 3, of its n0
      double v3[3];
      scatter_sum24(v24, v3, lane);
      {
        const int base = 12 * ((lane >> 5) & 1) + 6 * ((lane >> 4) & 1) + 3 * ((lane >> 3) & 1);
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const int slot = base + s;
          if (slot < 9) {
            F3A[slot * NMAX + n0] = v3[s];
          } else if (slot < 18) {
            if (doN) acc[L.nlc + (slot - 9) + 9 * n0] += wk * v3[s];
          } else if (slot < 21) {  // Omega_ij = -2 v, Omega_ji = 2 v for pair (i, j)
            const int c = slot - 18, i = (c == 2) ? 1 : 0, j = (c == 0) ? 1 : 2;
            OMA[(i + 3 * j) * NMAX + n0] = -2.0 * v3[s];
            OMA[(j + 3 * i) * NMAX + n0] = 2.0 * v3[s];
          } else {                 // Omega_ii = 0
            OMA[4 * (slot - 21) * NMAX + n0] = 0.0;
          }
        }
      }
      if (pmode) {
        scatter_sum24(ps, v3, lane);
        const int base = 12 * ((lane >> 5) & 1) + 6 * ((lane >> 4) & 1) + 3 * ((lane >> 3) & 1);
#pragma unroll
        for (int s = 0; s < 3; ++s) v3[s] = group8_sum(v3[s]);
        if ((lane & 7) == 0)
#pragma unroll
          for (int s = 0; s < 3; ++s) PW[wv * 24 + base + s] = v3[s];
      }
    } else if (pmode && lane < 24) {
      PW[wv * 24 + lane] = 0.0;
    }
    }
    // (the Hall and optical conductivities of several (hw, E_F) items run in
    // hwsum_kernel after this kernel)

    // ==== itinerant (IC) terms on MFMA ======================================
    // Wave item: one 16 x 16 block (tr, tc) of all three C_j when the
    // matrices hold several blocks (NT = 2), one C_j when they hold one.
    // The three products share their operands, V_c(row, m) and
    // B_c(m, col) = V_c(m, col) r(m, col), and give nine independent MFMA
    // chains.  Loads are unconditional: LDS rows and columns >= n are zero
    // and r is zero on the diagonal, so padded entries add nothing.
    if (doK || doNjob) {
      constexpr int NB = NT * NT;               // 16 x 16 blocks per matrix
      constexpr int JN = (NB >= 4) ? 3 : 1;     // components per wave item
      constexpr int NITEM = NB * (3 / JN);
      constexpr int NP = (JN == 3) ? 3 : 2;      // velocity planes an item reads
      for (int item = wv; item < NITEM; item += 4) {
        const int blk = item % NB, jsel = item / NB;  // jsel: the component when JN == 1
        const int tr = blk / NT, tc = blk % NT;
        const int r0 = 16 * tr, c0 = 16 * tc;
        const int ar = r0 + (lane & 15), bc = c0 + (lane & 15), kq = lane >> 4;
        // plane of operand slot c; component jj pairs slot ks(jj) (its k)
        // with slot ls(jj) (its l)
        int pl[NP];
#pragma unroll
        for (int c = 0; c < NP; ++c) pl[c] = (JN == 3) ? c : (jsel + 1 + c) % 3;
        auto ks = [](int jj) { return (JN == 3) ? (jj + 1) % 3 : 0; };
        auto ls = [](int jj) { return (JN == 3) ? (jj + 2) % 3 : 1; };
        // C = sum_m A(., m) B(m, .) with three real products per complex
        // one: P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi);
        // Re C = P1 - P2, Im C = P3 - P1 - P2
        mep_v4d p1[JN], p2[JN], p3[JN];
#pragma unroll
        for (int jj = 0; jj < JN; ++jj) p1[jj] = p2[jj] = p3[jj] = mep_v4d{0, 0, 0, 0};
#pragma unroll
        for (int k0 = 0; k0 < NMAX; k0 += 4) {
          const int m = k0 + kq;
          const double rr = Rs[m + LD * bc];
          zd a[NP], s[NP];
          double as[NP], ss[NP];
#pragma unroll
          for (int c = 0; c < NP; ++c) {
            a[c] = ldz(Vs + pl[c] * PL + ar + LD * m);
            s[c] = zscale(ldz(Vs + pl[c] * PL + m + LD * bc), rr);
            as[c] = a[c].re + a[c].im;
            ss[c] = s[c].re + s[c].im;
          }
          // C_j += V_k B_l, then C_j -= V_l B_k   ((j, k, l) cyclic)
#pragma unroll
          for (int jj = 0; jj < JN; ++jj) {
            const int kk = ks(jj), ll = ls(jj);
            p1[jj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kk].re, s[ll].re, p1[jj], 0, 0, 0);
            p2[jj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kk].im, s[ll].im, p2[jj], 0, 0, 0);
            p3[jj] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[kk], ss[ll], p3[jj], 0, 0, 0);
          }
#pragma unroll
          for (int jj = 0; jj < JN; ++jj) {
            const int kk = ks(jj), ll = ls(jj);
            p1[jj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ll].re, -s[kk].re, p1[jj], 0, 0, 0);
            p2[jj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ll].im, -s[kk].im, p2[jj], 0, 0, 0);
            p3[jj] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[ll], -ss[kk], p3[jj], 0, 0, 0);
          }
        }
        // contraction with A_i(n0, n) = V_i(n0, n) r(n, n0)^2, n0 = column;
        // r(n, n0) = 0 for n = n0 and outside n, so those rows add nothing.
        // f2[i + 3 jj]: component (i, j) with j = jj (JN == 3) or jsel
        const int n0 = bc;
        const bool colok = n0 < n;
        const bool doN = doNjob && colok && n0 < nval;
        double f2[3 * JN], f2m[3 * JN];
#pragma unroll
        for (int q = 0; q < 3 * JN; ++q) f2[q] = f2m[q] = 0.0;
        // mep_niu keeps, for (n0, n), the m outside one window [lo, hi]:
        // narrow windows take -Re[A C] here and add back the excluded terms
        // below; wide windows sum the kept terms directly below
        auto wide = [&](int lo, int hi) { return 2 * (hi - lo + 1) > n; };
        // rows whose window holds more than n0 itself (the others need no
        // second pass) and is not the whole band range
        bool wterm[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = r0 + kq + 4 * v;
          const double rr = Rs[row + LD * n0];
          const double rr2 = rr * rr;
          bool narrow = false;
          wterm[v] = false;
          if (doN && row < n && row != n0) {
            const int lo = win[2 * (n0 * NMAX + row)], hi = win[2 * (n0 * NMAX + row) + 1];
            const bool all = lo == 0 && hi == n - 1;
            narrow = !all && !wide(lo, hi);
            wterm[v] = !all && !(lo == n0 && hi == n0);
          }
          zd ai[3];
#pragma unroll
          for (int i = 0; i < 3; ++i) ai[i] = zscale(ldz(Vs + i * PL + n0 + LD * row), rr2);
#pragma unroll
          for (int jj = 0; jj < JN; ++jj) {
            const zd cv = zd{p1[jj][v] - p2[jj][v], p3[jj][v] - p1[jj][v] - p2[jj][v]};
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              const double c = ai[i].re * cv.re - ai[i].im * cv.im;
              f2[i + 3 * jj] -= c;
              f2m[i + 3 * jj] -= narrow ? c : 0.0;
            }
          }
        }
        // window terms, after the accumulators are dead
        if (doN)
#pragma unroll 1
          for (int v = 0; v < 4; ++v) {
            if (!wterm[v]) continue;
            const int row = r0 + kq + 4 * v;
            const int lo = win[2 * (n0 * NMAX + row)], hi = win[2 * (n0 * NMAX + row) + 1];
            const bool wd = wide(lo, hi);
            const double rr = Rs[row + LD * n0];
            const double rr2 = rr * rr;
            zd ai[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) ai[i] = zscale(ldz(Vs + i * PL + n0 + LD * row), rr2);
#pragma unroll
            for (int jj = 0; jj < JN; ++jj) {
              const zd* Vk_ = Vs + pl[ks(jj)] * PL;
              const zd* Vl_ = Vs + pl[ls(jj)] * PL;
              auto term = [&](int mm) {
                const double rm = Rs[mm + LD * n0];
                const zd bl = zscale(Vl_[mm + LD * n0], rm), bk = zscale(Vk_[mm + LD * n0], rm);
                return zsub(zmul(Vk_[row + LD * mm], bl), zmul(Vl_[row + LD * mm], bk));
              };
              zd x = zd{0.0, 0.0};
              if (wd) {  // kept terms: f2m -= Re(A x)
#pragma unroll 1
                for (int mm = 0; mm < lo; ++mm) x = zadd(x, term(mm));
#pragma unroll 1
                for (int mm = hi + 1; mm < n; ++mm) x = zadd(x, term(mm));
                x = zd{-x.re, -x.im};
              } else {   // excluded terms: f2m += Re(A x)
#pragma unroll 1
                for (int mm = lo; mm <= hi; ++mm)
                  if (mm != n0) x = zadd(x, term(mm));
              }
#pragma unroll
              for (int i = 0; i < 3; ++i) f2m[i + 3 * jj] += ai[i].re * x.re - ai[i].im * x.im;
            }
          }
        // sum the 4 row groups of the block (lanes l, l^16, l^32, l^48)
#pragma unroll
        for (int q = 0; q < 3 * JN; ++q) {
          f2[q] = rows4_sum(f2[q]);
          f2m[q] = rows4_sum(f2m[q]);
        }
        const int jb = (JN == 3) ? 0 : jsel;  // component of f2[3 * 0 ..]
        if (kq == 0 && colok)
#pragma unroll
          for (int q = 0; q < 3 * JN; ++q) QF[(tr * 9 + q + 3 * jb) * NMAX + n0] = f2m[q];
        // kubo_mep_IC: f-weighted sum over the block's 16 bands
        if (doK)
          for (int ef = 0; ef < nef; ++ef) {
            const double f = (colok && kq == 0) ? fd[ef + nef * n0] : 0.0;
#pragma unroll
            for (int q = 0; q < 3 * JN; ++q) {
              const int i = q % 3, j = q / 3 + jb;
              const double s = row16_sum(f * f2[q]);
              if (lane == 0) TS[((j * NB + blk) * 3 + i) * nef + ef] = s;
            }
          }
      }
    }
    __syncthreads();
    if constexpr (FUSED) fetch(k + gridDim.x);  // in flight during the sums below
    // ---- sums over the bands n0 and over the waves' partial records --------
    // Ne, kubo_ahc_tens (-sum f Omega) and kubo_mep_LC for every E_F
    // for every E_F; item (c, ef) is summed by 8 lanes, lane g taking the
    // bands m0 = g, g + 8, ...
    for (int base = 0; base < 8 * 19 * nef; base += BS_THREADS) {
      const int it = base + t, item = it >> 3, g = it & 7;
      const int c = item % 19, ef = item / 19;
      const bool valid = item < 19 * nef &&
                         (c == 0 || (c < 10 && doAHC) || (c >= 10 && doK));
      double s = 0.0;
      if (valid) {
        const double* src = (c == 0) ? nullptr : (c < 10) ? OMA + (c - 1) * NMAX : F3A + (c - 10) * NMAX;
        for (int m0 = g; m0 < n; m0 += 8) s += fd[ef + nef * m0] * (src ? src[m0] : 1.0);
      }
      s = group8_sum(s);  // the loop bounds are uniform over the wave
      if (valid && g == 0) {
        if (c == 0)
          acc[L.Ne + ef] += wk * s;
        else if (c < 10)
          acc[L.ahc + (c - 1) + 9 * ef] -= wk * s;
        else
          acc[L.klc + (c - 10) + 9 * ef] += wk * s;
      }
    }
    // velo_ahc_tens / kubo_ohc_tens at the single (hw, E_F): the last 18
    // threads, o = q (velo) and 9 + q (ohc)
    if (pmode && t >= BS_THREADS - 18) {
      const int o = t - (BS_THREADS - 18);  // 0..8 velo, 9..17 ohc
      const int q = o % 9, i = q % 3, j = q / 3, c = i + j - 1;
      auto u = [&](int e) { return ((PW[e] + PW[24 + e]) + PW[48 + e]) + PW[72 + e]; };
      zd val = zd{0.0, 0.0};
      if (!herm) {
        val = zd{PW[2 * o], PW[2 * o + 1]};
      } else if (o < 9) {  // velo: sum rv Im P_ij, antisymmetric in (i, j)
        if (i != j) val = (i < j) ? zd{u(2 * c), u(2 * c + 1)} : zd{-u(2 * c), -u(2 * c + 1)};
      } else {      // ohc: i sum ro conj(P_ij) = i conj(X - iY) for X = sum Re(ro) P, Y = sum Im(ro) P
        if (i == j) {
          val = zd{-u(21 + i), u(18 + i)};
        } else {
          const double xr = u(6 + 2 * c), xi = u(7 + 2 * c), yr = u(12 + 2 * c), yi = u(13 + 2 * c);
          val = (i < j) ? zd{xi - yr, xr + yi} : zd{-(xi + yr), xr - yi};
        }
      }
      double* dst = acc + (o < 9 ? L.velo : L.ohc) + 2 * q;
      dst[0] += wk * val.re;
      dst[1] += wk * val.im;
    }
    if (doK)
      for (int it = t; it < 9 * nef; it += BS_THREADS) {
        const int q = it % 9, ef = it / 9, i = q % 3, j = q / 3;
        double s = 0.0;
        for (int tt = 0; tt < NT * NT; ++tt) s += TS[((j * NT * NT + tt) * 3 + i) * nef + ef];
        acc[L.kic + q + 9 * ef] += wk * s;
      }
    if (doNjob)
      for (int it = t; it < 9 * nval; it += BS_THREADS) {
        const int q = it % 9, n0 = it / 9;
        double s = 0.0;
        for (int tr = 0; tr < NT; ++tr) s += QF[(tr * 9 + q) * NMAX + n0];
        acc[L.nic + q + 9 * n0] += wk * s;
      }
    // ---- Chern-Simons term (needs A and Omega) ------------------------------
    if ((doK || doNjob) && b.pos_op && A && Om) {
      __syncthreads();
      if (t < n) {
        const int64_t NN = (int64_t)n2;
        const zd* Ak = A + k * 3 * NN + 3 * ((int64_t)t + (int64_t)n * t);
        const zd* Ok = Om + k * 9 * NN + 9 * ((int64_t)t + (int64_t)n * t);
        zd omc[3];
        omc[0] = zsub(Ok[1 + 3 * 2], Ok[2 + 3 * 1]);
        omc[1] = zsub(Ok[2 + 3 * 0], Ok[0 + 3 * 2]);
        omc[2] = zsub(Ok[0 + 3 * 1], Ok[1 + 3 * 0]);
        zd dp = zd{0.0, 0.0};
#pragma unroll
        for (int a = 0; a < 3; ++a) dp = zadd(dp, zcmul(Ak[a], omc[a]));
        red[t] = dp.re;
      }
      __syncthreads();
      if (doK)
        for (int ef = t; ef < nef; ef += BS_THREADS) {
          double s = 0.0;
          for (int m0 = 0; m0 < n; ++m0) s += 0.5 * fd[ef + nef * m0] * red[m0];
          for (int i = 0; i < 3; ++i) acc[L.kcs + i + 3 * i + 9 * ef] += wk * s;
        }
      if (doNjob)
        for (int n0 = t; n0 < nval && n0 < n; n0 += BS_THREADS)
          for (int i = 0; i < 3; ++i) acc[L.ncs + i + 3 * i + 9 * n0] += wk * (0.5 * red[n0]);
    }
  }
  __syncthreads();
  // the record in HBM holds the velo / ohc block (gap entries from L.velo)
  // that this layout leaves out
  double* out = part + (int64_t)blockIdx.x * (L.total + gap);
  for (int64_t c = t; c < L.total; c += BS_THREADS) out[c < L.velo ? c : c + gap] += acc[c];
}

// ===========================================================================
// 32 < n <= 128 (no position operator).  Same sums and the same record
// layout; V and E are read from HBM / L2 instead of LDS, the workgroup's
// accumulator record lives in `part` (one writer per entry, fixed order).
//   per-band sums      thread (n0, part) over n = part, part + P, ...; parts
//                      combined through LDS
//   hw-resolved sums   a kernel of their own (hwsum_kernel below)
//   itinerant terms    C_j column tiles on FP64 MFMA, wave item (tc, j)
//                      accumulating over the row tiles in registers
// LDS: E n | fd nef*n | S 2*9*256 | F2, F2m 2*9*n | red 4*36 | windows
// ===========================================================================
static __device__ inline double rinv(const double* Es, int m, int c) {
  return (m == c) ? 0.0 : 1.0 / (Es[c] - Es[m]);  // 1 / (E_c - E_m)
}

constexpr int BW_CH = 4;  // four-deep MFMA steps per batch of operand loads
__global__ __launch_bounds__(256) void bandsum_wide_kernel(BandArgs b, double Tsmear, AccLayout L,
                                                           int64_t nk, const double* __restrict__ E,
                                                           const zd* __restrict__ V,
                                                           const double* __restrict__ w,
                                                           double* __restrict__ part,
                                                           zd* __restrict__ Bscr,
                                                           const zd* __restrict__ A,
                                                           const zd* __restrict__ Om) {
  extern __shared__ double smem[];
  const int n = b.n, nef = b.n_ef, nhw = b.n_hw, nval = b.n_val;
  double* Es = smem;
  double* fd = Es + n;
  double* S = fd + (int64_t)nef * n;  // [9][256]
  double* S2 = S + 9 * 256;           // [9][256]
  double* F2 = S2 + 9 * 256;          // [9][n]
  double* F2m = F2 + 9 * n;           // [9][n]
  double* red = F2m + 9 * n;          // [4][36]
  uint8_t* win = reinterpret_cast<uint8_t*>(red + 4 * 36);  // [nval][n] lo, hi
  const int t = threadIdx.x, lane = t & 63, wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int n2 = n * n;
  const bool finiteT = b.T_kelvin > 1e-2;
  const bool doAHC = (b.jobs & MEP_JOB_AHC) != 0;
  const bool doK = (b.jobs & MEP_JOB_KUBO) != 0;
  const bool doNjob = (b.jobs & MEP_JOB_MEP) != 0;
  const double tol = b.kubo_tol;
  double* acc = part + (int64_t)blockIdx.x * L.total;
  const int npad = n <= 64 ? 64 : 128, P = 256 / npad;
  const int p0 = t % npad, pp = t / npad;
  const bool pact = p0 < n;

  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    const double wk = w ? w[k] : 1.0;
    const zd* Vk = V + k * 3 * (int64_t)n2;
    __syncthreads();
    if (t < n) Es[t] = E[k * n + t];
    __syncthreads();
    if (doNjob)
      for (int idx = t; idx < nval * n; idx += 256) {
        const int nn = idx % n, n0 = idx / n;
        if (nn == n0) continue;
        const double E0 = Es[n0], dn = E0 - Es[nn];
        const double dn2 = dn * dn;
        int lo = n0, hi = n0;
        while (lo > 0 && !(fabs(dn2 * (E0 - Es[lo - 1])) > tol)) --lo;
        while (hi < n - 1 && !(fabs(dn2 * (E0 - Es[hi + 1])) > tol)) ++hi;
        win[2 * idx] = (uint8_t)lo;
        win[2 * idx + 1] = (uint8_t)hi;
      }
    for (int idx = t; idx < nef * n; idx += 256) {
      if (b.fd_in) {
        fd[idx] = b.fd_in[k * (int64_t)nef * n + idx];
      } else {
        const int ef = (nef == 1) ? 0 : idx % nef, nb = (nef == 1) ? idx : idx / nef;
        const double e = Es[nb], f = b.ef[ef];
        fd[idx] = finiteT ? 1.0 / (1.0 + exp((e - f) / Tsmear)) : ((e < f) ? 1.0 : 0.0);
      }
    }
    __syncthreads();
    for (int ef = t; ef < nef; ef += 256) {
      double sN = 0.0;
      for (int m0 = 0; m0 < n; ++m0) sN += fd[ef + nef * m0];
      acc[L.Ne + ef] += wk * sN;
    }
    // ---- Berry curvature per band -> kubo_ahc_tens --------------------------
    if (doAHC) {
      double om[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) om[q] = 0.0;
      if (pact)
        for (int nn = pp; nn < n; nn += P) {
          if (nn == p0) continue;
          zd x[3], y[3];
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            x[a] = Vk[a * (int64_t)n2 + p0 + (int64_t)n * nn];
            y[a] = Vk[a * (int64_t)n2 + nn + (int64_t)n * p0];
          }
          const double r = rinv(Es, nn, p0);  // 1 / (E_0 - E_n)
          const double r2 = r * r;
#pragma unroll
          for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) om[i + 3 * j] -= 2.0 * zmul(x[i], y[j]).im * r2;
        }
#pragma unroll
      for (int q = 0; q < 9; ++q) S[q * 256 + t] = om[q];
      __syncthreads();
      for (int it = t; it < 9 * nef; it += 256) {
        const int q = it % 9, ef = it / 9;
        double sA = 0.0;
        for (int m0 = 0; m0 < n; ++m0) {
          double o = 0.0;
          for (int u = 0; u < P; ++u) o += S[q * 256 + u * npad + m0];
          sA += fd[ef + nef * m0] * o;
        }
        acc[L.ahc + q + 9 * ef] -= wk * sA;
      }
      __syncthreads();
    }
    // ---- local (LC) MEP terms ----------------------------------------------
    if (doK || doNjob) {
      double f3[9], f3m[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) f3[q] = f3m[q] = 0.0;
      const bool doN = doNjob && pact && p0 < nval;
      if (pact) {
        const double E0 = Es[p0];
        zd z0[3];
#pragma unroll
        for (int l = 0; l < 3; ++l) z0[l] = Vk[l * (int64_t)n2 + p0 + (int64_t)n * p0];
        for (int nn = pp; nn < n; nn += P) {
          if (nn == p0) continue;
          zd x[3], y[3];
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            x[a] = Vk[a * (int64_t)n2 + p0 + (int64_t)n * nn];
            y[a] = Vk[a * (int64_t)n2 + nn + (int64_t)n * p0];
          }
          const double r = rinv(Es, nn, p0);  // 1 / (E_0 - E_n)
          const double r3 = r * r * r;
          zd cr[3];
          cr[0] = zsub(zmul(y[1], z0[2]), zmul(y[2], z0[1]));
          cr[1] = zsub(zmul(y[2], z0[0]), zmul(y[0], z0[2]));
          cr[2] = zsub(zmul(y[0], z0[1]), zmul(y[1], z0[0]));
          const double dn = E0 - Es[nn];
          const bool keep3 = doN && (fabs(dn * dn * dn) > tol);
#pragma unroll
          for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              const double v = zmul(x[i], cr[j]).re * r3;
              f3[i + 3 * j] += v;
              if (keep3) f3m[i + 3 * j] += v;
            }
        }
      }
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        S[q * 256 + t] = f3[q];
        S2[q * 256 + t] = f3m[q];
      }
      __syncthreads();
      if (doK)
        for (int it = t; it < 9 * nef; it += 256) {
          const int q = it % 9, ef = it / 9;
          double sL = 0.0;
          for (int m0 = 0; m0 < n; ++m0) {
            double o = 0.0;
            for (int u = 0; u < P; ++u) o += S[q * 256 + u * npad + m0];
            sL += fd[ef + nef * m0] * o;
          }
          acc[L.klc + q + 9 * ef] += wk * sL;
        }
      if (doNjob)
        for (int it = t; it < 9 * nval; it += 256) {
          const int q = it % 9, m0 = it / 9;
          double o = 0.0;
          for (int u = 0; u < P; ++u) o += S2[q * 256 + u * npad + m0];
          acc[L.nlc + q + 9 * m0] += wk * o;
        }
      __syncthreads();
    }
    // ---- itinerant (IC) terms on MFMA ----------------------------------------
    // B_a(m, n0) = V_a(m, n0) r(m, n0), r(m, n0) = 1 / (E_n0 - E_m) (0 on the
    // diagonal): one division per pair, into the workgroup's slot of Bscr,
    // so the MFMA loop below only loads (in chunks of BW_CH steps whose
    // loads are all issued before the chunk's products).  Bscr holds B
    // transposed (n0 + n m), so the 16 columns n0 of a tile load as one run
    if (doK || doNjob) {
      const int nt = (n + 15) / 16;
      const int lr = lane & 15, kq = lane >> 4;
      zd* Bw = Bscr + (int64_t)blockIdx.x * 3 * n2;
      for (int idx = t; idx < n2; idx += 256) {
        const int n0 = idx % n, m = idx / n;
        const double rr = rinv(Es, m, n0);
#pragma unroll
        for (int a = 0; a < 3; ++a)
          Bw[a * (int64_t)n2 + idx] = zscale(Vk[a * (int64_t)n2 + m + (int64_t)n * n0], rr);
      }
      __syncthreads();
      for (int item = wv; item < 3 * nt; item += 4) {
        const int j = item / nt, tc = item % nt;
        const int kk_ = (j + 1) % 3, ll_ = (j + 2) % 3;
        const zd* Vk_ = Vk + kk_ * (int64_t)n2;
        const zd* Vl_ = Vk + ll_ * (int64_t)n2;
        const zd* Bk_ = Bw + kk_ * (int64_t)n2;
        const zd* Bl_ = Bw + ll_ * (int64_t)n2;
        const int c0 = 16 * tc, n0 = c0 + lr;
        const bool colok = n0 < n;
        const bool doN = doNjob && colok && n0 < nval;
        double f2[3] = {0.0, 0.0, 0.0}, f2m[3] = {0.0, 0.0, 0.0};
        for (int tr = 0; tr < nt; ++tr) {
          const int r0 = 16 * tr, ar = r0 + lr;
          mep_v4d p1 = {0, 0, 0, 0}, p2 = {0, 0, 0, 0}, p3 = {0, 0, 0, 0};
          for (int m0 = 0; m0 < n; m0 += 4 * BW_CH) {
            zd a0[BW_CH], a1[BW_CH], b0[BW_CH], b1[BW_CH];
#pragma unroll
            for (int i = 0; i < BW_CH; ++i) {
              const int m = m0 + 4 * i + kq;
              a0[i] = a1[i] = b0[i] = b1[i] = zd{0.0, 0.0};
              if (m < n) {
                if (ar < n) {
                  a0[i] = Vk_[ar + (int64_t)n * m];
                  a1[i] = Vl_[ar + (int64_t)n * m];
                }
                if (colok) {
                  // uniform row base per step, one 32-bit lane offset
                  const int64_t rb = (int64_t)n * (m0 + 4 * i);
                  const int lo = n0 + n * kq;
                  b0[i] = (Bl_ + rb)[lo];
                  const zd bk = (Bk_ + rb)[lo];
                  b1[i] = zd{-bk.re, -bk.im};
                }
              }
            }
#pragma unroll
            for (int i = 0; i < BW_CH; ++i) {
              if (m0 + 4 * i >= n) break;  // uniform
              p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[i].re, b0[i].re, p1, 0, 0, 0);
              p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[i].im, b0[i].im, p2, 0, 0, 0);
              p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[i].re + a0[i].im, b0[i].re + b0[i].im, p3, 0, 0, 0);
              p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[i].re, b1[i].re, p1, 0, 0, 0);
              p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[i].im, b1[i].im, p2, 0, 0, 0);
              p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[i].re + a1[i].im, b1[i].re + b1[i].im, p3, 0, 0, 0);
            }
          }
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = r0 + kq + 4 * v;
            if (!colok || row >= n || row == n0) continue;
            const zd cv = zd{p1[v] - p2[v], p3[v] - p1[v] - p2[v]};
            const double rr = rinv(Es, row, n0);
            const double rr2 = rr * rr;
            zd ai[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              ai[i] = zscale(Vk[i * (int64_t)n2 + n0 + (int64_t)n * row], rr2);
              f2[i] -= ai[i].re * cv.re - ai[i].im * cv.im;
            }
            if (doN) {
              const int lo = win[2 * (n0 * n + row)], hi = win[2 * (n0 * n + row) + 1];
              if (lo == 0 && hi == n - 1) continue;  // every m excluded
              zd kept = zd{0.0, 0.0};
              auto term = [&](int mm) {
                const double rm = rinv(Es, mm, n0);
                const zd bl = zscale(Vl_[mm + (int64_t)n * n0], rm);
                const zd bk = zscale(Vk_[mm + (int64_t)n * n0], rm);
                return zsub(zmul(Vk_[row + (int64_t)n * mm], bl), zmul(Vl_[row + (int64_t)n * mm], bk));
              };
              if (2 * (hi - lo + 1) > n) {
                for (int mm = 0; mm < lo; ++mm) kept = zadd(kept, term(mm));
                for (int mm = hi + 1; mm < n; ++mm) kept = zadd(kept, term(mm));
              } else {
                zd ex = zd{0.0, 0.0};
                for (int mm = lo; mm <= hi; ++mm)
                  if (mm != n0) ex = zadd(ex, term(mm));
                kept = zsub(cv, ex);
              }
#pragma unroll
              for (int i = 0; i < 3; ++i) f2m[i] -= ai[i].re * kept.re - ai[i].im * kept.im;
            }
          }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          f2[i] += __shfl_xor(f2[i], 16, 64);
          f2[i] += __shfl_xor(f2[i], 32, 64);
          f2m[i] += __shfl_xor(f2m[i], 16, 64);
          f2m[i] += __shfl_xor(f2m[i], 32, 64);
        }
        if (kq == 0 && colok)
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            F2[(i + 3 * j) * n + n0] = f2[i];
            F2m[(i + 3 * j) * n + n0] = f2m[i];
          }
      }
      __syncthreads();
      if (doK)
        for (int it = t; it < 9 * nef; it += 256) {
          const int q = it % 9, ef = it / 9;
          double sI = 0.0;
          for (int m0 = 0; m0 < n; ++m0) sI += fd[ef + nef * m0] * F2[q * n + m0];
          acc[L.kic + q + 9 * ef] += wk * sI;
        }
      if (doNjob)
        for (int it = t; it < 9 * nval; it += 256) {
          const int q = it % 9, m0 = it / 9;
          acc[L.nic + q + 9 * m0] += wk * F2m[q * n + m0];
        }
    }
    // ---- Chern-Simons terms (position operator: A and Omega) ---------------
    // kubo_mep_CS / mep_niu_CS (kubo_mep.f90:16-47, mep_niu.f90:20-45):
    // red[n0] = Re sum_a conj(A_a(n0,n0)) omega_a(n0,n0), omega_c from Omega_ab
    if ((doK || doNjob) && b.pos_op && A && Om) {
      __syncthreads();
      for (int n0 = t; n0 < n; n0 += 256) {
        const int64_t NN = (int64_t)n2;
        const zd* Ak = A + k * 3 * NN + 3 * ((int64_t)n0 + (int64_t)n * n0);
        const zd* Ok = Om + k * 9 * NN + 9 * ((int64_t)n0 + (int64_t)n * n0);
        zd omc[3];
        omc[0] = zsub(Ok[1 + 3 * 2], Ok[2 + 3 * 1]);
        omc[1] = zsub(Ok[2 + 3 * 0], Ok[0 + 3 * 2]);
        omc[2] = zsub(Ok[0 + 3 * 1], Ok[1 + 3 * 0]);
        zd dp = zd{0.0, 0.0};
#pragma unroll
        for (int a = 0; a < 3; ++a) dp = zadd(dp, zcmul(Ak[a], omc[a]));
        red[n0] = dp.re;
      }
      __syncthreads();
      if (doK)
        for (int ef = t; ef < nef; ef += 256) {
          double s_ = 0.0;
          for (int m0 = 0; m0 < n; ++m0) s_ += 0.5 * fd[ef + nef * m0] * red[m0];
          for (int i = 0; i < 3; ++i) acc[L.kcs + i + 3 * i + 9 * ef] += wk * s_;
        }
      if (doNjob)
        for (int n0 = t; n0 < nval && n0 < n; n0 += 256)
          for (int i = 0; i < 3; ++i) acc[L.ncs + i + 3 * i + 9 * n0] += wk * (0.5 * red[n0]);
    }
  }
}

size_t bandsum_wide_lds_bytes(const BandArgs& b) {
  const size_t n = b.n;
  return sizeof(double) * (n + (size_t)b.n_ef * n + 2 * 9 * 256 + 2 * 9 * n + 4 * 36) +
         2 * (size_t)(b.n_val > 0 ? b.n_val : 1) * n;
}

// ===========================================================================
// velo_ahc_tens / kubo_ohc_tens for 32 < n <= 128: per k, the pairs (x, y),
// x != y, whose occupations differ at some E_F are compacted, in pair order,
// into LDS chunks of up to 256 records (dE, f_x - f_y per E_F, the 9
// Im(V_i(x,y) V_j(y,x)) and the 9 complex V_i(y,x) V_j(x,y)); thread
// (slot, split) walks every split-th record for its IPT items = (hw, E_F),
// accumulating in registers over all k; the splits are combined once at the
// end.  Items beyond 256 IPT are taken in further passes over the k-points.
// ===========================================================================
constexpr int HW_CH = 256;  // pair candidates per chunk (one per thread)
constexpr int HW_IPT = 2;   // items per thread, Hermitian V


// Records per pair (x, y) of differing occupation: dE, w_k (f_x - f_y) for
// every E_F, then the V products.  General V: 9 Im(V_i(x,y) V_j(y,x)) and 9
// complex V_i(y,x) V_j(x,y) (27 values).  V built here (Hermitian): with
// P_ij = x_i conj(x_j), x_a = V_a(x, y), only Im P_ij and P_ij (i < j) and
// |x_i|^2 (12 values), and the items accumulate sum rv Im P_ij,
// X = sum Re(ro) P_ij, Y = sum Im(ro) P_ij and the same of |x_i|^2.
template <bool HERM>
__host__ __device__ constexpr int hw_vals() { return HERM ? 12 : 27; }
template <bool HERM>
__host__ __device__ constexpr int hw_accs() { return HERM ? 24 : 36; }

// IPT items per thread: every record read from LDS feeds IPT items (the
// LDS return path, not the VALU, limits one item per thread).
template <bool HERM, int IPT>
__global__ __launch_bounds__(256, IPT > 1 ? 2 : 3) void hwsum_kernel(BandArgs b, double Tsmear, AccLayout L,
                                                       int64_t nk, const double* __restrict__ E,
                                                       const zd* __restrict__ V,
                                                       const double* __restrict__ w,
                                                       double* __restrict__ part) {
  constexpr int NV = hw_vals<HERM>(), NA = hw_accs<HERM>(), HA = NA / 2;
  extern __shared__ double smem[];
  const int n = b.n, nef = b.n_ef, nhw = b.n_hw;
  const int PS = 1 + nef + NV;  // doubles per record: dE, dfv[nef], the products
  double* Es = smem;
  double* fd = Es + n;
  int* cnt_w = reinterpret_cast<int*>(fd + (int64_t)nef * n);  // [4] per-wave counts
  double* st = fd + (int64_t)nef * n + 4;  // [HW_CH][PS]; reused for the split reduction
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int n2 = n * n;
  const bool finiteT = b.T_kelvin > 1e-2;
  const int items = nhw * nef;
  double* accp = part + (int64_t)blockIdx.x * L.total;
  for (int g0 = 0; g0 < items; g0 += 256 * IPT) {
    // this pass: IW items, IWT thread slots per split, each slot holding the
    // items slot + IWT u (u < IPT), S splits over the records
    const int IW = (items - g0 < 256 * IPT) ? items - g0 : 256 * IPT;
    const int IWT = (IW + IPT - 1) / IPT;
    const int S = 256 / IWT;
    const int slot = t % IWT, split = t / IWT;
    const bool act = split < S;
    bool iok[IPT];
    int ef[IPT];
    zd wh[IPT], wh2[IPT];
#pragma unroll
    for (int u = 0; u < IPT; ++u) {
      iok[u] = act && slot + IWT * u < IW;
      const int it = g0 + (iok[u] ? slot + IWT * u : 0);
      ef[u] = it / nhw;
      wh[u] = zd{b.hw[it % nhw], b.eta};
      wh2[u] = zmul(wh[u], wh[u]);
    }
    double acc[IPT][NA];
#pragma unroll
    for (int u = 0; u < IPT; ++u)
#pragma unroll
      for (int q = 0; q < NA; ++q) acc[u][q] = 0.0;
    for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
      const double wk = w ? w[k] : 1.0;
      const zd* Vk = V + k * 3 * (int64_t)n2;
      __syncthreads();
      if (t < n) Es[t] = E[k * n + t];
      __syncthreads();
      for (int idx = t; idx < nef * n; idx += 256) {
        if (b.fd_in) {
          fd[idx] = b.fd_in[k * (int64_t)nef * n + idx];
        } else {
          const int e_ = (nef == 1) ? 0 : idx % nef, nb = (nef == 1) ? idx : idx / nef;
          const double e = Es[nb], f = b.ef[e_];
          fd[idx] = finiteT ? 1.0 / (1.0 + exp((e - f) / Tsmear)) : ((e < f) ? 1.0 : 0.0);
        }
      }
      for (int c0 = 0; c0 < n2; c0 += HW_CH) {
        __syncthreads();  // fd ready; the previous chunk consumed
        const int pidx = c0 + t;
        const int x = pidx % n, y = pidx / n;
        // Hermitian V: one record per unordered pair x < y, summed for both
        // orders below; general V: every ordered pair
        bool keep = pidx < n2 && (HERM ? x < y : x != y);
        if (keep) {
          bool any = false;
          for (int e_ = 0; e_ < nef; ++e_) any |= fd[e_ + nef * x] != fd[e_ + nef * y];
          keep = any;  // pairs of equal occupation contribute exactly zero
        }
        const uint64_t ball = __ballot(keep);
        const int below = __popcll(ball & ((1ull << lane) - 1ull));
        if (lane == 0) cnt_w[wv] = __popcll(ball);
        __syncthreads();
        int base = 0;
        for (int u = 0; u < wv; ++u) base += cnt_w[u];
        const int cnt = cnt_w[0] + cnt_w[1] + cnt_w[2] + cnt_w[3];
        if (keep) {
          double* sp = st + (int64_t)(base + below) * PS;
          sp[0] = Es[x] - Es[y];
          for (int e_ = 0; e_ < nef; ++e_) sp[1 + e_] = wk * (fd[e_ + nef * x] - fd[e_ + nef * y]);
          double* pq = sp + 1 + nef;
          zd vxy[3];
#pragma unroll
          for (int a = 0; a < 3; ++a) vxy[a] = Vk[a * (int64_t)n2 + x + (int64_t)n * y];
          if constexpr (HERM) {
            // pq: Im P_c (3), P_c (3 complex), |x_a|^2 (3); c = (0,1), (0,2), (1,2)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const zd xi = vxy[c == 2 ? 1 : 0], xj = vxy[c == 0 ? 1 : 2];
              const zd P = zd{xi.re * xj.re + xi.im * xj.im, xi.im * xj.re - xi.re * xj.im};
              pq[c] = P.im;
              pq[3 + 2 * c] = P.re;
              pq[4 + 2 * c] = P.im;
              pq[9 + c] = zabs2(vxy[c]);
            }
          } else {
            zd vyx[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) vyx[a] = Vk[a * (int64_t)n2 + y + (int64_t)n * x];
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
              for (int i = 0; i < 3; ++i) {
                const int q = i + 3 * j;
                pq[q] = zmul(vxy[i], vyx[j]).im;
                const zd q2 = zmul(vyx[i], vxy[j]);
                pq[9 + 2 * q] = q2.re;
                pq[9 + 2 * q + 1] = q2.im;
              }
          }
        }
        __syncthreads();
        // one record (dE, the weight f of each item, the products pq) into
        // the items' sums
        auto consume = [&](const double dE, const double* fv, const double* pq) {
#pragma unroll
            for (int u = 0; u < IPT; ++u) {
              const double f = fv[u];
              if constexpr (HERM) {
                // the record of x < y stands for (x, y) and (y, x): (y, x) has
                // -dE, -f and conj P, so its velo term equals (x, y)'s and its
                // ohc weight is ro2 = -f / (dE (dE + w)).  One reciprocal for
                // the three denominators; f = 0 records (another E_F differs)
                // contribute exactly zero.
                // with p = dE - w, q = dE + w: denv = p q, deno = dE p and
                // the (y, x) denominator dE q, so R = f / (dE |p|^2 |q|^2)
                // gives rv = 2 f / denv = 2 dE R conj(p q),
                // ro = f / deno = |q|^2 R conj(p), ro2 = -|p|^2 R conj(q)
                const double wr = wh[u].re, wi = wh[u].im;
                const double pr = dE - wr, qr = dE + wr;
                const double ap = pr * pr + wi * wi, aq = qr * qr + wi * wi;
                const double R = f * rcp_nr(dE * ap * aq);
                const double Rq = R * aq, Rp = R * ap, Rd = 2.0 * R * dE;
                const zd rv = zd{(dE * dE - wh2[u].re) * Rd, wh2[u].im * Rd};
                const double t1 = pr * Rq, t2 = qr * Rp;
                const zd sp = zd{t1 - t2, wi * (Rq + Rp)};  // ro + ro2
                const zd sm = zd{t1 + t2, wi * (Rq - Rp)};  // ro - ro2
                // acc: [0..5] rv Im P_c, [6..11] Re(ro) |x_a|^2 then Im(ro) |x_a|^2,
                // [12..17] Re(ro) P_c, [18..23] Im(ro) P_c, over both orders
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                  acc[u][2 * c] += pq[c] * rv.re;
                  acc[u][2 * c + 1] += pq[c] * rv.im;
                  acc[u][6 + c] += sp.re * pq[9 + c];
                  acc[u][9 + c] += sp.im * pq[9 + c];
                  acc[u][12 + 2 * c] += sp.re * pq[3 + 2 * c];
                  acc[u][13 + 2 * c] += sm.re * pq[4 + 2 * c];
                  acc[u][18 + 2 * c] += sp.im * pq[3 + 2 * c];
                  acc[u][19 + 2 * c] += sm.im * pq[4 + 2 * c];
                }
              } else {
                const zd denv = zd{dE * dE - wh2[u].re, -wh2[u].im};
                const zd deno = zmul(zd{dE, 0.0}, zd{dE - wh[u].re, -wh[u].im});
                const double av = zabs2(denv), ao = zabs2(deno);
                // one reciprocal for both denominators
                const double fi = f * rcp_nr(av * ao);
                const double iv = fi * ao, io = fi * av;
                const zd rv = zd{denv.re * iv, -denv.im * iv};
                const zd ro = zd{deno.re * io, -deno.im * io};
                // acc: [0..17] velo_q (re, im), [18..35] ohc_q (re, im)
#pragma unroll
                for (int q = 0; q < 9; ++q) {
                  const double vim = pq[q];
                  acc[u][2 * q] += vim * rv.re;
                  acc[u][2 * q + 1] += vim * rv.im;
                  // i * (q2o * ro)
                  const double qr = pq[9 + 2 * q], qi = pq[9 + 2 * q + 1];
                  acc[u][18 + 2 * q] -= qr * ro.im + qi * ro.re;
                  acc[u][18 + 2 * q + 1] += qr * ro.re - qi * ro.im;
                }
              }
            }
        };
        if constexpr (HERM && IPT > 1) {
          // the m records j = split, split + S, ... taken in pairs in two
          // register sets, so one record's LDS loads are in flight while the
          // other is summed.  No exit between a load and its use (the
          // compiler would sink the loads to the use): past the last record
          // the loads repeat it, and an odd m's extra record adds with
          // weight 0 (exactly nothing: dE != 0 in every record)
          if (act && split < cnt) {
            constexpr int RW = 1 + IPT + NV;
            const int m = (cnt - 1 - split) / S + 1;
            auto rec = [&](int i) { return st + (int64_t)(split + (i < m ? i : m - 1) * S) * PS; };
            auto ld = [&](const double* p, double* r) {
              r[0] = p[0];
#pragma unroll
              for (int u = 0; u < IPT; ++u) r[1 + u] = p[1 + ef[u]];
#pragma unroll
              for (int q = 0; q < NV; ++q) r[1 + IPT + q] = p[1 + nef + q];
            };
            double ra[RW], rb[RW];
            ld(rec(0), ra);
            for (int i = 0; i < m; i += 2) {
              ld(rec(i + 1), rb);
              consume(ra[0], ra + 1, ra + 1 + IPT);
              ld(rec(i + 2), ra);
              double fb[IPT];
#pragma unroll
              for (int u = 0; u < IPT; ++u) fb[u] = (i + 1 < m) ? rb[1 + u] : 0.0;
              consume(rb[0], fb, rb + 1 + IPT);
            }
          }
        } else if (act) {
#pragma unroll 2
          for (int j = split; j < cnt; j += S) {  // several records' LDS loads in flight
            const double* sp = st + (int64_t)j * PS;
            double fv[IPT];
#pragma unroll
            for (int u = 0; u < IPT; ++u) fv[u] = sp[1 + ef[u]];
            consume(sp[0], fv, sp + 1 + nef);
          }
        }
      }
    }
    // combine the splits of each item in two halves of the accumulators
    // (red[split][item][HA] over the staging area); each half holds
    // everything its outputs need
    for (int half = 0; half < 2; ++half) {
      __syncthreads();
      if (act)
#pragma unroll
        for (int u = 0; u < IPT; ++u)
          if (iok[u])
#pragma unroll
            for (int q = 0; q < HA; ++q)
              st[(split * IW + slot + IWT * u) * HA + q] = acc[u][half * HA + q];
      __syncthreads();
      auto red = [&](int i2, int q) {
        double sR = 0.0;
        for (int u = 0; u < S; ++u) sR += st[(u * IW + i2) * HA + q];
        return sR;
      };
      if constexpr (!HERM) {
        for (int o = t; o < IW * 18; o += 256) {
          const int i2 = o / 18, q = o % 18;
          const int item = g0 + i2, hh = item % nhw, e2 = item / nhw;
          accp[(half ? L.ohc : L.velo) + 2 * 9 * (hh + (int64_t)nhw * e2) + q] += red(i2, q);
        }
      } else {
        // half 0: velo (antisymmetric, zero diagonal) and the ohc diagonal;
        // half 1: the ohc off-diagonal entries
        const int per = half ? 12 : 24;
        for (int o = t; o < IW * per; o += 256) {
          const int i2 = o / per, r = o % per;
          const int item = g0 + i2, hh = item % nhw, e2 = item / nhw;
          const int64_t rec = 2 * 9 * (hh + (int64_t)nhw * e2);
          if (half == 0) {
            if (r < 18) {  // velo_q, q = i + 3j, component r % 2
              const int q = r / 2, cpx = r % 2, i = q % 3, j = q / 3;
              if (i != j) {
                const double v = red(i2, 2 * (i + j - 1) + cpx);
                accp[L.velo + rec + r] += (i < j) ? v : -v;
              }
            } else {       // ohc_ii = i sum ro |x_i|^2 = (-Y, X)
              const int i = (r - 18) / 2, cpx = (r - 18) % 2;
              const double v = cpx ? red(i2, 6 + i) : -red(i2, 9 + i);
              accp[L.ohc + rec + 2 * (4 * i) + cpx] += v;
            }
          } else {         // ohc_ij (i < j) = (Xi - Yr, Xr + Yi); ohc_ji = (-(Xi + Yr), Xr - Yi)
            const int c = r / 4, sel = r % 4;  // sel: (ij re, ij im, ji re, ji im)
            const int i = (c == 2) ? 1 : 0, j = (c == 0) ? 1 : 2;
            const double xr = red(i2, 2 * c), xi = red(i2, 2 * c + 1);
            const double yr = red(i2, 6 + 2 * c), yi = red(i2, 6 + 2 * c + 1);
            double v;
            int q;
            if (sel == 0) { v = xi - yr; q = i + 3 * j; }
            else if (sel == 1) { v = xr + yi; q = i + 3 * j; }
            else if (sel == 2) { v = -(xi + yr); q = j + 3 * i; }
            else { v = xr - yi; q = j + 3 * i; }
            accp[L.ohc + rec + 2 * q + (sel & 1)] += v;
          }
        }
      }
    }
  }
}

// items per thread: two for Hermitian V with many items (C5's 256
// frequencies: 94.5 -> 83.4 ms per 32^3); few items (the single (hw, E_F)
// pair sums) keep one, at the higher occupancy
static int hwsum_ipt(const BandArgs& b) {
  return (b.v_herm && (int64_t)b.n_hw * b.n_ef >= 128) ? HW_IPT : 1;
}

size_t hwsum_lds_bytes(const BandArgs& b) {
  const size_t nv = b.v_herm ? hw_vals<true>() : hw_vals<false>();
  const size_t ha = (b.v_herm ? hw_accs<true>() : hw_accs<false>()) / 2;
  const size_t ps = 1 + (size_t)b.n_ef + nv;
  const size_t ipt = (size_t)hwsum_ipt(b);
  const size_t stage = (size_t)HW_CH * ps > 256 * ipt * ha ? (size_t)HW_CH * ps : 256 * ipt * ha;
  return sizeof(double) * ((size_t)b.n + (size_t)b.n_ef * b.n + 4 + stage);
}

// ===========================================================================
// Per-k Berry curvature at one Fermi level: out[k][i + 3j] = the k-point's
// term of kubo_ahc_tens (kubo.f90:37-78),
//   sum_n0 f(n0) sum_{n != n0} 2 Im(V_i(n0,n) V_j(n,n0)) / (E_n - E_n0)^2,
// one k-point per wavefront, the pairs spread over the lanes (any n).
// ===========================================================================
__global__ __launch_bounds__(64) void curv_kernel(int n, int64_t nk, const double* __restrict__ E,
                                                  const zd* __restrict__ V, double ef,
                                                  double Tsmear, int finiteT,
                                                  double* __restrict__ out) {
  const int lane = threadIdx.x;
  const int n2 = n * n;
  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    const double* Ek = E + k * n;
    const zd* Vk = V + k * 3 * (int64_t)n2;
    double om[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) om[q] = 0.0;
    for (int idx = lane; idx < n2; idx += 64) {
      const int n0 = idx % n, nn = idx / n;
      if (n0 == nn) continue;
      const double e0 = Ek[n0];
      const double f = finiteT ? 1.0 / (1.0 + exp((e0 - ef) / Tsmear)) : ((e0 < ef) ? 1.0 : 0.0);
      if (f == 0.0) continue;
      const double r = 1.0 / (e0 - Ek[nn]);
      const double w2 = 2.0 * f * r * r;
      zd x[3], y[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        x[a] = Vk[a * (int64_t)n2 + n0 + (int64_t)n * nn];
        y[a] = Vk[a * (int64_t)n2 + nn + (int64_t)n * n0];
      }
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) om[i + 3 * j] += w2 * zmul(x[i], y[j]).im;
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      const double v = wave_sum(om[q]);
      if (lane == 0) out[k * 9 + q] = v;
    }
  }
}

void launch_curv(hipStream_t s, int n, int64_t nk, const double* E, const zd* V, double ef,
                 double T_kelvin, double* out) {
  const double kB = 8.6173303 * 1e-5 / 27.211385;  // constants.f90:22
  unsigned grid = (unsigned)(nk < 65536 ? nk : 65536);
  hipLaunchKernelGGL(curv_kernel, dim3(grid), dim3(64), 0, s, n, nk, E, V, ef, kB * T_kelvin,
                     T_kelvin > 1e-2 ? 1 : 0, out);
}

// acc[c] += sum over the nblocks partial records of part[g][c].  A
// workgroup takes 64 consecutive c (one per lane, coalesced rows); its 16
// wavefronts take the records g = w, w + 16, ... in order, and the 16 wave
// sums are added in wave order: a fixed summation order, so the result is
// deterministic.
constexpr int RP_WAVES = 16;
__global__ __launch_bounds__(64 * RP_WAVES) void reduce_parts_kernel(int nblocks, int64_t len,
                                                                    const double* __restrict__ part,
                                                                    double* __restrict__ acc) {
  __shared__ double ws[RP_WAVES][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t c = blockIdx.x * (int64_t)64 + lane;
  double s = 0.0;
  if (c < len)
    for (int g = w; g < nblocks; g += RP_WAVES) s += part[(int64_t)g * len + c];
  ws[w][lane] = s;
  __syncthreads();
  if (w == 0 && c < len) {
    double t = 0.0;
#pragma unroll
    for (int v = 0; v < RP_WAVES; ++v) t += ws[v][lane];
    acc[c] += t;
  }
}

// n <= 32: the Hall and optical conductivities of several (hw, E_F) items
// (velo_ahc_tens / kubo_ohc_tens over a frequency list or several Fermi
// levels) run in hwsum_kernel, every item on its own threads, reading V
// from HBM; the single-item sums stay in the band-sum kernel's pair phase
static bool hw_many(const BandArgs& b) {
  return (b.jobs & MEP_JOB_AHC) && (int64_t)b.n_hw * b.n_ef > 1;
}

// The n <= 32 kernel's accumulators in LDS: the record's layout without the
// velo / ohc block when hwsum_kernel fills it (then N_hw does not bound the
// LDS); *gap is the length of the block left out
static AccLayout lds_layout(const BandArgs& b, const AccLayout& L, int64_t* gap) {
  AccLayout S = L;
  *gap = 0;
  if (hw_many(b)) {
    *gap = L.kic - L.velo;
    S.ohc = L.velo;
    S.kic -= *gap; S.klc -= *gap; S.kcs -= *gap;
    S.nic -= *gap; S.nlc -= *gap; S.ncs -= *gap;
    S.total -= *gap;
  }
  return S;
}

template <int NMAX, bool FUSED>
static void launch_bs(hipStream_t s, const BandArgs& b, double Tsmear, const AccLayout& L,
                      int64_t nk, const double* E, const zd* V, const zd* A, const zd* Om,
                      const double* w, int nblocks, double* part, const VeloFuse* f) {
  int64_t gap;
  const AccLayout S = lds_layout(b, L, &gap);
  size_t lds = BsLds<NMAX>::bytes(b.n_ef, S.total);
  (void)hipFuncSetAttribute((const void*)bandsum_kernel<NMAX, FUSED>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((bandsum_kernel<NMAX, FUSED>), dim3(nblocks), dim3(BS_THREADS), lds, s, b,
                     Tsmear, S, gap, nk, E, V, A, Om, w, part, f ? f->Hk : nullptr,
                     f ? f->strideHk : (int64_t)0, f ? f->U : nullptr);
}

size_t bandsum_lds_bytes(const BandArgs& b) {
  AccLayout L = acc_layout(b);
  if (b.n > 32) {
    const size_t a = bandsum_wide_lds_bytes(b), h = hwsum_lds_bytes(b);
    return a > h ? a : h;
  }
  int64_t gap;
  const AccLayout S = lds_layout(b, L, &gap);
  const size_t a = b.n <= 16 ? BsLds<16>::bytes(b.n_ef, S.total) : BsLds<32>::bytes(b.n_ef, S.total);
  const size_t h = hw_many(b) ? hwsum_lds_bytes(b) : 0;
  return a > h ? a : h;
}

// the velocities formed inside the band-sum kernel never reach HBM, which
// hwsum_kernel reads
bool bandsum_can_fuse(const BandArgs& b) {
  return b.n > 16 && b.n <= 32 && b.v_herm && !b.pos_op && !hw_many(b);
}

static void launch_hwsum(hipStream_t s, const BandArgs& b, double Tsmear, const AccLayout& L,
                         int64_t nk, const double* E, const zd* V, const double* w, int nblocks,
                         double* part) {
  const size_t lh = hwsum_lds_bytes(b);
  if (b.v_herm && hwsum_ipt(b) == HW_IPT) {
    (void)hipFuncSetAttribute((const void*)hwsum_kernel<true, HW_IPT>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lh);
    hipLaunchKernelGGL((hwsum_kernel<true, HW_IPT>), dim3(nblocks), dim3(256), lh, s, b, Tsmear, L,
                       nk, E, V, w, part);
  } else if (b.v_herm) {
    (void)hipFuncSetAttribute((const void*)hwsum_kernel<true, 1>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lh);
    hipLaunchKernelGGL((hwsum_kernel<true, 1>), dim3(nblocks), dim3(256), lh, s, b, Tsmear, L, nk,
                       E, V, w, part);
  } else {
    (void)hipFuncSetAttribute((const void*)hwsum_kernel<false, 1>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lh);
    hipLaunchKernelGGL((hwsum_kernel<false, 1>), dim3(nblocks), dim3(256), lh, s, b, Tsmear, L, nk,
                       E, V, w, part);
  }
}

void launch_bandsum(hipStream_t s, const BandArgs& b, int64_t nk, const double* E, const zd* V,
                    const zd* A, const zd* Om, const double* w, int nblocks, double* part,
                    const VeloFuse* fuse, zd* wide_scr) {
  AccLayout L = acc_layout(b);
  // kBoltz_Eh_K * T  (constants.f90:22, statistics.f90:77)
  const double kB = 8.6173303 * 1e-5 / 27.211385;
  double Tsmear = kB * b.T_kelvin;
  if (b.n > 32) {
    const size_t lds = bandsum_wide_lds_bytes(b);
    (void)hipFuncSetAttribute((const void*)bandsum_wide_kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(bandsum_wide_kernel, dim3(nblocks), dim3(256), lds, s, b, Tsmear, L, nk, E,
                       V, w, part, wide_scr, A, Om);
    if (b.jobs & MEP_JOB_AHC) launch_hwsum(s, b, Tsmear, L, nk, E, V, w, nblocks, part);
  } else {
    if (b.n <= 16)
      launch_bs<16, false>(s, b, Tsmear, L, nk, E, V, A, Om, w, nblocks, part, nullptr);
    else if (fuse)
      launch_bs<32, true>(s, b, Tsmear, L, nk, E, V, A, Om, w, nblocks, part, fuse);
    else
      launch_bs<32, false>(s, b, Tsmear, L, nk, E, V, A, Om, w, nblocks, part, nullptr);
    if (hw_many(b)) launch_hwsum(s, b, Tsmear, L, nk, E, V, w, nblocks, part);
  }
}

void launch_reduce_parts(hipStream_t s, int nblocks, int64_t len, const double* part, double* acc) {
  int64_t nb = (len + 63) / 64;
  hipLaunchKernelGGL(reduce_parts_kernel, dim3((unsigned)nb), dim3(64 * RP_WAVES), 0, s, nblocks,
                     len, part, acc);
}

__global__ void fill_zero_kernel(double* __restrict__ p, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

void launch_fill_zero(hipStream_t s, double* p, int64_t n) {
  int64_t nb = (n + 255) / 256;
  hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)nb), dim3(256), 0, s, p, n);
}
