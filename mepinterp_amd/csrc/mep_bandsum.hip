// mep_bandsum.hip -- Fermi-Dirac weights and the Berry / Kubo / MEP band
// sums of one batch of k-points, reduced into per-block partial records.
//
// One k-point per 256-thread workgroup (grid-stride over k; fixed k -> block
// map and fixed-order in-block reductions, so results are deterministic).
// V (planar, 3 x n x n) is staged in LDS as three padded column-major
// planes.  Rows of the reference it replaces:
//   d1_fd_stat, d1_fd_get_N_el          statistics.f90:65-124
//   kubo_ahc_tens, velo_ahc_tens,
//   kubo_ohc_tens                       kubo.f90:37-198
//   kubo_mep_CS / LC / IC               kubo_mep.f90:16-149
//   mep_niu_CS / LC / IC                mep_niu.f90:20-159
// The itinerant (IC) triple sums of every band n0 are diagonals of matrix
// products: sum_{n,m} a_i(n0,n) V_k(n,m) B_l(m,n0) = [A_i (V_k B_l)](n0,n0)
// with A_i(n0,n) = V_i(n0,n)/(E0-En)^2 and B_l(m,n0) = V_l(m,n0)/(E0-Em);
// the six products W_kl = V_k B_l (k != l) run on FP64 MFMA.  mep_niu's
// guard |(E0-En)^2 (E0-Em)| > kubo_tol excludes, for fixed (n0, n), one
// contiguous window of m around n0 (energies are sorted); those terms are
// subtracted from W (or the kept ones summed directly when the window
// covers most bands).
#include "mep_common.hpp"
#include "mep_kernels.hpp"
#include "../../include/mepinterp.h"
#include "mep_mfma.hpp"

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

// ordered Cartesian pairs (k,l), k != l, grouped so that for output index j
// pairs 2j (+1) and 2j+1 (-1) carry eps_jkl
__device__ constexpr int pk(int p) { return p == 0 ? 1 : p == 1 ? 2 : p == 2 ? 2 : p == 3 ? 0 : p == 4 ? 0 : 1; }
__device__ constexpr int pl(int p) { return p == 0 ? 2 : p == 1 ? 1 : p == 2 ? 0 : p == 3 ? 2 : p == 4 ? 1 : 0; }

// B_l(m, n0) = V_l(m, n0) / (E_n0 - E_m), zero for m = n0
struct BScaled {
  const zd* V;
  int ld;
  const double* E;
  __device__ inline zd operator()(int m, int n0) const {
    if (m == n0) return zd{0.0, 0.0};
    double d = E[n0] - E[m];
    zd v = V[m + ld * n0];
    return zd{v.re / d, v.im / d};
  }
};

constexpr int BS_THREADS = 256;

// LDS plan (doubles): V planes 2*3*PL | W 2*PL | E NMAX | fd nef*NMAX |
// red RED | windows (int16) 2*NMAX*NMAX/4 | acc L.total
template <int NMAX>
struct BsLds {
  static constexpr int LD = NMAX + 1;
  static constexpr int PL = NMAX * LD;
  static constexpr int RED = 27 * NMAX > 36 * (256 / 9) ? 27 * NMAX : 36 * (256 / 9);
  static size_t bytes(int nef, int64_t total) {
    return sizeof(double) * (2 * 3 * (size_t)PL + 2 * (size_t)PL + NMAX + (size_t)nef * NMAX + RED +
                             (2 * NMAX * NMAX + 3) / 4 + (size_t)total);
  }
};

template <int NMAX>
__global__ __launch_bounds__(BS_THREADS) void bandsum_kernel(BandArgs b, double Tsmear, AccLayout L,
                                                             int64_t nk,
                                                             const double* __restrict__ E,
                                                             const zd* __restrict__ V,
                                                             const zd* __restrict__ A,
                                                             const zd* __restrict__ Om,
                                                             const double* __restrict__ w,
                                                             double* __restrict__ part) {
  using S = BsLds<NMAX>;
  constexpr int LD = S::LD, PL = S::PL;
  extern __shared__ double smem[];
  const int n = b.n, nef = b.n_ef, nhw = b.n_hw, nval = b.n_val;
  zd* Vs = reinterpret_cast<zd*>(smem);
  zd* Ws = Vs + 3 * PL;
  double* Es = reinterpret_cast<double*>(Ws + PL);
  double* fd = Es + NMAX;
  double* red = fd + (int64_t)nef * NMAX;
  short* win = reinterpret_cast<short*>(red + S::RED);  // [n0][nn] lo, hi
  double* acc = red + S::RED + (2 * NMAX * NMAX + 3) / 4;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int n2 = n * n;
  const bool finiteT = b.T_kelvin > 1e-2;
  const bool doK = (b.jobs & MEP_JOB_KUBO) != 0;
  const bool doNjob = (b.jobs & MEP_JOB_MEP) != 0;
  for (int64_t c = t; c < L.total; c += BS_THREADS) acc[c] = 0.0;

  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    const double wk = w ? w[k] : 1.0;
    __syncthreads();
    const zd* Vk = V + k * 3 * (int64_t)n2;
    for (int idx = t; idx < 3 * n2; idx += BS_THREADS) {
      int a = idx / n2, r = idx - a * n2;
      Vs[a * PL + r % n + LD * (r / n)] = Vk[idx];
    }
    if (t < n) Es[t] = E[k * n + t];
    __syncthreads();
    for (int idx = t; idx < nef * n; idx += BS_THREADS) {
      if (b.fd_in) {
        fd[idx] = b.fd_in[k * (int64_t)nef * n + idx];
      } else {
        int ef = idx % nef, nb = idx / nef;
        double e = Es[nb], f = b.ef[ef];
        fd[idx] = finiteT ? 1.0 / (1.0 + exp((e - f) / Tsmear)) : ((e < f) ? 1.0 : 0.0);
      }
    }
    __syncthreads();
    // ---- electron count (serial over bands: fixed order) ---------------
    if (t < nef) {
      double s = 0.0;
      for (int nb = 0; nb < n; ++nb) s += fd[t + nef * nb];
      acc[L.Ne + t] += wk * s;
    }
    // ---- anomalous Hall conductivity, hw -> 0 -----------------------------
    if (b.jobs & MEP_JOB_AHC) {
      if (t < 9 * n) {
        const int nb = t % n, q = t / n, i = q % 3, j = q / 3;
        double cv = 0.0;
        for (int np = 0; np < n; ++np) {
          if (np == nb) continue;
          double de = Es[np] - Es[nb];
          zd p = zmul(Vs[i * PL + nb + LD * np], Vs[j * PL + np + LD * nb]);
          cv -= 2.0 * p.im / (de * de);
        }
        red[q * NMAX + nb] = cv;
      }
      __syncthreads();
      if (t < 9 * nef) {
        const int q = t % 9, ef = t / 9;
        double s = 0.0;
        for (int nb = 0; nb < n; ++nb) s -= red[q * NMAX + nb] * fd[ef + nef * nb];
        acc[L.ahc + q + 9 * ef] += wk * s;
      }
      __syncthreads();
      // ---- hw-resolved Hall (velo_ahc) and optical (kubo_ohc) -------------
      if (nhw * nef == 1) {
        // thread (q, g): component q = i + 3j, ordered pairs p = g, g + 28, ...
        constexpr int NG = BS_THREADS / 9;
        const int q = t % 9, g = t / 9;
        const int i = q % 3, j = q / 3;
        double vre = 0.0, vim_ = 0.0, ore = 0.0, oim = 0.0;
        if (g < NG) {
          const zd wq = zd{b.hw[0], b.eta};
          const zd w2 = zmul(wq, wq);
          for (int p = g; p < n2; p += NG) {
            const int x = p % n, y = p / n;
            if (x == y) continue;
            const double dE = Es[x] - Es[y];
            const double dfv = fd[x] - fd[y];
            const zd denv = zd{dE * dE - w2.re, -w2.im};
            const zd deno = zmul(zd{dE, 0.0}, zd{dE - b.hw[0], -b.eta});
            double vim = zmul(Vs[i * PL + x + LD * y], Vs[j * PL + y + LD * x]).im;
            zd q1 = zdiv(zd{vim, 0.0}, denv);
            vre += dfv * q1.re;
            vim_ += dfv * q1.im;
            zd vo = zmul(Vs[i * PL + y + LD * x], Vs[j * PL + x + LD * y]);
            zd q2 = zmuli(zscale(zdiv(vo, deno), dfv));
            ore += q2.re;
            oim += q2.im;
          }
          red[(4 * q + 0) * NG + g] = vre;
          red[(4 * q + 1) * NG + g] = vim_;
          red[(4 * q + 2) * NG + g] = ore;
          red[(4 * q + 3) * NG + g] = oim;
        }
        __syncthreads();
        if (t < 36) {
          const int qq = t / 4, c = t % 4;
          double s = 0.0;
          for (int gg = 0; gg < NG; ++gg) s += red[t * NG + gg];
          const int64_t dst = (c < 2 ? L.velo : L.ohc) + 2 * qq + (c & 1);
          acc[dst] += wk * s;
        }
      } else {
        // hw-parallel: thread h owns every accumulator entry of frequency h
        for (int h = t; h < nhw; h += BS_THREADS) {
          const zd wq = zd{b.hw[h], b.eta};
          const zd w2 = zmul(wq, wq);
          for (int x = 0; x < n; ++x)
            for (int y = 0; y < n; ++y) {
              if (x == y) continue;
              const double dE = Es[x] - Es[y];
              const zd denv = zd{dE * dE - w2.re, -w2.im};
              const zd deno = zmul(zd{dE, 0.0}, zd{dE - b.hw[h], -b.eta});
#pragma unroll
              for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                  const int q = i + 3 * j;
                  double vim = zmul(Vs[i * PL + x + LD * y], Vs[j * PL + y + LD * x]).im;
                  zd vo = zmul(Vs[i * PL + y + LD * x], Vs[j * PL + x + LD * y]);
                  zd q1 = zdiv(zd{vim, 0.0}, denv);
                  zd q2o = zdiv(vo, deno);
                  for (int ef = 0; ef < nef; ++ef) {
                    const double dfv = fd[ef + nef * x] - fd[ef + nef * y];
                    double* av = acc + L.velo + 2 * 9 * (h + (int64_t)nhw * ef);
                    double* ao = acc + L.ohc + 2 * 9 * (h + (int64_t)nhw * ef);
                    av[2 * q] += wk * (dfv * q1.re);
                    av[2 * q + 1] += wk * (dfv * q1.im);
                    zd q2 = zmuli(zscale(q2o, dfv));
                    ao[2 * q] += wk * q2.re;
                    ao[2 * q + 1] += wk * q2.im;
                  }
                }
            }
        }
      }
      __syncthreads();
    }
    // ---- magnetoelectric polarisability ----------------------------------
    if (doK || doNjob) {
      // exclusion windows of mep_niu's guard |(E0-En)^2 (E0-Em)| > kubo_tol:
      // energies are sorted, so for fixed (n0, n) the excluded m form one
      // contiguous index range around n0
      if (doNjob) {
        for (int idx = t; idx < nval * n; idx += BS_THREADS) {
          const int n0 = idx % nval, nn = idx / nval;
          const double E0 = Es[n0], dn = E0 - Es[nn];
          const double dn2 = dn * dn;
          int lo = n0, hi = n0;
          while (lo > 0 && !(fabs(dn2 * (E0 - Es[lo - 1])) > b.kubo_tol)) --lo;
          while (hi < n - 1 && !(fabs(dn2 * (E0 - Es[hi + 1])) > b.kubo_tol)) ++hi;
          win[2 * (n0 + NMAX * nn)] = (short)lo;
          win[2 * (n0 + NMAX * nn) + 1] = (short)hi;
        }
      }
      const int n0 = t % n, ii = t / n;  // thread (n0, i) for t < 3n
      const bool act = t < 3 * n;
      const bool doN = doNjob && act && n0 < nval;
      const double E0 = act ? Es[n0] : 0.0;
      // F2_ij = -sum_kl eps_jkl Re T_ikl: pair p contributes to j = p/2 with
      // sign -1 (even p) or +1 (odd p); LC likewise with the opposite sign
      double F2[3] = {0.0, 0.0, 0.0}, F2m[3] = {0.0, 0.0, 0.0};
      double F3[3] = {0.0, 0.0, 0.0}, F3m[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int p = 0; p < 6; ++p) {
        const double sg = (p & 1) ? 1.0 : -1.0;
        __syncthreads();
        // W = V_k B_l  (itinerant sum of all n0 at once)
        zgemm_tiles<false>(n, Vs + pk(p) * PL, LD, BScaled{Vs + pl(p) * PL, LD, Es}, Ws, LD);
        __syncthreads();
        if (act) {
          const zd* Vi = Vs + ii * PL;
          const zd* Vk_ = Vs + pk(p) * PL;
          const BScaled bl{Vs + pl(p) * PL, LD, Es};
          double tr = 0.0, trm = 0.0;
          for (int nn = 0; nn < n; ++nn) {
            if (nn == n0) continue;
            const double dn = E0 - Es[nn];
            const double dn2 = dn * dn;
            const zd vi = Vi[n0 + LD * nn];
            const zd a = zd{vi.re / dn2, vi.im / dn2};
            const zd wnm = Ws[nn + LD * n0];
            tr += a.re * wnm.re - a.im * wnm.im;
            if (doN) {
              const int lo = win[2 * (n0 + NMAX * nn)], hi = win[2 * (n0 + NMAX * nn) + 1];
              if (lo == 0 && hi == n - 1) continue;  // every m excluded
              zd kept = zd{0.0, 0.0};
              if (2 * (hi - lo + 1) > n) {
                for (int m = 0; m < lo; ++m) kept = zadd(kept, zmul(Vk_[nn + LD * m], bl(m, n0)));
                for (int m = hi + 1; m < n; ++m) kept = zadd(kept, zmul(Vk_[nn + LD * m], bl(m, n0)));
              } else {
                zd ex = zd{0.0, 0.0};
                for (int m = lo; m <= hi; ++m) ex = zadd(ex, zmul(Vk_[nn + LD * m], bl(m, n0)));
                kept = zsub(wnm, ex);
              }
              trm += a.re * kept.re - a.im * kept.im;
            }
          }
          F2[p / 2] += sg * tr;
          F2m[p / 2] += sg * trm;
        }
      }
      // local (LC) terms
      if (act) {
        zd z0[3];
#pragma unroll
        for (int l = 0; l < 3; ++l) z0[l] = Vs[l * PL + n0 + LD * n0];
        for (int nn = 0; nn < n; ++nn) {
          if (nn == n0) continue;
          const double dn = E0 - Es[nn];
          const double den3 = dn * dn * dn;
          const bool keep3 = doN && (fabs(den3) > b.kubo_tol);
          const zd xi = Vs[ii * PL + n0 + LD * nn];
          zd yk[3];
#pragma unroll
          for (int q = 0; q < 3; ++q) yk[q] = Vs[q * PL + nn + LD * n0];
#pragma unroll
          for (int p = 0; p < 6; ++p) {
            const double r = zmul(xi, zmul(yk[pk(p)], z0[pl(p)])).re / den3;
            const double sg = (p & 1) ? -1.0 : 1.0;
            F3[p / 2] += sg * r;
            if (keep3) F3m[p / 2] += sg * r;
          }
        }
      }
      __syncthreads();
      if (doK) {
        if (act)
          for (int j = 0; j < 3; ++j) {
            red[(ii + 3 * j) * NMAX + n0] = F2[j];
            red[(9 + ii + 3 * j) * NMAX + n0] = F3[j];
          }
        __syncthreads();
        if (t < 18 * nef) {
          const int q = t % 18, ef = t / 18;
          double s = 0.0;
          for (int m0 = 0; m0 < n; ++m0) s += red[q * NMAX + m0] * fd[ef + nef * m0];
          if (q < 9) acc[L.kic + q + 9 * ef] += wk * s;
          else acc[L.klc + (q - 9) + 9 * ef] += wk * s;
        }
      }
      if (doN) {
        for (int j = 0; j < 3; ++j) {
          acc[L.nic + ii + 3 * j + 9 * n0] += wk * F2m[j];
          acc[L.nlc + ii + 3 * j + 9 * n0] += wk * F3m[j];
        }
      }
      // ---- Chern-Simons term (needs A and Omega) --------------------------
      if (b.pos_op && A && Om) {
        __syncthreads();
        if (t < n) {
          const int64_t NN = (int64_t)n2;
          const zd* Ak = A + k * 3 * NN + 3 * ((int64_t)t + (int64_t)n * t);
          const zd* Ok = Om + k * 9 * NN + 9 * ((int64_t)t + (int64_t)n * t);
          zd om[3];
          om[0] = zsub(Ok[1 + 3 * 2], Ok[2 + 3 * 1]);
          om[1] = zsub(Ok[2 + 3 * 0], Ok[0 + 3 * 2]);
          om[2] = zsub(Ok[0 + 3 * 1], Ok[1 + 3 * 0]);
          zd dp = zd{0.0, 0.0};
#pragma unroll
          for (int a = 0; a < 3; ++a) dp = zadd(dp, zcmul(Ak[a], om[a]));
          red[t] = dp.re;
        }
        __syncthreads();
        if (doK && t < nef) {
          double s = 0.0;
          for (int m0 = 0; m0 < n; ++m0) s += 0.5 * fd[t + nef * m0] * red[m0];
          for (int i = 0; i < 3; ++i) acc[L.kcs + i + 3 * i + 9 * t] += wk * s;
        }
        if (doNjob && t < nval && t < n)
          for (int i = 0; i < 3; ++i) acc[L.ncs + i + 3 * i + 9 * t] += wk * (0.5 * red[t]);
      }
    }
  }
  __syncthreads();
  double* out = part + (int64_t)blockIdx.x * L.total;
  for (int64_t c = t; c < L.total; c += BS_THREADS) out[c] = acc[c];
}

__global__ void reduce_parts_kernel(int nblocks, int64_t len, const double* __restrict__ part,
                                    double* __restrict__ acc) {
  int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= len) return;
  double s = 0.0;
  for (int g = 0; g < nblocks; ++g) s += part[(int64_t)g * len + c];
  acc[c] += s;
}

template <int NMAX>
static void launch_bs(hipStream_t s, const BandArgs& b, double Tsmear, const AccLayout& L,
                      int64_t nk, const double* E, const zd* V, const zd* A, const zd* Om,
                      const double* w, int nblocks, double* part) {
  size_t lds = BsLds<NMAX>::bytes(b.n_ef, L.total);
  hipFuncSetAttribute((const void*)bandsum_kernel<NMAX>,
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(bandsum_kernel<NMAX>, dim3(nblocks), dim3(BS_THREADS), lds, s, b, Tsmear, L,
                     nk, E, V, A, Om, w, part);
}

size_t bandsum_lds_bytes(const BandArgs& b) {
  AccLayout L = acc_layout(b);
  return b.n <= 16 ? BsLds<16>::bytes(b.n_ef, L.total) : BsLds<32>::bytes(b.n_ef, L.total);
}

void launch_bandsum(hipStream_t s, const BandArgs& b, int64_t nk, const double* E, const zd* V,
                    const zd* A, const zd* Om, const double* w, int nblocks, double* part) {
  AccLayout L = acc_layout(b);
  // kBoltz_Eh_K * T  (constants.f90:22, statistics.f90:77)
  const double kB = 8.6173303 * 1e-5 / 27.211385;
  double Tsmear = kB * b.T_kelvin;
  if (b.n <= 16)
    launch_bs<16>(s, b, Tsmear, L, nk, E, V, A, Om, w, nblocks, part);
  else
    launch_bs<32>(s, b, Tsmear, L, nk, E, V, A, Om, w, nblocks, part);
}

void launch_reduce_parts(hipStream_t s, int nblocks, int64_t len, const double* part, double* acc) {
  int64_t nb = (len + 255) / 256;
  hipLaunchKernelGGL(reduce_parts_kernel, dim3((unsigned)nb), dim3(256), 0, s, nblocks, len, part,
                     acc);
}

__global__ void fill_zero_kernel(double* __restrict__ p, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

void launch_fill_zero(hipStream_t s, double* p, int64_t n) {
  int64_t nb = (n + 255) / 256;
  hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)nb), dim3(256), 0, s, p, n);
}
