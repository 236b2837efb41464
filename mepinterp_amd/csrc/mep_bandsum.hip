// mep_bandsum.hip -- Fermi-Dirac weights and the Berry / Kubo / MEP band
// sums of one batch of k-points, reduced into per-block partial records.
//
// One k-point per wavefront (grid-stride over k; fixed k -> block map, so
// the reduction order is deterministic).  V (3,n,n) is staged in LDS as
// three padded column-major planes.  Rows of the reference it replaces:
//   d1_fd_stat, d1_fd_get_N_el          statistics.f90:65-124
//   kubo_ahc_tens, velo_ahc_tens,
//   kubo_ohc_tens                       kubo.f90:37-198
//   kubo_mep_CS / LC / IC               kubo_mep.f90:16-149
//   mep_niu_CS / LC / IC                mep_niu.f90:20-159
// The IC triple sums are evaluated as sum_n a_i(n) sum_m V_k(n,m) b_l(m)
// with a_i(n) = V_i(n0,n)/(E0-En)^2 and b_l(m) = V_l(m,n0)/(E0-Em): 6n^2
// complex MACs per band instead of 27*3 n^2 triple products.  mep_niu's
// guard |(E0-En)^2 (E0-Em)| > kubo_tol couples n and m, so it is applied
// inside the m-sum.
#include "mep_common.hpp"
#include "mep_kernels.hpp"
#include "../../include/mepinterp.h"

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

template <int NMAX>
__global__ __launch_bounds__(64) void bandsum_kernel(BandArgs b, double Tsmear, AccLayout L,
                                                     int64_t nk, const double* __restrict__ E,
                                                     const zd* __restrict__ V,
                                                     const zd* __restrict__ A,
                                                     const zd* __restrict__ Om,
                                                     const double* __restrict__ w,
                                                     double* __restrict__ part) {
  constexpr int LD = NMAX + 1;
  constexpr int PL_ = NMAX * LD;  // plane size
  extern __shared__ double smem[];
  zd* Vs = reinterpret_cast<zd*>(smem);
  double* Es = smem + 2 * 3 * PL_;
  double* fd = Es + NMAX;
  double* acc = fd + (int64_t)b.n_ef * NMAX;
  const int t = threadIdx.x;
  const int n = b.n, nef = b.n_ef, nhw = b.n_hw;
  const bool finiteT = b.T_kelvin > 1e-2;
  for (int64_t c = t; c < L.total; c += 64) acc[c] = 0.0;

  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    const double wk = w ? w[k] : 1.0;
    __syncthreads();
    const zd* Vk = V + k * 3 * (int64_t)n * n;
    for (int idx = t; idx < 3 * n * n; idx += 64) {
      int a = idx % 3, r = idx / 3;
      int i = r % n, j = r / n;
      Vs[a * PL_ + i + LD * j] = Vk[idx];
    }
    if (t < n) Es[t] = E[k * n + t];
    __syncthreads();
    for (int idx = t; idx < nef * n; idx += 64) {
      if (b.fd_in) {
        fd[idx] = b.fd_in[k * (int64_t)nef * n + idx];
      } else {
        int ef = idx % nef, nb = idx / nef;
        double e = Es[nb], f = b.ef[ef];
        fd[idx] = finiteT ? 1.0 / (1.0 + exp((e - f) / Tsmear)) : ((e < f) ? 1.0 : 0.0);
      }
    }
    __syncthreads();
    // ---- electron count --------------------------------------------------
    for (int ef = 0; ef < nef; ++ef) {
      double s = (t < n) ? fd[ef + nef * t] : 0.0;
      s = wave_sum(s);
      if (t == 0) acc[L.Ne + ef] += wk * s;
    }
    // ---- anomalous Hall conductivity, hw -> 0 ----------------------------
    if (b.jobs & MEP_JOB_AHC) {
      double cv[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) cv[q] = 0.0;
      if (t < n) {
        for (int np = 0; np < n; ++np) {
          if (np == t) continue;
          double de = Es[np] - Es[t];
          double d2 = de * de;
#pragma unroll
          for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              zd p = zmul(Vs[i * PL_ + t + LD * np], Vs[j * PL_ + np + LD * t]);
              cv[i + 3 * j] -= 2.0 * p.im / d2;
            }
        }
      }
      for (int ef = 0; ef < nef; ++ef) {
        double f = (t < n) ? fd[ef + nef * t] : 0.0;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
          double s = wave_sum(-cv[q] * f);
          if (t == 0) acc[L.ahc + q + 9 * ef] += wk * s;
        }
      }
      // ---- hw-resolved Hall (velo_ahc) and optical (kubo_ohc) -----------
      if (nhw * nef == 1) {
        // pair-parallel, register accumulators
        double va[18], oa[18];
#pragma unroll
        for (int q = 0; q < 18; ++q) va[q] = oa[q] = 0.0;
        const zd wv = zd{b.hw[0], b.eta};
        const zd w2 = zmul(wv, wv);
        for (int p = t; p < n * n; p += 64) {
          int x = p % n, y = p / n;
          if (x == y) continue;
          double dE = Es[x] - Es[y];
          double dE2 = dE * dE;  // (E_n - E_np)^2 with n = x, np = y
          double dfv = fd[nef * x] - fd[nef * y];
          zd denv = zd{dE2 - w2.re, -w2.im};
          zd deno = zmul(zd{dE, 0.0}, zd{dE - b.hw[0], -b.eta});
#pragma unroll
          for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              double vim = zmul(Vs[i * PL_ + x + LD * y], Vs[j * PL_ + y + LD * x]).im;
              zd q1 = zdiv(zd{vim, 0.0}, denv);
              va[2 * (i + 3 * j)] += dfv * q1.re;
              va[2 * (i + 3 * j) + 1] += dfv * q1.im;
              // ohc with m = x, n = y: V_i(y,x) V_j(x,y) / (dE (dE - hw - i eta))
              zd vo = zmul(Vs[i * PL_ + y + LD * x], Vs[j * PL_ + x + LD * y]);
              zd q2 = zmuli(zscale(zdiv(vo, deno), dfv));
              oa[2 * (i + 3 * j)] += q2.re;
              oa[2 * (i + 3 * j) + 1] += q2.im;
            }
        }
#pragma unroll
        for (int q = 0; q < 18; ++q) {
          double s1 = wave_sum(va[q]);
          double s2 = wave_sum(oa[q]);
          if (t == 0) {
            acc[L.velo + q] += wk * s1;
            acc[L.ohc + q] += wk * s2;
          }
        }
      } else {
        // hw-parallel: lane h owns every accumulator entry of frequency h
        for (int x = 0; x < n; ++x)
          for (int y = 0; y < n; ++y) {
            if (x == y) continue;
            double dE = Es[x] - Es[y];
            double dE2 = dE * dE;
            double vim[9];
            zd vo[9];
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
              for (int i = 0; i < 3; ++i) {
                vim[i + 3 * j] = zmul(Vs[i * PL_ + x + LD * y], Vs[j * PL_ + y + LD * x]).im;
                vo[i + 3 * j] = zmul(Vs[i * PL_ + y + LD * x], Vs[j * PL_ + x + LD * y]);
              }
            for (int h = t; h < nhw; h += 64) {
              zd wv = zd{b.hw[h], b.eta};
              zd w2 = zmul(wv, wv);
              zd denv = zd{dE2 - w2.re, -w2.im};
              zd deno = zmul(zd{dE, 0.0}, zd{dE - b.hw[h], -b.eta});
              for (int ef = 0; ef < nef; ++ef) {
                double dfv = fd[ef + nef * x] - fd[ef + nef * y];
                double* av = acc + L.velo + 2 * 9 * (h + (int64_t)nhw * ef);
                double* ao = acc + L.ohc + 2 * 9 * (h + (int64_t)nhw * ef);
#pragma unroll
                for (int q = 0; q < 9; ++q) {
                  zd q1 = zdiv(zd{vim[q], 0.0}, denv);
                  av[2 * q] += wk * (dfv * q1.re);
                  av[2 * q + 1] += wk * (dfv * q1.im);
                  zd q2 = zmuli(zscale(zdiv(vo[q], deno), dfv));
                  ao[2 * q] += wk * q2.re;
                  ao[2 * q + 1] += wk * q2.im;
                }
              }
            }
          }
      }
    }
    // ---- magnetoelectric polarisability: IC and LC ------------------------
    const bool doK = (b.jobs & MEP_JOB_KUBO) != 0;
    const bool doNjob = (b.jobs & MEP_JOB_MEP) != 0;
    if (doK || doNjob) {
      const int n0 = t;
      const bool doN = doNjob && n0 < b.n_val && n0 < n;
      double F2[9], F2m[9], F3[9], F3m[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) F2[q] = F2m[q] = F3[q] = F3m[q] = 0.0;
      if (n0 < n && (doK || doN)) {
        const double E0 = Es[n0];
        zd T[18], TM[18];
#pragma unroll
        for (int q = 0; q < 18; ++q) T[q] = TM[q] = zd{0.0, 0.0};
        double R3[18], R3m[18];
#pragma unroll
        for (int q = 0; q < 18; ++q) R3[q] = R3m[q] = 0.0;
        zd z0[3];
#pragma unroll
        for (int l = 0; l < 3; ++l) z0[l] = Vs[l * PL_ + n0 + LD * n0];
        for (int nn = 0; nn < n; ++nn) {
          if (nn == n0) continue;
          const double dn = E0 - Es[nn];
          zd wv[6], wm[6];
#pragma unroll
          for (int p = 0; p < 6; ++p) wv[p] = wm[p] = zd{0.0, 0.0};
          for (int m = 0; m < n; ++m) {
            if (m == n0) continue;
            const double dm = E0 - Es[m];
            zd bl[3], vk[3];
#pragma unroll
            for (int l = 0; l < 3; ++l) {
              zd v = Vs[l * PL_ + m + LD * n0];
              bl[l] = zd{v.re / dm, v.im / dm};
              vk[l] = Vs[l * PL_ + nn + LD * m];
            }
            const bool keep = doN && (fabs((dn * dn) * dm) > b.kubo_tol);
#pragma unroll
            for (int p = 0; p < 6; ++p) {
              zd pr = zmul(vk[pk(p)], bl[pl(p)]);
              wv[p] = zadd(wv[p], pr);
              if (keep) wm[p] = zadd(wm[p], pr);
            }
          }
          const double dn2 = dn * dn;
          zd ai[3];
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            zd v = Vs[i * PL_ + n0 + LD * nn];
            ai[i] = zd{v.re / dn2, v.im / dn2};
          }
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int p = 0; p < 6; ++p) {
              T[i * 6 + p] = zadd(T[i * 6 + p], zmul(ai[i], wv[p]));
              if (doN) TM[i * 6 + p] = zadd(TM[i * 6 + p], zmul(ai[i], wm[p]));
            }
          // local (LC): eps_jkl Re(V_i(n0,nn) V_k(nn,n0) V_l(n0,n0)) / (E0-En)^3
          const double den3 = dn * dn * dn;
          const bool keep3 = doN && (fabs(den3) > b.kubo_tol);
          zd xi[3], yk[3];
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            xi[i] = Vs[i * PL_ + n0 + LD * nn];
            yk[i] = Vs[i * PL_ + nn + LD * n0];
          }
#pragma unroll
          for (int p = 0; p < 6; ++p) {
            zd yz = zmul(yk[pk(p)], z0[pl(p)]);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              double r = zmul(xi[i], yz).re / den3;
              R3[i * 6 + p] += r;
              if (keep3) R3m[i * 6 + p] += r;
            }
          }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            F2[i + 3 * j] = -(T[i * 6 + 2 * j].re - T[i * 6 + 2 * j + 1].re);
            F2m[i + 3 * j] = -(TM[i * 6 + 2 * j].re - TM[i * 6 + 2 * j + 1].re);
            F3[i + 3 * j] = R3[i * 6 + 2 * j] - R3[i * 6 + 2 * j + 1];
            F3m[i + 3 * j] = R3m[i * 6 + 2 * j] - R3m[i * 6 + 2 * j + 1];
          }
      }
      if (doK) {
        for (int ef = 0; ef < nef; ++ef) {
          double f = (n0 < n) ? fd[ef + nef * n0] : 0.0;
#pragma unroll
          for (int q = 0; q < 9; ++q) {
            double s2 = wave_sum(F2[q] * f);
            double s3 = wave_sum(F3[q] * f);
            if (t == 0) {
              acc[L.kic + q + 9 * ef] += wk * s2;
              acc[L.klc + q + 9 * ef] += wk * s3;
            }
          }
        }
      }
      if (doN) {
#pragma unroll
        for (int q = 0; q < 9; ++q) {
          acc[L.nic + q + 9 * n0] += wk * F2m[q];
          acc[L.nlc + q + 9 * n0] += wk * F3m[q];
        }
      }
      // ---- Chern-Simons term (needs A and Omega) ------------------------
      if (b.pos_op && A && Om) {
        double cs = 0.0;
        if (n0 < n) {
          const int64_t NN = (int64_t)n * n;
          const zd* Ak = A + k * 3 * NN + 3 * ((int64_t)n0 + (int64_t)n * n0);
          const zd* Ok = Om + k * 9 * NN + 9 * ((int64_t)n0 + (int64_t)n * n0);
          zd om[3];
          // om_c = sum_ab eps_abc Om_ab
          om[0] = zsub(Ok[1 + 3 * 2], Ok[2 + 3 * 1]);
          om[1] = zsub(Ok[2 + 3 * 0], Ok[0 + 3 * 2]);
          om[2] = zsub(Ok[0 + 3 * 1], Ok[1 + 3 * 0]);
          zd dp = zd{0.0, 0.0};
#pragma unroll
          for (int a = 0; a < 3; ++a) dp = zadd(dp, zcmul(Ak[a], om[a]));
          cs = dp.re;
        }
        if (doK) {
          for (int ef = 0; ef < nef; ++ef) {
            double f = (n0 < n) ? fd[ef + nef * n0] : 0.0;
            double s = wave_sum(0.5 * f * cs);
            if (t == 0)
              for (int i = 0; i < 3; ++i) acc[L.kcs + i + 3 * i + 9 * ef] += wk * s;
          }
        }
        if (doN)
          for (int i = 0; i < 3; ++i) acc[L.ncs + i + 3 * i + 9 * n0] += wk * (0.5 * cs);
      }
    }
  }
  __syncthreads();
  double* out = part + (int64_t)blockIdx.x * L.total;
  for (int64_t c = t; c < L.total; c += 64) out[c] = acc[c];
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
  size_t lds = sizeof(double) * (2 * 3 * (size_t)NMAX * (NMAX + 1) + NMAX +
                                 (size_t)b.n_ef * NMAX + (size_t)L.total);
  hipFuncSetAttribute((const void*)bandsum_kernel<NMAX>,
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(bandsum_kernel<NMAX>, dim3(nblocks), dim3(64), lds, s, b, Tsmear, L, nk, E,
                     V, A, Om, w, part);
}

size_t bandsum_lds_bytes(const BandArgs& b) {
  int nm = b.n <= 8 ? 8 : b.n <= 16 ? 16 : 32;
  AccLayout L = acc_layout(b);
  return sizeof(double) * (2 * 3 * (size_t)nm * (nm + 1) + nm + (size_t)b.n_ef * nm + (size_t)L.total);
}

void launch_bandsum(hipStream_t s, const BandArgs& b, int64_t nk, const double* E, const zd* V,
                    const zd* A, const zd* Om, const double* w, int nblocks, double* part) {
  AccLayout L = acc_layout(b);
  // kBoltz_Eh_K * T  (constants.f90:22, statistics.f90:77)
  const double kB = 8.6173303 * 1e-5 / 27.211385;
  double Tsmear = kB * b.T_kelvin;
  if (b.n <= 8)
    launch_bs<8>(s, b, Tsmear, L, nk, E, V, A, Om, w, nblocks, part);
  else if (b.n <= 16)
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
