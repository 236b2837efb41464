// mep_extsum.hip -- the optical, photogalvanic and gyrotropic band sums of
// the k-loop (core.f90:235-253), accumulated into per-workgroup partial
// records like the main band sums.
//   kubo_opt_tens        kubo.f90:204-405   (wann guide 12.10-12.15; needs A)
//   photo_2nd_cond       photo.f90:16-86    (Yang et al. PRB 97 241118 Eq. 1)
//   get_gyro_C / D / Dw  gyro.f90:153-300   (Souza et al. PRB 97 035158)
//   d1_fd_stat_deriv     statistics.f90:160-196 (1/(2 + 2 cosh) of the
//                        occupation itself, as the reference evaluates it)
// One k-point at a time per 256-thread workgroup; every output element is
// owned by one thread, which adds its per-k value to the workgroup's record
// (fixed order, deterministic).  For n <= 32 V is staged in LDS (planar,
// padded) where it fits beside the occupation tables; otherwise it is read
// from L2 / HBM.  photo_2nd_cond's m sum is a product per (b, c) shared by
// every (a, hw, E_F) item, kept in LDS (n <= 64 where it fits).  A and Omega are read in the
// reference layout (position operator only).  The occupations, their
// derivative and the per-band curvature sit in dynamic LDS, sized by n and
// n_eF (ext_lds_bytes).
#include "mep_common.hpp"
#include "mep_kernels.hpp"
#include "../../include/mepinterp.h"

ExtLayout ext_layout(const BandArgs& b) {
  ExtLayout X;
  const int64_t hf = (int64_t)b.n_hw * b.n_ef;
  int64_t o = 0;
  X.opt_s = o;   o += 18 * hf;
  X.opt_a = o;   o += 18 * hf;
  X.photo = o;   o += 54 * hf;
  X.gyro_C = o;  o += 18 * (int64_t)b.n_ef;
  X.gyro_D = o;  o += 18 * (int64_t)b.n_ef;
  X.gyro_Dw = o; o += 18 * hf;
  X.total = o;
  return X;
}

constexpr int XS_THREADS = 256;
constexpr int XS_NMAX = 32;
constexpr int XS_LD = XS_NMAX + 1;
constexpr int XS_PL = XS_NMAX * XS_LD;

__device__ inline int levi3(int a, int b, int c) {
  return (a == b || b == c || a == c) ? 0 : (((b - a + 3) % 3 == 1) ? 1 : -1);
}

// dynamic LDS: [V planes if LDSV] | E (128) | omega (3 x 128 complex) |
// fd (n_ef x n) | fd_stat_deriv (n_ef x n) | [with MEP_JOB_OPT: photo's W
// and P (XS_NMAX^2 complex each) if LDSV, else P (n^2 complex) where it
// fits; then split sums (XS_THREADS complex)]
__host__ __device__ inline size_t ext_lds_doubles(int n, int nef, bool ldsv, bool photo,
                                                  bool plds = true) {
  const size_t base = (ldsv ? 2 * 3 * (size_t)XS_PL : 0) + 128 + 2 * 3 * 128 + 2 * (size_t)nef * n;
  if (!photo) return base;
  return base + (ldsv ? 2 * 2 * XS_NMAX * XS_NMAX : plds ? 2 * (size_t)n * n : 0) +
         2 * XS_THREADS;
}
// V from L2: photo_2nd_cond's per-(b, c) product P (n^2 complex) in LDS
// where it fits beside the occupation tables (n <= 64)
constexpr int XS_PMAX = 64;
__host__ __device__ inline bool ext_photo_p(int n, int nef) {
  return n <= XS_PMAX && sizeof(double) * ext_lds_doubles(n, nef, false, true) <= 160 * 1024;
}

template <bool LDSV>
__global__ __launch_bounds__(XS_THREADS) void extsum_kernel(BandArgs b, double Tsmear, ExtLayout X,
                                                           int64_t nk, const double* __restrict__ E,
                                                           const zd* __restrict__ V,
                                                           const zd* __restrict__ A,
                                                           const zd* __restrict__ Om,
                                                           const double* __restrict__ w,
                                                           double* __restrict__ part,
                                                           zd* __restrict__ scr) {
  extern __shared__ __attribute__((aligned(16))) double xs_smem[];
  zd* Vs = reinterpret_cast<zd*>(xs_smem);                         // LDSV only
  double* Es = xs_smem + (LDSV ? 2 * 3 * XS_PL : 0);
  zd* omg = reinterpret_cast<zd*>(Es + 128);                         // omega_c(n) of get_gyro_D
  double* fd = reinterpret_cast<double*>(omg + 3 * 128);             // (n_ef, band)
  const int n = b.n, nef = b.n_ef, nhw = b.n_hw, n2 = n * n;
  double* fdd = fd + (int64_t)nef * n;                               // fd_stat_deriv
  const int t = threadIdx.x;
  const bool finiteT = b.T_kelvin > 1e-2;
  const bool doOpt = (b.jobs & MEP_JOB_OPT) != 0, doGyro = (b.jobs & MEP_JOB_GYRO) != 0;
  const bool pos = b.pos_op && A && Om;
  const double pi = 3.14159265358979323846;
  const double eta = b.eta;
  double* rec = part + (int64_t)blockIdx.x * X.total;
  auto Aat = [&](int64_t k, int a, int i, int j) {  // A_a(i, j) of k, reference layout
    return A[k * 3 * (int64_t)n2 + a + 3 * ((int64_t)i + (int64_t)n * j)];
  };
  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    const double wk = w ? w[k] : 1.0;
    __syncthreads();
    const zd* Vk = V + k * 3 * (int64_t)n2;
    // V_a(i, j): the LDS planes or the k-point's planar V
    auto Vat = [&](int a, int i, int j) {
      return LDSV ? Vs[a * XS_PL + i + XS_LD * j] : Vk[a * (int64_t)n2 + i + (int64_t)n * j];
    };
    if (LDSV)
      for (int idx = t; idx < 3 * n2; idx += XS_THREADS) {
        const int a = idx / n2, r = idx - a * n2;
        Vs[a * XS_PL + r % n + XS_LD * (r / n)] = Vk[idx];
      }
    for (int i = t; i < n; i += XS_THREADS) Es[i] = E[k * n + i];
    __syncthreads();
    for (int idx = t; idx < nef * n; idx += XS_THREADS) {
      double f;
      if (b.fd_in) {
        f = b.fd_in[k * (int64_t)nef * n + idx];
      } else {
        const int ef = idx % nef, nb = idx / nef;
        const double e = Es[nb], ff = b.ef[ef];
        f = finiteT ? 1.0 / (1.0 + exp((e - ff) / Tsmear)) : ((e < ff) ? 1.0 : 0.0);
      }
      fd[idx] = f;
      fdd[idx] = 1.0 / (2.0 + 2.0 * cosh(f));
    }
    // omega_c(n) = sum_ab eps_abc curv_ab(n): Omega_ab(n,n) with a position
    // operator, else get_kubo_curv (wann_interp.f90:113-146)
    for (int t0 = t; doGyro && t0 < n; t0 += XS_THREADS) {
      const int t = t0;  // the band of this pass
      zd curv[9];
      if (pos) {
        const zd* On = Om + k * 9 * (int64_t)n2 + 9 * ((int64_t)t + (int64_t)n * t);
        for (int q = 0; q < 9; ++q) curv[q] = On[q];
      } else {
        for (int q = 0; q < 9; ++q) curv[q] = zd{0.0, 0.0};
        for (int np = 0; np < n; ++np) {
          if (np == t) continue;
          const double dEE = (Es[np] - Es[t]) * (Es[np] - Es[t]);
          if (fabs(dEE) < b.kubo_tol) continue;
          for (int j = 0; j < 3; ++j)
            for (int a = 0; a < 3; ++a)
              curv[a + 3 * j].re +=
                  zmul(Vat(a, t, np), Vat(j, np, t)).im / dEE;
        }
        for (int q = 0; q < 9; ++q) curv[q] = zscale(curv[q], -2.0);
      }
      for (int c = 0; c < 3; ++c) {
        zd s = zd{0.0, 0.0};
        for (int bb = 0; bb < 3; ++bb)
          for (int a = 0; a < 3; ++a) {
            const int l = levi3(a, bb, c);
            if (l) s = zadd(s, zscale(curv[a + 3 * bb], (double)l));
          }
        omg[c + 3 * t] = s;
      }
    }
    __syncthreads();
    // ---- kubo_opt_tens: item (a, b, hw, ef) ----------------------------------
    if (doOpt && pos)
      for (int it = t; it < 9 * nhw * nef; it += XS_THREADS) {
        const int q = it % 9, h = (it / 9) % nhw, ef = it / (9 * nhw);
        const int a = q % 3, bb = q / 3;
        zd herm = zd{0.0, 0.0}, aherm = zd{0.0, 0.0};
        const double hwv = b.hw[h];
        for (int m = 0; m < n; ++m)
          for (int nn = 0; nn < n; ++nn) {
            if (nn == m) continue;
            const double dE = Es[m] - Es[nn];
            const double df = fd[ef + nef * m] - fd[ef + nef * nn];
            const zd amn = zscale(zmul(Aat(k, a, nn, m), Aat(k, bb, m, nn)), dE);
            // delta_broad: Im 1/(pi (dE - i eta) - pi hw)
            const zd dsm = zd{pi * (dE - hwv), -pi * eta};
            const double dbr = -dsm.im / zabs2(dsm);
            herm = zsub(herm, zscale(amn, pi * dbr * df));
            aherm = zadd(aherm, zmuli(zscale(amn, df / (dE - hwv))));
          }
        const int64_t o = 2 * (q + 9 * (h + (int64_t)nhw * ef));
        rec[X.opt_s + o] += wk * herm.re;       // (12.14) Re herm + i Im aherm
        rec[X.opt_s + o + 1] += wk * aherm.im;
        rec[X.opt_a + o] += wk * aherm.re;      // (12.15) Re aherm + i Im herm
        rec[X.opt_a + o + 1] += wk * herm.im;
      }
    // ---- photo_2nd_cond: item (a, b, c, hw, ef) ------------------------------
    // sum_{n, m, l != n} (f_l - f_n) / hw^2 V_a(n,l) V_b(l,m) V_c(m,n)
    //   / (E_n - E_m - i eta) sum_{+-} 1 / (E_n - E_l +- hw - 2 i eta)
    if (doOpt) {
      // P in LDS (V in LDS, or n <= 64), else in the workgroup's slot of scr
      const bool pl = LDSV || ext_photo_p(n, nef);
      if (pl || scr) {
        // the m sum does not depend on a, hw or E_F: per (b, c) the product
        // P(l, n) = sum_m V_b(l, m) W_c(m, n), W_c(m, n) = V_c(m, n) /
        // (E_n - E_m - i eta), in LDS; then each item (a, hw, E_F) is the
        // O(n^2) contraction with V_a(n, l) (n^3 per (b, c) instead of per
        // item).  Items are split over the threads like the omega-resolved
        // sums: thread (slot, split), splits combined in a fixed order.
        // V in LDS: W (XS_NMAX^2) | P (XS_NMAX^2) | red; V from L2: P (n^2) |
        // red, and W's factor 1 / (E_n - E_m - i eta) formed in the m loop
        zd* Wm = reinterpret_cast<zd*>(fd + 2 * (int64_t)nef * n);
        zd* Pm = LDSV ? Wm + XS_NMAX * XS_NMAX : pl ? Wm : scr + (int64_t)blockIdx.x * n2;
        zd* red = LDSV ? Pm + XS_NMAX * XS_NMAX : pl ? Pm + n2 : Wm;  // [XS_THREADS]
        const int NI = 3 * nhw * nef;
        for (int c = 0; c < 3; ++c) {
          __syncthreads();  // the previous (b, c)'s readers of W are done
          if (LDSV)
            for (int e = t; e < n2; e += XS_THREADS) {
              const int m = e % n, nn = e / n;
              const double x = Es[nn] - Es[m], iv = 1.0 / (x * x + eta * eta);
              Wm[e] = zmul(Vat(c, m, nn), zd{x * iv, eta * iv});
            }
          for (int bb = 0; bb < 3; ++bb) {
            __syncthreads();  // W written; the previous P's readers are done
            for (int e = t; e < n2; e += XS_THREADS) {
              const int l = e % n, nn = e / n;
              zd sp = zd{0.0, 0.0};
              if (LDSV) {
                for (int m = 0; m < n; ++m) sp = zadd(sp, zmul(Vat(bb, l, m), Wm[m + n * nn]));
              } else {
                for (int m = 0; m < n; ++m) {
                  const double x = Es[nn] - Es[m], iv = 1.0 / (x * x + eta * eta);
                  sp = zadd(sp, zmul(Vat(bb, l, m), zmul(Vat(c, m, nn), zd{x * iv, eta * iv})));
                }
              }
              Pm[e] = sp;
            }
            __syncthreads();
            for (int g0 = 0; g0 < NI; g0 += XS_THREADS) {
              const int IW = (NI - g0 < XS_THREADS) ? NI - g0 : XS_THREADS;
              const int S = XS_THREADS / IW, slot = t % IW, split = t / IW;
              const int it = g0 + slot, a = it % 3, h = (it / 3) % nhw, ef = it / (3 * nhw);
              const double hwv = b.hw[h];
              zd acc = zd{0.0, 0.0};
              if (split < S)
                for (int pp = split; pp < n2; pp += S) {
                  const int l = pp % n, nn = pp / n;
                  if (l == nn) continue;
                  const double df = fd[ef + nef * l] - fd[ef + nef * nn];
                  const double x = Es[nn] - Es[l];
                  const double xp = x + hwv, xm = x - hwv, e2 = 4.0 * eta * eta;
                  const double ip = 1.0 / (xp * xp + e2), im = 1.0 / (xm * xm + e2);
                  const zd sh = zd{xp * ip + xm * im, 2.0 * eta * (ip + im)};
                  acc = zadd(acc, zscale(zmul(zmul(Vat(a, nn, l), Pm[pp]), sh), df));
                }
              red[t] = acc;
              __syncthreads();
              if (split == 0) {
                zd sum = red[slot];
                for (int u = 1; u < S; ++u) sum = zadd(sum, red[slot + IW * u]);
                const int q = a + 3 * bb + 9 * c;
                const int64_t o = 2 * (q + 27 * (h + (int64_t)nhw * ef));
                rec[X.photo + o] += wk * (sum.re / (hwv * hwv));
                rec[X.photo + o + 1] += wk * (sum.im / (hwv * hwv));
              }
              __syncthreads();  // red is free again
            }
          }
        }
      } else {
        for (int it = t; it < 27 * nhw * nef; it += XS_THREADS) {
          const int q = it % 27, h = (it / 27) % nhw, ef = it / (27 * nhw);
          const int a = q % 3, bb = (q / 3) % 3, c = q / 9;
          const double hwv = b.hw[h];
          zd acc = zd{0.0, 0.0};
          for (int nn = 0; nn < n; ++nn)
            for (int m = 0; m < n; ++m) {
              const zd dEnm = zd{Es[nn] - Es[m], -eta};
              for (int l = 0; l < n; ++l) {
                if (l == nn) continue;
                const double df = fd[ef + nef * l] - fd[ef + nef * nn];
                const zd vvv = zmul(Vat(a, nn, l), zmul(Vat(bb, l, m), Vat(c, m, nn)));
                zd g = zd{0.0, 0.0};
                for (int om = -1; om <= 1; om += 2) {
                  const zd dEhw = zd{Es[nn] - Es[l] + om * hwv, -2.0 * eta};
                  g = zadd(g, zdiv(vvv, zmul(dEnm, dEhw)));
                }
                acc = zadd(acc, zscale(g, df / (hwv * hwv)));
              }
            }
          const int64_t o = 2 * (q + 27 * (h + (int64_t)nhw * ef));
          rec[X.photo + o] += wk * acc.re;
          rec[X.photo + o + 1] += wk * acc.im;
        }
      }
    }
    // ---- gyrotropic C, D: item (a, b, ef); Dw: item (a, b, hw, ef) -----------
    if (doGyro) {
      for (int it = t; it < 9 * nef; it += XS_THREADS) {
        const int q = it % 9, ef = it / 9, a = q % 3, bb = q / 3;
        zd c = zd{0.0, 0.0}, d = zd{0.0, 0.0};
        for (int nn = 0; nn < n; ++nn) {
          const zd va = Vat(a, nn, nn), vb = Vat(bb, nn, nn);
          const double f = fdd[ef + nef * nn];
          c = zadd(c, zscale(zmul(va, vb), f));
          d = zsub(d, zscale(zmul(va, omg[bb + 3 * nn]), f));
        }
        rec[X.gyro_C + 2 * it] += wk * c.re;
        rec[X.gyro_C + 2 * it + 1] += wk * c.im;
        rec[X.gyro_D + 2 * it] += wk * d.re;
        rec[X.gyro_D + 2 * it + 1] += wk * d.im;
      }
      if (pos)
        for (int it = t; it < 9 * nhw * nef; it += XS_THREADS) {
          const int q = it % 9, h = (it / 9) % nhw, ef = it / (9 * nhw);
          const int a = q % 3, bb = q / 3;
          const double hwv = b.hw[h];
          zd dw = zd{0.0, 0.0};
          for (int nn = 0; nn < n; ++nn) {
            // get_curvTilde_hw (gyro.f90:271-300), component bb
            double omt = 0.0;
            for (int m = 0; m < n; ++m) {
              if (m == nn) continue;
              const double ww = (Es[m] - Es[nn]) * (Es[m] - Es[nn]);
              const int b1 = (bb + 1) % 3, b2 = (bb + 2) % 3;
              const zd cr = zsub(zmul(Aat(k, b1, nn, m), Aat(k, b2, m, nn)),
                                 zmul(Aat(k, b2, nn, m), Aat(k, b1, m, nn)));
              omt -= cr.im * ww / (ww - hwv);
            }
            const zd va = Vat(a, nn, nn);
            dw = zsub(dw, zscale(va, omt * fdd[ef + nef * nn]));
          }
          const int64_t o = 2 * (q + 9 * (h + (int64_t)nhw * ef));
          rec[X.gyro_Dw + o] += wk * dw.re;
          rec[X.gyro_Dw + o + 1] += wk * dw.im;
        }
    }
  }
}

// V and photo's tables in LDS for n <= 32 where they fit beside the
// occupation tables; otherwise V is read from L2 / HBM
static bool ext_ldsv(const BandArgs& b) {
  const bool photo = (b.jobs & MEP_JOB_OPT) != 0;
  return b.n <= XS_NMAX && sizeof(double) * ext_lds_doubles(b.n, b.n_ef, true, photo) <= 160 * 1024;
}

// (with MEP_JOB_OPT and P outside LDS, the split sums still take
// XS_THREADS complex of LDS)
size_t ext_lds_bytes(const BandArgs& b) {
  const bool photo = (b.jobs & MEP_JOB_OPT) != 0;
  const bool ldsv = ext_ldsv(b);
  return sizeof(double) *
         ext_lds_doubles(b.n, b.n_ef, ldsv, photo, ldsv || ext_photo_p(b.n, b.n_ef));
}

void launch_extsum(hipStream_t s, const BandArgs& b, int64_t nk, const double* E, const zd* V,
                   const zd* A, const zd* Om, const double* w, int nblocks, double* part,
                   zd* scr) {
  ExtLayout X = ext_layout(b);
  const double kB = 8.6173303 * 1e-5 / 27.211385;  // kBoltz_Eh_K (constants.f90:22)
  const size_t lds = ext_lds_bytes(b);
  if (ext_ldsv(b)) {
    (void)hipFuncSetAttribute((const void*)extsum_kernel<true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(extsum_kernel<true>, dim3(nblocks), dim3(XS_THREADS), lds, s, b,
                       kB * b.T_kelvin, X, nk, E, V, A, Om, w, part, scr);
  } else {
    (void)hipFuncSetAttribute((const void*)extsum_kernel<false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(extsum_kernel<false>, dim3(nblocks), dim3(XS_THREADS), lds, s, b,
                       kB * b.T_kelvin, X, nk, E, V, A, Om, w, part, scr);
  }
}
