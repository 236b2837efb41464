// mep_eigh.hip -- batched Hermitian eigensolver, one k-point per wavefront.
//
// Replaces zheevd('V','U') of zheevd_wrapper (matrix_math.f90:192-226) inside
// get_wann_interp (wann_interp.f90:95).  Contract kept: only the upper
// triangle of H(k) is read, eigenvalues ascending, eigenvectors in columns.
//
// Supported sizes: n <= 32 here, 32 < n <= 128 in the wide kernels below
// (the caller checks).
// Three kernels per k-block:
//   eigh_house (a team of NMAX lanes per k, lane i <-> row i, matrix in LDS;
//      in FP64 for 16 < n <= 32 one k per wavefront with half a row per lane
//      in registers, eigh_house_half):
//      Householder reduction A = Q T Q^H with rank-2 trailing updates; the
//      reflectors are stored in U's slot of the k-point, the diagonal phases
//      P that make the subdiagonal real (T = P T_r P^H), d, e and tau in a
//      side record.
//   eigh_ql (one LANE per k-point, 64 k-points per wavefront): the implicit
//      QL recurrence on T_r (EISPACK tql2 method).  A first pass with d, e in
//      registers (eigh_ql_reg) gives the ascending eigenvalues and a
//      separation verdict; only k-points with close eigenvalues run the
//      second pass, which appends every sweep's Givens rotations (c_i, s_i)
//      to a per-k record in HBM.
//   eigh_vec (a team per k): the vectors of the real tridiagonal, directly
//      from twisted factorizations for separated k-points or by replaying
//      the rotations on Z (LDS, lane r <-> row r), then U = Q P Z: lane j
//      keeps column j of P Z in registers and applies the reflectors in
//      reverse order.
//   Gauge convention shared with the oracle: each eigenvector's
//   largest-modulus component is made real and positive.
// 32 < n <= 128: Householder with the matrix in one wavefront's LDS, a
// workgroup's LDS or a workgroup's registers (eigh_house_lds / _wide /
// _rows, MEP_VAR_EIGH), the QL passes (registers up to 48 bands, else LDS
// columns), the tridiagonal vectors (eigh_replay_wide) and the blocked
// compact-WY back-transformation on FP64 MFMA (eigh_back_mfma).
#include "mep_common.hpp"
#include "mep_kernels.hpp"
#include "mep_lane.hpp"

typedef double mep_v4d __attribute__((ext_vector_type(4)));

// NMAX lanes form a team that owns one k-point; a wavefront runs 64/NMAX
// teams side by side.  Team reductions and broadcasts run on the VALU
// (team_sum, team_bcast of mep_lane.hpp).

// side record per k-point (doubles): d[n] e[n] tau[n] ph[2n] count
__host__ __device__ inline int64_t side_len(int n) { return 5 * (int64_t)n + 1; }
// rotation record per k-point: capacity in rotations; each rotation is a
// (c, s) pair of doubles plus its column index (one byte)
__host__ __device__ inline int64_t rot_cap(int n) { return 4 * (int64_t)n * n + 4 * (int64_t)n + 8; }

template <int NMAX, class R>
__global__ __launch_bounds__(64) void eigh_house_kernel(int n, int64_t nk,
                                                        const typename cplx_of<R>::type* __restrict__ Hk,
                                                        int64_t strideHk, int packed,
                                                        const double* __restrict__ kpts,
                                                        const double* __restrict__ tau,
                                                        typename cplx_of<R>::type* __restrict__ U,
                                                        R* __restrict__ side) {
  using Z = typename cplx_of<R>::type;
  constexpr int TEAMS = 64 / NMAX;
  constexpr int LDA = NMAX;
  __shared__ Z a_s[TEAMS][NMAX * NMAX];
  __shared__ Z vv_s[TEAMS][NMAX], ww_s[TEAMS][NMAX], sub_s[TEAMS][NMAX];
  __shared__ R tauv_s[TEAMS][NMAX];
  const int lane = threadIdx.x;
  const int tm = lane / NMAX, t = lane % NMAX;
  Z* a = a_s[tm];
  Z *vv = vv_s[tm], *ww = ww_s[tm], *sub = sub_s[tm];
  R* tauv = tauv_s[tm];
  const R two_pi = R(6.283185307179586476925);

  for (int64_t kb = (int64_t)blockIdx.x * TEAMS; kb < nk; kb += (int64_t)gridDim.x * TEAMS) {
    const int64_t k = kb + tm;
    const bool valid = k < nk;
    // ---- load the upper triangle (tight-binding phase applied) ----------
    for (int idx = t; idx < n * n; idx += NMAX) {
      int i = idx % n, j = idx / n;
      if (i <= j) {
        const int64_t o = packed ? i + ((int64_t)j * (j + 1)) / 2 : i + (int64_t)n * j;
        Z h = valid ? Hk[k * strideHk + o] : Z{R(0), R(0)};
        if (valid && tau && i != j) {
          const R kx = kpts[3 * k], ky = kpts[3 * k + 1], kz = kpts[3 * k + 2];
          R ang = two_pi * (kx * (tau[3 * j] - tau[3 * i]) + ky * (tau[3 * j + 1] - tau[3 * i + 1]) +
                                 kz * (tau[3 * j + 2] - tau[3 * i + 2]));
          R sn, cs;
          sin_cos(ang, &sn, &cs);
          h = zmul(Z{cs, sn}, h);
        }
        if (i == j) {
          a[i + LDA * i] = Z{h.re, R(0)};
        } else {
          a[i + LDA * j] = h;
          a[j + LDA * i] = zconj(h);
        }
      }
    }
    __syncthreads();
    // ---- Householder (an already reduced column gets tau = 0, v = 0) ----
    for (int kc = 0; kc + 2 < n; ++kc) {
      const int m = n - kc - 1;
      const Z x = (t < m) ? a[(kc + 1 + t) + LDA * kc] : Z{R(0), R(0)};
      const R tail2 = team_sum<NMAX, R>((t >= 1 && t < m) ? zabs2(x) : R(0));
      const Z x0 = Z{team_bcast<NMAX>(x.re, 0), team_bcast<NMAX>(x.im, 0)};
      const bool refl = tail2 > R(0);
      const R ax02 = zabs2(x0);
      const R xn = sqrt(tail2 + ax02);
      const R ax0 = sqrt(ax02);
      const Z alpha = !refl ? x0 : ((ax0 > R(0)) ? zscale(x0, -xn / ax0) : Z{-xn, R(0)});
      Z v = refl ? x : Z{R(0), R(0)};
      if (t == 0 && refl) v = zsub(x, alpha);
      // |v|^2 = |x|^2 - 2 Re(conj(alpha) x0) + |alpha|^2 = 2 xn (xn + |x0|)
      const R vn2 = R(2) * xn * (xn + ax0);
      const R tk = refl ? R(2) / vn2 : R(0);
      vv[t] = (t < m) ? v : Z{R(0), R(0)};
      __syncthreads();
      Z p = Z{R(0), R(0)};
      if (t < m) {
        const Z* row = &a[kc + 1 + t + LDA * (kc + 1)];
        for (int j = 0; j < m; ++j) p = zadd(p, zmul(row[LDA * j], vv[j]));
        p = zscale(p, tk);
      }
      const R K = R(0.5) * tk * team_sum<NMAX, R>((t < m) ? (v.re * p.re + v.im * p.im) : R(0));
      const Z w = zsub(p, zscale(v, K));
      ww[t] = (t < m) ? w : Z{R(0), R(0)};
      __syncthreads();
      if (t < m && refl) {
        Z* row = &a[kc + 1 + t + LDA * (kc + 1)];
        for (int j = 0; j < m; ++j) {
          const Z vj = vv[j], wj = ww[j];
          const Z upd = zadd(zmul(v, zconj(wj)), zmul(w, zconj(vj)));
          row[LDA * j] = zsub(row[LDA * j], upd);
        }
      }
      if (t < m) a[(kc + 1 + t) + LDA * kc] = v;
      if (t == 0) {
        sub[kc] = alpha;
        tauv[kc] = tk;
      }
      __syncthreads();
    }
    if (t == 0 && n >= 2) {
      sub[n - 2] = a[(n - 1) + LDA * (n - 2)];
      tauv[n - 2] = R(0);
    }
    __syncthreads();
    if (valid) {
      R* sd = side + k * side_len(n);
      // reflectors (strict lower triangle) into U's slot
      Z* Uk = U + k * (int64_t)n * n;
      for (int idx = t; idx < n * n; idx += NMAX) Uk[idx] = a[idx % n + LDA * (idx / n)];
      if (t < n) {
        sd[t] = a[t + LDA * t].re;  // d
        sd[2 * n + t] = (t < n - 1) ? tauv[t] : R(0);
      }
      if (t + 1 < n) {
        const Z s_ = sub[t];
        const R as = sqrt(zabs2(s_));
        sd[n + t] = as;
        sub[t] = (as > R(0)) ? zscale(s_, R(1) / as) : Z{R(1), R(0)};
      }
    }
    // |s_i| and the unit phase u_i = s_i / |s_i| (1 for s_i = 0) of every i at
    // once, then the running product ph_{i+1} = ph_i u_i in one lane: the same
    // operations as one serial loop over i (ph u with u = 1 is exact)
    __syncthreads();
    if (valid && t == 0) {
      R* sd = side + k * side_len(n);
      Z ph = Z{R(1), R(0)};
      sd[3 * n] = R(1);
      sd[3 * n + 1] = R(0);
      for (int i = 0; i + 1 < n; ++i) {
        ph = zmul(ph, sub[i]);
        sd[3 * n + 2 * (i + 1)] = ph.re;
        sd[3 * n + 2 * (i + 1) + 1] = ph.im;
      }
      sd[n + n - 1] = R(0);
    }
    __syncthreads();
  }
}

// Householder tridiagonalisation in FP64 for 16 < n <= 32 with one k-point
// per wavefront: lane (t, h) = (lane % 32, lane / 32) keeps the columns
// j = 2 jj + h (jj < 16) of row t in registers (64 VGPRs, so four wavefronts
// share a SIMD).  Column kc lives in half kc % 2 at jj = kc / 2; its entries
// reach the other half through a permlane swap.  v_j = w_j = 0 for j <= kc,
// so at step kc the product and the rank-2 update skip every jj whose two
// columns are both <= kc (known at compile time in the unrolled loop: about
// half of the work) and run over the rest without a branch.  Same outputs as
// eigh_house_kernel.  TAU: Wannier centres given (tight-binding
// convention); without them the phase code is compiled out, which keeps the
// load of the row within the register budget.
template <bool TAU>
__global__ __launch_bounds__(64, 4) void eigh_house_half_kernel(int n, int64_t nk,
                                                             const zd* __restrict__ Hk,
                                                             int64_t strideHk, int packed,
                                                             const double* __restrict__ kpts,
                                                             const double* __restrict__ tau,
                                                             zd* __restrict__ U,
                                                             double* __restrict__ side) {
  constexpr int NMAX = 32, HC = 16;
  __shared__ zd vv[NMAX], ww[NMAX], sub[NMAX], fph[NMAX];
  __shared__ double tauv[NMAX];
  const double two_pi = 6.283185307179586476925;
  const zd zero = zd{0.0, 0.0};
  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    // the lane index is made opaque per k-point: otherwise the compiler
    // hoists every step's lane masks out of this loop and spills them
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane));
    const int t = lane & 31, h = lane >> 5;
    // the value of lane ^ 32 (the other half of the same row)
    auto other = [&](double v) {
      const unsigned long long u = (unsigned long long)__double_as_longlong(v);
      const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)u, (unsigned)u, false, false);
      const auto hi =
          __builtin_amdgcn_permlane32_swap((unsigned)(u >> 32), (unsigned)(u >> 32), false, false);
      return h == 0 ? join_d(lo[1], hi[1]) : join_d(lo[0], hi[0]);
    };
    // ---- row t, columns 16h + jj: upper triangle from H(k), lower as conjugates
    // tight-binding convention: element (t, j) takes e^{2 pi i k.(tau_j - tau_t)}
    // = f_j conj(f_t), f_j = e^{2 pi i k.tau_j}: one sincos per band, through LDS
    zd ft = zd{1.0, 0.0};
    if (TAU) {
      if (h == 0) {
        zd f = zd{1.0, 0.0};
        if (t < n) {
          const double ang = two_pi * (kpts[3 * k] * tau[3 * t] + kpts[3 * k + 1] * tau[3 * t + 1] +
                                       kpts[3 * k + 2] * tau[3 * t + 2]);
          sin_cos(ang, &f.im, &f.re);
        }
        fph[t] = f;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      ft = zconj(fph[t]);
    }
    zd row[HC];
#pragma unroll
    for (int jj = 0; jj < HC; ++jj) {
      const int j = 2 * jj + h;
      zd x = zero;
      if (t < n && j < n) {
        const int i0 = j >= t ? t : j, j0 = j >= t ? j : t;  // (i0 <= j0) element
        const int64_t o = packed ? i0 + ((int64_t)j0 * (j0 + 1)) / 2 : i0 + (int64_t)n * j0;
        x = Hk[k * strideHk + o];
        if (j < t) x = zconj(x);
        if (TAU && j != t) x = zmul(x, zmul(fph[j], ft));
        if (j == t) x.im = 0.0;
      }
      row[jj] = x;
    }
#pragma clang loop unroll(full)
    for (int kc = 0; kc < NMAX - 2; ++kc) {
      if (kc + 2 >= n) continue;  // uniform; every index below stays compile-time
      const int hk = kc % 2, jc = kc / 2;
      const int jj0 = (kc + 1) / 2;  // jj < jj0: columns 2 jj, 2 jj + 1 <= kc
      const bool mine = h == hk;   // this lane holds column kc
      const bool below = t > kc;   // rows kc+1 .. n-1 take part
      const zd xo = mine ? row[jc] : zero;
      const double tail2 = team_sum<32, double>((mine && t > kc + 1) ? zabs2(xo) : 0.0);
      // cross-lane reads stay outside conditionals: every lane takes part.
      // tail2 and xo are exactly zero in the half that does not own column
      // kc, so the sums over both halves are the owning half's values
      const double tail = swap32_sum(tail2, tail2);
      const zd x0 = zd{readlane_d(xo.re, 32 * hk + kc + 1), readlane_d(xo.im, 32 * hk + kc + 1)};
      const zd xt = zd{swap32_sum(xo.re, xo.re), swap32_sum(xo.im, xo.im)};  // x_t in both halves
      const bool refl = tail > 0.0;
      const double ax02 = zabs2(x0);
      const double xn = sqrt_nr(tail + ax02);
      const double ax0 = sqrt_nr(ax02);
      const zd alpha = !refl ? x0 : ((ax0 > 0.0) ? zscale(x0, -xn * rcp_nr(ax0)) : zd{-xn, 0.0});
      zd v = (refl && below) ? xt : zero;
      if (t == kc + 1 && refl) v = zsub(xt, alpha);
      // |v|^2 = |x|^2 - 2 Re(conj(alpha) x0) + |alpha|^2 = 2 xn (xn + |x0|)
      const double vn2 = 2.0 * xn * (xn + ax0);
      const double tk = refl ? 2.0 * rcp_nr(vn2) : 0.0;
      if (h == 0) vv[t] = v;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // p_t = tk sum_j A(t, j) v_j: this half's columns, then the other's
      zd p4[4] = {zero, zero, zero, zero};
#pragma unroll
      for (int jj = jj0; jj < HC; ++jj) {
        const zd a_ = row[jj], b_ = vv[2 * jj + h];
        zd& q = p4[(jj - jj0) & 3];
        if (jj - jj0 < 4) {  // the first product of each chain sets it
          q.re = fma(a_.re, b_.re, -a_.im * b_.im);
          q.im = fma(a_.re, b_.im, a_.im * b_.re);
        } else {
          q.re = fma(a_.re, b_.re, fma(-a_.im, b_.im, q.re));
          q.im = fma(a_.re, b_.im, fma(a_.im, b_.re, q.im));
        }
      }
      zd p = zadd(zadd(p4[0], p4[1]), zadd(p4[2], p4[3]));
      p = zd{swap32_sum(p.re, p.re), swap32_sum(p.im, p.im)};  // the same order in both halves
      p = below ? zscale(p, tk) : zero;
      const double K = 0.5 * tk * team_sum<32, double>(v.re * p.re + v.im * p.im);
      const zd w = zsub(p, zscale(v, K));
      if (h == 0) ww[t] = w;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (refl && below) {
#pragma unroll
        for (int jj = jj0; jj < HC; ++jj) {
          // row_j -= v conj(w_j) + w conj(v_j)  (zero for j <= kc)
          const zd vj = vv[2 * jj + h], wj = ww[2 * jj + h];
          zd r_ = row[jj];
          r_.re = fma(-v.re, wj.re, fma(-v.im, wj.im, fma(-w.re, vj.re, fma(-w.im, vj.im, r_.re))));
          r_.im = fma(-v.im, wj.re, fma(v.re, wj.im, fma(-w.im, vj.re, fma(w.re, vj.im, r_.im))));
          row[jj] = r_;
        }
      }
      if (mine && below) row[jc] = v;
      if (lane == 0) {
        sub[kc] = alpha;
        tauv[kc] = tk;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // ---- outputs -----------------------------------------------------------
    double dgl = 0.0;  // diagonal of row t, held by half t % 2
    zd last = zero;    // A(n-1, n-2)
#pragma unroll
    for (int jj = 0; jj < HC; ++jj) {
      if (2 * jj + h == t) dgl = row[jj].re;
      if (2 * jj + h == n - 2) last = row[jj];
    }
    const int hd = t % 2;
    const double dglo = other(dgl);
    dgl = (h == hd) ? dgl : dglo;
    {
      const int src = 32 * ((n - 2) % 2) + n - 1;
      const zd l2 = zd{readlane_d(last.re, src), readlane_d(last.im, src)};
      if (lane == 0) {
        sub[n - 2] = l2;
        tauv[n - 2] = 0.0;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double* sd = side + k * side_len(n);
    zd* Uk = U + k * (int64_t)n * n;
    // reflectors: row t keeps v_t of every column kc < t
#pragma unroll
    for (int jj = 0; jj < HC; ++jj) {
      const int kc = 2 * jj + h;
      if (kc < t && t < n) Uk[t + (int64_t)n * kc] = row[jj];
    }
    if (h == 0 && t < n) {
      sd[t] = dgl;
      sd[2 * n + t] = (t < n - 1) ? tauv[t] : 0.0;
    }
    // |s_i| and the unit phase u_i = s_i / |s_i| (1 for s_i = 0) of every i at
    // once, then the running product ph_{i+1} = ph_i u_i in one lane: the same
    // operations as one serial loop over i (ph u with u = 1 is exact)
    if (lane + 1 < n) {
      const zd s_ = sub[lane];
      const double as = sqrt(zabs2(s_));
      sd[n + lane] = as;
      sub[lane] = (as > 0.0) ? zscale(s_, 1.0 / as) : zd{1.0, 0.0};
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane == 0) {
      zd ph = zd{1.0, 0.0};
      sd[3 * n] = 1.0;
      sd[3 * n + 1] = 0.0;
      for (int i = 0; i + 1 < n; ++i) {
        ph = zmul(ph, sub[i]);
        sd[3 * n + 2 * (i + 1)] = ph.re;
        sd[3 * n + 2 * (i + 1) + 1] = ph.im;
      }
      sd[n + n - 1] = 0.0;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Householder tridiagonalisation in FP64 for 16 < n <= 32, two k-points per
// wavefront: lane (t, g) = (lane % 32, lane / 32) keeps the whole row t of
// k-point 2 b + g in registers.  The cross-lane work of a column (its norm,
// |x0|^2, v^H p) runs once for both k-points, and p = A v needs no lane swap
// (the row is whole).  Columns j <= kc are skipped at compile time (full
// unroll); rows t <= kc carry v = w = 0 through the update, which leaves them
// unchanged.  Same outputs as eigh_house_kernel.
template <bool TAU>
__global__ __launch_bounds__(64, 2) void eigh_house_row2_kernel(int n, int64_t nk,
                                                              const zd* __restrict__ Hk,
                                                              int64_t strideHk, int packed,
                                                              const double* __restrict__ kpts,
                                                              const double* __restrict__ tau,
                                                              zd* __restrict__ U,
                                                              double* __restrict__ side) {
  constexpr int NMAX = 32, LDV = NMAX + 1;  // LDV: the two k-points' vectors in other banks
  __shared__ __attribute__((aligned(16))) zd vv[2 * LDV], ww[2 * LDV];
  __shared__ zd sub[2 * NMAX], fph[2 * NMAX];
  __shared__ double tauv[2 * NMAX];
  const double two_pi = 6.283185307179586476925;
  const zd zero = zd{0.0, 0.0};
  for (int64_t kb = 2 * (int64_t)blockIdx.x; kb < nk; kb += 2 * (int64_t)gridDim.x) {
    // the lane index is made opaque per pair: otherwise the compiler hoists
    // every step's lane masks out of this loop and spills them
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane));
    const int t = lane & 31, g = lane >> 5;
    const int64_t k = kb + g;
    const bool valid = k < nk;  // the last pair may hold one k-point
    zd* vg = vv + g * LDV;
    zd* wg = ww + g * LDV;
    // ---- row t: upper triangle from H(k), lower as conjugates; with
    // centres, element (t, j) takes f_j conj(f_t), f_j = e^{2 pi i k.tau_j}
    zd ft = zd{1.0, 0.0};
    if (TAU) {
      zd f = zd{1.0, 0.0};
      if (valid && t < n) {
        const double ang = two_pi * (kpts[3 * k] * tau[3 * t] + kpts[3 * k + 1] * tau[3 * t + 1] +
                                     kpts[3 * k + 2] * tau[3 * t + 2]);
        sin_cos(ang, &f.im, &f.re);
      }
      fph[g * NMAX + t] = f;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      ft = zconj(fph[g * NMAX + t]);
    }
    zd row[NMAX];
#pragma unroll
    for (int j = 0; j < NMAX; ++j) {
      zd x = zero;
      if (valid && t < n && j < n) {
        const int i0 = j >= t ? t : j, j0 = j >= t ? j : t;  // (i0 <= j0) element
        const int64_t o = packed ? i0 + ((int64_t)j0 * (j0 + 1)) / 2 : i0 + (int64_t)n * j0;
        x = Hk[k * strideHk + o];
        if (j < t) x = zconj(x);
        if (TAU && j != t) x = zmul(x, zmul(fph[g * NMAX + j], ft));
        if (j == t) x.im = 0.0;
      }
      row[j] = x;
    }
#pragma clang loop unroll(full)
    for (int kc = 0; kc < NMAX - 2; ++kc) {
      if (kc + 2 >= n) continue;  // uniform; every index below stays compile-time
      const bool below = t > kc;  // rows kc+1 .. n-1 take part
      const zd xo = row[kc];      // A(t, kc)
      const double a2 = zabs2(xo);
      const double tail = team_sum<32, double>(t > kc + 1 ? a2 : 0.0);
      const double ax02 = team_bcast<32>(a2, kc + 1);  // |A(kc+1, kc)|^2 of this k-point
      const bool refl = tail > 0.0;
      const double xn = sqrt_nr(tail + ax02);
      const double ax0 = sqrt_nr(ax02);
      // alpha is formed in every lane from its own A(t, kc); lane kc + 1's
      // (x0 = A(kc+1, kc)) is the one kept
      const zd alpha = !refl ? xo : ((ax0 > 0.0) ? zscale(xo, -xn * rcp_nr(ax0)) : zd{-xn, 0.0});
      zd v = (refl && below) ? xo : zero;
      if (t == kc + 1 && refl) v = zsub(xo, alpha);
      // |v|^2 = |x|^2 - 2 Re(conj(alpha) x0) + |alpha|^2 = 2 xn (xn + |x0|)
      const double vn2 = 2.0 * xn * (xn + ax0);
      const double tk = refl ? 2.0 * rcp_nr(vn2) : 0.0;
      stz(vg + t, v);
      if (t == kc + 1) {
        sub[g * NMAX + kc] = alpha;
        tauv[g * NMAX + kc] = tk;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // p_t = tk sum_{j > kc} A(t, j) v_j, four interleaved chains
      zd p4[4] = {zero, zero, zero, zero};
#pragma unroll
      for (int j = 0; j < NMAX; ++j) {
        if (j <= kc) continue;
        const zd a_ = row[j], b_ = ldz(vg + j);
        zd& q = p4[(j - kc - 1) & 3];
        if (j - kc - 1 < 4) {  // the first product of each chain sets it
          q.re = fma(a_.re, b_.re, -a_.im * b_.im);
          q.im = fma(a_.re, b_.im, a_.im * b_.re);
        } else {
          q.re = fma(a_.re, b_.re, fma(-a_.im, b_.im, q.re));
          q.im = fma(a_.re, b_.im, fma(a_.im, b_.re, q.im));
        }
      }
      zd p = zadd(zadd(p4[0], p4[1]), zadd(p4[2], p4[3]));
      p = below ? zscale(p, tk) : zero;
      const double K = 0.5 * tk * team_sum<32, double>(v.re * p.re + v.im * p.im);
      const zd w = zsub(p, zscale(v, K));
      stz(wg + t, w);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int j = 0; j < NMAX; ++j) {
        if (j <= kc) continue;
        // row_j -= v conj(w_j) + w conj(v_j)
        const zd vj = ldz(vg + j), wj = ldz(wg + j);
        zd r_ = row[j];
        r_.re = fma(-v.re, wj.re, fma(-v.im, wj.im, fma(-w.re, vj.re, fma(-w.im, vj.im, r_.re))));
        r_.im = fma(-v.im, wj.re, fma(v.re, wj.im, fma(-w.im, vj.re, fma(w.re, vj.im, r_.im))));
        row[j] = r_;
      }
      if (below) row[kc] = v;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // ---- outputs -----------------------------------------------------------
    double dgl = 0.0;  // A(t, t)
#pragma unroll
    for (int j = 0; j < NMAX; ++j)
      if (j == t) dgl = row[j].re;
    {
      zd last = zero;  // A(n-1, n-2), held by lane n - 1
#pragma unroll
      for (int j = 0; j < NMAX; ++j)
        if (j == n - 2) last = row[j];
      if (t == n - 1) {
        sub[g * NMAX + n - 2] = last;
        tauv[g * NMAX + n - 2] = 0.0;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (valid) {
      double* sd = side + k * side_len(n);
      zd* Uk = U + k * (int64_t)n * n;
      // reflectors: row t keeps v_t of every column kc < t
#pragma unroll
      for (int j = 0; j < NMAX; ++j)
        if (j < t && t < n) Uk[t + (int64_t)n * j] = row[j];
      if (t < n) {
        sd[t] = dgl;
        sd[2 * n + t] = (t < n - 1) ? tauv[g * NMAX + t] : 0.0;
      }
      if (t + 1 < n) {
        const zd s_ = sub[g * NMAX + t];
        const double as = sqrt(zabs2(s_));
        sd[n + t] = as;
        sub[g * NMAX + t] = (as > 0.0) ? zscale(s_, 1.0 / as) : zd{1.0, 0.0};
      }
    }
    // |s_i| and the unit phase u_i = s_i / |s_i| (1 for s_i = 0) of every i at
    // once, then the running product ph_{i+1} = ph_i u_i in one lane: the same
    // operations as one serial loop over i (ph u with u = 1 is exact)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (valid && t == 0) {
      double* sd = side + k * side_len(n);
      const zd* ug = sub + g * NMAX;
      zd ph = zd{1.0, 0.0};
      sd[3 * n] = 1.0;
      sd[3 * n + 1] = 0.0;
      for (int i = 0; i + 1 < n; ++i) {
        ph = zmul(ph, ug[i]);
        sd[3 * n + 2 * (i + 1)] = ph.re;
        sd[3 * n + 2 * (i + 1) + 1] = ph.im;
      }
      sd[n + n - 1] = 0.0;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// The QL recurrence of one k-point, run by one lane; d and e live in LDS
// columns ds[i * LANES], es[i * LANES] (already offset to the lane).
// rec = false: eigenvalues only, no rotation stored; the side record keeps
// d and e and its count slot receives -1 when every gap between eigenvalues
// is at least 1e-5 |T| (1e-3 |T| in FP32; the vectors are then computed
// directly from T) or -2 otherwise (the caller reruns the k-point with
// rec = true).
template <int LANES, class R, int NMAX = 0>
__device__ __attribute__((always_inline)) inline void ql_one(int n, int64_t k, R* ds, R* es, R* __restrict__ side,
                              R* __restrict__ rot, uint8_t* __restrict__ rcol,
                              double* __restrict__ E, int* __restrict__ info, bool rec = true) {
  const R EPS = cplx_of<R>::eps;
  const int64_t cap = rot_cap(n);
  const R* sd = side + k * side_len(n);
  for (int i = 0; i < n; ++i) {
    ds[i * LANES] = sd[i];
    es[i * LANES] = sd[n + i];
  }
  R* rk = rot + k * 2 * cap;
  uint8_t* ck = rcol + k * cap;
  int64_t pos = 0;
  int fail = 0;
  for (int l = 0; l < n && !fail; ++l) {
    int iter = 0;
    for (;;) {
      int m = l;
      for (; m < n - 1; ++m) {
        const R dd = fabs(ds[m * LANES]) + fabs(ds[(m + 1) * LANES]);
        if (fabs(es[m * LANES]) <= EPS * dd) break;
      }
      if (m == l) break;
      if (++iter > 60 || (rec && pos + (m - l) > cap)) {
        fail = 1;
        break;
      }
      const R dl = ds[l * LANES], el = es[l * LANES];
      R g = (ds[(l + 1) * LANES] - dl) / (R(2) * el);
      R r = sqrt(g * g + R(1));
      g = ds[m * LANES] - dl + el / (g + copysign(r, g));
      R s = R(1), c = R(1), p = R(0);
      bool underflow = false;
      R ei = es[(m - 1) * LANES], di = ds[(m - 1) * LANES];
      R di1 = ds[m * LANES];
      for (int i = m - 1; i >= l; --i) {
        // operands of the next step are not touched by this one
        const R ein = (i > l) ? es[(i - 1) * LANES] : R(0);
        const R din = (i > l) ? ds[(i - 1) * LANES] : R(0);
        const R f = s * ei, b = c * ei;
        const R r2 = f * f + g * g;
        R rr;
        if constexpr (sizeof(R) == 8) {
          // r and 1/r from one reciprocal square root: a shorter serial chain
          rr = rsq_nr(r2);
          r = (r2 > R(0)) ? r2 * rr : R(0);
        } else {
          r = sqrt(r2);
        }
        es[(i + 1) * LANES] = r;
        if (r == R(0)) {
          ds[(i + 1) * LANES] = di1 - p;
          es[m * LANES] = R(0);
          underflow = true;
          break;
        }
        if constexpr (sizeof(R) != 8) rr = R(1) / r;
        s = f * rr;
        c = g * rr;
        g = di1 - p;
        r = (di - g) * s + R(2) * c * b;
        p = s * r;
        ds[(i + 1) * LANES] = g + p;
        g = c * r - b;
        if (rec) {
          rk[2 * pos] = c;
          rk[2 * pos + 1] = s;
          ck[pos] = (uint8_t)i;
        }
        pos += 1;
        ei = ein;
        di1 = di;
        di = din;
      }
      if (underflow) continue;
      ds[l * LANES] -= p;
      es[l * LANES] = g;
      es[m * LANES] = R(0);
    }
  }
  if (fail) atomicOr(info, 1);
  R gap = R(1e30), big = R(0);
  if constexpr (NMAX > 0) {
    // n <= NMAX: the ranks from registers (compile-time indices), then the
    // eigenvalues in ascending order through the LDS column, and the
    // smallest gap between neighbours
    R dv[NMAX];
#pragma unroll
    for (int i = 0; i < NMAX; ++i) dv[i] = (i < n) ? ds[i * LANES] : R(0);
#pragma unroll
    for (int j = 0; j < NMAX; ++j) {
      if (j >= n) continue;  // fixed trip count: every index stays compile-time
      int rnk = 0;
#pragma unroll
      for (int i = 0; i < NMAX; ++i)
        if (i < n && i != j) rnk += (dv[i] < dv[j]) || (dv[i] == dv[j] && i < j);
      big = fmax(big, fabs(dv[j]));
      es[j * LANES] = (R)rnk;  // e is no longer needed: keep the rank
      ds[rnk * LANES] = dv[j];
    }
    for (int i = 0; i < n; ++i) {
      const R v = ds[i * LANES];
      E[k * n + i] = v;
      if (i > 0) gap = fmin(gap, v - ds[(i - 1) * LANES]);
    }
  } else {
  // ascending order: rank of each eigenvalue (ties by index); smallest gap
  for (int j = 0; j < n; ++j) {
    const R dj = ds[j * LANES];
    int rnk = 0;
    for (int i = 0; i < n; ++i) {
      const R di = ds[i * LANES];
      rnk += (di < dj) || (di == dj && i < j);
      if (i != j) gap = fmin(gap, fabs(di - dj));
    }
    big = fmax(big, fabs(dj));
    E[k * n + rnk] = dj;
    es[j * LANES] = (R)rnk;  // e is no longer needed: keep the rank
  }
  }
  R* sdw = side + k * side_len(n);
  if (!rec) {
    // FP32 vectors are refined in FP64 afterwards: a looser separation suffices
    const R sep = (sizeof(R) == 8) ? R(1e-5) : R(1e-3);
    sdw[5 * n] = (gap >= sep * big) ? R(-1) : R(-2);
    return;
  }
  for (int j = 0; j < n; ++j) sdw[n + j] = es[j * LANES];  // rank over e's slot
  sdw[5 * n] = (R)pos;
}

// One lane per k-point: implicit QL on the real tridiagonal (d, e) held in
// LDS columns.  Every Givens rotation G(i, i+1; c, s) of every sweep is
// appended, in application order, to the k-point's flat stream
// rot[k][q] = (c, s), col[k][q] = i; the count goes to the side record.
// Ascending eigenvalues go to E, each band's rank to the side record.
// mode 0: every k-point records its rotations; mode 1: eigenvalues and the
// separation verdict only; mode 2: the k-points with verdict -2 record.
template <int NMAX, class R>
// mode 1: every k-point, eigenvalues and the separation verdict; the
// k-points that need rotation records (count slot -2) are appended to
// list[] (order immaterial: each is solved on its own).  mode 2: the listed
// k-points again, recording the rotations.
__global__ __launch_bounds__(64) void eigh_ql_kernel(int n, int64_t nk,
                                                     R* __restrict__ side,
                                                     R* __restrict__ rot,
                                                     uint8_t* __restrict__ rcol,
                                                     double* __restrict__ E,
                                                     int* __restrict__ info, int mode,
                                                     int* __restrict__ list,
                                                     int* __restrict__ count) {
  __shared__ R ds[NMAX * 64], es[NMAX * 64];
  const int lane = threadIdx.x;
  if (mode == 2) {
    const int64_t cnt = *count;
    for (int64_t i = (int64_t)blockIdx.x * 64 + lane; i < cnt; i += (int64_t)gridDim.x * 64)
      ql_one<64, R, NMAX>(n, list[i], ds + lane, es + lane, side, rot, rcol, E, info, true);
    return;
  }
  for (int64_t kb = (int64_t)blockIdx.x * 64; kb < nk; kb += (int64_t)gridDim.x * 64) {
    const int64_t k = kb + lane;
    bool again = false;
    if (k < nk) {
      ql_one<64, R, NMAX>(n, k, ds + lane, es + lane, side, rot, rcol, E, info, false);
      again = side[k * side_len(n) + 5 * n] == R(-2);
    }
    const unsigned long long m = __ballot(again);
    if (m) {
      const int leader = __ffsll((long long)m) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(count, __popcll(m));
      base = __shfl(base, leader, 64);
      if (again) list[base + __popcll(m & ((1ull << lane) - 1))] = (int)k;
    }
  }
}

// Pass 1 of the QL stage (eigenvalues and the separation verdict) with d and
// e in registers: one lane per k-point, NMAX <= 32, every array index a
// compile-time constant, no LDS.  The active block always starts at index
// 0: when e[0] is negligible, d[0] is an eigenvalue and both arrays rotate
// down by one (d[0] moves to the end), so after NMAX deflations d[j] holds
// the j-th eigenvalue found, where ql_one's LDS column holds it.  Entries
// n .. NMAX-1 are decoupled padding (e = 0) that deflates as soon as it
// reaches index 0.  Each trip of the wave loop first runs the pending
// deflations of every lane and then one QL sweep, unrolled over the NMAX-1
// rotation steps: the steps at or beyond a lane's m are masked off (and
// skipped by the whole wave once no lane needs them).  Same arithmetic in
// the same order as ql_one: identical eigenvalues and verdicts.
template <int NMAX, class R>
__global__ __launch_bounds__(64) void eigh_ql_reg_kernel(int n, int64_t nk,
                                                         R* __restrict__ side,
                                                         double* __restrict__ E,
                                                         int* __restrict__ info,
                                                         int* __restrict__ list,
                                                         int* __restrict__ count) {
  const R EPS = cplx_of<R>::eps;
  const int lane = threadIdx.x;
  for (int64_t kb = (int64_t)blockIdx.x * 64; kb < nk; kb += (int64_t)gridDim.x * 64) {
    const int64_t k = kb + lane;
    const bool valid = k < nk;
    R d[NMAX], e[NMAX];
    {
      const R* sd = side + (valid ? k : (int64_t)0) * side_len(n);
#pragma unroll
      for (int i = 0; i < NMAX; ++i) {
        d[i] = (valid && i < n) ? sd[i] : R(0);
        e[i] = (valid && i < n) ? sd[n + i] : R(0);  // e[n-1] = 0 in the record
      }
    }
    int ndef = valid ? 0 : NMAX;  // eigenvalues found (rotations done)
    int iter = 0;
    bool fail = false;
    while (ndef < NMAX) {
      // deflations: d[0] has converged when e[0] is negligible
      for (;;) {
        const bool dfl = ndef < NMAX && fabs(e[0]) <= EPS * (fabs(d[0]) + fabs(d[1]));
        if (!dfl) break;
        const R d0 = d[0];
#pragma unroll
        for (int i = 0; i + 1 < NMAX; ++i) {
          d[i] = d[i + 1];
          e[i] = e[i + 1];
        }
        d[NMAX - 1] = d0;
        e[NMAX - 1] = R(0);
        ++ndef;
        iter = 0;
      }
      if (ndef >= NMAX) break;
      // m: the first negligible subdiagonal (the end of the active block,
      // whose e is zero, at the latest)
      int m = NMAX - 1;
#pragma unroll
      for (int i = NMAX - 2; i >= 1; --i)
        if (fabs(e[i]) <= EPS * (fabs(d[i]) + fabs(d[i + 1]))) m = i;
      if (++iter > 60) {
        fail = true;
        break;
      }
      R dm = d[NMAX - 1];
#pragma unroll
      for (int j = NMAX - 2; j >= 1; --j) dm = (j == m) ? d[j] : dm;
      const R dl = d[0], el = e[0];
      R g = (d[1] - dl) / (R(2) * el);
      R r = sqrt(g * g + R(1));
      g = dm - dl + el / (g + copysign(r, g));
      R s = R(1), c = R(1), p = R(0);
      bool run = true, underflow = false;
#pragma unroll
      for (int i = NMAX - 2; i >= 0; --i) {
        if (run && i < m) {
          const R f = s * e[i], b = c * e[i];
          const R r2 = f * f + g * g;
          R rr;
          if constexpr (sizeof(R) == 8) {
            rr = rsq_nr(r2);
            r = (r2 > R(0)) ? r2 * rr : R(0);
          } else {
            r = sqrt(r2);
          }
          e[i + 1] = r;
          if (r == R(0)) {
            d[i + 1] = d[i + 1] - p;
            underflow = true;
            run = false;
          } else {
            if constexpr (sizeof(R) != 8) rr = R(1) / r;
            s = f * rr;
            c = g * rr;
            g = d[i + 1] - p;
            r = (d[i] - g) * s + R(2) * c * b;
            p = s * r;
            d[i + 1] = g + p;
            g = c * r - b;
          }
        }
      }
      if (!underflow) {
        d[0] -= p;
        e[0] = g;
      }
#pragma unroll
      for (int j = 1; j < NMAX; ++j)
        if (j == m) e[j] = R(0);
    }
    if (fail) atomicOr(info, 1);
    bool again = false;
    if (valid) {
      // ascending order by rank (ties by position), the largest modulus and
      // the smallest gap between eigenvalues
      R gap = R(1e30), big = R(0);
#pragma unroll
      for (int j = 0; j < NMAX; ++j) {
        if (j >= n) continue;
        int rnk = 0;
#pragma unroll
        for (int i = 0; i < NMAX; ++i)
          if (i < n && i != j) {
            rnk += (d[i] < d[j]) || (d[i] == d[j] && i < j);
            if (i > j) gap = fmin(gap, fabs(d[i] - d[j]));
          }
        big = fmax(big, fabs(d[j]));
        E[k * n + rnk] = d[j];
      }
      // FP32 vectors are refined in FP64 afterwards: a looser separation suffices
      const R sep = (sizeof(R) == 8) ? R(1e-5) : R(1e-3);
      const R verdict = (gap >= sep * big) ? R(-1) : R(-2);
      side[k * side_len(n) + 5 * n] = verdict;
      again = verdict == R(-2);
    }
    const unsigned long long msk = __ballot(again);
    if (msk) {
      const int leader = __ffsll((long long)msk) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(count, __popcll(msk));
      base = __shfl(base, leader, 64);
      if (again) list[base + __popcll(msk & ((1ull << lane) - 1))] = (int)k;
    }
  }
}

__global__ void zero_count_kernel(int* count) { *count = 0; }

// the same for 32 < n <= 128: 32 lanes per workgroup, d and e in dynamic LDS
// (n x 32 each); modes as eigh_ql_kernel (1: eigenvalues and the separation
// verdict, listing the k-points that need rotation records; 2: the listed
// k-points, recording)
template <class R>
__global__ __launch_bounds__(32) void eigh_ql_wide_kernel(int n, int64_t nk,
                                                          R* __restrict__ side,
                                                          R* __restrict__ rot,
                                                          uint8_t* __restrict__ rcol,
                                                          double* __restrict__ E,
                                                          int* __restrict__ info, int mode,
                                                          int* __restrict__ list,
                                                          int* __restrict__ count) {
  extern __shared__ double ql_smem[];
  R* ds = reinterpret_cast<R*>(ql_smem);
  R* es = ds + (int64_t)n * 32;
  const int lane = threadIdx.x;
  if (mode == 2) {
    const int64_t cnt = *count;
    for (int64_t i = (int64_t)blockIdx.x * 32 + lane; i < cnt; i += (int64_t)gridDim.x * 32)
      ql_one<32, R>(n, list[i], ds + lane, es + lane, side, rot, rcol, E, info, true);
    return;
  }
  for (int64_t kb = (int64_t)blockIdx.x * 32; kb < nk; kb += (int64_t)gridDim.x * 32) {
    const int64_t k = kb + lane;
    bool again = false;
    if (k < nk) {
      ql_one<32, R>(n, k, ds + lane, es + lane, side, rot, rcol, E, info, false);
      again = side[k * side_len(n) + 5 * n] == R(-2);
    }
    const unsigned long long m = __ballot(again);
    if (m) {
      const int leader = __ffsll((long long)m) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(count, __popcll(m));
      base = __shfl(base, leader, 64);
      if (again) list[base + __popcll(m & ((1ull << lane) - 1))] = (int)k;
    }
  }
}

// The eigenvectors of the k-point's real tridiagonal T into
// Zt[col * (NMAX + 1) + row]: the listed k-points (separation verdict not
// set) replay their QL rotations on the identity, lane t <-> row t; the
// separated ones take lane t <-> eigenvector t from T directly.  csb and ixb
// are the team's staging slots.
template <int NMAX, class R>
__device__ __attribute__((always_inline)) inline void tri_vectors(
    int n, int t, bool valid, bool direct, int64_t k, int64_t cap, const R* sd,
    const R* __restrict__ rot, const uint8_t* __restrict__ rcol, const double* __restrict__ E,
    R* Zt, R* csb, int* ixb) {
  constexpr int LDZ = NMAX + 1;
  if (!direct) {
#pragma unroll
    for (int c = 0; c < NMAX; ++c) Zt[c * LDZ + t] = (c == t) ? R(1) : R(0);
  }
  const int64_t Q = (valid && !direct) ? (int64_t)sd[5 * n] : 0;
  const R* rk = rot + (valid ? k : 0) * 2 * cap;
  const uint8_t* ck = rcol + (valid ? k : 0) * cap;
  R nc = R(0), ns = R(0);
  int ni = 0;
  if (t < Q) {
    nc = rk[2 * t];
    ns = rk[2 * t + 1];
    ni = ck[t];
  }
  int prev = -2;
  R carry = R(0);
  for (int64_t q0 = 0; q0 < Q; q0 += NMAX) {
    __builtin_amdgcn_wave_barrier();
    csb[2 * t] = nc;
    csb[2 * t + 1] = ns;
    ixb[t] = ni;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int64_t qn = q0 + NMAX + t;
    if (qn < Q) {
      nc = rk[2 * qn];
      ns = rk[2 * qn + 1];
      ni = ck[qn];
    }
    const int cnt = (Q - q0 < NMAX) ? (int)(Q - q0) : NMAX;
    for (int r = 0; r < cnt; ++r) {
      const R c = csb[2 * r], s = csb[2 * r + 1];
      const int i = ixb[r];
      const R zi = Zt[i * LDZ + t];
      const R zi1 = (prev == i + 1) ? carry : Zt[(i + 1) * LDZ + t];
      Zt[(i + 1) * LDZ + t] = s * zi + c * zi1;
      carry = c * zi - s * zi1;
      Zt[i * LDZ + t] = carry;
      prev = i;
    }
  }
  if (direct) {
    // lane t: eigenvector of T = tridiag(e, d, e) for lambda_t (ascending),
    // from the twisted factorization at the index r of smallest |gamma_r|:
    //   dp_i = d_i - l - e_{i-1}^2 / dp_{i-1}   (top down)
    //   um_i = d_i - l - e_i^2 / um_{i+1}       (bottom up)
    //   gamma_i = dp_i + um_i - (d_i - l)
    //   z_r = 1, z_i = -e_i z_{i+1} / dp_i (i < r), z_i = -e_{i-1} z_{i-1} / um_i (i > r)
    // (Dhillon and Parlett's twisted factorizations); pivots below
    // eps |T| are clamped to +-eps |T|.
    const R lam = (t < n) ? (R)E[k * n + t] : R(0);
    // d and e of the k-point in the team's rotation-staging slot (unused
    // on this path): one coalesced load, then LDS broadcasts
    csb[t] = (t < n) ? sd[t] : R(0);
    csb[NMAX + t] = (t < n) ? sd[n + t] : R(0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const R* dl = csb;
    const R* el = csb + NMAX;
    R tn = R(0);
#pragma unroll
    for (int i = 0; i < NMAX; ++i)
      if (i < n) tn = fmax(tn, fabs(dl[i]) + R(2) * fabs(el[i]));
    const R pivmin = cplx_of<R>::eps * (tn > R(0) ? tn : R(1));
    R um[NMAX], dp[NMAX];
    R u = R(1);
#pragma unroll
    for (int i = NMAX - 1; i >= 0; --i) {
      R v = R(1);
      if (i < n) {
        const R ei = (i < n - 1) ? el[i] : R(0);
        v = dl[i] - lam - ((i < n - 1) ? ei * ei * rcp_nr(u) : R(0));
        if (fabs(v) < pivmin) v = (v < R(0)) ? -pivmin : pivmin;
        u = v;
      }
      um[i] = v;
    }
    R dd = R(1);
#pragma unroll
    for (int i = 0; i < NMAX; ++i) {
      R v = R(1);
      if (i < n) {
        const R em = (i > 0) ? el[i - 1] : R(0);
        v = dl[i] - lam - ((i > 0) ? em * em * rcp_nr(dd) : R(0));
        if (fabs(v) < pivmin) v = (v < R(0)) ? -pivmin : pivmin;
        dd = v;
      }
      dp[i] = v;
    }
    int r = 0;
    R gmin = R(0);
#pragma unroll
    for (int i = 0; i < NMAX; ++i)
      if (i < n) {
        const R g = fabs(dp[i] + um[i] - (dl[i] - lam));
        if (i == 0 || g < gmin) {
          gmin = g;
          r = i;
        }
      }
    R zp = R(1);
#pragma unroll
    for (int i = NMAX - 1; i >= 0; --i)
      if (i < r) {
        zp = -(el[i] * rcp_nr(dp[i])) * zp;
        dp[i] = zp;
      }
    zp = R(1);
#pragma unroll
    for (int i = 0; i < NMAX; ++i)
      if (i > r && i < n) {
        zp = -(el[i - 1] * rcp_nr(um[i])) * zp;
        um[i] = zp;
      }
    R nz = R(0);
#pragma unroll
    for (int i = 0; i < NMAX; ++i) {
      const R zi = (i >= n) ? R(0) : (i < r ? dp[i] : (i == r ? R(1) : um[i]));
      dp[i] = zi;
      nz += zi * zi;
    }
    const R inv = R(1) / sqrt(nz);
#pragma unroll
    for (int i = 0; i < NMAX; ++i) Zt[t * LDZ + i] = dp[i] * inv;
  }
}

// A team per k-point.  Z starts as the identity in LDS (lane r <-> row r);
// the rotation stream is staged through LDS one team-width chunk at a time,
// the next chunk's loads in flight while the current one is applied.  Then
// U = Q P Z: lane j takes column j of P Z into registers, the reflectors are
// staged into the Z space eight at a time as full zero-padded columns (the
// next eight loads in flight), and applied in reverse order without branches.
template <int NMAX, class R>
__global__ __launch_bounds__(64) void eigh_vec_kernel(int n, int64_t nk,
                                                      const R* __restrict__ side,
                                                      const R* __restrict__ rot,
                                                      const uint8_t* __restrict__ rcol,
                                                      const double* __restrict__ E,
                                                      typename cplx_of<R>::type* __restrict__ U) {
  using Z = typename cplx_of<R>::type;
  constexpr int TEAMS = 64 / NMAX;
  constexpr int LDZ = NMAX + 1;
  // Zt[col*LDZ + row]; later RCH reflectors at a time as full columns,
  // RCH x NMAX (18 KB per workgroup for n <= 32 in FP64: two wavefronts per
  // SIMD fit in the LDS)
  constexpr int RCH = 8;  // reflectors staged per chunk
  constexpr int ZS = (NMAX * LDZ > 2 * RCH * NMAX) ? NMAX * LDZ : 2 * RCH * NMAX;
  __shared__ R Zt_s[TEAMS][ZS];
  __shared__ R cs_s[TEAMS][4 * NMAX];  // staging slots; [2 NMAX, 4 NMAX): the phases
  __shared__ int ix_s[TEAMS][NMAX];
  const int64_t cap = rot_cap(n);
  for (int64_t kb = (int64_t)blockIdx.x * TEAMS; kb < nk; kb += (int64_t)gridDim.x * TEAMS) {
    // the lane index is made opaque per k-point: otherwise the compiler
    // hoists lane-dependent values out of this loop into spill registers
    int lane = threadIdx.x;
    asm volatile("" : "+v"(lane));
    const int tm = lane / NMAX, t = lane % NMAX;
    R* Zt = Zt_s[tm];
    R* csb = cs_s[tm];
    int* ixb = ix_s[tm];
    const int64_t k = kb + tm;
    const bool valid = k < nk;
    const R* sd = side + (valid ? k : 0) * side_len(n);
    // count slot -1: eigenvalues separated, no rotations recorded; the
    // eigenvectors of the real tridiagonal come straight from T below
    const bool direct = valid && sd[5 * n] < R(0);
    // the phases of P (one coalesced load, then LDS broadcasts)
    R* phb = csb + 2 * NMAX;
    if (t < n) {
      phb[2 * t] = sd[3 * n + 2 * t];
      phb[2 * t + 1] = sd[3 * n + 2 * t + 1];
    }
    tri_vectors<NMAX, R>(n, t, valid, direct, k, cap, sd, rot, rcol, E, Zt, csb, ixb);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // ---- lane j <-> column j of P Z ----------------------------------------
    Z W[NMAX];
#pragma unroll
    for (int i = 0; i < NMAX; ++i) {
      const int ic = i < n ? i : n - 1;
      const Z ph = Z{phb[2 * ic], phb[2 * ic + 1]};
      W[i] = (valid && i < n && t < n) ? zscale(ph, Zt[t * LDZ + i]) : Z{R(0), R(0)};
    }
    // tau of the reflectors into the (now free) rotation staging area
    R* taus = csb;
    if (t < n) taus[t] = sd[2 * n + t];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // ---- reflectors, eight at a time, as full columns (zero at and above
    // the diagonal and beyond n) in the Z space, so the updates run
    // branch-free; applied last to first, the next eight loads in flight
    Z* Vf = reinterpret_cast<Z*>(Zt);
    const Z* Uk = U + (valid ? k : 0) * (int64_t)n * n;
    const int tc = t < n ? t : n - 1;
    const int nchunk = n > 2 ? (n - 2 + RCH - 1) / RCH : 0;
    Z vb[RCH];
    auto load_chunk = [&](int c) {
#pragma unroll
      for (int u = 0; u < RCH; ++u) {
        const int kc = RCH * c + u;
        vb[u] = Uk[tc + (int64_t)n * (kc < n ? kc : n - 1)];
      }
    };
    if (nchunk > 0) load_chunk(nchunk - 1);
    // chunk and reflector indexes are compile-time (full unroll), so each
    // reflector's products start below its diagonal: v_i = 0 for i <= kc
#pragma clang loop unroll(full)
    for (int c = (NMAX - 2 + RCH - 1) / RCH - 1; c >= 0; --c) {
      if (c >= nchunk) continue;  // uniform
      // the previous chunk's (or the W set-up's) reads of the Z space are done
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int u = 0; u < RCH; ++u) {
        const int kc = RCH * c + u;
        Vf[u * NMAX + t] = (kc + 2 < n && t > kc && t < n) ? vb[u] : Z{R(0), R(0)};
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (c > 0) load_chunk(c - 1);
      if (valid) {
#pragma clang loop unroll(full)
        for (int u = RCH - 1; u >= 0; --u) {
          const int kc = RCH * c + u;
          if (kc + 2 >= n || kc + 2 >= NMAX) continue;  // uniform
          const Z* v = Vf + u * NMAX;
          const R tk = taus[kc];
          // four interleaved partial sums of v^H W, fused multiply-adds
          Z s4[4] = {Z{R(0), R(0)}, Z{R(0), R(0)}, Z{R(0), R(0)}, Z{R(0), R(0)}};
#pragma unroll
          for (int i = 0; i < NMAX; ++i) {
            if (i <= kc) continue;
            const Z vi = v[i], wi = W[i];
            Z& q = s4[i & 3];
            q.re = fma(vi.re, wi.re, fma(vi.im, wi.im, q.re));
            q.im = fma(vi.re, wi.im, fma(-vi.im, wi.re, q.im));
          }
          const Z sacc = zscale(zadd(zadd(s4[0], s4[1]), zadd(s4[2], s4[3])), tk);
          // v is read again from LDS (a broadcast) rather than held beside W:
          // holding both would double the registers and halve the waves
          asm volatile("" ::: "memory");
#pragma unroll
          for (int i = 0; i < NMAX; ++i) {  // W -= v sacc
            if (i <= kc) continue;
            const Z vi = v[i];
            W[i].re = fma(-vi.re, sacc.re, fma(vi.im, sacc.im, W[i].re));
            W[i].im = fma(-vi.re, sacc.im, fma(-vi.im, sacc.re, W[i].im));
          }
          asm volatile("" ::: "memory");  // the next reflector's loads wait
        }
      }
    }
    // gauge convention: largest-modulus component (first on ties) real > 0
    {
      R best = -R(1);
      Z u = Z{R(1), R(0)};
#pragma unroll
      for (int i = 0; i < NMAX; ++i) {
        const R m2 = zabs2(W[i]);
        if (i < n && m2 > best) {
          best = m2;
          u = W[i];
        }
      }
      const R au = sqrt(zabs2(u));
      if (au > R(0)) {
        const Z phs = Z{u.re / au, -u.im / au};
#pragma unroll
        for (int i = 0; i < NMAX; ++i) W[i] = zmul(W[i], phs);
      }
    }
    // columns to U's column of their rank, through the team's LDS area in
    // two halves (NMAX / 2 columns, pitch NMAX + 1), so that each store
    // writes one contiguous column per team instead of one element per column
    const int rnk = (valid && t < n) ? (direct ? t : (int)sd[n + t]) : 0;
    ixb[t] = rnk;
    Z* Wt = reinterpret_cast<Z*>(Zt);
    constexpr int HC = NMAX / 2;
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (t >= hh * HC && t < (hh + 1) * HC)
#pragma unroll
        for (int i = 0; i < NMAX; ++i) Wt[(t - hh * HC) * (NMAX + 1) + i] = W[i];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (valid && t < n) {
        Z* Uk_ = U + k * (int64_t)n * n + t;
#pragma unroll
        for (int s2 = 0; s2 < HC; ++s2) {
          const int col = hh * HC + s2;
          if (col < n) Uk_[(int64_t)n * ixb[col]] = Wt[s2 * (NMAX + 1) + t];
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// ===========================================================================
// 32 < n <= 128: one 256-thread workgroup per k-point for the reduction and
// the back-transformation; the QL stage is shared (32 lanes per workgroup).
// ===========================================================================

// packed lower triangle (column-major): element (i >= j) at lo_off(j) + i
__device__ inline int lo_off(int j, int n) { return j * n - (j * (j + 1)) / 2; }

// block-wide sum of one double over 256 threads (every thread gets it);
// red holds 4 doubles; the call ends with a barrier
__device__ inline double block_sum256(double v, double* red) {
  v = rows4_sum(row16_sum(v));  // wave sum on DPP and permlane swaps
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// Householder reduction with the Hermitian matrix as a packed lower
// triangle in LDS (n(n+1)/2 complex, 132 KB at n = 128).  Output as the
// narrow kernel: reflectors in the strict lower triangle of U's slot, d, the
// subdiagonal phases and tau in the side record.
__global__ __launch_bounds__(256) void eigh_house_wide_kernel(int n, int64_t nk,
                                                              const zd* __restrict__ Hk,
                                                              int64_t strideHk, int packed,
                                                              const double* __restrict__ kpts,
                                                              const double* __restrict__ tau,
                                                              zd* __restrict__ U,
                                                              double* __restrict__ side) {
  extern __shared__ double hw_smem[];
  const int np = n * (n + 1) / 2;
  zd* Lw = reinterpret_cast<zd*>(hw_smem);
  zd* vv = Lw + np;
  zd* ww = vv + n;
  zd* pp = ww + n;    // [4][n] partial products
  zd* sub = pp + 4 * n;
  double* tauv = reinterpret_cast<double*>(sub + n);
  double* red = tauv + n;  // 4
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const double two_pi = 6.283185307179586476925;
  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    __syncthreads();
    // lower triangle: L(i, j) = conj(H(j, i)) with the tight-binding phase
    for (int idx = t; idx < n * n; idx += 256) {
      const int i = idx % n, j = idx / n;  // H(j, i), j <= i
      if (j > i) continue;
      const int64_t o = packed ? j + ((int64_t)i * (i + 1)) / 2 : j + (int64_t)n * i;
      zd h = Hk[k * strideHk + o];
      if (tau && i != j) {
        const double kx = kpts[3 * k], ky = kpts[3 * k + 1], kz = kpts[3 * k + 2];
        const double ang = two_pi * (kx * (tau[3 * i] - tau[3 * j]) +
                                     ky * (tau[3 * i + 1] - tau[3 * j + 1]) +
                                     kz * (tau[3 * i + 2] - tau[3 * j + 2]));
        double sn, cs;
        sincos(ang, &sn, &cs);
        h = zmul(zd{cs, sn}, h);
      }
      Lw[lo_off(j, n) + i] = (i == j) ? zd{h.re, 0.0} : zconj(h);
    }
    __syncthreads();
    for (int kc = 0; kc + 2 < n; ++kc) {
      const int m = n - kc - 1;
      zd* col = Lw + lo_off(kc, n) + kc + 1;  // x_r = col[r], r < m
      const zd x = (t < m) ? col[t] : zd{0.0, 0.0};
      const double tail2 = block_sum256((t >= 1 && t < m) ? zabs2(x) : 0.0, red);
      const zd x0 = col[0];
      const bool refl = tail2 > 0.0;
      const double ax02 = zabs2(x0);
      const double xn = sqrt_nr(tail2 + ax02);
      const double ax0 = sqrt_nr(ax02);
      const zd alpha = !refl ? x0 : ((ax0 > 0.0) ? zscale(x0, -xn / ax0) : zd{-xn, 0.0});
      const zd v0 = zsub(x0, alpha);
      const double vn2 = tail2 + zabs2(v0);
      const double tk = refl ? 2.0 / vn2 : 0.0;
      if (t < m) vv[t] = refl ? ((t == 0) ? v0 : x) : zd{0.0, 0.0};
      __syncthreads();
      // p = A v over the trailing block; wave wv takes columns c = wv mod 4
      const int base = kc + 1;
      for (int r = lane; r < m; r += 64) {
        const int row = base + r;
        const int lrow = lo_off(row, n);
        // A(row, cl): L(row, cl) below the diagonal, conj L(cl, row) above;
        // one load from a selected address (no branch), two chains
        auto elem = [&](int cl) {
          const bool lower = row >= cl;
          zd a = Lw[lower ? lo_off(cl, n) + row : lrow + cl];
          a.im = lower ? a.im : -a.im;
          return a;
        };
        zd acc0 = zd{0.0, 0.0}, acc1 = zd{0.0, 0.0};  // c = wv + 8j, wv + 4 + 8j
        zd acc2 = zd{0.0, 0.0}, acc3 = zd{0.0, 0.0};  // c = wv + 8 + 16j, wv + 12 + 16j
        int c = wv;
        for (; c + 12 < m; c += 16) {
          const zd a = elem(base + c), a1 = elem(base + c + 4);
          const zd a2 = elem(base + c + 8), a3 = elem(base + c + 12);
          acc0 = zadd(acc0, zmul(a, vv[c]));
          acc1 = zadd(acc1, zmul(a1, vv[c + 4]));
          acc2 = zadd(acc2, zmul(a2, vv[c + 8]));
          acc3 = zadd(acc3, zmul(a3, vv[c + 12]));
        }
        for (; c + 4 < m; c += 8) {
          const zd a = elem(base + c), a1 = elem(base + c + 4);
          acc0 = zadd(acc0, zmul(a, vv[c]));
          acc1 = zadd(acc1, zmul(a1, vv[c + 4]));
        }
        if (c < m) acc0 = zadd(acc0, zmul(elem(base + c), vv[c]));
        pp[wv * n + r] = zadd(zadd(acc0, acc1), zadd(acc2, acc3));
      }
      __syncthreads();
      zd p = zd{0.0, 0.0};
      if (t < m) {
        p = zadd(zadd(pp[t], pp[n + t]), zadd(pp[2 * n + t], pp[3 * n + t]));
        p = zscale(p, tk);
      }
      const zd vt = (t < m) ? vv[t] : zd{0.0, 0.0};
      const double K = 0.5 * tk * block_sum256((t < m) ? (vt.re * p.re + vt.im * p.im) : 0.0, red);
      if (t < m) ww[t] = zsub(p, zscale(vt, K));
      __syncthreads();
      if (refl)
        for (int c = wv; c < m; c += 4) {
          const zd vc = vv[c], wc = ww[c];
          zd* colc = Lw + lo_off(base + c, n) + base;
          for (int r = c + lane; r < m; r += 64) {
            const zd upd = zadd(zmul(vv[r], zconj(wc)), zmul(ww[r], zconj(vc)));
            colc[r] = zsub(colc[r], upd);
          }
        }
      __syncthreads();
      // no barrier after this: each thread rewrites only its own vv[t] next,
      // and the next step reads columns > kc only
      if (t < m) col[t] = vv[t];
      if (t == 0) {
        sub[kc] = alpha;
        tauv[kc] = tk;
      }
    }
    if (t == 0 && n >= 2) {
      sub[n - 2] = Lw[lo_off(n - 2, n) + n - 1];
      tauv[n - 2] = 0.0;
    }
    __syncthreads();
    double* sd = side + k * side_len(n);
    zd* Uk = U + k * (int64_t)n * n;
    for (int idx = t; idx < n * n; idx += 256) {
      const int i = idx % n, j = idx / n;
      if (i > j) Uk[idx] = Lw[lo_off(j, n) + i];
    }
    if (t < n) {
      sd[t] = Lw[lo_off(t, n) + t].re;
      sd[2 * n + t] = (t < n - 1) ? tauv[t] : 0.0;
    }
    // |s_i| and the unit phase u_i = s_i / |s_i| (1 for s_i = 0) of every i at
    // once, then the running product ph_{i+1} = ph_i u_i in one lane: the same
    // operations as one serial loop over i (ph u with u = 1 is exact)
    if (t + 1 < n) {
      const zd s_ = sub[t];
      const double as = sqrt(zabs2(s_));
      sd[n + t] = as;
      sub[t] = (as > 0.0) ? zscale(s_, 1.0 / as) : zd{1.0, 0.0};
    }
    __syncthreads();
    if (t == 0) {
      zd ph = zd{1.0, 0.0};
      sd[3 * n] = 1.0;
      sd[3 * n + 1] = 0.0;
      for (int i = 0; i + 1 < n; ++i) {
        ph = zmul(ph, sub[i]);
        sd[3 * n + 2 * (i + 1)] = ph.re;
        sd[3 * n + 2 * (i + 1) + 1] = ph.im;
      }
      sd[n + n - 1] = 0.0;
    }
  }
}

// Householder reduction for 32 < n <= 64 with one k-point per wavefront:
// lane t <-> row t, the full Hermitian matrix in LDS column-major (pitch n,
// so a column is one contiguous 16-byte-per-lane access), no workgroup
// barrier.  Per column kc: the tail norm and v^H p are wave sums, p = A v
// runs over columns j > kc in four chains, and the rank-2 update writes the
// rows t > kc of those columns.  Same outputs as eigh_house_wide_kernel.
__global__ __launch_bounds__(64) void eigh_house_lds_kernel(int n, int64_t nk,
                                                            const zd* __restrict__ Hk,
                                                            int64_t strideHk, int packed,
                                                            const double* __restrict__ kpts,
                                                            const double* __restrict__ tau,
                                                            zd* __restrict__ U,
                                                            double* __restrict__ side) {
  extern __shared__ __attribute__((aligned(16))) double hl_smem[];
  zd* A = reinterpret_cast<zd*>(hl_smem);  // A(r, c) at A[r + n c]
  zd* vv = A + (int64_t)n * n;
  zd* ww = vv + 64;
  zd* sub = ww + 64;
  zd* fph = sub + 64;
  double* tauv = reinterpret_cast<double*>(fph + 64);
  const int t = threadIdx.x;
  const double two_pi = 6.283185307179586476925;
  const zd zero = zd{0.0, 0.0};
  auto wave_sum = [](double x) { return rows4_sum(row16_sum(x)); };
  auto sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    // with centres, element (i, j) takes f_j conj(f_i), f_j = e^{2 pi i k.tau_j}
    if (tau) {
      zd f = zd{1.0, 0.0};
      if (t < n) {
        const double ang = two_pi * (kpts[3 * k] * tau[3 * t] + kpts[3 * k + 1] * tau[3 * t + 1] +
                                     kpts[3 * k + 2] * tau[3 * t + 2]);
        sin_cos(ang, &f.im, &f.re);
      }
      fph[t] = f;
      sync();
    }
    for (int idx = t; idx < n * n; idx += 64) {
      const int i = idx % n, j = idx / n;
      const int i0 = i <= j ? i : j, j0 = i <= j ? j : i;  // (i0 <= j0) element
      const int64_t o = packed ? i0 + ((int64_t)j0 * (j0 + 1)) / 2 : i0 + (int64_t)n * j0;
      zd x = Hk[k * strideHk + o];
      if (j < i) x = zconj(x);
      if (tau && i != j) x = zmul(x, zmul(fph[j], zconj(fph[i])));
      if (i == j) x.im = 0.0;
      stz(A + idx, x);
    }
    sync();
    for (int kc = 0; kc + 2 < n; ++kc) {
      const bool below = t > kc && t < n;  // rows kc+1 .. n-1 take part
      const zd xo = (t < n) ? ldz(A + t + (int64_t)n * kc) : zero;
      const double a2 = zabs2(xo);
      const double tail = wave_sum(t > kc + 1 ? a2 : 0.0);
      const double ax02 = readlane_d(a2, kc + 1);
      const bool refl = tail > 0.0;
      const double xn = sqrt_nr(tail + ax02);
      const double ax0 = sqrt_nr(ax02);
      // lane kc + 1 (x0 = A(kc+1, kc)) keeps its alpha
      const zd alpha = !refl ? xo : ((ax0 > 0.0) ? zscale(xo, -xn * rcp_nr(ax0)) : zd{-xn, 0.0});
      zd v = (refl && below) ? xo : zero;
      if (t == kc + 1 && refl) v = zsub(xo, alpha);
      // |v|^2 = |x|^2 - 2 Re(conj(alpha) x0) + |alpha|^2 = 2 xn (xn + |x0|)
      const double vn2 = 2.0 * xn * (xn + ax0);
      const double tk = refl ? 2.0 * rcp_nr(vn2) : 0.0;
      stz(vv + t, v);
      if (t == kc + 1) {
        sub[kc] = alpha;
        tauv[kc] = tk;
      }
      sync();
      zd p4[4] = {zero, zero, zero, zero};
      if (below) {
        const zd* col = A + t;
        int j = kc + 1;
        for (; j + 3 < n; j += 4) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const zd a_ = ldz(col + (int64_t)n * (j + u)), b_ = ldz(vv + j + u);
            p4[u].re = fma(a_.re, b_.re, fma(-a_.im, b_.im, p4[u].re));
            p4[u].im = fma(a_.re, b_.im, fma(a_.im, b_.re, p4[u].im));
          }
        }
        for (; j < n; ++j) {
          const zd a_ = ldz(col + (int64_t)n * j), b_ = ldz(vv + j);
          p4[0].re = fma(a_.re, b_.re, fma(-a_.im, b_.im, p4[0].re));
          p4[0].im = fma(a_.re, b_.im, fma(a_.im, b_.re, p4[0].im));
        }
      }
      zd p = zadd(zadd(p4[0], p4[1]), zadd(p4[2], p4[3]));
      p = below ? zscale(p, tk) : zero;
      const double K = 0.5 * tk * wave_sum(v.re * p.re + v.im * p.im);
      const zd w = zsub(p, zscale(v, K));
      stz(ww + t, w);
      sync();
      if (refl && below) {
        zd* col = A + t;
        for (int j = kc + 1; j < n; ++j) {
          // A(t, j) -= v conj(w_j) + w conj(v_j)
          const zd vj = ldz(vv + j), wj = ldz(ww + j);
          zd r_ = ldz(col + (int64_t)n * j);
          r_.re = fma(-v.re, wj.re, fma(-v.im, wj.im, fma(-w.re, vj.re, fma(-w.im, vj.im, r_.re))));
          r_.im = fma(-v.im, wj.re, fma(v.re, wj.im, fma(-w.im, vj.re, fma(w.re, vj.im, r_.im))));
          stz(col + (int64_t)n * j, r_);
        }
      }
      if (below) stz(A + t + (int64_t)n * kc, v);
      sync();
    }
    if (t == 0) {
      sub[n - 2] = ldz(A + (n - 1) + (int64_t)n * (n - 2));
      tauv[n - 2] = 0.0;
    }
    sync();
    double* sd = side + k * side_len(n);
    zd* Uk = U + k * (int64_t)n * n;
    // reflectors: the strict lower triangle
    for (int idx = t; idx < n * n; idx += 64) {
      const int i = idx % n, j = idx / n;
      if (i > j) Uk[idx] = ldz(A + idx);
    }
    if (t < n) {
      sd[t] = A[t + (int64_t)n * t].re;
      sd[2 * n + t] = (t < n - 1) ? tauv[t] : 0.0;
    }
    // |s_i| and the unit phase u_i = s_i / |s_i| (1 for s_i = 0) of every i at
    // once, then the running product ph_{i+1} = ph_i u_i in one lane: the same
    // operations as one serial loop over i (ph u with u = 1 is exact)
    if (t + 1 < n) {
      const zd s_ = sub[t];
      const double as = sqrt(zabs2(s_));
      sd[n + t] = as;
      sub[t] = (as > 0.0) ? zscale(s_, 1.0 / as) : zd{1.0, 0.0};
    }
    sync();
    if (t == 0) {
      zd ph = zd{1.0, 0.0};
      sd[3 * n] = 1.0;
      sd[3 * n + 1] = 0.0;
      for (int i = 0; i + 1 < n; ++i) {
        ph = zmul(ph, sub[i]);
        sd[3 * n + 2 * (i + 1)] = ph.re;
        sd[3 * n + 2 * (i + 1) + 1] = ph.im;
      }
      sd[n + n - 1] = 0.0;
    }
    sync();
  }
}

// a or b by value (a conditional on class lvalues selects addresses, which
// keeps register arrays in scratch)
__device__ inline zd zsel(bool c, zd a, zd b) { return zd{c ? a.re : b.re, c ? a.im : b.im}; }

// Householder reduction for 32 < n <= NMAX (NMAX = 64, 80, 96, 128) with the
// matrix in registers: a workgroup of NMAX / 16 wavefronts per k-point,
// lane 4 r + q of wave w holding row i = 16 w + r, columns j = 4 (jj + s) + q
// (jj < NMAX / 4), where s = kc / 4 for the column kc being reduced: every
// four columns the row arrays shift down one entry, so column kc always sits
// at entry 0 of the lanes q = kc % 4 and every register index stays a
// compile-time constant.  Per column: the owners write x = A(., kc) to LDS
// with the tail norm (block sum), every thread forms v and tau, the four
// lanes of a row sum their parts of p = tau A v (DPP), a block sum gives
// v^H p, and the update A -= v w^H + w v^H runs over the entries that still
// hold columns > kc - 4 (uniform skip); rows of finished waves skip it too.
// v goes to U's column kc as it is made.  Three barriers per column (LDS
// buffers alternate between columns).  Same outputs as the other kernels.
template <int NMAX>
__global__ __launch_bounds__(4 * NMAX, NMAX <= 96 ? 3 : 2) void eigh_house_rows_kernel(int n, int64_t nk,
                                                                  const zd* __restrict__ Hk,
                                                                  int64_t strideHk, int packed,
                                                                  const double* __restrict__ kpts,
                                                                  const double* __restrict__ tau,
                                                                  zd* __restrict__ U,
                                                                  double* __restrict__ side) {
  constexpr int NW = NMAX / 16, NC = NMAX / 4;
  __shared__ __attribute__((aligned(16))) zd vv[2][NMAX];
  __shared__ __attribute__((aligned(16))) zd ww[2][NMAX];
  __shared__ __attribute__((aligned(16))) zd sub[NMAX];
  __shared__ __attribute__((aligned(16))) zd fph[NMAX];
  __shared__ double red[2][2][NW];
  const double two_pi = 6.283185307179586476925;
  const zd zero = zd{0.0, 0.0};
  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    const int lane = t & 63, wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const int q = lane & 3, i = 16 * wv + (lane >> 2);
    double* sd = side + k * side_len(n);
    zd* Uk = U + k * (int64_t)n * n;
    __syncthreads();  // the previous k-point's reads of sub and fph are done
    if (tau) {
      if (t < n) {
        zd f;
        const double ang = two_pi * (kpts[3 * k] * tau[3 * t] + kpts[3 * k + 1] * tau[3 * t + 1] +
                                     kpts[3 * k + 2] * tau[3 * t + 2]);
        sin_cos(ang, &f.im, &f.re);
        fph[t] = f;
      }
      __syncthreads();
    }
    zd row[NC];
    {
      const zd fi = zsel(tau && i < n, zconj(fph[i < n ? i : 0]), zd{1.0, 0.0});
#pragma unroll
      for (int jj = 0; jj < NC; ++jj) {
        const int j = 4 * jj + q;
        zd x = zero;
        if (i < n && j < n) {
          const int i0 = j >= i ? i : j, j0 = j >= i ? j : i;  // (i0 <= j0) element
          const int64_t o = packed ? i0 + ((int64_t)j0 * (j0 + 1)) / 2 : i0 + (int64_t)n * j0;
          x = Hk[k * strideHk + o];
          if (j < i) x = zconj(x);
          if (tau && j != i) x = zmul(x, zmul(fph[j], fi));
          if (j == i) x.im = 0.0;
        }
        row[jj] = x;
      }
    }
    for (int kc = 0; kc < n; ++kc) {
      const int s = kc >> 2, m = kc & 3, b = kc & 1;
      if (m == 0 && kc > 0) {
        // the empty asm keeps the copies element-wise (as a block move the
        // array would leave the registers)
#pragma unroll
        for (int jj = 0; jj + 1 < NC; ++jj) {
          zd x = row[jj + 1];
          asm volatile("" : "+v"(x.re), "+v"(x.im));
          row[jj] = x;
        }
        row[NC - 1] = zero;
      }
      const bool own = q == m;  // this lane holds A(i, kc) in row[0]
      const zd xo = zsel(own, row[0], zero);
      if (own && i == kc) sd[kc] = xo.re;  // the diagonal is final here
      if (kc == n - 2 && own && i == n - 1) sub[n - 2] = xo;
      if (kc + 2 >= n) continue;  // uniform
      const bool below = i > kc && i < n;
      // x (rows > kc) to LDS, the tail norm over rows > kc + 1
      if (own) vv[b][i] = zsel(below, xo, zero);
      {
        double part = (own && i > kc + 1 && i < n) ? zabs2(xo) : 0.0;
        part = rows4_sum(row16_sum(part));
        if (lane == 0) red[b][0][wv] = part;
      }
      __syncthreads();  // (1)
      double tail = 0.0;
#pragma unroll
      for (int w2 = 0; w2 < NW; ++w2) tail += red[b][0][w2];
      const zd x0 = vv[b][kc + 1];
      const bool refl = tail > 0.0;
      const double ax02 = zabs2(x0);
      const double xn = sqrt_nr(tail + ax02);
      const double ax0 = sqrt_nr(ax02);
      const zd alpha = zsel(!refl, x0, zsel(ax0 > 0.0, zscale(x0, -xn * rcp_nr(ax0)), zd{-xn, 0.0}));
      const zd v0 = zsub(x0, alpha);
      const double vn2 = 2.0 * xn * (xn + ax0);
      const double tk = refl ? 2.0 * rcp_nr(vn2) : 0.0;
      auto vget = [&](int j) {  // v_j (zero at and above kc, and without a reflection)
        const zd x = vv[b][j];
        return zsel(!refl, zero, zsel(j == kc + 1, v0, x));
      };
      const zd vi = zsel(i < n, vget(i), zero);
      const bool active = 16 * wv + 15 > kc;  // uniform: the wave has rows > kc
      zd p = zero;
      if (active) {
        zd p2[2] = {zero, zero};
#pragma unroll
        for (int jj = 0; jj < NC; ++jj) {
          if (jj >= NC - s) continue;  // uniform: entries past the shifted end
          const int j = 4 * (jj + s) + q;
          const zd a_ = row[jj], b_ = vget(j);
          zd& c = p2[jj & 1];
          c.re = fma(a_.re, b_.re, fma(-a_.im, b_.im, c.re));
          c.im = fma(a_.re, b_.im, fma(a_.im, b_.re, c.im));
        }
        p = zadd(p2[0], p2[1]);
      }
      p.re += dpp_d<0xB1>(p.re);  // the four lanes of the row
      p.im += dpp_d<0xB1>(p.im);
      p.re += dpp_d<0x4E>(p.re);
      p.im += dpp_d<0x4E>(p.im);
      p = zsel(below, zscale(p, tk), zero);
      {
        double part = (q == 0) ? vi.re * p.re + vi.im * p.im : 0.0;
        part = rows4_sum(row16_sum(part));
        if (lane == 0) red[b][1][wv] = part;
      }
      __syncthreads();  // (2)
      double vp = 0.0;
#pragma unroll
      for (int w2 = 0; w2 < NW; ++w2) vp += red[b][1][w2];
      const double K = 0.5 * tk * vp;
      const zd wi_ = zsub(p, zscale(vi, K));
      if (q == 0) ww[b][i] = wi_;
      // v to U's column kc (rows > kc), as the other kernels leave it
      if (own && below) Uk[i + (int64_t)n * kc] = vi;
      if (t == 0) {
        sub[kc] = alpha;
        sd[2 * n + kc] = tk;
      }
      __syncthreads();  // (3)
      if (refl && active) {
#pragma unroll
        for (int jj = 0; jj < NC; ++jj) {
          if (jj >= NC - s) continue;  // uniform
          const int j = 4 * (jj + s) + q;
          // A(i, j) -= v_i conj(w_j) + w_i conj(v_j)  (zero for j <= kc or i <= kc)
          const zd vj = vget(j), wj = ww[b][j];
          zd r_ = row[jj];
          r_.re = fma(-vi.re, wj.re, fma(-vi.im, wj.im, fma(-wi_.re, vj.re, fma(-wi_.im, vj.im, r_.re))));
          r_.im = fma(-vi.im, wj.re, fma(vi.re, wj.im, fma(-wi_.im, vj.re, fma(wi_.re, vj.im, r_.im))));
          row[jj] = r_;
        }
      }
    }
    __syncthreads();  // sub complete
    if (t == 0) {
      sd[2 * n + n - 2] = 0.0;
      sd[2 * n + n - 1] = 0.0;
    }
    // |s_i| and the unit phase u_i = s_i / |s_i| (1 for s_i = 0) of every i at
    // once, then the running product ph_{i+1} = ph_i u_i in one lane: the same
    // operations as one serial loop over i (ph u with u = 1 is exact)
    if (t + 1 < n) {
      const zd s_ = sub[t];
      const double as = sqrt(zabs2(s_));
      sd[n + t] = as;
      sub[t] = (as > 0.0) ? zscale(s_, 1.0 / as) : zd{1.0, 0.0};
    }
    __syncthreads();
    if (t == 0) {
      zd ph = zd{1.0, 0.0};
      sd[3 * n] = 1.0;
      sd[3 * n + 1] = 0.0;
      for (int i2 = 0; i2 + 1 < n; ++i2) {
        ph = zmul(ph, sub[i2]);
        sd[3 * n + 2 * (i2 + 1)] = ph.re;
        sd[3 * n + 2 * (i2 + 1) + 1] = ph.im;
      }
      sd[n + n - 1] = 0.0;
    }
  }
}

size_t eigh_house_lds_bytes(int n) {
  return sizeof(zd) * ((size_t)n * n + 4 * 64) + sizeof(double) * 64;
}

size_t eigh_house_wide_lds(int n) {
  return sizeof(zd) * ((size_t)n * (n + 1) / 2 + 8 * (size_t)n) + sizeof(double) * ((size_t)n + 4);
}

// Eigenvectors Z of the real tridiagonal for 32 < n <= 128, Z in LDS
// (column-major with odd pitch).  Separated k-points (count slot -1): thread
// j computes column j, the vector of lambda_j, from the twisted factorization
// of T - lambda_j (as eigh_vec_kernel; its O(n) work arrays live in the
// thread's own column of Z), and the ranks become the identity.  Otherwise
// thread r <-> row r replays the recorded rotations, each wave staging the
// stream through its own LDS slots.  Then W = P Z is written row-major into
// the k-point's rotation record (consumed, or unused).
__global__ __launch_bounds__(128) void eigh_replay_wide_kernel(int n, int64_t nk,
                                                               double* __restrict__ side,
                                                               double* __restrict__ rot,
                                                               const uint8_t* __restrict__ rcol,
                                                               const double* __restrict__ E) {
  extern __shared__ double vw_smem[];
  const int LDZ = n | 1;
  double* Zt = vw_smem;                       // [n][LDZ]
  double* csb_all = Zt + (int64_t)n * LDZ;    // [waves][2*64]
  int* ixb_all = reinterpret_cast<int*>(csb_all + 2 * 64 * 2);
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  double* csb = csb_all + 2 * 64 * wv;
  int* ixb = ixb_all + 64 * wv;
  const int64_t cap = rot_cap(n);
  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    __syncthreads();
    double* sd = side + k * side_len(n);
    if (sd[5 * n] < 0.0) {  // uniform over the workgroup
      // d and e of the k-point in the staging slots (2n <= 256 doubles)
      double* dl = csb_all;
      double* el = csb_all + n;
      for (int i = t; i < n; i += blockDim.x) {
        dl[i] = sd[i];
        el[i] = sd[n + i];
      }
      __syncthreads();
      if (t < n) {
        double* z = Zt + (int64_t)t * LDZ;  // column t: work array, then the vector
        const double lam = E[k * n + t];
        double tn = 0.0;
        for (int i = 0; i < n; ++i) tn = fmax(tn, fabs(dl[i]) + 2.0 * fabs(el[i]));
        const double pivmin = cplx_of<double>::eps * (tn > 0.0 ? tn : 1.0);
        // dp_i = d_i - l - e_{i-1}^2 / dp_{i-1}  (top down), into z
        double dd = 1.0;
        for (int i = 0; i < n; ++i) {
          double v = dl[i] - lam - ((i > 0) ? el[i - 1] * el[i - 1] / dd : 0.0);
          if (fabs(v) < pivmin) v = (v < 0.0) ? -pivmin : pivmin;
          dd = v;
          z[i] = v;
        }
        // um_i = d_i - l - e_i^2 / um_{i+1} (bottom up); r = the first index
        // of smallest |gamma_i| = |dp_i + um_i - (d_i - l)|
        double u = 1.0, gmin = 0.0;
        int r = n - 1;
        for (int i = n - 1; i >= 0; --i) {
          double v = dl[i] - lam - ((i < n - 1) ? el[i] * el[i] / u : 0.0);
          if (fabs(v) < pivmin) v = (v < 0.0) ? -pivmin : pivmin;
          u = v;
          const double g = fabs(z[i] + v - (dl[i] - lam));
          if (i == n - 1 || g <= gmin) {
            gmin = g;
            r = i;
          }
        }
        // z_i = -e_i z_{i+1} / dp_i below r (dp_i still in z)
        double zp = 1.0, nz = 1.0;
        for (int i = r - 1; i >= 0; --i) {
          zp = -(el[i] / z[i]) * zp;
          z[i] = zp;
          nz += zp * zp;
        }
        // um_i again above r (into z), then z_i = -e_{i-1} z_{i-1} / um_i
        u = 1.0;
        for (int i = n - 1; i > r; --i) {
          double v = dl[i] - lam - ((i < n - 1) ? el[i] * el[i] / u : 0.0);
          if (fabs(v) < pivmin) v = (v < 0.0) ? -pivmin : pivmin;
          u = v;
          z[i] = v;
        }
        zp = 1.0;
        for (int i = r + 1; i < n; ++i) {
          zp = -(el[i - 1] / z[i]) * zp;
          z[i] = zp;
          nz += zp * zp;
        }
        z[r] = 1.0;
        const double inv = 1.0 / sqrt(nz);
        for (int i = 0; i < n; ++i) z[i] *= inv;
      }
      __syncthreads();  // every thread has read e: its slot takes the ranks
      if (t < n) sd[n + t] = (double)t;
    } else {
    if (t < n)
      for (int c = 0; c < n; ++c) Zt[c * LDZ + t] = (c == t) ? 1.0 : 0.0;
    const int64_t Q = (int64_t)sd[5 * n];
    const double* rk = rot + k * 2 * cap;
    const uint8_t* ck = rcol + k * cap;
    double nc = 0.0, ns = 0.0;
    int ni = 0;
    if (lane < Q) {
      nc = rk[2 * lane];
      ns = rk[2 * lane + 1];
      ni = ck[lane];
    }
    int prev = -2;
    double carry = 0.0;
    for (int64_t q0 = 0; q0 < Q; q0 += 64) {
      __builtin_amdgcn_wave_barrier();
      csb[2 * lane] = nc;
      csb[2 * lane + 1] = ns;
      ixb[lane] = ni;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int64_t qn = q0 + 64 + lane;
      if (qn < Q) {
        nc = rk[2 * qn];
        ns = rk[2 * qn + 1];
        ni = ck[qn];
      }
      const int cnt = (Q - q0 < 64) ? (int)(Q - q0) : 64;
      if (t < n)
        for (int r = 0; r < cnt; ++r) {
          const double c = csb[2 * r], s = csb[2 * r + 1];
          const int i = ixb[r];
          const double zi = Zt[i * LDZ + t];
          const double zi1 = (prev == i + 1) ? carry : Zt[(i + 1) * LDZ + t];
          Zt[(i + 1) * LDZ + t] = s * zi + c * zi1;
          carry = c * zi - s * zi1;
          Zt[i * LDZ + t] = carry;
          prev = i;
        }
    }
    }
    __syncthreads();
    // W = P Z, row-major in the rotation record
    zd* W = reinterpret_cast<zd*>(rot + k * 2 * cap);
    const int j = t;
    if (j < n)
      for (int i = 0; i < n; ++i)
        W[(int64_t)i * n + j] = zscale(zd{sd[3 * n + 2 * i], sd[3 * n + 2 * i + 1]}, Zt[j * LDZ + i]);
  }
}

size_t eigh_replay_wide_lds(int n) {
  return sizeof(double) * ((size_t)n * (n | 1) + 2 * 64 * 2) + sizeof(int) * 64 * 2;
}

// columns per LDS block of the back-transformation (about 73 KB of LDS, so
// that two workgroups fit a CU with the reflector buffers and tau)
// Back-transformation U = Q W for 32 < n <= 128 on FP64 MFMA.  The n - 2
// reflectors form blocks of 16; block J (reflectors 16J .. 16J + 15) is
// Q_J = H_16J ... H_16J+15 = I - V T V^H (compact WY: T upper triangular,
// T(i, i) = tau_i, T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i), G = V^H V), and
// W <- Q_J W = W - V (T (V^H W)) runs for J from the last block to the first.
// One wavefront per 16-column strip of W (NS = ceil(n / 16) waves per
// k-point); the strip stays in the wave's registers as NS accumulator tiles
// of 16 x 16 in the MFMA layout (lane l holds rows l / 16 + 4 v, v < 4, of
// column l % 16), and serves directly as the B operand of V^H W with the
// sum order over rows permuted to match.  Complex products take three real
// products each.  Per k-point: prologue, wave s computes T of the blocks
// J = s, s + NS, ... (G on MFMA from the reflectors in L2, the recurrence
// one row of T per lane); then per block the reflectors are staged in LDS
// (the next block's loads in flight) and every wave applies Q_J to its strip.
// Then each column takes the gauge convention (largest-modulus component,
// first on ties, real > 0) and goes to U's column of its eigenvalue rank.
template <int NSMAX>
__global__ __launch_bounds__(64 * NSMAX, NSMAX <= 4 ? 3 : 2) void eigh_back_mfma_kernel(int n, int64_t nk,
                                                                   const double* __restrict__ side,
                                                                   const double* __restrict__ rot,
                                                                   zd* __restrict__ U) {
  extern __shared__ __attribute__((aligned(16))) double bm_smem[];
  constexpr int VP = 17;  // LDS pitch (complex) of the staged block, G and T
  const int NS = (n + 15) >> 4;      // strips = waves = 16-row tiles
  const int NJ = (n - 2 + 15) >> 4;  // blocks of 16 reflectors
  zd* Vs = reinterpret_cast<zd*>(bm_smem);  // [16 NS][VP]; in the prologue G per wave
  zd* Ts = Vs + 16 * NS * VP;               // [NJ][16][VP]
  const int64_t cap = rot_cap(n);
  const zd zero = zd{0.0, 0.0};
  for (int64_t k = blockIdx.x; k < nk; k += gridDim.x) {
    // the thread index is made opaque per k-point: otherwise the compiler
    // hoists every lane-dependent address out of this loop and spills them
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    const int lane = t & 63;
    const int s = __builtin_amdgcn_readfirstlane(t >> 6);
    const int q = lane >> 4, c = lane & 15;
    const int j = 16 * s + c;  // the lane's column of W
    const double* sd = side + k * side_len(n);
    const zd* Wg = reinterpret_cast<const zd*>(rot + k * 2 * cap);
    const zd* Uk = U + k * (int64_t)n * n;
    mep_v4d wr[NSMAX], wi[NSMAX];
#pragma unroll
    for (int rt = 0; rt < NSMAX; ++rt)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * rt + q + 4 * v;
        const zd x = (rt < NS && row < n && j < n) ? Wg[(int64_t)row * n + j] : zero;
        wr[rt][v] = x.re;
        wi[rt][v] = x.im;
      }
    // ---- prologue: T of the blocks J = s, s + NS, ... ---------------------
    zd* G = Vs + s * 16 * VP;
    __syncthreads();  // the previous k-point's reads of Vs and Ts are done
    for (int J = s; J < NJ; J += NS) {
      const int kc = 16 * J + c;
      const bool colok = kc <= n - 3;
      mep_v4d g1 = {0, 0, 0, 0}, g2 = {0, 0, 0, 0}, g3 = {0, 0, 0, 0};
      for (int rt = J; rt < NS; ++rt) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 16 * rt + q + 4 * v;
          const zd x = (colok && row > kc && row < n) ? Uk[(int64_t)n * kc + row] : zero;
          // G = V^H V: A = conj(V)^T and B = V hold the same element here
          g1 = __builtin_amdgcn_mfma_f64_16x16x4f64(x.re, x.re, g1, 0, 0, 0);
          g2 = __builtin_amdgcn_mfma_f64_16x16x4f64(-x.im, x.im, g2, 0, 0, 0);
          g3 = __builtin_amdgcn_mfma_f64_16x16x4f64(x.re - x.im, x.re + x.im, g3, 0, 0, 0);
        }
      }
#pragma unroll
      for (int v = 0; v < 4; ++v)
        stz(G + (q + 4 * v) * VP + c, zd{g1[v] - g2[v], g3[v] - g1[v] - g2[v]});
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // lane r < 16: row r of T, T(r, i) = -tau_i sum_{r <= m < i} T(r, m) G(m, i),
      // the row kept in LDS (each lane reads back only what it wrote)
      zd* TJ = Ts + J * 16 * VP;
      if (lane < 16) {
        const int r = lane;
#pragma unroll 1
        for (int i = 0; i < 16; ++i) {
          const double ti = (16 * J + i <= n - 3) ? sd[2 * n + 16 * J + i] : 0.0;
          zd x0 = zero, x1 = zero;
          int m = r;
          for (; m + 1 < i; m += 2) {
            const zd g0 = ldz(G + m * VP + i), t0 = ldz(TJ + r * VP + m);
            const zd g1_ = ldz(G + (m + 1) * VP + i), t1 = ldz(TJ + r * VP + m + 1);
            x0.re = fma(t0.re, g0.re, fma(-t0.im, g0.im, x0.re));
            x0.im = fma(t0.re, g0.im, fma(t0.im, g0.re, x0.im));
            x1.re = fma(t1.re, g1_.re, fma(-t1.im, g1_.im, x1.re));
            x1.im = fma(t1.re, g1_.im, fma(t1.im, g1_.re, x1.im));
          }
          if (m < i) {
            const zd g0 = ldz(G + m * VP + i), t0 = ldz(TJ + r * VP + m);
            x0.re = fma(t0.re, g0.re, fma(-t0.im, g0.im, x0.re));
            x0.im = fma(t0.re, g0.im, fma(t0.im, g0.re, x0.im));
          }
          const zd x = zadd(x0, x1);
          stz(TJ + r * VP + i, (i < r) ? zero : (i == r) ? zd{ti, 0.0} : zd{-ti * x.re, -ti * x.im});
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // ---- blocks from the last to the first ---------------------------------
    // staging: thread t takes elements t + 64 NS u (u < 4) of the block,
    // consecutive threads on consecutive rows of one reflector
    const int nrow = 16 * NS;
    zd pv[4];
    auto fetch = [&](int J) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int idx = t + 64 * NS * u;
        const int row = idx % nrow, cc = idx / nrow, kc = 16 * J + cc;
        pv[u] = (kc <= n - 3 && row > kc && row < n) ? Uk[(int64_t)n * kc + row] : zero;
      }
    };
    fetch(NJ - 1);
    for (int J = NJ - 1; J >= 0; --J) {
      __syncthreads();  // T is complete; the previous block's reads of Vs are done
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int idx = t + 64 * NS * u;
        stz(Vs + (idx % nrow) * VP + idx / nrow, pv[u]);
      }
      __syncthreads();
      if (J > 0) fetch(J - 1);
      // Y = V^H W (rows above 16 J hold zeros in V)
      mep_v4d p1 = {0, 0, 0, 0}, p2 = {0, 0, 0, 0}, p3 = {0, 0, 0, 0};
#pragma unroll
      for (int rt = 0; rt < NSMAX; ++rt) {
        if (rt < J || rt >= NS) continue;  // uniform
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const zd a = ldz(Vs + (16 * rt + q + 4 * v) * VP + c);  // conj V(row, 16 J + c)
          p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, wr[rt][v], p1, 0, 0, 0);
          p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.im, wi[rt][v], p2, 0, 0, 0);
          p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re - a.im, wr[rt][v] + wi[rt][v], p3, 0, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);  // one row tile's operands in flight at a time
      }
      mep_v4d yr, yi;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        yr[v] = p1[v] - p2[v];
        yi[v] = p3[v] - p1[v] - p2[v];
      }
      // Y' = T Y
      const zd* TJ = Ts + J * 16 * VP;
      p1 = p2 = p3 = mep_v4d{0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const zd tt = ldz(TJ + c * VP + q + 4 * v);  // T(c, q + 4v)
        p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(tt.re, yr[v], p1, 0, 0, 0);
        p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(tt.im, yi[v], p2, 0, 0, 0);
        p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(tt.re + tt.im, yr[v] + yi[v], p3, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        yr[v] = p1[v] - p2[v];
        yi[v] = p3[v] - p1[v] - p2[v];
      }
      // W -= V Y'
#pragma unroll
      for (int rt = 0; rt < NSMAX; ++rt) {
        if (rt < J || rt >= NS) continue;  // uniform
        p1 = p2 = p3 = mep_v4d{0, 0, 0, 0};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const zd a = ldz(Vs + (16 * rt + c) * VP + q + 4 * v);  // V(16 rt + c, 16 J + q + 4v)
          p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, yr[v], p1, 0, 0, 0);
          p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, yi[v], p2, 0, 0, 0);
          p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re + a.im, yr[v] + yi[v], p3, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          wr[rt][v] -= p1[v] - p2[v];
          wi[rt][v] -= p3[v] - p1[v] - p2[v];
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    // ---- gauge: largest-modulus component (first on ties) real > 0 ---------
    double best = -1.0;
    int bi = n;
    double ur = 1.0, ui = 0.0;
#pragma unroll
    for (int rt = 0; rt < NSMAX; ++rt)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * rt + q + 4 * v;
        const double m2 = wr[rt][v] * wr[rt][v] + wi[rt][v] * wi[rt][v];
        if (rt < NS && row < n && m2 > best) {
          best = m2;
          bi = row;
          ur = wr[rt][v];
          ui = wi[rt][v];
        }
      }
#pragma unroll
    for (int off = 16; off <= 32; off *= 2) {
      const double ob = __shfl_xor(best, off, 64);
      const int oi = __shfl_xor(bi, off, 64);
      const double our = __shfl_xor(ur, off, 64), oui = __shfl_xor(ui, off, 64);
      if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
        ur = our;
        ui = oui;
      }
    }
    zd phs = zd{1.0, 0.0};
    {
      const double au = sqrt(ur * ur + ui * ui);
      if (bi < n && au > 0.0) phs = zd{ur / au, -ui / au};
    }
    // every wave has passed the last staging barrier, so no reflector of
    // this k-point is read after this point
    if (j < n) {
      const int rnk = (int)sd[n + j];
      zd* Uo = U + k * (int64_t)n * n + (int64_t)n * rnk;
#pragma unroll
      for (int rt = 0; rt < NSMAX; ++rt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 16 * rt + q + 4 * v;
          if (rt < NS && row < n) Uo[row] = zmul(zd{wr[rt][v], wi[rt][v]}, phs);
        }
    }
  }
}

size_t eigh_back_mfma_lds(int n) {
  const size_t ns = (size_t)(n + 15) / 16, nj = (size_t)(n - 2 + 15) / 16;
  return sizeof(zd) * 17 * 16 * (ns + nj);
}

static size_t eigh_list_offset(int n, int64_t nk);

void launch_eigh_wide(hipStream_t s, int n, int64_t nk, const zd* Hk, int64_t strideHk, int packed,
                      const KGeom& g, double* E, zd* U, int* info, double* scratch, int house) {
  double* side = scratch;
  double* rot = side + nk * side_len(n);
  uint8_t* rcol = reinterpret_cast<uint8_t*>(rot + nk * 2 * rot_cap(n));
  const size_t lh = eigh_house_wide_lds(n), lr = eigh_replay_wide_lds(n);
  const size_t lq = sizeof(double) * 2 * (size_t)n * 32;
  (void)hipFuncSetAttribute((const void*)eigh_house_wide_kernel,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lh);
  (void)hipFuncSetAttribute((const void*)eigh_replay_wide_kernel,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lr);
  (void)hipFuncSetAttribute((const void*)eigh_ql_wide_kernel<double>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lq);
  unsigned grid = (unsigned)(nk < 4096 ? nk : 4096);
  // measured crossovers (C3 models n = 36 ... 68, C5, seeded models n = 72 ... 128):
  // one wavefront with the matrix in LDS up to 44, registers to 64, the LDS
  // workgroup to 88 (five or six wavefronts of registers sit unevenly on the
  // four SIMDs), registers above
  if (house == 0) house = n <= 44 ? 2 : n <= 64 ? 3 : n <= 88 ? 1 : 3;
  if (house == 3) {
    // the matrix in registers, NMAX / 16 wavefronts per k-point
    const unsigned rgrid = (unsigned)(nk < 65536 ? nk : 65536);
    if (n <= 64)
      hipLaunchKernelGGL(eigh_house_rows_kernel<64>, dim3(rgrid), dim3(256), 0, s, n, nk, Hk,
                         strideHk, packed, g.kpts, g.tau, U, side);
    else if (n <= 80)
      hipLaunchKernelGGL(eigh_house_rows_kernel<80>, dim3(rgrid), dim3(320), 0, s, n, nk, Hk,
                         strideHk, packed, g.kpts, g.tau, U, side);
    else if (n <= 96)
      hipLaunchKernelGGL(eigh_house_rows_kernel<96>, dim3(rgrid), dim3(384), 0, s, n, nk, Hk,
                         strideHk, packed, g.kpts, g.tau, U, side);
    else
      hipLaunchKernelGGL(eigh_house_rows_kernel<128>, dim3(rgrid), dim3(512), 0, s, n, nk, Hk,
                         strideHk, packed, g.kpts, g.tau, U, side);
  } else if (house == 2 && n <= 64) {
    const size_t ll = eigh_house_lds_bytes(n);
    (void)hipFuncSetAttribute((const void*)eigh_house_lds_kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)ll);
    const unsigned lgrid = (unsigned)(nk < 65536 ? nk : 65536);
    hipLaunchKernelGGL(eigh_house_lds_kernel, dim3(lgrid), dim3(64), ll, s, n, nk, Hk, strideHk,
                       packed, g.kpts, g.tau, U, side);
  } else {
    hipLaunchKernelGGL(eigh_house_wide_kernel, dim3(grid), dim3(256), lh, s, n, nk, Hk, strideHk,
                       packed, g.kpts, g.tau, U, side);
  }
  int* list = reinterpret_cast<int*>(reinterpret_cast<uint8_t*>(scratch) + eigh_list_offset(n, nk));
  int* count = list + nk;
  int64_t qb = (nk + 31) / 32;
  unsigned qgrid = (unsigned)(qb < 65536 ? qb : 65536);
  hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(1), 0, s, count);
  // eigenvalues and the separation verdict: d and e in registers up to 48
  // bands (more k-points in flight than LDS columns allow), else in LDS
  const int64_t qb64 = (nk + 63) / 64;
  const unsigned qgrid64 = (unsigned)(qb64 < 65536 ? qb64 : 65536);
  if (n <= 40)
    hipLaunchKernelGGL((eigh_ql_reg_kernel<40, double>), dim3(qgrid64), dim3(64), 0, s, n, nk, side,
                       E, info, list, count);
  else if (n <= 48)
    hipLaunchKernelGGL((eigh_ql_reg_kernel<48, double>), dim3(qgrid64), dim3(64), 0, s, n, nk, side,
                       E, info, list, count);
  else
    hipLaunchKernelGGL(eigh_ql_wide_kernel<double>, dim3(qgrid), dim3(32), lq, s, n, nk, side, rot,
                       rcol, E, info, 1, list, count);
  hipLaunchKernelGGL(eigh_ql_wide_kernel<double>, dim3(qgrid), dim3(32), lq, s, n, nk, side, rot,
                     rcol, E, info, 2, list, count);
  const int rthreads = n <= 64 ? 64 : 128;
  hipLaunchKernelGGL(eigh_replay_wide_kernel, dim3(grid), dim3(rthreads), lr, s, n, nk, side, rot,
                     rcol, E);
  // back-transformation on MFMA: one wavefront per 16-column strip
  const size_t lm = eigh_back_mfma_lds(n);
  const int ns = (n + 15) / 16;
  const unsigned mgrid = (unsigned)(nk < 65536 ? nk : 65536);
#define MEP_BACKM_LAUNCH(NSM)                                                                     \
  (void)hipFuncSetAttribute((const void*)eigh_back_mfma_kernel<NSM>,                             \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lm);                \
  hipLaunchKernelGGL(eigh_back_mfma_kernel<NSM>, dim3(mgrid), dim3(64 * ns), lm, s, n, nk, side, \
                     rot, U);
  if (ns <= 3) {
    MEP_BACKM_LAUNCH(3)
  } else if (ns <= 4) {
    MEP_BACKM_LAUNCH(4)
  } else if (ns <= 6) {  // five strips: the six-strip build (no register spills)
    MEP_BACKM_LAUNCH(6)
  } else {
    MEP_BACKM_LAUNCH(8)
  }
#undef MEP_BACKM_LAUNCH
}

// side records, rotation records, rotation columns, then (256-byte
// aligned) the list of k-points rerun with records and its count
static size_t eigh_list_offset(int n, int64_t nk) {
  const size_t b =
      (size_t)nk * (sizeof(double) * (size_t)(side_len(n) + 2 * rot_cap(n)) + (size_t)rot_cap(n));
  return (b + 255) & ~(size_t)255;
}
size_t eigh_scratch_bytes(int n, int64_t nk) {
  return eigh_list_offset(n, nk) + sizeof(int) * (size_t)nk + 256;
}

template <class R>
static void launch_eigh_t(hipStream_t s, int n, int64_t nk, const typename cplx_of<R>::type* Hk,
                          int64_t strideHk, int packed, const KGeom& g, double* E,
                          typename cplx_of<R>::type* U, int* info, void* scratch, int house) {
  R* side = reinterpret_cast<R*>(scratch);
  R* rot = side + nk * side_len(n);
  uint8_t* rcol = reinterpret_cast<uint8_t*>(rot + nk * 2 * rot_cap(n));
  int* list = reinterpret_cast<int*>(reinterpret_cast<uint8_t*>(scratch) + eigh_list_offset(n, nk));
  int* count = list + nk;
  const int teams = n <= 8 ? 8 : n <= 16 ? 4 : 2;
  int64_t blocks = (nk + teams - 1) / teams;
  unsigned grid = (unsigned)(blocks < 65536 ? blocks : 65536);
  int64_t qb = (nk + 63) / 64;
  unsigned qgrid = (unsigned)(qb < 65536 ? qb : 65536);
  const unsigned hgrid = (unsigned)(nk < 65536 ? nk : 65536);  // one k per wavefront
  const int64_t h2 = (nk + 1) / 2;
  const unsigned h2grid = (unsigned)(h2 < 65536 ? h2 : 65536);  // two k per wavefront
#define MEP_EIGH_LAUNCH(NM)                                                                      \
  if constexpr (NM == 32 && sizeof(R) == 8) {                                                    \
    if (house == 2) {                                                                            \
      if (g.tau)                                                                                 \
        hipLaunchKernelGGL(eigh_house_row2_kernel<true>, dim3(h2grid), dim3(64), 0, s, n, nk,    \
                           Hk, strideHk, packed, g.kpts, g.tau, U, side);                        \
      else                                                                                       \
        hipLaunchKernelGGL(eigh_house_row2_kernel<false>, dim3(h2grid), dim3(64), 0, s, n, nk,   \
                           Hk, strideHk, packed, g.kpts, g.tau, U, side);                        \
    } else if (g.tau)                                                                            \
      hipLaunchKernelGGL(eigh_house_half_kernel<true>, dim3(hgrid), dim3(64), 0, s, n, nk, Hk,    \
                         strideHk, packed, g.kpts, g.tau, U, side);                              \
    else                                                                                         \
      hipLaunchKernelGGL(eigh_house_half_kernel<false>, dim3(hgrid), dim3(64), 0, s, n, nk, Hk,   \
                         strideHk, packed, g.kpts, g.tau, U, side);                              \
  } else                                                                                         \
    hipLaunchKernelGGL((eigh_house_kernel<NM, R>), dim3(grid), dim3(64), 0, s, n, nk, Hk, strideHk, \
                       packed, g.kpts, g.tau, U, side);                                          \
  hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(1), 0, s, count);                        \
  hipLaunchKernelGGL((eigh_ql_reg_kernel<NM, R>), dim3(qgrid), dim3(64), 0, s, n, nk, side, E,  \
                     info, list, count);                                                         \
  hipLaunchKernelGGL((eigh_ql_kernel<NM, R>), dim3(qgrid), dim3(64), 0, s, n, nk, side, rot,     \
                     rcol, E, info, 2, list, count);                                             \
  hipLaunchKernelGGL((eigh_vec_kernel<NM, R>), dim3(grid), dim3(64), 0, s, n, nk, side, rot,     \
                     rcol, E, U);
  if (n <= 8) {
    MEP_EIGH_LAUNCH(8)
  } else if (n <= 16) {
    MEP_EIGH_LAUNCH(16)
  } else {
    MEP_EIGH_LAUNCH(32)
  }
#undef MEP_EIGH_LAUNCH
}

void launch_eigh(hipStream_t s, int n, int64_t nk, const zd* Hk, int64_t strideHk, int packed,
                 const KGeom& g, double* E, zd* U, int* info, double* scratch, int house) {
  launch_eigh_t<double>(s, n, nk, Hk, strideHk, packed, g, E, U, info, scratch, house);
}

void launch_eigh_f32(hipStream_t s, int n, int64_t nk, const zf* Hk, int64_t strideHk, int packed,
                     const KGeom& g, double* E, zf* U, int* info, double* scratch) {
  launch_eigh_t<float>(s, n, nk, Hk, strideHk, packed, g, E, U, info, scratch, 0);
}
