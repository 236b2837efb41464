// mep_eigh.hip -- batched Hermitian eigensolver, one k-point per wavefront.
//
// Replaces zheevd('V','U') of zheevd_wrapper (matrix_math.f90:192-226) inside
// get_wann_interp (wann_interp.f90:95).  Contract kept: only the upper
// triangle of H(k) is read, eigenvalues ascending, eigenvectors in columns.
//
// Supported sizes: n <= 32 (the caller checks).
// Three kernels per k-block:
//   eigh_house (a team of NMAX lanes per k, lane i <-> row i, matrix in LDS):
//      Householder reduction A = Q T Q^H with rank-2 trailing updates; the
//      reflectors are stored in U's slot of the k-point, the diagonal phases
//      P that make the subdiagonal real (T = P T_r P^H), d, e and tau in a
//      side record.
//   eigh_ql (one LANE per k-point, 64 k-points per wavefront): the implicit
//      QL recurrence on T_r (EISPACK tql2 method) with d, e in LDS columns;
//      every sweep's Givens rotations (c_i, s_i) are appended to a per-k
//      record in HBM; eigenvalues are ranked and written in ascending order.
//   eigh_vec (a team per k): replays the rotations on Z (LDS, lane r <-> row
//      r), then U = Q P Z: lane j keeps column j of P Z in registers and
//      applies the reflectors in reverse order.
//   Gauge convention shared with the oracle: each eigenvector's
//   largest-modulus component is made real and positive.
#include "mep_common.hpp"
#include "mep_kernels.hpp"

// NMAX lanes form a team that owns one k-point; a wavefront runs 64/NMAX
// teams side by side.  Team reductions use width-NMAX shuffles.
template <int W, class T>
__device__ inline T team_sum(T v) {
#pragma unroll
  for (int o = W / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o, W);
  return v;
}

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
      const Z x0 = Z{__shfl(x.re, 0, NMAX), __shfl(x.im, 0, NMAX)};
      const bool refl = tail2 > R(0);
      const R ax02 = zabs2(x0);
      const R xn = sqrt(tail2 + ax02);
      const R ax0 = sqrt(ax02);
      const Z alpha = !refl ? x0 : ((ax0 > R(0)) ? zscale(x0, -xn / ax0) : Z{-xn, R(0)});
      Z v = refl ? x : Z{R(0), R(0)};
      if (t == 0 && refl) v = zsub(x, alpha);
      const R vn2 = team_sum<NMAX, R>((t < m) ? zabs2(v) : R(0));
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
      if (t == 0) {
        // phases and the real subdiagonal (serial recurrence)
        Z ph = Z{R(1), R(0)};
        sd[3 * n] = R(1);
        sd[3 * n + 1] = R(0);
        for (int i = 0; i + 1 < n; ++i) {
          const Z s_ = sub[i];
          const R as = sqrt(zabs2(s_));
          sd[n + i] = as;
          if (as > R(0)) ph = zmul(ph, zscale(s_, R(1) / as));
          sd[3 * n + 2 * (i + 1)] = ph.re;
          sd[3 * n + 2 * (i + 1) + 1] = ph.im;
        }
        sd[n + n - 1] = R(0);
      }
    }
    __syncthreads();
  }
}

// One lane per k-point: implicit QL on the real tridiagonal (d, e) held in
// LDS columns.  Every Givens rotation G(i, i+1; c, s) of every sweep is
// appended, in application order, to the k-point's flat stream
// rot[k][q] = (c, s), col[k][q] = i; the count goes to the side record.
// Ascending eigenvalues go to E, each band's rank to the side record.
template <int NMAX, class R>
__global__ __launch_bounds__(64) void eigh_ql_kernel(int n, int64_t nk,
                                                     R* __restrict__ side,
                                                     R* __restrict__ rot,
                                                     uint8_t* __restrict__ rcol,
                                                     double* __restrict__ E,
                                                     int* __restrict__ info) {
  using Z = typename cplx_of<R>::type;
  __shared__ R ds[NMAX * 64], es[NMAX * 64];
  const int lane = threadIdx.x;
  const R EPS = cplx_of<R>::eps;
  const int64_t cap = rot_cap(n);
  for (int64_t kb = (int64_t)blockIdx.x * 64; kb < nk; kb += (int64_t)gridDim.x * 64) {
    const int64_t k = kb + lane;
    if (k < nk) {
      const R* sd = side + k * side_len(n);
      for (int i = 0; i < n; ++i) {
        ds[i * 64 + lane] = sd[i];
        es[i * 64 + lane] = sd[n + i];
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
            const R dd = fabs(ds[m * 64 + lane]) + fabs(ds[(m + 1) * 64 + lane]);
            if (fabs(es[m * 64 + lane]) <= EPS * dd) break;
          }
          if (m == l) break;
          if (++iter > 60 || pos + (m - l) > cap) {
            fail = 1;
            break;
          }
          const R dl = ds[l * 64 + lane], el = es[l * 64 + lane];
          R g = (ds[(l + 1) * 64 + lane] - dl) / (R(2) * el);
          R r = sqrt(g * g + R(1));
          g = ds[m * 64 + lane] - dl + el / (g + copysign(r, g));
          R s = R(1), c = R(1), p = R(0);
          bool underflow = false;
          R ei = es[(m - 1) * 64 + lane], di = ds[(m - 1) * 64 + lane];
          R di1 = ds[m * 64 + lane];
          for (int i = m - 1; i >= l; --i) {
            // operands of the next step are not touched by this one
            const R ein = (i > l) ? es[(i - 1) * 64 + lane] : R(0);
            const R din = (i > l) ? ds[(i - 1) * 64 + lane] : R(0);
            const R f = s * ei, b = c * ei;
            r = sqrt(f * f + g * g);
            es[(i + 1) * 64 + lane] = r;
            if (r == R(0)) {
              ds[(i + 1) * 64 + lane] = di1 - p;
              es[m * 64 + lane] = R(0);
              underflow = true;
              break;
            }
            const R rr = R(1) / r;
            s = f * rr;
            c = g * rr;
            g = di1 - p;
            r = (di - g) * s + R(2) * c * b;
            p = s * r;
            ds[(i + 1) * 64 + lane] = g + p;
            g = c * r - b;
            rk[2 * pos] = c;
            rk[2 * pos + 1] = s;
            ck[pos] = (uint8_t)i;
            pos += 1;
            ei = ein;
            di1 = di;
            di = din;
          }
          if (underflow) continue;
          ds[l * 64 + lane] -= p;
          es[l * 64 + lane] = g;
          es[m * 64 + lane] = R(0);
        }
      }
      if (fail) atomicOr(info, 1);
      // ascending order: rank of each eigenvalue (ties by index)
      for (int j = 0; j < n; ++j) {
        const R dj = ds[j * 64 + lane];
        int rnk = 0;
        for (int i = 0; i < n; ++i) {
          const R di = ds[i * 64 + lane];
          rnk += (di < dj) || (di == dj && i < j);
        }
        E[k * n + rnk] = dj;
        es[j * 64 + lane] = (R)rnk;  // e is no longer needed: keep the rank
      }
      R* sdw = side + k * side_len(n);
      for (int j = 0; j < n; ++j) sdw[n + j] = es[j * 64 + lane];  // rank over e's slot
      sdw[5 * n] = (R)pos;
    }
  }
}

// A team per k-point.  Z starts as the identity in LDS (lane r <-> row r);
// the rotation stream is staged through LDS one team-width chunk at a time,
// the next chunk's loads in flight while the current one is applied.  Then
// U = Q P Z: lane j takes column j of P Z into registers, the packed
// reflectors are staged into the Z space, and applied in reverse order.
template <int NMAX, class R>
__global__ __launch_bounds__(64) void eigh_vec_kernel(int n, int64_t nk,
                                                      const R* __restrict__ side,
                                                      const R* __restrict__ rot,
                                                      const uint8_t* __restrict__ rcol,
                                                      typename cplx_of<R>::type* __restrict__ U) {
  using Z = typename cplx_of<R>::type;
  constexpr int TEAMS = 64 / NMAX;
  constexpr int LDZ = NMAX + 1;
  __shared__ R Zt_s[TEAMS][NMAX * LDZ];  // Zt[col*LDZ + row]; later the reflectors
  __shared__ R cs_s[TEAMS][2 * NMAX];
  __shared__ int ix_s[TEAMS][NMAX];
  const int lane = threadIdx.x;
  const int tm = lane / NMAX, t = lane % NMAX;
  R* Zt = Zt_s[tm];
  R* csb = cs_s[tm];
  int* ixb = ix_s[tm];
  const int64_t cap = rot_cap(n);
  for (int64_t kb = (int64_t)blockIdx.x * TEAMS; kb < nk; kb += (int64_t)gridDim.x * TEAMS) {
    const int64_t k = kb + tm;
    const bool valid = k < nk;
    const R* sd = side + (valid ? k : 0) * side_len(n);
#pragma unroll
    for (int c = 0; c < NMAX; ++c) Zt[c * LDZ + t] = (c == t) ? R(1) : R(0);
    const int64_t Q = valid ? (int64_t)sd[5 * n] : 0;
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
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // ---- lane j <-> column j of P Z ----------------------------------------
    Z W[NMAX];
#pragma unroll
    for (int i = 0; i < NMAX; ++i)
      W[i] = (valid && i < n && t < n)
                 ? zscale(Z{sd[3 * n + 2 * i], sd[3 * n + 2 * i + 1]}, Zt[t * LDZ + i])
                 : Z{R(0), R(0)};
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // ---- packed reflectors (strict lower triangle) into the Z space ---------
    Z* Vp = reinterpret_cast<Z*>(Zt);
    const Z* Uk = U + (valid ? k : 0) * (int64_t)n * n;
    if (valid)
      for (int kc = 0; kc + 2 < n; ++kc) {
        const int off = kc * (n - 1) - (kc * (kc - 1)) / 2;
        const int r = kc + 1 + t;
        if (r < n) Vp[off + t] = Uk[r + (int64_t)n * kc];
      }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (valid) {
      for (int kc = n - 3; kc >= 0; --kc) {
        const int off = kc * (n - 1) - (kc * (kc - 1)) / 2 - (kc + 1);  // Vp[off + i], i > kc
        const R tk = sd[2 * n + kc];
        Z sacc = Z{R(0), R(0)};
#pragma unroll
        for (int i = 0; i < NMAX; ++i)
          if (i > kc && i < n) sacc = zadd(sacc, zcmul(Vp[off + i], W[i]));
        sacc = zscale(sacc, tk);
#pragma unroll
        for (int i = 0; i < NMAX; ++i)
          if (i > kc && i < n) W[i] = zsub(W[i], zmul(Vp[off + i], sacc));
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
    const int rnk = (valid && t < n) ? (int)sd[n + t] : 0;
    if (valid && t < n) {
      Z* Uo = U + k * (int64_t)n * n + (int64_t)n * rnk;
#pragma unroll
      for (int i = 0; i < NMAX; ++i)
        if (i < n) Uo[i] = W[i];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

size_t eigh_scratch_bytes(int n, int64_t nk) {
  return (size_t)nk * (sizeof(double) * (size_t)(side_len(n) + 2 * rot_cap(n)) + (size_t)rot_cap(n));
}

template <class R>
static void launch_eigh_t(hipStream_t s, int n, int64_t nk, const typename cplx_of<R>::type* Hk,
                          int64_t strideHk, int packed, const KGeom& g, double* E,
                          typename cplx_of<R>::type* U, int* info, void* scratch) {
  R* side = reinterpret_cast<R*>(scratch);
  R* rot = side + nk * side_len(n);
  uint8_t* rcol = reinterpret_cast<uint8_t*>(rot + nk * 2 * rot_cap(n));
  const int teams = n <= 8 ? 8 : n <= 16 ? 4 : 2;
  int64_t blocks = (nk + teams - 1) / teams;
  unsigned grid = (unsigned)(blocks < 65536 ? blocks : 65536);
  int64_t qb = (nk + 63) / 64;
  unsigned qgrid = (unsigned)(qb < 65536 ? qb : 65536);
#define MEP_EIGH_LAUNCH(NM)                                                                      \
  hipLaunchKernelGGL((eigh_house_kernel<NM, R>), dim3(grid), dim3(64), 0, s, n, nk, Hk, strideHk, \
                     packed, g.kpts, g.tau, U, side);                                            \
  hipLaunchKernelGGL((eigh_ql_kernel<NM, R>), dim3(qgrid), dim3(64), 0, s, n, nk, side, rot,     \
                     rcol, E, info);                                                             \
  hipLaunchKernelGGL((eigh_vec_kernel<NM, R>), dim3(grid), dim3(64), 0, s, n, nk, side, rot,     \
                     rcol, U);
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
                 const KGeom& g, double* E, zd* U, int* info, double* scratch) {
  launch_eigh_t<double>(s, n, nk, Hk, strideHk, packed, g, E, U, info, scratch);
}

void launch_eigh_f32(hipStream_t s, int n, int64_t nk, const zf* Hk, int64_t strideHk, int packed,
                     const KGeom& g, double* E, zf* U, int* info, double* scratch) {
  launch_eigh_t<float>(s, n, nk, Hk, strideHk, packed, g, E, U, info, scratch);
}
