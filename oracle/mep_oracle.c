/*
 * mep_oracle.c -- CPU FP64 restatement of MEPinterp's k-space hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library or the Fortran host.
 *
 * Each function restates one reference routine and cites it.  Loop orders
 * follow the reference so that rounding stays close to its serial path.
 * Appendix-style semantic quirks of the reference are kept by default
 * (orc_model.textbook == 0):
 *   - rotate_gauge reuses the velocity temporary for the connection, so
 *     A_bar = (U^+ dH U) U                       (wann_interp.f90:529-536)
 *   - get_velo runs after conn_gaugeTrafo, so it uses A^(H) = A_bar + iD
 *     (wann_interp.f90:100-101, 406-444); V_nn is left unassigned there;
 *     here V_nn := Hbar_nn (the reference's value is uninitialised memory)
 *   - curv_gaugeTrafo uses A^(H), not A_bar       (wann_interp.f90:636-658)
 * None of these matter when no position operator is given (then V = Hbar).
 */
#include "mep_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define PI_DP (4.0 * atan(1.0))           /* constants.f90:13 */
#define AU_TO_EV 27.211385                /* constants.f90:19 */
#define KB_EH_K (8.6173303 * 1e-5 / AU_TO_EV) /* constants.f90:22 */
#define MIN_TEMP 1e-2                     /* statistics.f90:32 */

#define IX3(a, n, m, N) ((a) + 3 * ((n) + (N) * (m)))
#define IX33(a, b, n, m, N) ((a) + 3 * ((b) + 3 * ((n) + (N) * (m))))
#define IXM(n, m, N) ((n) + (N) * (m))

/* ------------------------------------------------------------------------ */
/* k_space.f90                                                              */
/* ------------------------------------------------------------------------ */

static void crossP(const double a[3], const double b[3], double c[3]) {
  /* matrix_math.f90:314-336 */
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

static double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

double orc_unit_vol(const double al[9]) {
  /* k_space.f90:271-278: a_i = a_latt(:,i) (columns); det is the same */
  double a1[3] = {al[0], al[1], al[2]}, a2[3] = {al[3], al[4], al[5]},
         a3[3] = {al[6], al[7], al[8]}, c[3];
  crossP(a1, a2, c);
  return dot3(c, a3);
}

void orc_recip_latt(const double al[9], double rl[9]) {
  /* k_space.f90:263-296: b1 = 2pi a2 x a3 / V, ... ; recip(:,i) = b_i */
  double a1[3] = {al[0], al[1], al[2]}, a2[3] = {al[3], al[4], al[5]},
         a3[3] = {al[6], al[7], al[8]}, c[3];
  double vol = orc_unit_vol(al);
  crossP(a2, a3, c);
  for (int i = 0; i < 3; ++i) rl[i] = 2.0 * PI_DP * c[i] / vol;
  crossP(a3, a1, c);
  for (int i = 0; i < 3; ++i) rl[3 + i] = 2.0 * PI_DP * c[i] / vol;
  crossP(a1, a2, c);
  for (int i = 0; i < 3; ++i) rl[6 + i] = 2.0 * PI_DP * c[i] / vol;
}

int64_t orc_kpt_idx(const int mp[3], int qix, int qiy, int qiz) {
  /* k_space.f90:341-348 (1-based; strides by mp(2), mp(3) as written) */
  return 1 + (int64_t)(qix - 1) +
         (int64_t)mp[1] * ((int64_t)(qiy - 1) + (int64_t)mp[2] * (qiz - 1));
}

void orc_rel_kpt(const int mp[3], int qix, int qiy, int qiz, double k[3]) {
  /* k_space.f90:360-362: k_i = (2 q_i - N_i - 1) / (2 N_i) */
  int q[3] = {qix, qiy, qiz};
  for (int i = 0; i < 3; ++i)
    k[i] = (2.0 * (double)q[i] - (double)mp[i] - 1.0) / (2.0 * (double)mp[i]);
}

void orc_mp_mesh(const int mp[3], double *kpts) {
  /* core.f90:165-167 visiting order (kiz outer, kix inner).  On one rank
     every (kix,kiy,kiz) is visited once whatever get_kpt_idx returns; that
     index only equals the visiting position for meshes with N1 = N2 = N3
     (k_space.f90:345 strides by mp(2), mp(3)). */
  int64_t pos = 0;
  for (int z = 1; z <= mp[2]; ++z)
    for (int y = 1; y <= mp[1]; ++y)
      for (int x = 1; x <= mp[0]; ++x) orc_rel_kpt(mp, x, y, z, &kpts[3 * pos++]);
}

void orc_linspace(double lo, double hi, int n, double *out) {
  /* matrix_math.f90:74-100 */
  if (n > 1) {
    double d = (hi - lo) / (double)(n - 1);
    for (int i = 0; i < n; ++i) out[i] = lo + (double)i * d;
  } else {
    out[0] = lo;
  }
}

/* ------------------------------------------------------------------------ */
/* Hermitian eigensolver: zheevd('V','U') contract (matrix_math.f90:192-226) */
/* Eigenvector phases are fixed canonically (largest component real > 0):    */
/* LAPACK leaves them unspecified, and with a position operator the          */
/* reference's connection (stale-tmp A) depends on them.                     */
/* Householder reduction to Hermitian tridiagonal form, a diagonal phase that */
/* makes the off-diagonal real, implicit-shift QL on the real tridiagonal,   */
/* back-transformation, ascending sort.                                      */
/* ------------------------------------------------------------------------ */

int orc_zheev(int64_t n, orc_z *A, double *w) {
  if (n <= 0) return 0;
  orc_z *a = malloc(sizeof(orc_z) * n * n);
  orc_z *Q = malloc(sizeof(orc_z) * n * n);
  orc_z *v = malloc(sizeof(orc_z) * n);
  orc_z *p = malloc(sizeof(orc_z) * n);
  double *d = malloc(sizeof(double) * n);
  double *e = malloc(sizeof(double) * n);
  double *Z = malloc(sizeof(double) * n * n);
  orc_z *ph = malloc(sizeof(orc_z) * n);
  int info = 0;

  /* the upper triangle defines the matrix; the diagonal's imaginary part is
     ignored, as LAPACK does for Hermitian input */
  for (int64_t j = 0; j < n; ++j) {
    for (int64_t i = 0; i < j; ++i) {
      a[IXM(i, j, n)] = A[IXM(i, j, n)];
      a[IXM(j, i, n)] = conj(A[IXM(i, j, n)]);
    }
    a[IXM(j, j, n)] = creal(A[IXM(j, j, n)]);
  }
  for (int64_t i = 0; i < n * n; ++i) Q[i] = 0;
  for (int64_t i = 0; i < n; ++i) Q[IXM(i, i, n)] = 1;

  /* Householder: column k, annihilate rows k+2..n-1 */
  for (int64_t k = 0; k + 2 < n; ++k) {
    int64_t m = n - k - 1;
    orc_z *x = &a[IXM(k + 1, k, n)];
    double xn2 = 0;
    for (int64_t i = 0; i < m; ++i) xn2 += creal(x[i] * conj(x[i]));
    double tail2 = xn2 - creal(x[0] * conj(x[0]));
    if (tail2 == 0.0) continue; /* already tridiagonal in this column */
    double xn = sqrt(xn2);
    double ax0 = cabs(x[0]);
    orc_z alpha = (ax0 > 0.0) ? -(x[0] / ax0) * xn : -xn;
    for (int64_t i = 0; i < m; ++i) v[i] = x[i];
    v[0] -= alpha;
    double vn2 = 0;
    for (int64_t i = 0; i < m; ++i) vn2 += creal(v[i] * conj(v[i]));
    double tau = 2.0 / vn2;
    /* p = tau * B v with B = a[k+1:, k+1:] */
    for (int64_t i = 0; i < m; ++i) {
      orc_z s = 0;
      for (int64_t j = 0; j < m; ++j) s += a[IXM(k + 1 + i, k + 1 + j, n)] * v[j];
      p[i] = tau * s;
    }
    orc_z vp = 0;
    for (int64_t i = 0; i < m; ++i) vp += conj(v[i]) * p[i];
    double K = 0.5 * tau * creal(vp);
    for (int64_t i = 0; i < m; ++i) p[i] -= K * v[i]; /* p := w */
    for (int64_t j = 0; j < m; ++j)
      for (int64_t i = 0; i < m; ++i)
        a[IXM(k + 1 + i, k + 1 + j, n)] -= v[i] * conj(p[j]) + p[i] * conj(v[j]);
    a[IXM(k + 1, k, n)] = alpha;
    a[IXM(k, k + 1, n)] = conj(alpha);
    for (int64_t i = 1; i < m; ++i) {
      a[IXM(k + 1 + i, k, n)] = 0;
      a[IXM(k, k + 1 + i, n)] = 0;
    }
    /* Q := Q H_k on columns k+1..n-1 */
    for (int64_t r = 0; r < n; ++r) {
      orc_z s = 0;
      for (int64_t j = 0; j < m; ++j) s += Q[IXM(r, k + 1 + j, n)] * v[j];
      s *= tau;
      for (int64_t j = 0; j < m; ++j) Q[IXM(r, k + 1 + j, n)] -= s * conj(v[j]);
    }
  }
  /* real tridiagonal via diagonal phases: T_real = P^+ T P */
  ph[0] = 1;
  for (int64_t i = 0; i < n; ++i) d[i] = creal(a[IXM(i, i, n)]);
  for (int64_t i = 0; i + 1 < n; ++i) {
    orc_z s = a[IXM(i + 1, i, n)];
    double as = cabs(s);
    e[i] = as;
    ph[i + 1] = (as > 0.0) ? ph[i] * (s / as) : ph[i];
  }
  e[n - 1] = 0.0;
  for (int64_t i = 0; i < n * n; ++i) Z[i] = 0;
  for (int64_t i = 0; i < n; ++i) Z[IXM(i, i, n)] = 1;

  /* implicit QL with Wilkinson-type shift (EISPACK tql2 method) */
  for (int64_t l = 0; l < n; ++l) {
    int iter = 0;
    for (;;) {
      int64_t m = l;
      for (; m < n - 1; ++m) {
        double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
      }
      if (m == l) break;
      if (++iter > 60) { info = 1; goto done; }
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + copysign(r, g));
      double s = 1.0, c = 1.0, pp = 0.0;
      int underflow = 0;
      for (int64_t i = m - 1; i >= l; --i) {
        double f = s * e[i], b = c * e[i];
        r = hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= pp;
          e[m] = 0.0;
          underflow = 1;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - pp;
        r = (d[i] - g) * s + 2.0 * c * b;
        pp = s * r;
        d[i + 1] = g + pp;
        g = c * r - b;
        for (int64_t k = 0; k < n; ++k) {
          double zf = Z[IXM(k, i + 1, n)];
          Z[IXM(k, i + 1, n)] = s * Z[IXM(k, i, n)] + c * zf;
          Z[IXM(k, i, n)] = c * Z[IXM(k, i, n)] - s * zf;
        }
      }
      if (underflow) continue;
      d[l] -= pp;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  /* ascending selection sort of (d, Z columns) */
  for (int64_t i = 0; i < n; ++i) {
    int64_t kmin = i;
    for (int64_t j = i + 1; j < n; ++j)
      if (d[j] < d[kmin]) kmin = j;
    if (kmin != i) {
      double t = d[i]; d[i] = d[kmin]; d[kmin] = t;
      for (int64_t r = 0; r < n; ++r) {
        double zt = Z[IXM(r, i, n)];
        Z[IXM(r, i, n)] = Z[IXM(r, kmin, n)];
        Z[IXM(r, kmin, n)] = zt;
      }
    }
  }
  /* U = Q diag(ph) Z */
  for (int64_t j = 0; j < n; ++j)
    for (int64_t r = 0; r < n; ++r) {
      orc_z s = 0;
      for (int64_t i = 0; i < n; ++i)
        s += Q[IXM(r, i, n)] * ph[i] * Z[IXM(i, j, n)];
      A[IXM(r, j, n)] = s;
    }
  /* gauge convention: the largest-modulus component of each eigenvector
     (first one on ties) is made real and positive */
  for (int64_t j = 0; j < n; ++j) {
    int64_t im = 0;
    double best = -1.0;
    for (int64_t r = 0; r < n; ++r) {
      double m2 = creal(A[IXM(r, j, n)] * conj(A[IXM(r, j, n)]));
      if (m2 > best) { best = m2; im = r; }
    }
    orc_z u = A[IXM(im, j, n)];
    double au = cabs(u);
    if (au > 0.0) {
      orc_z phs = conj(u) / au;
      for (int64_t r = 0; r < n; ++r) A[IXM(r, j, n)] *= phs;
    }
  }
  for (int64_t i = 0; i < n; ++i) w[i] = d[i];
done:
  free(a); free(Q); free(v); free(p); free(d); free(e); free(Z); free(ph);
  return info;
}

/* ------------------------------------------------------------------------ */
/* wann_interp.f90                                                          */
/* ------------------------------------------------------------------------ */

static void intern_to_cart(const double al[9], const double v[3], double c[3]) {
  /* wann_interp.f90:387-399: matmul(a_latt, vec) */
  for (int i = 0; i < 3; ++i)
    c[i] = al[i + 0] * v[0] + al[i + 3] * v[1] + al[i + 6] * v[2];
}

void orc_ft_R_to_k(const orc_model *md, const double kpt[3], orc_z *H_k,
                   orc_z *H_ka, orc_z *A_ka, orc_z *Om_kab) {
  /* wann_interp.f90:151-384 */
  const int64_t N = md->n_wf, NR = md->n_R;
  const double two_pi = 2.0 * PI_DP;
  const int use_pos = (A_ka && Om_kab && md->r_real);
  double recip[9], kc[3];
  orc_recip_latt(md->a_latt, recip);
  for (int i = 0; i < 3; ++i) /* get_cart_kpt: matmul(recip_latt, kpt) */
    kc[i] = recip[i] * kpt[0] + recip[i + 3] * kpt[1] + recip[i + 6] * kpt[2];

  for (int64_t i = 0; i < N * N; ++i) H_k[i] = 0;
  if (H_ka) for (int64_t i = 0; i < 3 * N * N; ++i) H_ka[i] = 0;
  if (use_pos) {
    for (int64_t i = 0; i < 3 * N * N; ++i) A_ka[i] = 0;
    for (int64_t i = 0; i < 9 * N * N; ++i) Om_kab[i] = 0;
  }
  const double *at = md->wf_centers;
  if (at) {
    /* TIGHT-BINDING convention: dR = R + tau_m - tau_n (:192-241) */
    for (int64_t s = 0; s < NR; ++s)
      for (int64_t m = 0; m < N; ++m)
        for (int64_t n = 0; n < N; ++n) {
          double dR[3], dRc[3], ang;
          for (int a = 0; a < 3; ++a)
            dR[a] = md->R_frac[a + 3 * s] + (at[a + 3 * m] - at[a + 3 * n]);
          if (md->use_cart_velo) {
            intern_to_cart(md->a_latt, dR, dRc);
            ang = dot3(kc, dRc);
          } else {
            ang = two_pi * dot3(kpt, dR);
          }
          orc_z ph = cos(ang) + I * sin(ang);
          orc_z h = md->H_real[n + N * (m + N * s)];
          H_k[IXM(n, m, N)] += ph * h;
          if (H_ka) {
            if (!md->use_cart_velo)
              for (int a = 0; a < 3; ++a) dRc[a] = two_pi * dR[a];
            for (int a = 0; a < 3; ++a)
              H_ka[IX3(a, n, m, N)] += ph * (I * dRc[a]) * h;
          }
        }
  } else {
    /* WANNIER convention (:245-280) */
    for (int64_t s = 0; s < NR; ++s) {
      const double *R = &md->R_frac[3 * s];
      double ang = two_pi * dot3(kpt, R), dRc[3];
      orc_z ph = cos(ang) + I * sin(ang);
      if (md->use_cart_velo) intern_to_cart(md->a_latt, R, dRc);
      else for (int a = 0; a < 3; ++a) dRc[a] = two_pi * R[a];
      for (int64_t m = 0; m < N; ++m)
        for (int64_t n = 0; n < N; ++n) {
          orc_z h = md->H_real[n + N * (m + N * s)];
          H_k[IXM(n, m, N)] += ph * h;
          if (H_ka)
            for (int a = 0; a < 3; ++a)
              H_ka[IX3(a, n, m, N)] += ph * I * dRc[a] * h;
        }
    }
  }
  if (!use_pos) return;
  /* connection and curvature (:288-375); TB phase is always 2pi k.dR (:306) */
  for (int64_t s = 0; s < NR; ++s)
    for (int64_t m = 0; m < N; ++m)
      for (int64_t n = 0; n < N; ++n) {
        double dR[3], dRc[3];
        for (int a = 0; a < 3; ++a)
          dR[a] = md->R_frac[a + 3 * s] + (at ? (at[a + 3 * m] - at[a + 3 * n]) : 0.0);
        double ang = two_pi * dot3(kpt, dR);
        orc_z ph = cos(ang) + I * sin(ang);
        if (md->use_cart_velo) intern_to_cart(md->a_latt, dR, dRc);
        else for (int a = 0; a < 3; ++a) dRc[a] = two_pi * dR[a];
        const orc_z *r = &md->r_real[3 * (n + N * (m + N * s))];
        for (int a = 0; a < 3; ++a) {
          A_ka[IX3(a, n, m, N)] += ph * r[a];
          for (int b = 0; b < 3; ++b) {
            Om_kab[IX33(a, b, n, m, N)] += ph * I * dRc[a] * r[b];
            Om_kab[IX33(a, b, n, m, N)] -= ph * I * dRc[b] * r[a];
          }
        }
      }
}

/* C = A B for n x n column-major (blas_matmul, matrix_math.f90:636-655) */
static void zmm(int64_t n, const orc_z *A, const orc_z *B, orc_z *C) {
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < n; ++i) {
      orc_z s = 0;
      for (int64_t l = 0; l < n; ++l) s += A[IXM(i, l, n)] * B[IXM(l, j, n)];
      C[IXM(i, j, n)] = s;
    }
}

/* gather / scatter component a of a (3,n,n) or (3,3,n,n) array */
static void get3(int64_t n, const orc_z *X, int a, orc_z *M) {
  for (int64_t i = 0; i < n * n; ++i) M[i] = X[a + 3 * i];
}
static void put3(int64_t n, orc_z *X, int a, const orc_z *M) {
  for (int64_t i = 0; i < n * n; ++i) X[a + 3 * i] = M[i];
}
static void get33(int64_t n, const orc_z *X, int a, int b, orc_z *M) {
  for (int64_t i = 0; i < n * n; ++i) M[i] = X[a + 3 * b + 9 * i];
}
static void put33(int64_t n, orc_z *X, int a, int b, const orc_z *M) {
  for (int64_t i = 0; i < n * n; ++i) X[a + 3 * b + 9 * i] = M[i];
}

int orc_get_wann_interp(const orc_model *md, const double kpt[3], double *e_k,
                        orc_z *V_ka, orc_z *A_ka, orc_z *Om_kab) {
  /* wann_interp.f90:56-108 */
  const int64_t N = md->n_wf, N2 = N * N;
  const int use_pos = (A_ka && Om_kab && md->r_real);
  orc_z *U = malloc(sizeof(orc_z) * N2);
  orc_z *Hka = V_ka ? malloc(sizeof(orc_z) * 3 * N2) : NULL;
  orc_ft_R_to_k(md, kpt, U, Hka, use_pos ? A_ka : NULL, use_pos ? Om_kab : NULL);
  int info = orc_zheev(N, U, e_k);
  if (info || !V_ka) { free(U); free(Hka); return info; }

  orc_z *Ud = malloc(sizeof(orc_z) * N2), *M = malloc(sizeof(orc_z) * N2),
        *T = malloc(sizeof(orc_z) * N2), *T2 = malloc(sizeof(orc_z) * N2);
  orc_z *D = use_pos ? malloc(sizeof(orc_z) * 3 * N2) : NULL;
  for (int64_t j = 0; j < N; ++j)
    for (int64_t i = 0; i < N; ++i) Ud[IXM(i, j, N)] = conj(U[IXM(j, i, N)]);

  if (md->do_gauge_trafo) {
    /* rotate_gauge (:502-550) */
    for (int a = 0; a < 3; ++a) {
      get3(N, Hka, a, M);
      zmm(N, M, U, T);  /* tmp = dH U */
      zmm(N, Ud, T, M); /* M = U^+ dH U */
      put3(N, Hka, a, M);
      if (use_pos) {
        if (md->textbook) {
          get3(N, A_ka, a, M);
          zmm(N, M, U, T2);
          zmm(N, Ud, T2, M);
          put3(N, A_ka, a, M);
        } else {
          zmm(N, Ud, T, T2);   /* tmp = U^+ tmp (stale tmp = dH U) */
          zmm(N, T2, U, M);    /* A = tmp U */
          put3(N, A_ka, a, M);
        }
        for (int b = 0; b < 3; ++b) {
          get33(N, Om_kab, a, b, T2);
          zmm(N, Ud, T2, M);
          zmm(N, M, U, T2);
          put33(N, Om_kab, a, b, T2);
        }
      }
    }
    if (use_pos) {
      orc_z *Abar = md->textbook ? malloc(sizeof(orc_z) * 3 * N2) : NULL;
      if (Abar) memcpy(Abar, A_ka, sizeof(orc_z) * 3 * N2);
      /* get_gauge_covar_deriv (:565-617) */
      for (int64_t i = 0; i < 3 * N2; ++i) D[i] = 0;
      for (int64_t m = 0; m < N; ++m)
        for (int64_t n = 0; n < N; ++n)
          if (n > m) {
            double de = e_k[m] - e_k[n];
            if (fabs(de) > md->kubo_tol)
              for (int a = 0; a < 3; ++a) {
                D[IX3(a, n, m, N)] = Hka[IX3(a, n, m, N)] / de;
                D[IX3(a, m, n, N)] = Hka[IX3(a, m, n, N)] / (-de);
              }
          }
      /* conn_gaugeTrafo (:620-633) */
      for (int64_t i = 0; i < 3 * N2; ++i) A_ka[i] = A_ka[i] + I * D[i];
      /* curv_gaugeTrafo (:636-658) */
      orc_z *Da = malloc(sizeof(orc_z) * N2), *Db = malloc(sizeof(orc_z) * N2),
            *Xa = malloc(sizeof(orc_z) * N2), *Xb = malloc(sizeof(orc_z) * N2),
            *C1 = malloc(sizeof(orc_z) * N2), *C2 = malloc(sizeof(orc_z) * N2);
      const orc_z *Asrc = md->textbook ? Abar : A_ka;
      for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a) {
          get3(N, D, a, Da); get3(N, D, b, Db);
          get3(N, Asrc, a, Xa); get3(N, Asrc, b, Xb);
          get33(N, Om_kab, a, b, M);
          zmm(N, Da, Xb, C1); zmm(N, Xb, Da, C2);      /* [D_a, A_b] */
          for (int64_t i = 0; i < N2; ++i) M[i] -= C1[i] - C2[i];
          zmm(N, Db, Xa, C1); zmm(N, Xa, Db, C2);      /* [D_b, A_a] */
          for (int64_t i = 0; i < N2; ++i) M[i] += C1[i] - C2[i];
          zmm(N, Da, Db, C1); zmm(N, Db, Da, C2);      /* [D_a, D_b] */
          for (int64_t i = 0; i < N2; ++i) M[i] -= I * (C1[i] - C2[i]);
          put33(N, Om_kab, a, b, M);
        }
      free(Da); free(Db); free(Xa); free(Xb); free(C1); free(C2);
      if (md->textbook) {
        /* textbook velocity uses A_bar, not A^(H) */
        for (int64_t m = 0; m < N; ++m)
          for (int64_t n = 0; n < N; ++n)
            for (int a = 0; a < 3; ++a) {
              orc_z v = Hka[IX3(a, n, m, N)];
              if (n != m) v -= I * (e_k[m] - e_k[n]) * Abar[IX3(a, n, m, N)];
              V_ka[IX3(a, n, m, N)] = v;
            }
        free(Abar);
        goto out;
      }
    }
  }
  /* get_velo (:406-444) */
  if (use_pos) {
    for (int64_t m = 0; m < N; ++m)
      for (int a = 0; a < 3; ++a) /* unassigned in the reference */
        V_ka[IX3(a, m, m, N)] = Hka[IX3(a, m, m, N)];
    for (int64_t m = 0; m < N; ++m)
      for (int64_t n = 0; n < N; ++n)
        if (n > m) {
          orc_z ce = I * (e_k[m] - e_k[n]);
          for (int a = 0; a < 3; ++a) {
            V_ka[IX3(a, n, m, N)] = Hka[IX3(a, n, m, N)] - ce * A_ka[IX3(a, n, m, N)];
            V_ka[IX3(a, m, n, N)] = Hka[IX3(a, m, n, N)] + ce * A_ka[IX3(a, m, n, N)];
          }
        }
  } else {
    memcpy(V_ka, Hka, sizeof(orc_z) * 3 * N2);
  }
out:
  free(U); free(Hka); free(Ud); free(M); free(T); free(T2); free(D);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* statistics.f90                                                           */
/* ------------------------------------------------------------------------ */

void orc_fd_stat(double e, const double *ef, int n_ef, double T, double *out) {
  /* statistics.f90:65-90 */
  if (T > MIN_TEMP) {
    double Ts = KB_EH_K * T;
    for (int i = 0; i < n_ef; ++i) out[i] = 1.0 / (1.0 + exp((e - ef[i]) / Ts));
  } else {
    for (int i = 0; i < n_ef; ++i) out[i] = (e < ef[i]) ? 1.0 : 0.0;
  }
}

/* ------------------------------------------------------------------------ */
/* kubo.f90                                                                 */
/* ------------------------------------------------------------------------ */

void orc_kubo_ahc_tens(int64_t N, const double *e, const orc_z *V, int n_ef,
                       const double *fd, double *o) {
  /* kubo.f90:37-78 (no degeneracy guard) */
  for (int i = 0; i < 9 * n_ef; ++i) o[i] = 0;
  for (int64_t n = 0; n < N; ++n) {
    double c[9] = {0};
    for (int64_t np = 0; np < N; ++np) {
      if (np == n) continue;
      double de = e[np] - e[n];
      double de2 = de * de;
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
          c[i + 3 * j] -= 2.0 * cimag(V[IX3(i, n, np, N)] * V[IX3(j, np, n, N)]) / de2;
    }
    for (int ef = 0; ef < n_ef; ++ef)
      for (int i = 0; i < 9; ++i) o[i + 9 * ef] -= c[i] * fd[ef + n_ef * n];
  }
}

void orc_velo_ahc_tens(int64_t N, const double *e, const orc_z *V, int n_hw,
                       const double *hw, int n_ef, const double *fd,
                       double eta, orc_z *o) {
  /* kubo.f90:86-140 */
  for (int64_t i = 0; i < 9LL * n_hw * n_ef; ++i) o[i] = 0;
  for (int64_t n = 0; n < N; ++n)
    for (int64_t np = 0; np < N; ++np) {
      if (np == n) continue;
      double dE2 = (e[n] - e[np]) * (e[n] - e[np]);
      double v[9];
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
          v[i + 3 * j] = cimag(V[IX3(i, n, np, N)] * V[IX3(j, np, n, N)]);
      for (int h = 0; h < n_hw; ++h) {
        orc_z w = hw[h] + I * eta;
        orc_z den = dE2 - w * w;
        for (int ef = 0; ef < n_ef; ++ef) {
          double df = fd[ef + n_ef * n] - fd[ef + n_ef * np];
          for (int i = 0; i < 9; ++i) o[i + 9 * (h + n_hw * ef)] += df * (v[i] / den);
        }
      }
    }
}

void orc_kubo_ohc_tens(int64_t N, const double *e, const orc_z *V, int n_hw,
                       const double *hw, int n_ef, const double *fd,
                       double eta, orc_z *o) {
  /* kubo.f90:145-198 */
  for (int64_t i = 0; i < 9LL * n_hw * n_ef; ++i) o[i] = 0;
  for (int64_t m = 0; m < N; ++m)
    for (int64_t n = 0; n < N; ++n) {
      if (m == n) continue;
      double dE = e[m] - e[n];
      orc_z v[9];
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
          v[i + 3 * j] = V[IX3(i, n, m, N)] * V[IX3(j, m, n, N)];
      for (int h = 0; h < n_hw; ++h) {
        orc_z den = dE * (dE - hw[h] - I * eta);
        for (int ef = 0; ef < n_ef; ++ef) {
          double df = fd[ef + n_ef * m] - fd[ef + n_ef * n];
          for (int i = 0; i < 9; ++i)
            o[i + 9 * (h + n_hw * ef)] += I * (v[i] / den) * df;
        }
      }
    }
}

/* ------------------------------------------------------------------------ */
/* kubo_mep.f90 / mep_niu.f90                                               */
/* ------------------------------------------------------------------------ */

static int levi(int i, int j, int k) {
  /* matrix_math.f90:125-139 */
  if ((i == 0 && j == 1 && k == 2) || (i == 1 && j == 2 && k == 0) ||
      (i == 2 && j == 0 && k == 1)) return 1;
  if ((i == 0 && j == 2 && k == 1) || (i == 2 && j == 1 && k == 0) ||
      (i == 1 && j == 0 && k == 2)) return -1;
  return 0;
}

void orc_kubo_mep_cs(int64_t N, const orc_z *A, const orc_z *Om, int n_ef,
                     const double *fd, double *o) {
  /* kubo_mep.f90:16-47 */
  for (int i = 0; i < 9 * n_ef; ++i) o[i] = 0;
  if (!A || !Om) return;
  for (int ef = 0; ef < n_ef; ++ef) {
    double cs = 0;
    for (int64_t n0 = 0; n0 < N; ++n0) {
      orc_z om[3] = {0, 0, 0};
      for (int c = 0; c < 3; ++c)
        for (int b = 0; b < 3; ++b)
          for (int a = 0; a < 3; ++a)
            om[c] += (double)levi(a, b, c) * Om[IX33(a, b, n0, n0, N)];
      orc_z dp = 0;
      for (int a = 0; a < 3; ++a) dp += conj(A[IX3(a, n0, n0, N)]) * om[a];
      cs += 0.5 * fd[ef + n_ef * n0] * creal(dp);
    }
    for (int i = 0; i < 3; ++i) o[i + 3 * i + 9 * ef] = cs;
  }
}

void orc_kubo_mep_lc(int64_t N, const double *e, const orc_z *V, int n_ef,
                     const double *fd, double *o) {
  /* kubo_mep.f90:51-97 */
  for (int i = 0; i < 9 * n_ef; ++i) o[i] = 0;
  for (int64_t n0 = 0; n0 < N; ++n0)
    for (int64_t n = 0; n < N; ++n) {
      if (n == n0) continue;
      double d = e[n0] - e[n];
      double den = d * d * d;
      double v[9] = {0};
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) {
            double lc = (double)levi(j, k, l);
            for (int i = 0; i < 3; ++i)
              v[i + 3 * j] += lc * creal(V[IX3(i, n0, n, N)] * V[IX3(k, n, n0, N)] *
                                          V[IX3(l, n0, n0, N)]);
          }
      for (int ef = 0; ef < n_ef; ++ef)
        for (int i = 0; i < 9; ++i) o[i + 9 * ef] += v[i] * fd[ef + n_ef * n0] / den;
    }
}

void orc_kubo_mep_ic(int64_t N, const double *e, const orc_z *V, int n_ef,
                     const double *fd, double *o) {
  /* kubo_mep.f90:100-149 */
  for (int i = 0; i < 9 * n_ef; ++i) o[i] = 0;
  for (int64_t n0 = 0; n0 < N; ++n0)
    for (int64_t m = 0; m < N; ++m)
      for (int64_t n = 0; n < N; ++n) {
        if (n == n0 || m == n0) continue;
        double d1 = e[n0] - e[m], d2 = e[n0] - e[n];
        double den = d1 * (d2 * d2);
        double v[9] = {0};
        for (int j = 0; j < 3; ++j)
          for (int l = 0; l < 3; ++l)
            for (int k = 0; k < 3; ++k) {
              double lc = (double)levi(j, k, l);
              for (int i = 0; i < 3; ++i)
                v[i + 3 * j] -= lc * creal(V[IX3(i, n0, n, N)] * V[IX3(k, n, m, N)] *
                                            V[IX3(l, m, n0, N)]);
            }
        for (int ef = 0; ef < n_ef; ++ef)
          for (int i = 0; i < 9; ++i) o[i + 9 * ef] += v[i] * fd[ef + n_ef * n0] / den;
      }
}

void orc_mep_niu_cs(int64_t N, const orc_z *A, const orc_z *Om, int n_val,
                    double *o) {
  /* mep_niu.f90:20-44 */
  for (int i = 0; i < 9 * n_val; ++i) o[i] = 0;
  if (!A || !Om) return;
  (void)N;
  for (int n0 = 0; n0 < n_val; ++n0) {
    orc_z om[3] = {0, 0, 0};
    for (int c = 0; c < 3; ++c)
      for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a)
          om[c] += (double)levi(a, b, c) * Om[IX33(a, b, n0, n0, N)];
    orc_z dp = 0;
    for (int a = 0; a < 3; ++a) dp += conj(A[IX3(a, n0, n0, N)]) * om[a];
    for (int i = 0; i < 3; ++i) o[i + 3 * i + 9 * n0] = 0.5 * creal(dp);
  }
}

void orc_mep_niu_lc(int64_t N, const orc_z *V, const double *e, int n_val,
                    double kubo_tol, double *o) {
  /* mep_niu.f90:47-99 */
  for (int i = 0; i < 9 * n_val; ++i) o[i] = 0;
  for (int n0 = 0; n0 < n_val; ++n0)
    for (int64_t n = 0; n < N; ++n) {
      if (n == n0) continue;
      double d = e[n0] - e[n];
      double den = d * d * d;
      if (!(fabs(den) > kubo_tol)) continue;
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) {
            double pf = (double)levi(j, k, l);
            if (fabs(pf) > 1e-1)
              for (int i = 0; i < 3; ++i)
                o[i + 3 * j + 9 * n0] +=
                    pf * creal(V[IX3(i, n0, n, N)] * V[IX3(k, n, n0, N)] *
                               V[IX3(l, n0, n0, N)]) / den;
          }
    }
}

void orc_mep_niu_ic(int64_t N, const orc_z *V, const double *e, int n_val,
                    double kubo_tol, double *o) {
  /* mep_niu.f90:102-159 */
  for (int i = 0; i < 9 * n_val; ++i) o[i] = 0;
  for (int n0 = 0; n0 < n_val; ++n0)
    for (int64_t m = 0; m < N; ++m)
      for (int64_t n = 0; n < N; ++n) {
        if (n == n0 || m == n0) continue;
        double d2 = e[n0] - e[n], d1 = e[n0] - e[m];
        double den = (d2 * d2) * d1;
        if (!(fabs(den) > kubo_tol)) continue;
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l) {
              double pf = -(double)levi(j, k, l);
              if (fabs(pf) > 1e-1)
                for (int i = 0; i < 3; ++i)
                  o[i + 3 * j + 9 * n0] +=
                      pf * creal(V[IX3(i, n0, n, N)] * V[IX3(k, n, m, N)] *
                                 V[IX3(l, m, n0, N)]) / den;
            }
      }
}

/* ------------------------------------------------------------------------ */
/* core.f90:165-270 -- the k-loop body, over an explicit k list             */
/* ------------------------------------------------------------------------ */
/* optical, photogalvanic and gyrotropic sums                               */
/* ------------------------------------------------------------------------ */

void orc_fd_stat_deriv(const double *fd, int n_ef, int64_t n, double *out) {
  /* statistics.f90:180-190: cosh of the occupation value */
  for (int64_t b = 0; b < n; ++b)
    for (int i = 0; i < n_ef; ++i) out[i + n_ef * b] = 1.0 / (2.0 + 2.0 * cosh(fd[i + n_ef * b]));
}

void orc_kubo_opt_tens(int64_t N, const double *e, const orc_z *A, int n_hw,
                       const double *hw, int n_ef, const double *fd, double eta,
                       orc_z *os, orc_z *oa) {
  const int64_t tot = 9 * (int64_t)n_hw * n_ef;
  memset(os, 0, sizeof(orc_z) * tot);
  memset(oa, 0, sizeof(orc_z) * tot);
  if (!A) return; /* kubo.f90:226: zero unless A_ka is allocated */
  orc_z *herm = calloc(tot, sizeof(orc_z)), *aherm = calloc(tot, sizeof(orc_z));
  const double pi = 3.14159265358979323846;
  for (int64_t m = 0; m < N; ++m)
    for (int64_t n = 0; n < N; ++n) {
      if (n == m) continue;
      const double dE = e[m] - e[n]; /* dE_mn = en_k(m) - en_k(n) */
      orc_z a_nm_mn[9];
      for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a)
          a_nm_mn[a + 3 * b] = dE * A[a + 3 * (n + N * m)] * A[b + 3 * (m + N * n)];
      for (int h = 0; h < n_hw; ++h) {
        /* delta_broad (kubo.f90:279-297) and the denominator of (12.11) */
        const orc_z dsm = pi * (dE - I * eta);
        const double dbr = cimag(1.0 / (dsm - pi * hw[h]));
        const double den = creal(dE - hw[h] - I * eta);
        for (int f = 0; f < n_ef; ++f) {
          const double df = fd[f + n_ef * m] - fd[f + n_ef * n];
          for (int q = 0; q < 9; ++q) {
            herm[q + 9 * (h + (int64_t)n_hw * f)] -= pi * a_nm_mn[q] * dbr * df;
            aherm[q + 9 * (h + (int64_t)n_hw * f)] += I * (a_nm_mn[q] / den) * df;
          }
        }
      }
    }
  for (int64_t q = 0; q < tot; ++q) {
    os[q] = creal(herm[q]) + I * cimag(aherm[q]); /* (12.14) */
    oa[q] = creal(aherm[q]) + I * cimag(herm[q]); /* (12.15) */
  }
  free(herm);
  free(aherm);
}

void orc_photo_2nd_cond(int64_t N, const double *e, const orc_z *V, int n_hw,
                        const double *hw, int n_ef, const double *fd, double eta,
                        orc_z *o) {
  memset(o, 0, sizeof(orc_z) * 27 * (int64_t)n_hw * n_ef);
  orc_z *tmp = malloc(sizeof(orc_z) * 27 * n_hw);
  for (int64_t n = 0; n < N; ++n)
    for (int64_t m = 0; m < N; ++m) {
      const orc_z dEnm = (e[n] - e[m]) - I * eta;
      for (int64_t l = 0; l < N; ++l) {
        if (l == n) continue;
        const orc_z dEnl = (e[n] - e[l]) - I * eta;
        orc_z vvv[27];
        for (int c = 0; c < 3; ++c)
          for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a)
              vvv[a + 3 * (b + 3 * c)] =
                  V[a + 3 * (n + N * l)] * (V[b + 3 * (l + N * m)] * V[c + 3 * (m + N * n)]);
        for (int h = 0; h < n_hw; ++h) {
          for (int q = 0; q < 27; ++q) tmp[q + 27 * h] = 0.0;
          for (int om = -1; om <= 1; om += 2) {
            const orc_z dEhw = dEnl + om * hw[h] - I * eta; /* photo.f90:64 */
            for (int q = 0; q < 27; ++q) tmp[q + 27 * h] += vvv[q] / (dEnm * dEhw);
          }
          for (int q = 0; q < 27; ++q) tmp[q + 27 * h] /= hw[h] * hw[h];
        }
        for (int f = 0; f < n_ef; ++f) {
          const double df = fd[f + n_ef * l] - fd[f + n_ef * n];
          for (int h = 0; h < n_hw; ++h)
            for (int q = 0; q < 27; ++q) o[q + 27 * (h + (int64_t)n_hw * f)] += tmp[q + 27 * h] * df;
        }
      }
    }
  free(tmp);
}

void orc_gyro_C(int64_t N, const orc_z *V, int n_ef, const double *fdd, orc_z *o) {
  memset(o, 0, sizeof(orc_z) * 9 * n_ef);
  for (int64_t n = 0; n < N; ++n)
    for (int f = 0; f < n_ef; ++f)
      for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a)
          o[a + 3 * (b + 3 * f)] +=
              V[a + 3 * (n + N * n)] * V[b + 3 * (n + N * n)] * fdd[f + n_ef * n];
}

void orc_gyro_D(int64_t N, const double *e, const orc_z *V, const orc_z *Om,
                double kubo_tol, int n_ef, const double *fdd, orc_z *o) {
  memset(o, 0, sizeof(orc_z) * 9 * n_ef);
  for (int64_t n = 0; n < N; ++n) {
    orc_z curv[9];
    if (Om) {
      for (int q = 0; q < 9; ++q) curv[q] = Om[q + 9 * (n + N * n)];
    } else { /* get_kubo_curv (wann_interp.f90:113-146) */
      for (int q = 0; q < 9; ++q) curv[q] = 0.0;
      for (int64_t np = 0; np < N; ++np) {
        if (np == n) continue;
        const double dEE = (e[np] - e[n]) * (e[np] - e[n]);
        if (fabs(dEE) < kubo_tol) continue;
        for (int j = 0; j < 3; ++j)
          for (int a = 0; a < 3; ++a)
            curv[a + 3 * j] += cimag(V[a + 3 * (n + N * np)] * V[j + 3 * (np + N * n)]) / dEE;
      }
      for (int q = 0; q < 9; ++q) curv[q] *= -2.0;
    }
    orc_z om[3] = {0.0, 0.0, 0.0}; /* convert_tens_to_vect */
    for (int c = 0; c < 3; ++c)
      for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a) om[c] += (double)levi(a, b, c) * curv[a + 3 * b];
    for (int f = 0; f < n_ef; ++f)
      for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a)
          o[a + 3 * (b + 3 * f)] -= V[a + 3 * (n + N * n)] * om[b] * fdd[f + n_ef * n];
  }
}

void orc_gyro_Dw(int64_t N, const double *e, const orc_z *V, const orc_z *A,
                 int n_hw, const double *hw, int n_ef, const double *fdd, orc_z *o) {
  memset(o, 0, sizeof(orc_z) * 9 * (int64_t)n_hw * n_ef);
  if (!A) return;
  /* get_curvTilde_hw (gyro.f90:271-300); the reference accumulates into an
     unzeroed allocation, zero-initialised here */
  double *omt = calloc(3 * (int64_t)n_hw * N, sizeof(double));
  for (int64_t n = 0; n < N; ++n)
    for (int64_t m = 0; m < N; ++m) {
      if (m == n) continue;
      const double ww = (e[m] - e[n]) * (e[m] - e[n]);
      const orc_z *a1 = &A[3 * (n + N * m)], *a2 = &A[3 * (m + N * n)];
      double ov[3];
      ov[0] = cimag(a1[1] * a2[2] - a1[2] * a2[1]) * ww;
      ov[1] = cimag(a1[2] * a2[0] - a1[0] * a2[2]) * ww;
      ov[2] = cimag(a1[0] * a2[1] - a1[1] * a2[0]) * ww;
      for (int h = 0; h < n_hw; ++h)
        for (int c = 0; c < 3; ++c) omt[c + 3 * (h + (int64_t)n_hw * n)] -= ov[c] / (ww - hw[h]);
    }
  for (int64_t n = 0; n < N; ++n)
    for (int f = 0; f < n_ef; ++f)
      for (int h = 0; h < n_hw; ++h)
        for (int b = 0; b < 3; ++b)
          for (int a = 0; a < 3; ++a)
            o[a + 3 * (b + 3 * (h + (int64_t)n_hw * f))] -=
                V[a + 3 * (n + N * n)] * omt[b + 3 * (h + (int64_t)n_hw * n)] * fdd[f + n_ef * n];
  free(omt);
}

/* ------------------------------------------------------------------------ */

static void sums_zero(const orc_params *p, orc_sums *s) {
  int nef = p->n_eF, nhw = p->n_hw, nv = p->valence_bands;
  memset(s->Ne, 0, sizeof(double) * nef);
  if (s->ahc) memset(s->ahc, 0, sizeof(double) * 9 * nef);
  if (s->ahc_velo) memset(s->ahc_velo, 0, sizeof(orc_z) * 9 * nhw * nef);
  if (s->ohc) memset(s->ohc, 0, sizeof(orc_z) * 9 * nhw * nef);
  if (s->kubo_ic) memset(s->kubo_ic, 0, sizeof(double) * 9 * nef);
  if (s->kubo_lc) memset(s->kubo_lc, 0, sizeof(double) * 9 * nef);
  if (s->kubo_cs) memset(s->kubo_cs, 0, sizeof(double) * 9 * nef);
  if (s->niu_ic) memset(s->niu_ic, 0, sizeof(double) * 9 * nv);
  if (s->niu_lc) memset(s->niu_lc, 0, sizeof(double) * 9 * nv);
  if (s->niu_cs) memset(s->niu_cs, 0, sizeof(double) * 9 * nv);
  if (s->opt_s) memset(s->opt_s, 0, sizeof(orc_z) * 9 * nhw * nef);
  if (s->opt_a) memset(s->opt_a, 0, sizeof(orc_z) * 9 * nhw * nef);
  if (s->photo) memset(s->photo, 0, sizeof(orc_z) * 27 * nhw * nef);
  if (s->gyro_C) memset(s->gyro_C, 0, sizeof(orc_z) * 9 * nef);
  if (s->gyro_D) memset(s->gyro_D, 0, sizeof(orc_z) * 9 * nef);
  if (s->gyro_Dw) memset(s->gyro_Dw, 0, sizeof(orc_z) * 9 * nhw * nef);
}

static void addv(double *a, const double *b, int64_t n) {
  if (a && b) for (int64_t i = 0; i < n; ++i) a[i] += b[i];
}
static void addz(orc_z *a, const orc_z *b, int64_t n) {
  if (a && b) for (int64_t i = 0; i < n; ++i) a[i] += b[i];
}

int orc_accumulate(const orc_model *md, const orc_params *p, int64_t nk,
                   const double *kpts, orc_sums *out) {
  const int64_t N = md->n_wf, N2 = N * N;
  const int nef = p->n_eF, nhw = p->n_hw, nv = p->valence_bands;
  const int use_pos = (md->r_real != NULL);
  int nthr = p->n_threads > 0 ? p->n_threads : omp_get_max_threads();
  double *ef = malloc(sizeof(double) * nef), *hw = malloc(sizeof(double) * nhw);
  orc_linspace(p->eF_min, p->eF_max, nef, ef); /* core.f90:124-125 */
  orc_linspace(p->hw_min, p->hw_max, nhw, hw);
  sums_zero(p, out);
  orc_sums *part = calloc(nthr, sizeof(orc_sums));
  int err = 0;

#pragma omp parallel num_threads(nthr)
  {
    int t = omp_get_thread_num();
    orc_sums *s = &part[t];
    s->Ne = calloc(nef, sizeof(double));
    s->ahc = calloc(9 * nef, sizeof(double));
    s->ahc_velo = calloc(9 * nhw * nef, sizeof(orc_z));
    s->ohc = calloc(9 * nhw * nef, sizeof(orc_z));
    s->kubo_ic = calloc(9 * nef, sizeof(double));
    s->kubo_lc = calloc(9 * nef, sizeof(double));
    s->kubo_cs = calloc(9 * nef, sizeof(double));
    s->niu_ic = calloc(9 * (nv > 0 ? nv : 1), sizeof(double));
    s->niu_lc = calloc(9 * (nv > 0 ? nv : 1), sizeof(double));
    s->niu_cs = calloc(9 * (nv > 0 ? nv : 1), sizeof(double));
    s->opt_s = calloc(9 * nhw * nef, sizeof(orc_z));
    s->opt_a = calloc(9 * nhw * nef, sizeof(orc_z));
    s->photo = calloc(27 * nhw * nef, sizeof(orc_z));
    s->gyro_C = calloc(9 * nef, sizeof(orc_z));
    s->gyro_D = calloc(9 * nef, sizeof(orc_z));
    s->gyro_Dw = calloc(9 * nhw * nef, sizeof(orc_z));
    orc_z *xa = malloc(sizeof(orc_z) * 27 * nhw * nef), *xb = malloc(sizeof(orc_z) * 9 * nhw * nef);
    double *fdd = malloc(sizeof(double) * nef * N);
    double *e = malloc(sizeof(double) * N), *fd = malloc(sizeof(double) * nef * N);
    double *tmp = malloc(sizeof(double) * 9 * (nef > nv ? nef : (nv > 0 ? nv : 1)));
    double *fl = malloc(sizeof(double) * nef);
    orc_z *ztmp = malloc(sizeof(orc_z) * 9 * nhw * nef);
    orc_z *V = malloc(sizeof(orc_z) * 3 * N2);
    orc_z *A = use_pos ? malloc(sizeof(orc_z) * 3 * N2) : NULL;
    orc_z *Om = use_pos ? malloc(sizeof(orc_z) * 9 * N2) : NULL;

#pragma omp for schedule(static)
    for (int64_t ik = 0; ik < nk; ++ik) {
      int info = orc_get_wann_interp(md, &kpts[3 * ik], e, V, A, Om);
      if (info) {
#pragma omp atomic write
        err = info;
        continue;
      }
      /* Fermi setup (core.f90:187-193) */
      for (int64_t n = 0; n < N; ++n) {
        orc_fd_stat(e[n], ef, nef, p->T_kelvin, fl);
        for (int i = 0; i < nef; ++i) {
          fd[i + nef * n] = fl[i];
        }
      }
      for (int64_t n = 0; n < N; ++n) /* d1_fd_get_N_el, statistics.f90:111-124 */
        for (int i = 0; i < nef; ++i) s->Ne[i] += fd[i + nef * n];
      if (p->do_mep) { /* core.f90:198-202 */
        orc_mep_niu_ic(N, V, e, nv, md->kubo_tol, tmp); addv(s->niu_ic, tmp, 9 * nv);
        orc_mep_niu_lc(N, V, e, nv, md->kubo_tol, tmp); addv(s->niu_lc, tmp, 9 * nv);
        orc_mep_niu_cs(N, A, Om, nv, tmp); addv(s->niu_cs, tmp, 9 * nv);
      }
      if (p->do_kubo) { /* core.f90:207-216 */
        orc_kubo_mep_ic(N, e, V, nef, fd, tmp); addv(s->kubo_ic, tmp, 9 * nef);
        orc_kubo_mep_lc(N, e, V, nef, fd, tmp); addv(s->kubo_lc, tmp, 9 * nef);
        orc_kubo_mep_cs(N, A, Om, nef, fd, tmp); addv(s->kubo_cs, tmp, 9 * nef);
      }
      if (p->do_ahc) { /* core.f90:221-230 */
        orc_kubo_ahc_tens(N, e, V, nef, fd, tmp); addv(s->ahc, tmp, 9 * nef);
        orc_velo_ahc_tens(N, e, V, nhw, hw, nef, fd, p->eta, ztmp);
        addz(s->ahc_velo, ztmp, 9 * nhw * nef);
        orc_kubo_ohc_tens(N, e, V, nhw, hw, nef, fd, p->eta, ztmp);
        addz(s->ohc, ztmp, 9 * nhw * nef);
      }
      if (p->do_opt) { /* core.f90:235-243 */
        orc_kubo_opt_tens(N, e, A, nhw, hw, nef, fd, p->eta, xa, xb);
        addz(s->opt_s, xa, 9 * nhw * nef);
        addz(s->opt_a, xb, 9 * nhw * nef);
        orc_photo_2nd_cond(N, e, V, nhw, hw, nef, fd, p->eta, xa);
        addz(s->photo, xa, 27 * nhw * nef);
      }
      if (p->do_gyro) { /* core.f90:247-253 */
        orc_fd_stat_deriv(fd, nef, N, fdd);
        orc_gyro_C(N, V, nef, fdd, xa); addz(s->gyro_C, xa, 9 * nef);
        orc_gyro_D(N, e, V, Om, md->kubo_tol, nef, fdd, xa); addz(s->gyro_D, xa, 9 * nef);
        orc_gyro_Dw(N, e, V, A, nhw, hw, nef, fdd, xa); addz(s->gyro_Dw, xa, 9 * nhw * nef);
      }
    }
    free(e); free(fd); free(tmp); free(fl); free(ztmp); free(V); free(A); free(Om);
    free(xa); free(xb); free(fdd);
  }
  for (int t = 0; t < nthr; ++t) {
    orc_sums *s = &part[t];
    addv(out->Ne, s->Ne, nef);
    addv(out->ahc, s->ahc, 9 * nef);
    addz(out->ahc_velo, s->ahc_velo, 9 * nhw * nef);
    addz(out->ohc, s->ohc, 9 * nhw * nef);
    addv(out->kubo_ic, s->kubo_ic, 9 * nef);
    addv(out->kubo_lc, s->kubo_lc, 9 * nef);
    addv(out->kubo_cs, s->kubo_cs, 9 * nef);
    addv(out->niu_ic, s->niu_ic, 9 * nv);
    addv(out->niu_lc, s->niu_lc, 9 * nv);
    addv(out->niu_cs, s->niu_cs, 9 * nv);
    addz(out->opt_s, s->opt_s, 9 * nhw * nef);
    addz(out->opt_a, s->opt_a, 9 * nhw * nef);
    addz(out->photo, s->photo, 27 * nhw * nef);
    addz(out->gyro_C, s->gyro_C, 9 * nef);
    addz(out->gyro_D, s->gyro_D, 9 * nef);
    addz(out->gyro_Dw, s->gyro_Dw, 9 * nhw * nef);
    free(s->Ne); free(s->ahc); free(s->ahc_velo); free(s->ohc);
    free(s->kubo_ic); free(s->kubo_lc); free(s->kubo_cs);
    free(s->niu_ic); free(s->niu_lc); free(s->niu_cs);
    free(s->opt_s); free(s->opt_a); free(s->photo);
    free(s->gyro_C); free(s->gyro_D); free(s->gyro_Dw);
  }
  free(part); free(ef); free(hw);
  return err;
}
