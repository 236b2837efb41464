/*
 * mep_oracle.h -- CPU FP64 restatement of MEPinterp's k-space hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (libmepinterp_amd.so, the Fortran host) never links it.
 *
 * Every array follows the reference's Fortran column-major layout with the
 * Cartesian index fastest (0-based here):
 *   H_real(n,m,s)    -> H[n + N*(m + N*s)]
 *   r_real(a,n,m,s)  -> r[a + 3*(n + N*(m + N*s))]
 *   R_frac(a,s)      -> R[a + 3*s]
 *   wf_centers(a,n)  -> c[a + 3*n]
 *   a_latt(i,j)      -> a[i + 3*j]     (a_latt(i,:) = a_i, input_paras.f90:213-216)
 *   V_ka(a,n,m)      -> V[a + 3*(n + N*m)]
 *   Om_kab(a,b,n,m)  -> Om[a + 3*(b + 3*(n + N*m))]
 *   fd_distrib(ef,n) -> fd[ef + nEF*n]
 *   o(i,j,ef)        -> o[i + 3*(j + 3*ef)]
 *   o(i,j,hw,ef)     -> o[i + 3*(j + 3*(hw + nHW*ef))]
 * Energies are Hartree, lengths Bohr, as inside the reference.
 *
 * Parity unpinned: the position-operator gauge transformation in the
 * reference's evaluation order (wann_interp.f90:502-658: stale-temporary
 * connection, A^(H) before get_velo, curvature from A^(H)), the Chern-Simons
 * terms built on it (kubo_mep_CS, mep_niu_CS) and V_nn with a position
 * operator (never assigned by the reference) are restated from the source
 * only; no fixture or identity fixes their values.  The optical, photo-
 * galvanic and gyrotropic sums (kubo_opt_tens, photo_2nd_cond, get_gyro_*)
 * are pinned only by symmetry identities (Hermitian / anti-Hermitian parts,
 * index symmetry) and by brute-force loops over a tiny model in the tests;
 * their absolute values are otherwise parity unpinned.  Every other
 * function is checked against LAPACK, finite differences, closed forms, the
 * FeMn fixtures or a second independent formula (tests/test_oracle_*.py).
 */
#ifndef MEP_ORACLE_H
#define MEP_ORACLE_H
#include <complex.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex orc_z;

typedef struct {
  int64_t n_wf, n_R;
  const double *R_frac;     /* (3, nR) fractional R                       */
  const orc_z *H_real;      /* (n, n, nR) Hartree                         */
  const orc_z *r_real;      /* (3, n, n, nR) Bohr, or NULL (no position op) */
  const double *wf_centers; /* (3, n) fractional, or NULL (Wannier conv.) */
  double a_latt[9];         /* (3,3) Bohr, rows = lattice vectors         */
  int use_cart_velo;
  int do_gauge_trafo;
  double kubo_tol;
  int textbook; /* 0: reference semantics (stale-tmp A, 2*Hbar velocity);
                   1: PRB 74 195118 Eqs. 21-31 as printed              */
} orc_model;

typedef struct {
  int do_mep, do_kubo, do_ahc;
  int n_eF;
  double eF_min, eF_max; /* Hartree */
  int n_hw;
  double hw_min, hw_max; /* Hartree */
  double T_kelvin;
  double eta; /* Hartree (i_eta_smr = i*eta) */
  int valence_bands;
  int n_threads;
  int do_opt, do_gyro; /* jobs%do_opt, jobs%do_gyro (core.f90:235-253) */
} orc_params;

/* Raw (un-normalised) k-sums, all Fermi levels kept; sizes as commented. */
typedef struct {
  double *Ne;        /* (n_eF)                  sum_k N_el(E_F)         */
  double *ahc;       /* (3,3,n_eF)              kubo_ahc_tens           */
  orc_z *ahc_velo;   /* (3,3,n_hw,n_eF)         velo_ahc_tens           */
  orc_z *ohc;        /* (3,3,n_hw,n_eF)         kubo_ohc_tens           */
  double *kubo_ic;   /* (3,3,n_eF)              kubo_mep_IC             */
  double *kubo_lc;   /* (3,3,n_eF)                                      */
  double *kubo_cs;   /* (3,3,n_eF)                                      */
  double *niu_ic;    /* (3,3,valence_bands)     mep_niu_IC              */
  double *niu_lc;    /* (3,3,valence_bands)                             */
  double *niu_cs;    /* (3,3,valence_bands)                             */
  orc_z *opt_s;      /* (3,3,n_hw,n_eF)         kubo_opt_tens opt_symm  */
  orc_z *opt_a;      /* (3,3,n_hw,n_eF)         kubo_opt_tens opt_asymm */
  orc_z *photo;      /* (3,3,3,n_hw,n_eF)       photo_2nd_cond          */
  orc_z *gyro_C;     /* (3,3,n_eF)              get_gyro_C              */
  orc_z *gyro_D;     /* (3,3,n_eF)              get_gyro_D              */
  orc_z *gyro_Dw;    /* (3,3,n_hw,n_eF)         get_gyro_Dw             */
} orc_sums;

/* k_space.f90 */
double orc_unit_vol(const double a_latt[9]);
void orc_recip_latt(const double a_latt[9], double recip[9]);
int64_t orc_kpt_idx(const int mp[3], int qix, int qiy, int qiz);
void orc_rel_kpt(const int mp[3], int qix, int qiy, int qiz, double kpt[3]);
/* kpts(3, N_k) in the reference's visiting order (kiz outer, kix inner) */
void orc_mp_mesh(const int mp[3], double *kpts);
void orc_linspace(double lo, double hi, int n, double *out);

/* matrix_math.f90 zheevd('V','U') contract: A (n,n) upper triangle in,
   eigenvectors out; w ascending.  Returns 0 on success. */
int orc_zheev(int64_t n, orc_z *A, double *w);

/* wann_interp.f90 pieces */
void orc_ft_R_to_k(const orc_model *m, const double kpt[3], orc_z *H_k,
                   orc_z *H_ka, orc_z *A_ka, orc_z *Om_kab);
int orc_get_wann_interp(const orc_model *m, const double kpt[3], double *e_k,
                        orc_z *V_ka, orc_z *A_ka, orc_z *Om_kab);

/* statistics.f90 */
void orc_fd_stat(double e, const double *ef, int n_ef, double T, double *out);

/* kubo.f90 / kubo_mep.f90 / mep_niu.f90 -- each overwrites its output */
void orc_kubo_ahc_tens(int64_t n, const double *e, const orc_z *V, int n_ef,
                       const double *fd, double *o);
void orc_velo_ahc_tens(int64_t n, const double *e, const orc_z *V, int n_hw,
                       const double *hw, int n_ef, const double *fd,
                       double eta, orc_z *o);
void orc_kubo_ohc_tens(int64_t n, const double *e, const orc_z *V, int n_hw,
                       const double *hw, int n_ef, const double *fd,
                       double eta, orc_z *o);
void orc_kubo_mep_ic(int64_t n, const double *e, const orc_z *V, int n_ef,
                     const double *fd, double *o);
void orc_kubo_mep_lc(int64_t n, const double *e, const orc_z *V, int n_ef,
                     const double *fd, double *o);
void orc_kubo_mep_cs(int64_t n, const orc_z *A, const orc_z *Om, int n_ef,
                     const double *fd, double *o);
void orc_mep_niu_ic(int64_t n, const orc_z *V, const double *e, int n_val,
                    double kubo_tol, double *o);
void orc_mep_niu_lc(int64_t n, const orc_z *V, const double *e, int n_val,
                    double kubo_tol, double *o);
void orc_mep_niu_cs(int64_t n, const orc_z *A, const orc_z *Om, int n_val,
                    double *o);

/* statistics.f90:160-196 d1_fd_stat_deriv: 1 / (2 + 2 cosh(f)) evaluated on
   the occupations f = fd(ef, n) themselves, as the reference does */
void orc_fd_stat_deriv(const double *fd, int n_ef, int64_t n, double *out);
/* kubo.f90:204-405 (wann guide 12.10, 12.11, 12.14, 12.15); zero without A */
void orc_kubo_opt_tens(int64_t n, const double *e, const orc_z *A, int n_hw,
                       const double *hw, int n_ef, const double *fd, double eta,
                       orc_z *opt_s, orc_z *opt_a);
/* photo.f90:16-86 (Yang et al. PRB 97 241118 Eq. 1) */
void orc_photo_2nd_cond(int64_t n, const double *e, const orc_z *V, int n_hw,
                        const double *hw, int n_ef, const double *fd, double eta,
                        orc_z *o);
/* gyro.f90:153-268 (Souza et al. PRB 97 035158 Eqs. 2, 12, C20) */
void orc_gyro_C(int64_t n, const orc_z *V, int n_ef, const double *fdd, orc_z *o);
void orc_gyro_D(int64_t n, const double *e, const orc_z *V, const orc_z *Om,
                double kubo_tol, int n_ef, const double *fdd, orc_z *o);
void orc_gyro_Dw(int64_t n, const double *e, const orc_z *V, const orc_z *A,
                 int n_hw, const double *hw, int n_ef, const double *fdd, orc_z *o);

/* core.f90 k-loop body over an explicit list of k-points (3, nk), summed
   into s (zeroed here).  OpenMP over k, per-thread partials reduced in
   thread order.  Returns 0, or the first failing eigensolve's code. */
int orc_accumulate(const orc_model *m, const orc_params *p, int64_t nk,
                   const double *kpts, orc_sums *s);

#ifdef __cplusplus
}
#endif
#endif
