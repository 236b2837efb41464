/*
 * mepinterp.h -- C-ABI of the MI355X (gfx950) k-space engine for MEPinterp.
 *
 * The reference (cMax347/MEPinterp) has no C-ABI: its boundary is a set of
 * Fortran module procedures.  Each entry point below replaces one of them;
 * the Fortran host in mepinterp_amd/fortran binds these through
 * iso_c_binding and keeps the reference's call signatures.
 *
 * Conventions
 *   - every array is column-major in the reference's index order, Cartesian
 *     index fastest: V(3,n,n), Om(3,3,n,n), e_k(n,nk), kpts(3,nk), ...
 *   - complex arrays are interleaved (re,im) doubles, binary-compatible with
 *     Fortran complex(c_double_complex) and C double _Complex
 *   - energies in Hartree, lengths in Bohr, k in fractional coordinates
 *   - every function returns MEP_OK (0) or a negative MEP_ERR_* code;
 *     mep_last_error() describes the last failure of the calling thread
 *   - one GPU, one HIP stream per context; a context is not re-entrant
 */
#ifndef MEPINTERP_H
#define MEPINTERP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MEP_OK 0
#define MEP_ERR_ARG (-1)         /* invalid argument / unsupported size     */
#define MEP_ERR_HIP (-2)         /* HIP runtime error                       */
#define MEP_ERR_EIG (-3)         /* eigensolver did not converge (zheevd info != 0) */
#define MEP_ERR_NOMEM (-4)       /* device allocation failed                */

/* model flags (mep_ctx_create) */
#define MEP_F_GAUGE_TRAFO 1      /* wannInterp%doGaugeTrafo                 */
#define MEP_F_CART_VELO 2        /* wannInterp%use_cart_velo                */
#define MEP_F_TEXTBOOK 4         /* PRB 74 195118 Eqs. 21-31 as printed instead
                                    of the reference's evaluation order (only
                                    differs with a position operator)       */
#define MEP_F_APPLY_DEG 8        /* divide X(R) by deg(R); the reference reads
                                    deg but never applies it                */
#define MEP_F_SEPARATE_VELO 16   /* runs (mep_run_mesh / _kpts / _batch): keep
                                    the velocities in their own kernel, V
                                    through HBM.  By default, for FP64,
                                    17 <= n_wf <= 32, no position operator,
                                    the Wannier convention and no optical /
                                    gyrotropic jobs, the band-sum kernel forms
                                    V itself (same arithmetic, same results) */

/* band-sum jobs (mep_run_params.jobs) */
#define MEP_JOB_AHC 1            /* jobs%do_ahc : kubo_ahc_tens, velo_ahc_tens, kubo_ohc_tens */
#define MEP_JOB_KUBO 2           /* jobs%do_kubo: kubo_mep_IC/LC/CS          */
#define MEP_JOB_MEP 4            /* jobs%do_mep : mep_niu_IC/LC/CS           */
#define MEP_JOB_OPT 8            /* jobs%do_opt : kubo_opt_tens (needs the position
                                    operator) and photo_2nd_cond            */
#define MEP_JOB_GYRO 16          /* jobs%do_gyro: get_gyro_C / D / Dw         */

/* arithmetic of the k-space pipeline */
#define MEP_PREC_FP64 0          /* FP64 Fourier sums (separable on meshes,
                                    MFMA GEMM for k lists), FP64 eigensolve */
#define MEP_PREC_MIXED 1         /* H(k) in FP32 (FP64 separable sums rounded
                                    on meshes, FP32 MFMA GEMM for k lists) and
                                    FP32 eigensolve;
                                    FP64 Rayleigh-quotient refinement of every
                                    eigenvalue and Rayleigh-Ritz of groups
                                    closer than 1e-3 eV; FP64 velocities, band
                                    sums and k-sums (no position operator)  */

typedef struct mep_ctx mep_ctx;

typedef struct {
  int32_t jobs;                  /* MEP_JOB_* bits                           */
  int32_t n_eF;                  /* Fermi%N_eF                               */
  double eF_min, eF_max;         /* Fermi%eF_min/max  (Hartree)              */
  int32_t n_hw;                  /* Laser%N_hw (>= 1)                        */
  double hw_min, hw_max;         /* Laser%hw_min/max  (Hartree)              */
  double T_kelvin;               /* Fermi%Tkelvin                            */
  double eta;                    /* Fermi%eta_smearing (Hartree)             */
  int32_t valence_bands;         /* MEP%valence_bands                        */
  int32_t precision;             /* MEP_PREC_*                               */
  const double *hw_lst;          /* (n_hw) frequencies [Hartree] as the caller's
                                    hw_lst argument of velo_ahc_tens /
                                    kubo_ohc_tens (kubo.f90:86-104, 145-164),
                                    any spacing; NULL = get_linspace(hw_min,
                                    hw_max, n_hw) (matrix_math.f90:74-100)   */
} mep_run_params;

/* Raw k-sums (not normalised; all Fermi levels kept).  Host buffers owned
   by the caller; a NULL member is skipped.                                  */
typedef struct {
  double *Ne;          /* (n_eF)              sum_k N_el(E_F)               */
  double *ahc;         /* (3,3,n_eF)          sum_k kubo_ahc_tens           */
  double *ahc_velo;    /* (2,3,3,n_hw,n_eF)   sum_k velo_ahc_tens (complex) */
  double *ohc;         /* (2,3,3,n_hw,n_eF)   sum_k kubo_ohc_tens (complex) */
  double *kubo_ic;     /* (3,3,n_eF)          sum_k kubo_mep_IC             */
  double *kubo_lc;     /* (3,3,n_eF)          sum_k kubo_mep_LC             */
  double *kubo_cs;     /* (3,3,n_eF)          sum_k kubo_mep_CS             */
  double *niu_ic;      /* (3,3,valence_bands) sum_k mep_niu_IC              */
  double *niu_lc;      /* (3,3,valence_bands) sum_k mep_niu_LC              */
  double *niu_cs;      /* (3,3,valence_bands) sum_k mep_niu_CS              */
  int64_t n_k;         /* out: number of k-points summed                    */
  double *opt_s;       /* (2,3,3,n_hw,n_eF)   sum_k kubo_opt_tens opt_symm  */
  double *opt_a;       /* (2,3,3,n_hw,n_eF)   sum_k kubo_opt_tens opt_asymm */
  double *photo;       /* (2,3,3,3,n_hw,n_eF) sum_k photo_2nd_cond          */
  double *gyro_C;      /* (2,3,3,n_eF)        sum_k get_gyro_C              */
  double *gyro_D;      /* (2,3,3,n_eF)        sum_k get_gyro_D              */
  double *gyro_Dw;     /* (2,3,3,n_hw,n_eF)   sum_k get_gyro_Dw             */
} mep_sums;

/* Upload the real-space model once.  Replaces the model half of
   read_tb_basis (file_io.f90:800-860) + kspace_allocator
   (k_space.f90:219-236).
     R_frac(3,n_R); H_real(n,n,n_R) [Hartree]; r_real(3,n,n,n_R) [Bohr] or
     NULL; wf_centers(3,n) or NULL (NULL = Wannier convention, given =
     tight-binding convention, wann_interp.f90:192-280); deg(n_R) or NULL;
     a_latt(3,3) with a_latt(i,:) = a_i (input_paras.f90:213-216).
   Sizes: n_wf <= 128, with or without a position operator.  The mixed-
   precision path takes n_wf <= 32; the optical / gyrotropic sums take
   n_wf x n_eF up to about 9,600.                                            */
int mep_ctx_create(int64_t n_wf, int64_t n_R, const double *R_frac, const double *wf_centers,
                   const double *H_real, const double *r_real, const int64_t *deg,
                   const double *a_latt, double kubo_tol, int32_t flags, mep_ctx **ctx);
int mep_ctx_destroy(mep_ctx *ctx);
const char *mep_last_error(void);

/* get_wann_interp (wann_interp.f90:56-108) for nk k-points at once.
     kpts(3,nk) -> e_k(n,nk), V_ka(3,n,n,nk) or NULL,
     A_ka(3,n,n,nk), Om_kab(3,3,n,n,nk) (only with a position operator;
     otherwise NULL).                                                        */
int mep_interp_batch(mep_ctx *ctx, int64_t nk, const double *kpts, double *e_k, double *V_ka,
                     double *A_ka, double *Om_kab);

/* The band sums of core.f90:187-230 over nk k-points whose e_k, V_ka (and
   A_ka, Om_kab) the caller provides: fd_stat + fd_get_N_el
   (statistics.f90:65-124), kubo_ahc_tens / velo_ahc_tens / kubo_ohc_tens
   (kubo.f90:37-198), kubo_mep_IC/LC/CS (kubo_mep.f90:16-149),
   mep_niu_IC/LC/CS (mep_niu.f90:20-159).  With nk = 1 the outputs are the
   reference functions' results.  fd_distrib(n_eF,n,nk) supplies the
   occupations as the reference's kubo_* arguments do; NULL = fd_stat of
   e_k at the p->n_eF Fermi levels and p->T_kelvin.                          */
int mep_bandsum_batch(mep_ctx *ctx, int64_t nk, const double *e_k, const double *V_ka,
                      const double *A_ka, const double *Om_kab, const double *fd_distrib,
                      const mep_run_params *p, mep_sums *out);

/* The whole k-loop of core_worker (core.f90:165-270) over the
   Monkhorst-Pack mesh mp_grid (k_space.f90:341-375), on the device; the
   raw sums come back to the host.  fermi_selector and normalize_k_int
   (core.f90:317-603) stay with the caller.                                  */
int mep_run_mesh(mep_ctx *ctx, const int32_t mp_grid[3], const mep_run_params *p, mep_sums *out);

/* The k-points mep_run_mesh generates on the device, k0 .. k0+nk-1 in the
   visiting order of core.f90:165-170, read back into kpts(3,nk): get_rel_kpt
   of get_kpt_idx(kix,kiy,kiz) (k_space.f90:341-375), x fastest.  Bit-exact
   with the reference's (2 q - N - 1) / (2 N).                              */
int mep_mesh_kpts(mep_ctx *ctx, const int32_t mp_grid[3], int64_t k0, int64_t nk, double *kpts);

/* Same over an explicit k list (3,nk) with optional weights(nk) (NULL = 1):
   sub-mesh refinement and sampled baselines.                                */
int mep_run_kpts(mep_ctx *ctx, int64_t nk, const double *kpts, const double *weights,
                 const mep_run_params *p, mep_sums *out);

/* mep_run_kpts without device allocation.  No counterpart in the reference;
   for callers that cannot allocate device memory: code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc.  Every work
   buffer of the call (k-points, phase tile, H(k), eigenvectors, velocities,
   eigensolver records, band-sum records) comes from one file-scope device
   array of MEP_STATIC_SCRATCH_BYTES on each GPU, held until the process
   ends; the model stays where mep_ctx_create put it.  Same kernels, same k
   -> workgroup map and same sums as mep_run_kpts on the same list.
   Limit: the buffers of the nk k-points must fit the array (about 190 KB per
   k-point at n_wf = 32, NR = 343, so about 350 k-points); a larger nk returns
   MEP_ERR_ARG and mep_last_error() names the largest nk that fits.  Every
   call on one GPU uses that GPU's array: calls that overlap on one GPU
   (other threads, other contexts) race.                                    */
#define MEP_STATIC_SCRATCH_BYTES (64ull << 20)
int mep_run_kpts_static(mep_ctx *ctx, int64_t nk, const double *kpts, const double *weights,
                        const mep_run_params *p, mep_sums *out);

/* Kernel timing with HIP events on the context's stream (off by default).
   name: "phase", "ft_gemm" (the Fourier sums: separable on meshes, GEMM for
   k lists), "eigh", "refine", "velo", "bandsum", "reduce", "kgen".          */
/* Several models in one call: mep_run_mesh (the k-loop of core.f90:165-270)
   for each context i on the mesh mp_grids(3,i) with params[i]; the k-blocks
   of all models are enqueued on their contexts' own streams, interleaved
   block by block so that the GPU runs them concurrently, and waited for
   once.  No counterpart in the reference, which runs one model per
   process.  outs[i] as for mep_run_mesh (outs may be NULL).                */
int mep_run_batch(int32_t n_models, mep_ctx *const *ctxs, const int32_t *mp_grids,
                  const mep_run_params *params, mep_sums *outs);
/* mep_run_batch with the batch's kernels (zeroing, every k-block of every
   model on the models' streams, the reductions) captured once into a HIP
   graph: the first call for a set of (contexts, meshes, parameters)
   captures, later calls with the same set replay the graph after the host
   set-up.  A reallocation, a variant change or mep_ctx_destroy of one of
   the contexts drops its graphs (the next call recaptures).  The same sums
   as mep_run_batch, bit for bit.  No counterpart in the reference.          */
int mep_run_batch_graph(int32_t n_models, mep_ctx *const *ctxs, const int32_t *mp_grids,
                        const mep_run_params *params, mep_sums *outs);

/* Per-k Berry curvature at one Fermi level: omega(9,nk), element i + 3j =
   sum_n f(n) sum_{m != n} 2 Im(V_i,nm V_j,mn) / (E_m - E_n)^2, i.e. the
   k-point's summand of kubo_ahc_tens (kubo.f90:37-78) before the k-sum, with
   the occupations of d1_fd_stat (statistics.f90:65-98).  It feeds the
   learned refinement below; the reference has no per-k output of it.       */
int mep_curv_batch(mep_ctx *ctx, int64_t nk, const double *kpts, double eF, double T_kelvin,
                   double *omega);

/* Learned adaptive refinement.  No counterpart in the reference, which
   integrates on uniform meshes only: a small MLP (sin/cos Fourier features
   of k, three tanh layers of `hidden` units, FP32 weights) learns
   log|Omega(k)| from a coarse mesh and ranks the cells to refine.  Training
   (mean squared error, Adam) and inference run in HIP kernels.
   Parameters: W1[6 n_freq padded to 16][hidden] b1 W2[hidden][hidden] b2
   W3[hidden][hidden] b3 W4[hidden][16] b4[16] (column 0 is the output).    */
typedef struct mep_mlp mep_mlp;
int mep_mlp_create(int32_t n_freq, int32_t hidden, uint64_t seed, mep_mlp **mlp);
int mep_mlp_destroy(mep_mlp *mlp);
int mep_mlp_num_params(const mep_mlp *mlp, int64_t *n_params);
int mep_mlp_get_params(mep_mlp *mlp, float *params);
int mep_mlp_set_params(mep_mlp *mlp, const float *params);
/* epochs over (kpts(3,n), target(n)) in shuffled batches (seeded); the mean
   loss of the last epoch in *final_loss */
int mep_mlp_train(mep_mlp *mlp, int64_t n, const double *kpts, const float *target,
                  int32_t epochs, int32_t batch, float lr, uint64_t seed, float *final_loss);
int mep_mlp_predict(mep_mlp *mlp, int64_t n, const double *kpts, float *out);
/* loss and gradient of one batch of all n points, parameters unchanged */
int mep_mlp_grad(mep_mlp *mlp, int64_t n, const double *kpts, const float *target, float *grad,
                 float *loss);

int mep_set_profiling(mep_ctx *ctx, int32_t on);
int mep_get_kernel_time(mep_ctx *ctx, const char *name, double *total_ms, int64_t *launches);
int mep_reset_kernel_times(mep_ctx *ctx);

/* Largest k-block the context processes per pass (device workspace bound). */
int mep_set_kblock(mep_ctx *ctx, int64_t kblock);
/* Upper bound on the bytes of per-k-block work buffers (default 64e9); k-blocks
   shrink to fit.  For several resident contexts (mep_run_batch) whose
   buffers must share the 288 GB of HBM.  No counterpart in the reference.  */
int mep_set_workspace_limit(mep_ctx *ctx, double bytes);
/* Kernel variant of one stage (value 0 restores the library's default).  A
   tuning knob with no counterpart in the reference; mepinterp_amd/tune.py
   times the variants and caches its choice.  Every variant gives the same
   sums to rounding.
     MEP_VAR_EIGH  FP64 Householder reduction.  16 < n <= 32:
                   1 = one k-point per wavefront, half a row per lane (default);
                   2 = two k-points per wavefront, one row per lane.
                   32 < n: 1 = one 256-thread workgroup per k-point with the
                   matrix in LDS (default for 64 < n <= 88); 2 = one
                   wavefront per k-point with the matrix in LDS (n <= 64;
                   default for n <= 44); 3 = n / 16 wavefronts per k-point
                   with the matrix in registers (default for 44 < n <= 64
                   and n > 88).
     MEP_VAR_GEMM32  the FP32 Fourier GEMM of the mixed path on k lists:
                   1 = FP32 MFMA (v_mfma_f32_16x16x4_f32); 2 = split bfloat16
                   (three bf16 parts per operand, six products accumulated in
                   FP32 on bf16 MFMA: FP32-equivalent) (default). */
#define MEP_VAR_EIGH 1
#define MEP_VAR_GEMM32 2
int mep_set_variant(mep_ctx *ctx, int32_t what, int32_t value);

#ifdef __cplusplus
}
#endif
#endif
