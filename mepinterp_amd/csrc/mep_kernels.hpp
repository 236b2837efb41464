// mep_kernels.hpp -- launchers of the k-space kernels (host side).
#pragma once
#include "mep_common.hpp"

// Per-k geometry handed to the consumers of the Fourier GEMM output.
struct KGeom {
  const double* kpts;  // (3, nk) fractional k of the block
  const double* tau;   // (3, n) Wannier centres (tight-binding convention) or null
  double cmat[9];      // c(dR) = cmat * dR, row-major 3x3 (2*pi*I, or a_latt)
};

void launch_mp_kpts(hipStream_t s, int64_t k0, int64_t nk, const int mp[3], double* kpts);
// real phase tile Pt[K][nk] (cos rows, sin rows of the Kh paired R, and a
// row of ones when has_zero) and the real GEMM C[M][N] = Pt^T X
void launch_phase(hipStream_t s, int64_t nk, int Kh, int has_zero, const double* kpts,
                  const double* Rh, double* Pt);
void launch_ft_gemm_f64(hipStream_t s, int M, int N, int K, const double* Pt, const double* X,
                        double* C);
// mixed precision: FP32 phase tile and FP32 GEMM (FP32 MFMA accumulation)
// Separable Fourier sums on a Monkhorst-Pack mesh (mep_ft.hip): box
// coefficients Xb [bx][by][bz][NC], per-mesh-index phase tables
// ex [n1][BX] (zero-padded to BX = 8 or 16), ey [n2][by], ez [n3][bz], and
// the z-stage workspace Y1 [planes][bx][by][NC] for the planes of a k-block.
struct SepFT {
  int64_t NC;
  int bx, by, bz;
  int n1, n2, n3;
  const zd* Xb;
  const zd* ex;
  const zd* ey;
  const zd* ez;
  zd* Y1;
};
// H(k) of the mesh k-points k0 .. k0+nk-1 into H64 (FP64) or, when H32 is
// not null, H32 (rounded to FP32), in the Fourier GEMM's output layout
void launch_sep_ft(hipStream_t s, const SepFT& f, int64_t k0, int64_t nk, zd* H64, zf* H32);
void launch_phase_f32(hipStream_t s, int64_t nk, int Kh, int has_zero, const double* kpts,
                      const double* Rh, float* Pt);
// the same product in split bfloat16 (three bf16 parts per operand, six
// products accumulated in FP32 on bf16 MFMA): FP32-equivalent
void launch_ft_gemm_bf16x3(hipStream_t s, int M, int N, int K, const float* Pt, const float* X,
                           float* C);
void launch_ft_gemm_f32(hipStream_t s, int M, int N, int K, const float* Pt, const float* X,
                        float* C);

// Batched Hermitian eigensolve of H(k) (upper triangle), n <= 32
// (launch_eigh_wide: 32 < n <= 128, FP64).
// E (nk, n) ascending; U (nk, n, n) column-major eigenvectors.  scratch
// holds eigh_scratch_bytes(n, nk) bytes (side records + rotation records).
// packed: H(k) holds the upper triangle only (element i <= j at i + j(j+1)/2).
size_t eigh_scratch_bytes(int n, int64_t nk);
// house: FP64 Householder variant for 16 < n <= 32 (1 = one k-point per
// wavefront, 2 = two; 0 = the default, 1)
void launch_eigh(hipStream_t s, int n, int64_t nk, const zd* Hk, int64_t strideHk, int packed,
                 const KGeom& g, double* E, zd* U, int* info, double* scratch, int house = 0);
// house (32 < n): 1 = one 256-thread workgroup per k-point, 2 = one wavefront
// per k-point with the matrix in LDS (n <= 64); 0 = the default (2 for n <= 52)
void launch_eigh_wide(hipStream_t s, int n, int64_t nk, const zd* Hk, int64_t strideHk, int packed,
                      const KGeom& g, double* E, zd* U, int* info, double* scratch, int house = 0);
// the same in FP32 arithmetic (FP32 H(k) and U; eigenvalues written as FP64)
void launch_eigh_f32(hipStream_t s, int n, int64_t nk, const zf* Hk, int64_t strideHk, int packed,
                     const KGeom& g, double* E, zf* U, int* info, double* scratch);

// Velocity (and, with a position operator, connection and curvature) in
// the Hamiltonian gauge.  V is written planar per k: V[k][a][i + n j].
struct VeloArgs {
  int n, nmat;
  int packed;  // upper-triangle matrices (no position operator)
  int gauge_trafo;
  int pos_op;
  int textbook;
  double kubo_tol;
  zd* scratch;  // n > 32: 3 n x n per workgroup (up to 2048 workgroups); with a
                // position operator and n > 16: 18 n x n per workgroup
  size_t scratch_bytes = 0;
};
// the scratch holds at least one position-operator slot (or none is needed)
bool velo_pos_scratch_ok(int n, size_t scratch_bytes);
void launch_velo(hipStream_t s, const VeloArgs& va, int64_t nk, const zd* Hk, int64_t strideHk,
                 const KGeom& g, const zd* U, const double* E, zd* V, zd* A, zd* Om);

// mixed precision: dH_a from the FP32 Fourier sum (no position operator),
// U and V in FP64
void launch_velo_h32(hipStream_t s, const VeloArgs& va, int64_t nk, const zf* Hk,
                     int64_t strideHk, const KGeom& g, const zd* U, zd* V);
// FP64 refinement of FP32 eigenpairs: Rayleigh quotients for every band,
// Rayleigh-Ritz inside groups closer than dE_thr (Hartree); U out in FP64.
void launch_refine(hipStream_t s, int n, int64_t nk, const zf* Hk, int64_t strideHk, int packed,
                   const KGeom& g, double dE_thr, const zf* U32, double* E, zd* U);

// V layout conversion: planar [a][i + n j] <-> reference (3,n,n), per k.
void launch_v_layout(hipStream_t s, int n, int64_t nk, const zd* src, zd* dst, bool to_ref);

// Band sums over a batch of k with per-block partial accumulators
// (V planar).
struct BandArgs {
  int n;
  int jobs;  // MEP_JOB_* bits
  int n_ef, n_hw, n_val;
  double T_kelvin, eta, kubo_tol;
  const double* ef;  // (n_ef) device
  const double* hw;  // (n_hw) device
  const double* fd_in;  // (n_ef, n, nk) device occupations, or null (fd_stat)
  int pos_op;
  int v_herm;  // V(n, m) = conj(V(m, n)) may be assumed (V built here, no position operator)
};
// layout of one accumulator record (doubles)
struct AccLayout {
  int64_t Ne, ahc, velo, ohc, kic, klc, kcs, nic, nlc, ncs, total;
};
AccLayout acc_layout(const BandArgs& b);
// V formed inside the band-sum kernel instead of read from HBM: the packed
// H(k) / dH_a(k) of the Fourier stage (per-k stride strideHk) and U.  Only
// where bandsum_can_fuse(b) (17 <= n <= 32, V built here, no position
// operator) and for the Wannier convention.
struct VeloFuse {
  const zd* Hk;
  int64_t strideHk;
  const zd* U;
};
bool bandsum_can_fuse(const BandArgs& b);
// wide_scr (n > 32): 3 n^2 complex per workgroup (nblocks of them), the
// columns of V scaled by 1 / (E_n0 - E_m) for the itinerant products
void launch_bandsum(hipStream_t s, const BandArgs& b, int64_t nk, const double* E, const zd* V,
                    const zd* A, const zd* Om, const double* w, int nblocks, double* part,
                    const VeloFuse* fuse = nullptr, zd* wide_scr = nullptr);
void launch_reduce_parts(hipStream_t s, int nblocks, int64_t len, const double* part, double* acc);

// optical / photogalvanic / gyrotropic sums (MEP_JOB_OPT, MEP_JOB_GYRO):
// record layout (complex interleaved doubles) and launcher; n_wf <= 128,
// n_eF x n_wf bounded by ext_lds_bytes
struct ExtLayout {
  int64_t opt_s, opt_a, photo, gyro_C, gyro_D, gyro_Dw, total;
};
ExtLayout ext_layout(const BandArgs& b);
// dynamic LDS of the optical / gyrotropic kernel (must stay within 160 KiB)
size_t ext_lds_bytes(const BandArgs& b);
// scr: nblocks slots of n^2 complex (photo_2nd_cond's per-(b, c) product
// where it does not fit in LDS), or null to keep the per-item loop there
void launch_extsum(hipStream_t s, const BandArgs& b, int64_t nk, const double* E, const zd* V,
                   const zd* A, const zd* Om, const double* w, int nblocks, double* part,
                   zd* scr);

// per-k Berry curvature summed over occupied bands at one E_F (V planar):
// out (nk, 9), element i + 3j
void launch_curv(hipStream_t s, int n, int64_t nk, const double* E, const zd* V, double ef,
                 double T_kelvin, double* out);
