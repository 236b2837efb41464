// mep_capi.hip -- the C-ABI of include/mepinterp.h: model upload, k-block
// workspace, the per-block pipeline (phase -> Fourier GEMM -> eigensolve ->
// velocity -> band sums -> reduction) and kernel timing.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mepinterp.h"
#include "mep_common.hpp"
#include "mep_kernels.hpp"

static thread_local std::string g_err;

// error text for the other translation units of the library
void mep_set_error(const std::string& msg) { g_err = msg; }

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                       \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess)                                                               \
      return fail(MEP_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));         \
  } while (0)

#define CHK(x)              \
  do {                      \
    int r_ = (x);           \
    if (r_ != MEP_OK) return r_; \
  } while (0)

enum Kind { K_KGEN = 0, K_PHASE, K_GEMM, K_EIGH, K_REFINE, K_VELO, K_BANDSUM, K_REDUCE, K_NKIND };
static const char* kKindName[K_NKIND] = {"kgen", "phase", "ft_gemm", "eigh", "refine", "velo",
                                         "bandsum", "reduce"};

struct EvPair {
  hipEvent_t a, b;
  int kind;
};

size_t bandsum_lds_bytes(const BandArgs& b);
void launch_fill_zero(hipStream_t s, double* p, int64_t n);

struct mep_ctx {
  int device = 0;  // the HIP device current at mep_ctx_create
  int64_t n = 0, nR = 0;
  int32_t flags = 0;
  bool pos = false, tb = false;
  double al[9];
  double kubo_tol = 0;
  double cmat[9];
  int nmat = 0;
  int packed = 0;       // upper triangles only (no position operator)
  int64_t np = 0;       // complex elements stored per matrix
  int64_t ncols = 0;    // complex columns of the GEMM output per k
  int64_t hk_per_k = 0; // complex elements of the Hk slot per k
  int Kh = 0, has_zero = 0;
  int64_t K = 0;        // GEMM depth: 2 Kh + has_zero
  double* dR = nullptr;  // (3, Kh) R of the paired entries
  double* dtau = nullptr;
  double* dX = nullptr;  // (K, 2 ncols) real coefficients
  float* dX32 = nullptr;  // the same in FP32 (mixed-precision path)
  // separable sums on Monkhorst-Pack meshes: the coefficients in a box of
  // integer R (origin sox, soy, soz; extent sbx x sby x sbz), the phase
  // tables of the current mesh and the z-stage workspace
  bool sep = false;
  int sox = 0, soy = 0, soz = 0, sbx = 0, sby = 0, sbz = 0;
  zd* dXb = nullptr;
  zd* dPh = nullptr;
  int64_t ph_cap = 0;
  int sep_mesh[3] = {0, 0, 0};
  zd* dY1 = nullptr;
  int64_t y1_cap = 0;
  // k-block workspace
  int64_t kcap = 0, kblock_req = 0;
  int var_eigh = 0;  // MEP_VAR_EIGH
  int var_gemm32 = 0;  // MEP_VAR_GEMM32
  // bumped whenever a device buffer is (re)allocated or a kernel variant
  // changes: a captured batch graph holds the pointers and kernels of one
  // generation
  uint64_t gen = 0;
  double* dk = nullptr;
  double* dw = nullptr;
  double* dP = nullptr;  // (K, kblock) real phase tile
  zd* dHk = nullptr;
  double* dE = nullptr;
  zd* dU = nullptr;
  zd* dV = nullptr;
  zd* dA = nullptr;
  zd* dOm = nullptr;
  int* dinfo = nullptr;
  double* dscr = nullptr;  // eigensolver records
  double* dcurv = nullptr;  // per-k curvature (mep_curv_batch)
  double ws_limit = 64.0e9;  // bytes of per-block work buffers (of 288 GB HBM)
  int64_t curv_cap = 0;
  double* dpart = nullptr;
  int64_t part_cap = 0;
  double* dacc = nullptr;
  int64_t acc_cap = 0;
  double* def = nullptr;
  double* dhw = nullptr;
  int64_t ef_cap = 0, hw_cap = 0;
  double* dfd = nullptr;
  int64_t fd_cap = 0;
  hipStream_t st = nullptr;
  // profiling
  bool prof = false;
  std::vector<EvPair> pending;
  std::vector<hipEvent_t> pool;
  double tms[K_NKIND] = {0};
  int64_t tcnt[K_NKIND] = {0};
};

// ---------------------------------------------------------------------------
// timing helpers
// ---------------------------------------------------------------------------
static hipEvent_t getev(mep_ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e;
  hipEventCreate(&e);
  return e;
}

template <class F>
static void timed(mep_ctx* c, int kind, F&& f) {
  if (!c->prof) {
    f();
    return;
  }
  hipEvent_t a = getev(c), b = getev(c);
  hipEventRecord(a, c->st);
  f();
  hipEventRecord(b, c->st);
  c->pending.push_back({a, b, kind});
}

static void resolve_times(mep_ctx* c) {
  for (auto& p : c->pending) {
    hipEventSynchronize(p.b);
    float ms = 0.f;
    hipEventElapsedTime(&ms, p.a, p.b);
    c->tms[p.kind] += ms;
    c->tcnt[p.kind] += 1;
    c->pool.push_back(p.a);
    c->pool.push_back(p.b);
  }
  c->pending.clear();
}

// ---------------------------------------------------------------------------
// workspace
// ---------------------------------------------------------------------------
static void free_ws(mep_ctx* c) {
  ++c->gen;
  hipFree(c->dk); hipFree(c->dw); hipFree(c->dP); hipFree(c->dHk); hipFree(c->dE);
  hipFree(c->dU); hipFree(c->dV); hipFree(c->dA); hipFree(c->dOm); hipFree(c->dscr);
  c->dscr = nullptr;
  c->dk = c->dw = c->dP = nullptr; c->dHk = c->dU = c->dV = c->dA = c->dOm = nullptr;
  c->dE = nullptr;
  c->kcap = 0;
}

static int64_t bytes_per_k(const mep_ctx* c) {
  int64_t n2 = c->n * c->n;
  int64_t b = 3 * 8 + 8 + c->K * 8 + c->hk_per_k * 16 + c->n * 8 + n2 * 16 + 3 * n2 * 16;
  if (c->pos) b += 12 * n2 * 16;
  b += (int64_t)eigh_scratch_bytes((int)c->n, 1);
  return b;
}

static int ensure_ws(mep_ctx* c, int64_t nk_total) {
  // default: about 16 GB of work buffers per k-block (HBM holds 288 GB; see
  // mep_set_workspace_limit), split evenly over the blocks of this run
  const int64_t fit = (int64_t)(c->ws_limit / (double)bytes_per_k(c));
  int64_t want = c->kblock_req > 0 ? c->kblock_req : fit;
  if (want > fit) want = fit;
  if (want < 1) want = 1;
  if (want > nk_total) want = nk_total;
  if (c->kblock_req <= 0 && want < nk_total) {
    const int64_t nblk = (nk_total + want - 1) / want;
    want = (nk_total + nblk - 1) / nblk;
  }
  want = ((want + 63) / 64) * 64;
  if (want < 64) want = 64;
  if (want <= c->kcap) return MEP_OK;
  free_ws(c);  // bumps the generation
  int64_t n2 = c->n * c->n;
  HIPCHK(hipMalloc(&c->dk, sizeof(double) * 3 * want));
  HIPCHK(hipMalloc(&c->dw, sizeof(double) * want));
  HIPCHK(hipMalloc(&c->dP, sizeof(double) * c->K * want));
  HIPCHK(hipMalloc(&c->dHk, sizeof(zd) * c->hk_per_k * want));
  HIPCHK(hipMalloc(&c->dE, sizeof(double) * c->n * want));
  HIPCHK(hipMalloc(&c->dU, sizeof(zd) * n2 * want));
  HIPCHK(hipMalloc(&c->dV, sizeof(zd) * 3 * n2 * want));
  HIPCHK(hipMalloc(&c->dscr, eigh_scratch_bytes((int)c->n, want)));
  if (c->pos) {
    HIPCHK(hipMalloc(&c->dA, sizeof(zd) * 3 * n2 * want));
    HIPCHK(hipMalloc(&c->dOm, sizeof(zd) * 9 * n2 * want));
  }
  c->kcap = want;
  return MEP_OK;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" const char* mep_last_error(void) { return g_err.c_str(); }

extern "C" int mep_ctx_create(int64_t n_wf, int64_t n_R, const double* R_frac,
                              const double* wf_centers, const double* H_real, const double* r_real,
                              const int64_t* deg, const double* a_latt, double kubo_tol,
                              int32_t flags, mep_ctx** out) {
  if (!out) return fail(MEP_ERR_ARG, "mep_ctx_create: ctx pointer is NULL");
  *out = nullptr;
  if (n_wf <= 0 || n_R <= 0 || !R_frac || !H_real || !a_latt)
    return fail(MEP_ERR_ARG, "mep_ctx_create: n_wf, n_R, R_frac, H_real and a_latt are required");
  if (n_wf > 128) return fail(MEP_ERR_ARG, "mep_ctx_create: n_wf > 128 is not supported");
  if (deg && (flags & MEP_F_APPLY_DEG))
    for (int64_t s = 0; s < n_R; ++s)
      if (deg[s] <= 0) return fail(MEP_ERR_ARG, "mep_ctx_create: deg(R) must be positive");
  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count <= 0)
    return fail(MEP_ERR_HIP, "mep_ctx_create: no HIP device");

  mep_ctx* c = new mep_ctx();
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    return fail(MEP_ERR_HIP, "mep_ctx_create: hipGetDevice failed");
  }
  c->n = n_wf;
  c->nR = n_R;
  c->flags = flags;
  c->pos = (r_real != nullptr);
  c->tb = (wf_centers != nullptr);
  c->kubo_tol = kubo_tol;
  for (int i = 0; i < 9; ++i) c->al[i] = a_latt[i];
  // c(dR) = 2 pi dR (internal) or intern_to_cart = matmul(a_latt, dR)
  // (wann_interp.f90:215-233, 387-399); row-major cmat[3i+j] = a_latt(i,j)
  const double two_pi = 2.0 * 3.14159265358979323846;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      c->cmat[3 * i + j] = (flags & MEP_F_CART_VELO) ? a_latt[i + 3 * j] : (i == j ? two_pi : 0.0);
  c->nmat = c->pos ? 10 : 4;
  c->packed = c->pos ? 0 : 1;
  const int64_t n = n_wf, n2 = n * n;
  c->np = c->packed ? n * (n + 1) / 2 : n2;
  c->ncols = c->nmat * c->np;
  // the Hk slot also serves as staging for V in the reference layout
  c->hk_per_k = c->ncols > 3 * n2 ? c->ncols : 3 * n2;

  // pair R with -R (exact match); R = 0 entries form the constant row
  std::vector<int64_t> partner(n_R, -1);
  std::vector<char> is_zero(n_R, 0);
  for (int64_t s = 0; s < n_R; ++s)
    is_zero[s] = (R_frac[3 * s] == 0.0 && R_frac[3 * s + 1] == 0.0 && R_frac[3 * s + 2] == 0.0);
  for (int64_t s = 0; s < n_R; ++s) {
    if (is_zero[s] || partner[s] >= 0) continue;
    for (int64_t t = s + 1; t < n_R; ++t)
      if (!is_zero[t] && partner[t] < 0 && R_frac[3 * t] == -R_frac[3 * s] &&
          R_frac[3 * t + 1] == -R_frac[3 * s + 1] && R_frac[3 * t + 2] == -R_frac[3 * s + 2]) {
        partner[s] = t;
        partner[t] = s;
        break;
      }
  }
  std::vector<int64_t> ent;  // entries: unpaired R, or the first of a pair
  bool anyzero = false;
  for (int64_t s = 0; s < n_R; ++s) {
    if (is_zero[s]) {
      anyzero = true;
      continue;
    }
    if (partner[s] < 0 || s < partner[s]) ent.push_back(s);
  }
  c->Kh = (int)ent.size();
  c->has_zero = anyzero ? 1 : 0;
  c->K = 2 * (int64_t)c->Kh + c->has_zero;

  // X_s: complex coefficients of R_s, columns [H | i c_a H | r_a |
  // i (c_a r_b - c_b r_a)], element (i, j) at p(i, j)
  const zd* H = reinterpret_cast<const zd*>(H_real);
  const zd* r = reinterpret_cast<const zd*>(r_real);
  auto build_x = [&](int64_t s, std::vector<zd>& Xs) {
    double cv[3];
    for (int a = 0; a < 3; ++a)
      cv[a] = c->cmat[3 * a] * R_frac[3 * s] + c->cmat[3 * a + 1] * R_frac[3 * s + 1] +
              c->cmat[3 * a + 2] * R_frac[3 * s + 2];
    double sc = (deg && (flags & MEP_F_APPLY_DEG)) ? 1.0 / (double)deg[s] : 1.0;
    for (int64_t o = 0; o < n2; ++o) {
      const int64_t i = o % n, j = o / n;
      if (c->packed && i > j) continue;
      const int64_t p = c->packed ? i + j * (j + 1) / 2 : o;
      zd h = zscale(H[o + n2 * s], sc);
      Xs[p] = h;
      for (int a = 0; a < 3; ++a) Xs[(1 + a) * c->np + p] = zmuli(zscale(h, cv[a]));
      if (c->pos) {
        const zd* rr = &r[3 * (o + n2 * s)];
        zd ra[3] = {zscale(rr[0], sc), zscale(rr[1], sc), zscale(rr[2], sc)};
        for (int a = 0; a < 3; ++a) Xs[(4 + a) * c->np + p] = ra[a];
        const int AA[3] = {1, 2, 0}, BB[3] = {2, 0, 1};
        for (int q = 0; q < 3; ++q) {
          int a = AA[q], b = BB[q];
          Xs[(7 + q) * c->np + p] = zmuli(zsub(zscale(ra[b], cv[a]), zscale(ra[a], cv[b])));
        }
      }
    }
  };
  // real rows: cos row e: X_s + X_t, sin row Kh + e: i (X_s - X_t), ones row
  // (R = 0): sum of the R = 0 coefficients; (re, im) interleaved columns
  const int64_t NC = c->ncols;
  std::vector<double> X((size_t)(c->K * 2 * NC), 0.0);
  std::vector<double> Rh((size_t)(3 * (c->Kh > 0 ? c->Kh : 1)), 0.0);
  std::vector<zd> xs(NC), xt(NC);
  for (int e = 0; e < c->Kh; ++e) {
    const int64_t s = ent[e], t = partner[s];
    for (int a = 0; a < 3; ++a) Rh[3 * e + a] = R_frac[3 * s + a];
    std::fill(xs.begin(), xs.end(), zd{0.0, 0.0});
    std::fill(xt.begin(), xt.end(), zd{0.0, 0.0});
    build_x(s, xs);
    if (t >= 0) build_x(t, xt);
    double* rc = &X[(size_t)(e * 2 * NC)];
    double* rs = &X[(size_t)((c->Kh + e) * 2 * NC)];
    for (int64_t q = 0; q < NC; ++q) {
      const zd sum = zadd(xs[q], xt[q]), dif = zmuli(zsub(xs[q], xt[q]));
      rc[2 * q] = sum.re;
      rc[2 * q + 1] = sum.im;
      rs[2 * q] = dif.re;
      rs[2 * q + 1] = dif.im;
    }
  }
  if (c->has_zero) {
    double* r0 = &X[(size_t)(2 * c->Kh * 2 * NC)];
    for (int64_t s = 0; s < n_R; ++s) {
      if (!is_zero[s]) continue;
      std::fill(xs.begin(), xs.end(), zd{0.0, 0.0});
      build_x(s, xs);
      for (int64_t q = 0; q < NC; ++q) {
        r0[2 * q] += xs[q].re;
        r0[2 * q + 1] += xs[q].im;
      }
    }
  }
  auto bad = [&](hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    mep_ctx_destroy(c);
    return MEP_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipStreamCreate(&c->st)) != hipSuccess) return bad(e, "hipStreamCreate");
  if ((e = hipMalloc(&c->dR, sizeof(double) * Rh.size())) != hipSuccess) return bad(e, "hipMalloc R");
  if ((e = hipMemcpy(c->dR, Rh.data(), sizeof(double) * Rh.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return bad(e, "upload R");
  if (c->tb) {
    if ((e = hipMalloc(&c->dtau, sizeof(double) * 3 * n)) != hipSuccess) return bad(e, "hipMalloc tau");
    if ((e = hipMemcpy(c->dtau, wf_centers, sizeof(double) * 3 * n, hipMemcpyHostToDevice)) != hipSuccess)
      return bad(e, "upload tau");
  }
  if ((e = hipMalloc(&c->dX, sizeof(double) * X.size())) != hipSuccess) return bad(e, "hipMalloc X");
  if ((e = hipMemcpy(c->dX, X.data(), sizeof(double) * X.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return bad(e, "upload X");
  if (!c->pos) {
    std::vector<float> X32(X.size());
    for (size_t q = 0; q < X.size(); ++q) X32[q] = (float)X[q];
    if ((e = hipMalloc(&c->dX32, sizeof(float) * X32.size())) != hipSuccess)
      return bad(e, "hipMalloc X32");
    if ((e = hipMemcpy(c->dX32, X32.data(), sizeof(float) * X32.size(), hipMemcpyHostToDevice)) !=
        hipSuccess)
      return bad(e, "upload X32");
  }
  if ((e = hipMalloc(&c->dinfo, sizeof(int))) != hipSuccess) return bad(e, "hipMalloc info");
  if ((e = hipMemset(c->dinfo, 0, sizeof(int))) != hipSuccess) return bad(e, "memset info");
  // coefficients in a box of integer R for the separable mesh sums, when
  // the box is compact (no more than 4 NR + 64 points, at most 16 along x)
  {
    bool integral = n_R > 0;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int64_t s2 = 0; s2 < n_R && integral; ++s2)
      for (int a = 0; a < 3; ++a) {
        const double v = R_frac[3 * s2 + a];
        if (v != std::floor(v) || std::fabs(v) > 1.0e6) integral = false;
        if (s2 == 0 || v < lo[a]) lo[a] = v;
        if (s2 == 0 || v > hi[a]) hi[a] = v;
      }
    if (integral) {
      const int bx = (int)(hi[0] - lo[0]) + 1, by = (int)(hi[1] - lo[1]) + 1,
                bz = (int)(hi[2] - lo[2]) + 1;
      const int64_t box = (int64_t)bx * by * bz;
      if (bx <= 16 && box <= 4 * n_R + 64) {
        std::vector<zd> Xb((size_t)(box * NC), zd{0.0, 0.0});
        for (int64_t s2 = 0; s2 < n_R; ++s2) {
          std::fill(xs.begin(), xs.end(), zd{0.0, 0.0});
          build_x(s2, xs);
          const int ix = (int)(R_frac[3 * s2] - lo[0]), iy = (int)(R_frac[3 * s2 + 1] - lo[1]),
                    iz = (int)(R_frac[3 * s2 + 2] - lo[2]);
          zd* dst = &Xb[(size_t)((((int64_t)ix * by + iy) * bz + iz) * NC)];
          for (int64_t q = 0; q < NC; ++q) dst[q] = zadd(dst[q], xs[q]);
        }
        if ((e = hipMalloc(&c->dXb, sizeof(zd) * Xb.size())) != hipSuccess)
          return bad(e, "hipMalloc box coefficients");
        if ((e = hipMemcpy(c->dXb, Xb.data(), sizeof(zd) * Xb.size(), hipMemcpyHostToDevice)) !=
            hipSuccess)
          return bad(e, "upload box coefficients");
        c->sep = true;
        c->sox = (int)lo[0];
        c->soy = (int)lo[1];
        c->soz = (int)lo[2];
        c->sbx = bx;
        c->sby = by;
        c->sbz = bz;
      }
    }
  }
  *out = c;
  return MEP_OK;
}

static void bg_drop_ctx_locked(mep_ctx* c);

extern "C" int mep_ctx_destroy(mep_ctx* c) {
  if (!c) return MEP_OK;
  bg_drop_ctx_locked(c);
  if (c->st) hipStreamSynchronize(c->st);
  resolve_times(c);
  free_ws(c);
  hipFree(c->dR); hipFree(c->dtau); hipFree(c->dX); hipFree(c->dX32); hipFree(c->dinfo);
  hipFree(c->dXb); hipFree(c->dPh); hipFree(c->dY1);
  hipFree(c->dpart); hipFree(c->dacc); hipFree(c->def); hipFree(c->dhw); hipFree(c->dfd);
  hipFree(c->dcurv);
  for (auto e : c->pool) hipEventDestroy(e);
  if (c->st) hipStreamDestroy(c->st);
  delete c;
  return MEP_OK;
}

extern "C" int mep_set_kblock(mep_ctx* c, int64_t kb) {
  if (!c || kb < 0) return fail(MEP_ERR_ARG, "mep_set_kblock: bad argument");
  c->kblock_req = kb;
  if (kb > 0 && kb < c->kcap) free_ws(c);
  return MEP_OK;
}

extern "C" int mep_set_variant(mep_ctx* c, int32_t what, int32_t value) {
  if (!c) return fail(MEP_ERR_ARG, "mep_set_variant: null context");
  if (what == MEP_VAR_EIGH) {
    if (value < 0 || value > 3) return fail(MEP_ERR_ARG, "mep_set_variant: MEP_VAR_EIGH takes 0 ... 3");
    if (value != c->var_eigh) ++c->gen;
    c->var_eigh = value;
    return MEP_OK;
  }
  if (what == MEP_VAR_GEMM32) {
    if (value < 0 || value > 2) return fail(MEP_ERR_ARG, "mep_set_variant: MEP_VAR_GEMM32 takes 0, 1 or 2");
    if (value != c->var_gemm32) ++c->gen;
    c->var_gemm32 = value;
    return MEP_OK;
  }
  return fail(MEP_ERR_ARG, "mep_set_variant: unknown stage");
}

extern "C" int mep_set_workspace_limit(mep_ctx* c, double bytes) {
  if (!c || !(bytes > 0.0)) return fail(MEP_ERR_ARG, "mep_set_workspace_limit: bad argument");
  c->ws_limit = bytes;
  if ((double)c->kcap * bytes_per_k(c) > bytes) free_ws(c);
  return MEP_OK;
}

extern "C" int mep_set_profiling(mep_ctx* c, int32_t on) {
  if (!c) return fail(MEP_ERR_ARG, "mep_set_profiling: ctx is NULL");
  c->prof = on != 0;
  return MEP_OK;
}

extern "C" int mep_reset_kernel_times(mep_ctx* c) {
  if (!c) return fail(MEP_ERR_ARG, "mep_reset_kernel_times: ctx is NULL");
  hipStreamSynchronize(c->st);
  resolve_times(c);
  for (int i = 0; i < K_NKIND; ++i) {
    c->tms[i] = 0;
    c->tcnt[i] = 0;
  }
  return MEP_OK;
}

extern "C" int mep_get_kernel_time(mep_ctx* c, const char* name, double* ms, int64_t* cnt) {
  if (!c || !name) return fail(MEP_ERR_ARG, "mep_get_kernel_time: bad argument");
  HIPCHK(hipStreamSynchronize(c->st));
  resolve_times(c);
  for (int i = 0; i < K_NKIND; ++i)
    if (strcmp(name, kKindName[i]) == 0) {
      if (ms) *ms = c->tms[i];
      if (cnt) *cnt = c->tcnt[i];
      return MEP_OK;
    }
  return fail(MEP_ERR_ARG, std::string("mep_get_kernel_time: unknown kernel ") + name);
}

// ---------------------------------------------------------------------------
// pipeline pieces
// ---------------------------------------------------------------------------
static KGeom geom(const mep_ctx* c) {
  KGeom g;
  g.kpts = c->dk;
  g.tau = c->tb ? c->dtau : nullptr;
  for (int i = 0; i < 9; ++i) g.cmat[i] = c->cmat[i];
  return g;
}

// near-degenerate threshold of the mixed-precision refinement: 1e-3 eV
static const double kMixedGroupHartree = 1e-3 / 27.211385;

// phase tables of the mesh mp and the z-stage workspace for the context's
// k-blocks (separable sums); no-op without a box
static int sep_prepare(mep_ctx* c, const int32_t* mp) {
  if (!c->sep) return MEP_OK;
  const int BX = c->sbx <= 8 ? 8 : 16;
  if (mp[0] != c->sep_mesh[0] || mp[1] != c->sep_mesh[1] || mp[2] != c->sep_mesh[2]) {
    const int64_t len = (int64_t)mp[0] * BX + (int64_t)mp[1] * c->sby + (int64_t)mp[2] * c->sbz;
    std::vector<zd> ph((size_t)len, zd{0.0, 0.0});
    const double two_pi = 2.0 * 3.14159265358979323846;
    // k_i = (2 (q + 1) - N - 1) / (2 N), as mp_kpts_kernel
    auto fill = [&](zd* dst, int nq, int stride, int nb, int origin) {
      for (int q = 0; q < nq; ++q) {
        const double kq = (2.0 * (double)(q + 1) - (double)nq - 1.0) / (2.0 * (double)nq);
        for (int i = 0; i < nb; ++i) {
          const double ang = two_pi * (kq * (double)(origin + i));
          dst[(size_t)q * stride + i] = zd{std::cos(ang), std::sin(ang)};
        }
      }
    };
    fill(ph.data(), mp[0], BX, c->sbx, c->sox);
    fill(ph.data() + (size_t)mp[0] * BX, mp[1], c->sby, c->sby, c->soy);
    fill(ph.data() + (size_t)mp[0] * BX + (size_t)mp[1] * c->sby, mp[2], c->sbz, c->sbz, c->soz);
    HIPCHK(hipStreamSynchronize(c->st));  // the previous mesh's tables may still be read
    if (len > c->ph_cap) {
      ++c->gen;
      hipFree(c->dPh);
      c->dPh = nullptr;
      c->ph_cap = 0;
      HIPCHK(hipMalloc(&c->dPh, sizeof(zd) * len));
      c->ph_cap = len;
    }
    HIPCHK(hipMemcpy(c->dPh, ph.data(), sizeof(zd) * len, hipMemcpyHostToDevice));
    for (int a = 0; a < 3; ++a) c->sep_mesh[a] = mp[a];
  }
  const int64_t plane = (int64_t)mp[0] * mp[1];
  int64_t nz = (c->kcap + plane - 1) / plane + 1;
  if (nz > mp[2]) nz = mp[2];
  const int64_t need = nz * c->sbx * c->sby * c->ncols;
  if (need > c->y1_cap) {
    HIPCHK(hipStreamSynchronize(c->st));
    ++c->gen;
    hipFree(c->dY1);
    c->dY1 = nullptr;
    c->y1_cap = 0;
    HIPCHK(hipMalloc(&c->dY1, sizeof(zd) * need));
    c->y1_cap = need;
  }
  return MEP_OK;
}

static SepFT sep_args(mep_ctx* c) {
  const int BX = c->sbx <= 8 ? 8 : 16;
  SepFT f;
  f.NC = c->ncols;
  f.bx = c->sbx;
  f.by = c->sby;
  f.bz = c->sbz;
  f.n1 = c->sep_mesh[0];
  f.n2 = c->sep_mesh[1];
  f.n3 = c->sep_mesh[2];
  f.Xb = c->dXb;
  f.ex = c->dPh;
  f.ey = c->dPh + (int64_t)f.n1 * BX;
  f.ez = f.ey + (int64_t)f.n2 * f.by;
  f.Y1 = c->dY1;
  return f;
}

// mp (optional): the k-block is k0 .. k0+nb-1 of that Monkhorst-Pack mesh,
// prepared by sep_prepare -> separable Fourier sums instead of the GEMM
static int interp_block(mep_ctx* c, int64_t nb, bool velo, int prec = MEP_PREC_FP64,
                        const int32_t* mp = nullptr, int64_t k0 = 0) {
  const bool sep = mp && c->sep && mp[0] == c->sep_mesh[0] && mp[1] == c->sep_mesh[1] &&
                   mp[2] == c->sep_mesh[2];
  const int64_t n = c->n;
  KGeom g = geom(c);
  VeloArgs va;
  va.n = (int)n;
  va.nmat = c->nmat;
  va.packed = c->packed;
  va.gauge_trafo = (c->flags & MEP_F_GAUGE_TRAFO) ? 1 : 0;
  va.pos_op = c->pos ? 1 : 0;
  va.textbook = (c->flags & MEP_F_TEXTBOOK) ? 1 : 0;
  va.kubo_tol = c->kubo_tol;
  va.scratch = reinterpret_cast<zd*>(c->dscr);  // free once the eigensolve is done
  va.scratch_bytes = eigh_scratch_bytes((int)n, c->kcap);
  if (velo && c->pos && !velo_pos_scratch_ok((int)n, va.scratch_bytes))
    return fail(MEP_ERR_ARG, "position operator: the k-block is too small for the velocity slots");
  if (prec == MEP_PREC_MIXED) {
    // FP32 phase tile, FP32 GEMM, FP32 eigensolve (U in V's slot), FP64
    // refinement (U -> dU), FP64 velocities from the FP32 dH
    float* P32 = reinterpret_cast<float*>(c->dP);
    zf* H32 = reinterpret_cast<zf*>(c->dHk);
    zf* U32 = reinterpret_cast<zf*>(c->dV);
    if (sep) {
      timed(c, K_GEMM, [&] { launch_sep_ft(c->st, sep_args(c), k0, nb, nullptr, H32); });
    } else {
      timed(c, K_PHASE,
            [&] { launch_phase_f32(c->st, nb, c->Kh, c->has_zero, c->dk, c->dR, P32); });
      timed(c, K_GEMM, [&] {
        if (c->var_gemm32 == 1)
          launch_ft_gemm_f32(c->st, (int)nb, (int)(2 * c->ncols), (int)c->K, P32, c->dX32,
                             reinterpret_cast<float*>(H32));
        else  // default: split bfloat16, FP32-equivalent
          launch_ft_gemm_bf16x3(c->st, (int)nb, (int)(2 * c->ncols), (int)c->K, P32, c->dX32,
                                reinterpret_cast<float*>(H32));
      });
    }
    timed(c, K_EIGH, [&] {
      launch_eigh_f32(c->st, (int)n, nb, H32, c->ncols, c->packed, g, c->dE, U32, c->dinfo,
                      c->dscr);
    });
    timed(c, K_REFINE, [&] {
      launch_refine(c->st, (int)n, nb, H32, c->ncols, c->packed, g, kMixedGroupHartree, U32, c->dE,
                    c->dU);
    });
    if (velo)
      timed(c, K_VELO, [&] { launch_velo_h32(c->st, va, nb, H32, c->ncols, g, c->dU, c->dV); });
    HIPCHK(hipGetLastError());
    return MEP_OK;
  }
  if (sep) {
    timed(c, K_GEMM, [&] { launch_sep_ft(c->st, sep_args(c), k0, nb, c->dHk, nullptr); });
  } else {
    timed(c, K_PHASE, [&] { launch_phase(c->st, nb, c->Kh, c->has_zero, c->dk, c->dR, c->dP); });
    timed(c, K_GEMM, [&] {
      launch_ft_gemm_f64(c->st, (int)nb, (int)(2 * c->ncols), (int)c->K, c->dP, c->dX,
                         reinterpret_cast<double*>(c->dHk));
    });
  }
  timed(c, K_EIGH, [&] {
    if (n > 32)
      launch_eigh_wide(c->st, (int)n, nb, c->dHk, c->ncols, c->packed, g, c->dE, c->dU, c->dinfo,
                       c->dscr, c->var_eigh);
    else
      launch_eigh(c->st, (int)n, nb, c->dHk, c->ncols, c->packed, g, c->dE, c->dU, c->dinfo,
                  c->dscr, c->var_eigh);
  });
  if (velo)
    timed(c, K_VELO, [&] {
      launch_velo(c->st, va, nb, c->dHk, c->ncols, g, c->dU, c->dE, c->dV, c->dA, c->dOm);
    });
  HIPCHK(hipGetLastError());
  return MEP_OK;
}

static int check_info(mep_ctx* c) {
  int info = 0;
  HIPCHK(hipMemcpyAsync(&info, c->dinfo, sizeof(int), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (info) {
    hipMemsetAsync(c->dinfo, 0, sizeof(int), c->st);
    hipStreamSynchronize(c->st);
    return fail(MEP_ERR_EIG, "eigensolver did not converge (zheevd info != 0 equivalent)");
  }
  return MEP_OK;
}

static void linspace(double lo, double hi, int n, std::vector<double>& out) {
  // get_linspace (matrix_math.f90:74-100)
  out.assign(n > 1 ? n : 1, lo);
  if (n > 1) {
    double d = (hi - lo) / (double)(n - 1);
    for (int i = 0; i < n; ++i) out[i] = lo + (double)i * d;
  }
}

struct BandSetup {
  BandArgs b;
  AccLayout L;
  ExtLayout X;  // optical / photogalvanic / gyrotropic records (when asked)
  bool ext = false;
  int nblocks;
};

// validation and the record layouts of one run (no device work)
static int band_plan(mep_ctx* c, const mep_run_params* p, int64_t kblock, BandSetup& bs) {
  if (!p) return fail(MEP_ERR_ARG, "run params are NULL");
  if (p->n_eF < 1 || p->n_hw < 1) return fail(MEP_ERR_ARG, "n_eF and n_hw must be >= 1");
  if ((p->jobs & MEP_JOB_MEP) && (p->valence_bands < 1 || p->valence_bands > c->n))
    return fail(MEP_ERR_ARG, "valence_bands must be in [1, n_wf] for the Niu MEP tensor");
  if (p->precision != MEP_PREC_FP64 && p->precision != MEP_PREC_MIXED)
    return fail(MEP_ERR_ARG, "unknown precision");
  if (p->precision == MEP_PREC_MIXED && c->pos)
    return fail(MEP_ERR_ARG, "the mixed-precision path does not take a position operator");
  if (p->precision == MEP_PREC_MIXED && c->n > 32)
    return fail(MEP_ERR_ARG, "the mixed-precision path supports n_wf <= 32");
  if (p->hw_lst)
    for (int32_t i = 0; i < p->n_hw; ++i)
      if (!std::isfinite(p->hw_lst[i]))
        return fail(MEP_ERR_ARG, "hw_lst holds a non-finite frequency");
  BandArgs& b = bs.b;
  b.n = (int)c->n;
  b.jobs = p->jobs;
  b.n_ef = p->n_eF;
  b.n_hw = p->n_hw;
  b.n_val = p->valence_bands > 0 ? p->valence_bands : 1;
  b.T_kelvin = p->T_kelvin;
  b.eta = p->eta;
  b.kubo_tol = c->kubo_tol;
  b.pos_op = c->pos ? 1 : 0;
  b.v_herm = c->pos ? 0 : 1;  // the library builds V itself
  b.ef = nullptr;
  b.hw = nullptr;
  b.fd_in = nullptr;
  if (bandsum_lds_bytes(b) > 160 * 1024)
    return fail(MEP_ERR_ARG, "band-sum tables exceed the 160 KiB LDS of one workgroup "
                             "(reduce N_eF)");
  bs.L = acc_layout(b);
  bs.ext = (p->jobs & (MEP_JOB_OPT | MEP_JOB_GYRO)) != 0;
  bs.X = ext_layout(b);
  if (bs.ext && ext_lds_bytes(b) > 160 * 1024)
    return fail(MEP_ERR_ARG, "optical / gyrotropic sums: n_eF x n_wf too large for the "
                             "occupation tables in LDS (about 9,600 above 32 bands)");
  bs.nblocks = (int)(kblock < 1024 ? kblock : 1024);
  return MEP_OK;
}

static int64_t band_record_len(const BandSetup& bs) { return bs.L.total + (bs.ext ? bs.X.total : 0); }

static int band_setup(mep_ctx* c, const mep_run_params* p, int64_t kblock, BandSetup& bs) {
  CHK(band_plan(c, p, kblock, bs));
  BandArgs& b = bs.b;
  std::vector<double> ef, hw;
  linspace(p->eF_min, p->eF_max, p->n_eF, ef);
  if (p->hw_lst) {
    // the caller's frequency list, any spacing (velo_ahc_tens / kubo_ohc_tens
    // take hw_lst as an argument, kubo.f90:86-104, 145-164)
    hw.assign(p->hw_lst, p->hw_lst + p->n_hw);
  } else {
    linspace(p->hw_min, p->hw_max, p->n_hw, hw);
  }
  if ((int64_t)ef.size() > c->ef_cap) {
    ++c->gen;
    hipFree(c->def);
    HIPCHK(hipMalloc(&c->def, sizeof(double) * ef.size()));
    c->ef_cap = ef.size();
  }
  if ((int64_t)hw.size() > c->hw_cap) {
    ++c->gen;
    hipFree(c->dhw);
    HIPCHK(hipMalloc(&c->dhw, sizeof(double) * hw.size()));
    c->hw_cap = hw.size();
  }
  HIPCHK(hipMemcpy(c->def, ef.data(), sizeof(double) * ef.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->dhw, hw.data(), sizeof(double) * hw.size(), hipMemcpyHostToDevice));
  b.ef = c->def;
  b.hw = c->dhw;
  const int64_t rec = band_record_len(bs);
  int64_t need = (int64_t)bs.nblocks * rec;
  if (need > c->part_cap) {
    ++c->gen;
    hipFree(c->dpart);
    HIPCHK(hipMalloc(&c->dpart, sizeof(double) * need));
    c->part_cap = need;
  }
  if (rec > c->acc_cap) {
    ++c->gen;
    hipFree(c->dacc);
    HIPCHK(hipMalloc(&c->dacc, sizeof(double) * rec));
    c->acc_cap = rec;
  }
  launch_fill_zero(c->st, c->dacc, rec);
  launch_fill_zero(c->st, c->dpart, need);
  HIPCHK(hipGetLastError());
  return MEP_OK;
}

// fuse: the band-sum kernel forms V itself from H(k)/dH(k) and U (the
// velocity kernel is not launched and V never reaches HBM)
static int band_block(mep_ctx* c, BandSetup& bs, int64_t nb, const double* w, bool fuse = false) {
  int nblk = (int)(nb < bs.nblocks ? nb : bs.nblocks);
  VeloFuse vf;
  vf.Hk = c->dHk;
  vf.strideHk = c->ncols;
  vf.U = c->dU;
  timed(c, K_BANDSUM, [&] {
    // n > 32: the eigensolver records (64 n^2 bytes or more per k-point of
    // the block, consumed by now) hold the 48 n^2 bytes per workgroup of
    // the wide kernel's scaled columns; nblk <= nb workgroups
    launch_bandsum(c->st, bs.b, nb, c->dE, c->dV, c->dA, c->dOm, w, nblk, c->dpart,
                   fuse ? &vf : nullptr, reinterpret_cast<zd*>(c->dscr));
    if (bs.ext)
      launch_extsum(c->st, bs.b, nb, c->dE, c->dV, c->dA, c->dOm, w, nblk,
                    c->dpart + (int64_t)bs.nblocks * bs.L.total,
                    c->n > 32 ? reinterpret_cast<zd*>(c->dscr) : nullptr);
  });
  HIPCHK(hipGetLastError());
  return MEP_OK;
}

// every output member of mep_sums with its offset and length in the
// reduced record: f(dst, offset, doubles)
template <class F>
static void for_each_output(const BandSetup& bs, const mep_run_params* p, mep_sums* out, F&& f) {
  const int64_t nef = p->n_eF, nhw = p->n_hw, nv = bs.b.n_val;
  f(out->Ne, bs.L.Ne, nef);
  f(out->ahc, bs.L.ahc, 9 * nef);
  f(out->ahc_velo, bs.L.velo, 18 * nhw * nef);
  f(out->ohc, bs.L.ohc, 18 * nhw * nef);
  f(out->kubo_ic, bs.L.kic, 9 * nef);
  f(out->kubo_lc, bs.L.klc, 9 * nef);
  f(out->kubo_cs, bs.L.kcs, 9 * nef);
  if (p->jobs & MEP_JOB_MEP) {
    f(out->niu_ic, bs.L.nic, 9 * nv);
    f(out->niu_lc, bs.L.nlc, 9 * nv);
    f(out->niu_cs, bs.L.ncs, 9 * nv);
  }
  if (p->jobs & MEP_JOB_OPT) {
    f(out->opt_s, bs.L.total + bs.X.opt_s, 18 * nhw * nef);
    f(out->opt_a, bs.L.total + bs.X.opt_a, 18 * nhw * nef);
    f(out->photo, bs.L.total + bs.X.photo, 54 * nhw * nef);
  }
  if (p->jobs & MEP_JOB_GYRO) {
    f(out->gyro_C, bs.L.total + bs.X.gyro_C, 18 * nef);
    f(out->gyro_D, bs.L.total + bs.X.gyro_D, 18 * nef);
    f(out->gyro_Dw, bs.L.total + bs.X.gyro_Dw, 18 * nhw * nef);
  }
}

static void reduce_records(mep_ctx* c, const BandSetup& bs) {
  // the partial records accumulate over every k-block of the run
  timed(c, K_REDUCE, [&] {
    launch_reduce_parts(c->st, bs.nblocks, bs.L.total, c->dpart, c->dacc);
    if (bs.ext)
      launch_reduce_parts(c->st, bs.nblocks, bs.X.total, c->dpart + (int64_t)bs.nblocks * bs.L.total,
                          c->dacc + bs.L.total);
  });
}

static int unpack(mep_ctx* c, const BandSetup& bs, const mep_run_params* p, int64_t nk,
                  mep_sums* out) {
  reduce_records(c, bs);
  HIPCHK(hipGetLastError());
  const int64_t rec = band_record_len(bs);
  std::vector<double> h(rec);
  HIPCHK(hipMemcpyAsync(h.data(), c->dacc, sizeof(double) * rec, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  resolve_times(c);
  if (!out) return MEP_OK;
  for_each_output(bs, p, out, [&](double* dst, int64_t off, int64_t len) {
    if (dst) memcpy(dst, h.data() + off, sizeof(double) * len);
  });
  out->n_k = nk;
  return MEP_OK;
}

// ---------------------------------------------------------------------------
// entry points
// ---------------------------------------------------------------------------
extern "C" int mep_interp_batch(mep_ctx* c, int64_t nk, const double* kpts, double* e_k,
                                double* V_ka, double* A_ka, double* Om_kab) {
  if (!c || nk < 0 || (nk > 0 && (!kpts || !e_k)))
    return fail(MEP_ERR_ARG, "mep_interp_batch: bad argument");
  if (nk == 0) return MEP_OK;
  if (c->pos && V_ka && (!A_ka || !Om_kab))
    return fail(MEP_ERR_ARG, "mep_interp_batch: a position operator needs A_ka and Om_kab");
  CHK(ensure_ws(c, nk));
  const int64_t n = c->n, n2 = n * n;
  for (int64_t k0 = 0; k0 < nk; k0 += c->kcap) {
    int64_t nb = (nk - k0 < c->kcap) ? nk - k0 : c->kcap;
    HIPCHK(hipMemcpyAsync(c->dk, kpts + 3 * k0, sizeof(double) * 3 * nb, hipMemcpyHostToDevice,
                          c->st));
    CHK(interp_block(c, nb, V_ka != nullptr));
    HIPCHK(hipMemcpyAsync(e_k + n * k0, c->dE, sizeof(double) * n * nb, hipMemcpyDeviceToHost,
                          c->st));
    if (V_ka) {
      // planar -> reference layout through the (now free) Fourier buffer
      launch_v_layout(c->st, (int)n, nb, c->dV, c->dHk, true);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(V_ka + 2 * 3 * n2 * k0, c->dHk, sizeof(zd) * 3 * n2 * nb,
                            hipMemcpyDeviceToHost, c->st));
    }
    if (V_ka && c->pos) {
      HIPCHK(hipMemcpyAsync(A_ka + 2 * 3 * n2 * k0, c->dA, sizeof(zd) * 3 * n2 * nb,
                            hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipMemcpyAsync(Om_kab + 2 * 9 * n2 * k0, c->dOm, sizeof(zd) * 9 * n2 * nb,
                            hipMemcpyDeviceToHost, c->st));
    }
    CHK(check_info(c));
  }
  HIPCHK(hipStreamSynchronize(c->st));
  resolve_times(c);
  return MEP_OK;
}

extern "C" int mep_curv_batch(mep_ctx* c, int64_t nk, const double* kpts, double ef,
                              double T_kelvin, double* omega) {
  if (!c || nk < 0 || (nk > 0 && (!kpts || !omega)))
    return fail(MEP_ERR_ARG, "mep_curv_batch: bad argument");
  if (nk == 0) return MEP_OK;
  CHK(ensure_ws(c, nk));
  if (9 * c->kcap > c->curv_cap) {
    hipFree(c->dcurv);
    c->dcurv = nullptr;
    HIPCHK(hipMalloc(&c->dcurv, sizeof(double) * 9 * c->kcap));
    c->curv_cap = 9 * c->kcap;
  }
  for (int64_t k0 = 0; k0 < nk; k0 += c->kcap) {
    int64_t nb = (nk - k0 < c->kcap) ? nk - k0 : c->kcap;
    HIPCHK(hipMemcpyAsync(c->dk, kpts + 3 * k0, sizeof(double) * 3 * nb, hipMemcpyHostToDevice,
                          c->st));
    CHK(interp_block(c, nb, true));
    timed(c, K_BANDSUM, [&] {
      launch_curv(c->st, (int)c->n, nb, c->dE, c->dV, ef, T_kelvin, c->dcurv);
    });
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(omega + 9 * k0, c->dcurv, sizeof(double) * 9 * nb,
                          hipMemcpyDeviceToHost, c->st));
    CHK(check_info(c));
  }
  HIPCHK(hipStreamSynchronize(c->st));
  resolve_times(c);
  return MEP_OK;
}

extern "C" int mep_bandsum_batch(mep_ctx* c, int64_t nk, const double* e_k, const double* V_ka,
                                 const double* A_ka, const double* Om_kab,
                                 const double* fd_distrib, const mep_run_params* p,
                                 mep_sums* out) {
  if (!c || nk < 0 || (nk > 0 && (!e_k || !V_ka)))
    return fail(MEP_ERR_ARG, "mep_bandsum_batch: bad argument");
  if (c->pos && (!A_ka || !Om_kab))
    return fail(MEP_ERR_ARG, "mep_bandsum_batch: a position operator needs A_ka and Om_kab");
  CHK(ensure_ws(c, nk > 0 ? nk : 1));
  BandSetup bs;
  CHK(band_setup(c, p, c->kcap, bs));
  bs.b.v_herm = 0;  // V comes from the caller: no symmetry assumed
  const int64_t n = c->n, n2 = n * n, nef = p->n_eF;
  if (fd_distrib && nef * n * c->kcap > c->fd_cap) {
    hipFree(c->dfd);
    c->dfd = nullptr;
    HIPCHK(hipMalloc(&c->dfd, sizeof(double) * nef * n * c->kcap));
    c->fd_cap = nef * n * c->kcap;
  }
  if (fd_distrib) bs.b.fd_in = c->dfd;
  for (int64_t k0 = 0; k0 < nk; k0 += c->kcap) {
    int64_t nb = (nk - k0 < c->kcap) ? nk - k0 : c->kcap;
    if (fd_distrib)
      HIPCHK(hipMemcpyAsync(c->dfd, fd_distrib + nef * n * k0, sizeof(double) * nef * n * nb,
                            hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->dE, e_k + n * k0, sizeof(double) * n * nb, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->dHk, V_ka + 2 * 3 * n2 * k0, sizeof(zd) * 3 * n2 * nb,
                          hipMemcpyHostToDevice, c->st));
    launch_v_layout(c->st, (int)n, nb, c->dHk, c->dV, false);
    HIPCHK(hipGetLastError());
    if (c->pos) {
      HIPCHK(hipMemcpyAsync(c->dA, A_ka + 2 * 3 * n2 * k0, sizeof(zd) * 3 * n2 * nb,
                            hipMemcpyHostToDevice, c->st));
      HIPCHK(hipMemcpyAsync(c->dOm, Om_kab + 2 * 9 * n2 * k0, sizeof(zd) * 9 * n2 * nb,
                            hipMemcpyHostToDevice, c->st));
    }
    CHK(band_block(c, bs, nb, nullptr));
  }
  return unpack(c, bs, p, nk, out);
}

// one k-block of a run: k-points (generated or uploaded), interpolation and
// band sums, all on the context's stream
static int run_block(mep_ctx* c, BandSetup& bs, int64_t k0, int64_t nb, const int32_t* mp,
                     const double* kpts, const double* weights, const mep_run_params* p) {
  if (mp) {
    int m3[3] = {mp[0], mp[1], mp[2]};
    timed(c, K_KGEN, [&] { launch_mp_kpts(c->st, k0, nb, m3, c->dk); });
  } else {
    HIPCHK(hipMemcpyAsync(c->dk, kpts + 3 * k0, sizeof(double) * 3 * nb, hipMemcpyHostToDevice,
                          c->st));
  }
  if (weights)
    HIPCHK(hipMemcpyAsync(c->dw, weights + k0, sizeof(double) * nb, hipMemcpyHostToDevice, c->st));
  // velocities inside the band-sum kernel where its conditions hold: FP64,
  // 17 <= n <= 32, V = U^H dH U (gauge transformation, no position
  // operator), Wannier convention, and no optical / gyrotropic sums (those
  // read V from HBM)
  const bool fuse = p->precision == MEP_PREC_FP64 && !bs.ext && c->packed && !c->tb &&
                    (c->flags & MEP_F_GAUGE_TRAFO) && !(c->flags & MEP_F_SEPARATE_VELO) &&
                    bandsum_can_fuse(bs.b);
  CHK(interp_block(c, nb, !fuse, p->precision, mp, k0));
  CHK(band_block(c, bs, nb, weights ? c->dw : nullptr, fuse));
  return MEP_OK;
}

static int run_common(mep_ctx* c, int64_t nk, const int32_t* mp, const double* kpts,
                      const double* weights, const mep_run_params* p, mep_sums* out) {
  CHK(ensure_ws(c, nk));
  if (mp) CHK(sep_prepare(c, mp));
  BandSetup bs;
  CHK(band_setup(c, p, c->kcap, bs));
  for (int64_t k0 = 0; k0 < nk; k0 += c->kcap) {
    int64_t nb = (nk - k0 < c->kcap) ? nk - k0 : c->kcap;
    CHK(run_block(c, bs, k0, nb, mp, kpts, weights, p));
  }
  CHK(check_info(c));
  return unpack(c, bs, p, nk, out);
}

extern "C" int mep_run_batch(int32_t n_models, mep_ctx* const* ctxs, const int32_t* mp_grids,
                             const mep_run_params* params, mep_sums* outs) {
  if (n_models < 1 || !ctxs || !mp_grids || !params)
    return fail(MEP_ERR_ARG, "mep_run_batch: bad argument");
  std::vector<BandSetup> bs((size_t)n_models);
  std::vector<int64_t> nk((size_t)n_models);
  int64_t rounds = 0;
  for (int32_t i = 0; i < n_models; ++i) {
    mep_ctx* c = ctxs[i];
    const int32_t* mp = mp_grids + 3 * i;
    if (!c) return fail(MEP_ERR_ARG, "mep_run_batch: a context is NULL");
    for (int32_t j = 0; j < i; ++j)
      if (ctxs[j] == c) return fail(MEP_ERR_ARG, "mep_run_batch: a context appears twice");
    if (mp[0] < 1 || mp[1] < 1 || mp[2] < 1)
      return fail(MEP_ERR_ARG, "mep_run_batch: mp_grid entries must be >= 1");
    nk[i] = (int64_t)mp[0] * mp[1] * mp[2];
    CHK(ensure_ws(c, nk[i]));
    CHK(sep_prepare(c, mp));
    CHK(band_setup(c, &params[i], c->kcap, bs[i]));
    const int64_t nblk = (nk[i] + c->kcap - 1) / c->kcap;
    if (nblk > rounds) rounds = nblk;
  }
  // block r of every model, model after model, each on its own stream
  for (int64_t r = 0; r < rounds; ++r)
    for (int32_t i = 0; i < n_models; ++i) {
      mep_ctx* c = ctxs[i];
      const int64_t k0 = r * c->kcap;
      if (k0 >= nk[i]) continue;
      const int64_t nb = (nk[i] - k0 < c->kcap) ? nk[i] - k0 : c->kcap;
      CHK(run_block(c, bs[i], k0, nb, mp_grids + 3 * i, nullptr, nullptr, &params[i]));
    }
  for (int32_t i = 0; i < n_models; ++i) {
    CHK(check_info(ctxs[i]));
    CHK(unpack(ctxs[i], bs[i], &params[i], nk[i], outs ? &outs[i] : nullptr));
  }
  return MEP_OK;
}

// ---------------------------------------------------------------------------
// mep_run_batch_graph: the kernels of a batch (accumulator zeroing, every
// k-block of every model on the models' own streams, the record reductions)
// captured once into a HIP graph and replayed by later calls with the same
// contexts, meshes and parameters.  The host set-up runs on every call (it
// re-uploads the E_F / frequency tables and the mesh phase tables); a
// reallocation or a variant change since the capture (the context's
// generation) or any difference in the parameters recaptures.
// ---------------------------------------------------------------------------
struct BatchGraph {
  std::vector<mep_ctx*> ctxs;
  std::vector<int32_t> grids;
  std::vector<mep_run_params> params;  // hw_lst compared through `hw`
  std::vector<std::vector<double>> hw;
  std::vector<uint64_t> gens;
  int device = 0;
  hipStream_t st = nullptr;
  hipGraphExec_t exec = nullptr;
};
static std::mutex g_bg_mu;
static std::vector<BatchGraph*> g_bg;

static void bg_free(BatchGraph* g) {
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->st) hipStreamDestroy(g->st);
  delete g;
}

// forget every graph that holds the context (called under g_bg_mu)
static void bg_drop_ctx(mep_ctx* c) {
  for (size_t i = 0; i < g_bg.size();) {
    BatchGraph* g = g_bg[i];
    if (std::find(g->ctxs.begin(), g->ctxs.end(), c) != g->ctxs.end()) {
      if (g->st) hipStreamSynchronize(g->st);
      bg_free(g);
      g_bg.erase(g_bg.begin() + (long)i);
    } else {
      ++i;
    }
  }
}

static void bg_drop_ctx_locked(mep_ctx* c) {
  std::lock_guard<std::mutex> lk(g_bg_mu);
  bg_drop_ctx(c);
}

static bool same_params(const mep_run_params& a, const mep_run_params& b) {
  return a.jobs == b.jobs && a.n_eF == b.n_eF && a.eF_min == b.eF_min && a.eF_max == b.eF_max &&
         a.n_hw == b.n_hw && a.hw_min == b.hw_min && a.hw_max == b.hw_max &&
         a.T_kelvin == b.T_kelvin && a.eta == b.eta && a.valence_bands == b.valence_bands &&
         a.precision == b.precision && (a.hw_lst == nullptr) == (b.hw_lst == nullptr);
}

static bool bg_matches(const BatchGraph* g, int32_t n_models, mep_ctx* const* ctxs,
                       const int32_t* mp_grids, const mep_run_params* params) {
  if ((int32_t)g->ctxs.size() != n_models) return false;
  for (int32_t i = 0; i < n_models; ++i) {
    if (g->ctxs[i] != ctxs[i] || g->gens[i] != ctxs[i]->gen) return false;
    for (int a = 0; a < 3; ++a)
      if (g->grids[3 * i + a] != mp_grids[3 * i + a]) return false;
    if (!same_params(g->params[i], params[i])) return false;
    if (params[i].hw_lst &&
        memcmp(g->hw[i].data(), params[i].hw_lst, sizeof(double) * (size_t)params[i].n_hw) != 0)
      return false;
  }
  return true;
}

// capture the batch into g->exec (streams of the contexts joined to g->st
// through events; profiling events are kept out of the capture)
static int bg_capture(BatchGraph* g, int32_t n_models, mep_ctx* const* ctxs,
                      const int32_t* mp_grids, const mep_run_params* params,
                      std::vector<BandSetup>& bs, const std::vector<int64_t>& nk, int64_t rounds) {
  std::vector<int> prof((size_t)n_models);
  for (int32_t i = 0; i < n_models; ++i) {
    prof[i] = ctxs[i]->prof;
    ctxs[i]->prof = 0;
    HIPCHK(hipStreamSynchronize(ctxs[i]->st));
  }
  std::vector<hipEvent_t> ev((size_t)n_models + 1, nullptr);
  for (auto& e : ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  int rc = MEP_OK;
  HIPCHK(hipStreamBeginCapture(g->st, hipStreamCaptureModeThreadLocal));
  {
    hipError_t e = hipEventRecord(ev[0], g->st);
    for (int32_t i = 0; i < n_models && e == hipSuccess; ++i) e = hipStreamWaitEvent(ctxs[i]->st, ev[0], 0);
    if (e != hipSuccess) rc = fail(MEP_ERR_HIP, std::string("mep_run_batch_graph: ") + hipGetErrorString(e));
    for (int32_t i = 0; i < n_models && rc == MEP_OK; ++i) {
      mep_ctx* c = ctxs[i];
      const int64_t rec = band_record_len(bs[i]);
      launch_fill_zero(c->st, c->dacc, rec);
      launch_fill_zero(c->st, c->dpart, (int64_t)bs[i].nblocks * rec);
    }
    for (int64_t r = 0; r < rounds && rc == MEP_OK; ++r)
      for (int32_t i = 0; i < n_models && rc == MEP_OK; ++i) {
        mep_ctx* c = ctxs[i];
        const int64_t k0 = r * c->kcap;
        if (k0 >= nk[i]) continue;
        const int64_t nb = (nk[i] - k0 < c->kcap) ? nk[i] - k0 : c->kcap;
        rc = run_block(c, bs[i], k0, nb, mp_grids + 3 * i, nullptr, nullptr, &params[i]);
      }
    for (int32_t i = 0; i < n_models && rc == MEP_OK; ++i) {
      reduce_records(ctxs[i], bs[i]);
      e = hipEventRecord(ev[1 + i], ctxs[i]->st);
      if (e == hipSuccess) e = hipStreamWaitEvent(g->st, ev[1 + i], 0);
      if (e != hipSuccess) rc = fail(MEP_ERR_HIP, std::string("mep_run_batch_graph: ") + hipGetErrorString(e));
    }
  }
  hipGraph_t graph = nullptr;
  const hipError_t ee = hipStreamEndCapture(g->st, &graph);
  if (rc == MEP_OK && ee != hipSuccess)
    rc = fail(MEP_ERR_HIP, std::string("mep_run_batch_graph: capture: ") + hipGetErrorString(ee));
  if (rc == MEP_OK) {
    const hipError_t ei = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess)
      rc = fail(MEP_ERR_HIP, std::string("mep_run_batch_graph: instantiate: ") + hipGetErrorString(ei));
  }
  if (graph) hipGraphDestroy(graph);
  for (auto& e : ev) hipEventDestroy(e);
  for (int32_t i = 0; i < n_models; ++i) ctxs[i]->prof = prof[i];
  (void)hipGetLastError();
  return rc;
}

extern "C" int mep_run_batch_graph(int32_t n_models, mep_ctx* const* ctxs,
                                   const int32_t* mp_grids, const mep_run_params* params,
                                   mep_sums* outs) {
  if (n_models < 1 || !ctxs || !mp_grids || !params)
    return fail(MEP_ERR_ARG, "mep_run_batch_graph: bad argument");
  std::vector<BandSetup> bs((size_t)n_models);
  std::vector<int64_t> nk((size_t)n_models);
  int64_t rounds = 0;
  for (int32_t i = 0; i < n_models; ++i) {
    mep_ctx* c = ctxs[i];
    const int32_t* mp = mp_grids + 3 * i;
    if (!c) return fail(MEP_ERR_ARG, "mep_run_batch_graph: a context is NULL");
    for (int32_t j = 0; j < i; ++j)
      if (ctxs[j] == c) return fail(MEP_ERR_ARG, "mep_run_batch_graph: a context appears twice");
    if (c->device != ctxs[0]->device)
      return fail(MEP_ERR_ARG, "mep_run_batch_graph: the contexts are on different devices");
    if (mp[0] < 1 || mp[1] < 1 || mp[2] < 1)
      return fail(MEP_ERR_ARG, "mep_run_batch_graph: mp_grid entries must be >= 1");
    if (params[i].hw_lst && params[i].n_hw < 1)
      return fail(MEP_ERR_ARG, "mep_run_batch_graph: n_hw must be >= 1");
    nk[i] = (int64_t)mp[0] * mp[1] * mp[2];
    CHK(ensure_ws(c, nk[i]));
    CHK(sep_prepare(c, mp));
    CHK(band_setup(c, &params[i], c->kcap, bs[i]));
    const int64_t nblk = (nk[i] + c->kcap - 1) / c->kcap;
    if (nblk > rounds) rounds = nblk;
  }
  // the uploads of the set-up precede the graph on every context's stream
  for (int32_t i = 0; i < n_models; ++i) HIPCHK(hipStreamSynchronize(ctxs[i]->st));
  BatchGraph* g = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_bg_mu);
    for (size_t i = 0; i < g_bg.size(); ++i)
      if (bg_matches(g_bg[i], n_models, ctxs, mp_grids, params)) g = g_bg[i];
    if (!g) {
      // a stale graph over the same contexts is replaced
      for (int32_t i = 0; i < n_models; ++i) bg_drop_ctx(ctxs[i]);
      g = new BatchGraph();
      g->device = ctxs[0]->device;
      g->ctxs.assign(ctxs, ctxs + n_models);
      g->grids.assign(mp_grids, mp_grids + 3 * n_models);
      g->params.assign(params, params + n_models);
      for (int32_t i = 0; i < n_models; ++i) {
        g->hw.emplace_back(params[i].hw_lst ? std::vector<double>(params[i].hw_lst,
                                                                  params[i].hw_lst + params[i].n_hw)
                                            : std::vector<double>());
        g->params[i].hw_lst = nullptr;
        if (params[i].hw_lst) g->params[i].hw_lst = g->hw[i].data();
      }
      hipError_t e = hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking);
      if (e != hipSuccess) {
        delete g;
        return fail(MEP_ERR_HIP, std::string("mep_run_batch_graph: ") + hipGetErrorString(e));
      }
      const int rc = bg_capture(g, n_models, ctxs, mp_grids, params, bs, nk, rounds);
      if (rc != MEP_OK) {
        bg_free(g);
        return rc;
      }
      for (int32_t i = 0; i < n_models; ++i) g->gens.push_back(ctxs[i]->gen);
      g_bg.push_back(g);
    }
    HIPCHK(hipGraphLaunch(g->exec, g->st));
    HIPCHK(hipStreamSynchronize(g->st));
  }
  for (int32_t i = 0; i < n_models; ++i) {
    mep_ctx* c = ctxs[i];
    CHK(check_info(c));
    const int64_t rec = band_record_len(bs[i]);
    std::vector<double> h(rec);
    HIPCHK(hipMemcpy(h.data(), c->dacc, sizeof(double) * rec, hipMemcpyDeviceToHost));
    if (outs) {
      for_each_output(bs[i], &params[i], &outs[i], [&](double* dst, int64_t off, int64_t len) {
        if (dst) memcpy(dst, h.data() + off, sizeof(double) * len);
      });
      outs[i].n_k = nk[i];
    }
  }
  return MEP_OK;
}

extern "C" int mep_run_mesh(mep_ctx* c, const int32_t mp_grid[3], const mep_run_params* p,
                            mep_sums* out) {
  if (!c || !mp_grid) return fail(MEP_ERR_ARG, "mep_run_mesh: bad argument");
  if (mp_grid[0] < 1 || mp_grid[1] < 1 || mp_grid[2] < 1)
    return fail(MEP_ERR_ARG, "mep_run_mesh: mp_grid entries must be >= 1");
  int64_t nk = (int64_t)mp_grid[0] * mp_grid[1] * mp_grid[2];
  return run_common(c, nk, mp_grid, nullptr, nullptr, p, out);
}

extern "C" int mep_mesh_kpts(mep_ctx* c, const int32_t mp_grid[3], int64_t k0, int64_t nk,
                             double* kpts) {
  if (!c || !mp_grid || !kpts || nk < 0 || k0 < 0)
    return fail(MEP_ERR_ARG, "mep_mesh_kpts: bad argument");
  if (mp_grid[0] < 1 || mp_grid[1] < 1 || mp_grid[2] < 1)
    return fail(MEP_ERR_ARG, "mep_mesh_kpts: mp_grid entries must be >= 1");
  const int64_t ntot = (int64_t)mp_grid[0] * mp_grid[1] * mp_grid[2];
  if (k0 + nk > ntot) return fail(MEP_ERR_ARG, "mep_mesh_kpts: k0 + nk exceeds the mesh");
  if (nk == 0) return MEP_OK;
  // the same kernel as the runs' k-point generation, into a buffer of its
  // own (the run workspace may not exist yet)
  const int64_t chunk = nk < (1 << 20) ? nk : (1 << 20);
  double* dk = nullptr;
  HIPCHK(hipMalloc(&dk, sizeof(double) * 3 * chunk));
  const int m3[3] = {mp_grid[0], mp_grid[1], mp_grid[2]};
  int rc = MEP_OK;
  for (int64_t q = 0; q < nk && rc == MEP_OK; q += chunk) {
    const int64_t nb = (nk - q < chunk) ? nk - q : chunk;
    timed(c, K_KGEN, [&] { launch_mp_kpts(c->st, k0 + q, nb, m3, dk); });
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(kpts + 3 * q, dk, sizeof(double) * 3 * nb, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) rc = fail(MEP_ERR_HIP, std::string("mep_mesh_kpts: ") + hipGetErrorString(e));
  }
  hipFree(dk);
  resolve_times(c);
  return rc;
}

extern "C" int mep_run_kpts(mep_ctx* c, int64_t nk, const double* kpts, const double* weights,
                            const mep_run_params* p, mep_sums* out) {
  if (!c || nk < 1 || !kpts) return fail(MEP_ERR_ARG, "mep_run_kpts: bad argument");
  return run_common(c, nk, nullptr, kpts, weights, p, out);
}

// ---------------------------------------------------------------------------
// mep_run_kpts_static: mep_run_kpts with every work buffer in a file-scope
// device array, for callers that cannot allocate device memory
// ---------------------------------------------------------------------------
// One copy per GPU, held until the process ends.  Host code never hands the
// array's name to a kernel or a copy: its device address comes from
// hipGetSymbolAddress, once per device, with that device current.
static __device__ __align__(256) unsigned char scratch[MEP_STATIC_SCRATCH_BYTES];

static const int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_dev[kMaxDevices] = {};

static int static_scratch(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices)
    return fail(MEP_ERR_ARG, "mep_run_kpts_static: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(MEP_ERR_HIP, std::string("mep_run_kpts_static: hipGetSymbolAddress: ") +
                                   hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch_dev[dev];
  return MEP_OK;
}

// get_linspace (matrix_math.f90:74-100) on the device, rounded as the host's
// lo + i * ((hi - lo) / (n - 1)) is (no fused multiply-add)
__global__ void linspace_kernel(double lo, double hi, int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (n == 1) {
    out[0] = lo;
    return;
  }
  const double d = (hi - lo) / (double)(n - 1);
  out[i] = __dadd_rn(lo, __dmul_rn((double)i, d));
}

static void launch_linspace(hipStream_t s, double lo, double hi, int n, double* out) {
  hipLaunchKernelGGL(linspace_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, lo, hi, n,
                     out);
}

// the work buffers of a context, as ensure_ws and band_setup size them
struct WsPtrs {
  double *dk = nullptr, *dw = nullptr, *dP = nullptr, *dE = nullptr, *dscr = nullptr;
  zd *dHk = nullptr, *dU = nullptr, *dV = nullptr, *dA = nullptr, *dOm = nullptr;
  double *def = nullptr, *dhw = nullptr, *dpart = nullptr, *dacc = nullptr;
};

// Carves the buffers of kc k-points and one band-sum record set out of
// base, each 256-byte aligned; returns the bytes used (with base == nullptr
// only the size is computed).
static size_t static_layout(const mep_ctx* c, int64_t kc, const BandSetup& bs,
                            const mep_run_params* p, unsigned char* base, WsPtrs* w) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    off = (off + 255) & ~(size_t)255;
    unsigned char* q = base ? base + off : nullptr;
    off += bytes;
    return q;
  };
  const int64_t n = c->n, n2 = n * n, rec = band_record_len(bs);
  WsPtrs t;
  t.dk = reinterpret_cast<double*>(take(sizeof(double) * 3 * kc));
  t.dw = reinterpret_cast<double*>(take(sizeof(double) * kc));
  t.dP = reinterpret_cast<double*>(take(sizeof(double) * c->K * kc));
  t.dHk = reinterpret_cast<zd*>(take(sizeof(zd) * c->hk_per_k * kc));
  t.dE = reinterpret_cast<double*>(take(sizeof(double) * n * kc));
  t.dU = reinterpret_cast<zd*>(take(sizeof(zd) * n2 * kc));
  t.dV = reinterpret_cast<zd*>(take(sizeof(zd) * 3 * n2 * kc));
  t.dscr = reinterpret_cast<double*>(take(eigh_scratch_bytes((int)n, kc)));
  if (c->pos) {
    t.dA = reinterpret_cast<zd*>(take(sizeof(zd) * 3 * n2 * kc));
    t.dOm = reinterpret_cast<zd*>(take(sizeof(zd) * 9 * n2 * kc));
  }
  t.def = reinterpret_cast<double*>(take(sizeof(double) * p->n_eF));
  t.dhw = reinterpret_cast<double*>(take(sizeof(double) * p->n_hw));
  t.dpart = reinterpret_cast<double*>(take(sizeof(double) * bs.nblocks * rec));
  t.dacc = reinterpret_cast<double*>(take(sizeof(double) * rec));
  if (w) *w = t;
  return off;
}

// Points the context's work buffers at the carved ones for the lifetime of
// the object; its own buffers (if any) are left untouched and restored.
struct WsSwap {
  mep_ctx* c;
  WsPtrs saved;
  int64_t kcap, ef_cap, hw_cap, part_cap, acc_cap;
  WsSwap(mep_ctx* c_, const WsPtrs& w, int64_t kc, const BandSetup& bs, const mep_run_params* p)
      : c(c_) {
    saved.dk = c->dk; saved.dw = c->dw; saved.dP = c->dP; saved.dHk = c->dHk; saved.dE = c->dE;
    saved.dU = c->dU; saved.dV = c->dV; saved.dA = c->dA; saved.dOm = c->dOm;
    saved.dscr = c->dscr; saved.def = c->def; saved.dhw = c->dhw; saved.dpart = c->dpart;
    saved.dacc = c->dacc;
    kcap = c->kcap; ef_cap = c->ef_cap; hw_cap = c->hw_cap; part_cap = c->part_cap;
    acc_cap = c->acc_cap;
    c->dk = w.dk; c->dw = w.dw; c->dP = w.dP; c->dHk = w.dHk; c->dE = w.dE; c->dU = w.dU;
    c->dV = w.dV; c->dA = w.dA; c->dOm = w.dOm; c->dscr = w.dscr; c->def = w.def;
    c->dhw = w.dhw; c->dpart = w.dpart; c->dacc = w.dacc;
    c->kcap = kc;
    c->ef_cap = p->n_eF;
    c->hw_cap = p->n_hw;
    c->part_cap = (int64_t)bs.nblocks * band_record_len(bs);
    c->acc_cap = band_record_len(bs);
  }
  ~WsSwap() {
    c->dk = saved.dk; c->dw = saved.dw; c->dP = saved.dP; c->dHk = saved.dHk; c->dE = saved.dE;
    c->dU = saved.dU; c->dV = saved.dV; c->dA = saved.dA; c->dOm = saved.dOm;
    c->dscr = saved.dscr; c->def = saved.def; c->dhw = saved.dhw; c->dpart = saved.dpart;
    c->dacc = saved.dacc;
    c->kcap = kcap; c->ef_cap = ef_cap; c->hw_cap = hw_cap; c->part_cap = part_cap;
    c->acc_cap = acc_cap;
  }
};

struct DeviceRestore {
  int dev;
  bool on;
  ~DeviceRestore() {
    if (on) hipSetDevice(dev);
  }
};

extern "C" int mep_run_kpts_static(mep_ctx* c, int64_t nk, const double* kpts,
                                   const double* weights, const mep_run_params* p, mep_sums* out) {
  if (!c || nk < 1 || !kpts || !p) return fail(MEP_ERR_ARG, "mep_run_kpts_static: bad argument");
  // one block of nk k-points, as mep_run_kpts takes a list that fits its
  // workspace: the same k -> workgroup map and record count, the same sums
  BandSetup bs;
  CHK(band_plan(c, p, nk, bs));
  const size_t need = static_layout(c, nk, bs, p, nullptr, nullptr);
  if (need > MEP_STATIC_SCRATCH_BYTES) {
    int64_t lo = 0, hi = nk - 1;  // the largest k-count that fits, for the message
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) / 2;
      BandSetup b2;
      if (band_plan(c, p, mid, b2) == MEP_OK &&
          static_layout(c, mid, b2, p, nullptr, nullptr) <= MEP_STATIC_SCRATCH_BYTES)
        lo = mid;
      else
        hi = mid - 1;
    }
    return fail(MEP_ERR_ARG, "mep_run_kpts_static: " + std::to_string(nk) + " k-points need " +
                                 std::to_string(need) + " bytes of work buffers; the static "
                                 "scratch holds " + std::to_string(MEP_STATIC_SCRATCH_BYTES) +
                                 " bytes (at most " + std::to_string(lo) +
                                 " k-points for this model and these parameters)");
  }
  int cur = 0;
  HIPCHK(hipGetDevice(&cur));
  DeviceRestore restore{cur, cur != c->device};
  if (cur != c->device) HIPCHK(hipSetDevice(c->device));
  unsigned char* base = nullptr;
  CHK(static_scratch(c->device, &base));
  WsPtrs w;
  static_layout(c, nk, bs, p, base, &w);
  WsSwap swap(c, w, nk, bs, p);
  // Fermi levels and frequencies (get_linspace, or the caller's hw_lst);
  // the records start from zero: the array keeps the previous call's data
  bs.b.ef = c->def;
  bs.b.hw = c->dhw;
  launch_linspace(c->st, p->eF_min, p->eF_max, p->n_eF, c->def);
  if (p->hw_lst)
    HIPCHK(hipMemcpyAsync(c->dhw, p->hw_lst, sizeof(double) * p->n_hw, hipMemcpyHostToDevice,
                          c->st));
  else
    launch_linspace(c->st, p->hw_min, p->hw_max, p->n_hw, c->dhw);
  const int64_t rec = band_record_len(bs);
  launch_fill_zero(c->st, c->dacc, rec);
  launch_fill_zero(c->st, c->dpart, (int64_t)bs.nblocks * rec);
  HIPCHK(hipGetLastError());
  CHK(run_block(c, bs, 0, nk, nullptr, kpts, weights, p));
  CHK(check_info(c));
  reduce_records(c, bs);
  HIPCHK(hipGetLastError());
  // each sum straight from the device record into the caller's array
  hipError_t e = hipSuccess;
  if (out)
    for_each_output(bs, p, out, [&](double* dst, int64_t off, int64_t len) {
      if (dst && e == hipSuccess)
        e = hipMemcpyAsync(dst, c->dacc + off, sizeof(double) * len, hipMemcpyDeviceToHost, c->st);
    });
  HIPCHK(e);
  HIPCHK(hipStreamSynchronize(c->st));
  resolve_times(c);
  if (out) out->n_k = nk;
  return MEP_OK;
}
