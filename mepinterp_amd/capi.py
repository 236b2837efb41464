"""ctypes binding of the C-ABI in include/mepinterp.h (libmepinterp_amd.so).

This is the Python-side counterpart of the Fortran iso_c_binding module: it
loads the in-tree gfx950 library and fails loudly when it is missing.  There
is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmepinterp_amd.so")

MEP_OK = 0
MEP_F_GAUGE_TRAFO = 1
MEP_F_CART_VELO = 2
MEP_F_TEXTBOOK = 4
MEP_F_APPLY_DEG = 8
MEP_F_SEPARATE_VELO = 16
MEP_JOB_AHC = 1
MEP_JOB_KUBO = 2
MEP_JOB_MEP = 4
MEP_JOB_OPT = 8
MEP_JOB_GYRO = 16
MEP_PREC_FP64 = 0
MEP_PREC_MIXED = 1
MEP_VAR_EIGH = 1
MEP_VAR_GEMM32 = 2

# every symbol include/mepinterp.h declares
EXPORTS = ("mep_ctx_create", "mep_ctx_destroy", "mep_last_error", "mep_interp_batch",
           "mep_bandsum_batch", "mep_run_mesh", "mep_run_kpts", "mep_set_profiling",
           "mep_get_kernel_time", "mep_reset_kernel_times", "mep_set_kblock", "mep_curv_batch",
           "mep_run_batch", "mep_set_workspace_limit", "mep_mesh_kpts", "mep_run_kpts_static",
           "mep_set_variant", "mep_run_batch_graph",
           "mep_mlp_create", "mep_mlp_destroy", "mep_mlp_num_params", "mep_mlp_get_params",
           "mep_mlp_set_params", "mep_mlp_train", "mep_mlp_predict", "mep_mlp_grad")

c_dp = C.POINTER(C.c_double)


class RunParams(C.Structure):
    _fields_ = [("jobs", C.c_int32), ("n_eF", C.c_int32), ("eF_min", C.c_double),
                ("eF_max", C.c_double), ("n_hw", C.c_int32), ("hw_min", C.c_double),
                ("hw_max", C.c_double), ("T_kelvin", C.c_double), ("eta", C.c_double),
                ("valence_bands", C.c_int32), ("precision", C.c_int32), ("hw_lst", c_dp)]


class Sums(C.Structure):
    _fields_ = [("Ne", c_dp), ("ahc", c_dp), ("ahc_velo", c_dp), ("ohc", c_dp),
                ("kubo_ic", c_dp), ("kubo_lc", c_dp), ("kubo_cs", c_dp),
                ("niu_ic", c_dp), ("niu_lc", c_dp), ("niu_cs", c_dp), ("n_k", C.c_int64),
                ("opt_s", c_dp), ("opt_a", c_dp), ("photo", c_dp), ("gyro_C", c_dp),
                ("gyro_D", c_dp), ("gyro_Dw", c_dp)]


_lib = None


class MepError(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Load libmepinterp_amd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MepError(f"{path} is missing: build it with `python -m mepinterp_amd.build`")
    lib = C.CDLL(path)
    lib.mep_last_error.restype = C.c_char_p
    for name in EXPORTS:
        if name != "mep_last_error":
            getattr(lib, name).restype = C.c_int
    lib.mep_ctx_create.argtypes = [C.c_int64, C.c_int64, c_dp, c_dp, c_dp, c_dp,
                                   C.POINTER(C.c_int64), c_dp, C.c_double, C.c_int32,
                                   C.POINTER(C.c_void_p)]
    lib.mep_ctx_destroy.argtypes = [C.c_void_p]
    lib.mep_interp_batch.argtypes = [C.c_void_p, C.c_int64, c_dp, c_dp, c_dp, c_dp, c_dp]
    lib.mep_bandsum_batch.argtypes = [C.c_void_p, C.c_int64, c_dp, c_dp, c_dp, c_dp, c_dp,
                                      C.POINTER(RunParams), C.POINTER(Sums)]
    lib.mep_run_mesh.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(RunParams),
                                 C.POINTER(Sums)]
    lib.mep_run_kpts.argtypes = [C.c_void_p, C.c_int64, c_dp, c_dp, C.POINTER(RunParams),
                                 C.POINTER(Sums)]
    lib.mep_mesh_kpts.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int64, C.c_int64, c_dp]
    lib.mep_run_kpts_static.argtypes = [C.c_void_p, C.c_int64, c_dp, c_dp, C.POINTER(RunParams),
                                        C.POINTER(Sums)]
    lib.mep_set_profiling.argtypes = [C.c_void_p, C.c_int32]
    lib.mep_get_kernel_time.argtypes = [C.c_void_p, C.c_char_p, c_dp, C.POINTER(C.c_int64)]
    lib.mep_reset_kernel_times.argtypes = [C.c_void_p]
    lib.mep_set_kblock.argtypes = [C.c_void_p, C.c_int64]
    lib.mep_set_workspace_limit.argtypes = [C.c_void_p, C.c_double]
    lib.mep_set_variant.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    lib.mep_curv_batch.argtypes = [C.c_void_p, C.c_int64, c_dp, C.c_double, C.c_double, c_dp]
    lib.mep_run_batch.argtypes = [C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_int32),
                                  C.POINTER(RunParams), C.POINTER(Sums)]
    lib.mep_run_batch_graph.argtypes = lib.mep_run_batch.argtypes
    c_fp = C.POINTER(C.c_float)
    lib.mep_mlp_create.argtypes = [C.c_int32, C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]
    lib.mep_mlp_destroy.argtypes = [C.c_void_p]
    lib.mep_mlp_num_params.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    lib.mep_mlp_get_params.argtypes = [C.c_void_p, c_fp]
    lib.mep_mlp_set_params.argtypes = [C.c_void_p, c_fp]
    lib.mep_mlp_train.argtypes = [C.c_void_p, C.c_int64, c_dp, c_fp, C.c_int32, C.c_int32,
                                  C.c_float, C.c_uint64, c_fp]
    lib.mep_mlp_predict.argtypes = [C.c_void_p, C.c_int64, c_dp, c_fp]
    lib.mep_mlp_grad.argtypes = [C.c_void_p, C.c_int64, c_dp, c_fp, c_fp, c_fp]
    _lib = lib
    return lib


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(c_dp)


def _pf(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _check(rc, what):
    if rc != MEP_OK:
        msg = load().mep_last_error().decode(errors="replace")
        raise MepError(f"{what} failed (rc={rc}): {msg}")


def _f(a, dtype):
    return None if a is None else np.asfortranarray(a, dtype=dtype)


class Engine:
    """A model resident on the GPU (one mep_ctx).

    Arguments follow the reference's arrays: R_frac (3,nR), H (n,n,nR)
    Hartree, r (3,n,n,nR) Bohr or None, wf_centers (3,n) or None, a_latt
    with rows a_i (Bohr).
    """

    def __init__(self, R_frac, H, r=None, wf_centers=None, a_latt=None, kubo_tol=1e-5,
                 gauge_trafo=True, cart_velo=False, textbook=False, deg=None, apply_deg=False,
                 separate_velo=False):
        lib = load()
        self.R = _f(R_frac, np.float64)
        self.H = _f(H, np.complex128)
        self.r = _f(r, np.complex128)
        self.tau = _f(wf_centers, np.float64)
        self.a_latt = np.asfortranarray(np.eye(3) if a_latt is None else a_latt, dtype=np.float64)
        self.n = self.H.shape[0]
        self.nR = self.H.shape[2]
        self.pos = self.r is not None
        flags = (MEP_F_GAUGE_TRAFO if gauge_trafo else 0) | (MEP_F_CART_VELO if cart_velo else 0) \
            | (MEP_F_TEXTBOOK if textbook else 0) | (MEP_F_APPLY_DEG if apply_deg else 0) \
            | (MEP_F_SEPARATE_VELO if separate_velo else 0)
        d = None if deg is None else np.ascontiguousarray(deg, dtype=np.int64)
        ctx = C.c_void_p()
        rc = lib.mep_ctx_create(self.n, self.nR, _p(self.R), _p(self.tau), _p(self.H), _p(self.r),
                                None if d is None else d.ctypes.data_as(C.POINTER(C.c_int64)),
                                _p(self.a_latt), float(kubo_tol), flags, C.byref(ctx))
        _check(rc, "mep_ctx_create")
        self.ctx = ctx

    def close(self):
        if getattr(self, "ctx", None):
            load().mep_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- get_wann_interp over a batch ------------------------------------
    def interp(self, kpts, velo=True):
        k = np.asfortranarray(kpts, dtype=np.float64).reshape(3, -1, order="F")
        nk, n = k.shape[1], self.n
        e = np.zeros((n, nk), order="F")
        V = np.zeros((3, n, n, nk), np.complex128, order="F") if velo else None
        A = np.zeros((3, n, n, nk), np.complex128, order="F") if (velo and self.pos) else None
        Om = np.zeros((3, 3, n, n, nk), np.complex128, order="F") if (velo and self.pos) else None
        rc = load().mep_interp_batch(self.ctx, nk, _p(k), _p(e), _p(V), _p(A), _p(Om))
        _check(rc, "mep_interp_batch")
        return e, V, A, Om

    @staticmethod
    def params(jobs=MEP_JOB_AHC, n_eF=1, eF_min=0.0, eF_max=0.0, n_hw=1, hw_min=0.0, hw_max=0.0,
               T_kelvin=0.0, eta=0.0, valence_bands=1, precision=MEP_PREC_FP64, hw_lst=None):
        """Run parameters; ``hw_lst`` (Hartree) replaces the linspace of
        (hw_min, hw_max, n_hw) with the caller's frequencies."""
        hw = None
        if hw_lst is not None:
            hw = np.ascontiguousarray(hw_lst, dtype=np.float64).ravel()
            n_hw = hw.size
        p = RunParams(int(jobs), int(n_eF), float(eF_min), float(eF_max), int(max(1, n_hw)),
                      float(hw_min), float(hw_max), float(T_kelvin), float(eta),
                      int(valence_bands), int(precision), _p(hw))
        p.hw_keep = hw  # the list lives as long as the parameters
        return p

    @staticmethod
    def _sums_buffers(p: RunParams):
        nef, nhw, nv = p.n_eF, p.n_hw, max(1, p.valence_bands)
        out = {"Ne": np.zeros(nef), "ahc": np.zeros((3, 3, nef), order="F"),
               "ahc_velo": np.zeros((3, 3, nhw, nef), np.complex128, order="F"),
               "ohc": np.zeros((3, 3, nhw, nef), np.complex128, order="F"),
               "kubo_ic": np.zeros((3, 3, nef), order="F"),
               "kubo_lc": np.zeros((3, 3, nef), order="F"),
               "kubo_cs": np.zeros((3, 3, nef), order="F"),
               "niu_ic": np.zeros((3, 3, nv), order="F"),
               "niu_lc": np.zeros((3, 3, nv), order="F"),
               "niu_cs": np.zeros((3, 3, nv), order="F"),
               "opt_s": np.zeros((3, 3, nhw, nef), np.complex128, order="F"),
               "opt_a": np.zeros((3, 3, nhw, nef), np.complex128, order="F"),
               "photo": np.zeros((3, 3, 3, nhw, nef), np.complex128, order="F"),
               "gyro_C": np.zeros((3, 3, nef), np.complex128, order="F"),
               "gyro_D": np.zeros((3, 3, nef), np.complex128, order="F"),
               "gyro_Dw": np.zeros((3, 3, nhw, nef), np.complex128, order="F")}
        s = Sums()
        for k in out:
            setattr(s, k, _p(out[k]))
        return out, s

    def bandsum(self, e, V, p: RunParams, A=None, Om=None, fd=None):
        e = np.asfortranarray(e, dtype=np.float64).reshape(self.n, -1, order="F")
        V = np.asfortranarray(V, dtype=np.complex128)
        out, s = self._sums_buffers(p)
        rc = load().mep_bandsum_batch(self.ctx, e.shape[1], _p(e), _p(V), _p(_f(A, np.complex128)),
                                      _p(_f(Om, np.complex128)), _p(_f(fd, np.float64)),
                                      C.byref(p), C.byref(s))
        _check(rc, "mep_bandsum_batch")
        out["n_k"] = s.n_k
        return out

    def run_mesh(self, mp_grid, p: RunParams):
        mp = (C.c_int32 * 3)(*[int(x) for x in mp_grid])
        out, s = self._sums_buffers(p)
        rc = load().mep_run_mesh(self.ctx, mp, C.byref(p), C.byref(s))
        _check(rc, "mep_run_mesh")
        out["n_k"] = s.n_k
        return out

    def mesh_kpts(self, mp_grid, k0=0, nk=None):
        """The device-generated k-points k0 .. k0+nk-1 of the MP mesh (3, nk)."""
        mp = (C.c_int32 * 3)(*[int(x) for x in mp_grid])
        nk = int(np.prod(mp_grid)) - int(k0) if nk is None else int(nk)
        k = np.zeros((3, nk), order="F")
        _check(load().mep_mesh_kpts(self.ctx, mp, int(k0), nk, _p(k)), "mep_mesh_kpts")
        return k

    def run_kpts(self, kpts, p: RunParams, weights=None):
        k = np.asfortranarray(kpts, dtype=np.float64).reshape(3, -1, order="F")
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        out, s = self._sums_buffers(p)
        rc = load().mep_run_kpts(self.ctx, k.shape[1], _p(k), _p(w), C.byref(p), C.byref(s))
        _check(rc, "mep_run_kpts")
        out["n_k"] = s.n_k
        return out

    def run_kpts_static(self, kpts, p: RunParams, weights=None):
        """mep_run_kpts_static: run_kpts with the work buffers in the library's
        static device array (no device allocation in the call)."""
        k = np.asfortranarray(kpts, dtype=np.float64).reshape(3, -1, order="F")
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        out, s = self._sums_buffers(p)
        rc = load().mep_run_kpts_static(self.ctx, k.shape[1], _p(k), _p(w), C.byref(p), C.byref(s))
        _check(rc, "mep_run_kpts_static")
        out["n_k"] = s.n_k
        return out

    # -- timing ------------------------------------------------------------
    def set_profiling(self, on=True):
        _check(load().mep_set_profiling(self.ctx, 1 if on else 0), "mep_set_profiling")

    def reset_times(self):
        _check(load().mep_reset_kernel_times(self.ctx), "mep_reset_kernel_times")

    def kernel_time(self, name):
        ms = C.c_double()
        cnt = C.c_int64()
        _check(load().mep_get_kernel_time(self.ctx, name.encode(), C.byref(ms), C.byref(cnt)),
               "mep_get_kernel_time")
        return ms.value, cnt.value

    def set_kblock(self, kb):
        _check(load().mep_set_kblock(self.ctx, int(kb)), "mep_set_kblock")

    def set_workspace_limit(self, nbytes):
        _check(load().mep_set_workspace_limit(self.ctx, float(nbytes)), "mep_set_workspace_limit")

    def set_variant(self, what, value):
        """Kernel variant of one stage (VAR_EIGH: 1 or 2; 0 = the default)."""
        _check(load().mep_set_variant(self.ctx, int(what), int(value)), "mep_set_variant")

    def curv(self, kpts, eF=0.0, T_kelvin=0.0):
        """Per-k Berry curvature summed over occupied bands, (3, 3, nk)."""
        k = np.asfortranarray(kpts, dtype=np.float64).reshape(3, -1, order="F")
        out = np.zeros((3, 3, k.shape[1]), order="F")
        _check(load().mep_curv_batch(self.ctx, k.shape[1], _p(k), float(eF), float(T_kelvin),
                                     _p(out)), "mep_curv_batch")
        return out


def run_batch(engines, mp_grids, params, graph=False):
    """mep_run_batch: several models' meshes in one call, run concurrently
    (one stream per engine).  Returns one sums dict per engine.  graph=True:
    mep_run_batch_graph (captured once into a HIP graph, replayed after)."""
    n = len(engines)
    ctxs = (C.c_void_p * n)(*[e.ctx for e in engines])
    mp = (C.c_int32 * (3 * n))(*[int(x) for g in mp_grids for x in g])
    ps = (RunParams * n)(*params)
    outs, sums = [], (Sums * n)()
    for i, p in enumerate(params):
        o, s_ = Engine._sums_buffers(p)
        outs.append(o)
        sums[i] = s_
    name = "mep_run_batch_graph" if graph else "mep_run_batch"
    _check(getattr(load(), name)(n, ctxs, mp, ps, sums), name)
    for i in range(n):
        outs[i]["n_k"] = sums[i].n_k
    return outs


class MLP:
    """The learned-refinement network on the GPU (one mep_mlp)."""

    def __init__(self, n_freq=4, hidden=128, seed=0):
        self.h = C.c_void_p()
        _check(load().mep_mlp_create(int(n_freq), int(hidden), int(seed), C.byref(self.h)),
               "mep_mlp_create")
        n = C.c_int64()
        _check(load().mep_mlp_num_params(self.h, C.byref(n)), "mep_mlp_num_params")
        self.n_params = n.value
        self.n_freq, self.hidden = int(n_freq), int(hidden)

    def close(self):
        if self.h:
            load().mep_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def params(self):
        p = np.zeros(self.n_params, np.float32)
        _check(load().mep_mlp_get_params(self.h, _pf(p)), "mep_mlp_get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        _check(load().mep_mlp_set_params(self.h, _pf(p)), "mep_mlp_set_params")

    def train(self, kpts, target, epochs=20, batch=4096, lr=1e-3, seed=0):
        k = np.asfortranarray(kpts, dtype=np.float64).reshape(3, -1, order="F")
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = np.zeros(1, np.float32)
        _check(load().mep_mlp_train(self.h, k.shape[1], _p(k), _pf(t), int(epochs), int(batch),
                                    float(lr), int(seed), _pf(loss)), "mep_mlp_train")
        return float(loss[0])

    def predict(self, kpts):
        k = np.asfortranarray(kpts, dtype=np.float64).reshape(3, -1, order="F")
        out = np.zeros(k.shape[1], np.float32)
        _check(load().mep_mlp_predict(self.h, k.shape[1], _p(k), _pf(out)), "mep_mlp_predict")
        return out

    def grad(self, kpts, target):
        k = np.asfortranarray(kpts, dtype=np.float64).reshape(3, -1, order="F")
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = np.zeros(1, np.float32)
        _check(load().mep_mlp_grad(self.h, k.shape[1], _p(k), _pf(t), _pf(g), _pf(loss)),
               "mep_mlp_grad")
        return g, float(loss[0])
