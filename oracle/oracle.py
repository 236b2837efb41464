"""ctypes front-end of the CPU FP64 oracle (oracle/mep_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.

Besides the per-routine wrappers it restates the reference's end-of-run
steps in numpy: ``fermi_selector`` (core.f90:317-415) and
``integrate_over_k_space`` / ``mep_sum_bands`` / ``normalize_k_int``
(core.f90:419-603, k_space.f90:81-202).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmep_oracle.so")
_lib = None

c_dp = C.POINTER(C.c_double)


class _Model(C.Structure):
    _fields_ = [("n_wf", C.c_int64), ("n_R", C.c_int64), ("R_frac", c_dp),
                ("H_real", c_dp), ("r_real", c_dp), ("wf_centers", c_dp),
                ("a_latt", C.c_double * 9), ("use_cart_velo", C.c_int),
                ("do_gauge_trafo", C.c_int), ("kubo_tol", C.c_double),
                ("textbook", C.c_int)]


class _Params(C.Structure):
    _fields_ = [("do_mep", C.c_int), ("do_kubo", C.c_int), ("do_ahc", C.c_int),
                ("n_eF", C.c_int), ("eF_min", C.c_double), ("eF_max", C.c_double),
                ("n_hw", C.c_int), ("hw_min", C.c_double), ("hw_max", C.c_double),
                ("T_kelvin", C.c_double), ("eta", C.c_double),
                ("valence_bands", C.c_int), ("n_threads", C.c_int),
                ("do_opt", C.c_int), ("do_gyro", C.c_int)]


class _Sums(C.Structure):
    _fields_ = [(k, c_dp) for k in ("Ne", "ahc", "ahc_velo", "ohc", "kubo_ic", "kubo_lc",
                                     "kubo_cs", "niu_ic", "niu_lc", "niu_cs", "opt_s",
                                     "opt_a", "photo", "gyro_C", "gyro_D", "gyro_Dw")]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _lib.orc_accumulate.restype = C.c_int
        _lib.orc_zheev.restype = C.c_int
        _lib.orc_get_wann_interp.restype = C.c_int
        _lib.orc_unit_vol.restype = C.c_double
        _lib.orc_kpt_idx.restype = C.c_int64
    return _lib


def _p(a):
    """Pointer to an array's data; strided views are refused (the C side
    reads contiguous memory)."""
    if a is None:
        return None
    if not (a.flags.c_contiguous or a.flags.f_contiguous):
        raise ValueError("oracle: array argument must be contiguous")
    return a.ctypes.data_as(c_dp)


class Model:
    """A tight-binding / Wannier model in the reference's layout.

    R_frac (3,nR), H (n,n,nR) Hartree, r (3,n,n,nR) Bohr or None,
    wf_centers (3,n) or None (Wannier convention), a_latt rows = a_i (Bohr).
    """

    def __init__(self, R_frac, H, r=None, wf_centers=None, a_latt=None,
                 use_cart_velo=False, do_gauge_trafo=True, kubo_tol=1e-5, textbook=False):
        self.R_frac = np.asfortranarray(R_frac, dtype=np.float64)
        self.H = np.asfortranarray(H, dtype=np.complex128)
        self.r = None if r is None else np.asfortranarray(r, dtype=np.complex128)
        self.wf_centers = None if wf_centers is None else np.asfortranarray(
            wf_centers, dtype=np.float64)
        self.a_latt = np.eye(3) if a_latt is None else np.asarray(a_latt, dtype=np.float64)
        self.use_cart_velo = bool(use_cart_velo)
        self.do_gauge_trafo = bool(do_gauge_trafo)
        self.kubo_tol = float(kubo_tol)
        self.textbook = bool(textbook)

    @property
    def n_wf(self):
        return self.H.shape[0]

    @property
    def n_R(self):
        return self.H.shape[2]

    def _c(self):
        m = _Model()
        m.n_wf, m.n_R = self.n_wf, self.n_R
        m.R_frac = _p(self.R_frac)
        m.H_real = _p(self.H)
        m.r_real = _p(self.r)
        m.wf_centers = _p(self.wf_centers)
        al = np.asfortranarray(self.a_latt)
        for i in range(9):
            m.a_latt[i] = al.ravel(order="F")[i]
        m.use_cart_velo = int(self.use_cart_velo)
        m.do_gauge_trafo = int(self.do_gauge_trafo)
        m.kubo_tol = self.kubo_tol
        m.textbook = int(self.textbook)
        return m


def unit_vol(a_latt):
    a = (C.c_double * 9)(*np.asfortranarray(a_latt, dtype=np.float64).ravel(order="F"))
    return lib().orc_unit_vol(a)


def mp_mesh(mp):
    mp = [int(x) for x in mp]
    k = np.zeros((3, mp[0] * mp[1] * mp[2]), dtype=np.float64, order="F")
    lib().orc_mp_mesh((C.c_int * 3)(*mp), _p(k))
    return k


def kpt_idx(mp, x, y, z):
    return lib().orc_kpt_idx((C.c_int * 3)(*mp), int(x), int(y), int(z))


def linspace(lo, hi, n):
    out = np.zeros(max(1, n))
    lib().orc_linspace(C.c_double(lo), C.c_double(hi), C.c_int(n), _p(out))
    return out


def zheev(A):
    A = np.asfortranarray(A, dtype=np.complex128).copy(order="F")
    n = A.shape[0]
    w = np.zeros(n)
    info = lib().orc_zheev(C.c_int64(n), _p(A), _p(w))
    if info:
        raise RuntimeError(f"orc_zheev info={info}")
    return w, A


def ft_R_to_k(model: Model, kpt, grad=True, pos=None):
    n = model.n_wf
    mc = model._c()
    Hk = np.zeros((n, n), np.complex128, order="F")
    Hka = np.zeros((3, n, n), np.complex128, order="F") if grad else None
    pos = (model.r is not None) if pos is None else pos
    A = np.zeros((3, n, n), np.complex128, order="F") if pos else None
    Om = np.zeros((3, 3, n, n), np.complex128, order="F") if pos else None
    k = np.ascontiguousarray(kpt, dtype=np.float64)
    lib().orc_ft_R_to_k(C.byref(mc), _p(k), _p(Hk), _p(Hka), _p(A), _p(Om))
    return Hk, Hka, A, Om


def get_wann_interp(model: Model, kpt, velo=True):
    """wann_interp.f90:56-108 for one k: returns (e_k, V_ka, A_ka, Om_kab)."""
    n = model.n_wf
    mc = model._c()
    e = np.zeros(n)
    V = np.zeros((3, n, n), np.complex128, order="F") if velo else None
    pos = model.r is not None
    A = np.zeros((3, n, n), np.complex128, order="F") if pos else None
    Om = np.zeros((3, 3, n, n), np.complex128, order="F") if pos else None
    k = np.ascontiguousarray(kpt, dtype=np.float64)
    info = lib().orc_get_wann_interp(C.byref(mc), _p(k), _p(e), _p(V), _p(A), _p(Om))
    if info:
        raise RuntimeError(f"eigensolver failed, info={info}")
    return e, V, A, Om


def fd_stat(e, ef_lst, T):
    ef = np.asarray(ef_lst, dtype=np.float64)
    out = np.zeros(ef.size)
    lib().orc_fd_stat(C.c_double(e), _p(ef), C.c_int(ef.size), C.c_double(T), _p(out))
    return out


def fd_distrib(e_k, ef_lst, T):
    return np.asfortranarray(np.stack([fd_stat(e, ef_lst, T) for e in e_k], axis=1))


def kubo_ahc_tens(e, V, fd):
    n_ef = fd.shape[0]
    o = np.zeros((3, 3, n_ef), order="F")
    lib().orc_kubo_ahc_tens(C.c_int64(e.size), _p(e), _p(V), C.c_int(n_ef),
                            _p(np.asfortranarray(fd)), _p(o))
    return o


def velo_ahc_tens(e, V, hw, fd, eta):
    n_ef, hw = fd.shape[0], np.asarray(hw, dtype=np.float64)
    o = np.zeros((3, 3, hw.size, n_ef), np.complex128, order="F")
    lib().orc_velo_ahc_tens(C.c_int64(e.size), _p(e), _p(V), C.c_int(hw.size), _p(hw),
                            C.c_int(n_ef), _p(np.asfortranarray(fd)), C.c_double(eta), _p(o))
    return o


def kubo_ohc_tens(e, V, hw, fd, eta):
    n_ef, hw = fd.shape[0], np.asarray(hw, dtype=np.float64)
    o = np.zeros((3, 3, hw.size, n_ef), np.complex128, order="F")
    lib().orc_kubo_ohc_tens(C.c_int64(e.size), _p(e), _p(V), C.c_int(hw.size), _p(hw),
                            C.c_int(n_ef), _p(np.asfortranarray(fd)), C.c_double(eta), _p(o))
    return o


def kubo_mep_ic(e, V, fd):
    o = np.zeros((3, 3, fd.shape[0]), order="F")
    lib().orc_kubo_mep_ic(C.c_int64(e.size), _p(e), _p(V), C.c_int(fd.shape[0]),
                          _p(np.asfortranarray(fd)), _p(o))
    return o


def kubo_mep_lc(e, V, fd):
    o = np.zeros((3, 3, fd.shape[0]), order="F")
    lib().orc_kubo_mep_lc(C.c_int64(e.size), _p(e), _p(V), C.c_int(fd.shape[0]),
                          _p(np.asfortranarray(fd)), _p(o))
    return o


def kubo_mep_cs(A, Om, fd):
    n = A.shape[1] if A is not None else 0
    o = np.zeros((3, 3, fd.shape[0]), order="F")
    lib().orc_kubo_mep_cs(C.c_int64(n), _p(A), _p(Om), C.c_int(fd.shape[0]),
                          _p(np.asfortranarray(fd)), _p(o))
    return o


def mep_niu_ic(V, e, n_val, kubo_tol):
    o = np.zeros((3, 3, n_val), order="F")
    lib().orc_mep_niu_ic(C.c_int64(e.size), _p(V), _p(e), C.c_int(n_val),
                         C.c_double(kubo_tol), _p(o))
    return o


def mep_niu_lc(V, e, n_val, kubo_tol):
    o = np.zeros((3, 3, n_val), order="F")
    lib().orc_mep_niu_lc(C.c_int64(e.size), _p(V), _p(e), C.c_int(n_val),
                         C.c_double(kubo_tol), _p(o))
    return o


def mep_niu_cs(A, Om, n_val):
    n = A.shape[1] if A is not None else 0
    o = np.zeros((3, 3, n_val), order="F")
    lib().orc_mep_niu_cs(C.c_int64(n), _p(A), _p(Om), C.c_int(n_val), _p(o))
    return o


class RunParams:
    """Job and Fermi/laser parameters of one run (energies in Hartree)."""

    def __init__(self, do_mep=False, do_kubo=False, do_ahc=True, n_eF=1, eF_min=0.0,
                 eF_max=0.0, n_hw=1, hw_min=0.0, hw_max=0.0, T_kelvin=0.0, eta=0.0,
                 valence_bands=1, n_threads=0, do_opt=False, do_gyro=False):
        self.__dict__.update(locals())
        del self.__dict__["self"]

    def _c(self):
        p = _Params()
        for k, _ in _Params._fields_:
            setattr(p, k, getattr(self, k))
        return p


def accumulate(model: Model, params: RunParams, kpts):
    """core.f90:165-270 over an explicit list kpts (3, nk): raw k-sums."""
    kpts = np.asfortranarray(kpts, dtype=np.float64)
    nef, nhw, nv = params.n_eF, params.n_hw, max(1, params.valence_bands)
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
    s = _Sums()
    for k, _ in _Sums._fields_:
        setattr(s, k, _p(out[k]))
    mc, pc = model._c(), params._c()
    rc = lib().orc_accumulate(C.byref(mc), C.byref(pc), C.c_int64(kpts.shape[1]),
                              _p(kpts), C.byref(s))
    if rc:
        raise RuntimeError(f"orc_accumulate failed rc={rc}")
    return out


def fermi_selector(Ne_sum, n_k, ef_lst, n_el_sys):
    """core.f90:317-415: first argmin |Ne(E_F)/N_k - valence_bands| (1-based)."""
    Ne = np.asarray(Ne_sum, dtype=np.float64) / n_k
    opt_idx, opt_err = 1, abs(Ne[0] - float(n_el_sys))
    for i in range(Ne.size):
        err = abs(Ne[i] - float(n_el_sys))
        if err < opt_err:
            opt_idx, opt_err = i + 1, err
    return opt_idx, Ne


def finalize(sums, params: RunParams, mp_grid, a_latt):
    """integrate_over_k_space (core.f90:419-561) + mep_sum_bands (570-603) +
    normalize_k_int (k_space.f90:81-202): the tensors the NPY writers emit."""
    n_k = int(np.prod(mp_grid))
    norm = abs(unit_vol(a_latt) * float(n_k))
    ef_lst = linspace(params.eF_min, params.eF_max, params.n_eF)
    ef_idx, Ne = fermi_selector(sums["Ne"], n_k, ef_lst, params.valence_bands)
    e = ef_idx - 1
    out = {"eF_idx": ef_idx, "Ne": Ne}
    if params.do_mep:
        nv = params.valence_bands
        ic = sums["niu_ic"][:, :, :nv].sum(axis=2) / norm
        lc = sums["niu_lc"][:, :, :nv].sum(axis=2) / norm
        cs = sums["niu_cs"][:, :, :nv].sum(axis=2) / norm
        out.update(mep_ic=ic, mep_lc=lc, mep_cs=cs, mep_tens=ic + lc + cs,
                   mep_bands=(sums["niu_ic"] + sums["niu_lc"] + sums["niu_cs"])[:, :, :nv] / norm)
    if params.do_kubo:
        ic = sums["kubo_ic"][:, :, e] / norm
        lc = sums["kubo_lc"][:, :, e] / norm
        cs = sums["kubo_cs"][:, :, e] / norm
        out.update(kubo_mep_ic=ic, kubo_mep_lc=lc, kubo_mep_cs=cs, kubo_mep_tens=ic + lc + cs)
    if params.do_ahc:
        out.update(ahc_tens=sums["ahc"][:, :, e] / norm,
                   ahcVELO=sums["ahc_velo"][:, :, :, e] / norm,
                   ohcVELO=sums["ohc"][:, :, :, e] / norm)
    if getattr(params, "do_opt", False):  # core.f90:512-520: opt at eF_idx, photo all E_F
        out.update(opt_Ssymm=sums["opt_s"][:, :, :, e] / norm,
                   opt_Asymm=sums["opt_a"][:, :, :, e] / norm,
                   photoC_2nd=sums["photo"] / norm)
    if getattr(params, "do_gyro", False):  # core.f90:523-531: every E_F
        out.update(gyro_C=sums["gyro_C"] / norm, gyro_D=sums["gyro_D"] / norm,
                   gyro_Dw=sums["gyro_Dw"] / norm)
    return out


def fd_stat_deriv(fd):
    """statistics.f90:160-196 on fd (n_eF, n): 1 / (2 + 2 cosh(fd))."""
    fd = np.asfortranarray(fd, dtype=np.float64)
    out = np.zeros_like(fd, order="F")
    lib().orc_fd_stat_deriv(_p(fd), C.c_int(fd.shape[0]), C.c_int64(fd.shape[1]), _p(out))
    return out


def kubo_opt_tens(e, A, hw, fd, eta):
    """kubo.f90:204-405 -> (opt_symm, opt_asymm), each (3,3,n_hw,n_eF)."""
    e = np.ascontiguousarray(e, dtype=np.float64)
    hw = np.ascontiguousarray(hw, dtype=np.float64)
    fd = np.asfortranarray(fd, dtype=np.float64)
    n, nhw, nef = e.size, hw.size, fd.shape[0]
    os_ = np.zeros((3, 3, nhw, nef), np.complex128, order="F")
    oa = np.zeros_like(os_)
    Ap = None if A is None else np.asfortranarray(A, dtype=np.complex128)
    lib().orc_kubo_opt_tens(C.c_int64(n), _p(e), _p(Ap), C.c_int(nhw), _p(hw), C.c_int(nef),
                            _p(fd), C.c_double(eta), _p(os_), _p(oa))
    return os_, oa


def photo_2nd_cond(e, V, hw, fd, eta):
    """photo.f90:16-86 -> (3,3,3,n_hw,n_eF)."""
    e = np.ascontiguousarray(e, dtype=np.float64)
    hw = np.ascontiguousarray(hw, dtype=np.float64)
    fd = np.asfortranarray(fd, dtype=np.float64)
    V = np.asfortranarray(V, dtype=np.complex128)
    o = np.zeros((3, 3, 3, hw.size, fd.shape[0]), np.complex128, order="F")
    lib().orc_photo_2nd_cond(C.c_int64(e.size), _p(e), _p(V), C.c_int(hw.size), _p(hw),
                             C.c_int(fd.shape[0]), _p(fd), C.c_double(eta), _p(o))
    return o


def gyro_C(V, fdd):
    V = np.asfortranarray(V, dtype=np.complex128)
    fdd = np.asfortranarray(fdd, dtype=np.float64)
    o = np.zeros((3, 3, fdd.shape[0]), np.complex128, order="F")
    lib().orc_gyro_C(C.c_int64(V.shape[1]), _p(V), C.c_int(fdd.shape[0]), _p(fdd), _p(o))
    return o


def gyro_D(e, V, Om, kubo_tol, fdd):
    e = np.ascontiguousarray(e, dtype=np.float64)
    V = np.asfortranarray(V, dtype=np.complex128)
    fdd = np.asfortranarray(fdd, dtype=np.float64)
    Omp = None if Om is None else np.asfortranarray(Om, dtype=np.complex128)
    o = np.zeros((3, 3, fdd.shape[0]), np.complex128, order="F")
    lib().orc_gyro_D(C.c_int64(e.size), _p(e), _p(V), _p(Omp), C.c_double(kubo_tol),
                     C.c_int(fdd.shape[0]), _p(fdd), _p(o))
    return o


def gyro_Dw(e, V, A, hw, fdd):
    e = np.ascontiguousarray(e, dtype=np.float64)
    V = np.asfortranarray(V, dtype=np.complex128)
    hw = np.ascontiguousarray(hw, dtype=np.float64)
    fdd = np.asfortranarray(fdd, dtype=np.float64)
    Ap = None if A is None else np.asfortranarray(A, dtype=np.complex128)
    o = np.zeros((3, 3, hw.size, fdd.shape[0]), np.complex128, order="F")
    lib().orc_gyro_Dw(C.c_int64(e.size), _p(e), _p(V), _p(Ap), C.c_int(hw.size), _p(hw),
                      C.c_int(fdd.shape[0]), _p(fdd), _p(o))
    return o
