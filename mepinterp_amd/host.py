"""Host-side driver of the k-space path, mirroring the reference's core_worker.

This is the Python counterpart of the Fortran host (mepinterp_amd/fortran):
same steps, same names, same outputs.

* ``fermi_selector``        core.f90:317-415
* ``finalize``              integrate_over_k_space + mep_sum_bands +
                            normalize_k_int (core.f90:419-603,
                            k_space.f90:81-202)
* ``write_outputs``         the NPY writers of file_io.f90:186-359
* ``core_worker``           core.f90:79-298: input.cfg -> w90files/<seed>_tb.dat or _hr.dat (+ _r.dat)
                            -> device k-loop -> out/*.npy
* ``band_worker``           band_calc.f90:30-121: ./kpts -> raw/enK.*,
                            out/eBands.dat (and out//velo/velo_Vka.*)
* ``get_wann_interp`` ...   per-k functions with the reference's argument
                            order, on the GPU (nk = 1 batches)
"""
from __future__ import annotations

import os

import numpy as np

from . import capi
from .formats import (AU_TO_EV, read_input_cfg, read_kpts, read_tb_basis,
                      write_en_binary, write_en_global, write_npy)


def unit_vol(a_latt):
    """(a1 x a2) . a3 with a_latt rows a_i (k_space.f90:278; same as columns)."""
    a = np.asarray(a_latt, dtype=np.float64)
    return float(np.dot(np.cross(a[0], a[1]), a[2]))


def linspace(lo, hi, n):
    """get_linspace (matrix_math.f90:74-100)."""
    if n > 1:
        d = (hi - lo) / float(n - 1)
        return np.array([lo + float(i) * d for i in range(n)])
    return np.array([lo])


def fermi_selector(Ne_sum, n_k, ef_lst, n_el_sys):
    """First E_F minimising |N_el(E_F) - valence_bands| (1-based index)."""
    Ne = np.asarray(Ne_sum, dtype=np.float64) / n_k
    opt_idx, opt_err = 1, abs(Ne[0] - float(n_el_sys))
    for i in range(Ne.size):
        err = abs(Ne[i] - float(n_el_sys))
        if err < opt_err:
            opt_idx, opt_err = i + 1, err
    return opt_idx, Ne


def finalize(sums, p: dict, mp_grid, a_latt, n_k=None):
    """Raw k-sums -> the normalised tensors the reference writes.

    ``p`` holds do_ahc/do_kubo/do_mep, n_eF, eF_min, eF_max, n_hw, hw_min,
    hw_max (Hartree) and valence_bands.  The k-integral is divided by
    |unit_vol * N1 N2 N3| as normalize_k_int does (k_space.f90:81-202)."""
    nk_mesh = int(np.prod(mp_grid))
    norm = abs(unit_vol(a_latt) * float(nk_mesh))
    n_k = nk_mesh if n_k is None else n_k
    ef = linspace(p.get("eF_min", 0.0), p.get("eF_max", 0.0), p.get("n_eF", 1))
    ef_idx, Ne = fermi_selector(sums["Ne"], n_k, ef, p.get("valence_bands", 0))
    e = ef_idx - 1
    out = {"eF_idx": ef_idx, "Ne": Ne, "ef_lst": ef}
    if p.get("do_mep"):
        nv = p["valence_bands"]
        ic = sums["niu_ic"][:, :, :nv].sum(axis=2) / norm
        lc = sums["niu_lc"][:, :, :nv].sum(axis=2) / norm
        cs = sums["niu_cs"][:, :, :nv].sum(axis=2) / norm
        out.update(mep_ic=ic, mep_lc=lc, mep_cs=cs, mep_tens=ic + lc + cs,
                   mep_bands=(sums["niu_ic"] + sums["niu_lc"] + sums["niu_cs"])[:, :, :nv] / norm)
    if p.get("do_kubo"):
        ic = sums["kubo_ic"][:, :, e] / norm
        lc = sums["kubo_lc"][:, :, e] / norm
        cs = sums["kubo_cs"][:, :, e] / norm
        out.update(kubo_mep_ic=ic, kubo_mep_lc=lc, kubo_mep_cs=cs, kubo_mep_tens=ic + lc + cs)
    if p.get("do_ahc"):
        out.update(ahc_tens=sums["ahc"][:, :, e] / norm,
                   ahcVELO=sums["ahc_velo"][:, :, :, e] / norm,
                   ohcVELO=sums["ohc"][:, :, :, e] / norm)
    if p.get("do_opt"):  # core.f90:512-520: opt tensors at eF_idx, photo at every E_F
        out.update(opt_Ssymm=sums["opt_s"][:, :, :, e] / norm,
                   opt_Asymm=sums["opt_a"][:, :, :, e] / norm,
                   photoC_2nd=sums["photo"] / norm)
    if p.get("do_gyro"):  # core.f90:523-531: every E_F
        out.update(gyro_C=sums["gyro_C"] / norm, gyro_D=sums["gyro_D"] / norm,
                   gyro_Dw=sums["gyro_Dw"] / norm)
    return out


def hw_list_ev(n_hw, hw_min, hw_max):
    """write_hw_list (file_io.f90:186-214); n_hw = 1 keeps hw_min in Hartree
    as the reference does."""
    if n_hw > 1:
        d = (hw_max - hw_min) / float(n_hw - 1)
        return np.array([(hw_min + float(i) * d) * AU_TO_EV for i in range(n_hw)])
    return np.array([hw_min])


def write_outputs(out: dict, out_dir: str, p: dict):
    """NPY files with the reference's names (file_io.f90:186-359)."""
    os.makedirs(out_dir, exist_ok=True)
    write_npy(os.path.join(out_dir, "hw_lst.npy"),
              hw_list_ev(p.get("n_hw", 1), p.get("hw_min", 0.0), p.get("hw_max", 0.0)))
    for key in ("mep_bands", "mep_ic", "mep_lc", "mep_cs", "mep_tens", "kubo_mep_ic",
                "kubo_mep_lc", "kubo_mep_cs", "kubo_mep_tens", "ahc_tens"):
        # mep_bands.npy only with MEP%do_write_mep_bands (core.f90:586-599)
        if key == "mep_bands" and not p.get("do_write_mep_bands", False):
            continue
        if key in out:
            write_npy(os.path.join(out_dir, key + ".npy"), out[key])
    if "ahcVELO" in out:
        write_npy(os.path.join(out_dir, "ahcVELO.npy"), out["ahcVELO"])
        write_npy(os.path.join(out_dir, "ohcVELO.npy"), out["ohcVELO"])
    # write_opt_tensors / write_gyro_tensors (file_io.f90:362-398)
    for key in ("opt_Ssymm", "opt_Asymm", "photoC_2nd", "gyro_C", "gyro_D", "gyro_Dw"):
        if key in out:
            write_npy(os.path.join(out_dir, key + ".npy"), out[key])
    with open(os.path.join(out_dir, "occ.dat"), "w") as f:
        for e, n in zip(out["ef_lst"], out["Ne"]):
            f.write(f"{e * AU_TO_EV:16.8f}\t{n:16.8f}\n")


def params_from_cfg(cfg: dict) -> dict:
    return dict(do_ahc=cfg["do_ahc"], do_kubo=cfg["do_kubo"], do_mep=cfg["do_mep"],
                do_opt=cfg.get("do_opt", False), do_gyro=cfg.get("do_gyro", False),
                n_eF=cfg["N_eF"], eF_min=cfg["eF_min"], eF_max=cfg["eF_max"],
                n_hw=cfg["N_hw"], hw_min=cfg["hw_min"], hw_max=cfg["hw_max"],
                T_kelvin=cfg["T_kelvin"], eta=cfg["eta"], valence_bands=cfg["valence_bands"],
                do_write_mep_bands=cfg.get("do_write_mep_bands", False))


def run_params(p: dict) -> capi.RunParams:
    jobs = (capi.MEP_JOB_AHC if p.get("do_ahc") else 0) | \
        (capi.MEP_JOB_KUBO if p.get("do_kubo") else 0) | (capi.MEP_JOB_MEP if p.get("do_mep") else 0) | \
        (capi.MEP_JOB_OPT if p.get("do_opt") else 0) | (capi.MEP_JOB_GYRO if p.get("do_gyro") else 0)
    return capi.Engine.params(jobs=jobs, n_eF=p.get("n_eF", 1), eF_min=p.get("eF_min", 0.0),
                              eF_max=p.get("eF_max", 0.0), n_hw=p.get("n_hw", 1),
                              hw_min=p.get("hw_min", 0.0), hw_max=p.get("hw_max", 0.0),
                              T_kelvin=p.get("T_kelvin", 0.0), eta=p.get("eta", 0.0),
                              valence_bands=max(1, p.get("valence_bands", 1)))


def _model_latt(cfg: dict, cell):
    """Lattice handed to the interpolation: the _tb.dat unit cell overrides
    the input.cfg lattice (file_io.f90:578-579); the k-sum normalisation keeps
    the input.cfg cell volume, as the reference's unit_vol is set from
    input.cfg only (input_paras.f90:281)."""
    if cell is None:
        return cfg["a_latt"]
    if np.max(np.abs(np.asarray(cell) - np.asarray(cfg["a_latt"]))) > 1e-6 * np.max(np.abs(cell)):
        print("[read_tb_basis]: WARNING the _tb.dat unit cell differs from the input.cfg lattice")
    return cell


def core_worker(run_dir: str = "."):
    """core.f90:79-298 on the GPU: read ./input.cfg and w90files/<seed>_hr.dat,
    integrate over the MP mesh, write out/*.npy and out/occ.dat."""
    cfg = read_input_cfg(os.path.join(run_dir, "input.cfg"))
    base = os.path.join(run_dir, "w90files", cfg["seed_name"])
    R, H, r, cell = read_tb_basis(base, R_float=cfg["R_vect_float"])
    p = params_from_cfg(cfg)
    with capi.Engine(R, H, r=r, wf_centers=cfg["wf_centers"], a_latt=_model_latt(cfg, cell),
                     kubo_tol=cfg["kubo_tol"], gauge_trafo=cfg["do_gauge_trafo"],
                     cart_velo=cfg["use_cart_velo"]) as eng:
        sums = eng.run_mesh(cfg["mp_grid"], run_params(p))
    out = finalize(sums, p, cfg["mp_grid"], cfg["a_latt"])
    write_outputs(out, os.path.join(run_dir, "out"), p)
    return out


def band_worker(run_dir: str = "."):
    """band_calc.f90:30-121 on the GPU: interpolate the k-points of ./kpts in
    one batch, write raw/enK.<k> and out/eBands.dat."""
    cfg = read_input_cfg(os.path.join(run_dir, "input.cfg"))
    base = os.path.join(run_dir, "w90files", cfg["seed_name"])
    R, H, r, cell = read_tb_basis(base, R_float=cfg["R_vect_float"])
    k = read_kpts(os.path.join(run_dir, "kpts"))
    with capi.Engine(R, H, r=r, wf_centers=cfg["wf_centers"], a_latt=_model_latt(cfg, cell),
                     kubo_tol=cfg["kubo_tol"], gauge_trafo=cfg["do_gauge_trafo"],
                     cart_velo=cfg["use_cart_velo"]) as eng:
        e, V, _, _ = eng.interp(k, velo=r is not None)
    os.makedirs(os.path.join(run_dir, "raw"), exist_ok=True)
    os.makedirs(os.path.join(run_dir, "out"), exist_ok=True)
    for q in range(k.shape[1]):
        write_en_binary(os.path.join(run_dir, "raw", f"enK.{q + 1:07d}"), e[:, q])
    write_en_global(os.path.join(run_dir, "out", "eBands.dat"), k, e)
    return k, e


# ---------------------------------------------------------------------------
# per-k functions with the reference's argument order (GPU, nk = 1)
# ---------------------------------------------------------------------------

def get_wann_interp(engine: capi.Engine, kpt):
    """wann_interp.f90:56-108 -> (e_k, V_ka, A_ka, Om_kab) for one k."""
    e, V, A, Om = engine.interp(np.asarray(kpt, dtype=np.float64).reshape(3, 1))
    return (e[:, 0], V[..., 0], None if A is None else A[..., 0],
            None if Om is None else Om[..., 0])
