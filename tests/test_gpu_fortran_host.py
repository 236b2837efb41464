"""The Fortran drop-in (mepinterp_amd/fortran) on the GPU against the oracle:
the mepInterp executable on an input.cfg run directory, and the per-k
interface (get_wann_interp, kubo_*, kubo_mep_*, mep_niu_*) called with the
reference's argument lists from dropin_check."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import AU_TO_EV, FEMN, FEMN_CENTERS, ROOT, oracle_model, rand_kpts

pytestmark = pytest.mark.gpu

FDIR = os.path.join(ROOT, "mepinterp_amd", "fortran", "_build")


def _need(exe):
    path = os.path.join(FDIR, exe)
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built (make -C mepinterp_amd/fortran)")
    return path


def _cfg(tmp, seed, a_latt, centers, mp, T, n_ef, n_hw, nval, eta_ev=0.1, jobs=("ahc", "mep", "kubo")):
    from mepinterp_amd.formats import write_input_cfg
    write_input_cfg(os.path.join(tmp, "input.cfg"), dict(
        do_ahc="ahc" in jobs, do_mep="mep" in jobs, do_kubo="kubo" in jobs,
        a_latt=a_latt, seed_name=seed, wf_centers=centers, mp_grid=mp, N_eF=n_ef,
        eF_min_eV=-0.2, eF_max_eV=0.2, T_kelvin=T, eta_eV=eta_ev, kubo_tol=1e-5,
        valence_bands=nval, do_write_mep_bands=True, N_hw=n_hw, hw_min_eV=0.0, hw_max_eV=3.0))


def test_mepinterp_executable_femn(tmp_path):
    import oracle as O
    from mepinterp_amd.formats import read_hr_dat, read_input_cfg, read_latt
    exe = _need("mepInterp")
    tmp = str(tmp_path)
    os.makedirs(os.path.join(tmp, "w90files"))
    shutil.copy(os.path.join(FEMN, "delta0.900", "FeMn_hr.dat"), os.path.join(tmp, "w90files"))
    latt = read_latt(os.path.join(FEMN, "delta0.900", "FeMn_d0.9.latt"))
    _cfg(tmp, "FeMn", latt, FEMN_CENTERS, (12, 12, 12), 0.0, 1, 25, 2, jobs=("ahc", "mep"))
    r = subprocess.run([exe], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    cfg = read_input_cfg(os.path.join(tmp, "input.cfg"))
    R, H, _ = read_hr_dat(os.path.join(tmp, "w90files", "FeMn_hr.dat"))
    m = O.Model(R, H, wf_centers=cfg["wf_centers"], a_latt=cfg["a_latt"], kubo_tol=1e-5)
    p = O.RunParams(do_ahc=True, do_mep=True, n_eF=1, eF_min=cfg["eF_min"], eF_max=cfg["eF_max"],
                    n_hw=cfg["N_hw"], hw_min=cfg["hw_min"], hw_max=cfg["hw_max"],
                    T_kelvin=0.0, eta=cfg["eta"], valence_bands=2)
    o = O.finalize(O.accumulate(m, p, O.mp_mesh((12, 12, 12))), p, (12, 12, 12), cfg["a_latt"])
    out = os.path.join(tmp, "out")
    ahc = np.load(os.path.join(out, "ahc_tens.npy"))
    assert abs(ahc[0, 1] - o["ahc_tens"][0, 1]) < 1e-8 * abs(o["ahc_tens"][0, 1])
    velo = np.load(os.path.join(out, "ahcVELO.npy"))
    assert np.max(np.abs(velo - o["ahcVELO"])) < 1e-8 * np.max(np.abs(o["ahcVELO"]))
    hw = np.load(os.path.join(out, "hw_lst.npy"))
    assert hw.shape == (25,) and abs(hw[-1] - 3.0) < 1e-12
    mep = np.load(os.path.join(out, "mep_tens.npy"))
    assert mep.shape == (3, 3) and np.max(np.abs(mep - o["mep_tens"])) < 1e-12
    assert np.load(os.path.join(out, "mep_bands.npy")).shape == (3, 3, 2)
    # header bytes exactly as the reference's files
    with open(os.path.join(out, "ahc_tens.npy"), "rb") as f:
        head = f.read(82)
    with open(os.path.join(FEMN, "delta0.900", "ahc_tens.npy"), "rb") as f:
        assert head == f.read(82)
    assert os.path.exists(os.path.join(out, "occ.dat"))


@pytest.mark.parametrize("kind", ["tb", "pos"])
def test_fortran_per_k_interface(tmp_path, kind):
    import oracle as O
    from mepinterp_amd import models
    from mepinterp_amd.formats import read_hr_dat, read_input_cfg, read_r_dat, write_hr_dat, \
        write_r_dat
    exe = _need("dropin_check")
    tmp = str(tmp_path)
    os.makedirs(os.path.join(tmp, "w90files"))
    if kind == "tb":
        mdl = models.make_random(seed=77, n=10, L=1, tb=True)
        nval, T = 4, 300.0
    else:
        mdl = models.make_c1(seed=5)
        nval, T = 4, 300.0
    write_hr_dat(os.path.join(tmp, "w90files", "m_hr.dat"), mdl["R_frac"], mdl["H"])
    if mdl["r"] is not None:
        write_r_dat(os.path.join(tmp, "w90files", "m_r.dat"), mdl["R_frac"], mdl["r"])
    _cfg(tmp, "m", mdl["a_latt"], mdl["wf_centers"], (2, 2, 2), T, 2, 3, nval, eta_ev=0.05)
    k = rand_kpts(5, seed=11)
    with open(os.path.join(tmp, "kpts.txt"), "w") as f:
        f.write(f"{k.shape[1]}\n")
        for i in range(k.shape[1]):
            f.write(" ".join(repr(float(x)) for x in k[:, i]) + "\n")
    r = subprocess.run([exe], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    cfg = read_input_cfg(os.path.join(tmp, "input.cfg"))
    R, H, _ = read_hr_dat(os.path.join(tmp, "w90files", "m_hr.dat"))
    rr = read_r_dat(os.path.join(tmp, "w90files", "m_r.dat"), R) if mdl["r"] is not None else None
    om = O.Model(R, H, r=rr, wf_centers=cfg["wf_centers"], a_latt=cfg["a_latt"], kubo_tol=1e-5)
    out = os.path.join(tmp, "out")
    n = H.shape[0]
    e_f = np.load(os.path.join(out, "dropin_e.npy"))
    V_f = np.load(os.path.join(out, "dropin_V.npy")).reshape(3, n, n, -1, order="F")
    ef = O.linspace(cfg["eF_min"], cfg["eF_max"], cfg["N_eF"])
    hw = O.linspace(cfg["hw_min"], cfg["hw_max"], cfg["N_hw"])
    got = {key: np.load(os.path.join(out, f"dropin_{key}.npy")) for key in
           ("ahc", "velo", "ohc", "kic", "klc", "kcs", "nic", "nlc", "ncs")}
    for ik in range(k.shape[1]):
        e, V, A, Om = O.get_wann_interp(om, k[:, ik])
        assert np.max(np.abs(e_f[:, ik] - e)) < 1e-10 / AU_TO_EV
        assert np.max(np.abs(V_f[..., ik] - V)) < 1e-8 * np.abs(V).max()
        fd = O.fd_distrib(e, ef, T)
        want = {"ahc": O.kubo_ahc_tens(e, V, fd),
                "velo": O.velo_ahc_tens(e, V, hw, fd, cfg["eta"]).reshape(3, 3, -1, order="F"),
                "ohc": O.kubo_ohc_tens(e, V, hw, fd, cfg["eta"]).reshape(3, 3, -1, order="F"),
                "kic": O.kubo_mep_ic(e, V, fd), "klc": O.kubo_mep_lc(e, V, fd),
                "kcs": O.kubo_mep_cs(A, Om, fd),
                "nic": O.mep_niu_ic(V, e, nval, 1e-5), "nlc": O.mep_niu_lc(V, e, nval, 1e-5),
                "ncs": O.mep_niu_cs(A, Om, nval)}
        for key, w in want.items():
            per = w.shape[2]
            g = got[key][:, :, ik * per:(ik + 1) * per] if key not in ("velo", "ohc") \
                else got[key][..., ik]
            scale = max(np.abs(w).max(), 1e-300)
            assert np.max(np.abs(g - w.reshape(g.shape, order="F"))) <= 1e-8 * scale, (key, ik)
    # the complex-conjugate model (same |H(R)|) is a different model: its
    # eigenvalues are those of H at -k, not a reuse of the resident context
    om_c = O.Model(R, np.conj(H), r=rr, wf_centers=cfg["wf_centers"], a_latt=cfg["a_latt"],
                   kubo_tol=1e-5)
    e_c = np.load(os.path.join(out, "dropin_e_conj.npy"))
    for ik in range(k.shape[1]):
        e, _, _, _ = O.get_wann_interp(om_c, k[:, ik])
        assert np.max(np.abs(e_c[:, ik] - e)) < 1e-10 / AU_TO_EV, ik
    assert np.max(np.abs(e_c - e_f)) > 1e-6
    # a non-uniform hw list reaches the band sums unchanged
    hw_nu = np.load(os.path.join(out, "dropin_hw_nu.npy"))
    assert np.max(np.abs(np.diff(hw_nu, 2))) > 1e-6  # not a linspace
    e, V, _, _ = O.get_wann_interp(om, k[:, 0])
    fd = O.fd_distrib(e, ef, T)
    for key, fn in (("velo_nu", O.velo_ahc_tens), ("ohc_nu", O.kubo_ohc_tens)):
        g = np.load(os.path.join(out, f"dropin_{key}.npy"))
        w = fn(e, V, hw_nu, fd, cfg["eta"]).reshape(g.shape, order="F")
        assert np.max(np.abs(g - w)) <= 1e-8 * np.abs(w).max(), key


def test_mepinterp_band_structure_mode(tmp_path):
    """jobs%plot_bands: ./kpts -> raw/enK.*, out/eBands.dat and (with
    jobs%do_write_velo) out//velo/velo_Vka.* against the oracle."""
    import oracle as O
    from mepinterp_amd import formats as F
    from mepinterp_amd import models
    exe = _need("mepInterp")
    tmp = str(tmp_path)
    os.makedirs(os.path.join(tmp, "w90files"))
    m = models.make_random(seed=31, n=10, L=1, tb=True)
    F.write_hr_dat(os.path.join(tmp, "w90files", "bnd_hr.dat"), m["R_frac"], m["H"])
    k = np.stack([np.linspace(-0.5, 0.5, 17), np.linspace(0.0, 0.3, 17), np.zeros(17)])
    F.write_kpts(os.path.join(tmp, "kpts"), k)
    F.write_input_cfg(os.path.join(tmp, "input.cfg"), dict(
        plot_bands=True, do_write_velo=True, a_latt=m["a_latt"], seed_name="bnd",
        wf_centers=m["wf_centers"], mp_grid=(1, 1, 1), T_kelvin=0.0, valence_bands=5))
    r = subprocess.run([exe], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    R, H, _ = F.read_hr_dat(os.path.join(tmp, "w90files", "bnd_hr.dat"))
    om = O.Model(R, H, wf_centers=m["wf_centers"], a_latt=m["a_latt"])
    idx, kk, ee = F.read_en_global(os.path.join(tmp, "out", "eBands.dat"))
    for q in range(k.shape[1]):
        eo, Vo, _, _ = O.get_wann_interp(om, k[:, q])
        e = F.read_en_binary(os.path.join(tmp, "raw", f"enK.{q + 1:07d}"))
        assert np.max(np.abs(e - eo)) < 1e-10 / AU_TO_EV
        assert np.allclose(ee[idx == q + 1], eo, atol=6e-9)
        V = F.read_velo(os.path.join(tmp, "out", "velo", f"velo_Vka.{q + 1:07d}"), 10)
        assert np.max(np.abs(V - Vo)) < 1e-7


def test_mepinterp_mesh_mode_writes_velocities(tmp_path):
    """jobs%do_write_velo in the k-space integration: out/velo/velo_Vka.<ki>
    for every k-point of the mesh, ki = get_kpt_idx(kix, kiy, kiz)
    (core.f90:258, file_io.f90:146-183), against the oracle's V at
    get_rel_kpt of the same index."""
    import oracle as O
    from mepinterp_amd import formats as F
    from mepinterp_amd import models
    exe = _need("mepInterp")
    tmp = str(tmp_path)
    os.makedirs(os.path.join(tmp, "w90files"))
    m = models.make_random(seed=32, n=8, L=1, tb=True)
    F.write_hr_dat(os.path.join(tmp, "w90files", "mv_hr.dat"), m["R_frac"], m["H"])
    mp = (3, 3, 2)
    F.write_input_cfg(os.path.join(tmp, "input.cfg"), dict(
        do_ahc=True, do_write_velo=True, a_latt=m["a_latt"], seed_name="mv",
        wf_centers=m["wf_centers"], mp_grid=mp, T_kelvin=0.0, valence_bands=4))
    r = subprocess.run([exe], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    R, H, _ = F.read_hr_dat(os.path.join(tmp, "w90files", "mv_hr.dat"))
    om = O.Model(R, H, wf_centers=m["wf_centers"], a_latt=m["a_latt"])
    # get_kpt_idx strides by N2 and N3 (k_space.f90:345): on this non-cubic
    # mesh some indices repeat, and the k-point visited last owns the file,
    # as in the reference's loop
    mesh = O.mp_mesh(mp)
    want = {}
    for kiz in range(1, mp[2] + 1):
        for kiy in range(1, mp[1] + 1):
            for kix in range(1, mp[0] + 1):
                ki = O.kpt_idx(mp, kix, kiy, kiz)
                want[ki] = mesh[:, (kix - 1) + mp[0] * ((kiy - 1) + mp[1] * (kiz - 1))]
    files = sorted(os.listdir(os.path.join(tmp, "out", "velo")))
    assert files == [f"velo_Vka.{ki:07d}" for ki in sorted(want)]
    for ki, kpt in want.items():
        _, Vo, _, _ = O.get_wann_interp(om, kpt)
        V = F.read_velo(os.path.join(tmp, "out", "velo", f"velo_Vka.{ki:07d}"), 8)
        assert np.max(np.abs(V - Vo)) < 1e-7, ki


def test_mepinterp_optical_and_gyrotropic_outputs(tmp_path):
    """jobs%do_opt / jobs%do_gyro: out/opt_Ssymm.npy, opt_Asymm.npy,
    photoC_2nd.npy, gyro_C/D/Dw.npy (file_io.f90:362-398) against the
    oracle on a model with a position operator."""
    import oracle as O
    from mepinterp_amd import models
    from mepinterp_amd.formats import read_hr_dat, read_input_cfg, read_r_dat, write_hr_dat, \
        write_input_cfg, write_r_dat
    exe = _need("mepInterp")
    tmp = str(tmp_path)
    os.makedirs(os.path.join(tmp, "w90files"))
    mdl = models.make_c1(seed=6)
    write_hr_dat(os.path.join(tmp, "w90files", "m_hr.dat"), mdl["R_frac"], mdl["H"])
    write_r_dat(os.path.join(tmp, "w90files", "m_r.dat"), mdl["R_frac"], mdl["r"])
    write_input_cfg(os.path.join(tmp, "input.cfg"), dict(
        do_ahc=True, do_opt=True, do_gyro=True, a_latt=mdl["a_latt"], seed_name="m",
        wf_centers=None, mp_grid=(3, 3, 3), N_eF=2, eF_min_eV=-0.2, eF_max_eV=0.2,
        T_kelvin=300.0, eta_eV=0.05, kubo_tol=1e-5, valence_bands=4, N_hw=3,
        hw_min_eV=0.1, hw_max_eV=0.5))
    r = subprocess.run([exe], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    cfg = read_input_cfg(os.path.join(tmp, "input.cfg"))
    R, H, _ = read_hr_dat(os.path.join(tmp, "w90files", "m_hr.dat"))
    rr = read_r_dat(os.path.join(tmp, "w90files", "m_r.dat"), R)
    om = O.Model(R, H, r=rr, a_latt=cfg["a_latt"], kubo_tol=1e-5)
    p = O.RunParams(do_ahc=True, do_opt=True, do_gyro=True, n_eF=2, eF_min=cfg["eF_min"],
                    eF_max=cfg["eF_max"], n_hw=3, hw_min=cfg["hw_min"], hw_max=cfg["hw_max"],
                    T_kelvin=300.0, eta=cfg["eta"], valence_bands=4)
    o = O.finalize(O.accumulate(om, p, O.mp_mesh((3, 3, 3))), p, (3, 3, 3), cfg["a_latt"])
    out = os.path.join(tmp, "out")
    for key, shape in (("opt_Ssymm", (3, 3, 3)), ("opt_Asymm", (3, 3, 3)),
                       ("photoC_2nd", (3, 3, 3, 3, 2)), ("gyro_C", (3, 3, 2)),
                       ("gyro_D", (3, 3, 2)), ("gyro_Dw", (3, 3, 3, 2))):
        a = np.load(os.path.join(out, key + ".npy"))
        assert a.shape == shape and a.dtype == np.complex128, key
        assert np.max(np.abs(a - o[key])) < 1e-9 * np.max(np.abs(o[key])), key


def test_mepinterp_reads_tb_dat(tmp_path):
    """w90files/<seed>_tb.dat (file_io.f90:555-658, preferred over _hr.dat by
    read_tb_basis :817-822): Hamiltonian, position operator and a non-cubic
    unit cell in one file, with Cartesian velocities, against the oracle on
    the values read back from the same file (R fractional)."""
    import oracle as O
    from mepinterp_amd import models
    from mepinterp_amd.formats import read_tb_dat, write_input_cfg, write_tb_dat
    exe = _need("mepInterp")
    tmp = str(tmp_path)
    os.makedirs(os.path.join(tmp, "w90files"))
    mdl = models.make_c1(seed=8)
    cell = np.array([[6.0, 0.0, 0.0], [-3.0, 5.196152422706632, 0.0], [0.0, 0.0, 8.0]])
    write_tb_dat(os.path.join(tmp, "w90files", "t_tb.dat"), mdl["R_frac"], mdl["H"], mdl["r"], cell)
    write_input_cfg(os.path.join(tmp, "input.cfg"), dict(
        do_ahc=True, do_mep=True, do_kubo=True, a_latt=cell, seed_name="t", wf_centers=None,
        use_cart_velo=True, mp_grid=(4, 4, 4), N_eF=3, eF_min_eV=-0.2, eF_max_eV=0.2,
        T_kelvin=300.0, eta_eV=0.05, kubo_tol=1e-5, valence_bands=4, N_hw=2,
        hw_min_eV=0.0, hw_max_eV=0.5))
    r = subprocess.run([exe], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "_tb.dat" in r.stdout
    R, H, rr, c = read_tb_dat(os.path.join(tmp, "w90files", "t_tb.dat"))
    from mepinterp_amd.formats import read_input_cfg
    cfg = read_input_cfg(os.path.join(tmp, "input.cfg"))
    om = O.Model(R, H, r=rr, a_latt=c, use_cart_velo=True, kubo_tol=1e-5)
    p = O.RunParams(do_ahc=True, do_mep=True, do_kubo=True, n_eF=3, eF_min=cfg["eF_min"],
                    eF_max=cfg["eF_max"], n_hw=2, hw_min=cfg["hw_min"], hw_max=cfg["hw_max"],
                    T_kelvin=300.0, eta=cfg["eta"], valence_bands=4)
    o = O.finalize(O.accumulate(om, p, O.mp_mesh((4, 4, 4))), p, (4, 4, 4), cfg["a_latt"])
    out = os.path.join(tmp, "out")
    for key in ("ahc_tens", "ahcVELO", "mep_tens", "mep_cs", "kubo_mep_tens", "kubo_mep_cs"):
        a = np.load(os.path.join(out, key + ".npy"))
        scale = np.max(np.abs(o[key]))
        assert scale > 0, key
        assert np.max(np.abs(a - o[key])) < 1e-9 * scale, key
    # the Python host reads the same file the same way
    from mepinterp_amd import host
    tmp2 = os.path.join(tmp, "py")
    os.makedirs(os.path.join(tmp2, "w90files"))
    shutil.copy(os.path.join(tmp, "w90files", "t_tb.dat"), os.path.join(tmp2, "w90files"))
    shutil.copy(os.path.join(tmp, "input.cfg"), tmp2)
    py = host.core_worker(tmp2)
    for key in ("ahc_tens", "mep_tens", "kubo_mep_tens"):
        assert np.max(np.abs(py[key] - o[key])) < 1e-9 * np.max(np.abs(o[key])), key
