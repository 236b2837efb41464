"""Boundary formats (NPY, _hr.dat, input.cfg) and the C-ABI library's
exports -- CPU only, no GPU call."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import FEMN, ROOT


def test_npy_header_bytes_match_fixtures():
    from mepinterp_amd.formats import npy_header
    with open(os.path.join(FEMN, "delta0.900", "ahc_tens.npy"), "rb") as f:
        raw = f.read()
    h = npy_header("<f8", (3, 3))
    assert raw[:len(h)] == h and len(h) == 82
    with open(os.path.join(FEMN, "delta0.900", "ahcVELO_sol09.npy"), "rb") as f:
        raw = f.read()
    h = npy_header("<c16", (3, 3, 121))
    assert raw[:len(h)] == h


@pytest.mark.parametrize("shape,cplx", [((3, 3), False), ((3, 3, 121), True), ((3, 3, 7), False),
                                        ((3, 3, 256, 4), True), ((5,), False)])
def test_npy_writer_roundtrip(tmp_path, shape, cplx):
    from mepinterp_amd.formats import write_npy
    rng = np.random.default_rng(0)
    a = rng.normal(size=shape) + (1j * rng.normal(size=shape) if cplx else 0)
    p = tmp_path / "x.npy"
    write_npy(str(p), a)
    b = np.load(str(p))
    np.testing.assert_array_equal(a, b)
    raw = p.read_bytes()
    hlen = int.from_bytes(raw[8:12], "little")
    assert raw[6:8] == b"\x02\x00" and (12 + hlen) % 16 == 2  # (10 + hlen) % 16 == 0


def test_fixture_written_back_identically(tmp_path):
    """Re-writing the fixture's array gives the fixture's bytes."""
    from mepinterp_amd.formats import write_npy
    src = os.path.join(FEMN, "delta0.900", "ahcVELO_sol09.npy")
    write_npy(str(tmp_path / "v.npy"), np.load(src))
    assert (tmp_path / "v.npy").read_bytes() == open(src, "rb").read()


def test_hr_roundtrip(tmp_path):
    from mepinterp_amd import models
    from mepinterp_amd.formats import read_hr_dat, write_hr_dat
    m = models.make_random(seed=1, n=4, L=1)
    p = tmp_path / "x_hr.dat"
    write_hr_dat(str(p), m["R_frac"], m["H"])
    R, H, deg = read_hr_dat(str(p))
    np.testing.assert_array_equal(R, m["R_frac"])
    big = np.abs(m["H"]) * 27.211385 >= 1e-7
    np.testing.assert_allclose(H[big], m["H"][big], rtol=1e-15)
    assert np.all(deg == 1)


CFG = """# input file for TB model from New J Physics 12, 053032 (2010)
# generated on 10:00AM on February 26, 2019

[jobs]
    plot_bands=False
    debug_mode=False
    R_vect_float=False
    do_write_velo=False
    do_mep=True
    do_kubo=False
    do_ahc=True
    do_opt=False
    do_gyro=False

[unitCell]
    a1= 5.90389713 0.0 3.56440913
    a2= -2.95194856 5.11292489 3.56440913
    a3= -2.95194856 -5.11292489 3.56440913
    a0= 1

[wannBase]
    seed_name= FeMn
    N_wf= 8
    wf_centers_x=  0.0 0.5 0.5 0.0 0.0 0.5 0.5 0.0
    wf_centers_y=  0.0 0.5 0.0 0.5 0.0 0.5 0.0 0.5
    wf_centers_z=  0.0 0.0 0.5 0.5 0.0 0.0 0.5 0.5

[wannInterp]
    use_cart_velo= False
    doGaugeTrafo= True
    mp_grid= 200 200 200

[Fermi]
    N_eF= 1
    eF_min= 0
    eF_max= 0
    Tkelvin= 0
    eta_smearing= 0.1
    kuboTol= 1e-05

[MEP]
    valence_bands= 2
    do_write_mep_bands= True

[Laser]
    N_hw= 121
    hw_min= 0
    hw_max= 6
    laser_phase= 1
"""


def test_input_cfg_reference_layout(tmp_path):
    """The layout tb_input_writer.py:107-188 writes."""
    from mepinterp_amd.formats import read_input_cfg
    p = tmp_path / "input.cfg"
    p.write_text(CFG)
    c = read_input_cfg(str(p))
    assert c["do_mep"] and c["do_ahc"] and not c["do_kubo"]
    assert c["seed_name"] == "FeMn" and c["N_wf"] == 8
    assert c["mp_grid"] == [200, 200, 200]
    assert c["N_hw"] == 121 and abs(c["hw_max"] * 27.211385 - 6.0) < 1e-14
    assert abs(c["eta"] * 27.211385 - 0.1) < 1e-15 and c["kubo_tol"] == 1e-5
    np.testing.assert_array_equal(c["a_latt"][1], [-2.95194856, 5.11292489, 3.56440913])
    np.testing.assert_array_equal(c["wf_centers"][:, 3], [0.0, 0.5, 0.5])


def test_input_cfg_fortran_logicals(tmp_path):
    from mepinterp_amd.formats import read_input_cfg
    text = CFG.replace("do_mep=True", "do_mep=.true.").replace("do_ahc=True", "do_ahc = T") \
        .replace("do_kubo=False", "do_kubo= .false.")
    p = tmp_path / "input.cfg"
    p.write_text(text)
    c = read_input_cfg(str(p))
    assert c["do_mep"] and c["do_ahc"] and not c["do_kubo"]


def _declared_functions():
    with open(os.path.join(ROOT, "include", "mepinterp.h")) as f:
        txt = f.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*\*?\s*(mep_\w+)\s*\(", txt, re.M)))


def test_capi_library_exports_every_declared_symbol():
    from mepinterp_amd import capi
    declared = _declared_functions()
    assert set(declared) == set(capi.EXPORTS)
    lib = capi.load()  # dlopen only; no GPU call
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for name in declared:
        assert name in syms, name


def test_capi_has_no_oracle_dependency():
    """The product library must not link or name the oracle."""
    from mepinterp_amd import capi
    out = subprocess.run(["nm", "-D", capi.LIB_PATH], capture_output=True, text=True).stdout
    assert "orc_" not in out
    ldd = subprocess.run(["ldd", capi.LIB_PATH], capture_output=True, text=True).stdout
    assert "mep_oracle" not in ldd


def test_band_structure_files_roundtrip(tmp_path):
    """kpts (kptsgen.pl), raw/enK.<k> and out/eBands.dat (file_io.f90:57-143,
    451-478): written and read back."""
    from mepinterp_amd import formats as F
    k = np.array([[0.0, 0.25, 0.5], [0.0, 0.0, 0.5], [0.0, 0.1, 0.5]])
    F.write_kpts(str(tmp_path / "kpts"), k)
    assert np.array_equal(F.read_kpts(str(tmp_path / "kpts")), k)
    e = np.array([[-0.1, -0.2, -0.3], [0.2, 0.3, 0.4]])
    F.write_en_binary(str(tmp_path / "enK.0000001"), e[:, 0])
    raw = (tmp_path / "enK.0000001").read_bytes()
    assert len(raw) == 16 and np.array_equal(np.frombuffer(raw, "<f8"), e[:, 0])
    F.write_en_global(str(tmp_path / "eBands.dat"), k, e)
    lines = (tmp_path / "eBands.dat").read_text().splitlines()
    assert len(lines) == 3 + 6 and lines[3].startswith("      1\t")
    # (i7,a) (3f16.8) (a,f18.8)
    assert len(lines[3]) == 7 + 1 + 48 + 1 + 18
    idx, kk, ee = F.read_en_global(str(tmp_path / "eBands.dat"))
    assert list(idx) == [1, 1, 2, 2, 3, 3]
    assert np.allclose(ee, e.T.ravel(), atol=5e-9)


# a two-orbital model in Wannier90's write_tb layout (unit cell in Angstrom,
# energies in eV, positions in Angstrom), typed out by hand
TB_DAT = """ written on 26Feb2019 at 10:00:00
    2.000000    0.000000    0.000000
    1.000000    1.732051    0.000000
    0.000000    0.000000    5.000000
           2
           2
    1    2

    0    0    0
    1    1   -0.50000000E+00    0.00000000E+00
    2    1    0.10000000E+00    0.20000000E+00
    1    2    0.10000000E+00   -0.20000000E+00
    2    2    0.50000000E+00    0.00000000E+00

    1   -1    0
    1    1    0.25000000E+00    0.00000000E+00
    2    1    0.00000000E+00    0.00000000E+00
    1    2    0.30000000E-07    0.00000000E+00
    2    2   -0.25000000E+00    0.10000000E+00

    0    0    0
    1    1    0.10000000E+00    0.00000000E+00    0.20000000E+00    0.00000000E+00    0.30000000E+00    0.00000000E+00
    2    1    0.00000000E+00    0.01000000E+00    0.00000000E+00    0.00000000E+00    0.00000000E+00    0.00000000E+00
    1    2    0.00000000E+00   -0.01000000E+00    0.00000000E+00    0.00000000E+00    0.00000000E+00    0.00000000E+00
    2    2    1.10000000E+00    0.00000000E+00    0.60000000E+00    0.00000000E+00    2.50000000E+00    0.00000000E+00

    1   -1    0
    1    1    0.00000000E+00    0.00000000E+00    0.00000000E+00    0.00000000E+00    0.00000000E+00    0.00000000E+00
    2    1    0.00000000E+00    0.00000000E+00    0.40000000E+00    0.00000000E+00    0.00000000E+00    0.00000000E+00
    1    2    0.00000000E+00    0.00000000E+00    0.00000000E+00    0.00000000E+00    0.00000000E+00    0.00000000E+00
    2    2    0.00000000E+00    0.00000000E+00    0.00000000E+00    0.00000000E+00    0.00000000E+00    0.00000000E+00
"""


def test_tb_dat_reads_wannier90_layout(tmp_path):
    """read_tb_file (file_io.f90:555-658) on a hand-typed file: unit cell
    rows a_i in Angstrom -> Bohr, R left as lattice coefficients, eV ->
    Hartree with the 1e-7 eV clamp, Angstrom -> Bohr positions."""
    from mepinterp_amd.formats import AU_TO_ANGSTROM, AU_TO_EV, read_tb_dat
    p = tmp_path / "m_tb.dat"
    p.write_text(TB_DAT)
    R, H, r, cell = read_tb_dat(str(p))
    assert R.shape == (3, 2) and H.shape == (2, 2, 2) and r.shape == (3, 2, 2, 2)
    np.testing.assert_array_equal(R, [[0, 1], [0, -1], [0, 0]])
    np.testing.assert_allclose(cell[1], np.array([1.0, 1.732051, 0.0]) / AU_TO_ANGSTROM, rtol=1e-15)
    assert H[1, 0, 0] == pytest.approx((0.1 + 0.2j) / AU_TO_EV, rel=1e-15)
    assert H[0, 1, 1] == 0.0  # 3e-8 eV is clamped
    assert H[1, 1, 1] == pytest.approx((-0.25 + 0.1j) / AU_TO_EV, rel=1e-15)
    np.testing.assert_allclose(r[:, 1, 1, 0], np.array([1.1, 0.6, 2.5]) / AU_TO_ANGSTROM, rtol=1e-15)
    assert r[1, 1, 0, 1] == pytest.approx(0.4 / AU_TO_ANGSTROM, rel=1e-15)
    assert r[0, 0, 1, 0] == pytest.approx(-0.01j / AU_TO_ANGSTROM, rel=1e-15)


def test_tb_dat_roundtrip_and_preferred_over_hr(tmp_path):
    """write_tb_dat -> read_tb_dat is exact to the printed digits, and
    read_tb_basis takes _tb.dat over _hr.dat/_r.dat (file_io.f90:817-822)."""
    from mepinterp_amd import models
    from mepinterp_amd.formats import read_tb_basis, write_hr_dat, write_tb_dat
    m = models.make_c1(seed=3)
    cell = np.array([[7.0, 0.0, 0.0], [3.5, 6.06, 0.0], [0.0, 0.0, 9.0]])
    base = str(tmp_path / "x")
    write_tb_dat(base + "_tb.dat", m["R_frac"], m["H"], m["r"], cell)
    write_hr_dat(base + "_hr.dat", m["R_frac"], 2.0 * m["H"])
    R, H, r, c = read_tb_basis(base)
    np.testing.assert_array_equal(R, m["R_frac"])
    big = np.abs(m["H"]) * 27.211385 >= 1e-7
    np.testing.assert_allclose(H[big], m["H"][big], rtol=1e-15)
    np.testing.assert_allclose(r, m["r"], rtol=1e-15, atol=1e-300)
    np.testing.assert_allclose(c, cell, rtol=1e-15)
    os.remove(base + "_tb.dat")
    R2, H2, r2, c2 = read_tb_basis(base)
    assert r2 is None and c2 is None
    np.testing.assert_allclose(H2[big], 2.0 * m["H"][big], rtol=1e-15)


def test_mep_bands_written_only_with_its_flag(tmp_path):
    """write_outputs writes mep_bands.npy only with MEP%do_write_mep_bands
    (core.f90:586-599); the other MEP tensors are written either way."""
    import numpy as np
    from mepinterp_amd import host
    out = dict(mep_bands=np.zeros((3, 3, 2)), mep_tens=np.ones((3, 3)), ef_lst=np.array([0.0]),
               Ne=np.array([2.0]))
    for flag in (False, True):
        d = tmp_path / f"flag{int(flag)}"
        host.write_outputs(out, str(d), dict(n_hw=1, hw_min=0.0, hw_max=0.0,
                                             do_write_mep_bands=flag))
        assert (d / "mep_tens.npy").exists()
        assert (d / "mep_bands.npy").exists() == flag
