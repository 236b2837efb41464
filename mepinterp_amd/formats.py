"""On-disk formats of MEPinterp's k-space path, host side.

Readers and writers for the three formats the drop-in keeps:

* ``<seed>_hr.dat``  -- Wannier90 real-space Hamiltonian (file_io.f90:663-746)
* ``input.cfg``      -- config_fortran INI-style input (input_paras.f90:79-293)
* ``*.npy``          -- NPY v2.0, Fortran order, as NPY-for-Fortran writes it
                        (file_io.f90:186-398; header bytes as in the fixtures)

The Fortran host (mepinterp_amd/fortran) carries its own implementation of
the same formats; this module serves Python callers, tests and bench.py.
"""
from __future__ import annotations

import os
import re
import struct

import numpy as np

AU_TO_EV = 27.211385          # constants.f90:19
AU_TO_ANGSTROM = 0.52917721092  # constants.f90:18


# --------------------------------------------------------------------------
# _hr.dat
# --------------------------------------------------------------------------

def read_hr_dat(path: str, R_float: bool = False):
    """Read ``<seed>_hr.dat`` like ``read_hr_file`` (file_io.f90:663-746).

    Returns ``(R_frac (3,nR) f8, H (n,n,nR) c16 in Hartree, deg (nR,) int)``.
    Entries with |x| < 1e-7 eV are clamped to 0 before the eV->Hartree
    division, as the reference does (file_io.f90:730-731, 741).  deg(R) is
    returned but, as in the reference, not applied.
    """
    with open(path) as f:
        f.readline()  # comment
        n_wf = int(f.readline().split()[0])
        n_R = int(f.readline().split()[0])
        deg = []
        while len(deg) < n_R:
            deg.extend(int(t) for t in f.readline().split())
        if len(deg) != n_R:
            raise ValueError("[read_hr_dat]: issue reading Wigner Seitz degeneracy")
        body = np.loadtxt(f, dtype=np.float64, ndmin=2)
    if body.shape[0] != n_R * n_wf * n_wf or body.shape[1] < 7:
        raise ValueError(f"[read_hr_dat]: expected {n_R * n_wf * n_wf} rows of 7 columns, "
                         f"got {body.shape}")
    R_frac = np.zeros((3, n_R), dtype=np.float64, order="F")
    H = np.zeros((n_wf, n_wf, n_R), dtype=np.complex128, order="F")
    rows = body.reshape(n_R, n_wf * n_wf, body.shape[1])
    for s in range(n_R):
        blk = rows[s]
        R = blk[0, 0:3] if R_float else np.trunc(blk[0, 0:3])
        R_frac[:, s] = R
        m = blk[:, 3].astype(np.int64) - 1
        n = blk[:, 4].astype(np.int64) - 1
        re_ = blk[:, 5].copy()
        im_ = blk[:, 6].copy()
        re_[np.abs(re_) < 1e-7] = 0.0
        im_[np.abs(im_) < 1e-7] = 0.0
        H[m, n, s] = re_ + 1j * im_
    H /= AU_TO_EV
    return R_frac, H, np.asarray(deg, dtype=np.int64)


def write_hr_dat(path: str, R_frac, H_hartree, deg=None, comment="written by mepinterp_amd"):
    """Write ``_hr.dat`` (energies converted back to eV), the layout the
    reference reads: R slowest, 4th column m slower than 5th column n."""
    R_frac = np.asarray(R_frac)
    H = np.asarray(H_hartree) * AU_TO_EV
    n_wf, _, n_R = H.shape
    if deg is None:
        deg = np.ones(n_R, dtype=np.int64)
    with open(path, "w") as f:
        f.write(comment.strip() + "\n")
        f.write(f"{n_wf:11d}\n{n_R:11d}\n")
        for i in range(0, n_R, 15):
            f.write("".join(f"{int(d):5d}" for d in deg[i:i + 15]) + "\n")
        for s in range(n_R):
            R = np.rint(R_frac[:, s]).astype(np.int64)
            for m in range(n_wf):
                for n in range(n_wf):
                    z = H[m, n, s]
                    f.write(f"{R[0]:5d}{R[1]:5d}{R[2]:5d}{m + 1:5d}{n + 1:5d}"
                            f"  {z.real: .17e}  {z.imag: .17e}\n")


def write_r_dat(path: str, R_frac, r_bohr, comment="written by mepinterp_amd"):
    """Position operator file read by the Fortran host's read_r_file: a
    comment line, n_wf, then rows ``R1 R2 R3 m n Re(x) Im(x) Re(y) Im(y)
    Re(z) Im(z)`` in Angstrom, R in the order of the _hr.dat file."""
    r = np.asarray(r_bohr) * AU_TO_ANGSTROM
    n = r.shape[1]
    with open(path, "w") as f:
        f.write(comment.strip() + "\n")
        f.write(f"{n:11d}\n")
        for s in range(r.shape[3]):
            R = np.rint(np.asarray(R_frac)[:, s]).astype(np.int64)
            for m in range(n):
                for nn in range(n):
                    v = r[:, m, nn, s]
                    f.write(f"{R[0]:5d}{R[1]:5d}{R[2]:5d}{m + 1:5d}{nn + 1:5d}" +
                            "".join(f"  {c.real: .17e}  {c.imag: .17e}" for c in v) + "\n")


def read_r_dat(path: str, R_frac, R_float: bool = False):
    """Read the file write_r_dat writes; returns r (3,n,n,nR) in Bohr."""
    with open(path) as f:
        f.readline()
        n = int(f.readline().split()[0])
        body = np.loadtxt(f, dtype=np.float64, ndmin=2)
    nR = np.asarray(R_frac).shape[1]
    if body.shape[0] != nR * n * n:
        raise ValueError("[read_r_dat]: wrong number of rows")
    r = np.zeros((3, n, n, nR), dtype=np.complex128, order="F")
    rows = body.reshape(nR, n * n, body.shape[1])
    for s in range(nR):
        blk = rows[s]
        R = blk[0, 0:3] if R_float else np.trunc(blk[0, 0:3])
        if np.max(np.abs(R - np.asarray(R_frac)[:, s])) > 1e-14:
            raise ValueError("[read_r_dat]: different R order in _hr.dat and _r.dat")
        m = blk[:, 3].astype(np.int64) - 1
        nn = blk[:, 4].astype(np.int64) - 1
        for a in range(3):
            r[a, m, nn, s] = blk[:, 5 + 2 * a] + 1j * blk[:, 6 + 2 * a]
    return r / AU_TO_ANGSTROM


# --------------------------------------------------------------------------
# _tb.dat  (Wannier90 write_tb: Hamiltonian and position operator together)
# --------------------------------------------------------------------------

def read_tb_dat(path: str, R_float: bool = False):
    """Read ``<seed>_tb.dat`` like ``read_tb_file`` (file_io.f90:555-658).

    Layout: a comment line; the unit cell in Angstrom, one lattice vector a_i
    per line; n_wf; n_R; the degeneracy list (15 per line, skipped); per R a
    blank line, the R line and n_wf**2 lines ``m n Re Im`` in eV; then per R a
    blank line, the R line and n_wf**2 lines ``m n Re(x) Im(x) Re(y) Im(y)
    Re(z) Im(z)`` in Angstrom.

    Returns ``(R_frac (3,nR), H (n,n,nR) Hartree, r (3,n,n,nR) Bohr,
    a_latt (3,3) Bohr with rows a_i)``.  Hamiltonian entries below 1e-7 eV
    are clamped to 0 as the reference does (:612-613).  R stays fractional:
    the reference stores matmul(unit_cell, R) (:605) and then uses it as the
    fractional R of the Fourier phase (wann_interp.f90:213-219), which breaks
    the interpolation of every model whose unit cell is not the identity;
    this reader keeps the lattice-vector coefficients the file holds.
    """
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines()]
    pos = 1

    def nxt():
        nonlocal pos
        while pos < len(lines) and not lines[pos].strip():
            pos += 1
        if pos >= len(lines):
            raise ValueError("[read_tb_dat]: unexpected end of file")
        ln = lines[pos]
        pos += 1
        return ln.split()

    cell = np.array([[float(t) for t in nxt()[:3]] for _ in range(3)], dtype=np.float64)
    n_wf = int(nxt()[0])
    n_R = int(nxt()[0])
    deg = []
    while len(deg) < n_R:
        deg.extend(int(t) for t in nxt())
    R_frac = np.zeros((3, n_R), dtype=np.float64, order="F")
    H = np.zeros((n_wf, n_wf, n_R), dtype=np.complex128, order="F")
    r = np.zeros((3, n_wf, n_wf, n_R), dtype=np.complex128, order="F")

    def read_R():
        t = nxt()
        v = np.array([float(x) for x in t[:3]])
        return v if R_float else np.trunc(v)

    for s in range(n_R):
        R_frac[:, s] = read_R()
        for _ in range(n_wf * n_wf):
            t = nxt()
            m, n = int(t[0]) - 1, int(t[1]) - 1
            re_, im_ = float(t[2]), float(t[3])
            re_ = 0.0 if abs(re_) < 1e-7 else re_
            im_ = 0.0 if abs(im_) < 1e-7 else im_
            H[m, n, s] = re_ + 1j * im_
    for s in range(n_R):
        R = read_R()
        if np.linalg.norm(R - R_frac[:, s]) > 1e-8:
            print("[read_tb_dat]: WARNING rHopp has different sc order then tHopp")
        for _ in range(n_wf * n_wf):
            t = nxt()
            m, n = int(t[0]) - 1, int(t[1]) - 1
            v = [float(x) for x in t[2:8]]
            for a in range(3):
                r[a, m, n, s] = v[2 * a] + 1j * v[2 * a + 1]
    return R_frac, H / AU_TO_EV, r / AU_TO_ANGSTROM, cell / AU_TO_ANGSTROM


def write_tb_dat(path: str, R_frac, H_hartree, r_bohr, a_latt_bohr, deg=None,
                 comment="written by mepinterp_amd"):
    """Write the ``_tb.dat`` layout read_tb_dat reads (Wannier90 write_tb):
    unit cell and positions in Angstrom, energies in eV, integer R."""
    R_frac = np.asarray(R_frac)
    H = np.asarray(H_hartree) * AU_TO_EV
    r = np.asarray(r_bohr) * AU_TO_ANGSTROM
    cell = np.asarray(a_latt_bohr) * AU_TO_ANGSTROM
    n_wf, _, n_R = H.shape
    if deg is None:
        deg = np.ones(n_R, dtype=np.int64)
    with open(path, "w") as f:
        f.write(comment.strip() + "\n")
        for i in range(3):
            f.write("".join(f"{c: 24.16e}" for c in cell[i]) + "\n")
        f.write(f"{n_wf:11d}\n{n_R:11d}\n")
        for i in range(0, n_R, 15):
            f.write("".join(f"{int(d):5d}" for d in deg[i:i + 15]) + "\n")
        for s in range(n_R):
            R = np.rint(R_frac[:, s]).astype(np.int64)
            f.write(f"\n{R[0]:5d}{R[1]:5d}{R[2]:5d}\n")
            for n in range(n_wf):
                for m in range(n_wf):
                    z = H[m, n, s]
                    f.write(f"{m + 1:5d}{n + 1:5d}  {z.real: .17e}  {z.imag: .17e}\n")
        for s in range(n_R):
            R = np.rint(R_frac[:, s]).astype(np.int64)
            f.write(f"\n{R[0]:5d}{R[1]:5d}{R[2]:5d}\n")
            for n in range(n_wf):
                for m in range(n_wf):
                    v = r[:, m, n, s]
                    f.write(f"{m + 1:5d}{n + 1:5d}" +
                            "".join(f"  {c.real: .17e}  {c.imag: .17e}" for c in v) + "\n")


def read_tb_basis(base: str, R_float: bool = False):
    """read_tb_basis (file_io.f90:800-860): ``<base>_tb.dat`` when present
    (Hamiltonian, position operator and unit cell), else ``<base>_hr.dat``
    plus an optional ``<base>_r.dat``.  Returns ``(R_frac, H, r or None,
    a_latt or None)``; a_latt is the _tb.dat unit cell that overrides the
    input.cfg lattice (file_io.f90:578-579)."""
    if os.path.exists(base + "_tb.dat"):
        R, H, r, cell = read_tb_dat(base + "_tb.dat", R_float=R_float)
        return R, H, r, cell
    if not os.path.exists(base + "_hr.dat"):
        raise FileNotFoundError(f'[read_tb_basis]: "{base}_hr.dat" does not exist')
    R, H, _ = read_hr_dat(base + "_hr.dat", R_float=R_float)
    r = read_r_dat(base + "_r.dat", R, R_float) if os.path.exists(base + "_r.dat") else None
    return R, H, r, None


def read_latt(path: str):
    """Read a ``.latt`` file (comment line + 3 rows a_i in Bohr)."""
    rows = []
    with open(path) as f:
        for line in f:
            if line.strip().startswith("#") or not line.strip():
                continue
            rows.append([float(t) for t in line.split()[:3]])
    return np.array(rows[:3], dtype=np.float64)


# --------------------------------------------------------------------------
# input.cfg (config_fortran grammar)
# --------------------------------------------------------------------------

_TRUE = {"t", "true", ".true.", "1", "y", "yes"}
_FALSE = {"f", "false", ".false.", "0", "n", "no"}


def parse_cfg_text(text: str) -> dict:
    """Parse config_fortran text into ``{"section%key": [tokens...]}``."""
    out = {}
    section = ""
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].split(";", 1)[0].strip()
        if not line:
            continue
        m = re.match(r"^\[(.+)\]$", line)
        if m:
            section = m.group(1).strip()
            continue
        if "=" not in line:
            continue
        key, val = line.split("=", 1)
        key = key.strip()
        full = f"{section}%{key}" if section and "%" not in key else key
        out[full] = val.replace(",", " ").split()
    return out


def _as_bool(tok):
    t = tok.strip().strip("'\"").lower()
    if t in _TRUE:
        return True
    if t in _FALSE:
        return False
    raise ValueError(f"not a logical: {tok!r}")


def read_input_cfg(path: str = "./input.cfg") -> dict:
    """Read ``input.cfg`` and apply ``init_parameters``' unit handling
    (input_paras.f90:106-181): energies eV -> Hartree, a_latt rows = a_i*a0,
    N_hw = max(1, N_hw), i_eta_smr = i*eta."""
    with open(path) as f:
        kv = parse_cfg_text(f.read())

    def g(key, conv, n=None, default=None):
        if key not in kv:
            if default is not None:
                return default
            raise KeyError(f"input.cfg: missing {key}")
        toks = kv[key]
        vals = [conv(t) for t in toks]
        return vals if n is not None else vals[0]

    p = {}
    for k in ("plot_bands", "debug_mode", "R_vect_float", "do_write_velo", "do_mep",
              "do_kubo", "do_ahc", "do_opt", "do_gyro"):
        p[k] = g(f"jobs%{k}", _as_bool, default=False)
    a1 = g("unitCell%a1", float, 3)
    a2 = g("unitCell%a2", float, 3)
    a3 = g("unitCell%a3", float, 3)
    a0 = g("unitCell%a0", float, default=1.0)
    p["a_latt"] = a0 * np.array([a1, a2, a3], dtype=np.float64)  # a_latt(i,:) = a_i
    p["seed_name"] = g("wannBase%seed_name", str).strip("'\"")
    p["N_wf"] = g("wannBase%N_wf", int, default=0)
    if p["N_wf"] > 0:
        wx = g("wannBase%wf_centers_x", float, 1)
        wy = g("wannBase%wf_centers_y", float, 1)
        wz = g("wannBase%wf_centers_z", float, 1)
        p["wf_centers"] = np.array([wx[:p["N_wf"]], wy[:p["N_wf"]], wz[:p["N_wf"]]],
                                   dtype=np.float64)
    else:
        p["wf_centers"] = None
    p["use_cart_velo"] = g("wannInterp%use_cart_velo", _as_bool, default=False)
    p["do_gauge_trafo"] = g("wannInterp%doGaugeTrafo", _as_bool, default=True)
    p["mp_grid"] = [int(x) for x in g("wannInterp%mp_grid", int, 3)]
    p["N_eF"] = g("Fermi%N_eF", int, default=1)
    p["eF_min"] = g("Fermi%eF_min", float, default=0.0) / AU_TO_EV
    p["eF_max"] = g("Fermi%eF_max", float, default=0.0) / AU_TO_EV
    p["T_kelvin"] = g("Fermi%Tkelvin", float, default=0.0)
    p["eta"] = g("Fermi%eta_smearing", float, default=0.0) / AU_TO_EV
    p["kubo_tol"] = g("Fermi%kuboTol", float, default=1e-5)
    p["valence_bands"] = g("MEP%valence_bands", int, default=0)
    p["do_write_mep_bands"] = g("MEP%do_write_mep_bands", _as_bool, default=False)
    p["N_hw"] = max(1, g("Laser%N_hw", int, default=1))
    p["hw_min"] = g("Laser%hw_min", float, default=0.0) / AU_TO_EV
    p["hw_max"] = g("Laser%hw_max", float, default=0.0) / AU_TO_EV
    return p


def write_input_cfg(path: str, p: dict):
    """Write ``input.cfg`` in the layout of tb_input_writer.py:107-188
    (energies given in eV in ``p``)."""
    def b(x):
        return "T" if x else "F"

    def arr(x):
        return " ".join(repr(float(v)) for v in x)

    wf = p.get("wf_centers")
    lines = ["# mepinterp_amd input", "", "[jobs]"]
    for k in ("plot_bands", "debug_mode", "R_vect_float", "do_write_velo", "do_mep",
              "do_kubo", "do_ahc", "do_opt", "do_gyro"):
        lines.append(f"    {k}= {b(p.get(k, False))}")
    a = np.asarray(p["a_latt"])
    lines += ["", "[unitCell]", f"    a1= {arr(a[0])}", f"    a2= {arr(a[1])}",
              f"    a3= {arr(a[2])}", f"    a0= {float(p.get('a0', 1.0))!r}", "", "[wannBase]",
              f"    seed_name= {p['seed_name']}",
              f"    N_wf= {0 if wf is None else np.asarray(wf).shape[1]}"]
    if wf is not None:
        wf = np.asarray(wf)
        lines += [f"    wf_centers_x= {arr(wf[0])}", f"    wf_centers_y= {arr(wf[1])}",
                  f"    wf_centers_z= {arr(wf[2])}"]
    mp = p["mp_grid"]
    lines += ["", "[wannInterp]", f"    use_cart_velo= {b(p.get('use_cart_velo', False))}",
              f"    doGaugeTrafo= {b(p.get('do_gauge_trafo', True))}",
              f"    mp_grid= {mp[0]} {mp[1]} {mp[2]}", "", "[Fermi]",
              f"    N_eF= {int(p.get('N_eF', 1))}", f"    eF_min= {float(p.get('eF_min_eV', 0.0))!r}",
              f"    eF_max= {float(p.get('eF_max_eV', 0.0))!r}",
              f"    Tkelvin= {float(p.get('T_kelvin', 0.0))!r}",
              f"    eta_smearing= {float(p.get('eta_eV', 0.0))!r}",
              f"    kuboTol= {float(p.get('kubo_tol', 1e-5))!r}", "", "[MEP]",
              f"    valence_bands= {int(p.get('valence_bands', 0))}",
              f"    do_write_mep_bands= {b(p.get('do_write_mep_bands', True))}", "", "[Laser]",
              f"    N_hw= {int(p.get('N_hw', 1))}", f"    hw_min= {float(p.get('hw_min_eV', 0.0))!r}",
              f"    hw_max= {float(p.get('hw_max_eV', 0.0))!r}", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))


# --------------------------------------------------------------------------
# band-structure mode files (file_io.f90:57-143, 146-185, 451-478)
# --------------------------------------------------------------------------

def read_kpts(path: str):
    """./kpts as written by kptsgen.pl: first line "num_kpts <x>", then one
    line "kx ky kz weight" per k-point (read_kptsgen_pl_file) -> (3, nk)."""
    with open(path) as f:
        first = f.readline().split()
        nk = int(first[0])
        k = np.zeros((3, nk))
        for i in range(nk):
            v = f.readline().split()
            k[:, i] = [float(x) for x in v[:3]]
    return k


def write_kpts(path: str, kpts):
    k = np.asarray(kpts, dtype=np.float64)
    with open(path, "w") as f:
        f.write(f"{k.shape[1]} 1.0\n")
        for i in range(k.shape[1]):
            f.write(f"{float(k[0, i])!r} {float(k[1, i])!r} {float(k[2, i])!r} 1.0\n")


def write_en_binary(path: str, e):
    """raw/enK.<k>: the band energies as raw little-endian doubles."""
    np.asarray(e, dtype="<f8").tofile(path)


def read_en_binary(path: str):
    return np.fromfile(path, dtype="<f8")


def write_en_global(path: str, kpts, e):
    """out/eBands.dat (write_en_global): three header lines, then per k and
    band "k-index  kx ky kz  E" with formats (i7,a) (3f16.8) (a,f18.8)."""
    k = np.asarray(kpts)
    with open(path, "w") as f:
        f.write(" # energies interpolated by MEPinterp program, written by mepInterp (MI355X)\n")
        f.write(" # first 3 columns give the relative k-mesh, 4th column are the enegies\n")
        f.write(" # Kpt_idx  K_x (frac)       K_y (frac)        K_z (frac)       Energy (Hartree)  \n")
        for q in range(k.shape[1]):
            for band in range(e.shape[0]):
                f.write(f"{q + 1:7d}\t" + "".join(f"{k[x, q]:16.8f}" for x in range(3)) +
                        f"\t{e[band, q]:18.8f}\n")


def read_en_global(path: str):
    """-> (k-index, kpts (nk*nb, 3), E) rows of out/eBands.dat."""
    rows = [ln.split() for ln in open(path) if ln.strip() and not ln.lstrip().startswith("#")]
    a = np.array([[float(x) for x in r] for r in rows])
    return a[:, 0].astype(int), a[:, 1:4], a[:, 4]


def read_velo(path: str, n: int):
    """out//velo/velo_Vka.<k> (write_velo) -> V (3, n, n) complex."""
    V = np.zeros((3, n, n), np.complex128)
    for ln in open(path):
        t = ln.split()
        if not t or t[0].startswith("#"):
            continue
        x, i, j = int(t[0]) - 1, int(t[1]) - 1, int(t[2]) - 1
        V[x, i, j] = float(t[3]) + 1j * float(t[4])
    return V


# --------------------------------------------------------------------------
# NPY v2.0 as written by NPY-for-Fortran (fixture header bytes)
# --------------------------------------------------------------------------

def npy_header(descr: str, shape) -> bytes:
    """Header of an NPY v2.0 Fortran-order file, byte-identical to the
    reference's fixtures: dict padded with spaces, then '\\n', so that
    (10 + hlen) % 16 == 0 with the 4-byte little-endian hlen of v2.0."""
    shp = "".join(f"{int(s)}," for s in shape)
    d = f"{{'descr': '{descr}', 'fortran_order': True, 'shape': ({shp}), }}"
    hlen = len(d) + 1
    while (10 + hlen) % 16:
        hlen += 1
    d = d + " " * (hlen - 1 - len(d)) + "\n"
    return b"\x93NUMPY\x02\x00" + struct.pack("<I", hlen) + d.encode("latin1")


def write_npy(path: str, arr):
    """Write ``arr`` as NPY v2.0, Fortran order, '<f8' or '<c16'."""
    a = np.asarray(arr)
    if np.iscomplexobj(a):
        a = a.astype("<c16")
        descr = "<c16"
    else:
        a = a.astype("<f8")
        descr = "<f8"
    with open(path, "wb") as f:
        f.write(npy_header(descr, a.shape))
        f.write(np.asfortranarray(a).tobytes(order="F"))


def ensure_dir(d: str):
    os.makedirs(d, exist_ok=True)
