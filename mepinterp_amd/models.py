"""Seeded synthetic tight-binding models for the k-space workloads.

All generators return ``dict(R_frac (3,nR), H (n,n,nR) Hartree, r (3,n,n,nR)
Bohr or None, wf_centers or None, a_latt (3,3) Bohr)`` with H(-R) = H(R)^+
enforced exactly, so H(k) is Hermitian for every k.  Energies are generated
in eV and divided by 27.211385 as the reference's reader does
(file_io.f90:741); positions in Angstrom divided by 0.52917721092
(file_io.f90:790).
"""
from __future__ import annotations

import numpy as np

AU_TO_EV = 27.211385
AU_TO_ANGSTROM = 0.52917721092


def cube_R(L: int):
    """R in {-L..L}^3, lexicographic with the first component slowest
    (the order of the fixtures' _hr.dat: (-1,-1,-1), (-1,-1,0), ...)."""
    g = np.arange(-L, L + 1)
    R = np.array([[x, y, z] for x in g for y in g for z in g], dtype=np.float64).T
    return np.asfortranarray(R)


def _hermitian_pairs(R, blocks_pos, zero_block):
    """Fill X(R) for all R from the values on R > 0 (lexicographic) and
    R = 0, with X(-R) = X(R)^+ (last two axes)."""
    nR = R.shape[1]
    idx = {tuple(R[:, s].astype(int)): s for s in range(nR)}
    out = np.zeros(zero_block.shape + (nR,), dtype=np.complex128)
    for s in range(nR):
        key = tuple(R[:, s].astype(int))
        if key == (0, 0, 0):
            out[..., s] = zero_block
        elif key > (0, 0, 0):
            out[..., s] = blocks_pos[s]
            out[..., idx[tuple(-x for x in key)]] = np.conj(np.swapaxes(blocks_pos[s], -1, -2))
    return out


def make_c1(seed: int = 1, n: int = 8, a: float = 6.0, with_r: bool = True):
    """n_wf=8, R in {-1,0,1}^3 (deg 1).  H(0) real diagonal U(-2,2) eV;
    hoppings complex N(0,1) exp(-|R|) eV; r_a(R) complex N(0,0.1) A,
    Hermitian-symmetrised the same way."""
    rng = np.random.default_rng(seed)
    R = cube_R(1)
    nR = R.shape[1]
    Hpos, rpos = {}, {}
    for s in range(nR):
        if tuple(R[:, s].astype(int)) > (0, 0, 0):
            rn = np.linalg.norm(R[:, s])
            Hpos[s] = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) * np.exp(-rn)
            rpos[s] = (rng.normal(scale=0.1, size=(3, n, n)) + 1j * rng.normal(scale=0.1, size=(3, n, n)))
    H0 = np.diag(rng.uniform(-2.0, 2.0, size=n)).astype(np.complex128)
    H = _hermitian_pairs(R, Hpos, H0) / AU_TO_EV
    r = None
    if with_r:
        X = rng.normal(scale=0.1, size=(3, n, n)) + 1j * rng.normal(scale=0.1, size=(3, n, n))
        r0 = 0.5 * (X + np.conj(np.swapaxes(X, -1, -2)))
        r = _hermitian_pairs(R, rpos, r0) / AU_TO_ANGSTROM
    return dict(R_frac=R, H=np.asfortranarray(H),
                r=None if r is None else np.asfortranarray(r), wf_centers=None,
                a_latt=a * np.eye(3), name="c1_n8_nr27")


_PAULI = np.array([[[0, 1], [1, 0]], [[0, -1j], [1j, 0]], [[1, 0], [0, -1]]], dtype=np.complex128)


def make_c2(seed: int = 2, n_sites: int = 8, n_orb: int = 2, L: int = 3, decay: float = 1.5,
            m_ex: float = 1.0, a: float = 6.0):
    """n_wf = n_sites x n_orb x 2 spins (32 by default) with a seeded
    non-collinear exchange -|m| m_s.sigma (x) P_s on every site, 3q-like
    (tetrahedral) directions; R in {-L..L}^3, hoppings N(0,1) exp(-|R|/decay)
    eV.  Basis index = (site*n_orb + orb)*2 + spin."""
    rng = np.random.default_rng(seed)
    n = n_sites * n_orb * 2
    R = cube_R(L)
    nR = R.shape[1]
    Hpos = {}
    for s in range(nR):
        if tuple(R[:, s].astype(int)) > (0, 0, 0):
            rn = np.linalg.norm(R[:, s])
            Hpos[s] = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) * np.exp(-rn / decay)
    X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H0 = 0.5 * (X + X.conj().T)
    tet = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=np.float64) / np.sqrt(3)
    for site in range(n_sites):
        m = tet[site % 4].copy()
        if site >= 4:  # rotate the second tetrahedron by a seeded angle
            th = rng.uniform(0, np.pi)
            c, s_ = np.cos(th), np.sin(th)
            m = np.array([c * m[0] - s_ * m[1], s_ * m[0] + c * m[1], m[2]])
        ms = -m_ex * np.einsum("a,aij->ij", m, _PAULI)
        for orb in range(n_orb):
            i0 = (site * n_orb + orb) * 2
            H0[i0:i0 + 2, i0:i0 + 2] += ms
    H = _hermitian_pairs(R, Hpos, H0) / AU_TO_EV
    return dict(R_frac=R, H=np.asfortranarray(H), r=None, wf_centers=None,
                a_latt=a * np.eye(3), name=f"c2_n{n}_nr{nR}")


def make_random(seed: int, n: int, L: int = 1, decay: float = 1.0, with_r: bool = False,
                tb: bool = False, a: float = 6.0):
    """Generic seeded Hermitian model for parity tests (any n, cube {-L..L}^3)."""
    rng = np.random.default_rng(seed)
    R = cube_R(L)
    nR = R.shape[1]
    Hpos, rpos = {}, {}
    for s in range(nR):
        if tuple(R[:, s].astype(int)) > (0, 0, 0):
            rn = np.linalg.norm(R[:, s])
            Hpos[s] = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) * np.exp(-rn / decay)
            rpos[s] = rng.normal(scale=0.1, size=(3, n, n)) + 1j * rng.normal(scale=0.1, size=(3, n, n))
    X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = _hermitian_pairs(R, Hpos, 0.5 * (X + X.conj().T)) / AU_TO_EV
    r = None
    if with_r:
        Y = rng.normal(scale=0.1, size=(3, n, n)) + 1j * rng.normal(scale=0.1, size=(3, n, n))
        r = _hermitian_pairs(R, rpos, 0.5 * (Y + np.conj(np.swapaxes(Y, -1, -2)))) / AU_TO_ANGSTROM
    tau = rng.uniform(0, 1, size=(3, n)) if tb else None
    return dict(R_frac=R, H=np.asfortranarray(H), r=None if r is None else np.asfortranarray(r),
                wf_centers=None if tau is None else np.asfortranarray(tau),
                a_latt=a * np.eye(3), name=f"rand_n{n}_L{L}")


def make_c5(seed: int = 5, n: int = 96, L: int = 5, decay: float = 2.0, a: float = 6.0):
    """Large model: n_wf = 96, R in {-5..5}^3 (NR = 1331), hoppings
    N(0,1) exp(-|R|/2) eV (complex, H(-R) = H(R)^dagger), Hermitian onsite."""
    m = make_random(seed=seed, n=n, L=L, decay=decay, a=a)
    m["name"] = f"c5_n{n}_nr{m['R_frac'].shape[1]}"
    return m


C3_N = tuple(range(8, 69, 4))        # 8, 12, ..., 68 (16 models)
C3_L = (1, 2, 3, 4)                  # NR = 27, 125, 343, 729


def make_c3(seed: int = 3):
    """The 16 models of the batched workload: n_wf = 8, 12, ..., 68 with NR
    cycling through 27, 125, 343, 729 and meshes from 24^3 to 48^3.
    Returns a list of (model, mp_grid)."""
    out = []
    for i, n in enumerate(C3_N):
        L = C3_L[i % 4]
        m = make_random(seed=seed * 1000 + i, n=n, L=L, decay=1.0 + 0.25 * L)
        m["name"] = f"c3_{i:02d}_n{n}_nr{(2 * L + 1) ** 3}"
        q = 24 + (24 * i) // 15
        out.append((m, (q, q, q)))
    return out


def make_c4(seed: int = 4, n: int = 24, L: int = 3, n_nodes: int = 8, gap_eV: float = 1e-3,
            v_eV: float = 0.005, m_eV: float = 0.02, mz_eV: float = 0.008,
            couple_eV: float = 0.002, bg_eV: float = 3.0, bg_hop_eV: float = 0.05,
            decay: float = 1.5, a: float = 6.0):
    """Learned-refinement workload: n_wf = 24, R in {-3..3}^3 (NR = 343).

    Orbitals (2w, 2w+1), w < n_nodes, form a gapped Chern layer with one
    planted node at a seeded k_w: with p = 2 pi (k - k_w),
        h_w(k) = v (sin p_x s_x + sin p_y s_y) + s_w M(p) s_z,
        M(p) = gap/2 + m_z (1 - cos p_z) - m (2 - cos p_x - cos p_y).
    M is gap/2 at p = 0 and negative at the other three (p_x, p_y) zeros of
    the sines, so every k_z slice of block w has Chern number -s_w, and the
    gap closes to `gap` only at k_w (a massive Dirac point whose Berry
    curvature peaks within ~gap / v of it).  The signs s_w are seeded with
    sum_w s_w = +-2, so sigma_xy is of order 2 / (2 pi a) a.u.  The other
    orbitals are background bands centred at -bg and +bg (half each) with
    hoppings N(0, bg_hop) exp(-|R|/decay) eV, and every node orbital couples
    to them with N(0, couple) exp(-|R|/decay) eV.  E_F = 0 lies in the gap of
    every node."""
    rng = np.random.default_rng(seed)
    R = cube_R(L)
    nR = R.shape[1]
    idx = {tuple(R[:, s].astype(int)): s for s in range(nR)}
    nw = 2 * n_nodes
    nb = n - nw
    H = np.zeros((n, n, nR), dtype=np.complex128)

    def rand_block():
        X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        X[:nw, :nw] = 0.0
        X[:nw, nw:] *= couple_eV
        X[nw:, :nw] *= couple_eV
        X[nw:, nw:] *= bg_hop_eV
        return X

    for s in range(nR):
        Rv = tuple(R[:, s].astype(int))
        if Rv > (0, 0, 0):
            X = rand_block() * np.exp(-np.linalg.norm(R[:, s]) / decay)
            H[:, :, s] = X
            H[:, :, idx[tuple(-np.array(Rv))]] = X.conj().T
    X0 = rand_block()
    X0 = 0.5 * (X0 + X0.conj().T)
    X0[nw:, nw:] += np.diag(np.where(np.arange(nb) < nb // 2, -bg_eV, bg_eV))
    H[:, :, idx[(0, 0, 0)]] += X0
    # planted Chern layers: (n_nodes + 2) / 2 blocks of one sign, the rest
    # of the other, in seeded order and overall sign
    sw = np.array([1.0] * ((n_nodes + 2) // 2) + [-1.0] * ((n_nodes - 2) // 2))
    sw = rng.permutation(sw) * rng.choice([-1.0, 1.0])
    sx, sy, sz = _PAULI
    kw = rng.uniform(-0.5, 0.5, size=(3, n_nodes))
    e = np.eye(3, dtype=int)
    # M = (gap/2 + m_z - 2 m) + m cos p_x + m cos p_y - m_z cos p_z
    cos_coef = (m_eV, m_eV, -mz_eV)
    for w in range(n_nodes):
        o = slice(2 * w, 2 * w + 2)
        H[o, o, idx[(0, 0, 0)]] += sw[w] * (0.5 * gap_eV + mz_eV - 2.0 * m_eV) * sz
        for a_ in range(3):
            ph = np.exp(-2j * np.pi * kw[a_, w])  # e^{-2 pi i k_w.a}
            plus, minus = idx[tuple(e[a_])], idx[tuple(-e[a_])]
            # c cos p_a s_z = c (e^{i p} + e^{-i p}) / 2 s_z
            H[o, o, plus] += 0.5 * sw[w] * cos_coef[a_] * ph * sz
            H[o, o, minus] += 0.5 * sw[w] * cos_coef[a_] * np.conj(ph) * sz
            if a_ < 2:
                s_ = sx if a_ == 0 else sy
                # v sin p_a s_a = v (e^{i p} - e^{-i p}) / (2i) s_a
                H[o, o, plus] += v_eV * ph / 2j * s_
                H[o, o, minus] += -v_eV * np.conj(ph) / 2j * s_
    return dict(R_frac=R, H=np.asfortranarray(H / AU_TO_EV), r=None, wf_centers=None,
                a_latt=a * np.eye(3), name=f"c4_n{n}_nr{nR}", nodes=kw, signs=sw)
