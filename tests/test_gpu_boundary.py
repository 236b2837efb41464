"""GPU checks of the drop-in boundary: the device k-mesh, the job switches
the reference selects by its inputs (doGaugeTrafo, deg(R), the caller's
hw_lst) and the eigensolver's direct-vector path just above its separation
threshold, each against the oracle or against an exact identity."""
import numpy as np
import pytest

from conftest import AU_TO_EV, oracle_model, rand_kpts

pytestmark = pytest.mark.gpu

E_TOL = 1e-10 / AU_TO_EV


def _engine(m, **kw):
    from mepinterp_amd import capi
    return capi.Engine(m["R_frac"], m["H"], r=m.get("r"), wf_centers=m.get("wf_centers"),
                       a_latt=m["a_latt"], **kw)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("mp", [(64, 64, 64), (5, 7, 3), (1, 1, 1), (2, 9, 4), (128, 1, 3)])
def test_device_kmesh_is_bit_exact(mp):
    """The k-points the runs generate on the device equal get_rel_kpt of the
    reference's visiting order (k_space.f90:341-375) bit for bit: the oracle's
    mesh and the closed form (2 q - N - 1) / (2 N) with numpy's correctly
    rounded division."""
    import oracle as O
    from mepinterp_amd import models
    m = models.make_random(seed=1, n=8, L=1)
    with _engine(m) as eng:
        k = eng.mesh_kpts(mp)
        part = eng.mesh_kpts(mp, k0=3, nk=min(11, int(np.prod(mp)) - 3)) if np.prod(mp) > 3 else None
    ko = O.mp_mesh(mp)
    assert k.shape == ko.shape
    assert np.array_equal(k.view(np.uint64), ko.view(np.uint64))
    axes = [(2.0 * np.arange(1, q + 1) - q - 1) / (2.0 * q) for q in mp]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    closed = np.stack([x.ravel(), y.ravel(), z.ravel()])
    assert np.array_equal(k.view(np.uint64), np.ascontiguousarray(closed).view(np.uint64))
    if part is not None:
        assert np.array_equal(part, k[:, 3:3 + part.shape[1]])


@pytest.mark.parametrize("tb", [False, True])
def test_gauge_trafo_off_matches_oracle(tb):
    """doGaugeTrafo = F (wann_interp.f90:100): V is the Wannier-gauge dH(k),
    not rotated; eigenvalues unchanged.  Per k and through a mesh run."""
    import oracle as O
    from mepinterp_amd import capi, models
    m = models.make_random(seed=70 + tb, n=12, L=1, tb=tb)
    k = rand_kpts(10, seed=4)
    om = oracle_model(m, do_gauge_trafo=False)
    with _engine(m, gauge_trafo=False) as eng:
        e, V, _, _ = eng.interp(k)
        kw = dict(n_eF=2, eF_min=-0.01, eF_max=0.01, n_hw=3, hw_min=0.0, hw_max=0.1,
                  T_kelvin=300.0, eta=0.004, valence_bands=6)
        g = eng.run_mesh((4, 3, 3), capi.Engine.params(
            jobs=capi.MEP_JOB_AHC | capi.MEP_JOB_KUBO | capi.MEP_JOB_MEP, **kw))
    for ik in range(k.shape[1]):
        eo, Vo, _, _ = O.get_wann_interp(om, k[:, ik])
        assert np.max(np.abs(e[:, ik] - eo)) < E_TOL
        assert _rel(V[..., ik], Vo) < 1e-9
    o = O.accumulate(om, O.RunParams(do_ahc=True, do_kubo=True, do_mep=True, **kw),
                     O.mp_mesh((4, 3, 3)))
    for key in ("Ne", "ahc", "ahc_velo", "ohc", "niu_ic", "niu_lc"):
        assert _rel(g[key], o[key]) < 1e-8, key


def test_apply_deg_divides_by_the_degeneracy():
    """MEP_F_APPLY_DEG divides X(R) by deg(R); without the flag deg is read
    and ignored, as the reference does (file_io.f90:681-700)."""
    import oracle as O
    from mepinterp_amd import capi, models
    m = models.make_random(seed=75, n=10, L=1)
    nR = m["H"].shape[2]
    rng = np.random.default_rng(5)
    deg = rng.integers(1, 4, size=nR)
    # keep H(-R) = H(R)^+ after the division: deg(-R) = deg(R)
    R = m["R_frac"].astype(int)
    for s in range(nR):
        t = int(np.where(np.all(R == -R[:, s:s + 1], axis=0))[0][0])
        deg[t] = deg[s] if t > s else deg[t]
    Hd = np.asfortranarray(m["H"] / deg[None, None, :])
    md = dict(m, H=Hd)
    k = rand_kpts(8, seed=6)
    with _engine(m, deg=deg, apply_deg=True) as eng:
        e1, V1, _, _ = eng.interp(k)
    with _engine(m, deg=deg) as eng:
        e0, V0, _, _ = eng.interp(k)
    with _engine(m) as eng:
        e00, V00, _, _ = eng.interp(k)
    assert np.array_equal(e0, e00) and np.array_equal(V0, V00)
    om = oracle_model(md)
    for ik in range(k.shape[1]):
        eo, Vo, _, _ = O.get_wann_interp(om, k[:, ik])
        assert np.max(np.abs(e1[:, ik] - eo)) < E_TOL
        assert _rel(V1[..., ik], Vo) < 1e-9
    assert np.max(np.abs(e1 - e0)) > 1e-6


def test_caller_hw_list_reaches_the_band_sums():
    """velo_ahc_tens / kubo_ohc_tens take hw_lst as an argument
    (kubo.f90:86-104, 145-164): a non-uniform list through run_kpts and
    through the band sums of caller-given (e, V)."""
    import oracle as O
    from mepinterp_amd import capi, models
    m = models.make_random(seed=76, n=16, L=1)
    k = rand_kpts(24, seed=7)
    hw = np.array([0.0, 0.001, 0.004, 0.02, 0.05, 0.11])
    ef = O.linspace(-0.01, 0.01, 2)
    T, eta = 300.0, 0.004
    p = capi.Engine.params(jobs=capi.MEP_JOB_AHC, n_eF=2, eF_min=-0.01, eF_max=0.01, T_kelvin=T,
                           eta=eta, valence_bands=8, hw_lst=hw)
    assert p.n_hw == hw.size
    om = oracle_model(m)
    want = {"ahc_velo": 0, "ohc": 0}
    es, Vs = [], []
    for ik in range(k.shape[1]):
        e, V, _, _ = O.get_wann_interp(om, k[:, ik])
        fd = O.fd_distrib(e, ef, T)
        want["ahc_velo"] = want["ahc_velo"] + O.velo_ahc_tens(e, V, hw, fd, eta)
        want["ohc"] = want["ohc"] + O.kubo_ohc_tens(e, V, hw, fd, eta)
        es.append(e)
        Vs.append(V)
    with _engine(m) as eng:
        g = eng.run_kpts(k, p)
        gb = eng.bandsum(np.stack(es, axis=1), np.stack(Vs, axis=-1), p)
    for key in want:
        assert _rel(g[key], want[key]) < 1e-9, key
        assert _rel(gb[key], want[key]) < 1e-9, key


@pytest.mark.parametrize("n", [24, 64])
def test_direct_vectors_just_above_the_separation_threshold(n):
    """Two uncoupled copies split by about 3e-5 |T|, just above the 1e-5 |T|
    threshold: every k-point takes the direct (twisted factorization) vector
    path, where the vector error grows as eps |T| / gap.  Eigenvalues,
    |V_nm|^2 and V_nn against the oracle."""
    import oracle as O
    from mepinterp_amd import models
    h = n // 2
    m1 = models.make_random(seed=62, n=h, L=1)
    k = rand_kpts(12, seed=10)
    om1 = oracle_model(m1)
    big = max(np.max(np.abs(O.get_wann_interp(om1, k[:, i], velo=False)[0]))
              for i in range(k.shape[1]))
    delta = 3e-5 * big
    R = m1["R_frac"]
    H = np.zeros((n, n, R.shape[1]), dtype=np.complex128)
    H[:h, :h] = m1["H"]
    H[h:, h:] = m1["H"]
    r0 = int(np.where(np.all(R == 0, axis=0))[0][0])
    H[h:, h:, r0] += delta * np.eye(h)
    m = dict(R_frac=R, H=np.asfortranarray(H), a_latt=m1["a_latt"])
    with _engine(m) as eng:
        e, V, _, _ = eng.interp(k)
    om = oracle_model(m)
    for ik in range(k.shape[1]):
        eo, Vo, _, _ = O.get_wann_interp(om, k[:, ik])
        gap = np.min(np.diff(eo))
        bigk = np.max(np.abs(eo))
        assert 1e-5 * bigk < gap < 1e-4 * bigk  # the direct path, near its threshold
        assert np.max(np.abs(e[:, ik] - eo)) < E_TOL
        assert _rel(np.abs(V[..., ik]) ** 2, np.abs(Vo) ** 2) < 1e-8
        assert _rel(np.einsum("ann->an", V[..., ik]), np.einsum("ann->an", Vo)) < 1e-9


def test_textbook_switch_matches_oracle():
    """MEP_F_TEXTBOOK (PRB 74 195118 as printed) with a position operator:
    A, Om and V per k, and the run sums including the Chern-Simons terms,
    against the oracle's textbook path (pinned by mathematics on the CPU)."""
    import oracle as O
    from mepinterp_amd import capi, models
    m = models.make_c1(seed=12)
    k = rand_kpts(12, seed=13)
    om = oracle_model(m, textbook=True)
    kw = dict(n_eF=2, eF_min=-0.01, eF_max=0.01, n_hw=2, hw_min=0.0, hw_max=0.1,
              T_kelvin=300.0, eta=0.004, valence_bands=4)
    with _engine(m, textbook=True) as eng:
        e, V, A, Om = eng.interp(k)
        g = eng.run_mesh((4, 4, 3), capi.Engine.params(
            jobs=capi.MEP_JOB_AHC | capi.MEP_JOB_KUBO | capi.MEP_JOB_MEP, **kw))
    for ik in range(k.shape[1]):
        eo, Vo, Ao, Omo = O.get_wann_interp(om, k[:, ik])
        assert np.max(np.abs(e[:, ik] - eo)) < E_TOL
        assert _rel(V[..., ik], Vo) < 1e-8
        assert _rel(A[..., ik], Ao) < 1e-8
        assert _rel(Om[..., ik], Omo) < 1e-8
    o = O.accumulate(om, O.RunParams(do_ahc=True, do_kubo=True, do_mep=True, **kw),
                     O.mp_mesh((4, 4, 3)))
    for key in ("Ne", "ahc", "ahc_velo", "ohc", "kubo_ic", "kubo_lc", "kubo_cs", "niu_ic",
                "niu_lc", "niu_cs"):
        assert np.max(np.abs(o[key])) > 0, key
        assert _rel(g[key], o[key]) < 1e-8, key
