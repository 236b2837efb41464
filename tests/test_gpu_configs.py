"""GPU parity on the workloads of the named configurations, against the CPU
FP64 oracle on the same k-points.

* C1: n_wf = 8, NR = 27 with a position operator, the full 32^3 mesh,
  T = 300 K, eta = 1e-3 eV (the reference-semantics gauge transform).
* C2: the n_wf = 32, NR = 343 exchange model: run_mesh on a 12^3 mesh (the
  separable Fourier sums and the fused velocity / band-sum kernel) and
  run_kpts on 2,048 seeded points of the 64^3 mesh (the Fourier GEMM).
* C3: all 16 models (n_wf = 8 ... 68, NR = 27 ... 729) in one run_batch,
  each model on a 2^3 or 3^3 mesh, each checked against the oracle.
* C5: n_wf = 96, NR = 1331 on a 2^3 mesh with n_hw = 256 in 0-5 eV,
  eta = 0.05 eV, T = 300 K: alpha (Kubo and Niu) and sigma_ab(omega).

The core_worker loop is core.f90:165-270; its band sums kubo.f90:37-198,
kubo_mep.f90:16-149 and mep_niu.f90:20-159.  Both sides are FP64 but sum in
different orders, so the raw k-sums are compared at 1e-8 relative to the
largest element of each quantity.
"""
import numpy as np
import pytest

from conftest import AU_TO_EV, oracle_model

pytestmark = pytest.mark.gpu

REL = 1e-8
KEYS = ("Ne", "ahc", "ahc_velo", "ohc", "kubo_ic", "kubo_lc", "kubo_cs", "niu_ic", "niu_lc",
        "niu_cs")
ALL_JOBS = dict(do_ahc=True, do_kubo=True, do_mep=True)


def _engine(m, **kw):
    from mepinterp_amd import capi
    return capi.Engine(m["R_frac"], m["H"], r=m.get("r"), wf_centers=m.get("wf_centers"),
                       a_latt=m["a_latt"], **kw)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _compare(g, o, keys=KEYS, rel=REL, what=""):
    for key in keys:
        b = o[key]
        if np.max(np.abs(b)) == 0.0:
            assert np.max(np.abs(g[key])) < 1e-14, f"{what} {key}: expected zero"
            continue
        err = _rel(g[key], b)
        print(f"{what} {key}: rel err {err:.2e}", flush=True)
        assert err < rel, f"{what} {key}: rel err {err:.3e}"


def _both(m, p, kpts=None, mp=None, weights=None):
    """GPU run (mesh or k list) and the oracle over the same k-points."""
    import oracle as O
    from mepinterp_amd import host
    with _engine(m) as eng:
        g = eng.run_mesh(mp, host.run_params(p)) if mp else eng.run_kpts(
            kpts, host.run_params(p), weights=weights)
    pk = {k: v for k, v in p.items() if k in O.RunParams.__init__.__code__.co_varnames}
    o = O.accumulate(oracle_model(m), O.RunParams(**pk), O.mp_mesh(mp) if mp else kpts)
    return g, o


def test_c1_full_mesh_matches_oracle():
    """C1 at its full size: 32^3 mesh, position operator (A1-A3 order)."""
    from mepinterp_amd import host, models
    m = models.make_c1(seed=1)
    mp = (32, 32, 32)
    p = dict(ALL_JOBS, n_eF=1, eF_min=0.0, eF_max=0.0, n_hw=1, hw_min=0.0, hw_max=0.0,
             T_kelvin=300.0, eta=1e-3 / AU_TO_EV, valence_bands=4)
    g, o = _both(m, p, mp=mp)
    assert g["n_k"] == 32 ** 3
    _compare(g, o, what="C1 32^3")
    gs = host.finalize(g, p, mp, m["a_latt"])
    os_ = host.finalize(o, p, mp, m["a_latt"])
    sxy = os_["ahc_tens"][0, 1]
    assert abs(sxy) > 0
    assert abs(gs["ahc_tens"][0, 1] - sxy) <= 1e-8 * abs(sxy)


def test_c2_mesh_matches_oracle():
    """C2 model on a 12^3 mesh: separable sums + fused velocity band sums."""
    from mepinterp_amd import host, models
    m = models.make_c2(seed=2)
    mp = (12, 12, 12)
    p = dict(ALL_JOBS, n_eF=1, eF_min=0.0, eF_max=0.0, n_hw=1, hw_min=0.0, hw_max=0.0,
             T_kelvin=300.0, eta=1e-3 / AU_TO_EV, valence_bands=16)
    g, o = _both(m, p, mp=mp)
    _compare(g, o, what="C2 12^3")
    gs = host.finalize(g, p, mp, m["a_latt"])
    os_ = host.finalize(o, p, mp, m["a_latt"])
    for key in ("ahc_tens", "mep_tens", "kubo_mep_tens"):
        assert _rel(gs[key], os_[key]) < REL, key


def test_c2_kpoint_list_matches_oracle():
    """C2 model on 2,048 seeded points of the 64^3 mesh: the Fourier GEMM."""
    from mepinterp_amd import models
    from mepinterp_amd.refine import mp_mesh
    m = models.make_c2(seed=2)
    mesh = mp_mesh((64, 64, 64))
    rng = np.random.default_rng(2048)
    k = np.asfortranarray(mesh[:, rng.choice(mesh.shape[1], size=2048, replace=False)])
    p = dict(ALL_JOBS, n_eF=3, eF_min=-0.01, eF_max=0.01, n_hw=4, hw_min=0.0,
             hw_max=0.5 / AU_TO_EV, T_kelvin=300.0, eta=1e-3 / AU_TO_EV, valence_bands=16)
    g, o = _both(m, p, kpts=k)
    assert g["n_k"] == 2048
    _compare(g, o, what="C2 k-list")


def test_c3_batch_matches_oracle():
    """All 16 C3 models in one mep_run_batch, each against the oracle."""
    import oracle as O
    from mepinterp_amd import capi, host, models
    work = models.make_c3()
    meshes, ps, engines = [], [], []
    base = dict(ALL_JOBS, n_eF=2, eF_min=-0.01, eF_max=0.01, n_hw=3, hw_min=0.0,
                hw_max=0.2 / AU_TO_EV, T_kelvin=300.0, eta=1e-3 / AU_TO_EV)
    for i, (m, _) in enumerate(work):
        q = 2 + (i % 2)
        meshes.append((q, q, q))
        ps.append(dict(base, valence_bands=m["H"].shape[0] // 2))
    try:
        for m, _ in work:
            engines.append(_engine(m))
            engines[-1].set_workspace_limit(4e9)
        batch = capi.run_batch(engines, meshes, [host.run_params(p) for p in ps])
    finally:
        for e in engines:
            e.close()
    for (m, _), mp, p, g in zip(work, meshes, ps, batch):
        pk = {k: v for k, v in p.items() if k in O.RunParams.__init__.__code__.co_varnames}
        o = O.accumulate(oracle_model(m), O.RunParams(**pk), O.mp_mesh(mp))
        assert g["n_k"] == int(np.prod(mp))
        _compare(g, o, what=m["name"])


def test_c5_small_mesh_matches_oracle():
    """C5 model (n_wf = 96, NR = 1331) on a 2^3 mesh with C5's 256
    frequencies in 0-5 eV, eta = 0.05 eV, T = 300 K."""
    from mepinterp_amd import host, models
    m = models.make_c5()
    mp = (2, 2, 2)
    p = dict(ALL_JOBS, n_eF=1, eF_min=0.0, eF_max=0.0, n_hw=256, hw_min=0.0,
             hw_max=5.0 / AU_TO_EV, T_kelvin=300.0, eta=0.05 / AU_TO_EV, valence_bands=48)
    g, o = _both(m, p, mp=mp)
    _compare(g, o, what="C5 2^3")
    gs = host.finalize(g, p, mp, m["a_latt"])
    os_ = host.finalize(o, p, mp, m["a_latt"])
    for key in ("mep_tens", "kubo_mep_tens", "ahcVELO", "ohcVELO"):
        assert _rel(gs[key], os_[key]) < REL, key


@pytest.mark.parametrize("centres", [False, True])
def test_c2_householder_variants_match_oracle(centres):
    """Both FP64 Householder kernels for 16 < n <= 32 (one k-point per
    wavefront with half rows per lane; two k-points per wavefront with whole
    rows) against the oracle, with and without tight-binding centres, on an
    odd k count (the two-k kernel's last pair holds one k-point)."""
    import oracle as O
    from mepinterp_amd import capi, host, models
    m = models.make_c2(seed=2)
    if centres:
        rng = np.random.default_rng(7)
        m = dict(m, wf_centers=np.asfortranarray(rng.uniform(0.0, 1.0, (3, m["H"].shape[0]))))
    rng = np.random.default_rng(11)
    k = np.asfortranarray(rng.uniform(-0.5, 0.5, (3, 301)))
    p = dict(ALL_JOBS, n_eF=1, eF_min=0.0, eF_max=0.0, n_hw=1, hw_min=0.0, hw_max=0.0,
             T_kelvin=300.0, eta=1e-3 / AU_TO_EV, valence_bands=16)
    pk = {kk: v for kk, v in p.items() if kk in O.RunParams.__init__.__code__.co_varnames}
    o = O.accumulate(oracle_model(m), O.RunParams(**pk), k)
    with _engine(m) as eng:
        for var in (1, 2):
            eng.set_variant(capi.MEP_VAR_EIGH, var)
            g = eng.run_kpts(k, host.run_params(p))
            assert g["n_k"] == 301
            _compare(g, o, what=f"C2 eigh variant {var} centres={centres}")
        with pytest.raises(RuntimeError):
            eng.set_variant(capi.MEP_VAR_EIGH, 4)
        with pytest.raises(RuntimeError):
            eng.set_variant(99, 1)


def test_tuner_picks_and_caches(tmp_path):
    """tune.tune times the eigensolver variants and k-blocks once, applies the
    choice, and reads it back from its cache on the next call."""
    import json
    from mepinterp_amd import host, models, tune
    m = models.make_c2(seed=2)
    p = dict(ALL_JOBS, n_eF=1, eF_min=0.0, eF_max=0.0, n_hw=1, hw_min=0.0, hw_max=0.0,
             T_kelvin=300.0, eta=1e-3 / AU_TO_EV, valence_bands=16)
    rp = host.run_params(p)
    cache = str(tmp_path / "tune.json")
    mp = (16, 16, 16)
    with _engine(m) as eng:
        c1 = tune.tune(eng, mp, rp, m["R_frac"].shape[1], reps=1, path=cache)
        assert c1["eigh"] in (1, 2) and c1["kblock"] in tune.KBLOCKS
        table = json.load(open(cache))
        entry = table[tune.key(32, m["R_frac"].shape[1], 0, 16 ** 3)]
        assert set(entry["seconds_eigh"]) == {"1", "2"}
        c2 = tune.tune(eng, mp, rp, m["R_frac"].shape[1], reps=1, path=cache)
        assert c2 == c1
        s = eng.run_mesh(mp, rp)
        assert s["n_k"] == 16 ** 3


@pytest.mark.parametrize("n", [40, 60, 72, 90, 100])
def test_wide_householder_variants_match_oracle(n):
    """Every Householder kernel for 32 < n <= 128 (the matrix in a 256-thread
    workgroup's LDS; in one wavefront's LDS for n <= 64; in the registers of
    n / 16 wavefronts, 64-, 80-, 96- and 128-row builds) against the oracle on
    random k-points."""
    import oracle as O
    from mepinterp_amd import capi, host, models
    m = models.make_random(seed=90 + n, n=n, L=1)
    rng = np.random.default_rng(n)
    k = np.asfortranarray(rng.uniform(-0.5, 0.5, (3, 37)))
    p = dict(ALL_JOBS, n_eF=1, eF_min=0.0, eF_max=0.0, n_hw=1, hw_min=0.0, hw_max=0.0,
             T_kelvin=300.0, eta=1e-3 / AU_TO_EV, valence_bands=n // 2)
    pk = {kk: v for kk, v in p.items() if kk in O.RunParams.__init__.__code__.co_varnames}
    o = O.accumulate(oracle_model(m), O.RunParams(**pk), k)
    from mepinterp_amd import tune
    with _engine(m) as eng:
        for var in tune.eigh_variants(n, capi.MEP_PREC_FP64):
            eng.set_variant(capi.MEP_VAR_EIGH, var)
            g = eng.run_kpts(k, host.run_params(p))
            _compare(g, o, what=f"n={n} eigh variant {var}")


def test_mixed_kpoint_list_gemm_variants():
    """The mixed path on an explicit k list (the FP32 Fourier GEMM route):
    the FP32-MFMA GEMM and the split-bfloat16 GEMM (MEP_VAR_GEMM32) against
    the FP64 oracle on 1,024 seeded points of C2's 64^3 mesh, and against
    each other.  Both carry FP32 rounding of H(k) accumulated over NR = 343
    terms; the split-bfloat16 products reproduce FP32 products to 2^-24."""
    import oracle as O
    from mepinterp_amd import capi, host, models
    from mepinterp_amd.refine import mp_mesh
    m = models.make_c2(seed=2)
    mesh = mp_mesh((64, 64, 64))
    rng = np.random.default_rng(1024)
    k = np.asfortranarray(mesh[:, rng.choice(mesh.shape[1], size=1024, replace=False)])
    p = dict(ALL_JOBS, n_eF=1, eF_min=0.0, eF_max=0.0, n_hw=1, hw_min=0.0, hw_max=0.0,
             T_kelvin=300.0, eta=1e-3 / AU_TO_EV, valence_bands=16)
    pk = {kk: v for kk, v in p.items() if kk in O.RunParams.__init__.__code__.co_varnames}
    o = O.accumulate(oracle_model(m), O.RunParams(**pk), k)
    rp = host.run_params(dict(p))
    rp.precision = capi.MEP_PREC_MIXED
    got = {}
    with _engine(m) as eng:
        for var in (1, 2):
            eng.set_variant(capi.MEP_VAR_GEMM32, var)
            got[var] = eng.run_kpts(k, rp)
            for key in ("ahc", "kubo_ic", "kubo_lc", "niu_ic", "niu_lc"):
                err = _rel(got[var][key], o[key])
                print(f"GEMM32 variant {var} {key}: rel err vs FP64 oracle {err:.2e}", flush=True)
                assert err < 2e-4, (var, key, err)
        with pytest.raises(RuntimeError):
            eng.set_variant(capi.MEP_VAR_GEMM32, 5)
    for key in ("ahc", "kubo_ic", "niu_ic"):
        assert _rel(got[2][key], got[1][key]) < 1e-4, key
