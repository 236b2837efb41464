"""mep_run_kpts_static (the work buffers in the library's static device
array, no device allocation in the call) against mep_run_kpts on the same
k lists: two calls in a row, the second smaller, for the narrow kernels
(with the velocities inside the band sums), the position operator with the
optical / gyrotropic records, and the wide kernels; plus the size limit."""
import numpy as np
import pytest

from conftest import AU_TO_EV, oracle_model, rand_kpts

pytestmark = pytest.mark.gpu

KEYS = ("Ne", "ahc", "ahc_velo", "ohc", "kubo_ic", "kubo_lc", "kubo_cs", "niu_ic", "niu_lc",
        "niu_cs", "opt_s", "opt_a", "photo", "gyro_C", "gyro_D", "gyro_Dw")


def _engine(m):
    from mepinterp_amd import capi
    return capi.Engine(m["R_frac"], m["H"], r=m.get("r"), wf_centers=m.get("wf_centers"),
                       a_latt=m["a_latt"])


@pytest.mark.parametrize("name", ["c2", "c1_opt", "wide"])
def test_static_scratch_equals_run_kpts(name):
    import oracle as O
    from mepinterp_amd import capi, models
    jobs = capi.MEP_JOB_AHC | capi.MEP_JOB_KUBO | capi.MEP_JOB_MEP
    if name == "c2":
        m, sizes, nval = models.make_c2(seed=2), (300, 97), 16
    elif name == "c1_opt":
        m, sizes, nval = models.make_c1(seed=3), (1200, 513), 4
        jobs |= capi.MEP_JOB_OPT | capi.MEP_JOB_GYRO
    else:
        m, sizes, nval = models.make_random(seed=9, n=48, L=1), (40, 11), 24
    # hw > 0: photo_2nd_cond divides by hw (photo.f90:16-86)
    p = capi.Engine.params(jobs=jobs, n_eF=2, eF_min=-0.01, eF_max=0.01, n_hw=3,
                           hw_min=0.01 / AU_TO_EV, hw_max=0.2 / AU_TO_EV, T_kelvin=300.0,
                           eta=1e-3 / AU_TO_EV, valence_bands=nval)
    with _engine(m) as eng:
        for i, nk in enumerate(sizes):
            k = rand_kpts(nk, seed=100 + i)
            w = np.linspace(0.5, 1.5, nk)
            print(f"{name}: call {i + 1}, nk = {nk}", flush=True)
            st = eng.run_kpts_static(k, p, weights=w)
            ref = eng.run_kpts(k, p, weights=w)
            assert st["n_k"] == nk
            for key in KEYS:
                assert np.all(np.isfinite(ref[key])), (name, nk, key)
                assert np.array_equal(st[key], ref[key]), (name, nk, key)
    if name == "c2":  # and the oracle on the last list
        po = O.RunParams(do_ahc=True, do_kubo=True, do_mep=True, n_eF=2, eF_min=-0.01,
                         eF_max=0.01, n_hw=3, hw_min=0.01 / AU_TO_EV, hw_max=0.2 / AU_TO_EV,
                         T_kelvin=300.0,
                         eta=1e-3 / AU_TO_EV, valence_bands=nval)
        k = rand_kpts(sizes[-1], seed=100 + len(sizes) - 1)
        o = O.accumulate(oracle_model(m), po, k)
        wsum = {key: 0 for key in ("ahc", "niu_ic")}
        with _engine(m) as eng:
            g = eng.run_kpts_static(k, p)
        for key in wsum:
            err = np.max(np.abs(g[key] - o[key])) / np.max(np.abs(o[key]))
            assert err < 1e-8, (key, err)


def test_static_scratch_rejects_lists_beyond_its_size():
    from mepinterp_amd import capi, models
    m = models.make_c2(seed=2)
    p = capi.Engine.params(jobs=capi.MEP_JOB_AHC, n_eF=1, T_kelvin=300.0, valence_bands=16)
    with _engine(m) as eng:
        with pytest.raises(capi.MepError, match="at most [0-9]+ k-points"):
            eng.run_kpts_static(rand_kpts(5000, seed=1), p)
        # the context is unchanged: the ordinary entry point still runs
        assert eng.run_kpts(rand_kpts(10, seed=2), p)["n_k"] == 10
