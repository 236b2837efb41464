"""mep_run_batch_graph: the batch captured into one HIP graph and replayed.

Every replay must give the stream version's sums bit for bit (the same
kernels in the same order on each context), and anything that changes what
the graph would have to launch (parameters, a context's buffers or kernel
variant, another mesh run on a context in between) must be picked up.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("Ne", "ahc", "ahc_velo", "ohc", "kubo_ic", "kubo_lc", "kubo_cs", "niu_ic", "niu_lc",
        "niu_cs")


def _setup():
    from mepinterp_amd import capi, models
    # one model of each kernel family: n <= 16, the fused 17..32, the wide n > 32
    ms = [models.make_random(seed=80 + n, n=n, L=L) for n, L in ((8, 1), (24, 2), (40, 1))]
    meshes = [(6, 5, 4), (5, 4, 4), (3, 3, 2)]

    def params(n_eF):
        return [capi.Engine.params(jobs=capi.MEP_JOB_AHC | capi.MEP_JOB_KUBO | capi.MEP_JOB_MEP,
                                   n_eF=n_eF, eF_min=-0.01, eF_max=0.01, n_hw=2, hw_max=0.05,
                                   T_kelvin=300.0, eta=0.004, valence_bands=m["H"].shape[0] // 2)
                for m in ms]
    engines = [capi.Engine(m["R_frac"], m["H"], a_latt=m["a_latt"]) for m in ms]
    return capi, engines, meshes, params


def _same(a_list, b_list, what):
    for i, (a, b) in enumerate(zip(a_list, b_list)):
        assert a["n_k"] == b["n_k"], (what, i)
        for key in KEYS:
            np.testing.assert_array_equal(a[key], b[key], err_msg=f"{what} model {i} {key}")


def test_graph_replays_equal_the_stream_batch():
    capi, engines, meshes, params = _setup()
    try:
        engines[0].set_kblock(64)  # several k-blocks for one model
        ps = params(2)
        ref = capi.run_batch(engines, meshes, ps)
        for call in range(3):  # capture, then two replays
            _same(capi.run_batch(engines, meshes, ps, graph=True), ref, f"replay {call}")
        # other parameters: recaptured
        ps3 = params(3)
        ref3 = capi.run_batch(engines, meshes, ps3)
        _same(capi.run_batch(engines, meshes, ps3, graph=True), ref3, "n_eF = 3")
        # a different mesh on one context in between (its phase tables and
        # parameter tables are rewritten), then the graph call again
        engines[1].run_mesh((7, 3, 2), ps[1])
        _same(capi.run_batch(engines, meshes, ps3, graph=True), ref3, "after another mesh")
        # a kernel variant change on the fused-size model: recaptured
        engines[1].set_variant(capi.MEP_VAR_EIGH, 2)
        ref_v = capi.run_batch(engines, meshes, ps3)
        _same(capi.run_batch(engines, meshes, ps3, graph=True), ref_v, "variant 2")
        # a reallocation (smaller k-blocks): recaptured
        engines[2].set_kblock(8)
        ref_k = capi.run_batch(engines, meshes, ps3)
        _same(capi.run_batch(engines, meshes, ps3, graph=True), ref_k, "k-block 8")
    finally:
        for e in engines:
            e.close()


def test_graph_of_a_destroyed_context_is_dropped():
    capi, engines, meshes, params = _setup()
    ps = params(1)
    try:
        capi.run_batch(engines, meshes, ps, graph=True)
        engines[2].close()  # drops the graph that holds it
        ref = capi.run_batch(engines[:2], meshes[:2], ps[:2])
        _same(capi.run_batch(engines[:2], meshes[:2], ps[:2], graph=True), ref, "two models")
    finally:
        for e in engines:
            e.close()
