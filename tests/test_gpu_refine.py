"""The learned-refinement pieces on the GPU: per-k curvature against the
oracle's kubo_ahc_tens, the network's forward pass and gradients against the
FP64 network oracle (and its finite differences), training, and what the
trained network gains over refinement without it."""
import numpy as np
import pytest

from conftest import oracle_model, rand_kpts

pytestmark = pytest.mark.gpu


def test_curv_batch_matches_oracle_per_k():
    import oracle as O
    from mepinterp_amd import capi, models
    m = models.make_random(seed=21, n=12, L=1)
    k = rand_kpts(20, seed=4)
    ef, T = 0.01, 300.0
    with capi.Engine(m["R_frac"], m["H"], a_latt=m["a_latt"]) as eng:
        om = eng.curv(k, ef, T)
    o = oracle_model(m)
    for ik in range(k.shape[1]):
        e, V, _, _ = O.get_wann_interp(o, k[:, ik])
        fd = O.fd_distrib(e, np.array([ef]), T)
        ref = O.kubo_ahc_tens(e, V, fd)[:, :, 0]
        assert np.max(np.abs(om[:, :, ik] - ref)) < 1e-9 * np.max(np.abs(ref)), ik


def test_mlp_forward_and_gradient_match_oracle():
    import mlp as M
    from mepinterp_amd import capi
    rng = np.random.default_rng(5)
    k = rng.uniform(-0.5, 0.5, size=(3, 200))
    t = rng.normal(size=200).astype(np.float32)
    with capi.MLP(n_freq=4, hidden=32, seed=7) as net:
        p = net.params.astype(np.float64)
        assert p.size == M.n_params(4, 32)
        y = net.predict(k)
        yo = M.forward(p, k, 4, 32)
        assert np.max(np.abs(y - yo)) < 1e-5 * max(1.0, np.max(np.abs(yo)))
        g, loss = net.grad(k, t)
        lo = M.loss(p, k, t, 4, 32)
        assert abs(loss - lo) < 1e-5 * lo
        go = M.grad(p, k, t, 4, 32)
        scale = np.max(np.abs(go))
        assert np.max(np.abs(g - go)) < 1e-4 * scale
        # the FP64 finite differences agree with the GPU gradient too
        idx = rng.choice(p.size, size=40, replace=False)
        fd = M.grad_fd(p, k, t, 4, 32, idx)
        assert np.max(np.abs(g[idx] - fd)) < 1e-4 * scale


def test_mlp_training_fits_a_smooth_field():
    from mepinterp_amd import capi
    rng = np.random.default_rng(6)
    k = rng.uniform(-0.5, 0.5, size=(3, 8192))
    t = (np.cos(2 * np.pi * k[0]) + 0.5 * np.sin(2 * np.pi * (k[1] - k[2]))).astype(np.float32)
    with capi.MLP(n_freq=2, hidden=64, seed=1) as net:
        g0, loss0 = net.grad(k, t)
        loss = net.train(k, t, epochs=40, batch=1024, lr=3e-3, seed=2)
        assert loss < 1e-2 * loss0, (loss0, loss)
        # deterministic: the same seeds give the same weights
        p1 = net.params
    with capi.MLP(n_freq=2, hidden=64, seed=1) as net2:
        net2.train(k, t, epochs=40, batch=1024, lr=3e-3, seed=2)
        np.testing.assert_array_equal(net2.params, p1)


def test_learned_refinement_at_full_size():
    """The learned-refinement workload at its full size: eight Chern layers
    with a planted node each (gap 1e-3 eV, net Chern number 2,
    sigma_xy = -0.0582 a.u.) on a 48^3 base mesh with 5 % of the cells
    refined by 4^3 sub-meshes, against a uniform 256^3 run (converged to
    1e-13: the 256^3 and 288^3 sums agree to 3e-14).  The windowed estimator
    with boxes placed by the trained network matches 256^3 to 1e-4 with more
    than 10x fewer k-points; the same budget placed by the coarse |Omega|
    itself and at random, and the uniform 48^3 mesh, are less accurate;
    replacing cells by their sub-meshes loses to the uniform mesh."""
    from mepinterp_amd import capi, models, refine
    m = models.make_c4()
    mp = (48, 48, 48)
    with capi.Engine(m["R_frac"], m["H"], a_latt=m["a_latt"]) as eng:
        ref = refine.sigma_uniform(eng, (256, 256, 256), m["a_latt"])
        base = refine.sigma_uniform(eng, mp, m["a_latt"])
        got = {sel: refine.refine_sigma(eng, mp, m["a_latt"], select=sel, epochs=40, batch=2048,
                                        seed=3) for sel in ("mlp", "coarse", "random")}
        rep = refine.refine_sigma(eng, mp, m["a_latt"], select="coarse", estimator="replace")
    rel = {sel: abs(v[0] - ref) / abs(ref) for sel, v in got.items()}
    rel_base = abs(base - ref) / abs(ref)
    rel_rep = abs(rep[0] - ref) / abs(ref)
    n_k = {sel: v[1]["n_k"] for sel, v in got.items()}
    print(f"sigma_xy 256^3 {ref:.10e}; rel error uniform 48^3 {rel_base:.2e}, network {rel['mlp']:.2e}, "
          f"coarse |Omega| {rel['coarse']:.2e}, random {rel['random']:.2e}, cells replaced "
          f"{rel_rep:.2e}; k-points {n_k}")
    assert 1e-3 < abs(ref) < 1e-1
    assert max(n_k.values()) * 10 < 256 ** 3
    assert rel["mlp"] <= 1e-4
    assert rel["mlp"] < rel["coarse"] and rel["mlp"] < 0.5 * rel["random"]
    assert rel_base > 3 * rel["mlp"]
    assert rel_rep > rel["mlp"]
