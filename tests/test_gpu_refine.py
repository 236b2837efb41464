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


def test_learned_refinement_beats_refinement_without_the_network():
    """Planted nodes (gap 1e-3 eV) in an insulator on a coarse 16^3 mesh: the
    top 5 % of cells, refined with a 4^3 sub-mesh, chosen by the trained
    network, against the same number of cells chosen by the coarse |Omega|
    itself and at random, with a uniform 192^3 mesh as the reference value."""
    from mepinterp_amd import capi, models, refine
    m = models.make_c4(seed=4)
    mp = (16, 16, 16)
    with capi.Engine(m["R_frac"], m["H"], a_latt=m["a_latt"]) as eng:
        ref = refine.sigma_uniform(eng, (192, 192, 192), m["a_latt"])
        coarse = refine.sigma_uniform(eng, mp, m["a_latt"])
        got = {sel: refine.refine_sigma(eng, mp, m["a_latt"], select=sel, epochs=40, batch=2048,
                                        seed=3) for sel in ("mlp", "coarse", "random")}
    n_k = {sel: v[1]["n_k"] for sel, v in got.items()}
    assert len(set(n_k.values())) == 1 and n_k["mlp"] < 0.01 * 192 ** 3
    err = {sel: abs(v[0] - ref) for sel, v in got.items()}
    info = got["mlp"][1]
    print(f"sigma_xy 192^3 {ref:.3e}; |error| uniform 16^3 {abs(coarse - ref):.3e}, "
          f"network {err['mlp']:.3e}, coarse |Omega| {err['coarse']:.3e}, random {err['random']:.3e}; "
          f"loss {info['loss']:.3f} of variance {info['target_var']:.3f}")
    assert info["loss"] < 0.5 * info["target_var"]
    assert err["mlp"] < 0.1 * err["coarse"]
    assert err["mlp"] < 0.1 * err["random"]
    assert err["mlp"] < 0.01 * abs(coarse - ref)
