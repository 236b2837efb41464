"""The FP64 network oracle (oracle/mlp.py) pinned by finite differences of
its own loss, and the refinement helpers (mesh, sub-mesh, C4 model)."""
import numpy as np

import mlp as M


def test_oracle_mlp_gradient_equals_finite_differences():
    rng = np.random.default_rng(3)
    nf, hid = 2, 16
    p = rng.normal(scale=0.3, size=M.n_params(nf, hid))
    k = rng.uniform(-0.5, 0.5, size=(3, 40))
    t = rng.normal(size=40)
    g = M.grad(p, k, t, nf, hid)
    idx = rng.choice(p.size, size=60, replace=False)
    fd = M.grad_fd(p, k, t, nf, hid, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-6, atol=1e-9)


def test_oracle_mlp_features_and_padding():
    x = M.features(np.array([[0.25], [0.0], [-0.125]]), 4)
    assert x.shape == (1, 32)
    assert x[0, 0] == 1.0 and abs(x[0, 1]) < 1e-15      # sin, cos of pi/2
    assert np.all(x[0, 24:] == 0.0)
    assert abs(x[0, 8 + 1] - 1.0) < 1e-15               # cos(0) for k_y


def test_refine_mesh_and_submesh_weights():
    import oracle as O
    from mepinterp_amd import refine
    mp = (4, 3, 2)
    np.testing.assert_array_equal(refine.mp_mesh(mp), O.mp_mesh(mp))
    off = refine.sub_offsets(4)
    assert off.shape == (3, 64)
    # sub-points tile the unit cell symmetrically, a 4^3 MP mesh of it
    np.testing.assert_allclose(off.mean(axis=1), 0.0, atol=1e-15)
    assert np.all(np.abs(off) < 0.5)


def test_c4_model_is_hermitian_with_planted_nodes():
    from mepinterp_amd import models
    m = models.make_c4()
    H, R = m["H"], m["R_frac"]
    assert H.shape == (24, 24, 343) and m["nodes"].shape == (3, 8)
    idx = {tuple(R[:, s].astype(int)): s for s in range(R.shape[1])}
    for s in range(R.shape[1]):
        t = idx[tuple(-R[:, s].astype(int))]
        np.testing.assert_array_equal(H[:, :, t], H[:, :, s].conj().T)
