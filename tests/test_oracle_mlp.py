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


def _chern_block_curvature(k, kw, v=0.005, m=0.02, mz=0.008, gap=1e-3):
    """Berry curvature (x 4 pi^2, per unit k^3) of the lower band of one C4
    block h = v (sin p_x s_x + sin p_y s_y) + M(p) s_z, p = 2 pi (k - kw),
    M = gap/2 + mz (1 - cos p_z) - m (2 - cos p_x - cos p_y): the closed form
    d.(d_x d x d_y d) / (2 |d|^3); its integral over the zone is 2 pi (Chern
    number 1 in every k_z slice)."""
    p = 2 * np.pi * (k - kw[:, None])
    px, py, pz = p
    M = gap / 2 + mz * (1 - np.cos(pz)) - m * (2 - np.cos(px) - np.cos(py))
    d = np.stack([v * np.sin(px), v * np.sin(py), M])
    dx = np.stack([v * np.cos(px), 0 * px, -m * np.sin(px)])
    dy = np.stack([0 * px, v * np.cos(py), -m * np.sin(py)])
    return 0.5 * np.einsum("ik,ik->k", d, np.cross(dx, dy, axis=0)) / \
        np.linalg.norm(d, axis=0) ** 3 * (2 * np.pi) ** 2


def test_windowed_estimator_beats_uniform_and_cell_replacement():
    """One planted node of the C4 model in closed form, at six positions:
    the uniform coarse rule's error swings with the node's offset from the
    mesh (3e-6 to 1e-4 here), the windowed estimator (a 9^3-cell box around
    the hottest coarse cell) stays near 1.4e-5, so it wins on average; and
    replacing the same budget of hottest cells by their sub-meshes is worse
    than the uniform mesh at every position."""
    from mepinterp_amd import refine
    N = 48
    mp = (N, N, N)
    kc = refine.mp_mesh(mp)
    off = refine.sub_offsets(4) / N
    exact = 2 * np.pi
    nodes = [np.array([0.231, -0.114, 0.402])] + list(
        np.random.default_rng(0).uniform(-0.5, 0.5, size=(5, 3)))
    err_u, err_w = [], []
    for kw in nodes:
        f = _chern_block_curvature(kc, kw)
        err_u.append(abs(f.mean() - exact) / exact)
        cents = refine.hot_boxes(np.log(np.abs(f) + 1e-9), mp, 9, 729, 3)
        assert cents.shape[1] == 1
        k, w = refine.window_points(mp, cents, 9, 1.0, 4)
        assert np.all(np.isfinite(w)) and k.shape[1] == w.size
        err_w.append(abs((_chern_block_curvature(k, kw) * w).sum() / N ** 3 - exact) / exact)
        top = np.argsort(-np.abs(f))[:729]
        fs = _chern_block_curvature((kc[:, top][:, :, None] + off[:, None, :]).reshape(3, -1), kw)
        err_r = abs((f.sum() - f[top].sum() + fs.reshape(729, -1).mean(axis=1).sum()) / N ** 3
                    - exact) / exact
        assert err_r > err_u[-1], (err_r, err_u[-1])
    assert max(err_w) < 2e-5, err_w
    assert np.mean(err_w) < np.mean(err_u) / 2, (err_w, err_u)


def test_window_on_smooth_integrands():
    """On a smooth periodic integrand, 1 + cos(2 pi k_x) cos(2 pi k_z), the
    windowed estimator keeps only the small bias of the coarse mesh's sum of
    the C-infinity window (about 0.36 of a cell volume per 9^3 box), with
    separate and with overlapping boxes (the union window)."""
    from mepinterp_amd import refine
    mp = (24, 20, 16)
    for cents in (np.array([[3, 15], [5, 2], [7, 15]]), np.array([[3, 5], [5, 6], [7, 8]])):
        k, w = refine.window_points(mp, cents, 9, 1.0, 4)
        assert abs(w.sum() - np.prod(mp)) < 0.5 * cents.shape[1]
        f = 1.0 + np.cos(2 * np.pi * k[0]) * np.cos(2 * np.pi * k[2])
        assert abs((f * w).sum() / np.prod(mp) - 1.0) < 1e-4
