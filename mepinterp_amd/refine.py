"""Learned adaptive k refinement of the anomalous Hall conductivity.

The reference integrates on uniform Monkhorst-Pack meshes only
(core.f90:165-270).  This module adds an adaptive alternative, on the GPU:

1. the per-k Berry curvature Omega_xy(k) (the summand of kubo_ahc_tens,
   ``Engine.curv``) on a coarse MP mesh;
2. an MLP (``capi.MLP``: sin/cos features of k, three tanh layers, FP32,
   trained and evaluated by the kernels of mep_mlp.hip) learns
   log(|Omega_xy| + floor) from those samples;
3. every coarse cell is scored by the largest prediction over its centre and
   the eight points at a quarter cell from it; hot spots are taken greedily
   in score order (a cell near an earlier centre is skipped), each claiming
   a box of B^3 cells, until the union of the boxes' cells reaches the
   budget fraction of the mesh;
4. sigma_xy from one weighted ``mep_run_kpts``.  The estimator splits the
   integrand with a smooth window chi (a product of C-infinity steps, 1 near
   the box centre and 0 at its faces; overlapping boxes combine as
   1 - prod (1 - chi_b)): f (1 - chi) on the coarse mesh and f chi on the
   boxes' sub^3 sub-meshes, which tile a uniform fine grid.
   Both parts are smooth periodic (or compactly supported) integrands on
   uniform grids, so neither loses the uniform rule's fast convergence; the
   peak itself is integrated on the fine grid.

``estimator="replace"`` instead replaces each chosen cell by its sub-mesh
(weights 1/sub^3).  That breaks the error cancellation of the periodic
uniform rule at the edges of the refined region, and on smooth-plus-peaked
integrands it is less accurate than the uniform coarse mesh; it is kept for
comparison.  ``select="coarse"`` ranks cells by the coarse |Omega_xy| itself
and ``select="random"`` picks them at random (seeded): the two methods
without the network, at the same k-point budget.
"""
from __future__ import annotations

import numpy as np

from . import capi, host


def mp_mesh(mp):
    """MP k-points (3, N) in the reference's visiting order (k_space.f90:341-375):
    k_i = (2 q_i - N_i - 1) / (2 N_i), x fastest."""
    axes = [(2.0 * np.arange(1, m + 1) - m - 1) / (2.0 * m) for m in mp]
    kz, ky, kx = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return np.asfortranarray(np.stack([kx.ravel(), ky.ravel(), kz.ravel()]))


def sub_offsets(sub: int):
    """Offsets of a sub^3 MP sub-mesh inside a cell of unit width."""
    o = (2.0 * np.arange(1, sub + 1) - sub - 1) / (2.0 * sub)
    oz, oy, ox = np.meshgrid(o, o, o, indexing="ij")
    return np.stack([ox.ravel(), oy.ravel(), oz.ravel()])


def cell_scores(mlp: capi.MLP, kc, mp):
    """Largest network prediction over each cell's centre and the eight
    points at (+-1/4, +-1/4, +-1/4) of a cell from it."""
    h = np.asarray(mp, dtype=np.float64).reshape(3, 1)
    pts = [kc] + [kc + np.array(s).reshape(3, 1) / (4.0 * h)
                  for s in np.array(np.meshgrid([-1, 1], [-1, 1], [-1, 1])).reshape(3, -1).T]
    return np.max(np.stack([mlp.predict(p) for p in pts]), axis=0)


def ahc_params(eF=0.0, T_kelvin=0.0):
    p = dict(do_ahc=True, n_eF=1, eF_min=eF, eF_max=eF, n_hw=1, hw_min=0.0, hw_max=0.0,
             T_kelvin=T_kelvin, eta=0.0, valence_bands=1)
    return p, host.run_params(p)


def sigma_uniform(engine: capi.Engine, mp, a_latt, eF=0.0, T_kelvin=0.0):
    """sigma_xy on a uniform MP mesh (the plain method)."""
    p, rp = ahc_params(eF, T_kelvin)
    out = host.finalize(engine.run_mesh(mp, rp), p, mp, a_latt)
    return float(out["ahc_tens"][0, 1])


def smooth_step(t):
    """C-infinity step: 0 for t <= 0, 1 for t >= 1."""
    t = np.clip(np.asarray(t, dtype=np.float64), 0.0, 1.0)
    with np.errstate(divide="ignore", over="ignore"):
        g0 = np.where(t > 0, np.exp(-1.0 / np.maximum(t, 1e-300)), 0.0)
        g1 = np.where(t < 1, np.exp(-1.0 / np.maximum(1.0 - t, 1e-300)), 0.0)
    return g0 / (g0 + g1)


def box_window(d, half, flat):
    """Product over the axes of 1 for |d| <= flat, 0 for |d| >= half (cell
    units from the box centre), C-infinity in between."""
    return np.prod(smooth_step((half - np.abs(d)) / (half - flat)), axis=0)


def _box_cells(cell, centre, N, r):
    """Mask of the mesh cells within r cells (per axis, periodic) of centre."""
    d = (cell - centre[:, None] + N[:, None] // 2) % N[:, None] - N[:, None] // 2
    return np.all(np.abs(d) <= r, axis=0)


def _mesh_cells(mp):
    """Integer cell (0 .. N-1 per axis) of every MP k-point, in mesh order."""
    N = np.asarray(mp, dtype=np.int64)
    kc = mp_mesh(mp)
    return kc, np.rint((kc * N[:, None] * 2 + N[:, None] - 1) / 2.0).astype(np.int64)


def hot_boxes(score, mp, box, budget, excl):
    """Greedy box centres in descending score: a cell becomes a centre unless
    it lies within `excl` cells (per axis) of an earlier centre; stops when
    the union of the boxes' cells would exceed `budget` cells.  Boxes may
    overlap.  Returns the centres as (3, nb) integer cells."""
    N = np.asarray(mp, dtype=np.int64)
    _, cell = _mesh_cells(mp)
    r = (box - 1) // 2
    order = np.argsort(-score, kind="stable")
    near = np.zeros(cell.shape[1], dtype=bool)   # within excl of a centre
    union = np.zeros(cell.shape[1], dtype=bool)  # cells of some box
    cents = []
    for q in order:
        if near[q]:
            continue
        c = cell[:, q]
        grown = union | _box_cells(cell, c, N, r)
        if grown.sum() > budget:
            break
        union = grown
        near |= _box_cells(cell, c, N, excl)
        cents.append(c)
    return np.array(cents, dtype=np.int64).T.reshape(3, -1)


def window_points(mp, cents, box, flat, sub):
    """The weighted k list of the windowed estimator.  The window is the
    union chi = 1 - prod_b (1 - chi_b) of one product window per box (1
    within `flat` cells of its centre, 0 at the box faces): never above 1,
    C-infinity, supported in the union of the box cells.  Coarse points
    weigh 1 - chi, the sub^3 sub-mesh points of every union cell chi / sub^3
    (points of zero weight dropped)."""
    N = np.asarray(mp, dtype=np.int64)
    kc, cell = _mesh_cells(mp)
    r = (box - 1) // 2
    half = box / 2.0
    union = np.zeros(cell.shape[1], dtype=bool)
    for b in range(cents.shape[1]):
        union |= _box_cells(cell, cents[:, b], N, r)

    def chi(pts):  # pts (3, P) in cell coordinates
        rest = np.ones(pts.shape[1])
        for b in range(cents.shape[1]):
            d = (pts - cents[:, b:b + 1] + N[:, None] / 2.0) % N[:, None] - N[:, None] / 2.0
            rest *= 1.0 - box_window(d, half, flat)
        return 1.0 - rest

    o = (2.0 * np.arange(1, sub + 1) - sub - 1) / (2.0 * sub)
    oz, oy, ox = np.meshgrid(o, o, o, indexing="ij")
    offc = np.stack([ox.ravel(), oy.ravel(), oz.ravel()])
    cu = cell[:, union].astype(np.float64)
    chi_c = np.zeros(cell.shape[1])
    chi_c[union] = chi(cu)
    wf = chi((cu[:, :, None] + offc[:, None, :]).reshape(3, -1)) / sub ** 3
    kf = (kc[:, union][:, :, None] + offc[:, None, :] / N[:, None, None]).reshape(3, -1)
    wc = 1.0 - chi_c
    kk, fk = wc != 0, wf != 0
    return (np.asfortranarray(np.concatenate([kc[:, kk], kf[:, fk]], axis=1)),
            np.concatenate([wc[kk], wf[fk]]))


def refine_sigma(engine: capi.Engine, mp, a_latt, frac=0.05, sub=4, eF=0.0, T_kelvin=0.0,
                 select="mlp", mlp=None, epochs=30, batch=4096, lr=2e-3, floor=None, seed=0,
                 estimator="window", box=9, flat=1.0, excl=None, n_freq=4, hidden=128):
    """sigma_xy with a fraction `frac` of the coarse cells refined by sub^3
    sub-meshes.

    estimator="window": hot-spot boxes of box^3 cells with the smooth
    partition f (1 - chi) + f chi (module docstring); "replace": the top
    cells replaced by their sub-meshes.  Returns (sigma_xy, info) where info
    holds the k-point count, the chosen cells or box centres and, for
    select="mlp", the final training loss."""
    kc = mp_mesh(mp)
    N = kc.shape[1]
    nref = max(1, int(round(frac * N)))
    info = {}
    if select == "random":
        score = np.random.default_rng(seed).random(N)
    else:
        om = np.abs(engine.curv(kc, eF, T_kelvin)[0, 1])
        fl = floor if floor is not None else max(1e-12, float(np.median(om)) * 1e-3)
        target = np.log(om + fl)
        if select == "coarse":
            score = target
        elif select == "mlp":
            own = mlp is None
            net = capi.MLP(n_freq=n_freq, hidden=hidden, seed=seed) if own else mlp
            info["loss"] = net.train(kc, target, epochs=epochs, batch=batch, lr=lr, seed=seed)
            info["target_var"] = float(np.var(target))
            score = cell_scores(net, kc, mp)
            if own:
                net.close()
        else:
            raise ValueError(f"unknown select {select!r}")
    p, rp = ahc_params(eF, T_kelvin)
    if estimator == "window":
        cents = hot_boxes(score, mp, box, nref, int(np.ceil(flat)) + 2 if excl is None else excl)
        k_all, w_all = window_points(mp, cents, box, flat, sub)
        info["boxes"] = cents
    elif estimator == "replace":
        # highest scores first; ties broken by the k index (stable sort)
        chosen = np.argsort(-score, kind="stable")[:nref]
        keep = np.ones(N, dtype=bool)
        keep[chosen] = False
        off = sub_offsets(sub) / np.asarray(mp, dtype=np.float64).reshape(3, 1)
        ks = (kc[:, chosen][:, :, None] + off[:, None, :]).reshape(3, -1)
        k_all = np.asfortranarray(np.concatenate([kc[:, keep], ks], axis=1))
        w_all = np.concatenate([np.ones(int(keep.sum())), np.full(ks.shape[1], 1.0 / sub ** 3)])
        info["cells"] = chosen
    else:
        raise ValueError(f"unknown estimator {estimator!r}")
    sums = engine.run_kpts(k_all, rp, weights=w_all)
    out = host.finalize(sums, p, mp, a_latt, n_k=N)
    info.update(n_k=int(k_all.shape[1]), weight_sum=float(w_all.sum()))
    return float(out["ahc_tens"][0, 1]), info
