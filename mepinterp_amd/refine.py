"""Learned adaptive k refinement of the anomalous Hall conductivity.

The reference integrates on uniform Monkhorst-Pack meshes only
(core.f90:165-270).  This module adds an adaptive alternative, on the GPU:

1. the per-k Berry curvature Omega_xy(k) (the summand of kubo_ahc_tens,
   ``Engine.curv``) on a coarse MP mesh;
2. an MLP (``capi.MLP``: sin/cos features of k, three tanh layers, FP32,
   trained and evaluated by the kernels of mep_mlp.hip) learns
   log(|Omega_xy| + floor) from those samples;
3. every coarse cell is scored by the largest prediction over its centre and
   the eight points at a quarter cell from it, and the top fraction of cells
   is refined with a sub^3 sub-mesh; each sub-point weighs 1/sub^3 of its
   parent, so the quadrature weights still sum to the coarse point count;
4. sigma_xy from one weighted ``mep_run_kpts`` over the kept coarse points
   and the sub-points, normalised as the uniform coarse mesh would be.

``select="coarse"`` ranks cells by the coarse |Omega_xy| itself and
``select="random"`` picks them at random (seeded): the two methods without
the network, at the same k-point count.
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


def refine_sigma(engine: capi.Engine, mp, a_latt, frac=0.05, sub=4, eF=0.0, T_kelvin=0.0,
                 select="mlp", mlp=None, epochs=30, batch=4096, lr=2e-3, floor=None, seed=0):
    """sigma_xy with the top `frac` of the coarse cells refined.

    Returns (sigma_xy, info) where info holds the k-point count, the cells
    chosen and, for select="mlp", the final training loss."""
    kc = mp_mesh(mp)
    N = kc.shape[1]
    nref = max(1, int(round(frac * N)))
    info = {}
    if select == "random":
        chosen = np.random.default_rng(seed).choice(N, size=nref, replace=False)
    else:
        om = np.abs(engine.curv(kc, eF, T_kelvin)[0, 1])
        fl = floor if floor is not None else max(1e-12, float(np.median(om)) * 1e-3)
        target = np.log(om + fl)
        if select == "coarse":
            score = target
        elif select == "mlp":
            own = mlp is None
            net = capi.MLP(n_freq=4, hidden=128, seed=seed) if own else mlp
            info["loss"] = net.train(kc, target, epochs=epochs, batch=batch, lr=lr, seed=seed)
            info["target_var"] = float(np.var(target))
            score = cell_scores(net, kc, mp)
            if own:
                net.close()
        else:
            raise ValueError(f"unknown select {select!r}")
        # highest scores first; ties broken by the k index (stable sort)
        chosen = np.argsort(-score, kind="stable")[:nref]
    keep = np.ones(N, dtype=bool)
    keep[chosen] = False
    off = sub_offsets(sub) / np.asarray(mp, dtype=np.float64).reshape(3, 1)
    ks = (kc[:, chosen][:, :, None] + off[:, None, :]).reshape(3, -1)
    k_all = np.asfortranarray(np.concatenate([kc[:, keep], ks], axis=1))
    w_all = np.concatenate([np.ones(int(keep.sum())), np.full(ks.shape[1], 1.0 / sub ** 3)])
    p, rp = ahc_params(eF, T_kelvin)
    sums = engine.run_kpts(k_all, rp, weights=w_all)
    out = host.finalize(sums, p, mp, a_latt, n_k=N)
    info.update(n_k=int(k_all.shape[1]), cells=chosen)
    return float(out["ahc_tens"][0, 1]), info
