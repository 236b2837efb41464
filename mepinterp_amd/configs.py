"""Run one of the named workloads on the GPU and print one JSON line.

    python -m mepinterp_amd.configs c1 [--mesh Q]
    python -m mepinterp_amd.configs c2k [--kpts N]
    python -m mepinterp_amd.configs c3 [--scale S]
    python -m mepinterp_amd.configs c4 [--mesh Q]
    python -m mepinterp_amd.configs c5 [--mesh Q]

* c1  n_wf = 8, NR = 27 with a position operator, 32^3 mesh, sigma_xy at
      E_F = 0, T = 300 K, eta = 1e-3 eV
* c2k n_wf = 32, NR = 343 (the bench model) on N seeded points of its 64^3
      mesh passed as a k list: the Fourier GEMM route
* c3  16 models, n_wf = 8 ... 68, NR = 27 ... 729, meshes 24^3 ... 48^3
      (each divided by --scale): aggregate k-pts/s of one mep_run_mesh per
      model, of one mep_run_batch over all of them with autotuned k-blocks
      and eigensolver variants, and of the same batch as one replayed HIP
      graph (mep_run_batch_graph)
* c4  n_wf = 24, NR = 343, eight Chern layers with a planted node each
      (gap 1e-3 eV, net Chern number 2): sigma_xy
      on a Q^3 base mesh (48 by default) with the top 5 % of cells refined by
      a 4^3 sub-mesh, cells chosen by the trained network, by the coarse
      |Omega| and at random, against the uniform base and 256^3 meshes
* c5  n_wf = 96, NR = 1331, Q^3 mesh (128 by default), alpha_ij (Kubo and
      Niu) plus sigma_ab(omega) on 256 frequencies in 0-5 eV, eta = 0.05 eV

Times are wall-clock per mep_run_mesh call (model upload excluded) and the
per-stage HIP-event split of the same call.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import capi, host, models

AU_TO_EV = 27.211385
STAGES = ("kgen", "phase", "ft_gemm", "eigh", "velo", "bandsum", "reduce")


def _run(m, mp, params, warm=True):
    with capi.Engine(m["R_frac"], m["H"], r=m.get("r"), wf_centers=m.get("wf_centers"),
                     a_latt=m["a_latt"]) as eng:
        if warm:
            eng.run_mesh((4, 4, 4), params)
        eng.set_profiling(True)
        eng.reset_times()
        t0 = time.perf_counter()
        sums = eng.run_mesh(mp, params)
        dt = time.perf_counter() - t0
        split = {s: eng.kernel_time(s)[0] for s in STAGES}
    return sums, dt, split


def run_c1(q: int = 32):
    m = models.make_c1(seed=1)
    p = dict(do_ahc=True, do_kubo=False, do_mep=False, n_eF=1, eF_min=0.0, eF_max=0.0, n_hw=1,
             hw_min=0.0, hw_max=0.0, T_kelvin=300.0, eta=1e-3 / AU_TO_EV, valence_bands=4)
    params = host.run_params(p)
    sums, dt, split = _run(m, (q, q, q), params)
    out = host.finalize(sums, p, (q, q, q), m["a_latt"])
    return dict(config="c1", workload=f"n_wf=8 NR=27 with r(R), {q}^3 MP mesh, sigma_xy",
                n_k=q ** 3, seconds=dt, k_pts_s=q ** 3 / dt, kernel_ms=split,
                sigma_xy=float(out["ahc_tens"][0, 1]))


def run_c2k(nk: int = 262144, mixed: bool = False):
    """C2's model (n_wf = 32, NR = 343) on nk seeded points of its 64^3 mesh
    as an explicit k list: the Fourier GEMM route (phase table + FP64 MFMA
    GEMM) instead of the separable mesh sums.  mixed: the mixed-precision
    path instead, timed with each FP32 GEMM (FP32 MFMA and split bfloat16)."""
    import numpy as np_
    from .refine import mp_mesh
    m = models.make_c2(seed=2)
    mesh = mp_mesh((64, 64, 64))
    rng = np_.random.default_rng(2048)
    k = np_.asfortranarray(mesh[:, rng.choice(mesh.shape[1], size=nk, replace=False)])
    p = dict(do_ahc=True, do_kubo=True, do_mep=True, n_eF=1, eF_min=0.0, eF_max=0.0, n_hw=1,
             hw_min=0.0, hw_max=0.0, T_kelvin=300.0, eta=1e-3 / AU_TO_EV, valence_bands=16)
    rp = host.run_params(p)
    if mixed:
        rp.precision = capi.MEP_PREC_MIXED
    res = {}
    with capi.Engine(m["R_frac"], m["H"], a_latt=m["a_latt"]) as eng:
        for var in ((1, 2) if mixed else (0,)):
            if mixed:
                eng.set_variant(capi.MEP_VAR_GEMM32, var)
            eng.run_kpts(k[:, :64], rp)
            eng.set_profiling(True)
            eng.reset_times()
            t0 = time.perf_counter()
            eng.run_kpts(k, rp)
            dt = time.perf_counter() - t0
            eng.set_profiling(False)
            split = {s_: eng.kernel_time(s_)[0] for s_ in STAGES + ("refine",)}
            res[{0: "fp64", 1: "mixed_fp32_gemm", 2: "mixed_bf16x3_gemm"}[var]] = dict(
                seconds=dt, k_pts_s=nk / dt, kernel_ms=split)
    return dict(config="c2k", workload=f"n_wf=32 NR=343, {nk} seeded points of the 64^3 mesh as a "
                "k list (Fourier GEMM)" + (", mixed precision" if mixed else ""), n_k=nk, **res)


def run_c3_models(only):
    """One run of each listed C3 model (indexes into models.make_c3()), each
    after a 4^3 warm-up and apart from the others by a pause, so that a
    kernel trace separates them."""
    jobs = dict(do_ahc=True, do_kubo=True, do_mep=True, n_eF=1, eF_min=0.0, eF_max=0.0, n_hw=1,
                hw_min=0.0, hw_max=0.0, T_kelvin=300.0, eta=1e-3 / AU_TO_EV)
    c3 = models.make_c3()
    out = []
    for i in only:
        m, mp = c3[i]
        rp = host.run_params(dict(jobs, valence_bands=m["H"].shape[0] // 2))
        eng = capi.Engine(m["R_frac"], m["H"], a_latt=m["a_latt"])
        try:
            eng.run_mesh((4, 4, 4), rp)
            time.sleep(0.3)
            t0 = time.perf_counter()
            eng.run_mesh(mp, rp)
            dt = time.perf_counter() - t0
        finally:
            eng.close()
        time.sleep(0.3)
        out.append(dict(model=m["name"], n_k=mp[0] ** 3, seconds=dt, k_pts_s=mp[0] ** 3 / dt))
    return dict(config="c3", workload="selected C3 models, one run each", models=out)


def run_c3(scale: int = 1):
    from . import tune
    jobs = dict(do_ahc=True, do_kubo=True, do_mep=True, n_eF=1, eF_min=0.0, eF_max=0.0, n_hw=1,
                hw_min=0.0, hw_max=0.0, T_kelvin=300.0, eta=1e-3 / AU_TO_EV)
    work = []
    for m, mp in models.make_c3():
        q = max(4, mp[0] // scale)
        p = dict(jobs, valence_bands=m["H"].shape[0] // 2)
        work.append((m, (q, q, q), host.run_params(p)))
    engines = [capi.Engine(m["R_frac"], m["H"], a_latt=m["a_latt"]) for m, _, _ in work]
    try:
        for e in engines:  # all sixteen stay resident: share about 160 GB of HBM
            e.set_workspace_limit(160e9 / len(engines))
        # per-model launches with the library's default k-blocks
        per, t_seq = [], 0.0
        seq_sums = []
        for eng, (m, mp, rp) in zip(engines, work):
            eng.run_mesh((4, 4, 4), rp)
            t0 = time.perf_counter()
            seq_sums.append(eng.run_mesh(mp, rp))
            dt = time.perf_counter() - t0
            t_seq += dt
            # the stage split from a second, event-timed run
            eng.set_profiling(True)
            eng.reset_times()
            eng.run_mesh(mp, rp)
            eng.set_profiling(False)
            split = {s_: round(eng.kernel_time(s_)[0], 3) for s_ in STAGES}
            per.append(dict(model=m["name"], n_k=mp[0] ** 3, seconds=dt, k_pts_s=mp[0] ** 3 / dt,
                            kernel_ms=split))
        # autotuned k-blocks (cached), then one batched call over all models
        tuned = [tune.tune(eng, mp, rp, m["R_frac"].shape[1]) for eng, (m, mp, rp) in zip(engines, work)]
        capi.run_batch(engines, [w[1] for w in work], [w[2] for w in work])
        t0 = time.perf_counter()
        bat = capi.run_batch(engines, [w[1] for w in work], [w[2] for w in work])
        t_bat = time.perf_counter() - t0
        dev = max(float(abs(b["ahc"][0, 1, 0] - s["ahc"][0, 1, 0]) /
                        max(abs(s["ahc"][0, 1, 0]), 1e-300)) for b, s in zip(bat, seq_sums))
        # the same batch captured into one HIP graph: capture, then replays
        t0 = time.perf_counter()
        capi.run_batch(engines, [w[1] for w in work], [w[2] for w in work], graph=True)
        t_cap = time.perf_counter() - t0
        t_rep = float("inf")
        for _ in range(2):
            t0 = time.perf_counter()
            gr = capi.run_batch(engines, [w[1] for w in work], [w[2] for w in work], graph=True)
            t_rep = min(t_rep, time.perf_counter() - t0)
        graph_equal = all(np.array_equal(g_["ahc"], b["ahc"]) and np.array_equal(g_["kubo_ic"], b["kubo_ic"])
                          for g_, b in zip(gr, bat))
    finally:
        for e in engines:
            e.close()
    tot_k = sum(w[1][0] ** 3 for w in work)
    return dict(config="c3", workload="16 models n_wf=8..68, NR=27..729, meshes 24^3..48^3"
                + (f" / {scale}" if scale > 1 else ""),
                n_k=tot_k, per_model=dict(seconds=t_seq, k_pts_s=tot_k / t_seq),
                batched=dict(seconds=t_bat, k_pts_s=tot_k / t_bat, tuned=tuned,
                             max_rel_dev_sigma_xy_vs_per_model=dev),
                graph=dict(capture_seconds=t_cap, seconds=t_rep, k_pts_s=tot_k / t_rep,
                           bit_equal_to_stream_batch=bool(graph_equal)),
                models=per)


def run_c4(q: int = 48, box: int = 9, flat: float = 1.0):
    from . import refine
    m = models.make_c4()
    mp = (q, q, q)
    res = {}
    with capi.Engine(m["R_frac"], m["H"], a_latt=m["a_latt"]) as eng:
        t0 = time.perf_counter()
        ref = refine.sigma_uniform(eng, (256, 256, 256), m["a_latt"])
        t_ref = time.perf_counter() - t0
        base = refine.sigma_uniform(eng, mp, m["a_latt"])
        for est in ("window", "replace"):
            for sel in ("mlp", "coarse", "random"):
                t0 = time.perf_counter()
                s, info = refine.refine_sigma(eng, mp, m["a_latt"], select=sel, epochs=40,
                                              batch=2048, seed=3, estimator=est, box=box,
                                              flat=flat)
                r = dict(sigma_xy=s, abs_err=abs(s - ref), rel_err=abs(s - ref) / abs(ref),
                         n_k=info["n_k"], seconds=time.perf_counter() - t0)
                if sel == "mlp":
                    r.update(train_loss=info["loss"], target_var=info["target_var"])
                if est == "window":
                    r.update(boxes=int(info["boxes"].shape[1]))
                res[f"{est}_{sel}"] = r
    return dict(config="c4", workload=f"n_wf=24 NR=343, 8 Chern layers with planted nodes (gap 1e-3 "
                f"eV, net Chern number 2), {q}^3 base mesh, 5 % of cells refined 4^3 (windowed: "
                f"{box}^3-cell boxes, flat {flat}), sigma_xy at E_F=0, T=0",
                sigma_xy_256=ref, n_k_256=256 ** 3, seconds_256=t_ref,
                uniform_base=dict(sigma_xy=base, abs_err=abs(base - ref),
                                  rel_err=abs(base - ref) / abs(ref), n_k=q ** 3),
                refined=res, k_ratio=256 ** 3 / res["window_mlp"]["n_k"])


def run_c5(q: int = 128):
    m = models.make_c5()
    p = dict(do_ahc=True, do_kubo=True, do_mep=True, n_eF=1, eF_min=0.0, eF_max=0.0, n_hw=256,
             hw_min=0.0, hw_max=5.0 / AU_TO_EV, T_kelvin=300.0, eta=0.05 / AU_TO_EV,
             valence_bands=48)
    sums, dt, split = _run(m, (q, q, q), host.run_params(p))
    out = host.finalize(sums, p, (q, q, q), m["a_latt"])
    return dict(config="c5", workload=f"n_wf=96 NR=1331, {q}^3 MP mesh, alpha_ij (Kubo, Niu), "
                "sigma_ab(omega) on 256 frequencies in 0-5 eV, eta=0.05 eV, FP64",
                n_k=q ** 3, seconds=dt, k_pts_s=q ** 3 / dt, kernel_ms=split,
                alpha_niu_xx=float(out["mep_tens"][0, 0]))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("config", choices=("c1", "c2k", "c3", "c4", "c5"))
    ap.add_argument("--kpts", type=int, default=262144, help="c2k: number of k-points")
    ap.add_argument("--mixed", action="store_true", help="c2k: the mixed-precision path")
    ap.add_argument("--mesh", type=int, default=None, help="Q for a Q^3 mesh (c1, c4, c5)")
    ap.add_argument("--scale", type=int, default=1, help="divide the c3 meshes by this")
    ap.add_argument("--only", type=str, default=None,
                    help="c3: comma-separated model indexes, one run each (for kernel traces)")
    ap.add_argument("--box", type=int, default=9, help="c4: cells per box edge (odd)")
    ap.add_argument("--flat", type=float, default=1.0, help="c4: half-width of the window's flat top")
    a = ap.parse_args(argv)
    if a.config == "c1":
        r = run_c1(a.mesh or 32)
    elif a.config == "c2k":
        r = run_c2k(a.kpts, a.mixed)
    elif a.config == "c3":
        r = run_c3_models([int(x) for x in a.only.split(",")]) if a.only else run_c3(a.scale)
    elif a.config == "c4":
        r = run_c4(a.mesh or 48, a.box, a.flat)
    else:
        r = run_c5(a.mesh or 128)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
