#!/usr/bin/env python3
"""k-space throughput of the MI355X engine on the n_wf=32 workload.

One step = the whole k-loop of core_worker (mep_run_mesh) over the 64^3
Monkhorst-Pack mesh of a seeded n_wf=32, NR=343 model (8 sites x 2 orbitals
x 2 spins, non-collinear exchange, hoppings N(0,1) exp(-|R|/1.5) eV): k-point
generation, separable Fourier sums (H(k), dH(k) on the mesh), batched eigensolves,
velocity rotation, Fermi-Dirac weights, AHC + velo/ohc, Kubo MEP, Niu MEP and
the device reduction, ending with the sums on the host.  X(R) is uploaded
before the timed region.

Prints ONE JSON line on stdout; diagnostics go to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

AU_TO_EV = 27.211385
PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0

# HBM bytes per k-point moved by the kernels on this workload, from the
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes of this script (steps 1,
# warmup 0, no CPU leg) on the current kernels, scaled by the counters'
# readings for a plain 1 GiB copy (x2.0 / x1.0), full-mesh dispatches
# (profiles/kspace_n32_64cubed_fp64.md).  Counters cannot be read from
# inside the timed run, so the figures are per-k constants of that profile;
# they match the algorithmic bytes (33,792 B/k written by the Fourier sums,
# 41,984 B/k read by the fused band sums) to 0.05 %.
SEP_XY_TRAFFIC_BYTES_PER_K = (0.873e9 + 8.862e9) / 262144.0
BANDSUM_TRAFFIC_BYTES_PER_K = (12.957e9 + 0.004e9) / 262144.0        # V read from HBM
BANDSUM_FUSED_TRAFFIC_BYTES_PER_K = (11.011e9 + 0.004e9) / 262144.0  # dH and U read


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload():
    from mepinterp_amd import models
    m = models.make_c2(seed=2)
    p = dict(do_ahc=True, do_kubo=True, do_mep=True, n_eF=1, eF_min=0.0, eF_max=0.0, n_hw=1,
             hw_min=0.0, hw_max=0.0, T_kelvin=300.0, eta=1e-3 / AU_TO_EV, valence_bands=16)
    return m, p, (64, 64, 64)


def _host_cpus():
    """CPUs this process may run on (the box's share), the model name and the
    physical cores the machine reports."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    model, phys = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            cur = {}
            for line in f:
                if ":" not in line:
                    if cur:
                        phys.add((cur.get("physical id"), cur.get("core id")))
                    cur = {}
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                cur[k] = v
                if k == "model name":
                    model = v
            if cur:
                phys.add((cur.get("physical id"), cur.get("core id")))
    except OSError:
        pass
    return allowed, model, len(phys)


def _oracle_params(O, p, threads):
    return O.RunParams(do_ahc=p["do_ahc"], do_kubo=p["do_kubo"], do_mep=p["do_mep"],
                       n_eF=p["n_eF"], eF_min=p["eF_min"], eF_max=p["eF_max"], n_hw=p["n_hw"],
                       hw_min=p.get("hw_min", 0.0), hw_max=p.get("hw_max", 0.0),
                       T_kelvin=p["T_kelvin"], eta=p["eta"], valence_bands=p["valence_bands"],
                       n_threads=threads)


def cpu_baseline(m, p, mp, seconds=12.0):
    """The oracle (FP64, reference loop structure, -O3) on the host cores,
    over a fixed seeded sample of the same mesh's k-points: once on every
    CPU of this process's share (at most 16, the GPU box's share per GPU)
    and once on one core.  Returns (record, sample k-points, oracle sums)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    allowed, model, phys = _host_cpus()
    threads = max(1, min(16, allowed))
    om = O.Model(m["R_frac"], m["H"], a_latt=m["a_latt"], kubo_tol=1e-5)
    rp = _oracle_params(O, p, threads)
    mesh = O.mp_mesh(mp)
    rng = np.random.default_rng(1234)
    probe = mesh[:, rng.choice(mesh.shape[1], size=threads * 2, replace=False)]
    t0 = time.perf_counter()
    O.accumulate(om, rp, probe)
    per_k = (time.perf_counter() - t0) / probe.shape[1]
    nk = int(max(threads * 4, min(mesh.shape[1], seconds / max(per_k, 1e-9))))
    nk = (nk // threads) * threads
    sample = np.asfortranarray(mesh[:, rng.choice(mesh.shape[1], size=nk, replace=False)])
    t0 = time.perf_counter()
    sums = O.accumulate(om, rp, sample)
    dt = time.perf_counter() - t0
    # one core, on the first points of the same sample (about a tenth of the time)
    n1 = max(4, int(0.1 * seconds / max(per_k * threads, 1e-9)))
    t0 = time.perf_counter()
    O.accumulate(om, _oracle_params(O, p, 1), np.asfortranarray(sample[:, :n1]))
    dt1 = time.perf_counter() - t0
    rec = {"value": nk / dt, "unit": "k-pts/s", "cores": threads, "kind": "port",
           "one_core_value": n1 / dt1, "cpu_model": model, "physical_cores_seen": phys,
           "cpus_allowed": allowed,
           "sample": f"{nk} seeded random k-points of the 64^3 mesh, same jobs, {dt:.1f} s on "
                     f"{threads} host threads (OpenMP over k); one core: the first {n1} of them "
                     f"in {dt1:.1f} s"}
    return rec, sample, sums


def oracle_error(eng, rp_gpu, p, sample, o):
    """GPU run_kpts on the oracle's sample against the oracle's sums: relative
    error of sigma_xy, the AHC tensor, the Kubo alpha tensor (IC + LC + CS at
    the first E_F) and the Niu alpha tensor (summed over the valence bands)."""
    g = eng.run_kpts(sample, rp_gpu)
    nv = p["valence_bands"]

    def tens(s):
        return {"sigma_xy": np.array([s["ahc"][0, 1, 0]]), "ahc_tens": s["ahc"][:, :, 0],
                "kubo_alpha": (s["kubo_ic"] + s["kubo_lc"] + s["kubo_cs"])[:, :, 0],
                "niu_alpha": (s["niu_ic"] + s["niu_lc"] + s["niu_cs"])[:, :, :nv].sum(axis=2)}
    tg, to = tens(g), tens(o)
    return {k: float(np.max(np.abs(tg[k] - to[k])) / max(np.max(np.abs(to[k])), 1e-300))
            for k in tg}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(k-blocks would be sharded across GPUs with one final allreduce of the sums)",
              file=sys.stderr)
        sys.exit(2)

    from mepinterp_amd import capi, host
    m, p, mp = workload()
    nk = int(np.prod(mp))
    n, nR = m["H"].shape[0], m["H"].shape[2]
    eng = capi.Engine(m["R_frac"], m["H"], a_latt=m["a_latt"], kubo_tol=1e-5)
    rp = host.run_params(p)
    for _ in range(args.warmup):
        eng.run_mesh(mp, rp)
    times = []
    sums = None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        sums = eng.run_mesh(mp, rp)
        times.append(time.perf_counter() - t0)
    t_med = statistics.median(times)
    value = nk / t_med
    log(f"[bench] step times (s): {['%.4f' % t for t in times]}")

    # kernel split from HIP events on the engine's stream (separate pass)
    eng.set_profiling(True)
    eng.reset_times()
    eng.run_mesh(mp, rp)
    split = {}
    for name in ("kgen", "phase", "ft_gemm", "eigh", "velo", "bandsum", "reduce"):
        ms, cnt = eng.kernel_time(name)
        split[name] = (ms, cnt)
        log(f"[bench] {name:8s} total {ms:9.3f} ms over {cnt} launches")
    eng.set_profiling(False)
    # per-stage rooflines (algorithmic work per k of the algorithm each kernel runs)
    launches = max(split["bandsum"][1], 1)
    k_per_launch = nk / launches
    n_up = n * (n + 1) // 2
    survey_flop_k = 8.0 * nR * 4 * n * n       # direct complex sum (8 NR n_mat n^2)
    ft_bytes_k = 16.0 * 4 * n_up                # H(k) + 3 dH(k) upper triangles written
    bs_flop_k = 48.0 * n ** 3                   # 3 products C_j = V_k B_l - V_l B_k (2 each)
    bs_bytes_k = 48.0 * n * n + 8.0 * n         # V (c128) + E (f64) read once
    velo_flop_k = 48.0 * n ** 3                 # U^H (dH_a U) for a = x, y, z
    # 17 <= n <= 32: the band-sum kernel forms V itself (no velocity launch);
    # it then reads dH_a (packed) and U instead of V
    fused = split["velo"][1] == 0
    if fused:
        bs_flop_k += velo_flop_k
        bs_bytes_k = 16.0 * 3 * n_up + 16.0 * n * n + 8.0 * n
    stages = {}
    for name, fl in (("velo", velo_flop_k), ("bandsum", bs_flop_k)):
        ms, cnt = split[name]
        if cnt == 0:
            stages[name] = {"fused_into": "bandsum_kernel<32, true>"}
            continue
        ach = fl * nk / (ms * 1e-3) / 1e12
        stages[name] = {"bound": "mfma", "flop_per_k": fl, "achieved_tflops": ach,
                        "frac": ach / PEAK_FP64_TFLOPS, "launch_ms": ms / max(cnt, 1)}
    ms, cnt = split["ft_gemm"]
    ach = ft_bytes_k * nk / (ms * 1e-3) / 1e9
    stages["fourier"] = {"bound": "hbm", "bytes_per_k": ft_bytes_k, "achieved_gbs": ach,
                         "frac": ach / PEAK_HBM_GBS, "launch_ms": ms / max(cnt, 1),
                         "direct_sum_flop_per_k": survey_flop_k,
                         "traffic_per_launch": SEP_XY_TRAFFIC_BYTES_PER_K * k_per_launch}
    stages["bandsum"]["hbm_bytes_per_k"] = bs_bytes_k
    stages["bandsum"]["achieved_gbs"] = bs_bytes_k * nk / (split["bandsum"][0] * 1e-3) / 1e9
    stages["bandsum"]["traffic_per_launch"] = (
        BANDSUM_FUSED_TRAFFIC_BYTES_PER_K if fused else BANDSUM_TRAFFIC_BYTES_PER_K) * k_per_launch
    stages["eigh"] = {"bound": "latency", "ms": split["eigh"][0]}
    # the dominant single kernel (eigh is three kernels)
    dom = max(("ft_gemm", "velo", "bandsum"), key=lambda nm: split[nm][0])
    d = stages["fourier" if dom == "ft_gemm" else dom]
    if dom == "ft_gemm":
        roofline = {"bound": "hbm", "achieved": d["achieved_gbs"], "peak": PEAK_HBM_GBS,
                    "unit": "GB/s", "frac": d["frac"], "traffic": d.get("traffic_per_launch"),
                    "kernel": "sep_xy_kernel<8, zd>", "launch_ms": d["launch_ms"],
                    "stages": stages}
    else:
        roofline = {"bound": "mfma", "achieved": d["achieved_tflops"], "peak": PEAK_FP64_TFLOPS,
                    "unit": "TFLOP/s", "frac": d["frac"], "traffic": d.get("traffic_per_launch"),
                    "frac_basis": "algorithmic complex-product flops (8 real flops per complex "
                                  "multiply-add, 48 n^3 velocity + 48 n^3 band sums per k); the "
                                  "kernel issues 3 real MFMA products per complex product and "
                                  "runs its pair sums on the VALU, so the MFMA pipe is busy for "
                                  "at most 3/4 of frac",
                    "mfma_issued_frac_bound": 0.75 * d["frac"],
                    "kernel": {"velo": "velo_mfma_kernel<zd>",
                               "bandsum": "bandsum_kernel<32, %s>" % ("true" if fused else "false")}[dom],
                    "launch_ms": d["launch_ms"], "stages": stages}
    out = host.finalize(sums, p, mp, m["a_latt"])
    # the mixed-precision path on the same mesh: throughput and error per
    # quantity against this FP64 run
    rpm = host.run_params(p)
    rpm.precision = capi.MEP_PREC_MIXED
    eng.run_mesh(mp, rpm)
    tm = []
    msums = None
    for _ in range(max(1, min(args.steps, 3))):
        t0 = time.perf_counter()
        msums = eng.run_mesh(mp, rpm)
        tm.append(time.perf_counter() - t0)
    mout = host.finalize(msums, p, mp, m["a_latt"])

    def relerr(a, b):
        a, b = np.asarray(a), np.asarray(b)
        return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
    mixed = {"k_pts_s": nk / statistics.median(tm), "ms_per_step": statistics.median(tm) * 1e3,
             "speedup_vs_fp64": t_med / statistics.median(tm),
             "rel_err_vs_fp64": {"sigma_xy": abs(mout["ahc_tens"][0, 1] - out["ahc_tens"][0, 1])
                                 / abs(out["ahc_tens"][0, 1]),
                                 "ahc_tens": relerr(mout["ahc_tens"], out["ahc_tens"]),
                                 "mep_tens_niu": relerr(mout["mep_tens"], out["mep_tens"]),
                                 "kubo_mep_tens": relerr(mout["kubo_mep_tens"], out["kubo_mep_tens"])}}
    log(f"[bench] mixed precision {mixed}")
    log(f"[bench] sigma_xy = {out['ahc_tens'][0, 1]:.10e}  alpha_Niu_xx = {out['mep_tens'][0, 0]:.6e}")

    cpu = None
    sample = osums = None
    if not args.no_cpu:
        try:
            cpu, sample, osums = cpu_baseline(m, p, mp, seconds=args.cpu_seconds)
            log(f"[bench] cpu baseline {cpu}")
        except Exception as e:  # the baseline is reported, never the product
            log(f"[bench] cpu baseline failed: {e}")
    if sample is None:  # no CPU leg: a small sample for the error check only
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        rng = np.random.default_rng(1234)
        mesh = O.mp_mesh(mp)
        sample = np.asfortranarray(mesh[:, rng.choice(mesh.shape[1], size=2048, replace=False)])
        osums = O.accumulate(O.Model(m["R_frac"], m["H"], a_latt=m["a_latt"], kubo_tol=1e-5),
                             _oracle_params(O, p, max(1, min(16, _host_cpus()[0]))), sample)
    # the metric's error half: GPU sums against the FP64 oracle on the same k-points
    err = {"n_k": int(sample.shape[1]), "fp64": oracle_error(eng, rp, p, sample, osums),
           "mixed": oracle_error(eng, rpm, p, sample, osums)}
    log(f"[bench] error vs the FP64 oracle on {sample.shape[1]} k-points: {err}")
    mixed["rel_err_vs_oracle"] = err["mixed"]
    res = {
        "metric": "k-pts/s (interp+eigh+Berry) at n_wf=32",
        "value": value,
        "unit": "k-pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_med * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "n_wf=32 NR=343 ({-3..3}^3) seeded exchange model, 64^3 MP mesh, "
                               "AHC+velo/ohc (n_hw=1), Kubo MEP, Niu MEP (16 valence), T=300 K, "
                               "E_F=0, FP64 path",
                   "n_wf": n, "n_R": nR, "mp_grid": list(mp), "n_k": nk,
                   "kernel_ms": {k: v[0] for k, v in split.items()}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "oracle_rel_err": err,
        "mixed_precision": mixed,
    }
    print(json.dumps(res), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
