"""k-block autotuner for mep_run_mesh / mep_run_batch.

For a model of (n_wf, NR) at a given precision, the tuner times the full run
of the model's own mesh for a few k-block sizes (0 = the library's default:
the largest block within the context's workspace limit) and keeps the fastest.  Choices are
cached in a JSON file keyed by "n_wf/NR/precision/mesh size", so a workload is
measured once: $MEPINTERP_TUNE_CACHE, else ~/.cache/mepinterp_amd/tune.json.
"""
from __future__ import annotations

import json
import os
import time

CANDIDATES = (0, 8192, 32768, 131072)


def cache_path() -> str:
    return os.environ.get("MEPINTERP_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "mepinterp_amd",
                                       "tune.json"))


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, table):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def key(n_wf: int, n_R: int, precision: int, n_k: int) -> str:
    return f"{n_wf}/{n_R}/{precision}/{n_k}"


def tune(engine, mp_grid, params, n_R: int, reps: int = 2, path: str | None = None) -> int:
    """Pick (and set on `engine`) the fastest k-block size for this workload."""
    path = path or cache_path()
    n_k = int(mp_grid[0]) * int(mp_grid[1]) * int(mp_grid[2])
    k = key(engine.n, n_R, int(params.precision), n_k)
    table = _load(path)
    if k in table:
        engine.set_kblock(int(table[k]["kblock"]))
        return int(table[k]["kblock"])
    times = {}
    for kb in CANDIDATES:
        if kb and kb >= 2 * n_k:
            continue  # one block either way: the default covers it
        engine.set_kblock(kb)
        engine.run_mesh(mp_grid, params)  # allocation and warm-up
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            engine.run_mesh(mp_grid, params)
            best = min(best, time.perf_counter() - t0)
        times[kb] = best
    kb = min(times, key=times.get)
    engine.set_kblock(kb)
    table = _load(path)
    table[k] = {"kblock": kb, "seconds": {str(a): b for a, b in times.items()}}
    _store(path, table)
    return kb
