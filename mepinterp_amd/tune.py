"""Autotuner for mep_run_mesh / mep_run_batch: k-block size and kernel variants.

For a model of (n_wf, NR) at a given precision, the tuner times the full run
of the model's own mesh
  1. for each eigensolver variant the library offers at this n_wf (the FP64
     Householder reduction, MEP_VAR_EIGH: for 16 < n_wf <= 32 one or two
     k-points per wavefront; for 32 < n_wf <= 128 the matrix in a workgroup's
     LDS, in one wavefront's LDS (n_wf <= 64) or in a workgroup's registers),
     at the default k-block, and keeps the faster;
  2. then for a few k-block sizes (0 = the library's default: the largest
     block within the context's workspace limit) with that variant.
Choices are cached in a JSON file keyed by "n_wf/NR/precision/mesh size", so a
workload is measured once: $MEPINTERP_TUNE_CACHE, else
~/.cache/mepinterp_amd/tune.json.  Every variant gives the same sums to
rounding, so the choice is a matter of time only.
"""
from __future__ import annotations

import json
import os
import time

from . import capi

KBLOCKS = (0, 8192, 32768, 131072)


def eigh_variants(n_wf: int, precision: int) -> tuple:
    """MEP_VAR_EIGH values that select different kernels at this size."""
    if precision != capi.MEP_PREC_FP64:
        return (0,)
    if 16 < n_wf <= 32:
        return (1, 2)
    if 32 < n_wf <= 64:
        return (1, 2, 3)
    if 64 < n_wf <= 128:
        return (1, 3)
    return (0,)


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


def _apply(engine, choice):
    engine.set_variant(capi.MEP_VAR_EIGH, int(choice.get("eigh", 0)))
    engine.set_kblock(int(choice["kblock"]))


def _time(engine, mp_grid, params, reps):
    engine.run_mesh(mp_grid, params)  # allocation and warm-up
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        engine.run_mesh(mp_grid, params)
        best = min(best, time.perf_counter() - t0)
    return best


def tune(engine, mp_grid, params, n_R: int, reps: int = 2, path: str | None = None) -> dict:
    """Pick (and set on `engine`) the fastest eigensolver variant and k-block
    size for this workload; returns {"kblock": .., "eigh": ..}."""
    path = path or cache_path()
    n_k = int(mp_grid[0]) * int(mp_grid[1]) * int(mp_grid[2])
    prec = int(params.precision)
    k = key(engine.n, n_R, prec, n_k)
    table = _load(path)
    if k in table and "eigh" in table[k]:
        _apply(engine, table[k])
        return {"kblock": int(table[k]["kblock"]), "eigh": int(table[k]["eigh"])}
    engine.set_kblock(0)
    t_var = {}
    for v in eigh_variants(engine.n, prec):
        engine.set_variant(capi.MEP_VAR_EIGH, v)
        t_var[v] = _time(engine, mp_grid, params, reps)
    var = min(t_var, key=t_var.get)
    engine.set_variant(capi.MEP_VAR_EIGH, var)
    t_kb = {}
    for kb in KBLOCKS:
        if kb and kb >= 2 * n_k:
            continue  # one block either way: the default covers it
        engine.set_kblock(kb)
        t_kb[kb] = t_var[var] if kb == 0 else _time(engine, mp_grid, params, reps)
    kb = min(t_kb, key=t_kb.get)
    choice = {"kblock": kb, "eigh": var}
    _apply(engine, choice)
    table = _load(path)
    table[k] = dict(choice, seconds_eigh={str(a): b for a, b in t_var.items()},
                    seconds_kblock={str(a): b for a, b in t_kb.items()})
    _store(path, table)
    return choice
