"""Launch-parameter autotuner for the transport kernels.

    best = tune_deck("deck.xml", precision=capi.PRECISION_FP32_TRACK)
    with capi.Driver("deck.xml", **best) as d: ...

The knobs are the persistent kernel's occupancy (blocks per CU, i.e. how many
photons are in flight at once), the block size, the refill threshold (idle
lanes of a wave that trigger a refill), the per-wave batch of photon indices
claimed at a time, and the tally form.  Each candidate is timed by the
transport kernel's own HIP events on a short sample of the workload (a few
steps at a reduced photon count); the search is coordinate descent from the
defaults, two sweeps, so the result is never slower than the defaults as
measured.  Results are cached per workload signature in a JSON file.
"""
from __future__ import annotations

import json
import os
from typing import Callable

from . import capi

CANDIDATES = {
    "blocks_per_cu": [0, 4, 8, 12],  # 0: the blocks that stay resident together
    "block_size": [128, 256],
    "refill_threshold": [4, 8, 16, 32],
    "batch_photons": [64, 256, 1024],
    "tally_mode": [0, 1, 4],
}
DEFAULTS = {"blocks_per_cu": 0, "block_size": 256, "refill_threshold": 16,
            "batch_photons": 256, "tally_mode": 0}


def default_cache_path() -> str:
    base = os.environ.get("BH_TUNE_CACHE")
    if base:
        return base
    root = os.environ.get("XDG_CACHE_HOME", os.path.join(os.path.expanduser("~"), ".cache"))
    return os.path.join(root, "branson_amd", "tuned.json")


def coordinate_descent(measure: Callable[[dict], float], candidates=None, start=None,
                       sweeps: int = 2) -> tuple[dict, float, list]:
    """Minimise measure(params) one knob at a time.  Returns (best params,
    best cost, [(params, cost), ...] of every evaluation)."""
    cand = candidates or CANDIDATES
    best = dict(start or DEFAULTS)
    seen: dict = {}
    log = []

    def cost(p):
        key = tuple(sorted(p.items()))
        if key not in seen:
            seen[key] = measure(dict(p))
            log.append((dict(p), seen[key]))
        return seen[key]

    best_cost = cost(best)
    for _ in range(sweeps):
        changed = False
        for knob, values in cand.items():
            for v in values:
                if v == best[knob]:
                    continue
                trial = dict(best, **{knob: v})
                c = cost(trial)
                if c < best_cost:
                    best, best_cost, changed = trial, c, True
        if not changed:
            break
    return best, best_cost, log


def _signature(deck, precision, algorithm, n_groups, n_photons):
    return (f"{os.path.basename(deck)}|G{n_groups}|prec{precision}|algo{algorithm}|"
            f"N{n_photons}")


def tune_deck(deck: str, precision: int = capi.PRECISION_FP64, n_groups: int = 1,
              n_photons: int = 0, sample_photons: int = 2_000_000, sample_steps: int = 2,
              cache: str | None = None, retune: bool = False, fp32_tol: float = 0.0) -> dict:
    """Best persistent-kernel options for a deck (keyword arguments for
    capi.Driver / capi.Engine).  Timed on `sample_steps` steps at
    `sample_photons` photons per step."""
    algorithm = capi.ALGO_PERSISTENT
    path = cache or default_cache_path()
    key = _signature(deck, precision, algorithm, n_groups, n_photons)
    table = {}
    if os.path.exists(path):
        with open(path) as f:
            table = json.load(f)
        if key in table and not retune:
            return dict(table[key]["params"], algorithm=algorithm, precision=precision)

    def measure(p):
        with capi.Driver(deck, precision=precision, algorithm=algorithm, n_groups=n_groups,
                         n_photons=sample_photons, fp32_tol=fp32_tol, **p) as d:
            ms = 0.0
            for _ in range(sample_steps):
                if d.finished():
                    break
                d.prepare()
                d.transport()
                ms += d.finish()["transport_ms"]
        return ms

    best, cost, log = coordinate_descent(measure)
    table[key] = {"params": best, "transport_ms": cost,
                  "default_ms": log[0][1], "evaluations": len(log)}
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1)
    os.replace(tmp, path)
    return dict(best, algorithm=algorithm, precision=precision)
