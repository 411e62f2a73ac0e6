"""Launch autotuner with a persistent cache.

``tune(solver)`` picks the launch configuration for a solver's problem
shape: the threads per block of the four kernel groups (gradient/limiter,
face flux, cell gather, sweep waves) and, in FP64, whether the face fluxes
and cell gathers run as two launches or as one fused launch whose face tile
is the cell group's block (64, 128 or 256 cells and every face they touch).
It looks the shape up in a JSON cache and, on a miss, runs the library's
timed search (``usg_autotune``: coordinate descent over 64, 128 and 256
threads, both schemes for the cell group, the state restored afterwards)
and stores the result as ``Solver.launch_config()``.
The key is (GPU arch, "launch-scheme", flow type, precision, number of
cases, cells rounded to 2 significant digits, faces/cells ratio rounded to
0.1).
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os

from . import usg

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "tmp_amd", "autotune.json")


def device_arch() -> str:
    buf = C.create_string_buffer(128)
    usg._check(usg.lib().usg_device_arch(buf, len(buf)))
    return buf.value.decode()


def _round2(n: int) -> int:
    if n <= 0:
        return 0
    p = 10 ** max(0, int(math.floor(math.log10(n))) - 1)
    return int(round(n / p) * p)


def shape_key(solver, precision: int, nf: int) -> str:
    return "|".join([device_arch(), "launch-scheme", str(solver.flow), str(precision), str(solver.ncase),
                     str(_round2(solver.nc)), f"{nf / solver.nc:.1f}"])


def load(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tune(solver, precision: int, nf: int, iters: int = 5, cache: str | None = DEFAULT_CACHE,
         force: bool = False) -> tuple[list, bool]:
    """Apply the cached (or freshly searched) launch configuration; returns
    (config, hit)."""
    key = shape_key(solver, precision, nf)
    table = load(cache) if cache else {}
    if not force and key in table:
        solver.set_launch_config(table[key])
        return table[key], True
    best = solver.autotune(iters)
    if cache:
        table = load(cache)
        table[key] = best
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, cache)
    return best, False
