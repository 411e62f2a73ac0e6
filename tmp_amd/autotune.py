"""Launch-size autotuner with a persistent cache.

``tune(solver)`` picks the threads per block of the four kernel groups
(gradient/limiter, face flux, cell gather, sweep waves) for a solver's
problem shape: it looks the shape up in a JSON cache and, on a miss, runs
the library's timed search (``usg_autotune``: coordinate descent over 64,
128 and 256 threads, the state restored afterwards) and stores the result.
In FP64 the face fluxes and cell gathers run as one fused launch whose
tiles are the cell group's blocks, so the cell group's size is the face
tile: 64, 128 or 256 cells and every face they touch.
The key is (GPU arch, the launch scheme, flow type, precision, number of cases, cells rounded
to 2 significant digits, faces/cells ratio rounded to 0.1).
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
    return "|".join([device_arch(), "fused-tiles", str(solver.flow), str(precision), str(solver.ncase),
                     str(_round2(solver.nc)), f"{nf / solver.nc:.1f}"])


def load(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tune(solver, precision: int, nf: int, iters: int = 5, cache: str | None = DEFAULT_CACHE,
         force: bool = False) -> tuple[list, bool]:
    """Apply the cached (or freshly searched) block sizes; returns (blocks, hit)."""
    key = shape_key(solver, precision, nf)
    table = load(cache) if cache else {}
    if not force and key in table:
        solver.set_blocks(table[key])
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
