"""Algorithmic byte counts used for the roofline figures.

Each count assumes every input value is read from HBM once per launch and
every output written once (ideal caching of neighbour gathers), FP64 = 8 B,
int32 = 4 B.  ``r`` is faces per cell (nf/nc) of the mesh.

Dominant kernel ``k_flux`` (residual of one cell), per cell:
  * own-cell inputs:  U 32, grad 64, phi 32, cell centre 16 (all flows)
                      + gradT 16 (viscous)
                      + tU 16, tgrad 32, F1 8, mut 8, vol 8, wdist 8 (SST)
  * face slots: 4 x (face id + neighbour id) = 32
  * face data: normal 16 + length 8 + midpoint 16 = 40 per face
  * outputs: R 32 (+ turbulence residual 16 for SST)

Whole LU-SGS iteration (the cell-iteration model of the reference path's
six inherent stages -- gradient, limiter, flux, forward sweep, backward
sweep, update): Euler 992 + 200 r, laminar 1064 + 200 r, SST 1536 + 200 r.
"""

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def flux_kernel_bytes_per_cell(flow: int, r: float) -> float:
    own = 32 + 64 + 32 + 16
    if flow >= 1:
        own += 16
    if flow == 2:
        own += 16 + 32 + 8 + 8 + 8 + 8
    out = 32 + (16 if flow == 2 else 0)
    return own + 32 + 40.0 * r + out


def iteration_bytes_per_cell(flow: int, r: float) -> float:
    base = {0: 992.0, 1: 1064.0, 2: 1536.0}[flow]
    return base + 200.0 * r
