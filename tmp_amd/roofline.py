"""Algorithmic byte counts used for the roofline figures.

Each count assumes every input value is read from HBM once per launch and
every output written once (ideal caching of neighbour gathers), FP64 = 8 B,
int32 = 4 B.  ``r`` is faces per cell (nf/nc) of the mesh.

Dominant kernel ``k_face`` (every flux of one face), per launch over the
whole mesh = nc x cell term + nf x face term:
  * cell inputs (each cell's values read once): U 32, grad 64, phi 32,
    cell centre 16 (all flows) + gradT 16 (viscous) + mut 8, tU 16, F1 8,
    tgrad 32 (SST)                      -> Euler 144, laminar 160, SST 224
  * face inputs: left/right cell 8, normal 16, length 8, midpoint 16 = 48
  * face outputs: flux 4 + spectral term 1 (Euler), + viscous flux 4 +
    viscous spectral term 1 (laminar), + SST flux 2: 40 / 80 / 96
  so per cell: Euler 144 + 88 r, laminar 160 + 128 r, SST 224 + 144 r.

Mixed precision (FP32 gradients, limiter, face geometry and face record;
FP64 state):
  * cell inputs: U 32, grad 32, phi 16 (+ gradT 8) (+ mut 8, tU 16, F1 4,
    tgrad 16)                            -> Euler 80, laminar 88, SST 132
  * face inputs: left/right cell 8, normal + length 12, offsets 16
    (+ centre-to-centre 8, viscous)      -> Euler 36, laminar/SST 44
  * face outputs: 5 / 10 / 12 floats     -> 20 / 40 / 48

Whole LU-SGS iteration (the cell-iteration model of the reference path's
six inherent stages -- gradient, limiter, flux, forward sweep, backward
sweep, update): Euler 992 + 200 r, laminar 1064 + 200 r, SST 1536 + 200 r.
"""

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def face_kernel_bytes_per_cell(flow: int, r: float, mixed: bool = False) -> float:
    if mixed:
        cell = {0: 80, 1: 88, 2: 132}[flow]
        face = {0: 36 + 20, 1: 44 + 40, 2: 44 + 48}[flow]
        return cell + face * r
    cell = 32 + 64 + 32 + 16
    if flow >= 1:
        cell += 16
    if flow == 2:
        cell += 8 + 16 + 8 + 32
    face_out = {0: 40, 1: 80, 2: 96}[flow]
    return cell + (48 + face_out) * r


def iteration_bytes_per_cell(flow: int, r: float) -> float:
    base = {0: 992.0, 1: 1064.0, 2: 1536.0}[flow]
    return base + 200.0 * r
