"""Algorithmic byte counts used for the roofline figures.

Each count assumes every input value is read from HBM once per launch and
every output written once (ideal caching of neighbour gathers), FP64 = 8 B,
int32 = 4 B.  ``r`` is faces per cell (nf/nc) of the mesh.

Dominant kernel ``k_face`` (every flux of one face), per launch over the
whole mesh = nc x cell term + nf x face term:
  * cell inputs (each cell's values read once): U 32, grad 64, phi 32,
    cell centre 16 (all flows) + gradT 16, mu 8 (viscous) + mut 8, tU 16,
    F1 8, tgrad 32 (SST)                -> Euler 144, laminar 168, SST 232
  * face inputs: left/right cell 8, normal 16, length 8, midpoint 16 = 48
    (+ centre distance 8, viscous)
  * face outputs: flux 4 + spectral term 1 (Euler), + viscous flux 4 +
    viscous spectral term 1 (laminar), + SST flux 2, all + the two LU-SGS
    r_A values 2: 56 / 96 / 112
  so per cell: Euler 144 + 104 r, laminar 168 + 152 r, SST 232 + 168 r.

Mixed precision (FP32 gradients, limiter, face geometry and face record;
FP64 state, and FP64 normal and r_A for the LU-SGS terms):
  * cell inputs: U 32, grad 32, phi 16 (+ gradT 8, mu 8) (+ mut 8, tU 16,
    F1 4, tgrad 16)                      -> Euler 80, laminar 96, SST 140
  * face inputs: left/right cell 8, normal + length 12, offsets 16, FP64
    normal 16 (+ centre-to-centre 8, centre distance 8, viscous)
                                         -> Euler 52, laminar/SST 68
  * face outputs: 5 / 10 / 12 floats + 2 doubles -> 36 / 56 / 64

Fused face-cell kernel ``k_facecell`` (FP64: every face of a tile of cells
into LDS, then each cell's ordered gather), per launch = nc x cell term +
nf x face term, each input once:
  * cell inputs: k_face's (144 / 168 / 232) + the cell's face slots 16,
    volume 8, local face indices 8                  -> 176 / 200 / 264
  * cell outputs: R 32, Lambda_c 8, dt 8, D 8 (+ Lambda_v 8) (+ the SST
    residual 16 and diagonal 16)                      -> 56 / 64 / 96
  * face inputs: k_face's 48 (+ 8 viscous) + the tile face list 4
  * face outputs: the two LU-SGS r_A values 16 (the face record stays in LDS)
  so per cell: Euler 232 + 68 r, laminar 264 + 76 r, SST 360 + 76 r.

Whole LU-SGS iteration (the cell-iteration model of the reference path's
six inherent stages -- gradient, limiter, flux, forward sweep, backward
sweep, update): Euler 992 + 200 r, laminar 1064 + 200 r, SST 1536 + 200 r.
"""

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def face_kernel_bytes_per_cell(flow: int, r: float, mixed: bool = False) -> float:
    if mixed:
        cell = {0: 80, 1: 96, 2: 140}[flow]
        face = {0: 52 + 36, 1: 68 + 56, 2: 68 + 64}[flow]
        return cell + face * r
    cell = 32 + 64 + 32 + 16
    face_in = 48
    if flow >= 1:
        cell += 16 + 8
        face_in += 8
    if flow == 2:
        cell += 8 + 16 + 8 + 32
    face_out = {0: 40, 1: 80, 2: 96}[flow] + 16
    return cell + (face_in + face_out) * r


def iteration_bytes_per_cell(flow: int, r: float) -> float:
    base = {0: 992.0, 1: 1064.0, 2: 1536.0}[flow]
    return base + 200.0 * r


def fused_kernel_bytes_per_cell(flow: int, r: float) -> float:
    cell = {0: 176 + 56, 1: 200 + 64, 2: 264 + 96}[flow]
    face = (48 if flow == 0 else 56) + 4 + 16
    return cell + face * r
