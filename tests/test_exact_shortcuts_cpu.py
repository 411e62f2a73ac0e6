"""Arithmetic shortcuts the kernels take, checked bit for bit on the CPU
(numpy rounds FP64 and FP32 division the way the gfx950 kernels do: IEEE,
round to nearest).

visc_spectral_max (tmp_amd/csrc/solver.hip) computes only gamma/rho for
rho > 0 in place of max(4/(3 rho), gamma/rho) (ViscousFlow2D.cpp:243-249,
TopologyHolderStrategy.cpp:217-323)."""
import numpy as np

GAMMA = 1.4


def _smax(a, b):
    return np.where(a < b, b, a)


def _densities(dtype):
    rng = np.random.default_rng(3)
    fi = np.finfo(dtype)
    vals = [rng.uniform(1e-6, 10.0, 200000), 10.0 ** rng.uniform(-300 if dtype == np.float64 else -37,
                                                                   300 if dtype == np.float64 else 37, 200000),
            np.array([fi.tiny, fi.tiny / 2 ** 10, fi.smallest_subnormal, fi.max, fi.max / 3, fi.max / 2.9,
                      np.inf, 1.0, 0.75, 4.0 / 3.0])]
    return np.concatenate(vals).astype(dtype)


def test_gamma_over_rho_is_the_larger_term_for_positive_density():
    with np.errstate(over="ignore", divide="ignore", under="ignore"):
        for dt in (np.float64, np.float32):
            rho = _densities(dt)
            g = dt(GAMMA)
            full = _smax(dt(4) / (dt(3.0) * rho), g / rho)
            short = g / rho
            assert np.array_equal(full.view(np.uint64 if dt == np.float64 else np.uint32),
                                  short.view(np.uint64 if dt == np.float64 else np.uint32))
