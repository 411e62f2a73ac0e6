"""The learned warm start's fallback and screen on the GPU: predictions
corrupted on the host -- a NaN, a negative pressure, a finite field far
from the solution -- are refused, and those cases then follow the cold
start (the free stream) bit for bit, while a sound prediction is kept."""
import numpy as np
import pytest

from tmp_amd import usg, warmstart

pytestmark = pytest.mark.gpu


def test_screen_keeps_sound_and_reverts_corrupted_predictions():
    cases = warmstart.family_cases(4, 11)
    ms = [warmstart.family_mesh(c) for c in cases]
    ref = warmstart.batch_solver(ms, cases)
    ref.lusgs(3000)                               # a field near the solution stands in for a prediction
    U = ref.U
    cell = ref.case_cell
    U[cell[1] + 7, 2] = np.nan                     # case 1: not finite
    U[cell[2] + 11, 1] = -0.3                      # case 2: negative pressure
    U[cell[3]:cell[4], 2:] *= -3.0                 # case 3: finite, reversed and tripled velocity
    s = warmstart.batch_solver(ms, cases)
    kept, r_cold, nprobe = warmstart.screened_start(s, U)
    print(f"kept {kept.tolist()} after {nprobe} probe iterations; cold first residuals {r_cold}", flush=True)
    assert kept.tolist() == [True, False, False, False]
    assert np.all(s.status() == 0)
    # the reverted cases now run exactly as cold starts do
    s.lusgs(40)
    c = warmstart.batch_solver(ms, cases)
    c.lusgs(40)
    for i in (1, 2, 3):
        sl = s.case_slice(i)
        assert np.array_equal(s.U[sl], c.U[sl])
    # the kept case stays ahead of the cold start
    assert s.res_l2()[0] < 0.1 * c.res_l2()[0]


def test_screen_on_a_single_case_context():
    """A single-case context fails the probe call when a check breaks; the
    screen then restarts it from the free stream."""
    case = warmstart.family_cases(1, 12)[0]
    m = warmstart.family_mesh(case)
    s = usg.Solver(m, flow=usg.LAMINAR, mach=case["mach"], aoa=case["aoa"], reynolds=case["re"])
    U = np.tile(s.freestream(), (m.nc, 1))
    U[:, 2:] *= -3.0
    kept, r_cold, _ = warmstart.screened_start(s, U)
    assert kept.tolist() == [False]
    assert np.array_equal(s.U, np.tile(s.freestream(), (m.nc, 1)))


def test_saved_network_predicts_the_same_bits(tmp_path):
    """save_net / load_net carry a network between processes: the loaded
    one predicts exactly what the saved one does, and the training figures
    come back with it."""
    net = warmstart._mlp.MLP(warmstart.NIN, 256, 8, warmstart.NOUT)
    net.set_params(*warmstart._init(warmstart.NIN, 256, 8, warmstart.NOUT, 77))
    case = warmstart.family_cases(1, 13)[0]
    x = warmstart.features(warmstart.family_mesh(case), case)
    path = str(tmp_path / "net.npz")
    warmstart.save_net(path, net, 10, 9, len(x), 3, [0.5, 0.25, 0.125], 1.5)
    loaded, losses, t_train, meta = warmstart.load_net(path)
    assert meta == (10, 9, len(x), 3) and losses == [0.5, 0.25, 0.125] and t_train == 1.5
    assert np.array_equal(net.predict(x), loaded.predict(x))
