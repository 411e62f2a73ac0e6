"""usg_stencil_apply_host_noalloc (the stencil apply staged in the
library's fixed device array, no allocation) against usg_stencil_apply_host
(which allocates its device copies): bit-identical on two calls in a row,
the second with a smaller input, in both layouts; oversized inputs are
refused before any device work."""
import numpy as np
import pytest

from tmp_amd import usg

pytestmark = pytest.mark.gpu

_dp = usg._dp
_ip = usg._ip


def _case(rng, ns, nd, nvar, nrecv):
    src = rng.normal(size=(ns, nvar))
    dst = rng.normal(size=(nd, nvar))
    recv = rng.choice(nd, size=nrecv, replace=False).astype(np.int32)
    nw = rng.choice([3, 4], size=nrecv).astype(np.int32)
    donor = rng.integers(0, ns, size=(nrecv, 4)).astype(np.int32)
    w = rng.uniform(0, 1, size=(nrecv, 4))
    w[nw == 3, 3] = 0.0
    w /= w.sum(axis=1, keepdims=True)
    return src, dst, recv, nw, donor, w


def _call(fn, layout, src, dst, recv, nw, donor, w):
    ns, nvar = src.shape
    nd = dst.shape[0]
    s = np.ascontiguousarray(src if layout == usg.ROW else src.T)
    d = np.ascontiguousarray(dst if layout == usg.ROW else dst.T)
    recv, nw = np.ascontiguousarray(recv, np.int32), np.ascontiguousarray(nw, np.int32)
    donor, w = np.ascontiguousarray(donor, np.int32), np.ascontiguousarray(w, np.float64)
    rc = fn(s.ctypes.data_as(_dp), d.ctypes.data_as(_dp), nvar, layout, ns, nd, len(recv),
            recv.ctypes.data_as(_ip), nw.ctypes.data_as(_ip), donor.ctypes.data_as(_ip), w.ctypes.data_as(_dp))
    assert rc == 0, usg.lib().usg_last_error()
    return d if layout == usg.ROW else d.T


@pytest.mark.parametrize("layout", [usg.ROW, usg.COLUMN], ids=["row", "column"])
def test_noalloc_equals_allocating_entry_point(layout):
    L = usg.lib()
    rng = np.random.default_rng(71)
    # two calls in a row, the second smaller: the staging area keeps the
    # first call's contents, which the second must not read
    for ns, nd, nvar, nrecv in [(60000, 40000, 6, 9000), (700, 500, 4, 90)]:
        case = _case(rng, ns, nd, nvar, nrecv)
        ref = _call(L.usg_stencil_apply_host, layout, *case)
        out = _call(L.usg_stencil_apply_host_noalloc, layout, *case)
        assert np.array_equal(out, ref)
        # receptors changed, every other point kept
        keep = np.setdiff1d(np.arange(nd), case[2])
        assert np.array_equal(out[keep], case[1][keep])


def test_noalloc_refuses_arrays_beyond_the_staging_area():
    L = usg.lib()
    ns, nvar = 2_000_000, 1  # 16 MB of source alone, plus the destination
    src = np.zeros(ns)
    dst = np.zeros(ns)
    recv, nw = np.zeros(1, np.int32), np.full(1, 3, np.int32)
    donor, w = np.zeros(4, np.int32), np.array([1.0, 0, 0, 0])
    rc = L.usg_stencil_apply_host_noalloc(src.ctypes.data_as(_dp), dst.ctypes.data_as(_dp), nvar, usg.ROW, ns, ns,
                                          1, recv.ctypes.data_as(_ip), nw.ctypes.data_as(_ip),
                                          donor.ctypes.data_as(_ip), w.ctypes.data_as(_dp))
    assert rc == -1
    assert b"16 MiB staging area" in L.usg_last_error()
