"""The allocation-free H-PLOC build (cwp_hploc_build_static: scratch in a fixed static device array) against the
allocating builder (cwp_hploc_build) through the C ABI: the same tree (canonical form, boxes and leaves bit for bit) on
two calls in a row, the second with fewer primitives, in both Morton modes; sizes above the limit are rejected."""
import ctypes

import numpy as np
import pytest

import cwp
import pyoracle as po

pytestmark = pytest.mark.gpu


def _boxes(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3)).astype(np.float32)
    return po._box4(c - np.float32(0.004)), po._box4(c + rng.random((n, 3)).astype(np.float32) * np.float32(0.03))


def _build_static(dlo, dhi, n, morton):
    out = cwp.DeviceArray((2 * n - 1) * 32)
    root = ctypes.c_int(-1)
    cwp._check(cwp.lib().cwp_hploc_build_static(dlo.ptr, dhi.ptr, n, morton, out.ptr, None, ctypes.byref(root)))
    nodes = out.to_host(np.float32, (2 * n - 1, 8))   # synchronous copy: the build on the default stream is done
    out.free()
    return nodes, root.value


def _build_allocating(dlo, dhi, n, morton):
    h = ctypes.c_void_p()
    cwp._check(cwp.lib().cwp_hploc_create(n, ctypes.byref(h)))
    try:
        root = ctypes.c_int(-1)
        cwp._check(cwp.lib().cwp_hploc_build_mode(h, dlo.ptr, dhi.ptr, n, morton, None, ctypes.byref(root)))
        p = ctypes.c_void_p()
        cwp._check(cwp.lib().cwp_hploc_nodes(h, ctypes.byref(p)))
        nodes = np.zeros((2 * n - 1, 8), np.float32)
        cwp._check(cwp.lib().cwp_memcpy(nodes.ctypes.data, p.value, nodes.nbytes, cwp.CWP_COPY_D2H))
        return nodes, root.value
    finally:
        cwp.lib().cwp_hploc_destroy(h)


@pytest.mark.parametrize("morton", [0, 30])
def test_static_build_equals_allocating_build(morton):
    for call, n in enumerate((65536, 5001)):   # the limit, then a smaller build on the same scratch
        print("case no-alloc morton=%d call=%d n=%d" % (morton, call, n), flush=True)
        lo, hi = _boxes(n, 100 + n + morton)
        dlo, dhi = cwp.DeviceArray.from_host(lo), cwp.DeviceArray.from_host(hi)
        s, sroot = _build_static(dlo, dhi, n, morton)
        a, aroot = _build_allocating(dlo, dhi, n, morton)
        dlo.free(); dhi.free()
        assert sroot == aroot == 2 * n - 2
        np.testing.assert_array_equal(s[:n].view(np.uint32), a[:n].view(np.uint32))
        assert po.hp_canonical(s, n, sroot) == po.hp_canonical(a, n, aroot)


def test_static_build_rejects_sizes_beyond_the_limit():
    lo, hi = _boxes(8, 1)
    dlo, dhi = cwp.DeviceArray.from_host(lo), cwp.DeviceArray.from_host(hi)
    out = cwp.DeviceArray(64)
    root = ctypes.c_int()
    for n in (0, 65537):
        rc = cwp.lib().cwp_hploc_build_static(dlo.ptr, dhi.ptr, n, 0, out.ptr, None, ctypes.byref(root))
        assert rc == 1 and b"primitives" in cwp.lib().cwp_last_error()
    assert cwp.lib().cwp_hploc_build_static(dlo.ptr, dhi.ptr, 8, 7, out.ptr, None, ctypes.byref(root)) == 1
