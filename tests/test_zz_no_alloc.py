"""The allocation-free removed-contacts download against the plain one, through the C ABI.

Two calls in a row on one GPU: a 3,000-body pile whose contacts come and go, then a smaller
heap in a second context, so the second call reuses the device's scratch array after a larger
one.  Each must give exactly the records of cc_download_removed_contacts (integer records, so
bit for bit), and a previous step larger than the scratch array is never reached here.
"""
import ctypes as C

import numpy as np
import pytest

from crashcat_amd import _abi
from crashcat_amd import scene as SC
from crashcat_amd._lib import REMOVED_DTYPE, Context, lib

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
P = C.POINTER(_abi.ContactRemoved)


def _download(fn, h):
    cnt = C.c_int32()
    rc = fn(h, None, 0, C.byref(cnt))
    assert rc == 0
    n = cnt.value
    out = np.zeros(max(n, 1), dtype=REMOVED_DTYPE)
    rc = fn(h, out.ctypes.data_as(P), n, C.byref(cnt))
    assert rc == 0 and cnt.value == n
    return out[:n]


def _step_until_removals(g, max_steps=80):
    for s in range(max_steps):
        g.step(DT, 1)
        g.sync()
        cnt = C.c_int32()
        assert lib().cc_download_removed_contacts(g.h, None, 0, C.byref(cnt)) == 0
        if cnt.value > 0:
            return cnt.value
    raise AssertionError("no contact was removed")


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_noalloc_removed_contacts_match_plain_download(precision):
    L = lib()
    big = Context(precision)
    big.upload(SC.mixed_pile(seed=1002, n=3000, extent=(20.0, 12.0, 20.0)))
    small = Context(precision)
    small.upload(SC.cube_heap(seed=1001, nx=3, ny=3, nz=3))
    try:
        n_big = _step_until_removals(big)
        n_small = _step_until_removals(small)
        assert n_small < n_big
        for g in (big, small):  # the larger input first, then the smaller one
            ref = _download(L.cc_download_removed_contacts, g.h)
            got = _download(L.cc_download_removed_contacts_noalloc, g.h)
            assert len(ref) > 0
            assert np.array_equal(ref, got)
        # a capacity below the count is refused, the count is still reported
        cnt = C.c_int32()
        out = np.zeros(1, dtype=REMOVED_DTYPE)
        if n_small > 1:
            assert L.cc_download_removed_contacts_noalloc(small.h, out.ctypes.data_as(P), 1, C.byref(cnt)) == \
                _abi.CC_E_ARG
            assert cnt.value == n_small
    finally:
        big.close()
        small.close()
