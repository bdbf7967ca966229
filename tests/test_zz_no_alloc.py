"""The no-allocation advectbyzx (mhd_advectbyzx_noalloc) through the C ABI:
bit-identical to mhd_advectbyzx on two calls in a row, the second on a
smaller box (the fixed scratch still holds the first call's field), and
boxes beyond its fixed scratch are refused before anything runs."""
import ctypes

import numpy as np
import pytest

from conftest import bitwise
from golden.states import random_state

pytestmark = pytest.mark.gpu
F32 = np.float32


def _run(M, op, u0, b0, shape, dt):
    du, db = M.DeviceArray.from_host(u0), M.DeviceArray.from_host(b0)
    ddt = M.DeviceArray.from_host(np.array([dt], F32))
    getattr(M, op)(du, db, *shape, ddt)
    return du.to_host(), db.to_host()


def test_noalloc_advectbyzx_equals_advectbyzx():
    import mhdgpu as M
    # the larger box first, then a smaller one over the same scratch
    for shape, seed, dt in (((64, 24, 16), 71, F32(0.29)), ((16, 8, 12), 72, F32(0.33))):
        print("advectbyzx_noalloc", shape, flush=True)
        u0, b0 = random_state(*shape, seed=seed, amp=0.5)
        ua, ba = _run(M, "advectbyzx", u0, b0, shape, dt)
        ub, bb = _run(M, "advectbyzx_noalloc", u0, b0, shape, dt)
        assert bitwise(ub, u0) and bitwise(ua, u0)
        assert bitwise(bb, ba)
        assert not bitwise(bb, b0)  # the sweep did move the field


def test_noalloc_advectbyzx_refuses_boxes_beyond_its_scratch():
    import mhdgpu as M
    tiny = M.DeviceArray(64)
    n = [ctypes.c_int(v) for v in (256, 128, 128)]  # 2^22 cells > 2^21: refused before any launch
    rc = M.lib().mhd_advectbyzx_noalloc(tiny.ptr, tiny.ptr, *(ctypes.byref(v) for v in n), tiny.ptr)
    assert rc == -2
    assert "scratch" in M.lib().mhd_last_error().decode()
