"""The C-ABI library loads, exports every symbol include/xs_hip.h declares,
and rejects bad arguments before touching a GPU."""
import ctypes as C
import os

import pytest

import xansons_hip as xh


def test_library_exports_every_declared_symbol():
    L = xh.lib()
    declared = xh.header_symbols()
    assert len(declared) >= 12
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_null_context_is_an_argument_error():
    L = xh.lib()
    rc = L.xs_synchronize(None)
    assert rc == -1
    assert b"null xs_ctx" in L.xs_last_error()


@pytest.mark.skipif(os.environ.get("HIP_VISIBLE_DEVICES") is None and xh.device_count() > 0,
                    reason="a GPU is visible")
def test_init_without_device_fails_cleanly():
    h = C.c_void_p()
    rc = xh.lib().xs_init(0, C.byref(h))
    assert rc == -4 and not h


def test_bad_device_id_rejected():
    h = C.c_void_p()
    rc = xh.lib().xs_init(10_000, C.byref(h))
    assert rc == -4


def test_batch_entry_points_reject_null_handles():
    L = xh.lib()
    plan = C.c_void_p()
    one = (C.c_uint * 1)(1)
    assert L.xs_batch_plan(None, 1, one, one, None, 0.01, None, 0, C.byref(plan)) == -1
    assert not plan
    assert L.xs_batch_run(None, None, None, 0, None, -1, None, None, None, 0) == -1
    assert b"null plan" in L.xs_last_error()
    assert L.xs_batch_info(None, None, None, None, None, None) == -1
    assert L.xs_batch_overflow(None, None) == -1
    assert L.xs_batch_free(None) == 0
