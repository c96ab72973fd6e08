"""C-ABI checks that need no GPU: the library loads, exports every entry
point include/cheddar_amd.h declares, and rejects invalid parameter sets
before touching the device (the reference exits in AssertTrue instead,
include/common/Assert.h:9-21)."""
import ctypes
import os
import re

from conftest import REPO

import cheddar_amd as ck

HEADER = os.path.join(REPO, "include", "cheddar_amd.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ck_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) >= 25
    lib = ctypes.CDLL(ck.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the whole header
    assert set(names) <= set(ck.EXPORTED), set(names) - set(ck.EXPORTED)


def test_version_and_error_string():
    assert ck.lib.ck_version() >= 100
    assert isinstance(ck.lib.ck_last_error(), bytes)


def _create(log_n, main, aux, ter=(), levels=None):
    levels = levels or [[i + 1, 0] for i in range(len(main))]
    flat = [v for p in levels for v in p]
    h = ctypes.c_void_p()
    rc = ck.lib.ck_context_create(
        log_n, (ctypes.c_uint32 * len(main))(*main), len(main),
        (ctypes.c_uint32 * max(1, len(ter)))(*ter) if ter else None, len(ter),
        (ctypes.c_uint32 * len(aux))(*aux), len(aux),
        (ctypes.c_int * len(flat))(*flat), len(levels), ctypes.byref(h))
    return rc, h


def test_invalid_parameters_return_status():
    main = [2147352577, 2146959361]
    aux = [2146041857]
    # only N = 2^16 on gfx950
    rc, h = _create(15, main, aux)
    assert rc == 2 and not h.value
    # not NTT friendly
    rc, h = _create(16, [2147483647, 2146959361], aux)
    assert rc == 1 and b"NTT-friendly" in ck.lib.ck_last_error()
    # level table must end with every prime
    rc, h = _create(16, main, aux, levels=[[1, 0]])
    assert rc == 1 and b"last level" in ck.lib.ck_last_error()
    # levels must grow
    rc, h = _create(16, main, aux, levels=[[2, 0], [1, 0], [2, 0]])
    assert rc == 1
    # null output pointer
    assert ck.lib.ck_context_create(16, None, 0, None, 0, None, 0, None, 0, None) == 1


def test_null_context_calls_fail_cleanly():
    assert ck.lib.ck_context_destroy(None) == 0
    assert ck.lib.ck_np_to_level(None, 1, 0) == -1
    assert ck.lib.ck_modup_beta(None, 0) == -1
    out = (ctypes.c_int * 10)()
    assert ck.lib.ck_context_info(None, out) == 1
    assert ck.lib.ck_ntt(None, None, None, 1, 0, 0, 0, 0, None) == 1
    assert ck.lib.ck_hmult(None, 1, None, None, None, None, None, None, None, None, 1, None) == 1


# entry points whose null argument is valid (a no-op) or that take no context
_NULL_OK = {"ck_context_destroy", "ck_plan_destroy", "ck_last_error", "ck_version",
            "ck_context_create"}


def test_every_entry_point_refuses_a_null_context():
    """Every entry point called with a null context (and zero / null for the
    other arguments) reports an error instead of touching memory or the GPU:
    status 1 (invalid argument), or -1 from the functions that return a count
    or a level."""
    for name, (res, args) in sorted(ck._SIGS.items()):
        if name in _NULL_OK:
            continue
        vals = [None if (a in (ctypes.c_void_p, ctypes.c_char_p) or issubclass(a, ctypes._Pointer))
                else a(0) for a in args]
        rc = getattr(ck.lib, name)(*vals)
        assert rc in (1, -1), (name, rc)
        assert ck.lib.ck_last_error() is not None
