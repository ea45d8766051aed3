"""The C ABI library loads and exports every symbol include/bh_abi.h declares
(no compute: this container has no GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "bh_abi.h")).read()
    return sorted(set(re.findall(r"\b(bh_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from gravity_amd import barnes_hut
    lib = barnes_hut.load_library()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [s for s in names if not hasattr(lib, s)]
    assert not missing, missing


def test_params_struct_matches_header():
    from gravity_amd import barnes_hut
    # 12 four-byte fields before flags... sizes must agree with the C struct
    assert ctypes.sizeof(barnes_hut.Params) == 15 * 4
    assert ctypes.sizeof(barnes_hut.Counters) == 4 * 4 + 2 * 8 + 6 * 4 + 8


def test_library_rejects_bad_arguments_without_gpu():
    from gravity_amd import barnes_hut
    lib = barnes_hut.load_library()
    out = ctypes.c_void_p()
    # n < 2 is rejected before any device call
    assert lib.bh_create(0, 1, None, None, ctypes.byref(out)) == -1
    assert lib.bh_sync(None) == -1
    assert lib.bh_last_error(None) == b"null context"
