"""The C ABI library loads and exports every symbol include/*.h declares
(no compute: this container has no GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"\b(bh_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from gravity_amd import barnes_hut
    lib = barnes_hut.load_library()
    names = declared_symbols()
    assert len(names) >= 40
    assert "bh_ensemble_autotune" in names and "bh_mlp_train" in names
    missing = [s for s in names if not hasattr(lib, s)]
    assert not missing, missing


def test_params_struct_matches_header():
    from gravity_amd import barnes_hut
    # 12 four-byte fields before flags... sizes must agree with the C struct
    assert ctypes.sizeof(barnes_hut.Params) == 16 * 4
    assert ctypes.sizeof(barnes_hut.Counters) == 4 * 4 + 2 * 8 + 6 * 4 + 8 + 8


def test_library_rejects_bad_arguments_without_gpu():
    from gravity_amd import barnes_hut
    lib = barnes_hut.load_library()
    out = ctypes.c_void_p()
    # n < 2 is rejected before any device call
    assert lib.bh_create(0, 1, None, None, ctypes.byref(out)) == -1
    assert lib.bh_sync(None) == -1
    assert lib.bh_last_error(None) == b"null context"


def test_ensemble_rejects_bad_arguments_without_gpu():
    from gravity_amd import barnes_hut
    from gravity_amd import ensemble
    lib = ensemble._lib()
    assert ctypes.sizeof(ensemble.Config) == 16
    out = ctypes.c_void_p()
    prm = barnes_hut.make_params()
    counts = (ctypes.c_uint32 * 2)(4096, 1)        # a simulation of one body is rejected
    bodies = (ctypes.c_float * (8 * 4097))()
    bounds = (ctypes.c_float * 12)()
    assert lib.bh_ensemble_create(0, 2, counts, bodies, bounds, ctypes.byref(prm), None, ctypes.byref(out)) == -1
    merging = barnes_hut.make_params(merging=barnes_hut.MERGING)
    counts2 = (ctypes.c_uint32 * 1)(4096)
    assert lib.bh_ensemble_create(0, 1, counts2, bodies, bounds, ctypes.byref(merging), None,
                                  ctypes.byref(out)) == -4
    bad = ensemble.Config(9, 4, 192, 1)
    assert lib.bh_ensemble_create(0, 1, counts2, bodies, bounds, ctypes.byref(prm), ctypes.byref(bad),
                                  ctypes.byref(out)) == -1
    assert lib.bh_ensemble_step(None, 1) == -1
    assert lib.bh_ensemble_last_error(None) == b"null ensemble"
