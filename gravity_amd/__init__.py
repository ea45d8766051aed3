"""MI355X-native Barnes-Hut step for gkane1234/gravity.

The product is ``libgravity_amd.so`` (HIP kernels for gfx950 behind the C ABI
of ``include/bh_abi.h``); :mod:`gravity_amd.barnes_hut` is its Python host
side, shaped like the reference's ``BarnesHut`` class.
"""
from .barnes_hut import BarnesHut, Params, Counters, load_library, LIB_PATH  # noqa: F401
