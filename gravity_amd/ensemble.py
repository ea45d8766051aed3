"""Ensembles of independent simulations stepped together on one GPU.

Wraps ``include/bh_ensemble.h`` (no reference counterpart: the reference runs
one simulation per process). The simulations' bodies are concatenated in one
set of device arrays; one step is one launch per stage over all of them (a
segmented (simulation, Morton code) radix sort, one tree build over all
segments, a persistent walk pulling 64-body tasks from per-XCD work queues), and
can be replayed as a captured hipGraph. Each simulation evolves bit for bit
as it would stepped alone through :class:`gravity_amd.barnes_hut.BarnesHut`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import barnes_hut as B


class Config(C.Structure):
    """``bh_ensemble_config``: sort digit width, walk workgroup and stack depth, graph replay."""

    _fields_ = [("radix_bits", C.c_uint32), ("walk_waves", C.c_uint32), ("stack_depth", C.c_uint32),
                ("use_graph", C.c_uint32)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


WALK_SHAPES = [(w, d) for w in (1, 2, 4, 8) for d in (64, 128, 192)]


def _lib():
    L = B.load_library()
    if getattr(L, "_ensemble_declared", False):
        return L
    P = C.c_void_p
    L.bh_ensemble_create.argtypes = [C.c_int, C.c_uint32, P, P, P, C.POINTER(B.Params), C.POINTER(Config),
                                     C.POINTER(C.c_void_p)]
    L.bh_ensemble_destroy.argtypes = [P]
    L.bh_ensemble_destroy.restype = None
    L.bh_ensemble_last_error.argtypes = [P]
    L.bh_ensemble_last_error.restype = C.c_char_p
    L.bh_ensemble_get_config.argtypes = [P, C.POINTER(Config)]
    L.bh_ensemble_set_config.argtypes = [P, C.POINTER(Config)]
    L.bh_ensemble_step.argtypes = [P, C.c_uint32]
    L.bh_ensemble_sync.argtypes = [P]
    L.bh_ensemble_download_bodies.argtypes = [P, P]
    L.bh_ensemble_download_bounds.argtypes = [P, P]
    L.bh_ensemble_download_sorted.argtypes = [P, P, P]
    L.bh_ensemble_download_internal_nodes.argtypes = [P, P]
    L.bh_ensemble_get_counters.argtypes = [P, C.POINTER(B.Counters)]
    L.bh_ensemble_autotune.argtypes = [P, C.POINTER(Config), P]
    L.bh_ensemble_attach_correction.argtypes = [P, P]
    L._ensemble_declared = True
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Ensemble:
    """``sims``: list of ``(n_s, 8)`` float32 Body arrays; ``bounds``: list of
    6-float initial bounds (tight bounds when omitted)."""

    def __init__(self, sims, params: B.Params, bounds=None, config: Config | None = None, device: int = 0):
        from .initial_conditions import tight_bounds
        self.L = _lib()
        self.counts = np.array([s.shape[0] for s in sims], dtype=np.uint32)
        self.offsets = np.concatenate([[0], np.cumsum(self.counts.astype(np.int64))[:-1]]).astype(np.int64)
        self.total = int(self.counts.astype(np.int64).sum())
        bodies = np.ascontiguousarray(np.concatenate(sims), dtype=np.float32)
        if bounds is None:
            bounds = [tight_bounds(s) for s in sims]
        bb = np.ascontiguousarray(np.stack(bounds), dtype=np.float32)
        self.params = params
        self.h = C.c_void_p()
        rc = self.L.bh_ensemble_create(device, len(sims), _ptr(self.counts), _ptr(bodies), _ptr(bb),
                                       C.byref(params), C.byref(config) if config is not None else None,
                                       C.byref(self.h))
        if rc != 0:
            raise RuntimeError(f"bh_ensemble_create: {B.ERRORS.get(rc, rc)}")

    def _check(self, rc, what):
        if rc != 0:
            msg = self.L.bh_ensemble_last_error(self.h)
            raise RuntimeError(f"{what}: {B.ERRORS.get(rc, rc)} {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.L.bh_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_sims(self) -> int:
        return int(self.counts.size)

    def config(self) -> Config:
        c = Config()
        self._check(self.L.bh_ensemble_get_config(self.h, C.byref(c)), "get_config")
        return c

    def set_config(self, c: Config) -> None:
        self._check(self.L.bh_ensemble_set_config(self.h, C.byref(c)), "set_config")

    def step(self, steps: int = 1) -> None:
        self._check(self.L.bh_ensemble_step(self.h, steps), "step")

    def sync(self) -> None:
        self._check(self.L.bh_ensemble_sync(self.h), "sync")

    def bodies(self) -> list[np.ndarray]:
        out = np.zeros((self.total, 8), np.float32)
        self._check(self.L.bh_ensemble_download_bodies(self.h, _ptr(out)), "download bodies")
        return [out[o:o + n] for o, n in zip(self.offsets, self.counts)]

    def bounds(self) -> np.ndarray:
        out = np.zeros((self.n_sims, 6), np.float32)
        self._check(self.L.bh_ensemble_download_bounds(self.h, _ptr(out)), "download bounds")
        return out

    def sorted(self):
        """Per simulation: (Morton codes, body ids local to the simulation) in sorted order."""
        k = np.zeros(self.total, np.uint64)
        i = np.zeros(self.total, np.uint32)
        self._check(self.L.bh_ensemble_download_sorted(self.h, _ptr(k), _ptr(i)), "download sorted")
        return [(k[o:o + n], (i[o:o + n].astype(np.int64) - o).astype(np.uint32))
                for o, n in zip(self.offsets, self.counts)]

    def internal_nodes(self) -> list[np.ndarray]:
        """Per simulation: its internal nodes in the 80-byte Node layout with
        node ids made local (leaves 0..n-1, internal n + k), as a single
        simulation of that size numbers them."""
        ni = self.total - self.n_sims
        raw = np.zeros((max(ni, 1), 20), np.uint32)
        self._check(self.L.bh_ensemble_download_internal_nodes(self.h, _ptr(raw)), "download nodes")
        out = []
        for s, (o, n) in enumerate(zip(self.offsets, self.counts)):
            ib = int(o) - s
            nodes = raw[ib:ib + int(n) - 1].copy()
            u = nodes
            for col in (10, 11, 15):    # childA, childB, parentId
                v = u[:, col].astype(np.int64)
                leaf = v < self.total
                empty = v == 0xFFFFFFFF
                local = np.where(leaf, v - o, int(n) + (v - self.total - ib))
                u[:, col] = np.where(empty, 0xFFFFFFFF, local).astype(np.uint32)
            out.append(nodes)
        return out

    def counters(self) -> dict:
        c = B.Counters()
        self._check(self.L.bh_ensemble_get_counters(self.h, C.byref(c)), "counters")
        return c.as_dict()

    def attach(self, mlp) -> None:
        """The network (gravity_amd.learned.Mlp) of an ensemble created with FLAG_LEARNED."""
        self._check(self.L.bh_ensemble_attach_correction(self.h, mlp.h if mlp is not None else None), "attach")
        self._mlp = mlp

    def autotune(self):
        """Measure the candidate shapes on the current state (which is left
        unchanged), keep the fastest; returns (config, timings in ms)."""
        c = Config()
        t = np.full(16, -1.0, np.float32)
        self._check(self.L.bh_ensemble_autotune(self.h, C.byref(c), _ptr(t)), "autotune")
        timings = {f"radix_bits_{5 + k}": float(t[k]) for k in range(4)}
        for k, (w, d) in enumerate(WALK_SHAPES):
            timings[f"walk_{w}x64_depth{d}"] = float(t[4 + k])
        return c, timings
