"""Python host side of the MI355X Barnes-Hut step.

Mirrors the reference's orchestration class ``BarnesHut``
(simulation/BarnesHut.java:99-579): :meth:`BarnesHut.init` and
:meth:`BarnesHut.step`, plus one method per private stage of the reference
(``reset_values``, ``partition_dead_bodies``, ``update_bounds``,
``generate_morton_codes``, ``radix_sort``, ``build_tree``, ``compute_force``,
``merge_bodies``) and the buffer access of ``GPU``/``SSBO`` keyed by the
reference's binding ids. Every call goes through the C ABI of
``include/bh_abi.h`` into ``libgravity_amd.so``; there is no other compute
path, and loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgravity_amd.so")

# binding ids (shaders/common/common.glsl:80-120)
LEAF_NODES, INTERNAL_NODES, SIMULATION_VALUES, BODIES_IN, BODIES_OUT = 0, 1, 2, 3, 4
MORTON_IN, MORTON_OUT, INDEX_IN, INDEX_OUT = 5, 6, 7, 8
MERGE_TASKS = 14
NODE_QUADRUPOLES = 16   # the corrected walk's records (no reference counterpart)

NEITHER, MERGING, COLLISION = 0, 1, 2
STATIC, DYNAMIC = 0, 1
FLAG_COLORS, FLAG_TIMING, FLAG_LEARNED, FLAG_GENERAL_WALK, FLAG_QUADRUPOLE = 1, 2, 4, 8, 16
FLAG_RELATIVE_MAC = 32
FLAG_GRAPH = 64
FLAG_FILL_CAP = 128
FLAG_CORRECTED = FLAG_LEARNED | FLAG_QUADRUPOLE
STAGES = ("reset", "dead", "bounds", "morton", "sort", "tree", "force", "merge")
NOALLOC_MAX_SAMPLE = 65536   # BH_NOALLOC_MAX_SAMPLE

ERRORS = {-1: "invalid argument", -2: "out of memory", -3: "HIP error", -4: "unsupported", -5: "too large",
          -6: "walk stack overflow", -7: "merge task overflow"}


class Params(C.Structure):
    """The step's uniforms (gpu/GPU.java:402-498) as ``bh_params``."""

    _fields_ = [("theta", C.c_float), ("dt", C.c_float), ("softening", C.c_float), ("elasticity", C.c_float),
                ("restitution", C.c_float), ("merging_collision_or_neither", C.c_uint32),
                ("static_or_dynamic", C.c_uint32), ("wrap_around", C.c_uint32), ("unit_mass", C.c_float),
                ("unit_density", C.c_float), ("unit_len", C.c_float), ("unit_time", C.c_float),
                ("gravitational_constant", C.c_float), ("body_length_constant", C.c_float),
                ("flags", C.c_uint32), ("force_tolerance", C.c_float)]


class Counters(C.Structure):
    _fields_ = [("num_bodies", C.c_uint32), ("initial_num_bodies", C.c_uint32), ("just_died", C.c_uint32),
                ("out_of_bounds", C.c_uint32), ("opened", C.c_uint64), ("interactions", C.c_uint64),
                ("merge_tasks", C.c_uint32), ("merges_applied", C.c_uint32), ("tree_height", C.c_uint32),
                ("max_stack", C.c_uint32), ("error_flags", C.c_uint32), ("reserved", C.c_uint32),
                ("corrected", C.c_uint64), ("quad_terms", C.c_uint64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libgravity_amd.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(path)
    P = C.c_void_p
    ctx = C.c_void_p
    L.bh_create.argtypes = [C.c_int, C.c_uint32, P, C.POINTER(Params), C.POINTER(ctx)]
    L.bh_destroy.argtypes = [ctx]
    L.bh_destroy.restype = None
    L.bh_last_error.argtypes = [ctx]
    L.bh_last_error.restype = C.c_char_p
    L.bh_binding_size.argtypes = [ctx, C.c_int]
    L.bh_binding_size.restype = C.c_size_t
    for name in ("bh_upload", "bh_download"):
        getattr(L, name).argtypes = [ctx, C.c_int, P, C.c_size_t, C.c_size_t]
    for name in ("bh_swap_body_buffers", "bh_swap_morton_and_index_buffers", "bh_init", "bh_partition_dead_bodies",
                 "bh_update_bounds", "bh_generate_morton_codes", "bh_radix_sort", "bh_merge_bodies", "bh_sync"):
        getattr(L, name).argtypes = [ctx]
    for name in ("bh_reset_values", "bh_build_tree", "bh_compute_force", "bh_step"):
        getattr(L, name).argtypes = [ctx, C.POINTER(Params)]
    L.bh_get_counters.argtypes = [ctx, C.POINTER(Counters)]
    L.bh_get_stage_times.argtypes = [ctx, C.POINTER(C.c_float), C.c_int]
    L.bh_direct_sum_f64.argtypes = [ctx, P, C.c_uint32, C.c_double, P]
    L.bh_direct_sum_f64_noalloc.argtypes = [ctx, P, C.c_uint32, C.c_double, P]
    L.bh_selftest_math.argtypes = [ctx, P]
    L.bh_render_nodes.argtypes = [ctx, C.c_uint32, C.c_uint32, P, C.c_uint32, C.POINTER(C.c_uint32)]
    L.bh_render_bodies.argtypes = [ctx, P, C.c_uint32, C.POINTER(C.c_uint32)]
    L.bh_host_alloc.argtypes = [C.c_size_t, C.POINTER(C.c_void_p)]
    L.bh_host_free.argtypes = [P]
    L.bh_host_free.restype = None
    _lib = L
    return L


# the render feed's records (include/bh_abi.h bh_render_node, bh_render_body)
RENDER_NODE = np.dtype([("com", "<f4", 3), ("mass", "<f4"), ("aabb_min", "<f4", 3), ("depth", "<u4"),
                        ("aabb_max", "<f4", 3), ("rgba", "<u4")])
RENDER_BODY = np.dtype([("pos", "<f4", 3), ("radius", "<f4"), ("rgba", "<u4")])


class PinnedBuffer:
    """Page-locked host memory from bh_host_alloc, viewed as a numpy array
    (`array`); freed by close() or when collected."""

    def __init__(self, shape, dtype):
        self.L = load_library()
        dt = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * dt.itemsize
        self.ptr = C.c_void_p()
        if self.L.bh_host_alloc(nbytes, C.byref(self.ptr)) != 0:
            raise MemoryError(f"bh_host_alloc({nbytes})")
        buf = (C.c_char * max(nbytes, 1)).from_address(self.ptr.value)
        self.array = np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape)

    def close(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            self.array = None
            self.L.bh_host_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_params(theta=0.5, dt=1e-3, softening=0.01, elasticity=0.5, restitution=0.0, merging=NEITHER,
                dynamic=DYNAMIC, wrap_around=0, units=None, G=0.0, body_length_constant=0.0, flags=0,
                force_tolerance=0.0) -> Params:
    if units is None:
        from .initial_conditions import g_one_units
        units = g_one_units()
    return Params(theta, dt, softening, elasticity, restitution, merging, dynamic, wrap_around, units.mass,
                  units.density, units.len, units.time, G, body_length_constant, flags, force_tolerance)


class BarnesHut:
    """One simulation on one GPU behind the C ABI.

    ``bodies`` is an ``(n, 8)`` float32 array in the reference's ``Body``
    layout; ``bounds`` is (min xyz, max xyz), the host-supplied initial bounds
    (simulation/GPUSimulation.java:261-263).
    """

    def __init__(self, bodies: np.ndarray, bounds, params: Params | None = None, device: int = 0):
        self.L = load_library()
        self.params = params if params is not None else make_params()
        b = np.ascontiguousarray(bodies, dtype=np.float32)
        if b.ndim != 2 or b.shape[1] != 8:
            raise ValueError("bodies must be (n, 8) float32")
        self.n = b.shape[0]
        bd = np.ascontiguousarray(bounds, dtype=np.float32)
        self.ctx = C.c_void_p()
        rc = self.L.bh_create(device, self.n, bd.ctypes.data_as(C.c_void_p), C.byref(self.params),
                              C.byref(self.ctx))
        if rc != 0:
            raise RuntimeError(f"bh_create failed: {ERRORS.get(rc, rc)}")
        self.upload(BODIES_IN, b)

    # -- plumbing
    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = self.L.bh_last_error(self.ctx)
            raise RuntimeError(f"{what}: {ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.L.bh_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def binding_size(self, binding: int) -> int:
        return int(self.L.bh_binding_size(self.ctx, binding))

    def upload(self, binding: int, data: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(data)
        self._check(self.L.bh_upload(self.ctx, binding, a.ctypes.data_as(C.c_void_p), offset, a.nbytes),
                    f"upload {binding}")

    def download(self, binding: int, dtype=np.uint8, count: int | None = None, offset: int = 0) -> np.ndarray:
        size = self.binding_size(binding) - offset if count is None else count * np.dtype(dtype).itemsize
        out = np.empty(size // np.dtype(dtype).itemsize, dtype=dtype)
        self._check(self.L.bh_download(self.ctx, binding, out.ctypes.data_as(C.c_void_p), offset, out.nbytes),
                    f"download {binding}")
        return out

    # -- typed views of the bindings
    def bodies(self, which: int = BODIES_IN) -> np.ndarray:
        return self.download(which, np.float32).reshape(self.n, 8)

    def morton(self, which: int = MORTON_IN) -> np.ndarray:
        return self.download(which, np.uint64)

    def index(self, which: int = INDEX_IN) -> np.ndarray:
        return self.download(which, np.uint32)

    def sim_values(self) -> np.ndarray:
        return self.download(SIMULATION_VALUES, np.uint32)

    def bounds(self) -> np.ndarray:
        s = self.sim_values().view(np.float32)
        return np.concatenate([s[8:11], s[12:15]])

    def nodes(self, which: int) -> np.ndarray:
        rows = self.n if which == LEAF_NODES else self.n - 1
        return self.download(which, np.uint32).reshape(rows, 20)

    def node_quadrupoles(self) -> np.ndarray:
        """The corrected walk's records of the internal nodes, (n - 1, 8) float32:
        G T xx xy xz yy yz zz, -1.5 G tr T, 0."""
        return self.download(NODE_QUADRUPOLES, np.float32).reshape(self.n - 1, 8)

    def render_nodes(self, depth_min: int = 0, depth_max: int = 0xFFFFFFFF, cap: int | None = None,
                     out: np.ndarray | None = None) -> np.ndarray:
        """The render feed's internal nodes (RENDER_NODE records) of the last
        tree, into `out` (RENDER_NODE array, e.g. a PinnedBuffer's) if given."""
        if out is None:
            cap = self.n if cap is None else cap
            out = np.zeros(max(cap, 1), RENDER_NODE)
        else:
            cap = out.shape[0] if cap is None else min(cap, out.shape[0])
        count = C.c_uint32()
        self._check(self.L.bh_render_nodes(self.ctx, depth_min, depth_max, out.ctypes.data_as(C.c_void_p), cap,
                                           C.byref(count)), "render_nodes")
        self.render_total = int(count.value)
        return out[: min(cap, count.value)]

    def render_bodies(self, cap: int | None = None, out: np.ndarray | None = None) -> np.ndarray:
        """The render feed's alive bodies (RENDER_BODY records) of BodiesIn,
        into `out` (RENDER_BODY array) if given."""
        if out is None:
            cap = self.n if cap is None else cap
            out = np.zeros(max(cap, 1), RENDER_BODY)
        else:
            cap = out.shape[0] if cap is None else min(cap, out.shape[0])
        count = C.c_uint32()
        self._check(self.L.bh_render_bodies(self.ctx, out.ctypes.data_as(C.c_void_p), cap, C.byref(count)),
                    "render_bodies")
        self.render_total = int(count.value)
        return out[: min(cap, count.value)]

    def merge_tasks(self) -> np.ndarray:
        raw = self.download(MERGE_TASKS, np.uint32)
        return raw[2: 2 + 2 * int(raw[1])].reshape(-1, 2)

    def counters(self) -> dict:
        c = Counters()
        self._check(self.L.bh_get_counters(self.ctx, C.byref(c)), "counters")
        return c.as_dict()

    def stage_times(self) -> dict:
        buf = (C.c_float * len(STAGES))()
        self._check(self.L.bh_get_stage_times(self.ctx, buf, len(STAGES)), "stage times")
        return dict(zip(STAGES, list(buf)))

    def direct_sum_f64(self, sample: np.ndarray, softening: float) -> np.ndarray:
        """FP64 direct-sum accelerations of the sampled body ids over BodiesIn."""
        s = np.ascontiguousarray(sample, dtype=np.uint32)
        out = np.zeros((s.shape[0], 3), np.float64)
        self._check(self.L.bh_direct_sum_f64(self.ctx, s.ctypes.data_as(C.c_void_p), s.shape[0], softening,
                                             out.ctypes.data_as(C.c_void_p)), "direct_sum_f64")
        return out

    def direct_sum_f64_noalloc(self, sample: np.ndarray, softening: float) -> np.ndarray:
        """direct_sum_f64 through the entry point that allocates no device memory
        (at most NOALLOC_MAX_SAMPLE samples per call)."""
        s = np.ascontiguousarray(sample, dtype=np.uint32)
        out = np.zeros((s.shape[0], 3), np.float64)
        self._check(self.L.bh_direct_sum_f64_noalloc(self.ctx, s.ctypes.data_as(C.c_void_p), s.shape[0], softening,
                                                     out.ctypes.data_as(C.c_void_p)), "direct_sum_f64_noalloc")
        return out

    def selftest_math(self) -> list:
        out = np.zeros(3, np.uint32)
        self._check(self.L.bh_selftest_math(self.ctx, out.ctypes.data_as(C.c_void_p)), "selftest_math")
        return out.tolist()

    def build_current_tree(self, params: Params | None = None) -> None:
        """Run the stages up to the tree on the current BodiesIn (no walk)."""
        p = params or self.params
        self.reset_values(p)
        self.partition_dead_bodies()
        if p.static_or_dynamic == DYNAMIC:
            self.update_bounds()
        self.generate_morton_codes()
        self.radix_sort()
        self.build_tree(p)

    def accelerations(self, params: Params | None = None) -> np.ndarray:
        """Walk accelerations of every body for the current BodiesIn, by the
        reference's own extraction: velocities set to 0 and dt = 1 make the
        integrated velocity equal the acceleration. BodiesIn is restored;
        BodiesOut, the tree and the sorted order are left as this pass made them."""
        saved = self.bodies(BODIES_IN)
        z = saved.copy()
        z[:, 4:7] = 0.0
        p1 = Params.from_buffer_copy(params or self.params)
        p1.dt = 1.0
        p1.flags &= FLAG_CORRECTED | FLAG_GENERAL_WALK | FLAG_RELATIVE_MAC
        p1.merging_collision_or_neither = NEITHER
        p1.static_or_dynamic = DYNAMIC
        self.upload(BODIES_IN, z)
        self.reset_values(p1)
        self.partition_dead_bodies()
        self.update_bounds()
        self.generate_morton_codes()
        self.radix_sort()
        self.build_tree(p1)
        self.compute_force(p1)
        acc = self.bodies(BODIES_OUT)[:, 4:7].copy()
        self.upload(BODIES_IN, saved)
        return acc

    def sync(self) -> None:
        self._check(self.L.bh_sync(self.ctx), "sync")

    # -- BarnesHut.java
    def init(self) -> None:
        """BarnesHut.init(): indexIn[i] = i."""
        self._check(self.L.bh_init(self.ctx), "init")

    def step(self, params: Params | None = None) -> None:
        """BarnesHut.step(): the whole pipeline, asynchronous on the context's stream."""
        self._check(self.L.bh_step(self.ctx, C.byref(params or self.params)), "step")

    def reset_values(self, params: Params | None = None) -> None:
        self._check(self.L.bh_reset_values(self.ctx, C.byref(params or self.params)), "reset_values")

    def partition_dead_bodies(self) -> None:
        self._check(self.L.bh_partition_dead_bodies(self.ctx), "partition_dead_bodies")

    def update_bounds(self) -> None:
        self._check(self.L.bh_update_bounds(self.ctx), "update_bounds")

    def generate_morton_codes(self) -> None:
        self._check(self.L.bh_generate_morton_codes(self.ctx), "generate_morton_codes")

    def radix_sort(self) -> None:
        self._check(self.L.bh_radix_sort(self.ctx), "radix_sort")

    def build_tree(self, params: Params | None = None) -> None:
        self._check(self.L.bh_build_tree(self.ctx, C.byref(params or self.params)), "build_tree")

    def compute_force(self, params: Params | None = None) -> None:
        self._check(self.L.bh_compute_force(self.ctx, C.byref(params or self.params)), "compute_force")

    def merge_bodies(self) -> None:
        self._check(self.L.bh_merge_bodies(self.ctx), "merge_bodies")

    def swap_body_buffers(self) -> None:
        self._check(self.L.bh_swap_body_buffers(self.ctx), "swap_body_buffers")

    def swap_morton_and_index_buffers(self) -> None:
        self._check(self.L.bh_swap_morton_and_index_buffers(self.ctx), "swap_morton_and_index_buffers")
