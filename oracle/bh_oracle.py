"""ctypes front end of the CPU oracle (liboracle ``libbh_oracle.so``).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path. Only the tests, the
smoke check and bench.py's CPU-baseline leg import this module; the product
package never does.

Arrays are numpy: bodies ``(n, 8) float32`` (the reference's ``Body`` std430
layout), nodes ``(n, 20) uint32`` (the 80-byte ``Node`` std430 layout; use
:func:`node_fields` to split them).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbh_oracle.so")
EMPTY = 0xFFFFFFFF

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class Params(C.Structure):
    _fields_ = [("theta", C.c_float), ("dt", C.c_float), ("softening", C.c_float),
                ("elasticity", C.c_float), ("restitution", C.c_float),
                ("merging", C.c_uint32), ("dynamic", C.c_uint32), ("wrap_around", C.c_uint32)]


class WalkStats(C.Structure):
    _fields_ = [("opened", C.c_uint64), ("interactions", C.c_uint64), ("mac_tests", C.c_uint64),
                ("out_of_bounds", C.c_uint32), ("max_stack", C.c_uint32),
                ("n_tasks", C.c_uint32), ("task_overflow", C.c_uint32)]


class Sim(C.Structure):
    _fields_ = [("numBodies", C.c_uint32), ("initialNumBodies", C.c_uint32), ("justDied", C.c_uint32),
                ("merged", C.c_uint32), ("outOfBounds", C.c_uint32), ("relativeTo", C.c_uint32),
                ("pad1", C.c_uint32), ("pad2", C.c_uint32),
                ("minCorner", C.c_float * 3), ("padA", C.c_float),
                ("maxCorner", C.c_float * 3), ("padB", C.c_float),
                ("unitMass", C.c_float), ("unitDensity", C.c_float), ("unitLen", C.c_float),
                ("unitTime", C.c_float), ("cameraScale", C.c_float),
                ("gravitationalConstant", C.c_float), ("bodyLengthConstant", C.c_float),
                ("uintDebug", C.c_uint32 * 100), ("floatDebug", C.c_float * 100), ("tail_pad", C.c_uint32)]


class State(C.Structure):
    _fields_ = [("n_init", C.c_uint32), ("bodies", C.c_void_p * 2), ("index", C.c_void_p * 2),
                ("morton", C.c_void_p), ("leaf", C.c_void_p), ("internal", C.c_void_p),
                ("tasks", C.c_void_p), ("task_cap", C.c_uint32), ("sim", Sim), ("cur", C.c_int),
                ("fill_cap", C.c_int), ("stack_cap", C.c_int), ("threads", C.c_int),
                ("last", WalkStats), ("last_height", C.c_uint32)]


P = C.c_void_p
U32 = C.c_uint32


def _declare(L: C.CDLL) -> None:
    L.bho_units.argtypes = [C.c_float] * 4 + [C.POINTER(C.c_float)] * 2
    L.bho_dead_partition.argtypes = [P, P, P, U32, P]
    L.bho_dead_partition.restype = U32
    L.bho_aabb.argtypes = [P, P, U32, P]
    L.bho_expand_bits21.argtypes = [U32]
    L.bho_expand_bits21.restype = C.c_uint64
    L.bho_morton3d.argtypes = [U32, U32, U32]
    L.bho_morton3d.restype = C.c_uint64
    L.bho_morton_encode.argtypes = [P, P, U32, P, P]
    L.bho_sort.argtypes = [P, P, U32]
    L.bho_safe_lcp.argtypes = [P, U32, C.c_int, C.c_int]
    L.bho_safe_lcp.restype = C.c_int
    L.bho_build.argtypes = [P, U32, U32, P, P]
    L.bho_fill.argtypes = [P, P, U32, U32, C.c_float, C.c_float, C.c_int, C.c_int, P, P]
    L.bho_fill.restype = U32
    L.bho_star_color.argtypes = [C.c_float, C.c_float, P]
    L.bho_force.argtypes = [P, P, P, U32, U32, P, P, C.POINTER(Params), C.c_float, C.c_float, P, C.c_int,
                            P, P, P, U32, C.POINTER(WalkStats), C.c_int]
    L.bho_force.restype = C.c_int
    L.bho_force_range.argtypes = L.bho_force.argtypes + [U32, U32]
    L.bho_force_range.restype = C.c_int
    L.bho_force_corrected.argtypes = [P, P, U32, U32, P, C.c_float, C.c_float, C.c_float, P, C.c_float, U32,
                                      C.c_int, P, P, P, C.POINTER(WalkStats), C.c_int]
    L.bho_force_corrected.restype = C.c_int
    L.bho_node_moments.argtypes = [P, P, U32, U32, P, C.c_int]
    L.bho_node_moments.restype = None
    L.bho_merge.argtypes = [P, P, U32]
    L.bho_merge.restype = U32
    L.bho_merge_pair.argtypes = [P, P, P]
    L.bho_state_init.argtypes = [C.POINTER(State), U32, P, P, C.c_float, C.c_float, C.c_float, C.c_float]
    L.bho_state_init.restype = C.c_int
    L.bho_state_free.argtypes = [C.POINTER(State)]
    L.bho_step.argtypes = [C.POINTER(State), C.POINTER(Params)]
    L.bho_step.restype = C.c_int
    L.bho_direct_sum_f64.argtypes = [P, U32, P, U32, C.c_double, C.c_double, P, C.c_int]
    L.mlp_num_params.restype = C.c_int
    L.mlp_forward_f64.argtypes = [P, P, C.c_int, P]
    L.mlp_quad_basis_f64.argtypes = [P, P, P]
    L.mlp_quad_basis_f64.restype = None
    L.mlp_loss_grad_f64.argtypes = [P, P, P, C.c_int, P]
    L.mlp_loss_grad_f64.restype = C.c_double


def _p(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def params(theta=0.5, dt=1e-3, softening=0.01, elasticity=0.5, restitution=0.0, merging=0, dynamic=1,
           wrap_around=0) -> Params:
    return Params(theta, dt, softening, elasticity, restitution, merging, dynamic, wrap_around)


def node_fields(nodes: np.ndarray) -> dict:
    """Split (n, 20) uint32 std430 nodes into named fields."""
    f = nodes.view(np.float32)
    return {"com": f[:, 0:3], "mass": f[:, 3], "aabb": f[:, 4:10], "childA": nodes[:, 10],
            "childB": nodes[:, 11], "nodeDepth": nodes[:, 12], "bodiesContained": nodes[:, 13],
            "readyChildren": nodes[:, 14], "parentId": nodes[:, 15], "avgColor": f[:, 16:20]}


def units(mass, density, length, time):
    g = C.c_float()
    b = C.c_float()
    lib().bho_units(mass, density, length, time, C.byref(g), C.byref(b))
    return np.float32(g.value), np.float32(b.value)


def dead_partition(src: np.ndarray, dst: np.ndarray | None, index: np.ndarray, num: int):
    out = index.copy()
    died = lib().bho_dead_partition(_p(src), _p(dst), _p(index), num, _p(out))
    return out, int(died)


def aabb(bodies: np.ndarray, index: np.ndarray, num: int, prev_bounds) -> np.ndarray:
    b = np.ascontiguousarray(prev_bounds, dtype=np.float32).copy()
    lib().bho_aabb(_p(bodies), _p(index), num, _p(b))
    return b


def morton_encode(bodies: np.ndarray, index: np.ndarray, num: int, bounds) -> np.ndarray:
    codes = np.zeros(num, dtype=np.uint64)
    b = np.ascontiguousarray(bounds, dtype=np.float32)
    lib().bho_morton_encode(_p(bodies), _p(index), num, _p(b), _p(codes))
    return codes


def sort(keys: np.ndarray, index: np.ndarray):
    k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    i = np.ascontiguousarray(index, dtype=np.uint32).copy()
    lib().bho_sort(_p(k), _p(i), k.shape[0])
    return k, i


def build(keys: np.ndarray, num: int, n_init: int, leaf=None, internal=None):
    leaf = np.zeros((n_init, 20), np.uint32) if leaf is None else leaf
    internal = np.zeros((max(n_init - 1, 1), 20), np.uint32) if internal is None else internal
    lib().bho_build(_p(np.ascontiguousarray(keys, dtype=np.uint64)), num, n_init, _p(leaf), _p(internal))
    return leaf, internal


def fill(bodies, index, num, n_init, leaf, internal, unit_mass=1.0, unit_density=1.0, max_passes=64,
         with_colors=True) -> int:
    return int(lib().bho_fill(_p(bodies), _p(index), num, n_init, unit_mass, unit_density, max_passes,
                              int(with_colors), _p(leaf), _p(internal)))


def force(src, index, num, n_init, leaf, internal, prm: Params, G, blc, bounds, dst=None, stack_cap=64,
          want_accel=True, task_cap=None, threads=0):
    accel = np.zeros((n_init, 3), np.float32) if want_accel else None
    inter = np.zeros(n_init, np.uint32) if want_accel else None
    task_cap = n_init if task_cap is None else task_cap
    tasks = np.zeros((max(task_cap, 1), 2), np.uint32)
    st = WalkStats()
    b = np.ascontiguousarray(bounds, dtype=np.float32)
    rc = lib().bho_force(_p(src), _p(dst), _p(index), num, n_init, _p(leaf), _p(internal), C.byref(prm),
                         float(G), float(blc), _p(b), stack_cap, _p(accel), _p(inter), _p(tasks), task_cap,
                         C.byref(st), threads)
    return {"rc": rc, "accel": accel, "interactions": inter, "tasks": tasks[: st.n_tasks].copy(), "stats": st}


def force_corrected(leaf, internal, num, n_init, quad, theta, softening, G=1.0, amag=None, alpha=0.0, group=64,
                    stack_cap=256, threads=0):
    """The corrected walk in FP64 over the nodes that the GPU's FP32 tests
    take, in groups of `group` sorted positions that take a node only
    together (64: the GPU's rule, 1: each body alone) (geometric test at theta, or with
    amag (|a| by sorted position) the relative test at alpha): per sorted
    position the acceleration (float64), the sum of the taken terms'
    magnitudes and the count of nodes taken; quad[k] holds internal node k's
    G T xx xy xz yy yz zz, -1.5 G tr T (8 floats, the last unused)."""
    q = np.ascontiguousarray(quad, dtype=np.float32)
    assert q.shape == (n_init - 1, 8)
    am = None if amag is None else np.ascontiguousarray(amag, dtype=np.float32)
    if am is not None:
        assert am.shape == (num,)
    accel = np.zeros((max(num, 1), 3), np.float64)
    absum = np.zeros(max(num, 1), np.float64)
    inter = np.zeros(max(num, 1), np.uint32)
    st = WalkStats()
    rc = lib().bho_force_corrected(_p(leaf), _p(internal), num, n_init, _p(q), float(theta), float(softening),
                                   float(G), None if am is None else _p(am), float(alpha), group, stack_cap, _p(accel),
                                   _p(absum), _p(inter), C.byref(st), threads)
    return {"rc": rc, "accel": accel[:num], "absum": absum[:num], "interactions": inter[:num],
            "opened": int(st.opened), "quad_terms": int(st.mac_tests), "max_stack": int(st.max_stack)}


def node_moments(leaf, internal, num, n_init, threads=0) -> np.ndarray:
    """FP64 second moment (xx xy xz yy yz zz) of each internal node's bodies about its centre of mass."""
    m2 = np.zeros((max(num - 1, 1), 6), np.float64)
    lib().bho_node_moments(_p(leaf), _p(internal), num, n_init, _p(m2), threads)
    return m2[: max(num - 1, 0)]


def force_range(src, index, num, n_init, leaf, internal, prm: Params, G, blc, bounds, g_begin, g_end,
                dst=None, stack_cap=64, threads=0):
    """Walk only the sorted positions [g_begin, g_end); returns the walk statistics."""
    st = WalkStats()
    b = np.ascontiguousarray(bounds, dtype=np.float32)
    rc = lib().bho_force_range(_p(src), _p(dst), _p(index), num, n_init, _p(leaf), _p(internal), C.byref(prm),
                               float(G), float(blc), _p(b), stack_cap, None, None, None, 0, C.byref(st), threads,
                               g_begin, g_end)
    return {"rc": rc, "stats": st}


def merge(dst: np.ndarray, tasks: np.ndarray) -> int:
    t = np.ascontiguousarray(tasks, dtype=np.uint32).copy()
    return int(lib().bho_merge(_p(dst), _p(t), t.shape[0]))


def direct_sum_f64(bodies: np.ndarray, sample: np.ndarray, G=1.0, softening=0.01, threads=0) -> np.ndarray:
    s = np.ascontiguousarray(sample, dtype=np.uint32)
    out = np.zeros((s.shape[0], 3), np.float64)
    lib().bho_direct_sum_f64(_p(bodies), bodies.shape[0], _p(s), s.shape[0], G, softening, _p(out), threads)
    return out


def mlp_num_params() -> int:
    return int(lib().mlp_num_params())


def mlp_forward(params: np.ndarray, x: np.ndarray) -> np.ndarray:
    """FP64 forward of the 12-64-64-6 SiLU network (oracle/mlp_oracle.c)."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros((xx.shape[0], 6), np.float64)
    lib().mlp_forward_f64(_p(p), _p(xx), xx.shape[0], _p(y))
    return y


def mlp_quad_basis(u, o) -> np.ndarray:
    """B(u) o = -3 o u - 1.5 tr(o) u + 7.5 (u.o.u) u for one direction u (3) and tensor o (6)."""
    uu = np.ascontiguousarray(u, dtype=np.float64)
    oo = np.ascontiguousarray(o, dtype=np.float64)
    out = np.zeros(3, np.float64)
    lib().mlp_quad_basis_f64(_p(uu), _p(oo), _p(out))
    return out


def mlp_loss_grad(params: np.ndarray, x: np.ndarray, t: np.ndarray):
    """FP64 loss through the quadrupole basis (targets t[rows][7] = (u, y, w)) and its gradient."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32)
    tt = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros(p.shape[0], np.float64)
    loss = lib().mlp_loss_grad_f64(_p(p), _p(xx), _p(tt), xx.shape[0], _p(g))
    return float(loss), g


class Simulation:
    """Host-buffer restatement of ``BarnesHut.step`` (simulation/BarnesHut.java:99-164)."""

    def __init__(self, bodies: np.ndarray, bounds, unit_mass, unit_density, unit_len=1.0, unit_time=1.0,
                 G=None, body_len_const=None, fill_cap=64, stack_cap=64, threads=0):
        self.st = State()
        b = np.ascontiguousarray(bodies, dtype=np.float32)
        bd = np.ascontiguousarray(bounds, dtype=np.float32)
        rc = lib().bho_state_init(C.byref(self.st), b.shape[0], _p(b), _p(bd), unit_mass, unit_density,
                                  unit_len, unit_time)
        if rc != 0:
            raise ValueError("bho_state_init failed")
        if G is not None:
            self.st.sim.gravitationalConstant = float(G)
        if body_len_const is not None:
            self.st.sim.bodyLengthConstant = float(body_len_const)
        self.st.fill_cap = fill_cap
        self.st.stack_cap = stack_cap
        self.st.threads = threads
        self.n = b.shape[0]

    def step(self, prm: Params) -> int:
        return lib().bho_step(C.byref(self.st), C.byref(prm))

    def _arr(self, ptr, dtype, count):
        buf = (C.c_char * (np.dtype(dtype).itemsize * count)).from_address(ptr)
        return np.frombuffer(buf, dtype=dtype, count=count).copy()

    def bodies(self) -> np.ndarray:
        return self._arr(self.st.bodies[self.st.cur], np.float32, self.n * 8).reshape(self.n, 8)

    def index(self) -> np.ndarray:
        return self._arr(self.st.index[0], np.uint32, self.n)

    def morton(self) -> np.ndarray:
        return self._arr(self.st.morton, np.uint64, self.n)

    def nodes(self):
        return (self._arr(self.st.leaf, np.uint32, self.n * 20).reshape(self.n, 20),
                self._arr(self.st.internal, np.uint32, max(self.n - 1, 1) * 20).reshape(max(self.n - 1, 1), 20))

    @property
    def num_bodies(self) -> int:
        return int(self.st.sim.numBodies)

    @property
    def bounds(self) -> np.ndarray:
        return np.array(list(self.st.sim.minCorner) + list(self.st.sim.maxCorner), np.float32)

    def __del__(self):
        try:
            lib().bho_state_free(C.byref(self.st))
        except Exception:
            pass
