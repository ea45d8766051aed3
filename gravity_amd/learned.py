"""Host side of the learned far-field correction (no reference counterpart).

The corrected walk adds, for every accepted internal node, the softened
quadrupole of T = M2 + M h^2 o: M2 the node's exact second moment about its
centre of mass, h half its longest side, and o the output of a 12-64-64-6
SiLU network evaluated once per node and step from the node's shape (AABB
extents, centre-of-mass offset, normalised second moment). The network learns
the residual of the node's exact field after its monopole and exact
quadrupole. This module wraps the ``bh_mlp_*`` and
``bh_collect_training_pairs`` entry points of ``include/bh_abi.h``: pair
generation with FP64 labels on the GPU, FP32 training with Adam on the GPU
through the quadrupole basis, BF16 MFMA inference, and attaching the network
to a simulation so that ``BH_FLAG_LEARNED`` steps use it
(``BH_FLAG_QUADRUPOLE`` is the same walk without a network).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import barnes_hut as B

N_FEATURES = 12
N_OUT = 6       # the residual tensor o: xx, xy, xz, yy, yz, zz
N_TARGET = 7    # a training row's target: u = r / s (3), the scaled residual (3), the row's weight


def _declare(L):
    P = C.c_void_p
    L.bh_mlp_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    L.bh_mlp_destroy.argtypes = [P]
    L.bh_mlp_destroy.restype = None
    L.bh_mlp_last_error.argtypes = [P]
    L.bh_mlp_last_error.restype = C.c_char_p
    L.bh_mlp_num_params.restype = C.c_uint32
    L.bh_mlp_set_params.argtypes = [P, P, C.c_uint32]
    L.bh_mlp_get_params.argtypes = [P, P, C.c_uint32]
    L.bh_mlp_forward.argtypes = [P, P, C.c_uint32, P, C.c_int]
    L.bh_mlp_loss_grad.argtypes = [P, P, P, C.c_uint32, P, C.POINTER(C.c_float)]
    L.bh_mlp_train.argtypes = [P, P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64,
                               C.POINTER(C.c_float)]
    L.bh_collect_training_pairs.argtypes = [P, C.c_float, C.c_uint32, C.c_float, C.c_uint32, C.c_uint32, P, P, P,
                                            C.POINTER(C.c_uint32)]
    L.bh_attach_correction.argtypes = [P, P]


def _lib():
    L = B.load_library()
    if not getattr(L, "_learned_declared", False):
        _declare(L)
        L._learned_declared = True
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def num_params() -> int:
    return int(_lib().bh_mlp_num_params())


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initialisation; the output layer starts at zero (no correction)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(num_params(), np.float32)
    o = 0
    for fan_out, fan_in, zero in ((64, 12, False), (64, 64, False), (N_OUT, 64, True)):
        w = np.zeros((fan_out, fan_in)) if zero else rng.normal(0, np.sqrt(2.0 / fan_in), (fan_out, fan_in))
        p[o:o + fan_out * fan_in] = w.ravel()
        o += fan_out * fan_in + fan_out
    return p


class Mlp:
    def __init__(self, device: int = 0, params: np.ndarray | None = None):
        self.L = _lib()
        self.h = C.c_void_p()
        rc = self.L.bh_mlp_create(device, C.byref(self.h))
        if rc != 0:
            raise RuntimeError(f"bh_mlp_create failed: {rc}")
        if params is not None:
            self.set_params(params)

    def _check(self, rc, what):
        if rc != 0:
            msg = self.L.bh_mlp_last_error(self.h)
            raise RuntimeError(f"{what}: {B.ERRORS.get(rc, rc)} {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.L.bh_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p: np.ndarray) -> None:
        a = np.ascontiguousarray(p, dtype=np.float32)
        self._check(self.L.bh_mlp_set_params(self.h, _ptr(a), a.size), "set_params")

    def params(self) -> np.ndarray:
        out = np.zeros(num_params(), np.float32)
        self._check(self.L.bh_mlp_get_params(self.h, _ptr(out), out.size), "get_params")
        return out

    def forward(self, x: np.ndarray, mfma: bool = False) -> np.ndarray:
        xx = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((xx.shape[0], N_OUT), np.float32)
        self._check(self.L.bh_mlp_forward(self.h, _ptr(xx), xx.shape[0], _ptr(y), int(mfma)), "forward")
        return y

    def loss_grad(self, x: np.ndarray, t: np.ndarray):
        xx = np.ascontiguousarray(x, dtype=np.float32)
        tt = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(num_params(), np.float32)
        loss = C.c_float()
        self._check(self.L.bh_mlp_loss_grad(self.h, _ptr(xx), _ptr(tt), xx.shape[0], _ptr(g), C.byref(loss)),
                    "loss_grad")
        return float(loss.value), g

    def train(self, x, t, epochs=20, batch=4096, lr=1e-3, seed=0) -> float:
        xx = np.ascontiguousarray(x, dtype=np.float32)
        tt = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        self._check(self.L.bh_mlp_train(self.h, _ptr(xx), _ptr(tt), xx.shape[0], epochs, batch, lr, seed,
                                        C.byref(loss)), "train")
        return float(loss.value)


def collect_training_pairs(bh: B.BarnesHut, theta: float, stride: int, softening: float,
                           max_node_bodies: int = 1 << 20, cap: int = 1 << 22, with_pairs: bool = False):
    """Features and FP64-labelled targets of the (body, accepted node) pairs
    of the walk at `theta` on bh's current tree (and the (sorted body,
    internal node) pairs). When the walkers at `stride` find more than `cap`
    pairs, the stride is raised until they fit, so the sample always spans the
    whole sorted order. Pairs whose node is too large to label, whose node
    holds the target itself (its residual is the target's own mass, not the
    node's far field), or whose label is not finite (a node of coincident
    bodies) are dropped.

    Each row's weight (target column 6) is (unit / |a|)^2, clipped at 1000
    times its median and normalised to mean 1: unit = (h / s)^2 G M / s^2 is
    the scale of the node's correction and |a| the target's total acceleration
    (FP64 direct sum), so the loss measures the rows' errors relative to the
    force they perturb."""
    L = _lib()
    f = np.zeros((cap, N_FEATURES), np.float32)
    t = np.zeros((cap, N_TARGET), np.float32)
    pr = np.zeros((cap, 2), np.uint32)
    n = C.c_uint32()
    for _ in range(8):
        rc = L.bh_collect_training_pairs(bh.ctx, theta, stride, softening, max_node_bodies, cap, _ptr(f), _ptr(t),
                                         _ptr(pr), C.byref(n))
        bh._check(rc, "collect_training_pairs")
        if n.value <= cap:
            break
        stride = int(np.ceil(stride * n.value / cap * 1.05))
    else:
        raise RuntimeError("collect_training_pairs: pairs do not fit in cap")
    f, t, pr = f[: n.value], t[: n.value], pr[: n.value]
    keep = np.isfinite(t).all(axis=1) & np.isfinite(f).all(axis=1)
    # the walkers append in whatever order their atomics land: sorted by
    # (body, node), the rows (and a training run on them) are reproducible
    order = np.lexsort((pr[keep, 1], pr[keep, 0]))
    f, t, pr = f[keep][order], t[keep][order], pr[keep][order]
    # |a| of every target (sorted positions -> body ids of the current order)
    targets, inverse = np.unique(pr[:, 0], return_inverse=True)
    ids = bh.index()[targets]
    a = np.linalg.norm(bh.direct_sum_f64(ids, softening), axis=1)
    w = (t[:, 6].astype(np.float64) / a[inverse]) ** 2
    if w.size:
        w = np.minimum(w, 1000.0 * np.median(w))
        t[:, 6] = (w / w.mean()).astype(np.float32)
    if with_pairs:
        return f, t, pr
    return f, t


def attach(bh: B.BarnesHut, mlp: Mlp | None) -> None:
    L = _lib()
    bh._check(L.bh_attach_correction(bh.ctx, mlp.h if mlp is not None else None), "attach_correction")
    bh._correction = mlp   # keep the network alive while attached


def scaled_params(params: np.ndarray, alpha: float) -> np.ndarray:
    """The parameters with the last layer (W3, b3) scaled by alpha: the
    network's residual shrunk by alpha (0: the exact quadrupole)."""
    p = np.array(params, dtype=np.float32, copy=True)
    p[-(N_OUT * 64 + N_OUT):] *= np.float32(alpha)
    return p


def calibrate(mlp: Mlp, bodies: np.ndarray, theta: float, softening: float, n_sample: int = 4096,
              alphas=(0.0, 0.25, 0.5, 0.75, 1.0), seed: int = 0,
              force_tolerance: float | None = None) -> tuple[float, list]:
    """Choose the output scale alpha on a validation snapshot: the alpha whose
    corrected walk (the geometric test at `theta`, or with `force_tolerance`
    the relative test) has the lowest p99 force error against the FP64 direct
    sum (then p50). The network is left with that scale. Returns (alpha,
    [(alpha, p50, p99), ...])."""
    base = mlp.params()
    bh = B.BarnesHut(bodies, _bounds(bodies), B.make_params(theta=theta, softening=softening))
    bh.init()
    rng = np.random.default_rng(seed)
    sample = np.sort(rng.choice(bodies.shape[0], min(n_sample, bodies.shape[0]), replace=False)).astype(np.uint32)
    exact = bh.direct_sum_f64(sample, softening)
    attach(bh, mlp)
    rows = []
    for a in alphas:
        mlp.set_params(scaled_params(base, a))
        if force_tolerance is None:
            prm = B.make_params(theta=theta, softening=softening, flags=B.FLAG_LEARNED)
        else:
            prm = B.make_params(theta=theta, softening=softening, flags=B.FLAG_LEARNED | B.FLAG_RELATIVE_MAC,
                                force_tolerance=force_tolerance)
        acc = bh.accelerations(prm)
        e = np.linalg.norm(acc[sample].astype(np.float64) - exact, axis=1) / np.linalg.norm(exact, axis=1)
        rows.append((float(a), float(np.percentile(e, 50)), float(np.percentile(e, 99))))
    attach(bh, None)
    bh.close()
    best = min(rows, key=lambda r: (r[2], r[1]))
    mlp.set_params(scaled_params(base, best[0]))
    return best[0], rows


def _bounds(bodies: np.ndarray) -> np.ndarray:
    alive = bodies[:, 3] != 0
    p = bodies[alive, 0:3]
    return np.concatenate([p.min(axis=0), p.max(axis=0)]).astype(np.float32)
