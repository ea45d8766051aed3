"""Host side of the learned far-field correction (no reference counterpart).

A 12-64-64-3 SiLU network predicts, for a (target body, tree node) pair that
the wide opening angle accepts but the plain angle would open, the residual
between the node's exact field and its monopole in units of G M / d^2. This
module wraps the ``bh_mlp_*`` and ``bh_collect_training_pairs`` entry points
of ``include/bh_abi.h``: pair generation with FP64 labels on the GPU, FP32
training with Adam on the GPU, BF16 MFMA inference, and attaching the network
to a simulation so that ``BH_FLAG_LEARNED`` steps use it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import barnes_hut as B

N_FEATURES = 12
N_OUT = 3


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
    L.bh_collect_training_pairs.argtypes = [P, C.c_float, C.c_float, C.c_uint32, C.c_float, C.c_uint32,
                                            C.c_uint32, P, P, P, C.POINTER(C.c_uint32)]
    L.bh_attach_correction.argtypes = [P, P, C.c_float]


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
    for fan_out, fan_in, zero in ((64, 12, False), (64, 64, False), (3, 64, True)):
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


def collect_training_pairs(bh: B.BarnesHut, theta: float, theta_plain: float, stride: int, softening: float,
                           max_node_bodies: int = 1 << 20, cap: int = 1 << 22, with_pairs: bool = False):
    """Features and FP64-labelled targets of the pairs on bh's current tree
    (and the (sorted body, internal node) pairs). When the walkers at
    `stride` find more than `cap` pairs, the stride is raised until they fit,
    so the sample always spans the whole sorted order. Pairs whose node is
    too large to label are dropped."""
    L = _lib()
    f = np.zeros((cap, N_FEATURES), np.float32)
    t = np.zeros((cap, N_OUT), np.float32)
    pr = np.zeros((cap, 2), np.uint32)
    n = C.c_uint32()
    for _ in range(8):
        rc = L.bh_collect_training_pairs(bh.ctx, theta, theta_plain, stride, softening, max_node_bodies, cap,
                                         _ptr(f), _ptr(t), _ptr(pr), C.byref(n))
        bh._check(rc, "collect_training_pairs")
        if n.value <= cap:
            break
        stride = int(np.ceil(stride * n.value / cap * 1.05))
    else:
        raise RuntimeError("collect_training_pairs: pairs do not fit in cap")
    f, t, pr = f[: n.value], t[: n.value], pr[: n.value]
    keep = np.isfinite(t).all(axis=1) & np.isfinite(f).all(axis=1)
    if with_pairs:
        return f[keep], t[keep], pr[keep]
    return f[keep], t[keep]


def attach(bh: B.BarnesHut, mlp: Mlp | None, theta_plain: float) -> None:
    L = _lib()
    bh._check(L.bh_attach_correction(bh.ctx, mlp.h if mlp is not None else None, theta_plain), "attach_correction")
    bh._correction = mlp   # keep the network alive while attached
