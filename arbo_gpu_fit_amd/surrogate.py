"""ctypes mirror of include/arbo_surrogate.h and the Python side of the
surrogate: the feature map of the MCMC parameter vector and a thin wrapper
over the HIP MLP (training, forward, gradients with respect to inputs)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .engine import ArboError

_u32 = C.c_uint32
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
SurPtr = C.c_void_p

PROTOTYPES = {
    "arbo_sur_create": (C.c_int, [C.c_int32, _u32, C.POINTER(C.c_uint32), _u32, C.c_uint64, C.POINTER(SurPtr)]),
    "arbo_sur_n_params": (C.c_int, [SurPtr, C.POINTER(C.c_uint64)]),
    "arbo_sur_shape": (C.c_int, [SurPtr, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "arbo_sur_get_params": (C.c_int, [SurPtr, _f32p]),
    "arbo_sur_set_params": (C.c_int, [SurPtr, _f32p]),
    "arbo_sur_forward": (C.c_int, [SurPtr, _u32, _f32p, _f32p]),
    "arbo_sur_forward_noalloc": (C.c_int, [SurPtr, _u32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "arbo_sur_input_grad": (C.c_int, [SurPtr, _u32, _f32p, _f32p, _f32p, _f32p]),
    "arbo_sur_loss_grad": (C.c_int, [SurPtr, _u32, _f32p, _f32p, _f64p, _f32p]),
    "arbo_sur_train": (C.c_int, [SurPtr, _u32, _f32p, _f32p, _u32, _u32, C.c_float, C.c_uint64, _f64p]),
    "arbo_sur_destroy": (None, [SurPtr]),
    "arbo_sur_last_error": (C.c_char_p, [SurPtr]),
}


def bind(lib: C.CDLL) -> C.CDLL:
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Surrogate:
    """The MLP of include/arbo_surrogate.h on one device."""

    def __init__(self, widths, max_batch: int = 4096, seed: int = 1, device: int = 0, lib: C.CDLL | None = None):
        self.lib = bind(lib if lib is not None else abi.load_hip_library())
        self.widths = [int(w) for w in widths]
        self.max_batch = max_batch
        w = (C.c_uint32 * len(self.widths))(*self.widths)
        self._h = SurPtr()
        rc = self.lib.arbo_sur_create(device, len(self.widths), w, max_batch, seed, C.byref(self._h))
        if rc:
            raise ArboError(rc, "arbo_sur_create")

    def _check(self, rc):
        if rc:
            msg = self.lib.arbo_sur_last_error(self._h)
            raise ArboError(rc, msg.decode() if msg else "")

    @property
    def n_params(self) -> int:
        n = C.c_uint64()
        self._check(self.lib.arbo_sur_n_params(self._h, C.byref(n)))
        return n.value

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        self._check(self.lib.arbo_sur_get_params(self._h, out.ctypes.data_as(_f32p)))
        return out

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.n_params
        self._check(self.lib.arbo_sur_set_params(self._h, p.ctypes.data_as(_f32p)))

    def layers(self, p=None):
        """[(W [out, in], b [out])] views of a flat parameter vector."""
        p = self.get_params() if p is None else p
        out, o = [], 0
        for k, n in zip(self.widths[:-1], self.widths[1:]):
            W = p[o:o + n * k].reshape(n, k)
            o += n * k
            b = p[o:o + n]
            o += n
            out.append((W, b))
        return out

    def forward(self, x) -> np.ndarray:
        x = _f32(x)
        n = x.shape[0]
        out = np.zeros((n, self.widths[-1]), np.float32)
        self._check(self.lib.arbo_sur_forward(self._h, n, x.ctypes.data_as(_f32p), out.ctypes.data_as(_f32p)))
        return out

    def input_grad(self, x, dout):
        x, dout = _f32(x), _f32(dout)
        n = x.shape[0]
        out = np.zeros((n, self.widths[-1]), np.float32)
        dx = np.zeros_like(x)
        self._check(self.lib.arbo_sur_input_grad(self._h, n, x.ctypes.data_as(_f32p), dout.ctypes.data_as(_f32p),
                                                 out.ctypes.data_as(_f32p), dx.ctypes.data_as(_f32p)))
        return out, dx

    def loss_grad(self, x, y):
        x, y = _f32(x), _f32(y)
        loss = C.c_double()
        g = np.zeros(self.n_params, np.float32)
        self._check(self.lib.arbo_sur_loss_grad(self._h, x.shape[0], x.ctypes.data_as(_f32p),
                                                y.ctypes.data_as(_f32p), C.byref(loss), g.ctypes.data_as(_f32p)))
        return loss.value, g

    def train(self, x, y, epochs: int, batch: int, lr: float, seed: int = 1) -> np.ndarray:
        x, y = _f32(x), _f32(y)
        hist = np.zeros(epochs)
        self._check(self.lib.arbo_sur_train(self._h, x.shape[0], x.ctypes.data_as(_f32p), y.ctypes.data_as(_f32p),
                                            epochs, batch, lr, seed, hist.ctypes.data_as(_f64p)))
        return hist

    def close(self):
        if self._h:
            self.lib.arbo_sur_destroy(self._h)
            self._h = SurPtr()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def features(theta, priors) -> np.ndarray:
    """Surrogate inputs of MCMC parameter vectors [n, 6]: parameters 0-4 in
    order, each by its prior (uniform scaled from its range to [-1, 1],
    circular as (cos, sin), truncated normal (x - a) / b, Beta standardised),
    zeros after.  The reporting rate is not an input (it scales the
    surrogate's mean in the likelihood)."""
    th = np.atleast_2d(np.asarray(theta, dtype=np.float64))
    x = np.zeros((th.shape[0], 6))
    j = 0
    for i in range(5):
        k, a, b = priors[i]
        if k == 1:      # circular
            x[:, j], x[:, j + 1] = np.cos(th[:, i]), np.sin(th[:, i])
            j += 2
            continue
        if k == 3:      # truncated normal
            c, s = a, b
        elif k == 2:    # Beta
            c, s = a / (a + b), np.sqrt(a * b / ((a + b) ** 2 * (a + b + 1.0)))
        else:           # uniform
            c, s = 0.5 * (a + b), 0.5 * (b - a)
        x[:, j] = (th[:, i] - c) / s
        j += 1
    return x
