"""ctypes binding of the warm-start network (include/lbm_surrogate.h) and the
feature/target encoding of 2D channel flows it is trained on.

Features per cell (channel order): solid flag, y across the channel in
[-1, 1], x along the channel in [0, 1], distance to the nearest solid cell
or wall / 16, log(Re) / log(200).  Targets: u_x / u_in, u_y / u_in,
(rho - 1) / u_in^2.
"""
import ctypes as ct
import math

import numpy as np

from . import capi

IN_CHANNELS = 5


class SurrogateError(RuntimeError):
    pass


def _bind(lib):
    if getattr(lib, "_sur_bound", False):
        return lib
    P = ct.c_void_p
    fp = ct.POINTER(ct.c_float)
    lib.lbm_sur_create.argtypes = [ct.c_int32, ct.c_int32, ct.c_int32, ct.c_int32, ct.c_uint64, ct.POINTER(P)]
    lib.lbm_sur_param_count.argtypes = [P, ct.POINTER(ct.c_int64)]
    lib.lbm_sur_get_params.argtypes = [P, fp]
    lib.lbm_sur_set_params.argtypes = [P, fp]
    lib.lbm_sur_forward.argtypes = [P, fp, ct.c_int32, fp]
    lib.lbm_sur_loss_grad.argtypes = [P, fp, fp, ct.POINTER(ct.c_uint8), ct.c_int32, ct.POINTER(ct.c_double), fp]
    lib.lbm_sur_set_dataset.argtypes = [P, fp, fp, ct.POINTER(ct.c_uint8), ct.c_int64]
    lib.lbm_sur_train.argtypes = [P, ct.POINTER(ct.c_int64), ct.c_int32, ct.c_float, ct.POINTER(ct.c_double)]
    lib.lbm_sur_predict_dataset.argtypes = [P, ct.c_int64, ct.c_int32, fp]
    lib.lbm_sur_last_error.argtypes = [P]
    lib.lbm_sur_last_error.restype = ct.c_char_p
    lib.lbm_sur_destroy.argtypes = [P]
    for n in EXPORTS:
        if n != "lbm_sur_last_error":
            getattr(lib, n).restype = ct.c_int
    lib._sur_bound = True
    return lib


EXPORTS = ["lbm_sur_create", "lbm_sur_param_count", "lbm_sur_get_params", "lbm_sur_set_params", "lbm_sur_forward",
           "lbm_sur_loss_grad", "lbm_sur_set_dataset", "lbm_sur_train", "lbm_sur_predict_dataset",
           "lbm_sur_last_error", "lbm_sur_destroy"]


def _f(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_float))


class Surrogate:
    """U-Net warm-start model on the GPU."""

    def __init__(self, batch, height, width, in_channels=IN_CHANNELS, seed=0):
        self.lib = _bind(capi.load())
        self.batch, self.H, self.W, self.cin = int(batch), int(height), int(width), int(in_channels)
        h = ct.c_void_p()
        rc = self.lib.lbm_sur_create(self.batch, self.H, self.W, self.cin, int(seed), ct.byref(h))
        if rc != 0:
            raise SurrogateError(self.lib.lbm_sur_last_error(None).decode())
        self.h = h
        n = ct.c_int64()
        self._check(self.lib.lbm_sur_param_count(self.h, ct.byref(n)))
        self.nparam = n.value

    def _check(self, rc):
        if rc != 0:
            raise SurrogateError(self.lib.lbm_sur_last_error(self.h).decode())

    def params(self):
        p = np.empty(self.nparam, np.float32)
        self._check(self.lib.lbm_sur_get_params(self.h, _f(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        self._check(self.lib.lbm_sur_set_params(self.h, _f(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = x.shape[0]
        y = np.empty((n, self.H, self.W, 3), np.float32)
        self._check(self.lib.lbm_sur_forward(self.h, _f(x), n, _f(y)))
        return y

    def loss_grad(self, x, t, mask, grad=True):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        loss = ct.c_double()
        g = np.empty(self.nparam, np.float32) if grad else None
        self._check(self.lib.lbm_sur_loss_grad(self.h, _f(x), _f(t), m.ctypes.data_as(ct.POINTER(ct.c_uint8)),
                                               x.shape[0], ct.byref(loss), _f(g) if grad else None))
        return loss.value, g

    def set_dataset(self, x, t, mask):
        self._x = np.ascontiguousarray(x, dtype=np.float32)
        self._t = np.ascontiguousarray(t, dtype=np.float32)
        self._m = np.ascontiguousarray(mask, dtype=np.uint8)
        self.ndata = self._x.shape[0]
        self._check(self.lib.lbm_sur_set_dataset(self.h, _f(self._x), _f(self._t),
                                                 self._m.ctypes.data_as(ct.POINTER(ct.c_uint8)), self.ndata))

    def train(self, idx, lr):
        idx = np.ascontiguousarray(idx, dtype=np.int64).reshape(-1)
        nsteps = idx.size // self.batch
        losses = np.empty(nsteps)
        self._check(self.lib.lbm_sur_train(self.h, idx.ctypes.data_as(ct.POINTER(ct.c_int64)), nsteps, float(lr),
                                           losses.ctypes.data_as(ct.POINTER(ct.c_double))))
        return losses

    def predict_dataset(self, i0, n):
        y = np.empty((n, self.H, self.W, 3), np.float32)
        self._check(self.lib.lbm_sur_predict_dataset(self.h, int(i0), int(n), _f(y)))
        return y

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_sur_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def features(solid, re):
    """Per-cell input features of a channel (solid[H][W] bool, walls at y = -1/2
    and y = H - 1/2) at Reynolds number re: [H][W][5] float32."""
    from scipy.ndimage import distance_transform_edt

    H, W = solid.shape
    fluid = ~solid
    padded = np.zeros((H + 2, W), bool)
    padded[1:-1] = fluid
    dist = distance_transform_edt(padded)[1:-1]
    y = (np.arange(H) + 0.5) / H * 2 - 1
    x = (np.arange(W) + 0.5) / W
    f = np.zeros((H, W, IN_CHANNELS), np.float32)
    f[..., 0] = solid
    f[..., 1] = y[:, None]
    f[..., 2] = x[None, :]
    f[..., 3] = dist / 16.0
    f[..., 4] = math.log(re) / math.log(200.0)
    return f


def encode_targets(rho, u, u_in):
    """(rho[H*W], u[H*W, 2]) of a converged field -> targets [H*W, 3]."""
    return np.stack([u[:, 0] / u_in, u[:, 1] / u_in, (rho - 1.0) / u_in ** 2], axis=1).astype(np.float32)


def decode_prediction(y, solid, u_in):
    """Network output [H][W][3] -> (rho[H*W], u[H*W, 2]) initial fields (solid cells at rest)."""
    y = np.asarray(y, np.float64).reshape(-1, 3)
    s = np.asarray(solid, bool).reshape(-1)
    u = np.stack([y[:, 0] * u_in, y[:, 1] * u_in], axis=1)
    rho = 1.0 + y[:, 2] * u_in ** 2
    u[s] = 0.0
    rho[s] = 1.0
    return rho, u
