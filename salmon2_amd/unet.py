"""ctypes binding of the learned potential predictor (include/salmon_unet.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import hip

EXPORTS = (
    "salmon_unet_create",
    "salmon_unet_destroy",
    "salmon_unet_set_stream",
    "salmon_unet_num_params",
    "salmon_unet_get_params",
    "salmon_unet_set_params",
    "salmon_unet_forward",
    "salmon_unet_train_sample",
    "salmon_unet_adam_step",
    "salmon_unet_get_grads",
    "salmon_unet_features",
    "salmon_unet_predict",
    "salmon_unet_target",
    "salmon_unet_residual",
    "salmon_unet_gate",
    "salmon_unet_get_activation",
)

_bound = False


def _lib():
    global _bound
    L = hip.load()
    if not _bound:
        vp = C.c_void_p
        fp = C.POINTER(C.c_float)
        L.salmon_unet_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.POINTER(vp)]
        L.salmon_unet_destroy.argtypes = [vp]
        L.salmon_unet_set_stream.argtypes = [vp, vp]
        L.salmon_unet_num_params.argtypes = [vp, C.POINTER(C.c_long)]
        L.salmon_unet_get_params.argtypes = [vp, fp]
        L.salmon_unet_set_params.argtypes = [vp, fp]
        L.salmon_unet_forward.argtypes = [vp, vp, vp]
        L.salmon_unet_train_sample.argtypes = [vp, vp, vp, vp]
        L.salmon_unet_adam_step.argtypes = [vp, C.c_float, C.c_int]
        L.salmon_unet_get_grads.argtypes = [vp, fp]
        L.salmon_unet_features.argtypes = [vp] + [vp] * 5 + [C.c_double, C.c_double, vp]
        L.salmon_unet_predict.argtypes = [vp, vp, vp, vp, C.c_double, vp]
        L.salmon_unet_target.argtypes = [vp, vp, vp, vp, C.c_double, vp]
        L.salmon_unet_residual.argtypes = [vp, vp, vp, vp, vp]
        L.salmon_unet_gate.argtypes = [vp, vp, C.c_double, vp, vp, vp, C.c_int]
        L.salmon_unet_get_activation.argtypes = [vp, C.c_int, fp, C.c_long]
        _bound = True
    return L


def _dp(t):
    return C.c_void_p(None if t is None else (t if isinstance(t, int) else t.data_ptr()))


class UNet:
    """Predictor network on an (nx, ny, nz) grid; tensors are torch device float32, channels-last."""

    def __init__(self, nx, ny, nz, seed=0, stream=None):
        L = _lib()
        h = C.c_void_p()
        hip._chk(L.salmon_unet_create(nx, ny, nz, seed, C.byref(h)), "salmon_unet_create")
        self.h = h
        self.shape = (nz, ny, nx)
        if stream is not None:
            self.set_stream(stream)
        n = C.c_long()
        hip._chk(L.salmon_unet_num_params(self.h, C.byref(n)), "salmon_unet_num_params")
        self.nparam = n.value

    def set_stream(self, stream):
        """Later launches go to `stream` (a torch stream or a raw handle; 0/None = the null stream)."""
        raw = getattr(stream, "cuda_stream", stream)
        hip._chk(_lib().salmon_unet_set_stream(self.h, C.c_void_p(raw or None)), "salmon_unet_set_stream")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            _lib().salmon_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        w = np.zeros(self.nparam, dtype=np.float32)
        hip._chk(_lib().salmon_unet_get_params(self.h, w.ctypes.data_as(C.POINTER(C.c_float))), "get_params")
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.nparam
        hip._chk(_lib().salmon_unet_set_params(self.h, w.ctypes.data_as(C.POINTER(C.c_float))), "set_params")

    def get_grads(self) -> np.ndarray:
        g = np.zeros(self.nparam, dtype=np.float32)
        hip._chk(_lib().salmon_unet_get_grads(self.h, g.ctypes.data_as(C.POINTER(C.c_float))), "get_grads")
        return g

    ACT_SHAPES = [(16, 0), (16, 0), (16, 1), (32, 1), (32, 1), (32, 2), (64, 2), (64, 2), (32, 1), (32, 1), (16, 0),
                  (16, 0)]

    def activation(self, index):
        """Activation of the last forward pass (index -1 = packed input), (npts_level, channels) float32."""
        nz, ny, nx = self.shape
        c, lv = (16, 0) if index < 0 else self.ACT_SHAPES[index]
        n = (nx >> lv) * (ny >> lv) * (nz >> lv)
        a = np.zeros((n, c), dtype=np.float32)
        hip._chk(_lib().salmon_unet_get_activation(self.h, int(index), a.ctypes.data_as(C.POINTER(C.c_float)), a.size),
                 "get_activation")
        return a

    def forward(self, x, out):
        """x: (npts, 4) float32 device tensor; out: (npts,) float32."""
        hip._chk(_lib().salmon_unet_forward(self.h, _dp(x), _dp(out)), "salmon_unet_forward")

    def train_sample(self, x, target, loss=None):
        hip._chk(_lib().salmon_unet_train_sample(self.h, _dp(x), _dp(target), _dp(loss)), "salmon_unet_train_sample")

    def features(self, v0, v1, v2, v3, vpsl, s_diff, s_abs, x):
        hip._chk(_lib().salmon_unet_features(self.h, _dp(v0), _dp(v1), _dp(v2), _dp(v3), _dp(vpsl), float(s_diff),
                                             float(s_abs), _dp(x)), "salmon_unet_features")

    def predict(self, v0, v1, nn, s_out, vpred):
        hip._chk(_lib().salmon_unet_predict(self.h, _dp(v0), _dp(v1), _dp(nn), float(s_out), _dp(vpred)),
                 "salmon_unet_predict")

    def target(self, vmid, v0, v1, s_out, t):
        hip._chk(_lib().salmon_unet_target(self.h, _dp(vmid), _dp(v0), _dp(v1), float(s_out), _dp(t)),
                 "salmon_unet_target")

    def residual(self, vpred, v0, vnew, out):
        hip._chk(_lib().salmon_unet_residual(self.h, _dp(vpred), _dp(v0), _dp(vnew), _dp(out)), "salmon_unet_residual")

    def gate(self, resid, tol, flag, stats, log=None):
        """flag = (resid <= tol) on the device; stats[0] accepted, stats[1] attempted (device int32);
        log[attempt] = resid while it has room."""
        cap = 0 if log is None else int(log.numel())
        hip._chk(_lib().salmon_unet_gate(self.h, _dp(resid), float(tol), _dp(flag), _dp(stats), _dp(log), cap),
                 "salmon_unet_gate")

    def adam_step(self, lr, nsamples):
        hip._chk(_lib().salmon_unet_adam_step(self.h, float(lr), int(nsamples)), "salmon_unet_adam_step")
