"""Learned closure of the shallow-water step (libtmgpu ``tm_mlp_*`` and
``tm_closure_*``, include/tmcore_gpu.h).

``Mlp`` is the per-edge network 40 -> 128 -> 128 -> 128 -> 2 (GELU, BF16
weights on MFMA, FP32 accumulation and FP32 master weights with Adam), run
and trained by the library's own kernels.  ``Closure`` binds an ``Mlp`` to a
``SwmGpu`` model: the inputs of the model's current state, attaching the
correction to every following step, a training set held on the device and
training/evaluation over it.  There is no host fallback: every call goes to
the HIP library and raises ``TmError`` when it fails.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from .swm import SwmGpu, gpu_lib

NIN, NHID, NOUT = 40, 128, 2
NPARAMS = NIN * NHID + NHID + 2 * (NHID * NHID + NHID) + NHID * NOUT + NOUT  # 38530
FIELDS = ("u", "gd_edge", "vorticity_edge", "divergence_edge")  # input j = field * 10 + slot

_vp = C.c_void_p
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_bound = False


def _lib():
    global _bound
    L = gpu_lib()
    if not _bound:
        L.tm_mlp_create.argtypes = [C.POINTER(_vp), C.c_int, C.c_uint64]
        L.tm_mlp_destroy.argtypes = [_vp]
        L.tm_mlp_last_error.argtypes = [_vp]
        L.tm_mlp_last_error.restype = C.c_char_p
        L.tm_mlp_get_params.argtypes = [_vp, _fp]
        L.tm_mlp_set_params.argtypes = [_vp, _fp]
        L.tm_mlp_set_scales.argtypes = [_vp, _fp, _fp, _fp]
        L.tm_mlp_get_scales.argtypes = [_vp, _fp, _fp, _fp]
        L.tm_mlp_forward.argtypes = [_vp, _fp, C.c_int, _fp]
        L.tm_mlp_backward.argtypes = [_vp, _fp, _fp, C.c_int, _fp, _fp]
        L.tm_mlp_adam_step.argtypes = [_vp, _fp, C.c_float]
        L.tm_closure_features.argtypes = [_vp, _vp, _fp]
        L.tm_closure_attach.argtypes = [_vp, _vp]
        L.tm_closure_set_interval.argtypes = [_vp, C.c_int, C.c_int]
        L.tm_closure_set_gain.argtypes = [_vp, C.c_double]
        L.tm_closure_dataset_add.argtypes = [_vp, _vp, _dp, _dp, C.c_double]
        L.tm_closure_dataset_size.argtypes = [_vp, C.POINTER(C.c_int)]
        L.tm_closure_dataset_clear.argtypes = [_vp]
        L.tm_closure_eval.argtypes = [_vp, _vp, C.c_int, C.c_int, _dp]
        L.tm_closure_train.argtypes = [_vp, _vp, C.c_int, C.c_int, C.c_float, C.c_uint64, _dp]
        _bound = True
    return L


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def _pf(a):
    return None if a is None else a.ctypes.data_as(_fp)


def _pd(a):
    return None if a is None else a.ctypes.data_as(_dp)


def unpack(params):
    """Flat parameter vector -> dict W1 [40,128], b1, W2, b2, W3, b3, W4 [128,2], b4."""
    p = np.asarray(params)
    out, o = {}, 0
    for name, shape in (("W1", (NIN, NHID)), ("b1", (NHID,)), ("W2", (NHID, NHID)), ("b2", (NHID,)),
                        ("W3", (NHID, NHID)), ("b3", (NHID,)), ("W4", (NHID, NOUT)), ("b4", (NOUT,))):
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


class Mlp:
    """The closure network on one device."""

    def __init__(self, seed: int = 0, device: int = 0):
        h = _vp()
        rc = _lib().tm_mlp_create(C.byref(h), device, seed)
        if rc != 0:
            raise _ffi.TmError(rc, "tm_mlp_create")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            _lib().tm_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, where):
        if rc != 0:
            msg = _lib().tm_mlp_last_error(self.h)
            raise _ffi.TmError(rc, where, msg.decode() if msg else "")

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(NPARAMS, np.float32)
        self._check(_lib().tm_mlp_get_params(self.h, _pf(p)), "tm_mlp_get_params")
        return p

    @params.setter
    def params(self, p):
        p = _f32(p, (NPARAMS,))
        self._check(_lib().tm_mlp_set_params(self.h, _pf(p)), "tm_mlp_set_params")

    def set_scales(self, in_shift, in_scale, out_scale):
        a, b, c = _f32(in_shift, (NIN,)), _f32(in_scale, (NIN,)), _f32(out_scale, (NOUT,))
        self._check(_lib().tm_mlp_set_scales(self.h, _pf(a), _pf(b), _pf(c)), "tm_mlp_set_scales")

    def get_scales(self):
        a, b, c = np.zeros(NIN, np.float32), np.zeros(NIN, np.float32), np.zeros(NOUT, np.float32)
        self._check(_lib().tm_mlp_get_scales(self.h, _pf(a), _pf(b), _pf(c)), "tm_mlp_get_scales")
        return a, b, c

    def forward(self, x) -> np.ndarray:
        x = _f32(x)
        n = x.shape[0]
        if x.shape != (n, NIN):
            raise ValueError("x must be [n, 40]")
        y = np.zeros((n, NOUT), np.float32)
        self._check(_lib().tm_mlp_forward(self.h, _pf(x), n, _pf(y)), "tm_mlp_forward")
        return y

    def backward(self, x, dy, want_dx: bool = True):
        """Gradients of sum(dy * MLP(x)): (grad[NPARAMS], dx[n, 40] or None)."""
        x = _f32(x)
        n = x.shape[0]
        dy = _f32(dy, (n, NOUT))
        g = np.zeros(NPARAMS, np.float32)
        dx = np.zeros((n, NIN), np.float32) if want_dx else None
        self._check(_lib().tm_mlp_backward(self.h, _pf(x), _pf(dy), n, _pf(g), _pf(dx)), "tm_mlp_backward")
        return g, dx

    def adam_step(self, grad, lr: float):
        g = _f32(grad, (NPARAMS,))
        self._check(_lib().tm_mlp_adam_step(self.h, _pf(g), lr), "tm_mlp_adam_step")


class Closure:
    """An ``Mlp`` bound to a ``SwmGpu`` model (same device)."""

    def __init__(self, model: SwmGpu, mlp: Mlp):
        self.g = model
        self.mlp = mlp

    def features(self) -> np.ndarray:
        """Normalised inputs of the model's current state, [nEdges, 40]."""
        x = np.zeros((self.g.nE, NIN), np.float32)
        self.g._check(_lib().tm_closure_features(self.g.h, self.mlp.h, _pf(x)), "tm_closure_features")
        return x

    def attach(self):
        self.g._check(_lib().tm_closure_attach(self.g.h, self.mlp.h), "tm_closure_attach")

    def detach(self):
        self.g._check(_lib().tm_closure_attach(self.g.h, None), "tm_closure_attach")

    def set_interval(self, nsteps: int, lumped: bool = False):
        """Run the MLP every nsteps steps; add its correction after every step
        (lumped False) or once per interval, times nsteps (lumped True)."""
        self.g._check(_lib().tm_closure_set_interval(self.g.h, nsteps, int(lumped)), "tm_closure_set_interval")

    def set_gain(self, gain: float):
        """Factor on every applied correction (default 1)."""
        self.g._check(_lib().tm_closure_set_gain(self.g.h, gain), "tm_closure_set_gain")

    def add_sample(self, t_u, t_gd, tg_scale: float):
        """Training sample: inputs of the current state plus targets
        t_u[nEdges] = (iap_target - iap_plain) / (dt sqrt(gd_edge)) and
        t_gd[nCells] = (gd_target - gd_plain) / dt."""
        t_u = np.ascontiguousarray(t_u, np.float64)
        t_gd = np.ascontiguousarray(t_gd, np.float64)
        if t_u.shape != (self.g.nE,) or t_gd.shape != (self.g.nC,):
            raise ValueError("target shapes must be [nEdges] and [nCells]")
        self.g._check(_lib().tm_closure_dataset_add(self.g.h, self.mlp.h, _pd(t_u), _pd(t_gd), tg_scale),
                      "tm_closure_dataset_add")

    def size(self) -> int:
        n = C.c_int()
        self.g._check(_lib().tm_closure_dataset_size(self.g.h, C.byref(n)), "tm_closure_dataset_size")
        return n.value

    def clear(self):
        self.g._check(_lib().tm_closure_dataset_clear(self.g.h), "tm_closure_dataset_clear")

    def eval(self, first: int = 0, count: int | None = None) -> np.ndarray:
        """Mean (velocity, thickness) loss over samples [first, first + count)."""
        count = self.size() - first if count is None else count
        out = np.zeros(2)
        self.g._check(_lib().tm_closure_eval(self.g.h, self.mlp.h, first, count, _pd(out)), "tm_closure_eval")
        return out

    def train(self, ntrain: int, epochs: int, lr: float, seed: int = 0) -> np.ndarray:
        """Adam over samples [0, ntrain); returns the per-epoch loss terms [epochs, 2]."""
        hist = np.zeros((epochs, 2))
        self.g._check(_lib().tm_closure_train(self.g.h, self.mlp.h, ntrain, epochs, lr, seed, _pd(hist)),
                      "tm_closure_train")
        return hist
