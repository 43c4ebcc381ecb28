"""Learned flux surrogate: a BF16 MLP that maps a heliostat's pose, the sun
and its aim offset to its receiver flux image, trained on images the GPU
tracer makes and served inside an aim-point sweep with the tracer as
fallback and judge.

The reference computes receiver flux only by tracing and binning
(app/src/project.cpp:1139-1436) and aims heliostats by a fixed rule
(app/deploy/api/examples/tower-demo.py:49-54); nothing here changes what a
trace computes.  The parts:

* :class:`Mlp` -- the network of soltrace_amd/csrc/st_mlp.hip through its
  C ABI (st_amd_mlp_*): device memory from torch tensors, every call on
  torch's current stream.
* :func:`field` / :func:`trace_images` -- the tower field with per-heliostat
  aim offsets, traced with per-source flux maps (st_amd_flux_by_source):
  one 64 x 64 receiver image per heliostat per trace.
* :class:`Surrogate` -- input scaling, target scaling and training.
* :func:`aim_sweep` -- greedy peak-flux flattening over aim offsets, with
  the tracer as fallback and judge.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field as dfield

import numpy as np

from . import stapi, workloads


# ----------------------------------------------------------------- network --

def _mlp_sigs(dll):
    vp = C.c_void_p
    sig = {
        "st_amd_mlp_create": (vp, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong]),
        "st_amd_mlp_free": (None, [vp]),
        "st_amd_mlp_error": (C.c_char_p, []),
        "st_amd_mlp_num_params": (C.c_longlong, [vp]),
        "st_amd_mlp_get_params": (C.c_int, [vp, C.POINTER(C.c_float), C.c_int]),
        "st_amd_mlp_set_params": (C.c_int, [vp, C.POINTER(C.c_float)]),
        "st_amd_mlp_forward": (C.c_int, [vp, vp, C.c_int, vp, vp]),
        "st_amd_mlp_train_step": (C.c_int, [vp, vp, vp, C.c_int, C.c_float, C.POINTER(C.c_double), vp]),
        "st_amd_mlp_gradients": (C.c_int, [vp, vp, vp, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_double),
                                           vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(dll, name)
        f.restype = res
        f.argtypes = args


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Mlp:
    """n_in -> depth x [width, ReLU] -> n_out, BF16 weights and activations,
    FP32 accumulation (st_amd_mlp_* in include/stapi_amd.h)."""

    def __init__(self, lib: stapi.StLib, n_in: int, width: int, depth: int, n_out: int, max_batch: int,
                 seed: int = 1):
        self.dll = lib.dll
        _mlp_sigs(self.dll)
        self.shape = (n_in, width, depth, n_out)
        self.max_batch = max_batch
        self.h = self.dll.st_amd_mlp_create(n_in, width, depth, n_out, max_batch, seed)
        if not self.h:
            raise RuntimeError("st_amd_mlp_create: " + self.dll.st_amd_mlp_error().decode())

    def close(self):
        if self.h:
            self.dll.st_amd_mlp_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ok(self, rc, what):
        if rc != 1:
            raise RuntimeError(f"{what}: " + self.dll.st_amd_mlp_error().decode())

    @property
    def num_params(self) -> int:
        return int(self.dll.st_amd_mlp_num_params(self.h))

    def get_params(self, which: int = 0) -> np.ndarray:
        out = np.zeros(self.num_params, dtype=np.float32)
        self._ok(self.dll.st_amd_mlp_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float)), which),
                 "get_params")
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params
        self._ok(self.dll.st_amd_mlp_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))), "set_params")

    @staticmethod
    def _dev(t, dtype):
        import torch
        assert t.is_cuda and t.dtype == dtype and t.is_contiguous(), (t.device, t.dtype)
        return C.c_void_p(t.data_ptr())

    def forward(self, x, out=None):
        """x: cuda float32 (b, n_in) -> cuda float32 (b, n_out)."""
        import torch
        b = x.shape[0]
        y = out if out is not None else torch.empty((b, self.shape[3]), device=x.device, dtype=torch.float32)
        self._ok(self.dll.st_amd_mlp_forward(self.h, self._dev(x, torch.float32), b, self._dev(y, torch.float32),
                                             _stream()), "forward")
        return y

    def train_step(self, x, t, lr: float, want_loss: bool = False):
        """One Adam step on the MSE to t (cuda bfloat16 (b, n_out)); the loss
        before the step when asked (that synchronises)."""
        import torch
        loss = C.c_double(0.0)
        self._ok(self.dll.st_amd_mlp_train_step(self.h, self._dev(x, torch.float32), self._dev(t, torch.bfloat16),
                                                x.shape[0], lr, C.byref(loss) if want_loss else None,
                                                _stream()), "train_step")
        return loss.value if want_loss else None

    def gradients(self, x, t):
        """(loss, gradient in exchange order) at the current weights."""
        import torch
        g = np.zeros(self.num_params, dtype=np.float32)
        loss = C.c_double(0.0)
        self._ok(self.dll.st_amd_mlp_gradients(self.h, self._dev(x, torch.float32), self._dev(t, torch.bfloat16),
                                               x.shape[0], g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss),
                                               _stream()), "gradients")
        return loss.value, g
