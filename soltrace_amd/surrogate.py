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
        "st_amd_mlp_fit": (C.c_int, [vp, vp, vp, C.c_longlong, vp, C.c_int, C.c_int, vp, C.POINTER(C.c_double),
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

    def fit(self, X, T, idx, lr, want_loss: bool = False):
        """len(lr) Adam steps replayed from a HIP graph: step k trains on rows
        idx[k] (cuda int32 (steps, batch)) of X (cuda float32 (rows, n_in))
        and T (cuda bfloat16 (rows, n_out)) at learning rate lr[k] (cuda
        float32).  Returns the last step's loss when asked."""
        import torch
        steps, batch = idx.shape
        assert lr.shape == (steps,) and X.shape[0] == T.shape[0]
        loss = C.c_double(0.0)
        self._ok(self.dll.st_amd_mlp_fit(self.h, self._dev(X, torch.float32), self._dev(T, torch.bfloat16),
                                         X.shape[0], self._dev(idx, torch.int32), batch, steps,
                                         self._dev(lr, torch.float32), C.byref(loss) if want_loss else None,
                                         _stream()), "fit")
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


# ------------------------------------------------------------------- field --

DNI = 1000.0           # W/m^2 on the sun plane; power per ray = DNI * sun-box area / sun rays
REC_HALF = 7.5         # the receiver is 15 x 15 m
IMG = 64               # receiver image side, bins
FLUX_SCALE = 2000.0    # W/m^2 per unit of the network's targets (a near heliostat's image peaks at ~1.5 units)
N_IN = 8


def features(pts: np.ndarray, sun_az: float, sun_el: float, offsets: np.ndarray, slope: float) -> np.ndarray:
    """Network inputs, one row per heliostat: x, y, z (m), sun azimuth and
    elevation (deg), aim offset du, dv (m, along the receiver's horizontal
    and upward axes) and slope error (mrad), each shifted and scaled to
    about [-1, 1] over the ranges the training data spans."""
    n = len(pts)
    f = np.empty((n, N_IN), dtype=np.float32)
    f[:, 0] = pts[:, 0] / 1000.0
    f[:, 1] = pts[:, 1] / 1000.0
    f[:, 2] = pts[:, 2] / 10.0
    f[:, 3] = (sun_az - 180.0) / 30.0
    f[:, 4] = (sun_el - 45.0) / 15.0
    f[:, 5] = offsets[:, 0] / 6.0
    f[:, 6] = offsets[:, 1] / 6.0
    f[:, 7] = (slope - 1.5) / 0.5
    return f


@dataclass
class FieldState:
    """What the sweep and the training draws change: the sun, the mirrors'
    slope error and every heliostat's aim offset on the receiver."""
    sun_az: float = 180.0
    sun_el: float = 45.0
    slope: float = 1.5
    offsets: np.ndarray = None  # (n, 2) metres

    def copy(self):
        return FieldState(self.sun_az, self.sun_el, self.slope, None if self.offsets is None else self.offsets.copy())


class Field:
    """The tower field of :func:`workloads.tower` in one context, re-aimed in
    place through the C API (st_sun_xyz, st_optic, st_element_aim,
    st_element_surface_params) and traced with one IMG x IMG absorbed-flux
    image of the receiver per heliostat (st_amd_flux_by_source)."""

    def __init__(self, lib: stapi.StLib, n_helio: int = 1000, bits: int = 32, r_max: float = 1000.0,
                 by_source: bool = True):
        self.lib, self.dll = lib, lib.dll
        self.bits, self.by_source = bits, by_source
        self.scene = workloads.tower(n_helio=n_helio, raycount=1000, r_max=r_max)
        els = self.scene.stages[0].elements
        self.n = len(els)
        self.pts = np.array([e.xyz for e in els], dtype=np.float64)
        self.rec, self.ax_u, self.ax_v = workloads.receiver_axes(self.pts[:, :2])
        d = self.dll
        d.st_amd_set_precision.argtypes = [C.c_void_p, C.c_int]
        d.st_amd_flux_request.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_double, C.c_double, C.c_double, C.c_double]
        d.st_amd_flux_result.argtypes = [C.c_void_p, C.POINTER(C.c_uint), C.POINTER(C.c_double)]
        d.st_amd_flux_by_source.argtypes = [C.c_void_p, C.c_int]
        d.st_amd_flux_source_result_device.argtypes = [C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p]
        d.st_amd_flux_source_result_device.restype = C.c_longlong
        self.ctx = lib.st_create_context()
        if not self.ctx:
            raise RuntimeError("st_create_context failed")
        self.scene.apply(lib, self.ctx)
        d.st_amd_set_precision(self.ctx, bits)
        if d.st_amd_flux_request(self.ctx, 1, 0, IMG, IMG, 1, -REC_HALF, REC_HALF, -REC_HALF, REC_HALF) != 1:
            raise RuntimeError("flux request failed")
        if by_source and d.st_amd_flux_by_source(self.ctx, 1) != 1:
            raise RuntimeError("st_amd_flux_by_source failed")
        self.state = FieldState(offsets=np.zeros((self.n, 2)))
        self._src = None
        self._map = (C.c_uint * (IMG * IMG))()
        self.bin_area = (2 * REC_HALF / IMG) ** 2

    def close(self):
        if self.ctx:
            self.lib.st_free_context(self.ctx)
            self.ctx = None

    def _set_optic(self, slope: float):
        o = self.scene.optics[0]
        for fb, face in ((1, o.front), (2, o.back)):
            g = (C.c_double * 4)(*face.grating)
            z = (C.c_double * 1)()
            rc = self.lib.st_optic(self.ctx, 0, fb, face.dist.encode(), face.optnum, face.apgr, face.order,
                                   face.rreal, face.rimag, face.refl, face.trans, g, float(slope), face.spec,
                                   0, 0, z, z, 0, 0, z, z)
            if rc < 0:
                raise RuntimeError("st_optic failed")

    def _aim(self, i: int, du: float, dv: float, sv: np.ndarray):
        p = self.pts[i]
        target = self.rec + du * self.ax_u + dv * self.ax_v
        nrm, slant = workloads.aim_heliostat(p, target, sv)
        a = p + 100.0 * nrm
        if self.lib.st_element_aim(self.ctx, 0, i, float(a[0]), float(a[1]), float(a[2])) < 0:
            raise RuntimeError("st_element_aim failed")
        sp = (C.c_double * 8)(1.0 / (2.0 * slant), 0, 0, 0, 0, 0, 0, 0)
        if self.lib.st_element_surface_params(self.ctx, 0, i, sp) < 0:
            raise RuntimeError("st_element_surface_params failed")

    def set_state(self, st: FieldState):
        """Re-aim every heliostat (and move the sun, change the slope error)
        for state st; only what differs from the current state is set."""
        sun_moved = (st.sun_az, st.sun_el) != (self.state.sun_az, self.state.sun_el)
        sv = workloads.sun_vector(st.sun_el, st.sun_az)
        if sun_moved:
            s = sv * 1000.0
            self.lib.st_sun_xyz(self.ctx, float(s[0]), float(s[1]), float(s[2]))
        if st.slope != self.state.slope:
            self._set_optic(st.slope)
        off = np.asarray(st.offsets, dtype=np.float64)
        todo = range(self.n) if sun_moved else np.nonzero((off != self.state.offsets).any(axis=1))[0]
        for i in todo:
            self._aim(int(i), off[i, 0], off[i, 1], sv)
        self.state = FieldState(st.sun_az, st.sun_el, st.slope, off.copy())

    def set_offset(self, i: int, du: float, dv: float):
        """Re-aim heliostat i alone."""
        self._aim(i, du, dv, workloads.sun_vector(self.state.sun_el, self.state.sun_az))
        self.state.offsets[i] = (du, dv)

    def trace(self, hits: int, seed: int, images: bool = True):
        """Trace `hits` ray hits.  Returns (images, total, p_ray): the
        per-heliostat receiver images (n, IMG, IMG) and the receiver map
        (IMG, IMG) in W/m^2 as float32 torch tensors on the GPU (images None
        when not asked for or not kept), and the power per ray in W."""
        import torch
        want = bool(images and self.by_source)
        self.dll.st_amd_flux_by_source(self.ctx, 1 if want else 0)
        self.lib.st_sim_params(self.ctx, int(hits), 2 ** 31 - 1, 1)
        n = stapi.run(self.lib, self.ctx, seed)
        if n < 0:
            raise RuntimeError("st_sim_run failed: " + "; ".join(stapi.messages(self.lib, self.ctx)))
        b = [C.c_double() for _ in range(4)]
        ns = C.c_int()
        self.lib.st_sun_stats(self.ctx, *[C.byref(v) for v in b], C.byref(ns))
        p_ray = DNI * (b[1].value - b[0].value) * (b[3].value - b[2].value) / ns.value
        w = p_ray / self.bin_area
        self.dll.st_amd_flux_result(self.ctx, self._map, None)
        total = torch.from_numpy(np.frombuffer(self._map, dtype=np.uint32).astype(np.float32).reshape(IMG, IMG))
        total = total.cuda() * w
        imgs = None
        if want:
            # device to device, on torch's stream: the maps never cross PCIe
            words = self.n * IMG * IMG
            if self._src is None:
                self._src = torch.empty(words, dtype=torch.int32, device="cuda")
            got = self.dll.st_amd_flux_source_result_device(self.ctx, C.c_void_p(self._src.data_ptr()), words,
                                                            _stream())
            if got != words:
                raise RuntimeError("st_amd_flux_source_result_device failed")
            imgs = self._src.view(self.n, IMG, IMG).to(torch.float32) * w
        return imgs, total, p_ray


# ---------------------------------------------------------------- training --

@dataclass
class Ranges:
    """Where the training draws come from: sun azimuth and elevation (deg),
    slope error (mrad) and aim offsets (m, uniform in a square)."""
    sun_az: tuple = (160.0, 200.0)
    sun_el: tuple = (35.0, 55.0)
    slope: tuple = (1.0, 2.0)
    offset: float = 6.0


def make_dataset(field: Field, traces: int, hits: int, seed: int, ranges: Ranges = Ranges(), repeats: int = 1,
                 log=None):
    """`traces` random states of the field (sun, slope error and
    per-heliostat aim offsets drawn from `ranges`), each traced `repeats`
    times at `hits` ray hits on different streams and averaged: field.n
    training pairs per state, with about hits * repeats / field.n rays per
    image.  Returns (X, T) on the GPU: inputs float32 (traces * n, 8) and
    targets bfloat16 (traces * n, IMG * IMG) = images in units of
    FLUX_SCALE."""
    import torch
    rng = np.random.default_rng(seed)
    n = field.n
    X = torch.empty((traces * n, N_IN), dtype=torch.float32, device="cuda")
    T = torch.empty((traces * n, IMG * IMG), dtype=torch.bfloat16, device="cuda")
    t0 = time.perf_counter()
    for k in range(traces):
        st = FieldState(float(rng.uniform(*ranges.sun_az)), float(rng.uniform(*ranges.sun_el)),
                        float(rng.uniform(*ranges.slope)), rng.uniform(-ranges.offset, ranges.offset, (n, 2)))
        field.set_state(st)
        acc = None
        for r in range(repeats):
            imgs, _, _ = field.trace(hits, (seed * 1000 + k) * 64 + r)
            acc = imgs if acc is None else acc.add_(imgs)
        X[k * n:(k + 1) * n] = torch.from_numpy(features(field.pts, st.sun_az, st.sun_el, st.offsets, st.slope)).cuda()
        T[k * n:(k + 1) * n] = (acc.view(n, -1) / (FLUX_SCALE * repeats)).to(torch.bfloat16)
        if log and (k % 20 == 0 or k == traces - 1):
            log(f"dataset: trace {k + 1}/{traces}, {time.perf_counter() - t0:.1f} s")
    return X, T


class Surrogate:
    """The network with its input/target scaling: predict() gives receiver
    images in W/m^2 for rows of :func:`features`."""

    def __init__(self, lib: stapi.StLib, width: int = 256, depth: int = 6, max_batch: int = 1024, seed: int = 1):
        self.net = Mlp(lib, N_IN, width, depth, IMG * IMG, max_batch=max_batch, seed=seed)
        self.max_batch = max_batch
        self.rho = None  # validation RMS error at bright bins over their RMS flux (set by fit)

    def fit(self, X, T, steps: int, batch: int = 512, lr: float = 1e-3, lr_end: float = 2e-5, val=None,
            seed: int = 0, log=None) -> dict:
        """Adam on the MSE with a cosine learning-rate decay from lr to
        lr_end, on random batches of the rows (X, T); val = (Xv, Tv), held-out
        states traced with more rays, gives the validation errors and the
        relative error rho the sweep's error estimate uses."""
        import torch
        n_tr = X.shape[0]
        n_val = 0 if val is None else val[0].shape[0]
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        t0 = time.perf_counter()
        hist = []
        # batches and the cosine schedule on the device; the steps run as
        # replays of one captured step, in chunks between loss readings
        sched = torch.tensor([lr_end + 0.5 * (lr - lr_end) * (1 + math.cos(math.pi * s / max(1, steps - 1)))
                              for s in range(steps)], dtype=torch.float32, device="cuda")
        chunk = 2000
        for s0 in range(0, steps, chunk):
            k = min(chunk, steps - s0)
            idx = torch.randint(0, n_tr, (k, batch), device="cuda", generator=g, dtype=torch.int32)
            loss = self.net.fit(X, T, idx, sched[s0:s0 + k].contiguous(), want_loss=log is not None)
            if log is not None:
                hist.append((s0 + k - 1, loss))
                log(f"train: step {s0 + k - 1}, loss {loss:.4e}, {time.perf_counter() - t0:.1f} s")
        torch.cuda.synchronize()
        train_s = time.perf_counter() - t0
        ev = self.evaluate(*(val if val is not None else (X[-2048:], T[-2048:])))
        self.rho = ev["rel_rms_bright"]
        return dict(ev, train_s=train_s, steps=steps, batch=batch, rows=n_tr, val_rows=n_val, loss_hist=hist)

    def predict(self, F):
        """Images (rows, IMG * IMG) in W/m^2 (float32, GPU, clamped at 0) for
        feature rows F (numpy or GPU tensor)."""
        import torch
        if isinstance(F, np.ndarray):
            F = torch.from_numpy(np.ascontiguousarray(F, dtype=np.float32)).cuda()
        outs = []
        for i in range(0, F.shape[0], self.max_batch):
            outs.append(self.net.forward(F[i:i + self.max_batch].contiguous()))
        y = outs[0] if len(outs) == 1 else torch.cat(outs)
        return y.clamp_(min=0.0).mul_(FLUX_SCALE)

    def evaluate(self, X, T) -> dict:
        """Errors on rows (X, T): the RMS over bins brighter than a tenth of
        the row's peak (absolute, and over the RMS flux of those bins), and
        the relative error of each image's peak and total power (medians and
        95th percentiles)."""
        import torch
        y = self.predict(X)
        t = T.to(torch.float32) * FLUX_SCALE
        pk = t.max(dim=1).values.clamp(min=1e-6)
        bright = t > 0.1 * pk[:, None]
        err = (y - t)
        rms_b = float(torch.sqrt((err[bright] ** 2).mean()).item())
        rel_b = rms_b / float(torch.sqrt((t[bright] ** 2).mean()).item())
        pk_rel = ((y.max(dim=1).values - pk).abs() / pk)
        tot = t.sum(dim=1).clamp(min=1e-6)
        tot_rel = ((y.sum(dim=1) - tot).abs() / tot)
        q = lambda v, p: float(torch.quantile(v.float(), p).item())
        return {"rms_bright_wm2": rms_b, "rel_rms_bright": rel_b, "peak_rel_median": q(pk_rel, 0.5), "peak_rel_p95": q(pk_rel, 0.95),
                "power_rel_median": q(tot_rel, 0.5), "power_rel_p95": q(tot_rel, 0.95),
                "mean_image_peak_wm2": float(pk.mean().item())}


# ------------------------------------------------------------------- sweep --

_DIRS = np.array([(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (1, -1), (-1, 1), (-1, -1)], dtype=np.float64)


@dataclass
class SweepResult:
    offsets: np.ndarray
    peak_start: float          # W/m^2, of the first anchor trace
    peak_pred: float           # the method's own final estimate of the peak
    steps: int
    accepted: int
    wall_s: float
    fallbacks: int = 0         # tracer re-anchors (surrogate sweep)
    judged_errors: list = dfield(default_factory=list)  # |pred - traced| / traced peak at each re-anchor
    traces: int = 0
    est_max: float = 0.0       # largest error estimate the sweep ran with


def _anchor(field: Field, offsets, hits, seed):
    st = field.state.copy()
    st.offsets = offsets
    field.set_state(st)
    return field.trace(hits, seed)


def aim_sweep(field: Field, sur: Surrogate, offsets0: np.ndarray, steps: int = 500, step_m: float = 0.5,
              max_off: float = 6.0, anchor_hits: int = 50_000_000, err_budget: float = 0.02,
              spill_tol: float = 0.005, top_k: int = 8, seed: int = 0) -> SweepResult:
    """Greedy peak-flux flattening over aim offsets with the network as the
    model and the tracer as fallback and judge.

    The tracer gives an anchor: the receiver map M0 and each heliostat's
    image at the current offsets.  Each step takes the bin of the predicted
    map's peak, picks one of the top_k heliostats that send most light into
    it, and scores its eight neighbouring offsets (step_m apart, within
    max_off) by one batched forward of the network: the predicted map is
    M0 plus, for every heliostat moved since the anchor, its predicted image
    now minus its predicted image at the anchor.  The best candidate is
    taken if it lowers the predicted peak without losing more than spill_tol
    of the anchor's intercepted power.  Only the moved heliostats' predicted
    images enter the map, so its error at the peak bin is bounded by the
    network's relative validation error rho times the light they put there
    (now and at the anchor): when that estimate passes err_budget of the
    peak, the tracer re-anchors at the current offsets (the fallback), and
    the gap between the prediction and that trace is recorded (the judge).
    The sweep ends with the prediction, not with a trace."""
    import torch
    assert sur.rho is not None, "fit the surrogate first"
    rng = np.random.default_rng(seed)
    off = np.array(offsets0, dtype=np.float64)
    st = field.state
    sun_az, sun_el, slope = st.sun_az, st.sun_el, st.slope
    t0 = time.perf_counter()
    res = SweepResult(off, 0.0, 0.0, steps, 0, 0.0)
    k_anchor = 0

    def anchor():
        nonlocal k_anchor
        imgs, total, _ = _anchor(field, off.copy(), anchor_hits, 10_007 * (seed + 1) + k_anchor)
        k_anchor += 1
        res.traces += 1
        s_anchor = sur.predict(features(field.pts, sun_az, sun_el, off, slope))
        return total.view(-1), imgs.view(field.n, -1), s_anchor

    M0, Tm, S_anchor = anchor()
    S_cur = S_anchor.clone()
    P = M0.clone()
    intercept0 = float(M0.sum().item())
    res.peak_start = float(M0.max().item())
    moved = set()
    for _ in range(steps):
        b = int(torch.argmax(P).item())
        contrib = Tm[:, b] + (S_cur[:, b] - S_anchor[:, b])
        top = torch.topk(contrib, top_k).indices.cpu().numpy()
        h = int(rng.choice(top))
        cand = np.clip(off[h][None, :] + step_m * _DIRS, -max_off, max_off)
        F = features(field.pts[[h] * len(cand)], sun_az, sun_el, cand, slope)
        Sc = sur.predict(F)
        Pc = P[None, :] - S_cur[h][None, :] + Sc
        peaks = Pc.max(dim=1).values
        ok = Pc.sum(dim=1) >= (1.0 - spill_tol) * intercept0
        score = torch.where(ok, peaks, torch.full_like(peaks, float("inf")))
        j = int(torch.argmin(score).item())
        if float(score[j].item()) < float(P.max().item()):
            P = Pc[j].clone()
            S_cur[h] = Sc[j]
            off[h] = cand[j]
            moved.add(h)
            res.accepted += 1
        if not moved:
            continue
        bp = int(torch.argmax(P).item())
        mv = torch.tensor(sorted(moved), device=P.device)
        est = sur.rho * float((S_cur[mv, bp] + S_anchor[mv, bp]).sum().item()) / max(1.0, float(P[bp].item()))
        res.est_max = max(res.est_max, est)
        if est > err_budget:
            pred_peak = float(P.max().item())
            M0, Tm, S_anchor = anchor()
            traced = float(M0.max().item())
            res.judged_errors.append(abs(pred_peak - traced) / traced)
            res.fallbacks += 1
            S_cur = S_anchor.clone()
            P = M0.clone()
            moved.clear()
    torch.cuda.synchronize()
    res.offsets = off
    res.peak_pred = float(P.max().item())
    res.wall_s = time.perf_counter() - t0
    return res


def aim_sweep_traced(field: Field, offsets0: np.ndarray, steps: int, step_m: float = 0.5, max_off: float = 6.0,
                     eval_hits: int = 10_000_000, spill_tol: float = 0.005, top_k: int = 8,
                     seed: int = 0) -> SweepResult:
    """The same greedy sweep without the network: each step traces the
    current offsets (with per-heliostat images, which pick the heliostat as
    above) and then the field with each of the eight candidate offsets, all
    nine traces on the same random streams (common random numbers, so the
    comparison sees the change and not the noise), eval_hits ray hits each;
    the best candidate is taken if its traced peak is lower."""
    import torch
    rng = np.random.default_rng(seed)
    off = np.array(offsets0, dtype=np.float64)
    t0 = time.perf_counter()
    res = SweepResult(off, 0.0, 0.0, steps, 0, 0.0)
    st = field.state.copy()
    st.offsets = off.copy()
    field.set_state(st)
    intercept0 = None
    P = None
    for s in range(steps):
        sd = 20_011 * (seed + 1) + s
        imgs, P, _ = field.trace(eval_hits, sd, images=True)
        res.traces += 1
        P = P.view(-1)
        if intercept0 is None:
            intercept0 = float(P.sum().item())
            res.peak_start = float(P.max().item())
        b = int(torch.argmax(P).item())
        top = torch.topk(imgs.view(field.n, -1)[:, b], top_k).indices.cpu().numpy()
        h = int(rng.choice(top))
        cand = np.clip(off[h][None, :] + step_m * _DIRS, -max_off, max_off)
        best = (float(P.max().item()), None)
        for c in cand:
            field.set_offset(h, float(c[0]), float(c[1]))
            _, total, _ = field.trace(eval_hits, sd, images=False)
            res.traces += 1
            pk = float(total.max().item())
            if pk < best[0] and float(total.sum().item()) >= (1.0 - spill_tol) * intercept0:
                best = (pk, c.copy())
        if best[1] is not None:
            off[h] = best[1]
            res.accepted += 1
        field.set_offset(h, float(off[h][0]), float(off[h][1]))
        res.peak_pred = best[0]
    torch.cuda.synchronize()
    res.offsets = off
    res.wall_s = time.perf_counter() - t0
    return res


def judge(field64: Field, offsets: np.ndarray, hits: int, seed: int):
    """The receiver map (W/m^2, GPU) of a trace at these offsets, for judging
    a sweep's result: the FP64 path at many rays."""
    st = field64.state.copy()
    st.offsets = np.array(offsets, dtype=np.float64)
    field64.set_state(st)
    _, total, _ = field64.trace(hits, seed, images=False)
    return total
