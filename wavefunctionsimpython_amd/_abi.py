"""ctypes binding of ``libwfsim.so`` (the C ABI declared in ``include/wfsim.h``).

The library is the only compute path of this package.  Loading fails loudly
when it has not been built (``python -c "import __graft_entry__ as g; g.build()"``
or ``make -C wavefunctionsimpython_amd/csrc``); there is no CPU fallback.

Every launch helper checks on the host that tensor devices, dtypes, shapes and
contiguity match what the kernels and their grids assume before any pointer
crosses the boundary.
"""

from __future__ import annotations

import ctypes
import math
import os
import threading

import numpy as np
import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libwfsim.so")

WFS_C64 = 0
WFS_C128 = 1
WFS_METHOD_FOURIER = 0
WFS_METHOD_POWER_SERIES = 1
SUPPORTED_LENGTHS = (8, 16, 32, 64, 96, 128, 192, 256, 512)

# Every function declared in include/wfsim.h, with its ctypes signature.
_c_int3 = ctypes.c_int * 3
_c_dbl3 = ctypes.c_double * 3
_vp = ctypes.c_void_p
# wfs_between_steps: int (*)(void* ctx, int step)
BETWEEN_STEPS = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int)


class WfsBox(ctypes.Structure):
    _fields_ = [("bottom", ctypes.c_int * 3), ("count", ctypes.c_int * 3)]


SIGNATURES = {
    "wfs_version": (ctypes.c_char_p, []),
    "wfs_last_error": (ctypes.c_char_p, []),
    "wfs_plan_create": (ctypes.c_int, [_c_int3, _c_dbl3, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "wfs_plan_create_batched": (ctypes.c_int, [_c_int3, _c_dbl3, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "wfs_plan_destroy": (ctypes.c_int, [_vp]),
    "wfs_gemm_bf16": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_long * 4, ctypes.c_long * 9, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_long, _vp, _vp, _vp]),
    "wfs_fno_bias_act": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_long, ctypes.c_long, ctypes.c_int, _vp]),
    "wfs_fno_gelu_backward": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_long, _vp]),
    "wfs_fno_axpby": (ctypes.c_int, [_vp, _vp, ctypes.c_float, ctypes.c_float, ctypes.c_long, _vp]),
    "wfs_fno_layer": (ctypes.c_int, [_vp] * 9 + [ctypes.c_int] * 5 + [_vp]),
    "wfs_fno_modes": (ctypes.c_int, [_vp] * 4 + [ctypes.c_int] * 2 + [_vp]),
    "wfs_fno_pack": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_float, ctypes.c_float, _vp, ctypes.c_long,
                                    _vp]),
    "wfs_fno_unpack": (ctypes.c_int, [_vp, ctypes.c_float, _vp, ctypes.c_long, _vp]),
    "wfs_fno_mse": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, ctypes.c_long, _vp, ctypes.c_long, _vp, _vp]),
    "wfs_fno_adam": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_long, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_float, ctypes.c_int, _vp]),
    "wfs_parareal_update": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_long, _vp, ctypes.c_long, _vp, _vp]),
    "wfs_narrow_c128": (ctypes.c_int, [_vp, _vp, ctypes.c_long, _vp]),
    "wfs_spectral_resample": (ctypes.c_int, [_vp, _c_int3, _vp, _c_int3, ctypes.c_long, ctypes.c_double, _vp, _vp]),
    "wfs_graph_begin": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "wfs_graph_end": (ctypes.c_int, [_vp, _vp]),
    "wfs_graph_launch": (ctypes.c_int, [_vp, _vp]),
    "wfs_graph_node_count": (ctypes.c_long, [_vp]),
    "wfs_graph_destroy": (ctypes.c_int, [_vp]),
    "wfs_step": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp]),
    "wfs_fft3d": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp]),
    "wfs_observables": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(WfsBox), ctypes.c_int, _vp, _vp]),
    "wfs_observables_noalloc": (ctypes.c_int, [_c_int3, _c_dbl3, ctypes.c_int, ctypes.c_int, _vp,
                                               ctypes.POINTER(WfsBox), ctypes.c_int, _vp, _vp]),
    "wfs_step_observe": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(WfsBox),
                                        ctypes.c_int, _vp, _vp]),
    "wfs_step_observe_between": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(WfsBox),
                                                ctypes.c_int, _vp, BETWEEN_STEPS, _vp, _vp]),
    "wfs_plane_current": (ctypes.c_int, [_vp, _vp, _c_dbl3, _c_dbl3, ctypes.c_double * 2, ctypes.c_int * 2, _vp,
                                         _vp, _vp]),
    "wfs_init_gaussian": (ctypes.c_int, [_c_int3, _c_dbl3, ctypes.c_int, _c_dbl3, _c_dbl3, _c_dbl3, _vp, _vp]),
    "wfs_add_cap": (ctypes.c_int, [_c_int3, _c_dbl3, _c_dbl3, _c_dbl3, ctypes.c_double, _c_dbl3, _c_dbl3,
                                   _c_dbl3, _c_dbl3, _vp, _vp]),
    "wfs_add_wall": (ctypes.c_int, [_c_int3, _c_dbl3, ctypes.c_double, _c_dbl3, _c_dbl3, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, _vp, _vp]),
    "wfs_add_cap_planes": (ctypes.c_int, [_c_int3, _c_dbl3, _c_dbl3, _c_dbl3, ctypes.c_double, _c_dbl3, _c_dbl3,
                                          _c_dbl3, _c_dbl3, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "wfs_add_wall_planes": (ctypes.c_int, [_c_int3, _c_dbl3, ctypes.c_double, _c_dbl3, _c_dbl3, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "wfs_add_harmonic": (ctypes.c_int, [_c_int3, _c_dbl3, ctypes.c_double, ctypes.c_double, _c_dbl3, _vp, _vp]),
    "wfs_add_lattice": (ctypes.c_int, [_c_int3, _c_dbl3, ctypes.c_double, ctypes.c_double, _vp, _vp]),
    "wfs_kinetic_operator": (ctypes.c_int, [_c_int3, _c_dbl3, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                            _vp, _vp]),
    "wfs_potential_operator": (ctypes.c_int, [ctypes.c_long, _vp, ctypes.c_double, ctypes.c_int, _vp, _vp]),
    "wfs_plan_set_profiling": (ctypes.c_int, [_vp, ctypes.c_int]),
    "wfs_plan_set_step_index": (ctypes.c_int, [_vp, ctypes.c_long]),
    "wfs_plan_get_step_index": (ctypes.c_long, [_vp]),
    "wfs_set_potential": (ctypes.c_int, [_vp, _vp, _vp]),
    "wfs_set_potential_planes": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    "wfs_set_potential_axes": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "wfs_plan_autotune": (ctypes.c_int, [_vp, _vp]),
    "wfs_plan_set_variant": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "wfs_plan_get_variant": (ctypes.c_int, [_vp, ctypes.c_int]),
    "wfs_plan_profile": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_long),
                                        ctypes.POINTER(ctypes.c_double)]),
}

# kernel classes of wfs_plan_profile
KERNEL_CLASSES = ("z_fwd", "z_inv", "z_inv_v_fwd", "y_fwd", "y_inv", "x_fwd_k_inv", "x_fwd", "x_inv",
                  "horner_stage", "observables", "yz_plane", "yz_slab", "x_slab", "slab_edge", "slab_observe",
                  "pv_planes")


class WfsError(RuntimeError):
    """A non-zero status returned by the HIP library."""


_lib = None
_lock = threading.Lock()


def lib() -> ctypes.CDLL:
    """Load libwfsim.so once; raise if it is missing (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} is not built: run `make -C wavefunctionsimpython_amd/csrc` "
                    "(the package has no CPU fallback)")
            handle = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().wfs_last_error().decode(errors="replace")
        raise WfsError(f"{what} failed with status {rc}: {msg}")


def i3(v) -> ctypes.Array:
    return _c_int3(*[int(x) for x in v])


def d3(v) -> ctypes.Array:
    return _c_dbl3(*[float(x) for x in v])


def stream_handle(stream=None) -> ctypes.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _require_device(t: torch.Tensor, name: str) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be a device tensor, got {t.device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous (C-order)")


def _same_grid(tshape, shape) -> bool:
    """Tensor shape equals the grid; a single packet may carry a leading 1."""
    tshape = tuple(int(s) for s in tshape)
    shape = tuple(int(s) for s in shape)
    if tshape == shape:
        return True
    if len(shape) == 3 and tshape == (1,) + shape:
        return True
    return len(shape) == 4 and shape[0] == 1 and tshape == shape[1:]


def check_state(t: torch.Tensor, shape, precision: int, name: str = "psi") -> None:
    _require_device(t, name)
    want = torch.complex128 if precision == WFS_C128 else torch.complex64
    if t.dtype != want:
        raise TypeError(f"{name} must be {want}, got {t.dtype}")
    if not _same_grid(t.shape, shape):
        raise ValueError(f"{name} shape {tuple(t.shape)} != grid {tuple(shape)}")


def check_potential(t: torch.Tensor, shape, name: str = "V") -> None:
    _require_device(t, name)
    if t.dtype != torch.complex64:
        raise TypeError(f"{name} must be complex64 (the potential is always complex64), got {t.dtype}")
    if not _same_grid(t.shape, shape):
        raise ValueError(f"{name} shape {tuple(t.shape)} != grid {tuple(shape)}")


def ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


class Plan:
    """Owner of one ``wfs_plan`` (twiddle and kinetic tables, scratch buffers)."""

    def __init__(self, n, dx, dt: float, mass: float, precision: int, method: int, order: int = 10,
                 batch: int = 1):
        n = tuple(int(v) for v in n)
        for v in n:
            if v not in SUPPORTED_LENGTHS:
                raise ValueError(f"grid length {v} not supported; choose from {SUPPORTED_LENGTHS}")
        self.n = n
        self.dx = tuple(float(v) for v in dx)
        self.dt = float(dt)
        self.mass = float(mass)
        self.precision = int(precision)
        self.method = int(method)
        self.order = int(order)
        self.batch = int(batch)
        if self.batch < 1:
            raise ValueError("batch must be >= 1")
        # tensor shape of the (stacked) state: (n0, n1, n2) or (batch, n0, n1, n2)
        self.shape = n if self.batch == 1 else (self.batch,) + n
        self._h = ctypes.c_void_p()
        self._potential_set = False
        self._V_ref = None
        check(lib().wfs_plan_create_batched(i3(n), d3(dx), self.dt, self.mass, self.precision, self.method,
                                            self.order, self.batch, ctypes.byref(self._h)), "wfs_plan_create")

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def step(self, psi: torch.Tensor, V: torch.Tensor | None, nsteps: int = 1, stream=None) -> None:
        """Advance psi by nsteps; V=None reuses the potential given to set_potential()."""
        check_state(psi, self.shape, self.precision)
        if V is None and self.method == WFS_METHOD_POWER_SERIES:
            V = self._V_ref
        if V is not None:
            check_potential(V, self.shape)
        elif not self._potential_set:
            raise ValueError("no potential: pass V or call set_potential() first")
        if nsteps < 0:
            raise ValueError("nsteps must be >= 0")
        check(lib().wfs_step(self._h, ptr(psi), ptr(V) if V is not None else None, int(nsteps),
                             stream_handle(stream)), "wfs_step")
        if V is not None:
            self._potential_set = True
            if self.method == WFS_METHOD_POWER_SERIES:
                self._V_ref = V

    def set_potential(self, V: torch.Tensor, stream=None) -> None:
        check_potential(V, self.shape)
        check(lib().wfs_set_potential(self._h, ptr(V), stream_handle(stream)), "wfs_set_potential")
        self._potential_set = True
        self._V_ref = V if self.method == WFS_METHOD_POWER_SERIES else None

    def set_potential_planes(self, V: torch.Tensor, x_begin: int, x_end: int, stream=None) -> None:
        """Rebuild x planes [x_begin, x_end) of the stored table from V, which must differ from the V
        of the table's last build in those planes only (one packet)."""
        check_potential(V, self.shape)
        check(lib().wfs_set_potential_planes(self._h, ptr(V), int(x_begin), int(x_end), stream_handle(stream)),
              "wfs_set_potential_planes")
        self._potential_set = True
        self._V_ref = V if self.method == WFS_METHOD_POWER_SERIES else None

    def set_potential_axes(self, fx, fy, fz) -> None:
        """Separable real potential V = fx(x) + fy(y) + fz(z): host arrays of shape (n_d,) or
        (batch, n_d), float64 Hartree.  Synchronises with the device (host upload)."""
        arrs = []
        for d, f in enumerate((fx, fy, fz)):
            a = np.ascontiguousarray(np.asarray(f, dtype=np.float64)).reshape(-1)
            if a.size != self.batch * self.n[d]:
                raise ValueError(f"axis {d}: expected {self.batch * self.n[d]} values, got {a.size}")
            arrs.append(a)
        torch.cuda.synchronize()
        check(lib().wfs_set_potential_axes(self._h, *(a.ctypes.data_as(ctypes.c_void_p) for a in arrs)),
              "wfs_set_potential_axes")
        self._potential_set = True
        self._V_ref = None

    def autotune(self, stream=None) -> dict:
        check(lib().wfs_plan_autotune(self._h, stream_handle(stream)), "wfs_plan_autotune")
        return {name: lib().wfs_plan_get_variant(self._h, k) for k, name in enumerate(KERNEL_CLASSES)}

    def variants(self) -> dict:
        """{kernel class: variant} in use."""
        return {name: lib().wfs_plan_get_variant(self._h, k) for k, name in enumerate(KERNEL_CLASSES)}

    def set_variant(self, kernel_class: str, variant: int) -> None:
        check(lib().wfs_plan_set_variant(self._h, KERNEL_CLASSES.index(kernel_class), int(variant)),
              "wfs_plan_set_variant")

    def fft3d(self, data: torch.Tensor, direction: int, stream=None) -> None:
        check_state(data, self.shape, self.precision, "data")
        check(lib().wfs_fft3d(self._h, ptr(data), int(direction), stream_handle(stream)), "wfs_fft3d")

    def _boxes(self, boxes):
        nb = len(boxes)
        arr = (WfsBox * max(nb, 1))()
        for i, (bottom, count) in enumerate(boxes):
            for a in range(3):
                b, c = int(bottom[a]), int(count[a])
                if b < 0 or c < 1 or b + c > self.n[a]:
                    raise ValueError(f"box {i} axis {a}: [{b}, {b + c}) lies outside the grid of {self.n[a]}")
                arr[i].bottom[a] = b
                arr[i].count[a] = c
        return nb, arr

    def step_observe(self, psi: torch.Tensor, V: torch.Tensor | None, nsteps: int, observe_every: int, boxes,
                     out: torch.Tensor | None = None, stream=None, between=None) -> torch.Tensor:
        """nsteps steps, measuring the boxes before steps 0, observe_every, ... (< nsteps):
        device float64 tensor (ceil(nsteps / observe_every), [batch,] nboxes, 7).  Same state
        and observables as alternating observables() and step(observe_every) calls.

        between(step): called after each step's kernels are queued, before the next step's
        (wfs_step_observe_between); it may rewrite V and the table on the current stream, must
        not touch psi or wait on the stream.  An exception in it stops the call and is re-raised."""
        check_state(psi, self.shape, self.precision)
        if V is None and self.method == WFS_METHOD_POWER_SERIES:
            V = self._V_ref
        if V is not None:
            check_potential(V, self.shape)
        elif not self._potential_set:
            raise ValueError("no potential: pass V or call set_potential() first")
        if nsteps < 0 or observe_every < 1:
            raise ValueError("nsteps must be >= 0 and observe_every >= 1")
        nb, arr = self._boxes(boxes)
        nobs = (nsteps + observe_every - 1) // observe_every
        per = (nb, 7) if self.batch == 1 else (self.batch, nb, 7)
        if out is None:
            out = torch.empty((nobs,) + per, dtype=torch.float64, device=psi.device)
        elif (out.dtype != torch.float64 or out.numel() < nobs * 7 * nb * self.batch or not out.is_contiguous()
              or out.device != psi.device):
            raise ValueError("out must be a contiguous float64 device tensor with 7 entries per box and observation")
        if between is None:
            check(lib().wfs_step_observe(self._h, ptr(psi), ptr(V) if V is not None else None, int(nsteps),
                                         int(observe_every), arr, nb, ptr(out) if nb else None,
                                         stream_handle(stream)), "wfs_step_observe")
        else:
            raised = []

            def hook(_ctx, step):
                try:
                    between(int(step))
                    return 0
                except BaseException as e:   # re-raised below, after the C call unwinds
                    raised.append(e)
                    return 1

            fn = BETWEEN_STEPS(hook)
            rc = lib().wfs_step_observe_between(self._h, ptr(psi), ptr(V) if V is not None else None, int(nsteps),
                                                int(observe_every), arr, nb, ptr(out) if nb else None, fn, None,
                                                stream_handle(stream))
            if raised:
                raise raised[0]
            check(rc, "wfs_step_observe_between")
        if V is not None:
            self._potential_set = True
            if self.method == WFS_METHOD_POWER_SERIES:
                self._V_ref = V
        return out

    def observables(self, psi: torch.Tensor, boxes, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        """Moments of |psi|^2 per box: device float64 tensor (nboxes, 7), or (batch, nboxes, 7)."""
        check_state(psi, self.shape, self.precision)
        nb, arr = self._boxes(boxes)
        if out is None:
            shape = (nb, 7) if self.batch == 1 else (self.batch, nb, 7)
            out = torch.empty(shape, dtype=torch.float64, device=psi.device)
        elif (out.dtype != torch.float64 or out.numel() < 7 * nb * self.batch or not out.is_contiguous()
              or out.device != psi.device):
            raise ValueError("out must be a contiguous float64 device tensor with 7 entries per box")
        if nb:
            check(lib().wfs_observables(self._h, ptr(psi), arr, nb, ptr(out), stream_handle(stream)),
                  "wfs_observables")
        return out

    def plane_current(self, psi: torch.Tensor, center, normal, size, count, out: torch.Tensor | None = None,
                      density: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        """Flux of j.n through a sampled plane into out[0] (device float64); optional j.n map."""
        if self.batch != 1:
            raise ValueError("plane currents need a single-grid plan")
        check_state(psi, self.shape, self.precision)
        cu, cv = int(count[0]), int(count[1])
        if cu < 2 or cv < 2:
            raise ValueError("a plane needs at least 2 x 2 samples")
        if out is None:
            out = torch.empty(1, dtype=torch.float64, device=psi.device)
        elif out.dtype != torch.float64 or out.numel() < 1 or out.device != psi.device:
            raise ValueError("out must be a float64 device tensor")
        if density is not None:
            want = torch.float64 if self.precision == WFS_C128 else torch.float32
            _require_device(density, "density")
            if density.dtype != want or tuple(density.shape) != (cu, cv):
                raise ValueError(f"density must be {want} of shape ({cu}, {cv})")
        check(lib().wfs_plane_current(self._h, ptr(psi), d3(center), d3(normal), (ctypes.c_double * 2)(*size),
                                      (ctypes.c_int * 2)(cu, cv), ptr(density) if density is not None else None,
                                      ptr(out), stream_handle(stream)), "wfs_plane_current")
        return out

    def set_step_index(self, index: int) -> None:
        """Table copy of the next step (the plan advances it by every call's steps)."""
        check(lib().wfs_plan_set_step_index(self._h, int(index)), "wfs_plan_set_step_index")

    def step_index(self) -> int:
        return int(lib().wfs_plan_get_step_index(self._h))

    def set_profiling(self, enable: bool) -> None:
        check(lib().wfs_plan_set_profiling(self._h, 1 if enable else 0), "wfs_plan_set_profiling")

    def profile(self, reset: bool = True) -> dict:
        """{kernel class: (launches, total ms)} of the event-timed launches so far."""
        out = {}
        for k, name in enumerate(KERNEL_CLASSES):
            n = ctypes.c_long()
            ms = ctypes.c_double()
            check(lib().wfs_plan_profile(self._h, k, 1 if reset else 0, ctypes.byref(n), ctypes.byref(ms)),
                  "wfs_plan_profile")
            if n.value:
                out[name] = (int(n.value), float(ms.value))
        return out

    def close(self) -> None:
        if self._h and self._h.value:
            lib().wfs_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


GEMM_SPLIT_A, GEMM_SPLIT_B, GEMM_NO_C = 1, 2, 4


def gemm_bf16(A: torch.Tensor, B: torch.Tensor, C: torch.Tensor, batch: int, M: int, N: int, K: int,
              sA, sB, sC, alpha: float = 1.0, beta: float = 0.0, splits: int = 1,
              work: torch.Tensor | None = None, stream=None, flags: int = 0,
              bias: torch.Tensor | None = None, act: torch.Tensor | None = None) -> None:
    """C[b](m,n) = alpha sum_k A[b](m,k) B[b](k,n) + beta C[b](m,n) (+ bias[m]) on the bf16 matrix cores.

    sA = (batch, m, k), sB = (batch, k, n), sC = (batch, m, n) element strides of float32
    device tensors.  Every addressed element is checked to lie inside its tensor.
    flags: GEMM_SPLIT_A / GEMM_SPLIT_B carry that operand as bf16 hi + lo parts; GEMM_NO_C
    (with act) leaves C unwritten.  act receives GELU of the result with C's strides.
    """
    tensors = [(A, "A"), (B, "B"), (C, "C")] + [(t, nm) for t, nm in ((bias, "bias"), (act, "act")) if t is not None]
    for t, name in tensors:
        _require_device_any(t, name)
        if t.dtype != torch.float32:
            raise TypeError(f"{name} must be float32")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous (strides address its flat storage)")

    def span(strides, d0, d1, d2):
        return (d0 - 1) * strides[0] + (d1 - 1) * strides[1] + (d2 - 1) * strides[2]

    if min(batch, M, N, K) < 1:
        raise ValueError("GEMM sizes must be >= 1")
    checks = [(A, sA, (batch, M, K), "A"), (B, sB, (batch, K, N), "B"), (C, sC, (batch, M, N), "C")]
    if act is not None:
        checks.append((act, sC, (batch, M, N), "act"))
    for t, st, dims, name in checks:
        if min(st) < 0 or span(st, *dims) >= t.numel():
            raise ValueError(f"GEMM operand {name} strides {st} reach outside its {t.numel()} elements")
    if bias is not None and bias.numel() < M:
        raise ValueError(f"bias has {bias.numel()} elements, the GEMM has {M} rows")
    if (flags & GEMM_NO_C) and act is None:
        raise ValueError("GEMM_NO_C needs an activation output")
    if splits > 1:
        need = batch * splits * M * N
        if work is None or work.numel() < need or work.dtype != torch.float32:
            raise ValueError(f"split-K needs a float32 workspace of {need} elements")
    dims = (ctypes.c_long * 4)(batch, M, N, K)
    st = (ctypes.c_long * 9)(*(list(sA) + list(sB) + list(sC)))
    check(lib().wfs_gemm_bf16(ptr(A), ptr(B), ptr(C), dims, st, float(alpha), float(beta), int(flags), int(splits),
                              ptr(work) if work is not None else None, int(work.numel()) if work is not None else 0,
                              ptr(bias) if bias is not None else None, ptr(act) if act is not None else None,
                              stream_handle(stream)), "wfs_gemm_bf16")


def fno_layer(W: torch.Tensor, v: torch.Tensor, Z: torch.Tensor | None, bias: torch.Tensor, Gz: torch.Tensor,
              Fz: torch.Tensor, u: torch.Tensor | None, vn: torch.Tensor | None, T: torch.Tensor | None,
              cin: int, C: int, n: int, tm: int, act: bool, stream=None) -> None:
    """One Fourier-neural-operator layer fused over grid lines (``wfs_fno_layer``):
    u = W v + Z Gz + b, vn = GELU(u) (act) or u, T = vn Fz, per line (x, y) of the n^3 grid.

    W [C][cin], v [cin][n^3], Z [C][n^2][tm] (or None), bias [C], Gz [tm][n], Fz [n][tm];
    outputs u, vn [C][n^3] and T [C][n^2][tm], each optional (T needs vn).  float32, contiguous.
    """
    P = n ** 3
    need = [(W, C * cin, "W"), (v, cin * P, "v"), (bias, C, "bias"), (Gz, tm * n, "Gz"), (Fz, n * tm, "Fz")]
    for t, size, name in need + [(Z, C * n * n * tm, "Z"), (u, C * P, "u"), (vn, C * P, "vn"),
                                 (T, C * n * n * tm, "T")]:
        if t is None:
            continue
        _require_device_any(t, name)
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous float32 tensor")
        if t.numel() < size:
            raise ValueError(f"{name} has {t.numel()} elements, needs {size}")
    check(lib().wfs_fno_layer(ptr(W), ptr(v), ptr(Z) if Z is not None else None, ptr(bias), ptr(Gz), ptr(Fz),
                              ptr(u) if u is not None else None, ptr(vn) if vn is not None else None,
                              ptr(T) if T is not None else None, int(cin), int(C), int(n), int(tm),
                              1 if act else 0, stream_handle(stream)), "wfs_fno_layer")


def fno_modes(Rr: torch.Tensor, Ri: torch.Tensor, X: torch.Tensor, O: torch.Tensor, m3: int, C: int,
              stream=None) -> None:
    """Complex per-mode channel mixing O[o][q] = sum_i R[q][i][o] X[i][q] (``wfs_fno_modes``):
    R [m3][C][C] real and imaginary parts, X and O [C][2][m3]; float32, contiguous."""
    for t, size, name in ((Rr, m3 * C * C, "Rr"), (Ri, m3 * C * C, "Ri"), (X, 2 * C * m3, "X"),
                          (O, 2 * C * m3, "O")):
        _require_device_any(t, name)
        if t.dtype != torch.float32 or not t.is_contiguous() or t.numel() < size:
            raise ValueError(f"{name} must be a contiguous float32 tensor of at least {size} elements")
    check(lib().wfs_fno_modes(ptr(Rr), ptr(Ri), ptr(X), ptr(O), int(m3), int(C), stream_handle(stream)),
          "wfs_fno_modes")


def spectral_resample(src: torch.Tensor, dst: torch.Tensor, scale: float | None = None, shift=None,
                      stream=None) -> None:
    """dst spectrum from src spectrum (``wfs_spectral_resample``): complex64 device tensors of shape
    (..., n0, n1, n2) with the same leading (batch) size; dst frequency k takes src frequency
    k + shift (shift: three integers, default none) times scale (default (|dst| / |src|)^(1/2), the
    ortho-transform value-preserving factor), 0 where src does not hold k + shift."""
    for t, name in ((src, "src"), (dst, "dst")):
        _require_device(t, name)
        if t.dtype != torch.complex64 or t.dim() < 3:
            raise ValueError(f"{name} must be a complex64 tensor of at least 3 dimensions")
    ns, nd = tuple(src.shape[-3:]), tuple(dst.shape[-3:])
    batch = src.numel() // (ns[0] * ns[1] * ns[2])
    if dst.numel() // (nd[0] * nd[1] * nd[2]) != batch:
        raise ValueError("src and dst hold different numbers of grids")
    if src.data_ptr() == dst.data_ptr():
        raise ValueError("src and dst must differ")
    if scale is None:
        scale = math.sqrt((nd[0] * nd[1] * nd[2]) / (ns[0] * ns[1] * ns[2]))
    sh = None if shift is None else i3(tuple(int(v) for v in shift))
    check(lib().wfs_spectral_resample(ptr(src), i3(ns), ptr(dst), i3(nd), int(batch), float(scale), sh,
                                      stream_handle(stream)), "wfs_spectral_resample")


def _require_device_any(t: torch.Tensor, name: str) -> None:
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise ValueError(f"{name} must be a device tensor")


class Graph:
    """hipGraph of library calls captured on ``stream`` (``with Graph(stream) as g: ...``)."""

    def __init__(self, stream: torch.cuda.Stream):
        if stream is None or stream.cuda_stream == 0:
            raise ValueError("capture needs a non-default stream")
        self.stream = stream
        self._h = ctypes.c_void_p()

    def __enter__(self):
        check(lib().wfs_graph_begin(ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(self._h)),
              "wfs_graph_begin")
        return self

    def __exit__(self, exc_type, exc, tb):
        rc = lib().wfs_graph_end(self._h, ctypes.c_void_p(self.stream.cuda_stream))
        if exc_type is None:
            check(rc, "wfs_graph_end")
        return False

    def launch(self, stream=None) -> None:
        s = stream if stream is not None else self.stream
        check(lib().wfs_graph_launch(self._h, ctypes.c_void_p(s.cuda_stream)), "wfs_graph_launch")

    def node_count(self) -> int:
        return int(lib().wfs_graph_node_count(self._h))

    def close(self) -> None:
        if self._h and self._h.value:
            lib().wfs_graph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _grid_numel(n) -> int:
    return int(n[0]) * int(n[1]) * int(n[2])


def init_gaussian(n, dx, precision: int, sigma, r0, k0, psi: torch.Tensor, stream=None) -> None:
    check_state(psi, n, precision)
    check(lib().wfs_init_gaussian(i3(n), d3(dx), int(precision), d3(sigma), d3(r0), d3(k0), ptr(psi),
                                  stream_handle(stream)), "wfs_init_gaussian")


def observables_noalloc(n, dx, precision: int, psi: torch.Tensor, boxes, out: torch.Tensor, batch: int = 1,
                        stream=None) -> torch.Tensor:
    """Box moments without a plan and without allocation (``wfs_observables_noalloc``): out is a
    caller-owned contiguous float64 device tensor of batch * len(boxes) * 7 entries."""
    n = tuple(int(v) for v in n)
    shape = n if batch == 1 else (int(batch),) + n
    check_state(psi, shape, precision)
    nb = len(boxes)
    arr = (WfsBox * max(nb, 1))()
    for i, (bottom, count) in enumerate(boxes):
        for a in range(3):
            arr[i].bottom[a] = int(bottom[a])
            arr[i].count[a] = int(count[a])
    if (out.dtype != torch.float64 or not out.is_contiguous() or out.device != psi.device
            or out.numel() < int(batch) * nb * 7):
        raise ValueError("out must be a contiguous float64 device tensor with 7 entries per box and packet")
    check(lib().wfs_observables_noalloc(i3(n), d3(dx), int(precision), int(batch), ptr(psi), arr, nb, ptr(out),
                                        stream_handle(stream)), "wfs_observables_noalloc")
    return out


def add_cap(n, dx, bottom, top, offset, pot_pos, pot_neg, exp_pos, exp_neg, V: torch.Tensor, stream=None,
            planes=None) -> None:
    """planes=(x_begin, x_end): only those x planes of V (wfs_add_cap_planes)."""
    check_potential(V, n)
    if planes is None:
        check(lib().wfs_add_cap(i3(n), d3(dx), d3(bottom), d3(top), float(offset), d3(pot_pos), d3(pot_neg),
                                d3(exp_pos), d3(exp_neg), ptr(V), stream_handle(stream)), "wfs_add_cap")
    else:
        check(lib().wfs_add_cap_planes(i3(n), d3(dx), d3(bottom), d3(top), float(offset), d3(pot_pos), d3(pot_neg),
                                       d3(exp_pos), d3(exp_neg), int(planes[0]), int(planes[1]), ptr(V),
                                       stream_handle(stream)), "wfs_add_cap_planes")


def add_wall(n, dx, potential, center, normal, plateau, slope, slope_exponent, slit_count, slit_spacing,
             slit_width, slit_rotation, V: torch.Tensor, stream=None, planes=None) -> None:
    """planes=(x_begin, x_end): only those x planes of V (wfs_add_wall_planes)."""
    check_potential(V, n)
    if planes is None:
        check(lib().wfs_add_wall(i3(n), d3(dx), float(potential), d3(center), d3(normal), float(plateau),
                                 float(slope), float(slope_exponent), int(slit_count), float(slit_spacing),
                                 float(slit_width), float(slit_rotation), ptr(V), stream_handle(stream)),
              "wfs_add_wall")
    else:
        check(lib().wfs_add_wall_planes(i3(n), d3(dx), float(potential), d3(center), d3(normal), float(plateau),
                                        float(slope), float(slope_exponent), int(slit_count), float(slit_spacing),
                                        float(slit_width), float(slit_rotation), int(planes[0]), int(planes[1]),
                                        ptr(V), stream_handle(stream)), "wfs_add_wall_planes")


def add_harmonic(n, dx, omega, mass, center, V: torch.Tensor, stream=None) -> None:
    check_potential(V, n)
    check(lib().wfs_add_harmonic(i3(n), d3(dx), float(omega), float(mass), d3(center), ptr(V),
                                 stream_handle(stream)), "wfs_add_harmonic")


def separable_axes(n, dx, omega: float = 0.0, mass: float = 1.0, center=(0.0, 0.0, 0.0), depth: float = 0.0,
                   period: float = 1.0):
    """Per-axis parts f_d of V = sum_d f_d(r_d) for the harmonic well 1/2 m omega^2 |r - c|^2 plus
    the lattice depth * sum_d cos^2(pi r_d / period), r_d(i) = dx_d (i - n_d / 2) — the same
    terms, in FP64, that add_harmonic and add_lattice add point by point."""
    out = []
    k = 0.5 * mass * omega * omega
    for d in range(3):
        r = float(dx[d]) * (np.arange(int(n[d]), dtype=np.float64) - int(n[d]) // 2)
        f = k * (r - float(center[d])) ** 2
        if depth:
            f = f + depth * np.cos(np.pi * r / period) ** 2
        out.append(f)
    return tuple(out)


def add_lattice(n, dx, depth, period, V: torch.Tensor, stream=None) -> None:
    check_potential(V, n)
    check(lib().wfs_add_lattice(i3(n), d3(dx), float(depth), float(period), ptr(V), stream_handle(stream)),
          "wfs_add_lattice")


def kinetic_operator(n, dx, dt, mass, precision: int, out: torch.Tensor, stream=None) -> None:
    check_state(out, n, precision, "out")
    check(lib().wfs_kinetic_operator(i3(n), d3(dx), float(dt), float(mass), int(precision), ptr(out),
                                     stream_handle(stream)), "wfs_kinetic_operator")


def potential_operator(V: torch.Tensor, dt, precision: int, out: torch.Tensor, stream=None) -> None:
    _require_device(V, "V")
    if V.dtype != torch.complex64:
        raise TypeError("V must be complex64")
    check_state(out, tuple(V.shape), precision, "out")
    check(lib().wfs_potential_operator(int(V.numel()), ptr(V), float(dt), int(precision), ptr(out),
                                       stream_handle(stream)), "wfs_potential_operator")
