"""ctypes binding of the glw C ABI (include/glw.h) and the device plan.

The library is built in-tree as ``_lib/libglw.so``.  Nothing here computes a
field: a missing library or device raises, there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
import threading
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_size_t, c_uint8, c_void_p
from pathlib import Path

import numpy as np

from .plan import HostPlan

#: the in-tree build; GLW_LIBRARY names another build of the same sources (A/B timing of build options)
LIB_PATH = Path(os.environ.get("GLW_LIBRARY") or Path(__file__).resolve().parent / "_lib" / "libglw.so")

_DP = POINTER(c_double)
_U8P = POINTER(c_uint8)


class GlwError(RuntimeError):
    """Raised when the native library reports an error."""


class CBox(ctypes.Structure):
    _fields_ = [("mx0", c_int32), ("my0", c_int32), ("wx", c_int32), ("wy", c_int32)]


class CBoxTables(ctypes.Structure):
    _fields_ = [(n, _DP) for n in ("qx", "qy", "nu", "inu", "q2", "a")]


class COrder2Desc(ctypes.Structure):
    _fields_ = [("eta_rank", c_int32), ("eta_damp", _DP), ("eta_cosh", _DP), ("eta_sinhr", _DP),
                ("eta_coef", _DP), ("channels", c_int32), ("psi_damp", _DP), ("psi_factor", _DP),
                ("psi_coef", _DP), ("psi_sign", _DP)]


class COrder4Desc(ctypes.Structure):
    _fields_ = [("rank", c_int32), ("p4", CBox), ("g4x", c_int32), ("g4y", c_int32), ("t4", CBoxTables),
                ("mask4", _U8P), ("damp_quartic", _DP), ("damp_third", _DP), ("damp_pair3", _DP),
                ("damp_pair4", _DP), ("pair_cosh", _DP), ("pair_sinhr", _DP), ("pair_coef", _DP),
                ("third_cosh", _DP), ("third_sinhr", _DP), ("third_coef", _DP), ("quartic_cosh", _DP),
                ("quartic_sinhr", _DP), ("quartic_coef", _DP), ("forcing_terms", c_int32)]


class CPlanDesc(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("depth", c_double), ("gravity", c_double),
                ("floor_level", c_double), ("precision", c_int32), ("max_instants", c_int32),
                ("order", c_int32), ("p1", CBox), ("p2", CBox), ("p3", CBox), ("g2x", c_int32),
                ("g2y", c_int32), ("g3x", c_int32), ("g3y", c_int32), ("t1", CBoxTables), ("t2", CBoxTables),
                ("t3", CBoxTables), ("omega", _DP), ("mask2", _U8P), ("mask3", _U8P), ("inner_rank", c_int32),
                ("outer_rank", c_int32), ("damp_outer", _DP), ("damp_inner", _DP), ("inner_cosh", _DP),
                ("inner_sinhr", _DP), ("inner_coef", _DP), ("outer_cosh", _DP), ("outer_sinhr", _DP),
                ("outer_coef", _DP), ("o2", COrder2Desc), ("with_order4", c_int32), ("o4", COrder4Desc)]


#: every symbol include/glw.h declares
EXPORTS = ("glw_plan_create", "glw_plan_destroy", "glw_linear_execute", "glw_order2_execute",
           "glw_order3_execute", "glw_order4_execute", "glw_wit_create", "glw_wit_destroy", "glw_wit_sector44", "glw_wit_sector44_range",
           "glw_wit_sector",
           "glw_adjoint_execute", "glw_plan_autotune", "glw_plan_get_schedule", "glw_plan_set_schedule", "glw_plan_hold",
           "glw_batch_create",
           "glw_batch_launch", "glw_batch_destroy", "glw_plan_profile", "glw_plan_stage_times", "glw_plan_buffer",
           "glw_fft_length_supported", "glw_hbm_probe", "glw_modal_length_supported", "glw_modal_workspace_size",
           "glw_modal_execute", "glw_modal_stage_times", "glw_modal_execute_noalloc", "glw_eta20_workspace_size", "glw_eta20_execute", "glw_last_error", "glw_version",
           "glw_abi_sizeof")

#: GLW_KIND_* of include/glw.h
KINDS = {"linear": 1, "order2": 2, "order3": 3, "order4": 4}


class CBatchItem(ctypes.Structure):
    _fields_ = [("plan", c_void_p), ("kind", c_int32), ("nt", c_int32), ("s0", c_void_p), ("times", c_void_p),
                ("a", c_void_p), ("b", c_void_p)]

_lib = None


def load_library(path: str | os.PathLike | None = None):
    """Load libglw.so (raises if it was not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise GlwError(f"native library {p} is missing: run __graft_entry__.build() (hipcc, gfx950)")
    lib = ctypes.CDLL(str(p))
    lib.glw_plan_create.argtypes = [POINTER(CPlanDesc), POINTER(c_void_p)]
    lib.glw_plan_create.restype = c_int
    lib.glw_plan_destroy.argtypes = [c_void_p]
    lib.glw_plan_destroy.restype = c_int
    for name in ("glw_linear_execute", "glw_order2_execute", "glw_order3_execute"):
        fn = getattr(lib, name)
        fn.argtypes = [c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p]
        fn.restype = c_int
    lib.glw_order4_execute.argtypes = [c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p]
    lib.glw_order4_execute.restype = c_int
    lib.glw_adjoint_execute.argtypes = [c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_void_p]
    lib.glw_adjoint_execute.restype = c_int
    lib.glw_plan_autotune.argtypes = [c_void_p, c_int32, c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_int32,
                                      c_void_p]
    lib.glw_plan_autotune.restype = c_int
    lib.glw_plan_get_schedule.argtypes = [c_void_p, c_char_p, c_size_t]
    lib.glw_plan_get_schedule.restype = c_int
    lib.glw_plan_set_schedule.argtypes = [c_void_p, c_char_p]
    lib.glw_plan_set_schedule.restype = c_int
    lib.glw_plan_hold.argtypes = [c_void_p, c_int32]
    lib.glw_plan_hold.restype = c_int
    lib.glw_batch_create.argtypes = [POINTER(CBatchItem), c_int32, c_int32, POINTER(c_void_p)]
    lib.glw_batch_create.restype = c_int
    lib.glw_batch_launch.argtypes = [c_void_p, c_void_p]
    lib.glw_batch_launch.restype = c_int
    lib.glw_batch_destroy.argtypes = [c_void_p]
    lib.glw_batch_destroy.restype = c_int
    lib.glw_plan_profile.argtypes = [c_void_p, c_int32]
    lib.glw_plan_profile.restype = c_int
    lib.glw_plan_stage_times.argtypes = [c_void_p, c_int32, c_char_p, c_size_t, _DP, POINTER(c_int32), _DP,
                                         POINTER(c_int32)]
    lib.glw_plan_stage_times.restype = c_int
    lib.glw_plan_buffer.argtypes = [c_void_p, c_char_p, POINTER(c_void_p), POINTER(c_size_t)]
    lib.glw_plan_buffer.restype = c_int
    lib.glw_fft_length_supported.argtypes = [c_int32]
    lib.glw_fft_length_supported.restype = c_int
    lib.glw_hbm_probe.argtypes = [c_int32, c_void_p, c_void_p, c_size_t, c_int32, _DP]
    lib.glw_hbm_probe.restype = c_int
    lib.glw_last_error.restype = c_char_p
    lib.glw_version.restype = c_char_p
    if path is None:
        _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        msg = load_library().glw_last_error().decode()
        raise GlwError(f"glw error {rc}: {msg}")


def _dp(a: np.ndarray | None, keep: list):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    keep.append(a)
    return a.ctypes.data_as(_DP)


def _u8p(a: np.ndarray | None, keep: list):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.uint8)
    keep.append(a)
    return a.ctypes.data_as(_U8P)


def _cbox(b) -> CBox:
    if b is None:
        return CBox(0, 0, 0, 0)
    return CBox(b.mx0, b.my0, b.wx, b.wy)


def _ctab(t, keep) -> CBoxTables:
    if t is None:
        return CBoxTables()
    return CBoxTables(*(_dp(getattr(t, n), keep) for n in ("qx", "qy", "nu", "inu", "q2", "a")))


def describe(plan: HostPlan, max_instants: int) -> tuple[CPlanDesc, list]:
    """Fill the C descriptor from a host plan (arrays kept alive in the list)."""
    keep: list = []
    d = CPlanDesc()
    d.nx, d.ny = plan.nx, plan.ny
    d.depth, d.gravity, d.floor_level = plan.depth, plan.gravity, plan.floor_level
    d.precision, d.max_instants, d.order = plan.precision, max_instants, plan.order
    d.p1, d.p2, d.p3 = _cbox(plan.box1), _cbox(plan.box2), _cbox(plan.box3)
    d.g2x, d.g2y = plan.g2 if plan.g2 else (0, 0)
    d.g3x, d.g3y = plan.g3 if plan.g3 else (0, 0)
    d.t1, d.t2, d.t3 = _ctab(plan.t1, keep), _ctab(plan.t2, keep), _ctab(plan.t3, keep)
    d.omega = _dp(plan.omega, keep)
    d.mask2 = _u8p(plan.mask2, keep)
    d.mask3 = _u8p(plan.mask3, keep)
    if plan.order >= 3:
        o3 = plan.o3
        d.inner_rank, d.outer_rank = o3["inner_rank"], o3["outer_rank"]
        for n in ("damp_outer", "damp_inner", "inner_cosh", "inner_sinhr", "inner_coef", "outer_cosh",
                  "outer_sinhr", "outer_coef"):
            setattr(d, n, _dp(o3[n], keep))
    if plan.order >= 2:
        o2 = plan.o2
        d.o2 = COrder2Desc(o2["eta_rank"], _dp(o2["eta_damp"], keep), _dp(o2["eta_cosh"], keep),
                           _dp(o2["eta_sinhr"], keep), _dp(o2["eta_coef"], keep), o2["channels"],
                           _dp(o2["psi_damp"], keep), _dp(o2["psi_factor"], keep), _dp(o2["psi_coef"], keep),
                           _dp(o2["psi_sign"], keep))
    if plan.o4:
        o = plan.o4
        d.with_order4 = 1
        d.o4 = COrder4Desc(o["rank"], _cbox(plan.box4), plan.g4[0], plan.g4[1], _ctab(plan.t4, keep),
                           _u8p(plan.mask4, keep), *(_dp(o[k], keep) for k in (
                               "damp_quartic", "damp_third", "damp_pair3", "damp_pair4", "pair_cosh", "pair_sinhr",
                               "pair_coef", "third_cosh", "third_sinhr", "third_coef", "quartic_cosh",
                               "quartic_sinhr", "quartic_coef")), int(o.get("forcing_terms", 0)))
    return d, keep


# Device -> host copies of large outputs.  torch copies a device tensor into
# pageable host memory at about 4 GB/s on MI355X (a 268 MB [64][512][512]
# complex128 field in 65 ms), against about 50 GB/s into pinned memory.  The
# outputs are staged through two reusable pinned chunks instead: the copy of
# chunk k + 1 runs while chunk k is copied into the caller's NumPy array
# (by torch's multi-threaded host copy: np.copyto runs on one core at ~16 GB/s).
_STAGE_BYTES = 32 << 20
_stage = {}
_stage_lock = threading.Lock()  # the chunks are shared by every caller in the process


def to_host(x) -> np.ndarray:
    """A contiguous NumPy copy of device tensor ``x`` (same dtype and shape)."""
    import torch

    if isinstance(x, np.ndarray):
        return x
    if x.device.type != "cuda" or x.numel() * x.element_size() <= _STAGE_BYTES:
        return x.cpu().numpy()
    with _stage_lock:
        return _staged_copy(torch, x.contiguous())


def _staged_copy(torch, x) -> np.ndarray:
    flat = x.view(-1)
    out = np.empty(tuple(x.shape), dtype=torch.empty((), dtype=x.dtype).numpy().dtype)
    dst = torch.from_numpy(out.reshape(-1))  # host copy-out by torch: multi-threaded, unlike np.copyto
    key = (x.device.index, x.dtype)
    if key not in _stage:
        n = _STAGE_BYTES // x.element_size()
        _stage[key] = ([torch.empty(n, dtype=x.dtype, pin_memory=True) for _ in range(2)],
                       [torch.cuda.Event() for _ in range(2)])
    bufs, events = _stage[key]
    chunk = bufs[0].numel()
    stream = torch.cuda.current_stream(x.device)
    pending = []  # (slot, offset, count) copied and not yet drained
    for i, off in enumerate(range(0, flat.numel(), chunk)):
        cnt = min(chunk, flat.numel() - off)
        slot = i % 2
        if len(pending) == 2:  # the slot about to be refilled holds the oldest chunk: drain it first
            s0, o0, c0 = pending.pop(0)
            events[s0].synchronize()
            dst[o0:o0 + c0].copy_(bufs[s0][:c0])
        bufs[slot][:cnt].copy_(flat[off:off + cnt], non_blocking=True)
        events[slot].record(stream)
        pending.append((slot, off, cnt))
    for s0, o0, c0 in pending:
        events[s0].synchronize()
        dst[o0:o0 + c0].copy_(bufs[s0][:c0])
    return out


class DevicePlan:
    """A glw plan on the current HIP device plus the t=0 parent spectrum.

    ``run(kind, times)`` evaluates ``kind`` in {"linear", "order2", "order3"}
    at the given instants and returns two complex128 torch tensors
    ``[nt, ny, nx]`` on the device.
    """

    def __init__(self, host: HostPlan, s0_box: np.ndarray, *, max_instants: int = 16):
        import torch

        if not torch.cuda.is_available():
            raise GlwError("a HIP device is required (torch.cuda.is_available() is False)")
        self.lib = load_library()
        self.host = host
        self.max_instants = int(max_instants)
        desc, keep = describe(host, self.max_instants)
        handle = c_void_p()
        _check(self.lib.glw_plan_create(ctypes.byref(desc), ctypes.byref(handle)))
        del keep
        self._handle = handle
        self.device = torch.device("cuda", torch.cuda.current_device())
        s0 = np.ascontiguousarray(s0_box, dtype=np.complex128)
        if s0.shape != (host.box1.wy, host.box1.wx):
            raise ValueError("parent spectrum must be given on the plan's parent box")
        self.s0 = torch.from_numpy(s0).to(self.device)

    def close(self) -> None:
        """Destroy the plan (refused while a captured graph holds it: close that first)."""
        if getattr(self, "_handle", None) is not None and self._handle.value:
            _check(self.lib.glw_plan_destroy(self._handle))
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def hold(self, on: bool = True) -> None:
        """Hold (or release) the plan for captured graphs that replay it (``glw_plan_hold``)."""
        _check(self.lib.glw_plan_hold(self._handle, int(bool(on))))

    def times_tensor(self, times):
        import torch

        if isinstance(times, torch.Tensor):
            return times.to(device=self.device, dtype=torch.float64).contiguous()
        return torch.as_tensor(np.atleast_1d(np.asarray(times, dtype=float)), dtype=torch.float64,
                               device=self.device)

    def run(self, kind: str, times, out=None):
        import torch

        t = self.times_tensor(times)
        nt = int(t.numel())
        shape = (nt, self.host.ny, self.host.nx)
        if out is None:
            out = (torch.empty(shape, dtype=torch.complex128, device=self.device),
                   torch.empty(shape, dtype=torch.complex128, device=self.device))
        a, b = out
        for x in (a, b):
            if x.shape != shape or x.dtype != torch.complex128 or not x.is_contiguous() or x.device != self.device:
                raise ValueError(f"outputs must be contiguous complex128 {shape} on {self.device}")
        stream = torch.cuda.current_stream(self.device).cuda_stream
        if kind == "order4":
            _check(self.lib.glw_order4_execute(self._handle, c_void_p(self.s0.data_ptr()), c_void_p(t.data_ptr()), nt,
                                               c_void_p(a.data_ptr()), c_void_p(stream)))
            return a, None
        fn = {"linear": self.lib.glw_linear_execute, "order2": self.lib.glw_order2_execute,
              "order3": self.lib.glw_order3_execute}[kind]
        _check(fn(self._handle, c_void_p(self.s0.data_ptr()), c_void_p(t.data_ptr()), nt,
                  c_void_p(a.data_ptr()), c_void_p(b.data_ptr()), c_void_p(stream)))
        return a, b

    def _outputs(self, nt: int, out):
        import torch

        shape = (nt, self.host.ny, self.host.nx)
        if out is None:
            out = (torch.empty(shape, dtype=torch.complex128, device=self.device),
                   torch.empty(shape, dtype=torch.complex128, device=self.device))
        for x in out:
            if x is not None and (x.shape != shape or x.dtype != torch.complex128 or not x.is_contiguous()
                                  or x.device != self.device):
                raise ValueError(f"outputs must be contiguous complex128 {shape} on {self.device}")
        return out

    def autotune(self, kind: str, times, out=None, reps: int = 3) -> str:
        """Pick the fastest FFT schedule per stage for ``kind`` at these instants; returns it as text."""
        import torch

        t = self.times_tensor(times)
        a, b = self._outputs(int(t.numel()), out)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _check(self.lib.glw_plan_autotune(self._handle, KINDS[kind], c_void_p(self.s0.data_ptr()),
                                          c_void_p(t.data_ptr()), int(t.numel()), c_void_p(a.data_ptr()),
                                          c_void_p(b.data_ptr() if b is not None else None), int(reps),
                                          c_void_p(stream)))
        return self.schedule

    def adjoint(self, grad_field, times=(0.0,), orders: int = 7, recompute: bool = True, out=None):
        """dL/dRe S0 + i dL/dIm S0 on the parent box for a field gradient dL/dRe eta+ (+ i dL/dIm eta+).

        ``grad_field``: complex128 (or real float64) [ny, nx] device tensor; ``orders``: bit mask
        1 eta11, 2 eta22, 4 eta33 of the components that make up eta+.
        """
        import torch

        t = self.times_tensor(times)
        if int(t.numel()) != 1:
            raise ValueError("the adjoint evaluates one instant")
        g = grad_field
        if not torch.is_complex(g):
            g = g.to(torch.float64).to(torch.complex128)
        g = g.to(device=self.device, dtype=torch.complex128).contiguous()
        if g.shape != (self.host.ny, self.host.nx):
            raise ValueError(f"grad_field must be [{self.host.ny}, {self.host.nx}]")
        shape = (self.host.box1.wy, self.host.box1.wx)
        if out is None:
            out = torch.empty(shape, dtype=torch.complex128, device=self.device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _check(self.lib.glw_adjoint_execute(self._handle, c_void_p(self.s0.data_ptr()), c_void_p(t.data_ptr()),
                                            int(orders), c_void_p(g.data_ptr()), c_void_p(out.data_ptr()),
                                            int(bool(recompute)), c_void_p(stream)))
        return out

    @property
    def schedule(self) -> str:
        buf = ctypes.create_string_buffer(8192)
        _check(self.lib.glw_plan_get_schedule(self._handle, buf, 8192))
        return buf.value.decode()

    @schedule.setter
    def schedule(self, text: str) -> None:
        _check(self.lib.glw_plan_set_schedule(self._handle, text.encode()))

    def profile(self, enable: bool = True) -> None:
        """Bracket every kernel launch of later runs with HIP events on the run's stream."""
        _check(self.lib.glw_plan_profile(self._handle, int(enable)))

    def stage_times(self) -> dict[str, tuple[float, int, float]]:
        """{stage: (total ms, launches, compulsory HBM bytes)} since the last call (waits for the last event)."""
        n = 64
        names = ctypes.create_string_buffer(4096)
        ms = (c_double * n)()
        launches = (c_int32 * n)()
        moved = (c_double * n)()
        count = c_int32()
        _check(self.lib.glw_plan_stage_times(self._handle, n, names, 4096, ms, launches, moved, ctypes.byref(count)))
        keys = names.value.decode().split(",") if count.value else []
        return {k: (ms[i], launches[i], moved[i]) for i, k in enumerate(keys)}

    def buffer(self, name: str):
        ptr, size = c_void_p(), c_size_t()
        _check(self.lib.glw_plan_buffer(self._handle, name.encode(), ctypes.byref(ptr), ctypes.byref(size)))
        return ptr.value, size.value


class GraphBatch:
    """Execute calls of many plans captured in one HIP graph (``glw_batch_*``).

    ``items``: sequence of (DevicePlan, kind, times tensor, a, b) with device
    tensors that stay alive as long as the batch; every ``launch`` re-reads
    the inputs and rewrites the outputs.
    """

    def __init__(self, items, *, branches: int = 4):
        self.lib = load_library()
        self._keep = list(items)
        arr = (CBatchItem * len(self._keep))()
        for i, (plan, kind, times, a, b) in enumerate(self._keep):
            arr[i] = CBatchItem(plan._handle, KINDS[kind], int(times.numel()), plan.s0.data_ptr(), times.data_ptr(),
                                a.data_ptr(), b.data_ptr() if b is not None else None)
        handle = c_void_p()
        _check(self.lib.glw_batch_create(arr, len(self._keep), int(branches), ctypes.byref(handle)))
        self._handle = handle

    def launch(self, stream=None) -> None:
        import torch

        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _check(self.lib.glw_batch_launch(self._handle, c_void_p(st)))

    def close(self) -> None:
        if getattr(self, "_handle", None) is not None and self._handle.value:
            self.lib.glw_batch_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
