"""ctypes binding of ``libbeamz_hip.so`` (declared in ``include/beamz_hip.h``).

The library is the only compute path: there is no CPU or PyTorch fallback.
When it is missing or its ABI does not match, :func:`library` raises
:class:`HipBackendUnavailable` (the counterpart of the reference's
``CudaBackendUnavailable``, ``beamz/simulation/backend.py:35-36``).
"""

from __future__ import annotations

import ctypes
import os
import threading

import torch

from . import _abi as abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("BEAMZ_HIP_LIBRARY", os.path.join(HERE, "libbeamz_hip.so"))

F32, S32, BF16, F16, F64 = 0, 1, 2, 3, 4
_DTYPE_CODE = {
    torch.float32: F32,
    torch.int32: S32,
    torch.bfloat16: BF16,
    torch.float16: F16,
    torch.float64: F64,
}


class HipBackendUnavailable(RuntimeError):
    """The HIP library could not be loaded or reports another ABI."""


class HipLaunchError(RuntimeError):
    """A native call rejected its arguments or failed to enqueue work."""


class Buffer(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("rank", ctypes.c_int32),
        ("element_type", ctypes.c_int32),
        ("dims", ctypes.c_int64 * 4),
    ]


class PhaseAttrs(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("phase", ctypes.c_int32),
        ("nterms", ctypes.c_int32),
        ("dt", ctypes.c_float),
        ("resolution", ctypes.c_float),
        ("boundary_code", ctypes.c_int32),
        ("metric_kind", ctypes.c_int32),
    ]


class ProgramAttrs(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("nsteps", ctypes.c_int32),
        ("dt", ctypes.c_float),
        ("resolution", ctypes.c_float),
        ("boundary_code", ctypes.c_int32),
        ("metric_kind", ctypes.c_int32),
        ("program_layout", ctypes.c_int32),
        ("cpml_enabled", ctypes.c_int32),
        ("monitor_count", ctypes.c_int32),
        ("coincident_source_group_mask", ctypes.c_int32),
    ]


class Sim2d(ctypes.Structure):
    """``bzh_sim2d``: one TMxy simulation of a batch (two field banks)."""
    _fields_ = [
        ("ez", Buffer * 2), ("hx", Buffer * 2), ("hy", Buffer * 2),
        ("e_decay", Buffer), ("e_source", Buffer), ("h_decay", Buffer), ("h_source", Buffer),
        ("cpml_table", Buffer), ("psi", (Buffer * 4) * 2),
        ("src_coeff", Buffer * 4), ("src_wave", Buffer * 4),
        ("src_start", (ctypes.c_int32 * 2) * 4), ("nsrc", ctypes.c_int32),
        ("cpml_cells", ctypes.c_int32), ("metallic", ctypes.c_int32),
        ("resolution", ctypes.c_float),
    ]


class BatchAttrs(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("nsteps", ctypes.c_int32),
        ("tile_rows", ctypes.c_int32),
        ("workgroups", ctypes.c_int32),
    ]


class Adjoint2d(ctypes.Structure):
    """``bzh_adjoint2d``: the reverse-time TMxy problem and its adjoint state."""
    _fields_ = [
        ("e_decay", Buffer), ("e_source", Buffer), ("h_decay", Buffer), ("h_source", Buffer),
        ("cpml_table", Buffer), ("cpml_cells", ctypes.c_int32), ("metallic", ctypes.c_int32),
        ("resolution", ctypes.c_float),
        ("a_ez", Buffer), ("a_hx", Buffer), ("a_hy", Buffer), ("a_psi", Buffer * 4),
        ("work", Buffer * 4), ("source_profile", Buffer), ("source_time", Buffer),
        ("design", ctypes.c_int32 * 4), ("record", Buffer), ("grad", Buffer),
    ]


class Launch(ctypes.Structure):
    """``bzh_launch``: one Yee phase by descriptor (``cuda/src/launch.h:19-35``)."""
    _fields_ = [
        ("abi_version", ctypes.c_int32), ("flags", ctypes.c_int32), ("phase", ctypes.c_int32),
        ("nterms", ctypes.c_int32), ("metric_kind", ctypes.c_int32), ("dt", ctypes.c_float),
        ("resolution", ctypes.c_float), ("inv_resolution", ctypes.c_float),
        ("dt_over_eps", ctypes.c_float), ("dt_over_mu", ctypes.c_float),
        ("metallic_edges", ctypes.c_int32), ("uniform_cpml_thickness", ctypes.c_int32),
        ("inputs", Buffer * 37), ("metrics", Buffer * 3), ("outputs", Buffer * 9),
    ]


class SourceGroup(ctypes.Structure):
    """``bzh_source_group`` (``cuda/src/launch.h:37-45``)."""
    _fields_ = [
        ("coefficients", Buffer), ("waveforms", Buffer), ("starts", Buffer),
        ("current_step", Buffer), ("component", ctypes.c_int32), ("timing", ctypes.c_int32),
        ("coincident", ctypes.c_int32),
    ]


class DftGroup(ctypes.Structure):
    """``bzh_dft_group`` (``cuda/src/launch.h:47-61``)."""
    _fields_ = [
        ("indices", Buffer), ("weights", Buffer), ("frequencies", Buffer),
        ("component_masks", Buffer), ("counts", Buffer), ("codes", Buffer), ("windows", Buffer),
        ("dft_re", Buffer), ("dft_im", Buffer), ("dft_weight", Buffer), ("time", Buffer),
        ("current_step", Buffer), ("monitor_count", ctypes.c_int32),
    ]


class ProgramLaunch(ctypes.Structure):
    """``bzh_program_launch`` (``cuda/src/launch.h:63-77``)."""
    _fields_ = [
        ("h_ab", Launch), ("e_ab", Launch), ("h_ba", Launch), ("e_ba", Launch),
        ("source_groups", ctypes.POINTER(SourceGroup)), ("source_group_count", ctypes.c_int32),
        ("monitors", ctypes.POINTER(DftGroup)), ("field_bank_count", ctypes.c_int32),
        ("nsteps", ctypes.c_int32),
    ]


# Every exported symbol and its (restype, argtypes); checked by the tests
# against the declarations of include/beamz_hip.h.
SIGNATURES = {
    "bzh_abi_version": (ctypes.c_int32, []),
    "bzh_component_version": (ctypes.c_char_p, []),
    "bzh_targets": (ctypes.c_char_p, []),
    "bzh_last_error": (ctypes.c_char_p, []),
    "bzh_last_program_fused": (ctypes.c_int32, []),
    "bzh_last_program_pairs": (ctypes.c_int32, []),
    "bzh_last_program_captured": (ctypes.c_int32, []),
    "bzh_launch_phase": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(Launch)]),
    "bzh_launch_program": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(ProgramLaunch)]),
    "bzh_streamed": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(Buffer), ctypes.c_int32,
                                      ctypes.POINTER(PhaseAttrs)]),
    "bzh_program": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(Buffer), ctypes.c_int32,
                                     ctypes.POINTER(ProgramAttrs)]),
    "bzh_program_noalloc": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(Buffer),
                                             ctypes.c_int32, ctypes.POINTER(ProgramAttrs)]),
    "bzh_batch2d": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(Sim2d), ctypes.c_int32,
                                     ctypes.POINTER(Buffer), ctypes.POINTER(BatchAttrs)]),
    "bzh_adjoint2d_record": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(Adjoint2d),
                                              ctypes.POINTER(Buffer), ctypes.POINTER(Buffer),
                                              ctypes.c_int32]),
    "bzh_adjoint2d_reverse": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(Adjoint2d),
                                               ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_int32]),
}

_LOCK = threading.Lock()
_LIB = None


def library():
    """Load the library once and check its ABI version."""
    global _LIB
    with _LOCK:
        if _LIB is not None:
            return _LIB
        if not os.path.exists(LIB_PATH):
            raise HipBackendUnavailable(
                f"{LIB_PATH} is missing; build it with `python -m beamz_amd.build`")
        try:
            lib = ctypes.CDLL(LIB_PATH)
        except OSError as exc:
            raise HipBackendUnavailable(f"cannot load {LIB_PATH}: {exc}") from exc
        for name, (restype, argtypes) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        version = int(lib.bzh_abi_version())
        if version != abi.ABI_VERSION:
            raise HipBackendUnavailable(
                f"libbeamz_hip ABI {version} does not match the host ABI {abi.ABI_VERSION}")
        _LIB = lib
        return lib


def status():
    """``(available, reason)`` without raising."""
    try:
        library()
    except HipBackendUnavailable as exc:
        return False, str(exc)
    return True, ""


def buffer(t) -> Buffer:
    """Describe a contiguous tensor (or ``None`` for an empty slot)."""
    b = Buffer()
    if t is None:
        b.data = None
        b.rank = 1
        b.element_type = F32
        b.dims[0] = 0
        return b
    if not t.is_contiguous():
        raise HipLaunchError("native buffers must be contiguous")
    if t.dim() > 4:
        raise HipLaunchError("native buffers have rank <= 4")
    b.data = t.data_ptr() if t.numel() else None
    b.rank = t.dim()
    b.element_type = _DTYPE_CODE[t.dtype]
    for d, n in enumerate(t.shape):
        b.dims[d] = int(n)
    return b


def buffers(tensors):
    arr = (Buffer * len(tensors))()
    for i, t in enumerate(tensors):
        arr[i] = buffer(t)
    return arr


def _check(rc, lib):
    if rc != 0:
        raise HipLaunchError(lib.bzh_last_error().decode())


def current_stream_handle():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def call_streamed(args, attrs: PhaseAttrs, stream=None):
    lib = library()
    arr = buffers(args)
    s = current_stream_handle() if stream is None else stream
    _check(lib.bzh_streamed(s, arr, len(args), ctypes.byref(attrs)), lib)


def last_program_fused() -> bool:
    """Did the last program call on this thread run the fused step kernels?"""
    return bool(library().bzh_last_program_fused())


def last_program_captured() -> bool:
    """Did the last program call's pair launches capture the monitored cells
    densely (rather than through the intermediate bank)?"""
    return bool(library().bzh_last_program_captured())


def last_program_pairs() -> int:
    """Two-step (pair) launches in the last program call on this thread."""
    return int(library().bzh_last_program_pairs())


def call_program(args, attrs: ProgramAttrs, stream=None, noalloc=False):
    """``bzh_program`` (or ``bzh_program_noalloc``, which allocates nothing)."""
    lib = library()
    arr = buffers(args)
    s = current_stream_handle() if stream is None else stream
    fn = lib.bzh_program_noalloc if noalloc else lib.bzh_program
    _check(fn(s, arr, len(args), ctypes.byref(attrs)), lib)


class PreparedProgram:
    """One ``bzh_program`` call whose buffer array is built once: calling the
    object issues the same call again on the current stream.  For loops that
    step one state many times in short chunks, where building the array
    (a few tens of tensors) costs far more host time than the launch.  It
    holds the tensors; they must keep their storage while it is used."""

    def __init__(self, args, attrs: ProgramAttrs):
        self._args = tuple(args)
        self._arr = buffers(self._args)
        self._attrs = attrs

    def __call__(self, stream=None):
        lib = library()
        s = current_stream_handle() if stream is None else stream
        _check(lib.bzh_program(s, self._arr, len(self._args), ctypes.byref(self._attrs)), lib)


def call_batch2d(sims, current_step, attrs: BatchAttrs, stream=None):
    """``sims``: a ctypes array of :class:`Sim2d`; tensors it points to must
    stay alive for the call."""
    lib = library()
    step = buffer(current_step)
    s = current_stream_handle() if stream is None else stream
    _check(lib.bzh_batch2d(s, sims, len(sims), ctypes.byref(step), ctypes.byref(attrs)), lib)


def call_adjoint_record(adj, hx, hy, slot, stream=None):
    lib = library()
    s = current_stream_handle() if stream is None else stream
    bx, by = buffer(hx), buffer(hy)
    _check(lib.bzh_adjoint2d_record(s, ctypes.byref(adj), ctypes.byref(bx), ctypes.byref(by),
                                    int(slot)), lib)


class PreparedAdjointRecord:
    """``bzh_adjoint2d_record`` on fixed Hx and Hy tensors, described once;
    calling the object with a slot records that slot."""

    def __init__(self, adj, hx, hy):
        self._adj = adj
        self._keep = (hx, hy)
        self._bx, self._by = buffer(hx), buffer(hy)

    def __call__(self, slot, stream=None):
        lib = library()
        s = current_stream_handle() if stream is None else stream
        _check(lib.bzh_adjoint2d_record(s, ctypes.byref(self._adj), ctypes.byref(self._bx),
                                        ctypes.byref(self._by), int(slot)), lib)


def call_adjoint_reverse(adj, n_last, count, slot_last, flags, stream=None):
    lib = library()
    s = current_stream_handle() if stream is None else stream
    _check(lib.bzh_adjoint2d_reverse(s, ctypes.byref(adj), int(n_last), int(count),
                                     int(slot_last), int(flags)), lib)


def call_launch_phase(launch: Launch, stream=None):
    """``bzh_launch_phase``: one phase described by a :class:`Launch`."""
    lib = library()
    s = current_stream_handle() if stream is None else stream
    _check(lib.bzh_launch_phase(s, ctypes.byref(launch)), lib)


def call_launch_program(program: ProgramLaunch, stream=None):
    """``bzh_launch_program``: a multi-step program described by a
    :class:`ProgramLaunch` (the tensors it points to must stay alive)."""
    lib = library()
    s = current_stream_handle() if stream is None else stream
    _check(lib.bzh_launch_program(s, ctypes.byref(program)), lib)
