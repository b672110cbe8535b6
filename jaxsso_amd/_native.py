"""ctypes binding of libjsso_hip.so (the C ABI declared in include/jsso.h).

The library is the only compute path: if it is missing, or no GPU is visible,
every compute call raises :class:`NativeUnavailable`.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libjsso_hip.so")

JSSO_OK = 0
JSSO_NOT_CONVERGED = 1
PREC = {"fp64": 0, "mixed": 1, "plain": 2, "mixed64": 3}

EXPORTS = ("jsso_version", "jsso_last_error", "jsso_plan_create", "jsso_plan_destroy", "jsso_plan_sizes",
           "jsso_plan_pattern", "jsso_element_ke_beam", "jsso_element_ke_quad", "jsso_assemble",
           "jsso_solve", "jsso_adjoint_grad", "jsso_spmv", "jsso_compliance", "jsso_plan_cg_tuning",
           "jsso_timing_enable", "jsso_timing_read", "jsso_plan_set_preconditioner",
           "jsso_plan_amg_levels", "jsso_timing_read_slot", "jsso_spmv_k", "jsso_dense_solve",
           "jsso_element_beam_local", "jsso_element_quad_local", "jsso_element_quad_gauss",
           "jsso_plan_symmetric", "jsso_csr_spmv", "jsso_adjoint_grad_noalloc", "jsso_plan_element_tuning",
           "jsso_plan_amg_reuse", "jsso_solve_transpose", "jsso_plan_nonsymmetric_route")
TIMING_SLOTS = {"solve": 0, "element": 1, "gather": 2, "adjoint": 3, "setup": 4}


class NativeUnavailable(RuntimeError):
    pass


class JssoError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"jsso error {code}: {msg}")
        self.code = code


class Info(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("replacements", ctypes.c_int32),
                ("converged", ctypes.c_int32), ("n_systems", ctypes.c_int32),
                ("rel_residual", ctypes.c_double), ("waves_per_slice", ctypes.c_int32),
                ("levels", ctypes.c_int32), ("dir_launches", ctypes.c_int32),
                ("full_iterations", ctypes.c_int32), ("at_floor", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the HIP library (no GPU needed to load it)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NativeUnavailable(f"{path} is not built: run jaxsso_amd.build_lib.build() "
                                "(or __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    P, I32, I64, D, V = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
    lib.jsso_version.restype = ctypes.c_int
    lib.jsso_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    lib.jsso_plan_create.argtypes = [I32, P, I32, P, I32, P, I32, P, ctypes.POINTER(ctypes.c_void_p)]
    lib.jsso_plan_destroy.argtypes = [P]
    lib.jsso_plan_sizes.argtypes = [P, P]
    lib.jsso_plan_pattern.argtypes = [P, P, P]
    lib.jsso_element_ke_beam.argtypes = [I32, P, P, P, P, V]
    lib.jsso_element_ke_quad.argtypes = [I32, P, P, P, V]
    lib.jsso_assemble.argtypes = [P, P, P, P, P, V]
    lib.jsso_solve.argtypes = [P, P, P, I32, I32, D, I32, ctypes.POINTER(Info), V]
    lib.jsso_adjoint_grad.argtypes = [P, P, P, P, P, P, D, I32, P, P, P, V]
    lib.jsso_spmv.argtypes = [P, P, P, V]
    lib.jsso_compliance.argtypes = [P, P, P, I32, D, P, V]
    lib.jsso_plan_cg_tuning.argtypes = [P, I32, I32, I32, P]
    lib.jsso_timing_enable.argtypes = [I32]
    lib.jsso_timing_read.argtypes = [P, P, I32]
    lib.jsso_timing_read_slot.argtypes = [I32, P, P, I32]
    lib.jsso_spmv_k.argtypes = [P, P, P, P, V]
    lib.jsso_dense_solve.argtypes = [P, P, P, P, I32, I32, V]
    lib.jsso_plan_symmetric.argtypes = [P, P, V]
    lib.jsso_csr_spmv.argtypes = [I32, P, P, P, P, P, P, V]
    lib.jsso_adjoint_grad_noalloc.argtypes = [P, P, P, P, P, P, D, I32, P, P, P, V]
    lib.jsso_plan_element_tuning.argtypes = [P, I32, P]
    lib.jsso_plan_amg_reuse.argtypes = [P, I32, P]
    lib.jsso_solve_transpose.argtypes = [P, P, P, I32, I32, D, I32, ctypes.POINTER(Info), V]
    lib.jsso_plan_nonsymmetric_route.argtypes = [P, I32, P]
    lib.jsso_element_beam_local.argtypes = [I32, P, P, P, P, V]
    lib.jsso_element_quad_local.argtypes = [I32, P, P, P, P, P, P, V]
    lib.jsso_element_quad_gauss.argtypes = [I32, P, P, D, D, P, P, P, P, P, P, P, V]
    lib.jsso_plan_set_preconditioner.argtypes = [P, I32]
    lib.jsso_plan_amg_levels.argtypes = [P, P, I32, P]
    for name in EXPORTS:
        if name not in ("jsso_version",):
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    load_library().jsso_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(code: int, allow_not_converged: bool = False) -> int:
    if code == JSSO_OK or (allow_not_converged and code == JSSO_NOT_CONVERGED):
        return code
    raise JssoError(code, last_error())


def lib() -> ctypes.CDLL:
    """The library, for compute calls: requires a visible GPU."""
    L = load_library()
    if not torch.cuda.is_available():
        raise NativeUnavailable("no GPU visible: the HIP path is the only compute path")
    return L


def device() -> torch.device:
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t) -> ctypes.c_void_p:
    """Device (or host) address of a contiguous tensor / numpy array, None allowed."""
    if t is None:
        return ctypes.c_void_p(0)
    if isinstance(t, torch.Tensor):
        assert t.is_contiguous(), "tensor must be contiguous"
        return ctypes.c_void_p(t.data_ptr())
    return ctypes.c_void_p(t.ctypes.data)
