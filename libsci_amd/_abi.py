"""ctypes binding of the C ABI (include/scihip.h) exported by lib/libscihip.so.

The library must exist: nothing here falls back to a CPU path. Build it with
``python -m libsci_amd.build`` (or ``__graft_entry__.build()``).
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "lib", "libscihip.so")

OK = 0
DECLINED = 1
ERRORS = {-1: "invalid argument", -2: "out of device memory", -3: "device error", -4: "compile error"}

# Every symbol include/scihip.h declares, with (restype, argtypes).
_vp, _sz, _u32, _u64, _i, _d = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
_pp = ctypes.POINTER(ctypes.c_void_p)
SIGNATURES = {
    "scihip_error": (ctypes.c_char_p, []),
    "scihip_open": (_i, [_u32, _pp]),
    "scihip_close": (_i, [_vp]),
    "scihip_name": (ctypes.c_char_p, [_vp]),
    "scihip_synchronize": (_i, [_vp]),
    "scihip_allocate": (_i, [_vp, _sz, _pp]),
    "scihip_dispose": (_i, [_vp, _vp]),
    "scihip_locate": (_i, [_vp, _vp]),
    "scihip_upload": (_i, [_vp, _vp, _vp, _sz]),
    "scihip_download": (_i, [_vp, _vp, _vp, _sz]),
    "scihip_copy": (_i, [_vp, _vp, _vp, _sz]),
    "scihip_execute": (_i, [_vp, _vp, _vp, _sz]),
    "scihip_reduce": (_i, [_vp, _vp, _u32, _sz, ctypes.POINTER(_d)]),
    "scihip_reduce_static": (_i, [_vp, _vp, _u32, _sz, _vp]),
    "scihip_set_policy": (_i, [_vp, _i]),
    "scihip_prepare": (_i, [_vp, _vp]),
    "scihip_statistics": (_i, [_vp, ctypes.POINTER(_u64), ctypes.POINTER(_u64), ctypes.POINTER(_u64)]),
    "scihip_product": (_i, [_vp, _u32, _vp, _vp, _vp, _sz, _sz, _sz]),
    "scihip_factor": (_i, [_vp, _u32, _vp, _sz, _vp]),
    "scihip_substitute": (_i, [_vp, _u32, _vp, _vp, _vp, _vp, _sz, _sz]),
    "scihip_solve": (_i, [_vp, _u32, _vp, _vp, _vp, _sz, _sz, _d, ctypes.c_int32, _vp]),
    "scihip_stream": (_i, [_vp, _pp]),
    "scihip_fill_uniform": (_i, [_vp, _vp, _u32, _sz, _u64, _d, _d, _u64, _d]),
    "scihip_adjoint": (_i, [_vp, _vp, _vp, _pp, _sz, ctypes.POINTER(_d)]),
    "scihip_adjoint_describe": (_i, [_vp, _i, _i, _vp]),
    "scihip_product_adjoint": (_i, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _sz, _sz, _sz]),
    "scihip_transpose": (_i, [_vp, _u32, _vp, _vp, _sz, _sz]),
    "scihip_reduce_axis": (_i, [_vp, _vp, _u32, _sz, _sz, _u32, _vp]),
    "scihip_capture_begin": (_i, [_vp]),
    "scihip_capture_end": (_i, [_vp, _pp]),
    "scihip_replay": (_i, [_vp, _vp]),
    "scihip_graph_release": (_i, [_vp, _vp]),
}


class AdjointProgram(ctypes.Structure):
    """scihip_adjoint_program (include/scihip.h)."""

    _fields_ = [("length", ctypes.c_uint32), ("arguments", ctypes.c_uint32), ("literals", ctypes.c_uint32),
                ("registers", ctypes.c_uint32), ("gradients", ctypes.c_uint32), ("loss", ctypes.c_uint32),
                ("contract", ctypes.c_uint32), ("left", ctypes.c_uint32), ("right", ctypes.c_uint32),
                ("word", ctypes.c_uint32 * 256), ("third", ctypes.c_uint8 * 256), ("slot", ctypes.c_uint8 * 64),
                ("constants", ctypes.c_double * 72)]



class SolveStats(ctypes.Structure):
    """scihip_solve_stats (include/scihip.h)."""

    _fields_ = [("iterations", ctypes.c_int32), ("converged", ctypes.c_int32), ("backward_error", ctypes.c_double),
                ("factor_ms", ctypes.c_double), ("solve_ms", ctypes.c_double)]


_lib = None


class BackendError(RuntimeError):
    """A scihip call failed; the message is the library's own."""


class Declined(BackendError):
    """The GPU cannot run the request as given (the C ABI's SCIHIP_DECLINED)."""


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            raise ImportError(f"{LIBRARY} is not built: run `python -m libsci_amd.build` first")
        handle = ctypes.CDLL(LIBRARY)
        for name, (restype, argtypes) in SIGNATURES.items():
            function = getattr(handle, name)
            function.restype = restype
            function.argtypes = argtypes
        _lib = handle
    return _lib


def check(rc: int, what: str = "scihip") -> None:
    if rc == OK:
        return
    message = lib().scihip_error().decode(errors="replace")
    if rc == DECLINED:
        raise Declined(f"{what}: declined: {message}")
    raise BackendError(f"{what}: {ERRORS.get(rc, rc)}: {message}")
