"""CPU oracle for the HIP backend. TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py import this
package; the backend (libsci_amd) never does. Two libraries:

* ``liboracle.so`` (oracle.cpp): a restatement of the reference's CPU path, each function citing
  the reference file:line it follows. It travels with the repository and is the checker on the GPU
  box, where the reference does not exist.
* ``_ref/libsciref.so`` (ref_harness.cpp + the reference's own kernel tables and pool, compiled from
  /root/reference by oracle/Makefile): used here, where the reference is mounted, to pin the
  restatement and to write the golden fixtures under tests/golden/.

Pinning status of each oracle function (what checks it besides the GPU comparison):

* chunk recording (libsci_amd.chunk): hand-derived bytecode KATs and the reference's own Chunk
  through ``_ref`` (tests/test_oracle.py, tests/golden/chunks.json).
* element kernels, every code: bit-for-bit against the reference's AVX-512 and fallback tables
  through ``_ref`` on seeded inputs and special values; closed forms (IEEE +, -, *, /, fma against
  numpy's correctly rounded float32 operations) and ulp bounds against float64 libm.
* reductions: bit-for-bit against ``_ref`` with the same part split; sums also within a stated
  bound of an exact float64 sum; min / max exact; NaN wins.
* product: against a float64 numpy product and the reference test's 37x53x29 KAT.
* LU and substitution: the reference tests' KATs (pivot rows [1, 2, 0] and x = [1, -2, 3]),
  P A = L U, and scipy's LAPACK getrf pivots (same first-maximum rule).
* synthetic data: closed form (every value exact in float, in [lower, upper)).
* adjoint (no reference counterpart): float64 central differences of the oracle's own forward
  pass in wide precision for every code and the C5 loss, and hand-derived gradients at chosen
  points (tests/test_oracle.py).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsciref.so")

_lib = None
_ref = None


def build(reference: bool = False) -> None:
    """Compiles liboracle.so, and _ref when asked and /root/reference exists."""
    targets = ["all"]
    if reference and os.path.isdir("/root/reference"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE, *targets], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
        vp, sz, i, d = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double
        L.oracle_execute.argtypes = [vp, vp, sz, i]
        L.oracle_reduce.argtypes = [vp, i, sz, i, i]
        L.oracle_reduce.restype = d
        L.oracle_cost.argtypes = [vp]
        L.oracle_cost.restype = ctypes.c_uint32
        L.oracle_apply_f32.argtypes = [i, ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.oracle_apply_f32.restype = ctypes.c_float
        L.oracle_apply_f64.argtypes = [i, d, d, d]
        L.oracle_apply_f64.restype = d
        L.oracle_map_f32.argtypes = [i, vp, vp, vp, vp, sz, i]
        L.oracle_product.argtypes = [i, vp, vp, vp, sz, sz, sz, i]
        L.oracle_factor.argtypes = [i, vp, vp, sz, i]
        L.oracle_substitute.argtypes = [i, vp, vp, vp, vp, sz, sz]
        L.oracle_linspace_f32.argtypes = [ctypes.c_float, ctypes.c_float, sz, vp]
        L.oracle_fill_uniform.argtypes = [vp, i, sz, ctypes.c_uint64, d, d, ctypes.c_uint64, d]
        L.oracle_adjoint.argtypes = [vp, sz, vp, ctypes.POINTER(vp), ctypes.POINTER(d)]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


class Step(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("code", ctypes.c_int32), ("arity", ctypes.c_int32),
                ("input", ctypes.c_int32), ("value", ctypes.c_double)]


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError("oracle/_ref/libsciref.so is not built (make -C oracle ref)")
        L = ctypes.CDLL(REF_SO)
        vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        L.ref_threads.argtypes = [i]
        L.ref_record.argtypes = [vp, i, i, i, vp, vp, vp, vp]
        L.ref_execute.argtypes = [vp, i, i, vp, vp, sz, i]
        L.ref_reduce.argtypes = [vp, i, i, vp, i, sz, i]
        L.ref_reduce.restype = ctypes.c_double
        L.ref_kernel.argtypes = [i, i, vp, vp, vp, vp, sz, i]
        _ref = L
    return _ref


def _ptr(array):
    return None if array is None else array.ctypes.data


def _steps(program):
    from libsci_amd.chunk import ARITY, Code

    steps = (Step * len(program))()
    for k, item in enumerate(program):
        if item[0] == "in":
            steps[k].kind, steps[k].input = 0, item[1]
        elif item[0] == "const":
            steps[k].kind, steps[k].value = 1, float(item[1])
        else:
            code = Code(item[1])
            steps[k].kind, steps[k].code, steps[k].arity = 2, int(code), ARITY[code]
    return steps


# ---------------------------------------------------------------- restatement entry points

def execute(chunk, count: int, threads: int = 1, dtype=np.float32) -> np.ndarray:
    """Runs a recorded chunk (inputs are host addresses) into a new array."""
    out = np.empty(count, dtype=dtype)
    lib().oracle_execute(ctypes.addressof(chunk.struct), out.ctypes.data, count, threads)
    return out


def reduce(chunk, fold: int, count: int, threads: int = 1, isa: int = 2) -> float:
    return lib().oracle_reduce(ctypes.addressof(chunk.struct), int(fold), count, threads, isa)


def map_f32(code: int, a, b=None, c=None, threads: int = 1) -> np.ndarray:
    out = np.empty_like(a)
    lib().oracle_map_f32(int(code), _ptr(a), _ptr(b), _ptr(c), out.ctypes.data, a.size, threads)
    return out


def apply_f64(code: int, a: float, b: float = 0.0, c: float = 0.0) -> float:
    return lib().oracle_apply_f64(int(code), a, b, c)


def product(left: np.ndarray, right: np.ndarray, threads: int = 1) -> np.ndarray:
    rows, depth = left.shape
    columns = right.shape[1]
    out = np.empty((rows, columns), dtype=left.dtype)
    precision = 0 if left.dtype == np.float32 else 1
    lib().oracle_product(precision, left.ctypes.data, right.ctypes.data, out.ctypes.data, rows, columns, depth, threads)
    return out


def factor(matrix: np.ndarray, threads: int = 1):
    """LU with partial pivoting in place on a copy: returns (packed factors, row order)."""
    work = np.ascontiguousarray(matrix).copy()
    rows = np.empty(work.shape[0], dtype=np.int64)
    precision = 0 if work.dtype == np.float32 else 1
    lib().oracle_factor(precision, work.ctypes.data, rows.ctypes.data, work.shape[0], threads)
    return work, rows


def substitute(factors: np.ndarray, rows: np.ndarray, right: np.ndarray) -> np.ndarray:
    order = factors.shape[0]
    width = 1 if right.ndim == 1 else right.shape[1]
    solution = np.empty_like(right)
    precision = 0 if factors.dtype == np.float32 else 1
    lib().oracle_substitute(precision, factors.ctypes.data, rows.ctypes.data, np.ascontiguousarray(right).ctypes.data,
                            solution.ctypes.data, order, width)
    return solution


def linspace(start: float, stop: float, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.float32)
    lib().oracle_linspace_f32(start, stop, count, out.ctypes.data)
    return out


def uniform(count: int, seed: int, lower: float, upper: float, dtype=np.float32, nan_seed: int = 0,
            nan_fraction: float = 0.0) -> np.ndarray:
    out = np.empty(count, dtype=dtype)
    precision = 0 if dtype == np.float32 else 1
    lib().oracle_fill_uniform(out.ctypes.data, precision, count, seed, lower, upper, nan_seed, nan_fraction)
    return out


def adjoint(chunk, count: int, seed=None, dtype=np.float32, loss: bool = True):
    """Gradients of sum(result) (or sum(seed * result)) of a recorded chunk whose inputs are host
    addresses, one array per chunk input, and the loss (None when not asked)."""
    arguments = chunk.struct.arguments
    gradients = [np.empty(count, dtype=dtype) for _ in range(arguments)]
    table = (ctypes.c_void_p * max(1, arguments))(*[g.ctypes.data for g in gradients])
    total = ctypes.c_double()
    lib().oracle_adjoint(ctypes.addressof(chunk.struct), count, _ptr(seed), table,
                         ctypes.byref(total) if loss else None)
    return (total.value if loss else None), gradients


# ---------------------------------------------------------------- the reference's own code (_ref)

def ref_record(program, precision: int = 0, finish: str = "output") -> dict:
    steps = _steps(program)
    header = (ctypes.c_int32 * 8)()
    code = (ctypes.c_uint8 * (5 * 64))()
    inputs = (ctypes.c_int32 * 64)()
    constants = (ctypes.c_double * 64)()
    rc = ref().ref_record(steps, len(program), precision, 0 if finish == "output" else 1, header, code, inputs,
                          constants)
    if rc != 0:
        raise ValueError("malformed program")
    length = header[1]
    return {
        "precision": header[0], "length": length, "arguments": header[2], "literals": header[3],
        "temporaries": header[4], "result": header[5], "valid": bool(header[6]), "cost": header[7],
        "instructions": [(code[5 * k], code[5 * k + 1], (code[5 * k + 2], code[5 * k + 3], code[5 * k + 4]))
                         for k in range(length)],
        "inputs": [inputs[k] for k in range(header[2])],
        "constants": [constants[k] for k in range(header[3])],
    }


def ref_execute(program, arrays, isa: int = 2) -> np.ndarray:
    out = np.empty_like(arrays[0])
    table = (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])
    precision = 0 if arrays[0].dtype == np.float32 else 1
    rc = ref().ref_execute(_steps(program), len(program), precision, table, out.ctypes.data, arrays[0].size, isa)
    if rc != 0:
        raise ValueError("malformed program")
    return out


def ref_reduce(program, arrays, fold: int, isa: int = 2) -> float:
    table = (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])
    precision = 0 if arrays[0].dtype == np.float32 else 1
    return ref().ref_reduce(_steps(program), len(program), precision, table, int(fold), arrays[0].size, isa)


def ref_kernel(code: int, a, b=None, c=None, isa: int = 2) -> np.ndarray:
    out = np.empty_like(a)
    precision = 0 if a.dtype == np.float32 else 1
    ref().ref_kernel(int(code), precision, _ptr(a), _ptr(b), _ptr(c), out.ctypes.data, a.size, isa)
    return out


def ref_threads(threads: int) -> int:
    return ref().ref_threads(threads)
