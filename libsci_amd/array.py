"""Arrays on a HIP device and lazy elementwise expressions over them.

The Python mirror of the reference's array layer: ``Device`` is ``sci::Device``
(include/core/drivers/device.hpp:27-89), ``Array`` is ``sci::Array`` (include/array/array.hpp:35-285)
for arrays that live on the device, and the expression nodes record a Chunk the way
``Node`` / ``Reference`` / ``Literal`` do (include/array/node.hpp:75-88, reference.hpp:45,
literal.hpp:34-36). Writing ``a * b + 2.0`` computes nothing; assigning it (``Array(expr)``,
``target.assign(expr)``) or reducing it (``sum(expr)``) records one Chunk and runs it in one pass
through the C ABI. Host data enters with ``to(device)`` and leaves with ``host()``.

There is no CPU engine here: every computation runs on the device, and an expression whose arrays
are not on one device raises.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np

from builtins import sum as builtins_sum

from . import _abi
from .chunk import Chunk, Code, Fold, Precision

_DTYPES = {np.dtype(np.float32): Precision.single, np.dtype(np.float64): Precision.wide}


class Device:
    """A HIP device (include/core/drivers/device.hpp:27-89). Raises when no GPU is usable."""

    def __init__(self, index: int = 0):
        handle = ctypes.c_void_p()
        _abi.check(_abi.lib().scihip_open(index, ctypes.byref(handle)), "scihip_open")
        self.handle = handle
        self.index = index

    @property
    def name(self) -> str:
        return _abi.lib().scihip_name(self.handle).decode()

    def synchronize(self) -> None:
        _abi.check(_abi.lib().scihip_synchronize(self.handle), "scihip_synchronize")

    def allocate(self, nbytes: int) -> int:
        pointer = ctypes.c_void_p()
        _abi.check(_abi.lib().scihip_allocate(self.handle, max(int(nbytes), 1), ctypes.byref(pointer)),
                   "scihip_allocate")
        return pointer.value

    def dispose(self, pointer: int) -> None:
        if self.handle:
            _abi.check(_abi.lib().scihip_dispose(self.handle, pointer), "scihip_dispose")

    def locate(self, pointer: int) -> bool:
        return bool(_abi.lib().scihip_locate(self.handle, pointer))

    def set_policy(self, policy: int) -> None:
        """0 automatic, 1 interpreter only, 2 specialised kernels only."""
        _abi.check(_abi.lib().scihip_set_policy(self.handle, int(policy)), "scihip_set_policy")

    def prepare(self, expression, precision=Precision.single, reduce: bool = False) -> None:
        """Compiles the specialised kernel of an expression's chunk now."""
        chunk = record(expression, reduce=reduce)
        _abi.check(_abi.lib().scihip_prepare(self.handle, ctypes.byref(chunk.struct)), "scihip_prepare")

    def statistics(self) -> dict:
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _abi.check(_abi.lib().scihip_statistics(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"interpreted": a.value, "specialised": b.value, "compiled": c.value}

    def begin_capture(self) -> None:
        """Records the device's asynchronous work (stores, products, adjoints without a loss, device
        copies) into a graph until end_capture(), instead of running it."""
        _abi.check(_abi.lib().scihip_capture_begin(self.handle), "scihip_capture_begin")

    def end_capture(self) -> "Graph":
        graph = ctypes.c_void_p()
        _abi.check(_abi.lib().scihip_capture_end(self.handle, ctypes.byref(graph)), "scihip_capture_end")
        return Graph(self, graph)

    def close(self) -> None:
        if self.handle:
            _abi.lib().scihip_close(self.handle)
            self.handle = None

    def __repr__(self) -> str:
        return f"Device({self.index}: {self.name})"


class Graph:
    """Captured device work, queued again as a whole by replay() (scihip_replay)."""

    def __init__(self, device: Device, handle: ctypes.c_void_p):
        self.device = device
        self.handle = handle

    def replay(self) -> None:
        _abi.check(_abi.lib().scihip_replay(self.device.handle, self.handle), "scihip_replay")

    def release(self) -> None:
        if self.handle and self.device.handle:
            _abi.check(_abi.lib().scihip_graph_release(self.device.handle, self.handle), "scihip_graph_release")
        self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


# ------------------------------------------------------------------------------------------ arrays

class _Storage:
    """A device allocation, returned to the device when the last array using it goes.

    Element access (``a[i]``, ``a[i, j] = v``) goes through a host copy of the allocation, since HBM
    is not mapped into the host as the reference's device memory is (include/array/array.hpp:186-217,
    include/array/storage.hpp:100-102). The copy is downloaded at the first access after device work
    (the download waits for that work), shared by the arrays viewing the allocation, and handed back
    at the next device use: ``Array.pointer``, which every device operation takes, uploads it first if
    an element was written and drops it, since the operation may write the allocation."""

    def __init__(self, device: Device, nbytes: int):
        self.device = device
        self.pointer = device.allocate(nbytes)
        self.nbytes = nbytes
        self.mirror = None   # the host copy (bytes), or None
        self.dirty = False   # an element was written through the copy since it was taken

    def host_bytes(self) -> np.ndarray:
        if self.mirror is None:
            mirror = np.empty(self.nbytes, dtype=np.uint8)
            if self.nbytes:
                _abi.check(_abi.lib().scihip_download(self.device.handle, mirror.ctypes.data, self.pointer,
                                                      self.nbytes), "scihip_download")
            self.mirror, self.dirty = mirror, False
        return self.mirror

    def release(self) -> None:
        """Before device work: elements written on the host go up, and the host copy is dropped."""
        if self.mirror is not None:
            if self.dirty and self.nbytes:
                _abi.check(_abi.lib().scihip_upload(self.device.handle, self.pointer, self.mirror.ctypes.data,
                                                    self.nbytes), "scihip_upload")
            self.mirror, self.dirty = None, False

    def __del__(self):
        try:
            if self.device.handle and self.pointer:
                self.device.dispose(self.pointer)
        except Exception:
            pass


class Array:
    """A contiguous row-major float32 / float64 array on a Device (include/array/array.hpp:35-285)."""

    def __init__(self, device: Device, shape, dtype=np.float32, storage: _Storage | None = None, offset: int = 0):
        self.device = device
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        if self.dtype not in _DTYPES:
            raise TypeError("arrays hold float32 or float64")
        self.size = int(np.prod(self.shape, dtype=np.int64)) if self.shape else 1
        self.storage = storage if storage is not None else _Storage(device, self.size * self.dtype.itemsize)
        self.offset = offset

    # -- identity
    @property
    def pointer(self) -> int:
        """The device address, for device work (hands the host copy of the elements back first)."""
        self.storage.release()
        return self.storage.pointer + self.offset

    @property
    def ndim(self) -> int:
        return len(self.shape)

    @property
    def precision(self) -> Precision:
        return _DTYPES[self.dtype]

    def __len__(self) -> int:
        return self.shape[0]

    # -- residency (array.hpp:223-237)
    @classmethod
    def from_numpy(cls, values, device: Device, dtype=None) -> "Array":
        data = np.ascontiguousarray(values, dtype=dtype if dtype is not None else None)
        if data.dtype not in _DTYPES:
            data = data.astype(np.float32)
        out = cls(device, data.shape, data.dtype)
        if data.size:
            _abi.check(_abi.lib().scihip_upload(device.handle, out.pointer, data.ctypes.data, data.nbytes),
                       "scihip_upload")
        return out

    def host(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        if out.size:
            _abi.check(_abi.lib().scihip_download(self.device.handle, out.ctypes.data, self.pointer, out.nbytes),
                       "scihip_download")
        return out

    numpy = host

    def to(self, device: Device) -> "Array":
        if device is self.device:
            out = Array(device, self.shape, self.dtype)
            _abi.check(_abi.lib().scihip_copy(device.handle, out.pointer, self.pointer, self.size * self.dtype.itemsize),
                       "scihip_copy")
            return out
        return Array.from_numpy(self.host(), device)

    # -- element access through the allocation's host copy (include/array/array.hpp:186-217)
    def _elements(self) -> np.ndarray:
        raw = self.storage.host_bytes()
        return raw[self.offset:self.offset + self.size * self.dtype.itemsize].view(self.dtype).reshape(self.shape)

    def __getitem__(self, index):
        value = self._elements()[index]
        return value.copy() if isinstance(value, np.ndarray) else value

    def __setitem__(self, index, value) -> None:
        self._elements()[index] = value
        self.storage.dirty = True

    def reshape(self, *shape) -> "Array":
        shape = shape[0] if len(shape) == 1 and isinstance(shape[0], (tuple, list)) else shape
        if int(np.prod(shape)) != self.size:
            raise ValueError("reshape keeps the element count")
        return Array(self.device, shape, self.dtype, self.storage, self.offset)

    def view(self, start: int, count: int) -> "Array":
        """Elements [start, start + count) of a flat array, sharing its memory (an interior pointer)."""
        if start < 0 or start + count > self.size:
            raise IndexError("view out of range")
        return Array(self.device, (count,), self.dtype, self.storage, self.offset + start * self.dtype.itemsize)

    # -- expression protocol
    def record(self, chunk: Chunk) -> int:
        return chunk.input(self.pointer)

    def devices(self):
        return [self.device]

    def leaf_dtype(self):
        return self.dtype

    def element_shape(self):
        return self.shape

    # -- evaluation
    def assign(self, expression) -> "Array":
        """Writes an expression into this array (include/array/expression.hpp:158-185)."""
        assign(self, expression)
        return self

    def sum(self):
        return sum(self)

    def mean(self):
        return mean(self)

    def minimum(self):
        return minimum(self)

    def maximum(self):
        return maximum(self)

    def __repr__(self) -> str:
        return f"Array(shape={self.shape}, dtype={self.dtype.name}, device={self.device.index})"

    # operators are shared with Node below
    __add__ = lambda self, other: _binary(Code.add, self, other)
    __radd__ = lambda self, other: _binary(Code.add, other, self)
    __sub__ = lambda self, other: _binary(Code.subtract, self, other)
    __rsub__ = lambda self, other: _binary(Code.subtract, other, self)
    __mul__ = lambda self, other: _binary(Code.multiply, self, other)
    __rmul__ = lambda self, other: _binary(Code.multiply, other, self)
    __truediv__ = lambda self, other: _binary(Code.divide, self, other)
    __rtruediv__ = lambda self, other: _binary(Code.divide, other, self)
    __neg__ = lambda self: Node(Code.negate, (self,))


class Literal:
    """A scalar leaf, already narrowed to the element type of the array side (expression.hpp:88-104)."""

    def __init__(self, value, dtype):
        self.value = float(np.dtype(dtype).type(value))
        self.dtype = np.dtype(dtype)

    def record(self, chunk: Chunk) -> int:
        return chunk.constant(self.value)

    def devices(self):
        return []

    def leaf_dtype(self):
        return self.dtype

    def element_shape(self):
        return None


class Node:
    """An operation over operands (include/array/node.hpp:26-88)."""

    def __init__(self, code: Code, children):
        self.code = Code(code)
        self.children = tuple(children)
        dtypes = {c.leaf_dtype() for c in self.children}
        if len(dtypes) != 1:
            raise TypeError("operands must share an element type")
        shapes = [c.element_shape() for c in self.children if c.element_shape() is not None]
        if not shapes:
            raise TypeError("an expression needs at least one array operand")
        if any(s != shapes[0] for s in shapes):
            raise ValueError("operands must share a shape (no broadcasting apart from scalars)")
        self.shape = shapes[0]
        self.dtype = dtypes.pop()

    def record(self, chunk: Chunk) -> int:
        # Operands first, left to right, then this operation (node.hpp:75-88).
        slots = [c.record(chunk) for c in self.children]
        return chunk.emit(self.code, *slots)

    def devices(self):
        found = []
        for c in self.children:
            for d in c.devices():
                if d not in found:
                    found.append(d)
        return found

    def leaf_dtype(self):
        return self.dtype

    def element_shape(self):
        return self.shape

    @property
    def size(self) -> int:
        return int(np.prod(self.shape)) if self.shape else 1

    def copy(self) -> Array:
        return evaluate(self)

    def sum(self):
        return sum(self)

    def mean(self):
        return mean(self)

    def minimum(self):
        return minimum(self)

    def maximum(self):
        return maximum(self)

    __add__ = Array.__add__
    __radd__ = Array.__radd__
    __sub__ = Array.__sub__
    __rsub__ = Array.__rsub__
    __mul__ = Array.__mul__
    __rmul__ = Array.__rmul__
    __truediv__ = Array.__truediv__
    __rtruediv__ = Array.__rtruediv__
    __neg__ = Array.__neg__


def _operand(x):
    return isinstance(x, (Array, Node))


def _binary(code: Code, left, right):
    if not _operand(left) and not _operand(right):
        return NotImplemented
    if not _operand(left):
        left = Literal(left, right.leaf_dtype())
    elif not _operand(right):
        right = Literal(right, left.leaf_dtype())
    return Node(code, (left, right))


# ------------------------------------------------------------------------------------------ running

def _device_of(expression) -> Device:
    devices = expression.devices()
    if not devices:
        raise TypeError("an expression needs an array operand")
    if len(devices) > 1:
        raise ValueError("every array of an expression must live on the same device")
    return devices[0]


def _record(expression, reduce: bool) -> Chunk:
    chunk = Chunk(_DTYPES[expression.leaf_dtype()])
    value = expression.record(chunk)
    if reduce:
        chunk.reduce(value)
    else:
        chunk.output(value)
    return chunk


def record(expression, reduce: bool = False) -> Chunk:
    """The one chunk of an expression; raises when it outgrows a chunk (see `stages` for the split)."""
    chunk = _record(expression, reduce)
    if not chunk.valid():
        raise ValueError("the expression outgrew a chunk (64 instructions, inputs or constants)")
    return chunk


class Staged:
    """A leaf standing for the result of an earlier stage of a split expression (a device temporary
    once the stages run). While the split is planned it records as an input of its own."""

    def __init__(self, index: int, dtype, shape):
        self.index = index
        self.dtype = np.dtype(dtype)
        self.shape = shape

    def record(self, chunk: Chunk) -> int:
        return chunk.input(0x10 + 16 * self.index)  # an address no allocation has; only identity matters

    def devices(self):
        return []

    def leaf_dtype(self):
        return self.dtype

    def element_shape(self):
        return self.shape


def _operations(expression) -> int:
    if not isinstance(expression, Node):
        return 0
    return 1 + builtins_sum(_operations(c) for c in expression.children)


def stages(expression, reduce: bool = False):
    """Splits an expression that outgrows one chunk (64 instructions, inputs, constants or
    temporaries, include/core/bytecode/chunk.hpp:31) into chunks run one after another on the device,
    each later one reading earlier results from device temporaries. The reference runs such an
    expression element by element on the host instead (include/array/expression.hpp:178-184,
    204-221). Returns (stages, final): stage k is a store whose Staged leaves name earlier stages;
    `final` is stored or reduced. The rule, shared with include/sci/hip.hpp: a node that fits one
    chunk stays whole; otherwise each child is made to fit on its own, then, while the node still
    outgrows a chunk, its operation child with the most operations (the leftmost of equals) becomes a
    stage of its own. A node over leaves always fits, so the split ends."""
    planned = []

    def fits(node, as_reduction):
        return _record(node, as_reduction).valid()

    def cut(node, as_reduction):
        if not isinstance(node, Node) or fits(node, as_reduction):
            return node
        children = [cut(c, False) for c in node.children]
        while True:
            candidate = Node(node.code, children)
            if fits(candidate, as_reduction):
                return candidate
            k = max((i for i, c in enumerate(children) if isinstance(c, Node)),
                    key=lambda i: (_operations(children[i]), -i))
            planned.append(children[k])
            children[k] = Staged(len(planned) - 1, children[k].leaf_dtype(), children[k].element_shape())

    final = cut(expression, reduce)
    return planned, final


def bind_stages(expression, results):
    """The expression with each Staged leaf replaced by results[index]."""
    if isinstance(expression, Staged):
        return results[expression.index]
    if isinstance(expression, Node):
        return Node(expression.code, [bind_stages(c, results) for c in expression.children])
    return expression


def _runnable(expression, reduce: bool):
    """The expression as one chunk can run it: unchanged when it fits, else its earlier stages are
    queued on the device (asynchronously, each into a temporary) and the last stage is returned."""
    if isinstance(expression, Node) and not _record(expression, reduce).valid():
        planned, final = stages(expression, reduce)
        results = []
        for stage in planned:
            results.append(evaluate(bind_stages(stage, results)))
        return bind_stages(final, results)
    return expression


def assign(target: Array, expression) -> None:
    """Runs an expression into `target` on the target's device (expression.hpp:158-185)."""
    if tuple(target.shape) != tuple(expression.element_shape()):
        raise ValueError("the target must have the expression's shape")
    if expression.leaf_dtype() != target.dtype:
        raise TypeError("the target must have the expression's element type")
    device = _device_of(expression)
    if device is not target.device:
        raise ValueError("the target and the expression's arrays must live on the same device")
    runnable = _runnable(expression, False)  # holds the stages' temporaries until the call is queued
    chunk = record(runnable)
    _abi.check(_abi.lib().scihip_execute(device.handle, ctypes.byref(chunk.struct), target.pointer, target.size),
               "scihip_execute")
    del runnable


def evaluate(expression) -> Array:
    """A new array holding the expression (Array(const source&), array.hpp:107-114)."""
    if isinstance(expression, Array):
        expression = Node(Code.copy, (expression,))
    device = _device_of(expression)
    out = Array(device, expression.element_shape(), expression.leaf_dtype())
    assign(out, expression)
    return out


def reduce_raw(expression, fold: Fold) -> float:
    """The fold of an expression in double, before the cast to the element type."""
    device = _device_of(expression)
    runnable = _runnable(expression, True)  # holds the stages' temporaries until the call is queued
    chunk = record(runnable, reduce=True)
    count = expression.size
    result = ctypes.c_double()
    _abi.check(_abi.lib().scihip_reduce(device.handle, ctypes.byref(chunk.struct), int(fold), count,
                                        ctypes.byref(result)), "scihip_reduce")
    del runnable
    return result.value


def reduce_axis(expression, fold: Fold, axis: int) -> Array:
    """Folds a matrix expression along one axis on the device (scihip_reduce_axis): axis 1 folds each
    row, axis 0 each column; each fold means what reduce means for that row or column alone."""
    shape = expression.element_shape()
    if shape is None or len(shape) != 2 or axis not in (0, 1):
        raise ValueError("axis reductions take a matrix expression and axis 0 or 1")
    device = _device_of(expression)
    runnable = _runnable(expression, True)  # holds the stages' temporaries until the call is queued
    chunk = record(runnable, reduce=True)
    rows, columns = shape
    out = Array(device, (rows if axis == 1 else columns,), expression.leaf_dtype())
    _abi.check(_abi.lib().scihip_reduce_axis(device.handle, ctypes.byref(chunk.struct), int(fold), rows, columns,
                                             axis, out.pointer), "scihip_reduce_axis")
    del runnable
    return out


def reduce(expression, fold: Fold, axis=None):
    """Sum, minimum or maximum, cast to the element type (expression.hpp:188-202); with an axis, a
    device array of the per-row (axis 1) or per-column (axis 0) folds."""
    if axis is not None:
        return reduce_axis(expression, fold, axis)
    return expression.leaf_dtype().type(reduce_raw(expression, fold))


def sum(expression, axis=None):
    return reduce(expression, Fold.sum, axis)


def mean(expression, axis=None):
    """float(sum) / float(count) in the element type (include/operations/reduce.hpp:24-28)."""
    dtype = expression.leaf_dtype().type
    if axis is not None:
        count = expression.element_shape()[1 - axis]
        sums = sum(expression, axis)
        return evaluate(sums / dtype(count)) if count else sums
    count = expression.size
    return dtype(0) if count == 0 else dtype(sum(expression) / dtype(count))


def minimum(left, right=None, axis=None):
    """minimum(a) reduces (reduce.hpp:36-38); minimum(a, b) is elementwise (binary.hpp:60-66)."""
    if right is None:
        return reduce(left, Fold.minimum, axis)
    return _binary(Code.minimum, left, right)


def maximum(left, right=None, axis=None):
    if right is None:
        return reduce(left, Fold.maximum, axis)
    return _binary(Code.maximum, left, right)


def save(path, array: Array) -> None:
    """Writes a device array to a .npy file (numpy's format, which include/sci/npy.hpp also reads and
    writes; the reference plans .npy I/O, README.md:155)."""
    np.save(path, array.host(), allow_pickle=False)


def load(path, device: Device) -> Array:
    """A device array holding a .npy file's float32 or float64 elements (nothing in the file is executed)."""
    values = np.load(path, allow_pickle=False)
    if values.dtype not in (np.float32, np.float64):
        raise TypeError(f"{path}: elements are {values.dtype}, not float32 or float64")
    return Array.from_numpy(values.reshape(values.shape or (1,)), device)


def transpose(array: Array) -> Array:
    """A new device array holding array^T (scihip_transpose; the reference's transpose,
    include/operations/manipulations.hpp:36-62, remaps on the host)."""
    if array.ndim != 2:
        raise ValueError("transpose takes a matrix")
    rows, columns = array.shape
    out = Array(array.device, (columns, rows), array.dtype)
    _abi.check(_abi.lib().scihip_transpose(array.device.handle, int(array.precision), array.pointer, out.pointer,
                                           rows, columns), "scihip_transpose")
    return out


def dot(left, right):
    """sum(a * b): contracted into a dot product (include/linear/product.hpp:72-76)."""
    return sum(_binary(Code.multiply, left, right))


def _leaves(expression, found):
    if isinstance(expression, Array):
        found.append(expression)
    elif isinstance(expression, Node):
        for child in expression.children:
            _leaves(child, found)
    return found


def gradient(expression, arrays, seed: Array | None = None, loss: bool = True):
    """Reverse-mode adjoint of an elementwise expression (scihip_adjoint), forward and backward in one
    pass: returns (loss, gradients), gradients[k] being d/d arrays[k] of sum(expression) or, with
    `seed`, of sum(seed * expression) (a vector-Jacobian product). The loss is the sum of the
    expression in double (None when `loss` is false); an array the expression does not use gets zeros.
    """
    device = _device_of(expression)
    chunk = record(_runnable(expression, True), reduce=True)
    count = expression.size
    inputs = [chunk.struct.inputs[i] for i in range(chunk.struct.arguments)]
    outputs = {}
    for array in arrays:
        if not isinstance(array, Array) or array.device is not device:
            raise ValueError("gradients are taken with respect to arrays on the expression's device")
        if array.pointer in inputs and array.pointer not in outputs:
            outputs[array.pointer] = Array(device, expression.element_shape(), expression.leaf_dtype())
    pointers = (ctypes.c_void_p * max(1, len(inputs)))()
    for i, address in enumerate(inputs):
        pointers[i] = outputs[address].pointer if address in outputs else None
    if seed is not None and (tuple(seed.shape) != tuple(expression.element_shape()) or
                             seed.dtype != expression.leaf_dtype() or seed.device is not device):
        raise ValueError("the seed must have the expression's shape, element type and device")
    total = ctypes.c_double()
    _abi.check(_abi.lib().scihip_adjoint(device.handle, ctypes.byref(chunk.struct),
                                         seed.pointer if seed is not None else None, pointers, count,
                                         ctypes.byref(total) if loss else None), "scihip_adjoint")
    gradients = []
    for array in arrays:
        if array.pointer in outputs:
            gradients.append(outputs[array.pointer])
        else:
            gradients.append(zeros(expression.element_shape(), device, expression.leaf_dtype()))
    return (total.value if loss else None), gradients


_PRODUCT_MODES = {"fp32": 0, "bf16": 1, "bf16x3": 2, "fp64": 3}


def product(left: Array, right: Array, mode: str = "fp32") -> Array:
    """left @ right on the matrix cores (scihip_product; include/linear/product.hpp:22-65)."""
    rows, depth = left.shape
    depth2, columns = right.shape
    if depth != depth2 or left.device is not right.device or left.dtype != right.dtype:
        raise ValueError("product needs conforming matrices of one element type on one device")
    out = Array(left.device, (rows, columns), left.dtype)
    _abi.check(_abi.lib().scihip_product(left.device.handle, _PRODUCT_MODES[mode], left.pointer, right.pointer,
                                         out.pointer, rows, columns, depth), "scihip_product")
    return out


def product_adjoint(left: Array, right: Array, seed: Array, mode: str = "fp32"):
    """(seed @ right^T, left^T @ seed): the gradients of sum(seed * (left @ right)) with respect to
    left and right, as two matrix-core products (scihip_product_adjoint)."""
    rows, depth = left.shape
    _, columns = right.shape
    if tuple(seed.shape) != (rows, columns):
        raise ValueError("the seed must have the product's shape")
    d_left = Array(left.device, (rows, depth), left.dtype)
    d_right = Array(left.device, (depth, columns), left.dtype)
    _abi.check(_abi.lib().scihip_product_adjoint(left.device.handle, _PRODUCT_MODES[mode], left.pointer,
                                                 right.pointer, seed.pointer, d_left.pointer, d_right.pointer,
                                                 rows, columns, depth), "scihip_product_adjoint")
    return d_left, d_right


def _unary(code: Code):
    def op(x):
        return Node(code, (x,))
    op.__name__ = code.name
    op.__doc__ = f"Elementwise {code.name} (include/operations/unary.hpp)."
    return op


def _named_binary(code: Code):
    def op(a, b):
        return _binary(code, a, b)
    op.__name__ = code.name
    return op


copy = _unary(Code.copy)
negate = _unary(Code.negate)
absolute = _unary(Code.absolute)
root = _unary(Code.root)
exponential = _unary(Code.exponential)
logarithm = _unary(Code.logarithm)
sin = _unary(Code.sin)
cos = _unary(Code.cos)
tan = _unary(Code.tan)
asin = _unary(Code.asin)
acos = _unary(Code.acos)
atan = _unary(Code.atan)
csc = _unary(Code.csc)
sec = _unary(Code.sec)
cot = _unary(Code.cot)
acsc = _unary(Code.acsc)
asec = _unary(Code.asec)
acot = _unary(Code.acot)
sinh = _unary(Code.sinh)
cosh = _unary(Code.cosh)
tanh = _unary(Code.tanh)
asinh = _unary(Code.asinh)
acosh = _unary(Code.acosh)
atanh = _unary(Code.atanh)
csch = _unary(Code.csch)
sech = _unary(Code.sech)
coth = _unary(Code.coth)
acsch = _unary(Code.acsch)
asech = _unary(Code.asech)
acoth = _unary(Code.acoth)
add = _named_binary(Code.add)
subtract = _named_binary(Code.subtract)
multiply = _named_binary(Code.multiply)
divide = _named_binary(Code.divide)


# ------------------------------------------------------------------------------------------ creation

def linspace(start: float, stop: float, count: int, device: Device, dtype=np.float32) -> Array:
    """x_i = T(start + i * step), step = double(stop - start) / (count - 1) (creation.hpp:65-84)."""
    t = np.dtype(dtype).type
    values = np.empty(count, dtype=dtype)
    if count == 1:
        values[0] = t(start)
    elif count > 1:
        step = float(t(t(stop) - t(start))) / (float(count) - 1.0)
        values[:] = (float(t(start)) + np.arange(count, dtype=np.float64) * step).astype(dtype)
    return Array.from_numpy(values, device)


def full(shape, value, device: Device, dtype=np.float32) -> Array:
    return Array.from_numpy(np.full(shape, value, dtype=dtype), device)


def ones(shape, device: Device, dtype=np.float32) -> Array:
    return full(shape, 1, device, dtype)


def zeros(shape, device: Device, dtype=np.float32) -> Array:
    return full(shape, 0, device, dtype)


def uniform(count: int, device: Device, seed: int, lower: float, upper: float, dtype=np.float32,
            nan_seed: int = 0, nan_fraction: float = 0.0) -> Array:
    """Seeded uniform data generated on the device (scihip_fill_uniform)."""
    out = Array(device, (count,), dtype)
    _abi.check(_abi.lib().scihip_fill_uniform(device.handle, out.pointer, int(_DTYPES[np.dtype(dtype)]), count,
                                              seed, lower, upper, nan_seed, nan_fraction), "scihip_fill_uniform")
    return out


__all__ = [name for name in dir() if not name.startswith("_") and name not in ("ctypes", "math", "np", "builtins_sum")]
