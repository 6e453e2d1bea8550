"""libsci_amd: an MI355X (gfx950) HIP backend for libsci's chunk executor, products and solvers.

The C ABI (include/scihip.h, built into lib/libscihip.so) is the drop-in boundary behind the
reference's ``sci::Device``. This package is its Python host mirror: Chunk recording
(``chunk``), device arrays and lazy expressions (``array``). The C++ host mirror lives in
``include/sci/``.
"""

from .chunk import Chunk, Code, Fold, Precision, record_postfix  # noqa: F401
from .array import (  # noqa: F401
    Array, Device, Graph, Literal, Node, assign, evaluate, record, reduce, reduce_raw,
    sum, mean, minimum, maximum, dot, gradient, product, product_adjoint, reduce_axis, transpose,
    copy, negate, absolute, root, exponential, logarithm, sin, cos, tan, asin, acos, atan, csc, sec, cot,
    acsc, asec, acot, sinh, cosh, tanh, asinh, acosh, atanh, csch, sech, coth, acsch, asech, acoth,
    add, subtract, multiply, divide,
    linspace, full, ones, zeros, uniform, save, load,
)
from ._abi import BackendError, Declined  # noqa: F401
