"""Helpers shared by the tests: fixtures on disk, float comparisons."""

from __future__ import annotations

import json
import math
import os
from fractions import Fraction

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_json(name: str):
    with open(os.path.join(GOLDEN, name)) as handle:
        return json.load(handle)


def golden_kernels():
    data = np.load(os.path.join(GOLDEN, "kernels.npz"), allow_pickle=False)
    return {k.replace("__", "/"): data[k] for k in data.files}


def program_tuples(program):
    return [(k, v) for k, v in program]


def ulp_distance(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Units in the last place between two float arrays of the same dtype, in exact integer
    arithmetic; NaN against NaN is 0, NaN against a number is a huge distance."""
    a = np.asarray(a)
    b = np.asarray(b)
    kind, mask = (np.int32, 0x7FFFFFFF) if a.dtype == np.float32 else (np.int64, 0x7FFFFFFFFFFFFFFF)
    ia = a.view(kind).astype(object)
    ib = b.view(kind).astype(object)
    order = np.vectorize(lambda i: -(i & mask) if i < 0 else i, otypes=[object])
    d = np.abs(order(ia) - order(ib)).astype(np.float64)
    nan_a, nan_b = np.isnan(a), np.isnan(b)
    return np.where(nan_a & nan_b, 0.0, np.where(nan_a | nan_b, float(1 << 40), d))


def ulps32(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Units in the last place between two float32 arrays, vectorised for large arrays: the
    sign-magnitude bit patterns mapped onto one ordered integer line. NaN against NaN is 0, NaN
    against a number 2^40."""
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    d = np.abs(ia - ib)
    nan_a, nan_b = np.isnan(a), np.isnan(b)
    return np.where(nan_a & nan_b, 0, np.where(nan_a | nan_b, 1 << 40, d))


def exact_sum(values) -> float:
    """The sum of float64 values, within 2^-52 |sum| + 2^-100 sum |x| of the exact sum: a pairwise
    tree of TwoSum steps (Knuth), whose rounding errors are kept and added back at the end. At 2^28
    terms it runs in seconds where math.fsum would take minutes."""
    x = np.asarray(values, dtype=np.float64).ravel()
    corrections = []
    while x.size > 1:
        if x.size % 2:
            x = np.append(x, 0.0)
        a, b = x[0::2], x[1::2]
        s = a + b
        bb = s - a
        corrections.append(float(np.sum((a - (s - bb)) + (b - bb))))
        x = s
    return math.fsum([float(x[0]) if x.size else 0.0] + corrections)


def same_bits(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Bit-for-bit equality, except that any NaN equals any NaN (payloads are not part of the contract)."""
    kind = np.uint32 if a.dtype == np.float32 else np.uint64
    return (a.view(kind) == b.view(kind)) | (np.isnan(a) & np.isnan(b))


def round_to_float32(value: Fraction) -> np.float32:
    """Correctly rounded (nearest, ties to even) float32 of an exact rational."""
    if value == 0:
        return np.float32(0.0)
    sign = -1 if value < 0 else 1
    v = abs(value)
    e = v.numerator.bit_length() - v.denominator.bit_length()
    if Fraction(2) ** e > v:
        e -= 1
    e = max(e, -126)  # subnormals share the smallest exponent
    if e > 127:
        return np.float32(sign * np.inf)
    scale = Fraction(2) ** (23 - e)
    scaled = v * scale
    q, r = divmod(scaled.numerator, scaled.denominator)
    twice = 2 * r
    if twice > scaled.denominator or (twice == scaled.denominator and q % 2 == 1):
        q += 1
    result = Fraction(q) / scale
    if result >= Fraction(2) ** 128:
        return np.float32(sign * np.inf)
    return np.float32(sign * float(result))


def fma32(a, b, c) -> np.float32:
    """fmaf(a, b, c) rounded once, from exact rationals."""
    return round_to_float32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


# ---------------------------------------------------------------- adjoint cases

def _signed(rng, n, lo, hi):
    return rng.uniform(lo, hi, n) * rng.choice([-1.0, 1.0], n)


def adjoint_operands(code, n: int, seed: int, dtype=np.float64):
    """Seeded operands inside the domain where `code` is differentiable and its derivative is
    moderate, one array per operand."""
    from libsci_amd.chunk import ARITY, Code

    rng = np.random.default_rng(seed)
    code = Code(code)
    if code in (Code.root, Code.logarithm):
        first = rng.uniform(0.1, 3.0, n)
    elif code in (Code.tan, Code.sec):
        first = rng.uniform(-1.2, 1.2, n)
    elif code in (Code.asin, Code.acos, Code.atanh):
        first = rng.uniform(-0.9, 0.9, n)
    elif code in (Code.csc, Code.cot, Code.csch, Code.coth, Code.acsch):
        first = _signed(rng, n, 0.2, 1.5)
    elif code in (Code.acsc, Code.asec, Code.acoth):
        first = _signed(rng, n, 1.2, 3.0)
    elif code == Code.acosh:
        first = rng.uniform(1.2, 3.0, n)
    elif code == Code.asech:
        first = rng.uniform(0.1, 0.9, n)
    else:
        first = rng.uniform(-2.0, 2.0, n)
    operands = [first]
    if ARITY[code] >= 2:
        operands.append(_signed(rng, n, 0.5, 2.0) if code == Code.divide else rng.uniform(-2.0, 2.0, n))
    if ARITY[code] == 3:
        operands.append(rng.uniform(-2.0, 2.0, n))
    return [np.ascontiguousarray(o.astype(dtype)) for o in operands]


def c5_numpy(a, b, c, d):
    """The differentiation workload's loss terms sinh(a*b + c) * exp(-(d*d)) and their gradients,
    in float64 closed form (d/da = cosh(ab+c) b e, d/db = cosh(ab+c) a e, d/dc = cosh(ab+c) e,
    d/dd = -2 d sinh(ab+c) e with e = exp(-d^2))."""
    a, b, c, d = (np.asarray(x, dtype=np.float64) for x in (a, b, c, d))
    z = a * b + c
    e = np.exp(-(d * d))
    return np.sinh(z) * e, [np.cosh(z) * b * e, np.cosh(z) * a * e, np.cosh(z) * e, -2.0 * d * np.sinh(z) * e]


# ---------------------------------------------------------------- split expressions

class HostLeaf:
    """A host array as an expression leaf: it records its own address, so that the oracle can run
    chunks recorded from expression trees (libsci_amd.array.Node) over host data."""

    def __init__(self, values: np.ndarray):
        self.values = values

    def record(self, chunk) -> int:
        return chunk.input(self.values.ctypes.data)

    def devices(self):
        return []

    def leaf_dtype(self):
        return self.values.dtype

    def element_shape(self):
        return self.values.shape


def deep_program(operations: int, inputs: int, seed: int, constants: bool = True):
    """A seeded postfix program of `operations` operations over `inputs` arrays (and constants),
    built bottom-up from a pool of values: each operation transforms a pool value or combines it with
    a fresh input, a constant or another pool value; whatever remains is added together at the end.
    Fresh inputs enter at every level, so the result keeps depending on every element's inputs (nested
    sines alone would converge to a constant). Only codes the GPU computes bit for bit like the CPU
    (IEEE arithmetic, min / max, sin, cos, exp, |x|, -x), arranged so that magnitudes stay near 1."""
    from libsci_amd.chunk import Code

    rng = np.random.default_rng(seed)
    pool = [[("in", k)] for k in range(inputs)]
    unary = [Code.sin, Code.cos, Code.negate, Code.absolute]
    binary = [Code.add, Code.subtract, Code.multiply, Code.minimum, Code.maximum]
    done = 0
    while done < operations:
        k = int(rng.integers(len(pool)))
        r = rng.random()
        if r < 0.25:
            code = unary[int(rng.integers(len(unary)))]
            if rng.random() < 0.15:  # exp of a bounded value: e^sin(x)
                pool[k] = pool[k] + [("op", Code.sin), ("op", Code.exponential)]
                done += 2
            else:
                pool[k] = pool[k] + [("op", code)]
                done += 1
            continue
        code = binary[int(rng.integers(len(binary)))]
        if r < 0.6:  # with a fresh input
            pool[k] = pool[k] + [("in", int(rng.integers(inputs))), ("op", code)]
        elif constants and r < 0.8:  # with a constant
            pool[k] = pool[k] + [("const", float(np.float32(rng.uniform(-1.5, 1.5)))), ("op", code)]
        elif len(pool) > 1:  # with another pool value
            j = int(rng.integers(len(pool) - 1))
            j += j >= k
            merged = pool[k] + pool[j] + [("op", code)]
            pool = [p for i, p in enumerate(pool) if i not in (k, j)] + [merged]
            k = len(pool) - 1
        else:
            pool[k] = pool[k] + [("in", int(rng.integers(inputs))), ("op", code)]
        done += 1
        if code in (Code.add, Code.subtract, Code.multiply) and rng.random() < 0.5:
            pool[k] = pool[k] + [("op", Code.sin)]  # back near 1
            done += 1
    program = pool[0]
    for rest in pool[1:]:
        program = program + rest + [("op", Code.add)]
    return program


def tree_of(program, leaves, dtype=np.float32):
    """The expression tree (libsci_amd.array nodes) of a postfix program over the given leaves."""
    from libsci_amd.array import Literal, Node
    from libsci_amd.chunk import ARITY, Code

    stack = []
    for kind, value in program:
        if kind == "in":
            stack.append(leaves[value])
        elif kind == "const":
            stack.append(Literal(value, dtype))
        else:
            arity = ARITY[Code(value)]
            operands = stack[-arity:]
            del stack[-arity:]
            stack.append(Node(Code(value), operands))
    assert len(stack) == 1
    return stack[0]


def float64_of(program, arrays):
    """The program evaluated in float64 by numpy: the exact-arithmetic reference of a split run."""
    from libsci_amd.chunk import Code

    f = {Code.sin: np.sin, Code.cos: np.cos, Code.negate: np.negative, Code.absolute: np.abs,
         Code.exponential: np.exp, Code.add: np.add, Code.subtract: np.subtract, Code.multiply: np.multiply,
         Code.minimum: np.minimum, Code.maximum: np.maximum}
    stack = []
    for kind, value in program:
        if kind == "in":
            stack.append(arrays[value].astype(np.float64))
        elif kind == "const":
            stack.append(np.float64(value))
        else:
            code = Code(value)
            if code in (Code.sin, Code.cos, Code.negate, Code.absolute, Code.exponential):
                stack.append(f[code](stack.pop()))
            else:
                b, a = stack.pop(), stack.pop()
                stack.append(f[code](a, b))
    return stack[0]


def run_stages_on_oracle(expression, reduce_fold=None):
    """Runs a split expression over host leaves the way the device runs it: every stage through the
    oracle into a host temporary, then the last stage stored (reduce_fold None) or reduced."""
    import oracle
    from libsci_amd.array import _record, bind_stages, stages

    planned, final = stages(expression, reduce_fold is not None)
    results = []
    for stage in planned:
        bound = bind_stages(stage, results)
        chunk = _record(bound, False)
        assert chunk.valid()
        results.append(HostLeaf(oracle.execute(chunk, int(np.prod(bound.element_shape())), threads=8,
                                               dtype=bound.leaf_dtype())))
    bound = bind_stages(final, results)
    count = int(np.prod(bound.element_shape()))
    if reduce_fold is None:
        chunk = _record(bound, False)
        assert chunk.valid()
        return oracle.execute(chunk, count, threads=8, dtype=bound.leaf_dtype()), len(planned)
    chunk = _record(bound, True)
    assert chunk.valid()
    return oracle.reduce(chunk, int(reduce_fold), count, threads=8), len(planned)


# ---------------------------------------------------------------- exhaustive float32 digests

DIGEST_CODES = ("exponential", "logarithm", "sin", "cos", "tan", "csc", "sec", "cot", "root")
DIGEST_TRIG = ("sin", "cos", "tan", "csc", "sec", "cot")
DIGEST_CHUNK = 1 << 24  # input patterns per chunk digest; 256 chunks cover all 2^32
_C1, _C2 = 0xBF58476D1CE4E5B9, 0x94D049BB133111EB
_REACH = 0x49800000  # |x| <= 2^20: the bit pattern of 1048576.0f


def digest_chunk(code: str, chunk: int, outputs: np.ndarray) -> int:
    """The digest of one chunk of a code's outputs over every float32 input pattern: the sum mod 2^64
    of SplitMix64's finaliser of (pattern << 32 | output bits) over the chunk's 2^24 patterns, NaN
    outputs counted as one pattern (0x7FC00000). For sin, cos and the codes built from them, inputs
    beyond the Cody-Waite reach (|x| > 2^20, infinities, NaN) are left out: there the reference calls
    glibc and the GPU the device library, which agree within 2 ulp, not bit for bit. A sum of mixed
    terms is order-free, so a GPU can form it in any order, yet any changed output changes it."""
    patterns = np.arange(chunk * DIGEST_CHUNK, (chunk + 1) * DIGEST_CHUNK, dtype=np.uint64)
    bits = outputs.view(np.uint32).astype(np.uint64)
    bits[np.isnan(outputs)] = 0x7FC00000
    z = (patterns << np.uint64(32)) | bits
    z ^= z >> np.uint64(30)
    z *= np.uint64(_C1)
    z ^= z >> np.uint64(27)
    z *= np.uint64(_C2)
    z ^= z >> np.uint64(31)
    if code in DIGEST_TRIG:
        z[(patterns & np.uint64(0x7FFFFFFF)) > np.uint64(_REACH)] = 0
    low = int((z & np.uint64(0xFFFFFFFF)).sum(dtype=np.uint64))
    high = int((z >> np.uint64(32)).sum(dtype=np.uint64))
    return (low + (high << 32)) % (1 << 64)


def digest_inputs(chunk: int) -> np.ndarray:
    """The float32 inputs of one digest chunk: the bit patterns chunk * 2^24 ... in order."""
    return np.arange(chunk * DIGEST_CHUNK, (chunk + 1) * DIGEST_CHUNK, dtype=np.uint64).astype(np.uint32).view(np.float32)
