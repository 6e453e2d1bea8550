"""Chunk bytecode: the recorded operations of one elementwise expression.

A Python restatement of the reference's recorder, ``core::Chunk``
(include/core/bytecode/chunk.hpp:28-232), with the same recording rules, because the rules decide
the results:

* inputs are deduplicated by address (chunk.hpp:40-47) and constants by their float64 bits
  (chunk.hpp:52-60);
* ``emit`` frees its sources before it picks its target, the lowest free temporary
  (chunk.hpp:69-93), so a target may reuse a source's register;
* an ``add`` whose second (then first) operand is the product just recorded turns that multiply
  into one ``fuse`` (a*b+c with one rounding) (chunk.hpp:73-76, 203-211);
* ``output`` retargets the last instruction to the output, or appends a copy (chunk.hpp:98-110);
* ``cost``, ``contracted`` and ``signature`` as in chunk.hpp:127-176.

The ctypes layout is ``scihip_chunk`` of include/scihip.h, the C ABI's chunk.
"""

from __future__ import annotations

import ctypes
import enum
import struct

CAPACITY = 64

INPUT = 0x00
LITERAL = 0x40
TEMPORARY = 0x80
OUTPUT = 0xC0
NONE = 0xFF
TYPES = 0xC0


class Code(enum.IntEnum):
    """Operation codes in the order of core::Code (include/core/bytecode/code.hpp:18-61)."""

    copy = 0
    negate = 1
    absolute = 2
    root = 3
    exponential = 4
    logarithm = 5
    sin = 6
    cos = 7
    tan = 8
    asin = 9
    acos = 10
    atan = 11
    csc = 12
    sec = 13
    cot = 14
    acsc = 15
    asec = 16
    acot = 17
    sinh = 18
    cosh = 19
    tanh = 20
    asinh = 21
    acosh = 22
    atanh = 23
    csch = 24
    sech = 25
    coth = 26
    acsch = 27
    asech = 28
    acoth = 29
    add = 30
    subtract = 31
    multiply = 32
    divide = 33
    minimum = 34
    maximum = 35
    fuse = 36


ARITY = {code: (3 if code == Code.fuse else 2 if code >= Code.add else 1) for code in Code}


class Fold(enum.IntEnum):
    """Reductions over a chunk's result (include/core/bytecode/fold.hpp:11)."""

    sum = 0
    minimum = 1
    maximum = 2


class Precision(enum.IntEnum):
    """Element types run natively (include/core/bytecode/precision.hpp:12)."""

    single = 0
    wide = 1


class Instruction(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("target", ctypes.c_uint8), ("sources", ctypes.c_uint8 * 3)]


class ChunkStruct(ctypes.Structure):
    """``scihip_chunk`` (include/scihip.h): the public fields of core::Chunk."""

    _fields_ = [
        ("precision", ctypes.c_uint8),
        ("length", ctypes.c_uint8),
        ("arguments", ctypes.c_uint8),
        ("literals", ctypes.c_uint8),
        ("temporaries", ctypes.c_uint8),
        ("result", ctypes.c_uint8),
        ("instructions", Instruction * CAPACITY),
        ("inputs", ctypes.c_void_p * CAPACITY),
        ("constants", ctypes.c_double * CAPACITY),
    ]


assert ctypes.sizeof(ChunkStruct) == 1352 and ChunkStruct.inputs.offset == 328 and ChunkStruct.constants.offset == 840


def _bits(value: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", float(value)))[0]


class Chunk:
    """Records one expression; ``struct`` is the C-ABI view passed to scihip_execute / scihip_reduce."""

    def __init__(self, precision: Precision = Precision.single):
        self.struct = ChunkStruct()
        self.struct.precision = int(precision)
        self.struct.result = NONE
        self.busy = 0
        self.overflow = False

    # -- the public fields, read the way the executors read them
    @property
    def precision(self) -> int:
        return self.struct.precision

    @property
    def length(self) -> int:
        return self.struct.length

    @property
    def arguments(self) -> int:
        return self.struct.arguments

    @property
    def literals(self) -> int:
        return self.struct.literals

    @property
    def temporaries(self) -> int:
        return self.struct.temporaries

    @property
    def result(self) -> int:
        return self.struct.result

    def instructions(self):
        s = self.struct
        return [(int(s.instructions[i].code), int(s.instructions[i].target), tuple(s.instructions[i].sources))
                for i in range(s.length)]

    def inputs(self):
        return [self.struct.inputs[i] or 0 for i in range(self.struct.arguments)]

    def constants(self):
        return [self.struct.constants[i] for i in range(self.struct.literals)]

    # -- recording (chunk.hpp:40-118)
    def _fail(self) -> int:
        self.overflow = True
        return NONE

    def input(self, address: int) -> int:
        s = self.struct
        for index in range(s.arguments):
            if (s.inputs[index] or 0) == address:
                return INPUT | index
        if s.arguments == CAPACITY:
            return self._fail()
        s.inputs[s.arguments] = address
        s.arguments += 1
        return INPUT | (s.arguments - 1)

    def constant(self, value: float) -> int:
        s = self.struct
        bits = _bits(value)
        for index in range(s.literals):
            if _bits(s.constants[index]) == bits:
                return LITERAL | index
        if s.literals == CAPACITY:
            return self._fail()
        s.constants[s.literals] = float(value)
        s.literals += 1
        return LITERAL | (s.literals - 1)

    def _release(self, value: int) -> None:
        if (value & TYPES) == TEMPORARY:
            self.busy &= ~(1 << (value & ~TYPES & 0x3F))

    def _contract(self, product: int, addend: int) -> int:
        s = self.struct
        if s.length == 0 or (product & TYPES) != TEMPORARY:
            return NONE
        last = s.instructions[s.length - 1]
        if last.code != Code.multiply or last.target != product:
            return NONE
        last.code = Code.fuse
        last.sources[2] = addend
        self._release(addend)
        return product

    def emit(self, code: Code, first: int, second: int = NONE, third: int = NONE) -> int:
        s = self.struct
        if self.overflow:
            return NONE
        if code == Code.add:
            merged = self._contract(second, first)
            if merged != NONE:
                return merged
            merged = self._contract(first, second)
            if merged != NONE:
                return merged
        if s.length == CAPACITY:
            return self._fail()
        self._release(first)
        self._release(second)
        self._release(third)
        if self.busy == (1 << 64) - 1:
            return self._fail()
        index = 0
        while (self.busy >> index) & 1:
            index += 1
        self.busy |= 1 << index
        s.temporaries = max(s.temporaries, index + 1)
        target = TEMPORARY | index
        ins = s.instructions[s.length]
        ins.code, ins.target = int(code), target
        ins.sources[0], ins.sources[1], ins.sources[2] = first, second, third
        s.length += 1
        return target

    def output(self, value: int) -> "Chunk":
        s = self.struct
        if self.overflow:
            return self
        if s.length > 0 and s.instructions[s.length - 1].target == value and (value & TYPES) == TEMPORARY:
            s.instructions[s.length - 1].target = OUTPUT
        elif s.length < CAPACITY:
            ins = s.instructions[s.length]
            ins.code, ins.target = int(Code.copy), OUTPUT
            ins.sources[0], ins.sources[1], ins.sources[2] = value, NONE, NONE
            s.length += 1
        else:
            self._fail()
        s.result = OUTPUT
        return self

    def reduce(self, value: int) -> "Chunk":
        self.struct.result = value
        return self

    def valid(self) -> bool:
        return not self.overflow

    def cost(self) -> int:
        s = self.struct
        total = s.arguments + 1
        for i in range(s.length):
            op = s.instructions[i].code
            if op in (Code.divide, Code.root):
                total += 4
            elif op in (Code.exponential, Code.logarithm):
                total += 12
            elif Code.sin <= op <= Code.tan:
                total += 16
            elif Code.asin <= op <= Code.acoth:
                total += 24
            else:
                total += 1
        return total

    def contracted(self, fold: Fold) -> bool:
        s = self.struct
        return (fold == Fold.sum and s.length > 0 and s.instructions[s.length - 1].code == Code.multiply
                and s.instructions[s.length - 1].target == s.result)

    def signature(self) -> bytes:
        s = self.struct
        key = bytearray([s.precision, s.arguments, s.literals, s.temporaries, s.result])
        for i in range(s.length):
            ins = s.instructions[i]
            key += bytes([ins.code, ins.target, *ins.sources])
        return bytes(key)

    def retarget(self, addresses) -> "Chunk":
        """Replaces input addresses in order (the recorded ids stay), for running one recording on other arrays."""
        for index, address in enumerate(addresses):
            self.struct.inputs[index] = address
        return self


def record_postfix(program, precision: Precision = Precision.single, finish: str = "output",
                   addresses=None) -> Chunk:
    """Records a postfix program the way Node::record walks a tree (include/array/node.hpp:75-88).

    ``program`` items: ("in", n) an input array numbered n, ("const", value), ("op", Code).
    Input n is recorded at ``addresses[n]`` (default: a distinct dummy address per n).
    """
    chunk = Chunk(precision)
    stack = []
    for item in program:
        kind = item[0]
        if kind == "in":
            address = addresses[item[1]] if addresses is not None else 0x1000 + 64 * item[1]
            stack.append(chunk.input(address))
        elif kind == "const":
            stack.append(chunk.constant(item[1]))
        else:
            code = Code(item[1])
            arity = ARITY[code]
            operands = stack[-arity:]
            del stack[-arity:]
            stack.append(chunk.emit(code, *operands))
    if len(stack) != 1:
        raise ValueError("a postfix program leaves exactly one value")
    if finish == "output":
        chunk.output(stack[0])
    else:
        chunk.reduce(stack[0])
    return chunk
