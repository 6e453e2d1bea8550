"""Adjoints on the CPU.

1. The oracle's adjoint (oracle.cpp, the chain rule over a chunk with textbook local derivatives) is
   pinned by float64 central differences of the oracle's own forward pass for every code, and by
   the closed-form gradient of the differentiation workload's loss sum(sinh(a*b + c) * exp(-(d*d))).
2. The backend's adjoint programs (the backward bytecode that the HIP kernels run), obtained through
   scihip_adjoint_describe without a GPU, are evaluated here element by element with the oracle's
   element kernels and must give the oracle's gradients: this checks the emitter's derivative
   rules, its dead-value removal and its register allocation on the CPU.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle
from helpers import adjoint_operands, c5_numpy
from libsci_amd import _abi
from libsci_amd.chunk import ARITY, Code, Precision, record_postfix

C5 = [("in", 0), ("in", 1), ("op", Code.multiply), ("in", 2), ("op", Code.add), ("op", Code.sinh),
      ("in", 3), ("in", 3), ("op", Code.multiply), ("op", Code.negate), ("op", Code.exponential),
      ("op", Code.multiply)]


def single_code(code):
    return [("in", k) for k in range(ARITY[Code(code)])] + [("op", Code(code))]


def chunk_for(program, arrays, precision, finish="reduce"):
    return record_postfix(program, precision, finish, addresses=[a.ctypes.data for a in arrays])


def forward(program, arrays, precision=Precision.wide):
    chunk = chunk_for(program, arrays, precision, "output")
    return oracle.execute(chunk, arrays[0].size, dtype=arrays[0].dtype)


def describe(chunk, seeded=False, loss=True):
    program = _abi.AdjointProgram()
    _abi.check(_abi.lib().scihip_adjoint_describe(ctypes.byref(chunk.struct), int(seeded), int(loss),
                                                  ctypes.byref(program)), "scihip_adjoint_describe")
    return program


def run_program(program, inputs, seed=None):
    """Evaluates an adjoint program in float64 on the CPU: returns (loss, gradients)."""
    n = inputs[0].size
    regs = np.zeros((max(program.registers, 1), n))
    for i, x in enumerate(list(inputs) + ([seed] if seed is not None else [])):
        regs[i] = x
    for k in range(program.literals):
        regs[program.arguments + k] = program.constants[k]
    for s in range(program.length):
        word = program.word[s]
        code, target = word & 0xFF, (word >> 8) & 0xFF
        a, b, c = regs[(word >> 16) & 0xFF], regs[word >> 24], regs[program.third[s]]
        if code < len(Code):
            out = np.array([oracle.apply_f64(code, a[e], b[e], c[e]) for e in range(n)])
        elif code == 37:  # a * sign(b)
            out = a * np.where(b > 0, 1.0, np.where(b < 0, -1.0, b))
        else:  # routing of minimum (38) / maximum (39): a where b wins, half on a tie
            wins, loses = (b < c, b > c) if code == 38 else (b > c, b < c)
            out = a * np.where(wins, 1.0, np.where(b == c, 0.5, np.where(loses, 0.0, 1.0)))
        regs[target] = out
    gradients = [regs[program.slot[i]].copy() for i in range(program.gradients)]
    if not program.loss:
        return None, gradients
    if program.contract:
        return float(np.sum(regs[program.left] * regs[program.right])), gradients
    return float(np.sum(regs[program.left])), gradients


# ---------------------------------------------------------------- the oracle, pinned

@pytest.mark.parametrize("code", list(Code), ids=lambda c: c.name)
def test_oracle_adjoint_matches_central_differences(code):
    arrays = adjoint_operands(code, 16, seed=100 + int(code))
    program = single_code(code)
    loss, gradients = oracle.adjoint(chunk_for(program, arrays, Precision.wide), 16, dtype=np.float64)
    assert abs(loss - forward(program, arrays).sum()) <= 1e-12 * max(1.0, abs(loss))
    for k, x in enumerate(arrays):
        h = 1e-6 * np.maximum(1.0, np.abs(x))
        up = [a.copy() for a in arrays]
        down = [a.copy() for a in arrays]
        up[k] = x + h
        down[k] = x - h
        difference = (forward(program, up) - forward(program, down)) / (up[k] - down[k])
        scale = np.maximum(1.0, np.abs(difference))
        assert (np.abs(gradients[k] - difference) <= 1e-6 * scale).all(), (code.name, k)


def test_oracle_adjoint_of_repeated_operands_and_ties():
    x = np.array([-1.5, -0.25, 0.5, 2.0])
    # x * x: d/dx = 2x; x / x: 0; maximum(x, x): the tie splits 1/2 + 1/2
    for program, want in (([("in", 0), ("in", 0), ("op", Code.multiply)], 2 * x),
                          ([("in", 0), ("in", 0), ("op", Code.divide)], 0 * x),
                          ([("in", 0), ("in", 0), ("op", Code.maximum)], np.ones_like(x)),
                          ([("in", 0), ("op", Code.absolute)], np.sign(x))):
        _, gradients = oracle.adjoint(chunk_for(program, [x, x], Precision.wide), 4, dtype=np.float64)
        assert np.allclose(gradients[0], want, rtol=0, atol=1e-15), program


@pytest.mark.parametrize("precision", [Precision.single, Precision.wide], ids=["fp32", "fp64"])
def test_oracle_adjoint_of_c5_matches_closed_form(precision):
    dtype = np.float32 if precision == Precision.single else np.float64
    n = 4096
    arrays = [oracle.uniform(n, 500 + k, -2.0, 2.0, dtype=dtype) for k in range(4)]
    loss, gradients = oracle.adjoint(chunk_for(C5, arrays, precision), n, dtype=dtype)
    terms, want = c5_numpy(*arrays)
    relative = 2e-6 if precision == Precision.single else 1e-13
    assert abs(loss - terms.sum()) <= relative * np.abs(terms).sum()
    for k in range(4):
        error = np.abs(gradients[k].astype(np.float64) - want[k])
        assert (error <= relative * (np.abs(want[k]) + np.abs(want[k]).max())).all(), k


# ---------------------------------------------------------------- the emitted adjoint programs

@pytest.mark.parametrize("code", list(Code), ids=lambda c: c.name)
@pytest.mark.parametrize("seeded", [False, True], ids=["loss", "seeded"])
def test_adjoint_program_matches_oracle(code, seeded):
    arrays = adjoint_operands(code, 8, seed=200 + int(code))
    chunk = chunk_for(single_code(code), arrays, Precision.wide)
    seed = np.linspace(-1.0, 2.0, 8) if seeded else None
    program = describe(chunk, seeded=seeded)
    loss, gradients = run_program(program, arrays, seed)
    want_loss, want = oracle.adjoint(chunk, 8, seed=seed, dtype=np.float64)
    if not seeded:
        assert abs(loss - want_loss) <= 1e-13 * max(1.0, abs(want_loss))
    for k in range(len(arrays)):
        assert np.allclose(gradients[k], want[k], rtol=1e-12, atol=1e-13), (code.name, k)


@pytest.mark.parametrize("case", ["c5", "square", "constants", "reuse", "minimum_literal", "chain"])
def test_adjoint_program_of_expressions(case):
    x, y, z, w = adjoint_operands(Code.fuse, 8, seed=7) + [np.linspace(0.5, 1.5, 8)]
    programs = {
        "c5": (C5, [x, y, z, w]),
        "square": ([("in", 0), ("in", 0), ("op", Code.multiply), ("op", Code.sin)], [x, x]),
        "constants": ([("in", 0), ("const", 2.5), ("op", Code.multiply), ("const", 1.0), ("op", Code.add)], [x]),
        "reuse": ([("in", 0), ("op", Code.exponential), ("in", 0), ("op", Code.multiply), ("in", 1), ("op", Code.divide),
                   ("in", 0), ("op", Code.tanh), ("op", Code.subtract)], [x, w, x]),
        "minimum_literal": ([("in", 0), ("const", 0.5), ("op", Code.minimum), ("in", 1), ("op", Code.maximum)], [x, y]),
        "chain": ([("in", 0), ("op", Code.sin), ("op", Code.cos), ("op", Code.atan), ("op", Code.sinh),
                   ("in", 1), ("op", Code.multiply), ("op", Code.absolute), ("op", Code.root)], [x, w]),
    }
    program, arrays = programs[case]
    chunk = chunk_for(program, arrays, Precision.wide)
    for seeded in (False, True):
        seed = np.linspace(-2.0, 1.0, 8) if seeded else None
        described = describe(chunk, seeded=seeded)
        inputs = [np.frombuffer((ctypes.c_double * 8).from_address(chunk.struct.inputs[i]), dtype=np.float64)
                  for i in range(chunk.struct.arguments)]
        loss, gradients = run_program(described, inputs, seed)
        want_loss, want = oracle.adjoint(chunk, 8, seed=seed, dtype=np.float64)
        if not seeded:
            assert abs(loss - want_loss) <= 1e-13 * max(1.0, abs(want_loss)), case
        for k in range(chunk.struct.arguments):
            assert np.allclose(gradients[k], want[k], rtol=1e-12, atol=1e-13), (case, seeded, k)


def test_c5_adjoint_program_shape():
    # The differentiation workload's chunk: contracted (the loss is sum(sinh(.) * exp(.)) as exact
    # products), 4 inputs, and a program that fits the 16-register interpreter with few instructions.
    arrays = [np.zeros(4) for _ in range(4)]
    chunk = chunk_for(C5, arrays, Precision.single)
    program = describe(chunk)
    assert program.contract == 1 and program.loss == 1
    assert program.arguments == 4 and program.gradients == 4
    assert program.registers <= 16
    assert program.length <= 14
    seeded = describe(chunk, seeded=True)
    assert seeded.contract == 0 and seeded.arguments == 5
    bare = describe(chunk, loss=False)
    assert bare.loss == 0 and bare.length <= program.length
