"""The C++ host mirror (include/sci/hip.hpp) through its test program tests/cpp/workload.cpp.

CPU: the mirror's Chunk records the workload expressions exactly as the Python recorder does, which
tests/test_oracle.py pins to the reference's own Chunk and to the hand-derived bytecode.
GPU: the program's small workload (c = a*b + 2 over linspace / ones, sum(sin(c)), a fused gradient,
a product, saved to .npy and loaded back, and the 3 x 3 solve KAT) agrees with the oracle: c bit for bit, the sum within the
reordering bound, the gradients within the adjoint tests' bound, the product exactly (its values are
exact in float) and x = [1, -2, 3] within 1e-12.
"""

from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import oracle
from libsci_amd.chunk import Code, Fold, Precision, record_postfix

HERE = os.path.dirname(os.path.abspath(__file__))
PROGRAM = os.path.join(HERE, "cpp", "workload")

PROGRAMS = {
    "c1_store": ([("in", 0), ("in", 1), ("op", Code.multiply), ("const", 2.0), ("op", Code.add)], "output"),
    "c1_reduce": ([("in", 0), ("op", Code.sin)], "reduce"),
    "c2_store": ([("in", 0), ("op", Code.exponential), ("const", 0.5), ("op", Code.multiply),
                  ("in", 1), ("op", Code.absolute), ("const", 1.0), ("op", Code.add), ("op", Code.logarithm),
                  ("in", 2), ("op", Code.multiply), ("op", Code.add)], "output"),
    "c2_reduce": ([("in", 0), ("op", Code.sin), ("in", 2), ("op", Code.cos), ("op", Code.multiply),
                   ("in", 1), ("op", Code.tanh), ("op", Code.add)], "reduce"),
    "c5_loss": ([("in", 0), ("in", 1), ("op", Code.multiply), ("in", 2), ("op", Code.add), ("op", Code.sinh),
                 ("in", 3), ("in", 3), ("op", Code.multiply), ("op", Code.negate), ("op", Code.exponential),
                 ("op", Code.multiply)], "reduce"),
}


def program():
    """The built test program (__graft_entry__.build() makes it; built here when missing)."""
    if not os.path.exists(PROGRAM):
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp")], check=True)
    return PROGRAM


def describe(chunk) -> str:
    s = chunk.struct
    words = [str(s.length), str(s.arguments), str(s.literals), str(s.temporaries), str(s.result)]
    for i in range(s.length):
        ins = s.instructions[i]
        words.append(f"{ins.code}:{ins.target}:{ins.sources[0]}:{ins.sources[1]}:{ins.sources[2]}")
    words += [f"L{s.constants[i]:.17g}" for i in range(s.literals)]
    return " ".join(words)


def test_cpp_recorder_matches_python_recorder():
    out = subprocess.run([program(), "record"], capture_output=True, text=True, check=True).stdout
    lines = dict(line.split(" ", 1) for line in out.strip().splitlines())
    assert set(lines) == set(PROGRAMS)
    for name, (steps, finish) in PROGRAMS.items():
        chunk = record_postfix(steps, Precision.single, finish, addresses=[0x1000 + 64 * k for k in range(4)])
        # the C++ program prints constants with %.17g, as describe() does; 2.0 prints as "2"
        assert lines[name] == describe(chunk).replace("L2.0000000000000000", "L2"), name


@pytest.mark.gpu
def test_cpp_workload_on_the_gpu(tmp_path):
    print("[test] tests/cpp/workload run", flush=True)
    result = subprocess.run([program(), "run", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert result.returncode == 0, result.stderr

    n = 1_000_000
    a = oracle.linspace(0.0, 1.0, n)
    b = np.ones(n, dtype=np.float32)
    want = oracle.execute(record_postfix(PROGRAMS["c1_store"][0], Precision.single, "output",
                                         [a.ctypes.data, b.ctypes.data]), n)
    c = np.fromfile(tmp_path / "c.f32", dtype=np.float32)
    assert (c.view(np.uint32) == want.view(np.uint32)).all()

    sines = oracle.map_f32(Code.sin, want).astype(np.float64)
    total = np.fromfile(tmp_path / "total.f64", dtype=np.float64)[0]
    assert abs(total - sines.sum()) <= n * 2 ** -24 * np.abs(sines).sum() / 64

    m = 1000
    inputs = np.fromfile(tmp_path / "inputs.f32", dtype=np.float32).reshape(4, m)
    gradients = np.fromfile(tmp_path / "gradients.f32", dtype=np.float32).reshape(4, m)
    arrays = [np.ascontiguousarray(inputs[k]) for k in range(4)]
    chunk = record_postfix(PROGRAMS["c5_loss"][0], Precision.single, "reduce", [x.ctypes.data for x in arrays])
    want_loss, want_gradients = oracle.adjoint(chunk, m)
    loss = np.fromfile(tmp_path / "loss.f64", dtype=np.float64)[0]
    assert abs(loss - want_loss) <= 1e-6 * max(1.0, abs(want_loss))
    for k in range(4):
        bound = 1e-5 * np.abs(want_gradients[k]) + 1e-6 * max(1.0, np.abs(want_gradients[k]).max())
        assert (np.abs(gradients[k] - want_gradients[k]) <= bound).all(), k

    left = ((np.arange(64 * 40) % 7) - 3).astype(np.float32).reshape(64, 40) * 0.25
    right = ((np.arange(40 * 48) % 5) - 2).astype(np.float32).reshape(40, 48) * 0.5
    product = np.fromfile(tmp_path / "product.f32", dtype=np.float32).reshape(64, 48)
    assert (product == (left.astype(np.float64) @ right.astype(np.float64))).all()
    saved = np.load(tmp_path / "product.npy", allow_pickle=False)  # Array::save
    assert saved.dtype == np.float32 and saved.shape == (64, 48) and (saved == product).all()
    assert "reloaded shape 64 x 48" in result.stdout  # Array::load
    assert (np.fromfile(tmp_path / "reloaded.f32", dtype=np.float32).reshape(64, 48) == product).all()

    # element access through the host copy: reads after device work, writes before it, shared by copies
    elements = np.fromfile(tmp_path / "elements.f32", dtype=np.float32)
    assert elements.tolist() == [6, 10, 20, 4, 6, 8, 40, 12, 40, 11, 21, 10, 2, 3, 4, 20, 6]
    x = np.fromfile(tmp_path / "solve.f64", dtype=np.float64)
    assert np.abs(x - [1.0, -2.0, 3.0]).max() < 1e-12
