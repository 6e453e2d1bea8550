"""Expressions too large for one chunk, run on the device as several chunks
(libsci_amd.array.stages, include/sci/hip.hpp split), against the oracle running the same stages.

Bars: stores bit for bit (the trees use only codes the GPU computes like the CPU); minimum and
maximum exact; sums within n * 2^-53 * sum |x| of the exact sum of the stored terms. The C++ mirror
must split the same tree into the same chunks as the Python mirror and give the same bits.
"""

from __future__ import annotations

import math
import os
import subprocess

import numpy as np
import pytest

import libsci_amd as sci
import oracle
from helpers import HostLeaf, deep_program, exact_sum, run_stages_on_oracle, same_bits, tree_of
from libsci_amd.array import _record, stages
from libsci_amd.chunk import Code, Fold

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("policy", [1, 2], ids=["interpreter", "specialised"])
def test_deep_tree_on_device(device, seed, policy):
    n = 300_001
    print(f"[test] split tree seed={seed} policy={policy}", flush=True)
    program = deep_program(200, 8, seed)
    hosts = [oracle.uniform(n, 10 + k, -2.0, 2.0) for k in range(8)]
    arrays = [sci.Array.from_numpy(h, device) for h in hosts]
    expression = tree_of(program, arrays)
    host_expression = tree_of(program, [HostLeaf(h) for h in hosts])
    want, planned = run_stages_on_oracle(host_expression)
    assert planned >= 2
    device.set_policy(policy)
    try:
        got = sci.evaluate(expression).host()
        assert same_bits(got, want).all(), np.flatnonzero(~same_bits(got, want))[:5]
        folds = {fold: sci.reduce_raw(expression, fold) for fold in Fold}
    finally:
        device.set_policy(0)
    terms = want.astype(np.float64)
    assert abs(folds[Fold.sum] - exact_sum(terms)) <= n * 2 ** -53 * float(np.abs(terms).sum())
    assert folds[Fold.minimum] == float(want.min()) and folds[Fold.maximum] == float(want.max())


def test_constant_chain_on_device(device):
    n = 100_003
    h = oracle.uniform(n, 3, -2.0, 2.0)
    values = [float(np.float32(0.37 * k - 11.0)) for k in range(1, 71)]
    expression = sci.Array.from_numpy(h, device)
    for v in values:
        expression = expression + v
    want = h.copy()
    for v in values:
        want = want + np.float32(v)
    got = sci.evaluate(expression).host()
    assert (got.view(np.uint32) == want.view(np.uint32)).all()


def describe(chunk) -> str:
    s = chunk.struct
    words = [str(s.length), str(s.arguments), str(s.literals), str(s.temporaries), str(s.result)]
    for i in range(s.length):
        ins = s.instructions[i]
        words.append(f"{ins.code}:{ins.target}:{ins.sources[0]}:{ins.sources[1]}:{ins.sources[2]}")
    words += [f"L{s.constants[i]:.17g}" for i in range(s.literals)]
    return " ".join(words)


def test_cpp_mirror_splits_alike(device, tmp_path):
    """tests/cpp/workload split: the C++ mirror's split of the same tree prints the same chunks as the
    Python planner's, and its store and folds equal the oracle's replay of those chunks."""
    n, inputs = 50_021, 8
    program = deep_program(200, inputs, 5)
    hosts = [oracle.uniform(n, 70 + k, -2.0, 2.0) for k in range(inputs)]
    for k, h in enumerate(hosts):
        h.tofile(tmp_path / f"input{k}.f32")
    with open(tmp_path / "program.txt", "w") as f:
        for kind, value in program:
            f.write(f"{kind} {float(value) if kind == 'const' else int(value)!r}\n")
    binary = os.path.join(HERE, "cpp", "workload")
    out = subprocess.run([binary, "split", str(tmp_path / "program.txt"), str(tmp_path), str(n), str(inputs)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    printed = [line.split(" ", 1) for line in out.stdout.strip().splitlines()]
    host_expression = tree_of(program, [HostLeaf(h) for h in hosts])
    expected = []
    for reduce in (False, True):
        planned, final = stages(host_expression, reduce)
        expected += [("reduce_stage" if reduce else "store_stage", describe(_record(s, False))) for s in planned]
        expected.append(("reduce_last" if reduce else "store_last", describe(_record(final, reduce))))
    assert [tuple(p) for p in printed] == expected
    want, _ = run_stages_on_oracle(host_expression)
    got = np.fromfile(tmp_path / "split.f32", dtype=np.float32)
    assert same_bits(got, want).all()
    folds = np.fromfile(tmp_path / "folds.f64", dtype=np.float64)
    terms = want.astype(np.float64)
    assert abs(folds[0] - exact_sum(terms)) <= n * 2 ** -53 * float(np.abs(terms).sum())
    assert folds[1] == float(want.min()) and folds[2] == float(want.max())
