"""Expressions too large for one chunk (64 instructions, inputs, constants or temporaries,
include/core/bytecode/chunk.hpp:31), split into chunks run one after another
(libsci_amd.array.stages; include/sci/hip.hpp's split follows the same rule). The reference runs
such an expression element by element on the host instead (include/array/expression.hpp:178-184).

CPU: the planner's structure, and its semantics through the oracle: each stage replayed as a chunk
over host arrays. Chains of additions, whose float32 result is fixed by IEEE rounding in order,
must come out bit for bit what numpy's sequential float32 additions give; seeded 200-operation trees
must agree with a float64 evaluation of the same tree to float32 rounding.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from helpers import HostLeaf, deep_program, float64_of, run_stages_on_oracle, tree_of
from libsci_amd.array import Literal, Node, Staged, _record, stages
from libsci_amd.chunk import Code, Fold


def leaves(count, n=4099, seed=50):
    hosts = [oracle.uniform(n, seed + k, -2.0, 2.0) for k in range(count)]
    return hosts, [HostLeaf(h) for h in hosts]


def test_a_chunk_that_fits_is_not_split():
    hosts, (a, b) = leaves(2)
    expression = Node(Code.add, [Node(Code.multiply, [a, b]), Literal(2.0, np.float32)])
    planned, final = stages(expression)
    assert planned == [] and final is expression


def test_constant_chain():
    """a + c1 + ... + c70 with 70 distinct constants: the chain of the first 64 becomes a stage (64
    instructions and constants), the last six additions the final chunk."""
    hosts, (a,) = leaves(1)
    values = [float(np.float32(0.37 * k - 11.0)) for k in range(1, 71)]
    expression = a
    for v in values:
        expression = Node(Code.add, [expression, Literal(v, np.float32)])
    assert not _record(expression, False).valid()
    planned, final = stages(expression)
    assert len(planned) == 1 and _record(planned[0], False).length == 64
    assert _record(final, False).length == 6 and _record(final, False).literals == 6
    got, _ = run_stages_on_oracle(expression)
    want = hosts[0].copy()
    for v in values:
        want = want + np.float32(v)  # float32 additions, in order
    assert (got.view(np.uint32) == want.view(np.uint32)).all()


def test_input_chain():
    """x0 + x1 + ... + x69 over 70 distinct arrays: 65 inputs overflow the chunk, so the chain of the
    first 64 becomes a stage; the sum is bit for bit numpy's sequential float32 sum."""
    hosts, xs = leaves(70, n=1031)
    expression = xs[0]
    for x in xs[1:]:
        expression = Node(Code.add, [expression, x])
    planned, final = stages(expression)
    assert len(planned) == 1 and _record(planned[0], False).arguments == 64
    assert _record(final, False).arguments == 7
    got, _ = run_stages_on_oracle(expression)
    want = hosts[0].copy()
    for h in hosts[1:]:
        want = want + h
    assert (got.view(np.uint32) == want.view(np.uint32)).all()


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_deep_trees(seed):
    """A seeded 200-operation tree over 8 arrays: every stage and the final chunk fit, each stage is a
    store the final chunk reads, and the replay agrees with float64 to float32 rounding."""
    program = deep_program(200, 8, seed)
    hosts, xs = leaves(8, n=20_011)
    expression = tree_of(program, xs)
    assert not _record(expression, False).valid()
    for reduce in (False, True):
        planned, final = stages(expression, reduce)
        assert len(planned) >= 2
        for stage in planned:
            assert _record(stage, False).valid()
        assert _record(final, reduce).valid()
        assert any(isinstance(leaf, Staged) for leaf in final.children) or isinstance(final, Node)
    got, count = run_stages_on_oracle(expression)
    want = float64_of(program, hosts)
    assert want.std() > 0.1  # the result depends on the inputs
    assert np.abs(got - want).max() <= 1e-5 * (1 + np.abs(want).max())
    total, _ = run_stages_on_oracle(expression, Fold.sum)
    assert abs(total - float(want.sum())) <= 1e-5 * float(np.abs(want).sum())
    low, _ = run_stages_on_oracle(expression, Fold.minimum)
    assert low == float(got.min())
