"""Device transpose and per-axis reductions.

Bars: the transpose is exact (a permutation) for every shape, float and double. Each row's (axis 1)
or column's (axis 0) fold agrees with the oracle's single-part reduction of that row alone
(Dispatch::reduce over the row, the way the reference reduces one contiguous array): minimum and
maximum exactly, NaN placement exactly, sums within 24 * 2^-24 * sum|terms| (float) of it, empty folds
0 or NaN. At the C2 2-D size (16384 x 16384 float32) the column folds of X equal the row folds of
the device transpose of X (min / max exactly, sums within the bound) and sampled rows agree with a
float64 sum.
"""

from __future__ import annotations

import numpy as np
import pytest

import libsci_amd as sci
import oracle
from libsci_amd.chunk import Code, Fold, Precision, record_postfix

pytestmark = pytest.mark.gpu

PLAIN = [("in", 0), ("op", Code.copy)]
FUSED = [("in", 0), ("op", Code.sin), ("in", 1), ("op", Code.multiply), ("const", 1.0), ("op", Code.add)]
DOT = [("in", 0), ("in", 1), ("op", Code.multiply)]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(64, 64), (100, 37), (1, 1000), (1000, 1), (129, 257), (2048, 4096)], ids=str)
def test_transpose_is_exact(device, dtype, shape):
    x = np.random.default_rng(shape[0] + shape[1]).uniform(-4, 4, shape).astype(dtype)
    t = sci.transpose(sci.Array.from_numpy(x, device)).host()
    assert t.shape == shape[::-1]
    assert np.array_equal(t, np.ascontiguousarray(x.T))


def oracle_rows(program, matrices, fold, precision):
    """The oracle's fold of each row alone (one part, as a single contiguous reduction)."""
    rows = matrices[0].shape[0]
    out = np.empty(rows, dtype=matrices[0].dtype)
    for r in range(rows):
        views = [np.ascontiguousarray(m[r]) for m in matrices]
        chunk = record_postfix(program, precision, "reduce", [v.ctypes.data for v in views])
        out[r] = oracle.reduce(chunk, int(fold), views[0].size, threads=1)
    return out


def magnitudes(program, matrices, precision):
    """Per row: the sum of |terms| the fold adds (the oracle's elementwise results, or products)."""
    if program is DOT:
        return np.abs(matrices[0].astype(np.float64) * matrices[1].astype(np.float64)).sum(axis=1)
    flat = [np.ascontiguousarray(m.reshape(-1)) for m in matrices]
    chunk = record_postfix(program, precision, "output", [f.ctypes.data for f in flat])
    terms = oracle.execute(chunk, flat[0].size, dtype=matrices[0].dtype).reshape(matrices[0].shape)
    return np.nansum(np.abs(terms.astype(np.float64)), axis=1)


def expression_of(program, arrays):
    if program is PLAIN:
        return sci.copy(arrays[0])
    if program is FUSED:
        return sci.sin(arrays[0]) * arrays[1] + 1.0
    return arrays[0] * arrays[1]


@pytest.mark.parametrize("policy", [1, 2], ids=["interpreter", "specialised"])
@pytest.mark.parametrize("name", ["plain", "fused", "dot"])
@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("shape", [(300, 257), (64, 1000), (1000, 64), (1, 5), (5, 1)], ids=str)
def test_axis_folds_match_oracle(device, name, axis, shape, policy):
    program = {"plain": PLAIN, "fused": FUSED, "dot": DOT}[name]
    rng = np.random.default_rng(shape[0] * 7 + shape[1] + axis)
    matrices = [rng.uniform(-4, 4, shape).astype(np.float32) for _ in range(2)]
    if shape[0] > 2 and shape[1] > 2:
        matrices[0][1, 2] = np.nan  # one NaN: its row and its column fold to NaN
    arrays = [sci.Array.from_numpy(m, device) for m in matrices]
    folds = [Fold.sum] if program is DOT else [Fold.sum, Fold.minimum, Fold.maximum]
    for fold in folds:
        device.set_policy(policy)
        try:
            got = sci.reduce_axis(expression_of(program, arrays), fold, axis).host()
        finally:
            device.set_policy(0)
        lined = [np.ascontiguousarray(m.T) if axis == 0 else m for m in matrices]
        want = oracle_rows(program, lined, fold, Precision.single)
        assert np.array_equal(np.isnan(got), np.isnan(want)), (fold, axis)
        finite = ~np.isnan(want)
        if fold == Fold.sum:
            bound = 24 * 2.0 ** -24 * magnitudes(program, lined, Precision.single)[finite] + 1e-30
            assert (np.abs(got[finite].astype(np.float64) - want[finite]) <= bound).all(), (fold, axis)
        else:
            assert np.array_equal(got[finite], want[finite]), (fold, axis)


def test_empty_axes(device):
    empty_rows = sci.Array(device, (0, 7), np.float32)
    empty_columns = sci.Array(device, (7, 0), np.float32)
    assert sci.reduce_axis(sci.copy(empty_rows), Fold.sum, 1).shape == (0,)
    assert (sci.reduce_axis(sci.copy(empty_rows), Fold.sum, 0).host() == 0).all()
    assert np.isnan(sci.reduce_axis(sci.copy(empty_columns), Fold.maximum, 1).host()).all()


def test_c2_two_dimensional_full_size(device):
    n = 16384
    print(f"[test] C2 2-D {n} x {n}", flush=True)
    x = sci.uniform(n * n, device, 61, -4.0, 4.0).reshape(n, n)
    t = sci.transpose(x)
    for fold in (Fold.sum, Fold.minimum, Fold.maximum):
        columns = sci.reduce_axis(sci.copy(x), fold, 0).host()
        rows_of_t = sci.reduce_axis(sci.copy(t), fold, 1).host()
        if fold == Fold.sum:
            assert np.abs(columns.astype(np.float64) - rows_of_t).max() <= 4 * 2.0 ** -24 * 4.0 * n
        else:
            assert np.array_equal(columns, rows_of_t)
    sample = np.random.default_rng(2).choice(n, 16, replace=False)
    host_t = t.host()
    host_x = x.host()
    assert np.array_equal(host_t[sample], host_x[:, sample].T)
    sums = sci.reduce_axis(sci.copy(x), Fold.sum, 1).host()
    for r in sample:
        exact = host_x[r].astype(np.float64).sum()
        assert abs(float(sums[r]) - exact) <= 4 * 2.0 ** -24 * np.abs(host_x[r]).astype(np.float64).sum() + 1e-3
