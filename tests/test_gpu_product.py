"""Matrix products on the matrix cores against a float64 product (the oracle for products).

Bars, per mode, on entries in [-1, 1]: max |C - C64| <= K * u * max-row-col magnitude bound with
u = 2^-23 (fp32 accumulation of fp32 products), 2^-8 (bf16 inputs), 2^-16 (bf16x3), 2^-52 (fp64),
and the reference test's 37 x 53 x 29 integer-valued KAT within 1e-3
(tests/linear/test_product.cpp:8-33).
"""

from __future__ import annotations

import numpy as np
import pytest

import libsci_amd as sci
import oracle
from libsci_amd import _abi

pytestmark = pytest.mark.gpu

MODES = {"fp32": 0, "bf16": 1, "bf16x3": 2, "fp64": 3}
UNIT = {"fp32": 2.0 ** -23, "bf16": 2.0 ** -8, "bf16x3": 2.0 ** -16, "fp64": 2.0 ** -52}


def gpu_product(device, left, right, mode):
    dtype = np.float64 if mode == "fp64" else np.float32
    a = sci.Array.from_numpy(left.astype(dtype), device)
    b = sci.Array.from_numpy(right.astype(dtype), device)
    c = sci.Array(device, (left.shape[0], right.shape[1]), dtype)
    _abi.check(_abi.lib().scihip_product(device.handle, MODES[mode], a.pointer, b.pointer, c.pointer,
                                         left.shape[0], right.shape[1], left.shape[1]), "scihip_product")
    return c.host()


@pytest.mark.parametrize("mode", list(MODES))
def test_reference_kat(device, mode):
    rows, inner, columns = 37, 53, 29
    left = np.array([[((r * 7 + c * 3) % 11) - 5.0 for c in range(inner)] for r in range(rows)])
    right = np.array([[((r * 5 + c) % 13) * 0.25 for c in range(columns)] for r in range(inner)])
    got = gpu_product(device, left, right, mode)
    want = left @ right
    assert got.shape == (rows, columns)
    assert np.abs(got - want).max() < 1e-3  # the entries are exact in bf16, so every mode is exact here


# (64, 4096, 64) and (200, 3000, 130) have few output tiles and a long shared dimension, so their
# shared dimension is split across workgroups and the slices folded afterwards.
@pytest.mark.parametrize("shape", [(1, 1, 1), (5, 3, 7), (129, 257, 130), (300, 64, 200), (512, 512, 512),
                                   (64, 4096, 64), (200, 3000, 130)])
@pytest.mark.parametrize("mode", list(MODES))
def test_against_float64(device, mode, shape):
    m, k, n = shape
    print(f"[test] product {mode} {shape}", flush=True)
    left = oracle.uniform(m * k, 1, -1.0, 1.0, dtype=np.float64).reshape(m, k)
    right = oracle.uniform(k * n, 2, -1.0, 1.0, dtype=np.float64).reshape(k, n)
    if mode != "fp64":
        left, right = left.astype(np.float32), right.astype(np.float32)
    got = gpu_product(device, left, right, mode)
    want = left.astype(np.float64) @ right.astype(np.float64)
    bound = 2 * k * UNIT[mode] * (np.abs(left).astype(np.float64) @ np.abs(right).astype(np.float64))
    assert (np.abs(got - want) <= bound + 1e-300).all(), np.abs(got - want).max()


@pytest.mark.parametrize("mode", list(MODES))
def test_split_depth_is_deterministic(device, mode):
    """A split product folds its slices in a fixed order: repeated calls agree bit for bit."""
    left = oracle.uniform(96 * 8192, 5, -1.0, 1.0, dtype=np.float64).reshape(96, 8192)
    right = oracle.uniform(8192 * 40, 6, -1.0, 1.0, dtype=np.float64).reshape(8192, 40)
    first = gpu_product(device, left, right, mode)
    for _ in range(3):
        assert np.array_equal(gpu_product(device, left, right, mode), first)


def test_zero_depth_writes_zeros(device):
    c = sci.full((4, 5), 7.0, device)
    a = sci.Array(device, (4, 1), np.float32)
    assert _abi.lib().scihip_product(device.handle, 0, a.pointer, a.pointer, c.pointer, 4, 5, 0) == 0
    assert (c.host() == 0).all()


def test_large_fp32_and_split(device):
    n = 2048
    left = oracle.uniform(n * n, 3, -1.0, 1.0).reshape(n, n)
    right = oracle.uniform(n * n, 4, -1.0, 1.0).reshape(n, n)
    want = left.astype(np.float64) @ right.astype(np.float64)
    for mode in ("fp32", "bf16x3", "bf16"):
        got = gpu_product(device, left, right, mode)
        relative = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"[test] 2048^3 {mode}: relative Frobenius error {relative:.3e}", flush=True)
        assert relative < {"fp32": 1e-6, "bf16x3": 1e-4, "bf16": 1e-2}[mode]


@pytest.mark.parametrize("shape", [(8192, 8192, 8192), (4096, 16384, 1024)], ids=["8192^3", "4096x16384x1024"])
def test_metric_sizes(device, shape):
    """The products the measurement names: 8192^3 and 4096 x 16384 times 16384 x 1024, entries seeded
    uniform[-1, 1] (generated on the device; the oracle's generator gives the same bits on the host).
    Against a float64 product of 64 seeded rows: every entry within the per-mode bound
    2k u (|A||B|), and the relative Frobenius error of those rows stated and bounded."""
    m, k, n = shape
    print(f"[test] product at {shape}", flush=True)
    left = sci.uniform(m * k, device, 31, -1.0, 1.0).reshape(m, k)
    right = sci.uniform(k * n, device, 32, -1.0, 1.0).reshape(k, n)
    output = sci.Array(device, (m, n), np.float32)
    left_h = oracle.uniform(m * k, 31, -1.0, 1.0).reshape(m, k)
    right_h = oracle.uniform(k * n, 32, -1.0, 1.0).reshape(k, n)
    rows = np.sort(np.random.default_rng(m + n).choice(m, 64, replace=False))
    assert np.array_equal(left.host()[rows], left_h[rows])
    right64 = right_h.astype(np.float64)
    want = left_h[rows].astype(np.float64) @ right64
    magnitude = np.abs(left_h[rows]).astype(np.float64) @ np.abs(right64)
    frobenius = {"fp32": 1e-5, "bf16x3": 1e-4, "bf16": 1e-2}
    for mode in ("fp32", "bf16x3", "bf16"):
        _abi.check(_abi.lib().scihip_product(device.handle, MODES[mode], left.pointer, right.pointer,
                                             output.pointer, m, n, k), "scihip_product")
        got = output.host()[rows].astype(np.float64)
        error = np.abs(got - want)
        relative = float(np.linalg.norm(got - want) / np.linalg.norm(want))
        print(f"[test] {shape} {mode}: max abs error {error.max():.3e}, relative Frobenius {relative:.3e}",
              flush=True)
        assert (error <= 2 * k * UNIT[mode] * magnitude).all()
        assert relative < frobenius[mode]


@pytest.mark.parametrize("shape", [(4100, 1000, 4200), (4100, 1003, 4200), (4352, 576, 4097), (2000, 8200, 2050)],
                         ids=["4100x1000x4200", "4100x1003x4200", "4352x576x4097", "2000x8200x2050"])
@pytest.mark.parametrize("mode", ["bf16", "bf16x3"])
def test_large_tiles_ragged(device, shape, mode):
    """bf16 and bf16x3 products with at least 256 output tiles of 256 x 256 take the large-tile
    kernel; these have partial last tiles in both directions and a shared dimension padded up to a
    multiple of 64 (1000 and 1003 -> 1024, 8200 -> 8256) or not (576); 1003 also ends inside a
    16-byte chunk of the packed rows. 2000 x 2050 has 72 tiles, so its 258 stages are split four
    ways (uneven slices) and the partial products folded. Sampled rows, the last rows among them,
    against float64."""
    m, k, n = shape
    print(f"[test] {mode} large tiles at {shape}", flush=True)
    left = sci.uniform(m * k, device, 41, -1.0, 1.0).reshape(m, k)
    right = sci.uniform(k * n, device, 42, -1.0, 1.0).reshape(k, n)
    output = sci.Array(device, (m, n), np.float32)
    left_h = oracle.uniform(m * k, 41, -1.0, 1.0).reshape(m, k)
    right_h = oracle.uniform(k * n, 42, -1.0, 1.0).reshape(k, n)
    chosen = np.random.default_rng(m + n).choice(m, 40, replace=False)
    rows = np.unique(np.concatenate([chosen, [0, 255, 256, m - 257, m - 256, m - 5, m - 1]]))
    assert np.array_equal(left.host()[rows], left_h[rows])
    right64 = right_h.astype(np.float64)
    want = left_h[rows].astype(np.float64) @ right64
    magnitude = np.abs(left_h[rows]).astype(np.float64) @ np.abs(right64)
    _abi.check(_abi.lib().scihip_product(device.handle, MODES[mode], left.pointer, right.pointer,
                                         output.pointer, m, n, k), "scihip_product")
    got = output.host()[rows].astype(np.float64)
    assert (np.abs(got - want) <= 2 * k * UNIT[mode] * magnitude).all()
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < {"bf16": 1e-2, "bf16x3": 1e-4}[mode]
