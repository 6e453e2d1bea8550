"""Parity of the HIP reductions with the CPU oracle, through the C ABI.

Bar: minimum and maximum equal the oracle's exactly; sums within a stated reordering bound of the
exact (math.fsum) sum of the same float terms: the GPU accumulates in double, so its error is at
most n * 2^-53 * sum |x_i|, and the CPU's float blocks at most 1024 * 2^-24 * sum |x_i|. Where a
chunk uses a code the CPU sends to glibc (tanh, sinh, ...), each GPU term may differ from the CPU
term by that code's ulp bound, so the sum gets LIBRARY_ULPS * 2^-24 * sum (|x_i| + 1) more. A
contracted sum (a dot product) folds the exact products of its float operands. NaN anywhere gives
NaN; the minimum and maximum of nothing are NaN and the sum of nothing 0
(src/core/drivers/device.cpp:266-268).
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

import libsci_amd as sci
import oracle
from libsci_amd import _abi
from libsci_amd.chunk import Code, Fold, Precision, record_postfix

pytestmark = pytest.mark.gpu

LIBRARY_ULPS = 4  # the measured maximum over every glibc-backed code (tests/test_gpu_chunks.py)

C1_REDUCE = [("in", 0), ("op", Code.sin)]
C2_REDUCE = [("in", 0), ("op", Code.sin), ("in", 2), ("op", Code.cos), ("op", Code.multiply),
             ("in", 1), ("op", Code.tanh), ("op", Code.add)]
C2_DOT = [("in", 0), ("in", 2), ("op", Code.multiply)]
C5_LOSS = [("in", 0), ("in", 1), ("op", Code.multiply), ("in", 2), ("op", Code.add), ("op", Code.sinh),
           ("in", 3), ("in", 3), ("op", Code.multiply), ("op", Code.negate), ("op", Code.exponential),
           ("op", Code.multiply)]


def gpu_reduce(device, program, device_arrays, fold, policy=0):
    device.set_policy(policy)
    chunk = record_postfix(program, Precision.single, "reduce", [a.pointer for a in device_arrays])
    result = ctypes.c_double()
    _abi.check(_abi.lib().scihip_reduce(device.handle, ctypes.byref(chunk.struct), int(fold),
                                        device_arrays[0].size, ctypes.byref(result)))
    device.set_policy(0)
    return result.value


def terms(program, arrays):
    """The float terms a (non-contracted) reduction folds: the chunk's result per element, from the oracle."""
    stored = record_postfix(program, Precision.single, "output", [a.ctypes.data for a in arrays])
    return oracle.execute(stored, arrays[0].size, threads=8)


@pytest.mark.parametrize("count", [1, 5, 1023, 4096, 300_001, (1 << 22) + 3])
@pytest.mark.parametrize("policy", [1, 2], ids=["interpreter", "specialised"])
def test_reductions_match_oracle(device, count, policy):
    print(f"[test] reductions count={count} policy={policy}", flush=True)
    arrays = [oracle.uniform(count, 100 + i, -4.0, 4.0) for i in range(3)]
    dev = [sci.Array.from_numpy(a, device) for a in arrays]
    for program, library in ((C1_REDUCE, False), (C2_REDUCE, True)):
        x = terms(program, arrays).astype(np.float64)
        magnitude = float(np.abs(x).sum())
        exact = math.fsum(x)
        got = gpu_reduce(device, program, dev, Fold.sum, policy)
        slack = LIBRARY_ULPS * 2 ** -24 * (magnitude + count) if library else 0.0
        assert abs(got - exact) <= count * 2 ** -53 * magnitude + slack
        chunk = record_postfix(program, Precision.single, "reduce", [a.ctypes.data for a in arrays])
        cpu = oracle.reduce(chunk, Fold.sum, count, threads=8)
        assert abs(cpu - exact) <= 1024 * 2 ** -24 * magnitude
        for fold in (Fold.minimum, Fold.maximum):
            got = gpu_reduce(device, program, dev, fold, policy)
            want = oracle.reduce(chunk, fold, count, threads=8)
            if library:
                assert abs(got - want) <= LIBRARY_ULPS * 2 ** -24 * (abs(want) + 1)
            else:
                assert got == want


@pytest.mark.parametrize("policy", [1, 2], ids=["interpreter", "specialised"])
def test_contracted_sums(device, policy):
    """sum(a * c) and the C5 loss are dot products: exact float products summed in double."""
    count = 1_000_003
    arrays = [oracle.uniform(count, 200 + i, -2.0, 2.0) for i in range(4)]
    dev = [sci.Array.from_numpy(a, device) for a in arrays]
    products = arrays[0].astype(np.float64) * arrays[2].astype(np.float64)
    exact = math.fsum(products)
    got = gpu_reduce(device, C2_DOT, dev, Fold.sum, policy)
    assert abs(got - exact) <= count * 2 ** -53 * float(np.abs(products).sum())
    # C5: sum(sinh(a*b+c) * exp(-d*d)); the two factors come from the oracle, the product is exact in double
    left = oracle.execute(record_postfix(C5_LOSS[:6], Precision.single, "output",
                                         [a.ctypes.data for a in arrays]), count, threads=8).astype(np.float64)
    right = oracle.execute(record_postfix([("in", 0), ("in", 0), ("op", Code.multiply), ("op", Code.negate),
                                           ("op", Code.exponential)], Precision.single, "output",
                                          [arrays[3].ctypes.data]), count, threads=8).astype(np.float64)
    exact = math.fsum(left * right)
    got = gpu_reduce(device, C5_LOSS, dev, Fold.sum, policy)
    # sinh goes to glibc on the CPU and the device library on the GPU: a few ulps per term
    magnitude = float(np.abs(left * right).sum())
    assert abs(got - exact) <= count * 2 ** -53 * magnitude + LIBRARY_ULPS * 2 ** -24 * magnitude
    cpu = oracle.reduce(record_postfix(C5_LOSS, Precision.single, "reduce", [a.ctypes.data for a in arrays]),
                        Fold.sum, count, threads=8)
    assert abs(cpu - exact) <= 1024 * 2 ** -24 * float(np.abs(left * right).sum())


def test_nan_wins_and_empty(device):
    a = oracle.uniform(100_000, 1, -1.0, 1.0)
    a[77_777] = np.nan
    dev = sci.Array.from_numpy(a, device)
    for fold in Fold:
        for policy in (1, 2):
            assert math.isnan(gpu_reduce(device, [("in", 0)], [dev], fold, policy))
    chunk = record_postfix([("in", 0)], Precision.single, "reduce", [dev.pointer])
    result = ctypes.c_double(7.0)
    assert _abi.lib().scihip_reduce(device.handle, ctypes.byref(chunk.struct), 0, 0, ctypes.byref(result)) == 0
    assert result.value == 0.0
    assert _abi.lib().scihip_reduce(device.handle, ctypes.byref(chunk.struct), 1, 0, ctypes.byref(result)) == 0
    assert math.isnan(result.value)
    # infinities without NaN: min / max see them; +inf + -inf sums to NaN as the arithmetic says
    b = np.array([1.0, np.inf, -np.inf, 2.0], dtype=np.float32)
    db = sci.Array.from_numpy(b, device)
    assert gpu_reduce(device, [("in", 0)], [db], Fold.minimum) == -np.inf
    assert gpu_reduce(device, [("in", 0)], [db], Fold.maximum) == np.inf
    assert math.isnan(gpu_reduce(device, [("in", 0)], [db], Fold.sum))


@pytest.mark.parametrize("policy", [1, 2], ids=["interpreter", "specialised"])
def test_one_workgroup_reductions(device, policy):
    """Reductions small enough for one workgroup leave their partial in pinned host memory and the
    host folds it (no fold kernel): the same bits as the oracle, signed zeros and NaN included, and as
    the static variant, whose fold runs on the device."""
    cases = [[-0.0], [-0.0, -0.0], [0.0, -0.0], [np.nan], [np.inf], [-np.inf], [5.0, -3.0, 0.25],
             [np.nan, 1.0, np.nan], [-np.inf, np.inf]]
    result_device = sci.Array(device, (1,), np.float64)
    for values in cases:
        a = np.array(values, dtype=np.float32)
        dev = sci.Array.from_numpy(a, device)
        chunk = record_postfix([("in", 0)], Precision.single, "reduce", [a.ctypes.data])
        for fold in Fold:
            got = gpu_reduce(device, [("in", 0)], [dev], fold, policy)
            want = oracle.reduce(chunk, fold, a.size, threads=1)
            assert np.array_equal(np.float64(got), np.float64(want), equal_nan=True), (values, fold, got, want)
            assert math.copysign(1.0, got) == math.copysign(1.0, want) or math.isnan(want), (values, fold, got, want)
            device.set_policy(policy)
            static = record_postfix([("in", 0)], Precision.single, "reduce", [dev.pointer])
            _abi.check(_abi.lib().scihip_reduce_static(device.handle, ctypes.byref(static.struct), int(fold), a.size,
                                                       ctypes.c_void_p(result_device.pointer)), "scihip_reduce_static")
            device.synchronize()
            device.set_policy(0)
            on_device = float(result_device.host()[0])
            assert np.array_equal(np.float64(got), np.float64(on_device), equal_nan=True), (values, fold, got, on_device)
            assert math.isnan(got) or math.copysign(1.0, got) == math.copysign(1.0, on_device)


def test_front_end_reductions(device):
    """tests/operations/test_reduce.cpp:20-27 on the device."""
    count = 1_000_000
    ramp_h = oracle.linspace(-1.0, 2.0, count)
    ramp = sci.Array.from_numpy(ramp_h, device)
    total = math.fsum(ramp_h.astype(np.float64))
    assert abs(float(sci.sum(ramp)) - total) < 1e-3 * count
    assert abs(float(sci.mean(ramp)) - total / count) < 1e-3
    assert sci.minimum(ramp) == ramp_h.min() and sci.maximum(ramp) == ramp_h.max()
    squared = sci.evaluate(ramp * ramp)
    assert abs(float(sci.sum(ramp * ramp)) - float(sci.sum(squared))) < 1e-2 * count
    assert sci.mean(ramp) == np.float32(sci.sum(ramp) / np.float32(count))
    ones = sci.ones((4_000_003,), device)
    for _ in range(5):  # tests/core/drivers/processor/test_pool.cpp:10-19: the doubled sum is exact
        assert float(sci.sum(ones + ones)) == 8_000_006.0


def test_c2_reduce_full_size(device):
    """N = 2^28: sum / min / max / mean of sin(a)*cos(c) + tanh(b) and the contracted sum of
    sin(a)*cos(c), with and without b's NaNs. Every link is checked against something independent:
    sin and cos against the oracle bit for bit, tanh (glibc on the CPU, the device library on the GPU)
    against glibc within LIBRARY_ULPS, the terms the GPU stores bit for bit against the same float
    operations on those values, and each GPU fold against the exact sum (or the exact extreme) of
    the terms with the tight double-accumulation bound n * 2^-53 * sum |x|, below 1% of |sum| (0.07% for
    the sum, 0.7% for the dot, whose terms nearly cancel)."""
    from helpers import exact_sum, ulps32

    n = 1 << 28
    print("[test] c2 reduce at 2^28: generating on the device", flush=True)
    a = sci.uniform(n, device, 1, -4.0, 4.0)
    c = sci.uniform(n, device, 3, -4.0, 4.0)
    b_nan = sci.uniform(n, device, 2, -4.0, 4.0, nan_seed=9, nan_fraction=0.01)
    b = sci.uniform(n, device, 2, -4.0, 4.0)
    expr_nan = sci.sin(a) * sci.cos(c) + sci.tanh(b_nan)
    for fold in Fold:
        assert math.isnan(sci.reduce_raw(expr_nan, fold))
    expr = sci.sin(a) * sci.cos(c) + sci.tanh(b)
    got = {fold: sci.reduce_raw(expr, fold) for fold in Fold}
    dot = sci.reduce_raw(sci.sin(a) * sci.cos(c), Fold.sum)
    product_extremes = [sci.reduce_raw(sci.sin(a) * sci.cos(c), fold) for fold in (Fold.minimum, Fold.maximum)]
    terms = sci.evaluate(expr).host()
    tanh_gpu = sci.evaluate(sci.tanh(b)).host()
    print("[test] c2 reduce at 2^28: oracle", flush=True)
    ah = oracle.uniform(n, 1, -4.0, 4.0)
    bh = oracle.uniform(n, 2, -4.0, 4.0)
    ch = oracle.uniform(n, 3, -4.0, 4.0)
    sines = oracle.map_f32(Code.sin, ah, threads=16)
    cosines = oracle.map_f32(Code.cos, ch, threads=16)
    tanh_cpu = oracle.map_f32(Code.tanh, bh, threads=16)
    tanh_ulps = int(ulps32(tanh_gpu, tanh_cpu).max())
    print(f"[test] tanh over 2^28 inputs in [-4, 4]: max {tanh_ulps} ulp between the GPU and glibc", flush=True)
    assert tanh_ulps <= LIBRARY_ULPS
    products = sines * cosines  # float32 multiply, correctly rounded, as the chunk's multiply
    assert (terms.view(np.uint32) == (products + tanh_gpu).view(np.uint32)).all()
    wide = terms.astype(np.float64)
    bound = n * 2 ** -53 * float(np.abs(wide).sum())
    exact = exact_sum(wide)
    print(f"[test] sum {got[Fold.sum]!r} exact {exact!r} bound {bound:.3g}", flush=True)
    assert abs(got[Fold.sum] - exact) <= bound and bound < 1e-2 * abs(exact)
    assert got[Fold.minimum] == float(terms.min()) and got[Fold.maximum] == float(terms.max())
    assert sci.mean(expr) == np.float32(np.float32(got[Fold.sum]) / np.float32(n))
    # no library code: sin(a)*cos(c) summed as exact products of the oracle's own sines and cosines
    exact_products = sines.astype(np.float64) * cosines.astype(np.float64)
    bound = n * 2 ** -53 * float(np.abs(exact_products).sum())
    exact = exact_sum(exact_products)
    print(f"[test] dot {dot!r} exact {exact!r} bound {bound:.3g}", flush=True)
    assert abs(dot - exact) <= bound and bound < 1e-2 * abs(exact)
    assert product_extremes == [float(products.min()), float(products.max())]
