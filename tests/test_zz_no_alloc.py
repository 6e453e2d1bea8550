"""scihip_reduce_static, the reduction that allocates nothing, against scihip_reduce.

Both run the same kernels over the same grid and fold the partials in the same order, so their
results must agree bit for bit (NaN against NaN). Each case makes two calls in a row without waiting
in between, the second over a smaller input, so the second call meets the first call's leftovers in
the static scratch. A graph captured around the call replays with the same result as the eager call.
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

C2_REDUCE = [("in", 0), ("op", Code.sin), ("in", 2), ("op", Code.cos), ("op", Code.multiply),
             ("in", 1), ("op", Code.tanh), ("op", Code.add)]
C2_DOT = [("in", 0), ("in", 2), ("op", Code.multiply)]
PLAIN = [("in", 0), ("in", 1), ("op", Code.multiply), ("const", 2.0), ("op", Code.add)]


def eager(device, chunk, fold, count):
    result = ctypes.c_double()
    _abi.check(_abi.lib().scihip_reduce(device.handle, ctypes.byref(chunk.struct), int(fold), count,
                                        ctypes.byref(result)), "scihip_reduce")
    return result.value


def static(device, chunk, fold, count, slot):
    _abi.check(_abi.lib().scihip_reduce_static(device.handle, ctypes.byref(chunk.struct), int(fold), count, slot),
               "scihip_reduce_static")


def same(a: float, b: float) -> bool:
    return (math.isnan(a) and math.isnan(b)) or a == b


def inputs(device, count, seed, precision):
    dtype = np.float32 if precision == Precision.single else np.float64
    hosts = [oracle.uniform(count, seed + k, -4.0, 4.0, dtype=dtype) for k in range(3)]
    return [sci.Array.from_numpy(h, device) for h in hosts]


@pytest.mark.parametrize("precision", [Precision.single, Precision.wide], ids=["f32", "f64"])
@pytest.mark.parametrize("policy", [1, 2], ids=["interpreter", "specialised"])
def test_static_reduce_matches_reduce(device, policy, precision):
    large, small = 3_000_017, 100_003
    print(f"[test] reduce_static policy={policy} precision={int(precision)}", flush=True)
    first = inputs(device, large, 10, precision)
    second = inputs(device, small, 20, precision)
    slots = sci.Array(device, (8,), np.float64)
    for program in (C2_REDUCE, C2_DOT, PLAIN):
        one = record_postfix(program, precision, "reduce", [a.pointer for a in first])
        two = record_postfix(program, precision, "reduce", [a.pointer for a in second])
        folds = [Fold.sum] if program is C2_DOT else list(Fold)
        if policy == 2:  # the static entry point runs a specialised kernel only once it is loaded
            for chunk in (one, two):
                _abi.check(_abi.lib().scihip_prepare(device.handle, ctypes.byref(chunk.struct)), "scihip_prepare")
        device.set_policy(policy)
        want = [(eager(device, one, f, large), eager(device, two, f, small)) for f in folds]
        for k, fold in enumerate(folds):
            static(device, one, fold, large, slots.pointer + 16 * k)
            static(device, two, fold, small, slots.pointer + 16 * k + 8)
        got = slots.host()
        device.set_policy(0)
        for k, fold in enumerate(folds):
            assert same(got[2 * k], want[k][0]), (program, fold, got[2 * k], want[k][0])
            assert same(got[2 * k + 1], want[k][1]), (program, fold, got[2 * k + 1], want[k][1])


def test_static_reduce_edges(device):
    a = sci.Array.from_numpy(oracle.uniform(1000, 3, -1.0, 1.0), device)
    chunk = record_postfix([("in", 0)], Precision.single, "reduce", [a.pointer])
    slots = sci.full((4,), 7.0, device, dtype=np.float64)
    for fold in Fold:  # the fold of nothing: 0 for a sum, NaN otherwise (device.cpp:266-268)
        static(device, chunk, fold, 0, slots.pointer + 8 * int(fold))
    got = slots.host()
    assert got[0] == 0.0 and math.isnan(got[1]) and math.isnan(got[2]) and got[3] == 7.0
    host_result = ctypes.c_double()
    rc = _abi.lib().scihip_reduce_static(device.handle, ctypes.byref(chunk.struct), 0, 1000, ctypes.byref(host_result))
    assert rc == _abi.DECLINED  # the result must be device memory
    rc = _abi.lib().scihip_reduce_static(device.handle, ctypes.byref(chunk.struct), 3, 1000, slots.pointer)
    assert rc == -1  # no such fold


def test_static_reduce_in_a_graph(device):
    """Captured into a HIP graph (where nothing may be allocated), the call replays with the eager
    result; the inputs change between replays and each replay is checked."""
    count = 1_000_003
    hosts = [oracle.uniform(count, 30 + k, -4.0, 4.0) for k in range(3)]
    arrays = [sci.Array.from_numpy(h, device) for h in hosts]
    slots = sci.Array(device, (2,), np.float64)
    chunk = record_postfix(C2_DOT, Precision.single, "reduce", [a.pointer for a in arrays])
    device.set_policy(1)
    device.begin_capture()
    try:
        static(device, chunk, Fold.sum, count, slots.pointer)
        static(device, chunk, Fold.maximum, count, slots.pointer + 8)
    finally:
        graph = device.end_capture()
    try:
        for replay in range(4):
            refreshed = oracle.uniform(count, 40 + replay, -4.0, 4.0)
            _abi.check(_abi.lib().scihip_upload(device.handle, arrays[0].pointer, refreshed.ctypes.data,
                                                refreshed.nbytes), "scihip_upload")
            want = (eager(device, chunk, Fold.sum, count), eager(device, chunk, Fold.maximum, count))
            graph.replay()
            got = slots.host()
            assert same(got[0], want[0]) and same(got[1], want[1]), (replay, got, want)
    finally:
        graph.release()
        device.set_policy(0)
