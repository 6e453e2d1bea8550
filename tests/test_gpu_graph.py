"""Graphs of repeated device work (scihip_capture_begin / scihip_capture_end / scihip_replay).

A graph recording two dependent chunk stores, a product and a fused adjoint without a loss is
replayed three times, with new input values uploaded eagerly between the replays; every replay must
give the eager path's results bit for bit (the same kernels on the same data). While a capture is
open, calls that wait for the device or return a value are refused, and the capture still ends
cleanly.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import libsci_amd as sci
from libsci_amd import _abi

pytestmark = pytest.mark.gpu


def same(got, want):
    return np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("policy", [1, 2], ids=["interpreter", "specialised"])
def test_replays_match_eager_runs(device, policy):
    n, m = 100_003, 96
    rng = np.random.default_rng(policy)
    a, b, c = (sci.Array.from_numpy(rng.uniform(-2, 2, n).astype(np.float32), device) for _ in range(3))
    left = sci.Array.from_numpy(rng.uniform(-1, 1, (m, m)).astype(np.float32), device)
    right = sci.Array.from_numpy(rng.uniform(-1, 1, (m, m)).astype(np.float32), device)
    first, second = sci.Array(device, (n,), np.float32), sci.Array(device, (n,), np.float32)
    product = sci.Array(device, (m, m), np.float32)
    gradients = [sci.Array(device, (n,), np.float32) for _ in range(3)]
    stage1 = a * b + 2.0
    stage2 = sci.sin(first) * c
    loss_expression = sci.sinh(a * b + c)
    adjoint_chunk = sci.record(loss_expression, reduce=True)
    pointers = (ctypes.c_void_p * 3)(*[g.pointer for g in gradients])
    device.set_policy(policy)
    try:
        for e in (stage1, stage2):
            device.prepare(e)
        device.begin_capture()
        sci.assign(first, stage1)
        sci.assign(second, stage2)  # reads what the first store wrote
        _abi.check(_abi.lib().scihip_product(device.handle, 0, left.pointer, right.pointer, product.pointer, m, m, m))
        _abi.check(_abi.lib().scihip_adjoint(device.handle, ctypes.byref(adjoint_chunk.struct), None, pointers, n, None))
        graph = device.end_capture()
        for replay in range(3):
            print(f"[test] graph replay {replay} policy={policy}", flush=True)
            for array in (a, b, c):  # new values, uploaded eagerly between replays
                _abi.check(_abi.lib().scihip_upload(device.handle, array.pointer,
                                                    rng.uniform(-2, 2, n).astype(np.float32).ctypes.data, 4 * n))
            graph.replay()
            device.synchronize()
            want_first = sci.evaluate(stage1).host()
            assert same(first.host(), want_first), replay
            assert same(second.host(), sci.evaluate(sci.sin(sci.Array.from_numpy(want_first, device)) * c).host())
            assert same(product.host(), sci.product(left, right).host())
            _, want = sci.gradient(loss_expression, [a, b, c], loss=False)
            for g, w in zip(gradients, want):
                assert same(g.host(), w.host()), replay
        graph.release()
    finally:
        device.set_policy(0)


def test_capture_refuses_waiting_calls(device):
    a = sci.Array.from_numpy(np.ones(1000, np.float32), device)
    device.begin_capture()
    with pytest.raises(sci.BackendError):
        sci.reduce_raw(a * 2.0, sci.Fold.sum)
    with pytest.raises(sci.BackendError):
        device.synchronize()
    out = sci.Array(device, (1000,), np.float32)
    sci.assign(out, a * 3.0)
    graph = device.end_capture()
    graph.replay()
    device.synchronize()
    assert (out.host() == 3.0).all()
    graph.release()
