"""Element access on device arrays (include/array/array.hpp:186-217 reads and writes device memory
the host can map; MI355X HBM is not host-mapped, so libsci_amd reads and writes a host copy of the
allocation, taken at the first access after device work and handed back at the next device use).

Each case checks what the reference's semantics fix: a read after device work sees that work's
result, a write reaches the next device operation that reads the array, a device write into an array
whose copy was read is seen by the next read, and views of one allocation share its elements.
"""

from __future__ import annotations

import numpy as np
import pytest

import libsci_amd as sci

pytestmark = pytest.mark.gpu


def test_reads_follow_device_work(device):
    a = sci.Array.from_numpy(np.arange(8, dtype=np.float32), device)
    assert a[3] == 3.0
    b = sci.evaluate(a * 2.0 + 1.0)
    assert b[3] == 7.0 and b[7] == 15.0
    sci.assign(b, a * 3.0)  # a device write into an array whose copy was read
    assert b[3] == 9.0
    assert np.array_equal(b[2:5], np.array([6.0, 9.0, 12.0], dtype=np.float32))


def test_writes_reach_the_next_device_use(device):
    values = np.linspace(-1.0, 1.0, 1000, dtype=np.float32)
    a = sci.Array.from_numpy(values, device)
    a[10] = 100.0
    a[20:23] = [1.0, 2.0, 3.0]
    want = values.copy()
    want[10] = 100.0
    want[20:23] = [1.0, 2.0, 3.0]
    got = sci.evaluate(a * 2.0).host()
    assert np.array_equal(got, want * np.float32(2.0))
    assert np.array_equal(a.host(), want)
    # a reduction reads the written elements too (sum in double of the float terms, exact here)
    a[:] = 0.5
    assert sci.reduce_raw(a + 0.0, sci.Fold.sum) == 500.0


def test_matrix_elements_and_products(device):
    m = sci.Array.from_numpy(np.zeros((4, 4), dtype=np.float32), device)
    for i in range(4):
        m[i, i] = 1.0
    m[0, 3] = 2.0
    x = sci.Array.from_numpy(np.arange(16, dtype=np.float32).reshape(4, 4), device)
    p = sci.product(m, x)
    want = np.eye(4, dtype=np.float64)
    want[0, 3] = 2.0
    assert np.array_equal(p.host(), (want @ np.arange(16, dtype=np.float64).reshape(4, 4)).astype(np.float32))
    assert m[0, 3] == 2.0 and m[1, 1] == 1.0


def test_views_share_the_host_copy(device):
    a = sci.Array.from_numpy(np.zeros(64, dtype=np.float64), device)
    v = a.view(16, 8)
    w = a.reshape(8, 8)
    v[0] = 5.0
    assert a[16] == 5.0 and w[2, 0] == 5.0
    w[7, 7] = -1.0
    total = sci.evaluate(a * 1.0).host()
    assert total[16] == 5.0 and total[63] == -1.0 and np.count_nonzero(total) == 2
