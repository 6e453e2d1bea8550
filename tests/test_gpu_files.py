"""libsci_amd.save / load: device arrays to and from .npy files (numpy's format), on the GPU."""

from __future__ import annotations

import numpy as np
import pytest

import libsci_amd as sci

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_save_and_load_round_trip(device, tmp_path, dtype):
    print(f"[test] npy round trip {np.dtype(dtype).name}", flush=True)
    values = np.random.default_rng(3).standard_normal((37, 5)).astype(dtype)
    values[4, 2] = np.nan
    array = sci.Array.from_numpy(values, device)
    path = tmp_path / "a.npy"
    sci.save(path, array)
    on_disk = np.load(path, allow_pickle=False)
    assert on_disk.dtype == dtype and on_disk.shape == values.shape
    assert on_disk.tobytes() == values.tobytes()
    back = sci.load(path, device)
    assert back.shape == values.shape and back.host().tobytes() == values.tobytes()


def test_load_refuses_other_element_types(device, tmp_path):
    path = tmp_path / "ints.npy"
    np.save(path, np.arange(5, dtype=np.int64))
    with pytest.raises(TypeError):
        sci.load(path, device)
