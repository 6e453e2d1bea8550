"""The .npy reader and writer of include/sci/npy.hpp against numpy's own (numpy.lib.format), through
the host-only program tests/cpp/npy_roundtrip.cpp: files numpy writes are read back bit for bit,
files the program writes load in numpy with the same shape and bits, and malformed or unsupported
files are refused with an error rather than misread. No GPU.
"""

from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PROGRAM = os.path.join(HERE, "cpp", "npy_roundtrip")


def program():
    if not os.path.exists(PROGRAM):
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp")], check=True)
    return PROGRAM


def run(*args):
    return subprocess.run([program(), *map(str, args)], capture_output=True, text=True)


CASES = [
    ("f4", np.arange(12, dtype=np.float32).reshape(3, 4) * np.float32(0.25) - 1),
    ("f8", np.linspace(-1.0, 1.0, 7)),
    ("f4", np.array([np.nan, -0.0, np.inf, -np.inf, 1e-45, 3.4028235e38], dtype=np.float32)),
    ("f8", np.zeros((0, 3))),
    ("f4", np.float32(2.5).reshape(())),
    ("f8", np.random.default_rng(7).standard_normal((2, 3, 5))),
]


@pytest.mark.parametrize("kind,values", CASES, ids=[f"{k}{v.shape}" for k, v in CASES])
def test_numpy_file_round_trips_bit_for_bit(tmp_path, kind, values):
    source, target = tmp_path / "in.npy", tmp_path / "out.npy"
    np.save(source, values)
    result = run("copy", source, target, kind)
    assert result.returncode == 0, result.stderr
    assert result.stdout.split() == ["shape", *map(str, values.shape)]
    back = np.load(target, allow_pickle=False)
    assert back.dtype == values.dtype and back.shape == values.shape
    assert back.tobytes() == values.tobytes()
    with open(target, "rb") as f:
        head = f.read(10)
    assert head[:6] == b"\x93NUMPY" and head[6] == 1
    assert (10 + int.from_bytes(head[8:10], "little")) % 64 == 0  # the data starts 64-byte aligned


def test_program_file_loads_in_numpy(tmp_path):
    target = tmp_path / "made.npy"
    result = run("make", target)
    assert result.returncode == 0, result.stderr
    got = np.load(target, allow_pickle=False)
    want = (np.arange(12, dtype=np.float32) * np.float32(0.5) - np.float32(1)).reshape(3, 4)
    assert got.dtype == np.float32 and got.shape == (3, 4)
    assert got.tobytes() == want.tobytes()


def test_version_two_header_is_read(tmp_path):
    values = np.arange(6, dtype=np.float64).reshape(2, 3)
    source, target = tmp_path / "v2.npy", tmp_path / "out.npy"
    with open(source, "wb") as f:
        np.lib.format.write_array(f, values, version=(2, 0))
    result = run("copy", source, target, "f8")
    assert result.returncode == 0, result.stderr
    assert np.load(target, allow_pickle=False).tobytes() == values.tobytes()


def test_wrong_element_type_is_refused(tmp_path):
    source = tmp_path / "ints.npy"
    np.save(source, np.arange(4, dtype=np.int32))
    result = run("copy", source, tmp_path / "out.npy", "f4")
    assert result.returncode == 2 and "<i4" in result.stderr
    np.save(source, np.arange(4, dtype=np.float64))
    result = run("copy", source, tmp_path / "out.npy", "f4")  # double elements asked for as float
    assert result.returncode == 2 and "<f8" in result.stderr


def test_fortran_order_is_refused(tmp_path):
    source = tmp_path / "fortran.npy"
    np.save(source, np.asfortranarray(np.ones((3, 2), dtype=np.float32)))
    result = run("copy", source, tmp_path / "out.npy", "f4")
    assert result.returncode == 2 and "Fortran" in result.stderr


def test_truncated_and_foreign_files_are_refused(tmp_path):
    source = tmp_path / "short.npy"
    np.save(source, np.ones(100, dtype=np.float32))
    data = source.read_bytes()
    source.write_bytes(data[:-4])
    result = run("copy", source, tmp_path / "out.npy", "f4")
    assert result.returncode == 2 and "truncated" in result.stderr
    source.write_bytes(b"not an npy file at all")
    result = run("copy", source, tmp_path / "out.npy", "f4")
    assert result.returncode == 2 and "not a .npy file" in result.stderr
