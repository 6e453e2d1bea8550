"""exp, log, sin, cos, tan, csc, sec, cot and sqrt over every float32 bit pattern, against the
reference's own AVX-512 table: one launch of 2^32 elements per code (16 GiB in, 16 GiB out, beyond
the 2^31 - 1 elements the reference's GPU path refuses, src/core/drivers/device.cpp:149), in the
interpreter and in the specialised kernel, whose outputs must reproduce the 256 chunk digests of
tests/golden/digests.json (the rule is tests/helpers.py digest_chunk; sin and cos beyond 2^20, where
both sides call a libm, are left out and checked within 2 ulp by test_gpu_chunks.py). The digests
are formed on the GPU with torch, as plumbing: the library's output is copied into a torch tensor.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import libsci_amd as sci
from helpers import DIGEST_CHUNK, DIGEST_CODES, DIGEST_TRIG, golden_json
from libsci_amd import _abi
from libsci_amd.chunk import Code, Precision, record_postfix

pytestmark = pytest.mark.gpu

N = 1 << 32
_REACH = 0x49800000


def _signed(value: int) -> int:
    return value - (1 << 64) if value >= 1 << 63 else value


C1, C2 = _signed(0xBF58476D1CE4E5B9), _signed(0x94D049BB133111EB)


def _shift_right(z, k):  # logical shift of int64 bit patterns
    return (z >> k) & ((1 << (64 - k)) - 1)


def digests_on_gpu(name: str, outputs: torch.Tensor, batch_chunks: int = 8):
    """digest_chunk of every chunk, computed with torch int64 arithmetic on the GPU."""
    result = []
    for first in range(0, N // DIGEST_CHUNK, batch_chunks):
        start, stop = first * DIGEST_CHUNK, (first + batch_chunks) * DIGEST_CHUNK
        patterns = torch.arange(start, stop, dtype=torch.int64, device=outputs.device)
        raw = outputs[start:stop]
        bits = raw.to(torch.int64) & 0xFFFFFFFF
        value = raw.view(torch.float32)
        bits = torch.where(value != value, torch.full_like(bits, 0x7FC00000), bits)
        z = (patterns << 32) | bits
        z = z ^ _shift_right(z, 30)
        z = z * C1
        z = z ^ _shift_right(z, 27)
        z = z * C2
        z = z ^ _shift_right(z, 31)
        if name in DIGEST_TRIG:
            z = torch.where((patterns & 0x7FFFFFFF) > _REACH, torch.zeros_like(z), z)
        z = z.view(batch_chunks, DIGEST_CHUNK)
        low = (z & 0xFFFFFFFF).sum(dim=1).tolist()
        high = _shift_right(z, 32).sum(dim=1).tolist()
        result += [(lo + (hi << 32)) % (1 << 64) for lo, hi in zip(low, high)]
    return result


def test_every_float32_pattern(device):
    fixture = golden_json("digests.json")["codes"]
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]

    def copy(destination, source):
        # a device-to-device hipMemcpy may return before the copy is done, and the library's stream
        # does not wait for the null stream: wait for the whole device before anyone reads the copy
        assert hip.hipMemcpy(destination, source, 4 * N, 3) == 0
        assert hip.hipDeviceSynchronize() == 0

    print("[test] 2^32 float32 patterns: building the input", flush=True)
    patterns = torch.cat([torch.arange(0, 1 << 31, dtype=torch.int32, device="cuda"),
                          torch.arange(-(1 << 31), 0, dtype=torch.int32, device="cuda")])
    x = sci.Array(device, (N,), np.float32)
    y = sci.Array(device, (N,), np.float32)
    torch.cuda.synchronize()
    copy(x.pointer, patterns.data_ptr())
    copied = patterns  # reused for the outputs
    failures = []
    for name in DIGEST_CODES:
        chunk = record_postfix([("in", 0), ("op", Code[name])], Precision.single, "output", [x.pointer])
        for policy in (1, 2):
            print(f"[test] {name} over 2^32 patterns, policy {policy}", flush=True)
            device.set_policy(policy)
            _abi.check(_abi.lib().scihip_execute(device.handle, ctypes.byref(chunk.struct), y.pointer, N))
            device.synchronize()
            copy(copied.data_ptr(), y.pointer)
            got = digests_on_gpu(name, copied)
            bad = [k for k, value in enumerate(got) if f"{value:016x}" != fixture[name][k]]
            if bad:
                failures.append((name, policy, [hex(k) for k in bad[:8]], len(bad)))
    device.set_policy(0)
    assert not failures, failures
