"""Parity of the HIP chunk executor with the CPU oracle, through the C ABI.

Bar: bit for bit for +, -, *, /, fma, sqrt, abs, negate, min / max (by value) and libsci's own
exp / log / sin / cos (and the codes built from them: tan, csc, sec, cot) for |x| <= 2^20; a
stated ulp bound where the CPU calls glibc and the GPU the device library (asin ... atanh and
their reciprocal forms, sin / cos beyond 2^20); NaN placement exact everywhere. Both the
interpreter and the specialised (hipRTC) kernels are checked.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

import libsci_amd as sci
import oracle
from libsci_amd import _abi
from libsci_amd.chunk import ARITY, Code, Fold, Precision, record_postfix
from tests.helpers import same_bits, ulp_distance, ulps32
from tests.test_oracle import kernel_inputs

pytestmark = pytest.mark.gpu

LIBRARY = {Code.asin, Code.acos, Code.atan, Code.sinh, Code.cosh, Code.tanh, Code.asinh, Code.acosh, Code.atanh,
           Code.acsc, Code.asec, Code.acot, Code.csch, Code.sech, Code.coth, Code.acsch, Code.asech, Code.acoth}
TRIG = {Code.sin, Code.cos, Code.tan, Code.csc, Code.sec, Code.cot}
LIBRARY_ULPS = 4  # measured: 2 for every code but coth (4), over 44 M float32 patterns (test_library_codes_ulps)


def upload(device, arrays):
    return [sci.Array.from_numpy(a, device) for a in arrays]


def run_chunk(device, program, arrays, dtype, policy):
    """Records `program` over device copies of `arrays`, runs it, returns the output."""
    device.set_policy(policy)
    inputs = upload(device, arrays)
    precision = Precision.single if dtype == np.float32 else Precision.wide
    chunk = record_postfix(program, precision, "output", [a.pointer for a in inputs])
    out = sci.Array(device, arrays[0].shape, dtype)
    _abi.check(_abi.lib().scihip_execute(device.handle, ctypes.byref(chunk.struct), out.pointer, out.size))
    return out.host()


def oracle_chunk(program, arrays, dtype):
    precision = Precision.single if dtype == np.float32 else Precision.wide
    chunk = record_postfix(program, precision, "output", [a.ctypes.data for a in arrays])
    return oracle.execute(chunk, arrays[0].size, threads=8, dtype=dtype)


def check_code(code, got, want, x):
    nan = np.isnan(want)
    assert (np.isnan(got) == nan).all(), f"{code.name}: NaN placement"
    if code in LIBRARY:
        d = ulp_distance(got, want)
        assert d.max() <= LIBRARY_ULPS, (code.name, d.max())
        return
    exact = same_bits(got, want)
    if code in (Code.minimum, Code.maximum):
        exact |= got == want
    if code in TRIG:
        reach = np.abs(x) <= 1048576.0
        assert exact[reach].all(), (code.name, np.flatnonzero(~exact & reach)[:5])
        assert ulp_distance(got[~reach], want[~reach]).max() <= 2
        return
    assert exact.all(), (code.name, np.flatnonzero(~exact)[:5], got[~exact][:3], want[~exact][:3])


@pytest.mark.parametrize("policy", [1, 2], ids=["interpreter", "specialised"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_every_code(device, dtype, policy):
    print(f"[test] every code dtype={np.dtype(dtype).name} policy={policy}", flush=True)
    inputs = kernel_inputs(dtype, 4096)
    for code in Code:
        args = inputs[: ARITY[code]]
        program = [("in", i) for i in range(len(args))] + [("op", code)]
        got = run_chunk(device, program, args, dtype, policy)
        want = oracle_chunk(program, args, dtype)
        check_code(code, got, want, args[0])
    device.set_policy(0)


def test_exhaustive_float32_exp_log_sin_cos(device):
    """Every float32 bit pattern in [0x00000000, 0xFFFFFFFF] step 97: exp, log, sin, cos bit for bit."""
    print("[test] strided sweep of float32 patterns", flush=True)
    x = np.arange(0, 2 ** 32, 97, dtype=np.uint64).astype(np.uint32).view(np.float32)
    for code in (Code.exponential, Code.logarithm, Code.sin, Code.cos, Code.root, Code.divide):
        args = [x, x[::-1].copy()][: ARITY[code]]
        program = [("in", i) for i in range(len(args))] + [("op", code)]
        got = run_chunk(device, program, args, np.float32, 1)
        want = oracle_chunk(program, args, np.float32)
        check_code(code, got, want, x)
    device.set_policy(0)


def test_library_codes_ulps(device):
    """Every float32 pattern in [0x00000000, 0xFFFFFFFF] step 97 through each code the CPU sends to
    glibc (asin ... atanh and their reciprocal forms), against the oracle (glibc): NaN placement exact,
    and the largest distance in ulp, reported per code, within LIBRARY_ULPS."""
    x = np.arange(0, 2 ** 32, 97, dtype=np.uint64).astype(np.uint32).view(np.float32)
    worst = {}
    for code in sorted(LIBRARY, key=int):
        program = [("in", 0), ("op", code)]
        got = run_chunk(device, program, [x], np.float32, 1)
        want = oracle_chunk(program, [x], np.float32)
        assert (np.isnan(got) == np.isnan(want)).all(), code.name
        worst[code.name] = int(ulps32(got, want).max())
    device.set_policy(0)
    print("[test] max ulp, GPU against glibc, over 44 M float32 patterns: " +
          ", ".join(f"{k} {v}" for k, v in worst.items()), flush=True)
    assert max(worst.values()) <= LIBRARY_ULPS, worst


C2_STORE = [("in", 0), ("op", Code.exponential), ("const", 0.5), ("op", Code.multiply),
            ("in", 1), ("op", Code.absolute), ("const", 1.0), ("op", Code.add), ("op", Code.logarithm),
            ("in", 2), ("op", Code.multiply), ("op", Code.add)]


@pytest.mark.parametrize("count", [0, 1, 3, 4, 5, 1023, 1025, 1 << 20, (1 << 20) + 7])
@pytest.mark.parametrize("policy", [1, 2], ids=["interpreter", "specialised"])
def test_c2_store_bit_exact(device, count, policy):
    print(f"[test] c2 store count={count} policy={policy}", flush=True)
    a = oracle.uniform(count, 1, -4.0, 4.0)
    b = oracle.uniform(count, 2, -4.0, 4.0, nan_seed=3, nan_fraction=0.01)
    c = oracle.uniform(count, 4, -4.0, 4.0)
    if count == 0:
        device.set_policy(policy)
        out = sci.Array(device, (1,), np.float32)
        chunk = record_postfix(C2_STORE, addresses=[out.pointer] * 3)
        assert _abi.lib().scihip_execute(device.handle, ctypes.byref(chunk.struct), out.pointer, 0) == 0
        return
    got = run_chunk(device, C2_STORE, [a, b, c], np.float32, policy)
    want = oracle_chunk(C2_STORE, [a, b, c], np.float32)
    assert same_bits(got, want).all()
    assert (np.isnan(got) == np.isnan(b)).all()
    device.set_policy(0)


def test_c1_store_is_one_fma(device):
    """c = a * b + 2 records as one fuse and rounds once (tests/operations/test_binary.cpp:40-42)."""
    a = oracle.linspace(0.0, 1.0, 1_000_000)
    b = np.ones(1_000_000, dtype=np.float32)
    program = [("in", 0), ("in", 1), ("op", Code.multiply), ("const", 2.0), ("op", Code.add)]
    got = run_chunk(device, program, [a, b], np.float32, 0)
    assert (got == a + np.float32(2.0)).all()
    # a = b = 1 + 2^-12, c = 2^-60: one rounding gives 0x1.002002p+0, two would give 0x1.002p+0.
    x = np.full(8, np.float32(1 + 2 ** -12))
    for policy in (1, 2):
        got = run_chunk(device, [("in", 0), ("in", 1), ("op", Code.multiply), ("const", 2.0 ** -60),
                                 ("op", Code.add)], [x, x], np.float32, policy)
        assert (got == np.float32(float.fromhex("0x1.002002p+0"))).all()
    device.set_policy(0)


def test_unaligned_interior_pointers(device):
    """Interior pointers off 16-byte alignment take the lane-by-lane path and stay exact."""
    a = oracle.uniform(10007, 1, -4.0, 4.0)
    b = oracle.uniform(10007, 2, -4.0, 4.0)
    da, db = upload(device, [a, b])
    va, vb = da.view(1, 10005), db.view(2, 10005)
    out = sci.Array(device, (10006,), np.float32)
    vo = out.view(1, 10005)
    chunk = record_postfix([("in", 0), ("in", 1), ("op", Code.divide), ("op", Code.sin)],
                           addresses=[va.pointer, vb.pointer])
    _abi.check(_abi.lib().scihip_execute(device.handle, ctypes.byref(chunk.struct), vo.pointer, 10005))
    got = vo.host()
    want = oracle.map_f32(Code.sin, oracle.map_f32(Code.divide, a[1:10006], b[2:10007]))
    assert same_bits(got, want).all()


def test_large_register_files(device):
    """Chunks beyond the VGPR register file (many inputs, constants, temporaries) run exactly."""
    n = 5000
    arrays = [oracle.uniform(n, 50 + i, 0.5, 2.0) for i in range(20)]
    # ((a0 + 0.5) * (a1 + 1.5)) ... a deep balanced tree of 20 inputs and 20 constants
    program = []
    for i in range(20):
        program += [("in", i), ("const", 0.25 * i + 0.5), ("op", Code.add), ("op", Code.sin)]
        if i:
            program += [("op", Code.multiply if i % 2 else Code.add)]
    for policy in (1, 2):
        got = run_chunk(device, program, arrays, np.float32, policy)
        want = oracle_chunk(program, arrays, np.float32)
        assert same_bits(got, want).all()
    # 32 temporaries at once: 32 values stay alive until the adds at the end (63 instructions)
    program = []
    for i in range(32):
        program += [("in", i % 20), ("op", Code.exponential)]
    program += [("op", Code.add)] * 31
    chunk = record_postfix(program, addresses=[a.ctypes.data for a in arrays])
    assert chunk.valid() and chunk.temporaries == 32
    for policy in (1, 2):
        got = run_chunk(device, program, arrays, np.float32, policy)
        want = oracle_chunk(program, arrays, np.float32)
        assert same_bits(got, want).all()
    device.set_policy(0)


def test_inputs_elsewhere_are_declined(device):
    """An input that is not device memory makes the call decline (the caller's CPU path runs it)."""
    host = np.ones(16, dtype=np.float32)
    out = sci.Array(device, (16,), np.float32)
    chunk = record_postfix([("in", 0), ("op", Code.sin)], addresses=[host.ctypes.data])
    assert _abi.lib().scihip_execute(device.handle, ctypes.byref(chunk.struct), out.pointer, 16) == _abi.DECLINED
    chunk = record_postfix([("in", 0), ("op", Code.sin)], addresses=[out.pointer])
    assert _abi.lib().scihip_execute(device.handle, ctypes.byref(chunk.struct), host.ctypes.data, 16) == _abi.DECLINED


def test_specialised_kernels_are_used(device):
    device.set_policy(0)
    a = sci.uniform(1 << 16, device, 1, -1.0, 1.0)
    expr = sci.cos(a) * 3.0 + a
    device.prepare(expr)
    before = device.statistics()
    sci.evaluate(expr)
    after = device.statistics()
    assert after["specialised"] == before["specialised"] + 1
    assert after["compiled"] >= 1


def test_front_end_expression(device):
    """The Python host mirror records what the reference records and runs it unchanged."""
    a_h = oracle.linspace(-3.0, 5.0, 300_000)
    b_h = oracle.linspace(2.0, 7.0, 300_000)
    a, b = sci.Array.from_numpy(a_h, device), sci.Array.from_numpy(b_h, device)
    assert (sci.evaluate(a + b).host() == a_h + b_h).all()
    assert (sci.evaluate(a - b).host() == a_h - b_h).all()
    assert (sci.evaluate(a * b).host() == a_h * b_h).all()
    assert (sci.evaluate(a / b).host() == a_h / b_h).all()
    assert (sci.evaluate(sci.minimum(a, b)).host() == np.minimum(a_h, b_h)).all()
    assert (sci.evaluate(2.0 * a).host() == np.float32(2.0) * a_h).all()
    fused = sci.evaluate(a * b + a).host()
    want = oracle.map_f32(Code.fuse, a_h, b_h, a_h)
    assert (fused == want).all()
    nan = sci.full((4,), np.nan, device)
    one = sci.ones((4,), device)
    assert np.isnan(sci.evaluate(sci.minimum(nan, one)).host()).all()
    assert np.isnan(sci.evaluate(sci.maximum(one, nan)).host()).all()
