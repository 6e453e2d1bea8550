"""Benchmark of the HIP backend on one MI355X. Prints one JSON line.

Headline (`value`): the fused chunk store exp(a)*0.5 + log(|b|+1)*c over N = 2^28 float32 elements
(a, b, c seeded uniform[-4, 4], 1% of b NaN), as GB/s of algorithmic traffic (3 arrays read, 1
written: 16 bytes per element) and as a fraction of the 8 TB/s HBM peak. A step is one pass of the
chunk over the whole arrays, inputs resident in HBM.

Beside it in the same line, every workload of the measurement plan:
- the streaming ceilings of the store's traffic mix (the same 16 B per element with one fma per
  element; a plain copy, 8 B per element);
- the C2 reductions sum / mean / minimum / maximum of sin(a)*cos(c) + tanh(b) and dot(a, c);
- the matrix products 8192^3 and 4096 x 16384 x 1024 in fp32, bf16 and bf16x3, with their errors
  against a float64 product of 64 sampled rows;
- the n = 16384 LU solves with 32 right-hand sides at fp64 accuracy, on the well-conditioned and the
  graded system, plain fp64 LU against the mixed factorisations refined in fp64, each with its
  backward error and its forward error against LAPACK's float64 solve;
- the differentiation workload (fused forward + backward of sum(sinh(a*b + c) * exp(-(d*d))) at
  N = 2^26 against its forward pass, and the product adjoint at 4096^2);
- launch-bound chains of small chunks issued call by call against one graph replay, the 16384 x
  16384 device transpose and axis folds, the fp64 store, small-array latency;
- the CPU baseline: the oracle (the reference's CPU path restated: AVX-512 packs, 1,024-element
  blocks, a fork-join pool on every host core) on the same workloads, median of 20 runs after 3
  warm-ups as the reference's demo times (src/main.cpp:21-22, 43-51).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-lu] [--no-products]
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK = 8.0e12  # bytes / s, MI355X


class Events:
    """HIP events on the library's own stream (torch's events would only see torch's streams)."""

    def __init__(self, stream: int):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.stream = ctypes.c_void_p(stream)
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]

    def create(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, event):
        assert self.hip.hipEventRecord(event, self.stream) == 0

    def elapsed_ms(self, start, stop) -> float:
        assert self.hip.hipEventSynchronize(stop) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), start, stop) == 0
        return ms.value


def time_launches(device, events, launch, steps: int, warmup: int):
    """Runs `launch` warmup + steps times; returns (wall ms per step, event ms per launch)."""
    for _ in range(warmup):
        launch()
    device.synchronize()
    start, stop = events.create(), events.create()
    t0 = time.perf_counter()
    events.record(start)
    for _ in range(steps):
        launch()
    events.record(stop)
    device.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return wall, events.elapsed_ms(start, stop) / steps


def median_time(run, repeats: int = 20, warmups: int = 3) -> float:
    """Seconds per run: the median of `repeats` timed runs after `warmups` untimed ones, as the
    reference's demo times its operations (src/main.cpp:21-22, 43-51)."""
    for _ in range(warmups):
        run()
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    return statistics.median(times)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(hosts, threads: int) -> dict:
    """The oracle (the reference's CPU path restated: 64-byte packs in the AVX-512 table's operation
    order, 1,024-element blocks, part split and fork-join pool, bit for bit the reference's results)
    on the host cores, on the workloads the GPU runs. `hosts` are the C2 inputs a, b (1% NaN), c
    and the NaN-free b as host arrays of the full size."""
    import numpy as np

    import oracle
    from libsci_amd.chunk import Code, Fold, Precision, record_postfix

    lib = oracle.lib()
    a, b, c, clean = hosts
    n = a.size
    store = [("in", 0), ("op", Code.exponential), ("const", 0.5), ("op", Code.multiply),
             ("in", 1), ("op", Code.absolute), ("const", 1.0), ("op", Code.add), ("op", Code.logarithm),
             ("in", 2), ("op", Code.multiply), ("op", Code.add)]
    chunk = record_postfix(store, Precision.single, "output", [a.ctypes.data, b.ctypes.data, c.ctypes.data])
    out = np.empty(n, dtype=np.float32)
    seconds = median_time(lambda: lib.oracle_execute(ctypes.addressof(chunk.struct), out.ctypes.data, n, threads))
    result = {"value": round(16 * n / seconds / 1e9, 2), "unit": "GB/s", "cores": threads, "kind": "port",
              "sample": f"the headline store itself, N=2^{n.bit_length() - 1} float32, median of 20 passes after "
                        f"3 warm-ups, oracle/oracle.cpp on {threads} host threads",
              "cpu_model": cpu_model()}
    paths = {}
    # C1: c = a*b + 2 stored, then sum(sin(c)), 1e6 elements (the reference's CPU-runnable workload)
    m = 1_000_000
    ramp = oracle.linspace(0.0, 1.0, m)
    ones = np.ones(m, dtype=np.float32)
    c1 = np.empty(m, dtype=np.float32)
    c1_store = record_postfix([("in", 0), ("in", 1), ("op", Code.multiply), ("const", 2.0), ("op", Code.add)],
                              Precision.single, "output", [ramp.ctypes.data, ones.ctypes.data])
    c1_sum = record_postfix([("in", 0), ("op", Code.sin)], Precision.single, "reduce", [c1.ctypes.data])
    store_s = median_time(lambda: lib.oracle_execute(ctypes.addressof(c1_store.struct), c1.ctypes.data, m, threads))
    sum_s = median_time(lambda: oracle.reduce(c1_sum, Fold.sum, m, threads=threads))
    paths["c1"] = {"store_us": round(store_s * 1e6, 1), "sum_sin_us": round(sum_s * 1e6, 1),
                   "total": oracle.reduce(c1_sum, Fold.sum, m, threads=threads)}
    # C2 reduction (glibc tanh per lane dominates): a 2^24-element sample; the dot at full size
    k = min(n, 1 << 24)
    reduction = record_postfix([("in", 0), ("op", Code.sin), ("in", 2), ("op", Code.cos), ("op", Code.multiply),
                                ("in", 1), ("op", Code.tanh), ("op", Code.add)], Precision.single, "reduce",
                               [a.ctypes.data, clean.ctypes.data, c.ctypes.data])
    seconds = median_time(lambda: oracle.reduce(reduction, Fold.sum, k, threads=threads))
    paths["c2_reduce_sum"] = {"GBps": round(12 * k / seconds / 1e9, 2), "sample": f"first 2^{k.bit_length() - 1} "
                                                                                   f"elements"}
    dot = record_postfix([("in", 0), ("in", 1), ("op", Code.multiply)], Precision.single, "reduce",
                         [a.ctypes.data, c.ctypes.data])
    seconds = median_time(lambda: oracle.reduce(dot, Fold.sum, n, threads=threads))
    paths["c2_dot"] = {"GBps": round(8 * n / seconds / 1e9, 2)}
    # C3: the register-blocked fp32 product (dispatch.cpp:171-224) at 1024^3
    q = 1024
    left = oracle.uniform(q * q, 31, -1.0, 1.0).reshape(q, q)
    right = oracle.uniform(q * q, 32, -1.0, 1.0).reshape(q, q)
    seconds = median_time(lambda: oracle.product(left, right, threads=threads), repeats=5, warmups=1)
    paths["c3_product_fp32"] = {"GFLOPs": round(2 * q ** 3 / seconds / 1e9, 1), "sample": f"{q}^3, median of 5"}
    # C4: the reference's unblocked LU with partial pivoting (decompose.hpp:29-72), fp64, n = 2048
    r = 2048
    matrix = oracle.uniform(r * r, 41, -1.0, 1.0, dtype=np.float64).reshape(r, r)
    matrix[np.diag_indices(r)] += r / 4
    seconds = median_time(lambda: oracle.factor(matrix, threads=threads), repeats=5, warmups=1)
    paths["c4_factor_fp64"] = {"GFLOPs": round(2 * r ** 3 / 3 / seconds / 1e9, 1), "ms": round(seconds * 1e3, 1),
                               "sample": f"n={r}, median of 5"}
    result["paths"] = paths
    return result


def measure_ceilings(sci, _abi, device, events, arrays, out, steps: int) -> dict:
    """What the store's traffic costs without its arithmetic, on the same arrays: the same 16 B per
    element with one fma per element (a*c + b), and a plain copy (8 B per element)."""
    a, b, c = arrays
    n = a.size
    result = {}
    for name, expression, traffic in (("fma_store", a * c + b, 16), ("copy", sci.copy(a), 8)):
        chunk = sci.record(expression)
        device.prepare(expression)
        device.set_policy(2)

        def run():
            _abi.check(_abi.lib().scihip_execute(device.handle, ctypes.byref(chunk.struct), out.pointer, n))

        _, ms = time_launches(device, events, run, steps, 10)
        device.set_policy(0)
        result[f"{name}_GBps"] = round(traffic * n / (ms * 1e-3) / 1e9, 1)
    return result


def measure_interpreter(sci, _abi, device, events, arrays, out, steps: int) -> dict:
    """The interpreter (policy 1: no specialised kernel) on the headline store (16 B per element) and on
    the C2 sum of sin(a)*cos(c) + tanh(b) (12 B per element): what a chunk runs at before its
    specialised kernel is compiled."""
    a, b, c = arrays
    n = a.size
    store_chunk = sci.record(sci.exponential(a) * 0.5 + sci.logarithm(sci.absolute(b) + 1.0) * c)
    sum_chunk = sci.record(sci.sin(a) * sci.cos(c) + sci.tanh(b), reduce=True)
    device.set_policy(1)
    total = ctypes.c_double()

    def store():
        _abi.check(_abi.lib().scihip_execute(device.handle, ctypes.byref(store_chunk.struct), out.pointer, n))

    def fold():
        _abi.check(_abi.lib().scihip_reduce(device.handle, ctypes.byref(sum_chunk.struct), 0, n, ctypes.byref(total)))

    _, store_ms = time_launches(device, events, store, steps, 2)
    sum_seconds = median_time(fold, steps, 2)
    device.set_policy(0)
    return {"store_GBps": round(16 * n / (store_ms * 1e-3) / 1e9, 1), "store_ms": round(store_ms, 4),
            "sum_call_GBps": round(12 * n / sum_seconds / 1e9, 1)}


def measure_reductions(sci, _abi, device, arrays, clean, steps: int) -> dict:
    """C2 reductions over N = 2^28: sum / mean / minimum / maximum of sin(a)*cos(c) + tanh(b) (12 B
    per element), with the 1%-NaN b of the workload (every result NaN) and the NaN-free b, and
    dot(a, c) (8 B per element). Each call waits for its result: the time includes both kernels,
    the 8-byte copy and the host wait."""
    from libsci_amd.chunk import Fold

    a, b, c = arrays
    n = a.size
    result = {}
    for label, bb in (("nan", b), ("clean", clean)):
        expression = sci.sin(a) * sci.cos(c) + sci.tanh(bb)
        chunk = sci.record(expression, reduce=True)
        device.prepare(expression, reduce=True)
        device.set_policy(2)
        total = ctypes.c_double()
        entry = {}
        for fold in Fold:
            def call():
                _abi.check(_abi.lib().scihip_reduce(device.handle, ctypes.byref(chunk.struct), int(fold), n,
                                                    ctypes.byref(total)))

            seconds = median_time(call, steps, 2)
            entry[fold.name] = {"value": total.value, "GBps": round(12 * n / seconds / 1e9, 1)}
        entry["mean"] = float(sci.mean(expression))
        device.set_policy(0)
        result[label] = entry
    expression = a * c
    chunk = sci.record(expression, reduce=True)
    device.prepare(expression, reduce=True)
    device.set_policy(2)
    total = ctypes.c_double()

    def dot():
        _abi.check(_abi.lib().scihip_reduce(device.handle, ctypes.byref(chunk.struct), 0, n, ctypes.byref(total)))

    seconds = median_time(dot, steps, 2)
    device.set_policy(0)
    result["dot"] = {"value": total.value, "GBps": round(8 * n / seconds / 1e9, 1)}
    return result


def measure_products(sci, _abi, device, events) -> dict:
    """C3: 8192^3 and 4096 x 16384 times 16384 x 1024, entries seeded uniform[-1, 1], in fp32, bf16
    (fp32 accumulation) and bf16x3, each as TFLOP/s of 2MNK with its max abs and relative Frobenius
    error against a float64 product of 64 seeded rows."""
    import numpy as np

    out = {}
    for m, k, n in ((8192, 8192, 8192), (4096, 16384, 1024)):
        left = sci.uniform(m * k, device, 31, -1.0, 1.0).reshape(m, k)
        right = sci.uniform(k * n, device, 32, -1.0, 1.0).reshape(k, n)
        product = sci.Array(device, (m, n), "float32")
        rows = np.sort(np.random.default_rng(m + n).choice(m, 64, replace=False))
        exact = left.host()[rows].astype(np.float64) @ right.host().astype(np.float64)
        for name, mode in (("fp32", 0), ("bf16", 1), ("bf16x3", 2)):
            def run():
                _abi.check(_abi.lib().scihip_product(device.handle, mode, left.pointer, right.pointer,
                                                     product.pointer, m, n, k))
            _, ms = time_launches(device, events, run, 5, 1)
            got = product.host()[rows].astype(np.float64)
            out[f"{m}x{k}x{n}_{name}"] = {
                "TFLOPs": round(2 * m * n * k / (ms * 1e-3) / 1e12, 1), "ms": round(ms, 3),
                "max_abs_error": float(np.abs(got - exact).max()),
                "relative_frobenius": float(np.linalg.norm(got - exact) / np.linalg.norm(exact))}
        del left, right, product
    # The same 8192^3 bf16 product on matrices of ones: the exact result (8192 everywhere) and the
    # same work, but operands whose bits barely toggle, which the matrix cores' power limit shows.
    m = k = n = 8192
    ones = sci.full((m, k), 1.0, device)
    product = sci.Array(device, (m, n), "float32")

    def run_ones():
        _abi.check(_abi.lib().scihip_product(device.handle, 1, ones.pointer, ones.pointer, product.pointer, m, n, k))

    _, ms = time_launches(device, events, run_ones, 5, 1)
    got = product.host()
    out["8192x8192x8192_bf16_ones"] = {"TFLOPs": round(2 * m * n * k / (ms * 1e-3) / 1e12, 1), "ms": round(ms, 3),
                                       "max_abs_error": float(np.abs(got.astype(np.float64) - k).max()),
                                       "relative_frobenius": 0.0 if np.all(got == k) else float("nan")}
    del ones, product
    return out


def solve_system(sci, device, order, width, graded):
    """A = U[-1,1] + (order/4) I, rows scaled by 10^(-6 i / order) when graded; B = U[-1,1]. Generated
    on the device (seeded), the diagonal shift and the row scaling applied on the host. Returns the
    device arrays and their host copies."""
    import numpy as np

    a_host = sci.uniform(order * order, device, 41, -1.0, 1.0, dtype=np.float64).host().reshape(order, order)
    a_host[np.diag_indices(order)] += order / 4
    if graded:
        a_host *= (10.0 ** (-6.0 * np.arange(order) / order))[:, None]
    a = sci.Array.from_numpy(a_host, device)
    b = sci.uniform(order * width, device, 42, -1.0, 1.0, dtype=np.float64).reshape(order, width)
    return a, b, a_host, b.host()


def measure_lu(sci, _abi, device, order=16384, width=32, graded=False, repeats=2) -> dict:
    """A X = B to backward error 1e-14: plain fp64 LU against the mixed-precision factorisations (the
    trailing update on fp32, bf16x3 or bf16 matrix cores, the panel in fp32) refined in fp64; best of
    `repeats` calls each. Every solve states its backward error (the library's, and recomputed on the
    host in float64) and its forward error against LAPACK's float64 LU solve of the same system.
    `mixed` names the fastest mixed variant that converged, `mixed_speedup` compares it with the
    plain solve."""
    import numpy as np
    import scipy.linalg

    a, b, a_host, b_host = solve_system(sci, device, order, width, graded)
    x = sci.Array(device, (order, width), "float64")
    out = {"order": order, "right_hand_sides": width, "tolerance": 1e-14,
           "system": "graded (rows scaled by 10^(-6i/n))" if graded else "shifted (U[-1,1] + n/4 I)"}
    solutions = {}
    for name, mode in (("plain_fp64", 0), ("mixed_fp32", 1), ("mixed_bf16x3", 2), ("mixed_bf16", 3)):
        best = None
        for _ in range(repeats):
            stats = _abi.SolveStats()
            _abi.check(_abi.lib().scihip_solve(device.handle, mode, a.pointer, b.pointer, x.pointer, order, width,
                                               1e-14, 30, ctypes.byref(stats)))
            if best is None or stats.solve_ms < best.solve_ms:
                best = stats
        solutions[name] = x.host()
        lu_flops = 2.0 * order ** 3 / 3.0
        out[name] = {"ms": round(best.solve_ms, 1), "factor_ms": round(best.factor_ms, 1),
                     "factor_TFLOPs": round(lu_flops / (best.factor_ms * 1e-3) / 1e12, 1),
                     "refinements": best.iterations, "backward_error": best.backward_error,
                     "converged": bool(best.converged)}
    t0 = time.perf_counter()
    reference = scipy.linalg.solve(a_host, b_host)
    out["lapack_seconds"] = round(time.perf_counter() - t0, 2)
    norm = np.abs(a_host).sum(axis=1).max()
    for name, solution in solutions.items():
        residual = np.abs(b_host - a_host @ solution).max(axis=0)
        backward = residual / (norm * np.abs(solution).max(axis=0) + np.abs(b_host).max(axis=0))
        out[name]["backward_error_host"] = float(backward.max())
        out[name]["forward_error"] = float(np.abs(solution - reference).max() / np.abs(reference).max())
    mixed = [name for name in ("mixed_fp32", "mixed_bf16x3", "mixed_bf16") if out[name]["converged"]]
    if mixed:
        best = min(mixed, key=lambda name: out[name]["ms"])
        out["mixed"] = best
        out["mixed_speedup"] = round(out["plain_fp64"]["ms"] / out[best]["ms"], 3)
    return out


def measure_two_dimensional(sci, _abi, device, events, n=16384, steps=10) -> dict:
    """The 2-D variant: X (n x n float32, seeded uniform[-4, 4]) transposed on the device (8 B per
    element), then summed along each axis (4 B per element), each as GB/s of that traffic."""
    from libsci_amd.chunk import Fold

    x = sci.uniform(n * n, device, 61, -4.0, 4.0).reshape(n, n)
    t = sci.Array(device, (n, n), "float32")
    out = sci.Array(device, (n,), "float32")

    def transpose():
        _abi.check(_abi.lib().scihip_transpose(device.handle, 0, x.pointer, t.pointer, n, n))

    result = {"workload": f"{n} x {n} float32"}
    _, ms = time_launches(device, events, transpose, steps, 2)
    result["transpose_ms"] = round(ms, 4)
    result["transpose_GBps"] = round(8 * n * n / (ms * 1e-3) / 1e9, 1)
    chunk = sci.record(sci.copy(x), reduce=True)
    device.set_policy(2)  # the specialised axis kernels (compiled on first use)
    for axis, name in ((1, "rows"), (0, "columns")):
        for fold, label in ((Fold.sum, "sum"), (Fold.maximum, "maximum")):
            def fold_axis():
                _abi.check(_abi.lib().scihip_reduce_axis(device.handle, ctypes.byref(chunk.struct), int(fold), n, n,
                                                         axis, out.pointer))
            _, ms = time_launches(device, events, fold_axis, steps, 2)
            result[f"{label}_{name}_GBps"] = round(4 * n * n / (ms * 1e-3) / 1e9, 1)
    device.set_policy(0)
    return result


def measure_wide_store(sci, _abi, device, events, count: int, steps: int) -> dict:
    """The C2 store expression over float64 arrays of `count` elements (the same bytes as the float32
    workload at twice the count): the precision the reference's GPU path refuses and hands to the CPU."""
    import numpy as np

    a = sci.uniform(count, device, 11, -4.0, 4.0, dtype=np.float64)
    b = sci.uniform(count, device, 12, -4.0, 4.0, dtype=np.float64, nan_seed=13, nan_fraction=0.01)
    c = sci.uniform(count, device, 14, -4.0, 4.0, dtype=np.float64)
    out = sci.Array(device, (count,), np.float64)
    expression = sci.exponential(a) * 0.5 + sci.logarithm(sci.absolute(b) + 1.0) * c
    chunk = sci.record(expression)
    device.prepare(expression)
    device.set_policy(2)

    def store():
        _abi.check(_abi.lib().scihip_execute(device.handle, ctypes.byref(chunk.struct), out.pointer, count))

    # fresh arrays: the first launches on them ran up to 20% slower, so warm up longer than 2 launches
    _, ms = time_launches(device, events, store, steps, 10)
    device.set_policy(0)
    return {"workload": f"exp(a)*0.5 + log(|b|+1)*c, N={count} float64", "ms": round(ms, 4),
            "GBps": round(32 * count / (ms * 1e-3) / 1e9, 1)}


def measure_small_arrays(sci, _abi, device, events, steps: int = 50) -> dict:
    """Latency of small chunks, the sizes below the reference's CPU hand-off threshold
    (src/core/drivers/device.cpp:100-104): a store (asynchronous; device time per launch) and a
    sum (synchronous; wall time per call, result on the host) of sin(a) * b."""
    out = {}
    for count in (1_000, 100_000, 1_000_000):
        a = sci.uniform(count, device, 21, -1.0, 1.0)
        b = sci.uniform(count, device, 22, -1.0, 1.0)
        target = sci.Array(device, (count,), "float32")
        expression = sci.sin(a) * b
        store_chunk = sci.record(expression)
        sum_chunk = sci.record(expression, reduce=True)
        device.prepare(expression)
        device.prepare(expression, reduce=True)
        total = ctypes.c_double()

        def store():
            _abi.check(_abi.lib().scihip_execute(device.handle, ctypes.byref(store_chunk.struct), target.pointer,
                                                 count))

        def total_of():
            _abi.check(_abi.lib().scihip_reduce(device.handle, ctypes.byref(sum_chunk.struct), 0, count,
                                                ctypes.byref(total)))

        _, store_ms = time_launches(device, events, store, steps, 3)
        sum_wall, _ = time_launches(device, events, total_of, steps, 3)
        out[str(count)] = {"store_us": round(store_ms * 1e3, 2), "sum_call_us": round(sum_wall * 1e3, 2)}
    return out


def measure_graph(sci, _abi, device, events, counts=(1_000, 10_000, 1_000_000), chain=32, steps=20) -> dict:
    """The launch-bound case: `chain` dependent chunk stores (c = a*b + 2, then alternately sin / cos
    of the previous result times b) over 1e3 and 1e4 float32 elements, where launching dominates, and
    over 1e6 (the CPU-runnable workload's size), issued call by call against one replay of the same
    calls captured in a graph."""
    out = {"workload": f"{chain} dependent chunk stores per replay, float32"}
    for count in counts:
        a = sci.linspace(0.0, 1.0, count, device)
        b = sci.ones(count, device)
        buffers = [sci.Array(device, (count,), "float32") for _ in range(2)]
        stages = [(a * b + 2.0, buffers[0])]
        for k in range(1, chain):
            source, target = buffers[(k - 1) % 2], buffers[k % 2]
            stages.append(((sci.sin(source) if k % 2 else sci.cos(source)) * b, target))
        chunks = [(sci.record(e), t) for e, t in stages]
        for e, _ in stages:
            device.prepare(e)
        device.set_policy(2)

        def eager():
            for chunk, target in chunks:
                _abi.check(_abi.lib().scihip_execute(device.handle, ctypes.byref(chunk.struct), target.pointer, count))

        device.begin_capture()
        eager()
        graph = device.end_capture()
        eager_wall, eager_ms = time_launches(device, events, eager, steps, 3)
        graph_wall, graph_ms = time_launches(device, events, graph.replay, steps, 3)
        graph.release()
        device.set_policy(0)
        out[str(count)] = {"eager_us_per_chunk": round(eager_wall * 1e3 / chain, 2),
                           "graph_us_per_chunk": round(graph_wall * 1e3 / chain, 2),
                           "eager_device_us_per_chunk": round(eager_ms * 1e3 / chain, 2),
                           "graph_device_us_per_chunk": round(graph_ms * 1e3 / chain, 2)}
    return out


def measure_adjoint(sci, _abi, device, events, log2n=26, steps=10) -> dict:
    """Fused forward + backward of sum(sinh(a*b + c) * exp(-(d*d))) over N = 2^26 float32 (a..d seeded
    uniform[-2, 2]) against the forward (the loss as one fused reduction), both specialised kernels;
    and the product adjoint (two matrix-core products) at 4096^2 fp32."""
    import numpy as np

    n = 1 << log2n
    operands = [sci.uniform(n, device, 40 + k, -2.0, 2.0) for k in range(4)]
    a, b, c, d = operands
    expression = sci.sinh(a * b + c) * sci.exponential(-(d * d))
    chunk = sci.record(expression, reduce=True)
    gradients = [sci.Array(device, (n,), "float32") for _ in range(4)]
    pointers = (ctypes.c_void_p * 4)(*[g.pointer for g in gradients])
    loss = ctypes.c_double()
    device.prepare(expression, reduce=True)
    device.set_policy(2)

    def forward():
        _abi.check(_abi.lib().scihip_reduce(device.handle, ctypes.byref(chunk.struct), 0, n, ctypes.byref(loss)))

    def both():
        _abi.check(_abi.lib().scihip_adjoint(device.handle, ctypes.byref(chunk.struct), None, pointers, n,
                                             ctypes.byref(loss)))

    both()  # compiles the specialised adjoint kernel (policy 2 waits for it)
    _, forward_ms = time_launches(device, events, forward, steps, 2)
    _, both_ms = time_launches(device, events, both, steps, 2)
    device.set_policy(0)
    out = {"workload": f"sum(sinh(a*b+c)*exp(-(d*d))), N=2^{log2n} float32",
           "forward_ms": round(forward_ms, 4), "forward_GBps": round(16 * n / (forward_ms * 1e-3) / 1e9, 1),
           "forward_backward_ms": round(both_ms, 4),
           "forward_backward_GBps": round(32 * n / (both_ms * 1e-3) / 1e9, 1),
           "ratio": round(both_ms / forward_ms, 3)}
    del operands, a, b, c, d, gradients

    m = 4096
    left = sci.uniform(m * m, device, 51, -1.0, 1.0).reshape(m, m)
    right = sci.uniform(m * m, device, 52, -1.0, 1.0).reshape(m, m)
    seed = sci.uniform(m * m, device, 53, -1.0, 1.0).reshape(m, m)
    d_left = sci.Array(device, (m, m), "float32")
    d_right = sci.Array(device, (m, m), "float32")

    def product_adjoint():
        _abi.check(_abi.lib().scihip_product_adjoint(device.handle, 0, left.pointer, right.pointer, seed.pointer,
                                                     d_left.pointer, d_right.pointer, m, m, m))

    _, ms = time_launches(device, events, product_adjoint, 5, 1)
    rows = np.arange(0, m, m // 32)
    want = seed.host()[rows].astype(np.float64) @ right.host().astype(np.float64).T
    got = d_left.host()[rows].astype(np.float64)
    out["product_adjoint_4096_fp32"] = {"ms": round(ms, 3), "TFLOPs": round(2 * 2 * m ** 3 / (ms * 1e-3) / 1e12, 1),
                                        "relative_frobenius_d_left": float(np.linalg.norm(got - want) /
                                                                           np.linalg.norm(want))}
    return out


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--log2n", type=int, default=28)
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    parser.add_argument("--no-lu", action="store_true", help="skip the n = 16384 LU solves")
    parser.add_argument("--no-products", action="store_true", help="skip the C3 products")
    parser.add_argument("--no-adjoint", action="store_true", help="skip the adjoint measurements")
    parser.add_argument("--solve-only", type=int, default=0, metavar="ORDER",
                        help="only time solves of this order (for profiling), in the --modes given")
    parser.add_argument("--modes", default="fp64,fp32",
                        help="with --solve-only: the factorisations to time (fp64, fp32, bf16x3, bf16)")
    parser.add_argument("--products-only", action="store_true", help="only the C3 products (for profiling)")
    parser.add_argument("--reductions-only", action="store_true", help="only the C2 reductions (for profiling)")
    parser.add_argument("--adjoint-only", action="store_true", help="only the adjoint measurements (for profiling)")
    parser.add_argument("--small-only", action="store_true", help="only the small-array latencies")
    parser.add_argument("--matrix-only", action="store_true", help="only the 2-D transpose and axis folds")
    parser.add_argument("--interpreter-only", action="store_true",
                        help="only the interpreter's store and sum (for profiling)")
    parser.add_argument("--store-only", action="store_true",
                        help="only the headline store, then a plain copy chunk of the same size (8 B per element) "
                             "that calibrates hardware byte counters; for profiling")
    args = parser.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this backend runs on one GPU; sharding across GPUs is designed but not built",
                          "n_gpus": args.gpus}))
        return 2

    import numpy as np

    import libsci_amd as sci
    from libsci_amd import _abi

    device = sci.Device(0)
    stream = ctypes.c_void_p()
    _abi.check(_abi.lib().scihip_stream(device.handle, ctypes.byref(stream)))
    events = Events(stream.value)

    if args.solve_only:
        codes = {"fp64": 0, "fp32": 1, "bf16x3": 2, "bf16": 3}
        a, b, _, _ = solve_system(sci, device, args.solve_only, 32, False)
        x = sci.Array(device, (args.solve_only, 32), "float64")
        out = {}
        for name in args.modes.split(","):
            stats = _abi.SolveStats()
            _abi.check(_abi.lib().scihip_solve(device.handle, codes[name], a.pointer, b.pointer, x.pointer,
                                               args.solve_only, 32, 1e-14, 30, ctypes.byref(stats)))
            out[name] = {"ms": round(stats.solve_ms, 1), "factor_ms": round(stats.factor_ms, 1),
                         "refinements": stats.iterations, "backward_error": stats.backward_error}
        print(json.dumps(out), flush=True)
        device.close()
        return 0
    if args.products_only:
        print(json.dumps(measure_products(sci, _abi, device, events)), flush=True)
        device.close()
        return 0
    if args.matrix_only:
        print(json.dumps(measure_two_dimensional(sci, _abi, device, events)), flush=True)
        device.close()
        return 0
    if args.small_only:
        print(json.dumps(measure_small_arrays(sci, _abi, device, events)), flush=True)
        device.close()
        return 0
    if args.adjoint_only:
        print(json.dumps(measure_adjoint(sci, _abi, device, events)), flush=True)
        device.close()
        return 0

    n = 1 << args.log2n
    a = sci.uniform(n, device, 1, -4.0, 4.0)
    b = sci.uniform(n, device, 2, -4.0, 4.0, nan_seed=3, nan_fraction=0.01)
    c = sci.uniform(n, device, 4, -4.0, 4.0)
    out = sci.Array(device, (n,), "float32")
    expression = sci.exponential(a) * 0.5 + sci.logarithm(sci.absolute(b) + 1.0) * c
    chunk = sci.record(expression)
    device.prepare(expression)  # steady state: the chunk's specialised kernel is compiled and ready
    device.set_policy(2)

    def store():
        _abi.check(_abi.lib().scihip_execute(device.handle, ctypes.byref(chunk.struct), out.pointer, n))

    if args.interpreter_only:
        print(json.dumps(measure_interpreter(sci, _abi, device, events, (a, b, c), out, args.steps)), flush=True)
        device.close()
        return 0
    if args.reductions_only:
        clean = sci.uniform(n, device, 2, -4.0, 4.0)
        print(json.dumps(measure_reductions(sci, _abi, device, (a, b, c), clean, args.steps)), flush=True)
        device.close()
        return 0
    bytes_per_step = 16 * n
    wall_ms, kernel_ms = time_launches(device, events, store, args.steps, args.warmup)
    achieved = bytes_per_step / (kernel_ms * 1e-3)
    value = bytes_per_step / (wall_ms * 1e-3) / 1e9
    device.set_policy(0)
    if args.store_only:
        copy_chunk = sci.record(sci.copy(a))
        device.prepare(sci.copy(a))
        device.set_policy(2)

        def copy():
            _abi.check(_abi.lib().scihip_execute(device.handle, ctypes.byref(copy_chunk.struct), out.pointer, n))

        _, copy_ms = time_launches(device, events, copy, args.steps, args.warmup)
        print(json.dumps({"store_GBps": round(achieved / 1e9, 1), "store_ms": round(kernel_ms, 4),
                          "copy_GBps": round(8 * n / (copy_ms * 1e-3) / 1e9, 1), "copy_ms": round(copy_ms, 4)}),
              flush=True)
        device.close()
        return 0

    # HBM bytes per launch of the store from rocprofv3 FETCH_SIZE / WRITE_SIZE, calibrated on a plain
    # copy of known size (the summary names the kernel and the command that measured it)
    traffic = None
    traffic_file = os.path.join(HERE, "profiles", "traffic_c2_store.txt")
    if os.path.exists(traffic_file):
        for line in open(traffic_file):
            if line.startswith("traffic_bytes:"):
                traffic = int(line.split(":")[1])

    def progress(stage):
        print(f"[bench] {stage}", file=sys.stderr, flush=True)

    progress("ceilings")
    ceilings = measure_ceilings(sci, _abi, device, events, (a, b, c), out, max(args.steps // 2, 5))
    progress("interpreter")
    interpreter = measure_interpreter(sci, _abi, device, events, (a, b, c), out, max(args.steps // 2, 5))
    clean = sci.uniform(n, device, 2, -4.0, 4.0)
    progress("reductions")
    reductions = measure_reductions(sci, _abi, device, (a, b, c), clean, max(args.steps // 2, 5))
    hosts = (a.host(), b.host(), c.host(), clean.host()) if not args.no_cpu else None
    del out
    products = None
    if not args.no_products:
        progress("products")
        products = measure_products(sci, _abi, device, events)
    lu = lu_graded = None
    if not args.no_lu:
        progress("lu (shifted)")
        lu = measure_lu(sci, _abi, device)
        progress("lu (graded)")
        lu_graded = measure_lu(sci, _abi, device, graded=True, repeats=1)
    adjoint = None
    if not args.no_adjoint:
        progress("adjoint")
        adjoint = measure_adjoint(sci, _abi, device, events)
    progress("graph, 2-D, fp64, small arrays")
    graph = measure_graph(sci, _abi, device, events)
    two_dimensional = measure_two_dimensional(sci, _abi, device, events)
    wide = measure_wide_store(sci, _abi, device, events, n // 2, max(args.steps // 2, 3))
    small = measure_small_arrays(sci, _abi, device, events)

    result = {
        "metric": "fused-chunk HBM GB/s (% peak); LU solve time at FP64 accuracy, mixed vs plain",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded uniform[-4,4] generated on the device, 1% NaN in b",
        "config": {"workload": f"fused chunk store exp(a)*0.5 + log(|b|+1)*c, N=2^{args.log2n} float32, "
                               f"specialised kernel", "elements": n, "bytes_per_step": bytes_per_step},
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic},
        "ceilings": ceilings,
        "interpreter": interpreter,
        "reductions": reductions,
    }
    if products:
        result["products"] = products
    if lu:
        result["lu_solve"] = lu
        result["lu_graded"] = lu_graded
    if adjoint:
        result["adjoint"] = adjoint
    result["graph"] = graph
    result["two_dimensional"] = two_dimensional
    result["fp64_store"] = wide
    result["small_arrays"] = small
    device.close()
    if hosts is not None:
        progress("cpu baseline")
        threads = len(os.sched_getaffinity(0))
        result["cpu_baseline"] = cpu_baseline(hosts, threads)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
