"""Benchmark of the HIP backend on one MI355X. Prints one JSON line.

Headline (`value`): the fused chunk store exp(a)*0.5 + log(|b|+1)*c over N = 2^28 float32 elements
(a, b, c seeded uniform[-4, 4], 1% of b NaN), as GB/s of algorithmic traffic (3 arrays read, 1
written: 16 bytes per element) and as a fraction of the 8 TB/s HBM peak. A step is one pass of the
chunk over the whole arrays, inputs resident in HBM.

Beside it in the same line: the other half of the metric, the n = 16384 LU solve with 32
right-hand sides at fp64 accuracy (plain fp64 LU against an fp32 factorisation refined in fp64),
the differentiation workload (fused forward + backward of sum(sinh(a*b + c) * exp(-(d*d))) at
N = 2^26 against its forward pass, and the product adjoint at 4096^2), launch-bound chains of small
chunks issued call by call against one graph replay, the 16384 x 16384 device transpose and axis
folds, and the CPU baseline.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--extra] [--linear]
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


def cpu_baseline(sample: int, threads: int) -> dict:
    """The oracle (a restatement of the reference's CPU path) on a bounded sample of the workload."""
    import numpy as np

    import oracle
    from libsci_amd.chunk import Code, Precision, record_postfix

    C2_STORE = [("in", 0), ("op", Code.exponential), ("const", 0.5), ("op", Code.multiply),
                ("in", 1), ("op", Code.absolute), ("const", 1.0), ("op", Code.add), ("op", Code.logarithm),
                ("in", 2), ("op", Code.multiply), ("op", Code.add)]

    a = oracle.uniform(sample, 1, -4.0, 4.0)
    b = oracle.uniform(sample, 2, -4.0, 4.0, nan_seed=3, nan_fraction=0.01)
    c = oracle.uniform(sample, 4, -4.0, 4.0)
    chunk = record_postfix(C2_STORE, Precision.single, "output", [a.ctypes.data, b.ctypes.data, c.ctypes.data])
    out = np.empty(sample, dtype=np.float32)
    lib = oracle.lib()
    lib.oracle_execute(ctypes.addressof(chunk.struct), out.ctypes.data, sample, threads)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        lib.oracle_execute(ctypes.addressof(chunk.struct), out.ctypes.data, sample, threads)
        times.append(time.perf_counter() - t0)
    seconds = statistics.median(times)
    return {"value": round(16 * sample / seconds / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"the same store over N=2^{sample.bit_length() - 1} float32 elements, median of 3 passes "
                      f"after 1 warm-up, oracle/oracle.cpp (block-wise restatement of dispatch.cpp) on "
                      f"{threads} host threads"}


def measure_linear(sci, _abi, device, events) -> dict:
    """8192^3 products in every mode (TFLOP/s, error against float64 on 64 sampled rows) and the
    n = 16384 solves in every mode and both matrix variants."""
    import numpy as np

    out = {}
    n = 8192
    left = sci.uniform(n * n, device, 31, -1.0, 1.0).reshape(n, n)
    right = sci.uniform(n * n, device, 32, -1.0, 1.0).reshape(n, n)
    product = sci.Array(device, (n, n), "float32")
    rows = np.arange(0, n, n // 64)
    exact = left.host()[rows].astype(np.float64) @ right.host().astype(np.float64)
    for name, mode in (("fp32", 0), ("bf16", 1), ("bf16x3", 2)):
        def run():
            _abi.check(_abi.lib().scihip_product(device.handle, mode, left.pointer, right.pointer, product.pointer,
                                                 n, n, n))
        _, ms = time_launches(device, events, run, 3, 1)
        got = product.host()[rows].astype(np.float64)
        out[f"product_8192_{name}"] = {"TFLOPs": round(2 * n ** 3 / (ms * 1e-3) / 1e12, 1), "ms": round(ms, 2),
                                       "max_abs_error": float(np.abs(got - exact).max()),
                                       "relative_frobenius": float(np.linalg.norm(got - exact) / np.linalg.norm(exact))}
    del left, right, product
    out.update(measure_solves(sci, _abi, device, 16384))
    return out


def solve_system(sci, device, order, width, graded):
    """A = U[-1,1] + (order/4) I, rows scaled by 10^(-6 i / order) when graded; B = U[-1,1]. Generated
    on the device (seeded), the diagonal shift and the row scaling applied on the host."""
    import numpy as np

    a_host = sci.uniform(order * order, device, 41, -1.0, 1.0, dtype=np.float64).host().reshape(order, order)
    a_host[np.diag_indices(order)] += order / 4
    if graded:
        a_host *= (10.0 ** (-6.0 * np.arange(order) / order))[:, None]
    a = sci.Array.from_numpy(a_host, device)
    b = sci.uniform(order * width, device, 42, -1.0, 1.0, dtype=np.float64).reshape(order, width)
    return a, b


def measure_solves(sci, _abi, device, order, width=32, graded_cases=(False, True),
                   modes=(("fp64", 0), ("fp32", 1), ("bf16x3", 2), ("bf16", 3))) -> dict:
    """A X = B to backward error 1e-14: plain fp64 LU against mixed factorisations + fp64 refinement."""
    out = {}
    for graded in graded_cases:
        a, b = solve_system(sci, device, order, width, graded)
        x = sci.Array(device, (order, width), "float64")
        label = "graded" if graded else "shifted"
        for name, mode in modes:
            stats = _abi.SolveStats()
            _abi.check(_abi.lib().scihip_solve(device.handle, mode, a.pointer, b.pointer, x.pointer, order, width,
                                               1e-14, 30, ctypes.byref(stats)))
            key = f"solve_{order}_{label}_{name}"
            out[key] = {"ms": round(stats.solve_ms, 1), "factor_ms": round(stats.factor_ms, 1),
                        "refinements": stats.iterations, "backward_error": stats.backward_error,
                        "converged": bool(stats.converged)}
            print(json.dumps({key: out[key]}), file=sys.stderr, flush=True)
    return out


def measure_lu_headline(sci, _abi, device, order=16384, width=32, repeats=2) -> dict:
    """The metric's second half: plain fp64 LU solve against the mixed-precision factorisations (the
    trailing update on fp32, bf16x3 or bf16 matrix cores, the panel in fp32) refined in fp64, all to
    backward error 1e-14, on the well-conditioned system; best of `repeats` calls each. `mixed` names
    the fastest mixed variant that converged, and `mixed_speedup` compares it with the plain solve."""
    a, b = solve_system(sci, device, order, width, False)
    x = sci.Array(device, (order, width), "float64")
    out = {"order": order, "right_hand_sides": width, "tolerance": 1e-14}
    for name, mode in (("plain_fp64", 0), ("mixed_fp32", 1), ("mixed_bf16x3", 2), ("mixed_bf16", 3)):
        best = None
        for _ in range(repeats):
            stats = _abi.SolveStats()
            _abi.check(_abi.lib().scihip_solve(device.handle, mode, a.pointer, b.pointer, x.pointer, order, width,
                                               1e-14, 30, ctypes.byref(stats)))
            if best is None or stats.solve_ms < best.solve_ms:
                best = stats
        lu_flops = 2.0 * order ** 3 / 3.0
        out[name] = {"ms": round(best.solve_ms, 1), "factor_ms": round(best.factor_ms, 1),
                     "factor_TFLOPs": round(lu_flops / (best.factor_ms * 1e-3) / 1e12, 1),
                     "refinements": best.iterations, "backward_error": best.backward_error,
                     "converged": bool(best.converged)}
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

    _, ms = time_launches(device, events, store, steps, 2)
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


def measure_graph(sci, _abi, device, events, count=1_000_000, chain=32, steps=20) -> dict:
    """The launch-bound case: `chain` dependent chunk stores over 1e6 float32 elements (the CPU-runnable
    workload's size: c = a*b + 2, then alternately sin / cos of the previous result times b), issued
    call by call against one replay of the same calls captured in a graph."""
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
    return {"workload": f"{chain} dependent chunk stores over {count} float32 elements",
            "eager_us_per_chunk": round(eager_wall * 1e3 / chain, 2),
            "graph_us_per_chunk": round(graph_wall * 1e3 / chain, 2),
            "eager_device_us_per_chunk": round(eager_ms * 1e3 / chain, 2),
            "graph_device_us_per_chunk": round(graph_ms * 1e3 / chain, 2)}


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
    parser.add_argument("--extra", action="store_true", help="also time reductions and the interpreter")
    parser.add_argument("--solve-only", type=int, default=0, metavar="ORDER",
                        help="only run one fp64 and one fp32-refined solve of this order (for profiling)")
    parser.add_argument("--modes", default="fp64,fp32",
                        help="with --solve-only: the factorisations to time (fp64, fp32, bf16x3, bf16)")
    parser.add_argument("--store-only", action="store_true",
                        help="only the headline store, then a plain copy chunk of the same size (8 B per element) "
                             "that calibrates hardware byte counters; for profiling")
    parser.add_argument("--no-lu", action="store_true", help="skip the n = 16384 LU solves")
    parser.add_argument("--no-adjoint", action="store_true", help="skip the adjoint measurements")
    parser.add_argument("--linear", action="store_true",
                        help="also time the 8192^3 products and the n = 16384 solves in every mode")
    args = parser.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this backend runs on one GPU; sharding across GPUs is designed but not built",
                          "n_gpus": args.gpus}))
        return 2

    import libsci_amd as sci
    from libsci_amd import _abi
    from libsci_amd.chunk import Fold

    if args.solve_only:
        device = sci.Device(0)
        codes = {"fp64": 0, "fp32": 1, "bf16x3": 2, "bf16": 3}
        modes = tuple((name, codes[name]) for name in args.modes.split(","))
        print(json.dumps(measure_solves(sci, _abi, device, args.solve_only, graded_cases=(False,), modes=modes)),
              flush=True)
        device.close()
        return 0

    n = 1 << args.log2n
    device = sci.Device(0)
    stream = ctypes.c_void_p()
    _abi.check(_abi.lib().scihip_stream(device.handle, ctypes.byref(stream)))
    events = Events(stream.value)

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

    bytes_per_step = 16 * n
    wall_ms, kernel_ms = time_launches(device, events, store, args.steps, args.warmup)
    achieved = bytes_per_step / (kernel_ms * 1e-3)
    value = bytes_per_step / (wall_ms * 1e-3) / 1e9
    traffic = None
    traffic_file = os.path.join(HERE, "profiles", "traffic_c2_store.txt")
    if os.path.exists(traffic_file):
        for line in open(traffic_file):
            if line.startswith("traffic_bytes:"):
                traffic = int(line.split(":")[1])

    extra = {}
    if args.extra:
        device.set_policy(1)
        w, k = time_launches(device, events, store, max(args.steps // 2, 3), 2)
        extra["interpreter_store_GBps"] = round(bytes_per_step / (k * 1e-3) / 1e9, 1)
        device.set_policy(0)
        total = ctypes.c_double()

        def reduction(expression):
            # A recorded reduction called through the C ABI; each call waits for its result (the
            # kernel, the 8-byte result copy and the host wait are all inside the time).
            chunk = sci.record(expression, reduce=True)
            device.prepare(expression, reduce=True)
            return lambda: _abi.check(_abi.lib().scihip_reduce(device.handle, ctypes.byref(chunk.struct),
                                                               int(Fold.sum), n, ctypes.byref(total)))

        reduce_call = reduction(sci.sin(a) * sci.cos(c) + sci.tanh(b))
        w, _ = time_launches(device, events, reduce_call, max(args.steps // 2, 5), 2)
        extra["c2_reduce_sum_GBps"] = round(12 * n / (w * 1e-3) / 1e9, 1)
        # the same traffic with one fma per element: the streaming ceiling of this kernel shape
        fma_expr = a * c + b
        fma_chunk = sci.record(fma_expr)
        device.prepare(fma_expr)
        device.set_policy(2)

        def fma_store():
            _abi.check(_abi.lib().scihip_execute(device.handle, ctypes.byref(fma_chunk.struct), out.pointer, n))

        _, k = time_launches(device, events, fma_store, max(args.steps // 2, 3), 2)
        extra["fma_store_GBps"] = round(bytes_per_step / (k * 1e-3) / 1e9, 1)
        device.set_policy(0)
        dot_call = reduction(a * c)
        w, _ = time_launches(device, events, dot_call, max(args.steps // 2, 5), 2)
        extra["c2_dot_GBps"] = round(8 * n / (w * 1e-3) / 1e9, 1)
        extra["fp64"] = measure_wide_store(sci, _abi, device, events, n // 2, max(args.steps // 2, 3))
        extra["small_arrays"] = measure_small_arrays(sci, _abi, device, events)

    if args.store_only:
        copy_chunk = sci.record(sci.copy(a))
        device.prepare(sci.copy(a))

        def copy():
            _abi.check(_abi.lib().scihip_execute(device.handle, ctypes.byref(copy_chunk.struct), out.pointer, n))

        _, copy_ms = time_launches(device, events, copy, args.steps, args.warmup)
        print(json.dumps({"store_GBps": round(achieved / 1e9, 1), "store_ms": round(kernel_ms, 4),
                          "copy_GBps": round(8 * n / (copy_ms * 1e-3) / 1e9, 1), "copy_ms": round(copy_ms, 4)}),
              flush=True)
        device.close()
        return 0

    device.set_policy(0)
    lu = measure_lu_headline(sci, _abi, device) if not args.no_lu else None
    adjoint = measure_adjoint(sci, _abi, device, events) if not args.no_adjoint else None
    graph = measure_graph(sci, _abi, device, events)
    two_dimensional = measure_two_dimensional(sci, _abi, device, events)

    linear = {}
    if args.linear:
        linear = measure_linear(sci, _abi, device, events)

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
    }
    if lu:
        result["lu_solve"] = lu
    if adjoint:
        result["adjoint"] = adjoint
    result["graph"] = graph
    result["two_dimensional"] = two_dimensional
    if extra:
        result["extra"] = extra
    if linear:
        result["linear"] = linear
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        result["cpu_baseline"] = cpu_baseline(min(n, 1 << 26), threads)
    print(json.dumps(result), flush=True)
    device.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
