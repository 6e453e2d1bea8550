"""Writes the golden fixtures of tests/golden/ from the reference's own compiled CPU code.

Run where /root/reference is mounted: ``python tests/golden/make_golden.py``. It builds
oracle/_ref (the reference's kernel tables, pool and Chunk, compiled from /root/reference by
oracle/Makefile) and records, through that library only:

* chunks.json     the bytecode the reference's Chunk records for a set of expression programs
                  (include/core/bytecode/chunk.hpp), with its cost and validity;
* kernels.npz     every code's output over seeded inputs plus special values, float32 and float64,
                  from the AVX-512 table (src/core/drivers/processor/kernels/avx512.cpp), after checking
                  that the fallback table built with a true fma gives the same values;
* reductions.json sum / minimum / maximum of chunks at several sizes, for 1 and 4 pool threads,
                  through the reference's Plan, folds and Accumulator;
* digests.json    (``--digests``, minutes) for exp, log, sin, cos, tan, csc, sec, cot and sqrt, the
                  digest of the AVX-512 table's outputs over every float32 bit pattern, one per chunk
                  of 2^24 patterns (tests/helpers.py digest_chunk gives the rule).

Inputs are regenerated in the tests from the recorded seeds with the oracle's generator, whose
values are fixed by a closed form (tests/test_oracle.py pins it), so only outputs are stored.
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from libsci_amd.chunk import ARITY, Code, Fold  # noqa: E402

SPECIALS32 = [0.0, -0.0, np.inf, -np.inf, np.nan, 1e-40, -1e-40, 1e-45, 1e30, -1e30, 3e6, -3e6, 1048576.0,
              1048577.0, 88.7, 89.1, -103.9, -104.5, 0.5, -0.5, 1.0, -1.0, 2.0, 1e-7, 0.70710677, 1.4142135,
              3.1415927, 1.5707964, 12582912.0, 1e38]
SPECIALS64 = [0.0, -0.0, np.inf, -np.inf, np.nan, 1e-310, -1e-310, 5e-324, 1e300, -1e300, 3e6, 1048577.0, 709.0,
              710.5, -745.0, -746.5, 0.5, -1.0, 2.0, 1e-17, 0.7071067811865476, 1.4142135623730951,
              3.141592653589793]

KERNEL_COUNT = 1024
KERNEL_SEEDS = (11, 12, 13)

# Expression programs, postfix, as Node::record walks them (include/array/node.hpp:75-88).
PROGRAMS = {
    "c1_store": [("in", 0), ("in", 1), ("op", Code.multiply), ("const", 2.0), ("op", Code.add)],
    "c1_reduce": [("in", 0), ("op", Code.sin)],
    "c2_store": [("in", 0), ("op", Code.exponential), ("const", 0.5), ("op", Code.multiply),
                 ("in", 1), ("op", Code.absolute), ("const", 1.0), ("op", Code.add), ("op", Code.logarithm),
                 ("in", 2), ("op", Code.multiply), ("op", Code.add)],
    "c2_reduce": [("in", 0), ("op", Code.sin), ("in", 2), ("op", Code.cos), ("op", Code.multiply),
                  ("in", 1), ("op", Code.tanh), ("op", Code.add)],
    "c2_dot": [("in", 0), ("in", 2), ("op", Code.multiply)],
    "c5_loss": [("in", 0), ("in", 1), ("op", Code.multiply), ("in", 2), ("op", Code.add), ("op", Code.sinh),
                ("in", 3), ("in", 3), ("op", Code.multiply), ("op", Code.negate), ("op", Code.exponential),
                ("op", Code.multiply)],
    "fused_left": [("in", 0), ("in", 0), ("in", 1), ("op", Code.multiply), ("op", Code.add)],
    "fused_right": [("in", 0), ("in", 1), ("op", Code.multiply), ("in", 0), ("op", Code.add)],
    "not_fused": [("in", 0), ("in", 1), ("op", Code.multiply), ("in", 2), ("op", Code.sin), ("op", Code.add)],
    "difference": [("in", 0), ("in", 1), ("op", Code.add), ("in", 0), ("in", 1), ("op", Code.subtract),
                   ("op", Code.multiply), ("const", 0.5), ("op", Code.add)],
    "copy": [("in", 0)],
    "constants": [("in", 0), ("const", 2.0), ("op", Code.multiply), ("const", 2.0), ("op", Code.add),
                  ("const", 3.0), ("op", Code.divide)],
    "minmax": [("in", 0), ("in", 1), ("op", Code.minimum), ("in", 2), ("op", Code.maximum)],
}


def random_program(rng: np.random.Generator, inputs: int, depth: int):
    """A random expression tree over `inputs` arrays, in postfix."""
    def node(d):
        if d == 0 or rng.random() < 0.25:
            if rng.random() < 0.2:
                return [("const", float(rng.choice([0.5, 1.0, 2.0, -3.0, 0.1])))]
            return [("in", int(rng.integers(inputs)))]
        code = Code(int(rng.choice([c for c in Code if c != Code.copy])))
        parts = []
        for _ in range(ARITY[code]):
            parts += node(d - 1)
        # A tree needs an array leaf; a constant-only subtree is fine as long as the whole has one.
        return parts + [("op", code)]
    while True:
        program = node(depth)
        if any(item[0] == "in" for item in program):
            return program


def program_json(program):
    return [[k, (int(v) if k == "op" or k == "in" else float(v))] for k, v in program]


def main() -> None:
    oracle.build(reference=True)
    oracle.ref_threads(4)
    rng = np.random.default_rng(2024)

    # ---- chunks
    programs = dict(PROGRAMS)
    for k in range(120):
        programs[f"random_{k}"] = random_program(rng, inputs=1 + k % 5, depth=2 + k % 6)
    # Deep trees overflow the 64-slot chunk (chunk.hpp:223-226).
    programs["overflow"] = [("in", 0)] + [("const", float(i)) if j == 0 else ("op", Code.add)
                                          for i in range(70) for j in range(2)]
    chunks = {}
    for name, program in programs.items():
        for finish in ("output", "reduce"):
            rec = oracle.ref_record(program, 0, finish)
            chunks[f"{name}/{finish}"] = {"program": program_json(program), "finish": finish, **rec}
    with open(os.path.join(HERE, "chunks.json"), "w") as handle:
        json.dump(chunks, handle, indent=0, sort_keys=True)

    # ---- element kernels
    arrays = {}
    for dtype, specials, tag in ((np.float32, SPECIALS32, "f32"), (np.float64, SPECIALS64, "f64")):
        inputs = []
        for seed in KERNEL_SEEDS:
            x = oracle.uniform(KERNEL_COUNT, seed, -8.0, 8.0, dtype=dtype)
            x[: len(specials)] = np.array(specials, dtype=dtype)[np.random.default_rng(seed).permutation(len(specials))]
            inputs.append(x)
        for code in Code:
            ar = ARITY[code]
            vector = oracle.ref_kernel(code, *inputs[:ar], isa=2)
            scalar = oracle.ref_kernel(code, *inputs[:ar], isa=0)
            # The vector table and the one-lane table agree on every value (NaN payloads and the sign
            # of a zero picked by minimum / maximum aside), so one output per code is kept.
            same = (vector == scalar) | (np.isnan(vector) & np.isnan(scalar))
            assert same.all(), (tag, code.name)
            arrays[f"{tag}/{code.name}"] = vector
    np.savez_compressed(os.path.join(HERE, "kernels.npz"), **{k.replace("/", "__"): v for k, v in arrays.items()})

    # ---- reductions
    reductions = {}
    for name in ("c1_reduce", "c2_reduce", "c2_dot", "c5_loss", "copy", "difference", "minmax"):
        program = PROGRAMS[name]
        inputs = max(v for k, v in program if k == "in") + 1
        for count in (1, 5, 1000, 1024, 3001, 70001, 300000):
            arrays_in = [oracle.uniform(count, 100 + i, -2.0, 2.0) for i in range(inputs)]
            for fold in Fold:
                for isa in (2, 0):
                    value = oracle.ref_reduce(program, arrays_in, fold, isa=isa)
                    reductions[f"{name}/{count}/{fold.name}/{isa}"] = {
                        "program": program_json(program), "count": count, "fold": int(fold), "isa": isa,
                        "seeds": [100 + i for i in range(inputs)], "threads": 4, "value": value}
    with open(os.path.join(HERE, "reductions.json"), "w") as handle:
        json.dump(reductions, handle, indent=0, sort_keys=True)
    print("wrote", len(chunks), "chunks,", len(arrays), "kernel outputs,", len(reductions), "reductions")


def write_digests(threads: int = 8) -> None:
    """digests.json: every chunk digest of every digested code, from the reference's AVX-512 table."""
    import concurrent.futures

    sys.path.insert(0, os.path.dirname(HERE))
    from helpers import DIGEST_CHUNK, DIGEST_CODES, digest_chunk, digest_inputs

    chunks = (1 << 32) // DIGEST_CHUNK

    def one(job):
        name, chunk = job
        x = digest_inputs(chunk)
        return name, chunk, digest_chunk(name, chunk, oracle.ref_kernel(int(Code[name]), x, isa=2))

    table = {name: [None] * chunks for name in DIGEST_CODES}
    with concurrent.futures.ThreadPoolExecutor(threads) as pool:
        for name, chunk, value in pool.map(one, [(n, c) for n in DIGEST_CODES for c in range(chunks)]):
            table[name][chunk] = f"{value:016x}"
    out = {"rule": "tests/helpers.py digest_chunk", "chunk": DIGEST_CHUNK,
           "source": "the reference's AVX-512 kernel table (src/core/drivers/processor/kernels/avx512.cpp) "
                     "through oracle/_ref", "codes": table}
    with open(os.path.join(HERE, "digests.json"), "w") as handle:
        json.dump(out, handle, indent=0)


if __name__ == "__main__":
    if "--digests" in sys.argv:
        oracle.build(reference=True)
        write_digests()
    else:
        main()
