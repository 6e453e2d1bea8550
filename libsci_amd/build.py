"""Builds libsci_amd/lib/libscihip.so for gfx950 with hipcc (no CMake, no JIT cache).

Usage: python -m libsci_amd.build   (or libsci_amd.build.build() from Python)

Every .hip file under csrc/ is compiled in parallel to an object, then linked with hipRTC into one
shared library exporting the C ABI of include/scihip.h. The device sources used by run-time
specialisation (device_math.hpp, interpreter.hpp, adjoint.hpp) are embedded into the library as text.
"""

from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib")
TARGET = os.path.join(LIB, "libscihip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    # Packed FP32 instructions stay off (the compiler would otherwise form them from float code).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-I" + BUILD,
]

EMBEDDED = ("device_math.hpp", "interpreter.hpp", "adjoint.hpp")


def _embed() -> None:
    parts = []
    for name in EMBEDDED:
        with open(os.path.join(CSRC, name)) as handle:
            for line in handle:
                if line.startswith("#pragma once") or line.startswith('#include "'):
                    continue
                parts.append(line)
    text = "".join(parts)
    assert ")SCIHIPSRC\"" not in text
    content = "static const char kEmbeddedSource[] = R\"SCIHIPSRC(" + text + ")SCIHIPSRC\";\n"
    path = os.path.join(BUILD, "embedded.inc")
    if not os.path.exists(path) or open(path).read() != content:
        with open(path, "w") as handle:
            handle.write(content)


def _compile(source: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(source) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", source, "-o", obj]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"hipcc failed for {source}:\n{result.stderr[-4000:]}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    _embed()
    sources = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    workers = min(len(sources), max(1, min(16, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        objects = list(pool.map(_compile, sources))
    temporary = TARGET + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objects, "-o", temporary,
           "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"link failed:\n{result.stderr[-4000:]}")
    os.replace(temporary, TARGET)
    if verbose:
        print(f"built {TARGET}")
    return TARGET


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
