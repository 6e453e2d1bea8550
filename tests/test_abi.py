"""The C-ABI library: it loads, exports exactly what include/scihip.h declares, and its structures
have the reference Chunk's layout. No GPU calls (this runs on the CPU)."""

from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

from libsci_amd import _abi
from libsci_amd.chunk import ChunkStruct

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for header in os.listdir(os.path.join(ROOT, "include")):
        if header.endswith(".h"):
            with open(os.path.join(ROOT, "include", header)) as handle:
                text = handle.read()
            names |= set(re.findall(r"\b(scihip_[a-z_]+)\s*\(", text))
    return names


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIBRARY], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if " T " in line and "scihip_" in line}


def test_library_loads_and_binds_every_declared_symbol():
    lib = _abi.lib()
    declared = declared_symbols()
    assert declared, "include/scihip.h declares the C ABI"
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(_abi.SIGNATURES), "the Python binding covers exactly the declared ABI"


def test_exports_match_header():
    exported = {s for s in exported_symbols() if s.startswith("scihip_")}
    assert exported == declared_symbols()


def test_chunk_layout_matches_reference_chunk():
    # core::Chunk's public fields (include/core/bytecode/chunk.hpp:180-196): six bytes, 64 five-byte
    # instructions, then 64 pointers and 64 doubles at 8-byte alignment. The C side is checked against
    # the reference's own header at compile time: oracle/ref_harness.cpp static_asserts every public
    # field's offset and size, and the numbering of codes, tags, folds and precisions; here the Python
    # binding is checked against the C layout.
    assert ChunkStruct.instructions.offset == 6
    assert ChunkStruct.inputs.offset == 328
    assert ChunkStruct.constants.offset == 840
    assert ctypes.sizeof(ChunkStruct) == 1352


def test_error_is_a_string():
    assert isinstance(_abi.lib().scihip_error(), bytes)


def test_null_device_is_rejected():
    lib = _abi.lib()
    assert lib.scihip_synchronize(None) == -1
    assert lib.scihip_execute(None, None, None, 0) == -1
    assert lib.scihip_name(None) == b"cpu"


def test_library_names_no_cpu_fallback():
    """The product path has no CPU engine: the library links no oracle and exports no CPU kernels."""
    out = subprocess.run(["nm", "-D", _abi.LIBRARY], capture_output=True, text=True, check=True).stdout
    assert "oracle_" not in out and "ref_" not in out.replace("pref", "")
