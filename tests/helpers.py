"""Helpers shared by the tests: fixtures on disk, float comparisons."""

from __future__ import annotations

import json
import os
from fractions import Fraction

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_json(name: str):
    with open(os.path.join(GOLDEN, name)) as handle:
        return json.load(handle)


def golden_kernels():
    data = np.load(os.path.join(GOLDEN, "kernels.npz"), allow_pickle=False)
    return {k.replace("__", "/"): data[k] for k in data.files}


def program_tuples(program):
    return [(k, v) for k, v in program]


def ulp_distance(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Units in the last place between two float arrays of the same dtype, in exact integer
    arithmetic; NaN against NaN is 0, NaN against a number is a huge distance."""
    a = np.asarray(a)
    b = np.asarray(b)
    kind, mask = (np.int32, 0x7FFFFFFF) if a.dtype == np.float32 else (np.int64, 0x7FFFFFFFFFFFFFFF)
    ia = a.view(kind).astype(object)
    ib = b.view(kind).astype(object)
    order = np.vectorize(lambda i: -(i & mask) if i < 0 else i, otypes=[object])
    d = np.abs(order(ia) - order(ib)).astype(np.float64)
    nan_a, nan_b = np.isnan(a), np.isnan(b)
    return np.where(nan_a & nan_b, 0.0, np.where(nan_a | nan_b, float(1 << 40), d))


def same_bits(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Bit-for-bit equality, except that any NaN equals any NaN (payloads are not part of the contract)."""
    kind = np.uint32 if a.dtype == np.float32 else np.uint64
    return (a.view(kind) == b.view(kind)) | (np.isnan(a) & np.isnan(b))


def round_to_float32(value: Fraction) -> np.float32:
    """Correctly rounded (nearest, ties to even) float32 of an exact rational."""
    if value == 0:
        return np.float32(0.0)
    sign = -1 if value < 0 else 1
    v = abs(value)
    e = v.numerator.bit_length() - v.denominator.bit_length()
    if Fraction(2) ** e > v:
        e -= 1
    e = max(e, -126)  # subnormals share the smallest exponent
    if e > 127:
        return np.float32(sign * np.inf)
    scale = Fraction(2) ** (23 - e)
    scaled = v * scale
    q, r = divmod(scaled.numerator, scaled.denominator)
    twice = 2 * r
    if twice > scaled.denominator or (twice == scaled.denominator and q % 2 == 1):
        q += 1
    result = Fraction(q) / scale
    if result >= Fraction(2) ** 128:
        return np.float32(sign * np.inf)
    return np.float32(sign * float(result))


def fma32(a, b, c) -> np.float32:
    """fmaf(a, b, c) rounded once, from exact rationals."""
    return round_to_float32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))
