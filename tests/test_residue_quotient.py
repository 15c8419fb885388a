"""The int8 residual's per-modulus residue step (oz_residue16 in
nextla.jl_amd/csrc/nla_ozaki.hip), restated with numpy's IEEE double
arithmetic, which is what the device's v_mul_f64 / round-to-nearest-even
compute: for every modulus of the table and every x the limb sums reach
(|x| < 2^25), the low byte of x - rint(x * RN(1/p)) * p is the low byte of
x's centred residue in [rhi - p + 1, rhi], rhi = floor((p - 1) / 2).

x / p can only round the wrong way next to a half-integer, so besides a
random sample the test takes every x whose residue is an end of the centred
range (the values closest to a half-integer), over the whole |x| < 2^25.
A float32 quotient, the planted mistake, fails the same check.
"""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "nextla.jl_amd", "csrc", "nla_internal.h")
XMAX = 1 << 25


def moduli():
    txt = open(HDR).read()
    mt = re.search(r"OZ_MODULI\[OZ_TMAX\]\s*=\s*\{([^}]*)\}", txt)
    assert mt, "moduli table not found"
    return [int(v) for v in mt.group(1).replace("\n", " ").split(",") if v.strip()]


def centred(x, p):
    rhi = (p - 1) // 2
    lo = rhi - p + 1
    return (x - lo) % p + lo


def candidates(p, rng):
    rhi = (p - 1) // 2
    q = np.arange(-(XMAX // p) - 1, XMAX // p + 2, dtype=np.int64)
    xs = [q * p + rhi, q * p + rhi - p + 1, q * p, rng.integers(-XMAX, XMAX, 1 << 20)]
    x = np.concatenate(xs)
    return x[(x > -XMAX) & (x < XMAX)]


def byte_of(x, p, dtype):
    inv = dtype(1.0) / dtype(p)
    q = np.rint(x.astype(dtype) * inv).astype(np.int64)
    return (x - q * p) & 255


def test_table_is_256_then_odd():
    P = moduli()
    assert P[0] == 256 and all(p % 2 == 1 and p < 256 for p in P[1:])
    assert len(set(P)) == len(P)


@pytest.mark.parametrize("p", moduli())
def test_double_quotient_gives_the_centred_byte(p):
    x = candidates(p, np.random.default_rng(p))
    assert np.array_equal(byte_of(x, p, np.float64), centred(x, p) & 255)


def test_float_quotient_is_caught():
    """The check itself: a float32 quotient without correction steps is off
    by one next to half-integers for the large x the limbs reach."""
    bad = 0
    for p in moduli()[1:]:
        x = candidates(p, np.random.default_rng(p))
        bad += int(np.count_nonzero(byte_of(x, p, np.float32) != (centred(x, p) & 255)))
    assert bad > 0
