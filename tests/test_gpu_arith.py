"""The fast correctly rounded sqrt and reciprocal used by the step kernels,
checked bit for bit against the compiler's correctly rounded sqrtf and
1.0f / x for EVERY float in their ranges (about 1.9e9 and 1.1e9 inputs),
on the GPU."""
import ctypes

import numpy as np
import pytest

import mercury_amd as mh
from mercury_amd._lib import check


def bits(x):
    return int(np.float32(x).view(np.uint32))


def mismatches(which, lo, hi):
    n = ctypes.c_uint64()
    check(mh.load().mh_check_arith(0, which, lo, hi - lo, ctypes.byref(n)))
    return n.value


@pytest.mark.gpu
def test_fast_sqrt_is_correctly_rounded_over_its_whole_range():
    # [2^-96, 2^128): every finite float from 2^-96 up (the force solver's
    # fast path takes all of them)
    print("arith: sqrt_rn_inrange over [2^-96, 2^128)", flush=True)
    lo, hi = bits(2.0 ** -96), 0x7f800000
    assert mismatches(1, lo, hi) == 0
    # informational: how often the uncorrected hardware sqrt differs
    print("hardware sqrt mismatches in range:", mismatches(0, lo, hi), flush=True)


@pytest.mark.gpu
def test_check_arith_counts_every_difference():
    print("arith: checker self-test", flush=True)
    lo, hi = bits(1.0), bits(4.0)
    assert mismatches(3, lo, hi) == hi - lo
    print("hardware rcp mismatches in [1, 4):", mismatches(2, lo, hi), flush=True)


@pytest.mark.gpu
def test_fast_division_is_correctly_rounded():
    """mh::div_rn_from_rcp (reciprocal + two fused corrections) against IEEE
    a / b on 2^31 pseudo-random pairs from the range the XPBD gather kernel
    admits to its short path (numerators 0 or of magnitude [2^-79, 2^63),
    divisors [2^-48, 2^40))."""
    print("arith: fast division, 2^31 pairs", flush=True)
    assert mismatches(9, 0, 1 << 31) == 0


@pytest.mark.gpu
def test_fast_reciprocal_is_correctly_rounded_over_its_whole_range():
    """mh::rcp_rn_inrange (hardware reciprocal + one fused Newton step) for
    every float in [2^-64, 2^64], which holds d + 1e-12 for every d the
    fast sqrt returns."""
    print("arith: rcp_rn_inrange over [2^-64, 2^64]", flush=True)
    lo, hi = bits(2.0 ** -64), bits(2.0 ** 64) + 1
    assert mismatches(4, lo, hi) == 0
