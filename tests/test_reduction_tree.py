"""The register-resident wave reductions (seg_sum in drot_device.h) add the
same pairs in the same order as the lane-xor butterfly they replace, so every
fp64 sum keeps its bits.  The lane movements are emulated here on 64-element
vectors:

  v_permlane32_swap / v_permlane16_swap with both operands equal: lane i holds
      {v_i, v_(i^32)} / {v_i, v_(i^16)};
  DPP row_ror:k: lane i reads lane 16*(i//16) + (i%16 + k) % 16 (the other
      rotation direction reads the same values for k = 8, and for k = 4 once
      the values repeat every 8 lanes, which is the only way it is used);
  DPP quad_perm [2,3,0,1] and [1,0,3,2]: xor 2 and xor 1 inside a quad.

row_ror:8 is xor 8 within a 16-lane row; row_ror:4 reads a different lane than
xor 4 in general, but after the xor-8 step the values repeat every 8 lanes, and
then both read the same value.  A plausible slip (rotating by 4 first, or
rotating before the row swaps) breaks the equality, which the last test shows.
"""
import numpy as np
import pytest

LANES = np.arange(64)


def xor_add(v, off):
    return v + v[LANES ^ off]


def swap_add(v, off):
    # {a, b} = {v_i, v_(i^off)}; the device adds a + b, which equals b + a
    a = np.where((LANES & off) == 0, v, v[LANES ^ off])
    b = np.where((LANES & off) == 0, v[LANES ^ off], v)
    return a + b


def ror_add(v, k):
    return v + v[(LANES // 16) * 16 + (LANES % 16 + k) % 16]


def quad_add(v, perm):
    return v + v[(LANES // 4) * 4 + np.asarray(perm)[LANES % 4]]


def butterfly(v, width):
    off = width // 2
    while off >= 1:
        v = xor_add(v, off)
        off //= 2
    return v


def register_sum(v, width):
    if width >= 64:
        v = swap_add(v, 32)
    if width >= 32:
        v = swap_add(v, 16)
    v = ror_add(v, 8)
    v = ror_add(v, 4)
    v = quad_add(v, [2, 3, 0, 1])
    return quad_add(v, [1, 0, 3, 2])


def _vectors(n, seed):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        yield rng.standard_normal(64) * 10.0 ** rng.uniform(-12, 12, 64)


@pytest.mark.parametrize("width", [16, 32, 64])
def test_register_sum_has_the_butterfly_bits(width):
    for v in _vectors(3000, width):
        a, b = butterfly(v.copy(), width), register_sum(v.copy(), width)
        assert np.array_equal(a.view(np.int64), b.view(np.int64))
        # every lane of a segment holds the same total
        seg = a.reshape(-1, width)
        assert (seg == seg[:, :1]).all()


def test_wrong_order_changes_the_bits():
    bad = 0
    for v in _vectors(300, 7):
        a = butterfly(v.copy(), 16)
        b = ror_add(v.copy(), 4)  # rotating by 4 before the xor-8 step
        b = ror_add(b, 8)
        b = quad_add(quad_add(b, [2, 3, 0, 1]), [1, 0, 3, 2])
        bad += not np.array_equal(a.view(np.int64), b.view(np.int64))
    assert bad > 0
