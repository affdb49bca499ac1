"""Key plan of the linearised (Z-order) layout, no device needed.

Pinned by the bit plans the reference documents for its Morton layout
(src/dynasor_morton.cpp:92-102: N=3 with I=(8192,8192,8192) gives 13 bits per
mode interleaved x,y,z from bit 0; N=4 with I=(16,16,4096,16) gives
b=(4,4,12,4), 24 bits), by its per-mode bit count (ceil(log2 I), at least 1,
src/dynasor_morton.cpp:83-89), and by the invariants of any such plan.  Beyond
64 bits the reference refuses the tensor; the device layout instead drops the
lowest rank bits of the largest modes until 64 remain.
"""
import numpy as np
import pytest

from dynasor_plus_amd.engine import linear_plan


def positions(m):
    return [b for b in range(64) if (m >> b) & 1]


def test_three_equal_modes_interleave_from_bit_zero():
    mask, bits, shift, total = linear_plan((8192, 8192, 8192))
    assert total == 39 and bits == [13, 13, 13] and shift == [0, 0, 0]
    assert positions(mask[0]) == list(range(0, 39, 3))
    assert positions(mask[1]) == list(range(1, 39, 3))
    assert positions(mask[2]) == list(range(2, 39, 3))


def test_unequal_modes_fill_the_top_bits_with_the_largest():
    mask, bits, shift, total = linear_plan((16, 16, 4096, 16))
    assert bits == [4, 4, 12, 4] and total == 24
    assert positions(mask[0]) == [0, 4, 8, 12]
    assert positions(mask[1]) == [1, 5, 9, 13]
    assert positions(mask[3]) == [3, 7, 11, 15]
    assert positions(mask[2]) == [2, 6, 10, 14] + list(range(16, 24))


@pytest.mark.parametrize("dims,want_bits", [((1, 2, 3), [1, 1, 2]), ((5, 8, 9), [3, 3, 4]),
                                            ((2 ** 20, 2 ** 20 + 1), [20, 21])])
def test_bits_are_ceil_log2_at_least_one(dims, want_bits):
    _, bits, shift, total = linear_plan(dims)
    assert bits == want_bits and sum(shift) == 0 and total == sum(want_bits)


def test_more_than_64_bits_shifts_the_largest_modes():
    # the 5-mode 500k x 100k x 50k x 5k x 100 workload: 19+17+16+13+7 = 72 bits
    mask, bits, shift, total = linear_plan((500_000, 100_000, 50_000, 5_000, 100))
    assert bits == [19, 17, 16, 13, 7]
    assert total == 64
    assert shift == [5, 2, 1, 0, 0]


@pytest.mark.parametrize("seed", range(6))
def test_plan_invariants(seed):
    rng = np.random.default_rng(seed)
    N = int(rng.integers(1, 9))
    dims = [int(x) for x in rng.integers(1, 2 ** 31, size=N)]
    mask, bits, shift, total = linear_plan(dims)
    assert total <= 64
    union = 0
    for n in range(N):
        assert bits[n] == max(1, (dims[n] - 1).bit_length())
        assert union & mask[n] == 0
        union |= mask[n]
        assert bin(mask[n]).count("1") == bits[n] - shift[n]
    assert union == (1 << total) - 1
    if sum(bits) <= 64:
        assert sum(shift) == 0 and total == sum(bits)
    else:
        # each removed bit came from a mode that then had the most bits left
        kept = [b - s for b, s in zip(bits, shift)]
        for n in range(N):
            if shift[n]:
                assert kept[n] >= max(kept) - 1


def test_invalid_arguments():
    with pytest.raises(ValueError):
        linear_plan((10, 0, 4))
    with pytest.raises(ValueError):
        linear_plan(tuple(range(1, 11)))
