"""The arithmetic the FP32 kernel uses on 1-byte bounded-density codes
(code = material * bins + bin over a material map of up to 64 entries,
bins = 255 // entries; code 255 marks mixed blocks): the material of a code
is (code * codeMul) >> 20 with codeMul = ceil(2^20 / bins), and the block
edge of a world is the smallest of 8, 16, 32 and 64 voxels whose table of
2-byte block codes fits 64 KiB. The GPU tests check the kernels against
every-step evaluation; these check the arithmetic over its whole domain."""
import numpy as np

SHIFT = 20


def code_mul(bins):
    return ((1 << SHIFT) + bins - 1) // bins


def test_material_of_code_by_multiply_and_shift():
    for entries in range(1, 65):
        bins = 255 // entries
        mul = code_mul(bins)
        assert mul < 1 << 24  # a 24-bit multiply operand
        codes = np.arange(entries * bins, dtype=np.uint64)
        assert np.array_equal((codes * mul) >> SHIFT, codes // bins)
        assert int(codes.max() * mul) < 1 << 32  # the product fits 32 bits
    # the largest code of a mixed block (0) names a valid row
    assert (0 * code_mul(1)) >> SHIFT == 0


def test_a_multiply_that_rounds_down_fails():
    # floor(2^20 / bins) instead of the ceiling misplaces the first code of
    # some material (what the ceiling's argument rules out)
    bad = [bins for bins in range(1, 256)
           if any(((m * bins) * ((1 << SHIFT) // bins)) >> SHIFT != m for m in range(1, 255 // bins))]
    assert bad, "the floor would have to fail for some bin count"


def block_edge(shape):
    for edge in (8, 16, 32, 64):
        if 2 * int(np.prod([-(-s // edge) for s in shape])) <= 65536:
            return edge
    return 64


def test_block_edge_rule():
    assert block_edge((300, 512, 512)) == 16  # CT spiral phantom
    assert block_edge((600, 1024, 1024)) == 32  # whole body
    assert block_edge((45, 50, 53)) == 8
    assert block_edge((64, 520, 530)) == 16
    assert block_edge((4096, 4096, 4096)) == 64
