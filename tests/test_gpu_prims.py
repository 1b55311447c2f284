"""The reference's GPU-primitive specs (UnitTests/cuda/scan.cpp:21-73,
UnitTests/cuda/reduce.cpp:20-51), restated against the library's wave64
workgroup scan and reduction (kifi_prims.hpp, through kifi_segmented_*):
ragged sizes 182,094 / 109,487 / 182,742 with segments of 256 / 1024 / 256,
values below 531 / 491 / 319, exact integer comparison with the CPU result
(std::partial_sum, util::partial_sum_exclusive, std::accumulate per segment).
Segments of one wave (64) and sums near 2^32 are added."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _seg_scan(x, seg, inclusive):
    out = np.empty_like(x)
    for i in range(0, len(x), seg):
        c = np.cumsum(x[i:i + seg], dtype=np.uint64).astype(np.uint32)
        out[i:i + seg] = c if inclusive else np.concatenate([[0], c[:-1]]).astype(np.uint32)
    return out


def test_segmented_inclusive_scan(ctx):                       # scan.cpp:23-48
    x = np.random.default_rng(1).integers(0, 531, 182094).astype(np.uint32)
    assert np.array_equal(ctx.segmented("inclusive", x, 256), _seg_scan(x, 256, True))


def test_segmented_exclusive_scan(ctx):                       # scan.cpp:50-72
    x = np.random.default_rng(2).integers(0, 491, 109487).astype(np.uint32)
    assert np.array_equal(ctx.segmented("exclusive", x, 1024), _seg_scan(x, 1024, False))


def test_segmented_reduce(ctx):                               # reduce.cpp:22-51
    x = np.random.default_rng(3).integers(0, 319, 182742).astype(np.uint32)
    want = np.add.reduceat(x.astype(np.uint64), np.arange(0, len(x), 256)).astype(np.uint32)
    assert np.array_equal(ctx.segmented("reduce", x, 256), want)


@pytest.mark.parametrize("seg", [64, 256, 1024])
def test_every_segment_size_and_wraparound(ctx, seg):
    rng = np.random.default_rng(seg)
    x = rng.integers(0, 2 ** 32, 3 * seg + 17, dtype=np.uint64).astype(np.uint32)   # sums wrap modulo 2^32
    assert np.array_equal(ctx.segmented("inclusive", x, seg), _seg_scan(x, seg, True))
    assert np.array_equal(ctx.segmented("exclusive", x, seg), _seg_scan(x, seg, False))
    want = np.add.reduceat(x.astype(np.uint64), np.arange(0, len(x), seg)).astype(np.uint32)
    assert np.array_equal(ctx.segmented("reduce", x, seg), want)
