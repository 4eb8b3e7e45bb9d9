"""Placement rule of nt2_amd.layout (host logic only)."""
import random

from nt2_amd.layout import COLORS, COLOR_STEP, color_of, slab_bytes, stream_offsets


def test_offsets_colour_alignment_and_fit():
    rng = random.Random(7)
    bases = [0, 512, 4096 * 5 + 512, 2 << 20] + [rng.randrange(0, 1 << 34, 256) for _ in range(300)]
    for base in bases:
        for nb in (64 << 20, 1 << 20, (3 << 20) + 12, 5_000_000):
            for first in (0, 7, 15):
                for cnt in (1, 2, 6, 12, 20):
                    o = stream_offsets(cnt, nb, base, first)
                    assert [color_of(base + x) for x in o] == [(first + k) % COLORS for k in range(cnt)]
                    assert all((base + x) % COLOR_STEP == 0 for x in o)
                    assert o[0] >= 0 and o[-1] + nb <= slab_bytes(cnt, nb)
                    assert all(o[i + 1] - o[i] >= nb for i in range(cnt - 1))  # no overlap


def test_pairwise_distinct_modulo_segment():
    # twelve 64 MiB tables (the Black-Scholes adjoint's streams): every pair
    # differs modulo 2 MiB, which is what separate allocations fail to give
    o = stream_offsets(12, 64 << 20, 0, 0)
    assert len({x % (2 << 20) for x in o}) == 12


def test_empty():
    assert stream_offsets(0, 1 << 20) == [] and slab_bytes(0, 1 << 20) == 0
