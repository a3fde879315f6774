"""The workload generators shared by bench.py, the fixtures and the GPU tests.

The batched-seas workload is drawn from PCG32; workloads.Pcg32 is pinned here
to the library's own generator (the oracle's restatement of
src/core/rng.hpp:35-62, itself pinned to the reference in test_oracle.py)
and to the published PCG32 demo output.
"""
import ctypes as C
from collections import Counter

import numpy as np

from oceanlib_amd.workloads import (CONFIG4_LEVEL_RATIO, Pcg32, checksum_weights,
                                    checksum_words, config4_cells, config4_levels,
                                    config4_seas)
from oracle_ffi import oracle


def test_pcg32_equals_library_generator():
    for seed, stream in ((42, 54), (0xC4, 1), (0, 0), (2**64 - 1, 12345)):
        want = np.zeros(64, np.uint32)
        oracle().orc_pcg32_draws(seed, stream, 64, want.ctypes.data_as(C.POINTER(C.c_uint32)))
        r = Pcg32(seed, stream)
        assert [r.next() for _ in range(64)] == want.tolist()


def test_pcg32_published_demo_values():
    # pcg32-demo (pcg-c-basic), seed 42, sequence 54: first six 32-bit outputs.
    r = Pcg32(42, 54)
    assert [r.next() for _ in range(6)] == [0xa15c02b7, 0x7b47f409, 0xba1d3330,
                                            0x83d2f293, 0xbfa4784b, 0xcbed606e]


def test_pcg32_below_is_in_range_and_covers():
    r = Pcg32(7, 3)
    draws = [r.below(3) for _ in range(3000)]
    assert set(draws) == {0, 1, 2}
    assert all(500 < c < 1500 for c in Counter(draws).values())


def test_config4_shape():
    seas = config4_seas()
    assert len(seas) == 256
    flat, labels = config4_levels()
    assert len(flat) == sum(len(s) for s in seas) == len(labels)
    sizes = Counter(kw["size"] for kw in flat)
    assert set(sizes) == {64, 128, 256, 512}
    for i, sea in enumerate(seas):
        assert 1 <= len(sea) <= 3
        assert all(kw["seed"] == i for kw in sea)
        assert 50.0 * 0.999 <= sea[0]["patch_length"] <= 1000.0 * 1.001
        assert 3.0 <= sea[0]["wind_speed"] <= 25.0
        d = sea[0]["water_depth"]
        assert d == 0.0 or 10.0 <= d <= 200.0
        for k in range(1, len(sea)):
            r = sea[k]["patch_length"] / sea[k - 1]["patch_length"]
            assert abs(r - CONFIG4_LEVEL_RATIO) < 1e-5
            assert sea[k]["size"] == sea[0]["size"]
    depths = [s[0]["water_depth"] == 0.0 for s in seas]
    assert 90 < sum(depths) < 166
    # about 256 seas x 2 levels x mean N^2 (87,040) cells per frame
    assert 35e6 < config4_cells() < 55e6
    assert config4_seas(16) == seas[:16]   # a prefix draws the same seas


def test_checksum_detects_any_single_word_change():
    rng = np.random.default_rng(1)
    w = rng.integers(0, 2**32, 4096, dtype=np.uint64).astype(np.uint32)
    k = checksum_weights(w.size)
    assert np.all(k & np.uint64(1) == 1)
    base = checksum_words(w, k)
    for i in rng.integers(0, w.size, 64):
        for bit in (0, 13, 31):
            v = w.copy()
            v[i] ^= np.uint32(1 << bit)
            assert checksum_words(v, k) != base
    # linear in the words: sum(w_i K_i) mod 2^64, checked against Python ints
    want = sum(int(a) * int(b) for a, b in zip(w[:100], k[:100])) % 2**64
    assert checksum_words(w[:100]) == want
