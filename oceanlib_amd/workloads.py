"""The benchmark workloads, by what they are and their sizes.

Shared by bench.py, the golden-fixture generator and the parity tests so that
all three run the same seas. Descriptor defaults are the reference's
(include/ocean/ocean.hpp:39-130).
"""
from __future__ import annotations

import numpy as np

from oceanlib_amd.capi import default_desc

# Plain single patch: 256^2, 200 m, 12 m/s JONSWAP, seed 1337, chop 1.0.
CONFIG1 = dict(size=256, patch_length=200.0, wind_speed=12.0, seed=1337, choppiness=1.0)

# Single heavy-chop patch: 512^2, 250 m, 22 m/s, chop 1.3, velocity (for the
# advected foam), 50 m depth, seed 7; 600 frames at 1/60 s with 4096 queries.
CONFIG2 = dict(size=512, patch_length=250.0, wind_speed=22.0, choppiness=1.3,
               compute_velocity=1, water_depth=50.0, seed=7)

# Three cascades of 256^2 over 800 / 150 / 25 m at 15 m/s, seed 42.
CONFIG3_LEVELS = [dict(size=256, patch_length=L, wind_speed=15.0, seed=42)
                  for L in (800.0, 150.0, 25.0)]

# Small raw-buffer cases (name, desc kwargs, times).
SMALL_CASES = [
    ("n16_default", dict(size=16, patch_length=20.0, wind_speed=10.0, seed=1337),
     (0.0, 1.0 / 60.0, 10.0)),
    ("n32_vel_depth", dict(size=32, patch_length=40.0, wind_speed=18.0, seed=9, choppiness=1.3,
                           compute_velocity=1, water_depth=5.0), (0.0, 2.5, 3600.0)),
    ("n32_swell_dir", dict(size=32, patch_length=64.0, wind_speed=7.0, seed=3, swell=0.8,
                           wind_direction=2.5, peak_enhancement=1.0, small_wave_cutoff=2.0),
     (0.25, 100.0)),
]


def desc_from(kw):
    return default_desc(None, **kw)


def query_points(frames: int, per_frame: int, extent: float, seed: int = 0xC2) -> np.ndarray:
    """(frames, per_frame, 2) float32 points uniform over [0, extent)^2, exact
    multiples of extent / 2^24 drawn from numpy's PCG64 with a fixed seed."""
    rng = np.random.default_rng(seed)
    u = rng.integers(0, 1 << 24, size=(frames, per_frame, 2)).astype(np.float32)
    return (u / np.float32(1 << 24)) * np.float32(extent)


# Interaction field of the three-cascade scene: 256^2 cells of 0.1 m, kernel
# radius 6 (13x13 taps), other fields at the reference defaults
# (include/ocean/interaction.hpp:64-139).
CONFIG3_INTERACTION = dict(size=256, extent=25.6, kernel_radius=6)


def config3_sources(count: int = 8, seed: int = 0x1A):
    """Eight continuous sources on circles around the grid centre: radius
    0.5-2 m, circle radius 5-20 m, speed 1-6 m/s (fixed numpy PCG64 draws)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        out.append(dict(radius=float(rng.uniform(0.5, 2.0)),
                        circle=float(rng.uniform(5.0, 20.0)),
                        speed=float(rng.uniform(1.0, 6.0)),
                        phase=float(rng.uniform(0.0, 2 * np.pi)),
                        strength=float(rng.uniform(0.02, 0.08))))
    return out


def config3_disturbances(sources, t: float, centre=(12.8, 12.8)):
    """World-space disturbance records for time t: position on the circle and
    the tangential velocity (so a substep sweep draws a continuous wake)."""
    recs = []
    for s in sources:
        w = s["speed"] / s["circle"]
        a = s["phase"] + w * t
        recs.append(dict(world_x=np.float32(centre[0] + s["circle"] * np.cos(a)),
                         world_z=np.float32(centre[1] + s["circle"] * np.sin(a)),
                         radius=np.float32(s["radius"]), strength=np.float32(s["strength"]),
                         velocity_y=np.float32(0.0),
                         velocity_x=np.float32(-s["speed"] * np.sin(a)),
                         velocity_z=np.float32(s["speed"] * np.cos(a)), kind=1))
    return recs


def config3_hull_mask(n: int = 256) -> np.ndarray:
    """One static rectangular hull: cells x in [100, 156), z in [120, 136)."""
    m = np.zeros((n, n), np.uint8)
    m[120 * n // 256:136 * n // 256, 100 * n // 256:156 * n // 256] = 1
    return m


class Pcg32:
    """PCG32 (XSH-RR 64/32, O'Neill 2014), the generator the library itself
    uses for h0 (src/core/rng.hpp:35-62): seeding as pcg32_srandom."""

    _MULT = 6364136223846793005
    _MASK = (1 << 64) - 1

    def __init__(self, seed: int, stream: int):
        self.state = 0
        self.inc = ((stream << 1) | 1) & self._MASK
        self.next()
        self.state = (self.state + seed) & self._MASK
        self.next()

    def next(self) -> int:
        old = self.state
        self.state = (old * self._MULT + self.inc) & self._MASK
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((32 - rot) & 31))) & 0xFFFFFFFF

    def uniform(self) -> float:
        """[0, 1) with 32 random bits."""
        return self.next() / 4294967296.0

    def below(self, n: int) -> int:
        """Integer in [0, n), rejection-sampled (no modulo bias)."""
        threshold = (1 << 32) % n
        while True:
            r = self.next()
            if r >= threshold:
                return r % n


# Batched uneven seas: 256 seas from PCG32(seed 0xC4, stream 1), drawn per sea
# in this order: N uniform over {64, 128, 256, 512}; patch length
# log-uniform over [50, 1000] m; wind uniform over [3, 25] m/s; depth infinite
# or uniform over [10, 200] m (50/50); 1-3 cascade levels. Level k of sea i
# has patch length L * CONFIG4_LEVEL_RATIO^k and seed i. 300 frames at 1/60 s.
CONFIG4_SEAS = 256
CONFIG4_FRAMES = 300
CONFIG4_LEVEL_RATIO = 0.19


def config4_seas(count: int = CONFIG4_SEAS, seed: int = 0xC4):
    """List of seas; each sea is a list of level kwargs for desc_from."""
    rng = Pcg32(seed, 1)
    seas = []
    for i in range(count):
        n = (64, 128, 256, 512)[rng.below(4)]
        length = 50.0 * 20.0 ** rng.uniform()
        wind = 3.0 + 22.0 * rng.uniform()
        depth = 0.0 if rng.uniform() < 0.5 else 10.0 + 190.0 * rng.uniform()
        levels = 1 + rng.below(3)
        seas.append([dict(size=n, patch_length=float(np.float32(length * CONFIG4_LEVEL_RATIO ** k)),
                          wind_speed=float(np.float32(wind)), water_depth=float(np.float32(depth)),
                          seed=i, thread_count=1)
                     for k in range(levels)])
    return seas


def config4_levels(count: int = CONFIG4_SEAS):
    """The seas' levels flattened in sea order, with (sea, level) labels."""
    flat, labels = [], []
    for i, sea in enumerate(config4_seas(count)):
        for k, kw in enumerate(sea):
            flat.append(kw)
            labels.append((i, k))
    return flat, labels


def config4_cells(count: int = CONFIG4_SEAS) -> int:
    return sum(kw["size"] ** 2 for kw in config4_levels(count)[0])


# Buffer checksum used for the batched-seas fixtures: the level's output
# buffers (displacement, then normal, then velocity when present) read as
# little-endian uint32 words w_i, summed as sum(w_i * K_i) mod 2^64 with odd
# K_i = (i * 0x9E3779B97F4A7C15 + 0x632BE59BD9B4E019) | 1. Any single changed
# word changes the sum (K_i is odd, so w * K_i == 0 mod 2^64 forces w == 0).
CHECKSUM_STEP = 0x9E3779B97F4A7C15
CHECKSUM_BASE = 0x632BE59BD9B4E019


def checksum_weights(words: int) -> np.ndarray:
    i = np.arange(words, dtype=np.uint64)
    return (i * np.uint64(CHECKSUM_STEP) + np.uint64(CHECKSUM_BASE)) | np.uint64(1)


def checksum_words(words: np.ndarray, weights: np.ndarray | None = None) -> int:
    w = np.ascontiguousarray(words).view(np.uint32).astype(np.uint64).ravel()
    k = checksum_weights(w.size) if weights is None else weights
    return int(np.sum(w * k, dtype=np.uint64))
