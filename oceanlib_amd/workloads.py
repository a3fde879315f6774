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
