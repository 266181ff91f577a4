"""Seeded inputs identical to GenDiL's FillRandom (include/gendil/Algebra/vector.hpp:632-645):
std::mt19937(seed) driving std::uniform_real_distribution<double>(min, max), whose
libstdc++ generate_canonical<double, 53> consumes two 32-bit draws per value:
u = (g1 + g2 * 2^32) / 2^64 (clamped below 1), value = (max - min) * u + min.
"""
from __future__ import annotations

import numpy as np


def fill_random(n: int, seed: int = 12345, min_value: float = -1.0, max_value: float = 1.0) -> np.ndarray:
    bg = np.random.MT19937()
    bg._legacy_seeding(int(seed) & 0xFFFFFFFF)  # init_genrand(seed), as std::mt19937(seed)
    raw = bg.random_raw(2 * int(n)).astype(np.float64)
    u = (raw[0::2] + raw[1::2] * 4294967296.0) / 18446744073709551616.0
    u = np.where(u >= 1.0, np.nextafter(1.0, 0.0), u)
    return (max_value - min_value) * u + min_value
