"""Host mirror of the device's synthetic-input generator.

TEST INFRASTRUCTURE ONLY (same rules as ``tenet_oracle``): the checker uses
it to regenerate, on the CPU, any slice of the seeded inputs that the
library's ``tenet_fill_gaussian`` wrote straight into device memory, so that
large workloads need no host-to-device upload and can still be checked.

Element i of a stream (seed s, offset o) is, with k = 2 (o + i),
    h(x)  = splitmix64 finaliser of x (mod 2^64)
    u1    = (h(k * G + s * M) >> 11) * 2^-53,  u2 = same with k + 1
    z     = sqrt(-2 ln(1 - u1)) * (cos 2 pi u2 + i sin 2 pi u2) / sqrt(2)
with G = 0x9E3779B97F4A7C15 and M = 0xD1B54A32D192ED03 (Box-Muller on a
counter-based uniform stream).  The integer part is bit-exact; the float
part agrees to a few ulp (libm vs device sin/cos/log).
"""
from __future__ import annotations

import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_M = np.uint64(0xD1B54A32D192ED03)


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniforms(seed: int, offset: int, count: int):
    """(u1, u2) float64 arrays in [0, 1) for elements offset .. offset+count-1."""
    with np.errstate(over="ignore"):
        i = np.arange(count, dtype=np.uint64) + np.uint64(offset)
        k = np.uint64(2) * i
        sm = np.uint64(seed) * _M
        a = _mix64(k * _G + sm) >> np.uint64(11)
        b = _mix64((k + np.uint64(1)) * _G + sm) >> np.uint64(11)
    return a.astype(np.float64) * 2.0 ** -53, b.astype(np.float64) * 2.0 ** -53


def gaussian(seed: int, count: int, offset: int = 0, dtype=np.complex128):
    """Standard complex Gaussians, element-for-element as on the device."""
    u1, u2 = uniforms(seed, offset, count)
    r = np.sqrt(-2.0 * np.log1p(-u1)) * np.sqrt(0.5)
    z = r * np.cos(2.0 * np.pi * u2) + 1j * (r * np.sin(2.0 * np.pi * u2))
    return z.astype(dtype)
