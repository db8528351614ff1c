"""Deterministic synthetic inputs for the LU factor-and-solve workloads.

The generator uses only IEEE-754 basic operations (+, -, *, /, sqrt, frexp) so
that the same seed yields the same bits on every x86-64 host, whatever SIMD
paths numpy's transcendental functions take there.  That is what lets a pivot
vector computed once by the CPU oracle be committed as a fixture and compared
with the GPU result on another machine.

* uniforms: SplitMix64 (Steele, Lea & Flood 2014) -> top 53 bits -> [0, 1);
* normals: Marsaglia's polar method with a series logarithm
  (log m = 2 atanh((m-1)/(m+1)), |s| <= 0.1716, 12 odd terms).
"""
import hashlib

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_LN2 = 0.6931471805599453


def splitmix_uniform(seed, count, offset=0):
    """count uniforms in [0, 1) from SplitMix64 state seed, starting after
    `offset` draws."""
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _log(x):
    """Natural log of positive finite doubles from basic operations only."""
    m, e = np.frexp(x)
    small = m < 0.7071067811865476
    m = np.where(small, m * 2.0, m)
    e = np.where(small, e - 1, e).astype(np.float64)
    s = (m - 1.0) / (m + 1.0)
    t = s * s
    p = np.full_like(s, 1.0 / 23.0)
    for k in range(21, 0, -2):
        p = p * t + 1.0 / k
    return e * _LN2 + 2.0 * s * p


def normal(seed, count):
    """count N(0,1) samples (Marsaglia polar method), deterministic."""
    out = np.empty(count, np.float64)
    filled = 0
    offset = 0
    while filled < count:
        need = count - filled
        pairs = int(need / 2 / 0.78) + 64
        u = splitmix_uniform(seed, 2 * pairs, offset)
        offset += 2 * pairs
        v1 = 2.0 * u[0::2] - 1.0
        v2 = 2.0 * u[1::2] - 1.0
        s = v1 * v1 + v2 * v2
        ok = (s > 0.0) & (s < 1.0)
        v1, v2, s = v1[ok], v2[ok], s[ok]
        f = np.sqrt(-2.0 * _log(s) / s)
        z = np.empty(2 * len(s), np.float64)
        z[0::2] = v1 * f
        z[1::2] = v2 * f
        take = min(need, len(z))
        out[filled:filled + take] = z[:take]
        filled += take
    return out


def gaussian_matrix(seed, rows, cols):
    """rows x cols row-major float64 matrix with i.i.d. N(0,1) entries."""
    return normal(seed, rows * cols).reshape(rows, cols)


def digest(a):
    """Short SHA-256 of an array's bytes (C order)."""
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]
