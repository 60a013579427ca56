"""Synthetic workloads for the ANQI reverse-kNN path (deterministic, numpy).

The generator is a counter-based SplitMix64 stream plus Box-Muller, written
out here so the bytes do not depend on numpy's Generator implementation.
Each workload is identified by (seed, stream); the SHA-256 of the float32
bytes is recorded next to every result.

Workloads (sizes as used by the tests and the benchmark):
  * gaussian(n, d, seed): i.i.d. N(0, 1) float32 rows (the 256 x 128 base and
    32 x 128 queries of the CPU-runnable sample, seeds 0 and 1).
  * mixture(n, d, centres, sigma, seed, ...): rows drawn from a Gaussian
    mixture whose centres are uniform in [0, 128)^d (centre seed 0), rounded to
    integers with round-half-even and clipped to [0, 255]; queries come from
    the same centres with a separate seed.  1,000 centres and sigma 20 give
    the million-row 1,000,000 x 128 base.
"""
from __future__ import annotations

import hashlib
import os
import struct

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_STREAM = np.uint64(0xD1B54A32D192ED03)


def _mix(z: np.ndarray) -> np.ndarray:
    z = z.copy()
    z ^= z >> np.uint64(30)
    z *= _M1
    z ^= z >> np.uint64(27)
    z *= _M2
    z ^= z >> np.uint64(31)
    return z


def u64(seed: int, stream: int, start: int, count: int) -> np.ndarray:
    """Counter-based SplitMix64 draws [start, start+count) of (seed, stream)."""
    with np.errstate(over="ignore"):
        key = _mix(np.array([np.uint64(seed) ^ (np.uint64(stream) * _STREAM)], dtype=np.uint64))[0]
        ctr = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        return _mix(key + ctr * _GOLDEN)


def uniform(seed: int, stream: int, start: int, count: int) -> np.ndarray:
    """Doubles in [0, 1) from the top 53 bits."""
    return (u64(seed, stream, start, count) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, stream: int, start: int, count: int) -> np.ndarray:
    """Box-Muller normals; draw pair p uses counters 2p, 2p+1 of the stream."""
    pairs = (count + 1) // 2
    first = start // 2
    raw = u64(seed, stream, 2 * first, 2 * pairs + 2)
    u1 = ((raw[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (2.0 ** -53)
    u2 = (raw[1::2] >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * len(r), dtype=np.float64)
    z[0::2] = r * np.cos(2.0 * np.pi * u2)
    z[1::2] = r * np.sin(2.0 * np.pi * u2)
    off = start - 2 * first
    return z[off:off + count]


def gaussian(n: int, d: int, seed: int) -> np.ndarray:
    return normal(seed, 1, 0, n * d).astype(np.float32).reshape(n, d)


def mixture_centres(centres: int, d: int, seed: int = 0) -> np.ndarray:
    return (uniform(seed, 2, 0, centres * d) * 128.0).reshape(centres, d)


def mixture(n: int, d: int, centres: int, sigma: float, seed: int,
            centre_seed: int = 0, chunk: int = 1 << 16, rounded: bool = True, threads: int = 0) -> np.ndarray:
    """n rows of clip(round(c_u + sigma * N(0, I)), 0, 255), u uniform.
    Chunks are independent counter ranges, so they are generated on a thread
    pool (threads 0: min(16, cores)); the bytes do not depend on it."""
    from concurrent.futures import ThreadPoolExecutor
    c = mixture_centres(centres, d, centre_seed)
    out = np.empty((n, d), dtype=np.float32)

    def fill(s):
        e = min(n, s + chunk)
        pick = (u64(seed, 3, s, e - s) % np.uint64(centres)).astype(np.int64)
        z = normal(seed, 4, s * d, (e - s) * d).reshape(e - s, d)
        v = c[pick] + sigma * z
        if rounded:
            v = np.rint(v)
        out[s:e] = (np.clip(v, 0.0, 255.0) + 0.0).astype(np.float32)

    workers = threads or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(fill, range(0, n, chunk)))
    return out


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def write_fbin(path: str, a: np.ndarray) -> None:
    a = np.ascontiguousarray(a, dtype=np.float32)
    with open(path, "wb") as f:
        f.write(struct.pack("<II", a.shape[0], a.shape[1]))
        f.write(a.tobytes())


def read_fbin(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        n, d = struct.unpack("<II", f.read(8))
        return np.frombuffer(f.read(n * d * 4), dtype=np.float32).reshape(n, d).copy()


# Named workloads.
def sample_base() -> np.ndarray:
    """256 x 128 N(0,1), seed 0."""
    return gaussian(256, 128, 0)


def sample_query() -> np.ndarray:
    """32 x 128 N(0,1), seed 1."""
    return gaussian(32, 128, 1)


def million_base(n: int = 1_000_000) -> np.ndarray:
    """1,000-centre mixture, sigma 20, seed 0 (first n rows)."""
    return mixture(n, 128, 1000, 20.0, seed=0)


def million_query(nq: int = 10_000) -> np.ndarray:
    """Queries from the same 1,000 centres, seed 1."""
    return mixture(nq, 128, 1000, 20.0, seed=1)
