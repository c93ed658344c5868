"""CPU restatement of the seeded Gaussian Omega -- TEST INFRASTRUCTURE ONLY.

The reference draws Omega from one Gaussian stream laid out row-major (Omega[j, c] = g[j Nl + c],
src/lib/libisvd/core/stage/@x@_sketch.h:124-132, src/lib/libisvd/la/vsl/@x@_rng_gaussian_driver.c:
16-41) with MKL's SFMT19937 + Box-Muller, which is not reproducible without MKL.  The library
keeps that layout (element e of the row-major n x Nl Omega, i.e. of the column-major Nl x n
Omega^T, is element e of the stream) and draws it with a counter-based generator so that any
element can be formed independently on the GPU:

  Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
  SC'11): counter (e mod 2^32, e div 2^32, 0, 0), key (seed mod 2^32, seed div 2^32), ten
  rounds with multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key increments 0x9E3779B9 /
  0xBB67AE85;
  Box-Muller: a = w0 w1 and b = w2 w3 as 64-bit words, u1 = (a >> 11 + 1) 2^-53,
  u2 = (b >> 11 + 1) 2^-53 (uniforms in (0, 1]), z = sqrt(-2 ln u1) cos(2 pi u2) in double;
  the single-precision stream is z rounded to float.

How it is pinned (tests/test_oracle_philox.py): philox4x32_10 against the published known-answer
vectors of the Random123 distribution and against rocRAND's philox4x32_10 engine (vectors in
tests/golden/philox_vectors.json, made by tests/golden/make_philox_vectors.py from the rocRAND
header's ten_rounds on the host); the Gaussian transform by its moments and its closed form at
chosen words.  Only tests/ may import this module.
"""
import numpy as np

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
_MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """Ten Philox-4x32 rounds.  ctr: uint32 array (..., 4); key: uint32 array (..., 2) (broadcast).
    Returns uint32 (..., 4)."""
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.broadcast_to(np.asarray(key, dtype=np.uint32), c.shape[:-1] + (2,))
    c0, c1, c2, c3 = (c[..., i].astype(np.uint64) for i in range(4))
    k0, k1 = k[..., 0].astype(np.uint64), k[..., 1].astype(np.uint64)
    for _ in range(10):
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK32
        n0 = hi1 ^ c1 ^ k0
        n2 = hi0 ^ c3 ^ k1
        c0, c1, c2, c3 = n0 & _MASK32, lo1, n2 & _MASK32, lo0
        k0 = (k0 + _W0) & _MASK32
        k1 = (k1 + _W1) & _MASK32
    return np.stack([c0, c1, c2, c3], axis=-1).astype(np.uint32)


def box_muller(w):
    """Gaussians (float64) from Philox words w (..., 4): z = sqrt(-2 ln u1) cos(2 pi u2)."""
    w = np.asarray(w, dtype=np.uint64)
    a = (w[..., 0] << np.uint64(32)) | w[..., 1]
    b = (w[..., 2] << np.uint64(32)) | w[..., 3]
    u1 = ((a >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = ((b >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586476925 * u2)


def gaussian_stream(seed, count, start=0):
    """Elements start .. start + count - 1 of the seeded stream (float64)."""
    e = np.arange(start, start + count, dtype=np.uint64)
    ctr = np.zeros((count, 4), dtype=np.uint32)
    ctr[:, 0] = (e & _MASK32).astype(np.uint32)
    ctr[:, 1] = (e >> np.uint64(32)).astype(np.uint32)
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    key = np.array([seed & 0xFFFFFFFF, seed >> 32], dtype=np.uint32)
    return box_muller(philox4x32_10(ctr, key))


def omega(seed, n, Nl, dtype=np.float64):
    """The seeded Omega (n x Nl, row-major: Omega[j, c] = stream[j Nl + c])."""
    return gaussian_stream(seed, n * Nl).reshape(n, Nl).astype(dtype)
