"""CPU oracle for the FP8 (OCP e4m3) sketch -- TEST INFRASTRUCTURE ONLY.

The reference (emfomy/isvd v1.3.2) has no reduced-precision path; this module states the
storage format and the product that the library's e family implements, so that its kernels
can be checked.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it, to check results against; the product path never calls it.

Format: OCP 8-bit floating point, E4M3 variant ("e4m3fn"): sign, 4 exponent bits (bias 7),
3 mantissa bits; exponent field 0 holds subnormals m * 2^-9; 0x7F / 0xFF are NaN; no
infinity; largest finite value 448 (0x7E).  A matrix is stored as codes plus one FP32 scale
per 128 x 128 block: A[i, j] = decode(a8[i, j]) * scale[i // 128, j // 128].

How each function is pinned (tests/test_oracle_fp8.py):
  E4M3_VALUES / decode -- all 256 codes against torch.float8_e4m3fn (an independent
                          implementation of the OCP format), NaNs included;
  encode               -- bit-exact against torch's FP32 -> float8_e4m3fn conversion (round
                          to nearest even) on random values over the whole finite range,
                          exact ties, subnormals and the largest finite value;
  block_scales         -- by property: a power of two, max|block| <= 448 s < 2 max|block|
                          (s is the smallest such power), 1 for a zero block;
  sketch               -- brute-force triple loop on tiny inputs.
"""
import numpy as np

BLOCK = 128


def _table():
    v = np.zeros(256)
    for c in range(256):
        s, e, m = (c >> 7) & 1, (c >> 3) & 15, c & 7
        if e == 15 and m == 7:
            v[c] = np.nan
        elif e == 0:
            v[c] = m * 2.0 ** -9
        else:
            v[c] = (1 + m / 8.0) * 2.0 ** (e - 7)
        if s:
            v[c] = -v[c]
    return v


E4M3_VALUES = _table()
_POS = E4M3_VALUES[:127]  # codes 0x00 .. 0x7E: the non-negative finite values, ascending


def decode(codes):
    """FP64 values of e4m3 codes (any integer array)."""
    return E4M3_VALUES[np.asarray(codes, dtype=np.int64) & 255]


def encode(x):
    """e4m3 codes of x (float array), round to nearest with ties to the even code; |x| > 448
    saturates to 448.  The nearest of the 127 non-negative finite values is looked up in the
    value table (no bit manipulation), so this is a separate statement from the kernels'."""
    x = np.asarray(x, dtype=np.float64)
    ax = np.minimum(np.abs(x), 448.0)
    hi = np.clip(np.searchsorted(_POS, ax, side="left"), 0, 126)
    lo = np.clip(hi - 1, 0, 126)
    dlo, dhi = ax - _POS[lo], _POS[hi] - ax
    pick_hi = (dhi < dlo) | ((dhi == dlo) & (hi % 2 == 0))
    code = np.where(pick_hi, hi, lo).astype(np.uint8)
    code = np.where(ax == _POS[hi], hi, code).astype(np.uint8)
    return np.where(np.signbit(x), code | 0x80, code).astype(np.uint8)


def block_scales(a):
    """Per 128 x 128 block of a (m x n): the smallest power of two s with max|A_b| <= 448 s
    (1 for an all-zero block)."""
    m, n = a.shape
    nbi, nbj = -(-m // BLOCK), -(-n // BLOCK)
    s = np.ones((nbi, nbj))
    for bi in range(nbi):
        for bj in range(nbj):
            amax = float(np.max(np.abs(a[bi * BLOCK:(bi + 1) * BLOCK, bj * BLOCK:(bj + 1) * BLOCK])))
            if amax > 0:
                f, q = np.frexp(amax)  # amax = f 2^q, f in [0.5, 1)
                e = q - 9 if f <= 0.875 else q - 8  # 448 2^(q-9) = 0.875 2^q
                s[bi, bj] = 2.0 ** max(-127, min(127, e))
    return s


def expand(scale, m, n):
    """The m x n matrix of per-element scales."""
    return np.repeat(np.repeat(scale, BLOCK, axis=0), BLOCK, axis=1)[:m, :n]


def quantize(a):
    """(codes m x n uint8, scales) of an FP64/FP32 matrix a (m x n)."""
    a = np.asarray(a, dtype=np.float64)
    s = block_scales(a)
    return encode(a / expand(s, *a.shape)), s


def dequantize(codes, scale):
    m, n = codes.shape
    return decode(codes) * expand(scale, m, n)


def sketch(a8, scale, omega8):
    """Y = A Omega in FP64 from the e4m3 A (m x n codes with block scales) and Omega (n x Nl
    codes, unscaled)."""
    return dequantize(a8, scale) @ decode(omega8)


def sketch_loops(a8, scale, omega8):
    """Triple-loop form of ``sketch`` for tiny inputs."""
    m, n = a8.shape
    Nl = omega8.shape[1]
    y = np.zeros((m, Nl))
    for i in range(m):
        for c in range(Nl):
            t = 0.0
            for j in range(n):
                t += E4M3_VALUES[a8[i, j]] * scale[i // BLOCK, j // BLOCK] * E4M3_VALUES[omega8[j, c]]
            y[i, c] = t
    return y
