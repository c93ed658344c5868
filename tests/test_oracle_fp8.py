"""Pins of the FP8 oracle (oracle/fp8.py): decode and encode against torch's float8_e4m3fn,
block scales by their defining property, the product against a triple loop."""
import numpy as np
import pytest

from oracle import fp8

torch = pytest.importorskip("torch")


def torch_codes(x):
    return torch.from_numpy(np.asarray(x, dtype=np.float32)).to(torch.float8_e4m3fn).view(
        torch.uint8).numpy()


def test_decode_table_matches_torch_float8_e4m3fn():
    ref = torch.arange(256, dtype=torch.uint8).view(torch.float8_e4m3fn).double().numpy()
    mine = fp8.decode(np.arange(256))
    assert np.array_equal(np.isnan(ref), np.isnan(mine))
    ok = ~np.isnan(ref)
    assert np.array_equal(ref[ok], mine[ok])
    assert fp8.decode(0x7E) == 448.0 and fp8.decode(0x01) == 2.0 ** -9


def test_encode_matches_torch_rounding():
    rng = np.random.default_rng(0)
    x = rng.standard_normal(200000) * np.exp2(rng.uniform(-12, 8.7, 200000))
    x = np.clip(x, -448, 448).astype(np.float32)
    # exact ties between neighbouring codes, subnormals, boundaries
    pos = fp8.E4M3_VALUES[:127]
    ties = ((pos[:-1] + pos[1:]) / 2).astype(np.float32)
    edge = np.array([0.0, -0.0, 2.0 ** -10, 2.0 ** -9, 3 * 2.0 ** -10, 7.5 * 2.0 ** -9, 2.0 ** -6,
                     440.0, 447.0, 448.0, -448.0, 1.0625, 1.1875], dtype=np.float32)
    for v in (x, ties, -ties, edge):
        assert np.array_equal(fp8.encode(v), torch_codes(v))


def test_block_scales_are_minimal_powers_of_two():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((300, 260)) * np.exp2(rng.integers(-30, 30, size=(300, 260)))
    a[128:256, 128:256] = 0.0
    s = fp8.block_scales(a)
    for bi in range(s.shape[0]):
        for bj in range(s.shape[1]):
            blk = a[bi * 128:(bi + 1) * 128, bj * 128:(bj + 1) * 128]
            amax = np.abs(blk).max()
            f, _ = np.frexp(s[bi, bj])
            assert f == 0.5  # a power of two
            if amax == 0:
                assert s[bi, bj] == 1.0
            else:
                assert amax <= 448 * s[bi, bj] and amax > 224 * s[bi, bj]


def test_quantize_round_trip_error_bound():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((200, 300)) * np.exp2(rng.integers(-6, 6, size=(200, 300)))
    codes, s = fp8.quantize(a)
    back = fp8.dequantize(codes, s)
    # relative half-ulp 2^-4 in the normal range, absolute 2^-10 s in the subnormal range
    sc = fp8.expand(s, *a.shape)
    assert np.all(np.abs(back - a) <= np.maximum(np.abs(a) * 2.0 ** -4, sc * 2.0 ** -10))


def test_sketch_matches_triple_loop():
    rng = np.random.default_rng(3)
    m, n, Nl = 130, 140, 5  # two scale blocks along each of m and n
    a8 = rng.integers(0, 256, size=(m, n)).astype(np.uint8)
    a8[(a8 & 0x7F) == 0x7F] = 0x3C
    om8 = rng.integers(0, 256, size=(n, Nl)).astype(np.uint8)
    om8[(om8 & 0x7F) == 0x7F] = 0x30
    scale = np.exp2(rng.integers(-5, 5, size=(2, 2))).astype(np.float64)
    y = fp8.sketch(a8, scale, om8)
    ref = fp8.sketch_loops(a8, scale, om8)
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-9 * np.abs(ref).max())
