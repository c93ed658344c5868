"""The CPU restatement of the seeded Omega stream (oracle/philox.py), pinned without a GPU:
Philox-4x32-10 against the known-answer vectors published with Random123 (Salmon et al., SC'11)
and against rocRAND's engine (tests/golden/philox_vectors.json); the Box-Muller transform by its
closed form at chosen words and by the moments of the stream; the layout by element order."""
import json
import os

import numpy as np

from conftest import ROOT
from oracle import philox as P

# Random123 kat_vectors, philox4x32 with 10 rounds: counter, key -> output
KAT = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
       ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
       ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
        (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]


def test_philox_known_answers():
    for ctr, key, out in KAT:
        got = P.philox4x32_10(np.array(ctr, dtype=np.uint32), np.array(key, dtype=np.uint32))
        assert tuple(int(x) for x in got) == out


def test_philox_matches_rocrand_engine():
    with open(os.path.join(ROOT, "tests", "golden", "philox_vectors.json")) as f:
        vec = json.load(f)["vectors"]
    ctr = np.array([v["counter"] for v in vec], dtype=np.uint32)
    key = np.array([v["key"] for v in vec], dtype=np.uint32)
    out = np.array([v["out"] for v in vec], dtype=np.uint32)
    assert np.array_equal(P.philox4x32_10(ctr, key), out)


def test_a_dropped_round_or_swapped_key_fails():
    """Plausible mistakes change every output word (the pins above would catch them)."""
    ctr, key, out = KAT[2]
    c, k = np.array(ctr, dtype=np.uint32), np.array(key, dtype=np.uint32)
    assert tuple(int(x) for x in P.philox4x32_10(c, k[::-1])) != out
    assert tuple(int(x) for x in P.philox4x32_10(c[[1, 0, 2, 3]], k)) != out


def test_box_muller_closed_form():
    # words giving u1 = 2^-1 (a >> 11 = 2^52 - 1) and u2 = 2^-2 (b >> 11 = 2^51 - 1)
    a = ((1 << 52) - 1) << 11
    b = ((1 << 51) - 1) << 11
    w = np.array([[a >> 32, a & 0xFFFFFFFF, b >> 32, b & 0xFFFFFFFF]], dtype=np.uint64)
    z = P.box_muller(w)[0]
    assert abs(z - np.sqrt(-2 * np.log(0.5)) * np.cos(np.pi / 2)) < 1e-15
    # u2 = 1: cos(2 pi) = 1, z = sqrt(2 ln 2)
    b = ((1 << 53) - 1) << 11
    w = np.array([[a >> 32, a & 0xFFFFFFFF, b >> 32, b & 0xFFFFFFFF]], dtype=np.uint64)
    assert abs(P.box_muller(w)[0] - np.sqrt(2 * np.log(2))) < 1e-15


def test_stream_moments_and_layout():
    z = P.gaussian_stream(7, 1 << 18)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01 and abs(np.mean(z ** 4) - 3) < 0.05
    om = P.omega(7, 64, 48)
    assert np.array_equal(om.ravel(), P.gaussian_stream(7, 64 * 48))  # row-major Omega
    assert np.array_equal(P.gaussian_stream(7, 10, start=100), P.gaussian_stream(7, 110)[100:])
    assert not np.array_equal(P.gaussian_stream(8, 10), P.gaussian_stream(7, 10))
