"""MatrixMarket array reader / writer (isvd_amd.mmio): parses the reference's format (banner,
% comments, size line, values column by column), round-trips, rejects what it does not read,
and reproduces the reference's data/a.mtx from its fixture."""
import numpy as np
import pytest

from isvd_amd import mmio

SAMPLE = """%%MatrixMarket matrix array real general
% a 3 x 2 matrix, values column by column
% second comment
3 2
1
2
3.5
-4e-3
5
6
"""


def test_reads_the_array_format(tmp_path):
    p = tmp_path / "s.mtx"
    p.write_text(SAMPLE)
    a, comments = mmio.read_array(str(p), with_comments=True)
    assert a.shape == (3, 2) and a.flags.f_contiguous
    assert np.array_equal(a, np.array([[1, -4e-3], [2, 5], [3.5, 6]]))
    assert comments == ["a 3 x 2 matrix, values column by column", "second comment"]


def test_round_trip_is_exact(tmp_path, golden):
    a = golden["a"]
    p = tmp_path / "a.mtx"
    mmio.write_array(str(p), a, comments=["the reference's data/a.mtx"])
    b = mmio.read_array(str(p))
    assert np.array_equal(a, b) and b.shape == (101, 1001)


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 1.0\n",
    "%%MatrixMarket matrix array complex general\n1 1\n1 0\n",
    "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n"])
def test_rejects_other_forms(tmp_path, text):
    p = tmp_path / "bad.mtx"
    p.write_text(text)
    with pytest.raises(ValueError):
        mmio.read_array(str(p))
