"""MatrixMarket array files (the format of the reference's data/*.mtx and of its demos' input:
ext/mmio/mmio.h:24-43; read by demo/cdemo.c:52-69): a banner line
``%%MatrixMarket matrix array real general``, comment lines starting with ``%``, a size line
``m n``, then the m*n values column by column.

read_array returns a column-major (Fortran-ordered) float64 array, as the C ABI takes A with
ordera 'C'; write_array writes one.  Host-side file I/O only: nothing here runs on the GPU."""
import numpy as np

BANNER = "%%MatrixMarket matrix array real general"


def read_array(path, with_comments=False):
    """(m, n) float64 array, Fortran order (and the comment lines when ``with_comments``)."""
    comments = []
    with open(path, "r") as f:
        banner = f.readline().strip().lower().split()
        if len(banner) < 5 or banner[:2] != ["%%matrixmarket", "matrix"] or banner[2] != "array":
            raise ValueError(f"{path}: not a MatrixMarket array file: {' '.join(banner)}")
        if banner[3] != "real" or banner[4] != "general":
            raise ValueError(f"{path}: unsupported field / symmetry {' '.join(banner[3:])}")
        line = f.readline()
        while line.startswith("%"):
            comments.append(line[1:].strip())
            line = f.readline()
        parts = line.split()
        if len(parts) != 2:
            raise ValueError(f"{path}: bad size line {line!r}")
        m, n = int(parts[0]), int(parts[1])
        vals = np.array(f.read().split(), dtype=np.float64)
    if vals.size != m * n:
        raise ValueError(f"{path}: expected {m * n} values, found {vals.size}")
    a = np.asfortranarray(vals.reshape(n, m).T)
    return (a, comments) if with_comments else a


def write_array(path, a, comments=()):
    """Write a 2-D array column by column (17 significant digits: read_array round-trips)."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim != 2:
        raise ValueError("write_array takes a 2-D array")
    with open(path, "w") as f:
        f.write(BANNER + "\n")
        for c in comments:
            f.write(f"% {c}\n")
        f.write(f"{a.shape[0]} {a.shape[1]}\n")
        for v in a.ravel(order="F"):
            f.write(f"{v:.17g}\n")
