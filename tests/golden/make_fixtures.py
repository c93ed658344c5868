"""Convert the reference's MatrixMarket golden vectors into one compact fixture file.

Reads ``data/*.mtx`` of the reference (emfomy/isvd v1.3.2, read from
/root/reference/data) and writes ``tests/golden/reference_data.npz``.
The ``.mtx`` files are MatrixMarket *array* files (real, general), values listed
column by column (format of ext/mmio/mmio.h:24-43; read by the reference tests,
e.g. check/check/libisvd/core/stage/@x@_integrate_kolmogorov_nagumo.cxx:36-55).

Provenance recorded per array (the ``% ...`` comment lines of each file), plus the
iteration counts the reference's own tests assert:
  q_kn: iter = 81 (data/q_kn.mtx:3-4; check/.../@x@_integrate_kolmogorov_nagumo.cxx:12-13)
  q_wy: iter = 21 (data/q_wy.mtx:3-4; check/.../@x@_integrate_wen_yin.cxx:12-13)

Run once in a container that has the reference mounted:
    python tests/golden/make_fixtures.py [/root/reference/data]
"""
import json
import os
import sys

import numpy as np

NAMES = ["a", "a_sym", "o_gp", "q", "q_hr", "q_kn", "q_wy", "qs", "s", "s_sym",
         "u", "u_sym", "v", "ys", "ys_gp"]


def read_mtx_array(path):
    """Parse a MatrixMarket 'array real general' file into an (m, n) float64 array."""
    comments = []
    with open(path, "r") as f:
        banner = f.readline().strip().lower().split()
        if banner[:2] != ["%%matrixmarket", "matrix"] or banner[2] != "array":
            raise ValueError(f"{path}: not a MatrixMarket array file: {banner}")
        if banner[3] != "real" or banner[4] != "general":
            raise ValueError(f"{path}: unsupported field/symmetry {banner[3:]}")
        line = f.readline()
        while line.startswith("%"):
            comments.append(line[1:].strip())
            line = f.readline()
        m, n = (int(t) for t in line.split())
        vals = np.array(f.read().split(), dtype=np.float64)
    if vals.size != m * n:
        raise ValueError(f"{path}: expected {m*n} values, found {vals.size}")
    return vals.reshape(n, m).T.copy(), comments


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/data"
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_data.npz")
    arrays, meta = {}, {}
    for name in NAMES:
        arr, comments = read_mtx_array(os.path.join(src, name + ".mtx"))
        arrays[name] = arr
        meta[name] = {"shape": list(arr.shape), "comments": comments}
    meta["_iterations"] = {"q_kn": {"tol": 1e-4, "maxit": 256, "iter": 81},
                           "q_wy": {"tol": 1e-3, "maxit": 256, "iter": 21}}
    meta["_source"] = "emfomy/isvd v1.3.2 data/*.mtx"
    arrays["_meta_json"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(out, **arrays)
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
