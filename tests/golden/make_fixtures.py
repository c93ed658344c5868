"""Convert the reference's MatrixMarket golden vectors into one compact fixture file.

Reads ``data/*.mtx`` of the reference (emfomy/isvd v1.3.2, read from
/root/reference/data) and writes ``tests/golden/reference_data.npz``.
The ``.mtx`` files are MatrixMarket *array* files (real, general), values listed
column by column, read by isvd_amd.mmio (format of ext/mmio/mmio.h:24-43; read by the reference tests,
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

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from isvd_amd.mmio import read_array  # noqa: E402

NAMES = ["a", "a_sym", "o_gp", "q", "q_hr", "q_kn", "q_wy", "qs", "s", "s_sym",
         "u", "u_sym", "v", "ys", "ys_gp"]


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/data"
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_data.npz")
    arrays, meta = {}, {}
    for name in NAMES:
        arr, comments = read_array(os.path.join(src, name + ".mtx"), with_comments=True)
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
