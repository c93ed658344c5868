"""Writes tests/golden/philox_vectors.json: Philox-4x32-10 outputs of rocRAND's engine for a set
of counters and keys, computed on the host by the ten_rounds() of
/opt/rocm/include/rocrand/rocrand_philox4x32_10.h (compiled here with hipcc; no GPU needed).
They cross-check oracle/philox.py, the CPU restatement of the seeded Omega stream.

    python tests/golden/make_philox_vectors.py
"""
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = r"""
#include <hip/hip_runtime.h>
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdlib>
struct probe : rocrand_device::philox4x32_10_engine {
  uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main(int argc, char **argv) {
  probe p;
  for (int i = 1; i + 5 < argc + 1; i += 6) {
    uint4 c = {(unsigned)strtoul(argv[i], 0, 16), (unsigned)strtoul(argv[i + 1], 0, 16),
               (unsigned)strtoul(argv[i + 2], 0, 16), (unsigned)strtoul(argv[i + 3], 0, 16)};
    uint2 k = {(unsigned)strtoul(argv[i + 4], 0, 16), (unsigned)strtoul(argv[i + 5], 0, 16)};
    uint4 r = p.run(c, k);
    printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
  }
  return 0;
}
"""


def main():
    rng = np.random.default_rng(123)
    cases = [[0, 0, 0, 0, 0, 0], [0xFFFFFFFF] * 6,
             [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344, 0xA4093822, 0x299F31D0],
             [5, 0, 0, 0, 42, 0], [0x89ABCDEF, 7, 0, 0, 0x12345678, 0x9ABCDEF0]]
    cases += [list(map(int, rng.integers(0, 2 ** 32, size=6, dtype=np.uint64))) for _ in range(11)]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "p.cpp"), os.path.join(d, "p")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-O1", src, "-o", exe],
                       check=True)
        args = [f"{x:08x}" for c in cases for x in c]
        out = subprocess.run([exe, *args], capture_output=True, text=True, check=True).stdout
    rows = [[int(x, 16) for x in ln.split()] for ln in out.strip().splitlines()]
    vec = [{"counter": c[:4], "key": c[4:], "out": r} for c, r in zip(cases, rows)]
    with open(os.path.join(HERE, "philox_vectors.json"), "w") as f:
        json.dump({"source": "rocRAND philox4x32_10_engine::ten_rounds (host), "
                             "rocrand_philox4x32_10.h of ROCm 7.2", "vectors": vec}, f, indent=1)
    print(f"wrote {len(vec)} vectors")


if __name__ == "__main__":
    main()
