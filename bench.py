"""Benchmark: seconds per rank-k iSVD on one MI355X.

Headline workload (FP64, synthetic): A = U diag(s) V^T + 1e-6 N(0,1), 131072 x 16384, stored
column-major (ordera 'C', as demo/cdemo.c passes it), s_i = 1/i for i < 2048, U / V seeded
orthonormal (CholeskyQR2 of Gaussian blocks on the GPU), Omega ~ N(0,1) 16384 x 4096;
k = 100, p = 28 (l = 128), N = 32; selectors GP / TS / KN / TS.  A and Omega are resident in
HBM before the timed region; one step is one full decomposition through the C ABI
(isvd_dIsvdOmega).  Reported beside it: the same decomposition on the row-major copy of A
(ordera 'R'), Wen-Yin integration, the mixed-precision (BF16) and FP8 families, the adjoint of
the differentiation workload, the per-stage roofline and the CPU oracle.

Prints ONE JSON line (rank 0).  --gpus N > 1 is rejected: the library runs on one GPU.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published peaks (dense matrix cores; HBM3E)
PEAK_FP64_MFMA = 78.6e12
PEAK_FP32_MFMA = 157.3e12
PEAK_BF16_MFMA = 2.5e15
PEAK_FP8_MFMA = 5.0e15
PEAK_HBM = 8.0e12


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def orth_factor(torch, g, rows, r):
    """Seeded orthonormal rows x r in FP64 on the GPU (CholeskyQR2: x := x R^-1, R^T R = x^T x)."""
    x = torch.randn(rows, r, generator=g, device="cuda", dtype=torch.float64)
    eye = torch.eye(r, device="cuda", dtype=torch.float64)
    for _ in range(2):
        c = torch.linalg.cholesky(x.T @ x)
        x = x @ torch.linalg.solve_triangular(c.T, eye, upper=True)
    return x


def make_problem(torch, m, n, r, seed, noise=1e-6, dtype=None, spectrum=None, ordera="C"):
    """A = U diag(s) V^T + noise, s_i = 1/i by default, generated on the GPU (torch; not part
    of the timed path).  Returns the array the C ABI takes for ``ordera``: the row-major n x m
    array A^T (column-major A) for 'C', the row-major m x n A for 'R'."""
    dtype = dtype or torch.float64
    g = torch.Generator(device="cuda").manual_seed(seed)
    u = orth_factor(torch, g, m, r)
    v = orth_factor(torch, g, n, r)
    if spectrum is None:
        s = 1.0 / torch.arange(1, r + 1, device="cuda", dtype=torch.float64)
    else:
        s = torch.as_tensor(spectrum, device="cuda", dtype=torch.float64)
    us = u * s[None, :]
    a = torch.empty((n, m) if ordera == "C" else (m, n), device="cuda", dtype=dtype)
    blk = 16384
    for r0 in range(0, m, blk):
        r1 = min(m, r0 + blk)
        t = us[r0:r1] @ v.T
        if noise:
            t.add_(torch.randn(r1 - r0, n, generator=g, device="cuda", dtype=torch.float64),
                   alpha=noise)
        if ordera == "C":
            a[:, r0:r1].copy_(t.T)
        else:
            a[r0:r1].copy_(t)
        del t
    del u, us
    return a, s, v


def a_matrix(a, ordera):
    """The m x n matrix A as a (possibly transposed) view of the array passed to the ABI."""
    return a.T if ordera == "C" else a


def host_info():
    """CPU model, logical cores, BLAS vendor / version / threads of the oracle's numpy."""
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    blas = {}
    try:
        from threadpoolctl import threadpool_info
        for i in threadpool_info():
            if i.get("user_api") == "blas":
                blas = {"blas": f"{i.get('internal_api')} {i.get('version')}",
                        "blas_threads": i.get("num_threads")}
                break
    except Exception:  # pragma: no cover
        pass
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), **blas}


def cpu_baseline(torch, a_dev, omega_dev, ordera, k, p, N, it_full, m_sample):
    """The CPU oracle (oracle/isvd_oracle.py: numpy + LAPACK, the reference's algorithm at
    P = 1).  Measured in full on the 2000 x 1000 CPU-runnable workload (k = 20, p = 12,
    N = 16, GP/GR/KN/GR), and on the first m_sample rows of the headline A with the same Omega
    (GP/TS/KN/TS): that sample's m-linear stages scaled by m / m_sample and its KN time per
    iteration by the headline's iteration count give the ESTIMATE reported as ``value``."""
    import numpy as np
    import oracle as O
    info = host_info()
    cores = info.get("blas_threads") or os.cpu_count()
    # the CPU-runnable workload in full (demo/cdemo.c:77-105 sizes)
    rng = np.random.default_rng(0)
    u0, _ = np.linalg.qr(rng.standard_normal((2000, 1000)))
    v0, _ = np.linalg.qr(rng.standard_normal((1000, 1000)))
    a0 = (u0 * 0.9 ** np.arange(1000)) @ v0.T
    om0 = np.random.default_rng(1).standard_normal((1000, 16 * 32))
    t0 = time.perf_counter()
    r0 = O.isvd(a0, om0, 20, 16, "GR", "KN", "GR")
    c0_sec = time.perf_counter() - t0
    # the headline workload on a row sample
    m = a_dev.shape[1] if ordera == "C" else a_dev.shape[0]
    rows = a_matrix(a_dev, ordera)[:m_sample]
    a = rows.cpu().numpy()
    om = omega_dev.cpu().numpy()
    t0 = time.perf_counter()
    y = O.sketch_gaussian_projection(a, om)
    t1 = time.perf_counter()
    qs = O.orthogonalize_tall_skinny_qr(y, N)
    t2 = time.perf_counter()
    qbar, it_s, _ = O.integrate_kolmogorov_nagumo(qs, N)
    t3 = time.perf_counter()
    O.postprocess_tall_skinny_qr(a, qbar, k)
    t4 = time.perf_counter()
    scale = m / m_sample
    est = (t1 - t0 + t2 - t1 + t4 - t3) * scale + (t3 - t2) / it_s * it_full * scale
    return {"value": est, "unit": "s", "cores": cores, "kind": "port",
            "sample": (f"ESTIMATE: oracle GP/TS/KN/TS on rows 0..{m_sample} of the same A and "
                       f"Omega ({t4 - t0:.1f} s of CPU work, {it_s} KN iterations); m-linear "
                       f"stages scaled x{scale:.0f}, KN time per iteration x{it_full} "
                       f"iterations.  Measured in full: the 2000x1000 k=20 p=12 N=16 "
                       f"GP/GR/KN/GR workload in {c0_sec:.3f} s ({r0['iter']} KN iterations)"),
            "stage_seconds_sample": [t1 - t0, t2 - t1, t3 - t2, t4 - t3],
            "cpu_runnable_workload_seconds": c0_sec,
            "cpu_runnable_workload_kn_iterations": r0["iter"], **info}


def cpu_runnable_workload_gpu(torch, capi, pipeline, reps=5):
    """The 2000 x 1000 CPU-runnable workload (k = 20, p = 12, N = 16, GP/GR/KN/GR; the inputs of
    the CPU baseline's measured run) through the C ABI on the GPU, inputs resident: best of reps."""
    import numpy as np
    rng = np.random.default_rng(0)
    u0, _ = np.linalg.qr(rng.standard_normal((2000, 1000)))
    v0, _ = np.linalg.qr(rng.standard_normal((1000, 1000)))
    a0 = torch.from_numpy(np.ascontiguousarray(((u0 * 0.9 ** np.arange(1000)) @ v0.T).T)).cuda()
    om0 = torch.from_numpy(np.random.default_rng(1).standard_normal((1000, 16 * 32))).cuda()
    prm = capi.create_param(2000, 1000, 20, 12, 16)
    best = float("inf")
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pipeline.isvd(prm, a0, omega=om0, algs=("GP", "GR", "KN", "GR"), ordera="C")
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def principal_angle_max(torch, a, b):
    """Largest principal angle (radians) between span(a) and span(b) (orthonormal columns)."""
    p = b - a @ (a.T @ b)
    return float(torch.arcsin(torch.clamp(torch.linalg.svdvals(p)[0], 0.0, 1.0)))


def stage_roofline(m, n, k, l, N, it, e_a, e_y, peaks, stage_sec, orth="TS", twice=False,
                   gram_factor=1.0):
    """Per-stage lower bound max(F / peak, B / HBM) from the algorithmic FLOP and bytes of this
    implementation, and its fraction of the measured stage time.  peaks: (sketch,
    orthogonalization, the integration's Gram, postprocess projection) FLOP/s.  Integration
    runs in sketch coordinates: the Nl x Nl Gram once (gram_factor m Nl^2 FLOP: 1 for the
    symmetric FP64 product, 4 for the stacked split-BF16 product), then per iteration M B
    (2 Nl^2 l), [D | Dg] and the update (12 Nl l^2) and ~30 l x l products (60 l^3) at the FP64
    peak, and Qbar = Qs X (2 m Nl l)."""
    Nl = N * l
    f_sk, b_sk = 2.0 * m * n * Nl, m * n * e_a + n * Nl * e_a + m * Nl * e_y
    if orth == "TS":
        f_or, b_or = N * (4.0 * m * l * l - 4.0 / 3.0 * l ** 3), 2.0 * m * Nl * e_y
    else:
        passes = 2 if twice else 1
        f_or, b_or = passes * N * 4.0 * m * l * l, passes * 2.0 * m * Nl * e_y
    f_gram = gram_factor * m * Nl * Nl + 2.0 * m * Nl * l
    f_iter = it * (2.0 * Nl * Nl * l + 12.0 * Nl * l * l + 60.0 * l ** 3)
    b_in = m * Nl * e_y * 2 + m * l * e_y
    f_po = 2.0 * m * n * l + 4.0 * n * l * l + 2.0 * (m + n) * l * k
    b_po = m * n * e_a + (m + n) * l * e_y
    out = []
    for name, f, b, pk, t in zip(("sketch", "orthogonalize", "integrate", "postprocess"),
                                 (f_sk, f_or, f_gram + f_iter, f_po), (b_sk, b_or, b_in, b_po),
                                 peaks, stage_sec):
        tf = f / pk if name != "integrate" else f_gram / pk + f_iter / PEAK_FP64_MFMA
        tb = max(tf, b / PEAK_HBM)
        out.append({"stage": name, "flop": f, "bytes": b,
                    "bound": "mfma" if tf >= b / PEAK_HBM else "hbm", "bound_seconds": tb,
                    "seconds": t, "frac": tb / t})
    return out


def traffic_of(kernel):
    """HBM bytes per launch of ``kernel`` from the committed counter summary, when it is the
    kernel that summary measured."""
    prof = os.path.join(ROOT, "profiles", "sketch_gemm_131072x16384_traffic.json")
    if not os.path.exists(prof):
        return None
    with open(prof) as f:
        d = json.load(f)
    for e in d.get("launches", []):
        if e.get("kernel") == kernel:
            return e.get("hbm_bytes_per_launch")
    return None


def timed(fn, steps, warmup, label):
    import numpy as np
    import torch
    res = None
    for i in range(warmup):
        res = fn()
    torch.cuda.synchronize()
    times, stages = [], []
    t_start = time.perf_counter()
    for i in range(steps):
        t0 = time.perf_counter()
        res = fn()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        stages.append(res["time"])
        log(f"{label} step {i + 1}/{steps}: {times[-1]:.4f} s, stages "
            f"{['%.4f' % x for x in res['time']]}, iter {res['iter']}")
    total = time.perf_counter() - t_start
    return res, total / steps, np.mean(np.array(stages), axis=0).tolist()


def mixed_case(torch, capi, pipeline, m, n, k, p, N, r, decay, steps, warmup):
    """BF16 A (spectrum 0.95^i or i^-0.5, + 1e-6 noise, column-major) through the mixed-precision
    driver (GP / GR twice in FP32 / KN / TS), timed, and compared with the FP64 driver (same
    selectors) on the identical BF16 values of A and Omega."""
    import numpy as np
    i = np.arange(r, dtype=np.float64)
    spectrum = 0.95 ** i if decay == "fast" else (i + 1.0) ** -0.5
    a64, s_true, _ = make_problem(torch, m, n, r, seed=4321, spectrum=spectrum, ordera="C")
    a_b = a64.to(torch.bfloat16)
    a64.copy_(a_b)  # FP64 copy of exactly the BF16 values
    gen = torch.Generator(device="cuda").manual_seed(77)
    om_b = torch.randn(n, (k + p) * N, generator=gen, device="cuda",
                       dtype=torch.float32).to(torch.bfloat16)
    torch.cuda.synchronize()
    prm = capi.create_param(m, n, k, p, N)
    algs = ("GP", "GR", "KN", "TS")  # the Gramian (applied twice in FP32) orthogonalizes
    rb, sec, st = timed(lambda: pipeline.isvd(prm, a_b, omega=om_b, algs=algs, ordera="C"),
                        steps, warmup, f"mixed ({decay})")
    om64 = om_b.double()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rd = pipeline.isvd(prm, a64, omega=om64, algs=algs, ordera="C")
    torch.cuda.synchronize()
    sec64 = time.perf_counter() - t0
    sb, sd = rb["s"][:k].double(), rd["s"][:k]
    flop = 2.0 * m * n * (k + p) * N
    out = {"decay": "s_i=0.95^i" if decay == "fast" else "s_i=i^-0.5", "seconds": sec,
           "stage_seconds": st, "kn_iterations": rb["iter"],
           "fp64_seconds": sec64, "fp64_stage_seconds": rd["time"], "speedup_vs_fp64": sec64 / sec,
           "fp64_kn_iterations": rd["iter"],
           "sigma_max_rel_err_vs_fp64": float(((sb - sd).abs() / sd).max()),
           "u_angle_max_vs_fp64": principal_angle_max(torch, rb["u"][:, :k].double(),
                                                      rd["u"][:, :k]),
           "v_angle_max_vs_fp64": principal_angle_max(torch, rb["v"][:, :k].double(),
                                                      rd["v"][:, :k]),
           "sigma_max_rel_err_vs_planted": float(((sb - s_true[:k]).abs() / s_true[:k]).max()),
           "roofline": {"bound": "mfma", "achieved": flop / st[0] / 1e12,
                        "peak": PEAK_BF16_MFMA / 1e12, "unit": "TFLOP/s",
                        "frac": flop / st[0] / PEAK_BF16_MFMA,
                        "traffic": traffic_of("isvd::bf16_gemm_kernel<false,false>"),
                        "kernel": "isvd::bf16_gemm_kernel<false,false> (BF16 sketch, A col-major)",
                        "flop_per_launch": flop, "avg_launch_seconds": st[0]},
           "stages_vs_roofline": stage_roofline(
               m, n, k, k + p, N, rb["iter"], 2, 4, st_peaks_mixed(), st, orth="GR",
               twice=True, gram_factor=4.0)}
    del a64, a_b, rb, rd
    torch.cuda.empty_cache()
    return out


def st_peaks_mixed():
    # sketch BF16; orthogonalization FP32; integration: split-BF16 Gram (the FLOP that
    # dominate), FP64 iterations; postprocess projection BF16
    return (PEAK_BF16_MFMA, PEAK_FP32_MFMA, PEAK_BF16_MFMA, PEAK_BF16_MFMA)


def fp8_case(torch, capi, pipeline, steps, warmup, m=262144, n=32768, k=64, p=64, N=48):
    """A = U diag(0.97^i) V^T + 1e-4 N(0,1) (rank-1024 planted part) held as OCP e4m3 with
    FP32 power-of-two scales per 128 x 128 block (isvd_eQuantize) plus a BF16 copy; the e4m3
    sketch on block-scaled FP8 matrix cores, then GR (twice, FP32) / KN / TS with the FP32
    Rayleigh-Ritz against the BF16 copy (isvd_eIsvd), timed.  Reference: the FP64 driver on
    the FP64 matrix with the identical Omega (the e4m3 values of the same seeded stream)."""
    import numpy as np
    r = 1024
    l, Nl = k + p, (k + p) * N
    algs = ("GP", "GR", "KN", "TS")
    seed = 2024
    log(f"fp8: generating A {m}x{n} (FP64, rank {r} + 1e-4 noise)")
    a64, s_true, _ = make_problem(torch, m, n, r, seed=555, noise=1e-4,
                                  spectrum=0.97 ** np.arange(r, dtype=np.float64), ordera="R")
    # Omega of the e family: e4m3 of the FP32 stream, read back through Y = I Omega
    eye = torch.eye(n, device="cuda", dtype=torch.float32)
    om32 = pipeline.sketch(capi.create_param(n, n, l, 0, N), eye, seed=seed)
    del eye
    om64 = om32.to(torch.float8_e4m3fn).double()
    del om32
    prm = capi.create_param(m, n, k, p, N)
    torch.cuda.synchronize()
    log("fp8: FP64 reference run")
    t0 = time.perf_counter()
    rd = pipeline.isvd(prm, a64, omega=om64, algs=algs)
    torch.cuda.synchronize()
    sec64 = time.perf_counter() - t0
    del om64
    # BF16 copy and e4m3 + block scales from FP32 row chunks (multiples of 128 rows)
    a_bf = torch.empty(m, n, device="cuda", dtype=torch.bfloat16)
    a8 = torch.empty(m, (n + 15) // 16 * 16, device="cuda", dtype=torch.uint8)
    scale = torch.empty((m + 127) // 128, (n + 127) // 128, device="cuda", dtype=torch.float32)
    chunk = 16384
    for r0 in range(0, m, chunk):
        r1 = min(m, r0 + chunk)
        a32 = a64[r0:r1].float()
        a_bf[r0:r1].copy_(a32)
        q8, qs = pipeline.quantize_e4m3(a32)
        a8[r0:r1].copy_(q8)
        scale[r0 // 128:(r1 + 127) // 128].copy_(qs)
        del a32, q8, qs
    del a64
    torch.cuda.empty_cache()
    torch.cuda.synchronize()
    re, sec, st = timed(lambda: pipeline.isvd_e(prm, a8, scale, a_bf, seed=seed, algs=algs),
                        steps, warmup, "fp8")
    se, sd = re["s"][:k].double(), rd["s"][:k]
    flop = 2.0 * m * n * Nl
    out = {"workload": f"A {m}x{n} as e4m3 + FP32 power-of-two scales per 128x128 block "
                       f"({a8.numel() / 1e9:.2f} GB + {scale.numel() * 4 / 2 ** 20:.1f} MiB) and "
                       f"a BF16 copy; U diag(0.97^i) V^T (rank {r}) + 1e-4 N(0,1); k={k} p={p} "
                       f"N={N}; GP/GR/KN/TS; errors vs the FP64 driver on the FP64 matrix with "
                       f"the same (e4m3-valued) Omega",
           "seconds": sec, "stage_seconds": st, "kn_iterations": re["iter"],
           "fp64_seconds": sec64, "fp64_stage_seconds": rd["time"],
           "fp64_kn_iterations": rd["iter"], "speedup_vs_fp64": sec64 / sec,
           "sigma_max_rel_err_vs_fp64": float(((se - sd).abs() / sd).max()),
           "u_angle_max_vs_fp64": principal_angle_max(torch, re["u"][:, :k].double(),
                                                      rd["u"][:, :k]),
           "v_angle_max_vs_fp64": principal_angle_max(torch, re["v"][:, :k].double(),
                                                      rd["v"][:, :k]),
           # the sketch stage is the Omega draw (e4m3) + scale check + the FP8 GEMM
           "roofline": {"bound": "mfma", "achieved": flop / st[0] / 1e12,
                        "peak": PEAK_FP8_MFMA / 1e12, "unit": "TFLOP/s",
                        "frac": flop / st[0] / PEAK_FP8_MFMA,
                        "traffic": traffic_of("isvd::f8_sketch_kernel<3>"),
                        "kernel": "isvd::f8_sketch_kernel<3> (sketch stage incl. Omega draw)",
                        "flop_per_launch": flop, "avg_launch_seconds": st[0]},
           "stages_vs_roofline": stage_roofline(
               m, n, k, l, N, re["iter"], 1, 4,
               (PEAK_FP8_MFMA, PEAK_FP32_MFMA, PEAK_BF16_MFMA, PEAK_BF16_MFMA), st, orth="GR",
               twice=True, gram_factor=4.0)}
    del a8, a_bf, scale, re, rd
    torch.cuda.empty_cache()
    return out


def adjoint_case(torch, capi, pipeline, steps, m=65536, n=8192, k=64, p=32, N=16):
    """The differentiation workload: A = U diag(2 * 0.9^i + 0.01) V^T, 65536 x 8192 FP32 (full
    rank, known factors), k = 64, l = 96, N = 16, loss sum w_i s_i + <G_U, U> + <G_V, V>.
    Times the forward (the four stages, Qbar kept) and the backward (isvd_sIsvdBackward: U, S,
    V recomputed from Qbar, complement series); checks the backward against the FP64 dense-SVD
    gradient computed from the known factors, with Qbar = the exact top-l subspace."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from adjoint_ref import dense_svd_grad  # the reference gradient of the adjoint's tests
    l = k + p
    g = torch.Generator(device="cuda").manual_seed(2025)
    log(f"adjoint: generating A {m}x{n} (full rank)")
    u = orth_factor(torch, g, m, n)
    v = orth_factor(torch, g, n, n)
    s = 2 * 0.9 ** torch.arange(n, device="cuda", dtype=torch.float64) + 0.01
    a32 = ((u * s[None, :]) @ v.T).float()
    w = torch.randn(k, generator=g, device="cuda", dtype=torch.float64)
    gU = torch.randn(m, k, generator=g, device="cuda", dtype=torch.float64)
    gV = torch.randn(n, k, generator=g, device="cuda", dtype=torch.float64)
    om = torch.randn(n, l * N, generator=g, device="cuda", dtype=torch.float32)
    prm = capi.create_param(m, n, k, p, N)
    fw, bw = [], []
    for i in range(steps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        y = pipeline.sketch(prm, a32, omega=om)
        pipeline.orthogonalize(prm, y, "TS")
        qbar, it, _ = pipeline.integrate(prm, y, "KN")
        del y
        pipeline.postprocess(prm, a32, qbar, "TS")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        pipeline.isvd_backward(prm, a32, qbar, w.float(), gU.float(), gV.float(), "TS")
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if i:  # the first round warms up
            fw.append(t1 - t0)
            bw.append(t2 - t1)
    # accuracy with the exact top-l subspace (signs of the forward's U applied to the factors)
    q_exact = u[:, :l].float().contiguous()
    da = pipeline.isvd_backward(prm, a32, q_exact, w.float(), gU.float(), gV.float(), "TS")
    _, u_fwd, _ = pipeline.postprocess(prm, a32, q_exact, "TS")
    sg = torch.sign((u_fwd[:, :k].double() * u[:, :k]).sum(0))
    u[:, :k] *= sg
    v[:, :k] *= sg
    ref = dense_svd_grad(torch, u, s, v, w, gU, gV, k)
    err = float(torch.linalg.norm(da.double() - ref) / torch.linalg.norm(ref))
    del u, v, a32, ref, da
    torch.cuda.empty_cache()
    return {"workload": f"A {m}x{n} FP32, s_i = 2*0.9^i + 0.01 (full rank), k={k} l={l} N={N}, "
                        f"GP/TS/KN/TS forward, loss sum w_i s_i + <G_U,U> + <G_V,V>",
            "forward_seconds": sum(fw) / len(fw), "backward_seconds": sum(bw) / len(bw),
            "kn_iterations": it,
            "backward_rel_err_vs_fp64_dense_svd_gradient_exact_subspace": err}


def sketch_only(torch, capi, pipeline, m, n, k, p, N, ordera, precision):
    """Counter runs: one sketch launch, then a plain device copy of 2^30 bytes whose
    FETCH_SIZE / WRITE_SIZE calibrate the counters."""
    Nl = (k + p) * N
    dt = torch.float64 if precision == "fp64" else torch.bfloat16
    shape = (n, m) if ordera == "C" else (m, n)
    a = torch.randn(shape, device="cuda", dtype=torch.float32).to(dt)
    omega = torch.randn(n, Nl, device="cuda", dtype=torch.float32).to(dt)
    torch.cuda.synchronize()
    prm = capi.create_param(m, n, k, p, N)
    t0 = time.perf_counter()
    pipeline.sketch(prm, a, omega=omega, ordera=ordera)
    t1 = time.perf_counter()
    src = torch.empty(2 ** 27, device="cuda", dtype=torch.float64).fill_(1.0)
    dst = src.clone()
    torch.cuda.synchronize()
    print(json.dumps({"sketch_seconds": t1 - t0, "ordera": ordera, "precision": precision,
                      "copy_bytes": 2 ** 30, "check": float(dst[0])}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--m", type=int, default=131072)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--k", type=int, default=100)
    ap.add_argument("--p", type=int, default=28)
    ap.add_argument("--N", type=int, default=32)
    ap.add_argument("--no-row-major", action="store_true")
    ap.add_argument("--no-wy", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp8", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--fp8-only", action="store_true", help="run only the FP8 (e4m3) case")
    ap.add_argument("--cpu-rows", type=int, default=8192)
    ap.add_argument("--sketch-only", choices=["fp64", "bf16"],
                    help="run one sketch GEMM and one 1 GiB device copy (for counter runs)")
    ap.add_argument("--ordera", choices=["C", "R"], default="C")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the iSVD library runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import numpy as np
    import torch

    from isvd_amd import capi, pipeline

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    m, n, k, p, N = args.m, args.n, args.k, args.p, args.N
    l, Nl = k + p, (k + p) * N
    if args.sketch_only:
        sketch_only(torch, capi, pipeline, m, n, k, p, N, args.ordera, args.sketch_only)
        return
    if args.fp8_only:
        print(json.dumps({"fp8": fp8_case(torch, capi, pipeline, args.steps, args.warmup)}))
        return
    r = min(2048, n)
    oa = args.ordera
    log(f"generating A {m}x{n} (FP64, ordera {oa}) and Omega {n}x{Nl} on the GPU")
    a, s_true, _ = make_problem(torch, m, n, r, seed=1234, ordera=oa)
    gen = torch.Generator(device="cuda").manual_seed(99)
    omega = torch.randn(n, Nl, generator=gen, device="cuda", dtype=torch.float64)
    torch.cuda.synchronize()
    prm = capi.create_param(m, n, k, p, N)
    algs = ("GP", "TS", "KN", "TS")

    def step(alg_i="KN", arr=None, order=oa):
        return pipeline.isvd(prm, a if arr is None else arr, omega=omega,
                             algs=("GP", "TS", alg_i, "TS"), ordera=order)

    res, sec, stage_avg = timed(step, args.steps, args.warmup, "fp64")

    # accuracy (size-independent properties): orthonormality, residual, planted spectrum
    A = a_matrix(a, oa)
    s_est = res["s"][:k].double()
    u = res["u"][:, :k]
    v = res["v"][:, :k]
    sigma_rel_err = float(((s_est - s_true[:k]).abs() / s_true[:k]).max())
    orth_u = float((u.T @ u - torch.eye(k, device="cuda", dtype=torch.float64)).abs().max())
    orth_v = float((v.T @ v - torch.eye(k, device="cuda", dtype=torch.float64)).abs().max())
    resid = float(torch.linalg.norm(A @ v - u * s_est[None, :]) / torch.linalg.norm(s_est))

    other = None
    if not args.no_row_major:
        o2 = "R" if oa == "C" else "C"
        log(f"the same decomposition with ordera {o2}")
        a2 = a.T.contiguous()
        r2, sec2, st2 = timed(lambda: step(arr=a2, order=o2), 2, 1, f"fp64 ordera {o2}")
        other = {"ordera": o2, "seconds": sec2, "stage_seconds": st2, "kn_iterations": r2["iter"],
                 "sigma_max_rel_diff": float(((r2["s"][:k] - res["s"][:k]).abs()
                                              / res["s"][:k]).max())}
        del a2, r2
        torch.cuda.empty_cache()

    wy = None
    if not args.no_wy:
        log("Wen-Yin integration run")
        t0 = time.perf_counter()
        rw = step("WY")
        torch.cuda.synchronize()
        wy = {"seconds": time.perf_counter() - t0, "stage_seconds": rw["time"],
              "iterations": rw["iter"],
              "sigma_max_rel_diff_vs_kn": float(((rw["s"][:k] - res["s"][:k]).abs()
                                                 / res["s"][:k]).max())}

    # dominant kernel: the sketch GEMM (2 m n Nl FLOP per launch); time[0] is the sketch stage
    # timed with HIP events on the library's stream, exactly one launch when Omega is supplied
    flop = 2.0 * m * n * Nl
    t_sketch = stage_avg[0]
    kname = "isvd::gemm_dl_kernel<false,false>" if oa == "C" else "isvd::gemm_dl_kernel<false,true>"
    roofline = {"bound": "mfma", "achieved": flop / t_sketch / 1e12,
                "peak": PEAK_FP64_MFMA / 1e12, "unit": "TFLOP/s",
                "frac": flop / t_sketch / PEAK_FP64_MFMA, "traffic": traffic_of(kname),
                "kernel": f"{kname} (sketch Y^T = Omega^T A^T, A {'col' if oa == 'C' else 'row'}-major)",
                "flop_per_launch": flop, "avg_launch_seconds": t_sketch,
                "algorithmic_bytes_per_launch": 8.0 * (m * n + n * Nl + m * Nl)}
    stages_rl = stage_roofline(m, n, k, l, N, res["iter"], 8, 8, (PEAK_FP64_MFMA,) * 4,
                               stage_avg, orth="TS")

    mixed = None
    if not args.no_mixed:
        del a, A
        torch.cuda.empty_cache()
        mixed = {"workload": f"A {m}x{n} stored in BF16, column-major (U diag(s) V^T + 1e-6 "
                             f"noise, rank {r}), k={k} p={p} N={N}, GP/GR/KN/TS; BF16 MFMA sketch "
                             f"and projection, Gramian twice in FP32, split-BF16 x3 Gram, FP64 "
                             f"l x l solves; errors vs the FP64 driver on the same BF16 values",
                 "cases": [mixed_case(torch, capi, pipeline, m, n, k, p, N, r, decay,
                                      args.steps, args.warmup) for decay in ("fast", "slow")]}
        mixed["speedup_vs_fp64_same_input"] = min(c["speedup_vs_fp64"] for c in mixed["cases"])
        a, _, _ = make_problem(torch, m, n, r, seed=1234, ordera=oa)

    fp8 = None
    if not args.no_fp8:
        del a
        torch.cuda.empty_cache()
        fp8 = fp8_case(torch, capi, pipeline, min(args.steps, 2), min(args.warmup, 1))
        a, _, _ = make_problem(torch, m, n, r, seed=1234, ordera=oa)

    adjoint = None
    if not args.no_adjoint:
        del a
        torch.cuda.empty_cache()
        adjoint = adjoint_case(torch, capi, pipeline, min(args.steps, 2))
        a, _, _ = make_problem(torch, m, n, r, seed=1234, ordera=oa)

    cpu = None
    if not args.no_cpu:
        log("CPU baseline (oracle: the CPU-runnable workload in full, the headline on a row sample)")
        cpu = cpu_baseline(torch, a, omega, oa, k, p, N, res["iter"], min(args.cpu_rows, m))
        cpu["cpu_runnable_workload_gpu_seconds"] = cpu_runnable_workload_gpu(torch, capi, pipeline)

    out = {
        "metric": "sec per rank-k iSVD & MFMA/HBM % roofline; mixed vs FP64 at stated σ error",
        "value": sec, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"iSVD FP64, A {m}x{n} {'column' if oa == 'C' else 'row'}-major "
                               f"(ordera '{oa}'), k={k} p={p} N={N}, GP/TS/KN/TS, s_i=1/i "
                               f"(i<{r}) + 1e-6 noise",
                   "m": m, "n": n, "k": k, "p": p, "N": N, "ordera": oa, "algs": list(algs)},
        "stage_seconds": stage_avg, "kn_iterations": res["iter"],
        "accuracy": {"sigma_max_rel_err_vs_planted": sigma_rel_err, "UtU_minus_I": orth_u,
                     "VtV_minus_I": orth_v, "residual_AV_minus_US_rel": resid},
        "other_order": other, "wen_yin": wy, "mixed": mixed, "fp8": fp8, "adjoint": adjoint,
        "roofline": roofline, "stages_vs_roofline": stages_rl, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
