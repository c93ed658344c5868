"""Benchmark: seconds per rank-k iSVD on one MI355X.

Workload (FP64, synthetic): A = U diag(s) V^T + 1e-6 N(0,1), 131072 x 16384, s_i = 1/i for
i < 2048, U / V seeded orthonormal (CholeskyQR2 of Gaussian blocks on the GPU), Omega ~ N(0,1)
16384 x 4096; k = 100, p = 28 (l = 128), N = 32; selectors GP / TS / KN / TS.  A and Omega are
resident in HBM before the timed region; one step is one full decomposition through the C ABI
(isvd_dIsvdOmega).  Wen-Yin integration is timed separately on the same input.

Prints ONE JSON line (rank 0).  --gpus N > 1 is rejected: the library runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA = 78.6e12   # MI355X dense FP64 matrix peak (FLOP/s)
PEAK_HBM = 8.0e12          # MI355X HBM3E (B/s)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_problem(torch, m, n, r, seed, noise=1e-6, dtype=None, spectrum=None):
    """A = U diag(s) V^T + noise, s_i = 1/i by default, generated on the GPU (torch; not part
    of the timed path)."""
    dtype = dtype or torch.float64
    g = torch.Generator(device="cuda").manual_seed(seed)

    def orth(rows):
        x = torch.randn(rows, r, generator=g, device="cuda", dtype=torch.float64)
        for _ in range(2):  # CholeskyQR2: x := x R^-1 with R^T R = x^T x
            c = torch.linalg.cholesky(x.T @ x)
            eye = torch.eye(r, device="cuda", dtype=torch.float64)
            rinv = torch.linalg.solve_triangular(c.T, eye, upper=True)
            x = x @ rinv
        return x

    u = orth(m)
    v = orth(n)
    if spectrum is None:
        s = 1.0 / torch.arange(1, r + 1, device="cuda", dtype=torch.float64)
    else:
        s = torch.as_tensor(spectrum, device="cuda", dtype=torch.float64)
    us = u * s[None, :]
    a = torch.empty(m, n, device="cuda", dtype=dtype)
    blk = 16384
    for r0 in range(0, m, blk):
        r1 = min(m, r0 + blk)
        t = us[r0:r1] @ v.T
        if noise:
            t.add_(torch.randn(r1 - r0, n, generator=g, device="cuda", dtype=torch.float64),
                   alpha=noise)
        a[r0:r1].copy_(t)
        del t
    del u, us
    return a, s, v


def cpu_baseline(a_dev, omega_dev, k, p, N, it_full, m_sample):
    """The CPU oracle (oracle/isvd_oracle.py, numpy + LAPACK) on the first m_sample rows of the
    same A with the same Omega; each m-linear stage is scaled by m / m_sample and the KN
    iteration cost by the full-size iteration count."""
    import numpy as np
    import oracle as O
    try:
        from threadpoolctl import threadpool_info
        cores = max((i.get("num_threads", 1) for i in threadpool_info()), default=1)
    except Exception:  # pragma: no cover
        cores = os.cpu_count()
    m = a_dev.shape[0]
    a = a_dev[:m_sample].cpu().numpy()
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
            "sample": (f"oracle GP/TS/KN/TS on rows 0..{m_sample} of the same A and Omega "
                       f"({t4 - t0:.1f} s of CPU work, {it_s} KN iterations); m-linear stages "
                       f"scaled x{scale:.0f}, KN per-iteration cost x{it_full} iterations"),
            "stage_seconds_sample": [t1 - t0, t2 - t1, t3 - t2, t4 - t3]}


def principal_angle_max(torch, a, b):
    """Largest principal angle (radians) between span(a) and span(b) (orthonormal columns)."""
    p = b - a @ (a.T @ b)
    return float(torch.arcsin(torch.clamp(torch.linalg.svdvals(p)[0], 0.0, 1.0)))


def mixed_case(torch, capi, pipeline, m, n, k, p, N, r, decay, steps, warmup):
    """BF16 A (spectrum 0.95^i or i^-0.5, + 1e-6 noise) through the mixed-precision driver
    (GP / GR twice in FP32 / KN / TS), timed, and compared with the FP64 driver (same selectors)
    on the identical BF16 values of A and Omega."""
    import numpy as np
    i = np.arange(r, dtype=np.float64)
    spectrum = 0.95 ** i if decay == "fast" else (i + 1.0) ** -0.5
    a64, s_true, _ = make_problem(torch, m, n, r, seed=4321, spectrum=spectrum)
    a_b = a64.to(torch.bfloat16)
    a64.copy_(a_b)  # FP64 copy of exactly the BF16 values
    gen = torch.Generator(device="cuda").manual_seed(77)
    om_b = torch.randn(n, (k + p) * N, generator=gen, device="cuda",
                       dtype=torch.float32).to(torch.bfloat16)
    torch.cuda.synchronize()
    prm = capi.create_param(m, n, k, p, N)
    algs = ("GP", "GR", "KN", "TS")  # the Gramian (applied twice in FP32) orthogonalizes
    for _ in range(warmup):
        rb = pipeline.isvd(prm, a_b, omega=om_b, algs=algs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stages = []
    for _ in range(steps):
        rb = pipeline.isvd(prm, a_b, omega=om_b, algs=algs)
        stages.append(rb["time"])
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / steps
    log(f"mixed ({decay} decay): {sec:.4f} s, stages {np.mean(stages, axis=0).round(4).tolist()}, "
        f"KN iter {rb['iter']}")
    om64 = om_b.double()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rd = pipeline.isvd(prm, a64, omega=om64, algs=algs)
    torch.cuda.synchronize()
    sec64 = time.perf_counter() - t0
    sb, sd = rb["s"][:k].double(), rd["s"][:k]
    out = {"decay": "s_i=0.95^i" if decay == "fast" else "s_i=i^-0.5", "seconds": sec,
           "stage_seconds": np.mean(stages, axis=0).tolist(), "kn_iterations": rb["iter"],
           "fp64_seconds": sec64, "fp64_stage_seconds": rd["time"], "speedup_vs_fp64": sec64 / sec,
           "fp64_kn_iterations": rd["iter"],
           "sigma_max_rel_err_vs_fp64": float(((sb - sd).abs() / sd).max()),
           "u_angle_max_vs_fp64": principal_angle_max(torch, rb["u"][:, :k].double(),
                                                      rd["u"][:, :k]),
           "v_angle_max_vs_fp64": principal_angle_max(torch, rb["v"][:, :k].double(),
                                                      rd["v"][:, :k]),
           "sigma_max_rel_err_vs_planted": float(((sb - s_true[:k]).abs() / s_true[:k]).max())}
    del a64, a_b, rb, rd
    torch.cuda.empty_cache()
    return out


PEAK_FP8_MFMA = 5.0e15  # MI355X dense FP8 matrix peak (FLOP/s)


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
                                  spectrum=0.97 ** np.arange(r, dtype=np.float64))
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
    for _ in range(warmup):
        re = pipeline.isvd_e(prm, a8, scale, a_bf, seed=seed, algs=algs)
    torch.cuda.synchronize()
    times, stages = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        re = pipeline.isvd_e(prm, a8, scale, a_bf, seed=seed, algs=algs)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        stages.append(re["time"])
    sec = float(np.mean(times))
    st = np.mean(stages, axis=0).tolist()
    log(f"fp8: {sec:.4f} s, stages {np.round(st, 4).tolist()}, KN iter {re['iter']}; "
        f"FP64 {sec64:.3f} s")
    se, sd = re["s"][:k].double(), rd["s"][:k]
    flop = 2.0 * m * n * Nl
    out = {"workload": f"A {m}x{n} as e4m3 + FP32 power-of-two scales per 128x128 block "
                       f"({a8.numel() / 1e9:.2f} GB + {scale.numel() * 4 / 2 ** 20:.1f} MiB) and "
                       f"a BF16 copy; U diag(0.97^i) V^T (rank {r}) + 1e-4 N(0,1); k={k} p={p} "
                       f"N={N}; GP/GR/KN/TS; errors vs the FP64 driver on the FP64 matrix with "
                       f"the same (e4m3-valued) Omega",
           "seconds": sec, "stage_seconds": st, "kn_iterations": re["iter"],
           "sketch_tflops": flop / st[0] / 1e12,
           "sketch_frac_of_fp8_peak": flop / st[0] / PEAK_FP8_MFMA,
           "fp64_seconds": sec64, "fp64_stage_seconds": rd["time"],
           "fp64_kn_iterations": rd["iter"], "speedup_vs_fp64": sec64 / sec,
           "sigma_max_rel_err_vs_fp64": float(((se - sd).abs() / sd).max()),
           "u_angle_max_vs_fp64": principal_angle_max(torch, re["u"][:, :k].double(),
                                                      rd["u"][:, :k]),
           "v_angle_max_vs_fp64": principal_angle_max(torch, re["v"][:, :k].double(),
                                                      rd["v"][:, :k])}
    del a8, a_bf, scale, re, rd
    torch.cuda.empty_cache()
    return out


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
    ap.add_argument("--no-wy", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp8", action="store_true")
    ap.add_argument("--fp8-only", action="store_true", help="run only the FP8 (e4m3) case")
    ap.add_argument("--cpu-rows", type=int, default=8192)
    ap.add_argument("--sketch-only", action="store_true",
                    help="run one sketch GEMM and one 1 GiB device copy (for counter runs)")
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
        # counter runs: the sketch GEMM once (HBM traffic per launch), then a plain copy of
        # 2^30 bytes whose FETCH_SIZE / WRITE_SIZE calibrate the counters
        a = torch.randn(m, n, device="cuda", dtype=torch.float64)
        omega = torch.randn(n, Nl, device="cuda", dtype=torch.float64)
        torch.cuda.synchronize()
        prm = capi.create_param(m, n, k, p, N)
        t0 = time.perf_counter()
        pipeline.sketch(prm, a, omega=omega)
        t1 = time.perf_counter()
        src = torch.empty(2 ** 27, device="cuda", dtype=torch.float64).fill_(1.0)
        dst = src.clone()
        torch.cuda.synchronize()
        print(json.dumps({"sketch_seconds": t1 - t0, "copy_bytes": 2 ** 30,
                          "check": float(dst[0])}))
        return
    if args.fp8_only:
        print(json.dumps({"fp8": fp8_case(torch, capi, pipeline, args.steps, args.warmup)}))
        return
    r = min(2048, n)
    log(f"generating A {m}x{n} (FP64) and Omega {n}x{Nl} on the GPU")
    a, s_true, _ = make_problem(torch, m, n, r, seed=1234)
    gen = torch.Generator(device="cuda").manual_seed(99)
    omega = torch.randn(n, Nl, generator=gen, device="cuda", dtype=torch.float64)
    torch.cuda.synchronize()
    prm = capi.create_param(m, n, k, p, N)
    algs = ("GP", "TS", "KN", "TS")

    def step(alg_i="KN"):
        return pipeline.isvd(prm, a, omega=omega, algs=("GP", "TS", alg_i, "TS"))

    for i in range(args.warmup):
        log(f"warmup {i + 1}/{args.warmup}")
        res = step()
    torch.cuda.synchronize()
    times, stage = [], []
    t_start = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        res = step()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        stage.append(res["time"])
        log(f"step {i + 1}/{args.steps}: {times[-1]:.4f} s, stages {['%.4f' % x for x in res['time']]}, KN iter {res['iter']}")
    torch.cuda.synchronize()
    total = time.perf_counter() - t_start
    sec = total / args.steps
    stage_avg = np.mean(np.array(stage), axis=0).tolist()

    # accuracy (size-independent properties): orthonormality, residual, planted spectrum
    s_est = res["s"][:k].double()
    u = res["u"][:, :k]
    v = res["v"][:, :k]
    sigma_rel_err = float(((s_est - s_true[:k]).abs() / s_true[:k]).max())
    orth_u = float((u.T @ u - torch.eye(k, device="cuda", dtype=torch.float64)).abs().max())
    orth_v = float((v.T @ v - torch.eye(k, device="cuda", dtype=torch.float64)).abs().max())
    resid = float(torch.linalg.norm(a @ v - u * s_est[None, :]) / torch.linalg.norm(s_est))

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

    # dominant kernel: the sketch GEMM (2 m n Nl FLOP per launch), timed with HIP events on the
    # library's stream by the driver (time[0])
    flop = 2.0 * m * n * Nl
    t_sketch = stage_avg[0]
    achieved = flop / t_sketch
    roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP64_MFMA / 1e12,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_MFMA, "traffic": None,
                "kernel": "isvd::gemm_dl_kernel<false,true> "
                          "(sketch Y^T = Omega^T A^T, A row-major)",
                "flop_per_launch": flop, "avg_launch_seconds": t_sketch}
    prof = os.path.join(ROOT, "profiles", "sketch_gemm_fp64_131072x16384_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            roofline["traffic"] = json.load(f).get("hbm_bytes_per_launch")

    mixed = None
    if not args.no_mixed:
        del a
        torch.cuda.empty_cache()
        mixed = {"workload": f"A {m}x{n} stored in BF16 (U diag(s) V^T + 1e-6 noise, rank {r}), "
                             f"k={k} p={p} N={N}, GP/GR/KN/TS; BF16 MFMA sketch and projection, "
                             f"Gramian twice in FP32, split-BF16 x3 Gram, FP64 l x l solves; errors vs the FP64 driver "
                             f"on the same BF16 values",
                 "cases": [mixed_case(torch, capi, pipeline, m, n, k, p, N, r, decay,
                                      args.steps, args.warmup) for decay in ("fast", "slow")]}
        mixed["speedup_vs_fp64_same_input"] = min(c["speedup_vs_fp64"] for c in mixed["cases"])
        a, _, _ = make_problem(torch, m, n, r, seed=1234)

    fp8 = None
    if not args.no_fp8:
        del a
        torch.cuda.empty_cache()
        fp8 = fp8_case(torch, capi, pipeline, min(args.steps, 2), min(args.warmup, 1))
        a, _, _ = make_problem(torch, m, n, r, seed=1234)

    cpu = None
    if not args.no_cpu:
        log("CPU baseline (oracle on a row sample)")
        cpu = cpu_baseline(a, omega, k, p, N, res["iter"], min(args.cpu_rows, m))

    out = {
        "metric": "sec per rank-k iSVD & MFMA/HBM % roofline; mixed vs FP64 at stated σ error",
        "value": sec, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"iSVD FP64 A {m}x{n} col-major-equivalent, k={k} p={p} N={N}, "
                               f"GP/TS/KN/TS, s_i=1/i (i<{r}) + 1e-6 noise",
                   "m": m, "n": n, "k": k, "p": p, "N": N, "algs": list(algs)},
        "stage_seconds": stage_avg, "kn_iterations": res["iter"],
        "accuracy": {"sigma_max_rel_err_vs_planted": sigma_rel_err, "UtU_minus_I": orth_u,
                     "VtV_minus_I": orth_v, "residual_AV_minus_US_rel": resid},
        "wen_yin": wy, "mixed": mixed, "fp8": fp8, "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
