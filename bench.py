#!/usr/bin/env python3
"""Swift SpMM benchmark on one MI355X.

Workload: the 10^6 x 10^6, 16M-nonzero matrix (even rows banded with
bandwidth 64 and 8..32 nonzeros, odd rows Zipf(1.8) lengths up to 20k with
uniform columns, values N(0,1)), FP32, K = 128, row-major B (uniform in
[-1,1], original column order) and C.  One step = one C = A*B (the row-view
walk over the regular and reside entries, then the tail launch that adds the
carry slots of rows cut by segment boundaries) with A's Swift format and B
already resident in HBM.

Prints one JSON line.  `roofline` is for the dominant kernel (the row-view
walk): algorithmic bytes per launch (CSR-equivalent A read once, B read once,
C written once) over its average launch time, timed with HIP events on the
launch stream.  `cpu_baseline` times the reference's own Swift_cpu
(oracle/_ref, compiled from the reference sources) on the same matrix.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "swift-high-performance-sparse-dense-matrix-multiplication-on-gpus_amd")
sys.path[:0] = [PKG, os.path.join(ROOT, "oracle")]

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(m, n, nnz, K, s_a=4, s_b=4, s_c=4):
    """nnz*(4 + s_A) + 4*(M+1) + N*K*s_B + M*K*s_C."""
    return nnz * (4 + s_a) + 4 * (m + 1) + n * K * s_b + m * K * s_c


ROCPROF = "/opt/rocm/bin/rocprofv3"
KERNEL_PREFIX = {"f32": "k_rowview_row<float, float, float,",
                 "bf16": "k_rowview_row<__hip_bfloat16, __hip_bfloat16, float,"}


def under_profiler():
    return "rocprof" in os.environ.get("LD_PRELOAD", "") or any(
        k.startswith("ROCPROF") for k in os.environ)


def live_traffic(storage):
    """HBM bytes per launch of the dominant kernel, measured in this run: two
    rocprofv3 --pmc child runs of this script's short step (--calibrate),
    FETCH_SIZE then WRITE_SIZE in KiB per dispatch, each scaled by what the
    same counter reads for the 256 MiB device copy that --calibrate runs
    first.  None when the profiler is missing or a run fails."""
    import csv
    import shutil
    import subprocess
    import tempfile
    if not os.access(ROCPROF, os.X_OK):
        return None
    got = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        out = tempfile.mkdtemp(prefix="swift_pmc_", dir="/tmp")
        cmd = [ROCPROF, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "run", "--",
               sys.executable, os.path.abspath(__file__), "--steps", "3", "--warmup", "1",
               "--no-cpu", "--no-vendor", "--no-configs", "--batch", "0", "--calibrate",
               "--storage", storage]
        kern, copy = [], []
        try:
            # the profiler may end with a signal after the results are written:
            # the files decide, not the exit code
            subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"),
                           capture_output=True, timeout=90)
            rows = []
            for root, _, files in os.walk(out):
                for f in files:
                    if f.endswith("counter_collection.csv"):
                        with open(os.path.join(root, f)) as fh:
                            rows += [r for r in csv.DictReader(fh) if r["Counter_Name"] == counter]
            rows.sort(key=lambda r: int(r.get("Dispatch_Id") or 0))
            for r in rows:
                name = r["Kernel_Name"]
                if "copyBuffer" in name:
                    copy.append(float(r["Counter_Value"]))
                elif KERNEL_PREFIX[storage] in name:
                    kern.append(float(r["Counter_Value"]))
        except (OSError, subprocess.TimeoutExpired, KeyError, ValueError):
            kern, copy = [], []
        finally:
            shutil.rmtree(out, ignore_errors=True)
        if not kern or not copy or copy[0] <= 0:
            log("live %s of the %s walk: no result" % (counter, storage))
            return None
        # the calibration copy is the run's first device copy
        got[counter] = (sum(kern) / len(kern), (256 << 10) / copy[0], len(kern))
    (fk, ff, nd), (wk, wf, _) = got["FETCH_SIZE"], got["WRITE_SIZE"]
    traffic = int(round(fk * 1024 * ff + wk * 1024 * wf))
    return {"traffic": traffic, "fetch_kib": round(fk, 1), "write_kib": round(wk, 1),
            "fetch_factor": round(ff, 3), "write_factor": round(wf, 3), "dispatches": nd,
            "source": "rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, child runs of "
                      "bench.py --calibrate --storage %s in this run" % storage}


def measured_traffic(path):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (already scaled by the copy calibration):
    the fallback when the live measurement is not available."""
    try:
        for line in open(path):
            f = line.split()
            if len(f) == 2 and f[0] == "traffic_bytes":
                return int(f[1])
    except OSError:
        pass
    return None


def keep_rows(d, parity):
    """The same matrix with only the rows i % 2 == parity kept (the others
    empty): the banded half (0) or the Zipf half (1), for diagnostics."""
    rp = d["rowptr"].astype(np.int64)
    m = d["m"]
    lens = np.where(np.arange(m) % 2 == parity, np.diff(rp), 0)
    nrp = np.zeros(m + 1, np.int64)
    np.cumsum(lens, out=nrp[1:])
    row = np.repeat(np.arange(m), lens)
    src = rp[row] + (np.arange(int(nrp[-1])) - nrp[:-1][row])
    out = dict(d)
    out.update(rowptr=nrp.astype(np.int32), colidx=d["colidx"][src], val=d["val"][src],
               nnz=int(nrp[-1]))
    return out


def host_cpu():
    """CPU model and logical CPU count of the host the baseline ran on."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu": model, "nproc": os.cpu_count()}


def cpu_baseline(d, K, seed):
    """Reference Swift_cpu (src/Swift_cpu.h:31-65), 1 thread, on the same matrix."""
    import oracle as O
    m, n, nnz = d["m"], d["n"], d["nnz"]
    vals = d["val"].astype(np.float32)
    f = O.build(m, n, d["rowptr"], d["colidx"], vals.astype(np.float64))
    B = np.random.default_rng(seed).uniform(-1, 1, (n, K)).astype(np.float32)
    sB = np.ascontiguousarray(B[f["perm"]])
    del B
    if O.Ref.available(np.float32):
        ref = O.Ref(np.float32)
        h, _ = ref.format(m, n, f["sortrowidx_tmp"], f["sortval_tmp"].astype(np.float32),
                          f["sortnnz"])
        _, ms = ref.swift_cpu(h, m, n, nnz, sB)
        ref.free(h)
        kind = "reference"
    else:
        _, ms = f.swift_cpu_f32_timed(K, sB)
        kind = "port"
    gflops = 2.0 * nnz * K / (ms * 1e6)
    return {"value": round(gflops, 3), "unit": "GFLOP/s", "cores": 1, "kind": kind,
            "sample": "whole workload once: %dx%d, %d nnz, K=%d, FP32, Swift_cpu loops "
                      "(%.0f ms, gettimeofday around the loops as the reference does)"
                      % (m, n, nnz, K, ms)}


def vendor_baseline(scale, K, nnz):
    """rocSPARSE rocsparse_spmm (CSR; the best of its CSR algorithms) on the
    same generated matrix, through the C driver's vendor mode (a child
    process): row-major and column-major B/C, each checked against the host
    golden."""
    import subprocess
    exe = os.path.join(PKG, "swift_driver")
    cmd = [exe, "--gen", "2", "--scale", repr(scale), "--fp32", "--seed", "0", "--reps", "10",
           "--mode", "vendor", str(K)]
    try:
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=600).stdout
    except (OSError, subprocess.TimeoutExpired) as e:
        return {"value": None, "error": str(e)}
    for line in out.splitlines():
        f = line.split()
        if f[:2] == ["vendor", "rocSPARSE"]:
            row, col = float(f[f.index("row-major") + 1]), float(f[f.index("col-major") + 1])
            ok = [t for t in (row, col) if t > 0]
            best = min(ok) if ok else None
            return {"value": round(2.0 * nnz * K / (best * 1e6), 2) if best else None,
                    "unit": "GFLOP/s",
                    "name": "rocSPARSE rocsparse_spmm (CSR), best of its CSR algorithms",
                    "row_major_ms": row, "col_major_ms": col,
                    "row_major_alg": f[f.index("row-alg") + 1],
                    "col_major_alg": f[f.index("col-alg") + 1],
                    "sample": "same matrix and K, B in [-0.5, 0.5), HIP events around the "
                              "compute stage, median of 10, each algorithm verified"}
    return {"value": None, "error": out[-300:]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 100 steps of ~0.7 ms: long enough that the GPU's clock ramp after the
    # format build (the first ~20 products run 2-3% slower) weighs little
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--scale", type=float, default=1.0, help="shrink the matrix (tests only)")
    ap.add_argument("--K", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-vendor", action="store_true", help="skip the rocSPARSE baseline")
    ap.add_argument("--extras", action="store_true", help="also time the other forms")
    ap.add_argument("--batch", type=int, default=512,
                    help="matrices of the batched form in the configuration table (0: skip)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the configuration table (the other benchmark configurations)")
    ap.add_argument("--calibrate", action="store_true",
                    help="first run a plain 256 MiB device copy (counter calibration) and "
                         "time the step only (the counter runs of the traffic measurement)")
    ap.add_argument("--no-live-traffic", action="store_true",
                    help="take roofline.traffic from the committed rocprofv3 summary instead "
                         "of measuring it with rocprofv3 child runs")
    ap.add_argument("--kernel", default="DEFAULT",
                    help="row-major kernel family (DEFAULT, ROWVIEW, WINDOW, FLAT, BALANCED)")
    ap.add_argument("--part", choices=("all", "banded", "zipf"), default="all",
                    help="diagnostics: keep only the banded (even) or the Zipf (odd) rows")
    ap.add_argument("--storage", choices=("f32", "bf16"), default="f32",
                    help="diagnostics: time the BF16 storage form (A's values and B in BF16, "
                         "FP32 accumulation and C) as the step")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the Swift SpMM path runs on a single MI355X "
              "(no multi-GPU path is built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import swift_amd as S

    dev = torch.device("cuda:0")
    torch.cuda.init()
    if args.calibrate:
        # known traffic for hardware-counter scaling: 256 MiB read + 256 MiB written
        src = torch.ones(64 << 20, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    K, seed = args.K, 0
    t0 = time.time()
    d = S.generate(2, scale=args.scale, seed=seed)
    if args.part != "all":
        d = keep_rows(d, 0 if args.part == "banded" else 1)
    m, n, nnz = d["m"], d["n"], d["nnz"]
    log("matrix %dx%d nnz=%d generated in %.1fs" % (m, n, nnz, time.time() - t0))
    vals32 = d["val"].astype(np.float32)

    torch.cuda.synchronize()
    t0 = time.time()
    M = S.SwiftMatrix(m, n, d["rowptr"], d["colidx"], vals32, dtype=torch.float32)
    torch.cuda.synchronize()
    build_ms = (time.time() - t0) * 1e3
    inf = M.info()
    log("swift_build %.1f ms: tilenum=%d reside_nnz=%d numblk=%d nsteps=%d" %
        (build_ms, inf.tilenum, inf.reside_nnz, inf.numblk, inf.nsteps))

    g = torch.Generator(device=dev).manual_seed(seed)
    B = torch.rand((n, K), device=dev, generator=g, dtype=torch.float32) * 2 - 1
    C = torch.empty((m, K), device=dev, dtype=torch.float32)
    stream = torch.cuda.current_stream()

    kern = getattr(S, "KERNEL_" + args.kernel.upper())
    bf16 = args.storage == "bf16"
    if bf16:
        M.prepare_values(S.BF16)
        Bs = S.convert_dense(B, S.BF16)
        step = lambda: M.spmm(Bs, out=C, dtype=S.BF16, kernel=kern)
    else:
        Bs = B
        step = lambda: M.spmm(B, out=C, kernel=kern)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # whole-job throughput: K steps bracketed by synchronisation
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_time(e1) / args.steps
    flops = 2.0 * nnz * K
    value = flops / (ms_step * 1e-3) / 1e9
    log("spmm: %.3f ms/step (events), wall %.3f ms/step, %.1f GFLOP/s" %
        (ms_step, wall * 1e3 / args.steps, value))

    # per-kernel launch times (HIP events between the launches, same stream)
    phases = np.array([M.spmm_timed(Bs, C, kernel=kern)
                       for _ in range(max(5, min(args.steps, 20)))])
    zero_ms, reg_ms, res_ms = [float(x) for x in np.median(phases, axis=0)]
    log("phases (median ms): zero %.4f regular %.4f tail %.4f" % (zero_ms, reg_ms, res_ms))
    s_lp = 2 if bf16 else 4
    alg = algorithmic_bytes(m, n, nnz, K, s_a=s_lp, s_b=s_lp)
    achieved = alg / (reg_ms * 1e-3)
    # the committed counter summaries cover the full-size K=128 workload only
    default = args.scale == 1.0 and K == 128 and args.part == "all"
    traffic = measured_traffic(os.path.join(
        ROOT, "profiles", "rowview_%s_k128_1Mx1M_16Mnnz_traffic.txt" % args.storage)
    ) if default else None
    out = {
        "metric": "SpMM GFLOP/s + % HBM roofline (K=128 fp32); bf16 & fwd+bwd vs plain fp32",
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.storage,
        "data": "synthetic (seeded generator, seed 0)",
        "config": {"workload": "Swift SpMM C=A*B, A %dx%d with %d nnz (even rows banded bw 64 "
                               "8..32 nnz, odd rows Zipf(1.8) <=20k), %s, K=%d, row-major B "
                               "U[-1,1] in original column order" % (
                                   m, n, nnz, "BF16 storage of A's values and B, FP32 "
                                   "accumulation and C" if bf16 else "FP32", K),
                   "rows": m, "cols": n, "nnz": nnz, "K": K, "layout": "row-major",
                   "rows_kept": args.part, "kernel": args.kernel.upper(),
                   # wall time of the constructor: CSR upload from pageable host memory,
                   # first kernel loads and swift_build; build_gpu_ms is swift_build's own
                   # phase timing on the GPU
                   "build_ms": round(build_ms, 1),
                   "build_gpu_ms": dict(zip(("column_sort", "blocking", "schedules", "total"),
                                            (round(float(x), 2) for x in inf.build_ms))),
                   # swift_spmm_timed's phases: zeroing (none for the row-view
                   # walks), the walk, the tail (carry slots of cut rows, reside
                   # blocks where the row view does not hold them)
                   "phase_ms": {"zero_C": round(zero_ms, 4), "regular": round(reg_ms, 4),
                                "tail": round(res_ms, 4)}},
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
                     "traffic_frac": (round(traffic / (reg_ms * 1e-3) / HBM_PEAK, 4)
                                      if traffic else None),
                     "kernel": "k_rowview_row (the row-view walk)",
                     "algorithmic_bytes": alg},
        "gpu": torch.cuda.get_device_name(0),
    }
    if args.part == "all" and not bf16 and not args.calibrate:
        t0 = time.time()
        out["forms"] = headline_forms(S, torch, M, d, B, K, ms_step)
        log("forms %.1fs: %s" % (time.time() - t0, out["forms"]))
    if args.extras:
        out["extras"] = extras(S, torch, M, d, B, K)
    table = (args.part == "all" and not bf16 and not args.no_configs and K == 128 and
             not args.calibrate)
    if table:
        t0 = time.time()
        out["configs"] = config_table(S, torch, M, d, B, ms_step)
        log("configs on the 10^6-row matrix %.1fs" % (time.time() - t0))
    M.close()
    del B, C
    torch.cuda.empty_cache()
    if table:
        t0 = time.time()
        out["configs"].update(gnn_config(S, torch, args.scale))
        log("graph configuration %.1fs" % (time.time() - t0))
        torch.cuda.empty_cache()
    if args.batch:
        t0 = time.time()
        res = batched(S, torch, args.batch, args.scale)
        out.setdefault("configs", {})["C5 batched %d matrices" % args.batch] = res
        log("batched %.1fs" % (time.time() - t0))
        torch.cuda.empty_cache()
    if not args.no_vendor and args.part == "all":
        t0 = time.time()
        out["vendor_baseline"] = vendor_baseline(args.scale, K, nnz)
        log("vendor baseline %.1fs: %s" % (time.time() - t0, out["vendor_baseline"]))
    if not args.no_cpu:
        t0 = time.time()
        out["cpu_baseline"] = cpu_baseline(d, K, seed)
        out["cpu_baseline"].update(host_cpu())
        log("cpu baseline %.1fs: %s" % (time.time() - t0, out["cpu_baseline"]))
    # HBM traffic of the dominant kernel, measured now (after every timed part)
    if default and not args.calibrate and not args.no_live_traffic and not under_profiler():
        forms = [("f32", out["roofline"], reg_ms)] if not bf16 else [("bf16", out["roofline"], reg_ms)]
        if "forms" in out:
            forms.append(("bf16", out["forms"]["bf16"]["roofline"], out["forms"]["bf16"]["regular_ms"]))
        failed = False
        for storage, roof, ms in forms:
            t0 = time.time()
            # one failed or slow profiler run ends the live measurement
            lt = None if failed else live_traffic(storage)
            failed = lt is None
            log("live traffic %s %.1fs: %s" % (storage, time.time() - t0, lt))
            if lt:
                roof["traffic"] = lt.pop("traffic")
                roof["traffic_frac"] = round(roof["traffic"] / (ms * 1e-3) / HBM_PEAK, 4)
                roof["traffic_measurement"] = lt
            elif roof.get("traffic"):
                roof["traffic_measurement"] = {"source": "committed rocprofv3 summary under "
                                                         "profiles/ (live run unavailable)"}
    print(json.dumps(out), flush=True)


def time_call(torch, fn, reps=10, warm=5):
    # the untimed calls first let the clocks settle after the builds between legs
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def headline_forms(S, torch, M, d, B, K, f32_ms):
    """The metric's two ratios on the same matrix and B.  BF16 storage (A's
    values and B in BF16, FP32 accumulation and C) against the plain FP32 step,
    with its error against FP64 and its regular kernel against the HBM roofline;
    and the fused backward (dB = A^T dC and dvals = SDDMM(dC, B)) against the
    forward alone."""
    m, n, nnz = d["m"], d["n"], d["nnz"]
    C = torch.empty((m, K), device="cuda", dtype=torch.float32)
    M.prepare_values(S.BF16)
    Bh = S.convert_dense(B, S.BF16)
    bf = time_call(torch, lambda: M.spmm(Bh, out=C, dtype=S.BF16), reps=20)
    reg = float(np.median([M.spmm_timed(Bh, C)[1] for _ in range(10)]))
    err = lowp_error(torch, d, B, C)
    alg = algorithmic_bytes(m, n, nnz, K, s_a=2, s_b=2)
    traffic = measured_traffic(os.path.join(
        ROOT, "profiles", "rowview_bf16_k128_1Mx1M_16Mnnz_traffic.txt")) if K == 128 else None
    del Bh
    torch.cuda.empty_cache()
    dC = torch.randn((m, K), device="cuda")
    bw = time_call(torch, lambda: M.backward_fused(dC, B), reps=10)
    del dC, C
    torch.cuda.empty_cache()
    return {
        "bf16": {"ms_per_step": round(bf, 4), "speedup_vs_f32": round(f32_ms / bf, 3),
                 "gflops": round(2.0 * nnz * K / bf / 1e6, 1),
                 "regular_ms": round(reg, 4),
                 "roofline": {"achieved": round(alg / (reg * 1e-3) / 1e9, 1),
                              "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                              "frac": round(alg / (reg * 1e-3) / HBM_PEAK, 4),
                              "algorithmic_bytes": alg, "traffic": traffic,
                              "traffic_frac": (round(traffic / (reg * 1e-3) / HBM_PEAK, 4)
                                               if traffic else None)},
                 "error": err},
        "fwd_bwd": {"fwd_ms": round(f32_ms, 4), "bwd_fused_ms": round(bw, 4),
                    "fwd_bwd_over_fwd": round((f32_ms + bw) / f32_ms, 3)},
    }


def entry(ms, flops, alg):
    """One configuration: time per product, GFLOP/s, algorithmic bytes and
    their fraction of the HBM roofline."""
    return {"ms": round(ms, 4), "gflops": round(flops / ms / 1e6, 1), "algorithmic_bytes": alg,
            "roofline_frac": round(alg / (ms * 1e-3) / HBM_PEAK, 4)}


def config_table(S, torch, M, d, B, f32_ms):
    """The configurations other than the headline step on the 10^6-row
    matrix: FP32 column-major and K=32, FP64 K=128/32, BF16 storage at K=256
    and F16 storage with per-column scales at K=128/256 (each with its error
    against FP64), and K=256 after the autotuner.  Times are whole products
    (HIP events over 10-20 calls)."""
    m, n, nnz = d["m"], d["n"], d["nnz"]
    res = {}
    C = torch.empty((m, 128), device="cuda")
    Bt = B.t().contiguous()
    Ct = torch.empty((128, m), device="cuda")
    ms = time_call(torch, lambda: M.spmm(Bt, out=Ct, layout=S.COL_MAJOR), reps=10)
    res["C2 f32 K=128 col-major"] = entry(ms, 2.0 * nnz * 128, algorithmic_bytes(m, n, nnz, 128))
    del Bt, Ct
    B32 = B[:, :32].contiguous()
    C32 = torch.empty((m, 32), device="cuda")
    ms = time_call(torch, lambda: M.spmm(B32, out=C32), reps=20)
    res["C2 f32 K=32"] = entry(ms, 2.0 * nnz * 32, algorithmic_bytes(m, n, nnz, 32))
    del B32, C32
    torch.cuda.empty_cache()
    # FP64 build of the same matrix
    M64 = S.SwiftMatrix(m, n, d["rowptr"], d["colidx"], d["val"], dtype=torch.float64)
    for K in (128, 32):
        Bd = B[:, :K].double().contiguous()
        Cd = torch.empty((m, K), device="cuda", dtype=torch.float64)
        ms = time_call(torch, lambda: M64.spmm(Bd, out=Cd), reps=10)
        res["C2 f64 K=%d" % K] = entry(ms, 2.0 * nnz * K,
                                       algorithmic_bytes(m, n, nnz, K, s_a=8, s_b=8, s_c=8))
        del Bd, Cd
    M64.close()
    torch.cuda.empty_cache()
    # 16-bit storage at K = 256 (K = 128 BF16 is in "forms"), F16 with scales
    g = torch.Generator(device="cuda").manual_seed(1)
    B256 = torch.rand((n, 256), device="cuda", generator=g) * 2 - 1
    C256 = torch.empty((m, 256), device="cuda")
    M.prepare_values(S.BF16)
    Bh = S.convert_dense(B256, S.BF16)
    ms = time_call(torch, lambda: M.spmm(Bh, out=C256, dtype=S.BF16), reps=10)
    e = entry(ms, 2.0 * nnz * 256, algorithmic_bytes(m, n, nnz, 256, s_a=2, s_b=2))
    e["error"] = lowp_error(torch, d, B256, C256)
    ms_f32 = time_call(torch, lambda: M.spmm(B256, out=C256), reps=10)
    e["f32_ms"] = round(ms_f32, 4)
    e["speedup_vs_f32"] = round(ms_f32 / ms, 3)
    res["C3 bf16 K=256"] = e
    del Bh
    M.prepare_values(S.F16)
    for K, Bk, Ck in ((128, B, C), (256, B256, C256)):
        Bf, sc = S.convert_dense(Bk, S.F16, col_scale=True)
        ms = time_call(torch, lambda: M.spmm_scaled(Bf, sc, out=Ck), reps=10)
        e = entry(ms, 2.0 * nnz * K, algorithmic_bytes(m, n, nnz, K, s_a=2, s_b=2))
        e["error"] = lowp_error(torch, d, Bk, Ck)
        res["C3 f16 column-scaled K=%d" % K] = e
        del Bf, sc
    # the autotuner's schedules at K = 256 (they pin the matrix's segment
    # size, so they come last)
    M.prepare_values(S.BF16)
    Bh = S.convert_dense(B256, S.BF16)
    for name, Bk, code, sa, base in (("C3 bf16 K=256 autotuned", Bh, S.BF16, 2,
                                      res["C3 bf16 K=256"]["ms"]),
                                     ("C2 f32 K=256 autotuned", B256, S.F32, 4, ms_f32)):
        M.autotune(Bk, C256)
        inf = M.info()
        ms = time_call(torch, lambda: M.spmm(Bk, out=C256, dtype=code), reps=10)
        res[name] = dict(entry(ms, 2.0 * nnz * 256, algorithmic_bytes(m, n, nnz, 256, s_a=sa,
                                                                        s_b=sa)),
                         default_ms=round(base, 4),
                         schedule={"segment_items": inf.segment_items, "slots": inf.tuned_slots,
                                   "kernel": inf.tuned_kernel, "max_slice_cols": inf.tuned_cols})
    del Bh
    del B256, C256, C
    torch.cuda.empty_cache()
    log("configs: %s" % res)
    return res


def backward_bytes(m, n, nnz, K, s=4, s_out=None):
    """Fused backward over A^T: per entry its column, value, CSR position and
    dvals once (4 + s + 4 + s_out bytes), A^T's row pointer, dC gathered once
    (m*K), B and dB once each (n*K); s is the stored element size of dC, B and
    the values, s_out that of dB and dvals (= s unless BF16 storage)."""
    s_out = s if s_out is None else s_out
    return nnz * (8 + s + s_out) + 4 * (n + 1) + (m + n) * K * s + n * K * s_out


def gnn_config(S, torch, scale):
    """The differentiation configuration: the symmetric 200,000-node power-law
    graph with D^-1/2 (A+I) D^-1/2 values, K = 128, FP32 forward, BF16
    forward, fused backward (dB = A^T dC and dvals = SDDMM(dC, B)) and the
    forward + backward to forward ratio."""
    d = S.generate(4, scale=scale, seed=0)
    m, n, nnz = d["m"], d["n"], d["nnz"]
    K = 128
    M = S.SwiftMatrix(m, n, d["rowptr"], d["colidx"], d["val"].astype(np.float32))
    g = torch.Generator(device="cuda").manual_seed(2)
    B = torch.randn((n, K), device="cuda", generator=g)
    dC = torch.randn((m, K), device="cuda", generator=g)
    C = torch.empty((m, K), device="cuda")
    fwd = min(time_call(torch, lambda: M.spmm(B, out=C), reps=20) for _ in range(3))
    M.backward_fused(dC, B)  # builds the transposed layout once
    bwd = time_call(torch, lambda: M.backward_fused(dC, B), reps=20)
    M.prepare_values(S.BF16)
    Bh = S.convert_dense(B, S.BF16)
    fbf = time_call(torch, lambda: M.spmm(Bh, out=C, dtype=S.BF16), reps=20)
    dCh = S.convert_dense(dC, S.BF16)
    M.backward_fused(dCh, Bh)  # prepares the transposed layout's BF16 values once
    bbf = time_call(torch, lambda: M.backward_fused(dCh, Bh), reps=20)
    del dCh
    # last: the autotuner pins the matrix's segment size for every later product
    M.autotune(B, C)
    inf = M.info()
    ftu = min(time_call(torch, lambda: M.spmm(B, out=C), reps=20) for _ in range(3))
    res = {
        "C4 graph f32 forward": dict(entry(fwd, 2.0 * nnz * K, algorithmic_bytes(m, n, nnz, K)),
                                     nodes=m, nnz=nnz),
        "C4 graph bf16 forward": entry(fbf, 2.0 * nnz * K,
                                       algorithmic_bytes(m, n, nnz, K, s_a=2, s_b=2)),
        "C4 graph fused backward": entry(bwd, 4.0 * nnz * K, backward_bytes(m, n, nnz, K)),
        "C4 graph fwd+bwd / fwd": round((fwd + bwd) / fwd, 3),
        "C4 graph bf16 fused backward": entry(bbf, 4.0 * nnz * K,
                                              backward_bytes(m, n, nnz, K, s=2, s_out=4)),
        "C4 graph bf16 fwd+bwd / fwd": round((fbf + bbf) / fbf, 3),
        "C4 graph bf16 fwd+bwd speedup vs f32": round((fwd + bwd) / (fbf + bbf), 3),
        "C4 graph f32 forward autotuned": dict(
            entry(ftu, 2.0 * nnz * K, algorithmic_bytes(m, n, nnz, K)), default_ms=round(fwd, 4),
            schedule={"segment_items": inf.segment_items, "slots": inf.tuned_slots,
                      "kernel": inf.tuned_kernel, "max_slice_cols": inf.tuned_cols}),
        "C4 graph fwd+bwd / fwd, autotuned forward": round((ftu + bwd) / ftu, 3),
    }
    M.close()
    del B, dC, C, Bh
    log("graph: %s" % res)
    return res


def extras(S, torch, M, d, B, K):
    """Other forms on the same matrix: the other row kernels, B in sorted row
    order, column-major B, host buffers over PCIe, the other kernels' BF16
    walks, and F16 storage with per-column scales."""
    m, n, nnz = d["m"], d["n"], d["nnz"]
    flops = 2.0 * nnz * K
    res = {}
    C = torch.empty((m, K), device="cuda", dtype=torch.float32)
    base = time_call(torch, lambda: M.spmm(B, out=C))
    res["f32_row_ms"] = round(base, 4)
    for name in ("ROWVIEW", "WINDOW"):
        k = getattr(S, "KERNEL_" + name)
        res["f32_row_%s_ms" % name.lower()] = round(
            time_call(torch, lambda: M.spmm(B, out=C, kernel=k)), 4)
    res["f32_row_tiles_scatter_ms"] = round(time_call(
        torch, lambda: M.spmm(B, out=C, kernel=S.KERNEL_BALANCED), reps=3, warm=1), 4)
    perm = torch.tensor(M.arrays(["perm"])["perm"].astype(np.int64), device="cuda")
    Bs = B[perm].contiguous()
    res["f32_row_Bsorted_ms"] = round(time_call(torch, lambda: M.spmm(Bs, out=C,
                                                                      flags=S.B_SORTED)), 4)
    Bt = B.t().contiguous()
    Ct = torch.empty((K, m), device="cuda", dtype=torch.float32)
    res["f32_col_ms"] = round(time_call(torch, lambda: M.spmm(Bt, out=Ct, layout=S.COL_MAJOR),
                                        reps=3, warm=1), 4)
    # host buffers handed over (pinned): B in over PCIe, SpMM, C back out
    Bhost = B.cpu().pin_memory()
    Chost = torch.empty((m, K), dtype=torch.float32).pin_memory()
    Bdev = torch.empty_like(B)

    def pcie():
        Bdev.copy_(Bhost, non_blocking=True)
        M.spmm(Bdev, out=C)
        Chost.copy_(C, non_blocking=True)

    res["f32_row_pcie_inclusive_ms"] = round(time_call(torch, pcie, reps=5, warm=1), 4)
    del Bhost, Chost, Bdev
    M.prepare_values(S.BF16)
    Bh = S.convert_dense(B, S.BF16)
    # (the default BF16 form and the backward are timed in the line's "forms")
    for name in ("ROWVIEW", "WINDOW"):
        k = getattr(S, "KERNEL_" + name)
        res["bf16_row_%s_ms" % name.lower()] = round(
            time_call(torch, lambda: M.spmm(Bh, out=C, dtype=S.BF16, kernel=k)), 4)
    M.prepare_values(S.F16)
    Bf, sc = S.convert_dense(B, S.F16, col_scale=True)
    f16 = time_call(torch, lambda: M.spmm_scaled(Bf, sc, out=C))
    res["f16_scaled_row_ms"] = round(f16, 4)
    res["f16_scaled_error"] = lowp_error(torch, d, B, C)
    del Bf, sc
    res["f32_row_gflops"] = round(flops / base / 1e6, 1)
    return res


def lowp_error(torch, d, B, C):
    """Error of a 16-bit-storage product C (FP32) against an FP64 product of the
    same matrix and B computed by torch's sparse CSR matmul on the GPU: the
    largest per-row max relative error (against the row's max |C64|) and the
    largest error normalised by (|A||B|)[i,k]."""
    try:
        return _lowp_error(torch, d, B, C)
    except (RuntimeError, ValueError) as e:  # the FP64 reference could not run
        return {"error": str(e)[:200]}


def _lowp_error(torch, d, B, C):
    dev = B.device
    rp = torch.tensor(d["rowptr"].astype(np.int64), device=dev)
    ci = torch.tensor(d["colidx"].astype(np.int64), device=dev)
    # the FP32 values the matrix was built from
    v = torch.tensor(d["val"].astype(np.float32).astype(np.float64), device=dev)
    m, n = d["m"], d["n"]
    A = torch.sparse_csr_tensor(rp, ci, v, (m, n))
    Aabs = torch.sparse_csr_tensor(rp, ci, v.abs(), (m, n))
    B64 = B.double()
    C64 = A @ B64
    bound = Aabs @ B64.abs()
    err = (C.double() - C64).abs()
    row = err.amax(1) / C64.abs().amax(1).clamp_min(1e-30)
    out = {"row_max_rel_max": float(row.max()),
           "row_max_rel_p99": float(torch.quantile(row.float(), 0.99)),
           "normalised_max": float((err / (bound + 1e-30)).max()),
           "reference": "torch sparse CSR matmul in FP64 on the GPU"}
    del A, Aabs, C64, bound, err
    return out


def batched(S, torch, count, scale=1.0, reps=10):
    """Batched form: `count` generated matrices (square, rows log-U[1k,64k],
    mean nonzeros per row log-U[4,128], Zipf(2.0) rows, K in {32,64,128},
    FP32 N(0,1)), all products in one persistent launch (plain and replayed
    from a HIP graph) against one swift_spmm call per matrix."""
    import ctypes as C
    L = S.lib()
    t0 = time.time()
    mats, Ks, Bs, Cs = [], [], [], []
    flops = 0.0
    nnz_total = 0
    alg = 0
    g = torch.Generator(device="cuda").manual_seed(0)
    import concurrent.futures as cf
    with cf.ThreadPoolExecutor(8) as ex:  # the generator releases the GIL
        gen = list(ex.map(lambda i: S.generate(5, scale=scale, seed=0, index=i), range(count)))
    for d in gen:
        M = S.SwiftMatrix(d["m"], d["n"], d["rowptr"], d["colidx"], d["val"].astype(np.float32),
                          dtype=torch.float32)
        K = d["k"]
        mats.append(M)
        Ks.append(K)
        Bs.append(torch.rand((d["n"], K), device="cuda", generator=g) * 2 - 1)
        Cs.append(torch.empty((d["m"], K), device="cuda"))
        flops += 2.0 * d["nnz"] * K
        nnz_total += d["nnz"]
        alg += algorithmic_bytes(d["m"], d["n"], d["nnz"], K)
    torch.cuda.synchronize()
    log("batched: %d matrices, %d nnz, built in %.1fs" % (count, nnz_total, time.time() - t0))
    stream = S._stream()
    handles = [m.h for m in mats]
    bptr = [S._ptr(b) for b in Bs]
    cptr = [S._ptr(c) for c in Cs]

    def separate():
        for h, K, b, c in zip(handles, Ks, bptr, cptr):
            L.swift_spmm(h, K, S.ROW_MAJOR, S.F32, S.KERNEL_DEFAULT, 0, b, c, stream)

    res = {"matrices": count, "nnz": nnz_total, "gflop": round(flops / 1e9, 3)}
    log("batched: algorithmic bytes %d" % alg)
    res["separate_default_ms"] = round(time_call(torch, separate, reps=reps), 4)
    t0 = time.time()
    batch = S.Batch(mats, Ks)
    torch.cuda.synchronize()
    res["batch_create_s"] = round(time.time() - t0, 2)
    res["params"] = batch.params()
    res["separate_tuned_ms"] = round(time_call(torch, separate, reps=reps), 4)
    res["batch_ms"] = round(time_call(torch, lambda: batch.spmm(Bs, Cs), reps=reps), 4)
    res["graph_ms"] = round(time_call(torch, lambda: batch.spmm(Bs, Cs, graph=True), reps=reps), 4)
    for k in ("separate_default", "separate_tuned", "batch", "graph"):
        res[k + "_gflops"] = round(flops / res[k + "_ms"] / 1e6, 1)
    res["graph_speedup_vs_separate"] = round(res["separate_default_ms"] / res["graph_ms"], 3)
    # every matrix's A, B and C once, over the replay's whole time
    res["algorithmic_bytes"] = alg
    res["graph_roofline_frac"] = round(alg / (res["graph_ms"] * 1e-3) / HBM_PEAK, 4)
    log("batched: %s" % res)
    batch.close()
    for M in mats:
        M.close()
    return res


if __name__ == "__main__":
    main()
