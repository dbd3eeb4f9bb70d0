#!/usr/bin/env python3
"""Benchmark of the dense 2-D filtering path on one MI355X.

Workload (a "step"): imfilter(Float32, img, centered(K), "reflect",
Algorithm.FIR()) with img a 4096x4096 Float32 image of seeded uniform [0,1)
values and K a 15x15 kernel of seeded N(0,1) taps normalised to sum 1 -- the
MFMA banded-Toeplitz kernel with FP16 hi/lo operand splitting.  Input and
output stay resident in
HBM; K steps are timed with HIP events on the library's stream.

Prints one JSON line: Gpixel/s, the roofline of the dominant kernel, the
FP64 CPU baseline (the oracle's single-threaded direct FIR on a row sample),
the max relative error of the timed output against FP64 on that sample, and
extra measurements of the other configurations.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "imagefiltering.jl_amd"))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def norm_kernel(rng, K):
    while True:
        k = rng.standard_normal((K, K))
        s = k.sum()
        if abs(s) >= 1:
            return k / s


class Timer:
    def __init__(self, L, ctx):
        self.L, self.ctx = L, ctx

    def run(self, fn, steps, warmup):
        for _ in range(warmup):
            fn()
        self.L.ifm_ctx_synchronize(self.ctx)
        self.L.ifm_timer_start(self.ctx)
        for _ in range(steps):
            fn()
        ms = C.c_float()
        self.L.ifm_timer_stop(self.ctx, C.byref(ms))
        self.L.ifm_ctx_synchronize(self.ctx)
        return ms.value / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="single-thread CPU time of the headline's oracle sample")
    ap.add_argument("--cpu-extra-seconds", type=float, default=1.5,
                    help="single-thread CPU time of each other configuration's oracle sample")
    ap.add_argument("--no-extras", action="store_true", help="skip the other configurations (serving runs unless --serve-requests 0)")
    ap.add_argument("--serve-requests", type=int, default=20000)
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the tile-shape autotuner (use the cached or default tile rules; profiling runs)")
    ap.add_argument("--profile-copy", action="store_true",
                    help="also run one plain copy of a known size (rocprof counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the filtering path runs on a single MI355X "
              "(multi-GPU sharding is a design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import imagefiltering as IF
    from imagefiltering import _lib
    from imagefiltering.device import ifm_array
    from imagefiltering.imfilter import _border_struct, _kernel_struct

    L = _lib.load()
    ctx = _lib.ctx()
    T = Timer(L, ctx)

    n = 4096
    rng = np.random.default_rng(20261204)
    img = rng.random((n, n), dtype=np.float32)
    k15 = norm_kernel(np.random.default_rng(20261214), 15)
    d_img = IF.DeviceArray.from_host(img)
    d_out = d_img.similar()
    a_in, a_out = ifm_array(d_img), ifm_array(d_out)
    b_reflect = _border_struct(IF.Pad("reflect"), IF.Float32)
    kern15 = IF.centered(k15)
    ks15 = _kernel_struct(kern15)
    keep15 = ks15

    def step():
        rc = L.ifm_imfilter_dense(ctx, a_in, a_out, None, keep15, b_reflect, _lib.PREC_AUTO)
        if rc:
            _lib.check(rc)

    log("bench: dense 15x15 reflect on 4096x4096 f32 (MFMA FP16 hi/lo split), warmup %d steps %d" % (args.warmup,
                                                                                                 args.steps))
    ms = T.run(step, args.steps, args.warmup)
    px = n * n
    gpix = px / (ms * 1e-3) / 1e9
    bytes_alg = px * (4 + 4)
    flops_alg = px * 2 * 15 * 15
    achieved_gbs = bytes_alg / (ms * 1e-3) / 1e9
    log("bench: %.4f ms/step, %.2f Gpix/s, %.0f GB/s" % (ms, gpix, achieved_gbs))

    # accuracy of the timed output against the FP64 oracle on a row sample; the
    # same oracle run is the CPU baseline (one thread, as CPU1(Algorithm.FIR()))
    out = d_out.to_host()
    from oracle import oracle as O
    cpu = CpuBaseline(O, args.cpu_extra_seconds)
    img64 = np.asfortranarray(img, dtype=np.float64)   # the oracle's own layout, outside the timing
    nrows0 = 2
    r0 = n // 2
    _, t_small = cpu.run(lambda: O.imfilter_dense_rows(img64, k15, [-7, -7], "reflect", r0, nrows0), 1)
    nrows = int(max(4, min(n, nrows0 * args.cpu_seconds / max(t_small, 1e-3))))
    r0 = max(0, n // 2 - nrows // 2)
    nrows = min(nrows, n - r0)
    want, t_cpu = cpu.run(lambda: O.imfilter_dense_rows(img64, k15, [-7, -7], "reflect", r0, nrows), 1)
    _, t_par = cpu.run(lambda: O.imfilter_dense_rows(img64, k15, [-7, -7], "reflect", r0, nrows), cpu.threads)
    cpu_gpix = n * nrows / t_cpu / 1e9
    got = out[:, r0:r0 + nrows].astype(np.float64)
    err = float(np.max(np.abs(got - want)) / np.max(np.abs(want)))
    log("bench: cpu baseline %.4f Gpix/s on %d rows (%.1f s, one thread; %.4f Gpix/s on %d threads, %s); "
        "max rel err vs FP64 %.2e" % (cpu_gpix, nrows, t_cpu, n * nrows / t_par / 1e9, cpu.threads, cpu.build, err))
    head = {"img64": img64, "r0": r0, "nrows": nrows, "want": want, "k15": k15}

    extras = {}
    serving = None
    del out
    tile_tune = None
    if (not args.no_extras or args.serve_requests > 0) and not args.no_autotune:
        # tile shapes of the fused separable kernel for this machine (the C1 extra
        # and the serving stream's separable requests use them)
        from imagefiltering import tiles
        t0 = time.time()
        tt = tiles.tune()
        try:
            tiles.save(tt)
        except OSError as e:  # an unwritable cache location only loses the cache
            log("bench: tile cache not saved: %s" % e)
        tile_tune = {"seconds": round(time.time() - t0, 2),
                     "rules": [[r["dtype"], r["kmax"], r["max_pixels"], r["tile_h"]] for r in tt["rules"]],
                     "rule_fields": ["dtype", "kmax", "max_pixels", "tile_h"], "trials": tt["trials"]}
        log("bench: tile autotune %.1f s: %s" % (tile_tune["seconds"],
                                                 ["%s k<=%d px<=%d -> %d" % tuple(r) for r in tile_tune["rules"]]))
    if not args.no_extras:
        extras = run_extras(IF, _lib, L, ctx, T, img, d_img, d_out, b_reflect, args, cpu, head)
    serving = run_serving(args)  # --serve-requests 0 skips it
    if args.profile_copy:
        # a plain f32 copy of a known size: 256 MiB read + 256 MiB written
        m = 1 << 26
        src = IF.DeviceArray((m,), IF.Float32)
        dst = IF.DeviceArray((m,), IF.Float32)
        b = _lib.IfmBorder()
        b.style = _lib.REPLICATE
        _lib.check(L.ifm_padarray(ctx, ifm_array(src), ifm_array(dst), b))
        L.ifm_ctx_synchronize(ctx)

    traffic = None
    tpath = os.path.join(REPO, "profiles", "traffic_dense_mfma_4096_k15.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("bytes_per_launch")

    rec = {
        "metric": "Gpix/s & %HBM peak; served p50/p99 vs plain imfilter at rel err<=1e-5",
        "value": round(gpix, 3),
        "unit": "Gpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded uniform [0,1) image, seeded N(0,1) taps normalised to sum 1)",
        "config": {"workload": "imfilter(Float32, img 4096x4096 Float32, centered(15x15 N(0,1)/sum), \"reflect\", "
                               "Algorithm.FIR()) -- dense 2-D, MFMA FP16 hi/lo split banded-Toeplitz",
                   "image": [n, n], "kernel": [15, 15], "border": "reflect", "out_eltype": "Float32"},
        "max_rel_err_vs_fp64": err,
        "roofline": {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_dense_mfma<1,2,false,15>",
                     "mfma_f16_tflops_issued": round(px * 3 * 2 * 32 * 15 / (ms * 1e-3) / 1e12, 1),
                     "algorithmic_tflops_fp32": round(flops_alg / (ms * 1e-3) / 1e12, 2)},
        "cpu_baseline": {"value": round(cpu_gpix, 6), "unit": "Gpix/s", "cores": 1, "kind": "port",
                         "sample": "rows %d..%d (%d x 4096 px) of the same image and kernel, FP64 direct FIR "
                                   "(oracle/ifm_oracle.c, %s, one thread like CPU1(Algorithm.FIR()))"
                                   % (r0, r0 + nrows - 1, nrows, cpu.build),
                         "parallel": {"value": round(n * nrows / t_par / 1e9, 6), "cores": cpu.threads,
                                      "note": "the same sample on all cores (parallel restatement)"}},
        "host": {"nproc": cpu.nproc, "oracle_build": cpu.build, "oracle_threads": cpu.threads},
        "serving": serving,
        "extras": extras,
        "tile_autotune": tile_tune,
    }
    print(json.dumps(rec), flush=True)


def run_serving(args):
    """The serving mix: the same seeded request stream through the batched
    request loop and through plain per-call dispatch, first closed-loop (all
    requests arrive at once: saturation throughput), then open-loop with
    Poisson arrivals at 0.8 x the batched saturation rate (latency).  The
    batch size and in-flight depth come from the schedule autotuner, run on a
    sample of the same stream first."""
    from imagefiltering import serving
    n = args.serve_requests
    if n <= 0:
        return None
    budget = 64 << 30
    res = {"requests": n, "hbm_budget_bytes": budget,
           "mix": "sides log-uniform [64,4096]; F32 or BF16; gaussian sigma~U[0.5,8] | sobel gradient pair | "
                  "median 7x7 | dense 15x15; five borders; inputs generated on the device"}
    mix = serving.request_mix(n, rate=None)
    sched, table = serving.autotune(mix, budget_bytes=budget)
    res["schedule"] = sched
    res["autotune_trials"] = table
    log("serving autotune: %s" % sched)
    for mode in ("batched", "percall"):
        _, done = serving.serve(mix, mode, budget_bytes=budget, **sched)
        res["closed_" + mode] = serving.stats(mix, done)
        log("serving closed-loop %-8s %s" % (mode, res["closed_" + mode]))
    sat = res["closed_batched"]["images_per_s"]
    rate = 0.8 * sat
    mix = serving.request_mix(n, rate=rate)
    res["open_rate_per_s"] = rate
    for mode in ("batched", "percall"):
        _, done = serving.serve(mix, mode, budget_bytes=budget, **sched)
        res["open_" + mode] = serving.stats(mix, done)
        log("serving open-loop %.0f/s %-8s %s" % (rate, mode, res["open_" + mode]))
    res["speedup_images_per_s"] = sat / res["closed_percall"]["images_per_s"]
    return res


class CpuBaseline:
    """The oracle's FP64 restatement timed on this host: built here with -O3
    -march=native and OpenMP (falling back to the portable one-thread build),
    run once with the reference's own threading for the configuration (one
    thread for CPU1 paths, all cores for FIRTiled) and once on all cores (the
    "parallel restatement").  Samples are sized to about `target` seconds."""

    def __init__(self, O, target):
        import subprocess
        self.O, self.target = O, max(0.05, target)
        r = subprocess.run(["make", "-f", "oracle/Makefile", "native"], cwd=REPO, capture_output=True, text=True)
        self.build = "portable -O2, one thread"
        self.threads = 1
        if r.returncode == 0 and os.path.exists(O.NATIVE_PATH):
            self.threads = O.use_library(O.NATIVE_PATH)
            self.build = "-O3 -march=native -fopenmp (built on this host)"
        try:
            self.nproc = len(os.sched_getaffinity(0))
        except AttributeError:
            self.nproc = os.cpu_count()
        self.threads = max(1, min(self.threads, self.nproc))

    def run(self, fn, threads):
        self.O.set_threads(threads)
        t0 = time.perf_counter()
        res = fn()
        return res, time.perf_counter() - t0

    def scaled(self, make, units_max, threads, units0=1):
        """make(units) -> (callable, pixels); picks units so one thread-count
        run takes about target seconds; returns (result, seconds, units, px)."""
        fn, px = make(units0)
        _, t = self.run(fn, threads)
        units = int(max(units0, min(units_max, units0 * self.target / max(t, 1e-4))))
        fn, px = make(units)
        res, t = self.run(fn, threads)
        return res, t, units, px

    def entry(self, px, secs, cores, sample, note):
        return {"value": round(px / secs / 1e9, 6), "unit": "Gpix/s", "cores": cores, "kind": "port",
                "sample": sample, "threading": note}


def relerr(got, want):
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    return float(np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1e-300))


def run_extras(IF, _lib, L, ctx, T, img, d_img, d_out, b_reflect, args, cpu, head):
    """The other configurations at their full sizes, a few steps each; every
    floating-point entry states its max relative error against the FP64 oracle
    on a sample and the order filters their bit-exact check; each
    configuration's oracle run is also its CPU baseline."""
    from imagefiltering.device import ifm_array
    from imagefiltering.imfilter import _border_struct, _kernel_struct
    from oracle import oracle as O
    ex = {}
    n = img.shape[0]
    steps = max(3, args.steps // 10)

    def rate(name, fn, px, bytes_px, steps=steps, **extra):
        ms = T.run(fn, steps, 2)
        gbs = px * bytes_px / (ms * 1e-3) / 1e9
        ex[name] = {"ms": round(ms, 4), "gpix_s": round(px / (ms * 1e-3) / 1e9, 3), "hbm_gbs": round(gbs, 1),
                    "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "alg_bytes_per_px": bytes_px}
        ex[name].update(extra)
        log("extra %-36s %8.3f ms  %8.2f Gpix/s  %7.0f GB/s" % (name, ms, ex[name]["gpix_s"], gbs))

    def note(name, **kv):
        ex[name].update(kv)
        log("      %-36s %s" % (name, ", ".join("%s=%s" % (k, v if not isinstance(v, dict) else v.get("value"))
                                                  for k, v in kv.items())))

    a_in, a_out = ifm_array(d_img), ifm_array(d_out)
    img64, rows_r0, rows_n, want15 = head["img64"], head["r0"], head["nrows"], head["want"]
    rows = slice(rows_r0, rows_r0 + rows_n)

    def out_rows(d):
        return np.asarray(d.to_host())[:, rows]

    rng31 = np.random.default_rng(20261215)
    k31 = norm_kernel(rng31, 31)
    ks31 = _kernel_struct(IF.centered(k31))
    rate("dense31_mfma_4096", lambda: L.ifm_imfilter_dense(ctx, a_in, a_out, None, ks31, b_reflect, 0), n * n, 8)
    (want31, t31, u31, px31) = cpu.scaled(
        lambda u: (lambda: O.imfilter_dense(img64, k31, [-15, -15], "reflect", [0, rows_r0], [n, u]), n * u),
        rows_n, 1, units0=2)
    _, tp31 = cpu.run(lambda: O.imfilter_dense(img64, k31, [-15, -15], "reflect", [0, rows_r0], [n, u31]),
                      cpu.threads)

    def out_rows31(d):
        return np.asarray(d.to_host())[:, rows_r0:rows_r0 + u31]
    note("dense31_mfma_4096", max_rel_err_vs_fp64=relerr(out_rows31(d_out), want31),
         cpu_baseline=cpu.entry(px31, t31, 1, "%d x %d px, direct FP64 FIR" % (n, u31), "one thread, as CPU1(FIR())"),
         cpu_parallel=cpu.entry(px31, tp31, cpu.threads, "same sample", "all cores (parallel restatement)"))

    # Algorithm.FFT(): the reference's default for these kernels (maxlen > 30)
    k15 = head["k15"]
    ks15 = _kernel_struct(IF.centered(k15))
    rate("dense15_fft_4096", lambda: L.ifm_imfilter_fft(ctx, a_in, a_out, None, ks15, b_reflect), n * n, 8, 10,
         kernel="k_fft_tile<64> (tiled overlap-save; kernel spectrum cached)")
    err15f = relerr(out_rows(d_out), want15)
    # the CPU baseline of the FFT path: numpy's FFT (pocketfft, one thread like
    # FFTW's default) over the whole padded 4374 x 4374 array
    fft64, tf = cpu.run(lambda: O.imfilter_fft(img64, k15, (-7, -7), "reflect"), 1)
    note("dense15_fft_4096", max_rel_err_vs_fp64=err15f,
         max_rel_err_vs_fp64_fft=relerr(d_out.to_host(), fft64),
         cpu_baseline=cpu.entry(n * n, tf, 1, "whole image, FP64 rfft/irfft over 4374 x 4374 (oracle.imfilter_fft)",
                                "one thread, as FFTW's default"))
    del fft64
    k15b = _kernel_struct(IF.centered(norm_kernel(np.random.default_rng(20261216), 15)))
    alt = [0]

    def fft_alternating():
        alt[0] ^= 1
        return L.ifm_imfilter_fft(ctx, a_in, a_out, None, ks15 if alt[0] else k15b, b_reflect)
    rate("dense15_fft_4096_new_kernel_each_call", fft_alternating, n * n, 8, 6)
    rate("dense31_fft_4096", lambda: L.ifm_imfilter_fft(ctx, a_in, a_out, None, ks31, b_reflect), n * n, 8, 10,
         kernel="k_fft_tile<128>")
    note("dense31_fft_4096", max_rel_err_vs_fp64=relerr(out_rows31(d_out), want31))
    rate("dense15_valu_fp32_4096", lambda: L.ifm_imfilter_dense(ctx, a_in, a_out, None, ks15, b_reflect, 1), n * n,
         8, 3)
    note("dense15_valu_fp32_4096", max_rel_err_vs_fp64=relerr(out_rows(d_out), want15))
    # small non-separable kernels: the register-blocked VALU stencils
    # (imfilter(Float32, img, centered(KxK), "reflect", Algorithm.FIR()))
    for K in (5, 7):
        kK = norm_kernel(np.random.default_rng(20261220 + K), K)
        ksK = _kernel_struct(IF.centered(kK))
        name = "dense%d_valu_4096" % K
        rate(name, lambda ksK=ksK: L.ifm_imfilter_dense(ctx, a_in, a_out, None, ksK, b_reflect, 0), n * n, 8, 10,
             kernel="k_dense2d_small<%d,%d>" % (K, K))
        (wantK, tK, uK, pxK) = cpu.scaled(
            lambda u, kK=kK, K=K: (lambda: O.imfilter_dense(img64, kK, [-(K // 2)] * 2, "reflect", [0, rows_r0],
                                                            [n, u]), n * u), rows_n, 1, units0=8)
        outK = np.asarray(d_out.to_host())[:, rows_r0:rows_r0 + uK]
        note(name, max_rel_err_vs_fp64=relerr(outK, wantK),
             cpu_baseline=cpu.entry(pxK, tK, 1, "%d x %d px, direct FP64 FIR" % (n, uK), "one thread, as CPU1(FIR())"))

    # the drop-in call through the public API with the default algorithm:
    # filter_algorithm's isfinite scan, then FFT
    kern15 = IF.centered(k15)
    d_drop = [None]

    def dropin():
        d_drop[0] = IF.imfilter(d_img, kern15, "reflect")
    rate("dropin_imfilter_default_4096", dropin, n * n, 8, 10,
         call="imfilter(img::DeviceArray, centered(15x15), \"reflect\"): isfinite scan + FFT + output allocation")
    note("dropin_imfilter_default_4096", max_rel_err_vs_fp64=relerr(out_rows(d_drop[0]), want15))
    d_drop[0] = None

    # BF16 storage in and out (error dominated by the BF16 rounding of the output)
    xbh = IF.bf16(img)
    xb = IF.DeviceArray.from_host(xbh)
    ob = xb.similar()
    a_b, a_ob = ifm_array(xb), ifm_array(ob)
    rate("dense15_mfma_bf16io_4096", lambda: L.ifm_imfilter_dense(ctx, a_b, a_ob, None, ks15, b_reflect, 0), n * n,
         4)
    xb64 = np.asfortranarray(IF.eltypes.bf16_to_float32(xbh), dtype=np.float64)
    wb = O.imfilter_dense(xb64, k15, [-7, -7], "reflect", [0, rows_r0], [n, 64])
    gotb = IF.eltypes.bf16_to_float32(np.asarray(ob.to_host()))[:, rows_r0:rows_r0 + 64]
    note("dense15_mfma_bf16io_4096", max_rel_err_vs_fp64=relerr(gotb, wb),
         err_note="BF16 output: the rounding of the stored result (2^-9 relative) dominates")
    del xb, ob, xb64
    # the same 15 x 15 call on host arrays: upload, filter, download (wall clock, PCIe included)
    out_h = np.empty_like(img)
    ah_in, ah_out = ifm_array(img), ifm_array(out_h)
    for _ in range(2):
        _lib.check(L.ifm_imfilter_dense(ctx, ah_in, ah_out, None, ks15, b_reflect, 0))
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        L.ifm_imfilter_dense(ctx, ah_in, ah_out, None, ks15, b_reflect, 0)
    ms_h = (time.perf_counter() - t0) * 1e3 / reps
    ex["dense15_mfma_4096_host_arrays"] = {"ms": round(ms_h, 4), "gpix_s": round(n * n / (ms_h * 1e-3) / 1e9, 3),
                                           "note": "host numpy arrays in and out; wall clock incl. PCIe copies"}
    log("extra %-36s %8.3f ms  %8.2f Gpix/s  (host arrays, PCIe included)" % ("dense15_mfma_4096_host_arrays", ms_h,
                                                                             n * n / (ms_h * 1e-3) / 1e9))
    del out_h

    # config 1: 1024^2 gaussian sigma 3 replicate, F32 -> F64 (FIRTiled: all threads in the reference)
    c1h = np.random.default_rng(1).random((1024, 1024), dtype=np.float32)
    c1 = IF.DeviceArray.from_host(c1h)
    o1 = c1.similar(IF.Float64)
    kern = IF.KernelFactors.gaussian((3.0, 3.0))
    fa = (_lib.IfmFactor * 2)()
    keep = []
    facs = []
    for i, f in enumerate(kern):
        t = np.ascontiguousarray(f.data.data, dtype=np.float64)
        keep.append(t)
        facs.append((f.Npre, t, f.data.first[0]))
        fa[i].axis, fa[i].len, fa[i].first = f.Npre, len(t), f.data.first[0]
        fa[i].taps = t.ctypes.data_as(C.POINTER(C.c_double))
    a1, ao1 = ifm_array(c1), ifm_array(o1)
    brep = _border_struct(IF.Pad("replicate"), IF.Float32)
    rate("gauss13_sep_1024_f64out", lambda: L.ifm_imfilter_fir(ctx, a1, ao1, None, fa, 2, brep), 1024 * 1024, 12, 20)
    c164 = np.asfortranarray(c1h, dtype=np.float64)
    w1, t1p = cpu.run(lambda: O.imfilter_factors(c164, facs, "replicate"), cpu.threads)
    _, t1s = cpu.run(lambda: O.imfilter_factors(c164, facs, "replicate"), 1)
    note("gauss13_sep_1024_f64out", max_rel_err_vs_fp64=relerr(o1.to_host(), w1),
         cpu_baseline=cpu.entry(1024 * 1024, t1p, cpu.threads, "whole image, FP64 factor cascade",
                                "all cores, as FIRTiled's threads"),
         cpu_one_thread=cpu.entry(1024 * 1024, t1s, 1, "whole image", "one thread"))
    del c1, o1

    from imagefiltering.imfilter import _processed, _run, factorkernel
    from imagefiltering import Kernel, KernelFactors

    def api(alg, out, x, kernel, border):
        comps = _processed(kernel)  # factoring (SVD) happens once, outside the timed loop
        return lambda: _run(alg, out, x, comps, IF.borderinstance(border), None)

    # sobel (SVD-factored to 3 + 3 taps) and LoG(2.0) (17 x 17) on the 4096^2 image, F64 out
    o64 = d_img.similar(IF.Float64)
    sob = Kernel.sobel()[0]
    rate("sobel1_4096_f64out", api(IF.Algorithm.FIR(), o64, d_img, sob, "replicate"), n * n, 12)
    sf = factorkernel(sob)
    sfacs = [(d, np.asarray(f.data, dtype=np.float64).ravel(), f.first[d]) for d, f in enumerate(sf)]
    ws, tss, us, pxs = cpu.scaled(lambda u: (lambda: O.imfilter_factors(img64, sfacs, "replicate", [0, rows_r0],
                                                                        [n, u]), n * u), rows_n, 1, 8)
    _, tsp = cpu.run(lambda: O.imfilter_factors(img64, sfacs, "replicate", [0, rows_r0], [n, us]), cpu.threads)
    got = np.asarray(o64.to_host())[:, rows_r0:rows_r0 + us]
    note("sobel1_4096_f64out", max_rel_err_vs_fp64=relerr(got, ws),
         cpu_baseline=cpu.entry(pxs, tsp, cpu.threads, "%d x %d px, FP64 factor cascade" % (n, us),
                                "all cores, as FIRTiled's threads"),
         cpu_one_thread=cpu.entry(pxs, tss, 1, "same sample", "one thread"))
    logk = Kernel.LoG(2.0)
    rate("log2_17x17_4096_f64out", api(IF.Algorithm.FIR(), o64, d_img, logk, "replicate"), n * n, 12,
         kernel="k_dense_mfma (Algorithm.FIR())")
    lk = np.asarray(logk.data, dtype=np.float64)
    wl, tl, ul, pxl = cpu.scaled(lambda u: (lambda: O.imfilter_dense(img64, lk, list(logk.first), "replicate",
                                                                     [0, rows_r0], [n, u]), n * u), rows_n, 1, 2)
    _, tlp = cpu.run(lambda: O.imfilter_dense(img64, lk, list(logk.first), "replicate", [0, rows_r0], [n, ul]),
                     cpu.threads)
    note("log2_17x17_4096_f64out", max_rel_err_vs_fp64=relerr(np.asarray(o64.to_host())[:, rows_r0:rows_r0 + ul], wl),
         cpu_baseline=cpu.entry(pxl, tl, 1, "%d x %d px, direct FP64 FIR" % (n, ul), "one thread, as CPU1(FIR())"),
         cpu_parallel=cpu.entry(pxl, tlp, cpu.threads, "same sample", "all cores (parallel restatement)"))
    rate("log2_17x17_fft_4096_f64out", api(IF.Algorithm.FFT(), o64, d_img, logk, "replicate"), n * n, 12,
         kernel="k_fft_tile<64> (Algorithm.FFT(), the reference's default for LoG(2.0))")
    note("log2_17x17_fft_4096_f64out",
         max_rel_err_vs_fp64=relerr(np.asarray(o64.to_host())[:, rows_r0:rows_r0 + ul], wl))
    del o64

    # 512 x 512 x 256 volume: separable FIR (symmetric, circular) and IIR (circular), F64 out
    vrng = np.random.default_rng(20261205)
    # column-major like a Julia Array: dimension 1 (512) is the contiguous one
    vol = np.asfortranarray(vrng.random((512, 512, 256), dtype=np.float32)
                            + vrng.normal(0, 0.05, (512, 512, 256)).astype(np.float32))
    d_vol = IF.DeviceArray.from_host(vol)
    vol64 = np.asfortranarray(vol, dtype=np.float64)
    del vol
    o_vol = d_vol.similar(IF.Float64)
    nv = 512 * 512 * 256
    g3 = KernelFactors.gaussian((2.0, 2.0, 1.5))
    gf = [(f.Npre, np.asarray(f.data.data, dtype=np.float64), f.data.first[0]) for f in g3]
    for style in ("symmetric", "circular"):
        name = "gauss_sep3d_512x512x256_" + style
        rate(name, api(IF.Algorithm.FIRTiled(), o_vol, d_vol, g3, style), nv, 12, 3)
        z0 = 120
        wv, tvs, uz, pxv = cpu.scaled(lambda u: (lambda: O.imfilter_factors(vol64, gf, style, [0, 0, z0],
                                                                            [512, 512, u]), 512 * 512 * u),
                                      16, 1, 1)
        _, tvp = cpu.run(lambda: O.imfilter_factors(vol64, gf, style, [0, 0, z0], [512, 512, uz]), cpu.threads)
        got = np.asarray(o_vol.to_host())[:, :, z0:z0 + uz]
        note(name, max_rel_err_vs_fp64=relerr(got, wv),
             cpu_baseline=cpu.entry(pxv, tvp, cpu.threads, "512 x 512 x %d planes, FP64 factor cascade" % uz,
                                    "all cores, as FIRTiled's threads"),
             cpu_one_thread=cpu.entry(pxv, tvs, 1, "same sample", "one thread"))
    i3 = KernelFactors.IIRGaussian((2.0, 2.0, 1.5))
    rate("iir_gauss3d_512x512x256_circular", api(IF.Algorithm.IIR(), o_vol, d_vol, i3, "circular"), nv, 12, 3)
    ik = [(f.Npre, f.data.a, f.data.b, f.data.scale, f.data.M) for f in i3]
    wi, tip = cpu.run(lambda: O.imfilter_iir(vol64, ik, "circular"), cpu.threads)
    sub = np.asfortranarray(vol64[:, :, :32])
    _, tis = cpu.run(lambda: O.imfilter_iir(sub, ik, "circular"), 1)
    note("iir_gauss3d_512x512x256_circular", max_rel_err_vs_fp64=relerr(o_vol.to_host(), wi),
         cpu_baseline=cpu.entry(512 * 512 * 32, tis, 1, "512 x 512 x 32 sub-volume, FP64 Triggs-Sdika (three axes)",
                                "one thread, as the reference's IIR"),
         cpu_parallel=cpu.entry(nv, tip, cpu.threads, "whole volume", "all cores (parallel restatement)"))
    # a dense 3-D stencil (register-blocked): imfilter(vol, centered(3x3x3), "reflect", Algorithm.FIR()), F64 out
    k3 = np.random.default_rng(20261230).standard_normal((3, 3, 3))
    k3 /= k3.sum()
    rate("dense3d_3x3x3_512x512x256", api(IF.Algorithm.FIR(), o_vol, d_vol, IF.centered(k3), "reflect"), nv, 12, 3,
         kernel="k_dense3d_small<3,3,3>")
    z3 = 120
    w3, t3, u3, px3 = cpu.scaled(lambda u: (lambda: O.imfilter_dense(vol64, k3, [-1, -1, -1], "reflect", [0, 0, z3],
                                                                     [512, 512, u]), 512 * 512 * u), 16, 1, 1)
    note("dense3d_3x3x3_512x512x256", max_rel_err_vs_fp64=relerr(np.asarray(o_vol.to_host())[:, :, z3:z3 + u3], w3),
         cpu_baseline=cpu.entry(px3, t3, 1, "512 x 512 x %d planes, direct FP64 FIR" % u3,
                                "one thread, as CPU1(FIR())"))
    del d_vol, o_vol, vol64, wi, sub

    # 8192^2 N0f8 with 5% salt and pepper: median 5x5 / 7x7, extrema 3x3 / 9x9, minimum 9x9
    crng = np.random.default_rng(20261206)
    codes = crng.integers(0, 256, (8192, 8192), dtype=np.uint8)
    u = crng.random((8192, 8192))
    codes[u < 0.025] = 0
    codes[(u >= 0.025) & (u < 0.05)] = 255
    del u
    d_c = IF.DeviceArray.from_host(IF.n0f8(codes))
    c64 = np.asfortranarray(codes, dtype=np.float64)
    oa, ob = d_c.similar(), d_c.similar()
    of32 = d_c.similar(IF.Float32)
    ac, aa, abb, af = ifm_array(d_c), ifm_array(oa), ifm_array(ob), ifm_array(of32)
    npx = 8192 * 8192
    r0c = 4000
    for w in (5, 7):
        wf, wl_ = _lib.i64([-(w // 2)] * 2), _lib.i64([w] * 2)
        name = "median%dx%d_8192_n0f8" % (w, w)
        rate(name, lambda wf=wf, wl_=wl_: L.ifm_mapwindow_median(ctx, ac, aa, None, wf, wl_, brep), npx, 2, 5,
             output="N0f8 codes")
        rate(name + "_dropin_f32", lambda wf=wf, wl_=wl_: L.ifm_mapwindow_median(ctx, ac, af, None, wf, wl_, brep),
             npx, 5, 5, output="Float32, the eltype of mapwindow(median!, ::N0f8)")
        h = w // 2
        wm, tm, um, pxm = cpu.scaled(lambda u: (lambda: O.median(c64, [-h, -h], [w, w], "replicate", [0, r0c],
                                                                 [8192, u]), 8192 * u), 64, 1, 1)
        got = np.asarray(oa.to_host())[:, r0c:r0c + um]
        gotf = np.asarray(of32.to_host())[:, r0c:r0c + um]
        note(name, bit_exact=bool(np.array_equal(got, wm)), checked_px=int(got.size),
             cpu_baseline=cpu.entry(pxm, tm, 1, "8192 x %d px, FP64 window sort" % um, "one thread, as mapwindow"))
        note(name + "_dropin_f32", bit_exact=bool(np.array_equal(gotf, (wm / 255).astype(np.float32))))
    for w in (3, 9):
        win = _lib.i64([w, w])
        name = "extrema%dx%d_8192_n0f8" % (w, w)
        rate(name, lambda win=win: L.ifm_extrema_filter(ctx, ac, aa, abb, win), npx, 3, 5)
        m = 8
        band = np.asfortranarray(c64[:, r0c - m:r0c + 64 + m])
        (mn, mx), te = cpu.run(lambda: O.extrema_lemire(band, [w, w]), 1)
        ok = (np.array_equal(np.asarray(oa.to_host())[:, r0c:r0c + 64], mn[:, m:m + 64]) and
              np.array_equal(np.asarray(ob.to_host())[:, r0c:r0c + 64], mx[:, m:m + 64]))
        note(name, bit_exact=bool(ok), checked_px=8192 * 64,
             cpu_baseline=cpu.entry(band.size, te, 1, "8192 x %d band, Lemire min/max" % band.shape[1],
                                    "one thread, as extrema_filter"))
    wf, wl_ = _lib.i64([-4, -4]), _lib.i64([9, 9])
    rate("minimum9x9_8192_n0f8", lambda: L.ifm_mapwindow_minmax(ctx, ac, aa, None, None, wf, wl_, brep), npx, 2, 5)
    (wmn, _), tmn, umn, pxmn = cpu.scaled(lambda u: (lambda: O.minmax(c64, [-4, -4], [9, 9], "replicate", [0, r0c],
                                                                      [8192, u]), 8192 * u), 64, 1, 1)
    note("minimum9x9_8192_n0f8", bit_exact=bool(np.array_equal(np.asarray(oa.to_host())[:, r0c:r0c + umn], wmn)),
         checked_px=8192 * umn,
         cpu_baseline=cpu.entry(pxmn, tmn, 1, "8192 x %d px, FP64 window min" % umn, "one thread, as mapwindow"))
    return ex


if __name__ == "__main__":
    main()
