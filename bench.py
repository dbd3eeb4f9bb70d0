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
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-extras", action="store_true", help="skip the other configurations (serving runs unless --serve-requests 0)")
    ap.add_argument("--serve-requests", type=int, default=20000)
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

    # accuracy of the timed output against the FP64 oracle on a row sample
    out = d_out.to_host()
    from oracle import oracle as O
    rows = 64
    r0 = n // 2 - rows // 2
    img64 = np.asfortranarray(img, dtype=np.float64)   # the oracle's own layout, outside the timing
    t0 = time.time()
    want = O.imfilter_dense_rows(img64, k15, [-7, -7], "reflect", r0, 4)
    t_small = time.time() - t0
    nrows = int(max(4, min(n, 4 * args.cpu_seconds / max(t_small, 1e-3))))
    r0 = max(0, n // 2 - nrows // 2)
    nrows = min(nrows, n - r0)
    t0 = time.time()
    want = O.imfilter_dense_rows(img64, k15, [-7, -7], "reflect", r0, nrows)
    t_cpu = time.time() - t0
    cpu_gpix = n * nrows / t_cpu / 1e9
    got = out[:, r0:r0 + nrows].astype(np.float64)
    err = float(np.max(np.abs(got - want)) / np.max(np.abs(want)))
    log("bench: cpu baseline %.4f Gpix/s on %d rows (%.1f s); max rel err vs FP64 %.2e" % (cpu_gpix, nrows, t_cpu,
                                                                                          err))

    extras = {}
    serving = None
    del out
    tile_tune = None
    if not args.no_extras or args.serve_requests > 0:
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
        extras = run_extras(IF, _lib, L, ctx, T, img, d_img, d_out, b_reflect, args)
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
                     "kernel": "k_dense_mfma<1,2>",
                     "mfma_f16_tflops_issued": round(px * 3 * 2 * 32 * 16 / (ms * 1e-3) / 1e12, 1),
                     "algorithmic_tflops_fp32": round(flops_alg / (ms * 1e-3) / 1e12, 2)},
        "cpu_baseline": {"value": round(cpu_gpix, 6), "unit": "Gpix/s", "cores": 1, "kind": "port",
                         "sample": "rows %d..%d (%d x 4096 px) of the same image and kernel, FP64 direct FIR "
                                   "(oracle/ifm_oracle.c, single thread like CPU1(Algorithm.FIR()))"
                                   % (r0, r0 + nrows - 1, nrows)},
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


def run_extras(IF, _lib, L, ctx, T, img, d_img, d_out, b_reflect, args):
    """The other configurations at their full sizes, a few steps each."""
    from imagefiltering.device import ifm_array
    from imagefiltering.imfilter import _border_struct, _kernel_struct
    ex = {}
    n = img.shape[0]
    steps = max(3, args.steps // 10)

    def rate(name, fn, px, bytes_px, steps=steps):
        ms = T.run(fn, steps, 2)
        gbs = px * bytes_px / (ms * 1e-3) / 1e9
        ex[name] = {"ms": round(ms, 4), "gpix_s": round(px / (ms * 1e-3) / 1e9, 3), "hbm_gbs": round(gbs, 1),
                    "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
        log("extra %-28s %8.3f ms  %8.2f Gpix/s  %7.0f GB/s" % (name, ms, ex[name]["gpix_s"], gbs))

    a_in, a_out = ifm_array(d_img), ifm_array(d_out)
    k31 = IF.centered(norm_kernel(np.random.default_rng(20261215), 31))
    ks31 = _kernel_struct(k31)
    rate("dense31_mfma_4096", lambda: L.ifm_imfilter_dense(ctx, a_in, a_out, None, ks31, b_reflect, 0), n * n, 8)
    k15 = IF.centered(norm_kernel(np.random.default_rng(20261214), 15))
    ks15 = _kernel_struct(k15)
    # repeated kernel: its spectrum stays in the FFT workspace between calls
    rate("dense15_fft_4096", lambda: L.ifm_imfilter_fft(ctx, a_in, a_out, None, ks15, b_reflect), n * n, 8, 3)
    # a new kernel every call (two kernels alternating): the kernel's transform is redone
    k15b = _kernel_struct(IF.centered(norm_kernel(np.random.default_rng(20261216), 15)))
    alt = [0]

    def fft_alternating():
        alt[0] ^= 1
        return L.ifm_imfilter_fft(ctx, a_in, a_out, None, ks15 if alt[0] else k15b, b_reflect)
    rate("dense15_fft_4096_new_kernel_each_call", fft_alternating, n * n, 8, 4)
    rate("dense15_valu_fp32_4096", lambda: L.ifm_imfilter_dense(ctx, a_in, a_out, None, ks15, b_reflect, 1), n * n, 8, 3)
    # BF16 storage in and out
    xb = IF.DeviceArray.from_host(IF.bf16(img))
    ob = xb.similar()
    a_b, a_ob = ifm_array(xb), ifm_array(ob)
    ks15 = _kernel_struct(k15)
    rate("dense15_mfma_bf16io_4096", lambda: L.ifm_imfilter_dense(ctx, a_b, a_ob, None, ks15, b_reflect, 0), n * n, 4)
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
    log("extra %-28s %8.3f ms  %8.2f Gpix/s  (host arrays, PCIe included)" % ("dense15_mfma_4096_host_arrays", ms_h,
                                                                             n * n / (ms_h * 1e-3) / 1e9))
    del out_h
    # config 1: 1024^2 gaussian sigma 3 replicate, F32 -> F64
    c1 = IF.DeviceArray.from_host(np.random.default_rng(1).random((1024, 1024), dtype=np.float32))
    o1 = c1.similar(IF.Float64)
    kern = IF.KernelFactors.gaussian((3.0, 3.0))
    fa = (_lib.IfmFactor * 2)()
    keep = []
    for i, f in enumerate(kern):
        t = np.ascontiguousarray(f.data.data, dtype=np.float64)
        keep.append(t)
        fa[i].axis, fa[i].len, fa[i].first = f.Npre, len(t), f.data.first[0]
        fa[i].taps = t.ctypes.data_as(C.POINTER(C.c_double))
    a1, ao1 = ifm_array(c1), ifm_array(o1)
    brep = _border_struct(IF.Pad("replicate"), IF.Float32)
    rate("gauss13_sep_1024_f64out", lambda: L.ifm_imfilter_fir(ctx, a1, ao1, None, fa, 2, brep), 1024 * 1024, 12, 20)
    del c1, o1

    from imagefiltering.imfilter import _processed, _run
    from imagefiltering import Kernel, KernelFactors

    def api(alg, out, x, kernel, border):
        comps = _processed(kernel)  # factoring (SVD) happens once, outside the timed loop
        return lambda: _run(alg, out, x, comps, IF.borderinstance(border), None)

    # sobel (SVD-factored to 3 + 3 taps) and LoG(2.0) (17 x 17, matrix cores) on the 4096^2 image, F64 out
    o64 = d_img.similar(IF.Float64)
    rate("sobel1_4096_f64out", api(IF.Algorithm.FIR(), o64, d_img, Kernel.sobel()[0], "replicate"), n * n, 12)
    rate("log2_17x17_4096_f64out", api(IF.Algorithm.FIR(), o64, d_img, Kernel.LoG(2.0), "replicate"), n * n, 12)
    del o64

    # 512 x 512 x 256 volume: separable FIR (symmetric, circular) and IIR (circular), F64 out
    vrng = np.random.default_rng(20261205)
    # column-major like a Julia Array: dimension 1 (512) is the contiguous one
    vol = np.asfortranarray(vrng.random((512, 512, 256), dtype=np.float32)
                            + vrng.normal(0, 0.05, (512, 512, 256)).astype(np.float32))
    d_vol = IF.DeviceArray.from_host(vol)
    del vol
    o_vol = d_vol.similar(IF.Float64)
    nv = 512 * 512 * 256
    g3 = KernelFactors.gaussian((2.0, 2.0, 1.5))
    rate("gauss_sep3d_512x512x256_symmetric", api(IF.Algorithm.FIRTiled(), o_vol, d_vol, g3, "symmetric"), nv, 12, 3)
    rate("gauss_sep3d_512x512x256_circular", api(IF.Algorithm.FIRTiled(), o_vol, d_vol, g3, "circular"), nv, 12, 3)
    i3 = KernelFactors.IIRGaussian((2.0, 2.0, 1.5))
    rate("iir_gauss3d_512x512x256_circular", api(IF.Algorithm.IIR(), o_vol, d_vol, i3, "circular"), nv, 12, 3)
    del d_vol, o_vol

    # 8192^2 N0f8 with 5% salt and pepper: median 5x5 / 7x7, extrema 3x3 / 9x9, minimum 9x9
    crng = np.random.default_rng(20261206)
    codes = crng.integers(0, 256, (8192, 8192), dtype=np.uint8)
    u = crng.random((8192, 8192))
    codes[u < 0.025] = 0
    codes[(u >= 0.025) & (u < 0.05)] = 255
    d_c = IF.DeviceArray.from_host(IF.n0f8(codes) if hasattr(IF, "n0f8") else codes)
    del codes, u
    oa, ob = d_c.similar(), d_c.similar()
    ac, aa, abb = ifm_array(d_c), ifm_array(oa), ifm_array(ob)
    npx = 8192 * 8192
    for w in (5, 7):
        wf, wl = _lib.i64([-(w // 2)] * 2), _lib.i64([w] * 2)
        rate("median%dx%d_8192_n0f8" % (w, w),
             lambda wf=wf, wl=wl: L.ifm_mapwindow_median(ctx, ac, aa, None, wf, wl, brep), npx, 2, 5)
    for w in (3, 9):
        win = _lib.i64([w, w])
        rate("extrema%dx%d_8192_n0f8" % (w, w), lambda win=win: L.ifm_extrema_filter(ctx, ac, aa, abb, win), npx, 3, 5)
    wf, wl = _lib.i64([-4, -4]), _lib.i64([9, 9])
    rate("minimum9x9_8192_n0f8", lambda: L.ifm_mapwindow_minmax(ctx, ac, aa, None, None, wf, wl, brep), npx, 2, 5)
    return ex


if __name__ == "__main__":
    main()
