"""Batched serving of the filtering path.

``request_mix`` draws the serving workload: 2-D images with log-uniform
sides, Float32 or BFloat16 pixels, one of four ops -- a Gaussian with
sigma ~ U[0.5, 8] (``imfilter(img, KernelFactors.gaussian((s, s)), border)``),
the sobel gradient pair (``imgradients(img, KernelFactors.sobel, border)``,
src/specialty.jl:39-53), a 7x7 median (``mapwindow(median!, img, (7, 7))``)
or a 15x15 dense kernel of N(0,1) taps normalised to sum 1 -- under one of
the five borders, arriving as a Poisson process.

``serve`` runs such a stream through libifm's device-resident request loop
(``ifm_serve``): mode "batched" admits arrived requests while their bytes fit
the HBM budget and launches them together, without waiting for the batches
already on the GPU; mode "percall" runs each request alone through the
single-image entry points, the plain drop-in dispatch.  Inputs are generated
on the device (seeded uniform [0,1)), so both modes do identical work.

Output eltypes follow filter_type (src/imfilter.jl:1131-1154): the linear
ops give Float64 (Float64 taps), the median the input eltype.
"""
import ctypes as C
import math

import numpy as np

from . import _lib
from .eltypes import BFloat16, Float32, Float64
from .kernelfactors import gaussian as _gaussian
from .kernelfactors import sobel as _sobel

OPS = ("gaussian", "sobel", "median7", "dense15")
BORDERS = ("replicate", "reflect", "symmetric", "circular", "fill0")
_STYLE = {"replicate": _lib.REPLICATE, "reflect": _lib.REFLECT, "symmetric": _lib.SYMMETRIC,
          "circular": _lib.CIRCULAR, "fill0": _lib.FILL}


def dense15(seed):
    """15 x 15 taps, N(0,1) normalised to sum 1 (redrawn while |sum| < 1)."""
    rng = np.random.default_rng(seed)
    while True:
        k = rng.standard_normal((15, 15))
        s = k.sum()
        if abs(s) >= 1:
            return k / s


def request_mix(n, rate=None, seed=20261207, side=(64, 4096)):
    """The serving stream: a dict of per-request arrays (seeded)."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(side[0]), math.log(side[1])
    w = np.rint(np.exp(rng.uniform(lo, hi, n))).astype(np.int64)
    h = np.rint(np.exp(rng.uniform(lo, hi, n))).astype(np.int64)
    dtype = rng.integers(0, 2, n)             # 0 Float32, 1 BFloat16
    op = rng.integers(0, len(OPS), n)
    border = rng.integers(0, len(BORDERS), n)
    sigma = rng.uniform(0.5, 8.0, n)
    kseed = rng.integers(0, 2 ** 62, n)
    arrival = np.cumsum(rng.exponential(1.0 / rate, n)) if rate else np.zeros(n)
    return {"n": n, "w": w, "h": h, "dtype": dtype, "op": op, "border": border, "sigma": sigma,
            "kseed": kseed, "arrival": arrival}


def _array(shape, eltype, data=None):
    a = _lib.IfmArray()
    a.ndim = 2
    a.dtype = eltype.code
    a.size[0], a.size[1] = shape
    a.stride[0], a.stride[1] = 1, shape[0]
    a.data = data
    a.device = 1
    return a


def _factors(dst, comps, keep):
    for j, f in enumerate(comps):
        t = np.ascontiguousarray(np.asarray(f.data.data, dtype=np.float64).ravel())
        keep.append(t)
        dst[j].axis, dst[j].len, dst[j].first = f.Npre, len(t), int(f.data.first[0])
        dst[j].taps = t.ctypes.data_as(C.POINTER(C.c_double))
    return len(comps)


def requests(mix, buffers=None):
    """ifm_request array of a mix.  buffers: optional list of (in, out, out2)
    device pointers per request (None: the serving loop allocates and
    generates).  Returns (array, keepalive)."""
    n = mix["n"]
    arr = (_lib.IfmRequest * n)()
    keep = []
    for i in range(n):
        q = arr[i]
        shape = (int(mix["w"][i]), int(mix["h"][i]))
        et = BFloat16 if mix["dtype"][i] else Float32
        op = OPS[mix["op"][i]]
        b = BORDERS[mix["border"][i]]
        q.border.style = _STYLE[b]
        q.border.fill = 0.0
        bufs = buffers[i] if buffers is not None else (None, None, None)
        q.in_ = _array(shape, et, bufs[0])
        if op == "median7":
            q.op = _lib.OP_MEDIAN
            q.window = 7
            q.out = _array(shape, et, bufs[1])
        elif op == "dense15":
            q.op = _lib.OP_DENSE
            k = np.asfortranarray(dense15(int(mix["kseed"][i])))
            flat = np.ascontiguousarray(k.ravel(order="F"))
            keep.append(flat)
            q.kernel.ndim = 2
            q.kernel.size[0], q.kernel.size[1] = 15, 15
            q.kernel.first[0], q.kernel.first[1] = -7, -7
            q.kernel.taps = flat.ctypes.data_as(C.POINTER(C.c_double))
            q.out = _array(shape, Float64, bufs[1])
        elif op == "gaussian":
            q.op = _lib.OP_SEPARABLE
            s = float(mix["sigma"][i])
            q.nfactors = _factors(q.factors, _gaussian((s, s)), keep)
            q.out = _array(shape, Float64, bufs[1])
        else:  # sobel gradient pair
            q.op = _lib.OP_SEPARABLE
            q.nfactors = _factors(q.factors, _sobel((True, True), 0), keep)
            q.nfactors2 = _factors(q.factors2, _sobel((True, True), 1), keep)
            q.out = _array(shape, Float64, bufs[1])
            q.out2 = _array(shape, Float64, bufs[2])
    return arr, keep


def request_pixels(mix):
    return mix["w"].astype(np.int64) * mix["h"].astype(np.int64)


def request_bytes(mix):
    """Compulsory HBM bytes per request: input read + outputs written."""
    px = request_pixels(mix)
    ein = np.where(mix["dtype"] == 1, 2, 4)
    op = mix["op"]
    eout = np.where(op == OPS.index("median7"), ein, 8)
    nout = np.where(op == OPS.index("sobel"), 2, 1)
    return px * (ein + nout * eout)


def arena_bytes(mix):
    """Arena bytes a generated request holds in ``serve``: its input and
    outputs, each rounded up to the arena's 256-byte granule."""
    px = request_pixels(mix)
    ein = np.where(mix["dtype"] == 1, 2, 4)
    op = mix["op"]
    eout = np.where(op == OPS.index("median7"), ein, 8)
    nout = np.where(op == OPS.index("sobel"), 2, 1)
    g = lambda b: (b + 255) // 256 * 256
    return g(px * ein) + nout * g(px * eout)


def serve(mix, mode="batched", budget_bytes=64 << 30, max_batch=1024, max_inflight=4, buffers=None):
    """Run a mix through ifm_serve; returns per-request admit and done times (s)."""
    L, ctx = _lib.load(), _lib.ctx()
    arr, keep = requests(mix, buffers)
    n = mix["n"]
    arrival = np.ascontiguousarray(mix["arrival"], dtype=np.float64)
    admit = np.zeros(n)
    done = np.zeros(n)
    cfg = _lib.IfmServeConfig()
    cfg.budget_bytes = int(budget_bytes)
    cfg.mode = 0 if mode == "batched" else 1
    cfg.max_batch = max_batch
    cfg.max_inflight = max_inflight
    dp = C.POINTER(C.c_double)
    _lib.check(L.ifm_serve(ctx, arr, arrival.ctypes.data_as(dp), n, C.byref(cfg), admit.ctypes.data_as(dp),
                           done.ctypes.data_as(dp)))
    return admit, done


class Server:
    """Non-blocking request server (ifm_server_*): ``submit`` queues one
    request (an ``IfmRequest`` over device arrays) and returns its ticket
    without waiting for the GPU; requests submitted while earlier batches run
    join the next batch.  ``poll(ticket)`` reports completion without
    waiting, ``wait(ticket)`` blocks (ticket 0: everything submitted)."""

    def __init__(self, budget_bytes=0, max_batch=1024, max_inflight=2):
        self._L, ctx = _lib.load(), _lib.ctx()
        cfg = _lib.IfmServerConfig()
        cfg.budget_bytes = int(budget_bytes)
        cfg.max_batch = int(max_batch)
        cfg.max_inflight = int(max_inflight)
        self._h = C.c_void_p()
        _lib.check(self._L.ifm_server_create(ctx, C.byref(cfg), C.byref(self._h)))

    def submit(self, req):
        t = C.c_uint64()
        _lib.check(self._L.ifm_server_submit(self._h, C.byref(req), C.byref(t)))
        return t.value

    def poll(self, ticket):
        d = C.c_int(0)
        _lib.check(self._L.ifm_server_poll(self._h, int(ticket), C.byref(d)))
        return bool(d.value)

    def wait(self, ticket=0):
        _lib.check(self._L.ifm_server_wait(self._h, int(ticket)))

    def close(self):
        if self._h:
            rc = self._L.ifm_server_destroy(self._h)
            self._h = C.c_void_p()
            _lib.check(rc)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def stats(mix, done):
    """images/s, Gpixel/s, HBM GB/s and latency percentiles (arrival to done)."""
    t_end = float(np.max(done)) if len(done) else 0.0
    lat = done - mix["arrival"]
    n = mix["n"]
    return {"requests": int(n), "seconds": t_end, "images_per_s": n / t_end if t_end > 0 else None,
            "gpix_per_s": float(request_pixels(mix).sum()) / t_end / 1e9 if t_end > 0 else None,
            "hbm_gbs": float(request_bytes(mix).sum()) / t_end / 1e9 if t_end > 0 else None,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3)}


def autotune(mix, budget_bytes=64 << 30, sample=3000, batches=(256, 512, 1024, 2048), inflight=(1, 2, 4, 7),
             seed=20261208, slack=0.02):
    """Schedule autotuner for the request loop: runs closed-loop trials of a
    seeded sample of the stream (requests drawn from `mix`, all arriving at
    once) for every (max_batch, max_inflight) pair and returns, among the pairs
    within `slack` of the highest images/s, the one with the fewest batches in
    flight and then the smallest batches (shallower queues wait less), with the
    trial table.  Each trial reuses the same sample, so the pairs are compared
    on identical work.  max_inflight is capped at 7 by the request loop (eight
    descriptor slots, one kept free for the batch being built)."""
    n = mix["n"]
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(n, size=min(sample, n), replace=False))
    sub = {k: (v[idx] if isinstance(v, np.ndarray) else v) for k, v in mix.items()}
    sub["n"] = len(idx)
    sub["arrival"] = np.zeros(len(idx))
    serve(sub, "batched", budget_bytes=budget_bytes)  # warm-up (allocations, code objects)
    table = []
    for mb in batches:
        for mi in inflight:
            _, done = serve(sub, "batched", budget_bytes=budget_bytes, max_batch=mb, max_inflight=mi)
            table.append({"max_batch": mb, "max_inflight": mi, "images_per_s": stats(sub, done)["images_per_s"]})
    top = max(r["images_per_s"] or 0.0 for r in table)
    near = [r for r in table if (r["images_per_s"] or 0.0) >= (1.0 - slack) * top]
    best = min(near, key=lambda r: (r["max_inflight"], r["max_batch"]))
    return {"max_batch": best["max_batch"], "max_inflight": best["max_inflight"]}, table
