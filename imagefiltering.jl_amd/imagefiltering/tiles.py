"""Tile-shape autotuner for the fused two-factor separable kernel.

The kernel behind imfilter with a pair of factors on axes 1 and 2 (and the
separable op of the batched serving form) computes 64 x 32 or 64 x 64 output
tiles; the library's tile rules (ifm_set_tile_rules) pick one per call.  A
taller tile stages fewer halo rows per output (the axis-0 factor runs over
TY + K - 1 staged rows for TY outputs), a shorter one makes twice as many
workgroups, which counts when an image has too few tiles to fill the 256
compute units.  Which wins depends on the kernel length, the image size and
the input type, so `tune` times both shapes on synthetic images for every
(input dtype, kernel-length bucket, image-size bucket), keeps the faster and
installs the result as rules.  `save` / `load` keep them as JSON, and the
process-wide context installs the cached rules when it is created, so a
machine tunes once.  Results do not depend on the tile shape: every output
sums the same taps in the same order.

ImageFiltering's own tiled FIR sizes its tiles from a cache heuristic
(src/imfilter.jl:1298-1312); this table is the device counterpart.
"""
import ctypes as C
import json
import os

import numpy as np

from . import _lib

TILE_HEIGHTS = (32, 64)
# (kmax, representative kernel length): the longer of the two factors.
KBUCKETS = ((9, 7), (17, 13), (33, 29), (64, 49))
# (max output pixels, representative side): square images of that side.
SIZEBUCKETS = ((384 ** 2, 256), (1536 ** 2, 1024), (3072 ** 2, 2048), (1 << 62, 4096))
DTYPES = ("Float32", "BFloat16")
_DT_CODE = {"Float32": 2, "Float64": 3, "BFloat16": 4, "Float16": 5, "N0f8": 1, "UInt8": 0}


def cache_path():
    """The JSON cache: $IFM_TILE_CACHE, else ~/.cache/imagefiltering_amd/tiles.json."""
    p = os.environ.get("IFM_TILE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "imagefiltering_amd", "tiles.json")


def rules_from_choices(choices):
    """Rules for a dict {(dtype name, kmax, max_pixels): tile_h}.  Rules are
    ordered by (dtype, kmax, max_pixels) so that the first rule a call matches
    is the one of its own buckets."""
    out = []
    for (dt, kmax, maxpx), th in sorted(choices.items(), key=lambda kv: (_DT_CODE[kv[0][0]], kv[0][1], kv[0][2])):
        if th not in TILE_HEIGHTS:
            raise ValueError("tile height must be 32 or 64, got %r" % (th,))
        out.append({"op": _lib.TILE_SEP2D, "dtype": _DT_CODE[dt], "kmax": int(kmax), "max_pixels": int(maxpx),
                    "tile_h": int(th)})
    if len(out) > _lib.MAX_TILE_RULES:
        raise ValueError("too many tile rules (%d)" % len(out))
    return out


def install(rules):
    """Replace the process-wide context's tile rules (a list of dicts as made
    by rules_from_choices; [] restores the default 64 x 32 tiles)."""
    L, ctx = _lib.load(), _lib.ctx()
    arr = (_lib.IfmTileRule * max(len(rules), 1))()
    for i, r in enumerate(rules):
        arr[i].op, arr[i].dtype, arr[i].kmax = r["op"], r["dtype"], r["kmax"]
        arr[i].tile_h, arr[i].max_pixels = r["tile_h"], r["max_pixels"]
    _lib.check(L.ifm_set_tile_rules(ctx, arr, len(rules)))


def tile_height(dtype, kmax, pixels):
    """The tile height a call with this input dtype, longest factor and output
    plane size would use under the installed rules."""
    L, ctx = _lib.load(), _lib.ctx()
    th = C.c_int()
    _lib.check(L.ifm_tile_height(ctx, _lib.TILE_SEP2D, _DT_CODE[dtype], int(kmax), int(pixels), C.byref(th)))
    return th.value


def save(result, path=None):
    path = path or cache_path()
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(result, f, indent=1)
    return path


def load(path=None):
    """Install the rules of a saved tuning result; returns it (None if there
    is no cache file, or if it cannot be read: the default tiles then stay,
    with a warning)."""
    path = path or cache_path()
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            result = json.load(f)
        install(result["rules"])
    except Exception as e:  # a stale or damaged cache must not break filtering
        import warnings
        warnings.warn("ignoring the tile cache %s: %s" % (path, e))
        return None
    return result


def _factors(k):
    """A normalised box pair of k taps each, centred (the taps' values do not
    change the kernel's time)."""
    t = np.full(k, 1.0 / k)
    fa = (_lib.IfmFactor * 2)()
    for axis in range(2):
        fa[axis].axis, fa[axis].len, fa[axis].first = axis, k, -(k // 2)
        fa[axis].taps = t.ctypes.data_as(C.POINTER(C.c_double))
    return fa, t


def tune(dtypes=DTYPES, kbuckets=KBUCKETS, sizebuckets=SIZEBUCKETS, rounds=3, target_ms=4.0):
    """Time every tile height on every bucket; install and return
    {"rules": [...], "trials": [...]}.  Each trial runs the fused kernel on a
    square image of the bucket's side with a kernel of the bucket's length,
    'reflect' border, F64 output (the eltype imfilter gives F32 or BF16 pixels
    with F64 taps); the time per call is the best of `rounds` runs of enough
    calls to fill about target_ms."""
    from .device import DeviceArray, ifm_array
    from .eltypes import BFloat16, Float32, Float64
    from .imfilter import _border_struct
    from .border import Pad
    et = {"Float32": Float32, "BFloat16": BFloat16}
    L, ctx = _lib.load(), _lib.ctx()
    trials, choices = [], {}
    try:
        for dt in dtypes:
            for maxpx, side in sizebuckets:
                img = DeviceArray((side, side), et[dt])
                _lib.check(L.ifm_fill_uniform(ctx, ifm_array(img), 1))
                out = img.similar(Float64)
                ai, ao = ifm_array(img), ifm_array(out)
                bord = _border_struct(Pad("reflect"), et[dt])
                for kmax, k in kbuckets:
                    fa, keep = _factors(k)
                    best = {}
                    for th in TILE_HEIGHTS:
                        install(rules_from_choices({(dt, 64, 1 << 62): th}))
                        _lib.check(L.ifm_imfilter_fir(ctx, ai, ao, None, fa, 2, bord))
                        _lib.check(L.ifm_ctx_synchronize(ctx))
                        # calls per run: about target_ms of work, from one timed call
                        ms = C.c_float()
                        L.ifm_timer_start(ctx)
                        L.ifm_imfilter_fir(ctx, ai, ao, None, fa, 2, bord)
                        _lib.check(L.ifm_timer_stop(ctx, C.byref(ms)))
                        reps = int(min(500, max(3, target_ms / max(ms.value, 1e-3))))
                        t = []
                        for _ in range(rounds):
                            L.ifm_timer_start(ctx)
                            for _ in range(reps):
                                L.ifm_imfilter_fir(ctx, ai, ao, None, fa, 2, bord)
                            _lib.check(L.ifm_timer_stop(ctx, C.byref(ms)))
                            t.append(ms.value / reps)
                        best[th] = min(t)
                        trials.append({"dtype": dt, "side": side, "k": k, "tile_h": th, "us": round(best[th] * 1e3, 3)})
                    choices[(dt, kmax, maxpx)] = min(best, key=best.get)
                    del keep
                del img, out
    except Exception:
        install([])
        raise
    rules = rules_from_choices(choices)
    install(rules)
    return {"rules": rules, "trials": trials}


if __name__ == "__main__":
    res = tune()
    print("wrote", save(res))
    for r in res["trials"]:
        print(r)
