"""ctypes front end of the CPU oracle (oracle/ifm_oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.  Arrays follow the same
convention as the product's Python host layer: numpy axis d is Julia
dimension d+1, 0-based indices, results in FP64 (or the input codes for the
order filters).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

STYLES = {"replicate": 0, "circular": 1, "symmetric": 2, "reflect": 3, "fill": 4}


NATIVE_PATH = os.path.join(_HERE, "_native", "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        _LIB = _open(path)
    return _LIB


def use_library(path):
    """Switch to another build of the same source (the -O3 -march=native
    OpenMP build the CPU baseline times); returns its thread count."""
    global _LIB
    _LIB = _open(path)
    return _LIB.orc_max_threads()


def set_threads(n):
    return lib().orc_set_threads(int(n))


def max_threads():
    return lib().orc_max_threads()


def _open(path):
    L = C.CDLL(path)
    i64p = C.POINTER(C.c_int64)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    L.orc_padmap.restype = C.c_int64
    L.orc_padmap.argtypes = [C.c_int, C.c_int64, C.c_int64]
    L.orc_padindex.argtypes = [C.c_int, C.c_int64, C.c_int64, C.c_int64, i64p]
    L.orc_padarray_f64.argtypes = [dp, C.c_int, i64p, C.c_int, C.c_double, i64p, i64p, dp]
    L.orc_imfilter_dense.argtypes = [dp, C.c_int, i64p, dp, i64p, i64p, C.c_int, C.c_double,
                                     i64p, i64p, dp]
    L.orc_imfilter_factors.argtypes = [dp, C.c_int, i64p, C.c_int, ip, ip, i64p, dp, C.c_int,
                                       C.c_double, i64p, i64p, dp]
    L.orc_imfilter_iir.argtypes = [dp, C.c_int, i64p, C.c_int, dp, ip, C.c_int, C.c_double, dp]
    L.orc_median_f64.argtypes = [dp, C.c_int, i64p, i64p, i64p, C.c_int, C.c_double, i64p, i64p, dp]
    L.orc_minmax_f64.argtypes = [dp, C.c_int, i64p, i64p, i64p, C.c_int, C.c_double, i64p, i64p,
                                 dp, dp]
    L.orc_extrema_lemire_f64.argtypes = [dp, C.c_int, i64p, i64p, dp, dp]
    L.orc_imfilter_dense_rows.argtypes = [dp, C.c_int, i64p, dp, i64p, i64p, C.c_int, C.c_double,
                                          C.c_int64, C.c_int64, dp]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_max_threads.argtypes = []
    return L


def _i64(v):
    a = (C.c_int64 * max(1, len(v)))(*[int(x) for x in v])
    return a


def _f64(a):
    a = np.asfortranarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def _outbuf(shape):
    o = np.zeros(shape, dtype=np.float64, order="F")
    return o, o.ctypes.data_as(C.POINTER(C.c_double))


def _style(border):
    if isinstance(border, str):
        return STYLES[border], 0.0
    kind, val = border  # ("fill", value)
    return 4, float(val)


def padindex(style, lo, n, hi):
    out = (C.c_int64 * (lo + n + hi))()
    if lib().orc_padindex(STYLES[style], lo, n, hi, out) != 0:
        raise ZeroDivisionError("reflect padding of a length-1 axis")
    return np.array(list(out), dtype=np.int64)


def padarray(img, border, lo, hi):
    a, ap = _f64(img)
    s, fill = _style(border)
    shape = tuple(lo[d] + a.shape[d] + hi[d] for d in range(a.ndim))
    o, op = _outbuf(shape)
    if lib().orc_padarray_f64(ap, a.ndim, _i64(a.shape), s, fill, _i64(lo), _i64(hi), op) != 0:
        raise ZeroDivisionError("reflect padding of a length-1 axis")
    return o


def _region(a, ofirst, osize):
    of = ofirst if ofirst is not None else [0] * a.ndim
    os_ = osize if osize is not None else list(a.shape)
    return of, os_


def imfilter_dense(img, k, kfirst, border, ofirst=None, osize=None):
    """Dense correlation of img with kernel k whose axis d starts at kfirst[d]
    (a kernel with fewer dimensions extends with unit axes at offset 0)."""
    a, ap = _f64(img)
    k = np.asarray(k, dtype=np.float64)
    k = k.reshape(k.shape + (1,) * (a.ndim - k.ndim))
    kfirst = list(kfirst) + [0] * (a.ndim - len(kfirst))
    kk, kp = _f64(k)
    s, fill = _style(border)
    of, os_ = _region(a, ofirst, osize)
    o, op = _outbuf(tuple(os_))
    if lib().orc_imfilter_dense(ap, a.ndim, _i64(a.shape), kp, _i64(kk.shape), _i64(kfirst), s, fill,
                                _i64(of), _i64(os_), op) != 0:
        raise ZeroDivisionError("reflect padding of a length-1 axis")
    return o


def imfilter_factors(img, factors, border, ofirst=None, osize=None):
    """Cascade of 1-D factors: list of (axis, taps, first)."""
    a, ap = _f64(img)
    s, fill = _style(border)
    nf = len(factors)
    faxis = (C.c_int * max(1, nf))(*[f[0] for f in factors])
    flen = (C.c_int * max(1, nf))(*[len(f[1]) for f in factors])
    ffirst = _i64([f[2] for f in factors])
    taps = np.ascontiguousarray(np.concatenate([np.asarray(f[1], dtype=np.float64) for f in factors])
                                if nf else np.zeros(1))
    of, os_ = _region(a, ofirst, osize)
    o, op = _outbuf(tuple(os_))
    rc = lib().orc_imfilter_factors(ap, a.ndim, _i64(a.shape), nf, faxis, flen, ffirst,
                                    taps.ctypes.data_as(C.POINTER(C.c_double)), s, fill,
                                    _i64(of), _i64(os_), op)
    if rc != 0:
        raise ZeroDivisionError("reflect padding of a length-1 axis")
    return o


def imfilter_iir(img, kernels, border):
    """kernels: list of (axis, a[3], b[3], scale, M 3x3)."""
    a, ap = _f64(img)
    nk = len(kernels)
    coef = np.zeros((max(1, nk), 16))
    for i, (ax, ka, kb, sc, M) in enumerate(kernels):
        coef[i, 0:3] = ka
        coef[i, 3:6] = kb
        coef[i, 6] = sc
        coef[i, 7:16] = np.asarray(M, dtype=np.float64).reshape(9)
    coef = np.ascontiguousarray(coef)
    axes = (C.c_int * max(1, nk))(*[k[0] for k in kernels])
    isfill = 0
    fill = 0.0
    if not isinstance(border, str):
        isfill, fill = 1, float(border[1])
    o, op = _outbuf(a.shape)
    rc = lib().orc_imfilter_iir(ap, a.ndim, _i64(a.shape), nk, coef.ctypes.data_as(C.POINTER(C.c_double)),
                                axes, isfill, fill, op)
    if rc != 0:
        raise ValueError("IIR axis too short")
    return o


def median(img, wfirst, wlen, border, ofirst=None, osize=None):
    a, ap = _f64(img)
    s, fill = _style(border)
    of, os_ = _region(a, ofirst, osize)
    o, op = _outbuf(tuple(os_))
    if lib().orc_median_f64(ap, a.ndim, _i64(a.shape), _i64(wfirst), _i64(wlen), s, fill, _i64(of),
                            _i64(os_), op) != 0:
        raise ZeroDivisionError("edge window with an empty or zero-period index map")
    return o


def minmax(img, wfirst, wlen, border, ofirst=None, osize=None):
    a, ap = _f64(img)
    s, fill = _style(border)
    of, os_ = _region(a, ofirst, osize)
    omin, pmin = _outbuf(tuple(os_))
    omax, pmax = _outbuf(tuple(os_))
    if lib().orc_minmax_f64(ap, a.ndim, _i64(a.shape), _i64(wfirst), _i64(wlen), s, fill, _i64(of),
                            _i64(os_), pmin, pmax) != 0:
        raise ZeroDivisionError("edge window with an empty or zero-period index map")
    return omin, omax


def extrema_lemire(img, window):
    a, ap = _f64(img)
    omin, pmin = _outbuf(a.shape)
    omax, pmax = _outbuf(a.shape)
    lib().orc_extrema_lemire_f64(ap, a.ndim, _i64(a.shape), _i64(window), pmin, pmax)
    return omin, omax


def imfilter_dense_rows(img, k, kfirst, border, row0, nrows):
    """Dense FIR over rows row0..row0+nrows-1 of axis 1 (CPU-baseline timing)."""
    a, ap = _f64(img)
    kk, kp = _f64(k)
    s, fill = _style(border)
    shape = list(a.shape)
    shape[1] = nrows
    o, op = _outbuf(tuple(shape))
    lib().orc_imfilter_dense_rows(ap, a.ndim, _i64(a.shape), kp, _i64(kk.shape), _i64(kfirst), s, fill,
                                  row0, nrows, op)
    return o


def _nextprod23(n):
    best = None
    p2 = 1
    while p2 < 2 * n + 2:
        p = p2
        while p < 2 * n + 2:
            if p >= n and (best is None or p < best):
                best = p
            p *= 3
        p2 *= 2
    return best


def imfilter_fft(img, k, kfirst, border):
    """Algorithm.FFT() restated with numpy's FFT in FP64 (src/imfilter.jl:796-866,
    src/border.jl:240-248): pad every axis with the border to
    nextprod([2,3], n + K) (the padding shifted by the kernel's first tap),
    wrap the kernel into an FFTView of the padded size, and read back
    irfft(rfft(Apad) .* conj(rfft(kpad))) at the image axes."""
    a = np.asarray(img, dtype=np.float64)
    k = np.asarray(k, dtype=np.float64).reshape(np.shape(k) + (1,) * (a.ndim - np.ndim(k)))
    kfirst = list(kfirst) + [0] * (a.ndim - len(kfirst))
    lo, hi, N = [], [], []
    for d in range(a.ndim):
        n, K, f = a.shape[d], k.shape[d], kfirst[d]
        # padded index u holds image index u - lo; output c reads
        # u = c + lo + f .. c + lo + f + K - 1, which must stay below N
        l = max(-f, 0)
        Nd = _nextprod23(max(n + K, n + max(f, 0) + K - 1)) if (K > 1 or f != 0) else n
        lo.append(l)
        hi.append(Nd - n - l)
        N.append(Nd)
    pad = padarray(a, border, lo, hi)
    kp = np.zeros(N)
    for idx in np.ndindex(*k.shape):
        u = tuple((idx[d] + kfirst[d]) % N[d] for d in range(a.ndim))
        kp[u] += k[idx]
    ax = list(range(a.ndim))
    R = np.fft.irfftn(np.fft.rfftn(pad, axes=ax) * np.conj(np.fft.rfftn(kp, axes=ax)), s=N, axes=ax)
    sl = tuple(slice(lo[d], lo[d] + a.shape[d]) for d in range(a.ndim))
    return np.asfortranarray(R[sl])


def imfilter_na(img, stage_fn):
    """NA() border restated from imfilter_na_inseparable! (src/imfilter.jl:
    1109-1122): stage_fn(x, border) filters x with the kernel under a border;
    the image with NaN set to zero and the mask of valid pixels are filtered
    with Fill(0) and divided."""
    a = np.asarray(img, dtype=np.float64)
    nan = np.isnan(a)
    num = stage_fn(np.where(nan, 0.0, a), ("fill", 0.0))
    den = stage_fn((~nan).astype(np.float64), ("fill", 0.0))
    with np.errstate(divide="ignore", invalid="ignore"):
        return num / den


def imfilter_mixed(img, stages, border):
    """A cascade of FIR and IIR stages as the reference's generic cascade runs
    it (src/imfilter.jl:372-457): pad once with the accumulated padding,
    doubled when IIR and FIR stages are mixed (src/border.jl:614-619); each
    stage filters the previous region shrunk by its own axes (IIR: 0:0), an
    IIR stage over its whole region with its edge initialised from that
    region's ends (or the Fill value); the last stage is cut to the image.
    stages: ("fir", axis, taps, first) | ("iir", axis, a, b, scale, M)."""
    a = np.asarray(img, dtype=np.float64)
    N = a.ndim
    ax = []
    for st in stages:
        t = [(0, 0)] * N
        if st[0] == "fir":
            t[st[1]] = (st[3], st[3] + len(st[2]) - 1)
        ax.append(t)
    pf = [sum(t[d][0] for t in ax) for d in range(N)]
    pl = [sum(t[d][1] for t in ax) for d in range(N)]
    if any(s[0] == "iir" for s in stages) and any(s[0] == "fir" for s in stages):
        pf = [2 * f if f < 0 else f for f in pf]
        pl = [2 * l if l > 0 else l for l in pl]
    lo = [max(-f, 0) for f in pf]
    hi = [max(l, 0) for l in pl]
    A = padarray(a, border, lo, hi)
    cov = [(-lo[d], a.shape[d] + hi[d] - 1) for d in range(N)]
    inds = [(0, a.shape[d] - 1) for d in range(N)]
    region = [(inds[d][0] + pf[d], inds[d][1] + pl[d]) for d in range(N)]
    for s, st in enumerate(stages):
        last = s == len(stages) - 1
        region = [(region[d][0] - ax[s][d][0], region[d][1] - ax[s][d][1]) for d in range(N)]
        if st[0] == "iir":
            sl = tuple(slice(region[d][0] - cov[d][0], region[d][1] - cov[d][0] + 1) for d in range(N))
            A = imfilter_iir(np.asfortranarray(A[sl]), [st[1:]], border if not isinstance(border, str) else "replicate")
            cov = region
            continue
        target = [(max(region[d][0], inds[d][0]), min(region[d][1], inds[d][1])) for d in range(N)] if last else region
        A = imfilter_factors(A, [(st[1], np.asarray(st[2], dtype=np.float64), st[3])], "replicate",
                             [target[d][0] - cov[d][0] for d in range(N)],
                             [target[d][1] - target[d][0] + 1 for d in range(N)])
        cov = target
    if stages and stages[-1][0] == "iir":
        sl = tuple(slice(inds[d][0] - cov[d][0], inds[d][1] - cov[d][0] + 1) for d in range(N))
        A = A[sl]
    return np.asfortranarray(A)


def _isless_key(x):
    """Sort key of Julia's isless on floats: NaN last, -0.0 before +0.0."""
    x = float(x)
    if np.isnan(x):
        return (2, 0.0)
    return (0 if (x < 0 or (x == 0 and np.signbit(x))) else 1, x)


def mapwindow_generic(img, fname, wfirst, wlen, border, first, step, shape, p=None):
    """mapwindow(f, img, window; border, indices) restated in numpy for the
    window functions sum, mean, minimum, maximum, first, median! (median_fast!),
    median (Statistics) and quantile(p) (src/mapwindow.jl:75-121, :270-333):
    f at the window centres first + i * step (0-based), edge windows mapped
    relative to their intersection with the image (copy_win! with
    padindex(border, inner, outer), :300-325), Fill substituting the value.
    Results in FP64 (selected elements keep their values)."""
    a = np.asarray(img, dtype=np.float64)
    N = a.ndim
    style, fill = (border, None) if isinstance(border, str) else ("fill", float(border[1]))
    out = np.empty(tuple(shape), dtype=np.float64)
    for oi in np.ndindex(*shape):
        axes_idx, axes_fill = [], []
        for d in range(N):
            q = first[d] + oi[d] * step[d]
            lo, hi = q + wfirst[d], q + wfirst[d] + wlen[d] - 1
            pos = np.arange(lo, hi + 1)
            if style == "fill":
                ok = (pos >= 0) & (pos < a.shape[d])
                axes_idx.append(np.where(ok, pos, 0))
                axes_fill.append(~ok)
            else:
                a0, b0 = max(lo, 0), min(hi, a.shape[d] - 1)
                idx = padindex(style, a0 - lo, b0 - a0 + 1, hi - b0) + a0
                axes_idx.append(idx)
                axes_fill.append(np.zeros(len(pos), bool))
        w = a[np.ix_(*axes_idx)]
        if style == "fill":
            mask = np.zeros(w.shape, bool)
            for d in range(N):
                sh = [1] * N
                sh[d] = -1
                mask |= axes_fill[d].reshape(sh)
            w = np.where(mask, fill, w)
        v = w.ravel(order="F")  # column-major window order, as the Julia buffer
        if fname == "sum":
            r = v.sum()
        elif fname == "mean":
            r = v.mean()
        elif fname == "minimum":
            r = np.nan if np.isnan(v).any() else min(v, key=_isless_key)
        elif fname == "maximum":
            r = np.nan if np.isnan(v).any() else max(v, key=_isless_key)
        elif fname == "first":
            r = v[0]
        elif fname == "median!":
            r = sorted(v, key=_isless_key)[(len(v) + 1) // 2 - 1]
        elif fname == "median":
            if np.isnan(v).any():
                r = np.nan
            else:
                s = np.sort(v)
                n = len(s)
                r = s[n // 2] if n % 2 else s[n // 2 - 1] / 2 + s[n // 2] / 2
        elif fname == "quantile":
            if np.isnan(v).any():
                raise ValueError("quantiles are undefined in presence of NaNs")
            r = np.quantile(v, p)
        else:
            raise ValueError(fname)
        out[oi] = r
    return out
