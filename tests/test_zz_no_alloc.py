"""ifm_imfilter_fir_noalloc (the cascade with its temporaries in a fixed
device array) against ifm_imfilter_fir on the same inputs, through the C ABI:
the same kernels in the same order, so the outputs must be bit-identical.  Two
calls in a row, the second on a smaller input, so a stale first-call
temporary would show."""
import ctypes as C

import numpy as np
import pytest

import imagefiltering as IF
from imagefiltering import _lib
from imagefiltering.device import DeviceArray, ifm_array
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _factors(spec):
    fa = (_lib.IfmFactor * len(spec))()
    keep = []
    for i, (axis, taps, first) in enumerate(spec):
        t = np.ascontiguousarray(taps, dtype=np.float64)
        keep.append(t)
        fa[i].axis, fa[i].len, fa[i].first = axis, len(t), first
        fa[i].taps = t.ctypes.data_as(C.POINTER(C.c_double))
    return fa, keep


def _run(entry, img, spec, style, fill=0.0):
    L = _lib.load()
    d_in = DeviceArray.from_host(img)
    d_out = DeviceArray(img.shape, IF.eltypes.Float32, d_in.order)
    b = _lib.IfmBorder()
    b.style = {"replicate": 0, "circular": 1, "symmetric": 2, "reflect": 3, "fill": 4}[style]
    b.fill = fill
    fa, keep = _factors(spec)
    rc = getattr(L, entry)(_lib.ctx(), ifm_array(d_in), ifm_array(d_out), None, fa, len(spec), b)
    msg = L.ifm_last_error(_lib.ctx()).decode()
    _lib.check(L.ifm_ctx_synchronize(_lib.ctx()))
    return rc, msg, d_out.to_host()


CASES = [
    # three factors whose axis-2 factor the fused 3-D kernel does not take:
    # a fused (0, 1) pass into a temporary, then the axis-2 pass
    ((37, 29, 23), [(0, [0.25, 0.5, 0.25], -1), (1, [1.0, -2.0, 1.0, 0.5], -2), (2, np.full(10, 0.1), -5)],
     "symmetric", 0.0),
    # axes 0 and 2 only: two single-axis passes through a temporary
    ((20, 31, 17), [(0, [0.2, 0.3, 0.5], -1), (2, [0.5, 0.25, 0.125, 0.125, 0.0], -2)], "fill", 0.7),
    # four axes: (0, 1) fused, then axes 2 and 3 through both temporaries
    ((9, 8, 7, 6), [(0, [1.0, 2.0], 0), (1, [0.5, 0.5], -1), (2, [0.25, 0.5, 0.25], -1), (3, [1.0, -1.0], -1)],
     "circular", 0.0),
]


def test_noalloc_matches_fir_twice_second_smaller():
    rng = np.random.default_rng(71)
    for shape, spec, style, fill in CASES:
        # column-major, so the library's axes are the factors' axes and every
        # case goes through cascade temporaries
        big = np.asfortranarray(rng.random(shape).astype(np.float32))
        small = np.asfortranarray(big[tuple(slice(0, max(2, s - 3)) for s in shape)])
        for img in (big, small):
            rc0, m0, ref = _run("ifm_imfilter_fir", img, spec, style, fill)
            rc1, m1, got = _run("ifm_imfilter_fir_noalloc", img, spec, style, fill)
            assert rc0 == 0, m0
            assert rc1 == 0, m1
            np.testing.assert_array_equal(got, ref)
            want = O.imfilter_factors(img.astype(np.float64),
                                      [(ax, np.asarray(t, dtype=np.float64), f) for ax, t, f in spec],
                                      ("fill", fill) if style == "fill" else style)
            assert np.max(np.abs(got - want)) / np.max(np.abs(want)) <= 1e-5


def test_noalloc_rejects_host_arrays_and_oversize_temporaries():
    L = _lib.load()
    img = np.zeros((8, 8, 8), dtype=np.float32)
    out = np.zeros_like(img)
    b = _lib.IfmBorder()
    fa, keep = _factors([(0, [1.0, 1.0], 0), (2, [1.0, 1.0], 0)])
    assert L.ifm_imfilter_fir_noalloc(_lib.ctx(), ifm_array(img), ifm_array(out), None, fa, 2, b) == _lib.E_ARG
    # a temporary of 2^24 floats (64 MiB) is over half the fixed array
    d_in = DeviceArray((256, 256, 256), IF.eltypes.Float32)
    d_out = DeviceArray((256, 256, 256), IF.eltypes.Float32)
    rc = L.ifm_imfilter_fir_noalloc(_lib.ctx(), ifm_array(d_in), ifm_array(d_out), None, fa, 2, b)
    assert rc == 6, L.ifm_last_error(_lib.ctx())
    assert b"no-alloc maximum" in L.ifm_last_error(_lib.ctx())
