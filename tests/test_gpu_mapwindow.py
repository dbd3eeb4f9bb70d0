"""mapwindow with the window-function catalogue (sum, mean, minimum, maximum,
first, median!, Statistics.median, quantile(p)) and strided `indices`, on the
GPU (ifm_mapwindow_generic and the fast order-filter kernels), against the
reference's own tests (test/mapwindow.jl:125-175) and the numpy restatement
oracle.mapwindow_generic.  Order statistics, first, minimum and maximum are
bit-exact; sums and means within 1e-12 (FP64) or 1e-6 (FP32) relative."""
import numpy as np
import pytest

import imagefiltering as IF
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def obord(b):
    return ("fill", float(b.value)) if isinstance(b, IF.Fill) else b


def test_reference_strided_groundtruth():
    """test/mapwindow.jl:135-151: f at strided indices equals the full
    result at those indices, for mapwindow and mapwindow!."""
    rng = np.random.default_rng(130)
    cases = [(IF.mean, rng.standard_normal(10), (1,), (range(0, 10, 2),)),
             (IF.median_, rng.standard_normal(10), (range(-1, 2),), (range(0, 8, 2),)),
             (IF.mean, rng.standard_normal(10), (range(-1, 2),), (range(0, 8, 2),)),
             (IF.mean, rng.standard_normal((10, 5)), (range(-1, 2), range(0, 1)), (range(0, 8, 2), range(0, 3))),
             (IF.mean, rng.standard_normal((10, 5)), (range(-1, 2), range(0, 1)), (range(0, 2), range(0, 3)))]
    for f, img, window, inds in cases:
        full = np.asarray(IF.mapwindow(f, img, window))
        expected = full[tuple(slice(r.start, r.stop, r.step) for r in inds)]
        got = IF.mapwindow(f, img, window, indices=inds)
        assert isinstance(got, np.ndarray) and not isinstance(got, IF.OffsetArray)
        np.testing.assert_array_equal(got, expected, err_msg=str((f, window, inds)))
        out = np.empty_like(expected)
        IF.mapwindow_(f, out, img, window, indices=inds)
        np.testing.assert_array_equal(out, expected)
    # output axes: the lengths of the index ranges (test/mapwindow.jl:152-157)
    x = rng.standard_normal(10)
    assert IF.mapwindow(IF.mean, x, (3,), border="replicate", indices=range(1, 7, 2)).shape == (3,)
    assert IF.mapwindow(IF.mean, x, (3,), indices=range(1, 7)).shape == (6,)
    assert IF.mapwindow(IF.mean, x, (3,)).shape == (10,)


def test_reference_inner_first_and_shifted_minimum():
    """test/mapwindow.jl:159-169."""
    img48 = 10.0 * np.arange(1, 11)
    np.testing.assert_array_equal(IF.mapwindow(IF.first, img48, (1,), border=IF.Inner()).data, img48)
    res = IF.mapwindow(IF.first, img48, (range(0, 2),), border=IF.Inner())
    assert res.first == (0,) and res.data.shape == (9,)
    np.testing.assert_array_equal(res.data, img48[:9])
    got = IF.mapwindow(IF.first, img48, (range(0, 3),), border=IF.Inner(), indices=range(1, 8, 2))
    np.testing.assert_array_equal(got, img48[1:8:2])
    shifted = IF.mapwindow(IF.minimum, img48, range(-2, 1), border=IF.Inner())
    assert shifted.first == (2,)
    np.testing.assert_array_equal(shifted.data, img48[0:8])


def test_reference_window_desugaring_and_median_vectors():
    """test/mapwindow.jl:105-124 and :170-178 (median over every window spelling)."""
    import json, os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))
    for c in g["median"]["cases"]:
        x = np.asarray(c["input"], dtype=np.float64)
        w = c["window"]
        window = range(w[0], w[1] + 1) if x.ndim == 1 else tuple(w)
        for f in (IF.median, IF.median_):
            np.testing.assert_array_equal(IF.mapwindow(f, x, window), np.asarray(c["expect"], dtype=np.float64))
    img58 = np.random.default_rng(58).random(10)
    truth = IF.mapwindow(IF.median, img58, (range(-1, 2),))
    for window in [3, (3,), [3], [range(-1, 2)], range(-1, 2)]:
        np.testing.assert_array_equal(IF.mapwindow(IF.median, img58, window), truth)


@pytest.mark.parametrize("border", ["replicate", "circular", "symmetric", "reflect", IF.Fill(0.25)])
def test_catalogue_against_oracle(border):
    """Every window function at strided centres of 2-D and 3-D images,
    windows with offsets, against oracle.mapwindow_generic."""
    rng = np.random.default_rng(140)
    cases = [((23, 17), (range(-1, 2), range(-2, 1)), (range(0, 23, 3), range(1, 17, 2))),
             ((12, 9, 5), (range(0, 2), range(-1, 2), range(-1, 1)), (range(1, 12, 4), range(0, 9), range(0, 5, 2)))]
    fns = [("sum", IF.sum_), ("mean", IF.mean), ("minimum", IF.minimum), ("maximum", IF.maximum),
           ("first", IF.first), ("median!", IF.median_), ("median", IF.median), ("quantile", IF.quantile(0.3))]
    for shape, window, inds in cases:
        x = rng.standard_normal(shape)
        wf = [w.start for w in window]
        wl = [len(w) for w in window]
        first = [r.start for r in inds]
        step = [r.step for r in inds]
        oshape = [len(r) for r in inds]
        for name, f in fns:
            try:
                want = O.mapwindow_generic(x, name, wf, wl, obord(border), first, step, oshape,
                                           p=getattr(f, "p", None))
            except ZeroDivisionError:
                # an edge window meeting one image element under "reflect"
                # (zero period: the reference's DivideError)
                with pytest.raises(IF.ArgumentError):
                    IF.mapwindow(f, x, window, border=border, indices=inds)
                continue
            got = IF.mapwindow(f, x, window, border=border, indices=inds)
            if name in ("sum", "mean", "quantile", "median"):
                np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12, err_msg=str((name, shape, border)))
            else:
                np.testing.assert_array_equal(got, want, err_msg=str((name, shape, border)))


def test_even_windows_float32_uint8_and_nan():
    """Statistics.median over even windows (the middle of the two central
    values, Float32 arithmetic for Float32), mean / sum / median / quantile
    eltypes of Float32 and UInt8 images, and NaN: median gives NaN,
    median! orders it last, minimum / maximum propagate it, quantile raises."""
    rng = np.random.default_rng(141)
    x32 = rng.standard_normal((30, 21)).astype(np.float32)
    w = (range(-1, 1), range(0, 2))  # 2 x 2
    got = IF.mapwindow(IF.median, x32, w)
    assert got.dtype == np.float32
    want = O.mapwindow_generic(x32, "median", [-1, 0], [2, 2], "replicate", [0, 0], [1, 1], list(x32.shape))
    # the middle x/2 + y/2 in Float32: within one FP32 rounding of the FP64 middle
    np.testing.assert_allclose(got, want, rtol=2 ** -23, atol=0)
    for name, f in [("mean", IF.mean), ("sum", IF.sum_)]:
        g32 = IF.mapwindow(f, x32, (3, 3))
        assert g32.dtype == np.float32
        want = O.mapwindow_generic(x32, name, [-1, -1], [3, 3], "replicate", [0, 0], [1, 1], list(x32.shape))
        np.testing.assert_allclose(g32, want, rtol=1e-6, atol=1e-6)
    u8 = rng.integers(0, 256, (19, 13)).astype(np.uint8)
    gm = IF.mapwindow(IF.mean, u8, (3, 5))
    assert gm.dtype == np.float64
    np.testing.assert_allclose(gm, O.mapwindow_generic(u8, "mean", [-1, -2], [3, 5], "replicate", [0, 0], [1, 1],
                                                       [19, 13]), rtol=1e-12)
    gs = IF.mapwindow(IF.sum_, u8, (3, 5))
    assert gs.dtype == np.uint64
    np.testing.assert_array_equal(gs, O.mapwindow_generic(u8, "sum", [-1, -2], [3, 5], "replicate", [0, 0], [1, 1],
                                                          [19, 13]).astype(np.uint64))
    ge = IF.mapwindow(IF.median, u8, (range(0, 2), range(0, 2)))
    assert ge.dtype == np.float64
    np.testing.assert_array_equal(ge, O.mapwindow_generic(u8, "median", [0, 0], [2, 2], "replicate", [0, 0], [1, 1],
                                                          [19, 13]))
    xn = rng.standard_normal((15, 11))
    xn[4, 5] = np.nan
    gmed = IF.mapwindow(IF.median, xn, (3, 3), indices=(range(0, 15, 2), range(0, 11, 2)))
    wmed = O.mapwindow_generic(xn, "median", [-1, -1], [3, 3], "replicate", [0, 0], [2, 2], [8, 6])
    np.testing.assert_array_equal(np.isnan(gmed), np.isnan(wmed))
    np.testing.assert_array_equal(gmed[~np.isnan(gmed)], wmed[~np.isnan(wmed)])
    for name, f in [("median!", IF.median_), ("minimum", IF.minimum), ("maximum", IF.maximum)]:
        g = IF.mapwindow(f, xn, (3, 3), indices=(range(1, 15, 2), range(0, 11, 3)))
        want = O.mapwindow_generic(xn, name, [-1, -1], [3, 3], "replicate", [1, 0], [2, 3], [7, 4])
        np.testing.assert_array_equal(g, want, err_msg=name)
    with pytest.raises(IF.ArgumentError):
        IF.mapwindow(IF.quantile(0.5), xn, (3, 3))


def test_numpy_and_builtin_functions_accepted():
    rng = np.random.default_rng(142)
    x = rng.standard_normal((16, 9))
    for ours, theirs in [(IF.mean, np.mean), (IF.sum_, sum), (IF.minimum, min), (IF.maximum, np.max),
                         (IF.median, np.median)]:
        np.testing.assert_array_equal(IF.mapwindow(theirs, x, (3, 3)), IF.mapwindow(ours, x, (3, 3)))
    with pytest.raises(NotImplementedError):
        IF.mapwindow(lambda v: v[0] * 2, x, (3, 3))
