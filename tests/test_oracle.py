"""Pins the CPU oracle (oracle/ifm_oracle.c) before it is trusted as the
checker of the GPU path: the reference's own known-answer tests (transcribed
in tests/golden/reference_fixtures.json), independent library routines
(numpy.pad, scipy.ndimage) and closed forms / invariants chosen so that a
dropped term, a wrong sign or a transposed index fails one of them."""
import json
import os

import numpy as np
import pytest
import scipy.ndimage as ndi

from oracle import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))

NP_MODE = {"replicate": "edge", "circular": "wrap", "symmetric": "symmetric", "reflect": "reflect"}
ND_MODE = {"replicate": "nearest", "circular": "wrap", "symmetric": "reflect", "reflect": "mirror"}


def julia_reshape(n, m):
    return np.arange(1, n * m + 1, dtype=np.float64).reshape(n, m, order="F")


# ---- border maps ------------------------------------------------------------------
@pytest.mark.parametrize("style", ["fill0", "replicate", "circular", "symmetric", "reflect"])
def test_padarray_reference_matrices_5x5(style):
    g = GOLD["padarray_5x5"]
    A = julia_reshape(5, 5)
    border = ("fill", 0) if style == "fill0" else style
    P = O.padarray(A, border, g["lo"], g["hi"])
    np.testing.assert_array_equal(P, np.array(g[style], dtype=np.float64))


@pytest.mark.parametrize("style", ["replicate", "circular", "symmetric", "reflect"])
def test_padarray_reference_pad_larger_than_image(style):
    g = GOLD["padarray_2x2_pad3"]
    P = O.padarray(np.array(g["input"], dtype=float), style, g["lo"], g["hi"])
    np.testing.assert_array_equal(P, np.array(g[style], dtype=np.float64))


def test_padarray_reference_asymmetric():
    for c in GOLD["padarray_asymmetric"]["cases"]:
        border = ("fill", c["fill"]) if c["style"] == "fill" else c["style"]
        P = O.padarray(np.array(c["input"], dtype=float), border, c["lo"], c["hi"])
        np.testing.assert_array_equal(P, np.array(c["expect"], dtype=np.float64), err_msg=str(c))


@pytest.mark.parametrize("style", ["replicate", "circular", "symmetric", "reflect"])
def test_padindex_matches_numpy_pad(style):
    rng = np.random.default_rng(1)
    for _ in range(200):
        n = int(rng.integers(1 if style != "reflect" else 2, 9))
        lo, hi = int(rng.integers(0, 25)), int(rng.integers(0, 25))
        got = O.padindex(style, lo, n, hi)
        want = np.pad(np.arange(n), (lo, hi), mode=NP_MODE[style])
        np.testing.assert_array_equal(got, want, err_msg="%s n=%d lo=%d hi=%d" % (style, n, lo, hi))


def test_reflect_length_one_is_an_error():
    with pytest.raises(ZeroDivisionError):
        O.padindex("reflect", 1, 1, 1)


# ---- FIR -------------------------------------------------------------------------------
@pytest.mark.parametrize("style", ["replicate", "circular", "symmetric", "reflect"])
def test_dense_matches_scipy_correlate(style):
    rng = np.random.default_rng(2)
    for shape, ks in [((17, 13), (5, 3)), ((9, 11), (7, 7)), ((6, 5, 7), (3, 3, 3)), ((4, 3), (5, 7))]:
        a = rng.random(shape)
        k = rng.standard_normal(ks)
        first = [-(s // 2) for s in ks]
        got = O.imfilter_dense(a, k, first, style)
        want = ndi.correlate(a, k, mode=ND_MODE[style])
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_dense_fill_matches_scipy_constant():
    rng = np.random.default_rng(3)
    a = rng.random((12, 10))
    k = rng.standard_normal((5, 3))
    got = O.imfilter_dense(a, k, [-2, -1], ("fill", 0.75))
    want = ndi.correlate(a, k, mode="constant", cval=0.75)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_cascade_known_vectors_nd_jl():
    g = GOLD["cascade_1d"]
    img = np.array(g["img"], dtype=float)
    k1, k2 = g["k1"], g["k2"]
    # A0 = padarray(img, Pad(:replicate, (2,), (4,))), A1 = imfilter(A0, k1, Inner())
    A0 = O.padarray(img, "replicate", [2], [4])
    A1 = O.imfilter_dense(A0, np.array(k1["taps"]), [k1["first"]], "replicate", ofirst=[1], osize=[12])
    np.testing.assert_allclose(A1, g["A1"]["values"], rtol=1e-14)
    A2 = O.imfilter_dense(A1, np.array(k2["taps"]), [k2["first"]], "replicate", ofirst=[0], osize=[10])
    np.testing.assert_allclose(A2, g["A2"]["values"], rtol=1e-14)
    A3 = O.imfilter_dense(A2, np.array(k1["taps"]), [k1["first"]], "replicate", ofirst=[1], osize=[8])
    casc = O.imfilter_factors(img, [(0, k1["taps"], -1), (0, k2["taps"], 1), (0, k1["taps"], -1)], "replicate")
    np.testing.assert_allclose(casc, A3, rtol=1e-14)
    np.testing.assert_allclose(O.imfilter_factors(img, [(0, k1["taps"], -1)], "replicate"), g["k1_only"], rtol=1e-14)
    # copy kernels in the cascade change nothing (test/nd.jl:71-75)
    np.testing.assert_allclose(O.imfilter_factors(img, [(0, [1.0], 0), (0, k1["taps"], -1)], "replicate"),
                               g["k1_only"], rtol=1e-14)


@pytest.mark.parametrize("style", ["replicate", "circular", "symmetric", "reflect", ("fill", 0.0)])
def test_factored_equals_convolved_kernel_cascade_jl(style):
    """test/cascade.jl:4-39: a cascade of factors equals filtering with their
    full convolution, for every border."""
    rng = np.random.default_rng(4)
    a = rng.random(15)
    k = [1.0, 1.0, 1.0]
    np.testing.assert_allclose(O.imfilter_factors(a, [(0, k, -1), (0, k, -1)], style),
                               O.imfilter_dense(a, np.array([1.0, 2, 3, 2, 1]), [-2], style), rtol=1e-13)
    a = rng.random((15, 15))
    c = np.array([1.0, 2, 3, 2, 1])
    kx, ky = (0, k, -1), (1, k, -1)
    kern2 = np.outer(c, c)
    np.testing.assert_allclose(O.imfilter_factors(a, [kx, ky, kx, ky], style),
                               O.imfilter_dense(a, kern2, [-2, -2], style), rtol=1e-13)
    np.testing.assert_allclose(O.imfilter_factors(a, [kx, kx, ky], style),
                               O.imfilter_dense(a, np.outer(c, np.ones(3)), [-2, -1], style), rtol=1e-13)


def test_prewitt_tiling_target_2d_jl():
    """test/2d.jl:7-27: the prewitt gradient of a diagonal band."""
    m = np.zeros((20, 20))
    for i in range(-2, 3):
        m += np.diag(np.full(20 - abs(i), 255.0), i)
    dv = GOLD["prewitt_tiling"]["dv"]
    target = np.zeros((20, 20))
    for i in range(1, 5):
        target += np.diag(np.full(20 - i, dv[i - 1]), i)
        target -= np.diag(np.full(20 - i, dv[i - 1]), -i)
    f = [(0, [-0.5, 0.0, 0.5], -1), (1, [1 / 3, 1 / 3, 1 / 3], -1)]
    mf = O.imfilter_factors(m, f, "replicate")
    np.testing.assert_allclose(mf[1:19, 1:19], target[1:19, 1:19], rtol=1e-12, atol=1e-12)


def _target1(kern, border):
    """test/2d.jl:154-173 for a unit impulse at (1,2) (0-based (0,1))."""
    ret = np.zeros((5, 7))
    if border in ("replicate", "symmetric"):
        ret[0, 0] = kern[0, 0] + kern[1, 0]
        ret[1, 0] = kern[0, 0]
    elif border == "circular":
        rk = kern[::-1, ::-1]
        for (i, j), v in np.ndenumerate(rk):
            ret[(0 + i) % 5, (-1 + j) % 7] = v
    else:
        ret[0, 0] = kern[1, 0]
        ret[1, 0] = kern[0, 0]
    return ret


@pytest.mark.parametrize("border", ["replicate", "circular", "symmetric", "reflect", "fill"])
def test_dense_edge_targets_2d_jl(border):
    kern = np.array([[0.1, 0.2], [0.4, 0.5]])
    img = np.zeros((5, 7))
    img[0, 1] = 1
    b = ("fill", 0.0) if border == "fill" else border
    got = O.imfilter_dense(img, kern, [-1, 1], b)
    np.testing.assert_allclose(got, _target1(kern, border), atol=1e-15)
    # factored form of the same kernel
    kf = [(0, [0.2, 0.8], -1), (1, [0.3, 0.6], 1)]
    kk = np.outer([0.2, 0.8], [0.3, 0.6])
    np.testing.assert_allclose(O.imfilter_factors(img, kf, b), _target1(kk, border), atol=1e-15)


def test_dense_interior_impulse_is_rot180():
    kern = np.array([[0.1, 0.2], [0.4, 0.5]])
    img = np.zeros((5, 7))
    img[2, 3] = 1
    for b in ["replicate", "circular", "symmetric", "reflect", ("fill", 0.0)]:
        got = O.imfilter_dense(img, kern, [-1, 1], b)
        want = np.zeros((5, 7))
        want[2:4, 1:3] = kern[::-1, ::-1]
        np.testing.assert_allclose(got, want, atol=1e-15)


def test_3d_box_closed_form_nd_jl():
    img = np.ones((10, 10, 10))
    k = np.full((3, 3, 3), 1 / 27)
    for b in ["replicate", "circular", "symmetric", "reflect"]:
        np.testing.assert_allclose(O.imfilter_dense(img, k, [-1, -1, -1], b), img, rtol=1e-14)
    got = O.imfilter_dense(img, k, [-1, -1, -1], ("fill", 0.0))
    edge = np.array([2 / 3] + [1.0] * 8 + [2 / 3])
    want = edge[:, None, None] * edge[None, :, None] * edge[None, None, :]
    np.testing.assert_allclose(got, want, rtol=1e-13)


# ---- IIR ---------------------------------------------------------------------------------
def _iir_coeffs(sigma):
    from imagefiltering import KernelFactors
    k = KernelFactors.IIRGaussian(sigma)
    return k


def _naive_recursion(x, k, pad):
    """Forward then backward recursion from zero state on a long constant
    extension (pad samples each side): the infinite-extension result that
    the Triggs-Sdika boundary conditions reproduce exactly."""
    xe = np.concatenate([np.full(pad, x[0]), x, np.full(pad, x[-1])])
    y = np.zeros_like(xe)
    for i in range(len(xe)):
        y[i] = xe[i] + sum(k.a[j] * (y[i - j - 1] if i - j - 1 >= 0 else 0.0) for j in range(3))
    z = np.zeros_like(xe)
    for i in range(len(xe) - 1, -1, -1):
        z[i] = y[i] + sum(k.b[j] * (z[i + j + 1] if i + j + 1 < len(xe) else 0.0) for j in range(3))
    return (z * k.scale)[pad:pad + len(x)]


@pytest.mark.parametrize("sigma", [1.5, 2.0, 5.0])
def test_iir_equals_infinite_constant_extension(sigma):
    k = _iir_coeffs(sigma)
    rng = np.random.default_rng(5)
    x = rng.random(40)
    got = O.imfilter_iir(x, [(0, k.a, k.b, k.scale, k.M)], "replicate")
    want = _naive_recursion(x, k, 4000)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12)


def test_iir_fill_equals_extension_with_fill_value():
    k = _iir_coeffs(3.0)
    rng = np.random.default_rng(6)
    x = rng.random(30)
    got = O.imfilter_iir(x, [(0, k.a, k.b, k.scale, k.M)], ("fill", 0.25))
    xe = np.concatenate([[0.25], x, [0.25]])
    want = _naive_recursion(xe, k, 4000)[1:-1]
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12)


def test_iir_constant_passes_through():
    k = _iir_coeffs(2.0)
    got = O.imfilter_iir(np.full((9, 7), 0.6), [(0, k.a, k.b, k.scale, k.M), (1, k.a, k.b, k.scale, k.M)],
                         "replicate")
    np.testing.assert_allclose(got, 0.6, rtol=1e-12)


@pytest.mark.parametrize("sigma", [3.1, 5, 10.0, 20, 50.0, 100.0])
def test_iir_impulse_close_to_gaussian_triggs_jl(sigma):
    """test/triggs.jl:10-27: impulse responses within 10% (L2) of the Gaussian."""
    k = _iir_coeffs(float(sigma))
    l = 1000
    for c in [1, 2, 3, 5, 10, 20, l >> 1, l - 19, l - 9, l - 4, l - 2, l - 1, l]:
        a = np.zeros(l)
        a[c - 1] = 1
        af = np.exp(-((np.arange(1, l + 1) - c) ** 2) / (2 * sigma ** 2)) / (sigma * np.sqrt(2 * np.pi))
        got = O.imfilter_iir(a, [(0, k.a, k.b, k.scale, k.M)], ("fill", 0.0))
        assert np.linalg.norm(got - af) < 0.1 * np.linalg.norm(af)


def test_iir_short_axis_is_an_error():
    k = _iir_coeffs(2.0)
    with pytest.raises(ValueError):
        O.imfilter_iir(np.array([1.0, 2.0]), [(0, k.a, k.b, k.scale, k.M)], "replicate")


# ---- median / min / max / extrema ------------------------------------------------------------
def test_median_reference_vectors():
    g = GOLD["median"]
    for c in g["cases"]:
        a = np.array(c["input"], dtype=float)
        w0, w1 = c["window"]
        got = O.median(a, [w0], [w1 - w0 + 1], "replicate")
        np.testing.assert_array_equal(got, c["expect"], err_msg=str(c))
    m = g["matrix"]
    got = O.median(np.array(m["input"], dtype=float), [-1, -1], [3, 3], "replicate")
    np.testing.assert_array_equal(got, np.array(m["expect"], dtype=float))


@pytest.mark.parametrize("w", [3, 5, 7])
def test_median_matches_scipy(w):
    rng = np.random.default_rng(7)
    a = rng.integers(0, 256, (23, 19)).astype(float)
    h = w // 2
    for style, mode in [("replicate", "nearest"), ("symmetric", "reflect"), ("reflect", "mirror")]:
        got = O.median(a, [-h, -h], [w, w], style)
        np.testing.assert_array_equal(got, ndi.median_filter(a, size=w, mode=mode), err_msg=style)
    got = O.median(a, [-h, -h], [w, w], ("fill", 7.0))
    np.testing.assert_array_equal(got, ndi.median_filter(a, size=w, mode="constant", cval=7.0))


def test_median_even_window_is_lower_median():
    a = np.array([4.0, 1.0, 3.0, 2.0])
    got = O.median(a, [0], [2], "replicate")   # windows {a_i, a_{i+1}} -> lower of the two
    np.testing.assert_array_equal(got, [1.0, 1.0, 2.0, 2.0])


def test_median_isless_order_nan_last_and_signed_zero():
    a = np.array([np.nan, -0.0, 0.0, np.nan, 1.0])
    got = O.median(a, [-1], [3], "replicate")
    # windows: [nan,nan,-0] -> sorted [-0, nan, nan] -> 2nd = nan ; [nan,-0,0] -> [-0,0,nan] -> 0.0
    assert np.isnan(got[0]) and got[1] == 0.0 and not np.signbit(got[1])


def _groundtruth(f, A, window):
    """test/mapwindow.jl:5-14 (clamped window, offsets -(w>>1) .. w-1-(w>>1))."""
    Aex = A.copy()
    hs = [w // 2 + 1 for w in window]
    for shift in np.ndindex(*window):
        idx = np.ix_(*[np.clip(np.arange(A.shape[d]) + shift[d] + 1 - hs[d], 0, A.shape[d] - 1)
                       for d in range(A.ndim)])
        Aex = f(Aex, A[idx])
    return Aex


def test_extrema_1d_reference_vectors():
    for c in GOLD["extrema_1d"]["cases"]:
        a = np.array(c["input"])
        mn, mx = O.extrema_lemire(a, [c["window"]])
        e = np.array(c["expect"])
        np.testing.assert_array_equal(mn, e[:, 0], err_msg=str(c))
        np.testing.assert_array_equal(mx, e[:, 1], err_msg=str(c))


@pytest.mark.parametrize("window", [(2, 2), (2, 3), (3, 2), (3, 3), (2, 5), (2, 2, 2), (2, 3, 2), (3, 3, 3), (2, 5, 3)])
def test_extrema_nd_vs_groundtruth_mapwindow_jl(window):
    rng = np.random.default_rng(8)
    A = rng.random((5,) * len(window)) / 10
    A[(1,) * len(window)] = 0.8
    mn, mx = O.extrema_lemire(A, list(window))
    np.testing.assert_array_equal(mx, _groundtruth(np.maximum, A, window))
    np.testing.assert_array_equal(mn, _groundtruth(np.minimum, A, window))


def test_minmax_generic_path_vs_scipy():
    rng = np.random.default_rng(9)
    a = rng.random((21, 17))
    for w in (3, 5, 9):
        h = w // 2
        mn, mx = O.minmax(a, [-h, -h], [w, w], "replicate")
        np.testing.assert_array_equal(mn, ndi.minimum_filter(a, size=w, mode="nearest"))
        np.testing.assert_array_equal(mx, ndi.maximum_filter(a, size=w, mode="nearest"))
        mn, mx = O.minmax(a, [-h, -h], [w, w], ("fill", 0.5))
        np.testing.assert_array_equal(mn, ndi.minimum_filter(a, size=w, mode="constant", cval=0.5))
        np.testing.assert_array_equal(mx, ndi.maximum_filter(a, size=w, mode="constant", cval=0.5))


# ---- FFT restatement -------------------------------------------------------------------
@pytest.mark.parametrize("style", ["replicate", "circular", "symmetric", "reflect", ("fill", 0.3)])
def test_fft_restatement_equals_direct_2d_jl(style):
    # test/2d.jl:69-75, :99-101 and test/nd.jl:77-81: FFT ≈ FIR for every
    # border, here to rtol 1e-12 over centred, offset and one-sided kernels
    rng = np.random.default_rng(61)
    img = rng.random((61, 47))
    for k, first in ((rng.standard_normal((15, 15)), (-7, -7)), (rng.standard_normal((3, 5)), (5, 2)),
                     (rng.standard_normal((31, 1)), (-15, 0)), (rng.standard_normal((2, 9)), (-3, 1))):
        got = O.imfilter_fft(img, k, first, style)
        want = O.imfilter_dense(img, k, list(first), style)
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), (k.shape, first)
    v = rng.random(100)
    kk = np.ones(31) / 31
    np.testing.assert_allclose(O.imfilter_fft(v, kk, (-15,), style), O.imfilter_dense(v, kk, [-15], style),
                               rtol=1e-12, atol=1e-14)


def test_fft_restatement_padded_size_is_nextprod():
    # Pad(kernel, img, FFT()) pads to nextprod([2,3], n + K) (src/border.jl:240-248)
    assert [O._nextprod23(n) for n in (4111, 4127, 9101, 27, 28)] == [4374, 4374, 9216, 27, 32]


# ---- the -O3 -march=native OpenMP build the CPU baseline times ---------------------------
def test_native_parallel_build_matches_portable_bit_for_bit():
    import subprocess
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(["make", "-f", "oracle/Makefile", "native"], cwd=repo, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("no OpenMP toolchain: " + r.stderr[-200:])
    rng = np.random.default_rng(62)
    img = rng.random((40, 33, 6))
    k = rng.standard_normal((5, 3, 2))
    iir = [(0, [1.07757, -0.427266, 0.0650941], [1.07757, -0.427266, 0.0650941], 0.0809994, np.eye(3).ravel())]

    def run():
        return [O.imfilter_dense(img, k, [-2, -1, 0], "reflect"),
                O.imfilter_factors(img, [(0, np.array([0.25, 0.5, 0.25]), -1), (2, np.array([1.0, -1.0]), 0)],
                                   ("fill", 0.5)),
                O.median(img, [-1, -2, 0], [3, 5, 1], "symmetric"),
                O.minmax(img, [-1, -1, 0], [3, 4, 1], "circular")[0],
                O.extrema_lemire(img, [3, 4, 1])[1],
                O.imfilter_iir(img, iir, "replicate")]
    base = run()
    old = O._LIB
    try:
        assert O.use_library(O.NATIVE_PATH) >= 1
        for n in (1, 3, 8):
            O.set_threads(n)
            for a, b in zip(base, run()):
                np.testing.assert_array_equal(a, b)
    finally:
        O._LIB = old


# ---- NA() and mixed IIR / FIR cascades -------------------------------------------------
def _factor_fn(taps, first, axis=0):
    return lambda x, b: O.imfilter_factors(x, [(axis, np.asarray(taps, dtype=np.float64), first)], b)


def test_na_reference_vectors_border_jl():
    # test/border.jl:272-284
    got = O.imfilter_na(np.arange(1, 11.0), _factor_fn([1 / 3] * 3, -1))
    np.testing.assert_allclose(got, [1.5] + list(range(2, 10)) + [9.5], rtol=1e-14)
    x = np.array([1, np.nan, np.inf, 0, -np.inf, np.inf, np.nan, np.nan, 1])
    got = O.imfilter_na(x, _factor_fn([1.0, 1.0], 0))
    want = np.array([1, np.inf, np.inf, -np.inf, np.nan, np.inf, np.nan, 1, 1])
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    np.testing.assert_array_equal(got[~np.isnan(want)], want[~np.isnan(want)])
    np.testing.assert_array_equal(O.imfilter_na(np.arange(1, 6.0), _factor_fn([1.0], 0)), np.arange(1, 6.0))


def test_na_target_2d_jl():
    # test/2d.jl:174-181: impulse at (1,2) with a kernel on axes (-1:0, 1:2)
    img = np.zeros((5, 7))
    img[0, 1] = 1
    k = np.array([[0.1, 0.2], [0.4, 0.5]])
    got = O.imfilter_na(img, lambda x, b: O.imfilter_dense(x, k, [-1, 1], b))
    want = np.zeros((5, 7))
    want[0, 0] = k[1, 0] / (k[1, 0] + k[1, 1])
    want[1, 0] = k[0, 0] / k.sum()
    want[:, 6] = np.nan
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    np.testing.assert_allclose(got[:, :6], want[:, :6], atol=1e-15)


def _iir_stage(axis, sigma):
    import imagefiltering as IF
    k = IF.KernelFactors.IIRGaussian(sigma)
    return ("iir", axis, k.a, k.b, k.scale, k.M)


@pytest.mark.parametrize("border", ["replicate", ("fill", 0.0)])
def test_mixed_cascade_commutes_triggs_jl(border):
    # test/triggs.jl:34-41: (IIRGaussian(2), box 3) ≈ (box 3, IIRGaussian(2)) on 1:8
    img = np.arange(1, 9.0)
    box = ("fir", 0, [1 / 3] * 3, -1)
    a = O.imfilter_mixed(img, [_iir_stage(0, 2.0), box], border)
    b = O.imfilter_mixed(img, [box, _iir_stage(0, 2.0)], border)
    np.testing.assert_allclose(a, b, rtol=1e-8)
    # and it is not the unpadded composition: the IIR ran over the doubled padding
    assert a.shape == img.shape


def test_mixed_cascade_replicate_equals_long_constant_extension():
    # with "replicate" the padded cascade equals the same stages on a long
    # edge-extended signal, cut back to the image (the IIR's init is exact for
    # a constant continuation)
    rng = np.random.default_rng(63)
    img = rng.random(40)
    stages = [_iir_stage(0, 1.5), ("fir", 0, [0.25, 0.5, 0.25], -1), ("fir", 0, [1.0, -1.0], 0)]
    got = O.imfilter_mixed(img, stages, "replicate")
    ext = np.concatenate([np.full(400, img[0]), img, np.full(400, img[-1])])
    x = O.imfilter_iir(ext, [stages[0][1:]], "replicate")
    x = O.imfilter_factors(x, [(0, np.array(stages[1][2]), -1), (0, np.array(stages[2][2]), 0)], "replicate")
    np.testing.assert_allclose(got, x[400:440], rtol=1e-9, atol=1e-12)


def test_dense_kernel_with_fewer_dims_extends():
    rng = np.random.default_rng(64)
    img = rng.random((12, 9))
    k = rng.standard_normal(3)
    np.testing.assert_array_equal(O.imfilter_dense(img, k, [-1], "reflect"),
                                  O.imfilter_dense(img, k.reshape(3, 1), [-1, 0], "reflect"))


# ---- generic mapwindow -----------------------------------------------------------------
def test_mapwindow_generic_reference_median_vectors():
    """test/mapwindow.jl:105-124 known answers for median and median!."""
    for c in GOLD["median"]["cases"]:
        x = np.asarray(c["input"], dtype=np.float64)
        w = c["window"]
        if x.ndim == 1:
            wf, wl = [w[0]], [w[1] - w[0] + 1]
        else:
            wf, wl = [-(w[0] // 2), -(w[1] // 2)], list(w)
        for f in ("median", "median!"):
            got = O.mapwindow_generic(x, f, wf, wl, "replicate", [0] * x.ndim, [1] * x.ndim, list(x.shape))
            np.testing.assert_array_equal(got, np.asarray(c["expect"], dtype=np.float64), err_msg=str((f, c)))


def test_mapwindow_generic_matches_scipy():
    """Unit-step windows against scipy.ndimage (uniform mean, sum via
    correlate with ones, min / max / median filters) for the Pad modes whose
    window-relative edge map equals the image-wide one on 3-wide windows
    (circular wraps inside the window's intersection instead, which
    orc_median_f64's own pins cover)."""
    rng = np.random.default_rng(5)
    a = rng.standard_normal((9, 7))
    for style in ("replicate", "symmetric", "reflect"):
        m = ND_MODE[style]
        for fname, ref in [("mean", ndi.uniform_filter(a, 3, mode=m)),
                           ("sum", ndi.correlate(a, np.ones((3, 3)), mode=m)),
                           ("minimum", ndi.minimum_filter(a, 3, mode=m)),
                           ("maximum", ndi.maximum_filter(a, 3, mode=m)),
                           ("median!", ndi.median_filter(a, 3, mode=m))]:
            got = O.mapwindow_generic(a, fname, [-1, -1], [3, 3], style, [0, 0], [1, 1], [9, 7])
            np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12, err_msg=str((style, fname)))


def test_mapwindow_generic_strided_is_subsample_and_inner_first():
    """test/mapwindow.jl:131-162: f at strided indices equals the full result
    subsampled; first over (0:2,) returns the image at the centres."""
    rng = np.random.default_rng(6)
    a = rng.standard_normal((10, 5))
    full = O.mapwindow_generic(a, "mean", [-1, 0], [3, 1], "replicate", [0, 0], [1, 1], [10, 5])
    sub = O.mapwindow_generic(a, "mean", [-1, 0], [3, 1], "replicate", [0, 0], [2, 1], [4, 3])
    np.testing.assert_array_equal(sub, full[0:8:2, 0:3])
    img48 = 10.0 * np.arange(1, 11)
    got = O.mapwindow_generic(img48, "first", [0], [3], "replicate", [1], [2], [4])
    np.testing.assert_array_equal(got, img48[1:9:2])
    got = O.mapwindow_generic(img48, "minimum", [-2], [3], "replicate", [2], [1], [8])
    np.testing.assert_array_equal(got, img48[0:8])


def test_mapwindow_generic_even_median_and_quantile():
    """Statistics.median of an even window is the middle of the two central
    values; quantile interpolates linearly (type 7, numpy's default); a
    4-element window pins both by hand."""
    v = np.array([4.0, 1.0, 3.0, 2.0])
    assert O.mapwindow_generic(v, "median", [0], [4], "replicate", [0], [1], [1])[0] == 2.5
    assert O.mapwindow_generic(v, "quantile", [0], [4], "replicate", [0], [1], [1], p=0.25)[0] == 1.75
    assert O.mapwindow_generic(v, "median!", [0], [4], "replicate", [0], [1], [1])[0] == 2.0
    nanv = np.array([1.0, np.nan, 3.0])
    assert np.isnan(O.mapwindow_generic(nanv, "median", [-1], [3], "replicate", [1], [1], [1])[0])
    assert O.mapwindow_generic(nanv, "median!", [-1], [3], "replicate", [1], [1], [1])[0] == 3.0
    # circular edge windows wrap inside their intersection with the image,
    # as the C oracle's median does
    rng = np.random.default_rng(7)
    c = rng.integers(0, 9, (7, 6)).astype(np.float64)
    np.testing.assert_array_equal(
        O.mapwindow_generic(c, "median!", [-1, -2], [3, 5], "circular", [0, 0], [1, 1], [7, 6]),
        O.median(c, [-1, -2], [3, 5], "circular"))
