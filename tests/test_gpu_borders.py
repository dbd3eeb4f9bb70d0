"""GPU parity of the NA() border, the NoPad sub-range form of imfilter! and
mixed IIR / FIR cascades (Algorithm.Mixed), against the oracle's restatements
and the reference's own vectors (test/border.jl:270-285, test/2d.jl:174-201,
test/triggs.jl:34-41, :70-86)."""
import numpy as np
import pytest

import imagefiltering as IF
from imagefiltering import KernelFactors
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def relerr(y, y64):
    y = np.asarray(y, dtype=np.float64)
    y64 = np.asarray(y64, dtype=np.float64)
    m = ~np.isnan(y64)
    scale = np.max(np.abs(y64[m])) if m.any() else 1.0
    return float(np.max(np.abs(y[m] - y64[m])) / (scale if scale > 0 else 1.0))


# ---- NA() ------------------------------------------------------------------------------
def test_na_reference_vectors():
    got = IF.imfilter(np.arange(1, 11), IF.centered(np.array([1, 1, 1]) / 3), IF.NA())
    np.testing.assert_allclose(got, [1.5] + list(range(2, 10)) + [9.5], rtol=1e-7)
    x = np.array([1, np.nan, np.inf, 0, -np.inf, np.inf, np.nan, np.nan, 1])
    got = np.asarray(IF.imfilter(x, IF.OffsetArray(np.array([1.0, 1.0]), range(0, 2)), IF.NA()))
    want = np.array([1, np.inf, np.inf, -np.inf, np.nan, np.inf, np.nan, 1, 1])
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    np.testing.assert_array_equal(got[~np.isnan(want)], want[~np.isnan(want)])
    np.testing.assert_allclose(IF.imfilter(np.arange(1, 6.0), IF.centered(np.array([1.0])), IF.NA()),
                               np.arange(1, 6.0))


@pytest.mark.parametrize("alg", [None, IF.Algorithm.FIR(), IF.Algorithm.FIRTiled(), IF.Algorithm.FFT()])
def test_na_target_2d(alg):
    img = np.zeros((5, 7))
    img[0, 1] = 1
    k = np.array([[0.1, 0.2], [0.4, 0.5]])
    kernel = IF.OffsetArray(k, range(-1, 1), range(1, 3))
    want = np.zeros((5, 7))
    want[0, 0] = k[1, 0] / (k[1, 0] + k[1, 1])
    want[1, 0] = k[0, 0] / k.sum()
    want[:, 6] = np.nan
    got = np.asarray(IF.imfilter(img, kernel, IF.NA()) if alg is None else IF.imfilter(img, kernel, IF.NA(), alg))
    if alg != IF.Algorithm.FFT():
        np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    m = ~np.isnan(want)
    np.testing.assert_allclose(got[m], want[m], atol=1e-6)
    # the factored form of the same kernel (test/2d.jl:203-221)
    kf = (IF.OffsetArray(np.array([0.2, 0.8]), range(-1, 1)), IF.OffsetArray(np.array([[0.3, 0.6]]), range(0, 1),
                                                                              range(1, 3)))
    got = np.asarray(IF.imfilter(img, kf, IF.NA()))
    kk = np.outer([0.2, 0.8], [0.3, 0.6])
    wantf = O.imfilter_na(img, lambda x, b: O.imfilter_dense(x, kk, [-1, 1], b))
    mf = ~np.isnan(wantf)
    np.testing.assert_allclose(got[mf], wantf[mf], atol=1e-6)


@pytest.mark.parametrize("kind", ["separable", "dense", "iir"])
def test_na_with_nans_matches_oracle(kind):
    rng = np.random.default_rng(91)
    img = rng.random((60, 45)).astype(np.float32)
    img[rng.random(img.shape) < 0.05] = np.nan
    img[:, 3] = np.nan                          # a whole column: windows with few valid values
    if kind == "separable":
        kern = KernelFactors.gaussian((1.5, 2.0))
        fn = lambda x, b: O.imfilter_factors(x, [(f.Npre, f.data.data.astype(np.float64), f.data.first[0])
                                                 for f in kern], b)
    elif kind == "dense":
        # NA() normalises by the filtered mask of valid pixels, so it is meant
        # for non-negative weights; with mixed-sign taps that mask can cancel
        # towards zero and the FP32 sums lose the relative accuracy of the quotient
        kd = np.abs(rng.standard_normal((5, 7))) + 0.05
        kern = IF.centered(kd)
        fn = lambda x, b: O.imfilter_dense(x, kd, [-2, -3], b)
    else:
        kern = KernelFactors.IIRGaussian((2.0, 2.0))
        fn = lambda x, b: O.imfilter_iir(x, [(f.Npre, f.data.a, f.data.b, f.data.scale, f.data.M) for f in kern], b)
    got = IF.imfilter(img, kern, IF.NA())
    want = O.imfilter_na(img.astype(np.float64), fn)
    g = np.asarray(got, dtype=np.float64)
    np.testing.assert_array_equal(np.isnan(g), np.isnan(want))
    assert relerr(g, want) <= TOL
    # imfilter! into a given output, on device arrays too
    d = IF.DeviceArray.from_host(img)
    out = IF.imfilter(d, kern, IF.NA()).to_host()
    np.testing.assert_array_equal(np.asarray(out), np.asarray(got))


def test_na_iir_nan_at_corner_triggs_jl():
    # test/triggs.jl:76-86
    rng = np.random.default_rng(92)
    imgf = rng.random((20, 22))
    kern = KernelFactors.IIRGaussian((2.0, 2.0))
    nan = imgf.copy()
    nan[0, 0] = np.nan
    num = imgf.copy()
    num[0, 0] = 0
    den = np.ones_like(imgf)
    den[0, 0] = 0
    retnum = IF.imfilter(num, kern, IF.Fill(0.0))
    retden = IF.imfilter(den, kern, IF.Fill(0.0))
    ret = np.asarray(IF.imfilter(nan, kern, IF.NA()))
    ret[0, 0] = retnum[0, 0] = 0
    np.testing.assert_allclose(ret, retnum / np.where(retden == 0, 1, retden), rtol=1e-9, atol=1e-12)


# ---- imfilter!(r, out, A, kernel, NoPad(), inds) ----------------------------------------
@pytest.mark.parametrize("kind", ["factors", "dense"])
def test_nopad_sub_range(kind):
    rng = np.random.default_rng(93)
    img = rng.random((40, 50))
    if kind == "factors":
        kern = KernelFactors.gaussian((1.0, 1.5))
        kd = np.outer(kern[0].data.data, kern[1].data.data)
        first = [kern[0].data.first[0], kern[1].data.first[0]]
    else:
        kd = rng.standard_normal((3, 5))
        kern = IF.centered(kd)
        first = [-1, -2]
    A = IF.padarray(img, IF.Pad("reflect", (4, 5), (4, 5)))
    out = IF.OffsetArray(np.full((40, 50), -7.0), (0, 0))
    inds = (range(3, 31), range(10, 44))
    IF.imfilter_(IF.CPU1(IF.Algorithm.FIR()), out, A, kern if kind == "dense" else kern, IF.NoPad(), inds)
    full = O.imfilter_dense(img, kd, first, "reflect")
    o = np.asarray(out.data)
    blk = (slice(3, 31), slice(10, 44))
    assert relerr(o[blk], full[blk]) <= TOL
    mask = np.ones(o.shape, bool)
    mask[blk] = False
    assert np.all(o[mask] == -7.0)          # the rest of out is untouched
    # the same on device arrays
    dout = IF.DeviceArray.from_host(np.full((40, 50), -7.0))
    IF.imfilter_(IF.CPU1(IF.Algorithm.FIR()), dout, IF.OffsetArray(IF.DeviceArray.from_host(A.data), A.first),
                 kern, IF.NoPad(), inds)
    np.testing.assert_array_equal(dout.to_host(), o)


# ---- mixed IIR / FIR cascades ----------------------------------------------------------
def _iir_stage(k, axis):
    return ("iir", axis, k.a, k.b, k.scale, k.M)


@pytest.mark.parametrize("border", [IF.Pad("replicate"), IF.Fill(0.0)])
def test_mixed_commutes_triggs_jl(border):
    img = np.arange(1, 9)
    k1 = KernelFactors.IIRGaussian(2)
    k2 = IF.centered(np.ones(3) / 3)
    a = np.asarray(IF.imfilter(img, (k1, k2), border))
    b = np.asarray(IF.imfilter(img, (k2, k1), border))
    np.testing.assert_allclose(a, b, rtol=1e-5)
    ob = "replicate" if isinstance(border, IF.Pad) else ("fill", 0.0)
    want = O.imfilter_mixed(img.astype(np.float64), [_iir_stage(k1, 0), ("fir", 0, np.ones(3) / 3, -1)], ob)
    assert relerr(a, want) <= TOL
    assert IF.filter_algorithm(img, (k1, k2)) == IF.Algorithm.Mixed()


@pytest.mark.parametrize("border", ["replicate", "circular", "symmetric", IF.Fill(0.4)])
def test_mixed_2d_cascades_match_oracle(border):
    rng = np.random.default_rng(94)
    img = rng.random((37, 52)).astype(np.float32)
    g = KernelFactors.IIRGaussian((1.5, 2.5))
    s = KernelFactors.sobel((True, True), 1)
    kern = (g[0], s[0], g[1], s[1])
    got = IF.imfilter(img, kern, border)
    stages = [_iir_stage(g[0].data, 0), ("fir", 0, s[0].data.data.astype(np.float64), s[0].data.first[0]),
              _iir_stage(g[1].data, 1), ("fir", 1, s[1].data.data.astype(np.float64), s[1].data.first[0])]
    ob = ("fill", 0.4) if isinstance(border, IF.Fill) else border
    want = O.imfilter_mixed(img.astype(np.float64), stages, ob)
    assert relerr(got, want) <= TOL, border
    # an IIR stage last, and an Inner() output
    kern2 = (s[0], g[0])
    got2 = IF.imfilter(img, kern2, border)
    want2 = O.imfilter_mixed(img.astype(np.float64), [stages[1], stages[0]], ob)
    assert relerr(got2, want2) <= TOL
