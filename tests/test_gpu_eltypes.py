"""GPU parity of the element-type rules beyond floats: integer outputs and
InexactError (src/imfilter.jl:233-254, :1131-1154), N0f8 linear filtering,
colour images through their channel view (src/imfilter.jl:868-883), and
tuples of bare TriggsSdika filters (test/triggs.jl:60-68)."""
import numpy as np
import pytest

import imagefiltering as IF
from imagefiltering import Kernel, KernelFactors
from imagefiltering.errors import InexactError
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dt,want_dt", [(np.int64, np.int64), (np.int32, np.int32), (np.uint8, np.int16),
                                        (np.uint16, np.int32), (np.bool_, np.int8)])
def test_laplacian_integer_impulses_specialty_jl(dt, want_dt):
    # test/specialty.jl:27-56: the Laplacian keeps signed eltypes and widens unsigned ones
    a = np.zeros(5, dtype=dt)
    a[2] = 1
    af = IF.imfilter(a, Kernel.Laplacian((True,)))
    assert af.dtype == want_dt
    np.testing.assert_array_equal(af, [0, 1, -2, 1, 0])
    a2 = np.zeros((5, 5), dtype=dt)
    a2[2, 2] = 1
    af2 = IF.imfilter(a2, Kernel.Laplacian())
    assert af2.dtype == want_dt
    want = np.array([[0, 0, 0, 0, 0], [0, 0, 1, 0, 0], [0, 1, -4, 1, 0], [0, 0, 1, 0, 0], [0, 0, 0, 0, 0]])
    np.testing.assert_array_equal(af2, want)


def test_integer_image_integer_kernel_exact():
    rng = np.random.default_rng(111)
    img = rng.integers(0, 256, (70, 53)).astype(np.uint8)
    # a separable integer matrix would be SVD-factored into Float64 factors
    # (factorkernel, src/imfilter.jl:1156-1184) and give Float64; this one is not
    k = np.array([[0, 1, 0], [1, -4, 2], [0, 1, 0]])
    assert len(IF.factorkernel(IF.centered(k))) == 1
    for b in ("replicate", "circular", "symmetric", "reflect"):
        got = IF.imfilter(img, IF.centered(k), b)
        assert got.dtype == np.int64                       # UInt8 * Int64 -> Int64
        want = O.imfilter_dense(img.astype(np.float64), k.astype(np.float64), [-1, -1], b)
        np.testing.assert_array_equal(got, want.astype(np.int64))
    got = IF.imfilter(img.astype(np.int16), IF.centered(np.array([1, -2, 1], dtype=np.int16)))
    assert got.dtype == np.int16
    want = O.imfilter_dense(img.astype(np.float64), np.array([1.0, -2.0, 1.0]), [-1], "replicate")
    np.testing.assert_array_equal(got, want.astype(np.int16))


def test_inexact_error_on_integer_outputs():
    rng = np.random.default_rng(112)
    img = rng.integers(0, 256, (20, 30)).astype(np.uint8)
    with pytest.raises(InexactError):          # a non-integer result in UInt8
        IF.imfilter(np.uint8, img, IF.centered(np.ones(3) / 3))
    with pytest.raises(InexactError):          # UInt8 * UInt8 stays UInt8 and overflows
        IF.imfilter(img | 128, IF.centered(np.array([1, 1, 1], dtype=np.uint8)))
    # representable results convert exactly
    out = IF.imfilter(np.int64, (img // 4) * 4, IF.centered(np.array([0.25, 0.5, 0.25])))
    want = O.imfilter_dense(((img // 4) * 4).astype(np.float64), np.array([0.25, 0.5, 0.25]), [-1], "replicate")
    if np.all(want == np.round(want)):
        np.testing.assert_array_equal(out, want.astype(np.int64))
    # imfilter! into an integer out
    out2 = np.zeros((20, 30), dtype=np.int32)
    IF.imfilter_(out2, img, IF.centered(np.array([1, 0, -1])))
    want2 = O.imfilter_dense(img.astype(np.float64), np.array([1.0, 0.0, -1.0]), [-1], "replicate")
    np.testing.assert_array_equal(out2, want2.astype(np.int32))


def test_n0f8_linear_filtering():
    rng = np.random.default_rng(113)
    codes = rng.integers(0, 256, (90, 77)).astype(np.uint8)
    img = IF.n0f8(codes)
    g = KernelFactors.gaussian((1.5, 2.5))
    for b in ("replicate", "reflect", IF.Fill(0.5)):
        out = IF.imfilter(img, g, b)
        assert out.dtype == np.float64           # N0f8 * Float64 -> Float64
        # Fill(0.5) pads an N0f8 image with convert(N0f8, 0.5) = 128/255
        ob = ("fill", 128 / 255) if isinstance(b, IF.Fill) else b
        want = O.imfilter_factors(codes / 255.0, [(f.Npre, f.data.data, f.data.first[0]) for f in g], ob)
        assert np.max(np.abs(out - want)) <= 1e-5 * np.max(np.abs(want))
    k = rng.standard_normal((15, 15))
    k /= k.sum()
    for alg in (IF.Algorithm.FIR(), IF.Algorithm.FFT()):
        out = IF.imfilter(np.float32, img, IF.centered(k), "symmetric", alg)
        want = O.imfilter_dense(codes / 255.0, k, [-7, -7], "symmetric")
        assert np.max(np.abs(out - want)) <= 1e-5 * np.max(np.abs(want)), alg
    # padarray(Float32, ::N0f8) holds the exactly rounded values code/255
    p = IF.padarray(np.float32, img, IF.Pad("replicate", (1, 1), (1, 1)))
    np.testing.assert_array_equal(np.asarray(p.data)[1:-1, 1:-1], (codes / 255.0).astype(np.float32))


def test_colour_images_filter_each_channel():
    rng = np.random.default_rng(114)
    ch = rng.random((3, 40, 51)).astype(np.float32)
    rgb = IF.colorview("RGB", ch)
    g = KernelFactors.gaussian((1.0, 2.0))
    out = IF.imfilter(rgb, g, "reflect")
    assert isinstance(out, IF.ColorImage) and out.shape == (40, 51)
    for c in range(3):
        np.testing.assert_array_equal(IF.channelview(out)[c], IF.imfilter(ch[c], g, "reflect"))
    k = IF.centered(rng.standard_normal((15, 15)))
    out = IF.imfilter(rgb, k, "replicate")                  # FFT by default, per channel
    for c in range(3):
        want = O.imfilter_dense(ch[c].astype(np.float64), np.asarray(k.data), [-7, -7], "replicate")
        assert np.max(np.abs(IF.channelview(out)[c] - want)) <= 1e-5 * np.max(np.abs(want))
    out = IF.imfilter(rgb, g, IF.Inner())
    assert isinstance(out, IF.OffsetArray) and out.data.shape == tuple(n - 2 * h for n, h in zip((40, 51), (2, 4)))
    n8 = IF.colorview("RGB", IF.n0f8((ch * 255).astype(np.uint8)))
    o8 = IF.imfilter(n8, g)
    assert IF.channelview(o8).dtype == np.float64


def test_bare_triggs_sdika_tuple_per_dimension_triggs_jl():
    # test/triggs.jl:60-68: (kern, kern) filters dimension d with kernel[d]
    rng = np.random.default_rng(115)
    img = rng.random((40, 33))
    k = KernelFactors.IIRGaussian(2.5)
    ret = np.asarray(IF.imfilter(img, (k, k)))
    out = np.empty_like(img)
    IF.imfilter_(IF.CPU1(IF.Algorithm.IIR()), out, img, k, 1, "replicate")
    IF.imfilter_(IF.CPU1(IF.Algorithm.IIR()), out, out, k, 2, "replicate")
    np.testing.assert_allclose(ret, out, rtol=1e-12, atol=1e-14)
    want = O.imfilter_iir(img, [(0, k.a, k.b, k.scale, k.M), (1, k.a, k.b, k.scale, k.M)], "replicate")
    np.testing.assert_allclose(ret, want, rtol=1e-9, atol=1e-12)
