"""GPU parity: the HIP path (through the public API and the C ABI) against the
CPU oracle on the same seeded inputs.  Integer / index / order-statistic work
must be bit-exact; floating-point filtering must stay within a normwise
relative error of 1e-5 of the FP64 oracle: max|y - y64| / max|y64| <= 1e-5."""
import numpy as np
import pytest

import imagefiltering as IF
from imagefiltering import Kernel, KernelFactors
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
STYLES = ["replicate", "circular", "symmetric", "reflect"]


def relerr(y, y64):
    y = np.asarray(y, dtype=np.float64)
    y64 = np.asarray(y64, dtype=np.float64)
    scale = np.max(np.abs(y64))
    return float(np.max(np.abs(y - y64)) / (scale if scale > 0 else 1.0))


def obord(b):
    if isinstance(b, IF.Fill):
        return ("fill", float(b.value))
    return b


def factors_of(kern):
    out = []
    for f in kern:
        out.append((f.Npre, f.data.data.astype(np.float64), f.data.first[0]))
    return out


# ---- padarray ----------------------------------------------------------------------------------
@pytest.mark.parametrize("style", STYLES + ["fill"])
def test_padarray_reference_matrices(style):
    import json, os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))
    A = np.arange(1, 26, dtype=np.float64).reshape(5, 5, order="F")
    b = IF.Fill(0, (2, 2), (2, 2)) if style == "fill" else IF.Pad(style, (2, 2), (2, 2))
    P = IF.padarray(A, b)
    assert P.first == (-2, -2)
    np.testing.assert_array_equal(P.data, np.array(g["padarray_5x5"]["fill0" if style == "fill" else style]))
    A2 = np.array([[1.0, 2.0], [3.0, 4.0]])
    if style != "fill":
        np.testing.assert_array_equal(IF.padarray(A2, IF.Pad(style, (3, 3), (3, 3))).data,
                                      np.array(g["padarray_2x2_pad3"][style]))


@pytest.mark.parametrize("dtype", ["f32", "f64", "u8", "n0f8", "bf16"])
def test_padarray_random_all_styles_bit_exact(dtype):
    rng = np.random.default_rng(11)
    for trial in range(12):
        nd = 1 + trial % 3
        shape = tuple(int(rng.integers(2, 9)) for _ in range(nd))
        lo = tuple(int(rng.integers(0, 12)) for _ in range(nd))
        hi = tuple(int(rng.integers(0, 12)) for _ in range(nd))
        x = rng.random(shape)
        if dtype == "f32":
            a, ref = x.astype(np.float32), x.astype(np.float32)
        elif dtype == "f64":
            a, ref = x, x
        elif dtype == "u8":
            a = (x * 255).astype(np.uint8)
            ref = a
        elif dtype == "n0f8":
            a = IF.n0f8((x * 255).astype(np.uint8))
            ref = np.asarray(a)
        else:
            a = IF.bf16(x)
            ref = np.asarray(a)
        for style in STYLES + ["fill"]:
            b = IF.Fill(0, lo, hi) if style == "fill" else IF.Pad(style, lo, hi)
            P = IF.padarray(a, b)
            want = O.padarray(ref.astype(np.float64), obord(b) if style == "fill" else style, list(lo), list(hi))
            np.testing.assert_array_equal(np.asarray(P.data).astype(np.float64), want, err_msg="%s %s" % (style, shape))


def test_padarray_long_rows_and_many_rows():
    # rows longer than one workgroup's span, and more rows than one grid column
    rng = np.random.default_rng(12)
    cases = [((5000,), (7,), (1030,)), ((1100, 3), (3, 2), (1, 4)), ((2, 300, 250), (1, 2, 3), (2, 1, 0))]
    for shape, lo, hi in cases:
        x = np.asfortranarray(rng.random(shape).astype(np.float32))
        for style in STYLES + ["fill"]:
            b = IF.Fill(0.25, lo, hi) if style == "fill" else IF.Pad(style, lo, hi)
            P = IF.padarray(x, b)
            want = O.padarray(x.astype(np.float64), obord(b) if style == "fill" else style, list(lo), list(hi))
            np.testing.assert_array_equal(np.asarray(P.data).astype(np.float64), want, err_msg="%s %s" % (style, shape))
        P = IF.padarray(IF.Float64, x, IF.Pad("circular", lo, hi))
        want = O.padarray(x.astype(np.float64), "circular", list(lo), list(hi))
        np.testing.assert_array_equal(P.data, want)


def test_padarray_conversion_and_layout():
    x = np.asfortranarray(np.random.default_rng(1).random((7, 5)).astype(np.float32))
    P = IF.padarray(IF.Float64, x, IF.Pad("symmetric", (3, 1), (2, 4)))
    want = O.padarray(x.astype(np.float64), "symmetric", [3, 1], [2, 4])
    np.testing.assert_array_equal(P.data, want)
    assert P.data.dtype == np.float64
    xc = np.ascontiguousarray(x)
    np.testing.assert_array_equal(IF.padarray(xc, IF.Pad("reflect", (3, 1), (2, 4))).data,
                                  O.padarray(xc.astype(np.float64), "reflect", [3, 1], [2, 4]))


# ---- separable FIR ---------------------------------------------------------------------------
def test_config1_gaussian_1024_replicate_f64_out():
    """imfilter(img, KernelFactors.gaussian((3.0,3.0)), "replicate") on a 1024x1024
    Float32 image: Float64 output within 1e-5 of the FP64 oracle."""
    rng = np.random.default_rng(20261203)
    img = rng.random((1024, 1024), dtype=np.float32)
    kern = KernelFactors.gaussian((3.0, 3.0))
    out = IF.imfilter(img, kern, "replicate")
    assert out.dtype == np.float64 and out.shape == img.shape
    want = O.imfilter_factors(img.astype(np.float64), factors_of(kern), "replicate")
    assert relerr(out, want) <= TOL


@pytest.mark.parametrize("border", STYLES + [IF.Fill(0.3)])
def test_separable_all_borders_random_shapes(border):
    rng = np.random.default_rng(12)
    for shape in [(37, 29), (5, 3), (64, 17), (130, 70)]:
        img = rng.random(shape).astype(np.float32)
        for sig in [(1.0, 2.0), (0.7, 3.3)]:
            kern = KernelFactors.gaussian(sig)
            out = IF.imfilter(img, kern, border)
            want = O.imfilter_factors(img.astype(np.float64), factors_of(kern), obord(border))
            assert relerr(out, want) <= TOL, (shape, sig, border)


def test_separable_offsets_repeats_and_inner():
    rng = np.random.default_rng(13)
    img = rng.random((40, 33))
    k1 = IF.OffsetArray(np.array([0.2, 0.8]), (-1,))
    k2 = IF.OffsetArray(np.array([0.3, 0.6, 0.1]), (1,))
    kern = (IF.KernelFactors.ReshapedOneD(2, 0, k1), IF.KernelFactors.ReshapedOneD(2, 1, k2),
            IF.KernelFactors.ReshapedOneD(2, 0, k1))
    for b in STYLES + [IF.Fill(-1.0)]:
        out = IF.imfilter(img, kern, b)
        want = O.imfilter_factors(img, [(0, [0.2, 0.8], -1), (1, [0.3, 0.6, 0.1], 1), (0, [0.2, 0.8], -1)], obord(b))
        assert relerr(out, want) <= TOL, b
    res = IF.imfilter(img, kern, IF.Inner())
    lo = (2, 0)
    hi = (0, 3)
    assert res.first == lo and res.data.shape == (40 - 2, 33 - 3)
    want = O.imfilter_factors(img, [(0, [0.2, 0.8], -1), (1, [0.3, 0.6, 0.1], 1), (0, [0.2, 0.8], -1)], "replicate",
                              ofirst=list(lo), osize=list(res.data.shape))
    assert relerr(res.data, want) <= TOL


def test_separable_3d_and_1d():
    rng = np.random.default_rng(14)
    vol = rng.random((31, 20, 17)).astype(np.float32)
    kern = KernelFactors.gaussian((2.0, 2.0, 1.5))
    for b in ["symmetric", "circular", IF.Fill(0.0)]:
        out = IF.imfilter(vol, kern, b)
        want = O.imfilter_factors(vol.astype(np.float64), factors_of(kern), obord(b))
        assert relerr(out, want) <= TOL, b
    v = rng.random(100)
    k = IF.centered(np.array([0.25, 0.5, 0.25]))
    out = IF.imfilter(v, k)
    want = O.imfilter_factors(v, [(0, [0.25, 0.5, 0.25], -1)], "replicate")
    assert relerr(out, want) <= TOL


def test_separable_3d_fused_many_tiles():
    # Several (x, y) tiles, three z chunks of 64 planes with a partial last
    # chunk, z extent not a multiple of the axis-2 factor length; F32 and F64
    # input; every border.
    rng = np.random.default_rng(41)
    for dt, sig in [(np.float32, (1.0, 2.0, 1.0)), (np.float64, (2.0, 1.0, 2.0))]:
        vol = np.asfortranarray(rng.random((70, 45, 139)).astype(dt))
        kern = KernelFactors.gaussian(sig)
        for b in ["replicate", "reflect", "symmetric", "circular", IF.Fill(0.5)]:
            out = IF.imfilter(vol, kern, b)
            want = O.imfilter_factors(vol.astype(np.float64), factors_of(kern), obord(b))
            assert relerr(out, want) <= TOL, (dt, b)


def test_single_axis_factors_long_rows():
    # one factor per call along axis 0, 1 or 2 (the per-axis pass), rows
    # longer than a workgroup's span, a 3-D cascade whose axis-2 factor is too
    # long for the fused pass, and the identity (a plain copy)
    rng = np.random.default_rng(42)
    sig = rng.random(5000).astype(np.float32)
    k7 = IF.centered(np.array([0.05, 0.1, 0.2, 0.3, 0.2, 0.1, 0.05]))
    for b in STYLES + [IF.Fill(0.5)]:
        out = IF.imfilter(sig, KernelFactors.kernelfactors((k7,)), b)
        want = O.imfilter_factors(sig.astype(np.float64), [(0, list(k7.data), -3)], obord(b))
        assert relerr(out, want) <= TOL, b
    vol = np.asfortranarray(rng.random((1500, 7, 9)).astype(np.float32))
    for ax in range(3):
        f = [np.array([1.0])] * 3
        f[ax] = np.array([0.25, 0.5, 0.25])
        kern = KernelFactors.kernelfactors(tuple(IF.centered(x) for x in f))
        for b in ["reflect", "circular"]:
            out = IF.imfilter(vol, kern, b)
            want = O.imfilter_factors(vol.astype(np.float64), factors_of(kern), obord(b))
            assert relerr(out, want) <= TOL, (ax, b)
    v3 = np.asfortranarray(rng.random((40, 30, 50)).astype(np.float32))
    kern = KernelFactors.gaussian((1.0, 1.5, 3.0))  # 5, 9 and 13 taps
    for b in ["symmetric", IF.Fill(0.0)]:
        out = IF.imfilter(v3, kern, b)
        want = O.imfilter_factors(v3.astype(np.float64), factors_of(kern), obord(b))
        assert relerr(out, want) <= TOL, b
    ident = KernelFactors.kernelfactors((IF.centered(np.array([1.0])), IF.centered(np.array([1.0]))))
    img = np.asfortranarray(rng.random((1300, 20)).astype(np.float32))
    np.testing.assert_array_equal(IF.imfilter(img, ident), img.astype(np.float64))


def test_row_major_input_same_as_column_major():
    rng = np.random.default_rng(15)
    img = rng.random((50, 70)).astype(np.float32)
    kern = KernelFactors.gaussian((1.5, 2.5))
    a = IF.imfilter(np.ascontiguousarray(img), kern, "reflect")
    b = IF.imfilter(np.asfortranarray(img), kern, "reflect")
    np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-7)


def test_prewitt_tiling_uint8():
    m = np.zeros((20, 20), dtype=np.uint8)
    for i in range(-2, 3):
        idx = np.arange(max(0, -i), min(20, 20 - i))
        m[idx, idx + i] = 255
    kernel = KernelFactors.prewitt((True, True), 0)
    mf = IF.imfilter(m, kernel)
    dv = [42.5, 85, 85, 42.5]
    target = np.zeros((20, 20))
    for i in range(1, 5):
        target += np.diag(np.full(20 - i, dv[i - 1]), i)
        target -= np.diag(np.full(20 - i, dv[i - 1]), -i)
    np.testing.assert_allclose(mf[1:19, 1:19], target[1:19, 1:19], atol=1e-3)


# ---- dense FIR ------------------------------------------------------------------------------
@pytest.mark.parametrize("ks", [(3, 3), (5, 7), (1, 9), (8, 6)])
def test_dense_valu_small_kernels(ks):
    rng = np.random.default_rng(16)
    img = rng.random((45, 38)).astype(np.float32)
    k = rng.standard_normal(ks)
    kern = IF.centered(k)
    for b in STYLES + [IF.Fill(0.5)]:
        out = IF.imfilter(img, kern, b, IF.Algorithm.FIR())
        want = O.imfilter_dense(img.astype(np.float64), k, list(kern.first), obord(b))
        assert relerr(out, want) <= TOL, (ks, b)


@pytest.mark.parametrize("ks", [(3, 3), (5, 5), (7, 7), (9, 13), (13, 11), (3, 13), (11, 5), (13, 13), (15, 15),
                                (15, 7), (3, 15)])
def test_dense_register_blocked_sizes(ks):
    """The register-blocked VALU kernels (odd sizes 3 .. 15 per axis):
    kernels off centre, images with interior and border tiles, a 3-D image
    (planes), F64 / BF16 / UInt8 inputs, every border."""
    from imagefiltering import _lib
    rng = np.random.default_rng(300 + 17 * ks[0] + ks[1])
    k = rng.standard_normal(ks)
    first = [-(ks[0] // 3), -(ks[1] - 2)]
    img = np.asfortranarray(rng.random((150, 70)).astype(np.float32))
    for b in STYLES + [IF.Fill(0.5)]:
        out = _dense_lib(img, k, first, b if isinstance(b, IF.Fill) else IF.Pad(b), _lib.PREC_FP32)
        want = O.imfilter_dense(img.astype(np.float64), k, first, obord(b))
        assert relerr(out, want) <= TOL, (ks, b)
    vol = np.asfortranarray(rng.random((70, 40, 3)))
    kern = IF.OffsetArray(k, *first)
    out = IF.imfilter(vol, kern, "reflect", IF.Algorithm.FIR())
    want = O.imfilter_dense(vol, k, first, "reflect")
    assert relerr(out, want) <= TOL
    codes = np.asfortranarray(rng.integers(0, 256, (67, 45)).astype(np.uint8))
    out = IF.imfilter(codes, kern, "symmetric", IF.Algorithm.FIR())
    assert relerr(out, O.imfilter_dense(codes.astype(np.float64), k, first, "symmetric")) <= TOL
    xb = IF.bf16(rng.random((66, 70)).astype(np.float32))
    out = IF.imfilter(IF.Float32, xb, kern, IF.Fill(0.0), IF.Algorithm.FIR())
    want = O.imfilter_dense(IF.bf16_to_float32(xb).astype(np.float64), k, first, ("fill", 0.0))
    assert relerr(out, want) <= TOL
    # the interior only (Inner()): output (0, 0) is image index -first
    res = IF.imfilter(img, kern, IF.Inner(), IF.Algorithm.FIR())
    lo = [-f for f in first]
    size = [img.shape[d] - ks[d] + 1 for d in range(2)]
    assert tuple(res.first) == tuple(lo) and res.data.shape == tuple(size)
    want = O.imfilter_dense(img.astype(np.float64), k, first, "replicate", lo, size)
    assert relerr(res.data, want) <= TOL


def norm_random_kernel(rng, K):
    while True:
        k = rng.standard_normal((K, K))
        s = k.sum()
        if abs(s) >= 1:
            return k / s


@pytest.mark.parametrize("K", [15, 17, 31])
def test_dense_mfma_split_bf16(K):
    rng = np.random.default_rng(17 + K)
    img = rng.random((300, 257)).astype(np.float32)
    k = norm_random_kernel(rng, K)
    kern = IF.centered(k)
    for b in ["reflect", "replicate", "circular", IF.Fill(0.0)]:
        out = IF.imfilter(img, kern, b, IF.Algorithm.FIR())
        want = O.imfilter_dense(img.astype(np.float64), k, list(kern.first), obord(b))
        e = relerr(out, want)
        assert e <= TOL, (K, b, e)


@pytest.mark.parametrize("K,dtype", [(15, "f32"), (15, "bf16"), (15, "u8"), (15, "f64"), (31, "f32")])
def test_dense_mfma_many_tiles_per_workgroup(K, dtype):
    """A stack of planes gives every persistent workgroup several tiles, so the
    prefetch of the next tile (and, for K = 31, the rotation of the Toeplitz
    register sets) is exercised; plane and border changes happen mid-run."""
    rng = np.random.default_rng(40 + K)
    shape = (700, 300, 14) if K == 15 else (400, 260, 8)
    x = rng.random(shape).astype(np.float32)
    if dtype == "bf16":
        img = IF.bf16(x)
        x64 = IF.bf16_to_float32(img).astype(np.float64)
    elif dtype == "u8":
        img = (x * 255).astype(np.uint8)
        x64 = img.astype(np.float64)
    elif dtype == "f64":
        img = x.astype(np.float64)
        x64 = img
    else:
        img = x
        x64 = x.astype(np.float64)
    k = norm_random_kernel(rng, K)[:, :, None]
    kern = IF.centered(k)
    fill = IF.Fill(3) if dtype == "u8" else IF.Fill(0.25)  # Fill converts to the image eltype
    for b in ["reflect", fill]:
        out = IF.imfilter(img, kern, b, IF.Algorithm.FIR())
        want = O.imfilter_dense(x64, k, list(kern.first), obord(b))
        e = relerr(out, want)
        assert e <= TOL, (K, dtype, b, e)


@pytest.mark.parametrize("K", [15, 17])
def test_dense_mfma_interior_tiles_2d(K):
    """Column-major 2-D images with many interior tiles (staged by 16-byte
    quads) and border tiles (staged element by element); widths that are and
    are not multiples of 4 (vector stores or element stores); F32 and F64
    outputs; every border."""
    rng = np.random.default_rng(60 + K)
    k = norm_random_kernel(rng, K)
    kern = IF.centered(k)
    for shape in [(700, 520), (701, 523)]:
        x = np.asfortranarray(rng.random(shape).astype(np.float32))
        for b in ["reflect", "circular", IF.Fill(0.25)]:
            for T in (None, IF.Float64):
                out = IF.imfilter(x, kern, b, IF.Algorithm.FIR()) if T is None else \
                    IF.imfilter(T, x, kern, b, IF.Algorithm.FIR())
                want = O.imfilter_dense(x.astype(np.float64), k, list(kern.first), obord(b))
                e = relerr(out, want)
                assert e <= TOL, (K, shape, b, T, e)


def _dense_lib(img, k, first, border, precision):
    """ifm_imfilter_dense with an explicit precision on device arrays."""
    from imagefiltering import _lib
    from imagefiltering.device import ifm_array
    from imagefiltering.imfilter import _border_struct, _kernel_struct
    L, ctx = _lib.load(), _lib.ctx()
    d_in = IF.DeviceArray.from_host(img)
    d_out = d_in.similar()
    rc = L.ifm_imfilter_dense(ctx, ifm_array(d_in), ifm_array(d_out), None,
                              _kernel_struct(IF.OffsetArray(k, *first)), _border_struct(border, IF.Float32),
                              precision)
    _lib.check(rc)
    return d_out.to_host()


@pytest.mark.parametrize("ks", [(3, 3), (5, 9), (25, 5), (17, 16), (49, 32), (33, 20), (31, 15), (40, 12)])
def test_dense_mfma_explicit_shapes(ks):
    from imagefiltering import _lib
    rng = np.random.default_rng(50 + ks[0] * 7 + ks[1])
    # column-major, so kernel axis 0 stays the fast image axis (a row-major
    # image transposes the kernel in the library's canonical layout)
    img = np.asfortranarray(rng.random((333, 290)).astype(np.float32))
    k = rng.standard_normal(ks)
    k /= np.abs(k).sum()
    first = [-(ks[0] // 2), -(ks[1] // 3)]
    for b in ["symmetric", "circular", IF.Fill(-1.0)]:
        out = _dense_lib(img, k, first, b if isinstance(b, IF.Fill) else IF.Pad(b), _lib.PREC_F16X3)
        want = O.imfilter_dense(img.astype(np.float64), k, first, obord(b))
        e = relerr(out, want)
        assert e <= TOL, (ks, b, e)


def test_dense_mfma_rejects_unsupported_shape():
    from imagefiltering import _lib
    img = np.asfortranarray(np.ones((64, 64), dtype=np.float32))
    with pytest.raises(NotImplementedError):
        _dense_lib(img, np.ones((3, 40)), [0, 0], IF.Pad("replicate"), _lib.PREC_F16X3)


def test_dense_mfma_nonfinite_tiles():
    """NaN and Inf reach exactly the outputs whose window holds them (tiles with
    a non-finite value take the direct FP32 path)."""
    rng = np.random.default_rng(60)
    img = rng.random((520, 400)).astype(np.float32)
    img[37, 211] = np.nan
    img[300, 5] = np.inf
    img[450, 390] = -np.inf
    k = norm_random_kernel(rng, 15)
    kern = IF.centered(k)
    out = IF.imfilter(img, kern, "replicate", IF.Algorithm.FIR())
    want = O.imfilter_dense(img.astype(np.float64), k, list(kern.first), "replicate")
    np.testing.assert_array_equal(np.isnan(out), np.isnan(want))
    np.testing.assert_array_equal(np.isinf(out), np.isinf(want))
    fin = np.isfinite(want)
    assert fin.sum() > 0.9 * want.size
    assert relerr(np.where(fin, out, 0), np.where(fin, want, 0)) <= TOL


def _abs_bound(img, k, first, border):
    """sum |k| |x| over each output's window: the scale the FP32 error of one
    output is measured against."""
    return O.imfilter_dense(np.abs(img.astype(np.float64)), np.abs(k), first, border)


@pytest.mark.parametrize("kscale,scales", [
    (1.0, [1e36, 1e-36, 3e-30, 1.0]),      # tile exponents beyond +-100 and inside
    (1e25, [1e9, 1e11]),                   # output scale 2^(e + ke) near and past 2^120
    (1e-25, [1e-10, 1e-12]),               # and near and past 2^-120
])
def test_dense_mfma_extreme_magnitudes(kscale, scales):
    """Images whose tiles span FP32's range.  Each 64 x 128 output tile takes
    the FP16 hi / lo split at its own power-of-two scale, or the direct FP32
    sum where that scale or the output's would leave the formats' range.
    An image of one magnitude keeps every output within 1e-5 of sum |k| |x|
    of the FP64 result; where magnitudes mix, an output is within 1e-5 of
    sum |k| times the largest |x| its tile reads (values more than 2^24 below
    that are beyond the split's reach), and nothing overflows."""
    from scipy.ndimage import maximum_filter
    from imagefiltering import _lib
    rng = np.random.default_rng(71)
    k = norm_random_kernel(rng, 15) * kscale
    first = [-7, -7]
    base = rng.random((512, 448)) + 0.5
    blocks = np.array(scales)[(np.arange(512)[:, None] // 128 + np.arange(448)[None, :] // 64) % len(scales)]
    img = np.asfortranarray((base * blocks).astype(np.float32))
    # a box around each output holding every input its tile reads (a tile
    # spans 64 outputs along axis 0 and 128 along axis 1, plus the 14-tap halo)
    tilemax = maximum_filter(np.abs(img.astype(np.float64)), size=(2 * 78 + 1, 2 * 142 + 1), mode="nearest")
    for b in ["replicate", IF.Fill(0.0)]:
        out = _dense_lib(img, k, first, b if isinstance(b, IF.Fill) else IF.Pad(b), _lib.PREC_F16X3)
        want = O.imfilter_dense(img.astype(np.float64), k, first, obord(b))
        assert np.all(np.isfinite(out)), (kscale, b)
        err = np.abs(out.astype(np.float64) - want) / (np.abs(k).sum() * tilemax)
        assert err.max() <= TOL, (kscale, b, float(err.max()))
    # one image per magnitude, through the explicit split and the default dispatch
    for s in scales:
        x = np.asfortranarray((base * s).astype(np.float32))
        out = _dense_lib(x, k, first, IF.Pad("symmetric"), _lib.PREC_F16X3)
        want = O.imfilter_dense(x.astype(np.float64), k, first, "symmetric")
        err = np.abs(out.astype(np.float64) - want) / _abs_bound(x, k, first, "symmetric")
        assert err.max() <= TOL, (kscale, s, float(err.max()))
    # The drop-in call first runs the reference's rank test, whose threshold
    # sqrt(eps) is absolute (src/imfilter.jl:1167-1174): taps of ~1e-25 have
    # every singular value below it, so the kernel applied is the rank-1
    # factor pair, as in the reference.
    from imagefiltering.imfilter import factorkernel
    fac = factorkernel(IF.OffsetArray(k, *first))
    assert len(fac) == (2 if kscale < 1e-10 else 1)
    kd = np.outer(fac[0].data.ravel(), fac[1].data.ravel()) if len(fac) == 2 else k
    for s in scales:
        x = np.asfortranarray((base * s).astype(np.float32))
        out = IF.imfilter(x, IF.OffsetArray(k, *first), "reflect", IF.Algorithm.FIR())
        want = O.imfilter_dense(x.astype(np.float64), kd, first, "reflect")
        err = np.abs(out.astype(np.float64) - want) / _abs_bound(x, kd, first, "reflect")
        assert err.max() <= TOL, (kscale, s, float(err.max()))


def test_dense_mfma_kernel_beyond_split_range():
    """Taps of magnitude 2^100 and beyond are outside the FP16 split's range:
    the explicit matrix-core precision refuses them and the default dispatch
    sums them in FP32."""
    from imagefiltering import _lib
    rng = np.random.default_rng(72)
    k = norm_random_kernel(rng, 15) * 1e31
    img = np.asfortranarray(rng.random((200, 180)).astype(np.float32))
    with pytest.raises(NotImplementedError):
        _dense_lib(img, k, [-7, -7], IF.Pad("replicate"), _lib.PREC_F16X3)
    out = IF.imfilter(img, IF.OffsetArray(k, -7, -7), "replicate", IF.Algorithm.FIR())
    want = O.imfilter_dense(img.astype(np.float64), k, [-7, -7], "replicate")
    err = np.abs(out.astype(np.float64) - want) / _abs_bound(img, k, [-7, -7], "replicate")
    assert err.max() <= TOL


def test_dense_mfma_bf16_storage_and_log():
    rng = np.random.default_rng(18)
    x = rng.random((200, 190)).astype(np.float32)
    xb = IF.bf16(x)
    xr = IF.bf16_to_float32(xb).astype(np.float64)
    kern = Kernel.LoG(2.0)
    out = IF.imfilter(IF.Float32, xb, kern, "reflect", IF.Algorithm.FIR())
    want = O.imfilter_dense(xr, kern.data, list(kern.first), "reflect")
    assert relerr(out, want) <= TOL
    out64 = IF.imfilter(x, kern, "reflect", IF.Algorithm.FIR())
    assert out64.dtype == np.float64
    want = O.imfilter_dense(x.astype(np.float64), kern.data, list(kern.first), "reflect")
    assert relerr(out64, want) <= TOL


@pytest.mark.parametrize("ks", [(3, 3, 3), (5, 5, 5), (3, 5, 3), (5, 3, 5)])
def test_dense_3d_register_blocked(ks):
    """The register-blocked 3-D stencils (3 or 5 taps per axis): volumes with
    interior and border tiles and several z chunks, kernels off centre, every
    border, F32 / F64 / UInt8 input, F32 and F64 output, a 4-D image."""
    rng = np.random.default_rng(500 + 25 * ks[0] + 5 * ks[1] + ks[2])
    k = rng.standard_normal(ks)
    first = [-(n // 3) for n in ks]
    kern = IF.OffsetArray(k, *first)
    vol = np.asfortranarray(rng.random((150, 70, 45)).astype(np.float32))
    for b in STYLES + [IF.Fill(0.5)]:
        out = IF.imfilter(vol, kern, b, IF.Algorithm.FIR())
        assert out.dtype == np.float64
        want = O.imfilter_dense(vol.astype(np.float64), k, first, obord(b))
        assert relerr(out, want) <= TOL, (ks, b)
    from imagefiltering import _lib
    out32 = _dense_lib(vol, k, first, IF.Pad("reflect"), _lib.PREC_FP32)
    assert relerr(out32, O.imfilter_dense(vol.astype(np.float64), k, first, "reflect")) <= TOL
    v64 = np.asfortranarray(rng.random((70, 33, 12)))
    assert relerr(IF.imfilter(v64, kern, "circular", IF.Algorithm.FIR()),
                  O.imfilter_dense(v64, k, first, "circular")) <= TOL
    codes = np.asfortranarray(rng.integers(0, 256, (66, 40, 9)).astype(np.uint8))
    assert relerr(IF.imfilter(codes, kern, "symmetric", IF.Algorithm.FIR()),
                  O.imfilter_dense(codes.astype(np.float64), k, first, "symmetric")) <= TOL
    res = IF.imfilter(vol, kern, IF.Inner(), IF.Algorithm.FIR())
    lo = [-f for f in first]
    size = [vol.shape[d] - ks[d] + 1 for d in range(3)]
    assert tuple(res.first) == tuple(lo) and res.data.shape == tuple(size)
    assert relerr(res.data, O.imfilter_dense(vol.astype(np.float64), k, first, "replicate", lo, size)) <= TOL
    v4 = np.asfortranarray(rng.random((60, 40, 20, 3)).astype(np.float32))
    k4 = IF.OffsetArray(k.reshape(ks + (1,)), *(first + [0]))
    assert relerr(IF.imfilter(v4, k4, "replicate", IF.Algorithm.FIR()),
                  O.imfilter_dense(v4.astype(np.float64), k.reshape(ks + (1,)), first + [0], "replicate")) <= TOL


def test_dense_3d_box_closed_form():
    img = np.ones((10, 10, 10))
    kern = IF.centered(np.full((3, 3, 3), 1 / 27))
    for b in STYLES:
        np.testing.assert_allclose(IF.imfilter(img, kern, b), img, rtol=1e-6)
    edge = np.array([2 / 3] + [1.0] * 8 + [2 / 3])
    want = edge[:, None, None] * edge[None, :, None] * edge[None, None, :]
    np.testing.assert_allclose(IF.imfilter(img, kern, IF.Fill(0)), want, rtol=1e-6)
    np.testing.assert_allclose(IF.imfilter(img, kern, IF.Fill(0), IF.Algorithm.FFT()), want, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("border", STYLES + ["fill"])
def test_dense_edge_targets(border):
    kern = np.array([[0.1, 0.2], [0.4, 0.5]])
    kernel = IF.OffsetArray(kern, range(-1, 1), range(1, 3))
    img = np.zeros((5, 7))
    img[0, 1] = 1
    b = IF.Fill(0.0) if border == "fill" else border
    want = O.imfilter_dense(img, kern, [-1, 1], obord(b))
    for alg in (IF.Algorithm.FIR(), IF.Algorithm.FIRTiled(), IF.Algorithm.FFT()):
        np.testing.assert_allclose(IF.imfilter(img, kernel, b, alg), want, atol=1e-6)


# ---- FFT -------------------------------------------------------------------------------------
@pytest.mark.parametrize("K", [15, 31])
def test_fft_matches_fir(K):
    rng = np.random.default_rng(19)
    img = rng.random((211, 187)).astype(np.float32)
    k = norm_random_kernel(rng, K)
    kern = IF.centered(k)
    for b in ["reflect", "symmetric", "circular", "replicate", IF.Fill(0.25)]:
        out = IF.imfilter(img, kern, b, IF.Algorithm.FFT())
        want = O.imfilter_dense(img.astype(np.float64), k, list(kern.first), obord(b))
        assert relerr(out, want) <= TOL, (K, b)
    # the default algorithm for a 15x15 kernel is FFT (maxlen > 30)
    assert IF.filter_algorithm(img, IF.factorkernel(kern)) == IF.Algorithm.FFT()


def test_fft_kernel_spectrum_reuse():
    # the library keeps the last kernel's spectrum: repeat it over different
    # images, switch kernels and padded sizes in between, and check every call
    rng = np.random.default_rng(25)
    k1, k2 = norm_random_kernel(rng, 15), norm_random_kernel(rng, 17)
    imgs = [rng.random(s).astype(np.float32) for s in [(120, 97), (120, 97), (64, 130), (120, 97)]]
    for i, (img, k) in enumerate(zip(imgs + imgs, [k1, k1, k1, k2, k1, k2, k2, k1])):
        kern = IF.centered(k)
        out = IF.imfilter(img, kern, "reflect", IF.Algorithm.FFT())
        want = O.imfilter_dense(img.astype(np.float64), k, list(kern.first), "reflect")
        assert relerr(out, want) <= TOL, i


def test_fft_1d_3d():
    rng = np.random.default_rng(20)
    v = rng.random(500)
    k = IF.centered(np.ones(31) / 31)
    out = IF.imfilter(v, k)            # FFT by default
    want = O.imfilter_dense(v, np.ones(31) / 31, [-15], "replicate")
    assert relerr(out, want) <= TOL
    vol = rng.random((24, 20, 18)).astype(np.float32)
    k3 = rng.standard_normal((5, 3, 7))
    out = IF.imfilter(vol, IF.centered(k3), "symmetric", IF.Algorithm.FFT())
    want = O.imfilter_dense(vol.astype(np.float64), k3, [-2, -1, -3], "symmetric")
    assert relerr(out, want) <= TOL


def _fft_check(img, k, kfirst, border, label):
    kern = IF.OffsetArray(k, *[range(f, f + n) for f, n in zip(kfirst, k.shape)])
    out = IF.imfilter(img, kern, border, IF.Algorithm.FFT())
    want = O.imfilter_dense(np.asarray(img, dtype=np.float64), k, list(kfirst), obord(border))
    assert out.shape == want.shape
    assert relerr(out, want) <= TOL, label


@pytest.mark.parametrize("order", ["C", "F"])
def test_fft_kernel_with_fewer_dims_than_image(order):
    # every line and plane is filtered when the kernel spans fewer axes than
    # the image: 41 taps run on the tiled path, 101 on the whole-array path
    # (the other axes are batch axes there)
    rng = np.random.default_rng(31)
    img = np.asarray(rng.random((70, 90)).astype(np.float32), order=order)
    for L in (41, 101):
        k = rng.random(L)
        k /= k.sum()
        for shape, first in (((L, 1), (-(L // 2), 0)), ((1, L), (0, -(L // 2)))):
            _fft_check(img, k.reshape(shape), first, "reflect", (order, L, shape))
        # the default choice for maxlen > 30 is FFT
        out = IF.imfilter(img, IF.centered(k.reshape(L, 1)), "replicate")
        want = O.imfilter_dense(img.astype(np.float64), k.reshape(L, 1), [-(L // 2), 0], "replicate")
        assert relerr(out, want) <= TOL
    vol = np.asarray(rng.random((13, 40, 37)).astype(np.float32), order=order)
    k15 = norm_random_kernel(rng, 15)
    for kshape, first in (((1, 15, 15), (0, -7, -7)), ((15, 15, 1), (-7, -7, 0)), ((15, 1, 15), (-7, 0, -7))):
        _fft_check(vol, k15.reshape(kshape), first, "symmetric", (order, kshape))
    kl = rng.standard_normal((1, 91, 5))
    _fft_check(vol, kl, (0, -45, -2), IF.Fill(0.5), (order, "91x5"))
    _fft_check(vol, kl.reshape(5, 91, 1), (-2, -45, 0), "circular", (order, "5x91"))


@pytest.mark.parametrize("border", STYLES + ["fill"])
def test_fft_kernel_offsets_beyond_centre(border):
    # kernels whose first tap lies at +2 or more: every output's window ends
    # past the image, so the padding must reach first + K - 1 beyond it
    rng = np.random.default_rng(32)
    b = IF.Fill(-0.5) if border == "fill" else border
    img = rng.random((61, 47)).astype(np.float32)
    for k, first in ((rng.standard_normal((3, 5)), (5, 2)), (rng.standard_normal((31, 4)), (2, -6)),
                     (rng.standard_normal((95, 3)), (5, 1)), (rng.standard_normal((3, 88)), (-1, 3))):
        _fft_check(img, k, first, b, (first, k.shape))
    v = rng.random(61)
    _fft_check(v, rng.standard_normal(3), (5,), b, "1-D")


def test_fft_long_axes():
    # padded lengths above 8192 take the two-level line transform (whole-array
    # path, kernels longer than the tiles take); the tiled path has no limit
    rng = np.random.default_rng(33)
    v = rng.random(9000)
    k = rng.random(101)
    k /= k.sum()
    _fft_check(v, k, (-50,), "reflect", "1-D 9000")
    img = rng.random((9000, 3)).astype(np.float32)
    _fft_check(img, k.reshape(101, 1), (-50, 0), "symmetric", "9000x3 axis 0")
    _fft_check(np.asfortranarray(img), k.reshape(101, 1), (-50, 0), IF.Fill(0.25), "9000x3 F axis 0")
    img2 = rng.random((2, 8300)).astype(np.float32)
    _fft_check(img2, rng.random((1, 41)), (0, -20), "replicate", "tiled 8300")


def test_fft_4d():
    rng = np.random.default_rng(34)
    a = rng.random((3, 4, 20, 22)).astype(np.float32)
    _fft_check(a, rng.standard_normal((1, 1, 9, 7)), (0, 0, -4, -3), "reflect", "4-D tiled")
    _fft_check(a, rng.standard_normal((1, 3, 5, 4)), (0, -1, -2, -2), "circular", "4-D whole-array")


def test_fft_nonfinite_gives_nan():
    # explicit Algorithm.FFT() on non-finite input: the transform spreads the
    # NaN over the whole output, as the reference's whole-array FFT does
    rng = np.random.default_rng(35)
    for bad in (np.nan, np.inf):
        img = rng.random((150, 170)).astype(np.float32)
        img[77, 12] = bad
        out = IF.imfilter(img, IF.centered(norm_random_kernel(rng, 15)), "reflect", IF.Algorithm.FFT())
        assert np.all(np.isnan(out))
    img = rng.random((150, 170)).astype(np.float32)
    out = IF.imfilter(img, IF.centered(norm_random_kernel(rng, 15)), "reflect", IF.Algorithm.FFT())
    assert np.all(np.isfinite(out))


@pytest.mark.parametrize("dtype", ["float32", "bfloat16", "n0f8", "float64"])
def test_fft_tiled_eltypes(dtype):
    rng = np.random.default_rng(36)
    x = rng.random((140, 133))
    if dtype == "bfloat16":
        img = IF.bf16(x.astype(np.float32))
        x64 = IF.eltypes.bf16_to_float32(img).astype(np.float64)
    elif dtype == "n0f8":
        codes = (x * 255).astype(np.uint8)
        img = IF.n0f8(codes)
        x64 = codes / 255.0
    else:
        img = x.astype(dtype)
        x64 = img.astype(np.float64)
    k = norm_random_kernel(rng, 17)
    out = IF.imfilter(img, IF.centered(k), "symmetric", IF.Algorithm.FFT())
    want = O.imfilter_dense(x64, k, [-8, -8], "symmetric")
    assert relerr(np.asarray(out, dtype=np.float64), want) <= TOL


# ---- IIR -------------------------------------------------------------------------------------
def iir_spec(kern):
    return [(f.Npre, f.data.a, f.data.b, f.data.scale, f.data.M) for f in kern]


@pytest.mark.parametrize("border", ["replicate", "circular", "symmetric", IF.Fill(0.0), IF.Fill(0.7)])
def test_iir_matches_oracle(border):
    rng = np.random.default_rng(21)
    vol = (rng.random((40, 33, 21)) + rng.normal(0, 0.05, (40, 33, 21))).astype(np.float32)
    kern = KernelFactors.IIRGaussian((2.0, 2.0, 1.5))
    out = IF.imfilter(vol, kern, border)
    assert out.dtype == np.float64
    # any Pad style initialises from the edge pixel (src/imfilter.jl:1030-1032)
    ob = ("fill", border.value) if isinstance(border, IF.Fill) else "replicate"
    want = O.imfilter_iir(vol.astype(np.float64), iir_spec(kern), ob)
    np.testing.assert_allclose(out, want, rtol=1e-9, atol=1e-12)


def test_iir_contiguous_lines_chunk_edges():
    # lines along the contiguous axis whose last three points straddle a chunk
    # boundary (n = 17, 18), the shortest accepted line (n = 4), a partial last
    # group of lines, F32 / F64 input, Pad and Fill borders
    rng = np.random.default_rng(23)
    for shape, dt in [((17, 300), np.float32), ((18, 5, 4), np.float64), ((4, 5), np.float32),
                      ((33, 7, 4), np.float32), ((35, 260), np.float64)]:
        x = rng.random(shape)
        x = x.astype(dt)
        x = np.asfortranarray(x)
        kern = KernelFactors.IIRGaussian(tuple(1.5 for _ in shape))
        for border in ["replicate", IF.Fill(0.25)]:
            out = IF.imfilter(x, kern, border)
            ob = ("fill", border.value) if isinstance(border, IF.Fill) else "replicate"
            want = O.imfilter_iir(x.astype(np.float64), iir_spec(kern), ob)
            np.testing.assert_allclose(out, want, rtol=1e-9, atol=1e-12, err_msg="%s %s" % (shape, border))


def test_iir_strided_lines_segments():
    # axes cut into segments of at most 32, two per wave: one segment
    # (n <= 32), one wave (33, 64), several waves, an odd count whose last
    # wave runs one segment (65, 97, 129, 200; last segment shorter), the
    # longest segmented axis (512) and the forms beyond it (513 strided, 600
    # contiguous), partial groups of 32 lines, long and short sigma, F32 / F64
    # input, contiguous and strided axes; 500 puts waves of full segments
    # (compile-time bounds) and the wave holding the short last segment in one
    # workgroup
    rng = np.random.default_rng(29)
    # (every axis is longer than 3: the reference rejects shorter ones)
    for shape, dt, sg in [((4, 65), np.float32, 2.0), ((70, 129), np.float64, 6.0), ((5, 64, 4), np.float32, 1.0),
                          ((4, 513), np.float32, 3.0), ((9, 200, 7), np.float64, 12.0), ((66, 512), np.float32, 2.0),
                          ((4, 4, 130), np.float32, 1.5), ((600, 5), np.float32, 2.0), ((512, 70), np.float64, 2.5),
                          ((33, 97, 5), np.float32, 2.0), ((97, 33), np.float64, 3.0), ((500, 31), np.float32, 4.0),
                          ((32, 63), np.float32, 1.2), ((5, 500), np.float64, 2.0), ((6, 500, 4), np.float32, 3.0)]:
        x = np.asfortranarray(rng.random(shape).astype(dt))
        kern = KernelFactors.IIRGaussian(tuple(sg for _ in shape))
        for border in ["replicate", IF.Fill(0.25)]:
            out = IF.imfilter(x, kern, border)
            ob = ("fill", border.value) if isinstance(border, IF.Fill) else "replicate"
            want = O.imfilter_iir(x.astype(np.float64), iir_spec(kern), ob)
            np.testing.assert_allclose(out, want, rtol=1e-9, atol=1e-12, err_msg="%s %s" % (shape, border))


def test_iir_per_dim_call_and_errors():
    imgf = np.zeros((5, 7))
    imgf[2, 3] = 1
    ret = IF.imfilter(imgf, KernelFactors.IIRGaussian((0, 5)), "replicate")
    out = np.zeros_like(ret)
    IF.imfilter_(IF.CPU1(IF.Algorithm.IIR()), out, imgf, KernelFactors.IIRGaussian(5), 2, "replicate")
    np.testing.assert_array_equal(out, ret)
    with pytest.raises(IF.DimensionMismatch):
        IF.imfilter(np.array([1.0, 2.0]), (KernelFactors.IIRGaussian(2.0),))


# ---- median / min / max / extrema -------------------------------------------------------------
@pytest.mark.parametrize("w", [3, 5, 7])
def test_median_u8_n0f8_bit_exact(w):
    rng = np.random.default_rng(22)
    codes = rng.integers(0, 256, (67, 93)).astype(np.uint8)
    sp = rng.random(codes.shape)
    codes[sp < 0.025] = 0
    codes[sp > 0.975] = 255
    h = w // 2
    for b in STYLES + [IF.Fill(0)]:
        want = O.median(codes.astype(np.float64), [-h, -h], [w, w], obord(b))
        got = IF.mapwindow(IF.median_, codes, (w, w), border=b)
        assert got.dtype == np.float64
        np.testing.assert_array_equal(got, want, err_msg=str(b))
        gotn = IF.mapwindow(IF.median_, IF.n0f8(codes), (w, w), border=b)
        assert gotn.dtype == np.float32
        np.testing.assert_array_equal(gotn, (want / 255).astype(np.float32))


def test_median_float_isless_and_generic_windows():
    rng = np.random.default_rng(23)
    a = rng.standard_normal((31, 26)).astype(np.float32)
    a[3, 4] = np.nan
    a[10, 10] = -0.0
    a[10, 11] = 0.0
    for w in [(3, 3), (5, 5), (7, 7), (2, 2), (1, 5), (4, 3)]:
        win = tuple(range(-(x // 2), x - x // 2) for x in w)
        for b in STYLES + [IF.Fill(0.5)]:
            try:
                want = O.median(a.astype(np.float64), [r.start for r in win], list(w), obord(b))
            except ZeroDivisionError:
                # reflect over a one-element window/image overlap: the reference's DivideError
                with pytest.raises(IF.ArgumentError):
                    IF.mapwindow(IF.median_, a, win, border=b)
                continue
            got = IF.mapwindow(IF.median_, a, win, border=b)
            np.testing.assert_array_equal(got, want.astype(np.float32),
                                          err_msg="%r %r" % (w, b))
            assert np.array_equal(np.signbit(got), np.signbit(want.astype(np.float32)))
    v = rng.standard_normal(40)
    np.testing.assert_array_equal(IF.mapwindow(IF.median_, v, 5), O.median(v, [-2], [5], "replicate"))


@pytest.mark.parametrize("et", ["bf16", "f16"])
def test_median_16bit_floats_packed_tiles(et):
    # 16-bit keys run two tiles per register (packed staging): several tiles
    # in both directions, NaN and signed zeros, every border
    rng = np.random.default_rng(24)
    x = rng.standard_normal((150, 141)).astype(np.float32)
    x[5, 7] = np.nan
    x[70, 71] = -0.0
    x[70, 72] = 0.0
    if et == "bf16":
        a = IF.bf16(x)
        xv = IF.bf16_to_float32(a)
    else:
        a = x.astype(np.float16)
        xv = a.astype(np.float32)
    for w in (3, 5, 7):
        h = w // 2
        for b in STYLES + [IF.Fill(0.5)]:
            want = O.median(xv.astype(np.float64), [-h, -h], [w, w], obord(b)).astype(np.float32)
            got = IF.mapwindow(IF.median_, a, (w, w), border=b)
            gv = IF.bf16_to_float32(got) if et == "bf16" else np.asarray(got).astype(np.float32)
            np.testing.assert_array_equal(gv, want, err_msg="%d %r" % (w, b))
            assert np.array_equal(np.signbit(gv), np.signbit(want))


def test_median_reference_vectors():
    b = np.array([1, 100, 1, 2, -1000, 2], dtype=np.float64)
    np.testing.assert_array_equal(IF.mapwindow(IF.median_, b, range(-1, 2)), [1, 1, 2, 1, 2, 2])
    np.testing.assert_array_equal(IF.mapwindow(IF.median, b, range(-2, 3)), [1, 1, 1, 2, 2, 2])
    A = np.array([[1, 5, -2, 3, 7], [2, 0, 3, 4, 4], [3, 3, 6, 2, 5], [1, -3, 5, 3, 0]], dtype=np.float64)
    np.testing.assert_array_equal(IF.mapwindow(IF.median_, A, (3, 3)),
                                  [[1, 1, 3, 3, 4], [2, 3, 3, 4, 4], [2, 3, 3, 4, 4], [1, 3, 3, 3, 2]])


@pytest.mark.parametrize("w", [3, 9])
def test_minmax_bit_exact(w):
    rng = np.random.default_rng(24)
    codes = rng.integers(0, 256, (71, 53)).astype(np.uint8)
    h = w // 2
    for b in ["replicate", "symmetric", "circular", IF.Fill(17)]:
        mn, mx = O.minmax(codes.astype(np.float64), [-h, -h], [w, w], obord(b))
        np.testing.assert_array_equal(IF.mapwindow(IF.minimum, codes, (w, w), border=b), mn)
        np.testing.assert_array_equal(IF.mapwindow(IF.maximum, codes, (w, w), border=b), mx)
    f = rng.standard_normal((33, 41)).astype(np.float32)
    mn, mx = O.minmax(f.astype(np.float64), [-h, -h], [w, w], "replicate")
    np.testing.assert_array_equal(IF.mapwindow(IF.minimum, f, (w, w)), mn.astype(np.float32))
    np.testing.assert_array_equal(IF.mapwindow(IF.maximum, f, (w, w)), mx.astype(np.float32))


@pytest.mark.parametrize("dtype", ["u8", "n0f8", "f32", "bf16"])
def test_minmax_tiled_many_tiles(dtype):
    """The one-pass 2-D min/max (several tiles and planes, square and
    non-square windows, compile-time and run-time window paths, NaN)."""
    rng = np.random.default_rng(27)
    codes = rng.integers(0, 256, (301, 170, 3)).astype(np.uint8)
    if dtype in ("u8", "n0f8"):
        img = IF.n0f8(codes) if dtype == "n0f8" else codes
        x64 = codes.astype(np.float64)
    else:
        f = rng.standard_normal(codes.shape).astype(np.float32)
        f[5, 7, 0] = np.nan
        f[200, 100, 2] = -0.0
        img = IF.bf16(f) if dtype == "bf16" else f
        x64 = (IF.bf16_to_float32(img) if dtype == "bf16" else f).astype(np.float64)
    for win in [(3, 3), (9, 9), (5, 7), (1, 11), (33, 3)]:
        wf = [-(w // 2) for w in win] + [0]
        for b in ["replicate", "reflect", IF.Fill(0)]:
            mn, mx = O.minmax(x64, wf, list(win) + [1], obord(b))
            gmn = IF.mapwindow(IF.minimum, img, win + (1,), border=b)
            gmx = IF.mapwindow(IF.maximum, img, win + (1,), border=b)
            if dtype == "bf16":
                gmn, gmx = IF.bf16_to_float32(gmn), IF.bf16_to_float32(gmx)
            gmn = np.asarray(gmn).astype(np.float64)  # N0f8 results are their uint8 codes
            gmx = np.asarray(gmx).astype(np.float64)
            np.testing.assert_array_equal(gmn, mn, err_msg=str((dtype, win, b)))
            np.testing.assert_array_equal(gmx, mx, err_msg=str((dtype, win, b)))
            if dtype == "f32":
                assert np.signbit(gmn[200, 100, 2]) == np.signbit(mn[200, 100, 2])


@pytest.mark.parametrize("w", [3, 5, 7, 9])
def test_minmax_u8_packed_tiles(w):
    """The packed 8-bit min/max tiles (UInt8 / N0f8 codes, square windows 3 ..
    9): rows of 4-byte multiples (word staging) and not, output rows of
    16-byte multiples (vector stores) and not, several planes, off-centre
    windows, Fill, minimum and maximum alone, and extrema (both at once)."""
    rng = np.random.default_rng(400 + w)
    for shape in [(300, 170, 2), (1024, 40), (600, 70), (135, 66), (4, 9)]:
        codes = rng.integers(0, 256, shape).astype(np.uint8)
        codes[rng.random(shape) < 0.03] = 0
        extra = [0] * (len(shape) - 2)
        for wf in ([-(w // 2), -(w // 2)], [-(w - 1), 0], [0, -(w - 1)]):
            window = tuple(range(f, f + w) for f in wf) + tuple(range(0, 1) for _ in extra)
            centred = wf[0] == -(w // 2) and wf[1] == -(w // 2)
            borders = ["replicate", IF.Fill(17)] + (["reflect"] if centred and min(shape[:2]) > w else [])
            for b in borders:
                mn, mx = O.minmax(codes.astype(np.float64), wf + extra, [w, w] + [1] * len(extra), obord(b))
                for f, want in ((IF.minimum, mn), (IF.maximum, mx)):
                    got = np.asarray(IF.mapwindow(f, codes, window, border=b))
                    np.testing.assert_array_equal(got, want.astype(np.uint8), err_msg=str((shape, wf, b, f)))
                if b == "replicate":
                    got = np.asarray(IF.mapwindow(IF.maximum, IF.n0f8(codes), window, border=b)).view(np.uint8)
                    np.testing.assert_array_equal(got, mx.astype(np.uint8), err_msg=str((shape, wf, "n0f8")))
        mn, mx = O.extrema_lemire(codes, [w, w] + [1] * len(extra))
        got = IF.mapwindow(IF.extrema, codes, (w, w) + (1,) * len(extra))
        np.testing.assert_array_equal(np.asarray(got[..., 0]), mn.astype(np.uint8), str(shape))
        np.testing.assert_array_equal(np.asarray(got[..., 1]), mx.astype(np.uint8), str(shape))
    # outputs of a sub-range (unit-step indices) and of the interior (Inner())
    h = w // 2
    codes = rng.integers(0, 256, (1024, 40)).astype(np.uint8)
    c64 = codes.astype(np.float64)
    inds = (range(5, 205), range(3, 34))
    got = np.asarray(IF.mapwindow(IF.minimum, codes, (w, w), indices=inds))
    want = O.minmax(c64, [-h, -h], [w, w], "replicate", [5, 3], [200, 31])[0]
    np.testing.assert_array_equal(got, want.astype(np.uint8))
    inner = IF.mapwindow(IF.maximum, codes, (w, w), border=IF.Inner())
    want = O.minmax(c64, [-h, -h], [w, w], "replicate", [h, h], [1024 - 2 * h, 40 - 2 * h])[1]
    np.testing.assert_array_equal(np.asarray(inner.data), want.astype(np.uint8))


def test_extrema_lemire_semantics():
    import json, os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))
    for c in g["extrema_1d"]["cases"]:
        got = IF.mapwindow(IF.extrema, np.array(c["input"]), c["window"])
        np.testing.assert_array_equal(got, np.array(c["expect"]), err_msg=str(c))
    rng = np.random.default_rng(25)
    for window in [(2, 2), (2, 3), (3, 3), (2, 5), (9, 9), (2, 5, 3)]:
        A = rng.random((11,) * len(window)).astype(np.float32)
        mn, mx = O.extrema_lemire(A.astype(np.float64), list(window))
        got = IF.mapwindow(IF.extrema, A, window)
        np.testing.assert_array_equal(got[..., 0], mn.astype(np.float32))
        np.testing.assert_array_equal(got[..., 1], mx.astype(np.float32))


@pytest.mark.parametrize("dtype", ["u8", "n0f8"])
def test_extrema_integer_types(dtype):
    """extrema_filter on UInt8 and N0f8 (the order-filter benchmark form):
    the reference's 1-D vectors as codes, and 2-D / 3-D windows, odd and
    even, against the Lemire oracle; results stay in the input eltype."""
    import json, os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))
    wrap = IF.n0f8 if dtype == "n0f8" else (lambda c: c)
    for c in g["extrema_1d"]["cases"]:
        # the vectors' values in tenths, as codes (x -> 50 x keeps their order)
        x = np.rint(np.asarray(c["input"]) * 50).astype(np.uint8)
        want = np.rint(np.asarray(c["expect"]) * 50).astype(np.uint8)
        got = IF.mapwindow(IF.extrema, wrap(x), c["window"])
        np.testing.assert_array_equal(np.asarray(got).view(np.uint8), want, err_msg=str(c))
    rng = np.random.default_rng(26)
    codes = rng.integers(0, 256, (301, 170)).astype(np.uint8)
    codes[rng.random(codes.shape) < 0.05] = 255
    for window in [(3, 3), (9, 9), (2, 5), (8, 3), (1, 17)]:
        mn, mx = O.extrema_lemire(codes, list(window))
        got = IF.mapwindow(IF.extrema, wrap(codes), window)
        if dtype == "n0f8":
            assert isinstance(got, IF.N0f8Array)
        else:
            assert got.dtype == np.uint8
        np.testing.assert_array_equal(np.asarray(got[..., 0]).view(np.uint8), mn.astype(np.uint8), str(window))
        np.testing.assert_array_equal(np.asarray(got[..., 1]).view(np.uint8), mx.astype(np.uint8), str(window))
    c3 = rng.integers(0, 256, (40, 33, 9)).astype(np.uint8)
    mn, mx = O.extrema_lemire(c3, [3, 4, 3])
    got = IF.mapwindow(IF.extrema, wrap(c3), (3, 4, 3))
    np.testing.assert_array_equal(np.asarray(got[..., 0]).view(np.uint8), mn.astype(np.uint8))
    np.testing.assert_array_equal(np.asarray(got[..., 1]).view(np.uint8), mx.astype(np.uint8))


def test_cascade_fill_non_unit_factors():
    """Cascades that run as separate passes (an axis-2 factor too long for the
    fused 3-D pass, an identity factor on axis 1, UInt8 input) with
    Fill(0.7) (Fill(7) for UInt8) and sobel factors whose sums are not 1 (0 for the derivative),
    so the fill value each later pass sees is the filtered fill, not 0.7."""
    rng = np.random.default_rng(43)
    vol = np.asfortranarray(rng.random((40, 31, 37)).astype(np.float32))
    zt = np.linspace(0.1, 0.3, 11)                                  # sum 2.2, 11 taps
    long_z = IF.OffsetArray(zt.reshape(1, 1, 11), 0, 0, -4)
    cases = []
    for d in range(3):
        sob = KernelFactors.sobel((True, True, True), d)
        cases.append((tuple(sob) + (long_z,), factors_of(sob) + [(2, zt, -4)]))
    for d in (0, 2):   # an identity factor on axis 1
        sob = KernelFactors.sobel((True, False, True), d)
        cases.append((sob, factors_of(sob)))
    for kern, fs in cases:
        for b in [IF.Fill(0.7), "reflect"]:
            out = IF.imfilter(vol, kern, b)
            want = O.imfilter_factors(vol.astype(np.float64), fs, obord(b))
            assert relerr(out, want) <= TOL, (len(kern), b)
    codes = np.asfortranarray(rng.integers(0, 256, (30, 26, 12)).astype(np.uint8))
    kern = KernelFactors.sobel((True, True, True), 1)
    out = IF.imfilter(codes, kern, IF.Fill(7))
    want = O.imfilter_factors(codes.astype(np.float64), factors_of(kern), ("fill", 7.0))
    assert relerr(out, want) <= TOL
    # the fill value converts to the image's eltype first (src/borderarray.jl:11-20)
    with pytest.raises(IF.ArgumentError):
        IF.imfilter(codes, kern, IF.Fill(0.7))


@pytest.mark.parametrize("border", STYLES + [IF.Fill(0.0), IF.Fill(0.7)])
def test_cascade_reference_identities(border):
    """test/cascade.jl:4-39: a tuple of factors equals the dense kernel of
    their convolution, for repeated factors along one axis and factors in
    any order, every border (here also Fill(0.7))."""
    rng = np.random.default_rng(44)
    a1 = rng.random(15)
    kern = IF.OffsetArray(np.ones(3), -1)
    kern2 = IF.OffsetArray(np.array([1.0, 2, 3, 2, 1]), -2)
    np.testing.assert_allclose(IF.imfilter(a1, (kern, kern), border), IF.imfilter(a1, kern2, border),
                               rtol=1e-6, atol=1e-6)
    a = rng.random((15, 15))
    kx, ky = IF.OffsetArray(np.ones((3, 1)), -1, 0), IF.OffsetArray(np.ones((1, 3)), 0, -1)
    c = np.array([1.0, 2, 3, 2, 1])
    k2 = IF.OffsetArray(np.outer(c, c), -2, -2)
    k2x = IF.OffsetArray(np.outer(c, np.ones(3)), -2, -1)
    k2y = IF.OffsetArray(np.outer(np.ones(3), c), -1, -2)
    for fac, dense in [((kx, ky, kx, ky), k2), ((kx, kx, ky, ky), k2), ((kx, kx, ky), k2x), ((ky, kx, ky), k2y)]:
        np.testing.assert_allclose(IF.imfilter(a, fac, border), IF.imfilter(a, dense, border), rtol=1e-6,
                                   atol=1e-5, err_msg=str(border))


def test_errors_from_the_library():
    with pytest.raises(IF.ArgumentError):
        IF.padarray(np.zeros((1, 4)), IF.Pad("reflect", (1, 1), (1, 1)))
    with pytest.raises(IF.DimensionMismatch):
        IF.imfilter(np.zeros((3, 9)), KernelFactors.IIRGaussian((2.0, 2.0)))


def test_device_resident_roundtrip():
    rng = np.random.default_rng(26)
    img = rng.random((128, 96)).astype(np.float32)
    d = IF.DeviceArray.from_host(img)
    kern = KernelFactors.gaussian((2.0, 2.0))
    out = IF.imfilter(IF.Float32, d, kern, "symmetric")
    assert isinstance(out, IF.DeviceArray)
    want = O.imfilter_factors(img.astype(np.float64), factors_of(kern), "symmetric")
    assert relerr(out.to_host(), want) <= TOL


# ---- filter_algorithm's isfinite scan on device arrays ----------------------------------
@pytest.mark.parametrize("dtype", ["float32", "float64", "bfloat16", "float16", "uint8"])
def test_all_finite_device_scan(dtype):
    from imagefiltering.imfilter import _device_all_finite
    rng = np.random.default_rng(81)
    base = rng.random(10007)   # odd length: a tail after the 4-wide loads
    for pos in (None, 0, 1, 5002, 10003, 10006):
        for bad in (np.nan, np.inf, -np.inf):
            x = base.copy()
            if pos is not None:
                x[pos] = bad
            if dtype == "bfloat16":
                host = IF.bf16(x.astype(np.float32))
            elif dtype == "uint8":
                host = (np.nan_to_num(x, nan=0, posinf=1, neginf=0) * 255).astype(np.uint8)
            else:
                host = x.astype(dtype)
            d = IF.DeviceArray.from_host(host)
            want = dtype == "uint8" or pos is None
            assert _device_all_finite(d) == want, (dtype, pos, bad)
            if pos is None:
                break
    # the default algorithm then takes FIR instead of FFT, as src/imfilter.jl:1200-1209
    img = rng.random((64, 64)).astype(np.float32)
    img[3, 4] = np.nan
    k = IF.centered(rng.standard_normal((7, 7)))   # 49 taps, not separable: FFT when finite
    assert IF.filter_algorithm(IF.DeviceArray.from_host(img), IF.factorkernel(k)) == IF.Algorithm.FIR()
    img[3, 4] = 0.5
    assert IF.filter_algorithm(IF.DeviceArray.from_host(img), IF.factorkernel(k)) == IF.Algorithm.FFT()
