"""Full-size parity: each benchmark workload (4096^2 dense 2-D, 512x512x256
volume, 8192^2 N0f8 order filters) run whole on the GPU and checked against
the oracle on bands that cover every edge and the middle.  The oracle takes
the whole image (its border maps see the full extents) and computes only the
band's outputs, so every checked output is compared with its true value.

Tolerances: 1e-5 normwise relative to the FP64 result for floating-point
FIR / FFT, rtol 1e-9 for the FP64 IIR recursion, bit-exact for order
filters."""
import numpy as np
import pytest

import imagefiltering as IF
from imagefiltering import Kernel, KernelFactors
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def relerr(y, y64):
    y = np.asarray(y, dtype=np.float64)
    y64 = np.asarray(y64, dtype=np.float64)
    scale = np.max(np.abs(y64))
    return float(np.max(np.abs(y - y64)) / (scale if scale > 0 else 1.0))


def bands2(n0, n1, w=64):
    """(first, size) regions: top, middle and bottom along axis 0, and the
    last columns along axis 1."""
    return [([0, 0], [w, n1]), ([n0 // 2 - w // 2, 0], [w, n1]), ([n0 - w, 0], [w, n1]),
            ([0, n1 - w], [n0, w])]


def cut(a, first, size):
    return a[tuple(slice(f, f + s) for f, s in zip(first, size))]


@pytest.fixture(scope="module")
def img4096():
    rng = np.random.default_rng(20261201)
    return rng.random((4096, 4096), dtype=np.float64).astype(np.float32)


def _k_sum1(rng, K):
    while True:
        k = rng.standard_normal((K, K))
        if abs(k.sum()) >= 1:
            return k / k.sum()


@pytest.mark.parametrize("K,alg", [(31, "fir"), (31, "fft"), (15, "fft")])
def test_dense_4096_fp32(img4096, K, alg):
    """4096^2 Float32, KxK N(0,1) taps normalised to sum 1, "reflect", as
    imfilter(Float32, img, centered(k), "reflect", alg)."""
    rng = np.random.default_rng(1000 + K)
    k = _k_sum1(rng, K)
    kern = IF.centered(k)
    a = IF.Algorithm.FIR() if alg == "fir" else IF.Algorithm.FFT()
    out = IF.imfilter(IF.Float32, img4096, kern, "reflect", a)
    assert out.shape == img4096.shape and out.dtype == np.float32
    for first, size in bands2(*img4096.shape):
        want = O.imfilter_dense(img4096, k, list(kern.first), "reflect", first, size)
        e = relerr(cut(out, first, size), want)
        assert e <= TOL, (K, alg, first, e)


@pytest.mark.parametrize("K", [15, 31])
def test_dense_4096_bf16_storage(img4096, K):
    """The same workload with BFloat16-stored pixels (FP32 accumulation),
    against the FP64 result on the BF16 values."""
    rng = np.random.default_rng(2000 + K)
    k = _k_sum1(rng, K)
    kern = IF.centered(k)
    xb = IF.bf16(img4096)
    x = IF.bf16_to_float32(xb)
    out = IF.imfilter(IF.Float32, xb, kern, "reflect", IF.Algorithm.FIR())
    for first, size in bands2(*x.shape):
        want = O.imfilter_dense(x, k, list(kern.first), "reflect", first, size)
        e = relerr(cut(out, first, size), want)
        assert e <= TOL, (K, first, e)


def test_sobel_log_4096(img4096):
    """Kernel.sobel (both directions, factored) and Kernel.LoG(2.0) at 4096^2."""
    for kern in list(Kernel.sobel()) + [Kernel.LoG(2.0)]:
        out = IF.imfilter(img4096, kern, "reflect")
        assert out.dtype == np.float64
        for first, size in bands2(*img4096.shape):
            want = O.imfilter_dense(img4096, kern.data, list(kern.first), "reflect", first, size)
            e = relerr(cut(out, first, size), want)
            assert e <= TOL, (kern.shape, first, e)


@pytest.fixture(scope="module")
def vol512():
    rng = np.random.default_rng(20261202)
    v = rng.random((512, 512, 256)) + rng.normal(0, 0.05, (512, 512, 256))
    return np.asfortranarray(v.astype(np.float32))


@pytest.mark.parametrize("border", ["symmetric", "circular"])
def test_volume_fir_gaussian(vol512, border):
    """512x512x256, KernelFactors.gaussian((2.0, 2.0, 1.5)) via FIR, checked
    on 16-plane slabs at both z ends and in the middle, and the x / y faces."""
    kern = KernelFactors.gaussian((2.0, 2.0, 1.5))
    fs = [(f.Npre, f.data.data.astype(np.float64), f.data.first[0]) for f in kern]
    out = IF.imfilter(vol512, kern, border, IF.Algorithm.FIR())
    assert out.shape == vol512.shape and out.dtype == np.float64
    regions = [([0, 0, 0], [512, 512, 16]), ([0, 0, 120], [512, 512, 16]), ([0, 0, 240], [512, 512, 16]),
               ([0, 0, 0], [8, 512, 256]), ([504, 0, 0], [8, 512, 256]), ([0, 0, 0], [512, 8, 256]),
               ([0, 504, 0], [512, 8, 256])]
    for first, size in regions:
        want = O.imfilter_factors(vol512, fs, border, first, size)
        e = relerr(cut(out, first, size), want)
        assert e <= TOL, (border, first, e)


def test_volume_iir_gaussian(vol512):
    """512x512x256, KernelFactors.IIRGaussian((2.0, 2.0, 1.5)) (Triggs-Sdika,
    FP64 recursion) with the "circular" border, which the IIR initialises
    from the edge pixel like every Pad style; the whole volume is checked."""
    kern = KernelFactors.IIRGaussian((2.0, 2.0, 1.5))
    spec = [(f.Npre, f.data.a, f.data.b, f.data.scale, f.data.M) for f in kern]
    out = IF.imfilter(vol512, kern, "circular")
    assert out.dtype == np.float64
    want = O.imfilter_iir(vol512, spec, "replicate")
    np.testing.assert_allclose(out, want, rtol=1e-9, atol=1e-12)


@pytest.fixture(scope="module")
def codes8192():
    rng = np.random.default_rng(20261203)
    c = rng.integers(0, 256, (8192, 8192), dtype=np.uint8)
    sp = rng.random(c.shape, dtype=np.float32)
    c[sp < 0.025] = 0
    c[sp > 0.975] = 255
    return c


@pytest.mark.parametrize("w", [5, 7])
def test_median_8192_n0f8(codes8192, w):
    """mapwindow(median!, img::N0f8, (w, w)) on 8192^2, "replicate": Float32
    results equal to the exact median codes / 255 on every band."""
    h = w // 2
    out = IF.mapwindow(IF.median_, IF.n0f8(codes8192), (w, w))
    assert out.dtype == np.float32
    for first, size in bands2(8192, 8192, 32):
        want = O.median(codes8192, [-h, -h], [w, w], "replicate", first, size)
        np.testing.assert_array_equal(cut(out, first, size), (want / 255).astype(np.float32),
                                      err_msg=str((w, first)))


@pytest.mark.parametrize("w", [3, 9])
def test_extrema_minmax_8192_n0f8(codes8192, w):
    """The order filters of the same image: extrema (truncated windows) and
    minimum / maximum with "replicate", bit-exact on bands."""
    img = IF.n0f8(codes8192)
    got = IF.mapwindow(IF.extrema, img, (w, w))
    m = w  # rows beyond the band that its truncated windows can reach
    for r0 in (0, 4096, 8192 - 64):
        lo, hi = max(r0 - m, 0), min(r0 + 64 + m, 8192)
        mn, mx = O.extrema_lemire(codes8192[lo:hi], [w, w])
        np.testing.assert_array_equal(np.asarray(got[r0:r0 + 64, :, 0]).view(np.uint8),
                                      mn[r0 - lo:r0 - lo + 64].astype(np.uint8), err_msg=str((w, r0)))
        np.testing.assert_array_equal(np.asarray(got[r0:r0 + 64, :, 1]).view(np.uint8),
                                      mx[r0 - lo:r0 - lo + 64].astype(np.uint8), err_msg=str((w, r0)))
    h = w // 2
    gmn = IF.mapwindow(IF.minimum, img, (w, w))
    for first, size in bands2(8192, 8192, 32):
        mn, _ = O.minmax(codes8192, [-h, -h], [w, w], "replicate", first, size)
        np.testing.assert_array_equal(np.asarray(cut(gmn, first, size)).view(np.uint8), mn.astype(np.uint8),
                                      err_msg=str((w, first)))
