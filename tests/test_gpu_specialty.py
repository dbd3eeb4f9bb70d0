"""GPU parity of the callers one step above imfilter: imgradients
(test/gradient.jl), the Laplacian stencil (test/specialty.jl), findlocalmaxima
/ findlocalminima and blob_LoG / multiLoG (test/extrema.jl)."""
import numpy as np
import pytest

import imagefiltering as IF
from imagefiltering import Kernel, KernelFactors
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GRADS = [KernelFactors.ando3, KernelFactors.sobel, KernelFactors.prewitt, KernelFactors.ando4, KernelFactors.ando5,
         KernelFactors.bickley, KernelFactors.scharr]


@pytest.mark.parametrize("kf", GRADS, ids=lambda f: f.__name__)
def test_imgradients_on_ramps_gradient_jl(kf):
    # test/gradient.jl:4-24: ramps along each axis have unit gradient there, zero elsewhere
    y = np.arange(1, 6.0)[:, None] * np.ones((1, 7))
    x = np.ones((5, 1)) * np.arange(1, 8.0)[None, :]
    for img, ey, ex in ((y, 1, 0), (x, 0, 1)):
        gy, gx = IF.imgradients(img, kf, IF.Inner())
        assert np.all(np.abs(np.asarray(gy) - ey) < 1e-4)
        assert np.all(np.abs(np.asarray(gx) - ex) < 1e-4)
        gy, gx = IF.imgradients(img, kf, IF.Pad("replicate"))
        assert gy.shape == gx.shape == img.shape


def test_imgradients_3d_and_vs_kernel_pairs():
    y, x, z = np.arange(1, 6.0), np.arange(1, 8.0), np.arange(1, 7.0)
    vol = y[:, None, None] * np.ones((1, 7, 6))
    for kf in (KernelFactors.sobel, KernelFactors.scharr, KernelFactors.bickley):
        g = IF.imgradients(vol, kf, IF.Inner())
        for d, e in enumerate((1, 0, 0)):
            assert np.all(np.abs(np.asarray(g[d]) - e) < 1e-4)
    # test/gradient.jl:26-37: Kernel.* pairs equal the factored gradients
    rng = np.random.default_rng(101)
    img = rng.random((40, 33))
    for fk, fkf in ((Kernel.sobel, KernelFactors.sobel), (Kernel.prewitt, KernelFactors.prewitt),
                    (Kernel.scharr, KernelFactors.scharr), (Kernel.bickley, KernelFactors.bickley)):
        ky, kx = fk()
        gy, gx = IF.imgradients(img, fkf)
        np.testing.assert_allclose(IF.imfilter(img, ky), gy, atol=1e-6)
        np.testing.assert_allclose(IF.imfilter(img, kx), gx, atol=1e-6)
    # on device arrays each gradient stays on the device
    d = IF.DeviceArray.from_host(img.astype(np.float32))
    gy, gx = IF.imgradients(d, KernelFactors.sobel, "reflect")
    want = O.imfilter_factors(img.astype(np.float32).astype(np.float64),
                              [(f.Npre, f.data.data, f.data.first[0]) for f in KernelFactors.sobel((True, True), 0)],
                              "reflect")
    np.testing.assert_allclose(gy.to_host(), want, atol=2e-6)


def test_laplacian_impulses_specialty_jl():
    # test/specialty.jl:27-120 for the float eltypes
    a = np.zeros(5)
    a[2] = 1
    np.testing.assert_array_equal(IF.imfilter(a, Kernel.Laplacian((True,))), [0, 1, -2, 1, 0])
    a = np.zeros((5, 5))
    a[2, 2] = 1
    want = np.array([[0, 0, 0, 0, 0], [0, 0, 1, 0, 0], [0, 1, -4, 1, 0], [0, 0, 1, 0, 0], [0, 0, 0, 0, 0]], float)
    np.testing.assert_array_equal(IF.imfilter(a, Kernel.Laplacian()), want)
    np.testing.assert_array_equal(IF.imfilter(a.astype(np.float32), Kernel.Laplacian()), want.astype(np.float32))
    wv = np.zeros((5, 5))
    wv[1:4, 2] = [1, -2, 1]
    np.testing.assert_array_equal(IF.imfilter(a, Kernel.Laplacian((True, False))), wv)
    rng = np.random.default_rng(102)
    img = rng.random((64, 48))
    for b in ("replicate", "circular", "symmetric", "reflect"):
        got = IF.imfilter(img, Kernel.Laplacian(), b)
        want = O.imfilter_dense(img, np.array([[0, 1, 0], [1, -4, 1], [0, 1, 0]], float), [-1, -1], b)
        np.testing.assert_allclose(got, want, atol=1e-6)


def test_findlocalextrema_extrema_jl():
    A = np.zeros((9, 9))
    A[[0, 1, 4], 4] = 1
    assert IF.findlocalmaxima(A) == [(4, 4)]
    assert IF.findlocalmaxima(A, window=(1, 3)) == [(0, 4), (1, 4), (4, 4)]
    assert IF.findlocalmaxima(A, window=(1, 3), edges=False) == [(1, 4), (4, 4)]
    A3 = np.zeros((9, 9, 9))
    A3[[0, 1, 4], 4, 4] = 1
    assert IF.findlocalmaxima(A3) == [(4, 4, 4)]
    assert IF.findlocalmaxima(A3, window=(1, 3, 1)) == [(0, 4, 4), (1, 4, 4), (4, 4, 4)]
    assert IF.findlocalmaxima(A3, window=(1, 3, 1), edges=False) == [(1, 4, 4), (4, 4, 4)]
    assert IF.findlocalminima(-A) == [(4, 4)]
    # against a brute-force scan on random data (ties excluded by the strict order)
    rng = np.random.default_rng(103)
    R = rng.integers(0, 6, (30, 41)).astype(np.float32)
    want = []
    for j in range(41):
        for i in range(30):
            nb = R[max(i - 1, 0):i + 2, max(j - 1, 0):j + 2]
            if (nb < R[i, j]).sum() == nb.size - 1:
                want.append((i, j))
    assert IF.findlocalmaxima(R) == want


def test_blob_log_extrema_jl():
    A = np.zeros((9, 9))
    A[4, 4] = 1
    blobs = IF.blob_LoG(A, 2.0 ** np.array([0.5, 0, 1]))
    assert len(blobs) == 1
    assert abs(blobs[0].amplitude - 0.3183098861837907) < 1e-6
    assert blobs[0].sigma == (1.0, 1.0)
    assert blobs[0].location == (4, 4)
    assert len(IF.blob_LoG(A, [1.0])) == 1
    assert len(IF.blob_LoG(A, [1.0], edges=(True, False, False))) == 1
    assert IF.blob_LoG(A, [1.0], edges=False) == []
    A = np.zeros((9, 9))
    A[0, 4] = 1
    assert all(b.amplitude < 1e-6 for b in IF.blob_LoG(A, 2.0 ** np.array([0.5, 0, 1])))
    blobs = [b for b in IF.blob_LoG(A, 2.0 ** np.array([0.5, 0, 1]), edges=True) if b.amplitude > 0.1]
    assert len(blobs) == 1 and blobs[0].location == (0, 4)
    assert IF.blob_LoG(A, 2.0 ** np.array([0, 1]), edges=(False, True, False)) == []
    A3 = np.zeros((9, 9, 9))
    A3[4, 4, 4] = 1
    blobs = IF.blob_LoG(A3, 2.0 ** np.array([0.5, 0, 1]))
    assert len(blobs) == 1 and blobs[0].location == (4, 4, 4)
    A3 = np.zeros((9, 9, 9))
    A3[4, 3:6, 4] = 1
    blobs = IF.blob_LoG(A3, 2.0 ** np.array([1.0, 0, 0.5]), sigma_shape=(1.0, 3.0, 1.0))
    assert len(blobs) == 1 and blobs[0].location == (4, 4, 4)
    v = np.concatenate([np.zeros(10), [1.0, 0.0]])
    assert len(IF.blob_LoG(v, [4], edges=True, rthresh=0)) > len(IF.blob_LoG(v, [4], edges=True))


def test_multilog_matches_oracle():
    rng = np.random.default_rng(104)
    img = rng.random((50, 61)).astype(np.float32)
    sig = [1.0, 1.5, 2.0]
    L = IF.multiLoG(img, sig, (1.0, 1.0))
    assert L.shape == (3, 50, 61) and L.dtype == np.float32
    for i, s in enumerate(sig):
        k = Kernel.LoG((s, s))
        want = -s * O.imfilter_dense(img.astype(np.float64), np.asarray(k.data), list(k.first), "reflect")
        assert np.max(np.abs(L[i] - want)) <= 1e-5 * np.max(np.abs(want))
