"""CPU tests of the host layer (no GPU): kernel constructors, border types,
eltype rules, dispatch decisions, and that libifm loads and exports every
entry point declared in include/ifm.h."""
import ctypes
import os
import re

import numpy as np
import pytest

import imagefiltering as IF
from imagefiltering import Kernel, KernelFactors, _lib
from imagefiltering.border import _range
from imagefiltering.eltypes import Float32, Float64, N0f8, UInt8, mul_type

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---- C ABI -----------------------------------------------------------------------------
def test_libifm_exports_every_declared_symbol():
    hdr = open(os.path.join(REPO, "include", "ifm.h")).read()
    declared = set(re.findall(r"\b(ifm_[a-z0-9_]+)\s*\(", hdr))
    assert len(declared) >= 20
    lib = ctypes.CDLL(_lib.LIBPATH)
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(_lib.EXPORTS) <= declared


def test_libifm_version_and_device_count_without_gpu():
    L = _lib.load()
    assert L.ifm_version().startswith(b"ifm ")
    n = ctypes.c_int(-1)
    assert L.ifm_device_count(ctypes.byref(n)) == 0
    assert n.value >= 0


def test_struct_layouts_match_header():
    assert ctypes.sizeof(_lib.IfmArray) == 88
    assert ctypes.sizeof(_lib.IfmBorder) == 80
    assert ctypes.sizeof(_lib.IfmFactor) == 24
    assert ctypes.sizeof(_lib.IfmKernel) == 80
    assert ctypes.sizeof(_lib.IfmIir) == 16 * 8 + 8


# ---- kernel factories ------------------------------------------------------------------
def gaussiancmp(sigma, xr):
    c = np.exp(-np.asarray(xr, dtype=float) ** 2 / (2 * sigma ** 2))
    return c / c.sum()


def test_gaussian_factors_specialty_jl():
    k = KernelFactors.gaussian(2, 9)
    assert k.first == (-4,)
    np.testing.assert_allclose(k.data, gaussiancmp(2, range(-4, 5)), rtol=1e-14)
    kk = KernelFactors.gaussian((2, 3), (9, 7))
    np.testing.assert_allclose(kk[0].data.data, gaussiancmp(2, range(-4, 5)), rtol=1e-14)
    np.testing.assert_allclose(kk[1].data.data, gaussiancmp(3, range(-3, 4)), rtol=1e-14)
    assert abs(KernelFactors.gaussian(5).data.sum() - 1) < 1e-14
    g3 = KernelFactors.gaussian((3.0, 3.0))
    assert [len(f) for f in g3] == [13, 13] and g3[0].Npre == 0 and g3[1].Npre == 1
    k = KernelFactors.gaussian((3, 3, 0))
    assert k[2].data.data.tolist() == [1.0] and k[2].data.first == (0,)


def test_kernel_gaussian_and_dog_log():
    k = Kernel.gaussian((2, 3), (9, 7))
    np.testing.assert_allclose(k.data, np.outer(gaussiancmp(2, range(-4, 5)), gaussiancmp(3, range(-3, 4))),
                               rtol=1e-14)
    assert abs(Kernel.gaussian(5).data.sum() - 1) < 1e-12
    d = Kernel.DoG((2, 3), (4, 3.5), (9, 7))
    k1 = np.outer(gaussiancmp(2, range(-4, 5)), gaussiancmp(3, range(-3, 4)))
    k2 = np.outer(gaussiancmp(4, range(-4, 5)), gaussiancmp(3.5, range(-3, 4)))
    np.testing.assert_allclose(d.data, k1 - k2, atol=1e-15)
    assert abs(Kernel.DoG(5).data.sum()) < 1e-8
    L = Kernel.LoG(2.0)
    assert L.shape == (17, 17) and L.first == (-8, -8)
    # the LoG closed form at the centre: C * (0 - 2/σ²) with C = 1/(σ² 2π)
    assert abs(L.data[8, 8] - (1 / (4 * 2 * np.pi)) * (-2 / 4)) < 1e-15
    assert L == Kernel.LoG((2.0, 2.0))


def test_sobel_kernel_equals_factor_product():
    ky, kx = Kernel.sobel()
    fy, fx = KernelFactors.sobel()
    np.testing.assert_allclose(ky.data, np.outer([-0.5, 0, 0.5], [0.25, 0.5, 0.25]))
    np.testing.assert_allclose(kx.data, np.outer([0.25, 0.5, 0.25], [-0.5, 0, 0.5]))
    assert ky.first == (-1, -1)
    np.testing.assert_allclose(fy[0].data.data, [-0.5, 0, 0.5])


def test_box_and_errors():
    assert [f.data.data.tolist() for f in KernelFactors.box(3, 5)] == [[1 / 3] * 3, [1 / 5] * 5]
    with pytest.raises(IF.ArgumentError):
        KernelFactors.box((3, 4, 5))


@pytest.mark.parametrize("sigma", [2.0, 1.5, 3.1, 10.0, 100.0])
def test_triggs_sdika_M_satisfies_eq8(sigma):
    """test/triggs.jl:19-20: M = I1 + B*M*A."""
    k = KernelFactors.IIRGaussian(sigma)
    A = np.vstack([k.a, np.eye(3)[:2]])
    B = np.vstack([k.b, np.eye(3)[:2]])
    I1 = np.zeros((3, 3))
    I1[0, 0] = 1
    np.testing.assert_allclose(k.M, I1 + B @ k.M @ A, rtol=1e-10, atol=1e-12)


def test_iirgaussian_coefficients_sigma2():
    k = KernelFactors.IIRGaussian(2.0)
    np.testing.assert_allclose(k.a, [1.07757, -0.427266, 0.0650941], rtol=1e-5)
    assert abs(k.scale - 0.0809994) < 1e-7
    k = KernelFactors.IIRGaussian(1.5)
    np.testing.assert_allclose(k.a, [0.751101, -0.211738, 0.0243229], rtol=1e-5)
    # unit DC gain: scale / ((1 - sum a)(1 - sum b)) = 1
    for s in (1.5, 2.0, 7.0):
        k = KernelFactors.IIRGaussian(s)
        assert abs(k.scale / ((1 - k.asum) * (1 - k.bsum)) - 1) < 1e-12


# ---- borders -------------------------------------------------------------------------------
def test_pad_fill_inner_constructors_border_jl():
    assert IF.Pad("replicate", [1, 2], [5, 3]) == IF.Pad("replicate", (1, 2), (5, 3))
    assert IF.Pad("circular")(np.random.rand(3, 5)) == IF.Pad("circular", (0, 0), (3, 5))
    assert IF.Pad("circular")(IF.centered(np.random.rand(3, 5))) == IF.Pad("circular", (1, 2), (1, 2))
    assert IF.Pad("symmetric")(Kernel.Laplacian()) == IF.Pad("symmetric", (1, 1), (1, 1))
    assert IF.Inner((1, 2)) == IF.Inner((1, 2), (1, 2))
    assert IF.Inner((1, 2), ()) == IF.Inner((1, 2), (0, 0))
    assert IF.Inner()(IF.centered(np.random.rand(3, 5))) == IF.Inner((1, 2), (1, 2))
    assert IF.Fill(1, (3, 2)) == IF.Fill(1, (3, 2), (3, 2))
    assert IF.Fill(-1)(IF.centered(np.random.rand(3, 5))) == IF.Fill(-1, (1, 2), (1, 2))
    assert IF.Fill(2)(Kernel.Laplacian(), np.zeros((5, 5)), IF.Algorithm.FIR()) == IF.Fill(2, (1, 1), (1, 1))


def test_borderinstance_errors():
    assert IF.borderinstance("reflect") == IF.Pad("reflect")
    with pytest.raises(IF.ArgumentError):
        IF.borderinstance("inner")
    with pytest.raises(IF.ArgumentError):
        IF.borderinstance("nonsense")


def test_padfft_lengths():
    img = np.zeros((4096, 4096))
    p = IF.Pad("reflect")(IF.centered(np.zeros((15, 15))), img, IF.Algorithm.FFT())
    assert p.lo == (7, 7) and 7 + 4096 + p.hi[0] == 4374


@pytest.mark.parametrize("style", ["replicate", "circular", "symmetric", "reflect"])
def test_padindex_host_matches_numpy(style):
    mode = {"replicate": "edge", "circular": "wrap", "symmetric": "symmetric", "reflect": "reflect"}[style]
    for n in range(2, 7):
        for lo, hi in [(0, 0), (3, 1), (9, 13)]:
            np.testing.assert_array_equal(IF.padindex(IF.Pad(style), lo, n, hi),
                                          np.pad(np.arange(n), (lo, hi), mode=mode))


@pytest.mark.parametrize("style", ["replicate", "circular", "symmetric", "reflect"])
def test_borderarray_lazy_getindex_single_reflection(style):
    from oracle import oracle as O
    A = np.random.default_rng(1).standard_normal((5, 5))
    B = IF.BorderArray(A, IF.Pad(style, (2, 3), (4, 1)))
    P = O.padarray(A, style, [2, 3], [4, 1])    # pad < n: one reflection == periodic table
    for i in range(-2, 9):
        for j in range(-3, 6):
            assert B[i, j] == P[i + 2, j + 3]


def test_borderarray_fill_convert_errors():
    with pytest.raises(IF.ArgumentError):
        IF.BorderArray(np.zeros(3), IF.Fill(0.0))          # Fill{0}: no padding sizes
    with pytest.raises(IF.ArgumentError):
        IF.BorderArray(np.zeros(3), IF.Fill(None, (1,)))   # nothing does not convert
    with pytest.raises(IF.ArgumentError):
        IF.BorderArray(np.zeros(3, dtype=np.uint8), IF.Fill(7.5, (1,)))


# ---- eltype rules and dispatch -----------------------------------------------------------------
def test_filter_type_rules():
    assert IF.filter_type(Float32, KernelFactors.gaussian((3.0, 3.0))) is Float64
    assert IF.filter_type(Float32, (IF.centered(np.ones(3, np.float32)),)) is Float32
    assert IF.filter_type(UInt8, (IF.centered(np.ones(3)),)) is Float64
    assert IF.filter_type(N0f8, (IF.centered(np.ones(3, np.float32)),)) is Float32
    assert mul_type(N0f8, Float64) is Float64
    assert IF.filter_type(Float32, (Kernel.Laplacian(),)) is Float32


def test_factorkernel_svd():
    k = IF.factorkernel(IF.centered(np.outer([1.0, 2, 1], [-1.0, 0, 1])))
    assert len(k) == 2 and k[0].shape == (3, 1) and k[1].shape == (1, 3)
    np.testing.assert_allclose(np.outer(k[0].data.ravel(), k[1].data.ravel()), np.outer([1.0, 2, 1], [-1.0, 0, 1]),
                               atol=1e-12)
    rng = np.random.default_rng(0)
    k = IF.factorkernel(IF.centered(rng.standard_normal((15, 15))))
    assert len(k) == 1 and k[0].shape == (15, 15)
    assert len(IF.factorkernel(Kernel.LoG(2.0))) == 1


def test_filter_algorithm_heuristic():
    v = np.random.rand(100)
    k31 = (IF.centered(np.ones(31)),)
    assert IF.filter_algorithm(v, k31) == IF.Algorithm.FFT()
    w = v.copy()
    w[5] = np.nan
    assert IF.filter_algorithm(w, k31) == IF.Algorithm.FIR()
    assert IF.filter_algorithm(v, KernelFactors.gaussian((1.0, 1.0))) == IF.Algorithm.FIRTiled()
    assert IF.filter_algorithm(v, KernelFactors.IIRGaussian((2.0, 2.0))) == IF.Algorithm.IIR()
    assert IF.filter_algorithm(v, (KernelFactors.IIRGaussian(2.0), IF.centered(np.ones(3)))) == IF.Algorithm.Mixed()


def test_method_errors_and_window_errors():
    r = IF.CPU1(IF.Algorithm.FIR())
    with pytest.raises(IF.MethodError):
        IF.imfilter(r, np.zeros(8), (IF.centered(np.ones(3)),), "replicate", IF.Algorithm.FIR())
    with pytest.raises(IF.ArgumentError):
        IF.imfilter(np.zeros(8), (IF.centered(np.ones(3)),), "inner")
    with pytest.raises(IF.ArgumentError):
        IF.resolve_window(())
    with pytest.raises(RuntimeError):
        IF.resolve_window((3, 4))
    assert IF.resolve_window(5) == (range(-2, 3),)
    assert IF.resolve_window((range(-1, 2), range(0, 1))) == (range(-1, 2), range(0, 1))


def test_explicit_padding_too_small_is_dimension_mismatch():
    A = np.ones((8, 8))
    kern = Kernel.gaussian((1, 1), (3, 3))
    with pytest.raises(IF.ArgumentError):
        IF.imfilter(A, kern, IF.Fill(0, (3,)))
    with pytest.raises(IF.DimensionMismatch):
        IF.imfilter(A, kern, IF.Fill(0, (1, 0)))
    with pytest.raises(IF.DimensionMismatch):
        IF.imfilter(A, kern, IF.Fill(0, (0, 0)))
