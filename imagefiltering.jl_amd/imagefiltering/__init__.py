"""MI355X-native drop-in for the filtering hot path of ImageFiltering.jl.

Host-side mirror of the reference's API (src/ImageFiltering.jl:24-33 export
list) over the C ABI of libifm (include/ifm.h): imfilter / imfilter!
(here imfilter_), padarray, BorderArray, Pad / Fill / Inner / NA / NoPad,
Kernel, KernelFactors, Algorithm, centered, kernelfactors, mapwindow /
mapwindow! (mapwindow_) for the median, minimum, maximum and extrema filters.

Conventions: numpy axis d is Julia dimension d+1; image indices are 0-based;
kernel offsets are the reference's (centered kernels start at -(n-1)÷2).
"""
from . import kernel as Kernel
from . import kernelfactors as KernelFactors
from .border import (BorderArray, Fill, Inner, NA, NoPad, Pad, borderinstance, calculate_padding, padarray,
                     padindex)
from .device import DeviceArray
from .eltypes import (BF16Array, BFloat16, Float16, Float32, Float64, N0f8, N0f8Array, UInt8, bf16,
                      bf16_to_float32, n0f8)
from .errors import ArgumentError, DimensionMismatch, InexactError, MethodError
from .imfilter import (CPU1, FFT, FIR, IIR, AbstractResource, Algorithm, ColorImage, CPUThreads, FIRTiled, HIPDevice,
                       Mixed, channelview, colorview, factorkernel, filter_algorithm, filter_type, imfilter, imfilter_,
                       kernelconv)
from .kernelfactors import kernelfactors
from .mapwindow import (extrema, first, mapwindow, mapwindow_, maximum, mean, median, median_, minimum, quantile,
                        resolve_window, sum_)
from .offsetarrays import OffsetArray, centered
from .specialty import BlobLoG, blob_LoG, findlocalmaxima, findlocalminima, imgradients, multiLoG

__all__ = [
    "Kernel", "KernelFactors", "Pad", "Fill", "Inner", "NA", "NoPad", "BorderArray", "Algorithm",
    "imfilter", "imfilter_", "mapwindow", "mapwindow_", "padarray", "centered", "kernelfactors",
    "OffsetArray", "DeviceArray", "median", "median_", "minimum", "maximum", "extrema", "mean", "sum_", "first",
    "quantile",
    "imgradients", "findlocalmaxima", "findlocalminima", "blob_LoG", "multiLoG", "BlobLoG",
    "ColorImage", "colorview", "channelview",
]
