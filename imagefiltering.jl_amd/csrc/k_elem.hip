// k_elem.hip — elementwise passes behind the NA() border and the NoPad
// sub-range form of imfilter! (gfx950): the NaN mask and the zero-filled image
// of imfilter_na_inseparable! (src/imfilter.jl:1109-1122), the division by
// the filtered mask, and copyto!(out, R, A, R) into a sub-block
// (src/imfilter.jl:372-375, :433-434).  One streaming pass each.
#include "ifm_internal.h"

namespace ifm {

struct NaPrepArgs {
  const void *in;
  void *tmp;    // in with NaN replaced by zero, in tmp's eltype
  void *valid;  // 1 where in is not NaN, else 0, in valid's eltype
  int dt_in, dt_tmp, dt_valid;
  int64_t count;
  int *hasna;
};

__device__ __forceinline__ bool is_nan_at(const void *p, int64_t i, int dt) {
  switch (dt) {
  case IFM_F32: return (((const uint32_t *)p)[i] & 0x7fffffffu) > 0x7f800000u;
  case IFM_F64: return (((const uint64_t *)p)[i] & 0x7fffffffffffffffull) > 0x7ff0000000000000ull;
  case IFM_BF16: return (((const uint16_t *)p)[i] & 0x7fffu) > 0x7f80u;
  case IFM_F16: return (((const uint16_t *)p)[i] & 0x7fffu) > 0x7c00u;
  default: return false;  // integers and N0f8 cannot be NaN
  }
}

__global__ void __launch_bounds__(256) k_na_prepare(NaPrepArgs a) {
  bool any = false;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.count;
       i += (int64_t)gridDim.x * blockDim.x) {
    const bool na = is_nan_at(a.in, i, a.dt_in);
    any |= na;
    if (a.dt_tmp == IFM_F64) store_d(a.tmp, i, IFM_F64, na ? 0.0 : load_d(a.in, i, a.dt_in));
    else store_f(a.tmp, i, a.dt_tmp, na ? 0.f : load_f(a.in, i, a.dt_in));
    store_d(a.valid, i, a.dt_valid, na ? 0.0 : 1.0);
  }
  if (any) atomicOr(a.hasna, 1);
}

__global__ void __launch_bounds__(256) k_divide(void *num, const void *den, int dt_num, int dt_den, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    store_d(num, i, dt_num, load_d(num, i, dt_num) / load_d(den, i, dt_den));
}

struct CopyIntoArgs {
  const void *src;
  void *dst;
  int64_t sn[MAXD];   // src extents
  int64_t dn[MAXD];   // dst extents
  int64_t df[MAXD];   // first dst index of the block
  int dt_src, dt_dst;
  int64_t count;
};

__global__ void __launch_bounds__(256) k_copy_into(CopyIntoArgs a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.count;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = i, o = 0, stride = 1;
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
      const int64_t c = r % a.sn[d];
      r /= a.sn[d];
      o += (c + a.df[d]) * stride;
      stride *= a.dn[d];
    }
    if (a.dt_src == a.dt_dst) {
      const int es = (a.dt_src == IFM_U8 || a.dt_src == IFM_N0F8) ? 1
                   : (a.dt_src == IFM_F32) ? 4 : (a.dt_src == IFM_F64) ? 8 : 2;
      switch (es) {
      case 1: ((uint8_t *)a.dst)[o] = ((const uint8_t *)a.src)[i]; break;
      case 2: ((uint16_t *)a.dst)[o] = ((const uint16_t *)a.src)[i]; break;
      case 4: ((uint32_t *)a.dst)[o] = ((const uint32_t *)a.src)[i]; break;
      default: ((uint64_t *)a.dst)[o] = ((const uint64_t *)a.src)[i]; break;
      }
    } else {
      store_d(a.dst, o, a.dt_dst, load_d(a.src, i, a.dt_src));
    }
  }
}

static unsigned elem_blocks(int64_t count) {
  const int64_t b = (count + 255) / 256;
  return (unsigned)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}

hipError_t launch_na_prepare(const Dense4 &in, Dense4 &tmp, Dense4 &valid, int *hasna, hipStream_t st) {
  NaPrepArgs a;
  a.in = in.data;
  a.tmp = tmp.data;
  a.valid = valid.data;
  a.dt_in = in.dtype;
  a.dt_tmp = tmp.dtype;
  a.dt_valid = valid.dtype;
  a.count = in.count();
  a.hasna = hasna;
  if (a.count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_na_prepare, dim3(elem_blocks(a.count)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_divide(Dense4 &num, const Dense4 &den, hipStream_t st) {
  const int64_t count = num.count();
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_divide, dim3(elem_blocks(count)), dim3(256), 0, st, num.data, (const void *)den.data, num.dtype,
                     den.dtype, count);
  return hipGetLastError();
}

hipError_t launch_copy_into(const Dense4 &src, Dense4 &dst, const int64_t *dst_first, hipStream_t st) {
  CopyIntoArgs a;
  a.src = src.data;
  a.dst = dst.data;
  for (int d = 0; d < MAXD; ++d) {
    a.sn[d] = src.n[d];
    a.dn[d] = dst.n[d];
    a.df[d] = dst_first[d];
  }
  a.dt_src = src.dtype;
  a.dt_dst = dst.dtype;
  a.count = src.count();
  if (a.count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_into, dim3(elem_blocks(a.count)), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace ifm

namespace ifm {

// findlocalmaxima / findlocalminima (src/extrema.jl:131-163): flag[i] = 1 when
// i lies in the clipped range (edges[d] == 0 drops the first and last index
// of axis d) and img[i] beats every in-bounds neighbour of its window (>
// for maxima, < for minima, compared as the element type's values).
struct LocalExtArgs {
  const void *in;
  uint8_t *flag;
  int64_t n[MAXD];
  int half[MAXD];
  int edges[MAXD];
  int dt;
  int sense;  // +1 maxima, -1 minima
  int64_t count;
};

__global__ void __launch_bounds__(256) k_local_extrema(LocalExtArgs a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.count;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t c[MAXD], r = i;
    bool ok = true;
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
      c[d] = r % a.n[d];
      r /= a.n[d];
      if (!a.edges[d] && (c[d] == 0 || c[d] == a.n[d] - 1)) ok = false;
    }
    if (ok) {
      const double v = load_d(a.in, i, a.dt);
      for (int j3 = -a.half[3]; j3 <= a.half[3] && ok; ++j3)
        for (int j2 = -a.half[2]; j2 <= a.half[2] && ok; ++j2)
          for (int j1 = -a.half[1]; j1 <= a.half[1] && ok; ++j1)
            for (int j0 = -a.half[0]; j0 <= a.half[0] && ok; ++j0) {
              if ((j0 | j1 | j2 | j3) == 0) continue;
              const int64_t q0 = c[0] + j0, q1 = c[1] + j1, q2 = c[2] + j2, q3 = c[3] + j3;
              if (q0 < 0 || q0 >= a.n[0] || q1 < 0 || q1 >= a.n[1] || q2 < 0 || q2 >= a.n[2] || q3 < 0 ||
                  q3 >= a.n[3])
                continue;
              const double w = load_d(a.in, q0 + a.n[0] * (q1 + a.n[1] * (q2 + a.n[2] * q3)), a.dt);
              ok = a.sense > 0 ? (v > w) : (v < w);
            }
    }
    a.flag[i] = ok ? 1 : 0;
  }
}

hipError_t launch_local_extrema(const Dense4 &in, uint8_t *flag, const int *half, const int *edges, int sense,
                                hipStream_t st) {
  LocalExtArgs a;
  a.in = in.data;
  a.flag = flag;
  for (int d = 0; d < MAXD; ++d) {
    a.n[d] = in.n[d];
    a.half[d] = half[d];
    a.edges[d] = edges[d];
  }
  a.dt = in.dtype;
  a.sense = sense;
  a.count = in.count();
  if (a.count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_local_extrema, dim3(elem_blocks(a.count)), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace ifm

namespace ifm {

// convert(T, x) of floating-point results into an integer eltype T of `bits`
// width (signed or not), scaled by `scale` first (255 for N0f8 codes): a
// value that is not an integer or lies outside T's range sets *inexact, as
// Julia's convert throws InexactError (src/imfilter.jl:233-254).
struct ToIntArgs {
  const void *src;
  void *dst;
  int dt_src;
  int bits, is_signed;
  double scale;
  int64_t count;
  int *inexact;
};

__global__ void __launch_bounds__(256) k_to_int(ToIntArgs a) {
  bool bad = false;
  double lo, hi;
  if (a.is_signed) {
    hi = ldexp(1.0, a.bits - 1) - 1.0;
    lo = -ldexp(1.0, a.bits - 1);
  } else {
    hi = ldexp(1.0, a.bits) - 1.0;
    lo = 0.0;
  }
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.count;
       i += (int64_t)gridDim.x * blockDim.x) {
    // integers must be exact; a fixed-point target (scale != 1, N0f8) rounds
    // to the nearest code as convert(N0f8, x) does and checks the range only
    const double s = load_d(a.src, i, a.dt_src) * a.scale;
    const double v = a.scale != 1.0 ? rint(s) : s;
    const bool ok = v == rint(v) && v >= lo && v <= hi;  // NaN fails every comparison
    bad |= !ok;
    const double c = ok ? v : 0.0;
    switch (a.bits) {
    case 8: ((uint8_t *)a.dst)[i] = (uint8_t)(int)(int64_t)c; break;
    case 16: ((uint16_t *)a.dst)[i] = (uint16_t)(int)(int64_t)c; break;
    case 32: ((uint32_t *)a.dst)[i] = (uint32_t)(int64_t)c; break;
    default:
      if (a.is_signed) ((int64_t *)a.dst)[i] = (int64_t)c;
      else ((uint64_t *)a.dst)[i] = (uint64_t)c;
      break;
    }
  }
  if (bad) atomicOr(a.inexact, 1);
}

hipError_t launch_to_int(const Dense4 &src, void *dst, int bits, int is_signed, double scale, int *inexact,
                         hipStream_t st) {
  ToIntArgs a;
  a.src = src.data;
  a.dst = dst;
  a.dt_src = src.dtype;
  a.bits = bits;
  a.is_signed = is_signed;
  a.scale = scale;
  a.count = src.count();
  a.inexact = inexact;
  if (a.count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_int, dim3(elem_blocks(a.count)), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace ifm
