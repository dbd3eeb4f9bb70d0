// k_iir.hip — Triggs-Sdika IIR filtering along one axis in FP64 (gfx950).
//
// Follows _imfilter_dim! (src/imfilter.jl:964-1024) with leftborder!
// (:1027-1048) and rightborder! (:1051-1092): forward recursion with the
// steady-state left initialisation, right-edge initialisation by the 3x3
// matrix M (Triggs & Sdika Eqs. 14-15), backward recursion and scaling.  A
// line is sequential, so each thread owns one line; lines along non-
// contiguous axes are adjacent in memory across the wave (coalesced).
#include "ifm_internal.h"

namespace ifm {

struct IirArgs {
  const void *src;
  double *dst;        // FP64 stage result (forward values, then the output)
  int dt_src;
  int n[MAXD];
  int axis;
  int64_t stride;     // element stride of the filtered axis
  int64_t lines;
  double a[3], b[3], scale, M[9];
  int isfill;
  double fill;
};

// The recursions are sequential along a line, so the inputs of the next
// IC steps are loaded into registers before those steps run (a step never
// waits for its own load).  Reading ahead is safe in place: the forward pass
// only writes positions it has already read, the backward pass likewise.
constexpr int IC = 16;

// Source element as FP64; DT < 0: element type read at run time.
template <int DT>
__device__ __forceinline__ double iir_ld(const IirArgs &p, int64_t i) {
  if constexpr (DT == IFM_F32) return (double)((const float *)p.src)[i];
  else if constexpr (DT == IFM_F64) return ((const double *)p.src)[i];
  else return load_d(p.src, i, (int)p.dt_src);
}

// Line accessor: element i of line `base` with stride `st` (ST1: st == 1, so
// a chunk is contiguous and the loads vectorise).
template <bool ST1, int DT>
__global__ void __launch_bounds__(256) k_iir_strided(IirArgs p) {
  const int64_t line = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (line >= p.lines) return;
  const int n = p.n[p.axis];
  const int64_t st = ST1 ? 1 : p.stride;
  const int64_t pre = line % st, post = line / st;
  const int64_t base = pre + post * st * n;
  const double a1 = p.a[0], a2 = p.a[1], a3 = p.a[2];
  const double b1 = p.b[0], b2 = p.b[1], b3 = p.b[2];
  const double asum = a1 + a2 + a3, bsum = b1 + b2 + b3;
  const double iminus = p.isfill ? p.fill : iir_ld<DT>(p, base);
  const double iplus = p.isfill ? p.fill : iir_ld<DT>(p, base + (int64_t)(n - 1) * st);
  // leftborder!
  const double uminus = iminus / (1 - asum);
  double y0 = iir_ld<DT>(p, base) + a1 * uminus + a2 * uminus + a3 * uminus;
  p.dst[base] = y0;
  double y1 = iir_ld<DT>(p, base + st) + a1 * y0 + a2 * uminus + a3 * uminus;
  p.dst[base + st] = y1;
  double y2 = iir_ld<DT>(p, base + 2 * st) + a1 * y1 + a2 * y0 + a3 * uminus;
  p.dst[base + 2 * st] = y2;
  // forward, omitting the final point
  double ym3 = y0, ym2 = y1, ym1 = y2;
  int i = 3;
  for (; i + IC - 1 <= n - 2; i += IC) {
    double x[IC];
#pragma unroll
    for (int j = 0; j < IC; ++j) x[j] = iir_ld<DT>(p, base + (int64_t)(i + j) * st);
#pragma unroll
    for (int j = 0; j < IC; ++j) {
      const double y = x[j] + a1 * ym1 + a2 * ym2 + a3 * ym3;
      p.dst[base + (int64_t)(i + j) * st] = y;
      ym3 = ym2;
      ym2 = ym1;
      ym1 = y;
    }
  }
  for (; i <= n - 2; ++i) {
    double y = iir_ld<DT>(p, base + (int64_t)i * st) + a1 * ym1 + a2 * ym2 + a3 * ym3;
    p.dst[base + (int64_t)i * st] = y;
    ym3 = ym2;
    ym2 = ym1;
    ym1 = y;
  }
  // rightborder!: the last forward value, then v at and beyond the edge
  const double ylast = iir_ld<DT>(p, base + (int64_t)(n - 1) * st) + a1 * ym1 + a2 * ym2 + a3 * ym3;
  const double uplus = iplus / (1 - asum);
  const double vplus = uplus / (1 - bsum);
  const double d0 = ylast - uplus, d1 = ym1 - uplus, d2 = ym2 - uplus;
  const double v0 = p.M[0] * d0 + p.M[1] * d1 + p.M[2] * d2 + vplus;
  const double v1 = p.M[3] * d0 + p.M[4] * d1 + p.M[5] * d2 + vplus;
  const double v2 = p.M[6] * d0 + p.M[7] * d1 + p.M[8] * d2 + vplus;
  // out[n-1] = v0; out[n-2] = y[n-2] + b1*out[n-1] + b2*v1 + b3*v2;
  // out[n-3] = y[n-3] + b1*out[n-2] + b2*out[n-1] + b3*v1
  const double z2 = v0;
  const double z1 = ym1 + b1 * z2 + b2 * v1 + b3 * v2;
  const double z0 = ym2 + b1 * z1 + b2 * z2 + b3 * v1;
  const double sc = p.scale;
  p.dst[base + (int64_t)(n - 1) * st] = z2 * sc;
  p.dst[base + (int64_t)(n - 2) * st] = z1 * sc;
  p.dst[base + (int64_t)(n - 3) * st] = z0 * sc;
  // backward
  double zp1 = z0, zp2 = z1, zp3 = z2;
  int k = n - 4;
  for (; k - (IC - 1) >= 0; k -= IC) {
    double y[IC];
#pragma unroll
    for (int j = 0; j < IC; ++j) y[j] = p.dst[base + (int64_t)(k - j) * st];
#pragma unroll
    for (int j = 0; j < IC; ++j) {
      const double z = y[j] + b1 * zp1 + b2 * zp2 + b3 * zp3;
      p.dst[base + (int64_t)(k - j) * st] = z * sc;
      zp3 = zp2;
      zp2 = zp1;
      zp1 = z;
    }
  }
  for (; k >= 0; --k) {
    double z = p.dst[base + (int64_t)k * st] + b1 * zp1 + b2 * zp2 + b3 * zp3;
    p.dst[base + (int64_t)k * st] = z * sc;
    zp3 = zp2;
    zp2 = zp1;
    zp1 = z;
  }
}

hipError_t launch_iir(const Dense4 &in, Dense4 &out, int axis, const IirCoef &c, int isfill, double fill,
                      hipStream_t st) {
  IirArgs p;
  p.src = in.data;
  p.dst = (double *)out.data;
  p.dt_src = in.dtype;
  int64_t stride = 1;
  for (int d = 0; d < MAXD; ++d) {
    p.n[d] = (int)in.n[d];
    if (d < axis) stride *= in.n[d];
  }
  p.axis = axis;
  p.stride = stride;
  p.lines = in.count() / in.n[axis];
  for (int i = 0; i < 3; ++i) {
    p.a[i] = c.a[i];
    p.b[i] = c.b[i];
  }
  p.scale = c.scale;
  for (int i = 0; i < 9; ++i) p.M[i] = c.M[i];
  p.isfill = isfill;
  p.fill = fill;
  if (p.lines == 0) return hipSuccess;
  int64_t blocks = (p.lines + 255) / 256;
  const bool s1 = stride == 1;
  if (in.dtype == IFM_F32) {
    if (s1) hipLaunchKernelGGL((k_iir_strided<true, IFM_F32>), dim3((unsigned)blocks), dim3(256), 0, st, p);
    else hipLaunchKernelGGL((k_iir_strided<false, IFM_F32>), dim3((unsigned)blocks), dim3(256), 0, st, p);
  } else if (in.dtype == IFM_F64) {
    if (s1) hipLaunchKernelGGL((k_iir_strided<true, IFM_F64>), dim3((unsigned)blocks), dim3(256), 0, st, p);
    else hipLaunchKernelGGL((k_iir_strided<false, IFM_F64>), dim3((unsigned)blocks), dim3(256), 0, st, p);
  } else {
    if (s1) hipLaunchKernelGGL((k_iir_strided<true, -1>), dim3((unsigned)blocks), dim3(256), 0, st, p);
    else hipLaunchKernelGGL((k_iir_strided<false, -1>), dim3((unsigned)blocks), dim3(256), 0, st, p);
  }
  return hipGetLastError();
}

}  // namespace ifm
