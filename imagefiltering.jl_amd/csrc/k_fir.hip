// k_fir.hip — padarray and separable FIR kernels (gfx950).
//
// padarray materialises BorderArray + copy! (src/border.jl:324-347).  The FIR
// kernels apply the border index map on load, so the reference's padded copy
// (src/imfilter.jl:336) is never built: a factor cascade reads the image once
// per fused group of axes.
#include "ifm_internal.h"

namespace ifm {

// ---------------------------------------------------------------------------
// padarray
// ---------------------------------------------------------------------------
struct PadArgs {
  const void *in;
  void *out;
  int n[MAXD];    // input extents
  int p[MAXD];    // output extents
  int lo[MAXD];
  int style;
  int dt_in, dt_out;
  int esize;      // element size when dt_in == dt_out (raw copy)
  uint64_t fill_bits;
  double fill;
  int64_t total;
};

__global__ void __launch_bounds__(256) k_pad(PadArgs a) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < a.total;
       q += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = q;
    int i0 = (int)(r % a.p[0]); r /= a.p[0];
    int i1 = (int)(r % a.p[1]); r /= a.p[1];
    int i2 = (int)(r % a.p[2]); r /= a.p[2];
    int i3 = (int)r;
    int s0 = bmap(a.style, i0 - a.lo[0], a.n[0]);
    int s1 = bmap(a.style, i1 - a.lo[1], a.n[1]);
    int s2 = bmap(a.style, i2 - a.lo[2], a.n[2]);
    int s3 = bmap(a.style, i3 - a.lo[3], a.n[3]);
    bool isfill = (s0 | s1 | s2 | s3) < 0;
    int64_t src = s0 + (int64_t)a.n[0] * (s1 + (int64_t)a.n[1] * (s2 + (int64_t)a.n[2] * s3));
    if (a.dt_in == a.dt_out) {
      switch (a.esize) {
      case 1: ((uint8_t *)a.out)[q] = isfill ? (uint8_t)a.fill_bits : ((const uint8_t *)a.in)[src]; break;
      case 2: ((uint16_t *)a.out)[q] = isfill ? (uint16_t)a.fill_bits : ((const uint16_t *)a.in)[src]; break;
      case 4: ((uint32_t *)a.out)[q] = isfill ? (uint32_t)a.fill_bits : ((const uint32_t *)a.in)[src]; break;
      default: ((uint64_t *)a.out)[q] = isfill ? a.fill_bits : ((const uint64_t *)a.in)[src]; break;
      }
    } else {
      double v = isfill ? a.fill : load_d(a.in, src, a.dt_in);
      store_d(a.out, q, a.dt_out, v);
    }
  }
}

static uint64_t fill_bits_for(int dt, double v) {
  uint64_t b = 0;
  switch (dt) {
  case IFM_U8:
  case IFM_N0F8: b = (uint64_t)(uint8_t)(int)v; break;  // caller passes the code
  case IFM_F32: { float f = (float)v; uint32_t u; memcpy(&u, &f, 4); b = u; break; }
  case IFM_F64: memcpy(&b, &v, 8); break;
  case IFM_BF16: { float f = (float)v; uint32_t u; memcpy(&u, &f, 4); b = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16; break; }
  default: { __half h = __float2half_rn((float)v); uint16_t u; memcpy(&u, &h, 2); b = u; break; }
  }
  return b;
}

hipError_t launch_pad(const Dense4 &in, Dense4 &out, const int64_t *lo, int style, double fill, hipStream_t st) {
  PadArgs a;
  a.in = in.data;
  a.out = out.data;
  for (int d = 0; d < MAXD; ++d) {
    a.n[d] = (int)in.n[d];
    a.p[d] = (int)out.n[d];
    a.lo[d] = (int)lo[d];
  }
  a.style = style;
  a.dt_in = in.dtype;
  a.dt_out = out.dtype;
  a.esize = (int)dtype_size(in.dtype);
  a.fill = fill;
  a.fill_bits = fill_bits_for(out.dtype, fill);
  a.total = out.count();
  if (a.total == 0) return hipSuccess;
  int64_t blocks = (a.total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_pad, dim3((unsigned)blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// fir1d: one factor along any canonical axis.  Input covers image indices
// in_first[d] .. in_first[d]+in.n[d]-1 on every axis but the filtered one,
// where it covers the whole image (the border map applies there).
// ---------------------------------------------------------------------------
struct Fir1dArgs {
  const void *in;
  void *out;
  int in_n[MAXD];
  int out_n[MAXD];
  int shift[MAXD];   // out coordinate + shift = in coordinate (non-filtered axes)
  int axis;
  int len;
  int first;         // tap offset + out_first along the axis
  int style;
  int dt_in, dt_out;
  float fill;
  int64_t total;
  float taps[MAXTAPS];
};

__global__ void __launch_bounds__(256) k_fir1d(Fir1dArgs a) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < a.total;
       q += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = q;
    int o[MAXD];
    o[0] = (int)(r % a.out_n[0]); r /= a.out_n[0];
    o[1] = (int)(r % a.out_n[1]); r /= a.out_n[1];
    o[2] = (int)(r % a.out_n[2]); r /= a.out_n[2];
    o[3] = (int)r;
    int64_t base = 0, stride = 1, astride = 1;
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
      if (d == a.axis) astride = stride;
      else base += (int64_t)(o[d] + a.shift[d]) * stride;
      stride *= a.in_n[d];
    }
    int x = o[a.axis] + a.first;
    int n = a.in_n[a.axis];
    float acc = 0.f;
    if (x >= 0 && x + a.len <= n) {
      const int64_t b0 = base + (int64_t)x * astride;
      for (int t = 0; t < a.len; ++t) acc = fmaf(a.taps[t], load_f(a.in, b0 + t * astride, a.dt_in), acc);
    } else {
      for (int t = 0; t < a.len; ++t) {
        int s = bmap(a.style, x + t, n);
        float v = s < 0 ? a.fill : load_f(a.in, base + (int64_t)s * astride, a.dt_in);
        acc = fmaf(a.taps[t], v, acc);
      }
    }
    store_f(a.out, q, a.dt_out, acc);
  }
}

hipError_t launch_fir1d(const Dense4 &in, const int64_t *in_first, Dense4 &out, const int64_t *out_first,
                        const int64_t *img_n, const FirPass &p, int style, float fill, hipStream_t st) {
  Fir1dArgs a;
  a.in = in.data;
  a.out = out.data;
  for (int d = 0; d < MAXD; ++d) {
    a.in_n[d] = (int)in.n[d];
    a.out_n[d] = (int)out.n[d];
    a.shift[d] = (int)(out_first[d] - in_first[d]);
  }
  (void)img_n;
  a.axis = p.axis;
  a.len = p.len;
  a.first = p.first + (int)out_first[p.axis];
  a.style = style;
  a.dt_in = in.dtype;
  a.dt_out = out.dtype;
  a.fill = fill;
  a.total = out.count();
  for (int t = 0; t < p.len; ++t) a.taps[t] = p.taps[t];
  if (a.total == 0) return hipSuccess;
  int64_t blocks = (a.total + 255) / 256;
  if (blocks > 1 << 20) blocks = 1 << 20;
  hipLaunchKernelGGL(k_fir1d, dim3((unsigned)blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// sep2d: fused pair of factors on canonical axes 0 and 1, batched over axes
// 2 and 3.  An output tile of 64 x 16 pixels loads its halo'd input tile into
// LDS once (border map on load), runs the axis-0 factor into a second LDS
// tile and the axis-1 factor from there, so HBM sees one read and one write.
// ---------------------------------------------------------------------------
constexpr int S2_TX = 64;
constexpr int S2_TY = 16;
constexpr int S2_MAXK = 64;

struct Sep2dArgs {
  const void *in;
  void *out;
  int n0, n1;          // image extents (axes 0, 1)
  int o0, o1;          // output extents
  int f0, f1;          // out_first
  int64_t in_plane, out_plane;
  int nplanes;
  int o2, n2, f2, f3;  // plane mapping: out plane (p2, p3) reads in plane (p2+f2, p3+f3)
  int k0, k1;          // tap counts
  int first0, first1;  // tap offsets
  int style;
  int dt_in, dt_out;
  float fill;
  float taps0[S2_MAXK];
  float taps1[S2_MAXK];
};

__global__ void __launch_bounds__(256) k_sep2d(Sep2dArgs a) {
  extern __shared__ float smem[];
  const int IN_W = S2_TX + a.k0 - 1;
  const int IN_H = S2_TY + a.k1 - 1;
  const int IN_WP = IN_W | 1;
  float *s_in = smem;                     // IN_H x IN_WP
  float *s_mid = smem + IN_H * IN_WP;     // IN_H x S2_TX
  const int tid = threadIdx.x;
  const int tx = tid & 63, ty = tid >> 6;
  const int x0 = a.f0 + blockIdx.x * S2_TX;   // image coordinate of the tile's first output
  const int y0 = a.f1 + blockIdx.y * S2_TY;
  const void *in = a.in;
  for (int plane = blockIdx.z; plane < a.nplanes; plane += gridDim.z) {
  const int p2 = plane % a.o2, p3 = plane / a.o2;
  const int64_t ibase = ((int64_t)(p2 + a.f2) + (int64_t)a.n2 * (p3 + a.f3)) * a.in_plane;

  // Load the input tile (border map on load).
  int sx[2];
#pragma unroll
  for (int cc = 0; cc < 2; ++cc) {
    int c = tx + 64 * cc;
    sx[cc] = (c < IN_W) ? bmap(a.style, x0 + a.first0 + c, a.n0) : 0;
  }
  for (int r = ty; r < IN_H; r += 4) {
    int sy = bmap(a.style, y0 + a.first1 + r, a.n1);
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
      int c = tx + 64 * cc;
      if (c < IN_W) {
        float v = (sy < 0 || sx[cc] < 0) ? a.fill : load_f(in, ibase + (int64_t)sy * a.n0 + sx[cc], a.dt_in);
        s_in[r * IN_WP + c] = v;
      }
    }
  }
  __syncthreads();
  // Axis-0 factor.
  for (int r = ty; r < IN_H; r += 4) {
    const float *row = s_in + r * IN_WP + tx;
    float acc = 0.f;
    for (int t = 0; t < a.k0; ++t) acc = fmaf(a.taps0[t], row[t], acc);
    s_mid[r * S2_TX + tx] = acc;
  }
  __syncthreads();
  // Axis-1 factor: each thread owns 4 consecutive rows of one column.
  const int r0 = ty * 4;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int u = 0; u < a.k1 + 3; ++u) {
    float v = s_mid[(r0 + u) * S2_TX + tx];
#pragma unroll
    for (int o = 0; o < 4; ++o) {
      int j = u - o;
      if (j >= 0 && j < a.k1) acc[o] = fmaf(a.taps1[j], v, acc[o]);
    }
  }
  const int ox = blockIdx.x * S2_TX + tx;
  if (ox < a.o0) {
    const int64_t obase = (int64_t)plane * a.out_plane;
#pragma unroll
    for (int o = 0; o < 4; ++o) {
      int oy = blockIdx.y * S2_TY + r0 + o;
      if (oy < a.o1) store_f(a.out, obase + (int64_t)oy * a.o0 + ox, a.dt_out, acc[o]);
    }
  }
  __syncthreads();  // the next plane reuses the LDS tiles
  }
}

hipError_t launch_sep2d(const Dense4 &in, Dense4 &out, const int64_t *out_first, const FirPass &p0,
                        const FirPass &p1, int style, float fill, hipStream_t st) {
  Sep2dArgs a;
  a.in = in.data;
  a.out = out.data;
  a.n0 = (int)in.n[0];
  a.n1 = (int)in.n[1];
  a.o0 = (int)out.n[0];
  a.o1 = (int)out.n[1];
  a.f0 = (int)out_first[0];
  a.f1 = (int)out_first[1];
  a.in_plane = in.n[0] * in.n[1];
  a.out_plane = out.n[0] * out.n[1];
  a.nplanes = (int)(out.n[2] * out.n[3]);
  a.o2 = (int)out.n[2];
  a.n2 = (int)in.n[2];
  a.f2 = (int)out_first[2];
  a.f3 = (int)out_first[3];
  a.k0 = p0.len;
  a.k1 = p1.len;
  a.first0 = p0.first;
  a.first1 = p1.first;
  a.style = style;
  a.dt_in = in.dtype;
  a.dt_out = out.dtype;
  a.fill = fill;
  for (int t = 0; t < p0.len; ++t) a.taps0[t] = p0.taps[t];
  for (int t = 0; t < p1.len; ++t) a.taps1[t] = p1.taps[t];
  if (out.count() == 0) return hipSuccess;
  const int IN_W = S2_TX + a.k0 - 1, IN_H = S2_TY + a.k1 - 1;
  size_t lds = sizeof(float) * ((size_t)IN_H * (IN_W | 1) + (size_t)IN_H * S2_TX);
  dim3 grid((a.o0 + S2_TX - 1) / S2_TX, (a.o1 + S2_TY - 1) / S2_TY, a.nplanes < 65535 ? a.nplanes : 65535);
  hipLaunchKernelGGL(k_sep2d, grid, dim3(256), lds, st, a);
  return hipGetLastError();
}

}  // namespace ifm
