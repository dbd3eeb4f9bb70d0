// k_dense.hip — dense (non-separable) FIR correlation (gfx950).
//
// Replaces __imfilter_inbounds! (src/imfilter.jl:592-669): out[i] =
// sum_J A[i+J] k[J] on the padded image.  The border map is applied while the
// input tile is staged into LDS.
//
//  * k_dense_nd     any rank <= 4, one output per thread (small kernels, 3-D)
//  * k_dense2d_valu 2-D, LDS tile, 8 rows per thread, FP32 FMA
//  * k_dense_mfma   2-D banded-Toeplitz on MFMA v_mfma_f32_16x16x32_bf16 with
//                   x and k split into bf16 hi+lo (3 products, FP32 accumulate)
#include "ifm_internal.h"

namespace ifm {

// ---------------------------------------------------------------------------
// Generic N-D (one thread per output, taps from global memory).
// ---------------------------------------------------------------------------
struct DenseNdArgs {
  const void *in;
  void *out;
  const float *taps;
  int n[MAXD];
  int o[MAXD];
  int of[MAXD];   // out_first + kernel first
  int kn[MAXD];
  int style;
  int dt_in, dt_out;
  float fill;
  int64_t total;
};

__global__ void __launch_bounds__(256) k_dense_nd(DenseNdArgs a) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < a.total;
       q += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = q;
    int c0 = (int)(r % a.o[0]); r /= a.o[0];
    int c1 = (int)(r % a.o[1]); r /= a.o[1];
    int c2 = (int)(r % a.o[2]); r /= a.o[2];
    int c3 = (int)r;
    float acc = 0.f;
    int t = 0;
    for (int j3 = 0; j3 < a.kn[3]; ++j3) {
      int s3 = bmap(a.style, c3 + a.of[3] + j3, a.n[3]);
      for (int j2 = 0; j2 < a.kn[2]; ++j2) {
        int s2 = bmap(a.style, c2 + a.of[2] + j2, a.n[2]);
        for (int j1 = 0; j1 < a.kn[1]; ++j1) {
          int s1 = bmap(a.style, c1 + a.of[1] + j1, a.n[1]);
          int64_t rowbase = (int64_t)a.n[0] * (s1 + (int64_t)a.n[1] * (s2 + (int64_t)a.n[2] * s3));
          bool rowfill = (s1 | s2 | s3) < 0;
          for (int j0 = 0; j0 < a.kn[0]; ++j0, ++t) {
            int s0 = bmap(a.style, c0 + a.of[0] + j0, a.n[0]);
            float v = (rowfill || s0 < 0) ? a.fill : load_f(a.in, rowbase + s0, a.dt_in);
            acc = fmaf(v, a.taps[t], acc);
          }
        }
      }
    }
    store_f(a.out, q, a.dt_out, acc);
  }
}

// ---------------------------------------------------------------------------
// 2-D VALU: output tile 64 x 32, 256 threads; thread = one column x 8 rows.
// The column of 8 + Ky - 1 tile values is swept once per kernel column so each
// LDS read feeds up to 8 FMAs; taps are wave-uniform scalar loads.
// ---------------------------------------------------------------------------
constexpr int DV_TX = 64;
constexpr int DV_TY = 32;
constexpr int DV_R = 8;

struct Dense2dArgs {
  const void *in;
  void *out;
  const float *taps;  // Kx * Ky, column-major (axis 0 fastest)
  int n0, n1;
  int o0, o1;
  int f0, f1;         // out_first
  int k0, k1;
  int first0, first1;
  int64_t in_plane, out_plane;
  int nplanes, o2, n2, f2, f3;
  int style;
  int dt_in, dt_out;
  float fill;
};

__global__ void __launch_bounds__(256) k_dense2d_valu(Dense2dArgs a) {
  extern __shared__ float smem[];
  const int IN_W = DV_TX + a.k0 - 1;
  const int IN_H = DV_TY + a.k1 - 1;
  const int IN_WP = IN_W | 1;
  const int tid = threadIdx.x;
  const int tx = tid & 63, ty = tid >> 6;
  const int x0 = a.f0 + blockIdx.x * DV_TX;
  const int y0 = a.f1 + blockIdx.y * DV_TY;
  const float *__restrict__ k = a.taps;
  for (int plane = blockIdx.z; plane < a.nplanes; plane += gridDim.z) {
    const int p2 = plane % a.o2, p3 = plane / a.o2;
    const int64_t ibase = ((int64_t)(p2 + a.f2) + (int64_t)a.n2 * (p3 + a.f3)) * a.in_plane;
    for (int r = ty; r < IN_H; r += 4) {
      int sy = bmap(a.style, y0 + a.first1 + r, a.n1);
      for (int c = tx; c < IN_W; c += 64) {
        int sx = bmap(a.style, x0 + a.first0 + c, a.n0);
        smem[r * IN_WP + c] = (sy < 0 || sx < 0) ? a.fill : load_f(a.in, ibase + (int64_t)sy * a.n0 + sx, a.dt_in);
      }
    }
    __syncthreads();
    float acc[DV_R];
#pragma unroll
    for (int o = 0; o < DV_R; ++o) acc[o] = 0.f;
    const int r0 = ty * DV_R;
    for (int t = 0; t < a.k0; ++t) {
      const float *col = smem + r0 * IN_WP + tx + t;
      for (int u = 0; u < a.k1 + DV_R - 1; ++u) {
        float v = col[u * IN_WP];
#pragma unroll
        for (int o = 0; o < DV_R; ++o) {
          int j = u - o;
          if (j >= 0 && j < a.k1) acc[o] = fmaf(k[t + j * a.k0], v, acc[o]);
        }
      }
    }
    const int ox = blockIdx.x * DV_TX + tx;
    if (ox < a.o0) {
      const int64_t obase = (int64_t)plane * a.out_plane;
#pragma unroll
      for (int o = 0; o < DV_R; ++o) {
        int oy = blockIdx.y * DV_TY + r0 + o;
        if (oy < a.o1) store_f(a.out, obase + (int64_t)oy * a.o0 + ox, a.dt_out, acc[o]);
      }
    }
    __syncthreads();
  }
}

hipError_t launch_dense_valu(const Dense4 &in, Dense4 &out, const int64_t *out_first, const DenseKernel &k,
                             const float *d_taps, int style, float fill, hipStream_t st) {
  if (out.count() == 0) return hipSuccess;
  bool twod = k.ndim <= 2 || (k.n[2] == 1 && k.n[3] == 1);
  if (twod && k.n[0] <= 64 && k.n[1] <= 64) {
    Dense2dArgs a;
    a.in = in.data;
    a.out = out.data;
    a.taps = d_taps;
    a.n0 = (int)in.n[0];
    a.n1 = (int)in.n[1];
    a.o0 = (int)out.n[0];
    a.o1 = (int)out.n[1];
    a.f0 = (int)out_first[0];
    a.f1 = (int)out_first[1];
    a.k0 = k.n[0];
    a.k1 = k.ndim > 1 ? k.n[1] : 1;
    a.first0 = k.first[0];
    a.first1 = k.ndim > 1 ? k.first[1] : 0;
    a.in_plane = in.n[0] * in.n[1];
    a.out_plane = out.n[0] * out.n[1];
    a.nplanes = (int)(out.n[2] * out.n[3]);
    a.o2 = (int)out.n[2];
    a.n2 = (int)in.n[2];
    a.f2 = (int)out_first[2] + (k.ndim > 2 ? k.first[2] : 0);
    a.f3 = (int)out_first[3] + (k.ndim > 3 ? k.first[3] : 0);
    a.style = style;
    a.dt_in = in.dtype;
    a.dt_out = out.dtype;
    a.fill = fill;
    const int IN_W = DV_TX + a.k0 - 1, IN_H = DV_TY + a.k1 - 1;
    size_t lds = sizeof(float) * (size_t)IN_H * (IN_W | 1);
    dim3 grid((a.o0 + DV_TX - 1) / DV_TX, (a.o1 + DV_TY - 1) / DV_TY, a.nplanes < 65535 ? a.nplanes : 65535);
    hipLaunchKernelGGL(k_dense2d_valu, grid, dim3(256), lds, st, a);
    return hipGetLastError();
  }
  DenseNdArgs a;
  a.in = in.data;
  a.out = out.data;
  a.taps = d_taps;
  for (int d = 0; d < MAXD; ++d) {
    a.n[d] = (int)in.n[d];
    a.o[d] = (int)out.n[d];
    a.kn[d] = d < k.ndim ? k.n[d] : 1;
    a.of[d] = (int)out_first[d] + (d < k.ndim ? k.first[d] : 0);
  }
  a.style = style;
  a.dt_in = in.dtype;
  a.dt_out = out.dtype;
  a.fill = fill;
  a.total = out.count();
  int64_t blocks = (a.total + 255) / 256;
  if (blocks > 1 << 20) blocks = 1 << 20;
  hipLaunchKernelGGL(k_dense_nd, dim3((unsigned)blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// MFMA banded-Toeplitz.
//
// For kernel row u, the correlation along x of 16 output columns a with the
// input columns s = a + t is a 16 x S product with the Toeplitz matrix
// T_u[s][a] = k[s - a, u] (zero outside 0 <= s-a < Kx), S = 16 + Kx - 1
// rounded up to 32-wide K blocks.  One MFMA 16x16x32 then computes
//     D[yy][a] += sum_s F[yy][s] * T_u[s][a]
// where F holds image rows.  A wave owns B accumulator blocks whose rows are
// interleaved (block b, lane-row yy -> output row y0 + b + B*yy), so the image
// fragment F_r (tile rows r + B*yy) serves every block b with u = r - b: one
// LDS fragment read feeds up to B MFMA groups.  The T_u fragments for a chunk
// of UCH kernel rows live in registers.  x = xh + xl and k = kh + kl in bf16;
// xh*kh + xl*kh + xh*kl accumulate in FP32 (xl*kl, ~2^-16 relative, is
// dropped; for bf16 input xl == 0 and its product is skipped).
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

constexpr int MF_WAVES = 4;
constexpr int MF_TX = 16 * MF_WAVES;  // output columns per workgroup

struct MfmaArgs {
  const void *in;
  void *out;
  const float *taps;  // Kx * Ky, column-major
  int n0, n1;
  int o0, o1;
  int f0, f1;
  int kx, ky;
  int first0, first1;
  int kb;             // K blocks of 32 input columns
  int twp;            // LDS row pitch (elements)
  int th;             // LDS rows
  int64_t in_plane, out_plane;
  int nplanes, o2, n2, f2, f3;
  int style;
  int dt_in, dt_out;
  int split_x;        // 0 when the input is bf16 (xl == 0)
  float fill;
};

__device__ __forceinline__ void split_bf16(float v, uint16_t &hi, uint16_t &lo) {
  hi = f32_to_bf16(v);
  lo = f32_to_bf16(v - bf16_to_f32(hi));
}

template <int B, int UCH>
__global__ void __launch_bounds__(256) k_dense_mfma(MfmaArgs a) {
  extern __shared__ __attribute__((aligned(16))) uint16_t lds16[];
  const int TWP = a.twp, TH = a.th;
  uint16_t *s_hi = lds16;
  uint16_t *s_lo = lds16 + (size_t)TH * TWP;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int x0 = a.f0 + blockIdx.x * MF_TX;
  const int y0 = a.f1 + blockIdx.y * (16 * B);
  const int TW = MF_TX + a.kx - 1;
  const int TH_valid = 16 * B + a.ky - 1;
  const int lrow = lane & 15, lgrp = lane >> 4;

  for (int plane = blockIdx.z; plane < a.nplanes; plane += gridDim.z) {
    const int p2 = plane % a.o2, p3 = plane / a.o2;
    const int64_t ibase = ((int64_t)(p2 + a.f2) + (int64_t)a.n2 * (p3 + a.f3)) * a.in_plane;
    // Stage the input tile as bf16 hi/lo planes (border map on load; padding
    // columns and rows are zero so they meet only zero Toeplitz entries).
    for (int r = wave; r < TH; r += MF_WAVES) {
      int sy = (r < TH_valid) ? bmap(a.style, y0 + a.first1 + r, a.n1) : -2;
      for (int c = lane; c < TWP; c += 64) {
        float v = 0.f;
        if (sy != -2 && c < TW) {
          int sx = bmap(a.style, x0 + a.first0 + c, a.n0);
          v = (sy < 0 || sx < 0) ? a.fill : load_f(a.in, ibase + (int64_t)sy * a.n0 + sx, a.dt_in);
        }
        uint16_t h, l;
        split_bf16(v, h, l);
        s_hi[r * TWP + c] = h;
        s_lo[r * TWP + c] = l;
      }
    }
    __syncthreads();

    f32x4 acc[B];
#pragma unroll
    for (int b = 0; b < B; ++b) acc[b] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int u0 = 0; u0 < a.ky; u0 += UCH) {
      for (int kb = 0; kb < a.kb; ++kb) {
        // Toeplitz fragments for this chunk: lane holds T[s][a] for
        // a = lane%16, s = 32kb + 8*(lane/16) + i.
        bf16x8 thi[UCH], tlo[UCH];
#pragma unroll
        for (int u = 0; u < UCH; ++u) {
          u16x8 h, l;
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            int s = 32 * kb + 8 * lgrp + i;
            int t = s - lrow;
            float kv = 0.f;
            if (t >= 0 && t < a.kx && u0 + u < a.ky) kv = a.taps[t + (u0 + u) * a.kx];
            uint16_t hh, ll;
            split_bf16(kv, hh, ll);
            h[i] = hh;
            l[i] = ll;
          }
          thi[u] = __builtin_bit_cast(bf16x8, h);
          tlo[u] = __builtin_bit_cast(bf16x8, l);
        }
        const int colbase = 16 * wave + 32 * kb + 8 * lgrp;
#pragma unroll
        for (int r = 0; r < B + UCH - 1; ++r) {
          const int row = u0 + r + B * lrow;
          const bf16x8 fhi = *(const bf16x8 *)(s_hi + row * TWP + colbase);
          const bf16x8 flo = *(const bf16x8 *)(s_lo + row * TWP + colbase);
#pragma unroll
          for (int b = 0; b < B; ++b) {
            const int u = r - b;
            if (u >= 0 && u < UCH && u0 + u < a.ky)
              acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fhi, thi[u], acc[b], 0, 0, 0);
          }
          if (a.split_x) {
#pragma unroll
            for (int b = 0; b < B; ++b) {
              const int u = r - b;
              if (u >= 0 && u < UCH && u0 + u < a.ky)
                acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(flo, thi[u], acc[b], 0, 0, 0);
            }
          }
#pragma unroll
          for (int b = 0; b < B; ++b) {
            const int u = r - b;
            if (u >= 0 && u < UCH && u0 + u < a.ky)
              acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fhi, tlo[u], acc[b], 0, 0, 0);
          }
        }
      }
    }
    // D layout: lane holds rows yy = 4*(lane/16) + i of column a = lane%16.
    const int ox = blockIdx.x * MF_TX + 16 * wave + lrow;
    if (ox < a.o0) {
      const int64_t obase = (int64_t)plane * a.out_plane;
#pragma unroll
      for (int b = 0; b < B; ++b) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          int oy = blockIdx.y * (16 * B) + b + B * (4 * lgrp + i);
          if (oy < a.o1) store_f(a.out, obase + (int64_t)oy * a.o0 + ox, a.dt_out, acc[b][i]);
        }
      }
    }
    __syncthreads();
  }
}

static int mfma_rows_per_block(const DenseKernel &k, int *twp_out, int *th_out, int *kb_out) {
  const int UCH = 16;
  int kx = k.n[0], ky = k.ndim > 1 ? k.n[1] : 1;
  int kb = (16 + kx - 1 + 31) / 32;
  int twp = 16 * (MF_WAVES - 1) + 32 * kb;  // covers every fragment column
  int tw = MF_TX + kx - 1;
  if (twp < tw) twp = tw;
  twp = (twp + 7) & ~7;
  int chunks = (ky + UCH - 1) / UCH;
  for (int B = 8; B >= 4; B /= 2) {
    int th = 16 * B + chunks * UCH - 1;
    size_t lds = (size_t)th * twp * 2 * sizeof(uint16_t);
    if (lds <= 64 * 1024) {
      *twp_out = twp;
      *th_out = th;
      *kb_out = kb;
      return B;
    }
  }
  return 0;
}

bool dense_mfma_supported(const DenseKernel &k, const Dense4 &in) {
  bool twod = k.ndim == 2 || (k.ndim > 2 && k.n[2] == 1 && (k.ndim < 4 || k.n[3] == 1));
  if (!twod || k.n[0] > 64 || k.n[1] > 64) return false;
  int twp, th, kb;
  return mfma_rows_per_block(k, &twp, &th, &kb) != 0 && in.dtype != IFM_F64;
}

hipError_t launch_dense_mfma(const Dense4 &in, Dense4 &out, const int64_t *out_first, const DenseKernel &k,
                             void *d_work, size_t work_bytes, int style, float fill, hipStream_t st) {
  (void)work_bytes;
  if (out.count() == 0) return hipSuccess;
  MfmaArgs a;
  int twp, th, kb;
  int B = mfma_rows_per_block(k, &twp, &th, &kb);
  a.in = in.data;
  a.out = out.data;
  a.taps = (const float *)d_work;
  a.n0 = (int)in.n[0];
  a.n1 = (int)in.n[1];
  a.o0 = (int)out.n[0];
  a.o1 = (int)out.n[1];
  a.f0 = (int)out_first[0];
  a.f1 = (int)out_first[1];
  a.kx = k.n[0];
  a.ky = k.n[1];
  a.first0 = k.first[0];
  a.first1 = k.first[1];
  a.kb = kb;
  a.twp = twp;
  a.th = th;
  a.in_plane = in.n[0] * in.n[1];
  a.out_plane = out.n[0] * out.n[1];
  a.nplanes = (int)(out.n[2] * out.n[3]);
  a.o2 = (int)out.n[2];
  a.n2 = (int)in.n[2];
  a.f2 = (int)out_first[2];
  a.f3 = (int)out_first[3];
  a.style = style;
  a.dt_in = in.dtype;
  a.dt_out = out.dtype;
  a.split_x = in.dtype == IFM_BF16 || in.dtype == IFM_U8 || in.dtype == IFM_N0F8 ? 0 : 1;
  if (in.dtype == IFM_N0F8) a.split_x = 1;  // code/255 is not a bf16 value
  a.fill = fill;
  size_t lds = (size_t)th * twp * 2 * sizeof(uint16_t);
  dim3 grid((a.o0 + MF_TX - 1) / MF_TX, (a.o1 + 16 * B - 1) / (16 * B), a.nplanes < 65535 ? a.nplanes : 65535);
  if (B == 8)
    hipLaunchKernelGGL((k_dense_mfma<8, 16>), grid, dim3(256), lds, st, a);
  else
    hipLaunchKernelGGL((k_dense_mfma<4, 16>), grid, dim3(256), lds, st, a);
  return hipGetLastError();
}

size_t dense_mfma_workspace(const DenseKernel &k) { return k.taps.size() * sizeof(float); }

}  // namespace ifm
