// k_dense.hip — dense (non-separable) FIR correlation (gfx950).
//
// Replaces __imfilter_inbounds! (src/imfilter.jl:592-669): out[i] =
// sum_J A[i+J] k[J] on the padded image.  The border map is applied while the
// input tile is staged into LDS.
//
//  * k_dense_nd     any rank <= 4, one output per thread (small kernels, 3-D)
//  * k_dense2d_valu 2-D, LDS tile, 8 rows per thread, FP32 FMA
//  * k_dense_mfma   2-D banded-Toeplitz on MFMA v_mfma_f32_16x16x32_f16 with
//                   scaled x and k split into FP16 hi+lo (3 products, FP32 accumulate)
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
// input columns s = a + t is a product with the Toeplitz matrix
// T_u[s][a] = k[s - a, u] (zero outside 0 <= s-a < Kx), s < S = 16 + Kx - 1
// taken in K blocks of 32.  One v_mfma_f32_16x16x32_f16 computes
//     D[yy][a] += sum_s F[yy][s] * T_u[s][a]
// with F holding image rows.  A wave owns B accumulator blocks whose rows are
// interleaved (block b, lane-row yy -> output row y0 + b + B*yy), so the image
// fragment F_r (tile rows r + B*yy) serves every block b with u = r - b: one
// LDS fragment read feeds up to B MFMA groups.
//
// Precision: x and k are split into FP16 hi + lo after scaling by powers of
// two (k by one factor per kernel, x by one factor per tile so the tile's
// largest |x| lies in [0.5, 1)); xh*kh + xl*kh + xh*kl accumulate in FP32.
// Each operand keeps ~22 significant bits, so a product carries a relative
// error near 2^-22 (the dropped xl*kl is below 2^-24), comparable to a plain
// FP32 sum.  Tiles holding Inf or NaN are computed by a direct FP32 loop
// instead, so non-finite values reach exactly the outputs whose window holds
// them, as in the reference's direct sum.
//
// The Toeplitz fragments are built on the host once per kernel: for each
// (K block, kernel row u) a lane holds T[s][a] for a = lane%16,
// s = 32*kb + 8*(lane/16) + i, i < 8, as FP16 hi and lo.
// ---------------------------------------------------------------------------
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int MF_WAVES = 4;
constexpr int MF_TX = 16 * MF_WAVES;  // output columns per workgroup
constexpr int MF_B = 8;               // interleaved row blocks per wave
constexpr int MF_UCH = 8;             // kernel rows whose fragments are held in registers
constexpr int MF_MAXCH = 8;           // kernel rows up to 64

struct MfmaArgs {
  const void *in;
  void *out;
  const uint16_t *ttab;  // [kb][u][hi/lo][lane][8] FP16 Toeplitz fragments (u padded to chunks)
  const float *taps;     // Kx * Ky FP32 taps (fallback path)
  float kinv;            // 1 / kernel scale
  int n0, n1;
  int o0, o1;
  int f0, f1;
  int kx, ky;
  int first0, first1;
  int nch;               // kernel-row chunks of MF_UCH
  int th;                // LDS rows used (16B + nch*UCH - 1)
  int64_t in_plane, out_plane;
  int nplanes, o2, n2, f2, f3;
  int style;
  int dt_in, dt_out;
  float fill;
};

template <int KB>
struct MfmaGeom {
  static constexpr int TWP = 48 + 32 * KB;  // LDS row pitch (32-bit words)
};

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

template <int KB>
__global__ void __launch_bounds__(256, 2) k_dense_mfma(MfmaArgs a) {
  using G = MfmaGeom<KB>;
  constexpr int TWP = G::TWP;
  // One 32-bit word per tile element: first the FP32 value, then (in place)
  // the packed FP16 pair hi | lo << 16.
  extern __shared__ __attribute__((aligned(16))) uint32_t lds32[];
  __shared__ float red[MF_WAVES];
  __shared__ int redflag[MF_WAVES];
  const int TH = a.th;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int lrow = lane & 15, lgrp = lane >> 4;
  const int x0 = a.f0 + blockIdx.x * MF_TX;
  const int y0 = a.f1 + blockIdx.y * (16 * MF_B);
  const int TW = MF_TX + a.kx - 1;
  const int TH_valid = 16 * MF_B + a.ky - 1;
  const int total = TH * TWP;

  for (int plane = blockIdx.z; plane < a.nplanes; plane += gridDim.z) {
    const int p2 = plane % a.o2, p3 = plane / a.o2;
    const int64_t ibase = ((int64_t)(p2 + a.f2) + (int64_t)a.n2 * (p3 + a.f3)) * a.in_plane;
    // Stage the tile as FP32 (border map on load; padding words are zero);
    // track max |x| and non-finite values.
    float amax = 0.f;
    bool bad = false;
    for (int idx = tid; idx < total; idx += 256) {
      const int r = idx / TWP, c = idx - r * TWP;
      float x = 0.f;
      if (r < TH_valid && c < TW) {
        int sy = bmap(a.style, y0 + a.first1 + r, a.n1);
        int sx = bmap(a.style, x0 + a.first0 + c, a.n0);
        x = (sy < 0 || sx < 0) ? a.fill : load_f(a.in, ibase + (int64_t)sy * a.n0 + sx, a.dt_in);
      }
      lds32[idx] = __float_as_uint(x);
      const float ax = fabsf(x);
      bad |= !(ax <= 3.0e38f);  // Inf or NaN
      amax = fmaxf(amax, ax);
    }
    amax = wave_max(amax);
    const bool wbad = __any(bad);
    if (lane == 0) {
      red[wave] = amax;
      redflag[wave] = wbad;
    }
    __syncthreads();
    amax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    const bool tilebad = redflag[0] | redflag[1] | redflag[2] | redflag[3];
    const int64_t obase = (int64_t)plane * a.out_plane;

    if (tilebad) {
      // Direct FP32 sums from the image (border map per tap) for this tile.
      for (int q = tid; q < MF_TX * 16 * MF_B; q += 256) {
        const int ox = blockIdx.x * MF_TX + (q % MF_TX);
        const int oy = blockIdx.y * (16 * MF_B) + q / MF_TX;
        if (ox >= a.o0 || oy >= a.o1) continue;
        float acc = 0.f;
        for (int u = 0; u < a.ky; ++u) {
          int sy = bmap(a.style, a.f1 + oy + a.first1 + u, a.n1);
          for (int t = 0; t < a.kx; ++t) {
            int sx = bmap(a.style, a.f0 + ox + a.first0 + t, a.n0);
            float x = (sy < 0 || sx < 0) ? a.fill : load_f(a.in, ibase + (int64_t)sy * a.n0 + sx, a.dt_in);
            acc = fmaf(x, a.taps[t + u * a.kx], acc);
          }
        }
        store_f(a.out, obase + (int64_t)oy * a.o0 + ox, a.dt_out, acc);
      }
      __syncthreads();
      continue;
    }

    // Power-of-two scale putting the tile's largest |x| in [0.5, 1).
    int e = 0;
    if (amax > 0.f) frexpf(amax, &e);
    e = e < -100 ? -100 : (e > 100 ? 100 : e);
    const float xs = ldexpf(1.0f, -e);
    const float xinv = ldexpf(1.0f, e);
    for (int idx = tid; idx < total; idx += 256) {  // each thread converts the words it wrote
      const float xv = __uint_as_float(lds32[idx]) * xs;
      const _Float16 h = (_Float16)xv;
      const _Float16 l = (_Float16)(xv - (float)h);
      lds32[idx] = (uint32_t)__builtin_bit_cast(uint16_t, h) | ((uint32_t)__builtin_bit_cast(uint16_t, l) << 16);
    }
    __syncthreads();

    f32x4 acc[MF_B];
#pragma unroll
    for (int b = 0; b < MF_B; ++b) acc[b] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int kb = 0; kb < KB; ++kb) {
      const int colbase = 16 * wave + 32 * kb + 8 * lgrp;
      for (int ch = 0; ch < a.nch; ++ch) {
        const int u0 = ch * MF_UCH;
        f16x8 thi[MF_UCH], tlo[MF_UCH];
        const u16x8 *tt = (const u16x8 *)a.ttab + ((size_t)(kb * a.nch * MF_UCH + u0) * 2) * 64 + lane;
#pragma unroll
        for (int u = 0; u < MF_UCH; ++u) {
          thi[u] = __builtin_bit_cast(f16x8, tt[(2 * u) * 64]);
          tlo[u] = __builtin_bit_cast(f16x8, tt[(2 * u + 1) * 64]);
        }
        // Rows beyond Ky have zero Toeplitz fragments, so every chunk runs the
        // same straight-line schedule.  Fragment r+1 is read while the MFMAs
        // of fragment r run; the scheduling barriers keep that pipeline (and
        // the register budget) instead of hoisting every LDS read.
        const u32x4 *src0 = (const u32x4 *)(lds32 + (u0 + MF_B * lrow) * TWP + colbase);
        u32x4 w0 = src0[0], w1 = src0[1];
#pragma unroll
        for (int r = 0; r < MF_B + MF_UCH - 1; ++r) {
          u32x4 n0 = w0, n1 = w1;
          if (r + 1 < MF_B + MF_UCH - 1) {
            const u32x4 *src = (const u32x4 *)(lds32 + (u0 + r + 1 + MF_B * lrow) * TWP + colbase);
            n0 = src[0];
            n1 = src[1];
          }
          // 8 packed words -> 8 hi halves and 8 lo halves (v_perm_b32)
          u32x4 ph, pl;
          ph[0] = __builtin_amdgcn_perm(w0[1], w0[0], 0x05040100u);
          ph[1] = __builtin_amdgcn_perm(w0[3], w0[2], 0x05040100u);
          ph[2] = __builtin_amdgcn_perm(w1[1], w1[0], 0x05040100u);
          ph[3] = __builtin_amdgcn_perm(w1[3], w1[2], 0x05040100u);
          pl[0] = __builtin_amdgcn_perm(w0[1], w0[0], 0x07060302u);
          pl[1] = __builtin_amdgcn_perm(w0[3], w0[2], 0x07060302u);
          pl[2] = __builtin_amdgcn_perm(w1[1], w1[0], 0x07060302u);
          pl[3] = __builtin_amdgcn_perm(w1[3], w1[2], 0x07060302u);
          const f16x8 fhi = __builtin_bit_cast(f16x8, ph);
          const f16x8 flo = __builtin_bit_cast(f16x8, pl);
#pragma unroll
          for (int b = 0; b < MF_B; ++b) {
            const int u = r - b;
            if (u >= 0 && u < MF_UCH)
              acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fhi, thi[u], acc[b], 0, 0, 0);
          }
#pragma unroll
          for (int b = 0; b < MF_B; ++b) {
            const int u = r - b;
            if (u >= 0 && u < MF_UCH)
              acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(flo, thi[u], acc[b], 0, 0, 0);
          }
#pragma unroll
          for (int b = 0; b < MF_B; ++b) {
            const int u = r - b;
            if (u >= 0 && u < MF_UCH)
              acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fhi, tlo[u], acc[b], 0, 0, 0);
          }
          w0 = n0;
          w1 = n1;
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
    // D layout: lane holds rows yy = 4*(lane/16) + i of column a = lane%16.
    const float oscale = xinv * a.kinv;
    const int ox = blockIdx.x * MF_TX + 16 * wave + lrow;
    if (ox < a.o0) {
#pragma unroll
      for (int b = 0; b < MF_B; ++b) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          int oy = blockIdx.y * (16 * MF_B) + b + MF_B * (4 * lgrp + i);
          if (oy < a.o1) store_f(a.out, obase + (int64_t)oy * a.o0 + ox, a.dt_out, acc[b][i] * oscale);
        }
      }
    }
    __syncthreads();
  }
}

static int mfma_kb(int kx) { return (16 + kx - 1 + 31) / 32; }

static size_t mfma_lds_bytes(const DenseKernel &k) {
  const int nch = (k.n[1] + MF_UCH - 1) / MF_UCH;
  const int th = 16 * MF_B + nch * MF_UCH - 1;
  return (size_t)th * (48 + 32 * mfma_kb(k.n[0])) * 2 * sizeof(uint16_t);
}

static int max_lds_per_block() {
  static int v = -1;
  if (v < 0) {
    int dev = 0, x = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&x, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      x = 64 * 1024;
    v = x;
  }
  return v;
}

bool dense_mfma_supported(const DenseKernel &k, const Dense4 &in) {
  (void)in;
  bool twod = k.n[2] == 1 && k.n[3] == 1;
  if (!(twod && k.n[0] <= 64 && k.n[1] <= MF_MAXCH * MF_UCH && mfma_kb(k.n[0]) <= 3)) return false;
  // LDS for the tile plus the block's static reduction scratch
  return mfma_lds_bytes(k) + 64 <= (size_t)max_lds_per_block();
}

// Host-side Toeplitz fragment table (FP16 hi/lo of the scaled taps).
void dense_mfma_table(const DenseKernel &k, std::vector<uint16_t> &tab, float *kinv) {
  const int kx = k.n[0], ky = k.n[1];
  const int kb = mfma_kb(kx), nch = (ky + MF_UCH - 1) / MF_UCH;
  float amax = 0.f;
  for (float t : k.taps) amax = fmaxf(amax, fabsf(t));
  int e = 0;
  if (amax > 0.f) frexpf(amax, &e);
  const float ks = ldexpf(1.0f, -e);
  *kinv = ldexpf(1.0f, e);
  const int urows = nch * MF_UCH;
  tab.assign((size_t)kb * urows * 2 * 64 * 8, 0);
  for (int b = 0; b < kb; ++b)
    for (int u = 0; u < urows; ++u)
      for (int lane = 0; lane < 64; ++lane)
        for (int i = 0; i < 8; ++i) {
          int s = 32 * b + 8 * (lane >> 4) + i;
          int t = s - (lane & 15);
          float kv = (t >= 0 && t < kx && u < ky) ? k.taps[t + u * kx] * ks : 0.f;
          _Float16 h = (_Float16)kv;
          _Float16 l = (_Float16)(kv - (float)h);
          size_t base = ((((size_t)b * urows + u) * 2) * 64 + lane) * 8 + i;
          tab[base] = __builtin_bit_cast(uint16_t, h);
          tab[base + 64 * 8] = __builtin_bit_cast(uint16_t, l);
        }
}

hipError_t launch_dense_mfma(const Dense4 &in, Dense4 &out, const int64_t *out_first, const DenseKernel &k,
                             const uint16_t *d_table, const float *d_taps, float kinv, int style, float fill,
                             hipStream_t st) {
  if (out.count() == 0) return hipSuccess;
  MfmaArgs a;
  const int KB = mfma_kb(k.n[0]);
  a.in = in.data;
  a.out = out.data;
  a.ttab = d_table;
  a.taps = d_taps;
  a.kinv = kinv;
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
  a.nch = (k.n[1] + MF_UCH - 1) / MF_UCH;
  a.th = 16 * MF_B + a.nch * MF_UCH - 1;
  a.in_plane = in.n[0] * in.n[1];
  a.out_plane = out.n[0] * out.n[1];
  a.nplanes = (int)(out.n[2] * out.n[3]);
  a.o2 = (int)out.n[2];
  a.n2 = (int)in.n[2];
  a.f2 = (int)out_first[2] + k.first[2];
  a.f3 = (int)out_first[3] + k.first[3];
  a.style = style;
  a.dt_in = in.dtype;
  a.dt_out = out.dtype;
  a.fill = fill;
  dim3 grid((a.o0 + MF_TX - 1) / MF_TX, (a.o1 + 16 * MF_B - 1) / (16 * MF_B),
            a.nplanes < 65535 ? a.nplanes : 65535);
  const int twp = 48 + 32 * KB;
  size_t lds = (size_t)a.th * twp * 2 * sizeof(uint16_t);
  if (KB == 1)
    hipLaunchKernelGGL(k_dense_mfma<1>, grid, dim3(256), lds, st, a);
  else if (KB == 2)
    hipLaunchKernelGGL(k_dense_mfma<2>, grid, dim3(256), lds, st, a);
  else
    hipLaunchKernelGGL(k_dense_mfma<3>, grid, dim3(256), lds, st, a);
  return hipGetLastError();
}

}  // namespace ifm
