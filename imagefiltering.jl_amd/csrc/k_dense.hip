// k_dense.hip — dense (non-separable) FIR correlation on the vector ALUs (gfx950).
//
// Replaces __imfilter_inbounds! (src/imfilter.jl:592-669): out[i] =
// sum_J A[i+J] k[J] on the padded image.  The border map is applied while the
// input tile is staged into LDS.
//
//  * k_dense_nd     any rank <= 4, one output per thread (other kernels)
//  * k_dense3d_small 3-D kernels of 3 or 5 taps per axis, register-blocked
//  * k_dense2d_valu 2-D, LDS tile, 8 rows per thread, FP32 FMA
//  * k_dense2d_small the same, register-blocked for odd kernel sizes 3 .. 15
// The matrix-core path for 2-D kernels is k_mfma.hip.
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

// ---------------------------------------------------------------------------
// 2-D VALU, register-blocked for compile-time kernel sizes KX x KY (odd, 3 ..
// 15): the same 64 x 32 tile and thread shape as k_dense2d_valu, but every
// loop is unrolled, the taps arrive as kernel arguments (scalar registers,
// v_fmac with an SGPR operand), and each LDS value read feeds up to KY * 8 /
// (KY + 7) FMAs through a sliding column of KX values.
//
// Staging.  Element e = tid + 256 i of the (64 + KX - 1) x (32 + KY - 1) input
// tile (row-major) is loaded by thread tid: a wave reads one or two
// contiguous runs, and for FP32 input every load of the tile is issued before
// the first LDS write.  A tile whose window reaches past the image edge maps
// its rows and columns once per workgroup into LDS (one bmap per row and per
// column, not per element), so border tiles cost about what interior tiles
// do.  Tiles are assigned XCD by XCD (xcd_tile), so the halo columns two
// neighbouring tiles share are fetched into one L2.
// ---------------------------------------------------------------------------
template <int KX, int KY>
struct SmallTaps {
  float k[KX * KY];  // column-major (axis 0 fastest)
};

template <int KX, int KY>
__global__ void __launch_bounds__(256) k_dense2d_small(Dense2dArgs a, SmallTaps<KX, KY> kt) {
  constexpr int IN_W = DV_TX + KX - 1, IN_H = DV_TY + KY - 1, IN_WP = IN_W | 1;
  constexpr int NE = IN_H * IN_W, NI = (NE + 255) / 256;
  __shared__ float smem[IN_H * IN_WP];
  __shared__ int smap[IN_H + IN_W];  // border tiles: source row of tile row r, then source column of tile column c
  const int tid = threadIdx.x;
  const int tx = tid & 63, ty = tid >> 6;
  const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int bx = tile % gridDim.x, by = tile / gridDim.x;
  const int gx = a.f0 + bx * DV_TX + a.first0;  // image index of staged (0, 0)
  const int gy = a.f1 + by * DV_TY + a.first1;
  const bool interior = gx >= 0 && gx + IN_W <= a.n0 && gy >= 0 && gy + IN_H <= a.n1;
  if (!interior) {
    for (int i = tid; i < IN_H + IN_W; i += 256)
      smap[i] = i < IN_H ? bmap(a.style, gy + i, a.n1) : bmap(a.style, gx + i - IN_H, a.n0);
    __syncthreads();
  }
  for (int plane = blockIdx.z; plane < a.nplanes; plane += gridDim.z) {
    const int p2 = plane % a.o2, p3 = plane / a.o2;
    const int64_t ibase = ((int64_t)(p2 + a.f2) + (int64_t)a.n2 * (p3 + a.f3)) * a.in_plane;
    if (a.dt_in == IFM_F32) {
      const float *src = (const float *)a.in + ibase;
      float x[NI];
      if (interior) {
        const float *t0 = src + (int64_t)gy * a.n0 + gx;
#pragma unroll
        for (int i = 0; i < NI; ++i) {
          const int e = tid + 256 * i;
          const int r = e / IN_W, c = e - r * IN_W;
          if (e < NE) x[i] = t0[(int64_t)r * a.n0 + c];
        }
      } else {
#pragma unroll
        for (int i = 0; i < NI; ++i) {
          const int e = tid + 256 * i;
          const int r = e / IN_W, c = e - r * IN_W;
          if (e < NE) {
            const int sy = smap[r], sx = smap[IN_H + c];
            x[i] = (sy < 0 || sx < 0) ? a.fill : src[(int64_t)sy * a.n0 + sx];
          }
        }
      }
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const int e = tid + 256 * i;
        const int r = e / IN_W, c = e - r * IN_W;
        if (e < NE) smem[r * IN_WP + c] = x[i];
      }
    } else {
      for (int e = tid; e < NE; e += 256) {
        const int r = e / IN_W, c = e - r * IN_W;
        const int sy = interior ? gy + r : smap[r], sx = interior ? gx + c : smap[IN_H + c];
        smem[r * IN_WP + c] = (sy < 0 || sx < 0) ? a.fill : load_f(a.in, ibase + (int64_t)sy * a.n0 + sx, a.dt_in);
      }
    }
    __syncthreads();
    float acc[DV_R];
#pragma unroll
    for (int o = 0; o < DV_R; ++o) acc[o] = 0.f;
    const float *base = smem + ty * DV_R * IN_WP + tx;
#pragma unroll
    for (int u = 0; u < KY + DV_R - 1; ++u) {
      float v[KX];
#pragma unroll
      for (int q = 0; q < KX; ++q) v[q] = base[u * IN_WP + q];
#pragma unroll
      for (int o = 0; o < DV_R; ++o) {
        const int j = u - o;
        if (j >= 0 && j < KY) {
#pragma unroll
          for (int q = 0; q < KX; ++q) acc[o] = fmaf(kt.k[q + j * KX], v[q], acc[o]);
        }
      }
    }
    const int ox = bx * DV_TX + tx;
    if (ox < a.o0) {
      const int64_t obase = (int64_t)plane * a.out_plane;
      if (a.dt_out == IFM_F32) {
        float *out = (float *)a.out + obase + ox;
#pragma unroll
        for (int o = 0; o < DV_R; ++o) {
          const int oy = by * DV_TY + ty * DV_R + o;
          if (oy < a.o1) out[(int64_t)oy * a.o0] = acc[o];
        }
      } else {
#pragma unroll
        for (int o = 0; o < DV_R; ++o) {
          const int oy = by * DV_TY + ty * DV_R + o;
          if (oy < a.o1) store_f(a.out, obase + (int64_t)oy * a.o0 + ox, a.dt_out, acc[o]);
        }
      }
    }
    __syncthreads();
  }
}

template <int KX, int KY>
static hipError_t launch_small(const Dense2dArgs &a, const float *taps, dim3 grid, hipStream_t st) {
  SmallTaps<KX, KY> kt;
  for (int i = 0; i < KX * KY; ++i) kt.k[i] = taps[i];
  hipLaunchKernelGGL((k_dense2d_small<KX, KY>), grid, dim3(256), 0, st, a, kt);
  return hipGetLastError();
}

template <int KX>
static hipError_t launch_small_ky(int ky, const Dense2dArgs &a, const float *taps, dim3 grid, hipStream_t st) {
  switch (ky) {
  case 3: return launch_small<KX, 3>(a, taps, grid, st);
  case 5: return launch_small<KX, 5>(a, taps, grid, st);
  case 7: return launch_small<KX, 7>(a, taps, grid, st);
  case 9: return launch_small<KX, 9>(a, taps, grid, st);
  case 11: return launch_small<KX, 11>(a, taps, grid, st);
  case 13: return launch_small<KX, 13>(a, taps, grid, st);
  default: return launch_small<KX, 15>(a, taps, grid, st);
  }
}

static bool small_size(int n) { return n >= 3 && n <= 15 && (n & 1); }

// ---------------------------------------------------------------------------
// 3-D VALU, register-blocked for compile-time kernel sizes KX x KY x KZ (3 or
// 5 per axis): a workgroup owns a 64 x 32 (x, y) tile and marches over D3_NZ
// output planes.  The KZ input planes an output plane needs sit in an LDS
// ring of staged planes (the k_dense2d_small layout); each step commits the
// plane its loads brought in, issues the loads of the next one, and sums the
// KZ plane slices with their kernel slices (taps in scalar registers), so
// every input plane is read from HBM once per workgroup.
// ---------------------------------------------------------------------------
constexpr int D3_NZ = 32;  // output planes per workgroup

struct Dense3dArgs {
  const void *in;
  void *out;
  int n0, n1, n2, n3;           // input extents
  int o0, o1, o2;               // output extents (axis 3 through the grid)
  int g0, g1, g2, g3;           // image index of output (0, 0, 0, 0)'s window start
  int64_t in_plane, out_plane;  // elements per (x, y) plane
  int nzc;                      // z chunks
  int style;
  int dt_in, dt_out;
  float fill;
};

template <int KX, int KY, int KZ>
struct SmallTaps3 {
  float k[KX * KY * KZ];  // column-major (axis 0 fastest)
};

template <int KX, int KY, int KZ>
__global__ void __launch_bounds__(256) k_dense3d_small(Dense3dArgs a, SmallTaps3<KX, KY, KZ> kt) {
  constexpr int IN_W = DV_TX + KX - 1, IN_H = DV_TY + KY - 1, IN_WP = IN_W | 1;
  constexpr int PL = IN_H * IN_WP;  // floats per staged plane
  constexpr int NE = IN_H * IN_W, NI = (NE + 255) / 256;
  __shared__ float smem[KZ * PL];
  __shared__ int smap[IN_H + IN_W];
  const int tid = threadIdx.x;
  const int tx = tid & 63, ty = tid >> 6;
  const int nxy = gridDim.x * gridDim.y;
  const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y + nxy * blockIdx.z, nxy * gridDim.z);
  const int bz = tile / nxy, bx = tile % gridDim.x, by = (tile % nxy) / gridDim.x;
  const int zc = bz % a.nzc, p3 = bz / a.nzc;
  const int oz0 = zc * D3_NZ, nz = min(D3_NZ, a.o2 - oz0);
  const int gx = a.g0 + bx * DV_TX, gy = a.g1 + by * DV_TY;
  const bool interior = gx >= 0 && gx + IN_W <= a.n0 && gy >= 0 && gy + IN_H <= a.n1;
  if (!interior) {
    for (int i = tid; i < IN_H + IN_W; i += 256)
      smap[i] = i < IN_H ? bmap(a.style, gy + i, a.n1) : bmap(a.style, gx + i - IN_H, a.n0);
    __syncthreads();
  }
  const int s3 = bmap(a.style, p3 + a.g3, a.n3);
  const int64_t vbase = s3 < 0 ? 0 : (int64_t)s3 * a.n2 * a.in_plane;
  const bool f32 = a.dt_in == IFM_F32;
  // input plane of relative index j (-1: a Fill plane)
  auto zin = [&](int j) {
    const int z = bmap(a.style, a.g2 + oz0 + j, a.n2);
    return s3 < 0 ? -1 : z;
  };
  float x[NI];
  auto fetch = [&](int zi) {  // FP32: the plane's loads into x
    if (!f32) return;
    const float *src = (const float *)a.in + vbase + (int64_t)(zi < 0 ? 0 : zi) * a.in_plane;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int e = tid + 256 * i;
      const int r = e / IN_W, c = e - r * IN_W;
      if (e < NE) {
        if (zi < 0) {
          x[i] = a.fill;
        } else if (interior) {
          x[i] = src[(int64_t)(gy + r) * a.n0 + gx + c];
        } else {
          const int sy = smap[r], sx = smap[IN_H + c];
          x[i] = (sy < 0 || sx < 0) ? a.fill : src[(int64_t)sy * a.n0 + sx];
        }
      }
    }
  };
  auto commit = [&](int zi, int slot) {
    float *dst = smem + slot * PL;
    if (f32) {
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const int e = tid + 256 * i;
        const int r = e / IN_W, c = e - r * IN_W;
        if (e < NE) dst[r * IN_WP + c] = x[i];
      }
    } else {
      for (int e = tid; e < NE; e += 256) {
        const int r = e / IN_W, c = e - r * IN_W;
        const int sy = interior ? gy + r : smap[r], sx = interior ? gx + c : smap[IN_H + c];
        dst[r * IN_WP + c] = (zi < 0 || sy < 0 || sx < 0)
                                 ? a.fill
                                 : load_f(a.in, vbase + (int64_t)zi * a.in_plane + (int64_t)sy * a.n0 + sx, a.dt_in);
      }
    }
  };
  // prologue: input planes 0 .. KZ-2 staged, plane KZ-1 in flight
  for (int j = 0; j < KZ - 1; ++j) {
    const int zi = zin(j);
    fetch(zi);
    commit(zi, j);
  }
  int zn = zin(KZ - 1);
  fetch(zn);
  const int ox = bx * DV_TX + tx;
  for (int q = 0; q < nz; ++q) {
    commit(zn, (q + KZ - 1) % KZ);
    __syncthreads();  // input planes q .. q + KZ - 1 staged
    if (q + 1 < nz) {
      zn = zin(q + KZ);
      fetch(zn);
    }
    float acc[DV_R];
#pragma unroll
    for (int o = 0; o < DV_R; ++o) acc[o] = 0.f;
#pragma unroll
    for (int w = 0; w < KZ; ++w) {
      const float *base = smem + ((q + w) % KZ) * PL + ty * DV_R * IN_WP + tx;
#pragma unroll
      for (int u = 0; u < KY + DV_R - 1; ++u) {
        float v[KX];
#pragma unroll
        for (int t = 0; t < KX; ++t) v[t] = base[u * IN_WP + t];
#pragma unroll
        for (int o = 0; o < DV_R; ++o) {
          const int j = u - o;
          if (j >= 0 && j < KY) {
#pragma unroll
            for (int t = 0; t < KX; ++t) acc[o] = fmaf(kt.k[t + j * KX + w * KX * KY], v[t], acc[o]);
          }
        }
      }
    }
    if (ox < a.o0) {
      const int64_t obase = ((int64_t)p3 * a.o2 + oz0 + q) * a.out_plane + ox;
#pragma unroll
      for (int o = 0; o < DV_R; ++o) {
        const int oy = by * DV_TY + ty * DV_R + o;
        if (oy < a.o1) {
          if (a.dt_out == IFM_F32) ((float *)a.out)[obase + (int64_t)oy * a.o0] = acc[o];
          else store_f(a.out, obase + (int64_t)oy * a.o0, a.dt_out, acc[o]);
        }
      }
    }
    __syncthreads();  // every wave is done with input plane q's slot
  }
}

template <int KX, int KY, int KZ>
static hipError_t launch_small3(const Dense3dArgs &a, const float *taps, dim3 grid, hipStream_t st) {
  SmallTaps3<KX, KY, KZ> kt;
  for (int i = 0; i < KX * KY * KZ; ++i) kt.k[i] = taps[i];
  hipLaunchKernelGGL((k_dense3d_small<KX, KY, KZ>), grid, dim3(256), 0, st, a, kt);
  return hipGetLastError();
}

template <int KX, int KY>
static hipError_t launch_small3_z(int kz, const Dense3dArgs &a, const float *taps, dim3 grid, hipStream_t st) {
  return kz == 3 ? launch_small3<KX, KY, 3>(a, taps, grid, st) : launch_small3<KX, KY, 5>(a, taps, grid, st);
}

static bool small3_size(int n) { return n == 3 || n == 5; }

static hipError_t launch_dense3d(const Dense4 &in, Dense4 &out, const int64_t *out_first, const DenseKernel &k,
                                 int style, float fill, hipStream_t st) {
  Dense3dArgs a;
  a.in = in.data;
  a.out = out.data;
  a.n0 = (int)in.n[0];
  a.n1 = (int)in.n[1];
  a.n2 = (int)in.n[2];
  a.n3 = (int)in.n[3];
  a.o0 = (int)out.n[0];
  a.o1 = (int)out.n[1];
  a.o2 = (int)out.n[2];
  a.g0 = (int)out_first[0] + k.first[0];
  a.g1 = (int)out_first[1] + k.first[1];
  a.g2 = (int)out_first[2] + k.first[2];
  a.g3 = (int)out_first[3] + (k.ndim > 3 ? k.first[3] : 0);
  a.in_plane = in.n[0] * in.n[1];
  a.out_plane = out.n[0] * out.n[1];
  a.nzc = (a.o2 + D3_NZ - 1) / D3_NZ;
  a.style = style;
  a.dt_in = in.dtype;
  a.dt_out = out.dtype;
  a.fill = fill;
  const dim3 grid((a.o0 + DV_TX - 1) / DV_TX, (a.o1 + DV_TY - 1) / DV_TY, (unsigned)(a.nzc * out.n[3]));
  const float *taps = k.taps.data();
  if (k.n[0] == 3) {
    if (k.n[1] == 3) return launch_small3_z<3, 3>(k.n[2], a, taps, grid, st);
    return launch_small3_z<3, 5>(k.n[2], a, taps, grid, st);
  }
  if (k.n[1] == 3) return launch_small3_z<5, 3>(k.n[2], a, taps, grid, st);
  return launch_small3_z<5, 5>(k.n[2], a, taps, grid, st);
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
    if (small_size(a.k0) && small_size(a.k1)) {
      // register-blocked form; taps from the host copy of the kernel
      const float *taps = k.taps.data();
      switch (a.k0) {
      case 3: return launch_small_ky<3>(a.k1, a, taps, grid, st);
      case 5: return launch_small_ky<5>(a.k1, a, taps, grid, st);
      case 7: return launch_small_ky<7>(a.k1, a, taps, grid, st);
      case 9: return launch_small_ky<9>(a.k1, a, taps, grid, st);
      case 11: return launch_small_ky<11>(a.k1, a, taps, grid, st);
      case 13: return launch_small_ky<13>(a.k1, a, taps, grid, st);
      default: return launch_small_ky<15>(a.k1, a, taps, grid, st);
      }
    }
    hipLaunchKernelGGL(k_dense2d_valu, grid, dim3(256), lds, st, a);
    return hipGetLastError();
  }
  const bool threed = k.ndim >= 3 && (k.ndim < 4 || k.n[3] == 1) && small3_size(k.n[0]) && small3_size(k.n[1]) &&
                      small3_size(k.n[2]);
  if (threed && (out.n[2] + D3_NZ - 1) / D3_NZ * out.n[3] <= 65535)
    return launch_dense3d(in, out, out_first, k, style, fill, st);
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

}  // namespace ifm
