// k_tiles.h — tile bodies shared by the single-image kernels and the batched
// serving kernels (gfx950).
#pragma once
#include "ifm_internal.h"

namespace ifm {

// ---------------------------------------------------------------------------
// Fused pair of 1-D factors on axes 0 and 1 (the two-factor cascade of
// src/imfilter.jl:372-457 with the factor loop of :671-739).  One workgroup of
// 256 threads computes a 64 x TY output tile, TY = 32 or 64 (the tile shape is
// a launch choice: a taller tile re-reads fewer halo rows per output, a
// shorter one makes more workgroups for small images):
//   1. the (TY + k1 - 1) x (64 + k0 - 1) input tile is loaded once into LDS,
//      border map on load; each thread owns one column and every other row.
//      Loads keep the raw element bits and are issued in groups of 8 before
//      any is used (issue / commit), so several are in flight per thread;
//   2. the axis-0 factor runs from registers: a thread slides over k0 + 3
//      LDS values to produce 4 consecutive outputs of one row (mid tile);
//   3. the axis-1 factor likewise produces 8 consecutive rows of one column,
//      once per 32-row band of the tile.
// Accumulation is FP32 in tap order, like the reference's loops.
// ---------------------------------------------------------------------------
constexpr int S2_TX = 64;
constexpr int S2_TY = 32;   // the default tile height; S2_TY2 is the tall one
constexpr int S2_TY2 = 64;
constexpr int S2_MAXK = 64;

struct Sep2dTile {
  const void *in;
  void *out;
  int64_t ibase, obase;  // plane offsets (elements)
  int n0, n1;            // input extents
  int o0, o1;            // output extents
  int x0, y0;            // input coordinate of the tile's first output
  int ox0, oy0;          // output index of the tile's first output
  int k0, k1, first0, first1;
  int style, dt_in, dt_out;
  float fill;
  int ty;                // output rows of the tile: S2_TY or S2_TY2
};

// LDS: input tile (odd pitch), the mid tile at a 16-byte boundary, then the
// row table (border-mapped input row of each tile row, -1 for Fill).
__host__ __device__ inline int sep2d_mid_offset(int k0, int k1, int ty = S2_TY) {
  const int in_h = ty + k1 - 1, in_w = S2_TX + k0 - 1;
  return (in_h * (in_w | 1) + 3) & ~3;
}
__host__ __device__ inline int sep2d_rows_offset(int k0, int k1, int ty = S2_TY) {
  return sep2d_mid_offset(k0, k1, ty) + (ty + k1 - 1) * S2_TX;
}
__host__ __device__ inline size_t sep2d_lds_floats(int k0, int k1, int ty = S2_TY) {
  return (size_t)sep2d_rows_offset(k0, k1, ty) + (size_t)(ty + k1 - 1);
}

// Raw element bits (F64 is narrowed at once) and their value: a load is not
// waited for until the value is formed.
template <int DT>
__device__ __forceinline__ uint32_t raw_ld(const void *p, int64_t i) {
  if constexpr (DT == IFM_F32) return ((const uint32_t *)p)[i];
  else if constexpr (DT == IFM_BF16 || DT == IFM_F16) return ((const uint16_t *)p)[i];
  else if constexpr (DT == IFM_U8 || DT == IFM_N0F8) return ((const uint8_t *)p)[i];
  else return __float_as_uint((float)((const double *)p)[i]);
}
template <int DT>
__device__ __forceinline__ float raw_val(uint32_t r) {
  if constexpr (DT == IFM_F32 || DT == IFM_F64) return __uint_as_float(r);
  else if constexpr (DT == IFM_BF16) return __uint_as_float(r << 16);
  else if constexpr (DT == IFM_F16) return (float)__builtin_bit_cast(_Float16, (uint16_t)r);
  else if constexpr (DT == IFM_U8) return (float)r;
  else return n0f8_to_f32(r);
}

// Thread's staging coordinates: column c (of in_w), rows rh, rh + 2, ...
struct StageLane {
  int c, rh, nrow, sx;
  bool colok;
};
__device__ __forceinline__ StageLane stage_lane(const Sep2dTile &a, int in_w, int in_h) {
  StageLane s;
  s.c = threadIdx.x & 127;
  s.rh = __builtin_amdgcn_readfirstlane(threadIdx.x >> 7);
  s.colok = s.c < in_w;
  s.sx = bmap(a.style, a.x0 + a.first0 + (s.colok ? s.c : in_w - 1), a.n0);
  s.nrow = (in_h - s.rh + 1) >> 1;
  return s;
}

// Fills the row table of the tile (the caller owns the barrier after it).
__device__ __forceinline__ void sep2d_rows(const Sep2dTile &a, int *s_rows) {
  const int in_h = a.ty + a.k1 - 1;
  for (int i = threadIdx.x; i < in_h; i += blockDim.x) s_rows[i] = bmap(a.style, a.y0 + a.first1 + i, a.n1);
}

// Loads of rows j0 .. j0+NV-1 of the thread.  The row table is read for all
// of them first (one LDS round trip), then every load is issued; Fill rows
// and columns load an element of row 0 / column 0 and set their bit in
// `fillrows` (wave-uniform), which the commit replaces by the Fill value.
template <int DT, int NV>
__device__ __forceinline__ void sep2d_issue(const Sep2dTile &a, const StageLane &s, const int *s_rows, int j0,
                                            uint32_t (&v)[NV], uint32_t &fillrows) {
  const int sxc = s.sx < 0 ? 0 : s.sx;
  const int last = 2 * s.nrow - 2 + s.rh;  // the thread's last row
  int sy[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) sy[j] = s_rows[min(s.rh + 2 * (j0 + j), last)];
  fillrows = 0;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const int y = __builtin_amdgcn_readfirstlane(sy[j]);
    fillrows |= (y < 0 ? 1u : 0u) << j;
    if (j0 + j < s.nrow)  // wave-uniform
      v[j] = raw_ld<DT>(a.in, a.ibase + (int64_t)(y < 0 ? 0 : y) * a.n0 + sxc);
  }
}

template <int DT, int NV>
__device__ __forceinline__ void sep2d_commit(const Sep2dTile &a, const StageLane &s, float *s_in, int in_wp, int j0,
                                             const uint32_t (&v)[NV], uint32_t fillrows) {
  if (!s.colok) return;
  const bool fillcol = s.sx < 0;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const int jj = j0 + j;
    if (jj < s.nrow) s_in[(s.rh + 2 * jj) * in_wp + s.c] = (((fillrows >> j) & 1u) || fillcol) ? a.fill : raw_val<DT>(v[j]);
  }
}

// Stages the tile (row table already filled and visible).
template <int DT>
__device__ __forceinline__ void sep2d_stage(const Sep2dTile &a, const int *s_rows, float *s_in, int in_w, int in_h,
                                            int in_wp) {
  const StageLane s = stage_lane(a, in_w, in_h);
  constexpr int NV = 8;  // loads in flight per thread
  for (int j0 = 0; j0 < s.nrow; j0 += NV) {
    uint32_t v[NV], fillrows;
    sep2d_issue<DT, NV>(a, s, s_rows, j0, v, fillrows);
    sep2d_commit<DT, NV>(a, s, s_in, in_wp, j0, v, fillrows);
  }
}

// Stores the thread's 8 outputs of the 32-row band starting at tile row h.
__device__ __forceinline__ void sep2d_store(const Sep2dTile &a, const float (&acc)[8], int h = 0) {
  const int tid = threadIdx.x;
  const int col = tid & 63, r0 = h + (tid >> 6) * 8;
  const int ox = a.ox0 + col;
  if (ox < a.o0) {
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int oy = a.oy0 + r0 + o;
      if (oy < a.o1) store_f(a.out, a.obase + (int64_t)oy * a.o0 + ox, a.dt_out, acc[o]);
    }
  }
}

// Axis-0 factor over every staged row of a staged tile (after the staging
// barrier) into the mid tile; ends with the barrier that publishes it.
__device__ __forceinline__ void sep2d_rowpass(const Sep2dTile &a, const float *taps0, float *smem) {
  const int in_w = S2_TX + a.k0 - 1;
  const int in_h = a.ty + a.k1 - 1;
  const int in_wp = in_w | 1;
  const float *s_in = smem;
  float *s_mid = smem + sep2d_mid_offset(a.k0, a.k1, a.ty);  // in_h x 64
  const int tid = threadIdx.x;
  // Axis-0 factor: item = (row, quad of 4 columns).
  for (int item = tid; item < in_h * 16; item += 256) {
    const int row = item >> 4, cq = (item & 15) * 4;
    const float *src = s_in + row * in_wp + cq;
    float q[4] = {0.f, 0.f, 0.f, 0.f};
    float w[4] = {src[0], src[1], src[2], 0.f};
    // unrolled so that several LDS reads are in flight
#pragma unroll 8
    for (int t = 0; t < a.k0; ++t) {
      w[3] = src[t + 3];
      const float k = taps0[t];
#pragma unroll
      for (int o = 0; o < 4; ++o) q[o] = fmaf(k, w[o], q[o]);
      w[0] = w[1];
      w[1] = w[2];
      w[2] = w[3];
    }
    *(float4 *)(s_mid + row * S2_TX + cq) = make_float4(q[0], q[1], q[2], q[3]);
  }
  __syncthreads();
}

// Axis-1 factor over the mid tile for the 32-row band starting at tile row h:
// each thread owns 8 consecutive rows (h + 8 (t / 64) ..) of column t % 64
// and slides an 8-value window down the column, one LDS read per tap.
__device__ __forceinline__ void sep2d_colpass(const Sep2dTile &a, const float *taps1, const float *smem, int h,
                                              float (&acc)[8]) {
  const float *s_mid = smem + sep2d_mid_offset(a.k0, a.k1, a.ty);
  const int tid = threadIdx.x;
  const int col = tid & 63, r0 = h + (tid >> 6) * 8;
#pragma unroll
  for (int o = 0; o < 8; ++o) acc[o] = 0.f;
  const float *src = s_mid + r0 * S2_TX + col;
  float w[8];
#pragma unroll
  for (int o = 0; o < 7; ++o) w[o] = src[o * S2_TX];
#pragma unroll 8
  for (int t = 0; t < a.k1; ++t) {
    w[7] = src[(t + 7) * S2_TX];
    const float k = taps1[t];
#pragma unroll
    for (int o = 0; o < 8; ++o) acc[o] = fmaf(k, w[o], acc[o]);
#pragma unroll
    for (int o = 0; o < 7; ++o) w[o] = w[o + 1];
  }
}

// Both factor passes over a staged tile of height S2_TY; on return thread t
// holds column t % 64, rows 8 (t / 64) .. +7 of the tile.
__device__ __forceinline__ void sep2d_filter(const Sep2dTile &a, const float *taps0, const float *taps1,
                                             float *smem, float (&acc)[8]) {
  sep2d_rowpass(a, taps0, smem);
  sep2d_colpass(a, taps1, smem, 0, acc);
}

// Stages, filters and stores one tile (the caller owns the barrier before
// LDS is reused).
__device__ __forceinline__ void sep2d_tile(const Sep2dTile &a, const float *taps0, const float *taps1,
                                           float *smem) {
  const int in_w = S2_TX + a.k0 - 1;
  const int in_h = a.ty + a.k1 - 1;
  const int in_wp = in_w | 1;
  int *s_rows = (int *)(smem + sep2d_rows_offset(a.k0, a.k1, a.ty));
  sep2d_rows(a, s_rows);
  __syncthreads();
  switch (a.dt_in) {
  case IFM_F32: sep2d_stage<IFM_F32>(a, s_rows, smem, in_w, in_h, in_wp); break;
  case IFM_BF16: sep2d_stage<IFM_BF16>(a, s_rows, smem, in_w, in_h, in_wp); break;
  case IFM_F16: sep2d_stage<IFM_F16>(a, s_rows, smem, in_w, in_h, in_wp); break;
  case IFM_F64: sep2d_stage<IFM_F64>(a, s_rows, smem, in_w, in_h, in_wp); break;
  case IFM_N0F8: sep2d_stage<IFM_N0F8>(a, s_rows, smem, in_w, in_h, in_wp); break;
  default: sep2d_stage<IFM_U8>(a, s_rows, smem, in_w, in_h, in_wp); break;
  }
  __syncthreads();
  sep2d_rowpass(a, taps0, smem);
  for (int h = 0; h < a.ty; h += S2_TY) {
    float acc[8];
    sep2d_colpass(a, taps1, smem, h, acc);
    sep2d_store(a, acc, h);
  }
}

}  // namespace ifm
