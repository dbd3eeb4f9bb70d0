// k_fir.hip — padarray and separable FIR kernels (gfx950).
//
// padarray materialises BorderArray + copy! (src/border.jl:324-347).  The FIR
// kernels apply the border index map on load, so the reference's padded copy
// (src/imfilter.jl:336) is never built: a factor cascade reads the image once
// per fused group of axes.
#include "k_tiles.h"

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

// Rows of the padded output (axis 0) are the unit of work: a workgroup maps
// its row's axes 1..3 through the border once (wave-uniform), and each thread
// then maps and copies PAD_EPT elements of the row, 256 apart, so reads and
// writes are coalesced and no per-element division is needed.
constexpr int PAD_EPT = 4;

template <int ES>  // raw copy of ES-byte elements (same eltype in and out); 0: converting copy
__global__ void __launch_bounds__(256) k_pad(PadArgs a) {
  using U = typename std::conditional<ES == 1, uint8_t,
            typename std::conditional<ES == 2, uint16_t,
            typename std::conditional<ES == 4, uint32_t, uint64_t>::type>::type>::type;
  const int rows = a.p[1] * a.p[2] * a.p[3];
  for (int row = blockIdx.y; row < rows; row += gridDim.y) {
    const int i1 = row % a.p[1], t = row / a.p[1];
    const int i2 = t % a.p[2], i3 = t / a.p[2];
    const int s1 = bmap(a.style, i1 - a.lo[1], a.n[1]);
    const int s2 = bmap(a.style, i2 - a.lo[2], a.n[2]);
    const int s3 = bmap(a.style, i3 - a.lo[3], a.n[3]);
    const bool rowfill = (s1 | s2 | s3) < 0;
    const int64_t srow = rowfill ? 0 : (s1 + (int64_t)a.n[1] * (s2 + (int64_t)a.n[2] * s3)) * a.n[0];
    const int64_t orow = (int64_t)row * a.p[0];
    const int x0 = blockIdx.x * (256 * PAD_EPT) + threadIdx.x;
#pragma unroll
    for (int k = 0; k < PAD_EPT; ++k) {
      const int x = x0 + 256 * k;
      if (x < a.p[0]) {
        const int s0 = bmap(a.style, x - a.lo[0], a.n[0]);
        const bool isfill = rowfill || s0 < 0;
        if constexpr (ES > 0) {
          ((U *)a.out)[orow + x] = isfill ? (U)a.fill_bits : ((const U *)a.in)[srow + s0];
        } else {
          const double v = isfill ? a.fill : load_d(a.in, srow + s0, a.dt_in);
          store_d(a.out, orow + x, a.dt_out, v);
        }
      }
    }
  }
}

static uint64_t fill_bits_for(int dt, double v) {
  uint64_t b = 0;
  switch (dt) {
  case IFM_U8: b = (uint64_t)(uint8_t)(int)v; break;
  case IFM_N0F8: b = (uint64_t)(uint8_t)(int)llrint(v * 255.0); break;  // the value code/255
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
  const int64_t rows = out.n[1] * out.n[2] * out.n[3];
  if (rows > INT32_MAX || out.n[0] > INT32_MAX) return hipErrorInvalidValue;
  dim3 grid((unsigned)((out.n[0] + 256 * PAD_EPT - 1) / (256 * PAD_EPT)), (unsigned)std::min<int64_t>(rows, 65535));
  if (in.dtype != out.dtype) hipLaunchKernelGGL(k_pad<0>, grid, dim3(256), 0, st, a);
  else switch (a.esize) {
  case 1: hipLaunchKernelGGL(k_pad<1>, grid, dim3(256), 0, st, a); break;
  case 2: hipLaunchKernelGGL(k_pad<2>, grid, dim3(256), 0, st, a); break;
  case 4: hipLaunchKernelGGL(k_pad<4>, grid, dim3(256), 0, st, a); break;
  default: hipLaunchKernelGGL(k_pad<8>, grid, dim3(256), 0, st, a); break;
  }
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

// Rows of the output along axis 0 are the unit of work (as in k_pad): the
// row's coordinates on axes 1..3 are formed once per workgroup, and a thread
// computes FIR_EPT outputs of the row, 256 apart.  Along axis 0 the taps read
// consecutive elements; along another axis they read one element per row of
// that axis, and the wave's lanes read consecutive addresses.
constexpr int FIR_EPT = 4;

template <int DT>  // input eltype; -1: read at run time
__device__ __forceinline__ float fir_ld(const Fir1dArgs &a, int64_t i) {
  if constexpr (DT == IFM_F32) return ((const float *)a.in)[i];
  else if constexpr (DT == IFM_F64) return (float)((const double *)a.in)[i];
  else return load_f(a.in, i, a.dt_in);
}

template <int DT>
__global__ void __launch_bounds__(256) k_fir1d(Fir1dArgs a) {
  const int rows = a.out_n[1] * a.out_n[2] * a.out_n[3];
  const int64_t st1 = a.in_n[0], st2 = st1 * a.in_n[1], st3 = st2 * a.in_n[2];
  const int64_t stride[MAXD] = {1, st1, st2, st3};
  const int n = a.in_n[a.axis];
  const int64_t astride = stride[a.axis];
  for (int row = blockIdx.y; row < rows; row += gridDim.y) {
    int o[MAXD];
    o[1] = row % a.out_n[1];
    const int t = row / a.out_n[1];
    o[2] = t % a.out_n[2];
    o[3] = t / a.out_n[2];
    int64_t base_u = 0;  // input offset of the row on the non-filtered axes 1..3
#pragma unroll
    for (int d = 1; d < MAXD; ++d)
      if (d != a.axis) base_u += (int64_t)(o[d] + a.shift[d]) * stride[d];
    const int xu = a.axis == 0 ? 0 : o[a.axis] + a.first;
    const int64_t orow = (int64_t)row * a.out_n[0];
    const int x0 = blockIdx.x * (256 * FIR_EPT) + threadIdx.x;
#pragma unroll
    for (int k = 0; k < FIR_EPT; ++k) {
      const int o0 = x0 + 256 * k;
      if (o0 >= a.out_n[0]) break;
      int64_t base;
      int x;
      if (a.axis == 0) {
        base = base_u;
        x = o0 + a.first;
      } else {
        base = base_u + o0 + a.shift[0];
        x = xu;
      }
      float acc = 0.f;
      if (x >= 0 && x + a.len <= n) {
        const int64_t b0 = base + (int64_t)x * astride;
        for (int tt = 0; tt < a.len; ++tt) acc = fmaf(a.taps[tt], fir_ld<DT>(a, b0 + tt * astride), acc);
      } else {
        for (int tt = 0; tt < a.len; ++tt) {
          const int s = bmap(a.style, x + tt, n);
          const float v = s < 0 ? a.fill : fir_ld<DT>(a, base + (int64_t)s * astride);
          acc = fmaf(a.taps[tt], v, acc);
        }
      }
      store_f(a.out, orow + o0, a.dt_out, acc);
    }
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
  const int64_t rows = out.n[1] * out.n[2] * out.n[3];
  if (rows > INT32_MAX) return hipErrorInvalidValue;
  dim3 grid((unsigned)((out.n[0] + 256 * FIR_EPT - 1) / (256 * FIR_EPT)), (unsigned)std::min<int64_t>(rows, 65535));
  if (in.dtype == IFM_F32) hipLaunchKernelGGL(k_fir1d<IFM_F32>, grid, dim3(256), 0, st, a);
  else if (in.dtype == IFM_F64) hipLaunchKernelGGL(k_fir1d<IFM_F64>, grid, dim3(256), 0, st, a);
  else hipLaunchKernelGGL(k_fir1d<-1>, grid, dim3(256), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// sep2d: fused pair of factors on canonical axes 0 and 1, batched over axes
// 2 and 3.  An output tile of 64 x 32 or 64 x 64 pixels loads its halo'd input tile into
// LDS once (border map on load), runs the axis-0 factor into a second LDS
// tile and the axis-1 factor from there, so HBM sees one read and one write.
// ---------------------------------------------------------------------------
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
  int ty;              // tile height (S2_TY or S2_TY2)
  float taps0[S2_MAXK];
  float taps1[S2_MAXK];
};

__global__ void __launch_bounds__(256) k_sep2d(Sep2dArgs a) {
  extern __shared__ float smem[];
  Sep2dTile t;
  t.in = a.in;
  t.out = a.out;
  t.n0 = a.n0;
  t.n1 = a.n1;
  t.o0 = a.o0;
  t.o1 = a.o1;
  t.k0 = a.k0;
  t.k1 = a.k1;
  t.first0 = a.first0;
  t.first1 = a.first1;
  t.style = a.style;
  t.dt_in = a.dt_in;
  t.dt_out = a.dt_out;
  t.fill = a.fill;
  t.ty = a.ty;
  // tiles XCD by XCD: row neighbours share their halo columns' cache lines
  const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  t.ox0 = (tile % gridDim.x) * S2_TX;
  t.oy0 = (tile / gridDim.x) * a.ty;
  t.x0 = a.f0 + t.ox0;
  t.y0 = a.f1 + t.oy0;
  for (int plane = blockIdx.z; plane < a.nplanes; plane += gridDim.z) {
    const int p2 = plane % a.o2, p3 = plane / a.o2;
    t.ibase = ((int64_t)(p2 + a.f2) + (int64_t)a.n2 * (p3 + a.f3)) * a.in_plane;
    t.obase = (int64_t)plane * a.out_plane;
    sep2d_tile(t, a.taps0, a.taps1, smem);
    __syncthreads();  // the next plane reuses the LDS tiles
  }
}

// ---------------------------------------------------------------------------
// sep3d: the three-factor cascade on axes 0, 1 and 2 in one pass.  A
// workgroup owns a 64 x 32 column of outputs in (x, y) and up to S3_NZ output
// planes; it marches over the S3_NZ + K2 - 1 input planes they need, filters
// each along axes 0 and 1 (sep2d_filter) and adds it into the K2 pending
// output planes kept in registers, writing a plane once its last input plane
// has arrived.  HBM sees the input once (plus halos) and the output once;
// there is no temporary.  Accumulation along axis 2 runs in tap order, like
// the cascade's third pass.
// ---------------------------------------------------------------------------
constexpr int S3_NZ = 64;  // 512 x 512 x 256: 512 workgroups, 1.12 x z halo (32: 1,024 and 8 % slower)
constexpr int S3_MAXK2 = 9;  // the plane loop is unrolled K2 times
constexpr int S3_NE = 16;  // staged elements per thread and plane

struct Sep3dArgs {
  const void *in;
  void *out;
  int n0, n1, n2;       // input extents (axes 0..2)
  int o0, o1, o2;       // output extents
  int f0, f1, f2, f3;   // out_first
  int64_t in_plane, out_plane;  // elements per (x, y) plane
  int n3;               // planes along axis 3 (output)
  int nzc;              // z chunks
  int k0, k1, first0, first1, first2;
  int style, dt_in, dt_out;
  float fill;           // the Fill value
  float fill01;         // the Fill value filtered by the axis-0 and axis-1 factors
  float taps0[S2_MAXK];
  float taps1[S2_MAXK];
  float taps2[S3_MAXK2];
};

// LDS words: the sep2d tile layout plus one word that absorbs the writes of
// staging slots with nothing to store.
inline size_t sep3d_lds_floats(int k0, int k1) { return sep2d_lds_floats(k0, k1) + 1; }

template <int K2, int DT>
__global__ void __launch_bounds__(256) k_sep3d(Sep3dArgs a) {
  extern __shared__ float smem[];
  // tiles XCD by XCD (x fastest, then y, then z chunk): neighbouring tiles
  // share halo cache lines, which one L2 then serves
  const int nxy = gridDim.x * gridDim.y;
  const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y + nxy * blockIdx.z, nxy * gridDim.z);
  const int bz = tile / nxy, bx = tile % gridDim.x, by = (tile % nxy) / gridDim.x;
  const int zc = bz % a.nzc, p3 = bz / a.nzc;
  const int oz0 = zc * S3_NZ, nz = min(S3_NZ, a.o2 - oz0);
  Sep2dTile t;
  t.in = a.in;
  t.out = a.out;
  t.n0 = a.n0;
  t.n1 = a.n1;
  t.o0 = a.o0;
  t.o1 = a.o1;
  t.k0 = a.k0;
  t.k1 = a.k1;
  t.first0 = a.first0;
  t.first1 = a.first1;
  t.style = a.style;
  t.dt_in = a.dt_in;
  t.dt_out = a.dt_out;
  t.fill = a.fill;
  t.ty = S2_TY;
  t.ox0 = bx * S2_TX;
  t.oy0 = by * S2_TY;
  t.x0 = a.f0 + t.ox0;
  t.y0 = a.f1 + t.oy0;
  const int in_w = S2_TX + a.k0 - 1, in_h = S2_TY + a.k1 - 1, in_wp = in_w | 1;
  int *s_rows = (int *)(smem + sep2d_rows_offset(a.k0, a.k1));
  const int dummy = sep2d_rows_offset(a.k0, a.k1) + in_h;
  sep2d_rows(t, s_rows);
  __syncthreads();
  // Staging map, the same for every plane: slot i of the thread is element
  // e = tid + 256 i of the in_h x in_w input tile, with its byte offset in the
  // plane (a plane is < 4 GiB) and its LDS word.  Fill elements get the Fill
  // value written once here and their slot writes go to the dummy word, as do
  // slots past the tile, so a plane is staged by plain loads and LDS writes.
  constexpr uint32_t ES = DT == IFM_F64 ? 8 : 4;
  uint32_t off[S3_NE];
  int lidx[S3_NE];
  {
    const int total = in_h * in_w;
#pragma unroll
    for (int i = 0; i < S3_NE; ++i) {
      const int e = (int)threadIdx.x + 256 * i;
      off[i] = 0;
      lidx[i] = dummy;
      if (e < total) {
        const int r = e / in_w, c = e - r * in_w;
        const int sy = s_rows[r];
        const int sx = bmap(a.style, t.x0 + a.first0 + c, a.n0);
        if (sy < 0 || sx < 0) {
          smem[r * in_wp + c] = a.fill;
        } else {
          off[i] = ((uint32_t)sy * (uint32_t)a.n0 + (uint32_t)sx) * ES;
          lidx[i] = r * in_wp + c;
        }
      }
    }
  }
  using Raw = typename std::conditional<DT == IFM_F64, double, float>::type;
  Raw v[S3_NE];
  auto issue = [&](int zp) {
    const char *pb = (const char *)a.in + ((int64_t)(p3 + a.f3) * a.n2 + zp) * a.in_plane * ES;
#pragma unroll
    for (int i = 0; i < S3_NE; ++i) v[i] = *(const Raw *)(pb + off[i]);
  };
  auto commit = [&]() {
#pragma unroll
    for (int i = 0; i < S3_NE; ++i) smem[lidx[i]] = (float)v[i];
  };
  auto store = [&](int od, const float (&acc)[8]) {
    t.obase = ((int64_t)p3 * a.o2 + oz0 + od) * a.out_plane;
    sep2d_store(t, acc);
  };
  // Output plane o accumulates input planes o .. o + K2 - 1 (taps 0 .. K2-1)
  // in slot o mod K2.  The plane loop is unrolled K2 times so that slots are
  // static registers: at input plane q (u = q mod K2) tap j goes to slot
  // (u - j) mod K2, tap 0 starting the sum; output q - K2 + 1 is then
  // complete in slot (u + 1) mod K2 and is written at the start of the next
  // step (after that step's loads have been waited for), the last one at once.
  float acc[K2][8];
  const int nq = nz + K2 - 1;
  auto plane_of = [&](int q) { return bmap(a.style, a.f2 + oz0 + a.first2 + q, a.n2); };
  int sz = plane_of(0);
  if (sz >= 0) issue(sz);
  for (int q0 = 0; q0 < nq; q0 += K2) {
#pragma unroll
    for (int u = 0; u < K2; ++u) {
      const int q = q0 + u;
      if (q < nq) {
        if (sz >= 0) commit();
        __syncthreads();  // plane q staged
        if (q >= K2) store(q - K2, acc[u]);
        const int szn = q + 1 < nq ? plane_of(q + 1) : -1;
        if (szn >= 0) issue(szn);
        float sl[8];
        if (sz < 0) {
#pragma unroll
          for (int r = 0; r < 8; ++r) sl[r] = a.fill01;
        } else {
          sep2d_filter(t, a.taps0, a.taps1, smem, sl);
        }
        __syncthreads();  // every wave is done with plane q's LDS tiles
        sz = szn;
#pragma unroll
        for (int r = 0; r < 8; ++r) acc[u][r] = a.taps2[0] * sl[r];
#pragma unroll
        for (int j = 1; j < K2; ++j)
#pragma unroll
          for (int r = 0; r < 8; ++r) acc[(u - j + K2) % K2][r] = fmaf(a.taps2[j], sl[r], acc[(u - j + K2) % K2][r]);
        if (q == nq - 1) store(q - K2 + 1, acc[(u + 1) % K2]);
      }
    }
  }
}

bool sep3d_supported(int k0, int k1, int k2, int dt_in, int64_t plane_elems) {
  return k0 <= S2_MAXK && k1 <= S2_MAXK && (S2_TY + k1 - 1) * (S2_TX + k0 - 1) <= 256 * S3_NE &&
         k2 <= S3_MAXK2 && (k2 & 1) && (dt_in == IFM_F32 || dt_in == IFM_F64) &&
         plane_elems * (dt_in == IFM_F64 ? 8 : 4) < (int64_t(1) << 32);
}

template <int K2>
static void sep3d_go(const Sep3dArgs &a, bool f64, dim3 grid, size_t lds, hipStream_t st) {
  if (f64)
    hipLaunchKernelGGL((k_sep3d<K2, IFM_F64>), grid, dim3(256), lds, st, a);
  else
    hipLaunchKernelGGL((k_sep3d<K2, IFM_F32>), grid, dim3(256), lds, st, a);
}

hipError_t launch_sep3d(const Dense4 &in, Dense4 &out, const int64_t *out_first, const FirPass &p0, const FirPass &p1,
                        const FirPass &p2, int style, float fill, hipStream_t st) {
  if (out.count() == 0) return hipSuccess;
  Sep3dArgs a;
  a.in = in.data;
  a.out = out.data;
  a.n0 = (int)in.n[0];
  a.n1 = (int)in.n[1];
  a.n2 = (int)in.n[2];
  a.o0 = (int)out.n[0];
  a.o1 = (int)out.n[1];
  a.o2 = (int)out.n[2];
  a.f0 = (int)out_first[0];
  a.f1 = (int)out_first[1];
  a.f2 = (int)out_first[2];
  a.f3 = (int)out_first[3];
  a.in_plane = in.n[0] * in.n[1];
  a.out_plane = out.n[0] * out.n[1];
  a.n3 = (int)out.n[3];
  a.nzc = (a.o2 + S3_NZ - 1) / S3_NZ;
  a.k0 = p0.len;
  a.k1 = p1.len;
  a.first0 = p0.first;
  a.first1 = p1.first;
  a.first2 = p2.first;
  a.style = style;
  a.dt_in = in.dtype;
  a.dt_out = out.dtype;
  a.fill = fill;
  float s0 = 0.f, s1 = 0.f;
  for (int t = 0; t < p0.len; ++t) s0 += p0.taps[t];
  for (int t = 0; t < p1.len; ++t) s1 += p1.taps[t];
  a.fill01 = fill * s0 * s1;
  for (int t = 0; t < p0.len; ++t) a.taps0[t] = p0.taps[t];
  for (int t = 0; t < p1.len; ++t) a.taps1[t] = p1.taps[t];
  for (int t = 0; t < p2.len; ++t) a.taps2[t] = p2.taps[t];
  const size_t lds = sizeof(float) * sep3d_lds_floats(a.k0, a.k1);
  const int64_t gz = (int64_t)a.nzc * a.n3;
  if (gz > 65535) return hipErrorInvalidValue;
  dim3 grid((a.o0 + S2_TX - 1) / S2_TX, (a.o1 + S2_TY - 1) / S2_TY, (unsigned)gz);
  const bool f64 = in.dtype == IFM_F64;
  switch (p2.len) {
  case 1: sep3d_go<1>(a, f64, grid, lds, st); break;
  case 3: sep3d_go<3>(a, f64, grid, lds, st); break;
  case 5: sep3d_go<5>(a, f64, grid, lds, st); break;
  case 7: sep3d_go<7>(a, f64, grid, lds, st); break;
  case 9: sep3d_go<9>(a, f64, grid, lds, st); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Batched serving: tile g of the launch is tile g - tile_first[r] of request
// r = tile_req[g] (row-major over its 64 x ty tiles).
__global__ void __launch_bounds__(256) k_sep_batch(const SepReq *__restrict__ reqs, const int *__restrict__ tile_req,
                                                   const int *__restrict__ tile_first) {
  extern __shared__ float smem[];
  const int g = blockIdx.x;
  const int r = tile_req[g];
  const SepReq &q = reqs[r];
  const int local = g - tile_first[r];
  const int ntx = (q.n0 + S2_TX - 1) / S2_TX;
  Sep2dTile t;
  t.in = q.in;
  t.out = q.out;
  t.ibase = 0;
  t.obase = 0;
  t.n0 = q.n0;
  t.n1 = q.n1;
  t.o0 = q.n0;
  t.o1 = q.n1;
  t.ty = q.ty;
  t.ox0 = (local % ntx) * S2_TX;
  t.oy0 = (local / ntx) * q.ty;
  t.x0 = t.ox0;
  t.y0 = t.oy0;
  t.k0 = q.k0;
  t.k1 = q.k1;
  t.first0 = q.first0;
  t.first1 = q.first1;
  t.style = q.style;
  t.dt_in = q.dt_in;
  t.dt_out = q.dt_out;
  t.fill = q.fill;
  sep2d_tile(t, q.taps0, q.taps1, smem);
}

hipError_t launch_sep_batch(const SepReq *reqs, const int *tile_req, const int *tile_first, int ntiles,
                            size_t lds_bytes, hipStream_t st) {
  if (ntiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sep_batch, dim3(ntiles), dim3(256), lds_bytes, st, reqs, tile_req, tile_first);
  return hipGetLastError();
}

hipError_t launch_sep2d(const Dense4 &in, Dense4 &out, const int64_t *out_first, const FirPass &p0,
                        const FirPass &p1, int style, float fill, int ty, hipStream_t st) {
  if (ty != S2_TY && ty != S2_TY2) return hipErrorInvalidValue;
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
  a.ty = ty;
  const size_t lds = sizeof(float) * sep2d_lds_floats(a.k0, a.k1, ty);
  dim3 grid((a.o0 + S2_TX - 1) / S2_TX, (a.o1 + ty - 1) / ty, a.nplanes < 65535 ? a.nplanes : 65535);
  hipLaunchKernelGGL(k_sep2d, grid, dim3(256), lds, st, a);
  return hipGetLastError();
}

}  // namespace ifm
