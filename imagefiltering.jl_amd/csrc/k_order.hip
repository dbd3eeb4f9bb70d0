// k_order.hip — mapwindow order filters (gfx950): median, minimum, maximum,
// extrema.  Bit-exact: all comparisons run on integer order keys (codes for
// UInt8/N0f8, isless-ordered keys for floats).
//
//  * k_median_net   k x k (k = 3, 5, 7) centred windows: every thread runs a
//                   generated selection network over a tile of outputs; keys
//                   of 16 bits or fewer are packed two to a register so each
//                   v_pk_min_u16 / v_pk_max_u16 serves two tiles.
//  * k_median_any   any window / border (incl. circular, whose edge windows
//                   wrap inside the window's intersection with the image,
//                   src/mapwindow.jl:310-316): counting selection.
//  * k_minmax2d     min / max / extrema in one pass over 2-D tiles; for 8-bit
//                   codes and square windows 3..9, k_minmax2d_u8 (packed keys).
//  * k_minmax_pass  one axis of the separable min/max over truncated windows.
#include "ifm_internal.h"

namespace ifm {

typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t kmin(uint32_t a, uint32_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint32_t kmax(uint32_t a, uint32_t b) { return a > b ? a : b; }
__device__ __forceinline__ u16x2 kmin(u16x2 a, u16x2 b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ u16x2 kmax(u16x2 a, u16x2 b) { return __builtin_elementwise_max(a, b); }

#include "median_nets.h"

template <int K, int TX, int TY>
struct MedianNet;
template <> struct MedianNet<3, 4, 2> {
  template <typename T> static __device__ __forceinline__ void run(const T (&v)[24], T (&o)[8]) { median_net_3_4_2(v, o); }
};
template <> struct MedianNet<5, 2, 2> {
  template <typename T> static __device__ __forceinline__ void run(const T (&v)[36], T (&o)[4]) { median_net_5_2_2(v, o); }
};
template <> struct MedianNet<7, 2, 2> {
  template <typename T> static __device__ __forceinline__ void run(const T (&v)[64], T (&o)[4]) { median_net_7_2_2(v, o); }
};


// Block = TXT x TYT threads (32 x 8, or 16 x 16 where 32 columns of threads
// would stage more than 128 columns); thread tile TX x TY outputs, PACK tiles
// stacked in rows (PACK = 2 packs two 16-bit keys per register).
template <int K, int TX, int TY, int PACK>
struct NetGeom {
  static constexpr int H = K / 2;
  // threads across x: 32 when the staged width still fits two columns per
  // lane (wider rows load and store in longer runs: 5x5 on 8192^2 N0f8 189 ->
  // 172 us), else 16
  static constexpr int TXT = (32 * TX + K - 1 <= 128) ? 32 : 16;
  static constexpr int TYT = 256 / TXT;
  static constexpr int BX = TXT * TX;
  static constexpr int BYH = TYT * TY;  // rows per packed half
  static constexpr int BY = BYH * PACK;
  static constexpr int LW = BX + K - 1;
  static constexpr int LH = BY + K - 1;
  static constexpr int VW = TX + K - 1, VH = TY + K - 1;
};

// Order key of an element of type DT (key32 with the type fixed).
template <int DT>
__device__ __forceinline__ uint32_t key_of(const void *p, int64_t i) {
  if constexpr (DT == IFM_U8) {
    return ((const uint8_t *)p)[i];
  } else if constexpr (DT == IFM_F32) {
    const uint32_t u = __float_as_uint(((const float *)p)[i]);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0xffffffffu;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  } else {
    const uint32_t u = ((const uint16_t *)p)[i];
    constexpr uint32_t expmask = DT == IFM_BF16 ? 0x7f80u : 0x7c00u;
    if ((u & 0x7fffu) > expmask) return 0xffffu;
    return (u & 0x8000u) ? (~u & 0xffffu) : (u | 0x8000u);
  }
}

// Stages the LH x LW key tile whose element (0, 0) is image position
// (gx, gy) before the border map.  Lane c of a wave owns tile columns c and
// c + 64 (their border maps computed once), wave w rows w, w + 4, ...: a row's
// map is wave-uniform, one load instruction reads a contiguous row piece, and
// every load of the thread is issued before the first is used.  BYH > 0
// (16-bit keys, two tiles per register): word (p, c) holds the keys of rows p
// and p + BYH in its low and high halves, the layout the packed network reads.
template <int DT, int LW, int LH, int BYH>
__device__ __forceinline__ void median_stage(const MedianArgs &a, int gx, int gy, int64_t ibase, uint32_t *tile) {
  static_assert(LW <= 128, "two columns per lane");
  constexpr int NR = (LH + 3) / 4;
  const int c = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int sx[2];
  bool cok[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    cok[h] = c + 64 * h < LW;
    sx[h] = cok[h] ? bmap(a.style, gx + c + 64 * h, a.n0) : -1;
  }
  uint32_t k[NR][2];
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const int r = w + 4 * j;
    const int sy = r < LH ? bmap(a.style, gy + r, a.n1) : -1;  // wave-uniform
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      k[j][h] = a.fillkey;
      if (sy >= 0 && sx[h] >= 0) k[j][h] = key_of<DT>(a.in, ibase + (int64_t)sy * a.n0 + sx[h]);
    }
  }
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const int r = w + 4 * j;
    if (r < LH) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (!cok[h]) continue;
        if constexpr (BYH > 0) {
          uint16_t *t16 = (uint16_t *)tile;
          if (r < LH - BYH) t16[2 * (r * LW + c + 64 * h)] = (uint16_t)k[j][h];
          if (r >= BYH) t16[2 * ((r - BYH) * LW + c + 64 * h) + 1] = (uint16_t)k[j][h];
        } else {
          tile[r * LW + c + 64 * h] = k[j][h];
        }
      }
    }
  }
}

// One BX x BY output tile whose first output is (bx, by) in output indices,
// image coordinate (a.f0 + bx, a.f1 + by), of the plane at ibase / obase.
template <int K, int TX, int TY, int PACK>
__device__ __forceinline__ void median_tile(const MedianArgs &a, int bx, int by, int64_t ibase, int64_t obase,
                                            uint32_t *tile) {
  using G = NetGeom<K, TX, TY, PACK>;
  constexpr int LW = G::LW, LH = G::LH, VW = G::VW, VH = G::VH, BYH = G::BYH;
  const int tid = threadIdx.x;
  const int tx = tid % G::TXT, ty = tid / G::TXT;
  const int x0 = a.f0 + bx, y0 = a.f1 + by;
  constexpr int PB = PACK == 2 ? BYH : 0;
  switch (a.dt) {
  case IFM_U8:
  case IFM_N0F8: median_stage<IFM_U8, LW, LH, PB>(a, x0 - G::H, y0 - G::H, ibase, tile); break;
  case IFM_F32: median_stage<IFM_F32, LW, LH, 0>(a, x0 - G::H, y0 - G::H, ibase, tile); break;
  case IFM_BF16: median_stage<IFM_BF16, LW, LH, PB>(a, x0 - G::H, y0 - G::H, ibase, tile); break;
  default: median_stage<IFM_F16, LW, LH, PB>(a, x0 - G::H, y0 - G::H, ibase, tile); break;
  }
  __syncthreads();
  const int cx = tx * TX, cy = ty * TY;
  if constexpr (PACK == 2) {
    u16x2 v[VW * VH];
#pragma unroll
    for (int r = 0; r < VH; ++r)
#pragma unroll
      for (int c = 0; c < VW; ++c) {
        v[c + VW * r] = __builtin_bit_cast(u16x2, tile[(cy + r) * LW + cx + c]);
      }
    u16x2 o[TX * TY];
    MedianNet<K, TX, TY>::run(v, o);
#pragma unroll
    for (int r = 0; r < TY; ++r)
#pragma unroll
      for (int c = 0; c < TX; ++c) {
        int ox = bx + cx + c;
        int oy = by + cy + r;
        if (ox < a.o0) {
          if (oy < a.o1) store_key_as(a.out, obase + (int64_t)oy * a.o0 + ox, a.dt, a.odt, o[c + TX * r].x);
          if (oy + BYH < a.o1)
            store_key_as(a.out, obase + (int64_t)(oy + BYH) * a.o0 + ox, a.dt, a.odt, o[c + TX * r].y);
        }
      }
  } else {
    uint32_t v[VW * VH];
#pragma unroll
    for (int r = 0; r < VH; ++r)
#pragma unroll
      for (int c = 0; c < VW; ++c) v[c + VW * r] = tile[(cy + r) * LW + cx + c];
    uint32_t o[TX * TY];
    MedianNet<K, TX, TY>::run(v, o);
#pragma unroll
    for (int r = 0; r < TY; ++r)
#pragma unroll
      for (int c = 0; c < TX; ++c) {
        int ox = bx + cx + c;
        int oy = by + cy + r;
        if (ox < a.o0 && oy < a.o1) store_key_as(a.out, obase + (int64_t)oy * a.o0 + ox, a.dt, a.odt, o[c + TX * r]);
      }
  }
}

template <int K, int TX, int TY, int PACK>
__global__ void __launch_bounds__(256) k_median_net(MedianArgs a) {
  using G = NetGeom<K, TX, TY, PACK>;
  __shared__ uint32_t tile[G::LH * G::LW];
  for (int plane = blockIdx.z; plane < a.nplanes; plane += gridDim.z) {
    const int p2 = plane % a.o2, p3 = plane / a.o2;
    const int64_t ibase = ((int64_t)(p2 + a.f2) + (int64_t)a.n2 * (p3 + a.f3)) * a.in_plane;
    median_tile<K, TX, TY, PACK>(a, blockIdx.x * G::BX, blockIdx.y * G::BY, ibase, (int64_t)plane * a.out_plane,
                                 tile);
    __syncthreads();
  }
}

// Batched serving: tile g belongs to request tile_req[g]; its requests are
// 2-D images whose output covers the whole image.
template <int K, int TX, int TY, int PACK>
__global__ void __launch_bounds__(256) k_median_batch(const MedianArgs *__restrict__ reqs,
                                                      const int *__restrict__ tile_req,
                                                      const int *__restrict__ tile_first) {
  using G = NetGeom<K, TX, TY, PACK>;
  __shared__ uint32_t tile[G::LH * G::LW];
  const int g = blockIdx.x;
  const int r = tile_req[g];
  const MedianArgs &a = reqs[r];
  const int local = g - tile_first[r];
  const int ntx = (a.o0 + G::BX - 1) / G::BX;
  median_tile<K, TX, TY, PACK>(a, (local % ntx) * G::BX, (local / ntx) * G::BY, 0, 0, tile);
}

// Generic window median: the thread enumerates its window twice (candidate,
// then rank count), computing each mapped position on the fly.
struct MedianAnyArgs {
  const void *in;
  void *out;
  int n[MAXD];
  int o[MAXD];
  int of[MAXD];
  int wf[MAXD];
  int wl[MAXD];
  int style;
  int dt, odt;
  double fill;
  int64_t total;
  int L;
  int64_t obase;             // output element of region index 0
  int64_t ostride[MAXD];     // output strides (elements) along the region's axes
};

template <typename KT>
__device__ __forceinline__ KT window_key(const MedianAnyArgs &a, const int *q, const int *ilo, const int *ilen, int j) {
  int s[MAXD];
  bool isfill = false;
#pragma unroll
  for (int d = 0; d < MAXD; ++d) {
    int jd = j % a.wl[d];
    j /= a.wl[d];
    int x = q[d] + a.wf[d] + jd;
    if ((unsigned)x < (unsigned)a.n[d]) {
      s[d] = x;
    } else if (a.style == IFM_FILL || ilen[d] <= 0) {
      isfill = true;
      s[d] = 0;
    } else {
      s[d] = ilo[d] + bmap(a.style, x - ilo[d], ilen[d]);
    }
  }
  if (isfill) {
    if constexpr (sizeof(KT) == 8) {
      double f = a.fill;
      uint64_t u = (uint64_t)__double_as_longlong(f);
      if ((u & 0x7fffffffffffffffull) > 0x7ff0000000000000ull) return ~0ull;
      return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
    } else {
      // fill already converted to the eltype; reuse the key function on a local
      float ff = (float)a.fill;
      uint32_t kk;
      if (a.dt == IFM_U8) kk = (uint32_t)(int)a.fill;
      else if (a.dt == IFM_N0F8) kk = (uint32_t)(int)rint(a.fill * 255.0);  // the value code/255
      else if (a.dt == IFM_F32) kk = key32(&ff, 0, IFM_F32);
      else if (a.dt == IFM_BF16) { uint16_t b = f32_to_bf16(ff); kk = key32(&b, 0, IFM_BF16); }
      else { __half h = __float2half_rn(ff); kk = key32(&h, 0, IFM_F16); }
      return kk;
    }
  }
  int64_t idx = s[0] + (int64_t)a.n[0] * (s[1] + (int64_t)a.n[1] * (s[2] + (int64_t)a.n[2] * s[3]));
  if constexpr (sizeof(KT) == 8) return key64(a.in, idx);
  else return key32(a.in, idx, a.dt);
}

template <typename KT>
__global__ void __launch_bounds__(256) k_median_any(MedianAnyArgs a) {
  const int rank = (a.L + 1) / 2 - 1;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < a.total;
       p += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = p, oidx = a.obase;
    int q[MAXD], ilo[MAXD], ilen[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
      const int ql = (int)(r % a.o[d]);
      oidx += ql * a.ostride[d];
      q[d] = ql + a.of[d];
      r /= a.o[d];
      int lo = q[d] + a.wf[d], hi = q[d] + a.wf[d] + a.wl[d] - 1;
      ilo[d] = lo < 0 ? 0 : lo;
      int h2 = hi > a.n[d] - 1 ? a.n[d] - 1 : hi;
      ilen[d] = h2 - ilo[d] + 1;
    }
    KT result = 0;
    for (int i = 0; i < a.L; ++i) {
      KT ki = window_key<KT>(a, q, ilo, ilen, i);
      int less = 0, eq = 0;
      for (int j = 0; j < a.L; ++j) {
        KT kj = window_key<KT>(a, q, ilo, ilen, j);
        less += kj < ki;
        eq += kj == ki;
      }
      if (less <= rank && rank < less + eq) {
        result = ki;
        break;
      }
    }
    if constexpr (sizeof(KT) == 8) store_key64(a.out, oidx, result);
    else store_key_as(a.out, oidx, a.dt, a.odt, result);
  }
}

static uint32_t fill_key32(int dt, double fill) {
  // host twin of key32 for the fill value
  switch (dt) {
  case IFM_U8: return (uint32_t)(int)fill;
  case IFM_N0F8: return (uint32_t)(int)llrint(fill * 255.0);  // the value code/255
  case IFM_F32: {
    float f = (float)fill;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0xffffffffu;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  }
  default: {
    uint32_t u;
    float f = (float)fill;
    if (dt == IFM_BF16) {
      uint32_t b;
      memcpy(&b, &f, 4);
      u = (b + 0x7fffu + ((b >> 16) & 1u)) >> 16;
    } else {
      __half h = __float2half_rn(f);
      uint16_t hb;
      memcpy(&hb, &h, 2);
      u = hb;
    }
    uint32_t expmask = (dt == IFM_BF16) ? 0x7f80u : 0x7c00u;
    if ((u & 0x7fffu) > expmask) return 0xffffu;
    return (u & 0x8000u) ? (~u & 0xffffu) : (u | 0x8000u);
  }
  }
}

template <int K, int TX, int TY, int PACK>
static void launch_net_on(const MedianArgs &a, hipStream_t st) {
  using G = NetGeom<K, TX, TY, PACK>;
  dim3 grid((a.o0 + G::BX - 1) / G::BX, (a.o1 + G::BY - 1) / G::BY, a.nplanes < 65535 ? a.nplanes : 65535);
  hipLaunchKernelGGL((k_median_net<K, TX, TY, PACK>), grid, dim3(256), 0, st, a);
}

// Batched median: tile geometry of the network kernel for window k.
void median_batch_tile(int k, bool pack, int *bx, int *by) {
  const int tx = k == 3 ? 4 : 2, ty = 2;
  const int txt = (32 * tx + k - 1 <= 128) ? 32 : 16;  // NetGeom::TXT
  *bx = txt * tx;
  *by = (256 / txt) * ty * (pack ? 2 : 1);
}

template <int K, int TX, int TY, int PACK>
static void launch_batch_on(const MedianArgs *reqs, const int *tile_req, const int *tile_first, int ntiles,
                            hipStream_t st) {
  hipLaunchKernelGGL((k_median_batch<K, TX, TY, PACK>), dim3(ntiles), dim3(256), 0, st, reqs, tile_req, tile_first);
}

hipError_t launch_median_batch(int k, bool pack, const MedianArgs *reqs, const int *tile_req, const int *tile_first,
                               int ntiles, hipStream_t st) {
  if (ntiles == 0) return hipSuccess;
  if (k == 3) {
    if (pack) launch_batch_on<3, 4, 2, 2>(reqs, tile_req, tile_first, ntiles, st);
    else launch_batch_on<3, 4, 2, 1>(reqs, tile_req, tile_first, ntiles, st);
  } else if (k == 5) {
    if (pack) launch_batch_on<5, 2, 2, 2>(reqs, tile_req, tile_first, ntiles, st);
    else launch_batch_on<5, 2, 2, 1>(reqs, tile_req, tile_first, ntiles, st);
  } else {
    if (pack) launch_batch_on<7, 2, 2, 2>(reqs, tile_req, tile_first, ntiles, st);
    else launch_batch_on<7, 2, 2, 1>(reqs, tile_req, tile_first, ntiles, st);
  }
  return hipGetLastError();
}

// The fill key of a median request (the fill value converted to the input
// eltype, as an order key).
uint32_t median_fill_key(int dt, double fill) { return fill_key32(dt, fill); }

// Generic median of the out region [rf, rf + rn) (out indices) of out.
static hipError_t median_any_region(const Dense4 &in, Dense4 &out, const int64_t *out_first, const int64_t *rf,
                                    const int64_t *rn, const int *wfirst, const int *wlen, int style, double fill,
                                    hipStream_t st) {
  MedianAnyArgs a;
  a.in = in.data;
  a.out = out.data;
  a.L = 1;
  a.total = 1;
  a.obase = 0;
  int64_t stride = 1;
  for (int d = 0; d < MAXD; ++d) {
    a.n[d] = (int)in.n[d];
    a.o[d] = (int)rn[d];
    a.of[d] = (int)(out_first[d] + rf[d]);
    a.wf[d] = wfirst[d];
    a.wl[d] = wlen[d];
    a.L *= wlen[d];
    a.total *= rn[d];
    a.ostride[d] = stride;
    a.obase += rf[d] * stride;
    stride *= out.n[d];
  }
  if (a.total == 0) return hipSuccess;
  a.style = style;
  a.dt = in.dtype;
  a.odt = out.dtype;
  a.fill = fill;
  int64_t blocks = (a.total + 255) / 256;
  if (blocks > 1 << 20) blocks = 1 << 20;
  if (in.dtype == IFM_F64)
    hipLaunchKernelGGL(k_median_any<uint64_t>, dim3((unsigned)blocks), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(k_median_any<uint32_t>, dim3((unsigned)blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

// Circular borders: mapwindow wraps an edge window inside the window's
// intersection with the image (src/mapwindow.jl:310-316, padindex of
// src/border.jl:592-596 on win ∩ img), so outputs whose k x k window leaves
// the image are recomputed after the network kernel has written every output
// (the interior is style-independent).  One thread per band output gathers
// its k*k keys into registers through the window-relative map and selects the
// ((L+1)÷2)-th smallest by counting.
template <int K>
__device__ __forceinline__ void band_one(const BandJob &j, int local, int plane) {
  constexpr int H = K / 2, L = K * K;
  if (local >= j.w * j.h) return;
  const int ox = j.x0 + local % j.w, oy = j.y0 + local / j.w;
  const int x = j.f0 + ox, y = j.f1 + oy;
  const int xlo = max(0, x - H), xm = min(j.n0 - 1, x + H) - xlo + 1;
  const int ylo = max(0, y - H), ym = min(j.n1 - 1, y + H) - ylo + 1;
  const int64_t ibase = (int64_t)plane * j.in_plane;
  uint32_t k[L];
#pragma unroll
  for (int v = 0; v < K; ++v) {
    int yy = y - H + v;
    if (yy < ylo || yy >= ylo + ym) yy = ylo + imod(yy - ylo, ym);
#pragma unroll
    for (int u = 0; u < K; ++u) {
      int xx = x - H + u;
      if (xx < xlo || xx >= xlo + xm) xx = xlo + imod(xx - xlo, xm);
      k[v * K + u] = key32(j.in, ibase + (int64_t)yy * j.n0 + xx, j.dt);
    }
  }
  constexpr int rank = (L + 1) / 2 - 1;
  uint32_t res = k[0];
#pragma unroll
  for (int i = 0; i < L; ++i) {
    int less = 0, eq = 0;
#pragma unroll
    for (int t = 0; t < L; ++t) {
      less += k[t] < k[i];
      eq += k[t] == k[i];
    }
    res = (less <= rank && rank < less + eq) ? k[i] : res;
  }
  store_key_as(j.out, (int64_t)plane * j.out_plane + (int64_t)oy * j.opitch + ox, j.dt, j.odt, res);
}

struct BandJobs4 {
  BandJob j[4];
  int first[5];  // first block of each job
};

template <int K>
__global__ void __launch_bounds__(256) k_median_band4(BandJobs4 jobs) {
  const int b = blockIdx.x;
  int s = 0;
  while (s < 3 && b >= jobs.first[s + 1]) ++s;
  for (int plane = blockIdx.y; plane < jobs.j[s].nplanes; plane += gridDim.y)
    band_one<K>(jobs.j[s], (b - jobs.first[s]) * 256 + threadIdx.x, plane);
}

template <int K>
__global__ void __launch_bounds__(256) k_median_band_batch(const BandJob *__restrict__ jobs,
                                                           const int *__restrict__ tile_req,
                                                           const int *__restrict__ tile_first) {
  const int g = blockIdx.x, r = tile_req[g];
  band_one<K>(jobs[r], (g - tile_first[r]) * 256 + threadIdx.x, 0);
}

// The (up to) four band strips of an out region, in out indices.
int median_band_jobs(const Dense4 &in, const Dense4 &out, const int64_t *out_first, int k, BandJob *jobs) {
  const int64_t h = k / 2, o0 = out.n[0], o1 = out.n[1];
  auto clampi = [](int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); };
  const int64_t top = clampi(h - out_first[1], 0, o1), bot = clampi(in.n[1] - h - out_first[1], top, o1);
  const int64_t left = clampi(h - out_first[0], 0, o0), right = clampi(in.n[0] - h - out_first[0], left, o0);
  const int64_t rect[4][4] = {{0, 0, o0, top}, {0, bot, o0, o1 - bot}, {0, top, left, bot - top},
                              {right, top, o0 - right, bot - top}};
  int n = 0;
  for (int s = 0; s < 4; ++s) {
    if (rect[s][2] <= 0 || rect[s][3] <= 0) continue;
    BandJob &j = jobs[n++];
    j.in = in.data;
    j.out = out.data;
    j.n0 = (int)in.n[0];
    j.n1 = (int)in.n[1];
    j.in_plane = in.n[0] * in.n[1];
    j.out_plane = out.n[0] * out.n[1];
    j.nplanes = (int)(out.n[2] * out.n[3]);
    j.opitch = (int)out.n[0];
    j.f0 = (int)out_first[0];
    j.f1 = (int)out_first[1];
    j.x0 = (int)rect[s][0];
    j.y0 = (int)rect[s][1];
    j.w = (int)rect[s][2];
    j.h = (int)rect[s][3];
    j.dt = in.dtype;
    j.odt = out.dtype;
  }
  return n;
}

hipError_t median_circular_band(const Dense4 &in, Dense4 &out, const int64_t *out_first, int k, double fill,
                                hipStream_t st) {
  (void)fill;  // circular windows never read the fill value
  BandJobs4 jobs = {};
  const int n = median_band_jobs(in, out, out_first, k, jobs.j);
  if (n == 0) return hipSuccess;
  int tot = 0;
  for (int s = 0; s < 4; ++s) {
    jobs.first[s] = tot;
    if (s < n) tot += (jobs.j[s].w * jobs.j[s].h + 255) / 256;
  }
  jobs.first[4] = tot;
  for (int s = n; s < 4; ++s) jobs.first[s] = tot;  // empty jobs are never selected
  const int np = jobs.j[0].nplanes < 65535 ? jobs.j[0].nplanes : 65535;
  dim3 grid(tot, np);
  if (k == 3) hipLaunchKernelGGL(k_median_band4<3>, grid, dim3(256), 0, st, jobs);
  else if (k == 5) hipLaunchKernelGGL(k_median_band4<5>, grid, dim3(256), 0, st, jobs);
  else hipLaunchKernelGGL(k_median_band4<7>, grid, dim3(256), 0, st, jobs);
  return hipGetLastError();
}

hipError_t launch_median_band_batch(int k, const BandJob *jobs, const int *tile_req, const int *tile_first, int ntiles,
                                    hipStream_t st) {
  if (ntiles == 0) return hipSuccess;
  if (k == 3) hipLaunchKernelGGL(k_median_band_batch<3>, dim3(ntiles), dim3(256), 0, st, jobs, tile_req, tile_first);
  else if (k == 5) hipLaunchKernelGGL(k_median_band_batch<5>, dim3(ntiles), dim3(256), 0, st, jobs, tile_req, tile_first);
  else hipLaunchKernelGGL(k_median_band_batch<7>, dim3(ntiles), dim3(256), 0, st, jobs, tile_req, tile_first);
  return hipGetLastError();
}

hipError_t launch_median(const Dense4 &in, Dense4 &out, const int64_t *out_first, const int *wfirst,
                         const int *wlen, int style, double fill, hipStream_t st) {
  if (out.count() == 0) return hipSuccess;
  // The tiled network path needs a centred odd square 2-D window; its
  // image-relative border map equals mapwindow's window-relative one for such
  // windows except for circular, whose edge band is recomputed afterwards.
  int k = wlen[0];
  bool square = (k == 3 || k == 5 || k == 7) && wlen[1] == k && wfirst[0] == -(k / 2) && wfirst[1] == -(k / 2) &&
                wlen[2] == 1 && wlen[3] == 1 && wfirst[2] == 0 && wfirst[3] == 0;
  bool small = in.n[0] >= k && in.n[1] >= k;  // mirrored maps stay within the window
  if (square && small && in.dtype != IFM_F64) {
    MedianArgs a;
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
    a.style = style;
    a.dt = in.dtype;
    a.odt = out.dtype;
    a.fillkey = fill_key32(in.dtype, fill);
    bool pack = in.dtype != IFM_F32;
    if (k == 3) {
      if (pack) launch_net_on<3, 4, 2, 2>(a, st); else launch_net_on<3, 4, 2, 1>(a, st);
    } else if (k == 5) {
      if (pack) launch_net_on<5, 2, 2, 2>(a, st); else launch_net_on<5, 2, 2, 1>(a, st);
    } else {
      if (pack) launch_net_on<7, 2, 2, 2>(a, st); else launch_net_on<7, 2, 2, 1>(a, st);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && style == IFM_CIRCULAR) e = median_circular_band(in, out, out_first, k, fill, st);
    return e;
  }
  const int64_t rf[MAXD] = {0, 0, 0, 0};
  return median_any_region(in, out, out_first, rf, out.n, wfirst, wlen, style, fill, st);
}

// ---------------------------------------------------------------------------
// min / max over truncated windows.  Julia's min/max on floats propagate NaN
// and order -0.0 below +0.0: the min pipeline maps NaN to key 0 (smallest),
// the max pipeline to the all-ones key (largest).
// ---------------------------------------------------------------------------
template <typename KT>
__device__ __forceinline__ KT minkey_of(KT k, int dt) {
  if constexpr (sizeof(KT) == 8) return k == ~0ull ? 0ull : k;
  else {
    if (dt == IFM_F32) return k == 0xffffffffu ? 0u : k;
    if (dt == IFM_BF16 || dt == IFM_F16) return k == 0xffffu ? 0u : k;
    return k;
  }
}
template <typename KT>
__device__ __forceinline__ KT minkey_back(KT k, int dt) {
  if constexpr (sizeof(KT) == 8) return k == 0ull ? ~0ull : k;
  else {
    if (dt == IFM_F32) return k == 0u ? 0xffffffffu : k;
    if (dt == IFM_BF16 || dt == IFM_F16) return k == 0u ? 0xffffu : k;
    return k;
  }
}
template <typename KT> __device__ __forceinline__ KT key_top(int dt) {
  if constexpr (sizeof(KT) == 8) return ~0ull;
  else return (dt == IFM_U8 || dt == IFM_N0F8) ? 0xffu : ((dt == IFM_F32) ? 0xffffffffu : 0xffffu);
}

template <typename KT>
__device__ __forceinline__ KT load_key(const void *p, int64_t i, int dt) {
  if constexpr (sizeof(KT) == 8) return key64(p, i);
  else return key32(p, i, dt);
}
template <typename KT>
__device__ __forceinline__ void put_key(void *p, int64_t i, int dt, KT k) {
  if constexpr (sizeof(KT) == 8) store_key64(p, i, k);
  else store_key32(p, i, dt, k);
}

// One separable pass along canonical axis `axis` over key arrays (first pass
// reads the image, later passes read the previous pass's keys).  The last
// pass writes the output eltype and merges the fill key where a window leaves
// the image on any axis.
struct MinmaxPassArgs {
  const void *src_img;   // first pass: image; else nullptr
  const void *src_min;   // key buffers of the previous pass
  const void *src_max;
  void *dst_min;         // keys (intermediate) or output (last pass)
  void *dst_max;
  int in_n[MAXD];        // extents of the source
  int out_n[MAXD];
  int shift[MAXD];       // out coord + shift = source coord (non-filtered axes)
  int img_n[MAXD];
  int of[MAXD];          // out_first (image coordinates of out index 0)
  int wf[MAXD], wl[MAXD];
  int axis;
  int first_pass, last_pass;
  int dt;
  int isfill;
  int64_t total;
};

template <typename KT>
__global__ void __launch_bounds__(256) k_minmax_pass(MinmaxPassArgs a, KT fillkey) {
  const KT top = key_top<KT>(a.dt);
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < a.total;
       q += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = q;
    int o[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
      o[d] = (int)(r % a.out_n[d]);
      r /= a.out_n[d];
    }
    int64_t base = 0, stride = 1, astride = 1;
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
      if (d == a.axis) astride = stride;
      else base += (int64_t)(o[d] + a.shift[d]) * stride;
      stride *= a.in_n[d];
    }
    const int n = a.in_n[a.axis];
    const int x = o[a.axis] + a.of[a.axis] + a.wf[a.axis];
    int lo = x < 0 ? 0 : x;
    int hi = x + a.wl[a.axis] - 1;
    if (hi > n - 1) hi = n - 1;
    KT mn = top, mx = 0;
    for (int s = lo; s <= hi; ++s) {
      int64_t idx = base + (int64_t)s * astride;
      KT kn, kx;
      if (a.first_pass) {
        KT k = load_key<KT>(a.src_img, idx, a.dt);
        kn = minkey_of<KT>(k, a.dt);
        kx = k;
      } else {
        kn = a.src_min ? ((const KT *)a.src_min)[idx] : top;
        kx = a.src_max ? ((const KT *)a.src_max)[idx] : 0;
      }
      mn = kn < mn ? kn : mn;
      mx = kx > mx ? kx : mx;
    }
    if (a.last_pass) {
      if (a.isfill) {
        bool outside = false;
#pragma unroll
        for (int d = 0; d < MAXD; ++d) {
          int xd = (d == a.axis ? o[d] : o[d]) + a.of[d] + a.wf[d];
          if (xd < 0 || xd + a.wl[d] - 1 > a.img_n[d] - 1) outside = true;
        }
        if (outside) {
          KT fn = minkey_of<KT>(fillkey, a.dt);
          mn = fn < mn ? fn : mn;
          mx = fillkey > mx ? fillkey : mx;
        }
      }
      if (a.dst_min) put_key<KT>(a.dst_min, q, a.dt, minkey_back<KT>(mn, a.dt));
      if (a.dst_max) put_key<KT>(a.dst_max, q, a.dt, mx);
    } else {
      if (a.dst_min) ((KT *)a.dst_min)[q] = mn;
      if (a.dst_max) ((KT *)a.dst_max)[q] = mx;
    }
  }
}

// ---------------------------------------------------------------------------
// 2-D min / max in one pass: a 64 x 32 output tile loads its (32 + w1 - 1) x
// (64 + w0 - 1) input keys into LDS with clamped (replicate) coordinates --
// the min / max over a clamped window equals the one over the window's
// intersection with the image whenever the window contains its own pixel --
// then reduces along axis 0 into two LDS tiles and along axis 1 into the
// output.  With a compile-time window the reductions are log-depth: level s
// turns m[x] = op(v[x .. x+s)) into op(v[x .. x+2s)), and the window result
// is op(m[x], m[x + w - p]) for the largest power of two p <= w.
// ---------------------------------------------------------------------------
constexpr int MM_TX = 64, MM_TY = 32, MM_MAXW = 33;
// Row pitch of the axis-0 results: odd, so the 8 rows one wave writes at once
// fall on different banks.
constexpr int MM_TXP = MM_TX + 1;

struct Minmax2dArgs {
  const void *in;
  void *omin, *omax;
  int n0, n1;            // image
  int o0, o1;            // output extents
  int f0, f1;            // image coordinates of output (0, 0)
  int wf0, wf1, wl0, wl1;
  int isfill;
  uint32_t fillkey;
  int64_t in_plane, out_plane;
  int nplanes;
};

template <int W, int NO>
__device__ __forceinline__ void win_minmax(uint32_t (&vn)[NO + W - 1], uint32_t (&vx)[NO + W - 1], uint32_t (&on)[NO],
                                           uint32_t (&ox)[NO]) {
  constexpr int L = NO + W - 1;
  constexpr int P = W >= 32 ? 32 : W >= 16 ? 16 : W >= 8 ? 8 : W >= 4 ? 4 : W >= 2 ? 2 : 1;
#pragma unroll
  for (int s = 1; 2 * s <= P; s *= 2) {
#pragma unroll
    for (int x = 0; x + s < L; ++x) {  // m[x+s] is still the previous level's value
      vn[x] = min(vn[x], vn[x + s]);
      vx[x] = max(vx[x], vx[x + s]);
    }
  }
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    on[o] = min(vn[o], vn[o + W - P]);
    ox[o] = max(vx[o], vx[o + W - P]);
  }
}

template <int DT>
__device__ __forceinline__ uint32_t mm_minkey(uint32_t k) {
  if constexpr (DT == IFM_F32) return k == 0xffffffffu ? 0u : k;
  else if constexpr (DT == IFM_BF16 || DT == IFM_F16) return k == 0xffffu ? 0u : k;
  else return k;
}

// DT: IFM_U8 (also N0f8), IFM_BF16, IFM_F16 or IFM_F32; W0 = W1 = 0: runtime window.
template <int DT, int W0, int W1>
__global__ void __launch_bounds__(256) k_minmax2d(Minmax2dArgs a) {
  extern __shared__ uint32_t mm_lds[];
  const int w0 = W0 ? W0 : a.wl0, w1 = W1 ? W1 : a.wl1;
  const int iw = MM_TX + w0 - 1, ih = MM_TY + w1 - 1, iwp = iw | 1;
  uint32_t *s_in = mm_lds;                 // ih x iwp raw keys
  uint32_t *s_n = mm_lds + ih * iwp;       // ih x 64 min keys after axis 0
  uint32_t *s_x = s_n + ih * MM_TXP;       // ih x 64 max keys after axis 0 (pitch 65)
  const int tid = threadIdx.x;
  const int ox0 = blockIdx.x * MM_TX, oy0 = blockIdx.y * MM_TY;
  const int gx0 = a.f0 + ox0 + a.wf0, gy0 = a.f1 + oy0 + a.wf1;
  for (int plane = blockIdx.z; plane < a.nplanes; plane += gridDim.z) {
    const int64_t ibase = (int64_t)plane * a.in_plane, obase = (int64_t)plane * a.out_plane;
    if constexpr (W0 > 0 && W1 > 0) {
      // Stage: lane c owns columns c and c + 64, wave w rows w, w + 4, ...
      // (a row's clamp is wave-uniform); every load is issued before the
      // first LDS write.
      constexpr int IW = MM_TX + W0 - 1, IH = MM_TY + W1 - 1, NR = (IH + 3) / 4;
      static_assert(IW <= 128, "two columns per lane");
      const int c = tid & 63;
      const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
      int sx[2];
      bool cok[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        cok[h] = c + 64 * h < IW;
        sx[h] = min(max(gx0 + c + 64 * h, 0), a.n0 - 1);
      }
      uint32_t v[NR][2];
#pragma unroll
      for (int j = 0; j < NR; ++j) {
        const int r = w + 4 * j;
        const int sy = min(max(gy0 + r, 0), a.n1 - 1);
#pragma unroll
        for (int h = 0; h < 2; ++h)
          if (r < IH && cok[h]) v[j][h] = key32(a.in, ibase + (int64_t)sy * a.n0 + sx[h], DT);
      }
#pragma unroll
      for (int j = 0; j < NR; ++j) {
        const int r = w + 4 * j;
#pragma unroll
        for (int h = 0; h < 2; ++h)
          if (r < IH && cok[h]) s_in[r * iwp + c + 64 * h] = v[j][h];
      }
    } else {
      // Stage: the thread's column, every other row, 8 loads in flight.
      const int c = tid & 127, rh = tid >> 7;
      const bool colok = c < iw;
      const int sx = min(max(gx0 + (colok ? c : iw - 1), 0), a.n0 - 1);
      const int nrow = (ih - rh + 1) >> 1;
      for (int j0 = 0; j0 < nrow; j0 += 8) {
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int jj = j0 + j < nrow ? j0 + j : nrow - 1;
          const int sy = min(max(gy0 + rh + 2 * jj, 0), a.n1 - 1);
          v[j] = key32(a.in, ibase + (int64_t)sy * a.n0 + sx, DT);
        }
        if (colok) {
#pragma unroll
          for (int j = 0; j < 8; ++j)
            if (j0 + j < nrow) s_in[(rh + 2 * (j0 + j)) * iwp + c] = v[j];
        }
      }
    }
    __syncthreads();
    // Axis 0: item = (row, group of 8 outputs).
    for (int item = tid; item < ih * 8; item += 256) {
      const int r = item >> 3, c0 = (item & 7) * 8;
      const uint32_t *src = s_in + r * iwp + c0;
      uint32_t on[8], ox[8];
      if constexpr (W0 > 0) {
        uint32_t vn[8 + W0 - 1], vx[8 + W0 - 1];
#pragma unroll
        for (int t = 0; t < 8 + W0 - 1; ++t) {
          vx[t] = src[t];
          vn[t] = mm_minkey<DT>(vx[t]);
        }
        win_minmax<W0, 8>(vn, vx, on, ox);
      } else {
#pragma unroll
        for (int o = 0; o < 8; ++o) {
          on[o] = 0xffffffffu;
          ox[o] = 0u;
        }
        for (int t = 0; t < w0 + 7; ++t) {
          const uint32_t k = src[t], kn = mm_minkey<DT>(k);
#pragma unroll
          for (int o = 0; o < 8; ++o)
            if (t - o >= 0 && t - o < w0) {
              on[o] = min(on[o], kn);
              ox[o] = max(ox[o], k);
            }
        }
      }
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        s_n[r * MM_TXP + c0 + o] = on[o];
        s_x[r * MM_TXP + c0 + o] = ox[o];
      }
    }
    __syncthreads();
    // Axis 1: thread = (column, group of 8 rows).
    {
      const int col = tid & 63, r0 = (tid >> 6) * 8;
      uint32_t on[8], ox[8];
      if constexpr (W1 > 0) {
        uint32_t vn[8 + W1 - 1], vx[8 + W1 - 1];
#pragma unroll
        for (int t = 0; t < 8 + W1 - 1; ++t) {
          vn[t] = s_n[(r0 + t) * MM_TXP + col];
          vx[t] = s_x[(r0 + t) * MM_TXP + col];
        }
        win_minmax<W1, 8>(vn, vx, on, ox);
      } else {
#pragma unroll
        for (int o = 0; o < 8; ++o) {
          on[o] = 0xffffffffu;
          ox[o] = 0u;
        }
        for (int t = 0; t < w1 + 7; ++t) {
          const uint32_t kn = s_n[(r0 + t) * MM_TXP + col], kx = s_x[(r0 + t) * MM_TXP + col];
#pragma unroll
          for (int o = 0; o < 8; ++o)
            if (t - o >= 0 && t - o < w1) {
              on[o] = min(on[o], kn);
              ox[o] = max(ox[o], kx);
            }
        }
      }
      const int ox_ = ox0 + col;
      if (ox_ < a.o0) {
        const int gx = a.f0 + ox_ + a.wf0;
        const bool xout = gx < 0 || gx + w0 - 1 > a.n0 - 1;
#pragma unroll
        for (int o = 0; o < 8; ++o) {
          const int oy = oy0 + r0 + o;
          if (oy >= a.o1) break;
          uint32_t mn = on[o], mx = ox[o];
          if (a.isfill) {
            const int gy = a.f1 + oy + a.wf1;
            if (xout || gy < 0 || gy + w1 - 1 > a.n1 - 1) {
              mn = min(mn, mm_minkey<DT>(a.fillkey));
              mx = max(mx, a.fillkey);
            }
          }
          const int64_t q = obase + (int64_t)oy * a.o0 + ox_;
          if (a.omin) store_key32(a.omin, q, DT, minkey_back<uint32_t>(mn, DT));
          if (a.omax) store_key32(a.omax, q, DT, mx);
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Min / max of 8-bit keys (UInt8 and N0f8 codes) over a W x W window, W = 3,
// 5, 7, 9: the separable doubling scheme of k_minmax2d on keys packed two to a
// register, so each v_pk_min_u16 / v_pk_max_u16 serves two output rows.
// Tile 64 x 64 outputs; packed row p holds input rows p and p + 32 (the two
// halves of the tile), so both passes run element-wise on the halves.
// Staging reads four codes per lane (aligned words) on tiles whose columns
// lie inside the image, and results leave through LDS as 16-byte stores.
// Lanes of the axis-0 pass run over packed rows (odd row pitch), so their LDS
// accesses fall on distinct banks.  Tiles are assigned XCD by XCD (xcd_tile):
// a 64-byte tile row shares its cache lines with its neighbours.  WANT: 1 minimum, 2 maximum, 3 both (one
// after the other through the same LDS buffers).
// ---------------------------------------------------------------------------
constexpr int M8_TX = 64, M8_TY = 64, M8_HY = 32, M8_NT = 256;

template <int W, int NO>
__device__ __forceinline__ void win_min2(u16x2 (&v)[NO + W - 1], u16x2 (&o)[NO]) {
  constexpr int L = NO + W - 1;
  constexpr int P = W >= 8 ? 8 : W >= 4 ? 4 : W >= 2 ? 2 : 1;
#pragma unroll
  for (int s = 1; 2 * s <= P; s *= 2) {
#pragma unroll
    for (int x = 0; x + s < L; ++x) v[x] = kmin(v[x], v[x + s]);
  }
#pragma unroll
  for (int i = 0; i < NO; ++i) o[i] = kmin(v[i], v[i + W - P]);
}

template <int W, int NO>
__device__ __forceinline__ void win_max2(u16x2 (&v)[NO + W - 1], u16x2 (&o)[NO]) {
  constexpr int L = NO + W - 1;
  constexpr int P = W >= 8 ? 8 : W >= 4 ? 4 : W >= 2 ? 2 : 1;
#pragma unroll
  for (int s = 1; 2 * s <= P; s *= 2) {
#pragma unroll
    for (int x = 0; x + s < L; ++x) v[x] = kmax(v[x], v[x + s]);
  }
#pragma unroll
  for (int i = 0; i < NO; ++i) o[i] = kmax(v[i], v[i + W - P]);
}

__device__ __forceinline__ u16x2 pack2(uint32_t lo, uint32_t hi) {
  u16x2 r;
  r.x = (uint16_t)lo;
  r.y = (uint16_t)hi;
  return r;
}

template <int W, int WANT>
__global__ void __launch_bounds__(M8_NT) k_minmax2d_u8(Minmax2dArgs a) {
  constexpr int IW = M8_TX + W - 1;        // input columns of a tile
  constexpr int NW = (IW + 6) / 4;         // aligned words covering them (any shift 0..3)
  constexpr int SIN = 4 * NW + 1;          // staged row pitch (packed keys), odd
  constexpr int NP = M8_HY + W - 1;        // packed rows
  constexpr int SHP = M8_TX + 1;           // pitch after the axis-0 pass
  constexpr int NI = (NP * NW + M8_NT - 1) / M8_NT;
  __shared__ u16x2 s_in[NP * SIN];
  __shared__ u16x2 s_h[NP * SHP];
  __shared__ uint32_t s_out[M8_TY * M8_TX / 4];
  const int tid = threadIdx.x;
  const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int tx_ = tile % gridDim.x, ty_ = tile / gridDim.x;
  const int ox0 = tx_ * M8_TX;
  const int gx0 = a.f0 + ox0 + a.wf0;
  const uint8_t *in = (const uint8_t *)a.in;
  // words: columns inside the image, rows of 4-byte multiples, aligned base
  const bool words = gx0 >= 0 && gx0 + IW <= a.n0 && (a.n0 & 3) == 0 && ((uintptr_t)a.in & 3) == 0;
  const int xa = words ? (gx0 & ~3) : gx0;  // image column of staged column 0
  const int sh = gx0 - xa;                  // staged column of tile column 0
  const int nwr = (sh + IW + 3) >> 2;       // words a row needs (ends at or before the row's end)
  const int oy0 = ty_ * M8_TY;
  const int gy0 = a.f1 + oy0 + a.wf1;
  for (int plane = blockIdx.z; plane < a.nplanes; plane += gridDim.z) {
    const int64_t ibase = (int64_t)plane * a.in_plane, obase = (int64_t)plane * a.out_plane;
    {
      if (words) {
        // every load of the tile issued before the first LDS write
        uint32_t w0[NI], w1[NI];
#pragma unroll
        for (int i = 0; i < NI; ++i) {
          const int e = tid + M8_NT * i, p = e / NW, j = e - p * NW;
          if (e < NP * NW && j < nwr) {
            const int sy0 = min(max(gy0 + p, 0), a.n1 - 1), sy1 = min(max(gy0 + p + M8_HY, 0), a.n1 - 1);
            w0[i] = *(const uint32_t *)(in + ibase + (int64_t)sy0 * a.n0 + xa + 4 * j);
            w1[i] = *(const uint32_t *)(in + ibase + (int64_t)sy1 * a.n0 + xa + 4 * j);
          }
        }
#pragma unroll
        for (int i = 0; i < NI; ++i) {
          const int e = tid + M8_NT * i, p = e / NW, j = e - p * NW;
          if (e < NP * NW && j < nwr) {
#pragma unroll
            for (int b = 0; b < 4; ++b)
              s_in[p * SIN + 4 * j + b] = pack2((w0[i] >> (8 * b)) & 0xff, (w1[i] >> (8 * b)) & 0xff);
          }
        }
      } else {
        for (int e = tid; e < NP * IW; e += M8_NT) {
          const int p = e / IW, c = e - p * IW;
          const int sx = min(max(gx0 + c, 0), a.n0 - 1);
          const int sy0 = min(max(gy0 + p, 0), a.n1 - 1), sy1 = min(max(gy0 + p + M8_HY, 0), a.n1 - 1);
          s_in[p * SIN + c] = pack2(in[ibase + (int64_t)sy0 * a.n0 + sx], in[ibase + (int64_t)sy1 * a.n0 + sx]);
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        if (!(WANT & (1 << q))) continue;
        const bool isMin = q == 0;
        // axis 0: item = (packed row, group of 8 outputs), packed rows across lanes
        for (int item = tid; item < NP * (M8_TX / 8); item += M8_NT) {
          const int p = item % NP, c0 = (item / NP) * 8;
          const u16x2 *src = s_in + p * SIN + sh + c0;
          u16x2 v[8 + W - 1], o[8];
#pragma unroll
          for (int k = 0; k < 8 + W - 1; ++k) v[k] = src[k];
          if (isMin) win_min2<W, 8>(v, o);
          else win_max2<W, 8>(v, o);
#pragma unroll
          for (int i = 0; i < 8; ++i) s_h[p * SHP + c0 + i] = o[i];
        }
        __syncthreads();
        // axis 1: thread = (column, group of 8 packed rows); bytes to s_out
        {
          const int col = tid % M8_TX, r0 = (tid / M8_TX) * 8;
          const int gx = a.f0 + ox0 + col + a.wf0;
          const bool xout = gx < 0 || gx + W - 1 > a.n0 - 1;
          u16x2 v[8 + W - 1], o[8];
#pragma unroll
          for (int k = 0; k < 8 + W - 1; ++k) v[k] = s_h[(r0 + k) * SHP + col];
          if (isMin) win_min2<W, 8>(v, o);
          else win_max2<W, 8>(v, o);
          uint8_t *so = (uint8_t *)s_out;
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            uint32_t lo = o[i].x, hi = o[i].y;
            if (a.isfill) {
              const int gyl = a.f1 + oy0 + r0 + i + a.wf1, gyh = gyl + M8_HY;
              const uint32_t fk = a.fillkey;
              if (xout || gyl < 0 || gyl + W - 1 > a.n1 - 1) lo = isMin ? min(lo, fk) : max(lo, fk);
              if (xout || gyh < 0 || gyh + W - 1 > a.n1 - 1) hi = isMin ? min(hi, fk) : max(hi, fk);
            }
            so[(r0 + i) * M8_TX + col] = (uint8_t)lo;
            so[(r0 + i + M8_HY) * M8_TX + col] = (uint8_t)hi;
          }
        }
        __syncthreads();
        // stores: thread = (row, 16 columns)
        {
          const int row = tid / (M8_TX / 16), c16 = (tid % (M8_TX / 16)) * 16;
          const int oy = oy0 + row;
          uint8_t *dst = (uint8_t *)(isMin ? a.omin : a.omax);
          const bool vec = ox0 + M8_TX <= a.o0 && (a.o0 & 15) == 0 && ((uintptr_t)dst & 15) == 0;
          if (oy < a.o1) {
            const uint32_t *srow = s_out + (row * M8_TX + c16) / 4;
            const int64_t off = obase + (int64_t)oy * a.o0 + ox0 + c16;
            if (vec) {
              *(uint4 *)(dst + off) = make_uint4(srow[0], srow[1], srow[2], srow[3]);
            } else {
              const uint8_t *sb = (const uint8_t *)srow;
              for (int b = 0; b < 16; ++b)
                if (ox0 + c16 + b < a.o0) dst[off + b] = sb[b];
            }
          }
        }
      }
      __syncthreads();
    }
  }
}

template <int W>
static void mm8_launch(const Minmax2dArgs &a, dim3 grid, hipStream_t st) {
  if (a.omin && a.omax) hipLaunchKernelGGL((k_minmax2d_u8<W, 3>), grid, dim3(M8_NT), 0, st, a);
  else if (a.omin) hipLaunchKernelGGL((k_minmax2d_u8<W, 1>), grid, dim3(M8_NT), 0, st, a);
  else hipLaunchKernelGGL((k_minmax2d_u8<W, 2>), grid, dim3(M8_NT), 0, st, a);
}

template <int DT, int W>
static void mm_launch(const Minmax2dArgs &a, dim3 grid, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((k_minmax2d<DT, W, W>), grid, dim3(256), lds, st, a);
}

template <int DT>
static void mm_dispatch(const Minmax2dArgs &a, dim3 grid, size_t lds, hipStream_t st) {
  const int w = a.wl0 == a.wl1 ? a.wl0 : 0;
  switch (w) {
  case 3: mm_launch<DT, 3>(a, grid, lds, st); break;
  case 5: mm_launch<DT, 5>(a, grid, lds, st); break;
  case 7: mm_launch<DT, 7>(a, grid, lds, st); break;
  case 9: mm_launch<DT, 9>(a, grid, lds, st); break;
  default: mm_launch<DT, 0>(a, grid, lds, st); break;
  }
}

// One-pass 2-D path; false when the shape needs the general passes.
static bool minmax2d(const Dense4 &in, Dense4 *omin, Dense4 *omax, const int64_t *out_first, const int *wfirst,
                     const int *wlen, int isfill, double fill, hipStream_t st, hipError_t *err) {
  Dense4 &ref = omin ? *omin : *omax;
  if (in.dtype == IFM_F64) return false;
  for (int d = 2; d < MAXD; ++d)
    if (wlen[d] != 1 || wfirst[d] != 0 || out_first[d] != 0 || ref.n[d] != in.n[d]) return false;
  for (int d = 0; d < 2; ++d)
    if (wlen[d] > MM_MAXW || wfirst[d] > 0 || wfirst[d] + wlen[d] - 1 < 0) return false;  // window holds its pixel
  Minmax2dArgs a;
  a.in = in.data;
  a.omin = omin ? omin->data : nullptr;
  a.omax = omax ? omax->data : nullptr;
  a.n0 = (int)in.n[0];
  a.n1 = (int)in.n[1];
  a.o0 = (int)ref.n[0];
  a.o1 = (int)ref.n[1];
  a.f0 = (int)out_first[0];
  a.f1 = (int)out_first[1];
  a.wf0 = wfirst[0];
  a.wf1 = wfirst[1];
  a.wl0 = wlen[0];
  a.wl1 = wlen[1];
  a.isfill = isfill;
  a.fillkey = fill_key32(in.dtype, fill);
  a.in_plane = in.n[0] * in.n[1];
  a.out_plane = ref.n[0] * ref.n[1];
  a.nplanes = (int)(ref.n[2] * ref.n[3]);
  const int iw = MM_TX + a.wl0 - 1, ih = MM_TY + a.wl1 - 1;
  const size_t lds = sizeof(uint32_t) * ((size_t)ih * (iw | 1) + 2 * (size_t)ih * MM_TXP);
  dim3 grid((a.o0 + MM_TX - 1) / MM_TX, (a.o1 + MM_TY - 1) / MM_TY, a.nplanes < 65535 ? a.nplanes : 65535);
  const bool u8 = in.dtype == IFM_U8 || in.dtype == IFM_N0F8;
  if (u8 && a.wl0 == a.wl1 && (a.wl0 == 3 || a.wl0 == 5 || a.wl0 == 7 || a.wl0 == 9)) {
    const dim3 g8((a.o0 + M8_TX - 1) / M8_TX, (a.o1 + M8_TY - 1) / M8_TY, grid.z);
    switch (a.wl0) {
    case 3: mm8_launch<3>(a, g8, st); break;
    case 5: mm8_launch<5>(a, g8, st); break;
    case 7: mm8_launch<7>(a, g8, st); break;
    default: mm8_launch<9>(a, g8, st); break;
    }
    *err = hipGetLastError();
    return true;
  }
  switch (in.dtype) {
  case IFM_F32: mm_dispatch<IFM_F32>(a, grid, lds, st); break;
  case IFM_BF16: mm_dispatch<IFM_BF16>(a, grid, lds, st); break;
  case IFM_F16: mm_dispatch<IFM_F16>(a, grid, lds, st); break;
  default: mm_dispatch<IFM_U8>(a, grid, lds, st); break;  // U8 and N0f8: the code is the key
  }
  *err = hipGetLastError();
  return true;
}

size_t minmax_workspace(const Dense4 &in, const int64_t *osize) {
  // two ping-pong key buffers per extreme, sized for the largest intermediate
  size_t ks = in.dtype == IFM_F64 ? 8 : 4;
  int64_t c = 1;
  for (int d = 0; d < MAXD; ++d) c *= (in.n[d] > osize[d] ? in.n[d] : osize[d]);
  return 4 * (size_t)c * ks;
}

hipError_t launch_minmax(const Dense4 &in, Dense4 *omin, Dense4 *omax, const int64_t *out_first,
                         const int *wfirst, const int *wlen, int isfill, double fill, void *work,
                         hipStream_t st) {
  Dense4 &ref = omin ? *omin : *omax;
  if (ref.count() == 0) return hipSuccess;
  hipError_t e2d;
  if (minmax2d(in, omin, omax, out_first, wfirst, wlen, isfill, fill, st, &e2d)) return e2d;
  bool k64 = in.dtype == IFM_F64;
  size_t ks = k64 ? 8 : 4;
  int64_t osz[MAXD];
  for (int d = 0; d < MAXD; ++d) osz[d] = ref.n[d];
  int64_t cap = 1;
  for (int d = 0; d < MAXD; ++d) cap *= (in.n[d] > osz[d] ? in.n[d] : osz[d]);
  char *w = (char *)work;
  void *bufmin[2] = {w, w + cap * ks};
  void *bufmax[2] = {w + 2 * cap * ks, w + 3 * cap * ks};
  // axes with a window longer than 1 (or offset) are filtered; at least one pass
  int axes[MAXD], na = 0;
  for (int d = 0; d < MAXD; ++d)
    if (wlen[d] != 1 || wfirst[d] != 0 || osz[d] != in.n[d] || out_first[d] != 0) axes[na++] = d;
  if (na == 0) axes[na++] = 0;
  // current source extents / offsets (image coordinates)
  int64_t cur_n[MAXD], cur_first[MAXD];
  for (int d = 0; d < MAXD; ++d) { cur_n[d] = in.n[d]; cur_first[d] = 0; }
  const void *smin = nullptr, *smax = nullptr;
  uint64_t fk64 = 0;
  uint32_t fk32 = fill_key32(in.dtype, fill);
  if (k64) {
    uint64_t u;
    memcpy(&u, &fill, 8);
    fk64 = ((u & 0x7fffffffffffffffull) > 0x7ff0000000000000ull) ? ~0ull
           : ((u >> 63) ? ~u : (u | 0x8000000000000000ull));
  }
  for (int p = 0; p < na; ++p) {
    int ax = axes[p];
    MinmaxPassArgs a;
    a.first_pass = p == 0;
    a.last_pass = p == na - 1;
    a.src_img = in.data;
    a.src_min = smin;
    a.src_max = smax;
    int64_t nn[MAXD], nf[MAXD];
    for (int d = 0; d < MAXD; ++d) {
      // after filtering axis d the buffer covers the output range on d
      bool done = false;
      for (int q = 0; q <= p; ++q) done |= axes[q] == d;
      nn[d] = done ? osz[d] : in.n[d];
      nf[d] = done ? out_first[d] : 0;
    }
    for (int d = 0; d < MAXD; ++d) {
      a.in_n[d] = (int)cur_n[d];
      a.out_n[d] = (int)nn[d];
      a.shift[d] = (int)(nf[d] - cur_first[d]);
      a.img_n[d] = (int)in.n[d];
      a.of[d] = (int)nf[d];
      a.wf[d] = wfirst[d];
      a.wl[d] = wlen[d];
    }
    a.shift[ax] = 0;
    a.axis = ax;
    a.dt = in.dtype;
    a.isfill = isfill;
    a.total = nn[0] * nn[1] * nn[2] * nn[3];
    if (a.last_pass) {
      a.dst_min = omin ? omin->data : nullptr;
      a.dst_max = omax ? omax->data : nullptr;
    } else {
      a.dst_min = omin ? bufmin[p & 1] : nullptr;
      a.dst_max = omax ? bufmax[p & 1] : nullptr;
    }
    int64_t blocks = (a.total + 255) / 256;
    if (blocks > 1 << 20) blocks = 1 << 20;
    if (k64)
      hipLaunchKernelGGL(k_minmax_pass<uint64_t>, dim3((unsigned)blocks), dim3(256), 0, st, a, fk64);
    else
      hipLaunchKernelGGL(k_minmax_pass<uint32_t>, dim3((unsigned)blocks), dim3(256), 0, st, a, fk32);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    smin = a.dst_min;
    smax = a.dst_max;
    for (int d = 0; d < MAXD; ++d) { cur_n[d] = nn[d]; cur_first[d] = nf[d]; }
  }
  return hipSuccess;
}

}  // namespace ifm

namespace ifm {

// ---------------------------------------------------------------------------
// Generic window functions at (possibly strided) window centres:
// mapwindow(f, img, window; border, indices) for f in {sum, mean, minimum,
// maximum, first, median!, Statistics.median, Statistics.quantile(., p)}
// (src/mapwindow.jl:75-121, :270-333; the centres are the affine index map
// of _IndexTransformer, :160-189).  One thread per output; the window is
// enumerated on the fly with the window-relative edge maps of copy_win!
// (window_key), and order statistics use counting selection on isless keys.
// ---------------------------------------------------------------------------
struct WinGenArgs {
  MedianAnyArgs m;          // image, window, border; m.of = first centre, m.o = output extents
  int step[MAXD];           // centre step per axis
  int op;
  double param;
  int *flag;                // set when quantile meets a NaN (the reference throws)
};

// value of an isless key of eltype dt (F64 keys are 64-bit)
template <typename KT>
__device__ __forceinline__ double key_value(KT k, int dt) {
  if constexpr (sizeof(KT) == 8) {
    const uint64_t u = (k == ~0ull) ? 0x7ff8000000000000ull : ((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k);
    return __longlong_as_double((long long)u);
  } else {
    switch (dt) {
    case IFM_U8: return (double)k;
    case IFM_N0F8: return (double)k / 255.0;
    case IFM_F32: {
      const uint32_t u = (k == 0xffffffffu) ? 0x7fc00000u : ((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
      return (double)__uint_as_float(u);
    }
    default: {
      uint16_t h;
      store_key32(&h, 0, dt, k);
      return dt == IFM_BF16 ? (double)__uint_as_float((uint32_t)h << 16)
                            : (double)(float)__builtin_bit_cast(_Float16, h);
    }
    }
  }
}

template <typename KT>
__device__ __forceinline__ bool key_isnan(KT k, int dt) {
  if constexpr (sizeof(KT) == 8) return k == ~0ull;
  else return dt == IFM_F32 ? k == 0xffffffffu : ((dt == IFM_BF16 || dt == IFM_F16) ? k == 0xffffu : false);
}

template <typename KT>
__global__ void __launch_bounds__(256) k_window_generic(WinGenArgs g) {
  const MedianAnyArgs &a = g.m;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < a.total;
       p += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = p, oidx = a.obase;
    int q[MAXD], ilo[MAXD], ilen[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
      const int ql = (int)(r % a.o[d]);
      oidx += ql * a.ostride[d];
      q[d] = a.of[d] + ql * g.step[d];
      r /= a.o[d];
      const int lo = q[d] + a.wf[d], hi = q[d] + a.wf[d] + a.wl[d] - 1;
      ilo[d] = lo < 0 ? 0 : lo;
      const int h2 = hi > a.n[d] - 1 ? a.n[d] - 1 : hi;
      ilen[d] = h2 - ilo[d] + 1;
    }
    const int L = a.L;
    switch (g.op) {
    case IFM_WIN_FIRST: {
      const KT k = window_key<KT>(a, q, ilo, ilen, 0);
      if constexpr (sizeof(KT) == 8) store_key64(a.out, oidx, k);
      else store_key_as(a.out, oidx, a.dt, a.odt, k);
      break;
    }
    case IFM_WIN_MIN:
    case IFM_WIN_MAX: {
      // min / max propagate NaN (Julia's min, max)
      KT best = window_key<KT>(a, q, ilo, ilen, 0);
      bool nan = key_isnan(best, a.dt);
      for (int j = 1; j < L; ++j) {
        const KT k = window_key<KT>(a, q, ilo, ilen, j);
        nan |= key_isnan(k, a.dt);
        best = g.op == IFM_WIN_MIN ? (k < best ? k : best) : (k > best ? k : best);
      }
      if (nan) best = (sizeof(KT) == 8 || a.dt == IFM_F32) ? (KT)~(KT)0 : (KT)0xffffu;  // the NaN key
      if constexpr (sizeof(KT) == 8) store_key64(a.out, oidx, best);
      else store_key_as(a.out, oidx, a.dt, a.odt, best);
      break;
    }
    case IFM_WIN_SUM:
    case IFM_WIN_MEAN: {
      double s = 0.0;
      for (int j = 0; j < L; ++j) s += key_value<KT>(window_key<KT>(a, q, ilo, ilen, j), a.dt);
      store_d(a.out, oidx, a.odt, g.op == IFM_WIN_MEAN ? s / (double)L : s);
      break;
    }
    default: {  // order statistics
      bool nan = false;
      if (g.op != IFM_WIN_MEDIAN_FAST)
        for (int j = 0; j < L; ++j) nan |= key_isnan(window_key<KT>(a, q, ilo, ilen, j), a.dt);
      if (nan && g.op == IFM_WIN_QUANTILE) {
        atomicOr(g.flag, 1);
        store_d(a.out, oidx, a.odt, __longlong_as_double(0x7ff8000000000000ll));
        break;
      }
      if (nan) {  // Statistics.median of a window holding NaN is NaN
        store_d(a.out, oidx, a.odt, __longlong_as_double(0x7ff8000000000000ll));
        break;
      }
      // ranks r1 <= r2 (0-based) and the interpolation weight between them
      int r1, r2;
      double gam = 0.0;
      if (g.op == IFM_WIN_MEDIAN_FAST) {
        r1 = r2 = (L + 1) / 2 - 1;
      } else if (g.op == IFM_WIN_MEDIAN) {
        r1 = (L - 1) / 2;
        r2 = L / 2;
      } else {
        const double h = (double)(L - 1) * g.param;
        r1 = (int)floor(h);
        if (r1 > L - 1) r1 = L - 1;
        if (r1 < 0) r1 = 0;
        r2 = r1 + 1 < L ? r1 + 1 : r1;
        gam = h - (double)r1;
      }
      KT k1 = 0, k2 = 0;
      bool f1 = false, f2 = false;
      for (int i = 0; i < L && !(f1 && f2); ++i) {
        const KT ki = window_key<KT>(a, q, ilo, ilen, i);
        int less = 0, eq = 0;
        for (int j = 0; j < L; ++j) {
          const KT kj = window_key<KT>(a, q, ilo, ilen, j);
          less += kj < ki;
          eq += kj == ki;
        }
        if (!f1 && less <= r1 && r1 < less + eq) {
          k1 = ki;
          f1 = true;
        }
        if (!f2 && less <= r2 && r2 < less + eq) {
          k2 = ki;
          f2 = true;
        }
      }
      if (g.op == IFM_WIN_MEDIAN_FAST || (g.op == IFM_WIN_MEDIAN && r1 == r2)) {
        if constexpr (sizeof(KT) == 8) store_key64(a.out, oidx, k1);
        else store_key_as(a.out, oidx, a.dt, a.odt, k1);
      } else if (g.op == IFM_WIN_MEDIAN) {
        // Statistics.middle(x, y) = x/2 + y/2, in Float32 for Float32 values
        const double x = key_value<KT>(k1, a.dt), y = key_value<KT>(k2, a.dt);
        if (a.odt == IFM_F32 || a.odt == IFM_BF16 || a.odt == IFM_F16)
          store_d(a.out, oidx, a.odt, (double)((float)x / 2.0f + (float)y / 2.0f));
        else store_d(a.out, oidx, a.odt, x / 2.0 + y / 2.0);
      } else {
        const double x = key_value<KT>(k1, a.dt), y = key_value<KT>(k2, a.dt);
        store_d(a.out, oidx, a.odt, x + gam * (y - x));
      }
    }
    }
  }
}

hipError_t launch_window_generic(const Dense4 &in, Dense4 &out, const int64_t *first, const int64_t *step,
                                 const int *wfirst, const int *wlen, int style, double fill, int op, double param,
                                 int *flag, hipStream_t st) {
  WinGenArgs g;
  MedianAnyArgs &a = g.m;
  a.in = in.data;
  a.out = out.data;
  a.L = 1;
  a.total = 1;
  a.obase = 0;
  int64_t stride = 1;
  for (int d = 0; d < MAXD; ++d) {
    a.n[d] = (int)in.n[d];
    a.o[d] = (int)out.n[d];
    a.of[d] = (int)first[d];
    g.step[d] = (int)step[d];
    a.wf[d] = wfirst[d];
    a.wl[d] = wlen[d];
    a.L *= wlen[d];
    a.total *= out.n[d];
    a.ostride[d] = stride;
    stride *= out.n[d];
  }
  if (a.total == 0) return hipSuccess;
  a.style = style;
  a.dt = in.dtype;
  a.odt = out.dtype;
  a.fill = fill;
  g.op = op;
  g.param = param;
  g.flag = flag;
  int64_t blocks = (a.total + 255) / 256;
  if (blocks > 1 << 20) blocks = 1 << 20;
  if (in.dtype == IFM_F64)
    hipLaunchKernelGGL(k_window_generic<uint64_t>, dim3((unsigned)blocks), dim3(256), 0, st, g);
  else
    hipLaunchKernelGGL(k_window_generic<uint32_t>, dim3((unsigned)blocks), dim3(256), 0, st, g);
  return hipGetLastError();
}

}  // namespace ifm

