// k_fft_tile.hip — tiled FFT correlation (gfx950) for the Algorithm.FFT()
// choice of imfilter (src/imfilter.jl:796-866) with kernels of up to 80 taps
// per axis on 1-D / 2-D planes.
//
// The reference pads the whole image to nextprod([2,3], n + K) and computes
// irfft(rfft(Apad) .* conj(rfft(kpad))).  Because the padded length covers
// the kernel's reach, that circular correlation equals the linear correlation
// of the border-extended image.  Here the same linear correlation is computed
// by overlap-save over F x F tiles (F = 64 or 128): a tile reads the
// border-mapped input block ext[origin + first + u], u in [0, F)^2, and its
// circular correlation with the kernel is exact for the T = F - K + 1 outputs
// per axis that do not wrap.  Each workgroup packs two tiles into one complex
// block (tile A real, tile B imaginary: the kernel is real, so the filter acts
// on both parts independently), transforms it, multiplies by conj of the
// kernel's F x F spectrum (computed once per kernel, on the device, in FP64)
// and transforms back.  The whole transform stays in LDS and registers: per
// tile, HBM sees one read of the input block and one write of the outputs.
//
// Each 1-D transform of length F = R1 * R2 is a two-stage decomposition
// (n = a + R2 p, k = k1 + R1 k2): radix-R1 DFTs over p in registers, twiddle
// W_F^(a k1), an LDS exchange, radix-R2 DFTs over a.  The inverse runs the
// stages in the opposite order, so the forward transform's last stage and the
// inverse's first stage work on the same registers with the kernel product in
// between.  Tasks are laid out line-fastest across lanes and the LDS pitch is
// F + 1 complex values, which keeps every exchange free of bank conflicts; the
// digit of the other stage (a or k1) is then wave-uniform, so twiddles are
// scalar loads.
#include <math.h>

#include "ifm_internal.h"

namespace ifm {

// cos and sin of 2 pi m / 16
__device__ __forceinline__ constexpr float cos16(int m) {
  return m == 0 ? 1.0f
       : m == 1 ? 0.92387953251128675613f
       : m == 2 ? 0.70710678118654752440f
       : m == 3 ? 0.38268343236508977173f
       : m == 4 ? 0.0f
       : m == 5 ? -0.38268343236508977173f
       : m == 6 ? -0.70710678118654752440f
       : -0.92387953251128675613f;  // m == 7
}
__device__ __forceinline__ constexpr float sin16(int m) { return cos16(m < 4 ? 4 - m : m - 4) * (m <= 8 ? 1.f : -1.f); }

__host__ __device__ constexpr int lg2c(int n) { return n <= 1 ? 0 : 1 + lg2c(n / 2); }
__host__ __device__ constexpr int brevc(int x, int bits) {
  int r = 0;
  for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1) << (bits - 1 - i);
  return r;
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// b * exp(SIGN 2 pi i m / 16) for a compile-time m in [0, 8).
template <int SIGN>
__device__ __forceinline__ float2 tw16(float2 b, int m) {
  if (m == 0) return b;
  if (m == 4) return SIGN < 0 ? make_float2(b.y, -b.x) : make_float2(-b.y, b.x);
  const float c = cos16(m), s = SIGN * sin16(m);
  if (m == 2 || m == 6) {  // |c| == |s|
    const float h = c;     // sqrt(1/2) with the sign of cos
    return SIGN < 0 ? (m == 2 ? make_float2((b.x + b.y) * h, (b.y - b.x) * h)
                              : make_float2((b.x - b.y) * h, (b.x + b.y) * h))
                    : (m == 2 ? make_float2((b.x - b.y) * h, (b.x + b.y) * h)
                              : make_float2((b.x + b.y) * h, (b.y - b.x) * h));
  }
  return make_float2(b.x * c - b.y * s, b.x * s + b.y * c);
}

// In-register DFT of N (2..16) points v[0], v[S], ..., natural order in and
// out: X[k] = sum_n x[n] exp(SIGN 2 pi i n k / N).  Radix-2, decimation in
// time; every index is a compile-time constant after unrolling.
template <int N, int SIGN, int S>
__device__ __forceinline__ void dft(float2 *v) {
  constexpr int L = lg2c(N);
  float2 t[N];
#pragma unroll
  for (int i = 0; i < N; ++i) t[i] = v[S * brevc(i, L)];
#pragma unroll
  for (int len = 2; len <= N; len <<= 1) {
    const int half = len >> 1;
#pragma unroll
    for (int s = 0; s < N; s += len) {
#pragma unroll
      for (int j = 0; j < half; ++j) {
        const float2 w = tw16<SIGN>(t[s + j + half], j * (16 / len));
        const float2 a = t[s + j];
        t[s + j] = cadd(a, w);
        t[s + j + half] = csub(a, w);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) v[S * i] = t[i];
}

struct FftTileArgs {
  const void *in;
  void *out;
  const float2 *kspec;  // conj(K^)/F^2 at [k1][k2][k0] (k = k1 + R1 k2 along axis 1)
  const float2 *tw;     // W_F^m = exp(-2 pi i m / F), m < F
  int *nanflag;         // set to gen when a non-finite value enters a tile
  int gen;
  int n0, n1;           // input plane
  int o0, o1;           // output plane
  int of0, of1;         // output region origin in the image (axes 0, 1)
  int kf0, kf1;         // kernel first offsets (axes 0, 1)
  int T0, T1;           // outputs per tile along axes 0, 1
  int ntx, nty;
  int64_t ntiles;       // over all planes
  int64_t in_plane, out_plane;
  int o2, n2, of2, of3; // plane mapping of axes 2 and 3
  int style, dt_in, dt_out;
  float fill;
};

template <int F>
struct TileCfg;
template <>
struct TileCfg<64> {
  static constexpr int R1 = 8, R2 = 8, NT = 256;
};
template <>
struct TileCfg<128> {
  static constexpr int R1 = 16, R2 = 8, NT = 1024;
};

template <int F>
constexpr size_t tile_lds_bytes() {
  return sizeof(float2) * (size_t)F * (F + 1);
}

template <int F>
__global__ void __launch_bounds__(TileCfg<F>::NT) k_fft_tile(FftTileArgs a) {
  constexpr int R1 = TileCfg<F>::R1, R2 = TileCfg<F>::R2, NT = TileCfg<F>::NT;
  constexpr int P = F + 1;           // LDS line pitch (complex values)
  constexpr int V = F * F / NT;      // complex values per thread
  constexpr int T1N = F * R2 / NT;   // radix-R1 tasks per thread
  constexpr int T2N = F * R1 / NT;   // radix-R2 tasks per thread
  static_assert(T1N * R1 == V && T2N * R2 == V, "task split");
  static_assert(NT % F == 0 || F % 64 == 0, "line-fastest lanes");
  extern __shared__ float2 lds[];
  const int t = threadIdx.x;
  float2 v[V];

  // ---- the two tiles of this workgroup ----
  const int64_t per_plane = (int64_t)a.ntx * a.nty;
  int64_t tile[2] = {2 * (int64_t)blockIdx.x, 2 * (int64_t)blockIdx.x + 1};
  const bool hasB = tile[1] < a.ntiles;
  int64_t ioff[2], ooff[2];
  int x0[2], y0[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int64_t ti = s == 1 && !hasB ? tile[0] : tile[s];
    const int64_t pl = ti / per_plane;
    const int64_t r = ti - pl * per_plane;
    const int ty = (int)(r / a.ntx), tx = (int)(r - (int64_t)ty * a.ntx);
    x0[s] = tx * a.T0;
    y0[s] = ty * a.T1;
    const int c2 = (int)(pl % a.o2), c3 = (int)(pl / a.o2);
    ioff[s] = ((int64_t)(c2 + a.of2) + (int64_t)a.n2 * (c3 + a.of3)) * a.in_plane;
    ooff[s] = pl * a.out_plane;
  }

  // ---- load both blocks (coalesced along axis 0) into LDS as (A, B) pairs ----
  bool bad = false;
  for (int idx = t; idx < F * F; idx += NT) {
    const int u0 = idx % F, u1 = idx / F;
    float val[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int c = bmap(a.style, a.of0 + x0[s] + a.kf0 + u0, a.n0);
      const int r = bmap(a.style, a.of1 + y0[s] + a.kf1 + u1, a.n1);
      val[s] = (c | r) < 0 ? a.fill : load_f(a.in, ioff[s] + c + (int64_t)a.n0 * r, a.dt_in);
      bad |= !(val[s] - val[s] == 0.f);
    }
    lds[u1 * P + u0] = make_float2(val[0], hasB ? val[1] : 0.f);
  }
  if (bad) atomicExch(a.nanflag, a.gen);
  __syncthreads();

  // Task coordinates: tau = t + j NT, line = tau % F, digit = tau / F (wave-uniform).
  // ---- axis 0, forward: radix-R1 over p (u0 = a + R2 p) ----
#pragma unroll
  for (int j = 0; j < T1N; ++j) {
    const int tau = t + j * NT, ln = tau % F, d = __builtin_amdgcn_readfirstlane(tau / F);
#pragma unroll
    for (int p = 0; p < R1; ++p) v[j * R1 + p] = lds[ln * P + d + R2 * p];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < T1N; ++j) {
    const int tau = t + j * NT, ln = tau % F, d = __builtin_amdgcn_readfirstlane(tau / F);
    dft<R1, -1, 1>(&v[j * R1]);
#pragma unroll
    for (int k1 = 1; k1 < R1; ++k1) v[j * R1 + k1] = cmul(v[j * R1 + k1], a.tw[d * k1]);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) lds[ln * P + k1 * R2 + d] = v[j * R1 + k1];
  }
  __syncthreads();
  // ---- axis 0, forward: radix-R2 over a; then to axis-1 lines (line k0, position u1) ----
#pragma unroll
  for (int j = 0; j < T2N; ++j) {
    const int tau = t + j * NT, ln = tau % F, k1 = __builtin_amdgcn_readfirstlane(tau / F);
#pragma unroll
    for (int q = 0; q < R2; ++q) v[j * R2 + q] = lds[ln * P + k1 * R2 + q];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < T2N; ++j) {
    const int tau = t + j * NT, ln = tau % F, k1 = __builtin_amdgcn_readfirstlane(tau / F);
    dft<R2, -1, 1>(&v[j * R2]);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) lds[(k1 + R1 * k2) * P + ln] = v[j * R2 + k2];
  }
  __syncthreads();
  // ---- axis 1, forward: radix-R1 over q (u1 = b + R2 q) ----
#pragma unroll
  for (int j = 0; j < T1N; ++j) {
    const int tau = t + j * NT, ln = tau % F, d = __builtin_amdgcn_readfirstlane(tau / F);
#pragma unroll
    for (int p = 0; p < R1; ++p) v[j * R1 + p] = lds[ln * P + d + R2 * p];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < T1N; ++j) {
    const int tau = t + j * NT, ln = tau % F, d = __builtin_amdgcn_readfirstlane(tau / F);
    dft<R1, -1, 1>(&v[j * R1]);
#pragma unroll
    for (int k1 = 1; k1 < R1; ++k1) v[j * R1 + k1] = cmul(v[j * R1 + k1], a.tw[d * k1]);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) lds[ln * P + k1 * R2 + d] = v[j * R1 + k1];
  }
  __syncthreads();
  // ---- axis 1: radix-R2 forward, kernel product, radix-R2 inverse ----
#pragma unroll
  for (int j = 0; j < T2N; ++j) {
    const int tau = t + j * NT, ln = tau % F, k1 = __builtin_amdgcn_readfirstlane(tau / F);
#pragma unroll
    for (int q = 0; q < R2; ++q) v[j * R2 + q] = lds[ln * P + k1 * R2 + q];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < T2N; ++j) {
    const int tau = t + j * NT, ln = tau % F, k1 = __builtin_amdgcn_readfirstlane(tau / F);
    dft<R2, -1, 1>(&v[j * R2]);
    const float2 *ks = a.kspec + (int64_t)k1 * R2 * F + ln;
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[j * R2 + k2] = cmul(v[j * R2 + k2], ks[k2 * F]);
    dft<R2, +1, 1>(&v[j * R2]);
#pragma unroll
    for (int b = 1; b < R2; ++b) v[j * R2 + b] = cmulc(v[j * R2 + b], a.tw[b * k1]);
#pragma unroll
    for (int b = 0; b < R2; ++b) lds[ln * P + k1 * R2 + b] = v[j * R2 + b];
  }
  __syncthreads();
  // ---- axis 1, inverse: radix-R1 over k1; back to axis-0 lines (line u1, position k0) ----
#pragma unroll
  for (int j = 0; j < T1N; ++j) {
    const int tau = t + j * NT, ln = tau % F, d = __builtin_amdgcn_readfirstlane(tau / F);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) v[j * R1 + k1] = lds[ln * P + k1 * R2 + d];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < T1N; ++j) {
    const int tau = t + j * NT, ln = tau % F, d = __builtin_amdgcn_readfirstlane(tau / F);
    dft<R1, +1, 1>(&v[j * R1]);
    // v[j R1 + q] = y[k0 = ln][u1 = d + R2 q]
#pragma unroll
    for (int q = 0; q < R1; ++q) lds[(d + R2 * q) * P + ln] = v[j * R1 + q];
  }
  __syncthreads();
  // ---- axis 0, inverse: radix-R2 over k2 (k0 = k1 + R1 k2) ----
#pragma unroll
  for (int j = 0; j < T2N; ++j) {
    const int tau = t + j * NT, ln = tau % F, k1 = __builtin_amdgcn_readfirstlane(tau / F);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[j * R2 + k2] = lds[ln * P + k1 + R1 * k2];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < T2N; ++j) {
    const int tau = t + j * NT, ln = tau % F, k1 = __builtin_amdgcn_readfirstlane(tau / F);
    dft<R2, +1, 1>(&v[j * R2]);
#pragma unroll
    for (int b = 1; b < R2; ++b) v[j * R2 + b] = cmulc(v[j * R2 + b], a.tw[b * k1]);
#pragma unroll
    for (int b = 0; b < R2; ++b) lds[ln * P + k1 * R2 + b] = v[j * R2 + b];
  }
  __syncthreads();
  // ---- axis 0, inverse: radix-R1 over k1 -> u0 = a + R2 p ----
#pragma unroll
  for (int j = 0; j < T1N; ++j) {
    const int tau = t + j * NT, ln = tau % F, d = __builtin_amdgcn_readfirstlane(tau / F);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) v[j * R1 + k1] = lds[ln * P + k1 * R2 + d];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < T1N; ++j) {
    const int tau = t + j * NT, ln = tau % F, d = __builtin_amdgcn_readfirstlane(tau / F);
    dft<R1, +1, 1>(&v[j * R1]);
#pragma unroll
    for (int p = 0; p < R1; ++p) lds[ln * P + d + R2 * p] = v[j * R1 + p];
  }
  __syncthreads();
  // ---- store the valid outputs of both tiles (coalesced along axis 0) ----
  for (int idx = t; idx < F * F; idx += NT) {
    const int u0 = idx % F, u1 = idx / F;
    if (u0 >= a.T0 || u1 >= a.T1) continue;
    const float2 r = lds[u1 * P + u0];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (s == 1 && !hasB) break;
      const int c0 = x0[s] + u0, c1 = y0[s] + u1;
      if (c0 < a.o0 && c1 < a.o1) store_f(a.out, ooff[s] + c0 + (int64_t)a.o0 * c1, a.dt_out, s == 0 ? r.x : r.y);
    }
  }
}

// Kernel spectrum: conj(K^[k0][k1]) / F^2 with K^ = sum_j k[j] W_F^(j0 k0 + j1 k1)
// over tap positions j relative to the kernel's first tap, in FP64, stored at
// [k1 % R1][k1 / R1][k0] (the order of the tile kernel's product stage).
struct FftSpecArgs {
  const float *taps;    // K0 x K1, axis 0 fastest
  const double2 *tw64;  // exp(-2 pi i m / F), m < F
  float2 *spec;
  int K0, K1, F, R1, R2;
};

__global__ void __launch_bounds__(256) k_fft_tile_spectrum(FftSpecArgs a) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= a.F * a.F) return;
  const int k0 = id % a.F, k1 = id / a.F;
  double re = 0.0, im = 0.0;
  for (int j1 = 0; j1 < a.K1; ++j1) {
    const int m1 = (j1 * k1) % a.F;
    int m = m1;
    for (int j0 = 0; j0 < a.K0; ++j0) {
      const double kv = (double)a.taps[j0 + a.K0 * j1];
      const double2 w = a.tw64[m];
      // conj(W) for conj(K^)
      re += kv * w.x;
      im -= kv * w.y;
      m += k0;
      if (m >= a.F) m -= a.F;
    }
  }
  const double sc = 1.0 / ((double)a.F * a.F);
  const int k1l = k1 % a.R1, k2 = k1 / a.R1;
  a.spec[((int64_t)k1l * a.R2 + k2) * a.F + k0] = make_float2((float)(re * sc), (float)(im * sc));
}

__global__ void __launch_bounds__(256) k_fft_nanfill(void *out, int64_t plane, int64_t nplanes, int dt,
                                                     const int *flag, int gen) {
  if (*flag != gen) return;
  const float nan = __int_as_float(0x7fc00000);
  const int64_t total = plane * nplanes;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
    store_f(out, i, dt, nan);
}

int fft_tile_size(const DenseKernel &k) {
  const bool twod = k.n[2] == 1 && k.n[3] == 1 && k.first[2] == 0 && k.first[3] == 0;
  if (!twod) return 0;
  const int km = k.n[0] > k.n[1] ? k.n[0] : k.n[1];
  if (km <= 24) return 64;
  if (km <= 80) return 128;
  return 0;
}

size_t fft_tile_spec_bytes(int F) { return sizeof(float2) * (size_t)F * F; }

void fft_tile_twiddles(int F, std::vector<float2> &tw32, std::vector<double2> &tw64) {
  tw32.resize(F);
  tw64.resize(F);
  for (int m = 0; m < F; ++m) {
    const double ang = -2.0 * M_PI * (double)m / (double)F;
    tw64[m] = make_double2(cos(ang), sin(ang));
    tw32[m] = make_float2((float)tw64[m].x, (float)tw64[m].y);
  }
}

hipError_t launch_fft_tile_spectrum(const float *d_taps, int K0, int K1, int F, const double2 *tw64, float2 *spec,
                                    hipStream_t st) {
  FftSpecArgs s;
  s.taps = d_taps;
  s.tw64 = tw64;
  s.spec = spec;
  s.K0 = K0;
  s.K1 = K1;
  s.F = F;
  s.R1 = F == 128 ? TileCfg<128>::R1 : TileCfg<64>::R1;
  s.R2 = F == 128 ? TileCfg<128>::R2 : TileCfg<64>::R2;
  hipLaunchKernelGGL(k_fft_tile_spectrum, dim3((F * F + 255) / 256), dim3(256), 0, st, s);
  return hipGetLastError();
}

hipError_t launch_fft_tile(const Dense4 &in, Dense4 &out, const int64_t *out_first, const DenseKernel &k, int F,
                           const float2 *spec, const float2 *tw32, int *nanflag, int gen, int style, float fill,
                           hipStream_t st) {
  FftTileArgs a;
  a.in = in.data;
  a.out = out.data;
  a.kspec = spec;
  a.tw = tw32;
  a.nanflag = nanflag;
  a.gen = gen;
  a.n0 = (int)in.n[0];
  a.n1 = (int)in.n[1];
  a.o0 = (int)out.n[0];
  a.o1 = (int)out.n[1];
  a.of0 = (int)out_first[0];
  a.of1 = (int)out_first[1];
  a.kf0 = k.first[0];
  a.kf1 = k.first[1];
  a.T0 = F - k.n[0] + 1;
  a.T1 = F - k.n[1] + 1;
  a.ntx = (a.o0 + a.T0 - 1) / a.T0;
  a.nty = (a.o1 + a.T1 - 1) / a.T1;
  const int64_t nplanes = out.n[2] * out.n[3];
  a.ntiles = (int64_t)a.ntx * a.nty * nplanes;
  a.in_plane = in.n[0] * in.n[1];
  a.out_plane = out.n[0] * out.n[1];
  a.o2 = (int)out.n[2];
  a.n2 = (int)in.n[2];
  a.of2 = (int)out_first[2];
  a.of3 = (int)out_first[3];
  a.style = style;
  a.dt_in = in.dtype;
  a.dt_out = out.dtype;
  a.fill = fill;
  if (a.ntiles == 0) return hipSuccess;
  const int64_t pairs = (a.ntiles + 1) / 2;
  if (pairs > 0x7fffffff) return hipErrorInvalidValue;
  if (F == 64)
    hipLaunchKernelGGL(k_fft_tile<64>, dim3((unsigned)pairs), dim3(TileCfg<64>::NT), tile_lds_bytes<64>(), st, a);
  else
    hipLaunchKernelGGL(k_fft_tile<128>, dim3((unsigned)pairs), dim3(TileCfg<128>::NT), tile_lds_bytes<128>(), st,
                       a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int64_t total = out.count();
  const int64_t blocks = std::min<int64_t>((total + 255) / 256, 1024);
  hipLaunchKernelGGL(k_fft_nanfill, dim3((unsigned)blocks), dim3(256), 0, st, out.data, out.n[0] * out.n[1], nplanes,
                     out.dtype, (const int *)nanflag, gen);
  return hipGetLastError();
}

}  // namespace ifm
