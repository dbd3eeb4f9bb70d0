// ifm_internal.h — shared host/device definitions of libifm (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include <string.h>
#include <string>
#include <vector>

#include "../../include/ifm.h"

namespace ifm {

constexpr int MAXD = IFM_MAXDIM;
constexpr int MAXTAPS = 128;  // per-axis tap limit of the FIR kernels (longer -> FFT)

// ---------------------------------------------------------------------------
// Border index map (src/border.jl:564-590, periodic form): the source index in
// [0, n) for padded coordinate i, or -1 for Fill.  Reflect needs n >= 2; the
// host rejects n == 1 as the reference does (zero period).
// ---------------------------------------------------------------------------
__host__ __device__ inline int imod(int x, int m) {
  int r = x % m;
  return r < 0 ? r + m : r;
}

__host__ __device__ inline int bmap(int style, int i, int n) {
  if ((unsigned)i < (unsigned)n) return i;
  switch (style) {
  case IFM_REPLICATE: return i < 0 ? 0 : n - 1;
  case IFM_CIRCULAR: return imod(i, n);
  case IFM_SYMMETRIC: { int j = imod(i, 2 * n); return j < n ? j : 2 * n - 1 - j; }
  case IFM_REFLECT: {
    if (n == 1) return 0;  // rejected on the host; keeps the index in bounds
    int j = imod(i, 2 * n - 2);
    return j < n ? j : 2 * n - 2 - j;
  }
  default: return -1;
  }
}

// Tile of workgroup `lin` (0 <= lin < n, the dispatch order) such that each
// of the 8 XCDs, which receive workgroups round-robin (lin % 8), walks one
// contiguous run of tiles: tiles that share cache lines (row neighbours,
// halos) then meet in the same L2 instead of being fetched once per XCD.
__device__ __forceinline__ int xcd_tile(int lin, int n) {
  const int x = lin & 7, k = lin >> 3, per = n >> 3, rem = n & 7;
  return x * per + min(x, rem) + k;
}

// N0f8 code / 255 rounded to nearest (Float32(x::N0f8)): the product with the
// rounded reciprocal corrected by one FMA step, which equals the correctly
// rounded quotient for every code 0..255 (checked exhaustively on the CPU).
__host__ __device__ __forceinline__ float n0f8_to_f32(uint32_t code) {
  const float x = (float)code, r = 1.0f / 255.0f;
  const float q = x * r;
  return fmaf(fmaf(-q, 255.0f, x), r, q);
}

// Element access with a runtime eltype: the branch is uniform over the grid.
__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float load_f(const void *p, int64_t i, int dt) {
  switch (dt) {
  case IFM_U8: return (float)((const uint8_t *)p)[i];
  case IFM_N0F8: return n0f8_to_f32(((const uint8_t *)p)[i]);
  case IFM_F32: return ((const float *)p)[i];
  case IFM_F64: return (float)((const double *)p)[i];
  case IFM_BF16: return bf16_to_f32(((const uint16_t *)p)[i]);
  default: return __half2float(((const __half *)p)[i]);
  }
}

__device__ __forceinline__ double load_d(const void *p, int64_t i, int dt) {
  switch (dt) {
  case IFM_U8: return (double)((const uint8_t *)p)[i];
  case IFM_N0F8: return (double)((const uint8_t *)p)[i] / 255.0;
  case IFM_F32: return (double)((const float *)p)[i];
  case IFM_F64: return ((const double *)p)[i];
  case IFM_BF16: return (double)bf16_to_f32(((const uint16_t *)p)[i]);
  default: return (double)__half2float(((const __half *)p)[i]);
  }
}

__device__ __forceinline__ void store_f(void *p, int64_t i, int dt, float v) {
  switch (dt) {
  case IFM_F32: ((float *)p)[i] = v; break;
  case IFM_F64: ((double *)p)[i] = (double)v; break;
  case IFM_BF16: ((uint16_t *)p)[i] = f32_to_bf16(v); break;
  case IFM_F16: ((__half *)p)[i] = __float2half_rn(v); break;
  default: break;  // integer outputs are produced by the order filters only
  }
}

__device__ __forceinline__ void store_d(void *p, int64_t i, int dt, double v) {
  switch (dt) {
  case IFM_F32: ((float *)p)[i] = (float)v; break;
  case IFM_F64: ((double *)p)[i] = v; break;
  case IFM_BF16: ((uint16_t *)p)[i] = f32_to_bf16((float)v); break;
  case IFM_F16: ((__half *)p)[i] = __float2half_rn((float)v); break;
  default: break;
  }
}

// Order keys: an unsigned integer whose order is Julia's isless on the
// element type (NaN above +Inf, -0.0 below +0.0).  Codes of UInt8 / N0f8 are
// their own keys.  key64 is used for Float64 only.
__device__ __forceinline__ uint32_t key32(const void *p, int64_t i, int dt) {
  switch (dt) {
  case IFM_U8:
  case IFM_N0F8: return ((const uint8_t *)p)[i];
  case IFM_F32: {
    uint32_t u = __float_as_uint(((const float *)p)[i]);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0xffffffffu;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  }
  default: {  // BF16 / F16: 16-bit float layout
    uint32_t u = ((const uint16_t *)p)[i];
    uint32_t expmask = (dt == IFM_BF16) ? 0x7f80u : 0x7c00u;
    if ((u & 0x7fffu) > expmask) return 0xffffu;
    return (u & 0x8000u) ? (~u & 0xffffu) : (u | 0x8000u);
  }
  }
}

__device__ __forceinline__ void store_key32(void *p, int64_t i, int dt, uint32_t k) {
  switch (dt) {
  case IFM_U8:
  case IFM_N0F8: ((uint8_t *)p)[i] = (uint8_t)k; break;
  case IFM_F32: {
    uint32_t u = (k == 0xffffffffu) ? 0x7fc00000u : ((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
    ((uint32_t *)p)[i] = u;
    break;
  }
  default: {
    uint32_t nan = (dt == IFM_BF16) ? 0x7fc0u : 0x7e00u;
    uint32_t u = (k == 0xffffu) ? nan : ((k & 0x8000u) ? (k & 0x7fffu) : (~k & 0xffffu));
    ((uint16_t *)p)[i] = (uint16_t)u;
    break;
  }
  }
}

// An order-statistic key stored in the output eltype odt: the input eltype
// itself, or for UInt8 / N0f8 codes the Float32 / Float64 value (the eltype of
// Statistics.middle of such an element).
__device__ __forceinline__ void store_key_as(void *p, int64_t i, int dt, int odt, uint32_t k) {
  if (odt == dt) {
    store_key32(p, i, dt, k);
  } else if (odt == IFM_F32) {
    ((float *)p)[i] = dt == IFM_N0F8 ? n0f8_to_f32(k) : (float)k;
  } else {
    ((double *)p)[i] = dt == IFM_N0F8 ? (double)k / 255.0 : (double)k;
  }
}

__device__ __forceinline__ uint64_t key64(const void *p, int64_t i) {
  uint64_t u = (uint64_t)__double_as_longlong(((const double *)p)[i]);
  if ((u & 0x7fffffffffffffffull) > 0x7ff0000000000000ull) return ~0ull;
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ void store_key64(void *p, int64_t i, uint64_t k) {
  uint64_t u = (k == ~0ull) ? 0x7ff8000000000000ull : ((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k);
  ((uint64_t *)p)[i] = u;
}

// ---------------------------------------------------------------------------
// Host side: canonical geometry.  Arrays are permuted so that axis 0 is the
// contiguous one; kernels see dense column-major 4-D arrays.
// ---------------------------------------------------------------------------
struct Dense4 {
  void *data = nullptr;
  int dtype = IFM_F32;
  int64_t n[MAXD] = {1, 1, 1, 1};
  int64_t count() const { return n[0] * n[1] * n[2] * n[3]; }
};

inline size_t dtype_size(int dt) {
  switch (dt) {
  case IFM_U8:
  case IFM_N0F8: return 1;
  case IFM_F32: return 4;
  case IFM_F64: return 8;
  default: return 2;
  }
}

// Kernel launch wrappers implemented in the .hip files.  All take canonical
// (memory-order) geometry; `st` is the stream.
struct FirPass {  // one 1-D factor applied along canonical axis `axis`
  int axis;
  int len;
  int first;
  float taps[MAXTAPS];
};

hipError_t launch_pad(const Dense4 &in, Dense4 &out, const int64_t *lo, int style, double fill, hipStream_t st);

// Separable FIR: out region starts at out_first (canonical axes).  `fill` is
// the Fill value already scaled for this pass (v * prod of earlier tap sums).
hipError_t launch_fir1d(const Dense4 &in, const int64_t *in_first, Dense4 &out, const int64_t *out_first,
                        const int64_t *img_n, const FirPass &p, int style, float fill, hipStream_t st);
hipError_t launch_sep2d(const Dense4 &in, Dense4 &out, const int64_t *out_first, const FirPass &p0,
                        const FirPass &p1, int style, float fill, int ty, hipStream_t st);
// Three factors on axes 0, 1, 2 in one pass (input tile <= 4096 elements,
// axis-2 factor odd and <= 9 taps, F32 or F64 input, an (x, y) plane < 4 GiB).
bool sep3d_supported(int k0, int k1, int k2, int dt_in, int64_t plane_elems);
hipError_t launch_sep3d(const Dense4 &in, Dense4 &out, const int64_t *out_first, const FirPass &p0, const FirPass &p1,
                        const FirPass &p2, int style, float fill, hipStream_t st);

struct DenseKernel {
  int ndim;
  int n[MAXD];
  int first[MAXD];
  std::vector<float> taps;  // column-major
};
hipError_t launch_dense_valu(const Dense4 &in, Dense4 &out, const int64_t *out_first, const DenseKernel &k,
                             const float *d_taps, int style, float fill, hipStream_t st);
// Matrix-core banded-Toeplitz path (2-D kernels up to 49 x 32).
struct MfmaArgs {
  const void *in;
  void *out;
  const uint16_t *ttab;  // [chunk][u][hi/lo][lane][8] FP16 Toeplitz fragments
  const float *taps;     // Kx * Ky FP32 taps (direct path of non-finite tiles)
  float kinv;            // 1 / kernel scale
  int n0, n1;            // input plane
  int o0, o1;            // output plane
  int f0, f1;            // out_first along axes 0 and 1
  int kx, ky;
  int first0, first1;    // kernel first offsets
  int ntx, nty, ntiles;  // tiles along x, along y, in total over all planes
  int64_t in_plane, out_plane;
  int o2, n2, f2, f3;    // plane mapping for axes 2 and 3
  int style;
  int dt_in, dt_out;
  float fill;
};
bool dense_mfma_supported(const DenseKernel &k, const Dense4 &in, const Dense4 &out);
void dense_mfma_table(const DenseKernel &k, std::vector<uint16_t> &table, float *kinv);
// Arguments of one image (ntiles = its tiles); false if the shape is unsupported.
bool dense_mfma_args(const Dense4 &in, Dense4 &out, const int64_t *out_first, const DenseKernel &k,
                     const uint16_t *d_table, const float *d_taps, float kinv, int style, float fill, MfmaArgs *a);
// Kernel variant of a kernel shape (equal for requests that can share a launch).
int dense_mfma_variant(const DenseKernel &k);
// Device-built Toeplitz tables (same bits as dense_mfma_table).
struct ToeplitzJob {
  const float *taps;  // device FP32 taps
  uint16_t *table;    // device table
  int kx, ky, kb, urows;
  float ks;
};
size_t dense_mfma_job(const DenseKernel &k, ToeplitzJob *job, float *kinv);  // returns the table bytes
hipError_t launch_toeplitz(const ToeplitzJob *jobs, int njobs, hipStream_t st);
hipError_t launch_dense_mfma(const Dense4 &in, Dense4 &out, const int64_t *out_first, const DenseKernel &k,
                             const uint16_t *d_table, const float *d_taps, float kinv, int style, float fill,
                             hipStream_t st);
hipError_t launch_dense_mfma_batch(int variant, const MfmaArgs *reqs, const int *tile_req, const int *tile_first,
                                   int ntiles, hipStream_t st);

// IIR: one cascade stage along canonical axis `axis`, FP64 recursion.
struct IirCoef {
  double a[3], b[3], scale, M[9];
};
hipError_t launch_iir(const Dense4 &in, Dense4 &out, int axis, const IirCoef &c, int isfill, double fill,
                      hipStream_t st);

// Order filters.
struct MedianArgs {
  const void *in;
  void *out;
  int n0, n1;
  int o0, o1;
  int f0, f1;
  int64_t in_plane, out_plane;
  int nplanes, o2, n2, f2, f3;
  int style;
  int dt;
  int odt;  // output eltype: dt, or F32 / F64 for UInt8 / N0f8 codes
  uint32_t fillkey;
};
hipError_t launch_window_generic(const Dense4 &in, Dense4 &out, const int64_t *first, const int64_t *step,
                                 const int *wfirst, const int *wlen, int style, double fill, int op, double param,
                                 int *flag, hipStream_t st);
hipError_t launch_median(const Dense4 &in, Dense4 &out, const int64_t *out_first, const int *wfirst,
                         const int *wlen, int style, double fill, hipStream_t st);
// The circular edge band of a centred k x k median (after the network kernel).
hipError_t median_circular_band(const Dense4 &in, Dense4 &out, const int64_t *out_first, int k, double fill,
                                hipStream_t st);
struct BandJob {  // one band strip (out indices) of a circular median
  const void *in;
  void *out;
  int64_t in_plane, out_plane;
  int n0, n1, nplanes, opitch;
  int f0, f1;
  int x0, y0, w, h;
  int dt, odt;
};
int median_band_jobs(const Dense4 &in, const Dense4 &out, const int64_t *out_first, int k, BandJob *jobs);
hipError_t launch_median_band_batch(int k, const BandJob *jobs, const int *tile_req, const int *tile_first, int ntiles,
                                    hipStream_t st);
// Separable min/max over truncated windows; with fill, the fill key joins any
// window that leaves the image.  key buffers: two arrays of count() uint32/64.
hipError_t launch_minmax(const Dense4 &in, Dense4 *omin, Dense4 *omax, const int64_t *out_first,
                         const int *wfirst, const int *wlen, int isfill, double fill, void *work,
                         hipStream_t st);
size_t minmax_workspace(const Dense4 &in, const int64_t *osize);

// FFT correlation over the whole padded array (up to 3 transformed axes; axes
// beyond nd are batch axes).  Lines longer than FFT_LINE_MAX are transformed
// in two levels.  The workspace holds two complex arrays of the padded size,
// the kernel spectrum (product of the transformed N), the twiddle tables and
// the kernel taps.
constexpr int FFT_LINE_MAX = 8192;
size_t fft_workspace(const int64_t *N, int nd);
// A padded axis length the whole-array transform can take (2^a 3^b, split in
// two levels beyond FFT_LINE_MAX).
bool fft_axis_supported(int64_t n);
// have_kernel: the workspace still holds this kernel's spectrum and twiddles
// from the previous call with the same kernel and sizes.
hipError_t launch_fft_filter(const Dense4 &in, Dense4 &out, const int64_t *out_first, const DenseKernel &k,
                             const int64_t *Npad, int nd, const int64_t *lo, int style, float fill, void *work,
                             bool have_kernel, hipStream_t st);
// Tiled (overlap-save) FFT correlation for 1-D / 2-D kernels of up to 80 taps
// per axis: F x F tiles, F = fft_tile_size(k) (0: the kernel does not fit).
int fft_tile_size(const DenseKernel &k);
size_t fft_tile_spec_bytes(int F);
void fft_tile_twiddles(int F, std::vector<float2> &tw32, std::vector<double2> &tw64);
hipError_t launch_fft_tile_spectrum(const float *d_taps, int K0, int K1, int F, const double2 *tw64, float2 *spec,
                                    hipStream_t st);
// nanflag / gen: a non-finite input value (or Fill value) sets *nanflag = gen
// and the whole output becomes NaN, as the whole-array transform gives.
hipError_t launch_fft_tile(const Dense4 &in, Dense4 &out, const int64_t *out_first, const DenseKernel &k, int F,
                           const float2 *spec, const float2 *tw32, int *nanflag, int gen, int style, float fill,
                           hipStream_t st);

// Elementwise passes (k_elem.hip): NaN mask and zero-filled copy (hasna set
// when a NaN is met), num /= den, and a dense src copied into dst at dst_first.
hipError_t launch_na_prepare(const Dense4 &in, Dense4 &tmp, Dense4 &valid, int *hasna, hipStream_t st);
hipError_t launch_divide(Dense4 &num, const Dense4 &den, hipStream_t st);
hipError_t launch_copy_into(const Dense4 &src, Dense4 &dst, const int64_t *dst_first, hipStream_t st);
// Floating-point values into an integer eltype (bits 8/16/32/64), *inexact set
// for a non-integer or out-of-range value.
hipError_t launch_to_int(const Dense4 &src, void *dst, int bits, int is_signed, double scale, int *inexact,
                         hipStream_t st);
// Local extrema flags (one byte per element, dense in canonical order).
hipError_t launch_local_extrema(const Dense4 &in, uint8_t *flag, const int *half, const int *edges, int sense,
                                hipStream_t st);

// Batched serving.  Each kernel takes an array of per-request descriptors in
// device memory, the request of every tile (tile_req) and the first tile of
// every request (tile_first); requests are 2-D images filtered over their
// whole extent.
struct SepReq {
  const void *in;
  void *out;
  const float *taps0, *taps1;  // device FP32 taps
  int n0, n1;
  int k0, k1, first0, first1;
  int style, dt_in, dt_out;
  float fill;
  int ty;  // tile height (S2_TY or S2_TY2)
};
hipError_t launch_sep_batch(const SepReq *reqs, const int *tile_req, const int *tile_first, int ntiles,
                            size_t lds_bytes, hipStream_t st);
void median_batch_tile(int k, bool pack, int *bx, int *by);
uint32_t median_fill_key(int dt, double fill);
hipError_t launch_median_batch(int k, bool pack, const MedianArgs *reqs, const int *tile_req, const int *tile_first,
                               int ntiles, hipStream_t st);
hipError_t launch_fill_uniform(void *data, int dtype, int64_t count, uint64_t seed, hipStream_t st);

}  // namespace ifm

// The context behind the opaque ifm_ctx of the C ABI.
struct ifm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<std::pair<void *, size_t>> slots;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // Device copies of dense kernels keyed by a hash of their FP32 taps, so a
  // kernel used call after call is uploaded once (no per-call copy or sync).
  struct TapEntry {
    uint64_t hash;
    size_t n;
    void *dev;                // FP32 taps
    void *table = nullptr;    // MFMA Toeplitz fragments (built on first MFMA use)
    float kinv = 1.f;
    int kx = 0, ky = 0, kz = 0;
    std::vector<float> host;  // the taps, compared exactly on lookup
  };
  std::vector<TapEntry> taps;
  // Batched runs: a ring of descriptor staging slots (pinned host + device),
  // each reusable once its event has completed.
  struct BatchSlot {
    void *host = nullptr;
    void *dev = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
  };
  std::vector<BatchSlot> bslots;
  int bnext = 0;
  std::vector<ifm_tile_rule> tile_rules;  // ifm_set_tile_rules
  // FFT workspace (its own slot) and the kernel spectrum it holds: valid when
  // fft_have is set and the kernel, its offsets and the padded sizes equal
  // the stored ones exactly.
  void *fft_work = nullptr;
  size_t fft_bytes = 0;
  bool fft_have = false;
  std::vector<float> fft_taps;
  int fft_kn[IFM_MAXDIM] = {0, 0, 0, 0}, fft_kf[IFM_MAXDIM] = {0, 0, 0, 0};
  int64_t fft_N[IFM_MAXDIM] = {0, 0, 0, 0};
  int fft_nd = 0;
  // Tiled FFT: per tile size F, the twiddle tables (FP32 and FP64), and a
  // small cache of kernel spectra keyed by the exact taps, shape and F.
  struct TileTw {
    int F = 0;
    void *tw32 = nullptr;
    void *tw64 = nullptr;
  };
  std::vector<TileTw> tile_tw;
  struct TileSpec {
    int F = 0, k0 = 0, k1 = 0;
    std::vector<float> taps;
    void *spec = nullptr;
  };
  std::vector<TileSpec> tile_spec;
  void *tile_taps = nullptr;  // device staging of the taps being transformed
  size_t tile_taps_bytes = 0;
  int *tile_flag = nullptr;   // non-finite flag of the tile kernel
  int tile_gen = 0;
  // device address of the no-alloc scratch array on this ctx's device
  void *noalloc_scratch = nullptr;
  int *finite_flag = nullptr;  // ifm_all_finite's result word
};

namespace ifm {
void batch_release(ifm_ctx *ctx);
// Tile height of the fused two-factor kernel for a call (the first matching
// tile rule, else S2_TY; tiles whose LDS would not fit fall back to S2_TY).
int sep2d_tile_h(const ifm_ctx *ctx, int dt_in, int k0, int k1, int64_t pixels);

}  // namespace ifm
