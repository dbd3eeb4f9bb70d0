// ifm_api.hip — the C ABI of libifm: argument validation (raising the
// reference's ArgumentError / DimensionMismatch cases as status codes), layout
// canonicalisation, host-array staging and kernel dispatch.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "ifm_internal.h"
#include "k_tiles.h"

using namespace ifm;

namespace {

enum Slot { S_IN = 0, S_OUT, S_OUT2, S_TMP1, S_TMP2, S_TAPS, S_WORK, S_COUNT };

int fail(ifm_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail(ctx, IFM_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

void *slot_buffer(ifm_ctx *ctx, int slot, size_t bytes) {
  if (ctx->slots.size() < S_COUNT) ctx->slots.resize(S_COUNT, {nullptr, 0});
  auto &s = ctx->slots[slot];
  if (bytes == 0) bytes = 16;
  if (s.second < bytes) {
    if (s.first) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipFree(s.first);
      s.first = nullptr;
      s.second = 0;
    }
    size_t want = bytes + bytes / 8;
    if (hipMalloc(&s.first, want) != hipSuccess) {
      s.first = nullptr;
      return nullptr;
    }
    s.second = want;
  }
  return s.first;
}

// Canonical geometry: perm[m] = user axis stored at canonical position m.
struct Geo {
  int nd;          // user ndim
  int perm[MAXD];  // -1 = padding axis
  int inv[MAXD];   // user axis -> canonical position
};

bool valid_dtype(int dt) { return dt >= IFM_U8 && dt <= IFM_F16; }

// Finds the axis permutation that makes `a` dense column-major.
bool layout_of(const ifm_array &a, Geo &g) {
  g.nd = a.ndim;
  std::vector<int> big, unit;
  for (int d = 0; d < a.ndim; ++d) (a.size[d] == 1 ? unit : big).push_back(d);
  std::stable_sort(big.begin(), big.end(), [&](int x, int y) { return a.stride[x] < a.stride[y]; });
  int m = 0;
  for (int d : big) g.perm[m++] = d;
  for (int d : unit) g.perm[m++] = d;
  for (; m < MAXD; ++m) g.perm[m] = -1;
  int64_t expect = 1;
  for (int d : big) {
    if (a.stride[d] != expect) return false;
    expect *= a.size[d];
  }
  for (int d = 0; d < MAXD; ++d) g.inv[d] = -1;
  for (int k = 0; k < MAXD; ++k)
    if (g.perm[k] >= 0) g.inv[g.perm[k]] = k;
  return true;
}

bool dense_under(const ifm_array &a, const Geo &g) {
  int64_t expect = 1;
  for (int m = 0; m < MAXD; ++m) {
    int d = g.perm[m];
    if (d < 0) continue;
    if (a.size[d] != 1 && a.stride[d] != expect) return false;
    expect *= a.size[d];
  }
  return true;
}

int check_array(ifm_ctx *ctx, const ifm_array *a, const char *name) {
  if (!a) return fail(ctx, IFM_E_ARG, std::string(name) + " is NULL");
  if (a->ndim < 1 || a->ndim > MAXD)
    return fail(ctx, IFM_E_UNSUPPORTED, std::string(name) + ": ndim must be 1..4");
  if (!valid_dtype(a->dtype)) return fail(ctx, IFM_E_ARG, std::string(name) + ": unknown eltype");
  for (int d = 0; d < a->ndim; ++d) {
    if (a->size[d] < 0) return fail(ctx, IFM_E_ARG, std::string(name) + ": negative size");
    if (a->size[d] > (int64_t)INT32_MAX / 2) return fail(ctx, IFM_E_UNSUPPORTED, std::string(name) + ": axis too long");
  }
  return IFM_OK;
}

int64_t count_of(const ifm_array &a) {
  int64_t c = 1;
  for (int d = 0; d < a.ndim; ++d) c *= a.size[d];
  return c;
}

// A device view of a user array in canonical order; host arrays are staged.
struct View {
  Dense4 d;
  const ifm_array *user = nullptr;
  bool staged = false;
  size_t bytes = 0;
};

int make_view(ifm_ctx *ctx, const ifm_array *a, const Geo &g, int slot, bool upload, View &v) {
  v.user = a;
  v.d.dtype = a->dtype;
  for (int m = 0; m < MAXD; ++m) v.d.n[m] = g.perm[m] >= 0 && g.perm[m] < a->ndim ? a->size[g.perm[m]] : 1;
  v.bytes = (size_t)count_of(*a) * dtype_size(a->dtype);
  if (a->device) {
    v.d.data = a->data;
    v.staged = false;
    return IFM_OK;
  }
  void *p = slot_buffer(ctx, slot, v.bytes);
  if (!p) return fail(ctx, IFM_E_NOMEM, "device allocation failed (staging)");
  if (upload && v.bytes) HIPCHK(hipMemcpyAsync(p, a->data, v.bytes, hipMemcpyHostToDevice, ctx->stream));
  v.d.data = p;
  v.staged = true;
  return IFM_OK;
}

int finish_view(ifm_ctx *ctx, View &v) {
  if (v.staged && v.bytes) {
    HIPCHK(hipMemcpyAsync(v.user->data, v.d.data, v.bytes, hipMemcpyDeviceToHost, ctx->stream));
  }
  return IFM_OK;
}

int sync_if_host(ifm_ctx *ctx, bool any_host) {
  HIPCHK(hipGetLastError());
  if (any_host) HIPCHK(hipStreamSynchronize(ctx->stream));
  return IFM_OK;
}

// Canonical per-axis values from user-axis arrays (NULL -> default).
void to_canon(const Geo &g, const int64_t *user, int64_t dflt, int64_t *out) {
  for (int m = 0; m < MAXD; ++m) {
    int d = g.perm[m];
    out[m] = (d >= 0 && user) ? user[d] : dflt;
  }
}

int check_region(ifm_ctx *ctx, const Dense4 &in, const Dense4 &out, const int64_t *of) {
  for (int m = 0; m < MAXD; ++m) {
    if (of[m] < 0 || of[m] + out.n[m] > in.n[m])
      return fail(ctx, IFM_E_DIM, "output indices lie outside the image axes");
  }
  return IFM_OK;
}

int common_setup(ifm_ctx *ctx, const ifm_array *in, const ifm_array *out, Geo &g) {
  int rc;
  if ((rc = check_array(ctx, in, "in"))) return rc;
  if ((rc = check_array(ctx, out, "out"))) return rc;
  if (in->ndim != out->ndim) return fail(ctx, IFM_E_DIM, "in and out must have the same number of dimensions");
  if (!layout_of(*in, g)) return fail(ctx, IFM_E_UNSUPPORTED, "in is not a dense array (copy it first)");
  if (!dense_under(*out, g)) return fail(ctx, IFM_E_UNSUPPORTED, "out must be dense with the same axis order as in");
  return IFM_OK;
}

int check_border(ifm_ctx *ctx, const ifm_border *b) {
  if (!b) return fail(ctx, IFM_E_ARG, "border is NULL");
  if (b->style < IFM_REPLICATE || b->style > IFM_FILL) return fail(ctx, IFM_E_ARG, "border style unrecognized");
  return IFM_OK;
}

// Reflect has period 2n-2; the reference fails (DivideError) for n == 1.
int check_reflect(ifm_ctx *ctx, const ifm_border *b, const Dense4 &in, const bool *padded_axis) {
  if (b->style != IFM_REFLECT) return IFM_OK;
  for (int m = 0; m < MAXD; ++m)
    if (padded_axis[m] && in.n[m] == 1)
      return fail(ctx, IFM_E_ARG, "reflect padding of an axis of length 1 (zero period)");
  return IFM_OK;
}

bool is_float_out(int dt) { return dt == IFM_F32 || dt == IFM_F64 || dt == IFM_BF16 || dt == IFM_F16; }

}  // namespace

// Scratch of ifm_imfilter_fir_noalloc: the cascade's FP32 temporaries (two
// halves, ping-pong).  Each device holds its own copy for the life of the
// process.  In host code `scratch` is only a symbol handle: kernels and copies
// use the device address that hipGetSymbolAddress returns for the current
// device (noalloc_scratch_of).
static __device__ __align__(256) unsigned char scratch[IFM_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
constexpr int SCRATCH_MAXDEV = 64;
void *g_scratch_ptr[SCRATCH_MAXDEV];
bool g_scratch_done[SCRATCH_MAXDEV];

// The device address of `scratch` on `device` (current), looked up once per
// device; nullptr if the lookup failed.
void *noalloc_scratch_of(int device) {
  if (device < 0 || device >= SCRATCH_MAXDEV) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_done[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
    g_scratch_done[device] = true;
    if (e != hipSuccess) (void)hipGetLastError();  // the failure is reported by the entry point
  }
  return g_scratch_ptr[device];
}

}  // namespace

namespace ifm {

int sep2d_tile_h(const ifm_ctx *ctx, int dt_in, int k0, int k1, int64_t pixels) {
  const int kmax = std::max(k0, k1);
  for (const ifm_tile_rule &r : ctx->tile_rules) {
    if (r.op != IFM_TILE_SEP2D || (r.dtype >= 0 && r.dtype != dt_in) || kmax > r.kmax || pixels > r.max_pixels)
      continue;
    if (r.tile_h == S2_TY) return S2_TY;
    static int max_lds = -1;
    if (max_lds < 0) {
      int dev = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        max_lds = 64 * 1024;
    }
    return sizeof(float) * sep2d_lds_floats(k0, k1, S2_TY2) <= (size_t)max_lds ? S2_TY2 : S2_TY;
  }
  return S2_TY;
}

}  // namespace ifm

// all(isfinite, a) for filter_algorithm (src/imfilter.jl:1200-1209): one
// streaming read, a flag set by any lane that meets an Inf or NaN.
__global__ void __launch_bounds__(256) k_all_finite(const void *p, int64_t count, int dt, int vec, int *flag) {
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  bool bad = false;
  if (dt == IFM_F32 && !vec) {
    for (int64_t i = tid; i < count; i += stride)
      bad |= (((const uint32_t *)p)[i] & 0x7f800000u) == 0x7f800000u;
  } else if (dt == IFM_F32) {  // 16-byte aligned: four values per load
    const uint4 *q = (const uint4 *)p;
    const int64_t n4 = count / 4;
    for (int64_t i = tid; i < n4; i += stride) {
      const uint4 v = q[i];
      bad |= ((v.x & 0x7f800000u) == 0x7f800000u) | ((v.y & 0x7f800000u) == 0x7f800000u) |
             ((v.z & 0x7f800000u) == 0x7f800000u) | ((v.w & 0x7f800000u) == 0x7f800000u);
    }
    for (int64_t i = 4 * n4 + tid; i < count; i += stride)
      bad |= (((const uint32_t *)p)[i] & 0x7f800000u) == 0x7f800000u;
  } else if (dt == IFM_F64) {
    for (int64_t i = tid; i < count; i += stride)
      bad |= (((const uint64_t *)p)[i] & 0x7ff0000000000000ull) == 0x7ff0000000000000ull;
  } else {  // BF16 / F16
    const uint32_t m = dt == IFM_BF16 ? 0x7f80u : 0x7c00u;
    for (int64_t i = tid; i < count; i += stride) bad |= (((const uint16_t *)p)[i] & m) == m;
  }
  if (bad) atomicOr(flag, 1);
}

// ---------------------------------------------------------------------------
extern "C" {

const char *ifm_version(void) { return "ifm 0.1 gfx950"; }

int ifm_device_count(int *count) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  if (count) *count = c;
  return IFM_OK;
}

int ifm_ctx_create(int device, void *stream, ifm_ctx **out) {
  if (!out) return IFM_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return IFM_E_DEVICE;
  if (device < 0 || device >= n) return IFM_E_ARG;
  if (hipSetDevice(device) != hipSuccess) return IFM_E_DEVICE;
  ifm_ctx *ctx = new ifm_ctx();
  ctx->device = device;
  ctx->stream = (hipStream_t)stream;
  ctx->noalloc_scratch = noalloc_scratch_of(device);
  ctx->slots.resize(S_COUNT, {nullptr, 0});
  *out = ctx;
  return IFM_OK;
}

int ifm_all_finite(ifm_ctx *ctx, const ifm_array *a, int *result) {
  if (!ctx) return IFM_E_ARG;
  int rc;
  if ((rc = check_array(ctx, a, "a"))) return rc;
  if (!result) return fail(ctx, IFM_E_ARG, "result is NULL");
  if (!a->device) return fail(ctx, IFM_E_ARG, "ifm_all_finite takes a device array");
  *result = 1;
  const int dt = a->dtype;
  if (!(dt == IFM_F32 || dt == IFM_F64 || dt == IFM_BF16 || dt == IFM_F16)) return IFM_OK;  // integers and N0f8
  Geo g;
  if (!layout_of(*a, g)) return fail(ctx, IFM_E_UNSUPPORTED, "a is not a dense array");
  const int64_t count = count_of(*a);
  if (count == 0) return IFM_OK;
  if (!ctx->finite_flag && hipMalloc(&ctx->finite_flag, sizeof(int)) != hipSuccess) {
    ctx->finite_flag = nullptr;
    return fail(ctx, IFM_E_NOMEM, "device allocation failed (finite flag)");
  }
  HIPCHK(hipMemsetAsync(ctx->finite_flag, 0, sizeof(int), ctx->stream));
  const bool vec = dt == IFM_F32 && ((uintptr_t)a->data & 15) == 0;
  const int64_t units = vec ? (count + 3) / 4 : count;
  const int64_t blocks = std::min<int64_t>((units + 255) / 256, 4096);
  hipLaunchKernelGGL(k_all_finite, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, a->data, count, dt,
                     vec ? 1 : 0, ctx->finite_flag);
  HIPCHK(hipGetLastError());
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, ctx->finite_flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *result = h ? 0 : 1;
  return IFM_OK;
}

// Shapes of two arrays agree axis by axis and they share one dense layout.
static int same_geometry(ifm_ctx *ctx, const ifm_array *a, const ifm_array *b, const Geo &g) {
  if (a->ndim != b->ndim) return fail(ctx, IFM_E_DIM, "arrays differ in dimensions");
  for (int d = 0; d < a->ndim; ++d)
    if (a->size[d] != b->size[d]) return fail(ctx, IFM_E_DIM, "arrays differ in size");
  if (!dense_under(*b, g)) return fail(ctx, IFM_E_UNSUPPORTED, "arrays must share one dense layout");
  return IFM_OK;
}

int ifm_na_prepare(ifm_ctx *ctx, const ifm_array *in, ifm_array *tmp, ifm_array *valid, int *hasna) {
  if (!ctx) return IFM_E_ARG;
  int rc;
  if ((rc = check_array(ctx, in, "in")) || (rc = check_array(ctx, tmp, "tmp")) || (rc = check_array(ctx, valid, "valid")))
    return rc;
  if (!hasna) return fail(ctx, IFM_E_ARG, "hasna is NULL");
  if (!is_float_out(tmp->dtype) || !is_float_out(valid->dtype))
    return fail(ctx, IFM_E_ARG, "tmp and valid must be floating point");
  Geo g;
  if (!layout_of(*in, g)) return fail(ctx, IFM_E_UNSUPPORTED, "in is not a dense array");
  if ((rc = same_geometry(ctx, in, tmp, g)) || (rc = same_geometry(ctx, in, valid, g))) return rc;
  View vi, vt, vv;
  if ((rc = make_view(ctx, in, g, S_IN, true, vi))) return rc;
  if ((rc = make_view(ctx, tmp, g, S_OUT, false, vt))) return rc;
  if ((rc = make_view(ctx, valid, g, S_OUT2, false, vv))) return rc;
  if (!ctx->finite_flag && hipMalloc(&ctx->finite_flag, sizeof(int)) != hipSuccess) {
    ctx->finite_flag = nullptr;
    return fail(ctx, IFM_E_NOMEM, "device allocation failed (flag)");
  }
  HIPCHK(hipMemsetAsync(ctx->finite_flag, 0, sizeof(int), ctx->stream));
  HIPCHK(launch_na_prepare(vi.d, vt.d, vv.d, ctx->finite_flag, ctx->stream));
  if ((rc = finish_view(ctx, vt)) || (rc = finish_view(ctx, vv))) return rc;
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, ctx->finite_flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *hasna = h ? 1 : 0;
  return IFM_OK;
}

int ifm_divide(ifm_ctx *ctx, ifm_array *num, const ifm_array *den) {
  if (!ctx) return IFM_E_ARG;
  int rc;
  if ((rc = check_array(ctx, num, "num")) || (rc = check_array(ctx, den, "den"))) return rc;
  if (!is_float_out(num->dtype)) return fail(ctx, IFM_E_ARG, "num must be floating point");
  Geo g;
  if (!layout_of(*num, g)) return fail(ctx, IFM_E_UNSUPPORTED, "num is not a dense array");
  if ((rc = same_geometry(ctx, num, den, g))) return rc;
  View vn, vd;
  if ((rc = make_view(ctx, num, g, S_OUT, true, vn))) return rc;
  if ((rc = make_view(ctx, den, g, S_IN, true, vd))) return rc;
  HIPCHK(launch_divide(vn.d, vd.d, ctx->stream));
  if ((rc = finish_view(ctx, vn))) return rc;
  return sync_if_host(ctx, !num->device || !den->device);
}

int ifm_copy_into(ifm_ctx *ctx, const ifm_array *src, ifm_array *dst, const int64_t *dst_first) {
  if (!ctx) return IFM_E_ARG;
  int rc;
  if ((rc = check_array(ctx, src, "src")) || (rc = check_array(ctx, dst, "dst"))) return rc;
  if (src->ndim != dst->ndim) return fail(ctx, IFM_E_DIM, "src and dst differ in dimensions");
  Geo g;
  if (!layout_of(*dst, g)) return fail(ctx, IFM_E_UNSUPPORTED, "dst is not a dense array");
  if (!dense_under(*src, g)) return fail(ctx, IFM_E_UNSUPPORTED, "src must share the layout of dst");
  for (int d = 0; d < src->ndim; ++d) {
    const int64_t f = dst_first ? dst_first[d] : 0;
    if (f < 0 || f + src->size[d] > dst->size[d]) return fail(ctx, IFM_E_DIM, "block lies outside dst");
  }
  View vs, vd;
  if ((rc = make_view(ctx, src, g, S_IN, true, vs))) return rc;
  if ((rc = make_view(ctx, dst, g, S_OUT, true, vd))) return rc;
  int64_t df[MAXD];
  to_canon(g, dst_first, 0, df);
  HIPCHK(launch_copy_into(vs.d, vd.d, df, ctx->stream));
  if ((rc = finish_view(ctx, vd))) return rc;
  return sync_if_host(ctx, !src->device || !dst->device);
}

int ifm_convert_int(ifm_ctx *ctx, const ifm_array *src, void *dst, int dst_device, int bits, int is_signed,
                    double scale, int *inexact) {
  if (!ctx) return IFM_E_ARG;
  int rc;
  if ((rc = check_array(ctx, src, "src"))) return rc;
  if (!dst || !inexact) return fail(ctx, IFM_E_ARG, "dst and inexact must not be NULL");
  if (bits != 8 && bits != 16 && bits != 32 && bits != 64) return fail(ctx, IFM_E_ARG, "bits must be 8, 16, 32 or 64");
  Geo g;
  if (!layout_of(*src, g)) return fail(ctx, IFM_E_UNSUPPORTED, "src is not a dense array");
  View vs;
  if ((rc = make_view(ctx, src, g, S_IN, true, vs))) return rc;
  const size_t bytes = (size_t)vs.d.count() * (size_t)(bits / 8);
  void *d = dst;
  if (!dst_device) {
    d = slot_buffer(ctx, S_OUT, bytes);
    if (!d) return fail(ctx, IFM_E_NOMEM, "device allocation failed (staging)");
  }
  if (!ctx->finite_flag && hipMalloc(&ctx->finite_flag, sizeof(int)) != hipSuccess) {
    ctx->finite_flag = nullptr;
    return fail(ctx, IFM_E_NOMEM, "device allocation failed (flag)");
  }
  HIPCHK(hipMemsetAsync(ctx->finite_flag, 0, sizeof(int), ctx->stream));
  HIPCHK(launch_to_int(vs.d, d, bits, is_signed, scale, ctx->finite_flag, ctx->stream));
  if (!dst_device && bytes) HIPCHK(hipMemcpyAsync(dst, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, ctx->finite_flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *inexact = h ? 1 : 0;
  return IFM_OK;
}

int ifm_local_extrema(ifm_ctx *ctx, const ifm_array *in, ifm_array *flags, const int64_t *window,
                      const int32_t *edges, int sense) {
  if (!ctx) return IFM_E_ARG;
  int rc;
  if ((rc = check_array(ctx, in, "in")) || (rc = check_array(ctx, flags, "flags"))) return rc;
  if (flags->dtype != IFM_U8) return fail(ctx, IFM_E_ARG, "flags must be UInt8");
  if (sense != 1 && sense != -1) return fail(ctx, IFM_E_ARG, "sense must be +1 or -1");
  Geo g;
  if (!layout_of(*in, g)) return fail(ctx, IFM_E_UNSUPPORTED, "in is not a dense array");
  if ((rc = same_geometry(ctx, in, flags, g))) return rc;
  int half[MAXD] = {0, 0, 0, 0}, edg[MAXD] = {1, 1, 1, 1};
  for (int m = 0; m < MAXD; ++m) {
    const int d = g.perm[m];
    if (d < 0 || d >= in->ndim) continue;
    const int64_t w = window ? window[d] : 3;
    if (w < 1 || w % 2 != 1 || w > 63) return fail(ctx, IFM_E_ARG, "window entries must be odd and at most 63");
    half[m] = (int)(w >> 1);
    edg[m] = edges ? (edges[d] != 0) : 1;
  }
  View vi, vf;
  if ((rc = make_view(ctx, in, g, S_IN, true, vi))) return rc;
  if ((rc = make_view(ctx, flags, g, S_OUT, false, vf))) return rc;
  HIPCHK(launch_local_extrema(vi.d, (uint8_t *)vf.d.data, half, edg, sense, ctx->stream));
  if ((rc = finish_view(ctx, vf))) return rc;
  return sync_if_host(ctx, !in->device || !flags->device);
}

int ifm_ctx_destroy(ifm_ctx *ctx) {
  if (!ctx) return IFM_OK;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &s : ctx->slots)
    if (s.first) (void)hipFree(s.first);
  if (ctx->fft_work) (void)hipFree(ctx->fft_work);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  for (auto &e : ctx->taps) {
    (void)hipFree(e.dev);
    if (e.table) (void)hipFree(e.table);
  }
  for (auto &e : ctx->tile_tw) {
    (void)hipFree(e.tw32);
    (void)hipFree(e.tw64);
  }
  for (auto &e : ctx->tile_spec) (void)hipFree(e.spec);
  if (ctx->tile_taps) (void)hipFree(ctx->tile_taps);
  if (ctx->tile_flag) (void)hipFree(ctx->tile_flag);
  if (ctx->finite_flag) (void)hipFree(ctx->finite_flag);
  batch_release(ctx);
  delete ctx;
  return IFM_OK;
}

int ifm_ctx_set_stream(ifm_ctx *ctx, void *stream) {
  if (!ctx) return IFM_E_ARG;
  // cached kernel spectra and tap tables were written on the previous stream
  if (ctx->stream != (hipStream_t)stream) HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->stream = (hipStream_t)stream;
  ctx->fft_have = false;
  return IFM_OK;
}

int ifm_ctx_synchronize(ifm_ctx *ctx) {
  if (!ctx) return IFM_E_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return IFM_OK;
}

const char *ifm_last_error(const ifm_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int ifm_device_alloc(ifm_ctx *ctx, uint64_t bytes, void **ptr) {
  if (!ctx || !ptr) return IFM_E_ARG;
  *ptr = nullptr;
  if (hipMalloc(ptr, bytes ? bytes : 16) != hipSuccess) return fail(ctx, IFM_E_NOMEM, "hipMalloc failed");
  return IFM_OK;
}

int ifm_device_free(ifm_ctx *ctx, void *ptr) {
  if (!ctx) return IFM_E_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (ptr) HIPCHK(hipFree(ptr));
  return IFM_OK;
}

int ifm_memcpy_h2d(ifm_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  if (!ctx) return IFM_E_ARG;
  if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return IFM_OK;
}

int ifm_memcpy_d2h(ifm_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  if (!ctx) return IFM_E_ARG;
  if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return IFM_OK;
}

int ifm_memset_d(ifm_ctx *ctx, void *dst, int value, uint64_t bytes) {
  if (!ctx) return IFM_E_ARG;
  if (bytes) HIPCHK(hipMemsetAsync(dst, value, bytes, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return IFM_OK;
}

int ifm_set_tile_rules(ifm_ctx *ctx, const ifm_tile_rule *rules, int n) {
  if (!ctx) return IFM_E_ARG;
  if (n < 0 || n > IFM_MAX_TILE_RULES || (n > 0 && !rules)) return fail(ctx, IFM_E_ARG, "bad tile rule count");
  for (int i = 0; i < n; ++i) {
    const ifm_tile_rule &r = rules[i];
    if (r.op != IFM_TILE_SEP2D) return fail(ctx, IFM_E_ARG, "tile rule: unknown op");
    if (r.tile_h != S2_TY && r.tile_h != S2_TY2) return fail(ctx, IFM_E_ARG, "tile rule: tile_h must be 32 or 64");
    if (r.dtype < -1 || r.dtype > IFM_F16) return fail(ctx, IFM_E_ARG, "tile rule: bad dtype");
  }
  ctx->tile_rules.assign(rules, rules + n);
  return IFM_OK;
}

int ifm_tile_height(ifm_ctx *ctx, int op, int dtype, int kmax, int64_t pixels, int *tile_h) {
  if (!ctx || !tile_h) return IFM_E_ARG;
  if (op != IFM_TILE_SEP2D) return fail(ctx, IFM_E_ARG, "unknown tile op");
  *tile_h = sep2d_tile_h(ctx, dtype, kmax, kmax, pixels);
  return IFM_OK;
}

int ifm_timer_start(ifm_ctx *ctx) {
  if (!ctx) return IFM_E_ARG;
  if (!ctx->ev0) HIPCHK(hipEventCreate(&ctx->ev0));
  if (!ctx->ev1) HIPCHK(hipEventCreate(&ctx->ev1));
  HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
  return IFM_OK;
}

int ifm_timer_stop(ifm_ctx *ctx, float *ms) {
  if (!ctx || !ctx->ev0) return IFM_E_ARG;
  HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
  HIPCHK(hipEventSynchronize(ctx->ev1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
  if (ms) *ms = t;
  return IFM_OK;
}

// ---- padarray --------------------------------------------------------------
int ifm_padarray(ifm_ctx *ctx, const ifm_array *in, ifm_array *out, const ifm_border *b) {
  if (!ctx) return IFM_E_ARG;
  Geo g;
  int rc;
  if ((rc = common_setup(ctx, in, out, g))) return rc;
  if ((rc = check_border(ctx, b))) return rc;
  for (int d = 0; d < in->ndim; ++d) {
    if (b->lo[d] < 0 || b->hi[d] < 0) return fail(ctx, IFM_E_ARG, "negative padding");
    if (out->size[d] != b->lo[d] + in->size[d] + b->hi[d])
      return fail(ctx, IFM_E_DIM, "axes(dst) == axes(src) must hold");
    if (b->style == IFM_REFLECT && in->size[d] == 1 && (b->lo[d] || b->hi[d]))
      return fail(ctx, IFM_E_ARG, "reflect padding of an axis of length 1 (zero period)");
    if (b->style != IFM_FILL && in->size[d] == 0 && (b->lo[d] || b->hi[d]))
      return fail(ctx, IFM_E_ARG, "cannot pad an empty axis");
  }
  if (in->dtype != out->dtype && !is_float_out(out->dtype))
    return fail(ctx, IFM_E_UNSUPPORTED, "padarray: conversion to an integer eltype");
  View vi, vo;
  if ((rc = make_view(ctx, in, g, S_IN, true, vi))) return rc;
  if ((rc = make_view(ctx, out, g, S_OUT, false, vo))) return rc;
  int64_t lo[MAXD];
  to_canon(g, b->lo, 0, lo);
  HIPCHK(launch_pad(vi.d, vo.d, lo, b->style, b->fill, ctx->stream));
  if ((rc = finish_view(ctx, vo))) return rc;
  return sync_if_host(ctx, !in->device || !out->device);
}

// ---- FIR, separable ------------------------------------------------------------
// The factor cascade.  fixed_tmp == nullptr: temporaries come from the ctx's
// slots (allocated on demand); otherwise from fixed_tmp, two halves of
// fixed_cap / 2 bytes each, and nothing is allocated.
static int fir_run(ifm_ctx *ctx, const ifm_array *in, ifm_array *out, const int64_t *out_first,
                   const ifm_factor *factors, int nfactors, const ifm_border *b, void *fixed_tmp, size_t fixed_cap) {
  Geo g;
  int rc;
  if ((rc = common_setup(ctx, in, out, g))) return rc;
  if ((rc = check_border(ctx, b))) return rc;
  if (!is_float_out(out->dtype)) return fail(ctx, IFM_E_UNSUPPORTED, "imfilter: output eltype must be floating point");
  if (nfactors < 0 || (nfactors > 0 && !factors)) return fail(ctx, IFM_E_ARG, "bad factor list");
  // Merge the factors per axis: consecutive correlations compose to the
  // correlation with their convolution (first offsets add).
  std::vector<double> taps[MAXD];
  int64_t first[MAXD] = {0, 0, 0, 0};
  bool has[MAXD] = {false, false, false, false};
  for (int f = 0; f < nfactors; ++f) {
    const ifm_factor &F = factors[f];
    if (F.axis < 0 || F.axis >= in->ndim) return fail(ctx, IFM_E_DIM, "factor axis beyond the image dimensions");
    if (F.len < 1 || !F.taps) return fail(ctx, IFM_E_ARG, "empty factor");
    if (F.len == 1 && F.first == 0 && F.taps[0] == 1.0) continue;  // iscopy
    int m = g.inv[F.axis];
    if (!has[m]) {
      taps[m].assign(F.taps, F.taps + F.len);
      first[m] = F.first;
      has[m] = true;
    } else {
      std::vector<double> c(taps[m].size() + F.len - 1, 0.0);
      for (size_t i = 0; i < taps[m].size(); ++i)
        for (int j = 0; j < F.len; ++j) c[i + j] += taps[m][i] * F.taps[j];
      taps[m].swap(c);
      first[m] += F.first;
    }
  }
  for (int m = 0; m < MAXD; ++m)
    if (has[m] && taps[m].size() > (size_t)MAXTAPS)
      return fail(ctx, IFM_E_UNSUPPORTED, "factor longer than 128 taps: use Algorithm.FFT()");
  View vi, vo;
  if ((rc = make_view(ctx, in, g, S_IN, true, vi))) return rc;
  if ((rc = make_view(ctx, out, g, S_OUT, false, vo))) return rc;
  int64_t of[MAXD];
  to_canon(g, out_first, 0, of);
  if ((rc = check_region(ctx, vi.d, vo.d, of))) return rc;
  if ((rc = check_reflect(ctx, b, vi.d, has))) return rc;
  for (int m = 0; m < MAXD; ++m)
    if (has[m] && vi.d.n[m] == 0 && b->style != IFM_FILL) return fail(ctx, IFM_E_ARG, "cannot pad an empty axis");

  std::vector<int> axes;
  for (int m = 0; m < MAXD; ++m)
    if (has[m]) axes.push_back(m);
  auto pass_of = [&](int m) {
    FirPass p;
    p.axis = m;
    p.len = (int)taps[m].size();
    p.first = (int)first[m];
    for (int t = 0; t < p.len; ++t) p.taps[t] = (float)taps[m][t];
    return p;
  };
  auto tapsum = [&](int m) {
    double s = 0;
    for (double t : taps[m]) s += t;
    return s;
  };
  const int style = b->style;
  double fill = b->fill;
  hipStream_t st = ctx->stream;
  if (vo.d.count() > 0) {
    if (axes.empty()) {
      FirPass p;
      p.axis = 0;
      p.len = 1;
      p.first = 0;
      p.taps[0] = 1.f;
      int64_t zero[MAXD] = {0, 0, 0, 0};
      HIPCHK(launch_fir1d(vi.d, zero, vo.d, of, vi.d.n, p, style, (float)fill, st));
    } else if (axes.size() == 3 && axes[0] == 0 && axes[1] == 1 && axes[2] == 2 &&
               sep3d_supported((int)taps[0].size(), (int)taps[1].size(), (int)taps[2].size(), vi.d.dtype,
                               vi.d.n[0] * vi.d.n[1])) {
      // all three factors in one pass, no temporaries
      HIPCHK(launch_sep3d(vi.d, vo.d, of, pass_of(0), pass_of(1), pass_of(2), style, (float)fill, st));
    } else {
      // Groups: a fused (0,1) pair when both are present and short enough,
      // then single passes for the remaining axes.
      std::vector<std::vector<int>> groups;
      size_t i = 0;
      if (axes.size() >= 2 && axes[0] == 0 && axes[1] == 1 && taps[0].size() <= 64 && taps[1].size() <= 64) {
        groups.push_back({0, 1});
        i = 2;
      }
      for (; i < axes.size(); ++i) groups.push_back({axes[i]});
      // Source of the current group: covers out region on filtered axes and the
      // whole image on the others.
      Dense4 src = vi.d;
      int64_t src_first[MAXD] = {0, 0, 0, 0};
      bool done[MAXD] = {false, false, false, false};
      for (size_t gi = 0; gi < groups.size(); ++gi) {
        const bool last = gi + 1 == groups.size();
        for (int m : groups[gi]) done[m] = true;
        Dense4 dst;
        int64_t dst_first[MAXD];
        if (last) {
          dst = vo.d;
          for (int m = 0; m < MAXD; ++m) dst_first[m] = of[m];
        } else {
          dst.dtype = IFM_F32;
          for (int m = 0; m < MAXD; ++m) {
            bool later = false;
            for (size_t gj = gi + 1; gj < groups.size(); ++gj)
              for (int x : groups[gj]) later |= x == m;
            dst.n[m] = later ? vi.d.n[m] : vo.d.n[m];
            dst_first[m] = later ? 0 : of[m];
          }
          const size_t need = (size_t)dst.count() * sizeof(float);
          if (fixed_tmp) {
            if (need > fixed_cap / 2)
              return fail(ctx, IFM_E_BUDGET,
                          "cascade temporary of " + std::to_string(need) + " bytes exceeds the no-alloc maximum of " +
                              std::to_string(fixed_cap / 2) + " bytes");
            dst.data = (char *)fixed_tmp + ((gi & 1) ? fixed_cap / 2 : 0);
          } else {
            dst.data = slot_buffer(ctx, (gi & 1) ? S_TMP2 : S_TMP1, need);
          }
          if (!dst.data) return fail(ctx, IFM_E_NOMEM, "device allocation failed (temporary)");
        }
        if (groups[gi].size() == 2) {
          // sep2d reads planes p2+f2, p3+f3 of its source; the source is the
          // input here (first group), so f2/f3 are the destination offsets.
          const int ty = sep2d_tile_h(ctx, src.dtype, pass_of(0).len, pass_of(1).len, dst.n[0] * dst.n[1]);
          HIPCHK(launch_sep2d(src, dst, dst_first, pass_of(0), pass_of(1), style, (float)fill, ty, st));
          fill *= tapsum(0) * tapsum(1);
        } else {
          int m = groups[gi][0];
          HIPCHK(launch_fir1d(src, src_first, dst, dst_first, vi.d.n, pass_of(m), style, (float)fill, st));
          fill *= tapsum(m);
        }
        src = dst;
        for (int m = 0; m < MAXD; ++m) src_first[m] = dst_first[m];
      }
    }
  }
  if ((rc = finish_view(ctx, vo))) return rc;
  return sync_if_host(ctx, !in->device || !out->device);
}

int ifm_imfilter_fir(ifm_ctx *ctx, const ifm_array *in, ifm_array *out, const int64_t *out_first,
                     const ifm_factor *factors, int nfactors, const ifm_border *b) {
  if (!ctx) return IFM_E_ARG;
  return fir_run(ctx, in, out, out_first, factors, nfactors, b, nullptr, 0);
}

int ifm_imfilter_fir_noalloc(ifm_ctx *ctx, const ifm_array *in, ifm_array *out, const int64_t *out_first,
                             const ifm_factor *factors, int nfactors, const ifm_border *b) {
  if (!ctx) return IFM_E_ARG;
  if (!in || !out || !in->device || !out->device)
    return fail(ctx, IFM_E_ARG, "ifm_imfilter_fir_noalloc takes device arrays only");
  if (!ctx->noalloc_scratch)
    return fail(ctx, IFM_E_DEVICE, "the no-alloc scratch array has no device address on this device");
  return fir_run(ctx, in, out, out_first, factors, nfactors, b, ctx->noalloc_scratch, IFM_NOALLOC_SCRATCH_BYTES);
}

// ---- FIR, dense --------------------------------------------------------------------
static int device_taps(ifm_ctx *ctx, const DenseKernel &dk, bool want_table, ifm_ctx::TapEntry **out) {
  const std::vector<float> &t = dk.taps;
  uint64_t h = 1469598103934665603ull;  // FNV-1a over the tap bytes and the kernel shape
  const unsigned char *p = (const unsigned char *)t.data();
  for (size_t i = 0; i < t.size() * sizeof(float); ++i) h = (h ^ p[i]) * 1099511628211ull;
  for (int d = 0; d < MAXD; ++d) h = (h ^ (uint64_t)dk.n[d]) * 1099511628211ull;
  ifm_ctx::TapEntry *hit = nullptr;
  for (auto &e : ctx->taps)
    if (e.hash == h && e.n == t.size() && e.kx == dk.n[0] && e.ky == dk.n[1] && e.kz == dk.n[2] * 1024 + dk.n[3] && e.host == t)
      hit = &e;
  if (!hit) {
    if (ctx->taps.size() >= 64) {  // bounded cache: drop the oldest entry
      HIPCHK(hipStreamSynchronize(ctx->stream));
      HIPCHK(hipFree(ctx->taps.front().dev));
      if (ctx->taps.front().table) HIPCHK(hipFree(ctx->taps.front().table));
      ctx->taps.erase(ctx->taps.begin());
    }
    void *d = nullptr;
    if (hipMalloc(&d, t.size() * sizeof(float)) != hipSuccess)
      return fail(ctx, IFM_E_NOMEM, "device allocation failed (taps)");
    HIPCHK(hipMemcpy(d, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
    ifm_ctx::TapEntry e;
    e.hash = h;
    e.n = t.size();
    e.dev = d;
    e.host = t;
    e.kx = dk.n[0];
    e.ky = dk.n[1];
    e.kz = dk.n[2] * 1024 + dk.n[3];  // axes 2 and 3 (each at most 1024 taps)
    ctx->taps.push_back(e);
    hit = &ctx->taps.back();
  }
  if (want_table && !hit->table) {
    std::vector<uint16_t> tab;
    dense_mfma_table(dk, tab, &hit->kinv);
    if (hipMalloc(&hit->table, tab.size() * sizeof(uint16_t)) != hipSuccess) {
      hit->table = nullptr;
      return fail(ctx, IFM_E_NOMEM, "device allocation failed (Toeplitz table)");
    }
    HIPCHK(hipMemcpy(hit->table, tab.data(), tab.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  }
  *out = hit;
  return IFM_OK;
}

static int dense_kernel_canon(ifm_ctx *ctx, const ifm_kernel *k, const Geo &g, int img_ndim, DenseKernel &dk) {
  if (!k || !k->taps) return fail(ctx, IFM_E_ARG, "kernel is NULL");
  if (k->ndim < 0 || k->ndim > img_ndim) return fail(ctx, IFM_E_DIM, "kernel has more dimensions than the image");
  int64_t ks[MAXD], kf[MAXD];
  for (int d = 0; d < MAXD; ++d) {
    ks[d] = d < k->ndim ? k->size[d] : 1;
    kf[d] = d < k->ndim ? k->first[d] : 0;
    if (ks[d] < 1) return fail(ctx, IFM_E_ARG, "empty kernel");
    if (ks[d] > 1024) return fail(ctx, IFM_E_UNSUPPORTED, "kernel axis longer than 1024");
  }
  dk.ndim = MAXD;
  for (int m = 0; m < MAXD; ++m) {
    int d = g.perm[m];
    dk.n[m] = d >= 0 ? (int)ks[d] : 1;
    dk.first[m] = d >= 0 ? (int)kf[d] : 0;
  }
  int64_t total = (int64_t)dk.n[0] * dk.n[1] * dk.n[2] * dk.n[3];
  dk.taps.resize(total);
  // user strides of the column-major kernel
  int64_t us[MAXD], acc = 1;
  for (int d = 0; d < MAXD; ++d) {
    us[d] = acc;
    acc *= ks[d];
  }
  int64_t q = 0;
  for (int i3 = 0; i3 < dk.n[3]; ++i3)
    for (int i2 = 0; i2 < dk.n[2]; ++i2)
      for (int i1 = 0; i1 < dk.n[1]; ++i1)
        for (int i0 = 0; i0 < dk.n[0]; ++i0, ++q) {
          int idx[MAXD] = {i0, i1, i2, i3};
          int64_t src = 0;
          for (int m = 0; m < MAXD; ++m)
            if (g.perm[m] >= 0) src += idx[m] * us[g.perm[m]];
          dk.taps[q] = (float)k->taps[src];
        }
  return IFM_OK;
}

int ifm_imfilter_dense(ifm_ctx *ctx, const ifm_array *in, ifm_array *out, const int64_t *out_first,
                       const ifm_kernel *kernel, const ifm_border *b, int precision) {
  if (!ctx) return IFM_E_ARG;
  Geo g;
  int rc;
  if ((rc = common_setup(ctx, in, out, g))) return rc;
  if ((rc = check_border(ctx, b))) return rc;
  if (!is_float_out(out->dtype)) return fail(ctx, IFM_E_UNSUPPORTED, "imfilter: output eltype must be floating point");
  DenseKernel dk;
  if ((rc = dense_kernel_canon(ctx, kernel, g, in->ndim, dk))) return rc;
  View vi, vo;
  if ((rc = make_view(ctx, in, g, S_IN, true, vi))) return rc;
  if ((rc = make_view(ctx, out, g, S_OUT, false, vo))) return rc;
  int64_t of[MAXD];
  to_canon(g, out_first, 0, of);
  if ((rc = check_region(ctx, vi.d, vo.d, of))) return rc;
  bool padded[MAXD];
  for (int m = 0; m < MAXD; ++m) padded[m] = dk.n[m] > 1 || dk.first[m] != 0;
  if ((rc = check_reflect(ctx, b, vi.d, padded))) return rc;
  bool twod = dk.n[2] == 1 && dk.n[3] == 1 && dk.first[2] == 0 && dk.first[3] == 0;
  bool mfma = false;
  if (precision == IFM_PREC_F16X3) {
    mfma = twod && dense_mfma_supported(dk, vi.d, vo.d);
    if (!mfma) return fail(ctx, IFM_E_UNSUPPORTED, "matrix-core path needs a 2-D kernel of at most 49 x 32 taps below 2^100 in magnitude");
  } else if (precision == IFM_PREC_AUTO)
    mfma = twod && dk.n[0] >= 15 && dk.n[1] >= 15 && dense_mfma_supported(dk, vi.d, vo.d);
  else if (precision != IFM_PREC_FP32) return fail(ctx, IFM_E_ARG, "unknown precision");
  ifm_ctx::TapEntry *te = nullptr;
  if ((rc = device_taps(ctx, dk, mfma, &te))) return rc;
  if (mfma)
    HIPCHK(launch_dense_mfma(vi.d, vo.d, of, dk, (const uint16_t *)te->table, (const float *)te->dev, te->kinv,
                             b->style, (float)b->fill, ctx->stream));
  else
    HIPCHK(launch_dense_valu(vi.d, vo.d, of, dk, (const float *)te->dev, b->style, (float)b->fill, ctx->stream));
  if ((rc = finish_view(ctx, vo))) return rc;
  return sync_if_host(ctx, !in->device || !out->device);
}

// ---- FFT ---------------------------------------------------------------------------
static int64_t nextprod23(int64_t n) {
  int64_t best = INT64_MAX;
  for (int64_t p2 = 1; p2 < 2 * n + 2; p2 *= 2)
    for (int64_t p = p2; p < 2 * n + 2; p *= 3)
      if (p >= n && p < best) best = p;
  return best;
}

// Twiddle tables of the tiled FFT for tile size F (uploaded once per ctx).
static int tile_twiddles(ifm_ctx *ctx, int F, const float2 **tw32, const double2 **tw64) {
  for (auto &e : ctx->tile_tw)
    if (e.F == F) {
      *tw32 = (const float2 *)e.tw32;
      *tw64 = (const double2 *)e.tw64;
      return IFM_OK;
    }
  std::vector<float2> h32;
  std::vector<double2> h64;
  fft_tile_twiddles(F, h32, h64);
  ifm_ctx::TileTw e;
  e.F = F;
  if (hipMalloc(&e.tw32, sizeof(float2) * F) != hipSuccess) return fail(ctx, IFM_E_NOMEM, "device allocation failed (FFT twiddles)");
  if (hipMalloc(&e.tw64, sizeof(double2) * F) != hipSuccess) {
    (void)hipFree(e.tw32);
    return fail(ctx, IFM_E_NOMEM, "device allocation failed (FFT twiddles)");
  }
  ctx->tile_tw.push_back(e);
  HIPCHK(hipMemcpy(e.tw32, h32.data(), sizeof(float2) * F, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e.tw64, h64.data(), sizeof(double2) * F, hipMemcpyHostToDevice));
  *tw32 = (const float2 *)e.tw32;
  *tw64 = (const double2 *)e.tw64;
  return IFM_OK;
}

// The kernel's F x F spectrum for the tiled FFT, cached per ctx and keyed by
// the exact FP32 taps, the kernel shape and F.
static int tile_spectrum(ifm_ctx *ctx, const DenseKernel &dk, int F, const double2 *tw64, const float2 **spec) {
  for (auto &e : ctx->tile_spec)
    if (e.F == F && e.k0 == dk.n[0] && e.k1 == dk.n[1] && e.taps == dk.taps) {
      *spec = (const float2 *)e.spec;
      return IFM_OK;
    }
  if (ctx->tile_spec.size() >= 16) {  // bounded cache: drop the oldest entry
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipFree(ctx->tile_spec.front().spec));
    ctx->tile_spec.erase(ctx->tile_spec.begin());
  }
  const size_t tb = dk.taps.size() * sizeof(float);
  if (ctx->tile_taps_bytes < tb) {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->tile_taps) HIPCHK(hipFree(ctx->tile_taps));
    ctx->tile_taps = nullptr;
    ctx->tile_taps_bytes = 0;
    if (hipMalloc(&ctx->tile_taps, tb) != hipSuccess) return fail(ctx, IFM_E_NOMEM, "device allocation failed (FFT taps)");
    ctx->tile_taps_bytes = tb;
  }
  ifm_ctx::TileSpec e;
  e.F = F;
  e.k0 = dk.n[0];
  e.k1 = dk.n[1];
  e.taps = dk.taps;
  if (hipMalloc(&e.spec, fft_tile_spec_bytes(F)) != hipSuccess)
    return fail(ctx, IFM_E_NOMEM, "device allocation failed (FFT kernel spectrum)");
  ctx->tile_spec.push_back(std::move(e));
  ifm_ctx::TileSpec &h = ctx->tile_spec.back();
  // the entry keeps the host taps alive until it is dropped (after a stream sync)
  HIPCHK(hipMemcpyAsync(ctx->tile_taps, h.taps.data(), tb, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(launch_fft_tile_spectrum((const float *)ctx->tile_taps, dk.n[0], dk.n[1], F, tw64, (float2 *)h.spec,
                                  ctx->stream));
  *spec = (const float2 *)h.spec;
  return IFM_OK;
}

int ifm_imfilter_fft(ifm_ctx *ctx, const ifm_array *in, ifm_array *out, const int64_t *out_first,
                     const ifm_kernel *kernel, const ifm_border *b) {
  if (!ctx) return IFM_E_ARG;
  Geo g;
  int rc;
  if ((rc = common_setup(ctx, in, out, g))) return rc;
  if ((rc = check_border(ctx, b))) return rc;
  if (!is_float_out(out->dtype)) return fail(ctx, IFM_E_UNSUPPORTED, "imfilter: output eltype must be floating point");
  DenseKernel dk;
  if ((rc = dense_kernel_canon(ctx, kernel, g, in->ndim, dk))) return rc;
  View vi, vo;
  if ((rc = make_view(ctx, in, g, S_IN, true, vi))) return rc;
  if ((rc = make_view(ctx, out, g, S_OUT, false, vo))) return rc;
  int64_t of[MAXD];
  to_canon(g, out_first, 0, of);
  if ((rc = check_region(ctx, vi.d, vo.d, of))) return rc;
  bool padded[MAXD];
  for (int m = 0; m < MAXD; ++m) padded[m] = dk.n[m] > 1 || dk.first[m] != 0;
  if ((rc = check_reflect(ctx, b, vi.d, padded))) return rc;

  // Tiled overlap-save transform: kernels of up to 80 taps on axes 0 and 1.
  const int F = fft_tile_size(dk);
  if (F > 0) {
    if (vo.d.count() > 0) {
      const float2 *tw32;
      const double2 *tw64;
      const float2 *spec;
      if ((rc = tile_twiddles(ctx, F, &tw32, &tw64))) return rc;
      if ((rc = tile_spectrum(ctx, dk, F, tw64, &spec))) return rc;
      if (!ctx->tile_flag) {
        if (hipMalloc(&ctx->tile_flag, sizeof(int)) != hipSuccess)
          return fail(ctx, IFM_E_NOMEM, "device allocation failed (FFT flag)");
        HIPCHK(hipMemsetAsync(ctx->tile_flag, 0, sizeof(int), ctx->stream));
      }
      if (++ctx->tile_gen <= 0) ctx->tile_gen = 1;
      HIPCHK(launch_fft_tile(vi.d, vo.d, of, dk, F, spec, tw32, ctx->tile_flag, ctx->tile_gen, b->style,
                             (float)b->fill, ctx->stream));
    }
    if ((rc = finish_view(ctx, vo))) return rc;
    return sync_if_host(ctx, !in->device || !out->device);
  }

  // Whole-array transform.  Axes with a kernel extent are padded to
  // nextprod([2,3], n + K) (Pad(kernel, img, FFT()), src/border.jl:240-248),
  // lengthened where needed so that every output's kernel window stays inside
  // the padded range: with lo = max(-first, 0) the window of output c reaches
  // c + lo + first + K - 1, so N >= of + on + max(first, 0) + K - 1.
  int64_t N[MAXD], lo[MAXD];
  for (int m = 0; m < MAXD; ++m) {
    if (padded[m]) {
      lo[m] = dk.first[m] < 0 ? -dk.first[m] : 0;
      const int64_t reach = of[m] + vo.d.n[m] + (dk.first[m] > 0 ? dk.first[m] : 0) + dk.n[m] - 1;
      N[m] = nextprod23(std::max<int64_t>(vi.d.n[m] + dk.n[m], reach));
    } else {
      lo[m] = 0;
      N[m] = vi.d.n[m];
    }
  }
  if (padded[3]) return fail(ctx, IFM_E_UNSUPPORTED, "FFT filtering transforms at most 3 dimensions");
  // The transform covers axes 0 .. nd-1; unpadded axes between padded ones are
  // transformed too (a length-n DFT round trip), padded if their length is
  // not a product of 2s and 3s.  Axes nd.. are batch axes.
  int nd = 1;
  for (int m = 0; m < 3; ++m)
    if (padded[m]) nd = m + 1;
  for (int m = 0; m < nd; ++m) {
    int64_t x = N[m];
    while (x % 2 == 0) x /= 2;
    while (x % 3 == 0) x /= 3;
    if (x != 1) {
      N[m] = nextprod23(vi.d.n[m] + dk.n[m]);
      padded[m] = true;
      lo[m] = dk.first[m] < 0 ? -dk.first[m] : 0;
    }
    if (!fft_axis_supported(N[m])) return fail(ctx, IFM_E_UNSUPPORTED, "FFT length not supported on one axis");
  }
  if ((double)N[0] * (double)N[1] * (double)N[2] >= 2147483648.0)
    return fail(ctx, IFM_E_UNSUPPORTED, "FFT padded array above 2^31 elements per 3-D block");
  // A 4th axis is a batch axis handled block by block.
  int64_t N3[MAXD] = {N[0], N[1], N[2], 1};
  size_t ws = fft_workspace(N3, nd) + dk.taps.size() * sizeof(float) + 256;
  if (ctx->fft_bytes < ws) {
    if (ctx->fft_work) {
      HIPCHK(hipStreamSynchronize(ctx->stream));
      HIPCHK(hipFree(ctx->fft_work));
    }
    ctx->fft_work = nullptr;
    ctx->fft_bytes = 0;
    ctx->fft_have = false;
    if (hipMalloc(&ctx->fft_work, ws + ws / 8) != hipSuccess) {
      ctx->fft_work = nullptr;
      return fail(ctx, IFM_E_NOMEM, "device allocation failed (FFT workspace)");
    }
    ctx->fft_bytes = ws + ws / 8;
  }
  bool have_kernel = ctx->fft_have && ctx->fft_nd == nd && ctx->fft_taps == dk.taps;
  for (int d = 0; d < MAXD && have_kernel; ++d)
    have_kernel = ctx->fft_kn[d] == dk.n[d] && ctx->fft_kf[d] == dk.first[d] && ctx->fft_N[d] == N3[d];
  ctx->fft_have = false;  // valid again once this call's kernel spectrum is in place
  if (vo.d.count() > 0) {
    const size_t ies = dtype_size(vi.d.dtype), oes = dtype_size(vo.d.dtype);
    for (int64_t c3 = 0; c3 < vo.d.n[3]; ++c3) {
      Dense4 bi = vi.d, bo = vo.d;
      bi.n[3] = bo.n[3] = 1;
      bi.data = (char *)vi.d.data + (size_t)((c3 + of[3]) * vi.d.n[0] * vi.d.n[1] * vi.d.n[2]) * ies;
      bo.data = (char *)vo.d.data + (size_t)(c3 * vo.d.n[0] * vo.d.n[1] * vo.d.n[2]) * oes;
      HIPCHK(launch_fft_filter(bi, bo, of, dk, N3, nd, lo, b->style, (float)b->fill, ctx->fft_work, have_kernel,
                               ctx->stream));
      have_kernel = true;
    }
    ctx->fft_have = true;
    ctx->fft_nd = nd;
    ctx->fft_taps = dk.taps;
    for (int d = 0; d < MAXD; ++d) {
      ctx->fft_kn[d] = dk.n[d];
      ctx->fft_kf[d] = dk.first[d];
      ctx->fft_N[d] = N3[d];
    }
  }
  if ((rc = finish_view(ctx, vo))) return rc;
  return sync_if_host(ctx, !in->device || !out->device);
}

// ---- IIR ------------------------------------------------------------------------------
int ifm_imfilter_iir(ifm_ctx *ctx, const ifm_array *in, ifm_array *out, const ifm_iir *factors, int nfactors,
                     const ifm_border *b) {
  if (!ctx) return IFM_E_ARG;
  Geo g;
  int rc;
  if ((rc = common_setup(ctx, in, out, g))) return rc;
  if ((rc = check_border(ctx, b))) return rc;
  if (!is_float_out(out->dtype)) return fail(ctx, IFM_E_UNSUPPORTED, "imfilter: output eltype must be floating point");
  for (int d = 0; d < in->ndim; ++d)
    if (in->size[d] != out->size[d]) return fail(ctx, IFM_E_DIM, "IIR output must have the image's axes");
  if (nfactors < 0 || (nfactors > 0 && !factors)) return fail(ctx, IFM_E_ARG, "bad factor list");
  for (int f = 0; f < nfactors; ++f) {
    if (factors[f].axis < 0 || factors[f].axis >= in->ndim)
      return fail(ctx, IFM_E_DIM, "cannot have more kernels than dimensions");
    if (in->size[factors[f].axis] <= 3)
      return fail(ctx, IFM_E_DIM, "size of img along the filtered dimension is too small for the IIR kernel of length 3");
  }
  View vi, vo;
  if ((rc = make_view(ctx, in, g, S_IN, true, vi))) return rc;
  if ((rc = make_view(ctx, out, g, S_OUT, false, vo))) return rc;
  hipStream_t st = ctx->stream;
  int isfill = b->style == IFM_FILL;
  std::vector<int> live;
  for (int f = 0; f < nfactors; ++f) {
    const ifm_iir &k = factors[f];
    bool copy = k.a[0] == 0 && k.a[1] == 0 && k.a[2] == 0 && k.b[0] == 0 && k.b[1] == 0 && k.b[2] == 0 &&
                k.scale == 1.0;  // iscopy(::TriggsSdika)
    if (!copy) live.push_back(f);
  }
  const int64_t cnt = vi.d.count();
  if (cnt > 0) {
    if (live.empty()) {
      FirPass p;
      p.axis = 0;
      p.len = 1;
      p.first = 0;
      p.taps[0] = 1.f;
      int64_t zero[MAXD] = {0, 0, 0, 0};
      if (vi.d.data != vo.d.data || vi.d.dtype != vo.d.dtype)
        HIPCHK(launch_fir1d(vi.d, zero, vo.d, zero, vi.d.n, p, b->style, (float)b->fill, st));
    } else {
      Dense4 src = vi.d;
      for (size_t s = 0; s < live.size(); ++s) {
        const ifm_iir &k = factors[live[s]];
        const bool last = s + 1 == live.size();
        Dense4 dst = vi.d;
        dst.dtype = IFM_F64;
        if (last && vo.d.dtype == IFM_F64) {
          dst.data = vo.d.data;
        } else {
          dst.data = slot_buffer(ctx, (s & 1) ? S_TMP2 : S_TMP1, (size_t)cnt * sizeof(double));
          if (!dst.data) return fail(ctx, IFM_E_NOMEM, "device allocation failed (temporary)");
        }
        IirCoef c;
        for (int i = 0; i < 3; ++i) {
          c.a[i] = k.a[i];
          c.b[i] = k.b[i];
        }
        c.scale = k.scale;
        for (int i = 0; i < 9; ++i) c.M[i] = k.M[i];
        HIPCHK(launch_iir(src, dst, g.inv[k.axis], c, isfill, b->fill, st));
        src = dst;
      }
      if (src.data != vo.d.data) {
        FirPass p;
        p.axis = 0;
        p.len = 1;
        p.first = 0;
        p.taps[0] = 1.f;
        int64_t zero[MAXD] = {0, 0, 0, 0};
        HIPCHK(launch_fir1d(src, zero, vo.d, zero, vi.d.n, p, IFM_REPLICATE, 0.f, st));
      }
    }
  }
  if ((rc = finish_view(ctx, vo))) return rc;
  return sync_if_host(ctx, !in->device || !out->device);
}

// ---- mapwindow ---------------------------------------------------------------------------
static int window_setup(ifm_ctx *ctx, const Geo &g, const int64_t *wf, const int64_t *wl, int *cwf, int *cwl) {
  if (!wf || !wl) return fail(ctx, IFM_E_ARG, "window is NULL");
  for (int m = 0; m < MAXD; ++m) {
    int d = g.perm[m];
    cwf[m] = d >= 0 ? (int)wf[d] : 0;
    cwl[m] = d >= 0 ? (int)wl[d] : 1;
    if (cwl[m] < 1) return fail(ctx, IFM_E_ARG, "empty window");
    if (cwl[m] > 4096) return fail(ctx, IFM_E_UNSUPPORTED, "window longer than 4096");
  }
  return IFM_OK;
}

// mapwindow maps edge windows relative to their intersection with the image
// (padindex(border, inner, outer), src/mapwindow.jl:310-316): an empty
// intersection fails for every Pad style, and reflect's period 2n-2 is zero
// for an intersection of one element (the reference's DivideError).
static int window_edge_check(ifm_ctx *ctx, int style, const Dense4 &in, const Dense4 &out, const int64_t *of,
                             const int *wf, const int *wl) {
  if (style == IFM_FILL) return IFM_OK;
  for (int m = 0; m < MAXD; ++m) {
    if (wf[m] == 0 && wl[m] == 1) continue;
    const int64_t n = in.n[m];
    auto bad = [&](int64_t q) {
      int64_t lo = q + wf[m], hi = q + wf[m] + wl[m] - 1;
      if (lo >= 0 && hi <= n - 1) return false;
      int64_t len = std::min(hi, n - 1) - std::max(lo, (int64_t)0) + 1;
      return len <= 0 || (style == IFM_REFLECT && len == 1);
    };
    const int64_t q0 = of[m], q1 = of[m] + out.n[m];
    for (int64_t q = q0; q < std::min(q1, q0 + wl[m] + 1); ++q)
      if (bad(q)) return fail(ctx, IFM_E_ARG, "window does not overlap the image enough for this border (empty or zero-period index map)");
    for (int64_t q = std::max(q0, q1 - wl[m] - 1); q < q1; ++q)
      if (bad(q)) return fail(ctx, IFM_E_ARG, "window does not overlap the image enough for this border (empty or zero-period index map)");
  }
  return IFM_OK;
}

int ifm_mapwindow_generic(ifm_ctx *ctx, const ifm_array *in, ifm_array *out, int op, double param,
                          const int64_t *idx_first, const int64_t *idx_step, const int64_t *win_first,
                          const int64_t *win_len, const ifm_border *b, int *nan_seen) {
  if (!ctx) return IFM_E_ARG;
  Geo g;
  int rc;
  if ((rc = common_setup(ctx, in, out, g))) return rc;
  if ((rc = check_border(ctx, b))) return rc;
  if (op < IFM_WIN_SUM || op > IFM_WIN_QUANTILE) return fail(ctx, IFM_E_ARG, "unknown window function");
  if (op == IFM_WIN_QUANTILE && (!nan_seen || !(param >= 0.0 && param <= 1.0)))
    return fail(ctx, IFM_E_ARG, "quantile needs p in [0, 1] and a nan_seen flag");
  if (!idx_first || !idx_step) return fail(ctx, IFM_E_ARG, "indices are NULL");
  const int dt = in->dtype, odt = out->dtype;
  const bool same = odt == dt;
  bool ok;
  switch (op) {
  case IFM_WIN_MIN:
  case IFM_WIN_MAX:
  case IFM_WIN_FIRST: ok = same; break;
  case IFM_WIN_MEDIAN_FAST: ok = same || ((dt == IFM_U8 || dt == IFM_N0F8) && (odt == IFM_F32 || odt == IFM_F64)); break;
  case IFM_WIN_QUANTILE: ok = odt == IFM_F64 && (dt == IFM_F32 || dt == IFM_F64 || dt == IFM_U8); break;
  case IFM_WIN_MEDIAN:
    ok = ((dt == IFM_F32 || dt == IFM_F64 || dt == IFM_BF16 || dt == IFM_F16) && same) || (dt == IFM_U8 && odt == IFM_F64);
    break;
  default: ok = ((dt == IFM_F32 || dt == IFM_F64) && same) || (dt == IFM_U8 && odt == IFM_F64); break;
  }
  if (!ok) return fail(ctx, IFM_E_UNSUPPORTED, "window function / eltype pair not provided on the GPU");
  int wf[MAXD], wl[MAXD];
  if ((rc = window_setup(ctx, g, win_first, win_len, wf, wl))) return rc;
  if ((int64_t)wl[0] * wl[1] * wl[2] * wl[3] > 4096) return fail(ctx, IFM_E_UNSUPPORTED, "window larger than 4096 elements");
  View vi, vo;
  if ((rc = make_view(ctx, in, g, S_IN, true, vi))) return rc;
  if ((rc = make_view(ctx, out, g, S_OUT, false, vo))) return rc;
  int64_t f[MAXD], st[MAXD];
  to_canon(g, idx_first, 0, f);
  to_canon(g, idx_step, 1, st);
  for (int m = 0; m < MAXD; ++m) {
    if (vo.d.n[m] == 0) continue;
    const int64_t last = f[m] + (vo.d.n[m] - 1) * st[m];
    if (st[m] == 0 || f[m] < 0 || f[m] >= vi.d.n[m] || last < 0 || last >= vi.d.n[m])
      return fail(ctx, IFM_E_DIM, "window centres lie outside the image axes");
    if (b->style == IFM_FILL || (wf[m] == 0 && wl[m] == 1)) continue;
    for (int64_t i = 0; i < vo.d.n[m]; ++i) {  // edge windows need a usable index map
      const int64_t q = f[m] + i * st[m], lo = q + wf[m], hi = q + wf[m] + wl[m] - 1;
      if (lo >= 0 && hi <= vi.d.n[m] - 1) continue;
      const int64_t len = std::min(hi, vi.d.n[m] - 1) - std::max(lo, (int64_t)0) + 1;
      if (len <= 0 || (b->style == IFM_REFLECT && len == 1))
        return fail(ctx, IFM_E_ARG, "window does not overlap the image enough for this border (empty or zero-period index map)");
    }
  }
  if (!ctx->finite_flag && hipMalloc(&ctx->finite_flag, sizeof(int)) != hipSuccess) {
    ctx->finite_flag = nullptr;
    return fail(ctx, IFM_E_NOMEM, "device allocation failed (flag)");
  }
  HIPCHK(hipMemsetAsync(ctx->finite_flag, 0, sizeof(int), ctx->stream));
  HIPCHK(launch_window_generic(vi.d, vo.d, f, st, wf, wl, b->style, b->fill, op, param, ctx->finite_flag,
                               ctx->stream));
  if ((rc = finish_view(ctx, vo))) return rc;
  if (nan_seen) {
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, ctx->finite_flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *nan_seen = h ? 1 : 0;
    return IFM_OK;
  }
  return sync_if_host(ctx, !in->device || !out->device);
}

int ifm_mapwindow_median(ifm_ctx *ctx, const ifm_array *in, ifm_array *out, const int64_t *out_first,
                         const int64_t *win_first, const int64_t *win_len, const ifm_border *b) {
  if (!ctx) return IFM_E_ARG;
  Geo g;
  int rc;
  if ((rc = common_setup(ctx, in, out, g))) return rc;
  if ((rc = check_border(ctx, b))) return rc;
  const bool codes_in = in->dtype == IFM_U8 || in->dtype == IFM_N0F8;
  if (in->dtype != out->dtype && !(codes_in && (out->dtype == IFM_F32 || out->dtype == IFM_F64)))
    return fail(ctx, IFM_E_ARG, "median output holds the input eltype (or Float32 / Float64 for UInt8 / N0f8)");
  int wf[MAXD], wl[MAXD];
  if ((rc = window_setup(ctx, g, win_first, win_len, wf, wl))) return rc;
  int64_t L = (int64_t)wl[0] * wl[1] * wl[2] * wl[3];
  if (L > 4096) return fail(ctx, IFM_E_UNSUPPORTED, "median window larger than 4096 elements");
  View vi, vo;
  if ((rc = make_view(ctx, in, g, S_IN, true, vi))) return rc;
  if ((rc = make_view(ctx, out, g, S_OUT, false, vo))) return rc;
  int64_t of[MAXD];
  to_canon(g, out_first, 0, of);
  if ((rc = check_region(ctx, vi.d, vo.d, of))) return rc;
  if ((rc = window_edge_check(ctx, b->style, vi.d, vo.d, of, wf, wl))) return rc;
  if (vi.d.count() == 0 && vo.d.count() > 0) return fail(ctx, IFM_E_ARG, "empty image");
  HIPCHK(launch_median(vi.d, vo.d, of, wf, wl, b->style, b->fill, ctx->stream));
  if ((rc = finish_view(ctx, vo))) return rc;
  return sync_if_host(ctx, !in->device || !out->device);
}

static int minmax_common(ifm_ctx *ctx, const ifm_array *in, ifm_array *omin, ifm_array *omax, const int64_t *out_first,
                         const int *wf_user_canon_unused, const int64_t *win_first, const int64_t *win_len, int style,
                         double fill, bool lemire) {
  (void)wf_user_canon_unused;
  Geo g;
  int rc;
  ifm_array *ref = omin ? omin : omax;
  if (!ref) return fail(ctx, IFM_E_ARG, "no output requested");
  if ((rc = common_setup(ctx, in, ref, g))) return rc;
  if (omin && omax) {
    if ((rc = check_array(ctx, omax, "omax"))) return rc;
    for (int d = 0; d < in->ndim; ++d)
      if (omin->size[d] != omax->size[d]) return fail(ctx, IFM_E_DIM, "omin and omax differ in size");
    if (!dense_under(*omax, g)) return fail(ctx, IFM_E_UNSUPPORTED, "omax must share the layout of in");
  }
  if ((omin && omin->dtype != in->dtype) || (omax && omax->dtype != in->dtype))
    return fail(ctx, IFM_E_ARG, "min/max outputs hold the input eltype");
  int wf[MAXD], wl[MAXD];
  if ((rc = window_setup(ctx, g, win_first, win_len, wf, wl))) return rc;
  View vi, vmin, vmax;
  if ((rc = make_view(ctx, in, g, S_IN, true, vi))) return rc;
  if (omin && (rc = make_view(ctx, omin, g, S_OUT, false, vmin))) return rc;
  if (omax && (rc = make_view(ctx, omax, g, S_OUT2, false, vmax))) return rc;
  int64_t of[MAXD];
  to_canon(g, out_first, 0, of);
  Dense4 &r = omin ? vmin.d : vmax.d;
  if ((rc = check_region(ctx, vi.d, r, of))) return rc;
  if (!lemire && (rc = window_edge_check(ctx, style, vi.d, r, of, wf, wl))) return rc;
  if (vi.d.count() == 0 && r.count() > 0) return fail(ctx, IFM_E_ARG, "empty image");
  size_t ws = minmax_workspace(vi.d, r.n);
  void *work = slot_buffer(ctx, S_WORK, ws);
  if (!work) return fail(ctx, IFM_E_NOMEM, "device allocation failed (min/max workspace)");
  HIPCHK(launch_minmax(vi.d, omin ? &vmin.d : nullptr, omax ? &vmax.d : nullptr, of, wf, wl,
                       style == IFM_FILL ? 1 : 0, fill, work, ctx->stream));
  if (omin && (rc = finish_view(ctx, vmin))) return rc;
  if (omax && (rc = finish_view(ctx, vmax))) return rc;
  bool host = !in->device || (omin && !omin->device) || (omax && !omax->device);
  return sync_if_host(ctx, host);
}

int ifm_mapwindow_minmax(ifm_ctx *ctx, const ifm_array *in, ifm_array *omin, ifm_array *omax, const int64_t *out_first,
                         const int64_t *win_first, const int64_t *win_len, const ifm_border *b) {
  if (!ctx) return IFM_E_ARG;
  int rc;
  if ((rc = check_border(ctx, b))) return rc;
  return minmax_common(ctx, in, omin, omax, out_first, nullptr, win_first, win_len, b->style, b->fill, false);
}

int ifm_extrema_filter(ifm_ctx *ctx, const ifm_array *in, ifm_array *omin, ifm_array *omax, const int64_t *window) {
  if (!ctx) return IFM_E_ARG;
  if (!window || !in) return fail(ctx, IFM_E_ARG, "window is NULL");
  int64_t wf[MAXD] = {0, 0, 0, 0}, wl[MAXD] = {1, 1, 1, 1};
  for (int d = 0; d < in->ndim && d < MAXD; ++d) {
    if (window[d] < 1) return fail(ctx, IFM_E_ARG, "window entries must be positive");
    // Lemire with the delayed cache covers offsets -(w>>1) .. w-1-(w>>1)
    wf[d] = -(window[d] >> 1);
    wl[d] = window[d];
  }
  return minmax_common(ctx, in, omin, omax, nullptr, nullptr, wf, wl, IFM_REPLICATE, 0.0, true);
}

}  // extern "C"
