/*
 * ifm.h — C ABI of libifm, the MI355X (gfx950) filtering library that backs the
 * ImageFiltering.jl hot path: imfilter / imfilter! (FIR, FFT, IIR), padarray and
 * the mapwindow median / min / max / extrema filters.
 *
 * Conventions
 *   - Axis d of an ifm_array is Julia dimension d+1.  Sizes and strides are in
 *     elements.  Any dense layout is accepted (column-major, row-major or any
 *     other axis permutation); input and output must share the permutation.
 *   - All indices are 0-based.  A Julia axis first(inds) is passed as the
 *     0-based offset of the output region (out_first) or of a kernel's first
 *     tap (ifm_factor.first, ifm_kernel.first).
 *   - data points to device memory when device == 1 and to host memory when
 *     device == 0 (the library then stages the array through HBM and the call
 *     returns after the result has been copied back).
 *   - Calls on one ctx are serialised on the ctx's stream.  Calls without a
 *     server return once the work is enqueued on the stream (device arrays) or
 *     finished (host arrays); ifm_ctx_synchronize waits for the stream.
 *   - Every entry point returns an ifm_status; ifm_last_error gives the message.
 *     IFM_E_ARG maps to Julia's ArgumentError and IFM_E_DIM to DimensionMismatch.
 *   - Border Fill values arrive already converted to the input eltype
 *     (src/borderarray.jl:11-24) and are passed as that value in a double:
 *     exact for every eltype but N0f8, whose value code/255 is passed (the
 *     library recovers the code as round(255 * fill)).
 */
#ifndef IFM_H
#define IFM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IFM_MAXDIM 4

typedef enum {
  IFM_OK = 0,
  IFM_E_ARG = 1,          /* ArgumentError */
  IFM_E_DIM = 2,          /* DimensionMismatch */
  IFM_E_NOMEM = 3,        /* device allocation failed */
  IFM_E_DEVICE = 4,       /* HIP runtime error */
  IFM_E_UNSUPPORTED = 5,  /* eltype / rank / layout not supported by a kernel */
  IFM_E_BUDGET = 6        /* serving: request exceeds the HBM budget */
} ifm_status;

/* Element types.  N0F8 is UInt8 storage read as code/255 by linear filters;
 * the order filters (median, min, max, extrema) work on the raw codes. */
typedef enum {
  IFM_U8 = 0, IFM_N0F8 = 1, IFM_F32 = 2, IFM_F64 = 3, IFM_BF16 = 4, IFM_F16 = 5
} ifm_dtype;

/* Border styles (src/border.jl:564-590, src/borderarray.jl:110-149). */
typedef enum {
  IFM_REPLICATE = 0, IFM_CIRCULAR = 1, IFM_SYMMETRIC = 2, IFM_REFLECT = 3, IFM_FILL = 4
} ifm_style;

/* Dense-kernel arithmetic for ifm_imfilter_dense. */
typedef enum {
  IFM_PREC_AUTO = 0,    /* library choice: matrix cores for 2-D kernels >= 15x15, else FP32 VALU */
  IFM_PREC_FP32 = 1,    /* FP32 FMA on the vector ALUs */
  IFM_PREC_F16X3 = 2    /* matrix cores: scaled x and k split in FP16 hi+lo, 3 products, FP32
                           accumulate (2-D kernels up to 49 x 32 of finite taps below 2^100;
                           error ~2^-22 relative) */
} ifm_precision;

typedef struct {
  int32_t ndim;                 /* 1..IFM_MAXDIM */
  int32_t dtype;                /* ifm_dtype */
  int64_t size[IFM_MAXDIM];
  int64_t stride[IFM_MAXDIM];
  void *data;
  int32_t device;               /* 1: device pointer, 0: host pointer */
  int32_t reserved;
} ifm_array;

typedef struct {
  int32_t style;                /* ifm_style */
  int32_t reserved;
  double fill;                  /* value for IFM_FILL */
  int64_t lo[IFM_MAXDIM];       /* explicit padding, used by ifm_padarray only */
  int64_t hi[IFM_MAXDIM];
} ifm_border;

/* One 1-D factor of a separable kernel: a ReshapedOneD along `axis`
 * (src/kernelfactors.jl:52-131) with taps k[first .. first+len-1]. */
typedef struct {
  int32_t axis;
  int32_t len;
  int64_t first;
  const double *taps;           /* host memory, len values */
} ifm_factor;

/* A dense N-D kernel with axes first[d] .. first[d]+size[d]-1, taps stored
 * column-major (axis 0 fastest) in host memory. */
typedef struct {
  int32_t ndim;
  int32_t reserved;
  int64_t size[IFM_MAXDIM];
  int64_t first[IFM_MAXDIM];
  const double *taps;
} ifm_kernel;

/* A Triggs-Sdika IIR factor (src/kernelfactors.jl:463-513): forward
 * coefficients a, backward b, final scale and the 3x3 boundary matrix M
 * (row-major, M[3*r + c]) applied along `axis`. */
typedef struct {
  double a[3];
  double b[3];
  double scale;
  double M[9];
  int32_t axis;
  int32_t reserved;
} ifm_iir;

typedef struct ifm_ctx ifm_ctx;

/* ---- context ------------------------------------------------------------ */
/* stream: a hipStream_t (may be NULL for the null stream). */
int ifm_ctx_create(int device, void *stream, ifm_ctx **out);
int ifm_ctx_destroy(ifm_ctx *ctx);
int ifm_ctx_set_stream(ifm_ctx *ctx, void *stream);
int ifm_ctx_synchronize(ifm_ctx *ctx);
const char *ifm_last_error(const ifm_ctx *ctx);
/* Library version string, e.g. "ifm 0.1 gfx950". */
const char *ifm_version(void);

/* ---- padarray ------------------------------------------------------------ */
/* Replaces padarray(T, img, border) -> BorderArray + copy!
 * (src/border.jl:324-347, src/borderarray.jl:157-196).  out.size[d] must be
 * lo[d] + in.size[d] + hi[d]; in and out may differ in eltype (convert). */
int ifm_padarray(ifm_ctx *ctx, const ifm_array *in, ifm_array *out, const ifm_border *border);

/* ---- imfilter: FIR ------------------------------------------------------- */
/* Replaces the cascade of 1-D factors imfilter!(r, out, padarray(img),
 * (f1, f2, ...), NoPad(border)) (src/imfilter.jl:331-457, :671-739).
 * out covers image indices out_first[d] .. out_first[d]+out.size[d]-1
 * (NULL out_first = 0, i.e. the full image as in imfilter; Inner() passes its
 * interior).  Factors may repeat an axis.  The border map is applied on load,
 * so no padded copy is made. */
int ifm_imfilter_fir(ifm_ctx *ctx, const ifm_array *in, ifm_array *out, const int64_t *out_first,
                     const ifm_factor *factors, int nfactors, const ifm_border *border);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ifm_imfilter_fir with the same arguments and results, but
 * the cascade's FP32 temporaries live in a fixed device array of
 * IFM_NOALLOC_SCRATCH_BYTES (two halves, one per ping-pong temporary; each GPU
 * holds its own copy until the process ends) and nothing is allocated.  in and
 * out must be device arrays.  A cascade whose temporary exceeds half the array
 * returns IFM_E_BUDGET.  Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (several ctx streams at once) race. */
#define IFM_NOALLOC_SCRATCH_BYTES (64u << 20)
int ifm_imfilter_fir_noalloc(ifm_ctx *ctx, const ifm_array *in, ifm_array *out, const int64_t *out_first,
                             const ifm_factor *factors, int nfactors, const ifm_border *border);

/* Replaces the dense N-D correlation __imfilter_inbounds!
 * (src/imfilter.jl:592-669) on the padded image. */
int ifm_imfilter_dense(ifm_ctx *ctx, const ifm_array *in, ifm_array *out, const int64_t *out_first,
                       const ifm_kernel *kernel, const ifm_border *border, int precision);

/* all(isfinite, a) for a device array (*result = 1 if every value is finite;
 * integer and N0f8 arrays are always finite): the scan filter_algorithm makes
 * before it picks FFT for kernels longer than 30 (src/imfilter.jl:1200-1209).
 * One streaming read of the array; returns after the answer is on the host. */
int ifm_all_finite(ifm_ctx *ctx, const ifm_array *a, int *result);

/* ---- imfilter: FFT ------------------------------------------------------- */
/* Replaces _imfilter_fft! / filtfft (src/imfilter.jl:796-866): the
 * correlation of the border-extended image with the kernel, computed by FFT.
 * Kernels spanning axes 0 and 1 only with at most 80 taps per axis run as
 * tiled overlap-save transforms (64 x 64 or 128 x 128 tiles, one HBM read and
 * write per pixel; further axes are batch axes).  Other kernels use a
 * whole-array transform padded to nextprod([2,3], n + K) per axis as
 * Pad(kernel, img, FFT()) (src/border.jl:240-248), up to 3 transformed axes
 * (a 4th is a batch axis), lines beyond 8192 in two levels.  The ctx caches
 * kernel spectra (compared exactly) and reuses them while the kernel and sizes
 * repeat.  A non-finite input value makes every output NaN, as the
 * whole-array transform does. */
int ifm_imfilter_fft(ifm_ctx *ctx, const ifm_array *in, ifm_array *out, const int64_t *out_first,
                     const ifm_kernel *kernel, const ifm_border *border);

/* ---- imfilter: IIR ------------------------------------------------------- */
/* Replaces _imfilter_dim! with leftborder!/rightborder!
 * (src/imfilter.jl:964-1092) applied once per factor, in order.  The
 * recursion runs in FP64.  A Pad-style border initialises from the edge
 * pixel, IFM_FILL from border->fill.  in and out may alias. */
int ifm_imfilter_iir(ifm_ctx *ctx, const ifm_array *in, ifm_array *out,
                     const ifm_iir *factors, int nfactors, const ifm_border *border);

/* ---- NA() borders and sub-ranges ------------------------------------------ */
/* The elementwise steps of imfilter(img, kernel, NA()) as
 * imfilter_na_inseparable! (src/imfilter.jl:1109-1122) takes them: tmp = img
 * with NaN replaced by zero, valid = 0 where img is NaN and 1 elsewhere (both
 * in their own floating-point eltype), *hasna = any(isnan, img).  The host
 * then filters tmp and valid with Fill(0) and divides. */
int ifm_na_prepare(ifm_ctx *ctx, const ifm_array *in, ifm_array *tmp, ifm_array *valid, int *hasna);
/* num ./= den elementwise in FP64 (out[I] /= vp[I], src/imfilter.jl:1119-1121). */
int ifm_divide(ifm_ctx *ctx, ifm_array *num, const ifm_array *den);
/* copyto!(dst, R, src, R) for the block R = dst_first .. dst_first + size(src)
 * - 1 of dst (src/imfilter.jl:372-375, :433-434): the NoPad sub-range form of
 * imfilter! and the IIR stages of mixed cascades.  Eltypes may differ (convert). */
int ifm_copy_into(ifm_ctx *ctx, const ifm_array *src, ifm_array *dst, const int64_t *dst_first);

/* convert(T, x) into an integer output eltype T (bits 8/16/32/64, signed or
 * not) of floating-point filter results, each first multiplied by scale (1,
 * or 255 for N0f8 codes, which round to the nearest code): the store into an
 * integer out array of __imfilter_inbounds! (src/imfilter.jl:650-669).  dst
 * has src's dense layout (device memory when dst_device, else host).
 * *inexact = 1 when a value is not an integer (scale 1) or lies outside T's
 * range; Julia's convert throws InexactError there (src/imfilter.jl:233-254). */
int ifm_convert_int(ifm_ctx *ctx, const ifm_array *src, void *dst, int dst_device, int bits, int is_signed,
                    double scale, int *inexact);

/* findlocalmaxima (sense = +1) / findlocalminima (sense = -1)
 * (src/extrema.jl:122-163): flags[i] = 1 where img[i] is strictly greater
 * (smaller) than every in-bounds element of the odd window around it
 * (window[d] <= 63; NULL: 3 per axis), excluding the first and last index of
 * axis d when edges[d] == 0 (NULL: all edges included).  The caller turns the
 * flags into the CartesianIndex list (the blob_LoG peaks). */
int ifm_local_extrema(ifm_ctx *ctx, const ifm_array *in, ifm_array *flags, const int64_t *window,
                      const int32_t *edges, int sense);

/* ---- mapwindow ------------------------------------------------------------ */
/* Window functions of ifm_mapwindow_generic. */
typedef enum {
  IFM_WIN_SUM = 0,          /* sum (FP64 accumulation, stored in out's eltype) */
  IFM_WIN_MEAN = 1,         /* Statistics.mean */
  IFM_WIN_MIN = 2,          /* minimum (NaN propagates) */
  IFM_WIN_MAX = 3,          /* maximum (NaN propagates) */
  IFM_WIN_FIRST = 4,        /* first (the window's first element, column-major) */
  IFM_WIN_MEDIAN_FAST = 5,  /* median! -> median_fast!: the ((L+1)/2)-th smallest under isless */
  IFM_WIN_MEDIAN = 6,       /* Statistics.median: middle of the two central values for even L,
                               NaN when the window holds a NaN */
  IFM_WIN_QUANTILE = 7      /* Statistics.quantile(v, param), linear interpolation between
                               order statistics; a window holding NaN sets *nan_seen */
} ifm_window_op;

/* Replaces mapwindow(f, img, window; border, indices) / mapwindow! with a
 * window function from ifm_window_op, src/mapwindow.jl:75-121, :270-333: out[i]
 * = f(window at centre idx_first + i * idx_step), per Julia axis, 0-based;
 * out's extents are the numbers of centres (the strided `indices` ranges;
 * _IndexTransformer, src/mapwindow.jl:156-189).  Edge windows map relative to
 * their intersection with the image (copy_win!, :300-325), Fill substitutes
 * the value.  out's eltype: in's for SUM / MEAN / MIN / MAX / FIRST / MEDIAN
 * of floats, Float64 for UInt8 SUM / MEAN / MEDIAN / QUANTILE, Float32 or
 * Float64 as typeof(middle(x)) for MEDIAN_FAST of UInt8 / N0f8, Float64 for
 * QUANTILE.  Windows up to 4096 elements; every centre must lie inside the
 * image.  nan_seen (may be NULL unless op is QUANTILE) reports a NaN met by
 * QUANTILE, where the reference throws ArgumentError. */
int ifm_mapwindow_generic(ifm_ctx *ctx, const ifm_array *in, ifm_array *out, int op, double param,
                          const int64_t *idx_first, const int64_t *idx_step, const int64_t *win_first,
                          const int64_t *win_len, const ifm_border *border, int *nan_seen);

/* Replaces mapwindow(median!, img, window; border) -> median_fast!
 * (src/mapwindow.jl:123-131, :270-333): the ((L+1)÷2)-th smallest of the
 * L window values under isless order.  win_first[d]/win_len[d] give the window
 * offsets.  out holds the selected element in the input eltype, or for UInt8 /
 * N0f8 input its Float32 / Float64 value (typeof(middle(x)), written directly). */
int ifm_mapwindow_median(ifm_ctx *ctx, const ifm_array *in, ifm_array *out, const int64_t *out_first,
                         const int64_t *win_first, const int64_t *win_len, const ifm_border *border);

/* Replaces mapwindow(minimum|maximum, img, window; border) and
 * mapwindow(extrema, img, window; border) (src/mapwindow.jl:270-333).
 * Either of omin / omax may be NULL. */
int ifm_mapwindow_minmax(ifm_ctx *ctx, const ifm_array *in, ifm_array *omin, ifm_array *omax,
                         const int64_t *out_first, const int64_t *win_first, const int64_t *win_len,
                         const ifm_border *border);

/* Replaces mapwindow(extrema, img, window::Dims) -> extrema_filter
 * (src/mapwindow.jl:337-481): truncated windows of width window[d] (even
 * widths allowed), independent of any border. */
int ifm_extrema_filter(ifm_ctx *ctx, const ifm_array *in, ifm_array *omin, ifm_array *omax,
                       const int64_t *window);

/* ---- batched serving ------------------------------------------------------- */
/* No counterpart in the reference: the serving form of the filters above, for
 * many 2-D images of uneven sizes in one launch per op class. */
typedef enum {
  IFM_OP_SEPARABLE = 0,  /* imfilter(img, (f1[, f2]), border): factors on axes 0 / 1, <= 64 taps each */
  IFM_OP_DENSE = 1,      /* imfilter(img, kernel, border, Algorithm.FIR()) on the matrix cores, <= 49 x 32 */
  IFM_OP_MEDIAN = 2      /* mapwindow(median!, img, (k, k); border), k = 3, 5 or 7 */
} ifm_op;

/* One request: a 2-D image filtered over its whole extent (out.size == in.size).
 * With in.data == NULL (ifm_serve only) the loop allocates in and out from its
 * HBM arena and fills in with seeded uniform [0,1) values (seed = the request's
 * index), the synthetic input of the serving benchmark. */
typedef struct {
  int32_t op;                   /* ifm_op */
  int32_t nfactors;             /* IFM_OP_SEPARABLE: 1 or 2 */
  ifm_array in;                 /* device, ndim 2 */
  ifm_array out;                /* device, ndim 2, same size */
  ifm_border border;
  ifm_factor factors[2];        /* IFM_OP_SEPARABLE (host taps) */
  ifm_kernel kernel;            /* IFM_OP_DENSE (host taps) */
  int64_t window;               /* IFM_OP_MEDIAN: k */
  int32_t nfactors2;            /* IFM_OP_SEPARABLE: > 0 adds out2 = imfilter(in, factors2), the */
  int32_t reserved;             /*   second component of imgradients (src/specialty.jl:39-53) */
  ifm_array out2;
  ifm_factor factors2[2];
} ifm_request;

/* Queues n requests on the ctx stream: one launch per op class (separable,
 * matrix-core dense per kernel shape, median per window and key width);
 * circular or smaller-than-window medians run the per-image median kernel.
 * Results equal those of the single-image entry points' kernels. */
int ifm_batch_run(ifm_ctx *ctx, const ifm_request *reqs, int n);

typedef struct {
  uint64_t budget_bytes;        /* HBM admission budget: in + out bytes of the requests in flight */
  int32_t mode;                 /* 0: continuous batching, 1: per-call dispatch (one synchronous call each) */
  int32_t max_batch;            /* most requests admitted into one batch (0: 1024) */
  int32_t max_inflight;         /* most batches queued on the device at once (0: 4) */
  int32_t reserved;
} ifm_serve_config;

/* Device-resident request loop.  Request i arrives arrival_s[i] seconds after
 * the call starts (non-decreasing).  Mode 0 admits arrived requests in arrival
 * order while their bytes fit the budget, launches them as a batch without
 * waiting for earlier batches, and retires batches as their events complete.
 * Mode 1 runs each request alone: allocate, generate, the single-image entry
 * point, synchronise.  On return admit_s[i] / done_s[i] hold the times (s,
 * since the call started) at which request i was launched / seen complete. */
int ifm_serve(ifm_ctx *ctx, const ifm_request *reqs, const double *arrival_s, int n,
              const ifm_serve_config *cfg, double *admit_s, double *done_s);

/* ---- request server: non-blocking submit / poll ------------------------------ */
/* No counterpart in the reference (its callers loop over imfilter /
 * mapwindow).  A server batches the requests submitted to it and launches
 * them on its ctx's stream without blocking the caller: a submitted request
 * joins the pending batch, which is launched as soon as fewer than
 * max_inflight batches are on the device and the bytes of the requests in
 * flight (in + out, plus out2) stay within budget_bytes; requests submitted
 * while others run join the next batch (continuous batching).  Tickets
 * complete in submission order.  Requests carry device arrays that the
 * caller keeps alive until their ticket completes; their taps are copied at
 * submit.  One thread per server; a server owns its ctx's stream while it has
 * requests in flight.  in.data == NULL (generated inputs) is not accepted. */
typedef struct ifm_server ifm_server;

typedef struct {
  uint64_t budget_bytes;        /* in + out bytes of the requests in flight (0: no limit) */
  int32_t max_batch;            /* most requests in one launch (0: 1024) */
  int32_t max_inflight;         /* most batches on the device at once (0: 2; at most 7) */
} ifm_server_config;

int ifm_server_create(ifm_ctx *ctx, const ifm_server_config *cfg, ifm_server **server);
/* Queues one request (checked as ifm_batch_run checks it) and returns its
 * ticket; launches what the limits allow.  Never waits for the device. */
int ifm_server_submit(ifm_server *server, const ifm_request *req, uint64_t *ticket);
/* Retires completed batches, launches what the limits allow, and sets *done
 * to 1 when the ticket's result is in its out array.  Never waits. */
int ifm_server_poll(ifm_server *server, uint64_t ticket, int *done);
/* Launches every pending request that fits the limits, then waits until the
 * ticket completes (0: until every submitted request completes). */
int ifm_server_wait(ifm_server *server, uint64_t ticket);
/* Waits for everything submitted, then frees the server. */
int ifm_server_destroy(ifm_server *server);

/* Fills a device array with seeded uniform [0,1) values: the top 24 bits of a
 * counter hash times 2^-24 (exact in FP32; BF16 / FP16 round to nearest even). */
int ifm_fill_uniform(ifm_ctx *ctx, ifm_array *a, uint64_t seed);

/* ---- tile shapes -------------------------------------------------------------
 * The fused two-factor separable kernel (single-image imfilter over axes 0 and 1,
 * and the separable op of ifm_batch_run / ifm_serve) computes 64 x tile_h output
 * tiles, tile_h = 32 or 64.  The host's autotuner times both per (input dtype,
 * longest factor, image size) bucket and installs the winners as rules; the
 * first rule that applies to a call picks its tile_h, and calls no rule covers
 * use 32.  Results do not depend on the tile shape (same taps, same order).
 * No reference counterpart: ImageFiltering's FIRTiled picks tile sizes from
 * TiledIteration's cache heuristic (src/imfilter.jl:1298-1312). */
typedef enum { IFM_TILE_SEP2D = 1 } ifm_tile_op;
typedef struct {
  int32_t op;          /* ifm_tile_op */
  int32_t dtype;       /* input ifm_dtype, or -1 for any */
  int32_t kmax;        /* applies when the longer of the two factors has <= kmax taps */
  int32_t tile_h;      /* 32 or 64 */
  int64_t max_pixels;  /* applies when the output plane has <= max_pixels pixels */
} ifm_tile_rule;
#define IFM_MAX_TILE_RULES 256
/* Replaces the ctx's rules (n = 0 clears them). */
int ifm_set_tile_rules(ifm_ctx *ctx, const ifm_tile_rule *rules, int n);
/* The tile height a call with these parameters would use. */
int ifm_tile_height(ifm_ctx *ctx, int op, int dtype, int kmax, int64_t pixels, int *tile_h);

/* ---- device memory plumbing ------------------------------------------------ */
/* Helpers for hosts without their own HIP allocator (the Python package and
 * the Julia binding use them for device-resident arrays).  Copies are
 * synchronous with respect to the ctx stream. */
int ifm_device_alloc(ifm_ctx *ctx, uint64_t bytes, void **ptr);
int ifm_device_free(ifm_ctx *ctx, void *ptr);
int ifm_memcpy_h2d(ifm_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int ifm_memcpy_d2h(ifm_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int ifm_memset_d(ifm_ctx *ctx, void *dst, int value, uint64_t bytes);
/* HIP events on the ctx stream: ms between the last start and stop. */
int ifm_timer_start(ifm_ctx *ctx);
int ifm_timer_stop(ifm_ctx *ctx, float *ms);
/* Number of visible devices (0 when no GPU); never initialises a context. */
int ifm_device_count(int *count);

#ifdef __cplusplus
}
#endif

#endif /* IFM_H */
