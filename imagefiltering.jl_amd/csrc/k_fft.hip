// k_fft.hip — FFT correlation (gfx950), the Algorithm.FFT() choice of
// imfilter (src/imfilter.jl:796-866).
//
// As in the reference, the image is padded per axis to N = nextprod([2,3],
// n + K) with the chosen border (src/border.jl:240-248), the kernel is
// wrapped into an N-periodic array (FFTView), and the result is
// ifft(fft(A) .* conj(fft(k))) read back at the output axes.  Transforms are
// complex FP32 Stockham autosort FFTs of radix 2, 3 and 4, one line per
// workgroup in LDS (N <= 8192); axes other than the contiguous one are
// transposed to it and back.  Twiddles come from an FP64-computed table.
#include <math.h>

#include "ifm_internal.h"

namespace ifm {

constexpr int FFT_MAXN = 8192;
constexpr int FFT_THREADS = 512;  // 8 waves per line: 7% faster than 4 (occupancy), 1024 slower
constexpr int FFT_MAXJ = FFT_MAXN / (2 * FFT_THREADS);  // radix-2 butterflies per thread
constexpr int FFT_MAXPASS = 16;

struct FftPlan {
  int N;
  int npass;
  int radix[FFT_MAXPASS];
};

static bool make_plan(int N, FftPlan &p) {
  p.N = N;
  p.npass = 0;
  int m = N;
  while (m % 3 == 0) { p.radix[p.npass++] = 3; m /= 3; }
  while (m % 4 == 0) { p.radix[p.npass++] = 4; m /= 4; }
  while (m % 2 == 0) { p.radix[p.npass++] = 2; m /= 2; }
  return m == 1 && N <= FFT_MAXN && p.npass <= FFT_MAXPASS;
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

struct FftLinesArgs {
  float2 *data;        // nlines x N, contiguous lines
  const float2 *tw;    // W_N^m = exp(-2 pi i m / N), m < N
  int N;
  int64_t nlines;
  int sign;            // -1 forward, +1 inverse (unnormalised)
  int npass;
  int radix[FFT_MAXPASS];
};

// j / Ns and j % Ns for j < 2^13 by a float reciprocal, corrected by one step
// (the product's error is far below one unit at these sizes).
__device__ __forceinline__ void divmod_small(int j, int ns, float rns, int &q, int &r) {
  q = (int)((float)j * rns);
  r = j - q * ns;
  if (r < 0) {
    --q;
    r += ns;
  } else if (r >= ns) {
    ++q;
    r -= ns;
  }
}

template <int R>
__device__ __forceinline__ void fft_pass(float2 *s, const float2 *tw, int N, int Ns, int sign) {
  const int NR = N / R;
  const float rns = 1.0f / (float)Ns;
  const int unit = N / (Ns * R);
  constexpr int J = (R == 2) ? FFT_MAXJ : (R == 3 ? (FFT_MAXN / 3 + FFT_THREADS - 1) / FFT_THREADS
                                                  : FFT_MAXN / (4 * FFT_THREADS));
  float2 v[J][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int c = 0; c < J; ++c) {
    const int j = tid + c * FFT_THREADS;
    if (j < NR) {
      int q, k;
      divmod_small(j, Ns, rns, q, k);
      const int step = unit * k;  // twiddle exponent unit: step * i < N
#pragma unroll
      for (int i = 0; i < R; ++i) {
        float2 x = s[j + i * NR];
        if (i > 0) {
          float2 w = tw[step * i];
          if (sign > 0) w.y = -w.y;
          x = cmulf(x, w);
        }
        v[c][i] = x;
      }
      // R-point DFT
      if (R == 2) {
        float2 a = v[c][0], b = v[c][1];
        v[c][0] = make_float2(a.x + b.x, a.y + b.y);
        v[c][1] = make_float2(a.x - b.x, a.y - b.y);
      } else if (R == 3) {
        const float c3 = -0.5f, s3 = 0.86602540378443864676f * (float)sign;
        float2 a = v[c][0], b = v[c][1], d = v[c][2];
        float2 t = make_float2(b.x + d.x, b.y + d.y);
        float2 u = make_float2(b.x - d.x, b.y - d.y);
        v[c][0] = make_float2(a.x + t.x, a.y + t.y);
        float2 m = make_float2(a.x + c3 * t.x, a.y + c3 * t.y);
        // -s3 * i * u = (s3 * u.y, -s3 * u.x) with s3 carrying the sign
        v[c][1] = make_float2(m.x - s3 * u.y, m.y + s3 * u.x);
        v[c][2] = make_float2(m.x + s3 * u.y, m.y - s3 * u.x);
      } else {  // R == 4
        float2 a = v[c][0], b = v[c][1], d = v[c][2], e = v[c][3];
        float2 t0 = make_float2(a.x + d.x, a.y + d.y), t1 = make_float2(a.x - d.x, a.y - d.y);
        float2 t2 = make_float2(b.x + e.x, b.y + e.y), t3 = make_float2(b.x - e.x, b.y - e.y);
        // multiply t3 by sign * i
        float2 t3i = (sign < 0) ? make_float2(t3.y, -t3.x) : make_float2(-t3.y, t3.x);
        v[c][0] = make_float2(t0.x + t2.x, t0.y + t2.y);
        v[c][2] = make_float2(t0.x - t2.x, t0.y - t2.y);
        v[c][1] = make_float2(t1.x + t3i.x, t1.y + t3i.y);
        v[c][3] = make_float2(t1.x - t3i.x, t1.y - t3i.y);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < J; ++c) {
    const int j = tid + c * FFT_THREADS;
    if (j < NR) {
      int q, k;
      divmod_small(j, Ns, rns, q, k);
      const int base = q * Ns * R + k;
#pragma unroll
      for (int i = 0; i < R; ++i) s[base + i * Ns] = v[c][i];
    }
  }
  __syncthreads();
}

__global__ void __launch_bounds__(FFT_THREADS) k_fft_lines(FftLinesArgs a) {
  extern __shared__ float2 sbuf[];
  for (int64_t line = blockIdx.x; line < a.nlines; line += gridDim.x) {
    float2 *g = a.data + line * a.N;
    for (int i = threadIdx.x; i < a.N; i += FFT_THREADS) sbuf[i] = g[i];
    __syncthreads();
    int Ns = 1;
    for (int p = 0; p < a.npass; ++p) {
      int R = a.radix[p];
      if (R == 2) fft_pass<2>(sbuf, a.tw, a.N, Ns, a.sign);
      else if (R == 3) fft_pass<3>(sbuf, a.tw, a.N, Ns, a.sign);
      else fft_pass<4>(sbuf, a.tw, a.N, Ns, a.sign);
      Ns *= R;
    }
    for (int i = threadIdx.x; i < a.N; i += FFT_THREADS) g[i] = sbuf[i];
    __syncthreads();
  }
}

// Swap axis 0 with axis `ax` (1 or 2) of a 3-D complex array n0 x n1 x n2.
struct TransArgs {
  const float2 *in;
  float2 *out;
  int n0, n1, n2;
  int ax;
};

__global__ void __launch_bounds__(256) k_transpose(TransArgs a) {
  __shared__ float2 tile[32][33];
  const int na = a.ax == 1 ? a.n1 : a.n2;   // extent of the swapped axis
  const int nm = a.ax == 1 ? a.n2 : a.n1;   // the untouched axis
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int m = blockIdx.z; m < nm; m += gridDim.z) {
    for (int r = ty; r < 32; r += 8) {
      int i = i0 + tx, j = j0 + r;  // i along axis 0, j along swapped axis
      if (i < a.n0 && j < na) {
        int64_t idx = a.ax == 1 ? (i + (int64_t)a.n0 * (j + (int64_t)a.n1 * m))
                                : (i + (int64_t)a.n0 * (m + (int64_t)a.n1 * j));
        tile[r][tx] = a.in[idx];
      }
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
      int j = j0 + tx, i = i0 + r;
      if (i < a.n0 && j < na) {
        // out dims: ax==1 -> n1 x n0 x n2 ; ax==2 -> n2 x n1 x n0
        int64_t idx = a.ax == 1 ? (j + (int64_t)a.n1 * (i + (int64_t)a.n0 * m))
                                : (j + (int64_t)a.n2 * (m + (int64_t)a.n1 * i));
        a.out[idx] = tile[tx][r];
      }
    }
    __syncthreads();
  }
}

struct FftFillArgs {
  const void *in;
  float2 *x;
  int n[3];    // image extents
  int N[3];    // padded extents
  int lo[3];
  int style;
  int dt;
  float fill;
  int64_t total;
};

// Grid: x over axis 0 of the padded array, y over its rows (axes 1 and 2);
// a row's border maps are computed once per row.
__global__ void __launch_bounds__(256) k_fft_fill(FftFillArgs a) {
  const int u0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (u0 >= a.N[0]) return;
  const int s0 = bmap(a.style, u0 - a.lo[0], a.n[0]);
  const int nrows = a.N[1] * a.N[2];
  for (int row = blockIdx.y; row < nrows; row += gridDim.y) {
    const int u1 = row % a.N[1], u2 = row / a.N[1];
    const int s1 = bmap(a.style, u1 - a.lo[1], a.n[1]);
    const int s2 = bmap(a.style, u2 - a.lo[2], a.n[2]);
    const float v = (s0 | s1 | s2) < 0 ? a.fill
                                       : load_f(a.in, s0 + (int64_t)a.n[0] * (s1 + (int64_t)a.n[1] * s2), a.dt);
    a.x[(int64_t)row * a.N[0] + u0] = make_float2(v, 0.f);
  }
}

struct FftKernArgs {
  const float *taps;
  float2 *x;
  int kn[3], kf[3];
  int N[3];
  int64_t ktotal;
};

__global__ void __launch_bounds__(256) k_fft_kernel_scatter(FftKernArgs a) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < a.ktotal;
       q += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = q;
    int j0 = (int)(r % a.kn[0]); r /= a.kn[0];
    int j1 = (int)(r % a.kn[1]); r /= a.kn[1];
    int j2 = (int)r;
    int u0 = imod(j0 + a.kf[0], a.N[0]);
    int u1 = imod(j1 + a.kf[1], a.N[1]);
    int u2 = imod(j2 + a.kf[2], a.N[2]);
    int64_t idx = u0 + (int64_t)a.N[0] * (u1 + (int64_t)a.N[1] * u2);
    a.x[idx].x = a.taps[q];  // N > K on every axis, so taps never share a cell
  }
}

// x[q] *= conj(k[q % ktot]): the kernel spectrum repeats over the batch axes.
__global__ void __launch_bounds__(256) k_cmul_conj(float2 *x, const float2 *k, int64_t total, int64_t ktot) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * blockDim.x) {
    float2 a = x[q], b = k[q % ktot];
    x[q] = make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
  }
}

// Two-level transform of a line of n = P Q values: element (k1, q) of the
// P x Q intermediate (k1 fastest) times W_n^(sign k1 q).
__global__ void __launch_bounds__(256) k_fft_twiddle4(float2 *x, int P, int Q, int64_t total, const float2 *tw,
                                                       int sign) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int k1 = (int)(i % P), q = (int)((i / P) % Q);
    float2 w = tw[(int64_t)k1 * q];
    if (sign > 0) w.y = -w.y;
    const float2 a = x[i];
    x[i] = make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
  }
}

struct FftExtractArgs {
  const float2 *x;
  void *out;
  int o[3];
  int off[3];   // out_first + lo
  int N[3];
  float scale;
  int dt;
  int64_t total;
};

// Grid: x over axis 0 of the output, y over its rows (axes 1 and 2).
__global__ void __launch_bounds__(256) k_fft_extract(FftExtractArgs a) {
  const int c0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (c0 >= a.o[0]) return;
  const int nrows = a.o[1] * a.o[2];
  for (int row = blockIdx.y; row < nrows; row += gridDim.y) {
    const int c1 = row % a.o[1], c2 = row / a.o[1];
    const int64_t idx = (c0 + a.off[0]) + (int64_t)a.N[0] * ((c1 + a.off[1]) + (int64_t)a.N[1] * (c2 + a.off[2]));
    store_f(a.out, (int64_t)row * a.o[0] + c0, a.dt, a.x[idx].x * a.scale);
  }
}

// Row blocks of a (x blocks, rows) launch: about 4096 workgroups in all, each
// looping over its rows, rather than one short workgroup per row.
static unsigned rows_grid(int rows, int xb) {
  const int want = 4096 / (xb > 0 ? xb : 1);
  return (unsigned)(rows < want ? rows : (want > 0 ? want : 1));
}

static unsigned grid_for(int64_t total) {
  int64_t b = (total + 255) / 256;
  return (unsigned)(b > (1 << 20) ? (1 << 20) : (b < 1 ? 1 : b));
}

// Per-axis transform plan: a line of n <= FFT_LINE_MAX runs in one LDS pass;
// a longer one (n = P Q, both <= FFT_LINE_MAX) runs in two levels and leaves
// its spectrum in a permuted order (position k2 + Q k1 holds X[k1 + P k2]).
// The image and the kernel pass through the same transform, so their product
// is taken in that order and the inverse restores the natural one.
struct AxisPlan {
  int n = 1, P = 0, Q = 0;
  size_t tw = 0, twp = 0, twq = 0;  // offsets of W_n, W_P, W_Q in the twiddle area
};

static bool is23(int64_t x) {
  if (x < 1) return false;
  while (x % 2 == 0) x /= 2;
  while (x % 3 == 0) x /= 3;
  return x == 1;
}

// Twiddle entries of one axis; fills the split of a long line.
static size_t axis_plan(int64_t n, AxisPlan &ap, size_t at) {
  ap.n = (int)n;
  ap.tw = at;
  size_t cnt = (size_t)n;
  ap.P = ap.Q = 0;
  if (n > FFT_LINE_MAX) {
    // the most balanced split into two 2^a 3^b factors <= FFT_LINE_MAX
    int64_t best = 0;
    for (int64_t p = 1; p <= FFT_LINE_MAX; ++p)
      if (n % p == 0 && is23(p) && n / p <= FFT_LINE_MAX && (best == 0 || llabs(p * p - n) < llabs(best * best - n)))
        best = p;
    ap.P = (int)best;
    ap.Q = best ? (int)(n / best) : 0;
    ap.twp = at + cnt;
    cnt += (size_t)ap.P;
    ap.twq = at + cnt;
    cnt += (size_t)ap.Q;
  }
  return cnt;
}

static size_t twiddle_count(const int64_t *N, int nd) {
  size_t at = 0;
  for (int d = 0; d < nd; ++d) {
    AxisPlan ap;
    at += axis_plan(N[d], ap, at);
  }
  return at;
}

bool fft_axis_supported(int64_t n) {
  if (!is23(n)) return false;
  if (n <= FFT_LINE_MAX) return true;
  AxisPlan ap;
  axis_plan(n, ap, 0);
  return ap.P > 0;
}

// Workspace: image arrays A and B (the padded size), the kernel spectrum C
// (the transformed axes only), the twiddle tables; the caller adds the taps.
size_t fft_workspace(const int64_t *N, int nd) {
  int64_t tot = N[0] * N[1] * N[2], ktot = 1;
  for (int d = 0; d < nd; ++d) ktot *= N[d];
  return (size_t)(2 * tot + ktot) * sizeof(float2) + twiddle_count(N, nd) * sizeof(float2) + 256;
}

static void fft_lines(float2 *data, int n, int64_t nlines, const float2 *tw, int sign, hipStream_t st) {
  FftPlan plan;
  make_plan(n, plan);
  FftLinesArgs la;
  la.data = data;
  la.tw = tw;
  la.N = n;
  la.sign = sign;
  la.npass = plan.npass;
  for (int p = 0; p < plan.npass; ++p) la.radix[p] = plan.radix[p];
  la.nlines = nlines;
  const size_t lds = sizeof(float2) * (size_t)n;
  const unsigned blocks = (unsigned)(nlines < 65536 ? nlines : 65536);
  hipLaunchKernelGGL(k_fft_lines, dim3(blocks), dim3(FFT_THREADS), lds, st, la);
}

// Swaps axis 0 with axis ax (1 or 2) of the n0 x n1 x n2 array in into out.
static void transpose3(const float2 *in, float2 *out, int n0, int n1, int64_t n2, int ax, hipStream_t st) {
  TransArgs t;
  t.in = in;
  t.out = out;
  t.n0 = n0;
  t.n1 = n1;
  t.n2 = (int)n2;
  t.ax = ax;
  const int64_t na = ax == 1 ? n1 : n2, nm = ax == 1 ? n2 : n1;
  dim3 g((n0 + 31) / 32, (unsigned)((na + 31) / 32), (unsigned)(nm < 65535 ? nm : 65535));
  hipLaunchKernelGGL(k_transpose, g, dim3(256), 0, st, t);
}

// nlines contiguous lines of n values, in place in data (scratch: a buffer of
// the same size).  sign -1: forward, natural -> plan order; +1: inverse,
// plan order -> natural, unnormalised.
static void line_fft(float2 *data, float2 *scratch, const AxisPlan &ap, int64_t nlines, const float2 *twbase,
                     int sign, hipStream_t st) {
  if (ap.n <= FFT_LINE_MAX) {
    fft_lines(data, ap.n, nlines, twbase + ap.tw, sign, st);
    return;
  }
  const int P = ap.P, Q = ap.Q;  // element q + Q p of a line
  const int64_t total = (int64_t)ap.n * nlines;
  const float2 *twn = twbase + ap.tw;
  if (sign < 0) {
    transpose3(data, scratch, Q, P, nlines, 1, st);         // index p + P q
    fft_lines(scratch, P, (int64_t)Q * nlines, twbase + ap.twp, -1, st);  // k1 + P q
    hipLaunchKernelGGL(k_fft_twiddle4, dim3(grid_for(total)), dim3(256), 0, st, scratch, P, Q, total, twn, -1);
    transpose3(scratch, data, P, Q, nlines, 1, st);         // q + Q k1
    fft_lines(data, Q, (int64_t)P * nlines, twbase + ap.twq, -1, st);     // k2 + Q k1
  } else {
    fft_lines(data, Q, (int64_t)P * nlines, twbase + ap.twq, +1, st);     // q + Q k1
    transpose3(data, scratch, Q, P, nlines, 1, st);         // k1 + P q
    hipLaunchKernelGGL(k_fft_twiddle4, dim3(grid_for(total)), dim3(256), 0, st, scratch, P, Q, total, twn, +1);
    fft_lines(scratch, P, (int64_t)Q * nlines, twbase + ap.twp, +1, st);  // p + P q
    transpose3(scratch, data, P, Q, nlines, 1, st);         // q + Q p
  }
}

// Dimensions of the n3 array with axes 0 and ax swapped.
static void swapped(const int *n3, int ax, int *s3) {
  s3[0] = n3[ax];
  s3[1] = ax == 1 ? n3[0] : n3[1];
  s3[2] = ax == 2 ? n3[0] : n3[2];
}

// Transform along axis ax of the n3 array in x.  ax == 0: in place (tmp is
// scratch).  Otherwise x is transposed (axis 0 <-> ax) into tmp and tmp's
// lines are transformed (x is then scratch); with `keep` the result stays in
// tmp in that swapped layout, else it is transposed back into x.
static void fft_axis(float2 *x, float2 *tmp, const int *n3, int ax, const AxisPlan &ap, const float2 *twbase,
                     int sign, bool keep, hipStream_t st) {
  const int64_t total = (int64_t)n3[0] * n3[1] * n3[2];
  if (ax == 0) {
    line_fft(x, tmp, ap, total / n3[0], twbase, sign, st);
    return;
  }
  transpose3(x, tmp, n3[0], n3[1], n3[2], ax, st);
  line_fft(tmp, x, ap, total / n3[ax], twbase, sign, st);
  if (!keep) {
    int s3[3];
    swapped(n3, ax, s3);
    transpose3(tmp, x, s3[0], s3[1], s3[2], ax, st);
  }
}

// Forward transform of axes 0 .. nd-1 of src (further axes are batch axes);
// the last transformed axis (if not 0) is left in its swapped layout in tmp.
// Returns the buffer holding the spectrum.
static float2 *fft_forward_keep(float2 *src, float2 *tmp, float2 *scratch0, const int *n3, int nd,
                                const AxisPlan *ap, const float2 *twbase, hipStream_t st) {
  for (int ax = 0; ax < nd; ++ax)
    fft_axis(src, ax == 0 ? scratch0 : tmp, n3, ax, ap[ax], twbase, -1, ax == nd - 1, st);
  return nd > 1 ? tmp : src;
}

// Inverse of fft_forward_keep: spec holds the spectrum in that layout, other
// is the second buffer; returns the buffer holding the (unscaled) result in
// the natural layout.
static float2 *fft_inverse_keep(float2 *spec, float2 *other, const int *n3, int nd, const AxisPlan *ap,
                                const float2 *twbase, hipStream_t st) {
  const int64_t total = (int64_t)n3[0] * n3[1] * n3[2];
  if (nd == 1) {
    line_fft(spec, other, ap[0], total / n3[0], twbase, +1, st);
    return spec;
  }
  const int last = nd - 1;
  line_fft(spec, other, ap[last], total / n3[last], twbase, +1, st);
  int s3[3];
  swapped(n3, last, s3);
  transpose3(spec, other, s3[0], s3[1], s3[2], last, st);
  for (int ax = 0; ax < last; ++ax) fft_axis(other, spec, n3, ax, ap[ax], twbase, +1, false, st);
  return other;
}

hipError_t launch_fft_filter(const Dense4 &in, Dense4 &out, const int64_t *out_first, const DenseKernel &k,
                             const int64_t *Npad, int nd, const int64_t *lo, int style, float fill, void *work,
                             bool have_kernel, hipStream_t st) {
  // Axes 0 .. nd-1 are transformed (each N a product of 2s and 3s); axes nd..2
  // are batch axes with N = n.  Work buffers A and B hold the padded image
  // (N0 N1 N2 complex values), C the kernel spectrum over the transformed axes,
  // then the twiddle tables and the kernel taps.  The kernel's spectrum is left
  // in C and reused by the next call with the same kernel and sizes
  // (have_kernel); the image uses A and B only.
  int N[3] = {(int)Npad[0], (int)Npad[1], (int)Npad[2]};
  int n3[3] = {N[0], N[1], N[2]};
  int nk3[3] = {N[0], nd > 1 ? N[1] : 1, nd > 2 ? N[2] : 1};
  const int64_t tot = (int64_t)N[0] * N[1] * N[2];
  const int64_t ktot = (int64_t)nk3[0] * nk3[1] * nk3[2];
  float2 *A = (float2 *)work;
  float2 *B = A + tot;
  float2 *Cb = B + tot;
  float2 *twbase = Cb + ktot;
  AxisPlan ap[3];
  size_t ntw = 0;
  for (int d = 0; d < nd; ++d) ntw += axis_plan(N[d], ap[d], ntw);
  float *taps_dev = (float *)(twbase + ntw);
  hipError_t e;
  if (!have_kernel) {
    // twiddle tables (computed in FP64 on the host) and the kernel taps
    std::vector<float2> h(ntw);
    for (int d = 0; d < nd; ++d) {
      const size_t offs[3] = {ap[d].tw, ap[d].twp, ap[d].twq};
      const int lens[3] = {ap[d].n, ap[d].P, ap[d].Q};
      for (int w = 0; w < 3; ++w)
        for (int m = 0; m < lens[w]; ++m) {
          double ang = -2.0 * M_PI * (double)m / (double)lens[w];
          h[offs[w] + m] = make_float2((float)cos(ang), (float)sin(ang));
        }
    }
    if ((e = hipMemcpyAsync(twbase, h.data(), sizeof(float2) * ntw, hipMemcpyHostToDevice, st)) != hipSuccess)
      return e;
    if ((e = hipMemcpyAsync(taps_dev, k.taps.data(), sizeof(float) * k.taps.size(), hipMemcpyHostToDevice, st)) !=
        hipSuccess)
      return e;
    // the host buffers must outlive their copies
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    // the kernel, wrapped into the padded array over the transformed axes
    // (FFTView), starts in B (nd > 1) or C (nd == 1) so that its spectrum ends in C
    float2 *ks = nd > 1 ? B : Cb;
    if ((e = hipMemsetAsync(ks, 0, sizeof(float2) * ktot, st)) != hipSuccess) return e;
    FftKernArgs ka;
    ka.taps = taps_dev;
    ka.x = ks;
    ka.ktotal = 1;
    for (int d = 0; d < 3; ++d) {
      ka.kn[d] = d < k.ndim ? k.n[d] : 1;
      ka.kf[d] = d < k.ndim ? k.first[d] : 0;
      ka.N[d] = nk3[d];
      ka.ktotal *= ka.kn[d];
    }
    hipLaunchKernelGGL(k_fft_kernel_scatter, dim3(grid_for(ka.ktotal)), dim3(256), 0, st, ka);
    fft_forward_keep(ks, Cb, nd > 1 ? Cb : A, nk3, nd, ap, twbase, st);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  FftFillArgs fa;
  fa.in = in.data;
  fa.x = A;
  for (int d = 0; d < 3; ++d) {
    fa.n[d] = (int)in.n[d];
    fa.N[d] = N[d];
    fa.lo[d] = (int)lo[d];
  }
  fa.style = style;
  fa.dt = in.dtype;
  fa.fill = fill;
  fa.total = tot;
  {
    const int rows = N[1] * N[2], xb = (N[0] + 255) / 256;
    hipLaunchKernelGGL(k_fft_fill, dim3(xb, rows_grid(rows, xb)), dim3(256), 0, st, fa);
  }
  float2 *spec = fft_forward_keep(A, B, B, n3, nd, ap, twbase, st);
  hipLaunchKernelGGL(k_cmul_conj, dim3(grid_for(tot)), dim3(256), 0, st, spec, (const float2 *)Cb, tot, ktot);
  float2 *res = fft_inverse_keep(spec, spec == A ? B : A, n3, nd, ap, twbase, st);
  FftExtractArgs xa;
  xa.x = res;
  xa.out = out.data;
  for (int d = 0; d < 3; ++d) {
    xa.o[d] = (int)out.n[d];
    xa.off[d] = (int)(out_first[d] + lo[d]);
    xa.N[d] = N[d];
  }
  xa.scale = (float)(1.0 / (double)ktot);
  xa.dt = out.dtype;
  xa.total = out.count();
  if (xa.total > 0) {
    const int rows = xa.o[1] * xa.o[2], xb = (xa.o[0] + 255) / 256;
    hipLaunchKernelGGL(k_fft_extract, dim3(xb, rows_grid(rows, xb)), dim3(256), 0, st, xa);
  }
  return hipGetLastError();
}

}  // namespace ifm
