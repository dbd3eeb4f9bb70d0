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

__global__ void __launch_bounds__(256) k_cmul_conj(float2 *x, const float2 *k, int64_t total) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * blockDim.x) {
    float2 a = x[q], b = k[q];
    x[q] = make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
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

// Twiddle tables live in the workspace after the three complex buffers.
size_t fft_workspace(const int64_t *N, int nd) {
  int64_t tot = 1;
  size_t tw = 0;
  for (int d = 0; d < nd; ++d) {
    tot *= N[d];
    tw += (size_t)N[d];
  }
  return (size_t)tot * sizeof(float2) * 3 + tw * sizeof(float2) + 256;  // + kernel taps (caller)
}

// Transforms along axis `ax` of the n3 array in x.  ax == 0: in place.
// Otherwise x is transposed (axis 0 <-> ax) into tmp and tmp's lines are
// transformed; with `keep` the result stays in tmp in that swapped layout,
// else it is transposed back into x.
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
static void transpose3(const float2 *in, float2 *out, int n0, int n1, int n2, int ax, hipStream_t st) {
  TransArgs t;
  t.in = in;
  t.out = out;
  t.n0 = n0;
  t.n1 = n1;
  t.n2 = n2;
  t.ax = ax;
  const int na = ax == 1 ? n1 : n2, nm = ax == 1 ? n2 : n1;
  dim3 g((n0 + 31) / 32, (na + 31) / 32, nm < 65535 ? nm : 65535);
  hipLaunchKernelGGL(k_transpose, g, dim3(256), 0, st, t);
}

// Dimensions of the n3 array with axes 0 and ax swapped.
static void swapped(const int *n3, int ax, int *s3) {
  s3[0] = n3[ax];
  s3[1] = ax == 1 ? n3[0] : n3[1];
  s3[2] = ax == 2 ? n3[0] : n3[2];
}

static void fft_axis(float2 *x, float2 *tmp, const int *n3, int ax, const float2 *tw, int sign, bool keep,
                     hipStream_t st) {
  const int64_t total = (int64_t)n3[0] * n3[1] * n3[2];
  if (ax == 0) {
    fft_lines(x, n3[0], total / n3[0], tw, sign, st);
    return;
  }
  transpose3(x, tmp, n3[0], n3[1], n3[2], ax, st);
  fft_lines(tmp, n3[ax], total / n3[ax], tw, sign, st);
  if (!keep) {
    int s3[3];
    swapped(n3, ax, s3);
    transpose3(tmp, x, s3[0], s3[1], s3[2], ax, st);
  }
}

// Forward transform of every axis of src; the last axis (if not 0) is left
// in its swapped layout in tmp.  Returns the buffer holding the spectrum.
static float2 *fft_forward_keep(float2 *src, float2 *tmp, const int *n3, int nd, float2 *const *tw, hipStream_t st) {
  for (int ax = 0; ax < nd; ++ax) fft_axis(src, tmp, n3, ax, tw[ax], -1, ax == nd - 1, st);
  return nd > 1 ? tmp : src;
}

// Inverse of fft_forward_keep: spec holds the spectrum in that layout, other
// is the second buffer; returns the buffer holding the (unscaled) result in
// the natural layout.
static float2 *fft_inverse_keep(float2 *spec, float2 *other, const int *n3, int nd, float2 *const *tw,
                                hipStream_t st) {
  if (nd == 1) {
    fft_lines(spec, n3[0], (int64_t)n3[0] * n3[1] * n3[2] / n3[0], tw[0], +1, st);
    return spec;
  }
  const int last = nd - 1;
  const int64_t total = (int64_t)n3[0] * n3[1] * n3[2];
  fft_lines(spec, n3[last], total / n3[last], tw[last], +1, st);
  int s3[3];
  swapped(n3, last, s3);
  transpose3(spec, other, s3[0], s3[1], s3[2], last, st);
  for (int ax = 0; ax < last; ++ax) fft_axis(other, spec, n3, ax, tw[ax], +1, false, st);
  return other;
}

hipError_t launch_fft_filter(const Dense4 &in, Dense4 &out, const int64_t *out_first, const DenseKernel &k,
                             const int64_t *Npad, int nd, const int64_t *lo, int style, float fill, void *work,
                             bool have_kernel, hipStream_t st) {
  // Axes 0 .. nd-1 are transformed (each N a product of 2s and 3s); further
  // axes are batch axes with N = n.  Work buffers A, B, C of N0 N1 N2 complex
  // values, then the twiddle tables, then the kernel taps.  The kernel's
  // spectrum is left in C and reused by the next call with the same kernel
  // and sizes (have_kernel); the image uses A and B only.
  int N[3] = {(int)Npad[0], (int)Npad[1], (int)Npad[2]};
  int n3[3] = {N[0], nd > 1 ? N[1] : 1, nd > 2 ? N[2] : 1};
  int64_t tot = (int64_t)N[0] * N[1] * N[2];
  float2 *A = (float2 *)work;
  float2 *B = A + tot;
  float2 *Cb = B + tot;
  float2 *twbase = Cb + tot;
  float2 *tw[3];
  size_t ntw = 0;
  for (int d = 0; d < nd; ++d) {
    tw[d] = twbase + ntw;
    ntw += (size_t)N[d];
  }
  float *taps_dev = (float *)(twbase + ntw);
  hipError_t e;
  if (!have_kernel) {
    // twiddle tables (computed in FP64 on the host) and the kernel taps
    std::vector<float2> h(ntw);
    size_t at = 0;
    for (int d = 0; d < nd; ++d) {
      for (int m = 0; m < N[d]; ++m) {
        double ang = -2.0 * M_PI * (double)m / (double)N[d];
        h[at + m] = make_float2((float)cos(ang), (float)sin(ang));
      }
      at += N[d];
    }
    if ((e = hipMemcpyAsync(twbase, h.data(), sizeof(float2) * ntw, hipMemcpyHostToDevice, st)) != hipSuccess)
      return e;
    if ((e = hipMemcpyAsync(taps_dev, k.taps.data(), sizeof(float) * k.taps.size(), hipMemcpyHostToDevice, st)) !=
        hipSuccess)
      return e;
    // the host buffers must outlive their copies
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    // the kernel, wrapped into the padded array (FFTView), starts in B (nd > 1)
    // or C (nd == 1) so that its spectrum ends in C
    float2 *ks = nd > 1 ? B : Cb;
    if ((e = hipMemsetAsync(ks, 0, sizeof(float2) * tot, st)) != hipSuccess) return e;
    FftKernArgs ka;
    ka.taps = taps_dev;
    ka.x = ks;
    ka.ktotal = 1;
    for (int d = 0; d < 3; ++d) {
      ka.kn[d] = d < k.ndim ? k.n[d] : 1;
      ka.kf[d] = d < k.ndim ? k.first[d] : 0;
      ka.N[d] = N[d];
      ka.ktotal *= ka.kn[d];
    }
    hipLaunchKernelGGL(k_fft_kernel_scatter, dim3(grid_for(ka.ktotal)), dim3(256), 0, st, ka);
    fft_forward_keep(ks, Cb, n3, nd, tw, st);
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
  float2 *spec = fft_forward_keep(A, B, n3, nd, tw, st);
  hipLaunchKernelGGL(k_cmul_conj, dim3(grid_for(tot)), dim3(256), 0, st, spec, (const float2 *)Cb, tot);
  float2 *res = fft_inverse_keep(spec, spec == A ? B : A, n3, nd, tw, st);
  FftExtractArgs xa;
  xa.x = res;
  xa.out = out.data;
  for (int d = 0; d < 3; ++d) {
    xa.o[d] = (int)out.n[d];
    xa.off[d] = (int)(out_first[d] + lo[d]);
    xa.N[d] = N[d];
  }
  xa.scale = (float)(1.0 / (double)tot);
  xa.dt = out.dtype;
  xa.total = out.count();
  if (xa.total > 0) {
    const int rows = xa.o[1] * xa.o[2], xb = (xa.o[0] + 255) / 256;
    hipLaunchKernelGGL(k_fft_extract, dim3(xb, rows_grid(rows, xb)), dim3(256), 0, st, xa);
  }
  return hipGetLastError();
}

}  // namespace ifm
