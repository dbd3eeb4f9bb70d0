// k_iir.hip — Triggs-Sdika IIR filtering along one axis in FP64 (gfx950).
//
// Follows _imfilter_dim! (src/imfilter.jl:964-1024) with leftborder!
// (:1027-1048) and rightborder! (:1051-1092): forward recursion with the
// steady-state left initialisation, right-edge initialisation by the 3x3
// matrix M (Triggs & Sdika Eqs. 14-15), backward recursion and scaling.
// Lines of up to 512 elements run k_iir_rows2 (contiguous axis, through an
// LDS transpose) or k_iir_seg2 (strided axes): whole lines in registers, the
// axis split into segments, two segments per wave.  Longer lines run
// k_iir_rows (contiguous axis, LDS-staged chunks) or k_iir_strided (one
// thread per line).  Lanes sit on adjacent lines, which are adjacent in memory for the
// strided axes (coalesced).
#include "ifm_internal.h"

namespace ifm {

struct IirArgs {
  const void *src;
  double *dst;        // FP64 stage result (forward values, then the output)
  int dt_src;
  int n[MAXD];
  int axis;
  int64_t stride;     // element stride of the filtered axis
  int64_t lines;
  double a[3], b[3], scale, M[9];
  int isfill;
  double fill;
};

// The recursions are sequential along a line, so the inputs of the next
// IC steps are loaded into registers before those steps run (a step never
// waits for its own load).  Reading ahead is safe in place: the forward pass
// only writes positions it has already read, the backward pass likewise.
constexpr int IC = 16;

// Source element as FP64; DT < 0: element type read at run time.
template <int DT>
__device__ __forceinline__ double iir_ld(const IirArgs &p, int64_t i) {
  if constexpr (DT == IFM_F32) return (double)((const float *)p.src)[i];
  else if constexpr (DT == IFM_F64) return ((const double *)p.src)[i];
  else return load_d(p.src, i, (int)p.dt_src);
}

// Line accessor: element i of line `base` with stride `st` (ST1: st == 1, so
// a chunk is contiguous and the loads vectorise).
template <bool ST1, int DT>
__global__ void __launch_bounds__(256) k_iir_strided(IirArgs p) {
  const int64_t line = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (line >= p.lines) return;
  const int n = p.n[p.axis];
  const int64_t st = ST1 ? 1 : p.stride;
  const int64_t pre = line % st, post = line / st;
  const int64_t base = pre + post * st * n;
  const double a1 = p.a[0], a2 = p.a[1], a3 = p.a[2];
  const double b1 = p.b[0], b2 = p.b[1], b3 = p.b[2];
  const double asum = a1 + a2 + a3, bsum = b1 + b2 + b3;
  const double iminus = p.isfill ? p.fill : iir_ld<DT>(p, base);
  const double iplus = p.isfill ? p.fill : iir_ld<DT>(p, base + (int64_t)(n - 1) * st);
  // leftborder!
  const double uminus = iminus / (1 - asum);
  double y0 = iir_ld<DT>(p, base) + a1 * uminus + a2 * uminus + a3 * uminus;
  p.dst[base] = y0;
  double y1 = iir_ld<DT>(p, base + st) + a1 * y0 + a2 * uminus + a3 * uminus;
  p.dst[base + st] = y1;
  double y2 = iir_ld<DT>(p, base + 2 * st) + a1 * y1 + a2 * y0 + a3 * uminus;
  p.dst[base + 2 * st] = y2;
  // forward, omitting the final point
  double ym3 = y0, ym2 = y1, ym1 = y2;
  int i = 3;
  for (; i + IC - 1 <= n - 2; i += IC) {
    double x[IC];
#pragma unroll
    for (int j = 0; j < IC; ++j) x[j] = iir_ld<DT>(p, base + (int64_t)(i + j) * st);
#pragma unroll
    for (int j = 0; j < IC; ++j) {
      const double y = x[j] + a1 * ym1 + a2 * ym2 + a3 * ym3;
      p.dst[base + (int64_t)(i + j) * st] = y;
      ym3 = ym2;
      ym2 = ym1;
      ym1 = y;
    }
  }
  for (; i <= n - 2; ++i) {
    double y = iir_ld<DT>(p, base + (int64_t)i * st) + a1 * ym1 + a2 * ym2 + a3 * ym3;
    p.dst[base + (int64_t)i * st] = y;
    ym3 = ym2;
    ym2 = ym1;
    ym1 = y;
  }
  // rightborder!: the last forward value, then v at and beyond the edge
  const double ylast = iir_ld<DT>(p, base + (int64_t)(n - 1) * st) + a1 * ym1 + a2 * ym2 + a3 * ym3;
  const double uplus = iplus / (1 - asum);
  const double vplus = uplus / (1 - bsum);
  const double d0 = ylast - uplus, d1 = ym1 - uplus, d2 = ym2 - uplus;
  const double v0 = p.M[0] * d0 + p.M[1] * d1 + p.M[2] * d2 + vplus;
  const double v1 = p.M[3] * d0 + p.M[4] * d1 + p.M[5] * d2 + vplus;
  const double v2 = p.M[6] * d0 + p.M[7] * d1 + p.M[8] * d2 + vplus;
  // out[n-1] = v0; out[n-2] = y[n-2] + b1*out[n-1] + b2*v1 + b3*v2;
  // out[n-3] = y[n-3] + b1*out[n-2] + b2*out[n-1] + b3*v1
  const double z2 = v0;
  const double z1 = ym1 + b1 * z2 + b2 * v1 + b3 * v2;
  const double z0 = ym2 + b1 * z1 + b2 * z2 + b3 * v1;
  const double sc = p.scale;
  p.dst[base + (int64_t)(n - 1) * st] = z2 * sc;
  p.dst[base + (int64_t)(n - 2) * st] = z1 * sc;
  p.dst[base + (int64_t)(n - 3) * st] = z0 * sc;
  // backward
  double zp1 = z0, zp2 = z1, zp3 = z2;
  int k = n - 4;
  for (; k - (IC - 1) >= 0; k -= IC) {
    double y[IC];
#pragma unroll
    for (int j = 0; j < IC; ++j) y[j] = p.dst[base + (int64_t)(k - j) * st];
#pragma unroll
    for (int j = 0; j < IC; ++j) {
      const double z = y[j] + b1 * zp1 + b2 * zp2 + b3 * zp3;
      p.dst[base + (int64_t)(k - j) * st] = z * sc;
      zp3 = zp2;
      zp2 = zp1;
      zp1 = z;
    }
  }
  for (; k >= 0; --k) {
    double z = p.dst[base + (int64_t)k * st] + b1 * zp1 + b2 * zp2 + b3 * zp3;
    p.dst[base + (int64_t)k * st] = z * sc;
    zp3 = zp2;
    zp2 = zp1;
    zp1 = z;
  }
}

// Lines along the contiguous axis.  One thread per line would read 64 lines
// per wave instruction, n elements apart; instead a workgroup owns 256 lines
// and walks them in chunks of IR_C elements staged through LDS: the chunk is
// loaded and stored with consecutive lanes on consecutive elements
// (coalesced), and each thread runs its line's recursion over the chunk from
// LDS.  The next chunk's loads are issued before the current chunk is
// filtered.  The arithmetic per element is that of k_iir_strided.
constexpr int IR_C = 16;                   // chunk length
constexpr int IR_L = 256;                  // lines per workgroup
constexpr int IR_P = IR_C + 1;             // LDS pitch (doubles)
constexpr int IR_E = IR_L * IR_C / 256;    // staged elements per thread

template <int DT>
__global__ void __launch_bounds__(256, 2) k_iir_rows(IirArgs p) {
  extern __shared__ double tile[];  // IR_L x IR_P
  const int t = threadIdx.x;
  const int64_t line0 = (int64_t)blockIdx.x * IR_L;
  const int n = p.n[0];
  const int nl = (int)min((int64_t)IR_L, p.lines - line0);
  const int nch = (n + IR_C - 1) / IR_C;
  const double a1 = p.a[0], a2 = p.a[1], a3 = p.a[2];
  const double b1 = p.b[0], b2 = p.b[1], b3 = p.b[2];
  const double asum = a1 + a2 + a3, bsum = b1 + b2 + b3;
  double reg[IR_E];
  // staged element j of the thread: line (t + 256 j) / IR_C, offset (t + 256 j) % IR_C
  auto fetch = [&](const void *src, int dtsrc, int c0, bool typed) {
#pragma unroll
    for (int j = 0; j < IR_E; ++j) {
      const int q = t + 256 * j, l = q / IR_C, e = q % IR_C;
      const int i = c0 + e;
      if (l < nl && i < n) {
        const int64_t g = (line0 + l) * n + i;
        reg[j] = typed ? iir_ld<DT>(p, g) : ((const double *)src)[g];
      }
    }
    (void)dtsrc;
  };
  auto put = [&]() {
#pragma unroll
    for (int j = 0; j < IR_E; ++j) {
      const int q = t + 256 * j;
      tile[(q / IR_C) * IR_P + q % IR_C] = reg[j];
    }
  };
  auto store = [&](int c0) {
#pragma unroll
    for (int j = 0; j < IR_E; ++j) {
      const int q = t + 256 * j, l = q / IR_C, e = q % IR_C;
      const int i = c0 + e;
      if (l < nl && i < n) p.dst[(line0 + l) * n + i] = tile[l * IR_P + e];
    }
  };
  double *row = tile + t * IR_P;
  // forward: y[i] = x[i] + a1 y[i-1] + a2 y[i-2] + a3 y[i-3], with y[-k] = u-
  double um = 0.0, ym1 = 0.0, ym2 = 0.0, ym3 = 0.0, iplus = p.fill;
  fetch(nullptr, 0, 0, true);
  for (int c = 0; c < nch; ++c) {
    const int c0 = c * IR_C;
    put();
    __syncthreads();
    if (c + 1 < nch) fetch(nullptr, 0, c0 + IR_C, true);
    if (t < nl) {
      const int e1 = min(IR_C, n - c0);
      if (c == 0) {
        const double iminus = p.isfill ? p.fill : row[0];
        um = iminus / (1 - asum);
      }
      if (c0 + e1 == n && !p.isfill) iplus = row[e1 - 1];
      for (int e = 0; e < e1; ++e) {
        const int i = c0 + e;  // wave-uniform
        const double x = row[e];
        double y;
        if (i == 0) y = x + a1 * um + a2 * um + a3 * um;
        else if (i == 1) y = x + a1 * ym1 + a2 * um + a3 * um;
        else if (i == 2) y = x + a1 * ym1 + a2 * ym2 + a3 * um;
        else y = x + a1 * ym1 + a2 * ym2 + a3 * ym3;
        row[e] = y;
        ym3 = ym2;
        ym2 = ym1;
        ym1 = y;
      }
    }
    __syncthreads();
    store(c0);
    __syncthreads();
  }
  // right edge (ym1 = y[n-1], ym2 = y[n-2], ym3 = y[n-3]), then backward:
  // z[k] = y[k] + b1 z[k+1] + b2 z[k+2] + b3 z[k+3], the last three from M
  const double sc = p.scale;
  const double uplus = iplus / (1 - asum);
  const double vplus = uplus / (1 - bsum);
  const double d0 = ym1 - uplus, d1 = ym2 - uplus, d2 = ym3 - uplus;
  const double v0 = p.M[0] * d0 + p.M[1] * d1 + p.M[2] * d2 + vplus;
  const double v1 = p.M[3] * d0 + p.M[4] * d1 + p.M[5] * d2 + vplus;
  const double v2 = p.M[6] * d0 + p.M[7] * d1 + p.M[8] * d2 + vplus;
  const double z2 = v0;
  const double z1 = ym2 + b1 * z2 + b2 * v1 + b3 * v2;
  const double z0 = ym3 + b1 * z1 + b2 * z2 + b3 * v1;
  double zp1 = 0.0, zp2 = 0.0, zp3 = 0.0;
  fetch(p.dst, IFM_F64, (nch - 1) * IR_C, false);
  for (int c = nch - 1; c >= 0; --c) {
    const int c0 = c * IR_C;
    put();
    __syncthreads();
    if (c > 0) fetch(p.dst, IFM_F64, c0 - IR_C, false);
    if (t < nl) {
      const int e1 = min(IR_C, n - c0);
      for (int e = e1 - 1; e >= 0; --e) {
        const int k = c0 + e;  // wave-uniform
        double z;
        if (k == n - 1) z = z2;
        else if (k == n - 2) z = z1;
        else if (k == n - 3) z = z0;
        else z = row[e] + b1 * zp1 + b2 * zp2 + b3 * zp3;
        row[e] = z * sc;
        zp3 = zp2;
        zp2 = zp1;
        zp1 = z;
      }
    }
    __syncthreads();
    store(c0);
    __syncthreads();
  }
}

// Lines of up to 512 elements, held in registers.  The axis is cut into nw
// segments of at most IS_S elements, each run by one half-wave lane, so
// every element is read once and written once (the one-thread-per-line form
// writes the forward pass out and reads it back).  The recursions are linear:
// a segment's values are its recursion from zero state plus the homogeneous
// response to the state entering it.  Each lane runs its segment from zero
// state; the lanes exchange their end states through LDS; each composes the
// state entering its segment from those with the 3x3 transfer matrix of a
// full segment, and adds sum_c h_c[j] s_c.  The first segment's forward pass
// and the last segment's backward pass start from the reference's border
// initialisation directly (the um / M start of k_iir_strided).
constexpr int IS_S = 32;     // most elements per segment (doubles in registers per lane)
constexpr int IS_MAXW = 16;  // most segments: n <= 512

struct IirSegTables {
  double h[3][IS_S];  // forward homogeneous responses: state (y[-1], y[-2], y[-3]) = e_c
  double g[3][IS_S];  // backward, by distance from the segment's right end
  int seg;            // segment length (the last may be shorter, >= 3)
  int nw;             // segments
};

// LDS (doubles): tables h, g (6 IS_S), then forward and backward end states
// [segment][component][line] (2 x nw x 3 x 32), then (lines along the
// contiguous axis) one transpose buffer per wave of 16 rows x IS_TP doubles,
// which also holds 32 rows x IS_TP floats.
constexpr int IS_TP = IS_S + 1;    // transpose buffer pitch (elements): conflict-free both ways

// Two segments per wave: a workgroup owns 32 adjacent lines and lane l of
// wave w runs segment 2w + (l >> 5) of line l & 31.  Segments of at most 32
// elements keep a lane's share of the line in 64 registers, so four waves
// share each SIMD (two workgroups of up to 8 waves per CU) and one's loads
// and stores overlap the others' recursions; 64-element segments (two waves
// per SIMD) measured 942 against 767 us on 512 x 512 x 256.
//  * k_iir_seg2  strided lines: each half-wave reads one contiguous run of 32
//                lines per element.
//  * k_iir_rows2 contiguous lines: a wave's two segments are one run of up
//                to 64 elements of a line, loaded and stored by its lanes;
//                they pass through a per-wave LDS transpose (rows = (line,
//                half) pairs), 16 lines per round for FP32 input, 8 for FP64
//                input and for the output.
constexpr int IS_SPW = 2;                 // segments per wave (4, with 16 lines per workgroup, measured slower)
constexpr int IS2_L = 64 / IS_SPW;        // lines per workgroup
constexpr int IS2_TB = 16 * IS_TP;        // transpose buffer per wave (doubles)
__host__ __device__ inline size_t iir_seg2_lds_doubles(int nw, bool rows) {
  return 6 * IS_S + 2 * (size_t)nw * 3 * IS2_L + (rows ? (size_t)((nw + IS_SPW - 1) / IS_SPW) * IS2_TB : 0);
}

struct Seg2 {
  int ln, sg, L;
  bool has, first, last;
};

__device__ __forceinline__ Seg2 seg2_lane(const IirSegTables &tb, int n) {
  Seg2 g;
  const int lane = threadIdx.x & 63;
  g.ln = lane % IS2_L;
  g.sg = IS_SPW * (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) + lane / IS2_L;
  g.has = g.sg < tb.nw;
  g.L = g.has ? min(tb.seg, n - g.sg * tb.seg) : 0;
  g.first = g.sg == 0;
  g.last = g.sg == tb.nw - 1;
  return g;
}

// The segmented recursion on v[0 .. L) of every lane: forward from zero state (the first segment from leftborder!),
// end states exchanged through LDS, the homogeneous response of the state
// entering the segment added; the same backward (the last segment from
// rightborder!).  Leaves v scaled by nothing; contains two barriers.  FULL:
// every lane of the wave holds a full segment of IS_S elements (the bounds
// are then compile-time constants and the per-element predicates vanish).
template <bool FULL>
__device__ __forceinline__ void seg2_filter(const IirArgs &p, const IirSegTables &tb, const Seg2 &g, double *sm,
                                            double (&v)[IS_S]) {
  double *hs = sm, *gs = sm + 3 * IS_S;
  double *stf = sm + 6 * IS_S, *stb = stf + tb.nw * 3 * IS2_L;
  const int nw = tb.nw, S = FULL ? IS_S : tb.seg, L = FULL ? IS_S : g.L, ln = g.ln, sg = g.sg;
  for (int i = threadIdx.x; i < 6 * IS_S; i += blockDim.x) sm[i] = i < 3 * IS_S ? tb.h[i / IS_S][i % IS_S]
                                                                                   : tb.g[i / IS_S - 3][i % IS_S];
  const double a1 = p.a[0], a2 = p.a[1], a3 = p.a[2];
  const double b1 = p.b[0], b2 = p.b[1], b3 = p.b[2];
  const double asum = a1 + a2 + a3, bsum = b1 + b2 + b3;
  double ym1 = 0.0, ym2 = 0.0, ym3 = 0.0, xlast = 0.0;
  if (g.first) {
    const double iminus = p.isfill ? p.fill : v[0];
    ym1 = ym2 = ym3 = iminus / (1 - asum);
  }
#pragma unroll
  for (int j = 0; j < IS_S; ++j) {
    if (j < L) {
      if (j == L - 1) xlast = v[j];
      const double y = v[j] + a1 * ym1 + a2 * ym2 + a3 * ym3;
      v[j] = y;
      ym3 = ym2;
      ym2 = ym1;
      ym1 = y;
    }
  }
  if (g.has) {
    stf[(sg * 3 + 0) * IS2_L + ln] = ym1;
    stf[(sg * 3 + 1) * IS2_L + ln] = ym2;
    stf[(sg * 3 + 2) * IS2_L + ln] = ym3;
  }
  __syncthreads();
  // state entering segment sg: T_0 = end of segment 0, T_u = Phi T_{u-1} + P_u
  // with Phi[r][c] = h_c[S - 1 - r]
  double e0 = ym1, e1 = ym2, e2 = ym3;  // true end state of this segment
  if (g.has && !g.first) {
    double t0 = stf[ln], t1 = stf[IS2_L + ln], t2 = stf[2 * IS2_L + ln];
    for (int u = 1; u < sg; ++u) {
      const double *q = stf + u * 3 * IS2_L;
      const double n0 = hs[S - 1] * t0 + hs[IS_S + S - 1] * t1 + hs[2 * IS_S + S - 1] * t2 + q[ln];
      const double n1 = hs[S - 2] * t0 + hs[IS_S + S - 2] * t1 + hs[2 * IS_S + S - 2] * t2 + q[IS2_L + ln];
      const double n2 = hs[S - 3] * t0 + hs[IS_S + S - 3] * t1 + hs[2 * IS_S + S - 3] * t2 + q[2 * IS2_L + ln];
      t0 = n0;
      t1 = n1;
      t2 = n2;
    }
#pragma unroll
    for (int j = 0; j < IS_S; ++j)
      if (j < L) v[j] += hs[j] * t0 + hs[IS_S + j] * t1 + hs[2 * IS_S + j] * t2;
    e0 = ym1 + hs[L - 1] * t0 + hs[IS_S + L - 1] * t1 + hs[2 * IS_S + L - 1] * t2;
    e1 = ym2 + hs[L - 2] * t0 + hs[IS_S + L - 2] * t1 + hs[2 * IS_S + L - 2] * t2;
    e2 = ym3 + hs[L - 3] * t0 + hs[IS_S + L - 3] * t1 + hs[2 * IS_S + L - 3] * t2;
  }
  double z0 = 0.0, z1 = 0.0, z2 = 0.0;
  if (g.last) {
    const double iplus = p.isfill ? p.fill : xlast;
    const double uplus = iplus / (1 - asum);
    const double vplus = uplus / (1 - bsum);
    const double d0 = e0 - uplus, d1 = e1 - uplus, d2 = e2 - uplus;
    const double m0 = p.M[0] * d0 + p.M[1] * d1 + p.M[2] * d2 + vplus;
    const double m1 = p.M[3] * d0 + p.M[4] * d1 + p.M[5] * d2 + vplus;
    const double m2 = p.M[6] * d0 + p.M[7] * d1 + p.M[8] * d2 + vplus;
    z2 = m0;
    z1 = e1 + b1 * z2 + b2 * m1 + b3 * m2;
    z0 = e2 + b1 * z1 + b2 * z2 + b3 * m1;
  }
  double zp1 = 0.0, zp2 = 0.0, zp3 = 0.0;
#pragma unroll
  for (int j = IS_S - 1; j >= 0; --j) {
    if (j < L) {
      double z;
      if (g.last && j == L - 1) z = z2;
      else if (g.last && j == L - 2) z = z1;
      else if (g.last && j == L - 3) z = z0;
      else z = v[j] + b1 * zp1 + b2 * zp2 + b3 * zp3;
      v[j] = z;
      zp3 = zp2;
      zp2 = zp1;
      zp1 = z;
    }
  }
  if (g.has) {
    stb[(sg * 3 + 0) * IS2_L + ln] = zp1;
    stb[(sg * 3 + 1) * IS2_L + ln] = zp2;
    stb[(sg * 3 + 2) * IS2_L + ln] = zp3;
  }
  __syncthreads();
  // state entering segment sg from the right: R_{nw-1} = left end of the last
  // segment, R_u = Psi R_{u+1} + Q_u with Psi[r][c] = g_c[S - 1 - r]
  if (g.has && !g.last) {
    const double *q0 = stb + (nw - 1) * 3 * IS2_L;
    double t0 = q0[ln], t1 = q0[IS2_L + ln], t2 = q0[2 * IS2_L + ln];
    for (int u = nw - 2; u > sg; --u) {
      const double *q = stb + u * 3 * IS2_L;
      const double n0 = gs[S - 1] * t0 + gs[IS_S + S - 1] * t1 + gs[2 * IS_S + S - 1] * t2 + q[ln];
      const double n1 = gs[S - 2] * t0 + gs[IS_S + S - 2] * t1 + gs[2 * IS_S + S - 2] * t2 + q[IS2_L + ln];
      const double n2 = gs[S - 3] * t0 + gs[IS_S + S - 3] * t1 + gs[2 * IS_S + S - 3] * t2 + q[2 * IS2_L + ln];
      t0 = n0;
      t1 = n1;
      t2 = n2;
    }
#pragma unroll
    for (int j = 0; j < IS_S; ++j)
      if (j < L) v[j] += gs[L - 1 - j] * t0 + gs[IS_S + L - 1 - j] * t1 + gs[2 * IS_S + L - 1 - j] * t2;
  }
}

template <int DT, bool FULL>
__device__ __forceinline__ void seg2_body(const IirArgs &p, const IirSegTables &tb, const Seg2 &g, double *sm) {
  const int n = p.n[p.axis];
  const int L = FULL ? IS_S : g.L;
  const int64_t line = (int64_t)blockIdx.x * IS2_L + g.ln;
  const bool ok = line < p.lines && g.has;
  const int64_t st = p.stride;
  const int64_t base = (line % st) + (line / st) * st * n + (int64_t)g.sg * tb.seg * st;
  double v[IS_S];
#pragma unroll
  for (int j = 0; j < IS_S; ++j)
    if (j < L && ok) v[j] = iir_ld<DT>(p, base + (int64_t)j * st);
  seg2_filter<FULL>(p, tb, g, sm, v);
  const double sc = p.scale;
  if (ok) {
#pragma unroll
    for (int j = 0; j < IS_S; ++j)
      if (j < L) p.dst[base + (int64_t)j * st] = v[j] * sc;
  }
}

template <int DT>
__global__ void __launch_bounds__(32 * IS_MAXW) __attribute__((amdgpu_waves_per_eu(4))) k_iir_seg2(IirArgs p, IirSegTables tb) {
  extern __shared__ double sm[];
  const Seg2 g = seg2_lane(tb, p.n[p.axis]);
  // a wave-uniform branch: both bodies pass the same two barriers.  Waves
  // whose segments all hold 32 elements (every wave when the line length is a
  // multiple of 32, e.g. 256 or 512) skip the per-element bounds: 766 -> 652
  // us on the 3-D Gaussian, at four waves per SIMD (128 registers, no spill)
  if (__all(g.L == IS_S)) seg2_body<DT, true>(p, tb, g, sm);
  else seg2_body<DT, false>(p, tb, g, sm);
}

template <int DT, bool FULL>
__device__ __forceinline__ void rows2_body(const IirArgs &p, const IirSegTables &tb, const Seg2 &g, double *sm) {
  using TE = typename std::conditional<DT == IFM_F32, float, double>::type;
  const int n = p.n[p.axis];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63, half = lane / IS2_L;
  const int S = FULL ? IS_S : tb.seg, nw = tb.nw, L = FULL ? IS_S : g.L;
  const int64_t line0 = (int64_t)blockIdx.x * IS2_L;
  double *tr = sm + iir_seg2_lds_doubles(nw, false) + w * IS2_TB;
  const TE *src = (const TE *)p.src;
  double v[IS_S];
  // wave w covers elements [e0, e0 + span) of every line (its two segments);
  // lane l holds element e0 + l, in transpose row 2 i + l / S, column l % S
  const int e0 = IS_SPW * w * S, span = min(IS_SPW * S, n - e0);
  const bool lok = lane < span;
  const int lh = lane >= S ? 1 : 0, lc = lane - lh * S;
  constexpr int RI = (DT == IFM_F32 ? 32 : 16) / IS_SPW;  // lines per round
  TE *tri = (TE *)tr;
  for (int r = 0; r < IS2_L / RI; ++r) {
    TE x[RI];
#pragma unroll
    for (int i = 0; i < RI; ++i) {
      const int64_t ln = line0 + r * RI + i;
      x[i] = (lok && ln < p.lines) ? src[ln * n + e0 + lane] : (TE)0;
    }
    if (lok) {
#pragma unroll
      for (int i = 0; i < RI; ++i) tri[(IS_SPW * i + lh) * IS_TP + lc] = x[i];
    }
    __syncthreads();
    if ((g.ln / RI) == r) {
      const int row = (g.ln - r * RI) * IS_SPW + half;
#pragma unroll
      for (int j = 0; j < IS_S; ++j)
        if (j < L) v[j] = (double)tri[row * IS_TP + j];
    }
    __syncthreads();
  }
  seg2_filter<FULL>(p, tb, g, sm, v);
  const double sc = p.scale;
  // stores: rounds of 8 lines (16 rows of doubles)
  constexpr int RO = 16 / IS_SPW;
  for (int r = 0; r < IS2_L / RO; ++r) {
    if ((g.ln / RO) == r) {
      const int row = (g.ln - r * RO) * IS_SPW + half;
#pragma unroll
      for (int j = 0; j < IS_S; ++j)
        if (j < L) tr[row * IS_TP + j] = v[j] * sc;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RO; ++i) {
      const int64_t ln = line0 + r * RO + i;
      if (lok && ln < p.lines) p.dst[ln * n + e0 + lane] = tr[(IS_SPW * i + lh) * IS_TP + lc];
    }
    __syncthreads();
  }
}

template <int DT>
__global__ void __launch_bounds__(32 * IS_MAXW) __attribute__((amdgpu_waves_per_eu(4))) k_iir_rows2(IirArgs p, IirSegTables tb) {
  static_assert(DT == IFM_F32 || DT == IFM_F64, "typed contiguous lines only");
  extern __shared__ double sm[];
  const Seg2 g = seg2_lane(tb, p.n[p.axis]);
  // a wave-uniform branch: both bodies pass the same barriers
  if (__all(g.L == IS_S)) rows2_body<DT, true>(p, tb, g, sm);
  else rows2_body<DT, false>(p, tb, g, sm);
}

// Segmentation of an axis of n elements: nw waves of seg elements (the last
// one shorter, at least 3); false if n is too long for one workgroup.
static bool iir_segments(int n, int *seg, int *nw) {
  const int w = (n + IS_S - 1) / IS_S;
  if (w > IS_MAXW || n < 4) return false;
  const int s = (n + w - 1) / w;
  if (n - (w - 1) * s < 3) return false;
  *seg = s;
  *nw = w;
  return true;
}

hipError_t launch_iir(const Dense4 &in, Dense4 &out, int axis, const IirCoef &c, int isfill, double fill,
                      hipStream_t st) {
  IirArgs p;
  p.src = in.data;
  p.dst = (double *)out.data;
  p.dt_src = in.dtype;
  int64_t stride = 1;
  for (int d = 0; d < MAXD; ++d) {
    p.n[d] = (int)in.n[d];
    if (d < axis) stride *= in.n[d];
  }
  p.axis = axis;
  p.stride = stride;
  p.lines = in.count() / in.n[axis];
  for (int i = 0; i < 3; ++i) {
    p.a[i] = c.a[i];
    p.b[i] = c.b[i];
  }
  p.scale = c.scale;
  for (int i = 0; i < 9; ++i) p.M[i] = c.M[i];
  p.isfill = isfill;
  p.fill = fill;
  if (p.lines == 0) return hipSuccess;
  int seg, nw;
  const bool typed = in.dtype == IFM_F32 || in.dtype == IFM_F64;
  // whole lines held in registers (contiguous lines: F32 / F64 input only; the
  // run-time element type's loads do not fit the transpose rounds' registers)
  if (iir_segments(p.n[axis], &seg, &nw) && (stride != 1 || typed)) {
    IirSegTables tb;
    tb.seg = seg;
    tb.nw = nw;
    for (int k = 0; k < 3; ++k) {
      // responses to the unit state e_k: y[-1 - k] = 1
      double hm[3] = {0.0, 0.0, 0.0}, gm[3] = {0.0, 0.0, 0.0};
      hm[k] = gm[k] = 1.0;
      for (int j = 0; j < IS_S; ++j) {
        const double h = c.a[0] * hm[0] + c.a[1] * hm[1] + c.a[2] * hm[2];
        const double g = c.b[0] * gm[0] + c.b[1] * gm[1] + c.b[2] * gm[2];
        tb.h[k][j] = h;
        tb.g[k][j] = g;
        hm[2] = hm[1];
        hm[1] = hm[0];
        hm[0] = h;
        gm[2] = gm[1];
        gm[1] = gm[0];
        gm[0] = g;
      }
    }
    // two segments per wave, 32 lines per workgroup
    const bool rows = stride == 1;
    const unsigned nb2 = (unsigned)((p.lines + IS2_L - 1) / IS2_L);
    const dim3 blk2(64 * ((nw + IS_SPW - 1) / IS_SPW));
    const size_t lds2 = sizeof(double) * iir_seg2_lds_doubles(nw, rows);
    if (rows) {
      if (in.dtype == IFM_F32) hipLaunchKernelGGL((k_iir_rows2<IFM_F32>), dim3(nb2), blk2, lds2, st, p, tb);
      else hipLaunchKernelGGL((k_iir_rows2<IFM_F64>), dim3(nb2), blk2, lds2, st, p, tb);
    } else {
      if (in.dtype == IFM_F32) hipLaunchKernelGGL((k_iir_seg2<IFM_F32>), dim3(nb2), blk2, lds2, st, p, tb);
      else if (in.dtype == IFM_F64) hipLaunchKernelGGL((k_iir_seg2<IFM_F64>), dim3(nb2), blk2, lds2, st, p, tb);
      else hipLaunchKernelGGL((k_iir_seg2<-1>), dim3(nb2), blk2, lds2, st, p, tb);
    }
    return hipGetLastError();
  }
  if (stride == 1) {  // lines along the contiguous axis: LDS-staged chunks
    const unsigned nb = (unsigned)((p.lines + IR_L - 1) / IR_L);
    const size_t lds = sizeof(double) * IR_L * IR_P;
    if (in.dtype == IFM_F32) hipLaunchKernelGGL(k_iir_rows<IFM_F32>, dim3(nb), dim3(256), lds, st, p);
    else if (in.dtype == IFM_F64) hipLaunchKernelGGL(k_iir_rows<IFM_F64>, dim3(nb), dim3(256), lds, st, p);
    else hipLaunchKernelGGL(k_iir_rows<-1>, dim3(nb), dim3(256), lds, st, p);
    return hipGetLastError();
  }
  int64_t blocks = (p.lines + 255) / 256;
  const bool s1 = stride == 1;
  if (in.dtype == IFM_F32) {
    if (s1) hipLaunchKernelGGL((k_iir_strided<true, IFM_F32>), dim3((unsigned)blocks), dim3(256), 0, st, p);
    else hipLaunchKernelGGL((k_iir_strided<false, IFM_F32>), dim3((unsigned)blocks), dim3(256), 0, st, p);
  } else if (in.dtype == IFM_F64) {
    if (s1) hipLaunchKernelGGL((k_iir_strided<true, IFM_F64>), dim3((unsigned)blocks), dim3(256), 0, st, p);
    else hipLaunchKernelGGL((k_iir_strided<false, IFM_F64>), dim3((unsigned)blocks), dim3(256), 0, st, p);
  } else {
    if (s1) hipLaunchKernelGGL((k_iir_strided<true, -1>), dim3((unsigned)blocks), dim3(256), 0, st, p);
    else hipLaunchKernelGGL((k_iir_strided<false, -1>), dim3((unsigned)blocks), dim3(256), 0, st, p);
  }
  return hipGetLastError();
}

}  // namespace ifm
