/*
 * ifm_oracle.c — CPU restatement of ImageFiltering.jl's filtering hot path, in
 * FP64, used ONLY as test infrastructure: the checker for the GPU library in
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Nothing in
 * the product (libifm, the Python host package) links, imports or calls it.
 *
 * Every function follows the reference operation for operation where that
 * matters for results (index maps, cascade regions, recursion order); file:line
 * citations are relative to the ImageFiltering.jl v0.7.12 source tree.
 *
 * Arrays are dense column-major (axis 0 fastest), ndim <= 4, sizes padded with
 * 1.  Indices are 0-based.  Border styles: 0 replicate, 1 circular,
 * 2 symmetric, 3 reflect, 4 fill.
 *
 * Parity pins (see tests/test_oracle.py): padindex/padarray against the
 * literal matrices of test/border.jl:36-206; FIR cascade against
 * test/nd.jl:58-75 and the closed forms of test/2d.jl:46-226, test/nd.jl:128-156;
 * median/min/max against test/mapwindow.jl:5-124 and brute force; Lemire
 * extrema against test/mapwindow.jl:17-73; IIR against Triggs-Sdika's Eq. 8
 * and impulse-response closeness (test/triggs.jl) plus the z-transform
 * identities checked in the tests.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXD 4

/* Threads of the loops below (the all-core "parallel restatement" of the CPU
 * baseline).  Every output element is computed by the same operations in the
 * same order whatever the thread count, so results do not depend on it.
 * Without OpenMP (the default build used by the tests) this is one thread. */
int orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
  return n > 0 ? n : 1;
#else
  (void)n;
  return 1;
#endif
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_num_procs();
#else
  return 1;
#endif
}

static int64_t jmod(int64_t x, int64_t m) { /* Julia mod for m > 0 */
  int64_t r = x % m;
  return r < 0 ? r + m : r;
}

/* padindex (src/border.jl:564-590), 0-based, for image axis 0..n-1.
 * Returns the source index of padded position i (image coordinates, may be
 * negative or >= n), or -1 for Fill.  Returns -2 for reflect with n == 1
 * (the reference divides by a zero period there). */
int64_t orc_padmap(int style, int64_t i, int64_t n) {
  if (i >= 0 && i < n) return i; /* every table is the identity inside the axis */
  switch (style) {
  case 0: /* :replicate -> clamp */
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
  case 1: /* :circular -> modrange(extend(lo, inds, hi), inds) */
    return jmod(i, n);
  case 2: { /* :symmetric -> table [1..n, n..1], period 2n */
    int64_t j = jmod(i, 2 * n);
    return j < n ? j : 2 * n - 1 - j;
  }
  case 3: { /* :reflect -> table [1..n, n-1..2], period 2n-2 */
    if (n == 1) return -2;
    int64_t j = jmod(i, 2 * n - 2);
    return j < n ? j : 2 * n - 2 - j;
  }
  default:
    return (i >= 0 && i < n) ? i : -1;
  }
}

int orc_padindex(int style, int64_t lo, int64_t n, int64_t hi, int64_t *idx) {
  for (int64_t t = 0; t < lo + n + hi; ++t) {
    int64_t v = orc_padmap(style, t - lo, n);
    if (v == -2) return -1;
    idx[t] = v;
  }
  return 0;
}

static void dims4(int ndim, const int64_t *n, int64_t *o) {
  for (int d = 0; d < MAXD; ++d) o[d] = d < ndim ? n[d] : 1;
}

/* padarray (src/border.jl:324-347) -> out of size lo+n+hi per axis. */
int orc_padarray_f64(const double *in, int ndim, const int64_t *n, int style, double fill,
                     const int64_t *lo, const int64_t *hi, double *out) {
  int64_t N[MAXD], L[MAXD], H[MAXD], P[MAXD];
  dims4(ndim, n, N);
  for (int d = 0; d < MAXD; ++d) {
    L[d] = d < ndim ? lo[d] : 0;
    H[d] = d < ndim ? hi[d] : 0;
    P[d] = L[d] + N[d] + H[d];
  }
  for (int d = 0; d < ndim; ++d)
    if (style == 3 && N[d] == 1 && (L[d] || H[d])) return -1;
  int64_t p = 0;
  for (int64_t i3 = 0; i3 < P[3]; ++i3)
    for (int64_t i2 = 0; i2 < P[2]; ++i2)
      for (int64_t i1 = 0; i1 < P[1]; ++i1)
        for (int64_t i0 = 0; i0 < P[0]; ++i0, ++p) {
          int64_t s0 = orc_padmap(style, i0 - L[0], N[0]), s1 = orc_padmap(style, i1 - L[1], N[1]);
          int64_t s2 = orc_padmap(style, i2 - L[2], N[2]), s3 = orc_padmap(style, i3 - L[3], N[3]);
          if (s0 < 0 || s1 < 0 || s2 < 0 || s3 < 0)
            out[p] = fill;
          else
            out[p] = in[s0 + N[0] * (s1 + N[1] * (s2 + N[2] * s3))];
        }
  return 0;
}

/* Value of the padded array at image coordinates x (any integers). */
static double padval(const double *img, const int64_t *N, int style, double fill, const int64_t *x) {
  int64_t s[MAXD];
  for (int d = 0; d < MAXD; ++d) {
    s[d] = orc_padmap(style, x[d], N[d]);
    if (s[d] < 0) return fill;
  }
  return img[s[0] + N[0] * (s[1] + N[1] * (s[2] + N[2] * s[3]))];
}

/* Dense correlation on the padded image (src/imfilter.jl:592-669):
 * out[i] = sum_J A[i+J] k[J], kernel dim 1 innermost, FP64 accumulation.
 * Output region: image indices ofirst[d] .. ofirst[d]+osize[d]-1. */
int orc_imfilter_dense(const double *img, int ndim, const int64_t *n, const double *k,
                       const int64_t *ksize, const int64_t *kfirst, int style, double fill,
                       const int64_t *ofirst, const int64_t *osize, double *out) {
  int64_t N[MAXD], KS[MAXD], KF[MAXD], OF[MAXD], OS[MAXD];
  dims4(ndim, n, N);
  for (int d = 0; d < MAXD; ++d) {
    KS[d] = d < ndim ? ksize[d] : 1;
    KF[d] = d < ndim ? kfirst[d] : 0;
    OF[d] = (d < ndim && ofirst) ? ofirst[d] : 0;
    OS[d] = d < ndim ? (osize ? osize[d] : N[d]) : 1;
  }
  for (int d = 0; d < ndim; ++d)
    if (style == 3 && N[d] == 1 && (KS[d] > 1 || KF[d] != 0)) return -1;
  /* padarray: materialise the padded block the outputs need
   * (src/imfilter.jl:336), then the direct correlation loop. */
  int64_t P[MAXD], PO[MAXD], PS[MAXD], cnt = 1;
  for (int d = 0; d < MAXD; ++d) {
    P[d] = OS[d] + KS[d] - 1;
    PO[d] = OF[d] + KF[d];  /* image coordinate of padded element 0 */
    PS[d] = cnt;
    cnt *= P[d];
  }
  double *pad = (double *)malloc(sizeof(double) * (size_t)(cnt ? cnt : 1));
  const int64_t prow = P[1] * P[2] * P[3];
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < prow; ++r) {
    int64_t x[MAXD];
    const int64_t i1 = r % P[1], i2 = (r / P[1]) % P[2], i3 = r / (P[1] * P[2]);
    for (int64_t i0 = 0; i0 < P[0]; ++i0) {
      x[0] = PO[0] + i0; x[1] = PO[1] + i1; x[2] = PO[2] + i2; x[3] = PO[3] + i3;
      pad[i0 + P[0] * r] = padval(img, N, style, fill, x);
    }
  }
  /* __imfilter_inbounds!: tmp += A[i+j, I+J] * k[j, J], j (dim 1) innermost */
  const int64_t orow = OS[1] * OS[2] * OS[3];
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < orow; ++r) {
    const int64_t o1 = r % OS[1], o2 = (r / OS[1]) % OS[2], o3 = r / (OS[1] * OS[2]);
    for (int64_t o0 = 0; o0 < OS[0]; ++o0) {
          const int64_t p = o0 + OS[0] * r;
          double tmp = 0.0;
          const double *kq = k;
          for (int64_t j3 = 0; j3 < KS[3]; ++j3)
            for (int64_t j2 = 0; j2 < KS[2]; ++j2)
              for (int64_t j1 = 0; j1 < KS[1]; ++j1) {
                const double *row = pad + o0 + (o1 + j1) * PS[1] + (o2 + j2) * PS[2] + (o3 + j3) * PS[3];
                for (int64_t j0 = 0; j0 < KS[0]; ++j0) tmp += row[j0] * kq[j0];
                kq += KS[0];
              }
          out[p] = tmp;
    }
  }
  free(pad);
  return 0;
}

/* Cascade of 1-D factors (src/imfilter.jl:331-457 with the FIR loops of
 * :671-739): pad once with the accumulated padding (src/border.jl:614-621,
 * :640-641), then filter stage by stage over the shrinking valid region
 * (indsstep, next_shrink).  factors: faxis[f], flen[f], ffirst[f], taps
 * concatenated in order.  Copy kernels (single tap 1 at offset 0) are skipped
 * as iscopy does (src/imfilter.jl:379, :386). */
int orc_imfilter_factors(const double *img, int ndim, const int64_t *n, int nf, const int *faxis,
                         const int *flen, const int64_t *ffirst, const double *taps, int style,
                         double fill, const int64_t *ofirst, const int64_t *osize, double *out) {
  int64_t N[MAXD], OF[MAXD], OS[MAXD];
  dims4(ndim, n, N);
  for (int d = 0; d < MAXD; ++d) {
    OF[d] = (d < ndim && ofirst) ? ofirst[d] : 0;
    OS[d] = d < ndim ? (osize ? osize[d] : N[d]) : 1;
  }
  for (int d = 0; d < ndim; ++d)
    if (style == 3 && N[d] == 1) return -1;
  /* tap offsets into the concatenated array; drop copy kernels */
  int *off = (int *)malloc(sizeof(int) * (nf + 1));
  int *keep = (int *)malloc(sizeof(int) * (nf + 1));
  int nk = 0, acc = 0;
  for (int f = 0; f < nf; ++f) {
    int iscopy = flen[f] == 1 && ffirst[f] == 0 && taps[acc] == 1.0;
    if (!iscopy) { keep[nk] = f; off[nk] = acc; ++nk; }
    acc += flen[f];
  }
  if (nk == 0) { /* trivial kernel: a copy (src/imfilter.jl:372-375) */
    int64_t p = 0, x[MAXD];
    for (int64_t o3 = 0; o3 < OS[3]; ++o3)
      for (int64_t o2 = 0; o2 < OS[2]; ++o2)
        for (int64_t o1 = 0; o1 < OS[1]; ++o1)
          for (int64_t o0 = 0; o0 < OS[0]; ++o0, ++p) {
            x[0] = OF[0] + o0; x[1] = OF[1] + o1; x[2] = OF[2] + o2; x[3] = OF[3] + o3;
            out[p] = padval(img, N, style, fill, x);
          }
    free(off); free(keep);
    return 0;
  }
  /* Region of stage s: the output region expanded by factors s+1..nk-1. */
  int64_t (*rlo)[MAXD] = malloc(sizeof(int64_t[MAXD]) * nk);
  int64_t (*rhi)[MAXD] = malloc(sizeof(int64_t[MAXD]) * nk);
  for (int s = 0; s < nk; ++s) {
    for (int d = 0; d < MAXD; ++d) { rlo[s][d] = OF[d]; rhi[s][d] = OF[d] + OS[d] - 1; }
    for (int t = s + 1; t < nk; ++t) {
      int f = keep[t];
      rlo[s][faxis[f]] += ffirst[f];
      rhi[s][faxis[f]] += ffirst[f] + flen[f] - 1;
    }
  }
  double *prev = NULL;
  int64_t plo[MAXD], pext[MAXD];
  for (int s = 0; s < nk; ++s) {
    int f = keep[s], a = faxis[f];
    int64_t ext[MAXD], cnt = 1;
    for (int d = 0; d < MAXD; ++d) { ext[d] = rhi[s][d] - rlo[s][d] + 1; cnt *= ext[d]; }
    double *cur = (s == nk - 1) ? out : (double *)malloc(sizeof(double) * (size_t)(cnt ? cnt : 1));
    const int64_t nrow = ext[1] * ext[2] * ext[3];
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < nrow; ++r) {
      int64_t x[MAXD];
      const int64_t i1 = r % ext[1], i2 = (r / ext[1]) % ext[2], i3 = r / (ext[1] * ext[2]);
          for (int64_t i0 = 0; i0 < ext[0]; ++i0) {
            const int64_t p = i0 + ext[0] * r;
            double tmp = 0.0;
            for (int j = 0; j < flen[f]; ++j) {
              x[0] = rlo[s][0] + i0; x[1] = rlo[s][1] + i1; x[2] = rlo[s][2] + i2; x[3] = rlo[s][3] + i3;
              x[a] += ffirst[f] + j;
              double v;
              if (s == 0) {
                v = padval(img, N, style, fill, x);
              } else {
                int64_t q = 0;
                for (int d = MAXD - 1; d >= 0; --d) q = q * pext[d] + (x[d] - plo[d]);
                v = prev[q];
              }
              tmp += v * taps[off[s] + j];
            }
            cur[p] = tmp;
          }
    }
    if (prev) free(prev);
    prev = (s == nk - 1) ? NULL : cur;
    for (int d = 0; d < MAXD; ++d) { plo[d] = rlo[s][d]; pext[d] = ext[d]; }
  }
  free(rlo); free(rhi); free(off); free(keep);
  return 0;
}

/* Triggs-Sdika IIR along one axis (src/imfilter.jl:964-1092): left border
 * init (Triggs & Sdika sec. II), forward recursion, right border init
 * (Eqs. 14-15), backward recursion, scaling.  `coef` = a[3], b[3], scale,
 * M[9] (row-major).  Pad borders initialise from the edge pixel of `src`,
 * Fill from `fill`. */
static void iir_axis(const double *src, double *dst, const int64_t *N, int axis, const double *coef,
                     int isfill, double fill) {
  const double *a = coef, *b = coef + 3, scale = coef[6];
  const double *M = coef + 7;
  double asum = a[0] + a[1] + a[2], bsum = b[0] + b[1] + b[2];
  int64_t n = N[axis], stride = 1;
  for (int d = 0; d < axis; ++d) stride *= N[d];
  int64_t total = N[0] * N[1] * N[2] * N[3], lines = total / n;
#pragma omp parallel for schedule(static)
  for (int64_t L = 0; L < lines; ++L) {
    int64_t pre = L % stride, post = L / stride;
    int64_t base = pre + post * stride * n;
#define IN(i) src[base + (int64_t)(i) * stride]
#define OUT(i) dst[base + (int64_t)(i) * stride]
    double iminus = isfill ? fill : IN(0);
    double iplus = isfill ? fill : IN(n - 1);
    /* _leftborder! */
    double uminus = iminus / (1 - asum);
    for (int c = 1; c <= 3; ++c) {
      int64_t i = c - 1;
      double tmp = IN(i);
      for (int j = 1; j <= c - 1; ++j) tmp += a[j - 1] * OUT(i - j);
      for (int j = c; j <= 3; ++j) tmp += a[j - 1] * uminus;
      OUT(i) = tmp;
    }
    /* forward, omitting the final point */
    for (int64_t i = 3; i <= n - 2; ++i) {
      double tmp = IN(i);
      for (int j = 1; j <= 3; ++j) tmp += a[j - 1] * OUT(i - j);
      OUT(i) = tmp;
    }
    /* _rightborder! */
    {
      int64_t i = n - 1;
      double tmp = IN(i);
      for (int j = 1; j <= 3; ++j) tmp += a[j - 1] * OUT(i - j);
      OUT(i) = tmp;
      double uplus = iplus / (1 - asum);
      double vplus = uplus / (1 - bsum);
      double du[3] = {OUT(n - 1) - uplus, OUT(n - 2) - uplus, OUT(n - 3) - uplus};
      double vr[3];
      for (int r = 0; r < 3; ++r) vr[r] = M[3 * r] * du[0] + M[3 * r + 1] * du[1] + M[3 * r + 2] * du[2] + vplus;
      OUT(n - 1) = vr[0];
      for (int c = 2; c <= 3; ++c) {
        int64_t ii = n - c;
        double t2 = OUT(ii);
        for (int j = 1; j <= c - 1; ++j) t2 += b[j - 1] * OUT(ii + j);
        for (int j = c; j <= 3; ++j) t2 += b[j - 1] * vr[j - c + 1];
        OUT(ii) = t2;
      }
    }
    /* backward */
    for (int64_t i = n - 4; i >= 0; --i) {
      double tmp = OUT(i);
      for (int j = 1; j <= 3; ++j) tmp += b[j - 1] * OUT(i + j);
      OUT(i) = tmp;
    }
    for (int64_t i = 0; i < n; ++i) OUT(i) *= scale;
#undef IN
#undef OUT
  }
}

/* Cascade of IIR factors (src/imfilter.jl:385-457 with _imfilter_dim!),
 * each stage reading the previous stage's result.  Returns -1 if an axis
 * has length <= 3 (the reference's DimensionMismatch). */
int orc_imfilter_iir(const double *img, int ndim, const int64_t *n, int nk, const double *coef,
                     const int *axes, int isfill, double fill, double *out) {
  int64_t N[MAXD];
  dims4(ndim, n, N);
  int64_t total = N[0] * N[1] * N[2] * N[3];
  for (int s = 0; s < nk; ++s)
    if (N[axes[s]] <= 3) return -1;
  double *a1 = (double *)malloc(sizeof(double) * (size_t)total);
  memcpy(a1, img, sizeof(double) * (size_t)total);
  double *a2 = (double *)malloc(sizeof(double) * (size_t)total);
  for (int s = 0; s < nk; ++s) {
    iir_axis(a1, a2, N, axes[s], coef + 16 * s, isfill, fill);
    double *t = a1; a1 = a2; a2 = t;
  }
  memcpy(out, a1, sizeof(double) * (size_t)total);
  free(a1); free(a2);
  return 0;
}

/* isless total order for floats as a sortable key: NaN above +Inf, -0.0 below +0.0. */
static uint64_t fkey(double v) {
  if (isnan(v)) return UINT64_MAX;
  uint64_t u;
  memcpy(&u, &v, 8);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
static int cmpkey(const void *x, const void *y) {
  uint64_t a = *(const uint64_t *)x, b = *(const uint64_t *)y;
  return a < b ? -1 : (a > b);
}

/* Gather one mapwindow window (src/mapwindow.jl:289-333): positions inside
 * the image are copied; at the edge, copy_win! maps every axis with
 * padindex(border, inner, outer) where inner is the window's intersection
 * with the image axes (:310-316), or fills with the Fill value (:323-331). */
/* *err is set when an edge window has an empty or zero-period index map. */
static int64_t gather_window(const double *img, const int64_t *N, int style, double fill,
                             const int64_t *q, const int64_t *WF, const int64_t *WL, double *buf, int *err) {
  int64_t ilo[MAXD], ilen[MAXD];
  for (int d = 0; d < MAXD; ++d) {
    int64_t a = q[d] + WF[d], b = q[d] + WF[d] + WL[d] - 1;
    ilo[d] = a < 0 ? 0 : a;
    int64_t hi = b > N[d] - 1 ? N[d] - 1 : b;
    ilen[d] = hi - ilo[d] + 1;
  }
  int64_t m = 0;
  for (int64_t j3 = 0; j3 < WL[3]; ++j3)
    for (int64_t j2 = 0; j2 < WL[2]; ++j2)
      for (int64_t j1 = 0; j1 < WL[1]; ++j1)
        for (int64_t j0 = 0; j0 < WL[0]; ++j0) {
          int64_t x[MAXD] = {q[0] + WF[0] + j0, q[1] + WF[1] + j1, q[2] + WF[2] + j2, q[3] + WF[3] + j3};
          int outside = 0, s[MAXD];
          for (int d = 0; d < MAXD; ++d) {
            if (x[d] >= 0 && x[d] < N[d]) { s[d] = x[d]; continue; }
            outside = 1;
            if (style == 4) { s[d] = -1; continue; }
            if (ilen[d] <= 0 || (style == 3 && ilen[d] == 1)) { *err = 1; s[d] = -1; continue; }
            s[d] = ilo[d] + orc_padmap(style, x[d] - ilo[d], ilen[d]);
          }
          int isfill = 0;
          for (int d = 0; d < MAXD; ++d) if (s[d] < 0) isfill = 1;
          (void)outside;
          buf[m++] = isfill ? fill : img[s[0] + N[0] * (s[1] + N[1] * (s[2] + N[2] * s[3]))];
        }
  return m;
}

/* mapwindow(median!, img, window; border) -> median_fast!
 * (src/mapwindow.jl:123-131): partialsort! to index (first+last)÷2 under
 * isless, i.e. the ((L+1)÷2)-th smallest of L values; middle(x) == x. */
int orc_median_f64(const double *img, int ndim, const int64_t *n, const int64_t *wfirst,
                   const int64_t *wlen, int style, double fill, const int64_t *ofirst,
                   const int64_t *osize, double *out) {
  int64_t N[MAXD], WF[MAXD], WL[MAXD], OF[MAXD], OS[MAXD];
  dims4(ndim, n, N);
  int64_t L = 1;
  for (int d = 0; d < MAXD; ++d) {
    WF[d] = d < ndim ? wfirst[d] : 0;
    WL[d] = d < ndim ? wlen[d] : 1;
    OF[d] = (d < ndim && ofirst) ? ofirst[d] : 0;
    OS[d] = d < ndim ? (osize ? osize[d] : N[d]) : 1;
    L *= WL[d];
  }
  int err = 0;
  const int64_t orow = OS[1] * OS[2] * OS[3];
#pragma omp parallel reduction(| : err)
  {
  double *buf = (double *)malloc(sizeof(double) * (size_t)L);
  uint64_t *key = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)L);
  int64_t q[MAXD];
#pragma omp for schedule(static)
  for (int64_t r = 0; r < orow; ++r) {
    const int64_t o1 = r % OS[1], o2 = (r / OS[1]) % OS[2], o3 = r / (OS[1] * OS[2]);
        for (int64_t o0 = 0; o0 < OS[0]; ++o0) {
          const int64_t p = o0 + OS[0] * r;
          q[0] = OF[0] + o0; q[1] = OF[1] + o1; q[2] = OF[2] + o2; q[3] = OF[3] + o3;
          int64_t m = gather_window(img, N, style, fill, q, WF, WL, buf, &err);
          for (int64_t t = 0; t < m; ++t) key[t] = fkey(buf[t]);
          qsort(key, (size_t)m, sizeof(uint64_t), cmpkey);
          uint64_t kk = key[(m + 1) / 2 - 1];
          double v;
          if (kk == UINT64_MAX) {
            v = NAN;
          } else {
            uint64_t u = (kk >> 63) ? (kk & 0x7fffffffffffffffull) : ~kk;
            memcpy(&v, &u, 8);
          }
          out[p] = v;
        }
  }
  free(buf); free(key);
  }
  return err ? -1 : 0;
}

/* Julia min/max on floats: NaN-propagating, -0.0 < +0.0. */
static double jmin(double x, double y) {
  if (isnan(x) || isnan(y)) return NAN;
  if (x == y) return signbit(x) ? x : y;
  return x < y ? x : y;
}
static double jmax(double x, double y) {
  if (isnan(x) || isnan(y)) return NAN;
  if (x == y) return signbit(x) ? y : x;
  return x > y ? x : y;
}

/* mapwindow(minimum|maximum, img, window; border) through the generic
 * window path (src/mapwindow.jl:270-333); either output may be NULL. */
int orc_minmax_f64(const double *img, int ndim, const int64_t *n, const int64_t *wfirst,
                   const int64_t *wlen, int style, double fill, const int64_t *ofirst,
                   const int64_t *osize, double *omin, double *omax) {
  int64_t N[MAXD], WF[MAXD], WL[MAXD], OF[MAXD], OS[MAXD];
  dims4(ndim, n, N);
  int64_t L = 1;
  for (int d = 0; d < MAXD; ++d) {
    WF[d] = d < ndim ? wfirst[d] : 0;
    WL[d] = d < ndim ? wlen[d] : 1;
    OF[d] = (d < ndim && ofirst) ? ofirst[d] : 0;
    OS[d] = d < ndim ? (osize ? osize[d] : N[d]) : 1;
    L *= WL[d];
  }
  int err = 0;
  const int64_t orow = OS[1] * OS[2] * OS[3];
#pragma omp parallel reduction(| : err)
  {
  double *buf = (double *)malloc(sizeof(double) * (size_t)L);
  int64_t q[MAXD];
#pragma omp for schedule(static)
  for (int64_t r = 0; r < orow; ++r) {
    const int64_t o1 = r % OS[1], o2 = (r / OS[1]) % OS[2], o3 = r / (OS[1] * OS[2]);
        for (int64_t o0 = 0; o0 < OS[0]; ++o0) {
          const int64_t p = o0 + OS[0] * r;
          q[0] = OF[0] + o0; q[1] = OF[1] + o1; q[2] = OF[2] + o2; q[3] = OF[3] + o3;
          int64_t m = gather_window(img, N, style, fill, q, WF, WL, buf, &err);
          double mn = buf[0], mx = buf[0];
          for (int64_t t = 1; t < m; ++t) { mn = jmin(mn, buf[t]); mx = jmax(mx, buf[t]); }
          if (omin) omin[p] = mn;
          if (omax) omax[p] = mx;
        }
  }
  free(buf);
  }
  return err ? -1 : 0;
}

/* Lemire streaming min/max along axis 0 of (mn, mx) pairs
 * (src/mapwindow.jl:426-481), including the delayed output cache and the
 * wedge pops, so even window widths get the reference's alignment. */
static void lemire_axis(double *mn, double *mx, const int64_t *N, int axis, int64_t w) {
  int64_t n = N[axis], stride = 1;
  for (int d = 0; d < axis; ++d) stride *= N[d];
  int64_t total = N[0] * N[1] * N[2] * N[3], lines = total / n;
  int64_t h = w >> 1;
  int64_t cap = w + 1;
  /* cache seed: first(A) of the whole array (:390-393) */
  const double a0n = mn[0], a0x = mx[0];
#pragma omp parallel
  {
  int64_t *L = (int64_t *)malloc(sizeof(int64_t) * (size_t)cap);
  int64_t *U = (int64_t *)malloc(sizeof(int64_t) * (size_t)cap);
  double *cmn = (double *)malloc(sizeof(double) * (size_t)(h + 1));
  double *cmx = (double *)malloc(sizeof(double) * (size_t)(h + 1));
#pragma omp for schedule(static)
  for (int64_t ln = 0; ln < lines; ++ln) {
    int64_t pre = ln % stride, post = ln / stride;
    int64_t base = pre + post * stride * n;
#define MN(i) mn[base + (int64_t)(i) * stride]
#define MX(i) mx[base + (int64_t)(i) * stride]
    /* deques as ring buffers over absolute positions */
    int64_t Lh = 0, Lt = 0, Uh = 0, Ut = 0; /* head index, tail (exclusive) count */
    int64_t ch = 0; /* cache head */
    for (int64_t t = 0; t < h; ++t) { cmn[t] = a0n; cmx[t] = a0x; }
    double cn = a0n, cx = a0x;
#define CYCLE()                                                                        \
    do {                                                                               \
      double gn = MN(L[Lh % cap]), gx = MX(U[Uh % cap]);                               \
      if (h == 0) { cn = gn; cx = gx; }                                                \
      else { cn = cmn[ch]; cx = cmx[ch]; cmn[ch] = gn; cmx[ch] = gx; ch = (ch + 1) % h; } \
    } while (0)
#define ADD(i)                                                                         \
    do {                                                                               \
      double vn = MN(i), vx = MX(i);                                                   \
      while (Lt > Lh && vn < MN(L[(Lt - 1) % cap])) --Lt;                              \
      while (Ut > Uh && vx > MX(U[(Ut - 1) % cap])) --Ut;                              \
      L[Lt % cap] = (i); ++Lt;                                                         \
      U[Ut % cap] = (i); ++Ut;                                                         \
    } while (0)
    int64_t iw = (n - 1) < (w - 1) ? (n - 1) : (w - 1);
    for (int64_t i = 0; i <= iw; ++i) { ADD(i); CYCLE(); }
    for (int64_t i = iw + 1; i <= n - 1; ++i) {
      MN(i - w) = cn; MX(i - w) = cx;
      if (i == w + U[Uh % cap]) ++Uh;
      if (i == w + L[Lh % cap]) ++Lh;
      ADD(i);
      CYCLE();
    }
    for (int64_t i = n - 1 - w + 1; i <= n - 2; ++i) {
      if (i >= 0) { MN(i) = cn; MX(i) = cx; }
      if (i == U[Uh % cap]) ++Uh;
      if (i == L[Lh % cap]) ++Lh;
      CYCLE();
    }
    MN(n - 1) = cn; MX(n - 1) = cx;
#undef CYCLE
#undef ADD
#undef MN
#undef MX
  }
  free(L); free(U); free(cmn); free(cmx);
  }
}

/* mapwindow(extrema, A, window::Dims) -> extrema_filter (src/mapwindow.jl:
 * 337-417): Lemire per axis in order 1..N; an axis with window <= 1 is left
 * unchanged. */
int orc_extrema_lemire_f64(const double *img, int ndim, const int64_t *n, const int64_t *window,
                           double *omin, double *omax) {
  int64_t N[MAXD];
  dims4(ndim, n, N);
  int64_t total = N[0] * N[1] * N[2] * N[3];
  memcpy(omin, img, sizeof(double) * (size_t)total);
  memcpy(omax, img, sizeof(double) * (size_t)total);
  for (int d = 0; d < ndim; ++d)
    if (window[d] > 1) lemire_axis(omin, omax, N, d, window[d]);
  return 0;
}

/* Timing helper for the CPU baseline: the dense FIR restatement over a
 * sub-block of rows of the image (same arithmetic as orc_imfilter_dense). */
int orc_imfilter_dense_rows(const double *img, int ndim, const int64_t *n, const double *k,
                            const int64_t *ksize, const int64_t *kfirst, int style, double fill,
                            int64_t row0, int64_t nrows, double *out) {
  int64_t of[MAXD] = {0, row0, 0, 0}, os[MAXD];
  for (int d = 0; d < ndim; ++d) os[d] = n[d];
  os[1] = nrows;
  return orc_imfilter_dense(img, ndim, n, k, ksize, kfirst, style, fill, of, os, out);
}
