// k_mfma.hip — 2-D dense correlation on the matrix cores (gfx950).
//
// Replaces __imfilter_inbounds! (src/imfilter.jl:592-669) for 2-D kernels of
// up to 49 x 32 taps: out[i] = sum_J A[i+J] k[J] on the padded image, the
// border map applied while the input tile is staged.
//
// Banded Toeplitz form.  For kernel row u, the correlation along x of 16
// output columns a with the input columns s = a + t is a product with the
// Toeplitz matrix T_u[s][a] = k[s - a, u] (zero outside 0 <= s-a < Kx),
// s < 16 + Kx - 1, taken in K blocks of 32.  One v_mfma_f32_16x16x32_f16
// computes
//     D[yy][a] += sum_s F[yy][s] * T_u[s][a]
// with F holding image rows.  A wave owns 8 accumulator blocks whose rows are
// interleaved (block b, lane-row yy -> output row y0 + b + 8*yy), so the image
// fragment F_r (tile rows r + 8*yy) serves every block b with u = r - b: one
// LDS fragment read feeds up to 8 MFMA groups.
//
// Precision.  x and k are split into FP16 hi + lo after scaling by powers of
// two (k by one factor per kernel, x by one factor per tile so the tile's
// largest |x| lies in [0.5, 1)); xh*kh + xl*kh + xh*kl accumulate in FP32.
// Each operand keeps ~22 significant bits, so a product carries a relative
// error near 2^-22 (the dropped xl*kl is below 2^-24), comparable to a plain
// FP32 sum.  Tiles holding Inf or NaN are computed by a direct FP32 loop
// instead, so non-finite values reach exactly the outputs whose window holds
// them, as in the reference's direct sum.
//
// Work decomposition.  A tile is 64 output columns (four waves of 16) by 128
// output rows.  The grid is persistent: a workgroup walks a run of vertically
// adjacent tiles, so the halo rows two consecutive tiles share come from L2.
// Kernel rows are processed in chunks of 8 per K block; a chunk's Toeplitz
// fragments (hi and lo) take 64 VGPRs and two register sets are kept.  With
// at most two chunks (Kx <= 17, Ky <= 16) both are loaded once per workgroup;
// with more, the fragments of chunk c + 2 are fetched while chunk c + 1's
// MFMAs run.  The next tile's global loads are issued before the current
// tile's MFMA phase and used after it, so HBM latency hides behind the matrix
// work.  Loads keep the raw element bits; conversion waits for the value.
//
// Staging map.  Tile columns 0..63 are read with lane = column, wave w taking
// rows w, w+4, ...; the remaining columns (16 or 48) in groups of 16 with
// thread t at column 64 + 16g + t%16 and rows t/16, t/16 + 16, ...  Every
// index is affine in the element number, so nothing per element stays live.
#include "ifm_internal.h"

namespace ifm {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int MF_WAVES = 4;
constexpr int MF_TX = 16 * MF_WAVES;  // output columns per tile
constexpr int MF_B = 8;               // interleaved row blocks per wave
constexpr int MF_TY = 16 * MF_B;      // output rows per tile
constexpr int MF_UCH = 8;             // kernel rows per chunk
constexpr int MF_MAXCH = 4;           // kernel rows up to 32
constexpr int MF_MAXKB = 2;           // kernel columns up to 49

// LDS layout of a tile.  A fragment read has lane (yy, g) read 8 words of
// tile row R0 + 8*yy, so rows are grouped by R mod 8: tile row R sits in LDS
// row (R % 8) * RB + R / 8, and the 16 rows one read touches are consecutive
// LDS rows.  The pitch P = 4q words with q odd puts them on 16 distinct
// 4-bank groups: a ds_read_b128 is free of bank conflicts.
template <int KB, int NCH>
struct MfmaGeom {
  static constexpr int TWP = 48 + 32 * KB;             // staged columns (>= 64 + Kx - 1)
  static constexpr int P = TWP + 4;                    // LDS row pitch in words
  static constexpr int TH = MF_TY + NCH * MF_UCH - 1;  // staged rows
  static constexpr int R16 = (TWP - 64) / 16;          // halo groups of 16 columns
  static constexpr int MR = (TH + 3) / 4;              // rows per wave in the lane = column part
  static constexpr int HR = (TH + 15) / 16;            // rows per thread in a halo group
  static constexpr int NEL = MR + R16 * HR;            // staged elements per thread
  // staging batches (register budget); four-chunk and two-block kernels
  // take two even when one would do: their fragments leave too few
  // registers (17 x 17: 80 -> 12 bytes of spill, 97 -> 94 us; 31 x 15:
  // 204 -> 68 bytes, 121 -> 106 us; 31 x 31 unchanged)
  static constexpr int NB =
      ((NEL + 63) / 64 > 1 || (NCH < 4 && KB < 2)) ? (NEL + 63) / 64 : 2;
  static constexpr int JB = (NEL + NB - 1) / NB;       // elements per thread and batch
  static constexpr int RB = 2 * HR;                    // LDS rows per residue class: 8 RB >= 16 HR
  static constexpr int RMAP = 16 * HR;                 // row map entries (every staged row)
  static constexpr int NC = KB * NCH;                  // chunks of 8 kernel rows
  static constexpr int LDS_WORDS = 8 * RB * P;
  static_assert(P % 4 == 0 && (P / 4) % 2 == 1, "pitch must be 4 * odd");
  static_assert(TWP >= 64 && TWP % 16 == 0 && TWP < 128, "one lane = column pass plus halo groups");
  static_assert(NC == 1 || NC % 2 == 0, "chunks are processed in pairs");
};

__device__ __forceinline__ constexpr int mf_slot(int r, int rb) { return (r & 7) * rb + (r >> 3); }

struct MfTile {
  int64_t ibase, obase;
  int ox0, oy0;   // output tile origin
  int gx0, gy0;   // input index of staged element (0, 0) before the border map
  bool interior;  // the staged region lies inside the image
};

__device__ __forceinline__ MfTile mf_tile(const MfmaArgs &a, int t) {
  MfTile T;
  const int ty = t % a.nty;
  const int rest = t / a.nty;
  const int tx = rest % a.ntx;
  const int plane = rest / a.ntx;
  const int p2 = plane % a.o2, p3 = plane / a.o2;
  T.ibase = ((int64_t)(p2 + a.f2) + (int64_t)a.n2 * (p3 + a.f3)) * a.in_plane;
  T.obase = (int64_t)plane * a.out_plane;
  T.ox0 = tx * MF_TX;
  T.oy0 = ty * MF_TY;
  T.gx0 = a.f0 + T.ox0 + a.first0;
  T.gy0 = a.f1 + T.oy0 + a.first1;
  const int tw = MF_TX + a.kx - 1, th = MF_TY + a.ky - 1;
  T.interior = T.gx0 >= 0 && T.gx0 + tw <= a.n0 && T.gy0 >= 0 && T.gy0 + th <= a.n1;
  return T;
}

// Element e of a thread: e < MR is row wave + 4e at column lane; otherwise
// halo group g, row t/16 + 16i, column 64 + 16g + t%16.
template <class G>
struct MfElem {
  static constexpr bool lanecol(int e) { return e < G::MR; }
  static constexpr int g(int e) { return (e - G::MR) / G::HR; }
  static constexpr int i(int e) { return e < G::MR ? e : (e - G::MR) % G::HR; }
};

__device__ __forceinline__ int mf_wave() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// The value of x, hidden from loop-invariance analysis: switches on the
// element type and border style stay inside the tile loop instead of the
// loop being copied once per case.
__device__ __forceinline__ int mf_opaque(int x) {
  asm volatile("" : "+v"(x));
  return __builtin_amdgcn_readfirstlane(x);
}

// Raw element bits of element v + s of a plane through its buffer descriptor:
// v is the per-lane offset (one VGPR), s a wave-uniform offset (an SGPR), both
// in elements.  F64 is narrowed at once.  Word 3 of the descriptor is the
// gfx9 raw-buffer format (32-bit data format, no swizzle).
constexpr int MF_RSRC_W3 = 0x00020000;

template <int W>
__device__ __forceinline__ uint32_t mf_ld(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
  if constexpr (W == 4) return __builtin_amdgcn_raw_buffer_load_b32(r, v * 4, s * 4, 0);
  else if constexpr (W == 2) return __builtin_amdgcn_raw_buffer_load_b16(r, v * 2, s * 2, 0);
  else if constexpr (W == 1) return __builtin_amdgcn_raw_buffer_load_b8(r, v, s, 0);
  else return __float_as_uint((float)__builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, v * 8, s * 8, 0)));
}

// Border maps of tile T, one bmap per entry: rmap[r] = source row of staged
// row r (clamped to the tile's last row), cmap[c] = source column of staged
// column c (clamped likewise); -1 where Fill applies.
template <class G>
__device__ __forceinline__ void mf_maps(const MfmaArgs &a, const MfTile &T, int *rmap, int *cmap) {
  const int tw = MF_TX + a.kx - 1, th = MF_TY + a.ky - 1;
  for (int q = threadIdx.x; q < G::RMAP + G::TWP; q += 256) {
    if (q < G::RMAP)
      rmap[q] = bmap(a.style, T.gy0 + min(q, th - 1), a.n1);
    else
      cmap[q - G::RMAP] = bmap(a.style, T.gx0 + min(q - G::RMAP, tw - 1), a.n0);
  }
}

// Issue the loads of batch BT of tile T through its border maps, so every load
// is in bounds and none is predicated; elements outside the staged region or
// under Fill are replaced when the values are formed.  A plane holds fewer
// than 2^32 bytes (checked on the host), so in-plane offsets are 32-bit.
template <class G, int W, int BT>
__device__ __forceinline__ void mf_issue(const MfmaArgs &a, const MfTile &T, const int *rmap, const int *cmap,
                                         uint32_t (&raw)[G::JB]) {
  using E = MfElem<G>;
  const int lane = threadIdx.x & 63, wv = mf_wave();
  const int hh = (threadIdx.x >> 4) & 15, cl = threadIdx.x & 15;
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      (void *)((const char *)a.in + T.ibase * W), (short)0, (int)(a.in_plane * W), MF_RSRC_W3);
  const int sxl = cmap[lane];
  const uint32_t xl = sxl < 0 ? 0u : (uint32_t)sxl;
#pragma unroll
  for (int j = 0; j < G::JB; ++j) {
    constexpr int e0 = BT * G::JB;
    const int e = e0 + j;
    if (e >= G::NEL) {
      raw[j] = 0;
    } else if (E::lanecol(e)) {
      const int sy = __builtin_amdgcn_readfirstlane(rmap[wv + 4 * E::i(e)]);  // wave-uniform row
      raw[j] = mf_ld<W>(rs, xl, (uint32_t)(sy < 0 ? 0 : sy) * (uint32_t)a.n0);
    } else {
      const int sy = rmap[hh + 16 * E::i(e)];
      const int sx = cmap[64 + 16 * E::g(e) + cl];
      const uint32_t off = (sy < 0 || sx < 0) ? 0u : (uint32_t)sy * (uint32_t)a.n0 + (uint32_t)sx;
      raw[j] = mf_ld<W>(rs, off, 0u);
    }
  }
}

template <class G, int BT>
__device__ __forceinline__ void mf_issue_dt(const MfmaArgs &a, const MfTile &T, const int *rmap, const int *cmap,
                                            uint32_t (&raw)[G::JB]) {
  switch (mf_opaque(a.dt_in)) {
  case IFM_F32: mf_issue<G, 4, BT>(a, T, rmap, cmap, raw); break;
  case IFM_BF16:
  case IFM_F16: mf_issue<G, 2, BT>(a, T, rmap, cmap, raw); break;
  case IFM_F64: mf_issue<G, 8, BT>(a, T, rmap, cmap, raw); break;
  default: mf_issue<G, 1, BT>(a, T, rmap, cmap, raw); break;
  }
}

// Raw bits -> FP32 values (fill and zero padding applied), max |x| and the
// non-finite flag.
template <class G, int BT>
__device__ __forceinline__ void mf_values(const MfmaArgs &a, const MfTile &T, uint32_t (&raw)[G::JB], float &amax,
                                          bool &bad) {
  using E = MfElem<G>;
  switch (mf_opaque(a.dt_in)) {
  case IFM_U8:
#pragma unroll
    for (int j = 0; j < G::JB; ++j) raw[j] = __float_as_uint((float)raw[j]);
    break;
  case IFM_N0F8:
#pragma unroll
    for (int j = 0; j < G::JB; ++j) raw[j] = __float_as_uint(n0f8_to_f32(raw[j]));
    break;
  case IFM_BF16:
#pragma unroll
    for (int j = 0; j < G::JB; ++j) raw[j] <<= 16;
    break;
  case IFM_F16:
#pragma unroll
    for (int j = 0; j < G::JB; ++j) raw[j] = __float_as_uint((float)__builtin_bit_cast(_Float16, (uint16_t)raw[j]));
    break;
  default: break;
  }
  const int lane = threadIdx.x & 63, wv = mf_wave();
  const int hh = (threadIdx.x >> 4) & 15, cl = threadIdx.x & 15;
  const int tw = MF_TX + a.kx - 1, th = MF_TY + a.ky - 1;
  const bool fillable = mf_opaque(a.style) == IFM_FILL && !T.interior;
#pragma unroll
  for (int j = 0; j < G::JB; ++j) {
    constexpr int e0 = BT * G::JB;
    const int e = e0 + j;
    if (e >= G::NEL) continue;
    int row, col;
    bool ok;
    if (E::lanecol(e)) {
      row = wv + 4 * E::i(e);
      col = lane;
      ok = row < th;  // lane < 64 <= tw
    } else {
      row = hh + 16 * E::i(e);
      col = 64 + 16 * E::g(e) + cl;
      ok = row < th && col < tw;
    }
    float x = __uint_as_float(raw[j]);
    if (fillable) {
      const int gy = T.gy0 + row, gx = T.gx0 + col;
      if (gy < 0 || gy >= a.n1 || gx < 0 || gx >= a.n0) x = a.fill;
    }
    x = ok ? x : 0.f;
    raw[j] = __float_as_uint(x);
    const float ax = fabsf(x);
    bad |= !(ax <= 3.0e38f);  // Inf or NaN
    amax = fmaxf(amax, ax);
  }
}

// (hi | lo << 16): FP16 hi and lo parts of x.
__device__ __forceinline__ uint32_t mf_pack(float x) {
  const _Float16 h = (_Float16)x;
  const _Float16 l = (_Float16)(x - (float)h);
  return (uint32_t)__builtin_bit_cast(uint16_t, h) | ((uint32_t)__builtin_bit_cast(uint16_t, l) << 16);
}

// Values of batch BT into their LDS words: packed pairs of x * xs when pack,
// else the FP32 bits.  Row wv + 4i sits in LDS row (wv + 4(i&1)) RB + i/2 and
// row hh + 16i in LDS row (hh&7) RB + hh/8 + 2i.
template <class G, int BT>
__device__ __forceinline__ void mf_put(uint32_t *lds, const uint32_t (&raw)[G::JB], float xs, bool pack) {
  using E = MfElem<G>;
  const int lane = threadIdx.x & 63, wv = mf_wave();
  const int hh = (threadIdx.x >> 4) & 15, cl = threadIdx.x & 15;
  uint32_t *lp = lds + wv * G::RB * G::P + lane;
  uint32_t *hp = lds + ((hh & 7) * G::RB + (hh >> 3)) * G::P + 64 + cl;
#pragma unroll
  for (int j = 0; j < G::JB; ++j) {
    constexpr int e0 = BT * G::JB;
    const int e = e0 + j;
    if (e >= G::NEL) continue;
    const uint32_t w = pack ? mf_pack(__uint_as_float(raw[j]) * xs) : raw[j];
    const int i = E::i(e);
    if (E::lanecol(e))
      lp[(4 * (i & 1) * G::RB + (i >> 1)) * G::P] = w;
    else
      hp[2 * i * G::P + 16 * E::g(e)] = w;
  }
}

// FP32 words -> packed pairs, in place (words of unused rows are converted
// too and never read).
template <class G>
__device__ __forceinline__ void mf_convert(uint32_t *lds, float xs) {
  u32x4 *w4 = (u32x4 *)lds;
  for (int q = threadIdx.x; q < G::LDS_WORDS / 4; q += 256) {
    u32x4 w = w4[q];
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = mf_pack(__uint_as_float(w[k]) * xs);
    w4[q] = w;
  }
}

// Toeplitz fragments of chunk c: thi[u] / tlo[u] for kernel row u0 + u.
__device__ __forceinline__ void mf_tload(const uint16_t *ttab, int c, f16x8 (&thi)[MF_UCH], f16x8 (&tlo)[MF_UCH]) {
  const u16x8 *tt = (const u16x8 *)ttab + (size_t)c * (MF_UCH * 2 * 64) + (threadIdx.x & 63);
#pragma unroll
  for (int u = 0; u < MF_UCH; ++u) {
    thi[u] = __builtin_bit_cast(f16x8, tt[(2 * u) * 64]);
    tlo[u] = __builtin_bit_cast(f16x8, tt[(2 * u + 1) * 64]);
  }
}

// The first nu rows of chunk c (the others are left zero).
__device__ __forceinline__ void mf_tload_rows(const uint16_t *ttab, int c, int nu, f16x8 (&thi)[MF_UCH],
                                              f16x8 (&tlo)[MF_UCH]) {
  const u16x8 *tt = (const u16x8 *)ttab + (size_t)c * (MF_UCH * 2 * 64) + (threadIdx.x & 63);
#pragma unroll
  for (int u = 0; u < MF_UCH; ++u) {
    if (u < nu) {
      thi[u] = __builtin_bit_cast(f16x8, tt[(2 * u) * 64]);
      tlo[u] = __builtin_bit_cast(f16x8, tt[(2 * u + 1) * 64]);
    } else {
      thi[u] = f16x8{};
      tlo[u] = f16x8{};
    }
  }
}

// Waits for every outstanding vector-memory operation of the wave.  Used
// right after the Toeplitz fragments are loaded: the sweep then carries no
// wait for them, which (the counter being in order) would also wait for the
// stores and loads issued after them.
__device__ __forceinline__ void mf_wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // vmcnt(0)

// One chunk: 8 kernel rows against the 8 interleaved blocks.  lbase points at
// the lane's first word of LDS row (chunk row offset + lane row) and column
// (K block + wave + lane group).
template <class G>
__device__ __forceinline__ void mf_chunk(const uint32_t *lbase, const f16x8 (&thi)[MF_UCH],
                                         const f16x8 (&tlo)[MF_UCH], f32x4 (&acc)[MF_B]) {
  // Fragment r+1 is read while the MFMAs of fragment r run; the scheduling
  // barrier keeps that pipeline (and the register budget) instead of hoisting
  // every LDS read.
  const u32x4 *src0 = (const u32x4 *)lbase;
  u32x4 w0 = src0[0], w1 = src0[1];
#pragma unroll
  for (int r = 0; r < MF_B + MF_UCH - 1; ++r) {
    u32x4 n0 = w0, n1 = w1;
    if (r + 1 < MF_B + MF_UCH - 1) {
      const u32x4 *src = (const u32x4 *)(lbase + mf_slot(r + 1, G::RB) * G::P);
      n0 = src[0];
      n1 = src[1];
    }
    // 8 packed words -> 8 hi halves and 8 lo halves (v_perm_b32)
    u32x4 ph, pl;
    ph[0] = __builtin_amdgcn_perm(w0[1], w0[0], 0x05040100u);
    ph[1] = __builtin_amdgcn_perm(w0[3], w0[2], 0x05040100u);
    ph[2] = __builtin_amdgcn_perm(w1[1], w1[0], 0x05040100u);
    ph[3] = __builtin_amdgcn_perm(w1[3], w1[2], 0x05040100u);
    pl[0] = __builtin_amdgcn_perm(w0[1], w0[0], 0x07060302u);
    pl[1] = __builtin_amdgcn_perm(w0[3], w0[2], 0x07060302u);
    pl[2] = __builtin_amdgcn_perm(w1[1], w1[0], 0x07060302u);
    pl[3] = __builtin_amdgcn_perm(w1[3], w1[2], 0x07060302u);
    const f16x8 fhi = __builtin_bit_cast(f16x8, ph);
    const f16x8 flo = __builtin_bit_cast(f16x8, pl);
#pragma unroll
    for (int b = 0; b < MF_B; ++b) {
      const int u = r - b;
      if (u >= 0 && u < MF_UCH) acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fhi, thi[u], acc[b], 0, 0, 0);
    }
#pragma unroll
    for (int b = 0; b < MF_B; ++b) {
      const int u = r - b;
      if (u >= 0 && u < MF_UCH) acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(flo, thi[u], acc[b], 0, 0, 0);
    }
#pragma unroll
    for (int b = 0; b < MF_B; ++b) {
      const int u = r - b;
      if (u >= 0 && u < MF_UCH) acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fhi, tlo[u], acc[b], 0, 0, 0);
    }
    w0 = n0;
    w1 = n1;
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Both chunks of a 16-row kernel (one K block) as one sweep: kernel row U of
// block b reads fragment row rho = b + U (chunk 1's row r is fragment row
// r + 8), so fragment rows 0 .. 22 are read once each and every one between 7
// and 15 feeds all 8 blocks; the pipeline fills and drains once instead of
// once per chunk.
template <class G, int NU>
__device__ __forceinline__ void mf_rows16(const uint32_t *lbase, const f16x8 (&ahi)[MF_UCH],
                                          const f16x8 (&alo)[MF_UCH], const f16x8 (&bhi)[MF_UCH],
                                          const f16x8 (&blo)[MF_UCH], f32x4 (&acc)[MF_B]) {
  // NU kernel rows (15 or 16; a 15-row kernel skips the zero 16th row)
  constexpr int NR = MF_B + NU - 1;
  const u32x4 *src0 = (const u32x4 *)lbase;
  u32x4 w0 = src0[0], w1 = src0[1];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    u32x4 n0 = w0, n1 = w1;
    if (r + 1 < NR) {
      const u32x4 *src = (const u32x4 *)(lbase + mf_slot(r + 1, G::RB) * G::P);
      n0 = src[0];
      n1 = src[1];
    }
    u32x4 ph, pl;
    ph[0] = __builtin_amdgcn_perm(w0[1], w0[0], 0x05040100u);
    ph[1] = __builtin_amdgcn_perm(w0[3], w0[2], 0x05040100u);
    ph[2] = __builtin_amdgcn_perm(w1[1], w1[0], 0x05040100u);
    ph[3] = __builtin_amdgcn_perm(w1[3], w1[2], 0x05040100u);
    pl[0] = __builtin_amdgcn_perm(w0[1], w0[0], 0x07060302u);
    pl[1] = __builtin_amdgcn_perm(w0[3], w0[2], 0x07060302u);
    pl[2] = __builtin_amdgcn_perm(w1[1], w1[0], 0x07060302u);
    pl[3] = __builtin_amdgcn_perm(w1[3], w1[2], 0x07060302u);
    const f16x8 fhi = __builtin_bit_cast(f16x8, ph);
    const f16x8 flo = __builtin_bit_cast(f16x8, pl);
#pragma unroll
    for (int b = 0; b < MF_B; ++b) {
      const int u = r - b;
      if (u >= 0 && u < NU)
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fhi, u < MF_UCH ? ahi[u] : bhi[u - MF_UCH], acc[b], 0, 0, 0);
    }
#pragma unroll
    for (int b = 0; b < MF_B; ++b) {
      const int u = r - b;
      if (u >= 0 && u < NU)
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(flo, u < MF_UCH ? ahi[u] : bhi[u - MF_UCH], acc[b], 0, 0, 0);
    }
#pragma unroll
    for (int b = 0; b < MF_B; ++b) {
      const int u = r - b;
      if (u >= 0 && u < NU)
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fhi, u < MF_UCH ? alo[u] : blo[u - MF_UCH], acc[b], 0, 0, 0);
    }
    w0 = n0;
    w1 = n1;
    __builtin_amdgcn_sched_barrier(0);
  }
}

// D layout: lane holds rows yy = 4*(lane/16) + i of column a = lane%16, i.e.
// output rows oy0 + b + 8*yy.  Stores go through a buffer descriptor of the
// output plane: an output outside the image (or every output, keep = false)
// gets an offset past the descriptor's range, which the hardware drops.  So
// every path issues exactly 32 stores per lane, and the wait for the next
// tile's loads (issued before these stores) need not wait for the stores.
constexpr uint32_t MF_DROP = 0x80000000u;  // beyond any plane (< 2^31 bytes)

template <int DT>
__device__ __forceinline__ void mf_store_t(const MfmaArgs &a, const MfTile &T, const f32x4 (&acc)[MF_B], float osc,
                                           int ox, int lgrp, bool keep) {
  constexpr int ES = DT == IFM_F64 ? 8 : (DT == IFM_F32 ? 4 : 2);
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      (void *)((char *)a.out + T.obase * ES), (short)0, (int)(a.out_plane * ES), MF_RSRC_W3);
  const bool colok = keep && ox < a.o0;
#pragma unroll
  for (int b = 0; b < MF_B; ++b) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int oy = T.oy0 + b + MF_B * (4 * lgrp + i);
      const uint32_t off = (colok && oy < a.o1) ? ((uint32_t)oy * (uint32_t)a.o0 + (uint32_t)ox) * ES : MF_DROP;
      const float v = acc[b][i] * osc;
      if constexpr (DT == IFM_F32) {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rs, off, 0, 0);
      } else if constexpr (DT == IFM_F64) {
        const uint64_t d = __builtin_bit_cast(uint64_t, (double)v);
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(rs, 0, 0, 0)), d), rs, off, 0, 0);
      } else if constexpr (DT == IFM_BF16) {
        __builtin_amdgcn_raw_buffer_store_b16(f32_to_bf16(v), rs, off, 0, 0);
      } else {
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, (_Float16)v), rs, off, 0, 0);
      }
    }
  }
}

__device__ __forceinline__ void mf_store(const MfmaArgs &a, const MfTile &T, const f32x4 (&acc)[MF_B], float osc,
                                         bool keep) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ox = T.ox0 + 16 * wave + (lane & 15);
  switch (mf_opaque(a.dt_out)) {
  case IFM_F32: mf_store_t<IFM_F32>(a, T, acc, osc, ox, lane >> 4, keep); break;
  case IFM_F64: mf_store_t<IFM_F64>(a, T, acc, osc, ox, lane >> 4, keep); break;
  case IFM_BF16: mf_store_t<IFM_BF16>(a, T, acc, osc, ox, lane >> 4, keep); break;
  default: mf_store_t<IFM_F16>(a, T, acc, osc, ox, lane >> 4, keep); break;
  }
}

// Direct FP32 sums from the image (border map per tap) for one tile.  Inlined:
// a call taking the argument block by reference would move it to scratch and
// turn every pointer of the kernel into a flat pointer.
__device__ __forceinline__ void mf_direct(const MfmaArgs &a, const MfTile &T) {
  for (int q = threadIdx.x; q < MF_TX * MF_TY; q += 256) {
    const int ox = T.ox0 + q % MF_TX, oy = T.oy0 + q / MF_TX;
    if (ox >= a.o0 || oy >= a.o1) continue;
    float acc = 0.f;
    for (int u = 0; u < a.ky; ++u) {
      const int sy = bmap(a.style, T.gy0 + (oy - T.oy0) + u, a.n1);
      for (int t = 0; t < a.kx; ++t) {
        const int sx = bmap(a.style, T.gx0 + (ox - T.ox0) + t, a.n0);
        const float x = (sy < 0 || sx < 0) ? a.fill : load_f(a.in, T.ibase + (int64_t)sy * a.n0 + sx, a.dt_in);
        acc = fmaf(x, a.taps[t + u * a.kx], acc);
      }
    }
    store_f(a.out, T.obase + (int64_t)oy * a.o0 + ox, a.dt_out, acc);
  }
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// Arguments of global tile t: the launch's own (one image), or those of
// request tile_req[t] (batched serving); local = the tile's index within its
// image, r = its request.
struct MfBatch {
  const MfmaArgs *reqs;
  const int *tile_req;
  const int *tile_first;
};

template <bool BATCH>
__device__ __forceinline__ const MfmaArgs &mf_args(const MfmaArgs &a0, const MfBatch &bt, int t, int &local,
                                                   int &r) {
  if constexpr (BATCH) {
    r = bt.tile_req[t];
    local = t - bt.tile_first[r];
    return bt.reqs[r];
  } else {
    r = 0;
    local = t;
    return a0;
  }
}

template <int KB, int NCH, bool BATCH, int NU = MF_UCH * NCH>
__global__ void __launch_bounds__(256, 2) k_dense_mfma(MfmaArgs a0, MfBatch bt) {
  using G = MfmaGeom<KB, NCH>;
  constexpr int NC = G::NC;
  // One 32-bit word per tile element: the packed FP16 pair hi | lo << 16
  // (with two staging batches, first the FP32 value).
  extern __shared__ __attribute__((aligned(16))) uint32_t lds32[];
  __shared__ float red[2][MF_WAVES];
  __shared__ int redflag[2][MF_WAVES];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int lrow = lane & 15, lgrp = lane >> 4;

  // A contiguous run of tiles per workgroup; run lengths differ by at most one.
  const int t0 = (int)((int64_t)blockIdx.x * a0.ntiles / gridDim.x);
  const int t1 = (int)((int64_t)(blockIdx.x + 1) * a0.ntiles / gridDim.x);
  if (t0 >= t1) return;  // the whole workgroup leaves together

  int lf, cur;
  const MfmaArgs &af = mf_args<BATCH>(a0, bt, t0, lf, cur);
  f16x8 ahi[MF_UCH], alo[MF_UCH], bhi[MF_UCH], blo[MF_UCH];
  mf_tload(af.ttab, 0, ahi, alo);
  if constexpr (NC > 1) mf_tload_rows(af.ttab, 1, NU - MF_UCH * (NCH - 1), bhi, blo);
  mf_wait_vmem();

  // Border maps: with one staging batch they are written between the two
  // barriers of the previous tile and read by the loads issued after the
  // second one; with two batches, between the barriers at the tile's start.
  __shared__ int rmap[G::RMAP], cmap[G::TWP];
  uint32_t raw[G::JB];
  if constexpr (G::NB == 1) {
    const MfTile T0 = mf_tile(af, lf);
    mf_maps<G>(af, T0, rmap, cmap);
    __syncthreads();
    mf_issue_dt<G, 0>(af, T0, rmap, cmap, raw);
  }

  const uint32_t *lw = lds32 + lrow * G::P + 16 * wave + 8 * lgrp;
  int par = 0;
  for (int t = t0; t < t1; ++t, par ^= 1) {
    int lt, rt;
    const MfmaArgs &a = mf_args<BATCH>(a0, bt, t, lt, rt);
    const MfTile T = mf_tile(a, lt);
    float amax = 0.f;
    bool bad = false;
    if constexpr (G::NB == 1) {
      mf_values<G, 0>(a, T, raw, amax, bad);
    } else {
      static_assert(G::NB == 2, "two staging batches at most");
      __syncthreads();  // every wave has read the previous tile's fragments and maps
      mf_maps<G>(a, T, rmap, cmap);
      __syncthreads();
      mf_issue_dt<G, 0>(a, T, rmap, cmap, raw);
      mf_values<G, 0>(a, T, raw, amax, bad);
      mf_put<G, 0>(lds32, raw, 1.f, false);
      mf_issue_dt<G, 1>(a, T, rmap, cmap, raw);
      mf_values<G, 1>(a, T, raw, amax, bad);
      mf_put<G, 1>(lds32, raw, 1.f, false);
    }
    amax = wave_max(amax);
    const bool wbad = __any(bad);
    if (lane == 0) {
      red[par][wave] = amax;
      redflag[par][wave] = wbad;
    }
    // Also orders this tile's LDS writes after the previous tile's reads.
    __syncthreads();
    amax = fmaxf(fmaxf(red[par][0], red[par][1]), fmaxf(red[par][2], red[par][3]));
    const bool tilebad = redflag[par][0] | redflag[par][1] | redflag[par][2] | redflag[par][3];
    // Power-of-two scale putting the tile's largest |x| in [0.5, 1).
    int e = 0;
    if (amax > 0.f) frexpf(amax, &e);
    // A tile holding Inf or NaN, or one whose largest |x| lies outside
    // [2^-100, 2^100) (its scaled hi / lo halves would leave FP16's range or
    // its scale factors FP32's), or whose output scale 2^e * kinv would not
    // be a normal FP32 number, is summed directly in FP32; it then follows
    // the same path as the others with its matrix work skipped and its
    // stores dropped.  The decision is uniform over the workgroup.
    int ke = 0;
    frexpf(a.kinv, &ke);  // kinv = 2^(ke - 1)
    const bool direct = tilebad || e > 100 || e < -100 || e + ke > 120 || e + ke < -120;
    if (direct) {
      mf_direct(a, T);
      e = 0;
    }
    const float xs = ldexpf(1.0f, -e);
    const float xinv = ldexpf(1.0f, e);
    if constexpr (G::NB == 1) {
      if (!direct) mf_put<G, 0>(lds32, raw, xs, true);
      if (t + 1 < t1) {
        int ln, rn;
        const MfmaArgs &an = mf_args<BATCH>(a0, bt, t + 1, ln, rn);
        mf_maps<G>(an, mf_tile(an, ln), rmap, cmap);
      }
    } else {
      if (!direct) mf_convert<G>(lds32, xs);
    }
    __syncthreads();
    if constexpr (G::NB == 1) {
      if (t + 1 < t1) {
        int ln, rn;
        const MfmaArgs &an = mf_args<BATCH>(a0, bt, t + 1, ln, rn);
        mf_issue_dt<G, 0>(an, mf_tile(an, ln), rmap, cmap, raw);
      }
    }
    if constexpr (BATCH) {
      if (rt != cur) {  // another request: its Toeplitz fragments
        mf_tload(a.ttab, 0, ahi, alo);
        if constexpr (NC > 1) mf_tload_rows(a.ttab, 1, NU - MF_UCH * (NCH - 1), bhi, blo);
        mf_wait_vmem();
        cur = rt;
      }
    }

    f32x4 acc[MF_B];
#pragma unroll
    for (int b = 0; b < MF_B; ++b) acc[b] = f32x4{0.f, 0.f, 0.f, 0.f};
    if (!direct) {
      // chunk c = kb * NCH + ch covers K block kb and kernel rows 8ch .. 8ch+7
      if constexpr (NC == 1) {
        mf_chunk<G>(lw, ahi, alo, acc);
      } else if constexpr (NC == 2 && NCH == 2) {
        mf_rows16<G, NU>(lw, ahi, alo, bhi, blo, acc);
      } else if constexpr (NC == 2) {
        mf_chunk<G>(lw, ahi, alo, acc);
        mf_chunk<G>(lw + 32, bhi, blo, acc);
      } else {
        for (int c = 0; c < NC; c += 2) {
          const int c1 = c + 1;
          mf_chunk<G>(lw + (c % NCH) * G::P + 32 * (c / NCH), ahi, alo, acc);
          mf_tload(a.ttab, (c + 2) % NC, ahi, alo);
          mf_chunk<G>(lw + (c1 % NCH) * G::P + 32 * (c1 / NCH), bhi, blo, acc);
          mf_tload(a.ttab, (c + 3) % NC, bhi, blo);
        }
      }
    }
    mf_store(a, T, acc, xinv * a.kinv, !direct);
  }
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
static int mfma_kb(int kx) { return (16 + kx - 1 + 31) / 32; }

// Kernel-row chunks, padded to 1, 2 or 4.
static int mfma_nch(int ky) {
  const int n = (ky + MF_UCH - 1) / MF_UCH;
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

typedef void (*MfKernel)(MfmaArgs, MfBatch);
struct MfVariant {
  MfKernel fn;   // one image
  MfKernel fnb;  // batched requests
  size_t lds;
};

template <int KB, int NCH, int NU = MF_UCH * NCH>
static MfVariant mf_variant() {
  return {k_dense_mfma<KB, NCH, false, NU>, k_dense_mfma<KB, NCH, true, NU>,
          sizeof(uint32_t) * (size_t)MfmaGeom<KB, NCH>::LDS_WORDS};
}

static bool mf_select(int kb, int nch, MfVariant *v) {
  switch (kb * 16 + nch) {
  case 16 + 1: *v = mf_variant<1, 1>(); return true;
  case 16 + 2: *v = mf_variant<1, 2>(); return true;
  case 16 + 4: *v = mf_variant<1, 4>(); return true;
  case 32 + 1: *v = mf_variant<2, 1>(); return true;
  case 32 + 2: *v = mf_variant<2, 2>(); return true;
  case 32 + 4: *v = mf_variant<2, 4>(); return true;
  default: return false;
  }
}

static int device_attr(hipDeviceAttribute_t attr, int dflt) {
  int dev = 0, x = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&x, attr, dev) != hipSuccess) return dflt;
  return x;
}

bool dense_mfma_supported(const DenseKernel &k, const Dense4 &in, const Dense4 &out) {
  const bool twod = k.n[2] == 1 && k.n[3] == 1;
  // non-finite taps, or taps whose largest |k| lies outside [2^-100, 2^100):
  // the scaled FP16 halves would not hold them (the FP32 VALU path takes
  // such kernels); an all-zero kernel is fine
  float kmax = 0.f;
  for (float t : k.taps) {
    if (!std::isfinite(t)) return false;
    kmax = fmaxf(kmax, fabsf(t));
  }
  if (kmax != 0.f && !(kmax >= ldexpf(1.f, -100) && kmax < ldexpf(1.f, 100))) return false;
  if (!(twod && mfma_kb(k.n[0]) <= MF_MAXKB && k.n[1] <= MF_MAXCH * MF_UCH)) return false;
  // buffer descriptors address one plane with 31-bit byte offsets
  if ((double)in.n[0] * (double)in.n[1] * (double)dtype_size(in.dtype) >= 2147483648.0) return false;
  if ((double)out.n[0] * (double)out.n[1] * (double)dtype_size(out.dtype) >= 2147483648.0) return false;
  MfVariant v;
  if (!mf_select(mfma_kb(k.n[0]), mfma_nch(k.n[1]), &v)) return false;
  static int max_lds = -1;
  if (max_lds < 0) max_lds = device_attr(hipDeviceAttributeMaxSharedMemoryPerBlock, 64 * 1024);
  // dynamic LDS plus the block's static words (reductions, border maps)
  return v.lds + 2048 <= (size_t)max_lds;
}

// Host-side Toeplitz fragment table (FP16 hi/lo of the scaled taps), padded
// to the chunk count the kernel runs.
// Power-of-two scale putting the kernel's largest |tap| in [0.5, 1).
static void mfma_scale(const DenseKernel &k, float *ks, float *kinv) {
  float amax = 0.f;
  for (float t : k.taps) amax = fmaxf(amax, fabsf(t));
  int e = 0;
  if (amax > 0.f) frexpf(amax, &e);
  *ks = ldexpf(1.0f, -e);
  *kinv = ldexpf(1.0f, e);
}

// Table entry e (of kb * urows * 64 * 8): FP16 hi and lo of the scaled tap
// T_u[s][a] = k[s - a, u] for K block b, kernel row u, lane (a = lane % 16,
// s = 32 b + 8 (lane / 16) + i) and element i; returns its position.
__host__ __device__ inline size_t toeplitz_entry(const float *taps, int kx, int ky, int urows, float ks, int e,
                                                 uint16_t *hi, uint16_t *lo) {
  const int i = e & 7, lane = (e >> 3) & 63, bu = e >> 9;
  const int u = bu % urows, b = bu / urows;
  const int s = 32 * b + 8 * (lane >> 4) + i;
  const int t = s - (lane & 15);
  const float kv = (t >= 0 && t < kx && u < ky) ? taps[t + u * kx] * ks : 0.f;
  const _Float16 h = (_Float16)kv;
  const _Float16 l = (_Float16)(kv - (float)h);
  *hi = __builtin_bit_cast(uint16_t, h);
  *lo = __builtin_bit_cast(uint16_t, l);
  return ((((size_t)b * urows + u) * 2) * 64 + lane) * 8 + i;
}

void dense_mfma_table(const DenseKernel &k, std::vector<uint16_t> &tab, float *kinv) {
  const int kx = k.n[0], ky = k.n[1];
  const int kb = mfma_kb(kx), urows = mfma_nch(ky) * MF_UCH;
  float ks;
  mfma_scale(k, &ks, kinv);
  const int entries = kb * urows * 64 * 8;
  tab.assign((size_t)entries * 2, 0);
  for (int e = 0; e < entries; ++e) {
    uint16_t h, l;
    const size_t pos = toeplitz_entry(k.taps.data(), kx, ky, urows, ks, e, &h, &l);
    tab[pos] = h;
    tab[pos + 64 * 8] = l;
  }
}

// The same tables built on the device, one workgroup per job (batched serving).
__global__ void __launch_bounds__(256) k_toeplitz(const ToeplitzJob *__restrict__ jobs) {
  const ToeplitzJob &j = jobs[blockIdx.x];
  const int entries = j.kb * j.urows * 64 * 8;
  for (int e = threadIdx.x; e < entries; e += 256) {
    uint16_t h, l;
    const size_t pos = toeplitz_entry(j.taps, j.kx, j.ky, j.urows, j.ks, e, &h, &l);
    j.table[pos] = h;
    j.table[pos + 64 * 8] = l;
  }
}

size_t dense_mfma_job(const DenseKernel &k, ToeplitzJob *job, float *kinv) {
  job->kx = k.n[0];
  job->ky = k.n[1];
  job->kb = mfma_kb(k.n[0]);
  job->urows = mfma_nch(k.n[1]) * MF_UCH;
  mfma_scale(k, &job->ks, kinv);
  return (size_t)job->kb * job->urows * 64 * 8 * 2 * sizeof(uint16_t);
}

hipError_t launch_toeplitz(const ToeplitzJob *jobs, int njobs, hipStream_t st) {
  if (njobs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_toeplitz, dim3(njobs), dim3(256), 0, st, jobs);
  return hipGetLastError();
}

int dense_mfma_variant(const DenseKernel &k) { return mfma_kb(k.n[0]) * 16 + mfma_nch(k.n[1]); }

bool dense_mfma_args(const Dense4 &in, Dense4 &out, const int64_t *out_first, const DenseKernel &k,
                     const uint16_t *d_table, const float *d_taps, float kinv, int style, float fill, MfmaArgs *ap) {
  MfVariant v;
  if (!mf_select(mfma_kb(k.n[0]), mfma_nch(k.n[1]), &v)) return false;
  MfmaArgs &a = *ap;
  a.in = in.data;
  a.out = out.data;
  a.ttab = d_table;
  a.taps = d_taps;
  a.kinv = kinv;
  a.n0 = (int)in.n[0];
  a.n1 = (int)in.n[1];
  a.o0 = (int)out.n[0];
  a.o1 = (int)out.n[1];
  a.f0 = (int)out_first[0];
  a.f1 = (int)out_first[1];
  a.kx = k.n[0];
  a.ky = k.n[1];
  a.first0 = k.first[0];
  a.first1 = k.first[1];
  a.in_plane = in.n[0] * in.n[1];
  a.out_plane = out.n[0] * out.n[1];
  a.o2 = (int)out.n[2];
  a.n2 = (int)in.n[2];
  a.f2 = (int)out_first[2] + k.first[2];
  a.f3 = (int)out_first[3] + k.first[3];
  a.style = style;
  a.dt_in = in.dtype;
  a.dt_out = out.dtype;
  a.fill = fill;
  a.ntx = (a.o0 + MF_TX - 1) / MF_TX;
  a.nty = (a.o1 + MF_TY - 1) / MF_TY;
  const int64_t ntiles = (int64_t)a.ntx * a.nty * out.n[2] * out.n[3];
  if (ntiles >= (int64_t)1 << 31) return false;
  a.ntiles = (int)ntiles;
  return true;
}

static int64_t mf_grid(int64_t ntiles) {
  static int cus = -1;
  if (cus < 0) cus = device_attr(hipDeviceAttributeMultiprocessorCount, 256);
  return ntiles < 2 * (int64_t)cus ? ntiles : 2 * (int64_t)cus;
}

hipError_t launch_dense_mfma(const Dense4 &in, Dense4 &out, const int64_t *out_first, const DenseKernel &k,
                             const uint16_t *d_table, const float *d_taps, float kinv, int style, float fill,
                             hipStream_t st) {
  if (out.count() == 0) return hipSuccess;
  MfVariant v;
  MfmaArgs a;
  if (!mf_select(mfma_kb(k.n[0]), mfma_nch(k.n[1]), &v) ||
      !dense_mfma_args(in, out, out_first, k, d_table, d_taps, kinv, style, fill, &a))
    return hipErrorInvalidValue;
  // a 15-row kernel in one K block sweeps 15 rows, not the padded 16
  if (mfma_kb(k.n[0]) == 1 && k.n[1] == 15) v = mf_variant<1, 2, 15>();
  MfBatch none = {nullptr, nullptr, nullptr};
  void *args[] = {&a, &none};
  return hipLaunchKernel((const void *)v.fn, dim3((unsigned)mf_grid(a.ntiles)), dim3(256), args, v.lds, st);
}

hipError_t launch_dense_mfma_batch(int variant, const MfmaArgs *reqs, const int *tile_req, const int *tile_first,
                                   int ntiles, hipStream_t st) {
  if (ntiles == 0) return hipSuccess;
  MfVariant v;
  if (!mf_select(variant / 16, variant % 16, &v)) return hipErrorInvalidValue;
  MfmaArgs a = {};
  a.ntiles = ntiles;  // the launch's tile count; everything else comes from reqs
  MfBatch bt = {reqs, tile_req, tile_first};
  void *args[] = {&a, &bt};
  return hipLaunchKernel((const void *)v.fnb, dim3((unsigned)mf_grid(ntiles)), dim3(256), args, v.lds, st);
}

}  // namespace ifm
