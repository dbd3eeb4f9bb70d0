// ifm_serve.hip — batched serving (gfx950): many 2-D images of uneven sizes
// and mixed ops per launch, and the device-resident request loop.
//
// ifm_batch_run turns n requests into one descriptor block: per-request
// argument structs, FP32 taps, matrix-core Toeplitz tables and the
// tile -> request tables.  The block is written into a pinned staging slot,
// copied to the device with one hipMemcpyAsync, and read by one launch per op
// class (k_sep_batch, k_median_batch<k>, k_dense_mfma<..., batched>).  A ring
// of slots, each guarded by an event, lets batches be queued while earlier
// ones still run.
//
// ifm_serve admits requests in arrival order while their bytes fit the HBM
// budget (an arena of that size holds generated inputs and outputs), launches
// each admitted group as a batch without waiting for the earlier batches, and
// retires batches as their events complete.  Its per-call mode runs the same
// requests one at a time through the single-image entry points.
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <thread>

#include "k_tiles.h"

using namespace ifm;

namespace ifm {

// ---------------------------------------------------------------------------
// Synthetic inputs: seeded uniform [0,1) (splitmix64 of seed and index, top
// 24 bits times 2^-24).
// ---------------------------------------------------------------------------
__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct FillReq {
  void *data;
  int64_t count;
  uint64_t key;  // splitmix64(seed)
  int dtype;
  int pad;
};

constexpr int FILL_TILE = 256 * 64;  // elements per workgroup

__device__ __forceinline__ void fill_range(const FillReq &f, int64_t i0) {
  for (int j = 0; j < FILL_TILE / 256; ++j) {
    const int64_t i = i0 + threadIdx.x + 256 * (int64_t)j;
    if (i >= f.count) return;
    const uint32_t u = (uint32_t)(splitmix64(f.key + (uint64_t)i) >> 40);  // 24 bits
    const float x = (float)u * (1.0f / 16777216.0f);
    switch (f.dtype) {
    case IFM_U8:
    case IFM_N0F8: ((uint8_t *)f.data)[i] = (uint8_t)(u >> 16); break;
    case IFM_F64: ((double *)f.data)[i] = (double)x; break;
    default: store_f(f.data, i, f.dtype, x); break;
    }
  }
}

__global__ void __launch_bounds__(256) k_fill_batch(const FillReq *__restrict__ reqs, const int *__restrict__ tile_req,
                                                    const int *__restrict__ tile_first) {
  const int g = blockIdx.x;
  const int r = tile_req[g];
  fill_range(reqs[r], (int64_t)(g - tile_first[r]) * FILL_TILE);
}

__global__ void __launch_bounds__(256) k_fill_one(FillReq f) {
  fill_range(f, (int64_t)blockIdx.x * FILL_TILE);
}

hipError_t launch_fill_uniform(void *data, int dtype, int64_t count, uint64_t seed, hipStream_t st) {
  if (count == 0) return hipSuccess;
  FillReq f = {data, count, splitmix64(seed), dtype, 0};
  const int64_t tiles = (count + FILL_TILE - 1) / FILL_TILE;
  if (tiles >= (int64_t)1 << 31) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fill_one, dim3((unsigned)tiles), dim3(256), 0, st, f);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Descriptor block builder: data is appended to a host byte vector; device
// pointers into the block are recorded as fixups and patched once the
// block's device address is known.
// ---------------------------------------------------------------------------
struct Block {
  std::vector<char> buf;
  std::vector<std::pair<size_t, size_t>> fix;  // (position of a pointer, target offset)
  size_t put(const void *p, size_t bytes, size_t align = 256) {
    size_t off = (buf.size() + align - 1) / align * align;
    buf.resize(off + bytes);
    if (bytes) memcpy(buf.data() + off, p, bytes);
    return off;
  }
  size_t reserve(size_t bytes, size_t align = 256) {
    size_t off = (buf.size() + align - 1) / align * align;
    buf.resize(off + bytes);
    return off;
  }
  void ptr(size_t pos, size_t target) { fix.push_back({pos, target}); }
  void patch(char *dev) {
    for (auto &f : fix) {
      void *p = dev + f.second;
      memcpy(buf.data() + f.first, &p, sizeof(p));
    }
  }
};

void batch_release(ifm_ctx *ctx) {
  for (auto &s : ctx->bslots) {
    if (s.ev) {
      (void)hipEventSynchronize(s.ev);
      (void)hipEventDestroy(s.ev);
    }
    if (s.host) (void)hipHostFree(s.host);
    if (s.dev) (void)hipFree(s.dev);
  }
  ctx->bslots.clear();
}

}  // namespace ifm

namespace {

constexpr int NSLOTS = 8;

int sfail(ifm_ctx *ctx, int code, const std::string &msg) {
  ctx->err = msg;
  return code;
}

#define SHIP(expr)                                                                                           \
  do {                                                                                                       \
    hipError_t e_ = (expr);                                                                                  \
    if (e_ != hipSuccess) return sfail(ctx, IFM_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

bool float_dtype(int dt) { return dt == IFM_F32 || dt == IFM_F64 || dt == IFM_BF16 || dt == IFM_F16; }

// A request's geometry as a canonical 2-D Dense4 (column-major, unit stride).
int req_view(ifm_ctx *ctx, const ifm_array &a, const char *name, Dense4 &d) {
  if (a.ndim != 2 || a.device != 1 || !a.data)
    return sfail(ctx, IFM_E_UNSUPPORTED, std::string(name) + ": batched requests are 2-D device arrays");
  if (a.dtype < IFM_U8 || a.dtype > IFM_F16) return sfail(ctx, IFM_E_ARG, std::string(name) + ": unknown dtype");
  if (a.size[0] <= 0 || a.size[1] <= 0 || a.stride[0] != 1 || a.stride[1] != a.size[0])
    return sfail(ctx, IFM_E_UNSUPPORTED, std::string(name) + ": batched requests are dense column-major");
  if (a.size[0] * a.size[1] >= ((int64_t)1 << 31) / 8)
    return sfail(ctx, IFM_E_UNSUPPORTED, std::string(name) + ": image too large for a batch");
  d.data = a.data;
  d.dtype = a.dtype;
  d.n[0] = a.size[0];
  d.n[1] = a.size[1];
  d.n[2] = d.n[3] = 1;
  return IFM_OK;
}

int check_request(ifm_ctx *ctx, const ifm_request &q, Dense4 &in, Dense4 &out) {
  int rc;
  if ((rc = req_view(ctx, q.in, "in", in)) || (rc = req_view(ctx, q.out, "out", out))) return rc;
  if (in.n[0] != out.n[0] || in.n[1] != out.n[1]) return sfail(ctx, IFM_E_DIM, "out must have the size of in");
  const int st = q.border.style;
  if (st < IFM_REPLICATE || st > IFM_FILL) return sfail(ctx, IFM_E_ARG, "unknown border style");
  if (st == IFM_REFLECT && (in.n[0] == 1 || in.n[1] == 1))
    return sfail(ctx, IFM_E_ARG, "reflect padding of a length-1 axis");
  switch (q.op) {
  case IFM_OP_SEPARABLE: {
    for (int part = 0; part < (q.nfactors2 > 0 ? 2 : 1); ++part) {
      const int nf = part ? q.nfactors2 : q.nfactors;
      const ifm_factor *fs = part ? q.factors2 : q.factors;
      if (nf < 1 || nf > 2) return sfail(ctx, IFM_E_ARG, "separable requests take 1 or 2 factors");
      int seen = 0;
      for (int f = 0; f < nf; ++f) {
        const ifm_factor &fa = fs[f];
        if (fa.axis < 0 || fa.axis > 1 || (seen >> fa.axis) & 1)
          return sfail(ctx, IFM_E_UNSUPPORTED, "separable requests take at most one factor per axis (0 or 1)");
        if (fa.len < 1 || fa.len > S2_MAXK || !fa.taps) return sfail(ctx, IFM_E_UNSUPPORTED, "factor of 1..64 taps");
        seen |= 1 << fa.axis;
      }
    }
    if (!float_dtype(out.dtype)) return sfail(ctx, IFM_E_ARG, "floating-point output eltype required");
    if (q.nfactors2 > 0) {
      Dense4 o2;
      if ((rc = req_view(ctx, q.out2, "out2", o2))) return rc;
      if (o2.n[0] != in.n[0] || o2.n[1] != in.n[1]) return sfail(ctx, IFM_E_DIM, "out2 must have the size of in");
      if (!float_dtype(o2.dtype)) return sfail(ctx, IFM_E_ARG, "floating-point output eltype required");
    }
    return IFM_OK;
  }
  case IFM_OP_DENSE:
    if (q.kernel.ndim != 2 || !q.kernel.taps) return sfail(ctx, IFM_E_UNSUPPORTED, "dense requests take 2-D kernels");
    if (out.dtype == IFM_U8 || out.dtype == IFM_N0F8) return sfail(ctx, IFM_E_ARG, "floating-point output eltype required");
    return IFM_OK;
  case IFM_OP_MEDIAN:
    if (q.window != 3 && q.window != 5 && q.window != 7)
      return sfail(ctx, IFM_E_UNSUPPORTED, "batched median windows are 3, 5 or 7 square");
    if (out.dtype != in.dtype) return sfail(ctx, IFM_E_ARG, "median output eltype must equal the input eltype");
    return IFM_OK;
  default: return sfail(ctx, IFM_E_ARG, "unknown op");
  }
}

DenseKernel dense_of(const ifm_kernel &k) {
  DenseKernel dk;
  dk.ndim = 2;
  for (int d = 0; d < MAXD; ++d) {
    dk.n[d] = d < 2 ? (int)k.size[d] : 1;
    dk.first[d] = d < 2 ? (int)k.first[d] : 0;
  }
  dk.taps.resize((size_t)dk.n[0] * dk.n[1]);
  for (size_t i = 0; i < dk.taps.size(); ++i) dk.taps[i] = (float)k.taps[i];
  return dk;
}

// Launch list of one batch, with offsets into the descriptor block.
struct Group {
  int kind;     // 0 fill, 1 separable, 2 median, 3 dense, 4 circular median band, 5 Toeplitz tables
  int param;    // median: k * 2 + pack; dense: variant
  size_t reqs, tile_req, tile_first;
  int ntiles;
  size_t lds;
};

// Builds the descriptor block of requests reqs[idx[0..n)] and their launch
// list.  gen[i] != 0: member i's input is generated first (seed idx[i]).
// Medians outside the batched launches: `alone` run the single-image kernel,
// `band` get their circular edge band recomputed after the batch.
struct Extra {
  std::vector<int> alone, band;
};

int build_batch(ifm_ctx *ctx, const ifm_request *reqs, const int *idx, int n, const char *gen, Block &blk,
                std::vector<Group> &groups, Extra &ex) {
  std::vector<int> &fallback = ex.alone, &band = ex.band;
  std::vector<Dense4> vin(n), vout(n);
  for (int i = 0; i < n; ++i) {
    int rc = check_request(ctx, reqs[idx[i]], vin[i], vout[i]);
    if (rc) return rc;
  }
  auto tiles_group = [&](int kind, int param, const std::vector<int> &members, const std::vector<int> &ntile,
                         size_t reqs_off, size_t lds) {
    std::vector<int> first(members.size()), treq;
    int tot = 0;
    for (size_t m = 0; m < members.size(); ++m) {
      first[m] = tot;
      for (int t = 0; t < ntile[m]; ++t) treq.push_back((int)m);
      tot += ntile[m];
    }
    Group g;
    g.kind = kind;
    g.param = param;
    g.reqs = reqs_off;
    g.tile_req = blk.put(treq.data(), treq.size() * sizeof(int));
    g.tile_first = blk.put(first.data(), first.size() * sizeof(int));
    g.ntiles = tot;
    g.lds = lds;
    if (tot > 0) groups.push_back(g);
  };
  // Inputs to generate (the first launch of the batch, on the same stream).
  if (gen) {
    std::vector<FillReq> fr;
    std::vector<int> nt, mem;
    for (int i = 0; i < n; ++i) {
      if (!gen[i]) continue;
      FillReq f = {vin[i].data, vin[i].count(), splitmix64((uint64_t)idx[i]), vin[i].dtype, 0};
      fr.push_back(f);
      nt.push_back((int)((f.count + FILL_TILE - 1) / FILL_TILE));
      mem.push_back(i);
    }
    if (!fr.empty()) tiles_group(0, 0, mem, nt, blk.put(fr.data(), fr.size() * sizeof(FillReq)), 0);
  }
  // Separable.
  {
    std::vector<int> mem, nt;
    std::vector<SepReq> sr;
    std::vector<std::pair<size_t, size_t>> tapoffs;  // (offset of taps0, taps1)
    size_t lds = 0;
    for (int i = 0; i < n; ++i)
    for (int part = 0; part < 2; ++part) {
      const ifm_request &q = reqs[idx[i]];
      if (q.op != IFM_OP_SEPARABLE || (part == 1 && q.nfactors2 <= 0)) continue;
      const int nf = part ? q.nfactors2 : q.nfactors;
      const ifm_factor *fs = part ? q.factors2 : q.factors;
      float t0[S2_MAXK] = {1.f}, t1[S2_MAXK] = {1.f};
      int k0 = 1, k1 = 1, f0 = 0, f1 = 0;
      for (int f = 0; f < nf; ++f) {
        const ifm_factor &fa = fs[f];
        float *t = fa.axis == 0 ? t0 : t1;
        for (int j = 0; j < fa.len; ++j) t[j] = (float)fa.taps[j];
        if (fa.axis == 0) {
          k0 = fa.len;
          f0 = (int)fa.first;
        } else {
          k1 = fa.len;
          f1 = (int)fa.first;
        }
      }
      SepReq s;
      s.in = vin[i].data;
      s.out = part ? q.out2.data : vout[i].data;
      s.taps0 = s.taps1 = nullptr;
      s.n0 = (int)vin[i].n[0];
      s.n1 = (int)vin[i].n[1];
      s.k0 = k0;
      s.k1 = k1;
      s.first0 = f0;
      s.first1 = f1;
      s.style = q.border.style;
      s.dt_in = vin[i].dtype;
      s.dt_out = part ? q.out2.dtype : vout[i].dtype;
      s.fill = (float)q.border.fill;
      s.ty = sep2d_tile_h(ctx, s.dt_in, k0, k1, (int64_t)s.n0 * s.n1);
      sr.push_back(s);
      tapoffs.push_back({blk.put(t0, sizeof(float) * k0, 16), blk.put(t1, sizeof(float) * k1, 16)});
      lds = std::max(lds, sizeof(float) * sep2d_lds_floats(k0, k1, s.ty));
      mem.push_back(i);
      nt.push_back((int)(((s.n0 + S2_TX - 1) / S2_TX) * ((s.n1 + s.ty - 1) / s.ty)));
    }
    if (!sr.empty()) {
      size_t off = blk.put(sr.data(), sr.size() * sizeof(SepReq));
      for (size_t m = 0; m < sr.size(); ++m) {
        blk.ptr(off + m * sizeof(SepReq) + offsetof(SepReq, taps0), tapoffs[m].first);
        blk.ptr(off + m * sizeof(SepReq) + offsetof(SepReq, taps1), tapoffs[m].second);
      }
      tiles_group(1, 0, mem, nt, off, lds);
    }
  }
  // Median, per (window, key width); small or F64 images run alone, and the
  // edge band of circular ones is recomputed after the batch (band list).
  for (int k = 3; k <= 7; k += 2)
    for (int pack = 0; pack < 2; ++pack) {
      std::vector<int> mem, nt;
      std::vector<MedianArgs> ma;
      for (int i = 0; i < n; ++i) {
        const ifm_request &q = reqs[idx[i]];
        if (q.op != IFM_OP_MEDIAN || q.window != k) continue;
        const bool net = vin[i].n[0] >= k && vin[i].n[1] >= k && vin[i].dtype != IFM_F64;
        if (!net) {
          if (pack == 0) fallback.push_back(i);
          continue;
        }
        if (pack == 0 && q.border.style == IFM_CIRCULAR) band.push_back(i);
        if ((vin[i].dtype != IFM_F32) != (pack == 1)) continue;
        MedianArgs a = {};
        a.in = vin[i].data;
        a.out = vout[i].data;
        a.n0 = a.o0 = (int)vin[i].n[0];
        a.n1 = a.o1 = (int)vin[i].n[1];
        a.in_plane = a.out_plane = vin[i].n[0] * vin[i].n[1];
        a.nplanes = a.o2 = a.n2 = 1;
        a.style = q.border.style;
        a.dt = vin[i].dtype;
        a.odt = vin[i].dtype;
        a.fillkey = median_fill_key(vin[i].dtype, q.border.fill);
        int bx, by;
        median_batch_tile(k, pack == 1, &bx, &by);
        ma.push_back(a);
        mem.push_back(i);
        nt.push_back((int)(((a.o0 + bx - 1) / bx) * ((a.o1 + by - 1) / by)));
      }
      if (!ma.empty()) tiles_group(2, k * 2 + pack, mem, nt, blk.put(ma.data(), ma.size() * sizeof(MedianArgs)), 0);
    }
  // Dense (matrix cores), per kernel variant.
  {
    std::map<int, std::vector<int>> byvar;
    std::vector<DenseKernel> dks(n);
    for (int i = 0; i < n; ++i) {
      const ifm_request &q = reqs[idx[i]];
      if (q.op != IFM_OP_DENSE) continue;
      dks[i] = dense_of(q.kernel);
      if (!dense_mfma_supported(dks[i], vin[i], vout[i]))
        return sfail(ctx, IFM_E_UNSUPPORTED, "dense requests need a 2-D kernel of at most 49 x 32 taps below 2^100 in magnitude");
      byvar[dense_mfma_variant(dks[i])].push_back(i);
    }
    // Toeplitz tables are built on the device by one launch ahead of the dense ones.
    std::vector<ToeplitzJob> tjobs;
    std::vector<std::pair<size_t, size_t>> tjoboffs;  // (taps, table) offsets
    for (auto &kv : byvar) {
      std::vector<MfmaArgs> ma;
      std::vector<int> nt;
      std::vector<std::pair<size_t, size_t>> offs;  // (table, taps)
      for (int i : kv.second) {
        ToeplitzJob tj = {};
        float kinv = 1.f;
        const size_t tbytes = dense_mfma_job(dks[i], &tj, &kinv);
        offs.push_back({blk.reserve(tbytes), blk.put(dks[i].taps.data(), dks[i].taps.size() * sizeof(float), 16)});
        tjobs.push_back(tj);
        tjoboffs.push_back({offs.back().second, offs.back().first});
        MfmaArgs a;
        const int64_t of[MAXD] = {0, 0, 0, 0};
        const ifm_request &q = reqs[idx[i]];
        if (!dense_mfma_args(vin[i], vout[i], of, dks[i], nullptr, nullptr, kinv, q.border.style,
                             (float)q.border.fill, &a))
          return sfail(ctx, IFM_E_UNSUPPORTED, "dense request shape");
        ma.push_back(a);
        nt.push_back(a.ntiles);
      }
      size_t off = blk.put(ma.data(), ma.size() * sizeof(MfmaArgs));
      for (size_t m = 0; m < ma.size(); ++m) {
        blk.ptr(off + m * sizeof(MfmaArgs) + offsetof(MfmaArgs, ttab), offs[m].first);
        blk.ptr(off + m * sizeof(MfmaArgs) + offsetof(MfmaArgs, taps), offs[m].second);
      }
      tiles_group(3, kv.first, kv.second, nt, off, 0);
    }
    if (!tjobs.empty()) {
      const size_t off = blk.put(tjobs.data(), tjobs.size() * sizeof(ToeplitzJob));
      for (size_t m = 0; m < tjobs.size(); ++m) {
        blk.ptr(off + m * sizeof(ToeplitzJob) + offsetof(ToeplitzJob, taps), tjoboffs[m].first);
        blk.ptr(off + m * sizeof(ToeplitzJob) + offsetof(ToeplitzJob, table), tjoboffs[m].second);
      }
      Group g = {5, 0, off, 0, 0, (int)tjobs.size(), 0};
      // before the first dense launch of this batch
      auto first_dense = std::find_if(groups.begin(), groups.end(), [](const Group &x) { return x.kind == 3; });
      groups.insert(first_dense, g);
    }
  }
  // Circular median edge bands, per window, after the network launches.
  for (int k = 3; k <= 7; k += 2) {
    std::vector<BandJob> jobs;
    std::vector<int> mem, nt;
    for (int i : band) {
      if (reqs[idx[i]].window != k) continue;
      BandJob bj[4];
      const int64_t of[MAXD] = {0, 0, 0, 0};
      const int nj = median_band_jobs(vin[i], vout[i], of, k, bj);
      for (int s = 0; s < nj; ++s) {
        mem.push_back((int)jobs.size());
        nt.push_back((bj[s].w * bj[s].h + 255) / 256);
        jobs.push_back(bj[s]);
      }
    }
    if (!jobs.empty()) {
      std::vector<int> ids(jobs.size());
      for (size_t m = 0; m < ids.size(); ++m) ids[m] = (int)m;
      tiles_group(4, k, ids, nt, blk.put(jobs.data(), jobs.size() * sizeof(BandJob)), 0);
    }
  }
  band.clear();  // handled by the launches above
  return IFM_OK;
}

// Copies the block into the next staging slot and queues every launch.
int launch_batch(ifm_ctx *ctx, const ifm_request *reqs, const int *idx, Block &blk, const std::vector<Group> &groups,
                 const Extra &ex, hipEvent_t *done) {
  if (ctx->bslots.size() < NSLOTS) ctx->bslots.resize(NSLOTS);
  ifm_ctx::BatchSlot &s = ctx->bslots[ctx->bnext];
  ctx->bnext = (ctx->bnext + 1) % NSLOTS;
  if (!s.ev) SHIP(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
  if (s.pending) SHIP(hipEventSynchronize(s.ev));  // the slot's previous copy and launches are done
  s.pending = false;
  const size_t bytes = blk.buf.size() ? blk.buf.size() : 16;
  if (s.cap < bytes) {
    if (s.host) SHIP(hipHostFree(s.host));
    if (s.dev) SHIP(hipFree(s.dev));
    s.host = s.dev = nullptr;
    s.cap = 0;
    size_t cap = 1 << 20;
    while (cap < bytes) cap <<= 1;
    SHIP(hipHostMalloc(&s.host, cap, hipHostMallocDefault));
    SHIP(hipMalloc(&s.dev, cap));
    s.cap = cap;
  }
  blk.patch((char *)s.dev);
  memcpy(s.host, blk.buf.data(), blk.buf.size());
  SHIP(hipMemcpyAsync(s.dev, s.host, blk.buf.size(), hipMemcpyHostToDevice, ctx->stream));
  char *d = (char *)s.dev;
  for (const Group &g : groups) {
    const int *treq = (const int *)(d + g.tile_req), *tfirst = (const int *)(d + g.tile_first);
    switch (g.kind) {
    case 0:
      hipLaunchKernelGGL(k_fill_batch, dim3(g.ntiles), dim3(256), 0, ctx->stream, (const FillReq *)(d + g.reqs), treq,
                         tfirst);
      SHIP(hipGetLastError());
      break;
    case 1: SHIP(launch_sep_batch((const SepReq *)(d + g.reqs), treq, tfirst, g.ntiles, g.lds, ctx->stream)); break;
    case 2:
      SHIP(launch_median_batch(g.param / 2, g.param % 2 == 1, (const MedianArgs *)(d + g.reqs), treq, tfirst,
                               g.ntiles, ctx->stream));
      break;
    case 3:
      SHIP(launch_dense_mfma_batch(g.param, (const MfmaArgs *)(d + g.reqs), treq, tfirst, g.ntiles, ctx->stream));
      break;
    case 5: SHIP(launch_toeplitz((const ToeplitzJob *)(d + g.reqs), g.ntiles, ctx->stream)); break;
    default:
      SHIP(launch_median_band_batch(g.param, (const BandJob *)(d + g.reqs), treq, tfirst, g.ntiles, ctx->stream));
      break;
    }
  }
  for (int pass = 0; pass < 2; ++pass)
    for (int i : pass ? ex.band : ex.alone) {
      const ifm_request &q = reqs[idx[i]];
      Dense4 in, out;
      int rc = check_request(ctx, q, in, out);
      if (rc) return rc;
      const int k = (int)q.window;
      const int64_t of[MAXD] = {0, 0, 0, 0};
      const int wf[MAXD] = {-(k / 2), -(k / 2), 0, 0}, wl[MAXD] = {k, k, 1, 1};
      if (pass)
        SHIP(median_circular_band(in, out, of, k, q.border.fill, ctx->stream));
      else
        SHIP(launch_median(in, out, of, wf, wl, q.border.style, q.border.fill, ctx->stream));
    }
  SHIP(hipEventRecord(s.ev, ctx->stream));
  s.pending = true;
  if (done) *done = s.ev;
  return IFM_OK;
}

// First-fit allocator over one device allocation (256-byte granules).
struct Arena {
  char *base = nullptr;
  uint64_t size = 0;
  std::map<uint64_t, uint64_t> free_;  // offset -> length
  void init(char *b, uint64_t n) {
    base = b;
    size = n;
    free_.clear();
    free_[0] = n;
  }
  bool alloc(uint64_t n, uint64_t *off) {
    n = (n + 255) & ~(uint64_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= n) {
        *off = it->first;
        const uint64_t rest = it->second - n, at = it->first + n;
        free_.erase(it);
        if (rest) free_[at] = rest;
        return true;
      }
    return false;
  }
  void release(uint64_t off, uint64_t n) {
    n = (n + 255) & ~(uint64_t)255;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

uint64_t array_bytes(const ifm_array &a) {
  return (uint64_t)a.size[0] * (uint64_t)a.size[1] * dtype_size(a.dtype);
}

}  // namespace

extern "C" {

int ifm_batch_run(ifm_ctx *ctx, const ifm_request *reqs, int n) {
  if (!ctx || (n > 0 && !reqs) || n < 0) return ctx ? sfail(ctx, IFM_E_ARG, "bad batch") : IFM_E_ARG;
  if (n == 0) return IFM_OK;
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  Block blk;
  std::vector<Group> groups;
  Extra ex;
  int rc = build_batch(ctx, reqs, idx.data(), n, nullptr, blk, groups, ex);
  if (rc) return rc;
  return launch_batch(ctx, reqs, idx.data(), blk, groups, ex, nullptr);
}

int ifm_fill_uniform(ifm_ctx *ctx, ifm_array *a, uint64_t seed) {
  if (!ctx || !a || !a->data || a->device != 1) return ctx ? sfail(ctx, IFM_E_ARG, "device array required") : IFM_E_ARG;
  int64_t count = 1;
  for (int d = 0; d < a->ndim; ++d) count *= a->size[d];
  SHIP(launch_fill_uniform(a->data, a->dtype, count, seed, ctx->stream));
  return IFM_OK;
}

int ifm_serve(ifm_ctx *ctx, const ifm_request *reqs, const double *arrival_s, int n, const ifm_serve_config *cfg,
              double *admit_s, double *done_s) {
  if (!ctx || !cfg || n < 0 || (n > 0 && (!reqs || !arrival_s || !admit_s || !done_s)))
    return ctx ? sfail(ctx, IFM_E_ARG, "bad serve arguments") : IFM_E_ARG;
  const int max_batch = cfg->max_batch > 0 ? cfg->max_batch : 1024;
  const int max_inflight = std::min(cfg->max_inflight > 0 ? cfg->max_inflight : 4, NSLOTS - 1);
  // Requests without buffers take theirs from the arena; its size is the budget.
  std::vector<ifm_request> rq(reqs, reqs + n);
  std::vector<uint64_t> need(n), offs(3 * (size_t)n, 0);
  bool any_gen = false;
  // need[i]: the arena bytes request i holds while in flight, each array
  // rounded up to the arena's 256-byte granule as Arena::alloc rounds it.
  auto granule = [](uint64_t b) { return (b + 255) & ~(uint64_t)255; };
  for (int i = 0; i < n; ++i) {
    need[i] = granule(array_bytes(rq[i].in)) + granule(array_bytes(rq[i].out)) +
              ((rq[i].op == IFM_OP_SEPARABLE && rq[i].nfactors2 > 0) ? granule(array_bytes(rq[i].out2)) : 0);
    if (!rq[i].in.data) any_gen = true;
    if (need[i] > cfg->budget_bytes) return sfail(ctx, IFM_E_BUDGET, "a request exceeds the HBM budget");
  }
  Arena arena;
  char *abase = nullptr;
  if (any_gen) {
    SHIP(hipMalloc((void **)&abase, cfg->budget_bytes));
    arena.init(abase, cfg->budget_bytes);
  }
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
  // The arrays of request i: in, out and (two-output separable) out2.
  auto arrays = [&](int i, ifm_array **a) {
    a[0] = &rq[i].in;
    a[1] = &rq[i].out;
    a[2] = (rq[i].op == IFM_OP_SEPARABLE && rq[i].nfactors2 > 0) ? &rq[i].out2 : nullptr;
  };
  auto give = [&](int i) {
    if (reqs[i].in.data) return;
    ifm_array *a[3];
    arrays(i, a);
    for (int j = 0; j < 3; ++j)
      if (a[j]) arena.release(offs[3 * i + j], array_bytes(*a[j]));
  };
  auto take = [&](int i) -> bool {  // buffers for request i, or false if the arena is full
    if (reqs[i].in.data) return true;
    ifm_array *a[3];
    arrays(i, a);
    for (int j = 0; j < 3; ++j) {
      if (!a[j]) continue;
      if (!arena.alloc(array_bytes(*a[j]), &offs[3 * i + j])) {
        for (int m = 0; m < j; ++m)
          if (a[m]) arena.release(offs[3 * i + m], array_bytes(*a[m]));
        return false;
      }
    }
    for (int j = 0; j < 3; ++j)
      if (a[j]) {
        a[j]->data = abase + offs[3 * i + j];
        a[j]->device = 1;
      }
    return true;
  };
  int rc = IFM_OK;
  uint64_t inflight_bytes = 0;
  if (cfg->mode == 1) {
    for (int i = 0; i < n && !rc; ++i) {
      while (now() < arrival_s[i]) std::this_thread::yield();
      if (!take(i)) {
        rc = sfail(ctx, IFM_E_BUDGET, "arena exhausted");
        break;
      }
      admit_s[i] = now();
      if (!reqs[i].in.data) SHIP(launch_fill_uniform(rq[i].in.data, rq[i].in.dtype,
                                                     rq[i].in.size[0] * rq[i].in.size[1], (uint64_t)i, ctx->stream));
      const ifm_request &q = rq[i];
      ifm_array in = q.in, out = q.out;
      switch (q.op) {
      case IFM_OP_SEPARABLE: {
        rc = ifm_imfilter_fir(ctx, &in, &out, nullptr, q.factors, q.nfactors, &q.border);
        ifm_array out2 = q.out2;
        if (!rc && q.nfactors2 > 0) rc = ifm_imfilter_fir(ctx, &in, &out2, nullptr, q.factors2, q.nfactors2, &q.border);
        break;
      }
      case IFM_OP_DENSE: rc = ifm_imfilter_dense(ctx, &in, &out, nullptr, &q.kernel, &q.border, IFM_PREC_AUTO); break;
      default: {
        const int64_t k = q.window, wf[2] = {-(k / 2), -(k / 2)}, wl[2] = {k, k};
        rc = ifm_mapwindow_median(ctx, &in, &out, nullptr, wf, wl, &q.border);
      }
      }
      if (!rc) SHIP(hipStreamSynchronize(ctx->stream));
      done_s[i] = now();
      give(i);
    }
  } else {
    struct Flight {
      hipEvent_t ev;
      std::vector<int> members;
    };
    std::vector<Flight> flights;
    int next = 0, ndone = 0;
    while (ndone < n && !rc) {
      bool progress = false;
      // Retire completed batches (their events complete in queue order).
      while (!flights.empty()) {
        hipError_t q = hipEventQuery(flights.front().ev);
        if (q == hipErrorNotReady) break;
        if (q != hipSuccess) {
          rc = sfail(ctx, IFM_E_DEVICE, std::string("batch: ") + hipGetErrorString(q));
          break;
        }
        const double t = now();
        for (int i : flights.front().members) {
          done_s[i] = t;
          inflight_bytes -= need[i];
          give(i);
          ++ndone;
        }
        flights.erase(flights.begin());
        progress = true;
      }
      if (rc) break;
      // Admit arrived requests into one batch.
      if ((int)flights.size() < max_inflight && next < n && arrival_s[next] <= now()) {
        std::vector<int> members;
        bool gen = false;
        while (next < n && (int)members.size() < max_batch && arrival_s[next] <= now() &&
               inflight_bytes + need[next] <= cfg->budget_bytes && take(next)) {
          gen = gen || !reqs[next].in.data;
          inflight_bytes += need[next];
          members.push_back(next++);
        }
        // With nothing in flight the arena is whole, so a request that fits
        // the budget always gets its buffers; report rather than spin if not.
        if (members.empty() && flights.empty() && next < n && arrival_s[next] <= now()) {
          rc = sfail(ctx, IFM_E_BUDGET, "a request does not fit the empty arena");
          break;
        }
        if (!members.empty()) {
          Block blk;
          std::vector<Group> groups;
          Extra ex;
          std::vector<char> genf(members.size());
          for (size_t m = 0; m < members.size(); ++m) genf[m] = !reqs[members[m]].in.data;
          rc = build_batch(ctx, rq.data(), members.data(), (int)members.size(), gen ? genf.data() : nullptr, blk,
                           groups, ex);
          if (!rc) {
            const double t = now();
            for (int i : members) admit_s[i] = t;
            Flight f;
            rc = launch_batch(ctx, rq.data(), members.data(), blk, groups, ex, &f.ev);
            f.members = members;
            if (!rc) flights.push_back(f);
          }
          progress = true;
        }
      }
      if (!progress) std::this_thread::yield();
    }
  }
  (void)hipStreamSynchronize(ctx->stream);
  if (abase) (void)hipFree(abase);
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Request server: non-blocking submit / poll over the batch machinery above.
// ---------------------------------------------------------------------------
struct ifm_server {
  ifm_ctx *ctx = nullptr;
  uint64_t budget = 0;
  int max_batch = 1024, max_inflight = 2;
  struct Req {
    ifm_request rq;
    uint64_t ticket = 0, bytes = 0;
    std::vector<double> taps;  // the request's taps, owned here
  };
  std::deque<Req> pending;
  struct Flight {
    hipEvent_t ev;
    uint64_t last_ticket, bytes;
  };
  std::deque<Flight> flights;
  uint64_t next_ticket = 1, done_below = 1, inflight_bytes = 0;
};

namespace {

int srv_fail(ifm_server *s, int code, const std::string &msg) { return sfail(s->ctx, code, msg); }

// Retires completed batches (they complete in launch order) and launches the
// pending requests the limits allow.  Never waits for the device.
int srv_pump(ifm_server *s) {
  while (!s->flights.empty()) {
    const hipError_t q = hipEventQuery(s->flights.front().ev);
    if (q == hipErrorNotReady) break;
    if (q != hipSuccess) return srv_fail(s, IFM_E_DEVICE, std::string("batch: ") + hipGetErrorString(q));
    s->done_below = s->flights.front().last_ticket + 1;
    s->inflight_bytes -= s->flights.front().bytes;
    s->flights.pop_front();
  }
  while (!s->pending.empty() && (int)s->flights.size() < s->max_inflight) {
    std::vector<ifm_request> rq;
    uint64_t bytes = 0;
    for (const auto &r : s->pending) {
      if ((int)rq.size() >= s->max_batch) break;
      if (s->budget && s->inflight_bytes + bytes + r.bytes > s->budget) break;
      rq.push_back(r.rq);
      bytes += r.bytes;
    }
    if (rq.empty()) break;  // waits for batches in flight to free budget
    std::vector<int> idx(rq.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
    Block blk;
    std::vector<Group> groups;
    Extra ex;
    int rc = build_batch(s->ctx, rq.data(), idx.data(), (int)rq.size(), nullptr, blk, groups, ex);
    if (rc) return rc;
    ifm_server::Flight f;
    rc = launch_batch(s->ctx, rq.data(), idx.data(), blk, groups, ex, &f.ev);
    if (rc) return rc;
    f.last_ticket = s->pending[rq.size() - 1].ticket;
    f.bytes = bytes;
    s->inflight_bytes += bytes;
    s->flights.push_back(f);
    for (size_t i = 0; i < rq.size(); ++i) s->pending.pop_front();
  }
  return IFM_OK;
}

}  // namespace

extern "C" {

int ifm_server_create(ifm_ctx *ctx, const ifm_server_config *cfg, ifm_server **server) {
  if (!ctx || !server) return ctx ? sfail(ctx, IFM_E_ARG, "server is NULL") : IFM_E_ARG;
  ifm_server *s = new ifm_server;
  s->ctx = ctx;
  if (cfg) {
    s->budget = cfg->budget_bytes;
    if (cfg->max_batch > 0) s->max_batch = cfg->max_batch;
    if (cfg->max_inflight > 0) s->max_inflight = std::min(cfg->max_inflight, NSLOTS - 1);
  }
  *server = s;
  return IFM_OK;
}

int ifm_server_submit(ifm_server *s, const ifm_request *req, uint64_t *ticket) {
  if (!s || !req || !ticket) return s ? srv_fail(s, IFM_E_ARG, "request or ticket is NULL") : IFM_E_ARG;
  if (!req->in.data || req->in.device != 1 || !req->out.data || req->out.device != 1)
    return srv_fail(s, IFM_E_ARG, "server requests take device arrays");
  ifm_server::Req r;
  r.rq = *req;
  // the taps move into the server (the caller's may not outlive the call)
  const bool sep = req->op == IFM_OP_SEPARABLE, dense = req->op == IFM_OP_DENSE;
  size_t nt = 0;
  if (sep) {
    if (req->nfactors < 1 || req->nfactors > 2 || req->nfactors2 < 0 || req->nfactors2 > 2)
      return srv_fail(s, IFM_E_ARG, "separable requests take 1 or 2 factors");
    for (int i = 0; i < req->nfactors; ++i) nt += (size_t)std::max(req->factors[i].len, 0);
    for (int i = 0; i < req->nfactors2; ++i) nt += (size_t)std::max(req->factors2[i].len, 0);
  } else if (dense) {
    size_t k = 1;
    for (int d = 0; d < req->kernel.ndim && d < IFM_MAXDIM; ++d) k *= (size_t)std::max<int64_t>(req->kernel.size[d], 0);
    nt = k;
  }
  r.taps.resize(nt);
  size_t off = 0;
  auto take = [&](const double *src, int64_t n) -> const double * {
    if (n <= 0) return nullptr;
    if (!src) return nullptr;
    std::memcpy(r.taps.data() + off, src, (size_t)n * sizeof(double));
    const double *p = r.taps.data() + off;
    off += (size_t)n;
    return p;
  };
  if (sep) {
    for (int i = 0; i < req->nfactors; ++i) r.rq.factors[i].taps = take(req->factors[i].taps, req->factors[i].len);
    for (int i = 0; i < req->nfactors2; ++i) r.rq.factors2[i].taps = take(req->factors2[i].taps, req->factors2[i].len);
  } else if (dense) {
    r.rq.kernel.taps = take(req->kernel.taps, (int64_t)nt);
  }
  // checked now, as a batch of one (no launch), so errors reach the caller
  {
    int i0 = 0;
    Block blk;
    std::vector<Group> groups;
    Extra ex;
    const int rc = build_batch(s->ctx, &r.rq, &i0, 1, nullptr, blk, groups, ex);
    if (rc) return rc;
  }
  r.bytes = array_bytes(req->in) + array_bytes(req->out) + ((sep && req->nfactors2 > 0) ? array_bytes(req->out2) : 0);
  if (s->budget && r.bytes > s->budget) return srv_fail(s, IFM_E_BUDGET, "a request exceeds the server's budget");
  r.ticket = s->next_ticket++;
  *ticket = r.ticket;
  s->pending.push_back(std::move(r));
  return srv_pump(s);
}

int ifm_server_poll(ifm_server *s, uint64_t ticket, int *done) {
  if (!s || !done) return s ? srv_fail(s, IFM_E_ARG, "done is NULL") : IFM_E_ARG;
  if (ticket == 0 || ticket >= s->next_ticket) return srv_fail(s, IFM_E_ARG, "unknown ticket");
  const int rc = srv_pump(s);
  if (rc) return rc;
  *done = ticket < s->done_below ? 1 : 0;
  return IFM_OK;
}

int ifm_server_wait(ifm_server *s, uint64_t ticket) {
  if (!s) return IFM_E_ARG;
  if (ticket >= s->next_ticket) return srv_fail(s, IFM_E_ARG, "unknown ticket");
  const uint64_t want = ticket ? ticket : s->next_ticket - 1;
  ifm_ctx *ctx = s->ctx;  // for SHIP
  for (;;) {
    int rc = srv_pump(s);
    if (rc) return rc;
    if (want < s->done_below) return IFM_OK;
    if (s->flights.empty()) return srv_fail(s, IFM_E_DEVICE, "nothing in flight for the ticket");
    SHIP(hipEventSynchronize(s->flights.front().ev));
  }
}

int ifm_server_destroy(ifm_server *s) {
  if (!s) return IFM_OK;
  const int rc = s->next_ticket > 1 ? ifm_server_wait(s, 0) : IFM_OK;
  delete s;
  return rc;
}

}  // extern "C"
