// Binning kernels: device radix sort, exclusive scan, intersection emission,
// tile ranges, and small fills.
//
// Sort contract (radix_sort.wgsl:21-26 and gpu_primitives/radix_sort/tests.rs:
// 17-43): stable ascending by key, ties keep input order.  LSD radix sort with
// 8-bit digits; each pass is
//   k_radix_hist    per-block digit histograms          hist[digit][block]
//   k_radix_spine   one block per digit: exclusive scan of its row over blocks
//   k_radix_scatter stable scatter; a block's items are ranked per 64-lane
//                   wave with ballots (bitwise match of the digit), in index order.
// Every kernel reads the item count from device memory (clamped to capacity),
// so the sequence is graph-capturable with no host readback.
//
// Scan contract (prefix_sum.wgsl, prefix_sum/tests.rs:17-59): the reference
// uses an inclusive scan; the binning consumes the exclusive form
// (cum[g-1], map_gaussian_to_intersects.wgsl:34).
#include "rgs_device.h"

namespace rgs {

constexpr int kSortThreads = 256;
constexpr int kSortTile = 1024;  // items per block iteration (4 waves x 4 slots x 64 lanes)

__device__ __forceinline__ uint32_t dev_count(const uint32_t* n_dev, uint32_t n_host, uint32_t cap) {
  const uint32_t n = n_dev ? *n_dev : n_host;
  return n < cap ? n : cap;
}

__device__ __forceinline__ void block_tiles(uint32_t n, uint32_t* t0, uint32_t* t1) {
  const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
  const uint32_t per = (tiles + gridDim.x - 1) / gridDim.x;
  const uint32_t a = blockIdx.x * per;
  *t0 = a < tiles ? a : tiles;
  const uint32_t b = *t0 + per;
  *t1 = b < tiles ? b : tiles;
}

// Exclusive scan of one value per thread over a 256-thread block; returns the
// block total through *total.  Uses `tmp` (>= 4 entries) of LDS.
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t* tmp, uint32_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) tmp[w] = x;
  __syncthreads();
  uint32_t wbase = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t t = tmp[i];
    if (i < w) wbase += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return wbase + x - v;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                             const uint32_t* __restrict__ n_dev, uint32_t n_host,
                                                             uint32_t cap, int shift, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[4][256];
  const int w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < 1024; i += kSortThreads) (&h[0][0])[i] = 0;
  __syncthreads();
  const uint32_t n = dev_count(n_dev, n_host, cap);
  uint32_t t0, t1;
  block_tiles(n, &t0, &t1);
  const uint32_t end = t1 * kSortTile < n ? t1 * kSortTile : n;
  for (uint32_t i = t0 * kSortTile + threadIdx.x; i < end; i += kSortThreads)
    atomicAdd(&h[w][(keys[i] >> shift) & 255u], 1u);
  __syncthreads();
  const uint32_t d = threadIdx.x;
  hist[(size_t)d * gridDim.x + blockIdx.x] = h[0][d] + h[1][d] + h[2][d] + h[3][d];
}

// grid = 256 blocks (one per digit); G = number of histogram blocks (<= 2048)
__global__ __launch_bounds__(kSortThreads) void k_radix_spine(uint32_t* __restrict__ hist, uint32_t G,
                                                              uint32_t* __restrict__ tot) {
  __shared__ uint32_t tmp[8];
  uint32_t* row = hist + (size_t)blockIdx.x * G;
  const uint32_t per = (G + kSortThreads - 1) / kSortThreads;  // <= 8
  const uint32_t a = threadIdx.x * per;
  uint32_t loc[8];
  uint32_t s = 0;
#pragma unroll
  for (uint32_t i = 0; i < 8; ++i) {
    const uint32_t v = (i < per && a + i < G) ? row[a + i] : 0u;
    loc[i] = v;
    s += v;
  }
  uint32_t total;
  uint32_t run = block_excl_scan256(s, tmp, &total);
#pragma unroll
  for (uint32_t i = 0; i < 8; ++i) {
    if (i < per && a + i < G) row[a + i] = run;
    run += loc[i];
  }
  if (threadIdx.x == 0) tot[blockIdx.x] = total;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, const uint32_t* __restrict__ n_dev, uint32_t n_host, uint32_t cap, int shift,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ tot) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t wh[4][256];
  __shared__ uint32_t tmp[8];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  {
    uint32_t total;
    const uint32_t db = block_excl_scan256(tot[threadIdx.x], tmp, &total);
    base[threadIdx.x] = db + hist[(size_t)threadIdx.x * gridDim.x + blockIdx.x];
  }
  const uint32_t n = dev_count(n_dev, n_host, cap);
  uint32_t t0, t1;
  block_tiles(n, &t0, &t1);
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (uint32_t t = t0; t < t1; ++t) {
    for (int i = threadIdx.x; i < 1024; i += kSortThreads) (&wh[0][0])[i] = 0;
    __syncthreads();
    uint32_t key[4], val[4], dig[4], loc[4];
    bool ok[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const uint32_t idx = t * kSortTile + w * 256 + s * 64 + lane;
      ok[s] = idx < n;
      key[s] = ok[s] ? kin[idx] : 0u;
      val[s] = ok[s] ? (vin ? vin[idx] : idx) : 0u;
      dig[s] = (key[s] >> shift) & 255u;
      uint64_t peers = __ballot(ok[s]);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const bool bit = (dig[s] >> b) & 1u;
        const uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
      }
      const uint32_t rank = (uint32_t)__popcll(peers & lt);
      uint32_t old = 0;
      if (ok[s]) old = wh[w][dig[s]];
      if (ok[s] && rank == 0) wh[w][dig[s]] = old + (uint32_t)__popcll(peers);
      loc[s] = old + rank;
    }
    __syncthreads();
    {
      const uint32_t d = threadIdx.x;
      uint32_t run = base[d];
#pragma unroll
      for (int ww = 0; ww < 4; ++ww) {
        const uint32_t c = wh[ww][d];
        wh[ww][d] = run;
        run += c;
      }
      base[d] = run;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (ok[s]) {
        const uint32_t dst = wh[w][dig[s]] + loc[s];
        kout[dst] = key[s];
        vout[dst] = val[s];
      }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ scan
// Exclusive scan of x[i] = in[gather ? gather[i] : i] for i < n.
__global__ __launch_bounds__(kSortThreads) void k_scan_reduce(const uint32_t* __restrict__ in,
                                                              const uint32_t* __restrict__ gather,
                                                              const uint32_t* __restrict__ n_dev, uint32_t n_host,
                                                              uint32_t cap, uint32_t* __restrict__ partial) {
  __shared__ uint32_t tmp[8];
  const uint32_t n = dev_count(n_dev, n_host, cap);
  uint32_t t0, t1;
  block_tiles(n, &t0, &t1);
  const uint32_t end = t1 * kSortTile < n ? t1 * kSortTile : n;
  uint32_t s = 0;
  for (uint32_t i = t0 * kSortTile + threadIdx.x; i < end; i += kSortThreads) s += in[gather ? gather[i] : i];
  uint32_t total;
  block_excl_scan256(s, tmp, &total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// single block: exclusive scan of partial[G] in place; *total_out = grand
// total, *max_out = max(*max_out, total) when non-null
__global__ __launch_bounds__(kSortThreads) void k_scan_spine(uint32_t* __restrict__ partial, uint32_t G,
                                                             uint32_t* __restrict__ total_out,
                                                             uint32_t* __restrict__ max_out) {
  __shared__ uint32_t tmp[8];
  const uint32_t per = (G + kSortThreads - 1) / kSortThreads;  // <= 8
  const uint32_t a = threadIdx.x * per;
  uint32_t loc[8];
  uint32_t s = 0;
#pragma unroll
  for (uint32_t i = 0; i < 8; ++i) {
    const uint32_t v = (i < per && a + i < G) ? partial[a + i] : 0u;
    loc[i] = v;
    s += v;
  }
  uint32_t total;
  uint32_t run = block_excl_scan256(s, tmp, &total);
#pragma unroll
  for (uint32_t i = 0; i < 8; ++i) {
    if (i < per && a + i < G) partial[a + i] = run;
    run += loc[i];
  }
  if (threadIdx.x == 0 && total_out) *total_out = total;
  if (threadIdx.x == 0 && max_out) atomicMax(max_out, total);
}

__global__ __launch_bounds__(kSortThreads) void k_scan_down(const uint32_t* __restrict__ in,
                                                            const uint32_t* __restrict__ gather,
                                                            const uint32_t* __restrict__ n_dev, uint32_t n_host,
                                                            uint32_t cap, const uint32_t* __restrict__ partial,
                                                            uint32_t* __restrict__ out, int inclusive) {
  __shared__ uint32_t tmp[8];
  const uint32_t n = dev_count(n_dev, n_host, cap);
  uint32_t t0, t1;
  block_tiles(n, &t0, &t1);
  uint32_t run = partial[blockIdx.x];
  for (uint32_t t = t0; t < t1; ++t) {
    const uint32_t i0 = t * kSortTile + threadIdx.x * 4;
    uint32_t x[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t i = i0 + k;
      x[k] = i < n ? in[gather ? gather[i] : i] : 0u;
      s += x[k];
    }
    uint32_t total;
    uint32_t e = block_excl_scan256(s, tmp, &total) + run;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t i = i0 + k;
      if (i < n) out[i] = inclusive ? e + x[k] : e;
      e += x[k];
    }
    run += total;
  }
}

// ------------------------------------------------------------- emission
// map_gaussian_to_intersects.wgsl:15-49: for each visible instance in depth
// order, write (global tile id, instance) for every tile of its bbox that
// passes the tile test, at offs[i] + k.  The 256 instances of a workgroup own
// one contiguous output range, which is assembled in LDS and written out
// coalesced (ranges too large for LDS are written directly).  Bboxes of at most
// 64 tiles reuse the pass mask of the projection pass instead of re-testing.
// Writes past `cap` are dropped and flagged; the step is then marked invalid
// so no parameter is updated.
constexpr uint32_t kEmitStage = 4096;

__global__ __launch_bounds__(256) void k_emit(const ProjSplat* __restrict__ proj, const uint32_t* __restrict__ order,
                                              const uint32_t* __restrict__ offs, const uint32_t* __restrict__ cnt,
                                              uint32_t* __restrict__ counters, uint32_t n_per_view,
                                              const ViewDev* __restrict__ views, int tw, uint32_t cap_inst,
                                              uint32_t cap, uint32_t* __restrict__ okeys,
                                              uint32_t* __restrict__ ovals) {
  __shared__ uint32_t sk[kEmitStage], sv[kEmitStage];
  __shared__ uint32_t s_base, s_total;
  __shared__ int s_over;
  const uint32_t V = min(counters[kCntVisible], cap_inst);
  const uint32_t i0 = blockIdx.x * blockDim.x;
  if (i0 >= V) return;  // whole workgroup
  if (threadIdx.x == 0) {
    const uint32_t last = min(i0 + blockDim.x, V) - 1;
    s_base = offs[i0];
    s_total = offs[last] + cnt[order[last]] - offs[i0];
    s_over = 0;
  }
  __syncthreads();
  const uint32_t base = s_base, total = s_total;
  const bool staged = total <= kEmitStage;
  const uint32_t i = i0 + threadIdx.x;
  bool over = false;
  if (i < V) {
    const uint32_t inst = order[i];
    const uint32_t v = inst / n_per_view;
    const ViewDev& VD = views[v];
    const ProjSplat s = proj[inst];
    uint32_t bb[4];
    unpack_u16x2(s.p2.z, &bb[0], &bb[1]);
    unpack_u16x2(s.p2.w, &bb[2], &bb[3]);
    const float a = 2.0f * s.p0.z, b = s.p0.w, c = 2.0f * s.p1.x, thr = s.p2.y;
    const uint32_t nb = (bb[3] - bb[1]) * (bb[2] - bb[0]);
    const uint64_t mask = (uint64_t)__float_as_uint(s.p3.x) | ((uint64_t)__float_as_uint(s.p3.y) << 32);
    uint32_t pos = offs[i];
    uint32_t k = 0;
    for (uint32_t ty = bb[1]; ty < bb[3]; ++ty)
      for (uint32_t tx = bb[0]; tx < bb[2]; ++tx, ++k) {
        const bool hit = nb <= 64 ? ((mask >> k) & 1ull) != 0
                                  : tile_contributes(tx, ty, tw, s.p0.x, s.p0.y, a, b, c, thr);
        if (!hit) continue;
        const uint32_t key = VD.tile_base + tx + ty * VD.tiles_x;
        if (staged) {
          sk[pos - base] = key;
          sv[pos - base] = inst;
        } else if (pos < cap) {
          okeys[pos] = key;
          ovals[pos] = inst;
        } else {
          over = true;
        }
        ++pos;
      }
  }
  if (staged) {
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
      const uint32_t p = base + e;
      if (p < cap) {
        okeys[p] = sk[e];
        ovals[p] = sv[e];
      } else {
        over = true;
      }
    }
  }
  if (over) {
    counters[kCntOverflow] = 1u;
    counters[kCntStepValid] = 0u;
  }
}

// get_tile_offsets.wgsl:12-30 over sorted keys: ranges[2t] = start, ranges[2t+1] = end.
__global__ __launch_bounds__(256) void k_tile_ranges(const uint32_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ counters, uint32_t cap,
                                                     uint32_t num_tiles, uint32_t* __restrict__ ranges) {
  const uint32_t m = min(counters[kCntIsect], cap);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t t = keys[i];
  if (t >= num_tiles) return;
  if (i == 0 || keys[i - 1] != t) ranges[2 * t] = i;
  if (i + 1 == m || keys[i + 1] != t) ranges[2 * t + 1] = i + 1;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// Step prologue: reset the per-step counters (a kernel, not a memset, so the
// step stays safe to capture and replay as a graph).  A step that found its
// predecessor invalid counts it as skipped first.
__global__ void k_step_begin(uint32_t* __restrict__ counters, float* __restrict__ fscal, int nviews) {
  const int t = threadIdx.x;
  if (t == 0) {
    counters[kCntOverflow] = 0;
    counters[kCntIsectMax] = 0;
    counters[kCntStepValid] = 1;
  }
  if (t < nviews) fscal[kFLossBase + t] = 0.0f;
}

// Forward prologue: per-forward counters.
__global__ void k_fwd_begin(uint32_t* __restrict__ counters) {
  if (threadIdx.x == 0) {
    counters[kCntVisible] = 0;
    counters[kCntIsect] = 0;
    counters[kCntQueueFwd] = 0;
    counters[kCntQueueBwd] = 0;
  }
}

// Step epilogue: count the step as skipped when it did not update.
__global__ void k_step_end(uint32_t* __restrict__ counters) {
  if (threadIdx.x == 0 && counters[kCntStepValid] == 0u) counters[kCntSkipped] += 1u;
}

// ------------------------------------------------------------------ host side
static uint32_t grid_for(uint64_t cap) {
  uint64_t tiles = (cap + kSortTile - 1) / kSortTile;
  if (tiles < 1) tiles = 1;
  return (uint32_t)(tiles < 2048 ? tiles : 2048);
}

// Stable sort of (key, value) pairs on key bits [0, key_bits).
// The count is read from n_dev when non-null (clamped to cap).
// Pass 0 reads (kin, vin) and writes (kb, vb); later passes ping-pong
// between (kb, vb) and (kc, vc).  (kc, vc) may alias (kin, vin) when the input
// may be clobbered.  vin == nullptr means values are the input indices.
// Returns 1 if the result is in (kb, vb), 2 if in (kc, vc), 0 if key_bits <= 0
// (nothing done; the input order is already sorted).
// scratch: 256 * G + 256 uint32 with G = radix_grid(cap).
uint32_t radix_grid(uint64_t cap) { return grid_for(cap); }

int radix_sort_pairs(hipStream_t s, const uint32_t* kin, const uint32_t* vin, uint32_t* kb, uint32_t* vb,
                     uint32_t* kc, uint32_t* vc, const uint32_t* n_dev, uint32_t n_host, uint32_t cap, int key_bits,
                     uint32_t* scratch) {
  if (key_bits <= 0) return 0;
  const uint32_t G = grid_for(cap);
  uint32_t* hist = scratch;
  uint32_t* tot = scratch + 256 * (size_t)G;
  const int passes = (key_bits + 7) / 8;
  const uint32_t* sk = kin;
  const uint32_t* sv = vin;
  int where = 0;
  for (int p = 0; p < passes; ++p) {
    uint32_t* dk = (p % 2 == 0) ? kb : kc;
    uint32_t* dv = (p % 2 == 0) ? vb : vc;
    k_radix_hist<<<G, kSortThreads, 0, s>>>(sk, n_dev, n_host, cap, 8 * p, hist);
    k_radix_spine<<<256, kSortThreads, 0, s>>>(hist, G, tot);
    k_radix_scatter<<<G, kSortThreads, 0, s>>>(sk, sv, dk, dv, n_dev, n_host, cap, 8 * p, hist, tot);
    sk = dk;
    sv = dv;
    where = (p % 2 == 0) ? 1 : 2;
  }
  return where;
}

// Exclusive (or inclusive) scan of in[gather[i]] (gather may be null) for
// i < n; the grand total is written to *total_dev when non-null.
// scratch: radix_grid(cap) uint32.
void scan_u32(hipStream_t s, const uint32_t* in, const uint32_t* gather, const uint32_t* n_dev, uint32_t n_host,
              uint32_t cap, uint32_t* out, int inclusive, uint32_t* total_dev, uint32_t* max_dev,
              uint32_t* scratch) {
  const uint32_t G = grid_for(cap);
  k_scan_reduce<<<G, kSortThreads, 0, s>>>(in, gather, n_dev, n_host, cap, scratch);
  k_scan_spine<<<1, kSortThreads, 0, s>>>(scratch, G, total_dev, max_dev);
  k_scan_down<<<G, kSortThreads, 0, s>>>(in, gather, n_dev, n_host, cap, scratch, out, inclusive);
}

void launch_emit(hipStream_t s, const ProjSplat* proj, const uint32_t* order, const uint32_t* offs,
                 const uint32_t* cnt, uint32_t* counters, uint32_t n_per_view, const ViewDev* views, int tw,
                 uint32_t cap_inst, uint32_t cap, uint32_t* okeys, uint32_t* ovals) {
  if (cap_inst == 0) return;
  k_emit<<<(cap_inst + 255) / 256, 256, 0, s>>>(proj, order, offs, cnt, counters, n_per_view, views, tw, cap_inst,
                                                 cap, okeys, ovals);
}

void launch_tile_ranges(hipStream_t s, const uint32_t* keys, const uint32_t* counters, uint32_t cap,
                        uint32_t num_tiles, uint32_t* ranges) {
  if (cap == 0) return;
  k_tile_ranges<<<(cap + 255) / 256, 256, 0, s>>>(keys, counters, cap, num_tiles, ranges);
}

void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, size_t n) {
  if (n == 0) return;
  size_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  k_fill_u32<<<(uint32_t)blocks, 256, 0, s>>>(p, v, n);
}

void launch_step_begin(hipStream_t s, uint32_t* counters, float* fscal, int nviews) {
  k_step_begin<<<1, 64, 0, s>>>(counters, fscal, nviews);
}

void launch_fwd_begin(hipStream_t s, uint32_t* counters) { k_fwd_begin<<<1, 64, 0, s>>>(counters); }

void launch_step_end(hipStream_t s, uint32_t* counters) { k_step_end<<<1, 64, 0, s>>>(counters); }

__global__ void k_set_u32(uint32_t* __restrict__ p, uint32_t v) {
  if (threadIdx.x == 0) *p = v;
}

void launch_set_u32(hipStream_t s, uint32_t* p, uint32_t v) { k_set_u32<<<1, 64, 0, s>>>(p, v); }

}  // namespace rgs
