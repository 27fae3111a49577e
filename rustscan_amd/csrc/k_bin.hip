// Binning kernels: device radix sort, exclusive scan, intersection emission,
// tile ranges, and small fills.
//
// Sort contract (radix_sort.wgsl:21-26 and gpu_primitives/radix_sort/tests.rs:
// 17-43): stable ascending by key, ties keep input order.  LSD radix sort with
// 8-bit digits; each pass is
//   k_radix_hist    per-block digit histograms          hist[digit][block]
//   k_radix_spine   one block per digit: exclusive scan of its row over blocks
//   k_radix_scatter per 4096-item tile: each wave ranks its 1024 contiguous
//                   items (equal digits found by ballots or by peer masks
//                   per digit in LDS), the tile
//                   is sorted by digit in LDS, and written out so that
//                   consecutive threads store consecutive slots of a bucket.
// Every kernel reads the item count from device memory (clamped to capacity),
// so the sequence is graph-capturable with no host readback.
//
// Scan contract (prefix_sum.wgsl, prefix_sum/tests.rs:17-59): the reference
// uses an inclusive scan; the binning consumes the exclusive form
// (cum[g-1], map_gaussian_to_intersects.wgsl:34).
#include "rgs_device.h"

namespace rgs {

constexpr int kSortThreads = 256;
constexpr int kSortTile = 1024;  // scan: items per block iteration (256 threads x 4)
constexpr int kRadixRows = 16;   // radix: items per lane per tile (pairs sorts)
constexpr int kRadixTile = kSortThreads * kRadixRows;  // 4096 items; each wave owns 1024 contiguous
// The packed tile-sort passes use 2048-item tiles: half the registers per
// lane, so more waves hide their memory latency (measured faster; the
// latency-bound depth sort is faster with 4096).
constexpr int kRadixRowsPacked = 8;
constexpr int kRadixTilePacked = kSortThreads * kRadixRowsPacked;

__device__ __forceinline__ uint32_t dev_count(const uint32_t* n_dev, uint32_t n_host, uint32_t cap) {
  const uint32_t n = n_dev ? *n_dev : n_host;
  return n < cap ? n : cap;
}

template <int TILE = kSortTile>
__device__ __forceinline__ void block_tiles(uint32_t n, uint32_t* t0, uint32_t* t1) {
  const uint32_t tiles = (n + TILE - 1) / TILE;
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

template <int TILE>
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                             const uint32_t* __restrict__ n_dev, uint32_t n_host,
                                                             uint32_t cap, int shift, uint32_t* __restrict__ hist) {
  // four copies per wave, picked by lane & 3: keys with few distinct digits
  // (the high digits of the sorts here) would otherwise pile all 64 lanes'
  // LDS atomics onto one or two bins
  constexpr int kCopies = 16;
  __shared__ uint32_t h[kCopies][256];
  const int c = ((threadIdx.x >> 6) << 2) | (threadIdx.x & 3);
  for (int i = threadIdx.x; i < kCopies * 256; i += kSortThreads) (&h[0][0])[i] = 0;
  __syncthreads();
  const uint32_t n = dev_count(n_dev, n_host, cap);
  uint32_t t0, t1;
  block_tiles<TILE>(n, &t0, &t1);
  const uint32_t begin = t0 * TILE;
  const uint32_t end = t1 * TILE < n ? t1 * TILE : n;
  // four keys per thread per step (vector loads), several steps in flight;
  // scalar loads when the range does not start 16-byte aligned
  const bool aligned = ((reinterpret_cast<uintptr_t>(keys) + 4ull * begin) & 15u) == 0;
  const uint32_t vend = aligned ? begin + ((end > begin ? end - begin : 0u) & ~3u) : begin;
  const uint4* kv = reinterpret_cast<const uint4*>(keys);
#pragma unroll 4
  for (uint32_t q = begin / 4 + threadIdx.x; q < vend / 4; q += kSortThreads) {
    const uint4 k = kv[q];
    atomicAdd(&h[c][(k.x >> shift) & 255u], 1u);
    atomicAdd(&h[c][(k.y >> shift) & 255u], 1u);
    atomicAdd(&h[c][(k.z >> shift) & 255u], 1u);
    atomicAdd(&h[c][(k.w >> shift) & 255u], 1u);
  }
  for (uint32_t i = vend + threadIdx.x; i < end; i += kSortThreads) atomicAdd(&h[c][(keys[i] >> shift) & 255u], 1u);
  __syncthreads();
  const uint32_t d = threadIdx.x;
  uint32_t t = 0;
#pragma unroll
  for (int k = 0; k < kCopies; ++k) t += h[k][d];
  hist[(size_t)d * gridDim.x + blockIdx.x] = t;
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

// Scatter modes.  kScatterPairs: (key, value) in and out.  The two passes of
// a packed two-pass sort (keys of 9..16 bits, value < 2^ib, key_bits - 8 +
// ib <= 32): kScatterPackOut sorts (key, value) pairs on the low 8 key bits
// and writes one word per item, (key >> 8) << ib | value; kScatterPackIn
// sorts those words on bits [ib, ib + 8) and writes (key, value) pairs again,
// taking the low key byte from the item's position (the first pass left the
// words grouped by it; tot1 holds its digit totals).  The middle array then
// moves 4 bytes per item instead of 8.
enum ScatterMode : int { kScatterPairs = 0, kScatterPackOut = 1, kScatterPackIn = 2 };

template <int MODE>
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, const uint32_t* __restrict__ n_dev, uint32_t n_host, uint32_t cap, int shift,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ tot, const uint32_t* __restrict__ gsrc,
    uint32_t* __restrict__ gout, int ib, const uint32_t* __restrict__ tot1) {
  __shared__ uint32_t base[256];     // next global slot of each digit for this block
  __shared__ uint32_t gbase[256];    // global slot of local position 0 of each digit's run in the tile
  __shared__ uint32_t wh[4][256];    // per-wave digit counts, then per-wave local bases
  constexpr int ROWS = MODE == kScatterPairs ? kRadixRows : kRadixRowsPacked;
  constexpr int TILE = kSortThreads * ROWS;
  __shared__ __align__(16) uint32_t sk[TILE];  // 8-byte peer masks during the ranking
  __shared__ uint32_t sv[TILE];
  __shared__ uint32_t lo_start[MODE == kScatterPackIn ? 257 : 1];  // first position of each low byte
  __shared__ uint32_t tmp[8];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  {
    uint32_t total;
    const uint32_t db = block_excl_scan256(tot[threadIdx.x], tmp, &total);
    base[threadIdx.x] = db + hist[(size_t)threadIdx.x * gridDim.x + blockIdx.x];
  }
  if (MODE == kScatterPackIn) {
    uint32_t total;
    lo_start[threadIdx.x] = block_excl_scan256(tot1[threadIdx.x], tmp, &total);
    if (threadIdx.x == 0) lo_start[256] = 0xffffffffu;
  }
  const uint32_t vmask = (ib <= 0 || ib >= 32) ? 0xffffffffu : ((1u << ib) - 1u);
  const uint32_t n = dev_count(n_dev, n_host, cap);
  uint32_t t0, t1;
  block_tiles<TILE>(n, &t0, &t1);
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  // In the tile sort's passes (several tiles per block) the next tile's keys
  // (and values) are loaded while this tile is ranked, sorted and written, so
  // their latency overlaps the block's LDS work.  The pairs mode (the depth
  // sort: about one tile per block) loads each tile when it starts and keeps
  // the registers (and occupancy) of one tile.
  constexpr bool kAhead = MODE != kScatterPairs;
  // Ranking: peer masks in LDS for the tile sort's passes (several tiles per
  // block, loads in flight), ballots for the depth sort (latency-bound: one
  // tile per block).
  constexpr bool kMaskRank = MODE != kScatterPairs;
  uint32_t nkey[ROWS], nval[ROWS];
  auto load_tile = [&](uint32_t tt) {
    const uint32_t wb = tt * TILE + (uint32_t)w * (64 * ROWS);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const uint32_t idx = wb + (uint32_t)r * 64 + lane;
      const bool ok = idx < n;
      nkey[r] = ok ? kin[idx] : 0u;
      if (MODE != kScatterPackIn) nval[r] = ok ? (vin ? vin[idx] : idx) : 0u;
    }
  };
  if (kAhead && t0 < t1) load_tile(t0);
  for (uint32_t t = t0; t < t1; ++t) {
    if (!kAhead) load_tile(t);
    uint32_t key[ROWS], val[ROWS], loc[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      key[r] = nkey[r];
      val[r] = MODE != kScatterPackIn ? nval[r] : 0u;
    }
    // per-wave peer masks of the ranking (in sk, unused until step 3)
    uint64_t* pm = reinterpret_cast<uint64_t*>(sk) + (size_t)w * 256;
    for (int i = threadIdx.x; i < 1024; i += kSortThreads) {
      (&wh[0][0])[i] = 0;
      if (kMaskRank) reinterpret_cast<uint64_t*>(sk)[i] = 0ull;
    }
    __syncthreads();
    // 1. each wave ranks its 1024 contiguous items row by row (64 per row),
    //    in index order, with a running count per digit
    const uint32_t wbase = t * TILE + (uint32_t)w * (64 * ROWS);
    // packed input: low key byte of position wbase (largest d with
    // lo_start[d] <= wbase), then advanced per item
    uint32_t lo_d = 0, lo_next = 0;
    if (MODE == kScatterPackIn) {
      uint32_t a = 0, b = 256;
      while (b - a > 1) {
        const uint32_t mid = (a + b) >> 1;
        if (lo_start[mid] <= wbase) a = mid; else b = mid;
      }
      lo_d = a;
      lo_next = lo_start[a + 1];
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const uint32_t idx = wbase + (uint32_t)r * 64 + lane;
      const bool ok = idx < n;
      if (MODE == kScatterPackIn) {
        while (lo_d < 255u && idx >= lo_next) lo_next = lo_start[++lo_d + 1];
        val[r] = lo_d;
      }
      const uint32_t d = (key[r] >> shift) & 255u;
      uint64_t peers;
      if (kMaskRank) {
        // lanes of the same digit: each ORs its bit into the digit's mask
        // (the result does not depend on the order) and reads it back; the
        // first of them clears it after the count update below
        if (ok) atomicOr(reinterpret_cast<unsigned long long*>(&pm[d]), 1ull << lane);
        wave_lds_fence();
        peers = ok ? pm[d] : 0ull;
      } else {
        // lanes of the same digit by a bitwise ballot match
        peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          const bool bit = (d >> b) & 1u;
          const uint64_t m = __ballot(bit);
          peers &= bit ? m : ~m;
        }
      }
      const uint32_t rank = (uint32_t)__popcll(peers & lt);
      uint32_t old = 0;
      if (ok) old = wh[w][d];
      if (kMaskRank) wave_lds_fence();
      if (ok && rank == 0) {
        wh[w][d] = old + (uint32_t)__popcll(peers);
        if (kMaskRank) pm[d] = 0ull;
      }
      if (kMaskRank) wave_lds_fence();
      loc[r] = old + rank;
    }
    if (kAhead && t + 1 < t1) load_tile(t + 1);
    __syncthreads();
    // 2. per digit: local start in the sorted tile (scan over digits), the
    //    base of each wave's run, and the global slot of the digit's run
    {
      const uint32_t d = threadIdx.x;
      const uint32_t c0 = wh[0][d], c1 = wh[1][d], c2 = wh[2][d], c3 = wh[3][d];
      const uint32_t cnt = c0 + c1 + c2 + c3;
      uint32_t total;
      const uint32_t start = block_excl_scan256(cnt, tmp, &total);
      wh[0][d] = start;
      wh[1][d] = start + c0;
      wh[2][d] = start + c0 + c1;
      wh[3][d] = start + c0 + c1 + c2;
      gbase[d] = base[d] - start;
      base[d] += cnt;
    }
    __syncthreads();
    // 3. sort the tile by digit in LDS (stable: wave order, then rank)
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const uint32_t idx = wbase + (uint32_t)r * 64 + lane;
      if (idx < n) {
        const uint32_t p = wh[w][(key[r] >> shift) & 255u] + loc[r];
        sk[p] = key[r];
        sv[p] = val[r];
      }
    }
    __syncthreads();
    // 4. write out: thread i stores local position i, so each digit's run
    //    goes to consecutive global slots from consecutive threads
    const uint32_t tn = min((uint32_t)TILE, n - t * TILE);
#pragma unroll 4
    for (uint32_t i = threadIdx.x; i < tn; i += kSortThreads) {
      const uint32_t k = sk[i];
      const uint32_t dst = gbase[(k >> shift) & 255u] + i;
      const uint32_t v = sv[i];
      if (MODE == kScatterPairs) {
        kout[dst] = k;
        if (ib > 0) {  // the value carries a count above bit ib; all ones = look it up in gsrc
          const uint32_t inst = v & vmask, c = v >> ib;
          vout[dst] = inst;
          gout[dst] = c == (0xffffffffu >> ib) ? gsrc[inst] : c;
        } else {
          vout[dst] = v;
          if (gout) gout[dst] = gsrc[v];
        }
      } else if (MODE == kScatterPackOut) {
        kout[dst] = ((k >> 8) << ib) | v;
      } else {
        kout[dst] = ((k >> ib) << 8) | v;
        vout[dst] = k & vmask;
      }
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
// order, write (global tile id, Gaussian index within its view) for every
// tile of its bbox that passes the tile test, at offs[i] + k, in row-major
// bbox order.  The tile id fixes the view, so the instance is
// view * n + index; the shorter value lets the packed tile sort carry it
// with the tile id in one word for larger batches.
//  * Bboxes of at most 64 tiles carry their pass mask from the projection
//    pass; the 256 instances of a workgroup own one contiguous output range,
//    which is assembled in LDS in windows of kEmitStage slots and written out
//    coalesced.
//  * Larger bboxes are tested by the whole wave, 64 tiles per step, and their
//    (long, contiguous) runs go straight to memory.
// Writes past `cap` are dropped and flagged; the step is then marked invalid
// so no parameter is updated.
constexpr uint32_t kEmitStage = 2048;  // staged output slots per window (8 KB + 8 KB of LDS)
constexpr uint32_t kEmptySlot = 0xffffffffu;

__global__ __launch_bounds__(256) void k_emit(const ProjSplat* __restrict__ proj, const uint32_t* __restrict__ order,
                                              const uint32_t* __restrict__ offs, const uint32_t* __restrict__ cnt,
                                              uint32_t* __restrict__ counters, uint32_t n_per_view,
                                              const ViewDev* __restrict__ views, int tw, uint32_t cap_inst,
                                              uint32_t cap, uint32_t* __restrict__ okeys,
                                              uint32_t* __restrict__ ovals) {
  __shared__ uint32_t sk[kEmitStage], sv[kEmitStage];
  __shared__ uint32_t s_base, s_total;
  const uint32_t V = min(counters[kCntVisible], cap_inst);
  const uint32_t i0 = blockIdx.x * blockDim.x;
  if (i0 >= V) return;  // whole workgroup
  if (threadIdx.x == 0) {
    const uint32_t last = min(i0 + blockDim.x, V) - 1;
    s_base = offs[i0];
    s_total = offs[last] + cnt[order[last]] - offs[i0];
  }
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64u - lane));
  const uint32_t i = i0 + threadIdx.x;
  const bool live = i < V;
  uint32_t inst = 0, gid = 0, pos = 0, bb[4] = {0u, 0u, 0u, 0u}, tbase = 0, tx_stride = 0;
  float mx = 0.f, my = 0.f, a = 0.f, b = 0.f, c = 0.f, thr = 0.f;
  uint64_t mask = 0;
  bool big = false;
  if (live) {
    inst = order[i];
    const uint32_t view = inst / n_per_view;
    gid = inst - view * n_per_view;
    const ViewDev& VD = views[view];
    tbase = VD.tile_base;
    tx_stride = VD.tiles_x;
    const ProjSplat s = proj[inst];
    unpack_u16x2(s.p2.z, &bb[0], &bb[1]);
    unpack_u16x2(s.p2.w, &bb[2], &bb[3]);
    mx = s.p0.x;
    my = s.p0.y;
    a = 2.0f * s.p0.z;
    b = s.p0.w;
    c = 2.0f * s.p1.x;
    thr = s.p2.y;
    mask = (uint64_t)__float_as_uint(s.p3.x) | ((uint64_t)__float_as_uint(s.p3.y) << 32);
    big = (bb[3] - bb[1]) * (bb[2] - bb[0]) > 64u;
    pos = offs[i];
  }
  bool over = false;
  // large bboxes: one wave per instance, direct coalesced runs
  uint64_t bigm = __ballot(big);
  while (bigm) {
    const int j = __builtin_ctzll(bigm);
    bigm &= bigm - 1;
    uint32_t jb[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) jb[e] = (uint32_t)__shfl((int)bb[e], j, 64);
    const float jx = __shfl(mx, j, 64), jy = __shfl(my, j, 64), ja = __shfl(a, j, 64), jbb = __shfl(b, j, 64),
                jc = __shfl(c, j, 64), jt = __shfl(thr, j, 64);
    const uint32_t jgid = (uint32_t)__shfl((int)gid, j, 64), jpos = (uint32_t)__shfl((int)pos, j, 64);
    const uint32_t jtb = (uint32_t)__shfl((int)tbase, j, 64), jts = (uint32_t)__shfl((int)tx_stride, j, 64);
    const uint32_t bw = jb[2] - jb[0], nb = (jb[3] - jb[1]) * bw;
    const float inv = 1.0f / (float)bw;
    uint32_t run = jpos;
    for (uint32_t k0 = 0; k0 < nb; k0 += 64) {
      const uint32_t k = k0 + lane;
      bool hit = false;
      uint32_t row = 0, col = 0;
      if (k < nb) {
        bbox_rc(k, bw, inv, &row, &col);
        hit = tile_contributes(jb[0] + col, jb[1] + row, tw, jx, jy, ja, jbb, jc, jt);
      }
      const uint64_t m = __ballot(hit);
      if (hit) {
        const uint32_t p = run + (uint32_t)__popcll(m & lt);
        if (p < cap) {
          okeys[p] = jtb + (jb[0] + col) + (jb[1] + row) * jts;
          ovals[p] = jgid;
        } else {
          over = true;
        }
      }
      run += (uint32_t)__popcll(m);
    }
  }
  // small bboxes: stage the block's range in windows, write out coalesced
  __syncthreads();
  const uint32_t base = s_base, total = s_total;
  const uint32_t bw = bb[2] - bb[0];
  const float inv = 1.0f / (float)(bw ? bw : 1u);
  for (uint32_t w0 = 0; w0 < total; w0 += kEmitStage) {
    for (uint32_t e = threadIdx.x; e < kEmitStage; e += blockDim.x) sv[e] = kEmptySlot;
    __syncthreads();
    if (live && !big) {
      const uint32_t lo = base + w0, hi = lo + kEmitStage;
      uint64_t m = mask;
      uint32_t p = pos;
      while (m && p < hi) {
        const uint32_t k = (uint32_t)__builtin_ctzll(m);
        m &= m - 1;
        if (p >= lo) {
          uint32_t row, col;
          bbox_rc(k, bw, inv, &row, &col);
          sk[p - lo] = tbase + (bb[0] + col) + (bb[1] + row) * tx_stride;
          sv[p - lo] = gid;
        }
        ++p;
      }
    }
    __syncthreads();
    const uint32_t wn = min(kEmitStage, total - w0);
    for (uint32_t e = threadIdx.x; e < wn; e += blockDim.x) {
      const uint32_t val = sv[e];
      if (val == kEmptySlot) continue;
      const uint32_t p = base + w0 + e;
      if (p < cap) {
        okeys[p] = sk[e];
        ovals[p] = val;
      } else {
        over = true;
      }
    }
    __syncthreads();
  }
  if (over) {
    counters[kCntOverflow] = 1u;
    counters[kCntStepValid] = 0u;
  }
}

// get_tile_offsets.wgsl:12-30 over sorted keys: ranges[2t] = start, ranges[2t+1] = end.
// Each thread scans 16 consecutive keys (vector loads) and writes the range
// boundaries that fall inside them.
constexpr uint32_t kRangeKeys = 16;
__global__ __launch_bounds__(256) void k_tile_ranges(const uint32_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ n_dev, uint32_t n_host,
                                                     uint32_t cap, uint32_t num_tiles,
                                                     uint32_t* __restrict__ ranges) {
  const uint32_t m = dev_count(n_dev, n_host, cap);
  const uint32_t i0 = (blockIdx.x * blockDim.x + threadIdx.x) * kRangeKeys;
  if (i0 >= m) return;
  uint32_t k[kRangeKeys];
  if (i0 + kRangeKeys <= m) {
    const uint4* src = reinterpret_cast<const uint4*>(keys + i0);
#pragma unroll
    for (uint32_t q = 0; q < kRangeKeys / 4; ++q) {
      const uint4 x = src[q];
      k[4 * q] = x.x;
      k[4 * q + 1] = x.y;
      k[4 * q + 2] = x.z;
      k[4 * q + 3] = x.w;
    }
  } else {
#pragma unroll
    for (uint32_t j = 0; j < kRangeKeys; ++j) k[j] = i0 + j < m ? keys[i0 + j] : 0xffffffffu;
  }
  uint32_t prev = i0 == 0 ? 0xffffffffu : keys[i0 - 1];
  const uint32_t after = i0 + kRangeKeys < m ? keys[i0 + kRangeKeys] : 0xffffffffu;
#pragma unroll
  for (uint32_t j = 0; j < kRangeKeys; ++j) {
    const uint32_t i = i0 + j;
    if (i >= m) break;
    const uint32_t t = k[j];
    const uint32_t next = i + 1 < m ? (j + 1 < kRangeKeys ? k[j + 1] : after) : 0xffffffffu;
    if (t < num_tiles) {
      if (t != prev) ranges[2 * t] = i;
      if (t != next) ranges[2 * t + 1] = i + 1;
    }
    prev = t;
  }
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

// Step epilogue: count the step as skipped when it did not update, and keep
// a capacity overflow (and the intersection count it needed) in the sticky
// counters until the host has re-sized.
__global__ void k_step_end(uint32_t* __restrict__ counters) {
  if (threadIdx.x != 0) return;
  if (counters[kCntStepValid] == 0u) counters[kCntSkipped] += 1u;
  if (counters[kCntOverflow] != 0u) {
    counters[kCntOverflowSticky] = 1u;
    counters[kCntIsectMaxSticky] = max(counters[kCntIsectMaxSticky], counters[kCntIsectMax]);
  }
}

// ------------------------------------------------------------------ host side
template <int TILE = kSortTile>
static uint32_t grid_for(uint64_t cap) {
  uint64_t tiles = (cap + TILE - 1) / TILE;
  if (tiles < 1) tiles = 1;
  return (uint32_t)(tiles < 2048 ? tiles : 2048);
}

// Stable sort of (key, value) pairs on key bits [0, key_bits).
// The count is read from n_dev when non-null (clamped to cap).
// Pass 0 reads (kin, vin) and writes (kb, vb); later passes ping-pong
// between (kb, vb) and (kc, vc).  (kc, vc) may alias (kin, vin) when the input
// may be clobbered.  vin == nullptr means values are the input indices.  When
// gout is non-null the last pass also writes gout[i] = gsrc[value[i]]; with
// count_ib > 0 the values instead carry a count in bits [count_ib, 32) (all
// ones: look it up), which the last pass splits off: value[i] & (2^count_ib -
// 1) goes to the value output, the count (or gsrc[value]) to gout.
// Returns 1 if the result is in (kb, vb), 2 if in (kc, vc), 0 if key_bits <= 0
// (nothing done; the input order is already sorted).
// scratch: radix_scratch_words(cap) uint32.
// (the larger of the two tile sizes' grids: scratch for either sort)
uint32_t radix_grid(uint64_t cap) { return grid_for<kRadixTilePacked>(cap); }
uint64_t radix_scratch_words(uint64_t cap) { return 256ull * radix_grid(cap) + 512; }

int radix_sort_pairs(hipStream_t s, const uint32_t* kin, const uint32_t* vin, uint32_t* kb, uint32_t* vb,
                     uint32_t* kc, uint32_t* vc, const uint32_t* n_dev, uint32_t n_host, uint32_t cap, int key_bits,
                     uint32_t* scratch, const uint32_t* gsrc, uint32_t* gout, int count_ib) {
  if (key_bits <= 0) return 0;
  const uint32_t G = grid_for<kRadixTile>(cap);
  uint32_t* hist = scratch;
  uint32_t* tot = scratch + 256 * (size_t)G;
  const int passes = (key_bits + 7) / 8;
  const uint32_t* sk = kin;
  const uint32_t* sv = vin;
  int where = 0;
  for (int p = 0; p < passes; ++p) {
    uint32_t* dk = (p % 2 == 0) ? kb : kc;
    uint32_t* dv = (p % 2 == 0) ? vb : vc;
    const bool last = p + 1 == passes;
    k_radix_hist<kRadixTile><<<G, kSortThreads, 0, s>>>(sk, n_dev, n_host, cap, 8 * p, hist);
    k_radix_spine<<<256, kSortThreads, 0, s>>>(hist, G, tot);
    k_radix_scatter<kScatterPairs><<<G, kSortThreads, 0, s>>>(sk, sv, dk, dv, n_dev, n_host, cap, 8 * p, hist, tot,
                                                              last ? gsrc : nullptr, last ? gout : nullptr,
                                                              last && gout ? count_ib : 0, nullptr);
    sk = dk;
    sv = dv;
    where = (p % 2 == 0) ? 1 : 2;
  }
  return where;
}

// Stable sort of (key, value) pairs with keys of key_bits <= 16 and values
// below 2^value_bits, result in (kout, vout); `mid` (cap words) holds the
// packed words between the two passes.  Falls back to radix_sort_pairs
// (result in kb/vb or kout/vout, as it reports) when the packing does not fit.
// Returns 1 if the result is in (kb, vb), 2 if in (kout, vout).
int radix_sort_pairs_packed(hipStream_t s, const uint32_t* kin, const uint32_t* vin, uint32_t* kb, uint32_t* vb,
                            uint32_t* kout, uint32_t* vout, const uint32_t* n_dev, uint32_t n_host, uint32_t cap,
                            int key_bits, int value_bits, uint32_t* scratch) {
  if (key_bits <= 8 || key_bits > 16 || key_bits - 8 + value_bits > 32)
    return radix_sort_pairs(s, kin, vin, kb, vb, kout, vout, n_dev, n_host, cap, key_bits, scratch, nullptr,
                            nullptr, 0);
  const uint32_t G = grid_for<kRadixTilePacked>(cap);
  uint32_t* hist = scratch;
  uint32_t* tot_a = scratch + 256 * (size_t)G;
  uint32_t* tot_b = tot_a + 256;
  uint32_t* mid = kb;
  k_radix_hist<kRadixTilePacked><<<G, kSortThreads, 0, s>>>(kin, n_dev, n_host, cap, 0, hist);
  k_radix_spine<<<256, kSortThreads, 0, s>>>(hist, G, tot_a);
  k_radix_scatter<kScatterPackOut><<<G, kSortThreads, 0, s>>>(kin, vin, mid, nullptr, n_dev, n_host, cap, 0, hist,
                                                              tot_a, nullptr, nullptr, value_bits, nullptr);
  k_radix_hist<kRadixTilePacked><<<G, kSortThreads, 0, s>>>(mid, n_dev, n_host, cap, value_bits, hist);
  k_radix_spine<<<256, kSortThreads, 0, s>>>(hist, G, tot_b);
  k_radix_scatter<kScatterPackIn><<<G, kSortThreads, 0, s>>>(mid, nullptr, kout, vout, n_dev, n_host, cap,
                                                             value_bits, hist, tot_b, nullptr, nullptr, value_bits,
                                                             tot_a);
  return 2;
}

// Exclusive (or inclusive) scan of in[gather[i]] (gather may be null) for
// i < n; the grand total is written to *total_dev when non-null.
// scratch: grid_for(cap) uint32 (<= 2048).
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
  const uint64_t threads = ((uint64_t)cap + kRangeKeys - 1) / kRangeKeys;
  k_tile_ranges<<<(uint32_t)((threads + 255) / 256), 256, 0, s>>>(keys, counters + kCntIsect, 0, cap, num_tiles,
                                                                   ranges);
}

// Same over a list of n keys (a host count).
void launch_tile_ranges_n(hipStream_t s, const uint32_t* keys, uint32_t n, uint32_t num_tiles, uint32_t* ranges) {
  if (n == 0) return;
  const uint64_t threads = ((uint64_t)n + kRangeKeys - 1) / kRangeKeys;
  k_tile_ranges<<<(uint32_t)((threads + 255) / 256), 256, 0, s>>>(keys, nullptr, n, n, num_tiles, ranges);
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
