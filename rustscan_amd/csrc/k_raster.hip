// Per-tile alpha compositing, forward (rasterize.wgsl:21-99, K7) and backward
// (rasterize_backwards.wgsl:24-169, K8).
//
// One workgroup composites one TWxTW tile with one thread per pixel; each
// 64-lane wave owns an 8x8 pixel block of the tile, so the lanes of a wave
// are spatially compact (a splat's footprint covers fewer partially-active
// waves than with row strips).  The tile's depth-sorted splats are staged
// through LDS in batches of BATCH.  Work items are (view, tile) pairs
// numbered by global tile id: either one workgroup per tile (grid launch) or
// a persistent grid pulling tiles from an atomic counter, which is how uneven
// views share one launch (MULTI: the tile's view is looked up per item).
//
// Forward changes versus the reference that do not change any result:
//  * a workgroup stops loading batches once every pixel is done
//    (__syncthreads_count), the reference keeps looping over its range;
//  * exp is skipped when sigma > log(255*opacity) + 1e-3: there alpha < 1/255
//    whatever exp returns, so the splat is rejected either way;
//  * visibility flags are gathered per batch in LDS and written once per splat.
// sigma is evaluated as ((a/2*dx)*dx + (b*dx)*dy) + (c/2*dy)*dy, which equals
// K7's 0.5*(a*dx*dx + 2*b*dx*dy + c*dy*dy) bit for bit (scaling by powers of
// two commutes with rounding); the backward uses K8's own grouping.
//
// Backward: per splat the 64 lanes of a wave reduce their 12 contributions
// with a transposed butterfly (permlane32 swap, permlane16 swap, then four
// DPP row shifts), which leaves the 12 totals in lanes 15/31/47/63 (the
// value-to-lane map of the butterfly is measured once per kernel by reducing
// tags); those lanes add them into an LDS accumulator for the batch.  After
// the batch the totals go to the per-instance accumulator with global
// atomics, one set per tile and splat, issued so that the lanes of one
// instruction hit one 64-byte line.
#include "rgs_device.h"

namespace rgs {

// A compositing workgroup covers at most 16x16 pixels: a 32x32 binning tile
// is composited by four workgroups (its 16x16 sub-tiles), each walking the
// tile's whole depth-sorted list and culling it per 8x8 block.  Work items
// are (tile, sub-tile) pairs, the sub-tiles of a tile consecutive.
template <int TW>
struct RasterShape {
  static constexpr int RW = TW > 16 ? 16 : TW;  // workgroup width in pixels
  static constexpr int SUB = TW / RW;           // sub-tiles per tile side
  static constexpr int NT = RW * RW;            // threads per workgroup
};

__device__ __forceinline__ bool next_item(const RasterArgs& A, uint32_t items, uint32_t* s_item, uint32_t& iter,
                                          uint32_t* item) {
  if (A.queue) {
    if (threadIdx.x == 0) *s_item = atomicAdd(A.queue, 1u);
    __syncthreads();
    *item = *s_item;
    __syncthreads();
  } else {
    if (iter++) return false;
    *item = blockIdx.x;
  }
  return *item < items;
}

template <bool MULTI>
__device__ __forceinline__ int view_of(const RasterArgs& A, uint32_t tile) {
  int v = A.v0;
  if (MULTI)
    for (int i = A.v0 + 1; i < A.v0 + A.nv; ++i)
      if (tile >= A.views[i].tile_base) v = i;
  return v;
}

// Pixel of thread `tid` in a TWxTW tile: wave w covers the 8x8 block
// (w % (TW/8), w / (TW/8)); lane l is pixel (l % 8, l / 8) of that block.
template <int TW>
__device__ __forceinline__ void pixel_in_tile(int tid, uint32_t* ox, uint32_t* oy) {
  constexpr int BPR = TW / 8;
  const int w = tid >> 6, l = tid & 63;
  *ox = (uint32_t)((w % BPR) * 8 + (l & 7));
  *oy = (uint32_t)((w / BPR) * 8 + (l >> 3));
}

// Wave vote straight from the lane-mask builtin: applied to a compare, the
// compare's own lane mask is the vote.  Every lane of the wave must be active.
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

constexpr float kInf = __builtin_huge_valf();

// max(c, 0) for a splat colour c, returned as -0.0 where c < 0 (or NaN): the
// value acts as 0 in every product and sum, and its sign bit carries the
// colour clamp's gradient gate (c >= 0).  c = sum + 0.5 is never -0.0.
__device__ __forceinline__ float clamp_colour(float c) { return c >= 0.0f ? c : -0.0f; }

#define RGS_DPP(v, ctrl, rm, bm) \
  __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, (v)), ctrl, rm, bm, true))

// Sum of v over the 16 lanes of each row, valid in the row's lane 15.
__device__ __forceinline__ float row_reduce_l15(float v) {
  v += RGS_DPP(v, 0x111, 0xf, 0xf);  // row_shr:1
  v += RGS_DPP(v, 0x112, 0xf, 0xf);  // row_shr:2
  v += RGS_DPP(v, 0x114, 0xf, 0xf);  // row_shr:4
  v += RGS_DPP(v, 0x118, 0xf, 0xf);  // row_shr:8
  return v;
}

// Exchanges one 32-lane half of a with the other half of b and adds: one half
// of the wave then holds pairwise sums of a's lanes, the other half b's.
// The empty asm makes the two swap results opaque: without it this compiler
// folds r[0] + r[1] into 2 * r[0].
__device__ __forceinline__ float fold32(float a, float b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, a), __builtin_bit_cast(uint32_t, b),
                                                  false, false);
  uint32_t r0 = r[0], r1 = r[1];
  asm volatile("" : "+v"(r0), "+v"(r1));
  return __builtin_bit_cast(float, r0) + __builtin_bit_cast(float, r1);
}

// Same with 16-lane rows: two rows then hold row-pair sums of a, two of b.
__device__ __forceinline__ float fold16(float a, float b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(uint32_t, a), __builtin_bit_cast(uint32_t, b),
                                                  false, false);
  uint32_t r0 = r[0], r1 = r[1];
  asm volatile("" : "+v"(r0), "+v"(r1));
  return __builtin_bit_cast(float, r0) + __builtin_bit_cast(float, r1);
}

// Wave-wide sums of val[0..11]: after the two folds each of the three
// registers holds partial sums of four different values, one per 16-lane
// row, and the row reductions leave the twelve totals in lanes 15/31/47/63
// (three per lane).  Which value lands in which (lane, register) is fixed by
// the swap pattern; wave_reduce12_slots() measures it once per kernel by
// reducing the tags 1..12.  Every lane of the wave must be active.
__device__ __forceinline__ void wave_reduce12(const float val[12], float out[3]) {
  float h[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) h[i] = fold32(val[i], val[i + 6]);
#pragma unroll
  for (int a = 0; a < 3; ++a) out[a] = row_reduce_l15(fold16(h[a], h[a + 3]));
}

// Index of the value whose total wave_reduce12 leaves in out[a] of this lane
// (meaningful in lanes 15/31/47/63): tag k+1 on every lane sums to 64(k+1).
__device__ __forceinline__ void wave_reduce12_slots(int slot[3]) {
  float tag[12], t[3];
#pragma unroll
  for (int k = 0; k < 12; ++k) tag[k] = (float)(k + 1);
  wave_reduce12(tag, t);
#pragma unroll
  for (int a = 0; a < 3; ++a) slot[a] = (int)(t[a] * (1.0f / 64.0f) + 0.5f) - 1;
}


// Publishes the block masks of staged splats j = j0 + tid as 64-bit words
// words[b][j / 64] (bit j % 64: splat j can reach block b).  Wave-uniform:
// every lane of a wave calls it.
template <int TW, int BATCH>
__device__ __forceinline__ void publish_block_masks(uint64_t (*words)[BATCH / 64], uint32_t j0, uint32_t bits) {
  constexpr int NB = (TW / 8) * (TW / 8);
  const uint32_t wj = j0 + (threadIdx.x & ~63u);  // first splat of this wave
  if (wj >= (uint32_t)BATCH) return;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const uint64_t m = __builtin_amdgcn_ballot_w64((bits >> b) & 1u);
    if ((threadIdx.x & 63u) == 0) words[b][wj / 64] = m;
  }
}

// Position of this thread's item among the flagged items of the workgroup
// (in thread order); *total = their number.  Every thread must call it.
// REUSE: s_wc may be written again before every thread has passed another
// barrier, so the function ends with one.
template <int NT, bool REUSE>
__device__ __forceinline__ uint32_t block_rank(bool flag, uint32_t* s_wc, uint32_t* total) {
  const uint64_t m = __builtin_amdgcn_ballot_w64(flag);
  const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t below = (uint32_t)__popcll(m & (lane ? (~0ull >> (64u - lane)) : 0ull));
  if (NT == 64) {
    *total = (uint32_t)__popcll(m);
    return below;
  }
  if (lane == 0) s_wc[w] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) {
    const uint32_t c = s_wc[i];
    off += (uint32_t)i < w ? c : 0u;
    tot += c;
  }
  if (REUSE) __syncthreads();
  *total = tot;
  return off + below;
}

// One staged splat in LDS: three 16-byte words at one address, so the
// per-splat reads of a wave share a single address register.
struct alignas(16) StagedSplat {
  float4 p0, p1, p2;
};

// Calls body(rec[j].p0, rec[j].p1, rec[j].p2, j) for the staged splats j < cnt whose bit
// is set in this wave's words, in increasing j, while alive() holds for some
// lane.  Two splats per trip, the next record fetched one splat ahead into
// alternating registers.
template <int BATCH, typename Alive, typename Body>
__device__ __forceinline__ void for_each_reaching(const uint64_t* words, uint32_t cnt, const StagedSplat* rec,
                                                  Alive&& alive, Body&& body) {
  const uint32_t nw = (cnt + 63) / 64;
  for (uint32_t wi = 0; wi < nw; ++wi) {
    const uint64_t wv = words[wi];
    uint64_t m = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(wv >> 32)) << 32) |
                 (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)wv);
    if (m == 0) continue;
    if (!wave_any(alive())) return;
    uint32_t ja = wi * 64 + (uint32_t)__builtin_ctzll(m);
    m &= m - 1;
    float4 p0 = rec[ja].p0, p1 = rec[ja].p1, p2 = rec[ja].p2;
    while (true) {
      const bool hb = m != 0;
      const uint32_t jb = hb ? wi * 64 + (uint32_t)__builtin_ctzll(m) : ja;
      m &= m - 1;
      const float4 q0 = rec[jb].p0, q1 = rec[jb].p1, q2 = rec[jb].p2;
      body(p0, p1, p2, ja);
      if (!hb || !wave_any(alive())) break;
      const bool ha = m != 0;
      ja = ha ? wi * 64 + (uint32_t)__builtin_ctzll(m) : jb;
      m &= m - 1;
      p0 = rec[ja].p0;
      p1 = rec[ja].p1;
      p2 = rec[ja].p2;
      body(q0, q1, q2, jb);
      if (!ha || !wave_any(alive())) break;
    }
  }
}

template <int TW, int BATCH, bool MULTI, bool COUNT>
__global__ __launch_bounds__(RasterShape<TW>::NT) void k_raster_fwd(RasterArgs A) {
  constexpr int RW = RasterShape<TW>::RW, SUB = RasterShape<TW>::SUB, NT = RasterShape<TW>::NT;
  // p0 = (x, y, a/2, b), p1 = (c/2, opacity, max(r,0), max(g,0)),
  // p2 = (max(b,0), exp-skip bound, blended flag (bits), unused)
  __shared__ StagedSplat rec[BATCH];
  __shared__ uint32_t sg[BATCH];
  __shared__ uint64_t swords[(RW / 8) * (RW / 8)][BATCH / 64];
  __shared__ uint32_t s_item;
  const int tid = threadIdx.x;
  uint32_t ox, oy;
  pixel_in_tile<RW>(tid, &ox, &oy);
  uint32_t iter = 0, item;
  while (next_item(A, A.tile_count * (SUB * SUB), &s_item, iter, &item)) {
    const uint32_t tile = A.tile_first + item / (SUB * SUB), sub = item % (SUB * SUB);
    const int v = view_of<MULTI>(A, tile);
    const ViewDev& V = A.views[v];
    const uint32_t local = tile - V.tile_base;
    const uint32_t tx = local % V.tiles_x, ty = local / V.tiles_x;
    const uint32_t tox = tx * TW + (sub % SUB) * RW, toy = ty * TW + (sub / SUB) * RW;
    const uint32_t px = tox + ox, py = toy + oy;
    const bool inside = px < V.W && py < V.H;
    const float pcx = (float)px + 0.5f, pcy = (float)py + 0.5f;
    const uint32_t rs = A.ranges[2 * tile], re = A.ranges[2 * tile + 1];
    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    // 0 while the pixel is being composited, +inf once it is done: the
    // per-splat test is then one add and one compare (sigma + pen <= bound).
    float pen = inside ? 0.0f : kInf;
    uint32_t n_eval = 0, n_blend = 0, n_iter = 0, n_iter_work = 0;
    for (uint32_t b = rs; b < re; b += BATCH) {
      if (__syncthreads_count(pen == 0.0f) == 0) break;
      const uint32_t cnt = min((uint32_t)BATCH, re - b);
      for (uint32_t j0 = 0; j0 < (uint32_t)BATCH; j0 += NT) {
        const uint32_t j = j0 + tid;
        uint32_t bits = 0;
        if (j < cnt) {
          const uint32_t gid = A.isect_vals[b + j];
          const ProjSplat q = A.proj[(uint32_t)v * A.n + gid];
          rec[j].p0 = q.p0;
          rec[j].p1 = make_float4(q.p1.x, q.p1.y, fmaxf_(q.p1.z, 0.0f), fmaxf_(q.p1.w, 0.0f));
          rec[j].p2 = make_float4(fmaxf_(q.p2.x, 0.0f), q.p2.y + kSkipMargin, 0.0f, 0.0f);
          sg[j] = gid;
          bits = A.cull ? block_touch_mask<RW>(q.p0, q.p1.x, q.p2.y + kSkipMargin, tox, toy)
                        : (1u << ((RW / 8) * (RW / 8))) - 1u;
        }
        publish_block_masks<RW, BATCH>(swords, j0, bits);
      }
      __syncthreads();
      // Wave-uniform loop: every lane walks the splats that can reach its
      // wave's 8x8 block until the whole wave is done; a finished lane only
      // stops updating its pixel.
      auto composite = [&](const float4 a0, const float4 a1, const float4 a2, uint32_t j) {
        const float dx = a0.x - pcx, dy = a0.y - pcy;
        const float sigma = ((a0.z * dx) * dx + (a0.w * dx) * dy) + (a1.x * dy) * dy;
        if (COUNT) {
          n_eval += pen == 0.0f ? 1u : 0u;
          ++n_iter;
        }
        const bool need = sigma + pen <= a2.y;
        if (!wave_any(need)) return;
        if (COUNT) ++n_iter_work;
        // v_min_f32: equal to the reference's min for every non-NaN product;
        // a NaN (only on lanes rejected by the sigma test) gives 0.999
        const float alpha = __builtin_fminf(RGS_F(0.999), a1.y * rgs_exp_core_fast(-sigma));
        const bool acc = need && sigma >= 0.0f && alpha >= RGS_F(1.0 / 255.0);
        if (acc) {
          if (COUNT) ++n_blend;
          const float vis = alpha * T;
          cr = cr + a1.z * vis;
          cg = cg + a1.w * vis;
          cb = cb + a2.x * vis;
          T = T * (1.0f - alpha);
          // flag in the record's third word; other waves read only its x, y
          reinterpret_cast<uint32_t*>(&rec[j].p2)[2] = 1u;
          pen = T < RGS_F(1e-4) ? kInf : pen;
        }
      };
      for_each_reaching<BATCH>(swords[tid >> 6], cnt, rec, [&] { return pen == 0.0f; }, composite);
      __syncthreads();
      if (A.visible)
        for (uint32_t j = tid; j < cnt; j += NT)
          if (__float_as_uint(rec[j].p2.z)) A.visible[sg[j]] = 1.0f;
      if (A.vis_inst)
        for (uint32_t j = tid; j < cnt; j += NT)
          if (__float_as_uint(rec[j].p2.z)) A.vis_inst[(size_t)v * A.n + sg[j]] = 1u;
    }
    float l1 = 0.0f;
    if (inside) {
      const size_t pix = (size_t)py * V.W + px;
      const float o0 = cr + T * V.bg[0], o1 = cg + T * V.bg[1], o2 = cb + T * V.bg[2];
      reinterpret_cast<float4*>(V.img)[pix] = make_float4(o0, o1, o2, 1.0f - T);
      if (A.loss && V.target) {
        const float* t = V.target + pix * 3;
        l1 = (fabsf(o0 - t[0]) + fabsf(o1 - t[1])) + fabsf(o2 - t[2]);
      }
    }
    if (A.loss) {
      const float s = wave_sum(l1);
      if ((tid & 63) == 0 && s != 0.0f) atomicAdd(&A.loss[v], s);
    }
    if (COUNT) {
      unsigned long long e = n_eval, bl = n_blend;
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) {
        e += __shfl_xor(e, o, 64);
        bl += __shfl_xor(bl, o, 64);
      }
      if ((tid & 63) == 0) {
        atomicAdd(&A.frag[0], e);
        atomicAdd(&A.frag[1], bl);
        atomicAdd(&A.frag[4], (unsigned long long)n_iter);
        atomicAdd(&A.frag[5], (unsigned long long)n_iter_work);
      }
    }
  }
}

template <int TW, int BATCH, bool REFINE, bool MULTI, bool COUNT>
__global__ __launch_bounds__(RasterShape<TW>::NT) void k_raster_bwd(RasterArgs A) {
  constexpr int RW = RasterShape<TW>::RW, SUB = RasterShape<TW>::SUB, NT = RasterShape<TW>::NT;
  // Accumulator slots per splat; the layout is that of the per-instance
  // accumulator: 0-1 v_xy, 2-4 v_conic, 5-7 v_rgb, 8 opacity term, 9 refine,
  // 10-11 |v_xy|, 12 coverage count.
  constexpr int NA = 13;
  constexpr int SA = NA;  // LDS stride per splat (13 floats: less LDS per workgroup, more resident waves)
  // p0 = (x, y, a/2, b), p1 = (c/2, opacity, r, g), p2 = (b, exp-skip bound,
  // c, instance id (bits)); colours as stored by clamp_colour(): max(col, 0),
  // with the sign bit set where col < 0 (the colour clamp's gradient gate)
  __shared__ StagedSplat rec[BATCH];
  __shared__ float sacc[BATCH * SA];  // all zero between batches
  __shared__ uint32_t sfl[BATCH];     // staged splats with contributions
  __shared__ uint32_t s_wc[NT / 64];
  // LDS accumulator position p of a splat row -> its slot in the per-instance
  // accumulator: lane 16g+15 adds its three totals at positions 3g..3g+2
  // (one address, three immediate offsets); s_pos_cov is the position of
  // the coverage count
  __shared__ uint32_t s_perm[SA];
  __shared__ uint32_t s_pos_cov;
  __shared__ uint64_t swords[(RW / 8) * (RW / 8)][BATCH / 64];
  __shared__ uint32_t s_item;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  uint32_t ox, oy;
  pixel_in_tile<RW>(tid, &ox, &oy);
  {
    int slot[3];
    wave_reduce12_slots(slot);
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (!REFINE && slot[a] == 9) slot[a] = 12;  // coverage count rides in the refine slot
    if (tid < 64 && (lane & 15) == 15)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        s_perm[3 * (lane >> 4) + a] = (uint32_t)slot[a];
        if (slot[a] == 12) s_pos_cov = 3u * (uint32_t)(lane >> 4) + (uint32_t)a;
      }
    if (tid == 0) {
      s_perm[12] = 12u;  // REFINE: the coverage count (from lane 63), else never written
      if (REFINE) s_pos_cov = 12u;
    }
  }
  for (uint32_t k = threadIdx.x; k < (uint32_t)(BATCH * SA); k += NT) sacc[k] = 0.0f;
  __syncthreads();
  const uint32_t pos_cov = s_pos_cov;
  uint32_t iter = 0, item;
  while (next_item(A, A.tile_count * (SUB * SUB), &s_item, iter, &item)) {
    const uint32_t tile = A.tile_first + item / (SUB * SUB), sub = item % (SUB * SUB);
    const int v = view_of<MULTI>(A, tile);
    const ViewDev& V = A.views[v];
    const uint32_t local = tile - V.tile_base;
    const uint32_t tx = local % V.tiles_x, ty = local / V.tiles_x;
    const uint32_t tox = tx * TW + (sub % SUB) * RW, toy = ty * TW + (sub / SUB) * RW;
    const uint32_t px = tox + ox, py = toy + oy;
    const bool inside = px < V.W && py < V.H;
    const float pcx = (float)px + 0.5f, pcy = (float)py + 0.5f;
    const uint32_t rs = A.ranges[2 * tile], re = A.ranges[2 * tile + 1];
    // rasterize_backwards.wgsl:46-60
    float vo[4] = {0.f, 0.f, 0.f, 0.f}, rgbf[4] = {0.f, 0.f, 0.f, 0.f};
    if (inside) {
      const size_t pix = (size_t)py * V.W + px;
      const float4 fc = reinterpret_cast<const float4*>(V.img)[pix];
      const float Tf = 1.0f - fc.w;
      rgbf[0] = fc.x - Tf * V.bg[0];
      rgbf[1] = fc.y - Tf * V.bg[1];
      rgbf[2] = fc.z - Tf * V.bg[2];
      rgbf[3] = fc.w;
      float g[4];
      if (V.vout) {
        const float4 go = reinterpret_cast<const float4*>(V.vout)[pix];
        g[0] = go.x; g[1] = go.y; g[2] = go.z; g[3] = go.w;
      } else {
        // L1 gradient (engine/loss.rs:144-196): sign(pred - target) * weight / (3HW); alpha gets none
        const float* t = V.target + pix * 3;
        const float d[3] = {fc.x - t[0], fc.y - t[1], fc.z - t[2]};
#pragma unroll
        for (int c = 0; c < 3; ++c) g[c] = d[c] > 0.0f ? V.loss_scale : (d[c] < 0.0f ? -V.loss_scale : 0.0f);
        g[3] = 0.0f;
      }
      vo[0] = g[0]; vo[1] = g[1]; vo[2] = g[2];
      vo[3] = (g[3] - ((V.bg[0] * g[0] + V.bg[1] * g[1]) + V.bg[2] * g[2])) * Tf;
    }
    float T = 1.0f;
    // v_alpha of a splat (rasterize_backwards.wgsl:115-127) is
    //   sum_c vo[c] * (T*clamp(col_c) + (po[c] - rgb[c]) / (1 - alpha)) + vo[3] / (1 - alpha)
    // with po the colour composited up to and including it; it is evaluated as
    //   T*Q + (P - R) / (1 - alpha),  Q = sum_c vo[c]*clamp(col_c),
    //   P = sum_c vo[c]*po[c] (running), R = sum_c vo[c]*rgb[c] - vo[3].
    const float R = ((vo[0] * rgbf[0] + vo[1] * rgbf[1]) + vo[2] * rgbf[2]) - vo[3];
    float P = 0.0f;
    float pen = inside ? 0.0f : kInf;  // +inf once the pixel is done
    uint32_t n_eval = 0, n_contrib = 0, n_iter = 0, n_iter_work = 0;
    const float Wf = (float)V.W, Hf = (float)V.H;
    for (uint32_t b = rs; b < re; b += BATCH) {
      if (__syncthreads_count(pen == 0.0f) == 0) break;
      const uint32_t cnt = min((uint32_t)BATCH, re - b);
      for (uint32_t j0 = 0; j0 < (uint32_t)BATCH; j0 += NT) {
        const uint32_t j = j0 + tid;
        uint32_t bits = 0;
        if (j < cnt) {
          const uint32_t inst = (uint32_t)v * A.n + A.isect_vals[b + j];
          const ProjSplat q = A.proj[inst];
          rec[j].p0 = q.p0;
          rec[j].p1 = make_float4(q.p1.x, q.p1.y, clamp_colour(q.p1.z), clamp_colour(q.p1.w));
          rec[j].p2 = make_float4(clamp_colour(q.p2.x), q.p2.y + kSkipMargin, 2.0f * q.p1.x, __uint_as_float(inst));
          bits = A.cull ? block_touch_mask<RW>(q.p0, q.p1.x, q.p2.y + kSkipMargin, tox, toy)
                        : (1u << ((RW / 8) * (RW / 8))) - 1u;
        }
        publish_block_masks<RW, BATCH>(swords, j0, bits);
      }
      __syncthreads();
      // Same walk as the forward: the splats that can reach the wave's block.
      auto accumulate = [&](const float4 a0, const float4 a1, const float4 a2, uint32_t j) {
        const float dx = a0.x - pcx, dy = a0.y - pcy;
        const float sigma = ((a0.z * dx) * dx + (a1.x * dy) * dy) + (a0.w * dx) * dy;
        if (COUNT) {
          n_eval += pen == 0.0f ? 1u : 0u;
          ++n_iter;
        }
        // The decisions are combined as lane masks: each compare's ballot is
        // its own result mask, the masks combine on the scalar unit, and the
        // selects below take them back per lane (inverse ballot) without
        // materialising the booleans in vector registers.
        const uint64_t m_need = __builtin_amdgcn_ballot_w64(sigma + pen <= a2.y);
        if (m_need == 0) return;
        const float gauss = rgs_exp_core_fast(-sigma);
        const float alpha = __builtin_fminf(RGS_F(0.999), a1.y * gauss);  // as in the forward
        const uint64_t m_pass = m_need & ~__builtin_amdgcn_ballot_w64(sigma < 0.0f) &
                                ~__builtin_amdgcn_ballot_w64(alpha < RGS_F(1.0 / 255.0));
        const float nT = T * (1.0f - alpha);
        const uint64_t m_stop = m_pass & __builtin_amdgcn_ballot_w64(nT <= RGS_F(1e-4));
        const uint64_t m = m_pass & ~m_stop;
        const bool contrib = __builtin_amdgcn_inverse_ballot_w64(m);
        pen = __builtin_amdgcn_inverse_ballot_w64(m_stop) ? kInf : pen;
        if (m == 0) return;
        if (COUNT) {
          n_contrib += contrib ? 1u : 0u;
          ++n_iter_work;
        }
        // The gradient terms are evaluated on every lane (no divergent
        // branch); lanes that do not contribute get vis = 0 and v_alpha = 0,
        // which zeroes every term they add and leaves P and T unchanged.
        // Their other intermediates are finite or masked by these selects.
        // The terms are fused into FMAs (they feed no compositing decision;
        // gradients are compared within a relative tolerance).  Factors that
        // are uniform over the splat are applied when the batch is flushed:
        // the colour clamp's gate on v_rgb and the 1/2 of v_conic[0] and [2].
        float val[12];
        {
#pragma clang fp contract(fast)
        const float cl[3] = {a1.z, a1.w, a2.x};
        const float vis = contrib ? alpha * T : 0.0f;
#pragma unroll
        for (int c = 0; c < 3; ++c) val[5 + c] = vis * vo[c];
        const float Q = (cl[0] * vo[0] + cl[1] * vo[1]) + cl[2] * vo[2];
        // 1/(1-alpha) with the hardware reciprocal (1 ulp); gradients are
        // compared within a relative tolerance, not bit for bit.
        const float ra = __builtin_amdgcn_rcpf(1.0f - alpha);
        P = P + vis * Q;  // po includes this splat's own colour
        const float v_alpha_all = T * Q + (P - R) * ra;
        // the alpha clamp gates the geometric terms (rasterize_backwards.wgsl:129-143)
        const bool geo = __builtin_amdgcn_inverse_ballot_w64(m & __builtin_amdgcn_ballot_w64(a1.y * gauss <= RGS_F(0.999)));
        const float v_alpha = geo ? v_alpha_all : 0.0f;
        const float v_sigma = -alpha * v_alpha;  // alpha is finite on every lane
        const float ca = 2.0f * a0.z, cb = a0.w, cc = a2.z;
        const float sdx = v_sigma * dx, sdy = v_sigma * dy;
        const float vx = ca * sdx + cb * sdy;
        const float vy = cb * sdx + cc * sdy;
        val[0] = vx;
        val[1] = vy;
        val[2] = sdx * dx;  // x 1/2 at the flush
        val[3] = sdx * dy;
        val[4] = sdy * dy;  // x 1/2 at the flush
        val[8] = v_alpha * (geo ? gauss : 0.0f);
        if (REFINE) {
          const float fa = fmaxf_(rgbf[3], RGS_F(1e-5));
          const float lx = vx * Wf, ly = vy * Hf;
          val[9] = rgs_sqrt(lx * lx + ly * ly) / fa;
        } else {
          val[9] = contrib ? 1.0f : 0.0f;  // coverage count rides in the refine slot
        }
        val[10] = fabsf(vx);
        val[11] = fabsf(vy);
        }
        T = contrib ? nT : T;
        float tot[3];
        wave_reduce12(val, tot);
        if ((lane & 15) == 15) {
          float* dst = sacc + j * SA + 3 * (lane >> 4);
#pragma unroll
          for (int a = 0; a < 3; ++a) atomicAdd(dst + a, tot[a]);
          if (REFINE && lane == 63) atomicAdd(sacc + j * SA + 12, (float)__popcll(m));
        }
      };
      for_each_reaching<BATCH>(swords[tid >> 6], cnt, rec, [&] { return pen == 0.0f; }, accumulate);
      __syncthreads();
      // Flush the splats that received contributions (the coverage count is
      // nonzero exactly then): 13 consecutive threads per splat, one float
      // each, so the atomics of a splat go to one 64-byte accumulator line.
      // Each position is cleared as it is read.
      uint32_t nf = 0;
      for (uint32_t j0 = 0; j0 < (uint32_t)BATCH; j0 += NT) {
        const uint32_t j = j0 + tid;
        const bool has = j < cnt && sacc[j * SA + pos_cov] != 0.0f;
        uint32_t tot;
        // one call per batch when BATCH <= NT: the next write of s_wc comes
        // after the barriers of the next batch
        const uint32_t r = nf + block_rank<NT, (BATCH > NT)>(has, s_wc, &tot);
        if (has) sfl[r] = j;
        nf += tot;
      }
      __syncthreads();
      for (uint32_t e = tid; e < nf * SA; e += NT) {
        const uint32_t j = sfl[e / SA], pk = e % SA, k = s_perm[pk];
        float x = sacc[j * SA + pk];
        sacc[j * SA + pk] = 0.0f;
        if (k == 2 || k == 4) x *= 0.5f;
        if (k >= 5 && k <= 7) {
          const float col = k == 5 ? rec[j].p1.z : (k == 6 ? rec[j].p1.w : rec[j].p2.x);
          if (__float_as_int(col) < 0) x = 0.0f;  // clamped colour: no gradient
        }
        if (x != 0.0f) atomicAdd(A.acc + (size_t)__float_as_uint(rec[j].p2.w) * kAcc + k, x);
      }
    }
    if (COUNT) {
      unsigned long long e = n_eval, cb2 = n_contrib;
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) {
        e += __shfl_xor(e, o, 64);
        cb2 += __shfl_xor(cb2, o, 64);
      }
      if ((tid & 63) == 0) {
        atomicAdd(&A.frag[2], e);
        atomicAdd(&A.frag[3], cb2);
        atomicAdd(&A.frag[6], (unsigned long long)n_iter);
        atomicAdd(&A.frag[7], (unsigned long long)n_iter_work);
      }
    }
  }
}

// ------------------------------------------------------------------ launch
// Variants: the per-view path (one view per launch) skips the view lookup;
// fragment counting and the refine statistic are instrumentation variants
// built only for the multi-view form.
template <int TW, int BATCH>
static void fwd_launch(hipStream_t s, const RasterArgs& A, uint32_t grid) {
  if (A.frag)
    k_raster_fwd<TW, BATCH, true, true><<<grid, RasterShape<TW>::NT, 0, s>>>(A);
  else if (A.nv > 1)
    k_raster_fwd<TW, BATCH, true, false><<<grid, RasterShape<TW>::NT, 0, s>>>(A);
  else
    k_raster_fwd<TW, BATCH, false, false><<<grid, RasterShape<TW>::NT, 0, s>>>(A);
}
template <int TW, int BATCH>
static void bwd_launch(hipStream_t s, const RasterArgs& A, uint32_t grid, bool refine) {
  if (refine)
    k_raster_bwd<TW, BATCH, true, true, false><<<grid, RasterShape<TW>::NT, 0, s>>>(A);
  else if (A.frag)
    k_raster_bwd<TW, BATCH, false, true, true><<<grid, RasterShape<TW>::NT, 0, s>>>(A);
  else if (A.nv > 1)
    k_raster_bwd<TW, BATCH, false, true, false><<<grid, RasterShape<TW>::NT, 0, s>>>(A);
  else
    k_raster_bwd<TW, BATCH, false, false, false><<<grid, RasterShape<TW>::NT, 0, s>>>(A);
}

#define RGS_DISPATCH(TWV, BV, CALL) \
  if (tw == TWV && batch == BV) {   \
    CALL;                           \
    return 0;                       \
  }

int launch_raster_fwd(hipStream_t s, int tw, int batch, const RasterArgs& A, uint32_t grid) {
  RGS_DISPATCH(8, 128, (fwd_launch<8, 128>(s, A, grid)))
  RGS_DISPATCH(8, 256, (fwd_launch<8, 256>(s, A, grid)))
  RGS_DISPATCH(8, 512, (fwd_launch<8, 512>(s, A, grid)))
  RGS_DISPATCH(16, 128, (fwd_launch<16, 128>(s, A, grid)))
  RGS_DISPATCH(16, 256, (fwd_launch<16, 256>(s, A, grid)))
  RGS_DISPATCH(16, 512, (fwd_launch<16, 512>(s, A, grid)))
  RGS_DISPATCH(32, 128, (fwd_launch<32, 128>(s, A, grid)))
  RGS_DISPATCH(32, 256, (fwd_launch<32, 256>(s, A, grid)))
  RGS_DISPATCH(32, 512, (fwd_launch<32, 512>(s, A, grid)))
  return 1;
}

int launch_raster_bwd(hipStream_t s, int tw, int batch, const RasterArgs& A, uint32_t grid, bool refine) {
  RGS_DISPATCH(8, 128, (bwd_launch<8, 128>(s, A, grid, refine)))
  RGS_DISPATCH(8, 256, (bwd_launch<8, 256>(s, A, grid, refine)))
  RGS_DISPATCH(8, 512, (bwd_launch<8, 512>(s, A, grid, refine)))
  RGS_DISPATCH(16, 128, (bwd_launch<16, 128>(s, A, grid, refine)))
  RGS_DISPATCH(16, 256, (bwd_launch<16, 256>(s, A, grid, refine)))
  RGS_DISPATCH(16, 512, (bwd_launch<16, 512>(s, A, grid, refine)))
  RGS_DISPATCH(32, 128, (bwd_launch<32, 128>(s, A, grid, refine)))
  RGS_DISPATCH(32, 256, (bwd_launch<32, 256>(s, A, grid, refine)))
  RGS_DISPATCH(32, 512, (bwd_launch<32, 512>(s, A, grid, refine)))
  return 1;
}

// Workgroup size and work items (tile, sub-tile) of a tile width.
int raster_threads(int tw) { return tw > 16 ? 256 : tw * tw; }
uint32_t raster_items(int tw, uint32_t tiles) { return tw > 16 ? tiles * (uint32_t)((tw / 16) * (tw / 16)) : tiles; }

// Occupancy-based persistent grid size for a configuration.
int raster_persistent_grid(int tw, int batch, bool bwd, int num_cus) {
  int per_cu = 0;
  const int threads = raster_threads(tw);
#define RGS_OCC(TWV, BV)                                                                                \
  if (tw == TWV && batch == BV) {                                                                       \
    if (bwd)                                                                                            \
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_raster_bwd<TWV, BV, false, true, false>, \
                                                   threads, 0);                                         \
    else                                                                                                \
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_raster_fwd<TWV, BV, true, false>, \
                                                   threads, 0);                                         \
  }
  RGS_OCC(8, 128) RGS_OCC(8, 256) RGS_OCC(8, 512)
  RGS_OCC(16, 128) RGS_OCC(16, 256) RGS_OCC(16, 512)
  RGS_OCC(32, 128) RGS_OCC(32, 256) RGS_OCC(32, 512)
#undef RGS_OCC
  if (per_cu < 1) per_cu = 1;
  return per_cu * num_cus;
}

}  // namespace rgs
